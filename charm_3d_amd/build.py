"""Build libcharm3d.so in-tree: host C (gcc, C99) + gfx950 HIP kernels (hipcc).

    python -m charm_3d_amd.build      # or charm_3d_amd.build.build()
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
SRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(ROOT, "_build", "charm3d")
LIB = os.path.join(PKG, "libcharm3d.so")
EXE = os.path.join(PKG, "charm_3d")          # the charm_3d program (task.yaml driver)

C_SOURCES = ["charm_forest.c", "charm_geom.c", "charm_model.c", "charm_amr.c", "charm_output.c", "charm_yaml.c", "charm_conn.c",
             "charm_pointwise.c", "charm_gmesh.c"]
HIP_SOURCES = ["charm_stage_f32_ph0.hip", "charm_stage_f32_ph1.hip", "charm_stage_f64_ph0.hip",
               "charm_stage_f64_ph1.hip", "charm_stage_mix_ph0.hip", "charm_kernels.hip", "charm_gpu.hip", "charm_amr.hip",
               "charm_ns.hip", "charm_adj.hip", "charm_gmesh.hip"]
ARCH = os.environ.get("CHARM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

C_FLAGS = ["-O2", "-fPIC", "-pthread", "-std=gnu99", "-Wall", "-Wno-unused-function", "-I", INC, "-I", SRC]
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
             # packed FP32 VALU forms are kept out of the generated code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             # FP32 '/' and sqrt lower to the hardware approximations (FP32 bulk path)
             "-fno-hip-fp32-correctly-rounded-divide-sqrt",
             "-Wno-unused-result", "-I", INC, "-I", SRC]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for s in C_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        jobs.append((["gcc"] + C_FLAGS + ["-c", os.path.join(SRC, s), "-o", o], o))
    for s in HIP_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        jobs.append(([HIPCC] + HIP_FLAGS + ["-c", os.path.join(SRC, s), "-o", o], o))
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for r in ex.map(lambda j: _run(j[0]), jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout, r.stderr, file=sys.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-pthread", "-o", tmp] + [o for _, o in jobs])
    os.replace(tmp, LIB)
    # the program links the library from its own directory
    main_o = os.path.join(OBJ, "charm_main.c.o")
    _run(["gcc"] + C_FLAGS + ["-c", os.path.join(SRC, "charm_main.c"), "-o", main_o])
    _run(["gcc", "-o", EXE + ".tmp", main_o, "-L", PKG, "-lcharm3d", "-Wl,-rpath,$ORIGIN"])
    os.replace(EXE + ".tmp", EXE)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
