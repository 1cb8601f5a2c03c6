"""Writes tests/golden/klin_mesh.npz from the reference's sample
samples/klin_amr (klin.msh and task.yaml: a Mach-2 wedge channel of 800
hexahedra), read with this library's own Gmsh reader and task loader: the
connectivity arrays, each tree face's boundary slot, and the task settings the
multi-tree tests use.  Run from the repository root where the reference tree
is present (python3 tests/golden/make_klin_fixture.py)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import charm_3d_amd as CA  # noqa: E402

SAMPLE = "/root/reference/samples/klin_amr"


def main():
    task = CA.load_task(os.path.join(SAMPLE, "task.yaml"))
    conn = CA.Connectivity(os.path.join(SAMPLE, "klin.msh"))
    tree_bnd, tree_reg = conn.bind_task(task)
    cfg = task.cfg
    reg = task.region[0]
    out = dict(
        vertices=conn.vertices, tree_to_vertex=conn.tree_to_vertex, tree_bnd=tree_bnd,
        bnd_type=np.array([cfg.bnd_type[i] for i in range(6)], np.int32),
        bnd_param=np.array([[cfg.bnd_param[i][j] for j in range(16)] for i in range(6)]),
        bnd_name=np.array([task.bnd_name[i].value.decode() for i in range(6)]),
        comp_m=np.array([cfg.comp[0].m]), comp_cp=np.array([cfg.comp[0].cp[0]]),
        flux=np.int32(cfg.flux), limiter=np.int32(cfg.limiter), model=np.int32(cfg.model),
        cfl=np.float64(cfg.cfl), tau=np.float64(cfg.dt),
        region_v=np.array(list(reg.v)), region_t=np.float64(reg.t), region_p=np.float64(reg.p))
    path = os.path.join(ROOT, "tests", "golden", "klin_mesh.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, conn.num_trees, "trees", conn.num_vertices, "vertices")


if __name__ == "__main__":
    main()
