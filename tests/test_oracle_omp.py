"""The OpenMP builds of the oracle (oracle/Makefile: liboracle_omp.so for the
CPU baseline, liboracle_faithful.so with the reference's 128-species records
and loops) reproduce the serial checker bit for bit: their face-parallel
convective pass adds each leaf's contributions in the serial order.  Each
build runs in a child process (CHARM_ORACLE_LIB, 4 OpenMP threads) on a 2:1
mesh with hanging faces, boundaries and the limiter."""
import os
import subprocess
import sys

import numpy as np
import pytest

import mesh_util as MU
import oracle_lib as OL

HERE = os.path.dirname(os.path.abspath(__file__))

CHILD = r"""
import sys, numpy as np
sys.path.insert(0, %r)
import oracle_lib as OL, mesh_util as MU
from test_oracle_omp import case
leaves, cfg, u0 = case()
s = OL.OracleSim(cfg, leaves)
s.set_state(u0)
s.run(2e-3, 3)
np.save(sys.argv[1], s.get_state())
"""


def case():
    leaves = MU.random_adapted(1, 3, seed=11)
    bnd = ("OUTLET", "WALL_SLIP", "OUTLET", "INLET", "WALL_NO_SLIP", "OUTLET")
    params = [None, None, None, (0.1, -0.2, 0.05, 1.1, 1.2), (0.9,), None]
    cfg = OL.make_config(flux="HLLC", limiter=True, model="NS", bnd=bnd, bnd_param=params)
    rng = np.random.default_rng(8)
    n = len(leaves)
    u = np.zeros((n, 5, 4))
    r = 1.0 + 0.2 * rng.uniform(-1, 1, n)
    vel = 0.3 * rng.normal(size=(n, 3))
    p = 1.0 + 0.2 * rng.uniform(-1, 1, n)
    u[:, 0:3, 0] = r[:, None] * vel
    u[:, 3, 0] = p / 0.4 + 0.5 * r * (vel ** 2).sum(1)
    u[:, 4, 0] = r
    u[:, :, 1:] = 0.05 * rng.normal(size=(n, 5, 3)) * np.abs(u[:, :, :1])
    return leaves, cfg, u


@pytest.mark.parametrize("build", ["liboracle_omp.so", "liboracle_faithful.so"])
def test_openmp_build_reproduces_serial_oracle(tmp_path, build):
    lib = os.path.join(OL.ORACLE_DIR, build)
    if not os.path.exists(lib):
        OL.build()
    out = str(tmp_path / "state.npy")
    env = dict(os.environ, CHARM_ORACLE_LIB=lib, OMP_NUM_THREADS="4")
    subprocess.run([sys.executable, "-c", CHILD % HERE, out], env=env, check=True, cwd=HERE)
    leaves, cfg, u0 = case()
    s = OL.OracleSim(cfg, leaves)
    s.set_state(u0)
    s.run(2e-3, 3)
    ref = s.get_state()
    got = np.load(out)
    assert np.array_equal(got, ref), np.abs(got - ref).max()
