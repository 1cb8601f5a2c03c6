"""The charm_3d program on a Gmsh task (src/charm_3d.c:11-50 with a gmsh_msh
mesh) on MI355X: a box of 3 x 2 x 2 hexahedral trees with rotated corner
labels and two regions, MIN_LEVEL 1, HLLC + BJ, inlet / outlet / both walls.

The program's solution files at steps 0 and 5 hold the state of the same
task driven through the C-ABI (charm_task_gmesh, charm_task_initial_state_
gmesh, charm_gpu_run) to 1e-13, its log lines fall on the steps
LOG_OUTPUT_STEP divides, and it warns that REFINE_PERIOD is not acted on.
"""
import os
import re

import numpy as np
import pytest

import charm_3d_amd as CA
from test_gmsh_task import setup_case
from test_output import read_vtu
from test_task_yaml import run_program

pytestmark = pytest.mark.gpu


def cell_fields(u, cfg):
    """cell-centre R, U, V, W, P of coefficient-0 states [n][V][4]"""
    m, cp = cfg.comp[0].m, cfg.comp[0].cp[0]
    gam = cp / (cp - CA.GAS_R / m)
    r = u[:, 4, 0]
    vel = u[:, 0:3, 0] / r[:, None]
    e = u[:, 3, 0] / r - 0.5 * (vel ** 2).sum(1)
    return {"R": r, "U": vel[:, 0], "V": vel[:, 1], "W": vel[:, 2], "P": r * e * (gam - 1)}


def test_program_runs_a_gmsh_task_as_the_api_does(tmp_path):
    # TAU 2e-4 to TMAX 1.3e-3: steps 0-6; output before steps 0 and 5, log
    # after steps 0, 2, 4 and 6
    d, t = setup_case(tmp_path)
    print("[gmsh task] program run", flush=True)
    r = run_program(d)
    assert r.returncode == 0, r.stderr
    assert "adaptation across trees is not built" in r.stderr
    steps = [int(x) for x in re.findall(r"STEP =\s+(\d+)", r.stdout)]
    assert steps == [0, 2, 4, 6], r.stdout
    out = {}
    for s in (0, 5):
        base = d / ("charm_3d_%08d" % s)
        assert os.path.exists(str(base) + ".pvtu")
        out[s] = read_vtu(str(base) + "_0000.vtu")[4]
    assert not os.path.exists(str(d / "charm_3d_00000010_0000.vtu"))

    print("[gmsh task] API run", flush=True)
    gm = CA.GMesh.from_task(t, str(d))
    u0 = gm.initial_state(t)
    ctx = CA.GpuContext(t.cfg)
    ctx.upload_gmesh(gm)
    ctx.set_state(u0)
    for s, nsteps in ((0, 0), (5, 5)):
        if nsteps:
            ctx.run(t.cfg.dt, nsteps)
        f = cell_fields(ctx.get_state(), t.cfg)
        # the same states (FP64 on both sides); the fields agree to their
        # host-side rounding
        for k, v in f.items():
            np.testing.assert_allclose(out[s][k], v, rtol=1e-13, atol=1e-13 * np.abs(v).max(),
                                       err_msg="%s at step %d" % (k, s))
    # the flow has moved: the inlet and the region interface change the state
    assert np.abs(out[5]["R"] - out[0]["R"]).max() > 1e-6 * out[0]["R"].max()
    ctx.close()
