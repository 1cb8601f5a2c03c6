"""Gmsh tasks (mesh.files_type gmsh_msh) on the CPU.

* task.yaml's boundaries of a Gmsh mesh take the configuration's slots 0-5 in
  the order of the file, whatever their names (a unit cube binds them to its
  sides by name); more than six are refused;
* charm_task_gmesh reads mesh.name relative to the task's directory, binds
  patches and regions by name, and equals the multi-tree mesh built from the
  connectivity's own arrays at MIN_LEVEL;
* charm_task_initial_state_gmesh gives every leaf the state of its tree's
  region (src/charm_init.c:12-73: EOS flag 3 from (T, p) into coefficient 0);
* charm_gmesh_write_vtk writes each leaf's 8 trilinear corners, its tree and
  level and the cell-centre fields;
* the charm_3d program reports a Gmsh task's missing mesh file.
"""
import os

import numpy as np
import pytest

import charm_3d_amd as CA
from test_conn_msh import write_box
from test_output import read_vtu
from test_task_yaml import run_program

# boundaries in an order of their own: the slots follow it
ORDER = ("TOP", "LEFT", "BACK", "RIGHT", "FRONT", "BOTTOM")
TYPES = {"TOP": "BOUND_WALL_SLIP", "LEFT": "BOUND_INLET", "BACK": "BOUND_WALL_NO_SLIP", "RIGHT": "BOUND_OUTLET",
         "FRONT": "BOUND_WALL_SLIP", "BOTTOM": "BOUND_WALL_SLIP"}


def gmsh_task(mesh="box.msh", min_level=1, tau=2e-4, tmax=1.3e-3, out_step=5, log_step=2, extra_bnd=()):
    bnd = []
    for name in ORDER + tuple(extra_bnd):
        b = "  -\n    name: %s\n    type: %s\n" % (name, TYPES.get(name, "BOUND_OUTLET"))
        if TYPES.get(name) == "BOUND_INLET":
            b += "    parameters:\n      V: {x: 120.0, y: 5.0, z: 0}\n      T: 310\n      P: 101325.0\n"
        elif TYPES.get(name) == "BOUND_WALL_NO_SLIP":
            b += "    parameters:\n      T: 290\n"
        bnd.append(b)
    return """\
method: CHARM_3D
version: 0.1.2
control:
  FLUX_TYPE: HLLC
  LIMITER: BJ
  MAX_ERROR: 0.01
  REFINE_PERIOD: 10
  REPARTITION_PERIOD: 10
  MIN_LEVEL: %d
  MAX_LEVEL: 2
  CFL: 0.1
  TAU: %g
  TMAX: %g
  FILE_OUTPUT_STEP: %d
  LOG_OUTPUT_STEP: %d
  MODEL: {name: EULER, use_visc: 0, use_diffusion: 0, t_ref: 298.15}
mesh:
  files_type: gmsh_msh
  name: %s
components:
  -
    id: 0
    name: AIR
    cp_type: CONST
    kp_type: CONST
    ml_type: CONST
    M: 0.02869409
    Cp: [1014.16]
    KP0: 0
    ML0: 0
    T0: 273.15
    TS: 218.0
    Sig: 3.512
    ek: 139.8
    h0: -8.386329e+4
materials:
  -
    id: 0
    name: MIXTURE
    eof_type: IDEAL
regions:
  -
    id: 0
    name: BOX
    cell_type: 0
    material_id: 0
    components:
      - {id: 0, concentration: 1}
    parameters:
      V: {x: 100.0, y: 0, z: 0}
      T: 300
      P: 101325.0
      G: {x: 0, y: 0, z: 0}
  -
    id: 1
    name: BOX2
    cell_type: 0
    material_id: 0
    components:
      - {id: 0, concentration: 1}
    parameters:
      V: {x: 60.0, y: -10.0, z: 5.0}
      T: 400
      P: 90000.0
      G: {x: 0, y: 0, z: 0}
boundaries:
%s""" % (min_level, tau, tmax, out_step, log_step, mesh, "".join(bnd))


def setup_case(tmp_path, shape=(3, 2, 2), seed=4, **kw):
    d = tmp_path / "case"
    d.mkdir()
    write_box(d / "box.msh", shape, np.random.default_rng(seed), rotate=True, split_x=1)
    (d / "task.yaml").write_text(gmsh_task(**kw))
    return d, CA.load_task(str(d / "task.yaml"))


def region_cons(v, t, p):
    m, cp = 0.02869409, 1014.16
    cv = cp - CA.GAS_R / m
    rho = p * m / (CA.GAS_R * t)
    v = np.asarray(v, float)
    return np.array([*(rho * v), rho * (cv * t + 0.5 * (v ** 2).sum()), rho])


def test_boundaries_take_slots_in_file_order(tmp_path):
    d, t = setup_case(tmp_path)
    assert [t.bnd_name[i].value.decode() for i in range(6)] == list(ORDER)
    want = [CA.BND[TYPES[n].replace("BOUND_", "")] for n in ORDER]
    assert [t.cfg.bnd_type[i] for i in range(6)] == want
    s = ORDER.index("LEFT")
    assert [t.cfg.bnd_param[s][j] for j in range(5)] == [120.0, 5.0, 0.0, 310.0, 101325.0]
    assert t.cfg.bnd_param[ORDER.index("BACK")][0] == 290.0
    (d / "seven.yaml").write_text(gmsh_task(extra_bnd=("LID",)))
    with pytest.raises(CA.CharmError, match="7 boundaries"):
        CA.load_task(str(d / "seven.yaml"))


def test_task_gmesh_equals_the_connectivity_mesh(tmp_path):
    d, t = setup_case(tmp_path)
    gm = CA.GMesh.from_task(t, str(d))
    c = CA.Connectivity(d / "box.msh")
    tb, tr = c.bind_task(t)
    ref = CA.GMesh(c.vertices, c.tree_to_vertex, tb, t.cfg.min_level)
    assert gm.n == ref.n == 12 * 8 and gm.nf == ref.nf
    np.testing.assert_array_equal(gm.leaves(), ref.leaves())
    np.testing.assert_array_equal(gm.faces(), ref.faces())
    ga, gb = gm.geometry(), ref.geometry()
    for k in ga:
        np.testing.assert_array_equal(ga[k], gb[k])
    # boundary faces carry the slot of their side's name
    fc = gm.faces()
    assert set(int(-1 - x) for x in fc[:, 1] if x < 0) == set(range(6))
    # a mesh path relative to the working directory fails from elsewhere
    with pytest.raises(CA.CharmError, match="open"):
        CA.GMesh.from_task(t, str(tmp_path))


def test_initial_state_takes_each_trees_region(tmp_path):
    d, t = setup_case(tmp_path)
    gm = CA.GMesh.from_task(t, str(d))
    u = gm.initial_state(t)
    c = CA.Connectivity(d / "box.msh")
    _, tr = c.bind_task(t)
    lv = gm.leaves()
    a = region_cons((100.0, 0, 0), 300.0, 101325.0)
    b = region_cons((60.0, -10.0, 5.0), 400.0, 90000.0)
    assert set(tr.tolist()) == {0, 1}
    for i in range(gm.n):
        want = a if tr[lv[i, 0]] == 0 else b
        np.testing.assert_allclose(u[i, :, 0], want, rtol=1e-14)
        assert (u[i, :, 1:] == 0).all()
    # the trees of BOX2 are those at x >= 1 (write_box's split)
    cen = gm.geometry()["c"]
    assert ((cen[:, 0] >= 1.0) == (tr[lv[:, 0]] == 1)).all()


def test_vtk_of_a_multi_tree_mesh(tmp_path):
    d, t = setup_case(tmp_path)
    gm = CA.GMesh.from_task(t, str(d))
    u = gm.initial_state(t)
    base = str(tmp_path / "mt")
    gm.write_vtk(base, u, t.cfg)
    assert os.path.exists(base + ".pvtu")
    npts, ncell, pts, cells, cdata = read_vtu(base + "_0000.vtu")
    assert ncell == gm.n and npts == 8 * gm.n
    cor = gm.corners()
    zk = [0, 1, 3, 2, 4, 5, 7, 6]
    for i in range(gm.n):
        np.testing.assert_array_equal(pts[cells["connectivity"][8 * i: 8 * i + 8]], cor[i, zk])
    # the leaf corners are the trilinear images of the tree corners: on this
    # undistorted box the leaf centre is their mean
    np.testing.assert_allclose(cor.mean(1), gm.geometry()["c"], atol=1e-14)
    lv = gm.leaves()
    np.testing.assert_array_equal(cdata["treeid"], lv[:, 0])
    np.testing.assert_array_equal(cdata["level"], lv[:, 4])
    c = CA.Connectivity(d / "box.msh")
    _, tr = c.bind_task(t)
    reg = tr[lv[:, 0]]
    np.testing.assert_allclose(cdata["T"], np.where(reg == 0, 300.0, 400.0), rtol=1e-12)
    np.testing.assert_allclose(cdata["P"], np.where(reg == 0, 101325.0, 90000.0), rtol=1e-12)
    np.testing.assert_allclose(cdata["U"], np.where(reg == 0, 100.0, 60.0), rtol=1e-14)


def test_program_reports_a_missing_gmsh_mesh(tmp_path):
    (tmp_path / "task.yaml").write_text(gmsh_task(mesh="absent.msh"))
    r = run_program(tmp_path)
    assert r.returncode == 1 and "mesh" in r.stderr and "absent.msh" in r.stderr
