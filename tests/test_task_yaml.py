"""task.yaml loader (charm_task_load_yaml; the keys of charm_init_context_yaml,
src/charm_init_yaml.cpp:333-426, and charm_model_ns_init,
src/models/ns/charm_model_ns_init.cpp:14-88), the unit-cube forest at
MIN_LEVEL (src/charm_3d.c:36-37) and the initial condition
(charm_init_initial_condition, src/charm_init.c:12-73).  CPU only.

The fixtures are written here in the layout of the reference's sample task
(samples/klin_amr/task.yaml): block mappings and sequences, "-" items on
their own line, flow mappings and sequences, and trailing comments."""
import textwrap

import numpy as np
import pytest

import charm_3d_amd as CA

R = CA.GAS_R

EULER_TASK = """\
method: CHARM_3D
version: 0.1.2
control:
  FLUX_TYPE: HLLC
  LIMITER: BJ
  MAX_ERROR: 0.01
  REFINE_PERIOD: 10
  REPARTITION_PERIOD: 10
  MIN_LEVEL: 3
  MAX_LEVEL: 5
  CFL: 0.1                    # Courant number
  TAU: 1.0e-2                 # time step
  TMAX: 50                    # end time
  FILE_OUTPUT_STEP: 100       # steps between output files
  LOG_OUTPUT_STEP: 10
  MODEL: {name: EULER, use_visc: 0, use_diffusion: 0, t_ref: 298.15}
mesh:
  files_type: unit_cube
  name: cube
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
    eof_type: IDEAL # IDEAL, MIX
regions:
  -
    id: 0
    name: CUBE
    cell_type: 0
    material_id: 0
    components:
      - {id: 0, concentration: 1}
    parameters:
      V: {x: 697.7082, y: 0, z: 0}
      T: 300
      P: 101325.0
      G: {x: 0, y: 0, z: -9.81}
boundaries:
  -
    name: BACK
    type: BOUND_WALL_SLIP
  -
    name: FRONT
    type: BOUND_WALL_SLIP
  -
    name: LEFT
    type: BOUND_INLET
    parameters:
      V: {x: 697.7082, y: 0, z: 0}
      T: 300
      P: 101325.0
  -
    name: RIGHT
    type: BOUND_OUTLET
  - name: TOP
    type: BOUND_WALL_NO_SLIP
    parameters: {T: 310.5}
  - {name: BOTTOM, type: BOUND_WALL_SLIP}
"""

NS_TASK = """\
method: CHARM_3D
version: 0.2.0
control:
  FLUX_TYPE: LF
  LIMITER: NONE
  MAX_ERROR: 0
  REFINE_PERIOD: 0
  REPARTITION_PERIOD: 0
  MIN_LEVEL: 2
  MAX_LEVEL: 2
  CFL: 0.3
  TAU: 1.0e-6
  TMAX: 1.0e-5
  FILE_OUTPUT_STEP: 0
  LOG_OUTPUT_STEP: 1
  PRECISION: FP32
  MODEL:
    name: NS
    use_visc: 1
    use_diffusion: 1
    t_ref: 298.15
mesh:
  files_type: unit_cube
  periodic: [1, 0, 1]
components:
  - id: 7
    name: H2
    cp_type: POLYNOM
    kp_type: SATHERLAND
    ml_type: SATHERLAND
    M: 0.002016
    Cp: [14000.0, 0.5, 1.0e-4]
    KP0: 0.17
    ML0: 8.4e-6
    T0: 273.0
    TS: 72.0
    Sig: 2.92
    ek: 38.0
    h0: 0
  - id: 3
    name: "O2"
    cp_type: POLYNOM
    kp_type: SATHERLAND
    ml_type: SATHERLAND
    M: 0.032
    Cp: [900.0, 0.2]
    KP0: 0.024
    ML0: 1.9e-5
    T0: 273.0
    TS: 127.0
    Sig: 3.458
    ek: 107.4
    h0: 0
materials:
  - {id: 4, name: GAS, eof_type: MIX}
regions:
  - id: 0
    name: MIXTURE
    cell_type: 0
    material_id: 4
    components:
      - {id: 3, concentration: 0.75}
      - {id: 7, concentration: 0.25}
    parameters:
      V: {x: 1, y: 2, z: 3}
      T: 400
      P: 2.0e5
      G: {x: 0, y: 0, z: 0}
boundaries:
  - {name: FRONT, type: BOUND_OUTLET}
  - {name: BACK, type: BOUND_OUTLET}
reactions:
  - {a: 1.0e9, e: 1.2e5, n: 0.5, left: [0, 1], right: [1]}
"""


def write(tmp_path, text, name="task.yaml"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_euler_task_fields(tmp_path):
    t = CA.load_task(write(tmp_path, EULER_TASK))
    c = t.cfg
    assert (c.flux, c.limiter, c.model, c.eos) == (CA.FLUX["HLLC"], 1, CA.MODEL["EULER"], CA.EOS["IDEAL"])
    assert (c.max_err, c.refine_period, t.repartition_period) == (0.01, 10, 10)
    assert (c.min_level, c.max_level, c.write_period, c.log_period) == (3, 5, 100, 10)
    assert (c.cfl, c.dt, c.tmax, c.precision) == (0.1, 1.0e-2, 50.0, CA.FP64)
    assert c.n_comp == 1 and c.comp_name[0].value == b"AIR"
    q = c.comp[0]
    assert (q.m, q.n_cp, q.cp[0], q.cp_type, q.ml_type, q.kp_type) == (0.02869409, 1, 1014.16, 0, 0, 0)
    assert (q.t0, q.ts, q.sig, q.ek, q.h0) == (273.15, 218.0, 3.512, 139.8, -8.386329e4)
    # boundaries by cube side: LEFT -x, RIGHT +x, FRONT -y, BACK +y, BOTTOM -z, TOP +z
    B = CA.BND
    assert list(c.bnd_type) == [B["INLET"], B["OUTLET"], B["WALL_SLIP"], B["WALL_SLIP"], B["WALL_SLIP"],
                                B["WALL_NO_SLIP"]]
    assert list(c.bnd_param[0])[:5] == [697.7082, 0.0, 0.0, 300.0, 101325.0]
    assert c.bnd_param[5][0] == 310.5
    assert [t.bnd_name[i].value for i in range(6)] == [b"LEFT", b"RIGHT", b"FRONT", b"BACK", b"BOTTOM", b"TOP"]
    assert t.n_regions == 1 and t.region[0].name == b"CUBE"
    g = t.region[0]
    assert (list(g.v), g.t, g.p, list(g.grav), g.c[0]) == ([697.7082, 0.0, 0.0], 300.0, 101325.0, [0.0, 0.0, -9.81], 1.0)
    assert list(c.grav) == [0.0, 0.0, -9.81]
    assert t.mesh_files_type == b"unit_cube" and t.mesh_name == b"cube"


def test_euler_task_forest_and_initial_state(tmp_path):
    t = CA.load_task(write(tmp_path, EULER_TASK))
    f = CA.task_forest(t)
    assert f.n == 8 ** 3 and (f.leaves()[:, 3] == 3).all()
    u = CA.task_initial_state(t, f)
    # EOS flag 3 (T, p) -> r, e with the constant cp of the component
    m, cp = 0.02869409, 1014.16
    cv = cp - R / m
    r = 101325.0 * m / (300.0 * R)
    e = 101325.0 / (r * (cp / cv - 1.0))
    ref = np.zeros((5, 4))
    ref[0, 0] = r * 697.7082
    ref[3, 0] = r * (e + 0.5 * 697.7082 ** 2)
    ref[4, 0] = r
    np.testing.assert_allclose(u, np.broadcast_to(ref, u.shape), rtol=1e-14, atol=0)


def test_ns_task_fields(tmp_path):
    t = CA.load_task(write(tmp_path, NS_TASK))
    c = t.cfg
    assert (c.model, c.use_visc, c.use_diff, c.t_ref) == (CA.MODEL["NS"], 1, 1, 298.15)
    assert (c.flux, c.limiter, c.eos, c.precision) == (CA.FLUX["LF"], 0, CA.EOS["MIX"], CA.FP32)
    assert c.n_comp == 2 and [c.comp_name[i].value for i in range(2)] == [b"H2", b"O2"]
    assert (c.comp[0].cp_type, c.comp[0].ml_type, c.comp[0].kp_type) == (1, 1, 1)
    assert (c.comp[0].n_cp, list(c.comp[0].cp)[:3]) == (3, [14000.0, 0.5, 1.0e-4])
    assert (c.comp[1].n_cp, list(c.comp[1].cp)[:2]) == (2, [900.0, 0.2])
    # concentrations by component id (O2 has id 3, H2 id 7)
    assert list(t.region[0].c)[:2] == [0.25, 0.75]
    assert list(t.mesh_periodic) == [1, 0, 1]
    P, B = CA.BND["PERIODIC"], CA.BND["OUTLET"]
    assert list(c.bnd_type) == [P, P, B, B, P, P]
    assert c.n_react == 1
    rx = c.react[0]
    assert (rx.a, rx.e, rx.n, list(rx.left), list(rx.right)) == (1e9, 1.2e5, 0.5, [0, 1, -1], [1, -1, -1])
    f = CA.task_forest(t)
    assert f.n == 64 and f.periodic == (True, False, True)
    u = CA.task_initial_state(t, f)
    # mixture EOS at T = 400 K: cp = sum c_i cp_i(T) (polynomial cp with
    # Sutherland viscosity), M = 1 / sum c_i / M_i
    y = np.array([0.25, 0.75])
    cps = np.array([14000.0 + 0.5 * 400 + 1e-4 * 400 ** 2, 900.0 + 0.2 * 400])
    mm = 1.0 / (y / np.array([0.002016, 0.032])).sum()
    cp = (y * cps).sum()
    r = 2.0e5 * mm / (400.0 * R)
    e = 2.0e5 / (r * (cp / (cp - R / mm) - 1.0))
    np.testing.assert_allclose(u[0, :, 0], [r, 2 * r, 3 * r, r * (e + 7.0), 0.25 * r, 0.75 * r], rtol=1e-14)
    assert (u[:, :, 1:] == 0).all()


@pytest.mark.parametrize("edit,msg", [
    (("method: CHARM_3D", "method: OTHER"), "CHARM_3D"),
    (("version: 0.1.2", "version: 0.0.9"), "version"),
    (("FLUX_TYPE: HLLC", "FLUX_TYPE: ROE"), "flux type 'ROE'"),
    (("LIMITER: BJ", "LIMITER: WENO"), "limiter type 'WENO'"),
    (("eof_type: IDEAL # IDEAL, MIX", "eof_type: TABLE"), "TABLE"),
    (("- {id: 0, concentration: 1}", "- {id: 0, concentration: 1.5}"), "sum of concentrations"),
    (("- {id: 0, concentration: 1}", "- {id: 5, concentration: 1}"), "unknown component id"),
    (("name: BACK", "name: SIDE"), "not a cube side"),
    (("type: BOUND_OUTLET", "type: BOUND_OPEN"), "boundary type 'BOUND_OPEN'"),
    (("TMAX: 50 ", "TMAX: fifty "), "'TMAX' in control is not a number"),
    (("  CFL: 0.1", "  CFL: 0.1\n      bad: 1"), "task.yaml line 12: unexpected indentation"),
    (("MODEL: {name: EULER", "MODEL: {name: LES"), "model type 'LES'"),
    (("  -\n    name: RIGHT\n    type: BOUND_OUTLET\n", ""), "no boundary for cube side RIGHT"),
])
def test_task_errors(tmp_path, edit, msg):
    text = EULER_TASK.replace(*edit)
    assert text != EULER_TASK
    with pytest.raises(CA.CharmError, match=msg):
        CA.load_task(write(tmp_path, text))


def test_gmsh_mesh_is_unsupported(tmp_path):
    text = EULER_TASK.replace("files_type: unit_cube\n  name: cube", "files_type: gmsh_msh\n  name: klin.msh")
    t = CA.load_task(write(tmp_path, text))
    assert t.mesh_files_type == b"gmsh_msh" and t.mesh_name == b"klin.msh"
    with pytest.raises(CA.CharmError, match="not supported"):
        CA.task_forest(t)


def test_missing_file():
    with pytest.raises(CA.CharmError, match="cannot open"):
        CA.load_task("/nonexistent/task.yaml")


def test_quoted_scalars_and_comments(tmp_path):
    text = EULER_TASK.replace("name: AIR", "name: 'AIR # not a comment'").replace(
        "FLUX_TYPE: HLLC", 'FLUX_TYPE: "HLLC"   # quoted')
    t = CA.load_task(write(tmp_path, text))
    assert t.cfg.comp_name[0].value == b"AIR # not a comment"
    assert t.cfg.flux == CA.FLUX["HLLC"]


def test_yaml_round_trip_against_python_yaml(tmp_path):
    """the subset parser reads the fixtures as PyYAML does (checked on the
    values the loader maps)"""
    yaml = pytest.importorskip("yaml")
    for text in (EULER_TASK, NS_TASK):
        d = yaml.safe_load(textwrap.dedent(text))
        t = CA.load_task(write(tmp_path, text))
        assert t.cfg.cfl == d["control"]["CFL"] and t.cfg.dt == float(d["control"]["TAU"])
        assert t.cfg.n_comp == len(d["components"])
        for i, q in enumerate(d["components"]):
            assert [t.cfg.comp[i].cp[k] for k in range(len(q["Cp"]))] == [float(v) for v in q["Cp"]]
            assert t.cfg.comp_name[i].value.decode() == q["name"]
        g = d["regions"][0]["parameters"]
        assert list(t.region[0].v) == [g["V"]["x"], g["V"]["y"], g["V"]["z"]]


def run_program(cwd, *args):
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(CA.__file__), "charm_3d")
    return subprocess.run([exe, *args], cwd=str(cwd), capture_output=True, text=True, timeout=120)


def test_program_reports_task_errors(tmp_path):
    """the charm_3d program (src/charm_3d.c:11-50) stops with status 1 and
    the loader's message before any device work"""
    r = run_program(tmp_path)
    assert r.returncode == 1 and "cannot open task.yaml" in r.stderr
    write(tmp_path, EULER_TASK.replace("files_type: unit_cube", "files_type: gmsh_msh"))
    r = run_program(tmp_path)
    assert r.returncode == 1 and "charm_3d: mesh: Failed to open" in r.stderr
    p = write(tmp_path, EULER_TASK.replace("LIMITER: BJ", "LIMITER: MINMOD"), "other.yaml")
    r = run_program(tmp_path, p)
    assert r.returncode == 1 and "unknown limiter type 'MINMOD'" in r.stderr
