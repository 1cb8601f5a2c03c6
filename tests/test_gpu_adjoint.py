"""GPU discrete adjoint of the SSP-RK3 time step (the Gaussian-pulse case):
gradients of J = sum_n dt int_box p against central finite differences of
the forward objective (FP64), invariance under the checkpoint period, and
the FP32 adjoint against the FP64 one."""
import numpy as np
import pytest

import charm_3d_amd as CA
import charm_3d_amd.workloads as W
import mesh_util as MU
from test_gpu_parity import device

pytestmark = pytest.mark.gpu

# a box next to the inlet, so that the inlet parameters act on J within the
# short horizon of these checks (the benchmark's box is the one of the case)
LO, HI = (0.0, 0.0, 0.0), (0.45, 1.0, 1.0)
DT, STEPS = 0.01, 16


def case(precision=CA.FP64, inlet=(0.5, 1.0, 1.0), adapted=False):
    lv = MU.random_adapted(2, 3, seed=8, frac=0.3) if adapted else CA.Forest.uniform(3).leaves()
    cfg, lv, u0, per = W.c5_pulse(precision=precision, inlet=inlet, leaves=lv)
    return cfg, lv, u0, per


def objective(precision, inlet, u0=None, adapted=False):
    cfg, lv, u, per = case(precision, inlet, adapted)
    ctx = device(cfg, lv, per, u if u0 is None else u0)
    return ctx.objective(DT, STEPS, LO, HI)


@pytest.mark.parametrize("adapted", [False, True])
def test_inlet_gradient_matches_central_differences(adapted):
    inlet = np.array([0.5, 1.0, 1.0])
    cfg, lv, u0, per = case(inlet=tuple(inlet), adapted=adapted)
    ctx = device(cfg, lv, per, u0)
    J, g, th, _, _ = ctx.adjoint(DT, STEPS, LO, HI, ckpt=5, theta_face=0)
    assert J > 0 and np.isfinite(g).all()
    for k in range(3):
        h = 1e-6 * max(abs(inlet[k]), 1.0)
        ip, im = inlet.copy(), inlet.copy()
        ip[k] += h
        im[k] -= h
        fd = (objective(CA.FP64, tuple(ip), adapted=adapted) - objective(CA.FP64, tuple(im), adapted=adapted)) / (2 * h)
        assert abs(fd) > 1e-4 and abs(th[k] - fd) <= 1e-5 * abs(fd), (k, th[k], fd)


@pytest.mark.parametrize("adapted", [False, True])
def test_initial_state_gradient_matches_directional_differences(adapted):
    cfg, lv, u0, per = case(adapted=adapted)
    ctx = device(cfg, lv, per, u0)
    _, g, _, _, _ = ctx.adjoint(DT, STEPS, LO, HI, ckpt=5)
    rng = np.random.default_rng(5)
    for _ in range(3):
        v = rng.normal(size=u0.shape) * np.abs(u0).max(axis=0, keepdims=True)
        eps = 1e-7
        jp = objective(CA.FP64, (0.5, 1.0, 1.0), u0 + eps * v, adapted)
        jm = objective(CA.FP64, (0.5, 1.0, 1.0), u0 - eps * v, adapted)
        fd = (jp - jm) / (2 * eps)
        ad = float((g * v).sum())
        assert abs(ad - fd) <= 1e-5 * abs(fd), (ad, fd)


def test_checkpoint_period_does_not_change_the_gradient():
    out = []
    # 0: every stage state kept, no recompute; 3 * STEPS: a period beyond the
    # run is one segment of the whole run
    for ckpt in (1, 7, STEPS, 3 * STEPS, 0):
        cfg, lv, u0, per = case()
        out.append(device(cfg, lv, per, u0).adjoint(DT, STEPS, LO, HI, ckpt=ckpt, theta_face=0))
    for o in out[1:]:
        assert o[0] == out[0][0]
        assert np.array_equal(o[1], out[0][1]) and np.array_equal(o[2], out[0][2])


@pytest.mark.parametrize("precision", [CA.FP64, CA.FP32])
def test_adjoint_leaves_the_final_state(precision):
    # the forward pass ends at u^N whether the last segment's stage states
    # are kept in the segment buffers or the steps run in place: the same
    # bits as the objective's plain forward pass, for every checkpoint period
    cfg, lv, u0, per = case(precision)
    ref = device(cfg, lv, per, u0)
    J = ref.objective(DT, STEPS, LO, HI)
    u_end = ref.get_state()
    assert not np.array_equal(u_end, np.asarray(u0, dtype=np.float64).reshape(u_end.shape))   # the steps moved it
    for ckpt in (1, 5, 7, STEPS, 3 * STEPS, 0):
        cfg, lv, u0, per = case(precision)
        ctx = device(cfg, lv, per, u0)
        Ja = ctx.adjoint(DT, STEPS, LO, HI, ckpt=ckpt)[0]
        assert Ja == J, (ckpt, Ja, J)
        assert np.array_equal(ctx.get_state(), u_end), ckpt


def test_fp32_adjoint_close_to_fp64():
    res = {}
    for prec in (CA.FP64, CA.FP32):
        cfg, lv, u0, per = case(prec)
        res[prec] = device(cfg, lv, per, u0).adjoint(DT, STEPS, LO, HI, ckpt=4, theta_face=0)
    g64, g32 = res[CA.FP64][1], res[CA.FP32][1]
    assert np.sqrt(((g32 - g64) ** 2).sum() / (g64 ** 2).sum()) < 1e-4
    np.testing.assert_allclose(res[CA.FP32][2], res[CA.FP64][2], rtol=1e-4)


def _mixture_cons(rho, vel, p, y, ms, cps):
    """conservative state of a constant-cp mixture (cp = sum y cp_i,
    M = 1 / sum y / M_i, e = cv T, p = rho R T / M)"""
    cp = sum(y[:, i] * cps[i] for i in range(len(ms)))
    rm = sum(y[:, i] / ms[i] for i in range(len(ms))) * CA.GAS_R      # R / M
    cv = cp - rm
    e = cv * p / (rho * rm)
    out = np.zeros((len(rho), 4 + len(ms)))
    out[:, 0:3] = rho[:, None] * vel
    out[:, 3] = rho * (e + 0.5 * (vel ** 2).sum(1))
    out[:, 4:] = rho[:, None] * y
    return out


def _branch_case(name):
    """configurations that drive the stage VJPs through their branches:
    the LF and central fluxes, HLLC in supersonic flow (outer wave-speed
    regions), a strong pressure jump (shock branches of the wave-speed
    estimate) and a two-species mixture (gamma depending on the state)"""
    lv = MU.random_adapted(2, 3, seed=8, frac=0.3)
    flux, vel0, jump, nc = {"lf": ("LF", 0.2, False, 1), "central": ("CD", 0.2, False, 1),
                            "supersonic": ("HLLC", 2.0, False, 1), "jump": ("HLLC", 0.1, True, 1),
                            "mixture": ("HLLC", 0.3, False, 2)}[name]
    ms = (CA.GAS_R, 2.0 * CA.GAS_R)[:nc]
    cps = (3.5, 1.67 / 0.67 * 0.5)[:nc]

    def fn(x):
        r2 = ((x - 0.4) ** 2).sum(1)
        rho = 1 + 0.1 * np.exp(-r2 / 0.02)
        p = np.ones(len(x))
        if jump:
            s = 0.5 * (1 + np.tanh((x[:, 0] - 0.5) / 0.05))
            rho = rho * (1 - 0.6 * s)
            p = 1 - 0.8 * s
        vel = np.column_stack([vel0 + 0.05 * np.sin(2 * np.pi * x[:, 1]), 0.1 * np.cos(2 * np.pi * x[:, 0]),
                               -0.05 * np.ones(len(x))])
        y = np.column_stack([0.7 + 0.2 * np.sin(2 * np.pi * x[:, 2]), 0.3 - 0.2 * np.sin(2 * np.pi * x[:, 2])])[:, :nc]
        y = y / y.sum(1, keepdims=True)
        return _mixture_cons(rho, vel, p, y, ms, cps)

    t_in = 1.0 / ((CA.GAS_R / ms[0]) if nc == 1 else 1.0)
    cfg = CA.make_config(precision=CA.FP64, model="NS", flux=flux, limiter=False,
                         eos="IDEAL" if nc == 1 else "MIX",
                         bnd=("INLET",) + ("OUTLET",) * 5,
                         bnd_param=[(vel0, 0.0, 0.0, t_in, 1.0)] + [None] * 5,
                         comp_m=ms, comp_cp=cps, cfl=0.1)
    return cfg, lv, W.project_p1(lv, fn, 4 + nc)


@pytest.mark.parametrize("name", ["lf", "central", "supersonic", "jump", "mixture"])
def test_stage_vjp_branches_match_directional_differences(name):
    cfg, lv, u0 = _branch_case(name)
    per = (False,) * 3
    ctx = device(cfg, lv, per, u0)
    dt = 0.5 * ctx.get_dt(1)
    steps = 8
    _, g, _, _, _ = ctx.adjoint(dt, steps, LO, HI, ckpt=3)
    assert np.isfinite(g).all()
    rng = np.random.default_rng(11)
    for _ in range(2):
        v = rng.normal(size=u0.shape) * np.abs(u0).max(axis=0, keepdims=True)
        eps = 1e-7
        jp = device(cfg, lv, per, u0 + eps * v).objective(dt, steps, LO, HI)
        jm = device(cfg, lv, per, u0 - eps * v).objective(dt, steps, LO, HI)
        fd = (jp - jm) / (2 * eps)
        ad = float((g * v).sum())
        print("%s: adjoint %.12e, central difference %.12e" % (name, ad, fd))
        assert abs(ad - fd) <= 1e-5 * abs(fd), (name, ad, fd)


@pytest.mark.parametrize("name", ["pulse_adapted", "lf", "central", "supersonic", "jump", "mixture"])
def test_fp32_chunked_adjoint_close_to_fp64(name):
    """The FP32 adjoint (chunked face-centric kernel: each interior face
    evaluated once, both sides' gradients from one reverse sweep) against
    the FP64 one (element-centric kernel), hanging faces included:
    relative L2 of dJ/du0 below 1e-4."""
    res = {}
    for prec in (CA.FP64, CA.FP32):
        if name == "pulse_adapted":
            cfg, lv, u0, per = case(prec, adapted=True)
            dt, steps = DT, STEPS
        else:
            cfg, lv, u0 = _branch_case(name)
            cfg.precision = prec
            per = (False,) * 3
            dt, steps = None, 8
        ctx = device(cfg, lv, per, u0)
        if dt is None:
            dt = 0.5 * ctx.get_dt(1)
        res[prec] = ctx.adjoint(dt, steps, LO, HI, ckpt=3)[1]
    g64, g32 = res[CA.FP64], res[CA.FP32]
    err = np.sqrt(((g32 - g64) ** 2).sum() / (g64 ** 2).sum())
    print("%s: FP32 vs FP64 gradient, relative L2 %.2e" % (name, err))
    assert err < 1e-4, err


def test_hllc_branch_flips_count_points_across_a_branch():
    """charm_gpu_hllc_branch_flips on a random subsonic state (no point near
    a branch switch): equal states flip no point; a leaf made supersonic
    flips points only on its own six faces; a 1e-13 relative perturbation
    flips none"""
    cfg, leaves, _, per = W.c5_pulse(3, CA.FP64)
    rng = np.random.default_rng(5)
    n = len(leaves)
    u0 = np.zeros((n, 5, 4))
    u0[:, :, 0] = W.prim_to_cons(1 + 0.1 * rng.uniform(-1, 1, n), 0.3 * rng.normal(size=(n, 3)),
                                 1 + 0.1 * rng.uniform(-1, 1, n))
    ce, cf = CA.recommended_chunks(cfg)
    mesh = CA.Forest.from_leaves(leaves, per).flatten(ce, cf)
    a, b = CA.GpuContext(cfg), CA.GpuContext(cfg)
    for c in (a, b):
        c.upload_mesh(mesh)
        c.set_state(u0)
    pts, flips = a.hllc_branch_flips(b)
    assert pts == 6 * 3 * 8 * 8 * 7 and flips == 0          # every interior face of the level-3 box, once
    u1 = u0.copy()
    e = 200
    r = u1[e, 4, 0]
    u1[e, 3, 0] -= 0.5 * (u1[e, 0:3, 0] ** 2).sum() / r
    u1[e, 0:3, 0] = [3.0 * r * np.sqrt(1.4 * 1.1 / 0.9), 0.0, 0.0]    # x velocity above 3 c: supersonic
    u1[e, 3, 0] += 0.5 * u1[e, 0, 0] ** 2 / r
    b.set_state(u1)
    pts1, flips1 = a.hllc_branch_flips(b)
    assert pts1 == pts and 1 <= flips1 <= 36, flips1
    b.set_state(u0 * (1 + 1e-13 * rng.uniform(-1, 1, u0.shape)))
    assert a.hllc_branch_flips(b)[1] == 0
    a.close()
    b.close()
