"""The reference's own entry points, called through their reference
signatures with a forest whose user pointer is the context (the way a
reference caller binds them): charm_calc_flux_{lf,hllc,godunov,cd}
(include/charm_fluxes.h:19-32), charm_limiter (include/charm_limiter.h:13-15)
and charm_model_{euler,ns}_get_dt (include/charm_models.h:15-23).  The
fluxes are checked against the oracle and HLLC also against the independent
Toro star-state form of tests/test_oracle_pins.py; the limiter against the
oracle and the hand-computed vertex psi."""
import ctypes as C

import numpy as np
import pytest

import charm_3d_amd as CA
import mesh_util as MU
import oracle_lib as OL
from test_gpu_parity import oracle_cfg, random_state
from test_oracle_pins import toro_hllc
from test_pointwise import ChPrim, FIELDS

pytestmark = pytest.mark.gpu

FLUXES = ("lf", "hllc", "godunov", "cd")


def bind():
    L = CA.lib()
    dp = C.POINTER(C.c_double)
    for nm in FLUXES:
        fn = getattr(L, "charm_calc_flux_" + nm)
        fn.restype = None
        fn.argtypes = [C.c_void_p, C.POINTER(ChPrim), dp, dp, dp, dp, dp, dp]
    L.charm_limiter.restype = None
    L.charm_limiter.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    for nm in ("charm_model_euler_get_dt", "charm_model_ns_get_dt"):
        getattr(L, nm).restype = C.c_double
        getattr(L, nm).argtypes = [C.c_void_p]
    return L


def ch_prim(p, nc):
    q = ChPrim()
    for k in FIELDS:
        setattr(q, k, getattr(p, k))
    for i in range(nc):
        q.c[i] = p.c[i]
    return q


@pytest.mark.parametrize("nc", [1, 2])
@pytest.mark.parametrize("name", FLUXES)
def test_calc_flux_reference_signature(name, nc):
    if nc == 1:
        cfg = CA.make_config(precision=CA.FP64, flux=name.upper())
    else:
        R = CA.GAS_R
        cfg = CA.make_config(precision=CA.FP64, flux=name.upper(), eos="MIX", comp_m=(0.029, 0.004),
                             comp_cp=(1.4 * R / (0.4 * 0.029), 1.67 * R / (0.67 * 0.004)))
    ocfg = oracle_cfg(cfg)
    L = bind()
    f = CA.Forest.uniform(1)
    s = CA.Solver(f, cfg)
    fn = getattr(L, "charm_calc_flux_" + name)
    rng = np.random.default_rng(17)
    pscale = 1.0 if nc == 1 else 1e5
    for _ in range(24):
        c = (None, None) if nc == 1 else (rng.dirichlet((1, 1)), rng.dirichlet((1, 1)))
        a = OL.prim_from(ocfg, rng.uniform(0.3, 2), *rng.normal(size=3), pscale * rng.uniform(0.3, 3), c=c[0])
        b = OL.prim_from(ocfg, rng.uniform(0.3, 2), *rng.normal(size=3), pscale * rng.uniform(0.3, 3), c=c[1])
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        prim = (ChPrim * 2)(ch_prim(a, nc), ch_prim(b, nc))
        qu, qv, qw, qe = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        qc = np.zeros(nc)
        nn = np.ascontiguousarray(n)
        fn(f.h, prim, C.byref(qu), C.byref(qv), C.byref(qw), C.byref(qe),
           qc.ctypes.data_as(C.POINTER(C.c_double)), nn.ctypes.data_as(C.POINTER(C.c_double)))
        assert s and CA.lib().charm_ctx_error(s.h) == 0
        got = np.array([qu.value, qv.value, qw.value, qe.value, *qc])
        ref = OL.flux(ocfg, name.upper(), a, b, n)
        tol = 1e-9 if name == "godunov" else 1e-12
        assert np.allclose(got, ref, rtol=tol, atol=tol * np.abs(ref).max()), (got, ref)
        if name == "hllc":
            toro, w = toro_hllc(a, b, n, nc)
            # the star region (the supersonic branches are the upwind flux)
            if w["SL"] < 0 < w["SR"] and w["SL"] < w["SS"] < w["SR"]:
                assert np.allclose(got, toro, rtol=1e-12, atol=1e-12 * np.abs(toro).max())
    s.close()


def test_limiter_and_get_dt_reference_signatures():
    """charm_limiter on a 2:1 mesh with boundaries against the oracle; the
    hand-computed vertex psi of tests/test_oracle_pins.py through the same
    call; get_dt returns TAU (src/models/euler/charm_model.c:43-53)."""
    L = bind()
    leaves = MU.random_adapted(1, 3, seed=3)
    bnd = ("OUTLET", "WALL_SLIP", "OUTLET", "INLET", "OUTLET", "OUTLET")
    params = [None, None, None, (0.1, -0.2, 0.05, 1.1, 1.2), None, None]
    cfg = CA.make_config(precision=CA.FP64, limiter=True, bnd=bnd, bnd_param=params, dt=1.25e-3)
    u0 = random_state(leaves, np.random.default_rng(4), amp=0.3)
    f = CA.Forest.from_leaves(leaves)
    s = CA.Solver(f, cfg)
    s.set_state(u0)
    L.charm_limiter(f.h, None, None)
    assert CA.lib().charm_ctx_error(s.h) == 0
    sim = OL.OracleSim(oracle_cfg(cfg), leaves)
    sim.set_state(u0)
    sim.limiter()
    ref = sim.get_state()
    got = s.get_state()
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    assert L.charm_model_euler_get_dt(f.h) == 1.25e-3
    assert L.charm_model_ns_get_dt(f.h) == 1.25e-3
    s.close()
    # the hand value: slopes on two axes, neighbour means above and below
    level = 2
    leaves = MU.uniform_leaves(level)
    h = MU.ROOT >> level
    key = {tuple(q[:3]): i for i, q in enumerate(leaves.tolist())}
    e, ex_p, ex_m = key[(h, h, h)], key[(2 * h, h, h)], key[(0, h, h)]
    base = np.array([0.1, 0.0, 0.0, 2.5, 1.0])
    scale = np.maximum(np.abs(base), 0.1)
    u = np.zeros((len(leaves), 5, 4))
    u[:, :, 0] = base
    a = (0.030 + 0.004 * np.arange(5)) * scale
    b = (0.045 + 0.003 * np.arange(5)) * scale
    sx = 0.08 * scale * np.array([1, -1, 1, 1, -1])
    sy = 0.06 * scale * np.array([-1, 1, 1, -1, 1])
    u[ex_p, :, 0] += a
    u[ex_m, :, 0] -= b
    u[e, :, 1], u[e, :, 2] = sx, sy
    f = CA.Forest.from_leaves(leaves)
    s = CA.Solver(f, CA.make_config(precision=CA.FP64, limiter=True))
    s.set_state(u)
    L.charm_limiter(f.h, None, None)
    out = s.get_state()
    psi = np.minimum(1.0, np.minimum(a, b) / (0.5 * (np.abs(sx) + np.abs(sy))))
    assert np.allclose(out[e, :, 1], psi * sx, rtol=1e-13, atol=1e-16)
    assert np.allclose(out[e, :, 2], psi * sy, rtol=1e-13, atol=1e-16)
    s.close()
