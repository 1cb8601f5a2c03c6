"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

The oracle is the checker for the product library; only tests, the smoke
check and bench.py's cpu_baseline leg use this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# CHARM_ORACLE_LIB points the checks at another build of the same sources
# (tests/test_oracle_mutants.py loads builds with a planted mistake)
LIB_PATH = os.environ.get("CHARM_ORACLE_LIB") or os.path.join(ORACLE_DIR, "liboracle.so")

NC_MAX = 8
NV_MAX = 4 + NC_MAX

FLUX = {"LF": 0, "HLLC": 1, "GODUNOV": 2, "CD": 3}
BND = {"INLET": 0, "OUTLET": 1, "WALL_SLIP": 2, "WALL_NO_SLIP": 3,
       "MASS_FLOW": 4, "PERIODIC": 100}
GAS_R = 8.31446261815324


class Prim(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("r", "u", "v", "w", "e", "e_tot", "p", "t", "cz", "gam", "cp", "cv", "m")] + \
               [("mat_id", C.c_int), ("c", C.c_double * NC_MAX)]


class Cons(C.Structure):
    _fields_ = [("ru", C.c_double), ("rv", C.c_double), ("rw", C.c_double), ("re", C.c_double),
                ("rc", C.c_double * NC_MAX), ("mat_id", C.c_int)]


class Comp(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("m", "ml0", "kp0", "t0", "ts", "sig", "ek", "h0")] + \
               [("cp_type", C.c_int), ("kp_type", C.c_int), ("ml_type", C.c_int),
                ("n_cp", C.c_int), ("cp", C.c_double * 8)]


class Reaction(C.Structure):
    _fields_ = [("left", C.c_int * 3), ("right", C.c_int * 3),
                ("a", C.c_double), ("e", C.c_double), ("n", C.c_double)]


class Config(C.Structure):
    _fields_ = [("n_comp", C.c_int), ("comp", Comp * NC_MAX), ("eos", C.c_int),
                ("flux", C.c_int), ("limiter", C.c_int), ("model", C.c_int),
                ("bnd_type", C.c_int * 6), ("bnd_param", (C.c_double * 16) * 6),
                ("grav", C.c_double * 3), ("use_visc", C.c_int), ("use_diff", C.c_int),
                ("t_ref", C.c_double), ("n_react", C.c_int), ("react", Reaction * 4),
                ("cfl", C.c_double), ("chem_parity", C.c_int), ("chem_energy_molar", C.c_int),
                ("visc_br2_eta", C.c_double)]


def build():
    """Compile the oracle with its Makefile (gcc; seconds)."""
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i64, dp, ip = C.c_void_p, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.or_create.restype = vp
        L.or_create.argtypes = [C.POINTER(Config), i64, ip]
        L.or_create_trees.restype = vp
        L.or_create_trees.argtypes = [C.POINTER(Config), i64, dp, ip, C.c_int]
        L.or_destroy.argtypes = [vp]
        L.or_nvar.argtypes = [vp]
        L.or_n_faces.restype = i64
        L.or_n_faces.argtypes = [vp]
        L.or_error.restype = C.c_char_p
        L.or_error.argtypes = [vp]
        for fn in ("or_set_state", "or_get_state", "or_conv_residual", "or_conservation_sums"):
            getattr(L, fn).argtypes = [vp, dp]
        for fn in ("or_limiter",):
            getattr(L, fn).argtypes = [vp]
        for fn in ("or_euler_step", "or_ns_step"):
            getattr(L, fn).argtypes = [vp, C.c_double]
        L.or_run.argtypes = [vp, C.c_double, C.c_int]
        L.or_faces.argtypes = [vp, ip]
        L.or_geom.argtypes = [vp, i64, dp]
        L.or_cfl_dt.restype = C.c_double
        L.or_cfl_dt.argtypes = [vp, C.c_int]
        L.or_cons_to_prim.argtypes = [C.POINTER(Config), C.POINTER(Cons), C.POINTER(Prim)]
        L.or_eos.argtypes = [C.POINTER(Config), C.POINTER(Prim), C.c_int]
        L.or_flux.argtypes = [C.POINTER(Config), C.c_int, C.POINTER(Prim), dp, dp]
        L.or_bnd.argtypes = [C.POINTER(Config), C.c_int, C.POINTER(Prim), C.POINTER(Prim), dp]
        L.or_matr_inv.argtypes = [dp, dp]
        L.or_rim.argtypes = [dp] * 6 + [C.c_double] * 12
        L.or_amr_indicator.argtypes = [vp, dp, dp]
        L.or_amr_flags.argtypes = [vp, dp, dp, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int8)]
        L.or_amr_project.argtypes = [C.c_int, i64, ip, dp, i64, ip, C.POINTER(C.c_int64), dp]
        for fn in ("or_ns_residual", "or_ns_aux", "or_ns_aux_own", "or_ns_visc", "or_ns_diff_coeffs"):
            getattr(L, fn).argtypes = [vp, dp]
        L.or_ns_chem_apply.argtypes = [vp, C.c_double, dp]
        L.or_gas_r.restype = C.c_double
        for fn in ("or_prim_size", "or_cons_size", "or_config_size"):
            getattr(L, fn).restype = C.c_size_t
        assert L.or_prim_size() == C.sizeof(Prim)
        assert L.or_cons_size() == C.sizeof(Cons)
        assert L.or_config_size() == C.sizeof(Config)
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_config(n_comp=1, comp_m=(GAS_R,), comp_cp=((3.5,),), eos="IDEAL", flux="LF",
                limiter=False, model="NS", bnd=("OUTLET",) * 6, bnd_param=None,
                grav=(0, 0, 0), cfl=0.3, transport=None, use_visc=False, use_diff=False,
                t_ref=298.15, reactions=(), chem_parity=False, chem_energy_molar=False,
                visc_br2_eta=0.0):
    """transport: per species dict with ml0, kp0, t0, ts, sig, ek, h0, ml_type,
    kp_type (Sutherland = 1); reactions: (left[3], right[3], a, e, n) tuples;
    visc_br2_eta > 0 selects the BR2 viscous flux with that penalty factor."""
    cfg = Config()
    cfg.n_comp = n_comp
    for i in range(n_comp):
        cfg.comp[i].m = comp_m[i]
        cps = comp_cp[i]
        cfg.comp[i].n_cp = len(cps)
        for j, v in enumerate(cps):
            cfg.comp[i].cp[j] = v
        cfg.comp[i].cp_type = 0 if len(cps) == 1 else 1
        cfg.comp[i].ml_type = 0 if len(cps) == 1 else 1
    cfg.eos = 0 if eos == "IDEAL" else 1
    cfg.flux = FLUX[flux]
    cfg.limiter = 1 if limiter else 0
    cfg.model = 1 if model == "NS" else 0
    for i in range(6):
        cfg.bnd_type[i] = BND[bnd[i]]
        if bnd_param is not None and bnd_param[i] is not None:
            for j, v in enumerate(bnd_param[i]):
                cfg.bnd_param[i][j] = v
    for i in range(3):
        cfg.grav[i] = grav[i]
    cfg.cfl = cfl
    if transport is not None:
        for i, tr in enumerate(transport):
            for k, v in tr.items():
                setattr(cfg.comp[i], k, v)
    cfg.use_visc, cfg.use_diff, cfg.t_ref = int(use_visc), int(use_diff), t_ref
    cfg.n_react = len(reactions)
    for j, (left, right, a, e, n) in enumerate(reactions):
        for m in range(3):
            cfg.react[j].left[m] = left[m]
            cfg.react[j].right[m] = right[m]
        cfg.react[j].a, cfg.react[j].e, cfg.react[j].n = a, e, n
    cfg.chem_parity = int(chem_parity)
    cfg.chem_energy_molar = int(chem_energy_molar)
    cfg.visc_br2_eta = visc_br2_eta
    return cfg


class OracleSim:
    """One oracle simulation over a fixed leaf set."""

    def __init__(self, cfg, leaves, trees=None):
        """leaves: int32 [n][4] of the unit-cube forest; or trees = (tree_vertices
        [nt][8][3], tree_bnd [nt][6], level) for a multi-tree forest refined
        uniformly (leaves is then ignored)."""
        self.cfg = cfg
        L = lib()
        if trees is None:
            self.leaves = np.ascontiguousarray(leaves, dtype=np.int32)
            self.n = len(self.leaves)
            self.h = L.or_create(C.byref(cfg), self.n,
                                 self.leaves.ctypes.data_as(C.POINTER(C.c_int32)))
        else:
            tv, tb, level = trees
            self._tv = np.ascontiguousarray(tv, dtype=np.float64)
            self._tb = np.ascontiguousarray(tb, dtype=np.int32)
            self.h = L.or_create_trees(C.byref(cfg), len(self._tv), _dptr(self._tv),
                                       self._tb.ctypes.data_as(C.POINTER(C.c_int32)), level)
            self.n = len(self._tv) * 8 ** level
            self.leaves = None
        self.nv = L.or_nvar(self.h)
        err = L.or_error(self.h)
        if err:
            raise RuntimeError(err.decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def check(self):
        err = lib().or_error(self.h)
        if err:
            raise RuntimeError(err.decode())

    def set_state(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.n, self.nv, 4)
        lib().or_set_state(self.h, _dptr(u))

    def get_state(self):
        u = np.zeros((self.n, self.nv, 4))
        lib().or_get_state(self.h, _dptr(u))
        return u

    def residual(self):
        out = np.zeros((self.n, self.nv, 4))
        lib().or_conv_residual(self.h, _dptr(out))
        self.check()
        return out

    def ns_residual(self):
        out = np.zeros((self.n, self.nv, 4))
        lib().or_ns_residual(self.h, _dptr(out))
        self.check()
        return out

    def ns_aux(self):
        """tau (xx yy zz xy xz yz) and q (x y z) coefficients: [n][9][4]"""
        out = np.zeros((self.n, 9, 4))
        lib().or_ns_aux(self.h, _dptr(out))
        return out

    def ns_aux_own(self):
        """BR2: the auxiliary field with the interior face terms from the own trace, [n][9][4]"""
        out = np.zeros((self.n, 9, 4))
        lib().or_ns_aux_own(self.h, _dptr(out))
        return out

    def ns_visc(self):
        """the viscous integrals alone: ru rv rw re, [n][4][4]"""
        out = np.zeros((self.n, 4, 4))
        lib().or_ns_visc(self.h, _dptr(out))
        self.check()
        return out

    def ns_diff_coeffs(self):
        out = np.zeros((self.n, self.nv - 4))
        lib().or_ns_diff_coeffs(self.h, _dptr(out))
        return out

    def ns_chem_apply(self, dt):
        out = np.zeros(self.n)
        lib().or_ns_chem_apply(self.h, dt, _dptr(out))
        return out

    def limiter(self):
        lib().or_limiter(self.h)
        self.check()

    def run(self, dt, nsteps):
        lib().or_run(self.h, dt, nsteps)
        self.check()

    def faces(self):
        nf = lib().or_n_faces(self.h)
        out = np.zeros((nf, 16), dtype=np.int32)
        lib().or_faces(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)))
        return out

    def geom(self, leaf):
        out = np.zeros(580)
        lib().or_geom(self.h, leaf, _dptr(out))
        o = 0
        res = {}
        for name, shape in (("n", (6, 3)), ("fc", (6, 3)), ("face_gp", (6, 6, 3)),
                            ("face_gw", (6, 6)), ("face_gj", (6, 6)), ("quad_gp", (8, 3)),
                            ("quad_gw", (8,)), ("quad_gj", (8,)), ("area", (6,)),
                            ("volume", ()), ("c", (3,)), ("dh", (3,)), ("a", (4, 4)),
                            ("a_inv", (4, 4))):
            size = int(np.prod(shape)) if shape else 1
            res[name] = out[o:o + size].reshape(shape) if shape else out[o]
            o += size
        return res

    def amr_indicator(self):
        err2, rho = np.zeros(self.n), np.zeros(self.n)
        lib().or_amr_indicator(self.h, _dptr(err2), _dptr(rho))
        self.check()
        return err2, rho

    def amr_flags(self, err2, rho, max_err, min_level, max_level):
        fl = np.zeros(self.n, dtype=np.int8)
        lib().or_amr_flags(self.h, _dptr(np.ascontiguousarray(err2)), _dptr(np.ascontiguousarray(rho)),
                           max_err, min_level, max_level, fl.ctypes.data_as(C.POINTER(C.c_int8)))
        return fl

    def conservation_sums(self):
        out = np.zeros(self.nv)
        lib().or_conservation_sums(self.h, _dptr(out))
        return out

    def cfl_dt(self, mode=1):
        return lib().or_cfl_dt(self.h, mode)


def prim_from(cfg, r, u, v, w, p, c=None):
    """Primitive state (r, u, v, w, p) completed through the oracle's EOS."""
    P = Prim()
    P.r, P.u, P.v, P.w = r, u, v, w
    nc = cfg.n_comp
    cc = [1.0] if c is None else list(c)
    for i in range(nc):
        P.c[i] = cc[i]
    # e from (r, p): flag 1 needs gam/cv, which the EOS computes from t
    P.t = 300.0
    P.p = p
    lib().or_eos(C.byref(cfg), C.byref(P), 1)
    P.e_tot = P.e + 0.5 * (u * u + v * v + w * w)
    lib().or_eos(C.byref(cfg), C.byref(P), 4)
    return P


def flux(cfg, kind, pl, pr, n):
    arr = (Prim * 2)(pl, pr)
    nn = np.ascontiguousarray(n, dtype=np.float64)
    q = np.zeros(4 + cfg.n_comp)
    lib().or_flux(C.byref(cfg), FLUX[kind], arr, _dptr(nn), _dptr(q))
    return q


def cons_to_prim(cfg, u):
    c = Cons()
    c.ru, c.rv, c.rw, c.re = u[0], u[1], u[2], u[3]
    for i in range(cfg.n_comp):
        c.rc[i] = u[4 + i]
    p = Prim()
    lib().or_cons_to_prim(C.byref(cfg), C.byref(c), C.byref(p))
    return p


def matr_inv(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros((4, 4))
    lib().or_matr_inv(_dptr(a), _dptr(out))
    return out


def rim(rb, pb, ub, vb, wb, re, pe, ue, ve, we, gb, ge=None):
    ge = gb if ge is None else ge
    outs = [C.c_double() for _ in range(6)]
    lib().or_rim(*[C.byref(o) for o in outs], rb, pb, ub, vb, wb, re, pe, ue, ve, we, gb, ge)
    return [o.value for o in outs]   # RI, EI, PI, UI, VI, WI


def amr_project(nv, old_leaves, u_old, new_leaves, mp):
    ol = np.ascontiguousarray(old_leaves, dtype=np.int32)
    nl = np.ascontiguousarray(new_leaves, dtype=np.int32)
    uo = np.ascontiguousarray(u_old, dtype=np.float64)
    mp = np.ascontiguousarray(mp, dtype=np.int64)
    out = np.zeros((len(nl), nv, 4))
    lib().or_amr_project(nv, len(ol), ol.ctypes.data_as(C.POINTER(C.c_int32)), _dptr(uo), len(nl),
                         nl.ctypes.data_as(C.POINTER(C.c_int32)), mp.ctypes.data_as(C.POINTER(C.c_int64)),
                         _dptr(out))
    return out
