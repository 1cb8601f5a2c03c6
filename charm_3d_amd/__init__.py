"""charm_3d_amd -- MI355X-native DG residual / time-step path of CHARM_3D.

The product is the C-ABI library libcharm3d.so (include/charm_gpu.h): a C
host octree with its flattening plus hand-written gfx950 HIP kernels.  This
module is a thin ctypes binding over that ABI for tests and benchmarks; it
never computes the path itself and raises if the library is missing.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libcharm3d.so")

MAX_COMP = 8
FP64, FP32, FP32_MIX = 0, 1, 2
FLUX = {"LF": 0, "HLLC": 1, "GODUNOV": 2, "CD": 3}
BND = {"INLET": 0, "OUTLET": 1, "WALL_SLIP": 2, "WALL_NO_SLIP": 3, "MASS_FLOW": 4,
       "PERIODIC": 100}
MODEL = {"EULER": 0, "NS": 1}
EOS = {"IDEAL": 0, "MIX": 1}
GAS_R = 8.31446261815324
ROOT_LEN = 1 << 19


class CharmError(RuntimeError):
    pass


class CompDesc(C.Structure):
    _fields_ = [("m", C.c_double), ("cp_type", C.c_int), ("ml_type", C.c_int),
                ("n_cp", C.c_int), ("cp", C.c_double * 8), ("ml0", C.c_double),
                ("kp0", C.c_double), ("t0", C.c_double), ("ts", C.c_double), ("h0", C.c_double),
                ("kp_type", C.c_int), ("sig", C.c_double), ("ek", C.c_double)]


class ReactionDesc(C.Structure):
    _fields_ = [("left", C.c_int * 3), ("right", C.c_int * 3), ("a", C.c_double),
                ("e", C.c_double), ("n", C.c_double)]


class Config(C.Structure):
    _fields_ = [("precision", C.c_int), ("model", C.c_int), ("flux", C.c_int),
                ("limiter", C.c_int), ("eos", C.c_int), ("n_comp", C.c_int),
                ("comp", CompDesc * MAX_COMP), ("bnd_type", C.c_int * 6),
                ("bnd_param", (C.c_double * 16) * 6), ("grav", C.c_double * 3),
                ("cfl", C.c_double), ("dt", C.c_double), ("tmax", C.c_double),
                ("log_period", C.c_int), ("chunk_elems", C.c_int), ("chunk_faces", C.c_int),
                ("max_err", C.c_double), ("min_level", C.c_int), ("max_level", C.c_int),
                ("refine_period", C.c_int), ("use_visc", C.c_int), ("use_diff", C.c_int),
                ("t_ref", C.c_double), ("n_react", C.c_int), ("react", ReactionDesc * 4),
                ("chem_parity", C.c_int), ("chem_energy_molar", C.c_int),
                ("visc_br2_eta", C.c_double), ("write_period", C.c_int),
                ("comp_name", (C.c_char * 32) * MAX_COMP), ("out_prefix", C.c_char * 256)]


class RegionDesc(C.Structure):
    _fields_ = [("id", C.c_int), ("cell_type", C.c_int), ("mat_id", C.c_int),
                ("name", C.c_char * 64), ("v", C.c_double * 3), ("t", C.c_double),
                ("p", C.c_double), ("grav", C.c_double * 3), ("c", C.c_double * MAX_COMP)]


class Task(C.Structure):
    """charm_task_t: what task.yaml holds."""
    _fields_ = [("cfg", Config), ("repartition_period", C.c_int),
                ("bnd_name", (C.c_char * 64) * 6), ("n_regions", C.c_int),
                ("region", RegionDesc * 16), ("mesh_files_type", C.c_char * 32),
                ("mesh_name", C.c_char * 256), ("mesh_periodic", C.c_int * 3)]


class ConnPatch(C.Structure):
    """charm_conn_patch_t: one $PhysicalNames entry."""
    _fields_ = [("dim", C.c_int), ("id", C.c_int), ("name", C.c_char * 64)]


_lib = None


def lib():
    """Load libcharm3d.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CharmError("libcharm3d.so not built: run `python -m charm_3d_amd.build` "
                         "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, i64, i32p, dp = C.c_void_p, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    sig = {
        "charm_gpu_last_error": (C.c_char_p, []),
        "charm_gpu_device_count": (C.c_int, [ip]),
        "charm_level_mass_inverse": (C.c_int, [C.c_int, dp, dp]),
        "charm_forest_new_uniform": (C.c_int, [C.c_int, ip, C.POINTER(vp)]),
        "charm_forest_new_leaves": (C.c_int, [i64, i32p, ip, C.POINTER(vp)]),
        "charm_forest_destroy": (C.c_int, [vp]),
        "charm_forest_num_leaves": (i64, [vp]),
        "charm_forest_get_leaves": (C.c_int, [vp, i32p]),
        "charm_forest_adapt": (C.c_int, [vp, C.POINTER(C.c_int8), C.c_int, C.c_int,
                                         C.POINTER(C.POINTER(C.c_int64)), C.POINTER(C.c_int64)]),
        "charm_forest_balance": (C.c_int, [vp]),
        "charm_forest_flatten": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(vp)]),
        "charm_mesh_destroy": (C.c_int, [vp]),
        "charm_mesh_num_faces": (i64, [vp]),
        "charm_mesh_num_chunks": (i64, [vp]),
        "charm_mesh_num_chunk_faces": (i64, [vp]),
        "charm_mesh_get_chunks": (C.c_int, [vp, i32p, i32p, i32p, i32p, C.POINTER(C.c_uint16)]),
        "charm_mesh_get_faces": (C.c_int, [vp, i32p]),
        "charm_mesh_get_neighbors": (C.c_int, [vp, i32p]),
        "charm_gpu_create": (C.c_int, [C.POINTER(Config), C.POINTER(vp)]),
        "charm_gpu_destroy": (C.c_int, [vp]),
        "charm_gpu_upload_mesh": (C.c_int, [vp, vp]),
        "charm_gpu_set_state": (C.c_int, [vp, dp, i64]),
        "charm_gpu_get_state": (C.c_int, [vp, dp, i64]),
        "charm_gpu_residual": (C.c_int, [vp, dp, i64]),
        "charm_gpu_limiter_bj": (C.c_int, [vp]),
        "charm_gpu_run": (C.c_int, [vp, C.c_double, C.c_int]),
        "charm_gpu_step": (C.c_int, [vp, C.c_int, C.c_double, C.c_int]),
        "charm_gpu_recommended_chunks": (C.c_int, [C.POINTER(Config), ip, ip]),
        "charm_gpu_time_kernels": (C.c_int, [vp, C.c_double, C.c_int, dp, dp]),
        "charm_gpu_time_copy": (C.c_int, [vp, C.c_int, dp, dp]),
        "charm_gpu_get_dt": (C.c_int, [vp, C.c_int, dp]),
        "charm_gpu_conservation_sums": (C.c_int, [vp, dp]),
        "charm_gpu_state_flag": (C.c_int, [vp, ip]),
        "charm_gpu_synchronize": (C.c_int, [vp]),
        "charm_gpu_last_timing": (C.c_int, [vp, dp, dp]),
        "charm_gpu_flux_batch": (C.c_int, [vp, C.c_int, i64, dp, dp, dp, dp]),
        "charm_write_vtk": (C.c_int, [C.c_char_p, i64, i32p, dp, C.POINTER(Config)]),
        "charm_write_solution": (None, [vp]),
        "charm_task_load_yaml": (C.c_int, [C.c_char_p, C.POINTER(Task)]),
        "charm_conn_read_msh": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
        "charm_conn_destroy": (None, [vp]),
        "charm_conn_info": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
        "charm_conn_get": (C.c_int, [vp, dp, i32p, i32p, C.POINTER(C.c_int8), i32p, i32p]),
        "charm_conn_patch": (C.c_int, [vp, C.c_int, C.POINTER(ConnPatch)]),
        "charm_conn_bind_task": (C.c_int, [vp, C.POINTER(Task), i32p, i32p]),
        "charm_task_forest": (C.c_int, [C.POINTER(Task), C.POINTER(vp)]),
        "charm_gmesh_build": (C.c_int, [i64, dp, i64, i32p, i32p, C.c_int, C.POINTER(vp)]),
        "charm_gmesh_destroy": (None, [vp]),
        "charm_gmesh_info": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "charm_gmesh_leaves": (C.c_int, [vp, i32p]),
        "charm_gmesh_geometry": (C.c_int, [vp, dp]),
        "charm_gmesh_faces": (C.c_int, [vp, i32p]),
        "charm_gmesh_leaf_corners": (C.c_int, [vp, dp]),
        "charm_gmesh_write_vtk": (C.c_int, [C.c_char_p, vp, dp, C.POINTER(Config)]),
        "charm_task_gmesh": (C.c_int, [C.POINTER(Task), C.c_char_p, C.POINTER(vp)]),
        "charm_task_initial_state_gmesh": (C.c_int, [C.POINTER(Task), vp, dp, i64]),
        "charm_gmesh_timesteps": (C.c_int, [vp, vp, C.POINTER(Config)]),
        "charm_gpu_hllc_branch_flips": (C.c_int, [vp, vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                                  C.POINTER(C.c_int64)]),
        "charm_gpu_upload_gmesh": (C.c_int, [vp, vp]),
        "charm_task_initial_state": (C.c_int, [C.POINTER(Task), vp, dp, i64]),
        "charm_gpu_amr_indicator": (C.c_int, [vp, dp, dp, i64]),
        "charm_gpu_ns_aux": (C.c_int, [vp, dp, i64]),
        "charm_gpu_objective": (C.c_int, [vp, C.c_double, C.c_int, dp, dp, dp]),
        "charm_gpu_stage_vjp": (C.c_int, [vp, C.c_double, dp, dp, dp, i64]),
        "charm_gpu_flux_vjp_batch": (C.c_int, [vp, C.c_int, i64, dp, dp, dp, dp, dp, dp]),
        "charm_gpu_adjoint": (C.c_int, [vp, C.c_double, C.c_int, C.c_int, dp, dp, C.c_int, dp, dp, dp,
                                        dp, dp]),
        "charm_gpu_ns_chem": (C.c_int, [vp, C.c_double, dp, i64]),
        "charm_gpu_remesh": (C.c_int, [vp, vp, C.POINTER(C.c_int64), i64]),
        "charm_amr_flags": (C.c_int, [i64, i32p, dp, dp, C.c_double, C.c_int, C.c_int,
                                      C.POINTER(C.c_int8)]),
        "charm_amr_adjoint_weights": (C.c_int, [i64, C.c_int, i32p, dp, dp]),
        "charm_adapt_weighted": (C.c_int, [vp, dp, i64]),
        "charm_adapt": (None, [vp, vp, vp]),
        "charm_ctx_error": (C.c_int, [vp]),
        "charm_ctx_create": (C.c_int, [vp, C.POINTER(Config), C.POINTER(vp)]),
        "charm_ctx_destroy": (C.c_int, [vp]),
        "charm_ctx_gpu": (vp, [vp]),
        "charm_model_ns_timestep_single": (None, [vp, dp, vp, vp]),
        "charm_model_euler_timestep_single": (None, [vp, dp, vp, vp]),
        "charm_model_ns_get_dt": (C.c_double, [vp]),
        "charm_model_euler_get_dt": (C.c_double, [vp]),
        "charm_timesteps": (None, [vp]),
        "charm_limiter": (None, [vp, vp, vp]),
        "charm_limiter_bj": (None, [vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().charm_gpu_last_error()
        raise CharmError("charm error %d: %s" % (rc, msg.decode() if msg else "?"))


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i32ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _periodic(per):
    arr = (C.c_int * 3)(*[1 if p else 0 for p in per])
    return arr


def make_config(precision=FP32, model="NS", flux="HLLC", limiter=False, eos="IDEAL",
                comp_m=(GAS_R,), comp_cp=(3.5,), bnd=("OUTLET",) * 6, bnd_param=None,
                grav=(0.0, 0.0, 0.0), cfl=0.3, dt=0.0, tmax=0.0, chunk_elems=0, chunk_faces=0,
                max_err=0.0, min_level=0, max_level=0, refine_period=0, transport=None,
                use_visc=False, use_diff=False, t_ref=298.15, reactions=(), chem_parity=False,
                chem_energy_molar=False, visc_br2_eta=0.0, write_period=0, comp_names=(),
                out_prefix=""):
    """charm_config_t from task.yaml-like keys.  comp_cp entries are a constant
    cp or a tuple of power-series coefficients (polynomial cp, which as in
    the reference takes effect with Sutherland viscosity, ml_type 1);
    transport: per species dict of charm_comp_desc_t fields (ml0, kp0, t0,
    ts, h0, ml_type, kp_type, sig, ek); reactions: (left[3], right[3], a, e, n);
    visc_br2_eta > 0: BR2 viscous flux with that penalty factor (0: BR1)."""
    cfg = Config()
    cfg.precision = precision
    cfg.model = MODEL[model]
    cfg.flux = FLUX[flux]
    cfg.limiter = 1 if limiter else 0
    cfg.eos = EOS[eos]
    cfg.n_comp = len(comp_m)
    for i, (m, cp) in enumerate(zip(comp_m, comp_cp)):
        cfg.comp[i].m = m
        cps = tuple(cp) if isinstance(cp, (tuple, list)) else (cp,)
        cfg.comp[i].n_cp = len(cps)
        for j, v in enumerate(cps):
            cfg.comp[i].cp[j] = v
        cfg.comp[i].cp_type = 0 if len(cps) == 1 else 1
        cfg.comp[i].ml_type = 0 if len(cps) == 1 else 1
    for i in range(6):
        cfg.bnd_type[i] = BND[bnd[i]]
        if bnd_param is not None and bnd_param[i] is not None:
            for j, v in enumerate(bnd_param[i]):
                cfg.bnd_param[i][j] = v
    for i in range(3):
        cfg.grav[i] = grav[i]
    cfg.cfl, cfg.dt, cfg.tmax = cfl, dt, tmax
    cfg.chunk_elems, cfg.chunk_faces = chunk_elems, chunk_faces
    cfg.max_err, cfg.min_level, cfg.max_level = max_err, min_level, max_level
    cfg.refine_period = refine_period
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
    cfg.write_period = write_period
    for i, nm in enumerate(comp_names):
        cfg.comp_name[i].value = nm.encode()[:31]
    cfg.out_prefix = out_prefix.encode()[:255]
    return cfg


def load_task(path):
    """charm_task_load_yaml: a Task from a task.yaml file."""
    t = Task()
    check(lib().charm_task_load_yaml(str(path).encode(), C.byref(t)))
    return t


def task_forest(task):
    """charm_task_forest: the unit-cube forest at MIN_LEVEL."""
    h = C.c_void_p()
    check(lib().charm_task_forest(C.byref(task), C.byref(h)))
    return Forest(h, tuple(task.mesh_periodic))


def task_initial_state(task, forest):
    """charm_task_initial_state: FP64 [n][V][4]."""
    n = forest.n
    u = np.zeros((n, 4 + task.cfg.n_comp, 4))
    check(lib().charm_task_initial_state(C.byref(task), forest.h, _dptr(u), n))
    return u


class GMesh:
    """charm_gmesh_build: a conforming multi-tree forest refined uniformly to
    `level`, with every leaf's trilinear geometry (FP64, host).  vertices
    [nv][3], tree_to_vertex [nt][8] (z-order corners), tree_bnd [nt][6] (the
    boundary slot of a tree face, -1 where it meets another tree)."""

    def __init__(self, vertices, tree_to_vertex, tree_bnd, level, _handle=None):
        if _handle is not None:
            self.h = _handle
            n, nf = C.c_int64(), C.c_int64()
            check(lib().charm_gmesh_info(self.h, C.byref(n), C.byref(nf)))
            self.n, self.nf = n.value, nf.value
            return
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        t2v = np.ascontiguousarray(tree_to_vertex, dtype=np.int32)
        tb = np.ascontiguousarray(tree_bnd, dtype=np.int32)
        self.h = C.c_void_p()
        check(lib().charm_gmesh_build(len(v), _dptr(v), len(t2v), _i32ptr(t2v), _i32ptr(tb), level,
                                      C.byref(self.h)))
        n, nf = C.c_int64(), C.c_int64()
        check(lib().charm_gmesh_info(self.h, C.byref(n), C.byref(nf)))
        self.n, self.nf = n.value, nf.value

    def leaves(self):
        """[n][5] tree, x, y, z, level"""
        out = np.zeros((self.n, 5), dtype=np.int32)
        check(lib().charm_gmesh_leaves(self.h, _i32ptr(out)))
        return out

    def geometry(self):
        """dict of centre [n][3], dh [n][3], volume [n], a_inv [n][4][4]"""
        g = np.zeros((self.n, 23))
        check(lib().charm_gmesh_geometry(self.h, _dptr(g)))
        return {"c": g[:, 0:3], "dh": g[:, 3:6], "volume": g[:, 6], "a_inv": g[:, 7:23].reshape(-1, 4, 4)}

    @classmethod
    def from_task(cls, task, directory):
        """charm_task_gmesh: a Gmsh task's mesh (mesh.name relative to
        `directory`) bound to the task's boundaries and regions, at
        MIN_LEVEL."""
        h = C.c_void_p()
        check(lib().charm_task_gmesh(C.byref(task), str(directory).encode(), C.byref(h)))
        return cls(None, None, None, None, _handle=h)

    def corners(self):
        """[n][8][3] leaf corners in z-order"""
        out = np.zeros((self.n, 8, 3))
        check(lib().charm_gmesh_leaf_corners(self.h, _dptr(out)))
        return out

    def write_vtk(self, name, u, cfg):
        """charm_gmesh_write_vtk: <name>_0000.vtu + <name>.pvtu"""
        uu = np.ascontiguousarray(u, dtype=np.float64)
        check(lib().charm_gmesh_write_vtk(str(name).encode(), self.h, _dptr(uu), C.byref(cfg)))

    def initial_state(self, task):
        """charm_task_initial_state_gmesh: FP64 [n][V][4], each leaf the
        region of its tree"""
        u = np.zeros((self.n, 4 + task.cfg.n_comp, 4))
        check(lib().charm_task_initial_state_gmesh(C.byref(task), self.h, _dptr(u), self.n))
        return u

    def faces(self):
        """[nf][4]: side-0 leaf, side-1 leaf or -1 - slot, side-0 face, side-1 face"""
        out = np.zeros((self.nf, 4), dtype=np.int32)
        check(lib().charm_gmesh_faces(self.h, _i32ptr(out)))
        return out

    def close(self):
        if self.h:
            lib().charm_gmesh_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Connectivity:
    """charm_conn_read_msh: a Gmsh 2.x hexahedral mesh as multi-tree
    connectivity (vertices, z-order tree corners, tree-to-tree face links with
    p4est's face + 6 * orientation encoding, boundary patch and region ids)."""

    def __init__(self, path):
        self.h = C.c_void_p()
        check(lib().charm_conn_read_msh(str(path).encode(), C.byref(self.h)))
        nv, nt, npt = C.c_int64(), C.c_int64(), C.c_int()
        check(lib().charm_conn_info(self.h, C.byref(nv), C.byref(nt), C.byref(npt)))
        self.num_vertices, self.num_trees = nv.value, nt.value
        self.vertices = np.zeros((nv.value, 3))
        self.tree_to_vertex = np.zeros((nt.value, 8), dtype=np.int32)
        self.tree_to_tree = np.zeros((nt.value, 6), dtype=np.int32)
        self.tree_to_face = np.zeros((nt.value, 6), dtype=np.int8)
        self.tree_to_patch = np.zeros((nt.value, 6), dtype=np.int32)
        self.tree_region = np.zeros(nt.value, dtype=np.int32)
        check(lib().charm_conn_get(self.h, _dptr(self.vertices), _i32ptr(self.tree_to_vertex),
                                   _i32ptr(self.tree_to_tree),
                                   self.tree_to_face.ctypes.data_as(C.POINTER(C.c_int8)),
                                   _i32ptr(self.tree_to_patch), _i32ptr(self.tree_region)))
        self.patches = []
        for i in range(npt.value):
            p = ConnPatch()
            check(lib().charm_conn_patch(self.h, i, C.byref(p)))
            self.patches.append((p.dim, p.id, p.name.decode()))

    def bind_task(self, task):
        """(tree_bnd [nt][6] boundary slot or -1, tree_reg [nt] region index)"""
        tb = np.zeros((self.num_trees, 6), dtype=np.int32)
        tr = np.zeros(self.num_trees, dtype=np.int32)
        check(lib().charm_conn_bind_task(self.h, C.byref(task), _i32ptr(tb), _i32ptr(tr)))
        return tb, tr

    def close(self):
        if self.h:
            lib().charm_conn_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_vtk(name, leaves, u, cfg):
    """charm_write_vtk: <name>_0000.vtu + <name>.pvtu of leaves [n][4] and
    the FP64 state u [n][V][4]."""
    leaves = np.ascontiguousarray(leaves, dtype=np.int32)
    u = np.ascontiguousarray(u, dtype=np.float64)
    n = len(leaves)
    check(lib().charm_write_vtk(name.encode(), n, leaves.ctypes.data_as(C.POINTER(C.c_int32)),
                                u.ctypes.data_as(C.POINTER(C.c_double)), C.byref(cfg)))


def level_mass_inverse(level):
    """FP64 mass matrix of a level-`level` cube and its Gauss-Jordan inverse."""
    a, ai = np.zeros((4, 4)), np.zeros((4, 4))
    check(lib().charm_level_mass_inverse(level, _dptr(a), _dptr(ai)))
    return a, ai


def recommended_chunks(cfg):
    ce, cf = C.c_int(), C.c_int()
    check(lib().charm_gpu_recommended_chunks(C.byref(cfg), C.byref(ce), C.byref(cf)))
    return ce.value, cf.value


class Mesh:
    """Flattened leaves (charm_mesh_t)."""

    def __init__(self, handle, n):
        self.h = handle
        self.n = n

    def __del__(self):
        if getattr(self, "h", None):
            lib().charm_mesh_destroy(self.h)
            self.h = None

    @property
    def num_faces(self):
        return lib().charm_mesh_num_faces(self.h)

    @property
    def num_chunks(self):
        return lib().charm_mesh_num_chunks(self.h)

    @property
    def num_chunk_faces(self):
        return lib().charm_mesh_num_chunk_faces(self.h)

    def chunks(self):
        """(ch_ebeg, ch_fbeg, cf4, slot_ptr, slots) of charm_mesh_get_chunks"""
        nch, ncf = self.num_chunks, self.num_chunk_faces
        n = self.n
        eb = np.zeros(nch + 1, dtype=np.int32)
        fb = np.zeros(nch + 1, dtype=np.int32)
        cf4 = np.zeros((ncf, 4), dtype=np.int32)
        sp = np.zeros(n + 1, dtype=np.int32)
        L = lib()
        # slot count = slot_ptr[n]; read the pointers first through a full call
        sl = np.zeros(24 * n + 1, dtype=np.uint16)
        check(L.charm_mesh_get_chunks(self.h, _i32ptr(eb), _i32ptr(fb), _i32ptr(cf4), _i32ptr(sp),
                                      sl.ctypes.data_as(C.POINTER(C.c_uint16))))
        return eb, fb, cf4, sp, sl[:sp[-1]]

    def faces(self):
        out = np.zeros((self.num_faces, 4), dtype=np.int32)
        check(lib().charm_mesh_get_faces(self.h, _i32ptr(out)))
        return out

    def neighbors(self):
        out = np.zeros((self.n, 6), dtype=np.int32)
        check(lib().charm_mesh_get_neighbors(self.h, _i32ptr(out)))
        return out


class Forest:
    """Host linear octree on the unit cube (charm_forest_t)."""

    def __init__(self, handle, periodic):
        self.h = handle
        self.periodic = tuple(bool(p) for p in periodic)

    @classmethod
    def uniform(cls, level, periodic=(False, False, False)):
        h = C.c_void_p()
        check(lib().charm_forest_new_uniform(level, _periodic(periodic), C.byref(h)))
        return cls(h, periodic)

    @classmethod
    def from_leaves(cls, leaves, periodic=(False, False, False)):
        lv = np.ascontiguousarray(leaves, dtype=np.int32)
        h = C.c_void_p()
        check(lib().charm_forest_new_leaves(len(lv), _i32ptr(lv), _periodic(periodic), C.byref(h)))
        return cls(h, periodic)

    def __del__(self):
        if getattr(self, "h", None):
            lib().charm_forest_destroy(self.h)
            self.h = None

    @property
    def n(self):
        return lib().charm_forest_num_leaves(self.h)

    def leaves(self):
        out = np.zeros((self.n, 4), dtype=np.int32)
        check(lib().charm_forest_get_leaves(self.h, _i32ptr(out)))
        return out

    def balance(self):
        check(lib().charm_forest_balance(self.h))

    def adapt(self, flags, max_level, min_level=0):
        fl = np.ascontiguousarray(flags, dtype=np.int8)
        mp = C.POINTER(C.c_int64)()
        nm = C.c_int64()
        check(lib().charm_forest_adapt(self.h, fl.ctypes.data_as(C.POINTER(C.c_int8)),
                                       max_level, min_level, C.byref(mp), C.byref(nm)))
        out = np.ctypeslib.as_array(mp, shape=(nm.value,)).copy()
        C.CDLL(None).free(C.cast(mp, C.c_void_p))
        return out

    def flatten(self, chunk_elems=128, chunk_faces=512):
        h = C.c_void_p()
        check(lib().charm_forest_flatten(self.h, chunk_elems, chunk_faces, C.byref(h)))
        return Mesh(h, self.n)


class GpuContext:
    """charm_gpu_ctx_t: one HIP stream, mesh and state on the device."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.nv = 4 + cfg.n_comp
        self.h = C.c_void_p()
        check(lib().charm_gpu_create(C.byref(cfg), C.byref(self.h)))
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            lib().charm_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def upload_mesh(self, mesh):
        check(lib().charm_gpu_upload_mesh(self.h, mesh.h))
        self.n = mesh.n

    def upload_gmesh(self, gmesh):
        """a multi-tree mesh (GMesh) in place of the octree"""
        check(lib().charm_gpu_upload_gmesh(self.h, gmesh.h))
        self.n = gmesh.n

    def hllc_branch_flips(self, other, per_bit=False):
        """charm_gpu_hllc_branch_flips: (interior face points, points whose
        HLLC branch differs between this context's state and other's[, the
        points where each of the 9 branch conditions differs])"""
        a, b = C.c_int64(), C.c_int64()
        bits = (C.c_int64 * 9)()
        check(lib().charm_gpu_hllc_branch_flips(self.h, other.h, C.byref(a), C.byref(b), bits))
        return (a.value, b.value, list(bits)) if per_bit else (a.value, b.value)

    def gmesh_timesteps(self, gmesh, cfg):
        """charm_gmesh_timesteps: the time loop to cfg.tmax on the uploaded
        multi-tree mesh (dt = cfg.dt, VTK every cfg.write_period steps)"""
        check(lib().charm_gmesh_timesteps(self.h, gmesh.h, C.byref(cfg)))

    def set_state(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.shape == (self.n, self.nv, 4), u.shape
        check(lib().charm_gpu_set_state(self.h, _dptr(u), self.n))

    def get_state(self):
        out = np.zeros((self.n, self.nv, 4))
        check(lib().charm_gpu_get_state(self.h, _dptr(out), self.n))
        return out

    def residual(self):
        out = np.zeros((self.n, self.nv, 4))
        check(lib().charm_gpu_residual(self.h, _dptr(out), self.n))
        return out

    def limiter_bj(self):
        check(lib().charm_gpu_limiter_bj(self.h))

    def run(self, dt, nsteps):
        check(lib().charm_gpu_run(self.h, dt, nsteps))

    def step(self, model, dt, nsteps):
        check(lib().charm_gpu_step(self.h, MODEL[model], dt, nsteps))

    def get_dt(self, mode=1):
        dt = C.c_double()
        check(lib().charm_gpu_get_dt(self.h, mode, C.byref(dt)))
        return dt.value

    def conservation_sums(self):
        out = np.zeros(self.nv)
        check(lib().charm_gpu_conservation_sums(self.h, _dptr(out)))
        return out

    def time_kernels(self, dt, reps):
        a, b = C.c_double(), C.c_double()
        check(lib().charm_gpu_time_kernels(self.h, dt, reps, C.byref(a), C.byref(b)))
        return a.value, b.value

    def time_copy(self, reps):
        a, b = C.c_double(), C.c_double()
        check(lib().charm_gpu_time_copy(self.h, reps, C.byref(a), C.byref(b)))
        return a.value, b.value

    def last_ms(self):
        a, b = C.c_double(), C.c_double()
        check(lib().charm_gpu_last_timing(self.h, C.byref(a), C.byref(b)))
        return a.value

    def synchronize(self):
        check(lib().charm_gpu_synchronize(self.h))

    def flux_batch(self, kind, prim_l, prim_r, normals):
        pl = np.ascontiguousarray(prim_l, dtype=np.float64)
        pr = np.ascontiguousarray(prim_r, dtype=np.float64)
        nn = np.ascontiguousarray(normals, dtype=np.float64)
        n = len(pl)
        out = np.zeros((n, self.nv))
        check(lib().charm_gpu_flux_batch(self.h, FLUX[kind], n, _dptr(pl), _dptr(pr),
                                         _dptr(nn), _dptr(out)))
        return out

    def flux_vjp_batch(self, kind, prim_l, prim_r, normals, qbar):
        """charm_gpu_flux_vjp_batch: (d(qbar . q)/d prim_l, d(qbar . q)/d prim_r)"""
        pl = np.ascontiguousarray(prim_l, dtype=np.float64)
        pr = np.ascontiguousarray(prim_r, dtype=np.float64)
        nn = np.ascontiguousarray(normals, dtype=np.float64)
        qb = np.ascontiguousarray(qbar, dtype=np.float64)
        bl, br = np.zeros_like(pl), np.zeros_like(pr)
        check(lib().charm_gpu_flux_vjp_batch(self.h, FLUX[kind], len(pl), _dptr(pl), _dptr(pr), _dptr(nn),
                                             _dptr(qb), _dptr(bl), _dptr(br)))
        return bl, br


    def amr_indicator(self):
        """(err2, rho_c) per leaf: the adaptation's FP64 error estimate."""
        err2, rho_c = np.zeros(self.n), np.zeros(self.n)
        check(lib().charm_gpu_amr_indicator(self.h, _dptr(err2), _dptr(rho_c), self.n))
        return err2, rho_c

    def ns_aux(self):
        """BR1 auxiliary tau (xx yy zz xy xz yz) and q (x y z): [n][9][4]."""
        out = np.zeros((self.n, 9, 4))
        check(lib().charm_gpu_ns_aux(self.h, _dptr(out), self.n))
        return out

    def ns_chem(self, dt):
        """Chemistry applied to the state; returns the energy source [n]."""
        out = np.zeros(self.n)
        check(lib().charm_gpu_ns_chem(self.h, dt, _dptr(out), self.n))
        return out

    def objective(self, dt, nsteps, lo, hi):
        """J = sum over the steps of dt * int_box p (forward pass)."""
        lo_, hi_ = np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64)
        j = C.c_double()
        check(lib().charm_gpu_objective(self.h, dt, nsteps, _dptr(lo_), _dptr(hi_), C.byref(j)))
        return j.value

    def stage_vjp(self, dt, u, x):
        """charm_gpu_stage_vjp: S_u^T x for S(u) = u - dt A^-1 int(u)."""
        u_ = np.ascontiguousarray(u, dtype=np.float64)
        x_ = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros_like(u_)
        check(lib().charm_gpu_stage_vjp(self.h, dt, _dptr(u_), _dptr(x_), _dptr(out), self.n))
        return out

    def adjoint(self, dt, nsteps, lo, hi, ckpt=20, theta_face=-1):
        """(J, dJ/du0 [n][V][4], dJ/d(inlet u, T, p), forward ms, adjoint ms)"""
        lo_, hi_ = np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64)
        j, fm, am = C.c_double(), C.c_double(), C.c_double()
        g = np.zeros((self.n, self.nv, 4))
        th = np.zeros(3)
        check(lib().charm_gpu_adjoint(self.h, dt, nsteps, ckpt, _dptr(lo_), _dptr(hi_), theta_face,
                                      C.byref(j), _dptr(g), _dptr(th), C.byref(fm), C.byref(am)))
        return j.value, g, th, fm.value, am.value

    def remesh(self, mesh, amap):
        """Upload the adapted mesh and project the state with the old->new map."""
        mp = np.ascontiguousarray(amap, dtype=np.int64)
        check(lib().charm_gpu_remesh(self.h, mesh.h, mp.ctypes.data_as(C.POINTER(C.c_int64)), len(mp)))
        self.n = mesh.n


def amr_flags(leaves, err2, rho_c, max_err, min_level, max_level):
    """Refine / coarsen / keep flags of the reference's criteria (host C)."""
    lv = np.ascontiguousarray(leaves, dtype=np.int32)
    e2 = np.ascontiguousarray(err2, dtype=np.float64)
    rc = np.ascontiguousarray(rho_c, dtype=np.float64)
    out = np.zeros(len(lv), dtype=np.int8)
    check(lib().charm_amr_flags(len(lv), _i32ptr(lv), _dptr(e2), _dptr(rc), max_err, min_level,
                                max_level, out.ctypes.data_as(C.POINTER(C.c_int8))))
    return out


def amr_adjoint_weights(leaves, grad):
    """charm_amr_adjoint_weights: per-leaf refinement weights from dJ/du0."""
    lv = np.ascontiguousarray(leaves, dtype=np.int32)
    g = np.ascontiguousarray(grad, dtype=np.float64)
    w = np.zeros(len(lv))
    check(lib().charm_amr_adjoint_weights(len(lv), g.shape[1], _i32ptr(lv), _dptr(g), _dptr(w)))
    return w


class Solver:
    """charm_ctx_t over a forest: the reference-named entry points
    (charm_timesteps, charm_model_*_timestep_single, charm_adapt)."""

    def __init__(self, forest, cfg):
        self.forest, self.cfg = forest, cfg
        self.nv = 4 + cfg.n_comp
        self.h = C.c_void_p()
        check(lib().charm_ctx_create(forest.h, C.byref(cfg), C.byref(self.h)))
        self.gpu = lib().charm_ctx_gpu(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().charm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_state(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.shape == (self.forest.n, self.nv, 4), u.shape
        check(lib().charm_gpu_set_state(self.gpu, _dptr(u), len(u)))

    def get_state(self):
        out = np.zeros((self.forest.n, self.nv, 4))
        check(lib().charm_gpu_get_state(self.gpu, _dptr(out), len(out)))
        return out

    def write_solution(self):
        """charm_write_solution: VTK files of the current state."""
        lib().charm_write_solution(self.forest.h)
        err = lib().charm_ctx_error(self.h)
        if err:
            check(err)

    def run(self, dt, nsteps):
        """nsteps device steps with a fixed dt (no adaptation)."""
        check(lib().charm_gpu_run(self.gpu, dt, nsteps))

    def last_ms(self):
        a, b = C.c_double(), C.c_double()
        check(lib().charm_gpu_last_timing(self.gpu, C.byref(a), C.byref(b)))
        return a.value

    def get_dt(self, mode=1):
        dt = C.c_double()
        check(lib().charm_gpu_get_dt(self.gpu, mode, C.byref(dt)))
        return dt.value

    def conservation_sums(self):
        out = np.zeros(self.nv)
        check(lib().charm_gpu_conservation_sums(self.gpu, _dptr(out)))
        return out

    def timestep(self):
        """One charm_model_{euler,ns}_timestep_single; returns dt."""
        dt = C.c_double()
        fn = (lib().charm_model_euler_timestep_single if self.cfg.model == MODEL["EULER"]
              else lib().charm_model_ns_timestep_single)
        fn(self.forest.h, C.byref(dt), None, None)
        return dt.value

    def adapt(self):
        """charm_adapt; raises if the adaptation failed."""
        lib().charm_adapt(self.forest.h, None, None)
        err = lib().charm_ctx_error(self.h)
        if err:
            check(err)

    def adapt_weighted(self, weight):
        """charm_adapt_weighted: the estimator scaled by weight [n] per leaf."""
        w = np.ascontiguousarray(weight, dtype=np.float64)
        check(lib().charm_adapt_weighted(self.forest.h, _dptr(w), len(w)))

    def adjoint(self, dt, nsteps, lo, hi, ckpt=20, theta_face=-1):
        """charm_gpu_adjoint on this context's state: (J, dJ/du0, dJ/dtheta)"""
        lo_, hi_ = np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64)
        n, nv = self.forest.n, 4 + self.cfg.n_comp
        j, fm, am = C.c_double(), C.c_double(), C.c_double()
        g, th = np.zeros((n, nv, 4)), np.zeros(3)
        check(lib().charm_gpu_adjoint(self.gpu, dt, nsteps, ckpt, _dptr(lo_), _dptr(hi_), theta_face,
                                      C.byref(j), _dptr(g), _dptr(th), C.byref(fm), C.byref(am)))
        return j.value, g, th

    def timesteps(self):
        lib().charm_timesteps(self.forest.h)
        err = lib().charm_ctx_error(self.h)
        if err:
            check(err)


def device_count():
    n = C.c_int()
    rc = lib().charm_gpu_device_count(C.byref(n))
    return n.value if rc == 0 else 0
