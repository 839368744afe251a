"""ctypes bindings for the CPU oracle (liborc.so) and the reference-pinning
harness (_ref/libcorofem_ref.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker and never as the thing
measured or shipped.
"""
import ctypes as C
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
REF_PATH = os.path.join(HERE, "_ref", "libcorofem_ref.so")

RM_ID, RM_QR, RM_MSVD, RM_DAPD = 0, 1, 2, 3
IT_QIE_V, IT_FIE_DX = 0, 1
MM_C_WRP, MM_C_LCM, MM_C_LCMH = 1, 2, 3
DM_TRUNCATED, DM_EXACT = 0, 1

_f = C.POINTER(C.c_float)
_u = C.POINTER(C.c_uint32)
_b = C.POINTER(C.c_uint8)
_d = C.POINTER(C.c_double)
DOC_INVALID = 12  # direction-of-collapse byte 4*vit0 + vit1; (3,0) is invalid


def fp(a):
    return a.ctypes.data_as(_f)


def up(a):
    return a.ctypes.data_as(_u)


class SolidParams(C.Structure):
    _fields_ = [
        ("fixed_dt", C.c_float), ("density", C.c_float), ("young_modulus", C.c_float),
        ("poisson_ratio", C.c_float), ("damping_ratio", C.c_float),
        ("rayleigh_freq1", C.c_float), ("rayleigh_freq2", C.c_float),
        ("gravity", C.c_float * 3), ("cg_rel_epsilon", C.c_float), ("cg_abs_epsilon", C.c_float),
        ("cg_max_iter", C.c_uint32), ("rm", C.c_int), ("degenerate_threshold", C.c_float),
        ("warm_start", C.c_int), ("reduce_fp64", C.c_int),
        ("integrator", C.c_int), ("nr_max_iter", C.c_uint32), ("nr_abs_epsilon", C.c_float),
        ("mm", C.c_int), ("ecie_k_factor", C.c_float), ("dm", C.c_int), ("time_scale", C.c_float),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("num_fixed_steps", C.c_uint32), ("cg_iters", C.c_int32), ("cg_rel", C.c_float),
        ("cg_abs", C.c_float), ("num_degenerate", C.c_uint32), ("sum_degenerate_detF", C.c_float),
        ("elastic_energy", C.c_float), ("min_detF", C.c_float),
        ("nr_iters", C.c_int32), ("nr_rel", C.c_float), ("nr_abs", C.c_float),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "orc_box_counts": (None, [C.c_uint32] * 3 + [_u, _u]),
        "orc_mesh_box": (None, [C.c_float] * 3 + [C.c_uint32] * 3 + [_f, _u]),
        "orc_box2_counts": (None, [C.c_uint32] * 2 + [_u, _u]),
        "orc_mesh_box2": (None, [C.c_float] * 2 + [C.c_uint32] * 2 + [_f, _u]),
        "orc_det3": (C.c_float, [_f]),
        "orc_inverse3": (None, [_f, _f]),
        "orc_pd3": (None, [_f, C.c_float, _f, _u]),
        "orc_svd3": (None, [_f, _f, _f, _f]),
        "orc_svd_rotation3": (None, [_f, _f]),
        "orc_svd3_sqrt3": (None, [_f, _f, _f, _f]),
        "orc_pd2": (None, [_f, C.c_float, _f]),
        "orc_svd_rotation2": (None, [_f, C.c_float, C.c_float, _f]),
        "orc_compute_K": (C.c_float, [_f, C.c_float, C.c_float, _f]),
        "orc_compute_K2": (C.c_float, [_f, C.c_float, C.c_float, _f]),
        "orc_sse_dot": (C.c_float, [_f, _f, C.c_uint32]),
        "orc_tetset_create": (C.c_void_p, [C.c_uint32, _f, C.c_uint32, _u, C.c_float, C.c_float]),
        "orc_tetset_destroy": (None, [C.c_void_p]),
        "orc_tetset_total_volume": (C.c_float, [C.c_void_p]),
        "orc_tetset_element": (None, [C.c_void_p, C.c_uint32, _f, _f, _f]),
        "orc_begin_evaluation": (None, [C.c_void_p, _f, _f, C.c_int, C.c_float, _f, _f, _u, _f]),
        "orc_tetset_get_doc": (None, [C.c_void_p, _b]),
        "orc_tetset_set_doc": (None, [C.c_void_p, _b]),
        "orc_tetset_set_dapd": (None, [C.c_void_p, C.c_float, C.c_float, C.c_float]),
        "orc_tetset_num_fallback": (C.c_uint32, [C.c_void_p]),
        "orc_gsl_cubic": (C.c_int, [C.c_double, C.c_double, C.c_double, _d]),
        "orc_solve_eq3": (C.c_int, [C.c_float] * 4 + [_f]),
        "orc_compute_doc": (C.c_uint8, [_f, _f, C.c_float, C.c_float, _f]),
        "orc_r_dapd_degenerate": (C.c_int, [C.c_float, C.c_float, C.c_uint8, _f, C.c_float, _f,
                                            C.c_float, C.c_float, C.c_float, _f]),
        "orc_set_rotations": (None, [C.c_void_p, _f]),
        "orc_f_e": (None, [C.c_void_p, _f, _f]),
        "orc_df_x": (None, [C.c_void_p, _f, _f]),
        "orc_V": (C.c_float, [C.c_void_p, _f]),
        "orc_solid_create": (C.c_void_p, [C.c_uint32, _f, C.c_uint32, _u, C.POINTER(SolidParams)]),
        "orc_solid_destroy": (None, [C.c_void_p]),
        "orc_solid_set_knc": (None, [C.c_void_p, C.c_uint32, _u]),
        "orc_solid_set_knc_targets": (None, [C.c_void_p, C.c_uint32, _u, _f]),
        "orc_solid_fixed_step": (None, [C.c_void_p, C.c_float, C.POINTER(StepStats)]),
        "orc_solid_step": (None, [C.c_void_p, C.c_float, C.POINTER(StepStats)]),
        "orc_solid_total_time": (C.c_float, [C.c_void_p]),
        "orc_rayleigh": (None, [C.c_float, C.c_float, C.c_float, _f]),
        "orc_solid2_step": (None, [C.c_void_p, C.c_float, C.POINTER(StepStats)]),
        "orc_solid2_total_time": (C.c_float, [C.c_void_p]),
        "orc_solids_fixed_step_parallel": (None, [C.POINTER(C.c_void_p), C.c_uint32, C.c_uint32]),
        "orc_solids2_fixed_step_parallel": (None, [C.POINTER(C.c_void_p), C.c_uint32, C.c_uint32]),
        "orc_solid_get_state": (None, [C.c_void_p, _f, _f]),
        "orc_solid_set_state": (None, [C.c_void_p, _f, _f]),
        "orc_solid_node_mass": (C.c_float, [C.c_void_p]),
        "orc_solid_rayleigh": (None, [C.c_void_p, _f]),
        "orc_tetset_set_material": (None, [C.c_void_p, C.c_int, C.c_float]),
        "orc_tetset_set_dm": (None, [C.c_void_p, C.c_int]),
        "orc_dR": (C.c_int, [_f, _f, _f, _f]),
        "orc_f_d": (None, [C.c_void_p, _f, _f, C.c_float, C.c_float, _f]),
        "orc_multi_update": (C.c_uint32, [C.POINTER(C.c_void_p), C.c_uint32, C.c_float, _f,
                                          C.POINTER(StepStats)]),
        "orc_solid2_create": (C.c_void_p, [C.c_uint32, _f, C.c_uint32, _u, C.POINTER(SolidParams), C.c_float]),
        "orc_solid2_destroy": (None, [C.c_void_p]),
        "orc_solid2_set_knc": (None, [C.c_void_p, C.c_uint32, _u]),
        "orc_solid2_fixed_step": (None, [C.c_void_p, C.c_float, C.POINTER(StepStats)]),
        "orc_solid2_get_state": (None, [C.c_void_p, _f, _f]),
        "orc_solid2_set_state": (None, [C.c_void_p, _f, _f]),
        "orc_solid2_node_masses": (None, [C.c_void_p, _f]),
        "orc_solid2_total_area": (C.c_float, [C.c_void_p]),
        "orc_fan_weighted_areas": (None, [C.c_uint32, C.c_uint32, _u, _f, _f]),
        "orc_solid2_begin_evaluation": (None, [C.c_void_p, _f, _f, _f, _f]),
        "orc_solid2_get_noc": (None, [C.c_void_p, _b]),
        "orc_solid2_set_dapd": (None, [C.c_void_p, C.c_float, C.c_float, C.c_float]),
        "orc_solid2_num_fallback": (C.c_uint32, [C.c_void_p]),
        "orc_compute_noc2": (C.c_uint8, [_f, _f, C.c_float, C.c_float, _f]),
        "orc_r_dapd_degenerate2": (None, [C.c_float, C.c_uint8, _f, C.c_float, _f, C.c_float, C.c_float,
                                          C.c_float, _f]),
        "orc_solid2_f_e": (None, [C.c_void_p, _f, _f]),
        "orc_solid2_df_x": (None, [C.c_void_p, _f, _f]),
        "orc_solid2_V": (C.c_float, [C.c_void_p, _f]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def _load_ref():
    if not os.path.exists(REF_PATH):
        return None
    lib = C.CDLL(REF_PATH)
    sig = {
        "ref_compute_K": (C.c_float, [_f, C.c_float, C.c_float, _f]),
        "ref_compute_volume": (C.c_float, [_f]),
        "ref_inv_dm": (None, [_f, _f]),
        "ref_compute_K2": (C.c_float, [_f, C.c_float, C.c_float, _f]),
        "ref_det3": (C.c_float, [_f]),
        "ref_inverse3": (None, [_f, _f]),
        "ref_pd3": (None, [_f, C.c_float, _f]),
        "ref_svd3": (None, [_f, _f, _f, _f]),
        "ref_R_svd3": (None, [_f, C.c_float, C.c_float, _f]),
        "ref_pd2": (None, [_f, C.c_float, _f]),
        "ref_R_svd2": (None, [_f, C.c_float, C.c_float, _f]),
        "ref_dot": (C.c_float, [_f, _f, C.c_uint32]),
        "ref_has_sse_path": (C.c_int, []),
        "ref_rayleigh": (None, [C.c_float, C.c_float, C.c_float, _f]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB = None
_REF = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def ref():
    """The reference-pinning harness, or None where it was not built."""
    global _REF
    if _REF is None:
        _REF = _load_ref()
    return _REF


# ---------------------------------------------------------------- helpers
def mesh_box(half, counts):
    nx, ny, nz = counts
    L = lib()
    nn = C.c_uint32()
    nt = C.c_uint32()
    L.orc_box_counts(nx, ny, nz, C.byref(nn), C.byref(nt))
    pos = np.zeros(3 * nn.value, np.float32)
    tets = np.zeros(4 * nt.value, np.uint32)
    L.orc_mesh_box(half[0], half[1], half[2], nx, ny, nz, fp(pos), up(tets))
    return pos, tets


def mesh_box2(half, counts):
    nx, ny = counts
    L = lib()
    nn = C.c_uint32()
    nt = C.c_uint32()
    L.orc_box2_counts(nx, ny, C.byref(nn), C.byref(nt))
    pos = np.zeros(2 * nn.value, np.float32)
    tris = np.zeros(3 * nt.value, np.uint32)
    L.orc_mesh_box2(half[0], half[1], nx, ny, fp(pos), up(tris))
    return pos, tris


class TetSet:
    """Oracle element set (EditableTetrahedronElementSet3 + its cache)."""

    def __init__(self, ref_pos, tets, young, poisson):
        self.ref_pos = np.ascontiguousarray(ref_pos, np.float32)
        self.tets = np.ascontiguousarray(tets, np.uint32)
        self.N = self.ref_pos.size // 3
        self.T = self.tets.size // 4
        self.h = lib().orc_tetset_create(self.N, fp(self.ref_pos), self.T, up(self.tets), young, poisson)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tetset_destroy(self.h)
            self.h = None

    def total_volume(self):
        return lib().orc_tetset_total_volume(self.h)

    def element(self, e):
        K = np.zeros(144, np.float32)
        D = np.zeros(9, np.float32)
        v = np.zeros(1, np.float32)
        lib().orc_tetset_element(self.h, e, fp(K), fp(D), fp(v))
        return K.reshape(12, 12), D.reshape(3, 3), float(v[0])

    def begin_evaluation(self, x, rm=RM_MSVD, thr=0.1, prev_x=None):
        x = np.ascontiguousarray(x, np.float32)
        px = None if prev_x is None else np.ascontiguousarray(prev_x, np.float32)
        R = np.zeros(9 * self.T, np.float32)
        dF = np.zeros(self.T, np.float32)
        nd = C.c_uint32()
        sd = np.zeros(1, np.float32)
        lib().orc_begin_evaluation(self.h, None if px is None else fp(px), fp(x), rm, thr, fp(R), fp(dF),
                                   C.byref(nd), fp(sd))
        return R, dF, nd.value, float(sd[0])

    def doc(self):
        out = np.zeros(self.T, np.uint8)
        lib().orc_tetset_get_doc(self.h, out.ctypes.data_as(_b))
        return out

    def set_doc(self, doc):
        doc = np.ascontiguousarray(doc, np.uint8)
        lib().orc_tetset_set_doc(self.h, doc.ctypes.data_as(_b))

    def set_dapd(self, l_factor=1.0, nl_factor=0.0, nl_exponent=2.0):
        lib().orc_tetset_set_dapd(self.h, l_factor, nl_factor, nl_exponent)

    def num_fallback(self):
        return lib().orc_tetset_num_fallback(self.h)

    def set_rotations(self, R):
        R = np.ascontiguousarray(R, np.float32)
        lib().orc_set_rotations(self.h, fp(R))

    def set_dm(self, dm):
        lib().orc_tetset_set_dm(self.h, dm)

    def set_material(self, mm, ecie_k_factor=1.0):
        lib().orc_tetset_set_material(self.h, mm, ecie_k_factor)

    def f_d(self, v, inv_mass, c0, c1):
        v = np.ascontiguousarray(v, np.float32)
        im = np.ascontiguousarray(inv_mass, np.float32)
        f = np.zeros(3 * self.N, np.float32)
        lib().orc_f_d(self.h, fp(v), fp(im), c0, c1, fp(f))
        return f

    def f_e(self, x, f=None):
        x = np.ascontiguousarray(x, np.float32)
        f = np.zeros(3 * self.N, np.float32) if f is None else f
        lib().orc_f_e(self.h, fp(x), fp(f))
        return f

    def df_x(self, dx, df=None):
        dx = np.ascontiguousarray(dx, np.float32)
        df = np.zeros(3 * self.N, np.float32) if df is None else df
        lib().orc_df_x(self.h, fp(dx), fp(df))
        return df

    def V(self, x):
        x = np.ascontiguousarray(x, np.float32)
        return lib().orc_V(self.h, fp(x))


def rayleigh(f1, f2, zeta):
    """Rayleigh coefficients (c0, c1) of the leaves (LeafDSH_Solid3D_FEM.h:347-354)."""
    c = np.zeros(2, np.float32)
    lib().orc_rayleigh(f1, f2, zeta, fp(c))
    return c


def gsl_cubic(a, b, c):
    """Real roots of x^3 + a x^2 + b x + c (ascending), GSL's method in double."""
    r = np.zeros(3, np.float64)
    n = lib().orc_gsl_cubic(a, b, c, r.ctypes.data_as(_d))
    return r[:n]


def solve_eq3(a, b, c, d):
    """mal::GSolvePolynomialEq3 in float: (number of roots, the three slots)."""
    r = np.zeros(3, np.float32)
    n = lib().orc_solve_eq3(a, b, c, d, fp(r))
    return n, r


def compute_doc(a, b, volume, thr=0.1, with_toc=False):
    """ComputeDoC for the linear path a -> b (4x3 each): the DoC byte, and
    the time of collapse (NaN when none was found) with with_toc."""
    a = np.ascontiguousarray(a, np.float32).ravel()
    b = np.ascontiguousarray(b, np.float32).ravel()
    toc = np.full(1, np.nan, np.float32)
    doc = int(lib().orc_compute_doc(fp(a), fp(b), volume, thr, fp(toc)))
    return (doc, float(toc[0])) if with_toc else doc


def r_dapd_degenerate(det_f, thr, doc, p, volume, inv_dm, l_factor=1.0, nl_factor=0.0, nl_exponent=2.0):
    p = np.ascontiguousarray(p, np.float32).ravel()
    inv_dm = np.ascontiguousarray(inv_dm, np.float32).ravel()
    R = np.zeros(9, np.float32)
    ok = lib().orc_r_dapd_degenerate(det_f, thr, doc, fp(p), volume, fp(inv_dm), l_factor, nl_factor,
                                     nl_exponent, fp(R))
    return bool(ok), R.reshape(3, 3)


NOC_INVALID = 3  # 2D node of collapse: 0, 1, 2 or invalid


def compute_noc2(a, b, area, thr=0.1, with_toc=False):
    """ComputeNoC for the linear path a -> b (3x2 each)."""
    a = np.ascontiguousarray(a, np.float32).ravel()
    b = np.ascontiguousarray(b, np.float32).ravel()
    toc = np.full(1, np.nan, np.float32)
    noc = int(lib().orc_compute_noc2(fp(a), fp(b), area, thr, fp(toc)))
    return (noc, float(toc[0])) if with_toc else noc


def r_dapd_degenerate2(thr, noc, p, area, inv_dm, l_factor=1.0, nl_factor=0.0, nl_exponent=2.0):
    p = np.ascontiguousarray(p, np.float32).ravel()
    inv_dm = np.ascontiguousarray(inv_dm, np.float32).ravel()
    R = np.zeros(4, np.float32)
    lib().orc_r_dapd_degenerate2(thr, noc, fp(p), area, fp(inv_dm), l_factor, nl_factor, nl_exponent, fp(R))
    return R.reshape(2, 2)


def default_params(**kw):
    p = SolidParams()
    p.fixed_dt = 1.0 / 60.0
    p.density = 1000.0
    p.young_modulus = 1e5
    p.poisson_ratio = 0.3
    p.damping_ratio = 0.0
    p.rayleigh_freq1 = 0.1
    p.rayleigh_freq2 = 10000.0
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, -9.81, 0.0
    p.cg_rel_epsilon = 1e-4
    p.cg_abs_epsilon = 1e-6
    p.cg_max_iter = 50
    p.rm = RM_MSVD
    p.degenerate_threshold = 0.1
    p.warm_start = 1
    p.reduce_fp64 = 0
    p.integrator = IT_QIE_V
    p.nr_max_iter = 3
    p.nr_abs_epsilon = 1e-6
    p.mm = MM_C_WRP
    p.ecie_k_factor = 1.0
    p.dm = DM_TRUNCATED
    p.time_scale = 1.0
    for k, v in kw.items():
        if k == "gravity":
            for i in range(3):
                p.gravity[i] = v[i]
        else:
            setattr(p, k, v)
    return p


def step_parallel(solids, threads):
    """One fixed step of each of `solids` (all Solid or all Solid2) over `threads` host threads."""
    hs = (C.c_void_p * len(solids))(*[s.h for s in solids])
    if isinstance(solids[0], Solid2):
        lib().orc_solids2_fixed_step_parallel(hs, len(solids), int(threads))
    else:
        lib().orc_solids_fixed_step_parallel(hs, len(solids), int(threads))


class Multi:
    """Oracle multi-solid universe step (SS_AggregateDSH_Basic::MultiDSH_Solid3D_Update)."""

    def __init__(self, solids):
        self.solids = list(solids)
        self.acc = np.zeros(1, np.float32)

    def update(self, dt):
        hs = (C.c_void_p * len(self.solids))(*[s.h for s in self.solids])
        st = (StepStats * len(self.solids))()
        n = lib().orc_multi_update(hs, len(self.solids), dt, fp(self.acc), st)
        return n, list(st)


class Solid:
    """Oracle leaf solid (LeafDSH_Solid3D_FEM: QIE-v, or FIE-dx by params.integrator)."""

    def __init__(self, ref_pos, tets, params, knc=None, knc_targets=None):
        self.ref_pos = np.ascontiguousarray(ref_pos, np.float32)
        self.tets = np.ascontiguousarray(tets, np.uint32)
        self.N = self.ref_pos.size // 3
        self.params = params
        self.h = lib().orc_solid_create(self.N, fp(self.ref_pos), self.tets.size // 4, up(self.tets),
                                        C.byref(params))
        if knc is not None:
            k = np.ascontiguousarray(knc, np.uint32)
            if knc_targets is None:
                lib().orc_solid_set_knc(self.h, k.size, up(k))
            else:
                t = np.ascontiguousarray(knc_targets, np.float32)
                lib().orc_solid_set_knc_targets(self.h, k.size, up(k), fp(t))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_solid_destroy(self.h)
            self.h = None

    def fixed_step(self, dt=None):
        st = StepStats()
        lib().orc_solid_fixed_step(self.h, self.params.fixed_dt if dt is None else dt, C.byref(st))
        return st

    def step(self, dt):
        st = StepStats()
        lib().orc_solid_step(self.h, dt, C.byref(st))
        return st

    def total_time(self):
        return lib().orc_solid_total_time(self.h)

    def rayleigh(self):
        c = np.zeros(2, np.float32)
        lib().orc_solid_rayleigh(self.h, fp(c))
        return c

    def state(self):
        x = np.zeros(3 * self.N, np.float32)
        v = np.zeros(3 * self.N, np.float32)
        lib().orc_solid_get_state(self.h, fp(x), fp(v))
        return x, v

    def set_state(self, x, v):
        x = np.ascontiguousarray(x, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        lib().orc_solid_set_state(self.h, fp(x), fp(v))

    def node_mass(self):
        return lib().orc_solid_node_mass(self.h)


class Solid2:
    """Oracle 2D leaf solid (TriangleElementSet2 + LeafDSH_Solid2D_FEM, QIE-v,
    lumped masses)."""

    def __init__(self, ref_pos, tris, params, knc=None, thickness=1.0):
        self.ref_pos = np.ascontiguousarray(ref_pos, np.float32)
        self.tris = np.ascontiguousarray(tris, np.uint32)
        self.N = self.ref_pos.size // 2
        self.T = self.tris.size // 3
        self.params = params
        self.h = lib().orc_solid2_create(self.N, fp(self.ref_pos), self.T, up(self.tris), C.byref(params),
                                         float(thickness))
        if knc is not None:
            k = np.ascontiguousarray(knc, np.uint32)
            lib().orc_solid2_set_knc(self.h, k.size, up(k))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_solid2_destroy(self.h)
            self.h = None

    def fixed_step(self, dt=None):
        st = StepStats()
        lib().orc_solid2_fixed_step(self.h, self.params.fixed_dt if dt is None else dt, C.byref(st))
        return st

    def step(self, dt):
        st = StepStats()
        lib().orc_solid2_step(self.h, dt, C.byref(st))
        return st

    def total_time(self):
        return lib().orc_solid2_total_time(self.h)

    def state(self):
        x = np.zeros(2 * self.N, np.float32)
        v = np.zeros(2 * self.N, np.float32)
        lib().orc_solid2_get_state(self.h, fp(x), fp(v))
        return x, v

    def set_state(self, x, v):
        x = np.ascontiguousarray(x, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        lib().orc_solid2_set_state(self.h, fp(x), fp(v))

    def node_masses(self):
        m = np.zeros(self.N, np.float32)
        lib().orc_solid2_node_masses(self.h, fp(m))
        return m

    def total_area(self):
        return lib().orc_solid2_total_area(self.h)

    def begin_evaluation(self, x, prev_x=None):
        x = np.ascontiguousarray(x, np.float32)
        px = None if prev_x is None else np.ascontiguousarray(prev_x, np.float32)
        R = np.zeros(4 * self.T, np.float32)
        dF = np.zeros(self.T, np.float32)
        lib().orc_solid2_begin_evaluation(self.h, None if px is None else fp(px), fp(x), fp(R), fp(dF))
        return R, dF

    def noc(self):
        out = np.zeros(self.T, np.uint8)
        lib().orc_solid2_get_noc(self.h, out.ctypes.data_as(_b))
        return out

    def set_dapd(self, l_factor=1.0, nl_factor=0.0, nl_exponent=2.0):
        lib().orc_solid2_set_dapd(self.h, l_factor, nl_factor, nl_exponent)

    def num_fallback(self):
        return lib().orc_solid2_num_fallback(self.h)

    def f_e(self, x):
        x = np.ascontiguousarray(x, np.float32)
        f = np.zeros(2 * self.N, np.float32)
        lib().orc_solid2_f_e(self.h, fp(x), fp(f))
        return f

    def df_x(self, dx):
        dx = np.ascontiguousarray(dx, np.float32)
        df = np.zeros(2 * self.N, np.float32)
        lib().orc_solid2_df_x(self.h, fp(dx), fp(df))
        return df

    def V(self, x):
        x = np.ascontiguousarray(x, np.float32)
        return lib().orc_solid2_V(self.h, fp(x))
