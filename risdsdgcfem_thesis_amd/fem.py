"""Python mirror of the reference's element-set / solver / leaf interfaces,
over the corofem C-ABI.

  TetrahedronElementSet3  <-> ms::fem::EditableTetrahedronElementSet3 + its
                              IForce3 cache (ms/fem/TetrahedronElementSet3.h:29-261)
  ILSS_CG                 <-> ns::ILSS_CG (ns/ILSS_CG.h:19-151)
  LeafSolid3D             <-> ds::LeafDSH_Solid3D_FEM (ds/dsh/LeafDSH_Solid3D_FEM.h:68-81)
  Batch                   <-> many LeafDSH_Solid3D_FEM leaves stepped together (no
                              reference counterpart; the reference loops over leaves)
  LeafSolid2D, Batch2     <-> ds::LeafDSH_Solid2D_FEM + ms::fem::TriangleElementSet2, one
                              or many 2D solids (2N-float node vectors)

Node vectors are flat float32 arrays of 3N values. Device vectors are
DeviceArray objects (memory owned by libcorofem's allocator).
"""
import ctypes as C
import numpy as np

from ._capi import (lib, check, FemParams, SolidParams, StepStats, CgResult, BatchInfo, AV_FN, CONS_FN,
                    RM_MSVD, MM_C_WRP, DM_TRUNCATED)

_f = C.POINTER(C.c_float)
_u = C.POINTER(C.c_uint32)


def _fp(a):
    return a.ctypes.data_as(_f)


def _up(a):
    return a.ctypes.data_as(_u)


def device_count():
    n = C.c_int(0)
    rc = lib.cf_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    check(lib.cf_device_synchronize(), "cf_device_synchronize")


class DeviceArray:
    """A float32/uint32 buffer in device memory."""

    def __init__(self, n, dtype=np.float32):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.ptr = C.c_void_p()
        check(lib.cf_device_alloc(max(1, self.n) * self.dtype.itemsize, C.byref(self.ptr)), "cf_device_alloc")

    @classmethod
    def from_host(cls, a, dtype=np.float32):
        a = np.ascontiguousarray(a, dtype)
        d = cls(a.size, dtype)
        d.upload(a)
        return d

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size == self.n
        check(lib.cf_memcpy_htod(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes), "cf_memcpy_htod")

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        check(lib.cf_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes), "cf_memcpy_dtoh")
        return out

    def zero(self):
        check(lib.cf_memset_device(self.ptr, 0, self.nbytes), "cf_memset_device")

    def free(self):
        if self.ptr:
            lib.cf_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class TetShape:
    """A tet shape after EditableTetSolidShape3::EndEdition: vertices that some
    tetrahedron uses, positively oriented tetrahedra in layer order (the
    reference's element order), face neighbours (-1 = boundary), outward
    boundary faces, layers (first, count) and, per baked tetrahedron, its index
    in the input."""

    def __init__(self, handle):
        self.h = handle
        try:
            nv, nt, nbf, nl = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
            check(lib.cf_tetshape_info(handle, C.byref(nv), C.byref(nt), C.byref(nbf), C.byref(nl)),
                  "cf_tetshape_info")
            self.pos = np.empty(3 * nv.value, np.float32)
            self.tets = np.empty(4 * nt.value, np.uint32)
            ntid = np.empty(4 * nt.value, np.uint32)
            self.boundary_faces = np.empty(3 * nbf.value, np.uint32)
            self.layers = np.empty(2 * nl.value, np.uint32).reshape(-1, 2)
            self.src = np.empty(nt.value, np.uint32)
            check(lib.cf_tetshape_get(handle, _fp(self.pos), _up(self.tets), _up(ntid), _up(self.boundary_faces),
                                      _up(self.layers), _up(self.src)), "cf_tetshape_get")
            self.ntid = ntid.astype(np.int64)
            self.ntid[ntid == 0xFFFFFFFF] = -1
        except Exception:
            self.close()
            raise

    def close(self):
        if getattr(self, "h", None):
            lib.cf_tetshape_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def save(self, path, binary=True):
        """ShapeLibrary::Save of a one-shape library (binary or text)."""
        check(lib.cf_tetshape_save(self.h, str(path).encode(), int(bool(binary))), "cf_tetshape_save")

    @classmethod
    def load(cls, path):
        """The first TetSolidShape3 item of a shape library file (binary or text)."""
        h = C.c_void_p()
        check(lib.cf_tetshape_load(str(path).encode(), C.byref(h)), "cf_tetshape_load")
        return cls(h)

    @classmethod
    def bake(cls, pos, tets):
        p = np.ascontiguousarray(pos, np.float32)
        t = np.ascontiguousarray(tets, np.uint32)
        h = C.c_void_p()
        check(lib.cf_tetshape_bake(p.size // 3, _fp(p), t.size // 4, _up(t), C.byref(h)), "cf_tetshape_bake")
        return cls(h)

    @classmethod
    def from_tetgen(cls, base_name, natural_scale=0.0):
        h = C.c_void_p()
        check(lib.cf_tetshape_load_tetgen(str(base_name).encode(), float(natural_scale), C.byref(h)),
              "cf_tetshape_load_tetgen")
        return cls(h)


def mesh_box_baked(half_sizes, counts):
    """Make_TetSolidShape3_Box as the reference's scenes see it: the box mesh
    after EndEdition (tetrahedra in layer order)."""
    return TetShape.bake(*mesh_box(half_sizes, counts))


def mesh_box(half_sizes, counts):
    """The vertices and tetrahedra Make_TetSolidShape3_Box adds, in the order it
    adds them (before EndEdition reorders the tetrahedra by layer, which
    mesh_box_baked applies); counts = vertices per axis (cubes + 1)."""
    nx, ny, nz = (int(c) for c in counts)
    nn, nt = C.c_uint32(), C.c_uint32()
    check(lib.cf_mesh_box_counts(nx, ny, nz, C.byref(nn), C.byref(nt)), "cf_mesh_box_counts")
    pos = np.empty(3 * nn.value, np.float32)
    tets = np.empty(4 * nt.value, np.uint32)
    h = np.asarray(half_sizes, np.float32)
    check(lib.cf_mesh_box(_fp(h), nx, ny, nz, _fp(pos), _up(tets)), "cf_mesh_box")
    return pos, tets


def mesh_box2(half_sizes, counts):
    """Make_MeshSolidShape2_Box: counts = vertices per axis."""
    nx, ny = (int(c) for c in counts)
    nn, nt = C.c_uint32(), C.c_uint32()
    check(lib.cf_mesh_box2_counts(nx, ny, C.byref(nn), C.byref(nt)), "cf_mesh_box2_counts")
    pos = np.empty(2 * nn.value, np.float32)
    tris = np.empty(3 * nt.value, np.uint32)
    h = np.asarray(half_sizes, np.float32)
    check(lib.cf_mesh_box2(_fp(h), nx, ny, _fp(pos), _up(tris)), "cf_mesh_box2")
    return pos, tris


def rayleigh_coefficients(f1, f2, zeta):
    """(c0, c1) of LeafDSH_Solid3D_FEM::RebuildRayleighCoeffsFromFreqs."""
    c = np.zeros(2, np.float32)
    lib.cf_rayleigh_coeffs(float(f1), float(f2), float(zeta), _fp(c))
    return float(c[0]), float(c[1])


def default_fem_params(**kw):
    p = FemParams()
    lib.cf_default_fem_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def default_solid_params(fem=None, **kw):
    p = SolidParams()
    lib.cf_default_solid_params(C.byref(p))
    if fem is not None:
        p.fem = fem
    for k, v in kw.items():
        if k == "gravity":
            for i in range(3):
                p.gravity[i] = v[i]
        elif k in ("young_modulus", "poisson_ratio", "damping_ratio", "rm", "degenerate_threshold_det_f", "mm",
                   "ecie_k_factor", "dm", "rayleigh_freq1", "rayleigh_freq2", "dapd_l_factor", "dapd_nl_factor",
                   "dapd_nl_exponent"):
            setattr(p.fem, k, v)
        else:
            setattr(p, k, v)
    return p


def set_tile_elems(n):
    check(lib.cf_set_tile_elems(int(n)), "cf_set_tile_elems")


def set_tile_threads(n):
    check(lib.cf_set_tile_threads(int(n)), "cf_set_tile_threads")


def set_tile_persistent(mode):
    check(lib.cf_set_tile_persistent(int(mode)), "cf_set_tile_persistent")


def set_tile_alternate(level):
    check(lib.cf_set_tile_alternate(int(level)), "cf_set_tile_alternate")


def set_fused_front(enable):
    """1 (default): QIE-v leaf steps open with the fused BeginEvaluation + f_e + K y tile pass; 0: separate passes."""
    check(lib.cf_set_fused_front(int(enable)), "cf_set_fused_front")


def get_tile_elems():
    v = C.c_uint32()
    check(lib.cf_get_tile_elems(C.byref(v)), "cf_get_tile_elems")
    return v.value


def set_cg_workgroups(n):
    check(lib.cf_set_cg_workgroups(int(n)), "cf_set_cg_workgroups")


def get_cg_workgroups():
    v = C.c_uint32()
    check(lib.cf_get_cg_workgroups(C.byref(v)), "cf_get_cg_workgroups")
    return v.value


class TetrahedronElementSet3:
    """Element set + IForce3 cache. Vectors passed to the operator methods are
    DeviceArray objects of 3N floats; outputs accumulate (+=)."""

    def __init__(self, ref_pos, tets, params=None, inv_mass=None, incremental=False):
        """incremental: build through BeginEdition / AddElement per tet /
        EndEdition (TetrahedronElementSet3.cpp:2084-2166) instead of the bulk
        cf_tetset_create; the two give the same set."""
        self.ref_pos = np.ascontiguousarray(ref_pos, np.float32)
        self.tets = np.ascontiguousarray(tets, np.uint32)
        self.N = self.ref_pos.size // 3
        self.T = self.tets.size // 4
        self.params = params if params is not None else default_fem_params()
        self.h = C.c_void_p()
        if incremental:
            check(lib.cf_tetset_begin_edition(self.N, _fp(self.ref_pos), self.T, C.byref(self.params),
                                              C.byref(self.h)), "cf_tetset_begin_edition")
            for i, j, k, l in self.tets.reshape(-1, 4).tolist():
                check(lib.cf_tetset_add_element(self.h, i, j, k, l), "cf_tetset_add_element")
            check(lib.cf_tetset_end_edition(self.h), "cf_tetset_end_edition")
        else:
            check(lib.cf_tetset_create(self.N, _fp(self.ref_pos), self.T, _up(self.tets), C.byref(self.params),
                                       C.byref(self.h)), "cf_tetset_create")
        self.c = C.c_void_p()
        im = None if inv_mass is None else np.ascontiguousarray(inv_mass, np.float32)
        check(lib.cf_cache_create(self.h, None if im is None else _fp(im), C.byref(self.c)), "cf_cache_create")

    def close(self):
        if getattr(self, "c", None):
            lib.cf_cache_destroy(self.c)
            self.c = None
        if getattr(self, "h", None):
            lib.cf_tetset_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        nn, nt, v = C.c_uint32(), C.c_uint32(), C.c_float()
        check(lib.cf_tetset_info(self.h, C.byref(nn), C.byref(nt), C.byref(v)), "cf_tetset_info")
        return nn.value, nt.value, v.value

    def total_volume(self):
        return self.info()[2]

    def elements(self):
        dinv = np.empty(9 * self.T, np.float32)
        vol = np.empty(self.T, np.float32)
        check(lib.cf_tetset_get_elements(self.h, _fp(dinv), _fp(vol)), "cf_tetset_get_elements")
        return dinv.reshape(self.T, 3, 3), vol

    def set_params(self, params):
        self.params = params
        check(lib.cf_tetset_set_params(self.h, C.byref(params)), "cf_tetset_set_params")

    def begin_evaluation(self, x, prev_x=None, v=None, t=0.0):
        self._x, self._v, self._prev_x = x, v, prev_x  # the cache keeps these device pointers
        check(lib.cf_begin_evaluation(self.c, prev_x.ptr if prev_x else None, x.ptr, v.ptr if v else None,
                                      float(t)), "cf_begin_evaluation")

    def f_e(self, f):
        check(lib.cf_f_e(self.c, f.ptr), "cf_f_e")

    def df_x(self, dx, df):
        check(lib.cf_df_x(self.c, dx.ptr, df.ptr), "cf_df_x")

    def f_d(self, f):
        """+= the Rayleigh damping force of the evaluation's velocities (needs inv_mass)."""
        check(lib.cf_f_d(self.c, f.ptr), "cf_f_d")

    def V(self):
        e = C.c_double()
        check(lib.cf_V(self.c, C.byref(e)), "cf_V")
        return e.value

    def end_evaluation(self):
        check(lib.cf_end_evaluation(self.c), "cf_end_evaluation")

    def rotations(self):
        R = np.empty(9 * self.T, np.float32)
        d = np.empty(self.T, np.float32)
        check(lib.cf_cache_get_rotations(self.c, _fp(R), _fp(d)), "cf_cache_get_rotations")
        return R.reshape(self.T, 3, 3), d

    def stats(self):
        n, s, m = C.c_uint32(), C.c_float(), C.c_float()
        check(lib.cf_cache_stats(self.c, C.byref(n), C.byref(s), C.byref(m)), "cf_cache_stats")
        return n.value, s.value, m.value

    def doc(self):
        """Per-element direction-of-collapse bytes (4*vit0 + vit1; 12 = invalid)."""
        out = np.empty(self.T, np.uint8)
        check(lib.cf_cache_get_doc(self.c, out.ctypes.data_as(C.POINTER(C.c_uint8))), "cf_cache_get_doc")
        return out

    def set_doc(self, doc):
        doc = np.ascontiguousarray(doc, np.uint8)
        check(lib.cf_cache_set_doc(self.c, doc.ctypes.data_as(C.POINTER(C.c_uint8))), "cf_cache_set_doc")

    def num_fallback(self):
        """DAPD elements of the last evaluation left at R = Id."""
        n = C.c_uint32()
        check(lib.cf_cache_num_fallback(self.c, C.byref(n)), "cf_cache_num_fallback")
        return n.value

    def synchronize(self):
        check(lib.cf_cache_synchronize(self.c), "cf_cache_synchronize")


class ILSS_CG:
    """IIterativeLinearSystemSolver::Solve(D_Av, D_ApplyConstraints, x, b, rel, abs, max)
    with device vectors. `av(q_ptr, v_ptr)` and `apply_constraints(v_ptr)` receive raw
    device pointers (ctypes float pointers) and must enqueue work on `self.stream`."""

    def __init__(self, n):
        self.n = int(n)
        self.h = C.c_void_p()
        check(lib.cf_cg_create(self.n, C.byref(self.h)), "cf_cg_create")
        self.stream = lib.cf_cg_stream(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib.cf_cg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, av, apply_constraints, x, b, rel=1e-3, abs_eps=1e-6, max_iterations=20):
        cav = AV_FN(lambda user, q, v: av(q, v))
        ccons = CONS_FN(lambda user, v: apply_constraints(v)) if apply_constraints else CONS_FN(0)
        res = CgResult()
        check(lib.cf_cg_solve(self.h, cav, ccons, None, x.ptr, b.ptr, float(rel), float(abs_eps),
                              int(max_iterations), C.byref(res)), "cf_cg_solve")
        return res


def cg_solve_static(av, apply_constraints, x, b, rel=1e-3, abs_eps=1e-6, max_iterations=20, stream=None):
    """ILSS_CG::Solve as ILSS_CG.solve, without a solver handle: the work
    buffers are the library's static device scratch (cf_cg_solve_static),
    so nothing is allocated. The callbacks enqueue on `stream` (None: the
    null stream)."""
    cav = AV_FN(lambda user, q, v: av(q, v))
    ccons = CONS_FN(lambda user, v: apply_constraints(v)) if apply_constraints else CONS_FN(0)
    res = CgResult()
    check(lib.cf_cg_solve_static(stream, int(x.n), cav, ccons, None, x.ptr, b.ptr, float(rel), float(abs_eps),
                                 int(max_iterations), C.byref(res)), "cf_cg_solve_static")
    return res


class LeafSolid3D:
    """LeafDSH_Solid3D_FEM with device-resident state (QIE-v, or FIE-dx with
    params.integrator = IT_FIE_DX)."""

    def __init__(self, ref_pos, tets, params=None, knc=None, knc_targets=None):
        self.ref_pos = np.ascontiguousarray(ref_pos, np.float32)
        self.tets = np.ascontiguousarray(tets, np.uint32)
        self.N = self.ref_pos.size // 3
        self.T = self.tets.size // 4
        self.params = params if params is not None else default_solid_params()
        self.h = C.c_void_p()
        check(lib.cf_solid_create(self.N, _fp(self.ref_pos), self.T, _up(self.tets), C.byref(self.params),
                                  C.byref(self.h)), "cf_solid_create")
        if knc is not None:
            self.set_knc(knc, knc_targets)

    def close(self):
        if getattr(self, "h", None):
            lib.cf_solid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_knc(self, nids, targets=None):
        k = np.ascontiguousarray(nids, np.uint32)
        t = None if targets is None else np.ascontiguousarray(targets, np.float32)
        check(lib.cf_solid_set_knc(self.h, k.size, _up(k), None if t is None else _fp(t)), "cf_solid_set_knc")

    def fixed_step(self, dt=None):
        st = StepStats()
        check(lib.cf_solid_fixed_step(self.h, float(self.params.fixed_dt if dt is None else dt), C.byref(st)),
              "cf_solid_fixed_step")
        return st

    def step(self, dt):
        st = StepStats()
        check(lib.cf_solid_step(self.h, float(dt), C.byref(st)), "cf_solid_step")
        return st

    def state(self):
        x = np.empty(3 * self.N, np.float32)
        v = np.empty(3 * self.N, np.float32)
        check(lib.cf_solid_get_state(self.h, _fp(x), _fp(v)), "cf_solid_get_state")
        return x, v

    def set_state(self, x, v):
        x = np.ascontiguousarray(x, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        check(lib.cf_solid_set_state(self.h, _fp(x), _fp(v)), "cf_solid_set_state")

    def total_time(self):
        t = C.c_float()
        check(lib.cf_solid_total_time(self.h, C.byref(t)), "cf_solid_total_time")
        return t.value

    def set_graphs(self, on):
        """CG chunks replayed as recorded HIP graphs (default) or launched
        directly (cf_solid_set_graphs); results are the same."""
        check(lib.cf_solid_set_graphs(self.h, int(bool(on))), "cf_solid_set_graphs")

    def info(self):
        nn, nt, m = C.c_uint32(), C.c_uint32(), C.c_float()
        r = np.zeros(2, np.float32)
        check(lib.cf_solid_info(self.h, C.byref(nn), C.byref(nt), C.byref(m), _fp(r)), "cf_solid_info")
        return {"nodes": nn.value, "tets": nt.value, "node_mass": m.value, "rayleigh": r}

    def assemble_rhs(self, dt=None):
        b = np.empty(3 * self.N, np.float32)
        y = np.empty(3 * self.N, np.float32)
        check(lib.cf_solid_assemble_rhs(self.h, float(self.params.fixed_dt if dt is None else dt), _fp(b), _fp(y)),
              "cf_solid_assemble_rhs")
        return b, y

    def apply_A(self, d, factor_m, factor_k):
        d = np.ascontiguousarray(d, np.float32)
        q = np.empty(3 * self.N, np.float32)
        check(lib.cf_solid_apply_A(self.h, float(factor_m), float(factor_k), _fp(d), _fp(q)), "cf_solid_apply_A")
        return q

    def time_kx(self, reps=20):
        """(ms per tile-kernel launch, ms per gather launch) of one K*x."""
        a, b = C.c_double(), C.c_double()
        check(lib.cf_solid_time_kx(self.h, int(reps), C.byref(a), C.byref(b)), "cf_solid_time_kx")
        return a.value, b.value

    def time_df_x(self, reps=20):
        ms = C.c_double()
        check(lib.cf_solid_time_df_x(self.h, int(reps), C.byref(ms)), "cf_solid_time_df_x")
        return ms.value


class Batch:
    """Many solids stepped together (cf_batch_*): meshes are shared
    topologies, simulations carry their own parameters, KNC and state. Each
    simulation's step equals LeafSolid3D.fixed_step on the same inputs."""

    def __init__(self, hbm_budget_bytes=0, tile_threads=256, tile_elems=0, cg_chunk=8, use_graphs=True,
                 step_mode=None):
        self.h = C.c_void_p()
        check(lib.cf_batch_create(int(hbm_budget_bytes), C.byref(self.h)), "cf_batch_create")
        check(lib.cf_batch_set_tiles(self.h, int(tile_threads), int(tile_elems)), "cf_batch_set_tiles")
        check(lib.cf_batch_set_schedule(self.h, int(cg_chunk), int(bool(use_graphs))), "cf_batch_set_schedule")
        if step_mode is not None:
            self.set_step_mode(step_mode)
        self.mesh_nodes = {}

    def close(self):
        if getattr(self, "h", None):
            lib.cf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_schedule(self, cg_chunk=8, use_graphs=True):
        check(lib.cf_batch_set_schedule(self.h, int(cg_chunk), int(bool(use_graphs))), "cf_batch_set_schedule")

    def set_step_mode(self, mode):
        """BATCH_STEP_HOST or BATCH_STEP_DEVICE (default); results are the same."""
        check(lib.cf_batch_set_step_mode(self.h, int(mode)), "cf_batch_set_step_mode")

    def add_mesh(self, ref_pos, tets):
        r = np.ascontiguousarray(ref_pos, np.float32)
        t = np.ascontiguousarray(tets, np.uint32)
        mid = C.c_uint32()
        check(lib.cf_batch_add_mesh(self.h, r.size // 3, _fp(r), t.size // 4, _up(t), C.byref(mid)),
              "cf_batch_add_mesh")
        self.mesh_nodes[mid.value] = r.size // 3
        return mid.value

    def add_meshes(self, meshes, threads=0):
        """meshes: list of (ref_pos, tets); tiles built on host threads."""
        rs = [np.ascontiguousarray(r, np.float32) for r, _ in meshes]
        ts = [np.ascontiguousarray(t, np.uint32) for _, t in meshes]
        n = len(rs)
        nn = np.array([r.size // 3 for r in rs], np.uint32)
        nt = np.array([t.size // 4 for t in ts], np.uint32)
        rp = (_f * n)(*[_fp(r) for r in rs])
        tp = (_u * n)(*[_up(t) for t in ts])
        ids = np.zeros(n, np.uint32)
        check(lib.cf_batch_add_meshes(self.h, n, _up(nn), rp, _up(nt), tp, int(threads), _up(ids)),
              "cf_batch_add_meshes")
        for i, mid in enumerate(ids):
            self.mesh_nodes[int(mid)] = int(nn[i])
        return [int(i) for i in ids]

    def remove_mesh(self, mesh_id):
        check(lib.cf_batch_remove_mesh(self.h, int(mesh_id)), "cf_batch_remove_mesh")
        self.mesh_nodes.pop(mesh_id, None)

    def sim_bytes(self, mesh_id):
        b = C.c_uint64()
        check(lib.cf_batch_sim_bytes(self.h, int(mesh_id), C.byref(b)), "cf_batch_sim_bytes")
        return b.value

    def add_sim(self, mesh_id, params=None, knc=None, knc_targets=None, x=None, v=None):
        p = params if params is not None else default_solid_params()
        k = np.ascontiguousarray(knc if knc is not None else [], np.uint32)
        kt = None if knc_targets is None else np.ascontiguousarray(knc_targets, np.float32)
        xa = None if x is None else np.ascontiguousarray(x, np.float32)
        va = None if v is None else np.ascontiguousarray(v, np.float32)
        sid = C.c_uint32()
        check(lib.cf_batch_add_sim(self.h, int(mesh_id), C.byref(p), k.size, _up(k) if k.size else None,
                                   None if kt is None else _fp(kt), None if xa is None else _fp(xa),
                                   None if va is None else _fp(va), C.byref(sid)), "cf_batch_add_sim")
        return sid.value

    def remove_sim(self, sim_id):
        check(lib.cf_batch_remove_sim(self.h, int(sim_id)), "cf_batch_remove_sim")

    def set_active(self, sim_id, active=True):
        check(lib.cf_batch_set_active(self.h, int(sim_id), int(bool(active))), "cf_batch_set_active")

    def info(self):
        i = BatchInfo()
        check(lib.cf_batch_info_get(self.h, C.byref(i)), "cf_batch_info_get")
        return {k: getattr(i, k) for k, _ in BatchInfo._fields_}

    def active_ids(self):
        n = self.info()["num_active"]
        ids = np.zeros(max(n, 1), np.uint32)
        check(lib.cf_batch_active_ids(self.h, _up(ids)), "cf_batch_active_ids")
        return [int(i) for i in ids[:n]]

    def fixed_step(self, stats=True):
        """One fixed step of every active simulation; returns {sim_id: StepStats}."""
        if not stats:
            check(lib.cf_batch_fixed_step(self.h, None), "cf_batch_fixed_step")
            return None
        ids = self.active_ids()
        st = (StepStats * max(len(ids), 1))()
        check(lib.cf_batch_fixed_step(self.h, st), "cf_batch_fixed_step")
        return {sid: st[i] for i, sid in enumerate(ids)}

    def request_steps(self, sim_id, steps):
        check(lib.cf_batch_request_steps(self.h, int(sim_id), int(steps)), "cf_batch_request_steps")

    def round(self, iters=8):
        """One round of iteration-level batching; returns ({sim_id: steps done}, busy)."""
        ids = self.active_ids()
        done = np.zeros(max(len(ids), 1), np.uint32)
        busy = C.c_uint32()
        check(lib.cf_batch_round(self.h, int(iters), _up(done), C.byref(busy)), "cf_batch_round")
        return {sid: int(done[i]) for i, sid in enumerate(ids)}, busy.value

    def state(self, sim_id, n_nodes):
        x = np.empty(3 * n_nodes, np.float32)
        v = np.empty(3 * n_nodes, np.float32)
        check(lib.cf_batch_get_state(self.h, int(sim_id), _fp(x), _fp(v)), "cf_batch_get_state")
        return x, v

    def set_state(self, sim_id, x, v):
        x = np.ascontiguousarray(x, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        check(lib.cf_batch_set_state(self.h, int(sim_id), _fp(x), _fp(v)), "cf_batch_set_state")

    def time_kx(self, reps=20):
        ms = C.c_double()
        check(lib.cf_batch_time_kx(self.h, int(reps), C.byref(ms)), "cf_batch_time_kx")
        return ms.value

    def synchronize(self):
        check(lib.cf_batch_synchronize(self.h), "cf_batch_synchronize")


def serve(meshes, requests, params, hbm_budget_bytes, duration_s, drain=False, max_tick_tets=0, tile_threads=256,
          cg_chunk=8, final_positions=None, use_graphs=True, target_round_ms=0.0):
    """Runs the real-time serving loop (cf_serve_trace).

    meshes: list of (ref_pos, tets); requests: list of dicts with arrival (s),
    mesh (index), steps, E, nu, knc (node ids held at rest), sorted by
    arrival; params: base SolidParams (dt, density, gravity, CG settings).
    target_round_ms > 0: the tets cap starts at max_tick_tets and adapts so
    that a round takes about that long.
    final_positions: None or a list (one entry per request, None to skip) of
    float32 arrays that receive the final positions.
    Returns (result dict, per-step latencies in ms)."""
    from ._capi import ServeMesh, ServeRequest, ServeConfig, ServeResult
    keep = []
    ms = (ServeMesh * len(meshes))()
    for i, (r, t) in enumerate(meshes):
        r = np.ascontiguousarray(r, np.float32)
        t = np.ascontiguousarray(t, np.uint32)
        keep += [r, t]
        ms[i].n_nodes, ms[i].ref_pos, ms[i].n_tets, ms[i].tets = r.size // 3, _fp(r), t.size // 4, _up(t)
    rq = (ServeRequest * max(len(requests), 1))()
    for i, q in enumerate(requests):
        k = np.ascontiguousarray(q.get("knc", []), np.uint32)
        keep.append(k)
        rq[i].arrival_s, rq[i].mesh, rq[i].steps = float(q["arrival"]), int(q["mesh"]), int(q["steps"])
        rq[i].young_modulus, rq[i].poisson_ratio = float(q["E"]), float(q["nu"])
        rq[i].n_knc, rq[i].knc_nids = k.size, (_up(k) if k.size else None)
    cfg = ServeConfig()
    cfg.base = params
    cfg.hbm_budget_bytes = int(hbm_budget_bytes)
    cfg.max_tick_tets = int(max_tick_tets)
    cfg.duration_s = float(duration_s)
    cfg.drain = int(bool(drain))
    cfg.tile_threads = int(tile_threads)
    cfg.cg_chunk = int(cg_chunk)
    cfg.use_graphs = int(bool(use_graphs))
    cfg.target_round_ms = float(target_round_ms)
    fp = None
    if final_positions is not None:
        fp = (_f * len(requests))(*[(_fp(a) if a is not None else None) for a in final_positions])
    total_steps = sum(int(q["steps"]) for q in requests)
    lat = np.zeros(max(total_steps, 1), np.float64)
    res = ServeResult()
    check(lib.cf_serve_trace(ms, len(meshes), rq, len(requests), C.byref(cfg), fp,
                             lat.ctypes.data_as(C.POINTER(C.c_double)), C.byref(res)), "cf_serve_trace")
    out = {k: getattr(res, k) for k, _ in ServeResult._fields_}
    return out, lat[:res.steps].copy()


def rotations2(F, det, rm, threshold=0.1):
    """Triangle rotations from deformation gradients F (n x 4, row-major) and
    det F, as the 2D step computes them (Compute_R_SVD for RM_MSVD,
    Compute_R_PD with the R = Id fallback below the threshold for RM_DAPD).
    Returns (R as n x 4, fallback flags)."""
    F = np.ascontiguousarray(F, np.float32).reshape(-1, 4)
    det = np.ascontiguousarray(det, np.float32).ravel()
    if det.size != F.shape[0]:
        raise ValueError("one det F per F")
    R = np.zeros_like(F)
    fb = np.zeros(det.size, np.uint32)
    check(lib.cf_rotations2(_fp(F), _fp(det), det.size, int(rm), float(threshold), _fp(R), _up(fb)), "cf_rotations2")
    return R, fb


class MultiSolid3D:
    """Universe step of several LeafSolid3D (SS_AggregateDSH_Basic::MultiDSH_Solid3D_Update
    without contacts): shared fixed dt (the smallest), Newton passes (the largest
    nr_max_iter) and time accumulator. The solids stay owned by the caller."""

    def __init__(self, solids, use_graphs=True):
        self.solids = list(solids)
        hs = (C.c_void_p * len(self.solids))(*[sd.h for sd in self.solids])
        self.h = C.c_void_p()
        check(lib.cf_multi_create(hs, len(self.solids), C.byref(self.h)), "cf_multi_create")
        check(lib.cf_multi_set_graphs(self.h, int(bool(use_graphs))), "cf_multi_set_graphs")

    def close(self):
        if getattr(self, "h", None):
            lib.cf_multi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, dt):
        """Returns (fixed steps taken, per-solid stats of the last one)."""
        n = C.c_uint32()
        st = (StepStats * len(self.solids))()
        check(lib.cf_multi_update(self.h, float(dt), C.byref(n), st), "cf_multi_update")
        return n.value, list(st)


class Batch2:
    """Many 2D solids (TriangleElementSet2 + LeafDSH_Solid2D_FEM) stepped in one
    launch; node vectors are 2N floats."""

    def __init__(self):
        self.h = C.c_void_p()
        check(lib.cf_batch2_create(C.byref(self.h)), "cf_batch2_create")
        self.nodes = []
        self.tris = []

    def close(self):
        if getattr(self, "h", None):
            lib.cf_batch2_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_sim(self, ref_pos, tris, params, knc=None):
        r = np.ascontiguousarray(ref_pos, np.float32)
        t = np.ascontiguousarray(tris, np.uint32)
        k = np.ascontiguousarray(knc if knc is not None else [], np.uint32)
        sid = C.c_uint32()
        check(lib.cf_batch2_add_sim(self.h, r.size // 2, _fp(r), t.size // 3, _up(t), C.byref(params), k.size,
                                    _up(k) if k.size else None, C.byref(sid)), "cf_batch2_add_sim")
        self.nodes.append(r.size // 2)
        self.tris.append(t.size // 3)
        return sid.value

    def fixed_step(self, stats=True):
        if not stats:
            check(lib.cf_batch2_fixed_step(self.h, None), "cf_batch2_fixed_step")
            return None
        st = (StepStats * max(len(self.nodes), 1))()
        check(lib.cf_batch2_fixed_step(self.h, st), "cf_batch2_fixed_step")
        return [st[i] for i in range(len(self.nodes))]

    def state(self, sim_id):
        n = self.nodes[sim_id]
        x = np.empty(2 * n, np.float32)
        v = np.empty(2 * n, np.float32)
        check(lib.cf_batch2_get_state(self.h, int(sim_id), _fp(x), _fp(v)), "cf_batch2_get_state")
        return x, v

    def set_state(self, sim_id, x, v):
        x = np.ascontiguousarray(x, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        check(lib.cf_batch2_set_state(self.h, int(sim_id), _fp(x), _fp(v)), "cf_batch2_set_state")

    def node_masses(self, sim_id):
        m = np.empty(self.nodes[sim_id], np.float32)
        check(lib.cf_batch2_node_masses(self.h, int(sim_id), _fp(m)), "cf_batch2_node_masses")
        return m

    def num_fallback(self, sim_id):
        """DAPD R = Id fallbacks of the simulation's last fixed step."""
        n = C.c_uint32()
        check(lib.cf_batch2_fallbacks(self.h, int(sim_id), C.byref(n)), "cf_batch2_fallbacks")
        return n.value

    def noc(self, sim_id):
        """Node of collapse per triangle (3 = none), DAPD simulations."""
        out = np.empty(self.tris[sim_id], np.uint8)
        check(lib.cf_batch2_get_noc(self.h, int(sim_id), out.ctypes.data_as(C.POINTER(C.c_uint8))),
              "cf_batch2_get_noc")
        return out

    def info(self):
        n, w = C.c_uint32(), C.c_uint32()
        t, nn = C.c_uint64(), C.c_uint64()
        check(lib.cf_batch2_info(self.h, C.byref(n), C.byref(t), C.byref(nn), C.byref(w)), "cf_batch2_info")
        return {"sims": n.value, "tris": t.value, "nodes": nn.value, "workgroups": w.value}


class LeafSolid2D:
    """LeafDSH_Solid2D_FEM drop-in: one 2D solid (the batched kernel on one mesh)."""

    def __init__(self, ref_pos, tris, params, knc=None):
        r = np.ascontiguousarray(ref_pos, np.float32)
        t = np.ascontiguousarray(tris, np.uint32)
        k = np.ascontiguousarray(knc if knc is not None else [], np.uint32)
        self.N = r.size // 2
        self.T = t.size // 3
        self.params = params
        self.h = C.c_void_p()
        check(lib.cf_solid2_create(self.N, _fp(r), t.size // 3, _up(t), C.byref(params), k.size,
                                   _up(k) if k.size else None, C.byref(self.h)), "cf_solid2_create")

    def close(self):
        if getattr(self, "h", None):
            lib.cf_solid2_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fixed_step(self):
        st = StepStats()
        check(lib.cf_solid2_fixed_step(self.h, C.byref(st)), "cf_solid2_fixed_step")
        return st

    def step(self, dt):
        st = StepStats()
        check(lib.cf_solid2_step(self.h, float(dt), C.byref(st)), "cf_solid2_step")
        return st

    def state(self):
        x = np.empty(2 * self.N, np.float32)
        v = np.empty(2 * self.N, np.float32)
        check(lib.cf_solid2_get_state(self.h, _fp(x), _fp(v)), "cf_solid2_get_state")
        return x, v

    def set_state(self, x, v):
        x = np.ascontiguousarray(x, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        check(lib.cf_solid2_set_state(self.h, _fp(x), _fp(v)), "cf_solid2_set_state")

    def num_fallback(self):
        """DAPD R = Id fallbacks of the last fixed step."""
        n = C.c_uint32()
        check(lib.cf_solid2_fallbacks(self.h, C.byref(n)), "cf_solid2_fallbacks")
        return n.value

    def noc(self):
        """Node of collapse per triangle (3 = none), DAPD only."""
        out = np.empty(self.T, np.uint8)
        check(lib.cf_solid2_get_noc(self.h, out.ctypes.data_as(C.POINTER(C.c_uint8))), "cf_solid2_get_noc")
        return out
