"""ctypes declarations of include/corofem.h."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcorofem.so")

CF_OK = 0
RM_ID, RM_QR, RM_MSVD, RM_DAPD = 0, 1, 2, 3
MM_C_WRP, MM_C_LCM, MM_C_LCMH = 1, 2, 3
DM_TRUNCATED, DM_EXACT = 0, 1
IT_QIE_V, IT_FIE_DX = 1, 3
BATCH_STEP_HOST, BATCH_STEP_DEVICE = 0, 1
CG_CONVERGED_REL, CG_CONVERGED_ABS, CG_MAX_ITER, CG_BREAKDOWN, CG_NONFINITE = 1, 2, 3, 4, 5
CF_CG_STATIC_MAX_N = 1 << 22  # unknowns of cf_cg_solve_static's static scratch (include/corofem.h)
ERR_INVALID_ARG, ERR_HIP, ERR_OUT_OF_MEMORY, ERR_NAN, ERR_UNSUPPORTED, ERR_NO_DEVICE, ERR_CAPACITY = (
    -1, -2, -3, -5, -6, -7, -8)

_f = C.POINTER(C.c_float)
_u = C.POINTER(C.c_uint32)
_vp = C.c_void_p


class FemParams(C.Structure):
    _fields_ = [
        ("young_modulus", C.c_float), ("poisson_ratio", C.c_float), ("damping_ratio", C.c_float),
        ("rayleigh_freq1", C.c_float), ("rayleigh_freq2", C.c_float),
        ("degenerate_threshold_det_f", C.c_float), ("mm", C.c_int32), ("rm", C.c_int32), ("dm", C.c_int32),
        ("ecie_k_factor", C.c_float), ("dapd_l_factor", C.c_float), ("dapd_nl_factor", C.c_float),
        ("dapd_nl_exponent", C.c_float),
    ]


class SolidParams(C.Structure):
    _fields_ = [
        ("fem", FemParams), ("fixed_dt", C.c_float), ("time_scale", C.c_float), ("density", C.c_float),
        ("gravity", C.c_float * 3), ("air_drag_coeff", C.c_float), ("cg_rel_epsilon", C.c_float),
        ("cg_abs_epsilon", C.c_float), ("cg_max_iter", C.c_uint32), ("warm_start", C.c_int32),
        ("compute_energy", C.c_int32), ("thickness", C.c_float),
        ("integrator", C.c_int32), ("nr_max_iter", C.c_uint32), ("nr_abs_epsilon", C.c_float),
        ("reference_order", C.c_int32),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("num_fixed_steps", C.c_uint32), ("cg_iters", C.c_int32), ("cg_rel_residual", C.c_float),
        ("cg_abs_residual", C.c_float), ("cg_reason", C.c_int32), ("num_degenerate", C.c_uint32),
        ("sum_degenerate_det_f", C.c_float), ("elastic_energy", C.c_float), ("min_det_f", C.c_float),
        ("step_ms", C.c_double),
        ("nr_iters", C.c_int32), ("nr_rel_residual", C.c_float), ("nr_abs_residual", C.c_float),
    ]


class CgResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("rel_residual", C.c_float), ("abs_residual", C.c_float),
                ("reason", C.c_int32)]


class BatchInfo(C.Structure):
    _fields_ = [
        ("num_meshes", C.c_uint32), ("num_sims", C.c_uint32), ("num_active", C.c_uint32),
        ("bytes_in_use", C.c_uint64), ("bytes_reserved", C.c_uint64), ("budget_bytes", C.c_uint64),
        ("active_tets", C.c_uint64), ("active_nodes", C.c_uint64),
        ("tile_threads", C.c_uint32), ("tile_elems", C.c_uint32), ("cg_chunk", C.c_uint32),
        ("use_graphs", C.c_uint32), ("step_mode", C.c_uint32), ("next_chunks", C.c_uint32),
        ("last_continuations", C.c_uint32), ("chunk_graphs", C.c_uint32),
    ]


class ServeMesh(C.Structure):
    _fields_ = [("n_nodes", C.c_uint32), ("ref_pos", _f), ("n_tets", C.c_uint32), ("tets", _u)]


class ServeRequest(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("mesh", C.c_uint32), ("steps", C.c_uint32),
                ("young_modulus", C.c_float), ("poisson_ratio", C.c_float), ("n_knc", C.c_uint32),
                ("knc_nids", _u)]


class ServeConfig(C.Structure):
    _fields_ = [("base", SolidParams), ("hbm_budget_bytes", C.c_uint64), ("max_tick_tets", C.c_uint64),
                ("duration_s", C.c_double), ("drain", C.c_int32), ("tile_threads", C.c_uint32),
                ("cg_chunk", C.c_uint32), ("use_graphs", C.c_int32), ("target_round_ms", C.c_double)]


class ServeResult(C.Structure):
    _fields_ = [
        ("requests", C.c_uint32), ("admitted", C.c_uint32), ("completed", C.c_uint32), ("max_active", C.c_uint32),
        ("ticks", C.c_uint32), ("mesh_uploads", C.c_uint32), ("steps", C.c_uint64), ("tet_steps", C.c_uint64),
        ("peak_bytes", C.c_uint64), ("wall_s", C.c_double), ("steps_per_s", C.c_double),
        ("tick_mean_ms", C.c_double), ("step_p50_ms", C.c_double), ("step_p99_ms", C.c_double),
        ("step_mean_ms", C.c_double), ("step_max_ms", C.c_double), ("first_p50_ms", C.c_double),
        ("first_p99_ms", C.c_double), ("final_tick_tets", C.c_uint64),
    ]


AV_FN = C.CFUNCTYPE(None, _vp, _f, _f)
CONS_FN = C.CFUNCTYPE(None, _vp, _f)

# name -> (restype, argtypes); every function declared in include/corofem.h
SIGNATURES = {
    "cf_default_fem_params": (None, [C.POINTER(FemParams)]),
    "cf_default_solid_params": (None, [C.POINTER(SolidParams)]),
    "cf_last_error": (C.c_char_p, []),
    "cf_version": (C.c_int, []),
    "cf_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "cf_set_device": (C.c_int, [C.c_int]),
    "cf_device_alloc": (C.c_int, [C.c_size_t, C.POINTER(_vp)]),
    "cf_device_free": (C.c_int, [_vp]),
    "cf_memcpy_htod": (C.c_int, [_vp, _vp, C.c_size_t]),
    "cf_memcpy_dtoh": (C.c_int, [_vp, _vp, C.c_size_t]),
    "cf_memcpy_dtod": (C.c_int, [_vp, _vp, C.c_size_t]),
    "cf_memset_device": (C.c_int, [_vp, C.c_int, C.c_size_t]),
    "cf_device_synchronize": (C.c_int, []),
    "cf_mesh_box_counts": (C.c_int, [C.c_uint32] * 3 + [_u, _u]),
    "cf_mesh_box": (C.c_int, [_f, C.c_uint32, C.c_uint32, C.c_uint32, _f, _u]),
    "cf_tetshape_bake": (C.c_int, [C.c_uint32, _f, C.c_uint32, _u, C.POINTER(_vp)]),
    "cf_tetshape_load_tetgen": (C.c_int, [C.c_char_p, C.c_float, C.POINTER(_vp)]),
    "cf_tetshape_info": (C.c_int, [_vp, _u, _u, _u, _u]),
    "cf_tetshape_get": (C.c_int, [_vp, _f, _u, _u, _u, _u, _u]),
    "cf_tetshape_destroy": (C.c_int, [_vp]),
    "cf_tetshape_load": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "cf_tetshape_save": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "cf_mesh_box2_counts": (C.c_int, [C.c_uint32] * 2 + [_u, _u]),
    "cf_mesh_box2": (C.c_int, [_f, C.c_uint32, C.c_uint32, _f, _u]),
    "cf_tetset_create": (C.c_int, [C.c_uint32, _f, C.c_uint32, _u, C.POINTER(FemParams), C.POINTER(_vp)]),
    "cf_tetset_begin_edition": (C.c_int, [C.c_uint32, _f, C.c_uint32, C.POINTER(FemParams), C.POINTER(_vp)]),
    "cf_tetset_add_element": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    "cf_tetset_end_edition": (C.c_int, [_vp]),
    "cf_tetset_set_params": (C.c_int, [_vp, C.POINTER(FemParams)]),
    "cf_tetset_destroy": (C.c_int, [_vp]),
    "cf_tetset_info": (C.c_int, [_vp, _u, _u, _f]),
    "cf_tetset_get_elements": (C.c_int, [_vp, _f, _f]),
    "cf_tet_compute_k": (C.c_int, [_f, C.c_float, C.c_float, _f, _f]),
    "cf_cache_create": (C.c_int, [_vp, _f, C.POINTER(_vp)]),
    "cf_cache_destroy": (C.c_int, [_vp]),
    "cf_begin_evaluation": (C.c_int, [_vp, _vp, _vp, _vp, C.c_float]),
    "cf_f": (C.c_int, [_vp, _vp]),
    "cf_f_e": (C.c_int, [_vp, _vp]),
    "cf_df_x": (C.c_int, [_vp, _vp, _vp]),
    "cf_f_d": (C.c_int, [_vp, _vp]),
    "cf_rayleigh_coeffs": (None, [C.c_float, C.c_float, C.c_float, _f]),
    "cf_V": (C.c_int, [_vp, C.POINTER(C.c_double)]),
    "cf_end_evaluation": (C.c_int, [_vp]),
    "cf_cache_stats": (C.c_int, [_vp, _u, _f, _f]),
    "cf_cache_get_rotations": (C.c_int, [_vp, _f, _f]),
    "cf_cache_get_doc": (C.c_int, [_vp, C.POINTER(C.c_uint8)]),
    "cf_cache_set_doc": (C.c_int, [_vp, C.POINTER(C.c_uint8)]),
    "cf_cache_num_fallback": (C.c_int, [_vp, _u]),
    "cf_cache_synchronize": (C.c_int, [_vp]),
    "cf_cg_create": (C.c_int, [C.c_uint32, C.POINTER(_vp)]),
    "cf_cg_destroy": (C.c_int, [_vp]),
    "cf_cg_stream": (_vp, [_vp]),
    "cf_cg_solve": (C.c_int, [_vp, AV_FN, CONS_FN, _vp, _vp, _vp, C.c_float, C.c_float, C.c_uint32,
                              C.POINTER(CgResult)]),
    "cf_cg_solve_static": (C.c_int, [_vp, C.c_uint32, AV_FN, CONS_FN, _vp, _vp, _vp, C.c_float, C.c_float,
                                     C.c_uint32, C.POINTER(CgResult)]),
    "cf_solid_create": (C.c_int, [C.c_uint32, _f, C.c_uint32, _u, C.POINTER(SolidParams), C.POINTER(_vp)]),
    "cf_solid_destroy": (C.c_int, [_vp]),
    "cf_solid_set_knc": (C.c_int, [_vp, C.c_uint32, _u, _f]),
    "cf_solid_step": (C.c_int, [_vp, C.c_float, C.POINTER(StepStats)]),
    "cf_solid_fixed_step": (C.c_int, [_vp, C.c_float, C.POINTER(StepStats)]),
    "cf_solid_get_state": (C.c_int, [_vp, _f, _f]),
    "cf_solid_set_state": (C.c_int, [_vp, _f, _f]),
    "cf_solid_device_state": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp)]),
    "cf_solid_info": (C.c_int, [_vp, _u, _u, _f, _f]),
    "cf_solid_total_time": (C.c_int, [_vp, _f]),
    "cf_solid_set_graphs": (C.c_int, [_vp, C.c_int]),
    "cf_solid_assemble_rhs": (C.c_int, [_vp, C.c_float, _f, _f]),
    "cf_solid_apply_A": (C.c_int, [_vp, C.c_float, C.c_float, _f, _f]),
    "cf_solid_time_df_x": (C.c_int, [_vp, C.c_uint32, C.POINTER(C.c_double)]),
    "cf_solid_time_kx": (C.c_int, [_vp, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "cf_multi_create": (C.c_int, [C.POINTER(_vp), C.c_uint32, C.POINTER(_vp)]),
    "cf_multi_destroy": (C.c_int, [_vp]),
    "cf_multi_update": (C.c_int, [_vp, C.c_float, _u, C.POINTER(StepStats)]),
    "cf_multi_set_graphs": (C.c_int, [_vp, C.c_int]),
    "cf_set_tile_elems": (C.c_int, [C.c_uint32]),
    "cf_get_tile_elems": (C.c_int, [_u]),
    "cf_set_tile_threads": (C.c_int, [C.c_uint32]),
    "cf_set_cg_workgroups": (C.c_int, [C.c_uint32]),
    "cf_get_cg_workgroups": (C.c_int, [_u]),
    "cf_set_tile_persistent": (C.c_int, [C.c_int]),
    "cf_set_tile_alternate": (C.c_int, [C.c_int]),
    "cf_set_fused_front": (C.c_int, [C.c_int]),
    "cf_batch_create": (C.c_int, [C.c_uint64, C.POINTER(_vp)]),
    "cf_batch_destroy": (C.c_int, [_vp]),
    "cf_batch_set_tiles": (C.c_int, [_vp, C.c_uint32, C.c_uint32]),
    "cf_batch_set_schedule": (C.c_int, [_vp, C.c_uint32, C.c_int]),
    "cf_batch_set_step_mode": (C.c_int, [_vp, C.c_int]),
    "cf_batch_add_mesh": (C.c_int, [_vp, C.c_uint32, _f, C.c_uint32, _u, _u]),
    "cf_batch_add_meshes": (C.c_int, [_vp, C.c_uint32, _u, C.POINTER(_f), _u, C.POINTER(_u), C.c_uint32, _u]),
    "cf_batch_remove_mesh": (C.c_int, [_vp, C.c_uint32]),
    "cf_batch_sim_bytes": (C.c_int, [_vp, C.c_uint32, C.POINTER(C.c_uint64)]),
    "cf_batch_add_sim": (C.c_int, [_vp, C.c_uint32, C.POINTER(SolidParams), C.c_uint32, _u, _f, _f, _f, _u]),
    "cf_batch_remove_sim": (C.c_int, [_vp, C.c_uint32]),
    "cf_batch_set_active": (C.c_int, [_vp, C.c_uint32, C.c_int]),
    "cf_batch_get_state": (C.c_int, [_vp, C.c_uint32, _f, _f]),
    "cf_batch_set_state": (C.c_int, [_vp, C.c_uint32, _f, _f]),
    "cf_batch_fixed_step": (C.c_int, [_vp, C.POINTER(StepStats)]),
    "cf_batch_active_ids": (C.c_int, [_vp, _u]),
    "cf_batch_request_steps": (C.c_int, [_vp, C.c_uint32, C.c_uint32]),
    "cf_batch_round": (C.c_int, [_vp, C.c_uint32, _u, _u]),
    "cf_batch_info_get": (C.c_int, [_vp, C.POINTER(BatchInfo)]),
    "cf_batch_time_kx": (C.c_int, [_vp, C.c_uint32, C.POINTER(C.c_double)]),
    "cf_batch_synchronize": (C.c_int, [_vp]),
    "cf_rotations2": (C.c_int, [_f, _f, C.c_uint32, C.c_int32, C.c_float, _f, _u]),
    "cf_batch2_create": (C.c_int, [C.POINTER(_vp)]),
    "cf_batch2_destroy": (C.c_int, [_vp]),
    "cf_batch2_add_sim": (C.c_int, [_vp, C.c_uint32, _f, C.c_uint32, _u, C.POINTER(SolidParams), C.c_uint32, _u, _u]),
    "cf_batch2_fixed_step": (C.c_int, [_vp, C.POINTER(StepStats)]),
    "cf_batch2_get_state": (C.c_int, [_vp, C.c_uint32, _f, _f]),
    "cf_batch2_set_state": (C.c_int, [_vp, C.c_uint32, _f, _f]),
    "cf_batch2_node_masses": (C.c_int, [_vp, C.c_uint32, _f]),
    "cf_batch2_fallbacks": (C.c_int, [_vp, C.c_uint32, _u]),
    "cf_solid2_fallbacks": (C.c_int, [_vp, _u]),
    "cf_batch2_get_noc": (C.c_int, [_vp, C.c_uint32, C.POINTER(C.c_uint8)]),
    "cf_solid2_get_noc": (C.c_int, [_vp, C.POINTER(C.c_uint8)]),
    "cf_batch2_info": (C.c_int, [_vp, _u, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), _u]),
    "cf_solid2_create": (C.c_int, [C.c_uint32, _f, C.c_uint32, _u, C.POINTER(SolidParams), C.c_uint32, _u,
                                   C.POINTER(_vp)]),
    "cf_solid2_destroy": (C.c_int, [_vp]),
    "cf_solid2_fixed_step": (C.c_int, [_vp, C.POINTER(StepStats)]),
    "cf_solid2_step": (C.c_int, [_vp, C.c_float, C.POINTER(StepStats)]),
    "cf_solid2_get_state": (C.c_int, [_vp, _f, _f]),
    "cf_solid2_set_state": (C.c_int, [_vp, _f, _f]),
    "cf_serve_trace": (C.c_int, [C.POINTER(ServeMesh), C.c_uint32, C.POINTER(ServeRequest), C.c_uint32,
                                 C.POINTER(ServeConfig), C.POINTER(_f), C.POINTER(C.c_double),
                                 C.POINTER(ServeResult)]),
}


class CorofemError(RuntimeError):
    """A non-zero status from the C-ABI; .status holds the code (ERR_*)."""

    def __init__(self, msg, status=0):
        super().__init__(msg)
        self.status = status


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libcorofem.so not found at {LIB_PATH}: build it with `python risdsdgcfem_thesis_amd/build.py` "
            "(or __graft_entry__.build()); there is no fallback path")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


lib = _load()


def check(rc, what=""):
    if rc != CF_OK:
        msg = lib.cf_last_error().decode(errors="replace")
        raise CorofemError(f"{what} failed with status {rc}: {msg}", rc)
    return rc
