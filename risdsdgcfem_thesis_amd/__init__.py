"""corofem for MI355X: the reference's corotational tet FEM implicit step
(Saphyre2 LeafDSH_Solid3D_FEM + TetrahedronElementSet3 + ILSS_CG) as hand-written
HIP kernels for gfx950 behind a C-ABI (include/corofem.h).

This package is the Python mirror of that C-ABI (ctypes). It holds no compute
of its own: every operation runs in libcorofem.so, and importing the package
fails loudly when the library has not been built.
"""
from ._capi import (  # noqa: F401
    lib, LIB_PATH, CorofemError, check, FemParams, SolidParams, StepStats, CgResult, BatchInfo,
    RM_ID, RM_QR, RM_MSVD, RM_DAPD, MM_C_WRP, MM_C_LCM, MM_C_LCMH, DM_TRUNCATED, DM_EXACT, IT_QIE_V, IT_FIE_DX,
    BATCH_STEP_HOST, BATCH_STEP_DEVICE,
    CG_CONVERGED_REL, CG_CONVERGED_ABS, CG_MAX_ITER, CG_BREAKDOWN, CG_NONFINITE,
    ERR_INVALID_ARG, ERR_HIP, ERR_OUT_OF_MEMORY, ERR_NAN, ERR_UNSUPPORTED, ERR_NO_DEVICE, ERR_CAPACITY,
)
from .fem import (  # noqa: F401
    DeviceArray, device_count, mesh_box, mesh_box2, mesh_box_baked, TetShape, TetrahedronElementSet3, ILSS_CG, LeafSolid3D, MultiSolid3D, Batch,
    Batch2, LeafSolid2D,
    default_fem_params, default_solid_params, rayleigh_coefficients, serve, cg_solve_static, rotations2, set_tile_elems, set_tile_threads, set_tile_persistent, set_tile_alternate, set_fused_front, get_tile_elems, set_cg_workgroups,
    get_cg_workgroups, synchronize,
)
