/*
 * corofem -- MI355X (gfx950) corotational tetrahedral FEM implicit step.
 *
 * C-ABI drop-in for the reference's hot path (Saphyre2):
 *   element set + cache   <-> ms::fem::EditableTetrahedronElementSet3 / IForce3
 *                             (ms/fem/TetrahedronElementSet3.h:29-57,70-261)
 *   linear solver         <-> ns::IIterativeLinearSystemSolver / ILSS_CG
 *                             (ns/IIterativeLinearSystemSolver.h:22-48, ns/ILSS_CG.h:19-151)
 *   leaf step             <-> ds::LeafDSH_Solid3D_FEM::Create/Step/FixedStep/QueryStats
 *                             (ds/dsh/LeafDSH_Solid3D_FEM.h:68-81, .cpp:125-256,898-990)
 *   2D twin               <-> ms::fem::TriangleElementSet2 + ds::LeafDSH_Solid2D_FEM
 *
 * Conventions
 *   - Every entry point returns an int status (CF_OK = 0, negative on error);
 *     no C++ exception crosses the ABI. cf_last_error() describes the last error.
 *   - Node vectors are flat FP32 arrays of 3*N values, node i at [3i, 3i+2]
 *     (the reference's Vec3 arrays reinterpreted as Real*, as ILSS_CG sees them).
 *   - Indices are uint32 (the reference's uint16 limit is not kept).
 *   - Pointers named d_* are device pointers (from cf_device_alloc or any HIP
 *     allocation on the current device); pointers named h_* are host pointers.
 *   - Operator outputs ACCUMULATE (+=) like IForce3 (TetrahedronElementSet3.h:20-21).
 *   - Each handle owns one HIP stream; calls on one handle are not thread-safe,
 *     different handles are independent.
 */
#ifndef COROFEM_H
#define COROFEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define CF_OK 0
#define CF_ERR_INVALID_ARG (-1)
#define CF_ERR_HIP (-2)
#define CF_ERR_OUT_OF_MEMORY (-3)
#define CF_ERR_NAN (-5)            /* a CG or Newton reduction was NaN or Inf (see CF_CG_NONFINITE) */
#define CF_ERR_UNSUPPORTED (-6)
#define CF_ERR_NO_DEVICE (-7)
#define CF_ERR_CAPACITY (-8)

/* CG termination reasons (cf_cg_result.reason) */
#define CF_CG_CONVERGED_REL 1   /* delta <= rel^2 * delta_0          (ILSS_CG.h:86-88) */
#define CF_CG_CONVERGED_ABS 2   /* delta <= abs^2                    (ILSS_CG.h:86)    */
#define CF_CG_MAX_ITER 3        /* k == max_iterations               (ILSS_CG.h:88)    */
#define CF_CG_BREAKDOWN 4       /* |d^T q| < abs^2                   (ILSS_CG.h:97)    */
/* r^T r or d^T q is NaN or Inf: the state, the right-hand side or the operator
   holds non-finite values (the reference asserts they do not,
   TetrahedronElementSet3.cpp:253-255). The step still runs to its end; the
   entry point returns CF_ERR_NAN. */
#define CF_CG_NONFINITE 5

/* material methods (ms/fem/Params.h:50-62): C_WRP (the stored-Ke corotational
   method, default), C_LCM and C_LCMH are implemented; CF_MM_L and the
   hyperelastic methods are not (CF_ERR_UNSUPPORTED) */
#define CF_MM_L 0
#define CF_MM_C_WRP 1
/* LCM and the LCM/WRP hybrid (ms/fem/Params.h:54-55): Piola forces
   C_LCM_f_e / C_LCM_df_x / C_LCMH_df_x / C_LCM_V
   (ms/fem/TetrahedronElementSet3.cpp:1155-1178,1489-1508,1838-1929), from the
   F and R of the last BeginEvaluation; single element sets and leaves only
   (batches take C_WRP). */
#define CF_MM_C_LCM 2
#define CF_MM_C_LCMH 3
/* rotation methods (ms/fem/Params.h:63-67) */
#define CF_RM_ID 0
#define CF_RM_QR 1
#define CF_RM_MSVD 2
#define CF_RM_DAPD 3
/* differential methods (ms/fem/Params.h:68-72): Truncated (default) and Exact
   are implemented; Exact_Inv and Numerical are not (CF_ERR_UNSUPPORTED) */
#define CF_DM_TRUNCATED 0
/* Exact differentials (ms/fem/Params.h:69): K*x adds the McAdams rotation
   differential dR (ms/fem/Rotation.cpp:622-644; C_df_x dR terms
   TetrahedronElementSet3.cpp:1804-1834, LCM :1862-1864) for MSVD, and for
   DAPD above the degeneration threshold; single element sets and leaves. */
#define CF_DM_EXACT 1

/* ---------------------------------------------------------------- params */
/* Mirrors ms::fem::Params (ms/fem/Params.h:10-99). */
typedef struct cf_fem_params {
    float young_modulus;              /* default 1000     */
    float poisson_ratio;              /* default 0.25     */
    float damping_ratio;              /* default 1.0      */
    float rayleigh_freq1;             /* default 0.1      */
    float rayleigh_freq2;             /* default 10000    */
    float degenerate_threshold_det_f; /* default 0.1      */
    int32_t mm;                       /* default CF_MM_C_WRP */
    int32_t rm;                       /* default CF_RM_MSVD  */
    int32_t dm;                       /* default CF_DM_TRUNCATED */
    float ecie_k_factor;              /* LCMH blend radius k (m_ECIE_k_factor); default 1 */
    float dapd_l_factor;              /* DAPD degenerate projection: m_DAPD_L_Factor, default 1 */
    float dapd_nl_factor;             /* m_DAPD_NL_Factor, default 0 */
    float dapd_nl_exponent;           /* m_DAPD_NL_Exponent, default 2 (ms/fem/Params.h:29,88) */
} cf_fem_params;

/* Mirrors the leaf's integrator/solver knobs (ds/dsh/LeafDSH_Solid3D_FEM.h:93-274). */
typedef struct cf_solid_params {
    cf_fem_params fem;
    float fixed_dt;            /* default 0.001 (leaf default); configs use 1/60 */
    float time_scale;          /* default 1 */
    float density;             /* default 1000 */
    float gravity[3];          /* default (0,0,0) */
    float air_drag_coeff;      /* default 0 */
    float cg_rel_epsilon;      /* default 1e-3 */
    float cg_abs_epsilon;      /* default 1e-6 */
    uint32_t cg_max_iter;      /* default 20   */
    int32_t warm_start;        /* default 1    */
    int32_t compute_energy;    /* 1: elastic energy each step like FixedStep (:977); default 1 */
    float thickness;           /* 2D only (LeafDSH_Solid2D_FEM.h:217); default 1 */
    int32_t integrator;        /* CF_IT_QIE_V (default) or CF_IT_FIE_DX (.h:109, dispatch .cpp:944-963) */
    uint32_t nr_max_iter;      /* FIE-dx Newton iterations (m_SolverNR_MaxIter, .h:233); default 3 */
    float nr_abs_epsilon;      /* FIE-dx stop: |b| < this (m_SolverNR_AbsEpsilon, .h:236-238); default 1e-6 */
    /* 1: 3D leaves compute each fixed step in the reference's FP32 order, bit for
       bit: stored 16-block Ke per element (TetrahedronElement3::Compute_K), the
       element-ordered serial scatter of C_f_e / C_df_x
       (TetrahedronElementSet3.cpp:1053-1096,1761-1836), ns::real_array's SSE dot
       order with FP32 CG scalars (ILSS_CG.h:38-151) and the QIE-v vector terms
       (LeafDSH_Solid3D_FEM.cpp:1205-1328). QIE-v, C_WRP, DM Truncated only
       (CF_ERR_UNSUPPORTED otherwise); rotations by DAPD are bit-exact, by MSVD
       within the SVD's ulps. For parity checks: 576 B of Ke per element and one
       host round trip per CG dot product. Default 0 (the fast path: FP64
       reductions, Piola-form element forces). */
    int32_t reference_order;
} cf_solid_params;

/* Integrators (LeafDSH_Solid3D_FEM.h:109). CF_IT_FIE_DX is
   Step_FullyImplicitEuler_dx (.cpp:1357-1552): Newton iterations in position
   increments with A = (1/dt^2 + c0/dt) M + (1 + c1/dt) K, a BeginEvaluation per
   iteration and KNC rows in increment form; 3D leaves and the multi-solid
   step only (batches and 2D solids take CF_IT_QIE_V). */
#define CF_IT_QIE_V 1
#define CF_IT_FIE_DX 3

/* Per-step statistics (LeafDSH_Solid3D_FEM::QueryStats fields, .cpp:698-770). */
typedef struct cf_step_stats {
    uint32_t num_fixed_steps;     /* fixed steps taken by this Step() call */
    int32_t cg_iters;             /* last fixed step */
    float cg_rel_residual;
    float cg_abs_residual;
    int32_t cg_reason;
    uint32_t num_degenerate;      /* elements with det F <= 0 */
    float sum_degenerate_det_f;
    float elastic_energy;
    float min_det_f;
    double step_ms;               /* device time of the last fixed step */
    int32_t nr_iters;             /* FIE-dx: Newton iterations of the last fixed step */
    float nr_rel_residual;        /* FIE-dx: last |b| or |y| over the first |b| (m_SolverNR_RelPrec) */
    float nr_abs_residual;        /* FIE-dx: last |b| or |y| (m_SolverNR_AbsPrec) */
} cf_step_stats;

typedef struct cf_cg_result {
    int32_t iterations;
    float rel_residual;
    float abs_residual;
    int32_t reason;
} cf_cg_result;

void cf_default_fem_params(cf_fem_params* p);
void cf_default_solid_params(cf_solid_params* p);
/* Rayleigh coefficients (c0, c1) from damping ratio zeta and frequencies f1, f2
   (LeafDSH_Solid3D_FEM::RebuildRayleighCoeffsFromFreqs, LeafDSH_Solid3D_FEM.h:347-354). */
void cf_rayleigh_coeffs(float f1, float f2, float zeta, float c[2]);

/* ---------------------------------------------------------------- runtime */
const char* cf_last_error(void);
int cf_version(void);
int cf_device_count(int* n);
int cf_set_device(int device);
int cf_device_alloc(size_t bytes, void** d_ptr);
int cf_device_free(void* d_ptr);
int cf_memcpy_htod(void* d_dst, const void* h_src, size_t bytes);
int cf_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes);
int cf_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes);
int cf_memset_device(void* d_ptr, int value, size_t bytes);
int cf_device_synchronize(void);

/* ---------------------------------------------------------------- mesh input
 * Make_TetSolidShape3_Box (Geo/shape/TetSolidShape3.cpp:921-999): nx,ny,nz are
 * vertex counts per axis (cubes + 1); 5 tets per cube with mirrored parity.
 * h_pos receives 3*nx*ny*nz floats, h_tets 4*5*(nx-1)(ny-1)(nz-1) indices. */
int cf_mesh_box_counts(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t* n_nodes, uint32_t* n_tets);
int cf_mesh_box(const float half_sizes[3], uint32_t nx, uint32_t ny, uint32_t nz,
                float* h_pos, uint32_t* h_tets);
/* Make_MeshSolidShape2_Box (Geo/shape/MeshSolidShape2.cpp:1651-1675). */
/* ---------------------------------------------------------------- tet shapes
 * EditableTetSolidShape3's edit-then-bake pipeline (Geo/shape/TetSolidShape3.cpp):
 * EndEdition (:262-268, :558-566) drops tetrahedra with det Ds <= 0
 * (AddTetrahedron :278-307), vertices no tetrahedron uses (FixDegenerateVertices
 * :809-845), builds face neighbours and oriented boundary faces (RebuildTopology
 * :619-749) and sorts the tetrahedra stably by layer, the least number of
 * tetrahedron hops from the boundary (RebuildLayers :752-807). The baked order is
 * the element order of the reference's element sets. Make_TetSolidShape3_Box
 * (:921-999) is cf_mesh_box followed by cf_tetshape_bake. Host memory; indices
 * uint32 (the reference stores uint16). */
typedef struct cf_tetshape cf_tetshape;
int cf_tetshape_bake(uint32_t n_vertices, const float* h_pos, uint32_t n_tets, const uint32_t* h_tets,
                     cf_tetshape** out);
/* TetGen <base>.node / <base>.ele as examples/tetgen2geo/main.cpp:80-180 reads them,
   then baked; natural_scale > 0: its -nsac option (centre, largest extent = natural_scale). */
int cf_tetshape_load_tetgen(const char* base_name, float natural_scale, cf_tetshape** out);
int cf_tetshape_info(const cf_tetshape* shape, uint32_t* n_vertices, uint32_t* n_tets,
                     uint32_t* n_boundary_faces, uint32_t* n_layers);
/* Any output may be NULL: positions 3V, tetrahedra 4T, face neighbours 4T
   (0xFFFFFFFF = boundary; entry n is the tet across the face opposite vertex n),
   boundary faces 3BF (outward), layers 2L (first tet, count), and per baked tet
   its index in the input (before dropping and sorting). */
int cf_tetshape_get(const cf_tetshape* shape, float* h_pos, uint32_t* h_tets, uint32_t* h_ntid, uint32_t* h_bf,
                    uint32_t* h_layers, uint32_t* h_src);
int cf_tetshape_destroy(cf_tetshape* shape);
/* Shape library files (ShapeLibrary::Load/Save, Geo/shape/ShapeLibrary.cpp:332-359,446-468;
   the ItemStream container of util/ItemStream.cpp and util/ItemStreamSerialization.cpp).
   load: the first TetSolidShape3 (version 4) item of a binary or text file (told
   apart by the magic number); its arrays are already baked (h_src = identity).
   save: a one-shape library, binary (binary != 0) or text; 16-bit indices, so
   at most 65,535 vertices and tetrahedra (CF_ERR_UNSUPPORTED beyond). */
int cf_tetshape_load(const char* path, cf_tetshape** out);
int cf_tetshape_save(const cf_tetshape* shape, const char* path, int binary);

int cf_mesh_box2_counts(uint32_t nx, uint32_t ny, uint32_t* n_nodes, uint32_t* n_tris);
int cf_mesh_box2(const float half_sizes[2], uint32_t nx, uint32_t ny, float* h_pos, uint32_t* h_tris);

/* ---------------------------------------------------------------- element set
 * EditableTetrahedronElementSet3::BeginEdition/AddElement/EndEdition
 * (ms/fem/TetrahedronElementSet3.cpp:2084-2166). cf_tetset_create is the bulk
 * form of BeginEdition + AddElement x T + EndEdition. */
typedef struct cf_tetset cf_tetset;
int cf_tetset_create(uint32_t n_nodes, const float* h_ref_pos, uint32_t n_tets, const uint32_t* h_tets,
                     const cf_fem_params* params, cf_tetset** out);
int cf_tetset_begin_edition(uint32_t n_nodes, const float* h_ref_pos, uint32_t n_tets,
                            const cf_fem_params* params, cf_tetset** out);
int cf_tetset_add_element(cf_tetset* set, uint32_t i, uint32_t j, uint32_t k, uint32_t l);
int cf_tetset_end_edition(cf_tetset* set);
int cf_tetset_set_params(cf_tetset* set, const cf_fem_params* params);      /* SetParams (:2168-2212) */
int cf_tetset_destroy(cf_tetset* set);
int cf_tetset_info(const cf_tetset* set, uint32_t* n_nodes, uint32_t* n_tets, float* total_volume);
/* Per-element precompute in the caller's element order: h_inv_dm (9T, row-major), h_vol (T). */
int cf_tetset_get_elements(const cf_tetset* set, float* h_inv_dm, float* h_vol);
/* TetrahedronElement3::Compute_K (ms/fem/TetrahedronElement3.h:85-173) for one
   element with corners r (12 floats): the 12x12 stiffness matrix, row-major, in
   the reference's FP32 order, and its volume. Host code; needs no device. */
int cf_tet_compute_k(const float h_r[12], float young_modulus, float poisson_ratio, float* h_K144, float* volume);

/* ---------------------------------------------------------------- IForce3 cache
 * CreateCache(inv_mass) / BeginEvaluation / f / df_x / V / EndEvaluation
 * (TetrahedronElementSet3.h:48-55,86-97). x, prev_x, v, f, dx, df: device,
 * 3N floats. BeginEvaluation computes F, det F and the rotation R per element
 * and updates each element's direction of collapse (DoC) from prev_x -> x
 * (UpdateDegeneration, TetrahedronElementSet3.cpp:477-743; prev_x may be NULL:
 * x). With CF_RM_DAPD an element with det F <= threshold takes the projected
 * rotation of Compute_R_DAPD_Degenerate (Rotation.cpp:364-545) when it has a
 * DoC, else R = Id (the reference's invalid-DoC path, :885-890), counted as a
 * fallback. */
typedef struct cf_cache cf_cache;
int cf_cache_create(cf_tetset* set, const float* h_inv_mass /* N, may be NULL */, cf_cache** out);
int cf_cache_destroy(cf_cache* cache);
int cf_begin_evaluation(cf_cache* cache, const float* d_prev_x, const float* d_x, const float* d_v, float t);
/* The cache's per-element DoC bytes (ElementCache3::m_DoC, TetrahedronElement3.h:28-38):
   4*vit0 + vit1, vertex-face when vit0 == vit1, edge-edge (0, vit1) when vit0 < vit1,
   12 = invalid. T bytes, host memory. set rejects other codes. */
int cf_cache_get_doc(cf_cache* cache, uint8_t* h_doc);
int cf_cache_set_doc(cf_cache* cache, const uint8_t* h_doc);
int cf_f(cf_cache* cache, float* d_f);                       /* += f_e (+ f_p, a no-op) */
int cf_f_e(cf_cache* cache, float* d_f);                     /* += elastic force (C_f_e :1053-1096) */
int cf_df_x(cf_cache* cache, const float* d_dx, float* d_df); /* += K(x) dx (C_df_x :1761-1836)    */
/* += C_f_d (:1098-1150): corotational Rayleigh damping of the evaluation's
   velocities v, -R (c0 m R^T v + c1 Ke R^T v) per element, with c from the
   set's damping ratio and frequencies (LeafDSH_Solid3D_FEM.h:347-354) and
   m = 1 / inv_mass (the cache must have been created with inv_mass). The
   reference's f() leaves it out (:294-300); its symplectic Euler adds it. */
int cf_f_d(cf_cache* cache, float* d_f);
int cf_V(cf_cache* cache, double* h_energy);                 /* C_V (:1456-1488)                  */
int cf_end_evaluation(cf_cache* cache);
int cf_cache_stats(const cf_cache* cache, uint32_t* num_degenerate, float* sum_degenerate_det_f,
                   float* min_det_f);
/* Elements of the last BeginEvaluation that DAPD left at R = Id (det F <= threshold
   without a DoC or with a collapsed axis; TetrahedronElementSet3.cpp:885-890). */
int cf_cache_num_fallback(const cf_cache* cache, uint32_t* num_fallback);
/* R per element in the caller's element order (9T row-major) and det F (T). */
int cf_cache_get_rotations(const cf_cache* cache, float* h_R, float* h_det_f);
int cf_cache_synchronize(cf_cache* cache);

/* ---------------------------------------------------------------- linear solver
 * IIterativeLinearSystemSolver::Solve(D_Av, D_ApplyConstraints, x, b, rel, abs, max)
 * (ns/IIterativeLinearSystemSolver.h:22-48) with the ILSS_CG algorithm
 * (ns/ILSS_CG.h:38-151). The callbacks receive DEVICE pointers and must
 * enqueue their work on the stream returned by cf_cg_stream(); x (in/out, warm
 * start) and b are device pointers of length n. Dot products accumulate in
 * FP64. */
typedef void (*cf_av_fn)(void* user, float* d_Av, const float* d_v);
typedef void (*cf_constraints_fn)(void* user, float* d_v);
typedef struct cf_cg cf_cg;
int cf_cg_create(uint32_t n, cf_cg** out);                   /* ILSS_CG::Init (:30-36) */
int cf_cg_destroy(cf_cg* cg);
void* cf_cg_stream(cf_cg* cg);
int cf_cg_solve(cf_cg* cg, cf_av_fn av, cf_constraints_fn apply_constraints, void* user,
                float* d_x, const float* d_b, float rel_epsilon, float abs_epsilon, uint32_t max_iterations,
                cf_cg_result* result);
/* No counterpart in the reference: cf_cg_solve for callers that cannot allocate
   device memory (code under HIP graph capture, tight serving loops, harnesses
   that forbid hipMalloc). It needs no cf_cg handle and allocates and frees
   nothing: r, d, q, the scalar record and the FP64 partials live in a static
   device array of the library, CF_CG_STATIC_BYTES per GPU (room for
   CF_CG_STATIC_MAX_N unknowns), which each GPU holds from the library's first
   use until the process ends. n > CF_CG_STATIC_MAX_N returns CF_ERR_CAPACITY.
   stream: the hipStream_t the callbacks must enqueue on (NULL: the null
   stream). Results are bit-identical to cf_cg_solve's on the same inputs.
   Every call on one GPU shares that GPU's array, which keeps the last call's
   contents: calls that overlap on one GPU race, so serialize them. */
#define CF_CG_STATIC_MAX_N (1u << 22)
#define CF_CG_STATIC_BYTES ((size_t)50495744)
int cf_cg_solve_static(void* stream, uint32_t n, cf_av_fn av, cf_constraints_fn apply_constraints, void* user,
                       float* d_x, const float* d_b, float rel_epsilon, float abs_epsilon, uint32_t max_iterations,
                       cf_cg_result* result);

/* ---------------------------------------------------------------- leaf solid (3D)
 * LeafDSH_Solid3D_FEM: Create (.cpp:125-223), Step (:225-256: fixed-dt
 * accumulator, strict '>'), FixedStep (:898-990) with QIE-v (:1205-1328),
 * uniform node mass (RebuildMass :1591-1611), kinematic node constraints
 * (KNC, :877-896,1045-1053,1268-1295). State stays resident on the device. */
typedef struct cf_solid cf_solid;
int cf_solid_create(uint32_t n_nodes, const float* h_ref_pos, uint32_t n_tets, const uint32_t* h_tets,
                    const cf_solid_params* params, cf_solid** out);
int cf_solid_destroy(cf_solid* s);
/* Position KNC at the given targets (NULL: rest positions), like
   CreateKinematicPointConstraint + eKM_Static. Replaces the current set. */
int cf_solid_set_knc(cf_solid* s, uint32_t n, const uint32_t* h_nids, const float* h_targets);
/* Step(dt) (:225-256): the total time grows by time_scale * dt once per call
   (:229-230); fixed steps of time_scale * fixed_dt run while the accumulator is
   strictly greater than that (:239). stats: the last fixed step, with
   num_fixed_steps = how many ran. */
int cf_solid_step(cf_solid* s, float dt, cf_step_stats* stats);
/* One FixedStep(dt) (:898-990); advances the total time by dt. Both return
   CF_ERR_NAN (after completing the step, stats->cg_reason = CF_CG_NONFINITE)
   when a CG or Newton reduction was NaN or Inf: the state is not finite. */
int cf_solid_fixed_step(cf_solid* s, float dt, cf_step_stats* stats);
/* m_TotalTime: the t handed to BeginEvaluation (:927). */
int cf_solid_total_time(const cf_solid* s, float* t);
/* Execution control with no counterpart in the reference. The leaf's CG runs
   in chunks of eight iterations; with graphs on (the default) each chunk is
   replayed as a HIP graph recorded on first use and kept until the linear
   system's coefficients, the iteration cap, the fixed nodes, the tile
   alternation or the CG grid change; off, the same kernels are launched
   directly (bit-identical; for kernel-trace profiling). */
int cf_solid_set_graphs(cf_solid* s, int use_graphs);
int cf_solid_get_state(const cf_solid* s, float* h_pos, float* h_vel);     /* QueryState (GLeafDSH_Model.h:34-43) */
int cf_solid_set_state(cf_solid* s, const float* h_pos, const float* h_vel);
int cf_solid_device_state(cf_solid* s, float** d_pos, float** d_vel);
int cf_solid_info(const cf_solid* s, uint32_t* n_nodes, uint32_t* n_tets, float* node_mass,
                  float rayleigh[2]);
/* Inspection of the step's stages on the current state (no state change):
   BeginEvaluation + QIE-v right-hand side b and warm start y0 (:1209-1295),
   and q = S(fM M d - fK K d) on the current rotations (LS_A_times_d + LS_ZeroKNC,
   :1045-1079). Host arrays of 3N floats. */
int cf_solid_assemble_rhs(cf_solid* s, float dt, float* h_b, float* h_y0);
int cf_solid_apply_A(cf_solid* s, float factor_m, float factor_k, const float* h_d, float* h_q);
/* Time one K*x (df_x) on the solid's current rotations: mean device ms over reps. */
int cf_solid_time_df_x(cf_solid* s, uint32_t reps, double* ms_per_call);
/* Mean device ms of the two kernels of one K*x, each timed with HIP events on
   the solid's stream over reps back-to-back launches: the LDS-tile element
   kernel and the node gather (matvec form with the mass term and d.q). */
int cf_solid_time_kx(cf_solid* s, uint32_t reps, double* ms_tile, double* ms_gather);

/* ---------------------------------------------------------------- multi-solid
 * SS_AggregateDSH_Basic::MultiDSH_Solid3D_Update (ds/ss/SS_AggregateDSH_Basic.cpp:229-300;
 * the universe step of several Solid3D leaves, __ENABLE_MULTI_SOLID3D), without
 * the contact passes: fixed dt = the smallest of the solids' fixed_dt, Newton
 * passes = the largest nr_max_iter, one time accumulator with strict '>'. Each
 * fixed step runs the leaves' NR_Init / NR_PreStep / NR_MidStep / NR_PostStep /
 * NR_End (ds/dsh/LeafDSH_Solid3D_FEM.cpp:2824-3039) with the frame dt, whatever
 * each leaf's integrator. The solids stay owned by the caller. */
typedef struct cf_multi cf_multi;
int cf_multi_create(cf_solid* const* solids, uint32_t n, cf_multi** out);
int cf_multi_destroy(cf_multi* m);
/* stats: NULL or one per solid (its last fixed step; elastic_energy 0, as the
   NR_End hook does not evaluate V). */
int cf_multi_update(cf_multi* m, float dt, uint32_t* num_fixed_steps, cf_step_stats* stats);
/* Each leaf's CG runs in chunks of eight iterations, every leaf's chunk
   enqueued on its own stream before the host waits on any. With graphs on (the
   default) each chunk is a HIP graph recorded per leaf on first use and kept
   with the leaf until its linear-system coefficients, iteration cap or fixed
   nodes change; off, the same kernels are launched directly (bit-identical). */
int cf_multi_set_graphs(cf_multi* m, int use_graphs);

/* ---------------------------------------------------------------- tuning */
/* Elements per LDS tile of the node-gather K*x (0 = library default 256).
   Takes effect for sets created afterwards. */
int cf_set_tile_elems(uint32_t tile_elems);
/* Workgroup size of the LDS-tile element kernel: 128, 256, 512 or 1024 (0 = default
   256). A tile holds at most this many elements and nodes. */
int cf_set_tile_threads(uint32_t threads);
/* Tile schedule. 0 (default): one tile per workgroup; 1: persistent grid of
   resident workgroups that walk their tiles with the next tile's loads in
   flight; 2: the pipelined kernel of 1 with one tile per workgroup. */
int cf_set_tile_persistent(int persistent);
/* 1: the leaf CG's odd iterations walk each XCD's tiles from the end (one
   tile per workgroup), so a K*x pass starts where the previous one ended, on
   data the memory-side cache may still hold; 0 (default): always forwards
   (the element stream is read non-temporally, which leaves the caches to the
   node data and gains more). Results are the same either way. Takes effect
   for solves enqueued or recorded afterwards. */
int cf_set_tile_alternate(int alternate);
/* 1 (default): a QIE-v leaf step of a C_WRP solid without the exact rotation
   differential opens its evaluation with one tile pass that computes each
   element's F, det F and rotation and, from the rotation in registers, the
   elastic forces f_e(x) and the CG start's K y, instead of BeginEvaluation
   followed by two tile passes; 0: the separate passes. Results are the same
   bit for bit (the sum of the degenerate elements' det F is summed per tile
   instead of per block, which can move its last bit). Takes effect from the
   next step. */
int cf_set_fused_front(int enable);
int cf_get_tile_elems(uint32_t* tile_elems);
/* Workgroups of the two CG iteration kernels of the leaf solvers (alpha with
   x, r; beta and the stop test with d), each summing the iteration's FP64
   partials itself (0 = default 1,024; at most 8,192). Results do not depend
   on it. */
int cf_set_cg_workgroups(uint32_t workgroups);
int cf_get_cg_workgroups(uint32_t* workgroups);

/* ---------------------------------------------------------------- batches
 * Many independent solids stepped together: each batched fixed step runs the
 * step of cf_solid_fixed_step (QIE-v, LeafDSH_Solid3D_FEM.cpp:898-990,
 * 1205-1328) for every active simulation of the batch in the same kernel
 * launches. No reference counterpart: the reference steps one leaf at a time
 * (SS_AggregateDSH_Basic.cpp:229-300 loops over leaves); this is the batching
 * of many leaves of that loop.
 *   - meshes (element-set topology, rest shape, tiles) are added once and
 *     shared by any number of simulations;
 *   - a simulation = its mesh + its own parameters, KNC and device state;
 *   - each simulation's results are bit-identical to cf_solid_fixed_step on
 *     the same mesh, parameters and state, whatever else is in the batch;
 *   - device memory comes from the batch's own arena; simulations are
 *     admitted while their bytes fit the HBM budget (CF_ERR_CAPACITY otherwise).
 * The step is recorded as HIP graphs (prologue, a chunk of CG iterations,
 * epilogue) re-recorded when the set of active simulations changes; the
 * host reads one counter per chunk to stop when every solve has ended. */
typedef struct cf_batch cf_batch;
typedef struct cf_batch_info {
    uint32_t num_meshes, num_sims, num_active;
    uint64_t bytes_in_use;     /* meshes + simulations, against the budget */
    uint64_t bytes_reserved;   /* device memory held by the arena */
    uint64_t budget_bytes;
    uint64_t active_tets, active_nodes;
    uint32_t tile_threads, tile_elems, cg_chunk, use_graphs;
    /* step schedule (CF_BATCH_STEP_*); under the device schedule, the chunks
       the next step is submitted with, the chunks the last step added after
       it ran, and the chunk graphs recorded (one per grid class) */
    uint32_t step_mode, next_chunks, last_continuations, chunk_graphs;
} cf_batch_info;

/* How cf_batch_fixed_step runs its CG chunks.
   HOST: the host reads the iterating count after every chunk (one chunk
   behind) and, once at most half of the simulations still iterate, compacts
   the tables and launches the rest directly.
   DEVICE (default): the whole step is submitted at once; before every chunk
   the device compacts the lists of the simulations still iterating, and the
   chunk's kernels (grids sized from the previous step's list lengths) walk
   them; the host reads nothing until the step has run. */
enum { CF_BATCH_STEP_HOST = 0, CF_BATCH_STEP_DEVICE = 1 };

int cf_batch_create(uint64_t hbm_budget_bytes /* 0: no limit */, cf_batch** out);
int cf_batch_destroy(cf_batch* b);
/* Tiles of meshes added afterwards: threads 128/256/512/1024, elems <= threads (0: = threads). */
int cf_batch_set_tiles(cf_batch* b, uint32_t tile_threads, uint32_t tile_elems);
/* CG iterations per recorded chunk (1..1024; default 8) and graph use (default 1). */
int cf_batch_set_schedule(cf_batch* b, uint32_t cg_chunk, int use_graphs);
/* CF_BATCH_STEP_* (results are bit-identical under both). */
int cf_batch_set_step_mode(cf_batch* b, int mode);
int cf_batch_add_mesh(cf_batch* b, uint32_t n_nodes, const float* h_ref_pos, uint32_t n_tets,
                      const uint32_t* h_tets, uint32_t* mesh_id);
/* count meshes at once, tiles built on up to `threads` host threads (0: hardware). */
int cf_batch_add_meshes(cf_batch* b, uint32_t count, const uint32_t* n_nodes, const float* const* h_ref_pos,
                        const uint32_t* n_tets, const uint32_t* const* h_tets, uint32_t threads,
                        uint32_t* mesh_ids);
/* Fails while a simulation uses the mesh. */
int cf_batch_remove_mesh(cf_batch* b, uint32_t mesh_id);
/* Bytes a simulation of this mesh takes from the budget. */
int cf_batch_sim_bytes(const cf_batch* b, uint32_t mesh_id, uint64_t* bytes);
/* New active simulation at rest (x = rest positions, v = 0) unless h_pos/h_vel
   are given; KNC as cf_solid_set_knc (h_knc_targets NULL: rest positions). */
int cf_batch_add_sim(cf_batch* b, uint32_t mesh_id, const cf_solid_params* params, uint32_t n_knc,
                     const uint32_t* h_knc_nids, const float* h_knc_targets, const float* h_pos,
                     const float* h_vel, uint32_t* sim_id);
int cf_batch_remove_sim(cf_batch* b, uint32_t sim_id);
/* Active simulations are stepped by cf_batch_fixed_step (new ones start active). */
int cf_batch_set_active(cf_batch* b, uint32_t sim_id, int active);
int cf_batch_get_state(cf_batch* b, uint32_t sim_id, float* h_pos, float* h_vel);
int cf_batch_set_state(cf_batch* b, uint32_t sim_id, const float* h_pos, const float* h_vel);
/* One fixed step (each simulation's own fixed_dt * time_scale) of every active
   simulation. stats: NULL, or one entry per active simulation in the order of
   cf_batch_active_ids (step_ms: device time of the whole batched step). */
int cf_batch_fixed_step(cf_batch* b, cf_step_stats* stats);
int cf_batch_active_ids(const cf_batch* b, uint32_t* ids /* num_active */);
/* Iteration-level batching (the serving loop's engine): queue `steps` more
   fixed steps for a simulation; a round then runs, in the same launches, the
   prologue of every simulation starting a step, `iters` CG iterations of every
   simulation iterating, and the epilogue of every simulation whose solve
   stopped, which starts its next queued step in the following round. Each
   step's arithmetic is the same as in cf_batch_fixed_step (bit-identical
   results). steps_done: NULL or, per active simulation in the order of
   cf_batch_active_ids, its completed steps since it was added; busy: NULL or
   the simulations with work left. Do not mix with cf_batch_fixed_step while
   steps are queued or running. */
int cf_batch_request_steps(cf_batch* b, uint32_t sim_id, uint32_t steps);
int cf_batch_round(cf_batch* b, uint32_t iters, uint32_t* steps_done, uint32_t* busy);
int cf_batch_info_get(const cf_batch* b, cf_batch_info* info);
/* Mean device ms of one batched K*x over the active simulations (element
   stage + node gather on the current rotations), reps launches. */
int cf_batch_time_kx(cf_batch* b, uint32_t reps, double* ms);
int cf_batch_synchronize(cf_batch* b);

/* ---------------------------------------------------------------- 2D solids
 * The 2D twin: TriangleElementSet2 + LeafDSH_Solid2D_FEM (QIE-v, plane-strain
 * linear triangles, Ke = A B^T E B, ms/fem/TriangleElement2.h:71-139;
 * BeginEvaluation ms/fem/TriangleElementSet2.cpp:183-262; C_f_e :1253-1286;
 * C_df_x :1859-1915; lumped masses ds/dsh/LeafDSH_Solid2D_FEM.cpp:1842-1863;
 * Step_QuasiImplicitEuler_v :1287-1403). Node vectors are 2N floats.
 * Rotations: the closed-form 2x2 polar factor for det F > threshold (the
 * reference's MSVD and DAPD both take it there); below it MSVD takes the 2x2
 * SVD of ms/fem/Rotation.cpp:27-36 (Mal/GMatDecomposition.h:255-345) and DAPD
 * the projected rotation of the triangle's node of collapse (NoC; ComputeNoC
 * TriangleElementSet2.cpp:739-835, Compute_R_DAPD_Degenerate Rotation.cpp:300-362),
 * kept per triangle and updated from the previous positions as
 * UpdateDegeneration (:583-640) does; without a NoC R = Id (:1100-1105), counted
 * in num_fallback. Previous positions start at the rest pose; set_state leaves them.
 * A batch steps each mesh inside one workgroup by a group of 1, 2, 4 or 8
 * wavefronts (up to 2048 triangles, 1536 nodes and 16 colours of a greedy
 * triangle colouring; cf_batch2_add_sim returns CF_ERR_UNSUPPORTED beyond),
 * all its simulations in ONE launch; each simulation's result is independent
 * of the others in the batch. cf_solid2 takes any size: a mesh beyond one
 * workgroup is stepped by multi-workgroup kernels (one thread per triangle or
 * node per pass, nodal forces summed in the reference's serial triangle order,
 * the CG on the device with the 3D leaf's stop flags). */
typedef struct cf_batch2 cf_batch2;
/* Rotations of n triangles from their deformation gradients F (4 floats each,
   row-major) and det F, as the 2D step computes them: ms/fem/Rotation.cpp:19-37
   (Compute_R_SVD, rm = CF_RM_MSVD) and Rotation.h:20 (Compute_R_PD, rm =
   CF_RM_DAPD) as called by TriangleElementSet2::ComputeRotations_MSVD/DAPD
   (TriangleElementSet2.cpp:1007-1018,1078-1110). Host arrays; R is 4n floats.
   Stateless: no node of collapse, so DAPD below the threshold gives R = Id and
   fallback (may be NULL) gets 1 there. */
int cf_rotations2(const float* h_F, const float* h_detF, uint32_t n, int32_t rm, float threshold, float* h_R,
                  uint32_t* h_fallback);
int cf_batch2_create(cf_batch2** out);
int cf_batch2_destroy(cf_batch2* b);
/* A simulation at rest (x = rest positions, v = 0), KNC nodes held at rest.
   Simulation ids are 0, 1, 2, ... in order of addition. */
int cf_batch2_add_sim(cf_batch2* b, uint32_t n_nodes, const float* h_ref_pos2, uint32_t n_tris,
                      const uint32_t* h_tris, const cf_solid_params* params, uint32_t n_knc,
                      const uint32_t* h_knc_nids, uint32_t* sim_id);
/* One fixed step of every simulation; stats: NULL or one per simulation (id order). */
int cf_batch2_fixed_step(cf_batch2* b, cf_step_stats* stats);
int cf_batch2_get_state(cf_batch2* b, uint32_t sim_id, float* h_pos2, float* h_vel2);
int cf_batch2_set_state(cf_batch2* b, uint32_t sim_id, const float* h_pos2, const float* h_vel2);
int cf_batch2_node_masses(const cf_batch2* b, uint32_t sim_id, float* h_mass);
/* Elements of the simulation's last fixed step whose DAPD rotation fell back
   to R = Id (det F <= threshold without a NoC; TriangleElementSet2.cpp:1100-1105). */
int cf_batch2_fallbacks(cf_batch2* b, uint32_t sim_id, uint32_t* num_fallback);
/* The simulation's node of collapse per triangle (0, 1, 2; 3 = none), in the
   caller's triangle order; tracked for DAPD simulations only. */
int cf_batch2_get_noc(cf_batch2* b, uint32_t sim_id, uint8_t* h_noc);
int cf_batch2_info(const cf_batch2* b, uint32_t* num_sims, uint64_t* tris, uint64_t* nodes, uint32_t* workgroups);

/* LeafDSH_Solid2D_FEM drop-in (Create / Step / FixedStep / QueryState): one
   simulation; the batched kernel launched for this mesh alone when it fits
   one workgroup, the multi-workgroup kernels otherwise. */
typedef struct cf_solid2 cf_solid2;
int cf_solid2_create(uint32_t n_nodes, const float* h_ref_pos2, uint32_t n_tris, const uint32_t* h_tris,
                     const cf_solid_params* params, uint32_t n_knc, const uint32_t* h_knc_nids, cf_solid2** out);
int cf_solid2_destroy(cf_solid2* s);
int cf_solid2_fixed_step(cf_solid2* s, cf_step_stats* stats);
int cf_solid2_step(cf_solid2* s, float dt, cf_step_stats* stats);
int cf_solid2_get_state(const cf_solid2* s, float* h_pos2, float* h_vel2);
int cf_solid2_set_state(cf_solid2* s, const float* h_pos2, const float* h_vel2);
int cf_solid2_fallbacks(cf_solid2* s, uint32_t* num_fallback);  /* as cf_batch2_fallbacks */
int cf_solid2_get_noc(cf_solid2* s, uint8_t* h_noc);               /* as cf_batch2_get_noc */

/* ---------------------------------------------------------------- serving
 * A real-time request loop over a batch (no reference counterpart: the
 * reference steps the leaves of one scene; this serves many independent
 * leaves). Requests arrive at given times; each asks for `steps` fixed steps
 * of a solid on one of the given meshes. The loop admits waiting requests in
 * arrival order while their bytes fit the HBM budget and the tets in flight
 * stay under max_tick_tets, keeps their state on the device, and runs rounds
 * (cf_batch_round with cg_chunk iterations): every round advances every
 * admitted request, and a step completes in the round its own CG stops
 * (continuous batching at the iteration level: a request admitted while others
 * run joins the next round; a finished one leaves and frees its memory).
 * Meshes are uploaded on first use and evicted, least recently used first,
 * when unused and memory is needed. A step's latency runs from the moment it
 * was requested (arrival for the first step, the previous step's result for
 * the others) to the end of the round that completes it. */
typedef struct cf_serve_mesh {
    uint32_t n_nodes;
    const float* ref_pos;
    uint32_t n_tets;
    const uint32_t* tets;
} cf_serve_mesh;
typedef struct cf_serve_request {
    double arrival_s;          /* seconds after the start of the run */
    uint32_t mesh;             /* index into the mesh list */
    uint32_t steps;            /* fixed steps requested (>= 1) */
    float young_modulus, poisson_ratio;
    uint32_t n_knc;            /* nodes held at their rest positions */
    const uint32_t* knc_nids;
} cf_serve_request;
typedef struct cf_serve_config {
    cf_solid_params base;      /* dt, density, gravity, CG settings (material from each request) */
    uint64_t hbm_budget_bytes; /* meshes + simulations resident at once */
    uint64_t max_tick_tets;    /* 0: no cap */
    double duration_s;         /* arrivals after this are not served */
    int32_t drain;             /* 1: run until every admitted and waiting request is done */
    uint32_t tile_threads, cg_chunk;
    int32_t use_graphs;        /* 1: record the batched step as HIP graphs */
    double target_round_ms;    /* > 0: the tets cap adapts from max_tick_tets (the starting cap) so that a
                                  round takes about this long: after a round that admission hit the cap
                                  and that took under 0.8 of it the cap grows by 1.25x; after a round
                                  over 1.25 of it the cap shrinks by 0.8x (not below 1/64 of the start).
                                  0: max_tick_tets is fixed. */
} cf_serve_config;
typedef struct cf_serve_result {
    uint32_t requests, admitted, completed, max_active, ticks, mesh_uploads;
    uint64_t steps, tet_steps, peak_bytes;
    double wall_s, steps_per_s, tick_mean_ms;
    double step_p50_ms, step_p99_ms, step_mean_ms, step_max_ms;
    double first_p50_ms, first_p99_ms;  /* arrival -> first step result */
    uint64_t final_tick_tets;           /* the tets cap at the end of the run (0: none) */
} cf_serve_result;
/* h_final_pos: NULL or one pointer per request (NULL entries skipped) that
   receives the final positions (3 n_nodes floats) of completed requests.
   step_latency_ms: NULL or room for sum(steps) values, filled in completion order. */
int cf_serve_trace(const cf_serve_mesh* meshes, uint32_t n_meshes, const cf_serve_request* reqs, uint32_t n_reqs,
                   const cf_serve_config* cfg, float* const* h_final_pos, double* step_latency_ms,
                   cf_serve_result* result);

#ifdef __cplusplus
}
#endif
#endif /* COROFEM_H */
