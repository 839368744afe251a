/*
 * corofem oracle: CPU restatement of the reference's corotational tet FEM
 * implicit step, used ONLY as a checker (tests, smoke, the CPU baseline leg of
 * bench.py). Nothing in the shipped library links or calls it.
 *
 * It follows the reference's FP32 semantics operation by operation:
 *   - box mesher        Geo/shape/TetSolidShape3.cpp:921-999
 *   - element precompute ms/fem/TetrahedronElement3.h:85-173,
 *                        ms/fem/TetrahedronElementSet3.cpp:2113-2166
 *   - BeginEvaluation   ms/fem/TetrahedronElementSet3.cpp:219-281
 *   - rotations         Mal/GMatDecomposition.h:411-432 (Newton polar),
 *                        ms/fem/Rotation.cpp:39-57 + Mal/GMatDecomposition_SVD3x3_PhysBAM.h:246-280
 *   - C_f_e / C_df_x / C_V  ms/fem/TetrahedronElementSet3.cpp:1053-1096,1761-1836,1456-1488
 *     (stored 3x3 Ke blocks, serial scatter in element order)
 *   - CG                ns/ILSS_CG.h:38-151 with ns/real_array.h:34-78 (SSE
 *                        4-lane FP32 reduction order, zero padding lanes)
 *   - QIE-v leaf step   ds/dsh/LeafDSH_Solid3D_FEM.cpp:225-256,898-990,1045-1079,1205-1328,1591-1611
 *   - 2D twin           ms/fem/TriangleElement2.h:71-139, TriangleElementSet2.cpp,
 *                        ds/dsh/LeafDSH_Solid2D_FEM.cpp:1101-1132,1287-1403,1842-1863
 * Build with -ffp-contract=off (no FMA) to keep the reference's rounding.
 */
#ifndef COROFEM_ORACLE_H
#define COROFEM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* rotation methods (ms/fem/Params.h:63-67) */
#define ORC_RM_ID 0
#define ORC_RM_QR 1
#define ORC_RM_MSVD 2
#define ORC_RM_DAPD 3

/* ---- mesh (Geo/shape/TetSolidShape3.cpp:921-999) ---- */
void orc_box_counts(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t* n_nodes, uint32_t* n_tets);
void orc_mesh_box(float hx, float hy, float hz, uint32_t nx, uint32_t ny, uint32_t nz,
                  float* pos, uint32_t* tets);
/* 2D grid (Geo/shape/MeshSolidShape2.cpp:1651-1675) */
void orc_box2_counts(uint32_t nx, uint32_t ny, uint32_t* n_nodes, uint32_t* n_tris);
void orc_mesh_box2(float hx, float hy, uint32_t nx, uint32_t ny, float* pos, uint32_t* tris);

/* ---- math kernels restated ---- */
float orc_det3(const float m[9]);
void orc_inverse3(const float m[9], float out[9]);
void orc_pd3(const float F[9], float detF, float R[9], uint32_t* iters);
void orc_svd3(const float F[9], float U[9], float s[3], float Vt[9]);
void orc_svd_rotation3(const float F[9], float R[9]);
void orc_svd3_sqrt3(const float F[9], float U[9], float s[3], float Vt[9]);
void orc_pd2(const float F[4], float detF, float R[4]);
void orc_svd_rotation2(const float F[4], float detF, float thr, float R[4]);
float orc_compute_K(const float r[12], float E, float nu, float K[144]);
float orc_compute_K2(const float r[6], float E, float nu, float K[36]);
float orc_sse_dot(const float* a, const float* b, uint32_t n);

/* ---- 3D element set + cache ---- */
typedef struct orc_tetset orc_tetset;
orc_tetset* orc_tetset_create(uint32_t n_nodes, const float* ref_pos, uint32_t n_tets,
                              const uint32_t* tets, float young, float poisson);
void orc_tetset_destroy(orc_tetset* s);
float orc_tetset_total_volume(const orc_tetset* s);
void orc_tetset_element(const orc_tetset* s, uint32_t e, float K144[144], float Dinv[9], float* vol);
/* BeginEvaluation: computes F, det F, R per element into the set's cache and
   updates the per-element direction of collapse (DoC) from prev_x -> x
   (prev_x may be NULL: x). R_out (9*T, row-major per element) and detF_out (T)
   may be NULL. */
void orc_begin_evaluation(orc_tetset* s, const float* prev_x, const float* x, int rm,
                          float degenerate_threshold, float* R_out, float* detF_out,
                          uint32_t* num_degenerate, float* sum_degenerate_detF);
void orc_set_rotations(orc_tetset* s, const float* R);  /* load R (9*T) directly */
/* Degenerate elements (DAPD). DoC byte = 4*vit0 + vit1; 12 = invalid. */
void orc_tetset_get_doc(const orc_tetset* s, uint8_t* doc);
void orc_tetset_set_doc(orc_tetset* s, const uint8_t* doc);
void orc_tetset_set_dapd(orc_tetset* s, float l_factor, float nl_factor, float nl_exponent);
uint32_t orc_tetset_num_fallback(const orc_tetset* s);  /* last BeginEvaluation's R = Id count */
int orc_gsl_cubic(double a, double b, double c, double roots[3]);          /* x^3 + a x^2 + b x + c */
int orc_solve_eq3(float a, float b, float c, float d, float roots[3]);     /* a x^3 + b x^2 + c x + d */
uint8_t orc_compute_doc(const float a[12], const float b[12], float volume, float thr,
                        float* toc /* may be NULL; set when a crossing is found */);
int orc_r_dapd_degenerate(float detF, float thr, uint8_t doc, const float p[12], float volume,
                          const float inv_dm[9], float l_factor, float nl_factor,
                          float nl_exponent, float R[9]);
/* Material model (ms/fem/Params.h:51-61): ORC_MM_C_WRP (default, stored Ke),
   ORC_MM_C_LCM or ORC_MM_C_LCMH (Piola forms from the F of the last
   BeginEvaluation), and m_ECIE_k_factor (LCMH blend radius, default 1). */
#define ORC_MM_C_WRP 1
#define ORC_MM_C_LCM 2
#define ORC_MM_C_LCMH 3
void orc_tetset_set_material(orc_tetset* s, int mm, float ecie_k_factor);
/* Differentials method (Params.h:68-72): ORC_DM_TRUNCATED (default) or
   ORC_DM_EXACT (the McAdams rotation differential dR in df_x for MSVD, and for
   DAPD above the degeneration threshold). */
#define ORC_DM_TRUNCATED 0
#define ORC_DM_EXACT 1
void orc_tetset_set_dm(orc_tetset* s, int dm);
/* Compute_dR for one element with Dm^-1 = Id (dDs = dF); returns 0 where the
   reference skips dR (|det(S - tr(S) I)| <= 1e-4). */
int orc_dR(const float F[9], const float dF[9], const float R[9], float dR[9]);
void orc_f_e(const orc_tetset* s, const float* x, float* f);      /* f  += f_e(x)  */
void orc_df_x(const orc_tetset* s, const float* dx, float* df);   /* df += K'(dx) */
float orc_V(const orc_tetset* s, const float* x);
/* f += C_f_d(v): corotational Rayleigh damping with coefficients c0 (mass) and
   c1 (stiffness) and per-node inverse masses (N). */
void orc_f_d(const orc_tetset* s, const float* v, const float* inv_mass, float c0, float c1, float* f);

/* ---- 3D leaf solid (QIE-v, or FIE-dx with params.integrator) ---- */
typedef struct orc_solid_params {
    float fixed_dt;
    float density;
    float young_modulus;
    float poisson_ratio;
    float damping_ratio;
    float rayleigh_freq1, rayleigh_freq2;
    float gravity[3];
    float cg_rel_epsilon, cg_abs_epsilon;
    uint32_t cg_max_iter;
    int rm;
    float degenerate_threshold;
    int warm_start;
    int reduce_fp64;   /* 0: reference FP32 SSE reductions; 1: FP64 reductions */
    int integrator;    /* ORC_IT_QIE_V (0, default) or ORC_IT_FIE_DX (1) */
    uint32_t nr_max_iter;  /* FIE-dx Newton iterations (reference default 3) */
    float nr_abs_epsilon;  /* FIE-dx: stop when |b| < this (reference default 1e-6) */
    int mm;                /* material model, ORC_MM_C_* (0 = C_WRP) */
    float ecie_k_factor;   /* LCMH blend radius (reference default 1) */
    int dm;                /* differentials method, ORC_DM_* (0 = Truncated) */
    float time_scale;      /* m_TimeScale (reference default 1): Step scales dt and fixed_dt by it */
} orc_solid_params;
#define ORC_IT_QIE_V 0
#define ORC_IT_FIE_DX 1

typedef struct orc_step_stats {
    uint32_t num_fixed_steps;
    int32_t cg_iters;
    float cg_rel;
    float cg_abs;
    uint32_t num_degenerate;
    float sum_degenerate_detF;
    float elastic_energy;
    float min_detF;
    int32_t nr_iters;  /* FIE-dx: Newton iterations taken */
    float nr_rel;      /* FIE-dx: last |b| or |y| over the first |b| */
    float nr_abs;      /* FIE-dx: last |b| or |y| */
} orc_step_stats;

typedef struct orc_solid orc_solid;
orc_solid* orc_solid_create(uint32_t n_nodes, const float* ref_pos, uint32_t n_tets,
                            const uint32_t* tets, const orc_solid_params* p);
void orc_solid_destroy(orc_solid* s);
void orc_solid_set_knc(orc_solid* s, uint32_t n, const uint32_t* nids);  /* held at the rest pose */
/* KNC held at given targets (3 floats per listed node), like SetPosKNC */
void orc_solid_set_knc_targets(orc_solid* s, uint32_t n, const uint32_t* nids, const float* targets);
/* One FixedStep(dt); advances the solid's total time by dt. */
void orc_solid_fixed_step(orc_solid* s, float dt, orc_step_stats* st);
/* Step(dt) (LeafDSH_Solid3D_FEM.cpp:225-256): total time += time_scale dt,
   then FixedStep(time_scale fixed_dt) while the accumulator is strictly
   greater than time_scale fixed_dt; num_fixed_steps counts them. */
void orc_solid_step(orc_solid* s, float dt, orc_step_stats* st);
float orc_solid_total_time(const orc_solid* s);
/* One fixed step (each solid's own fixed_dt) of n independent solids over
   `threads` host threads: the mesh-parallel CPU baseline of batches. */
void orc_solids_fixed_step_parallel(orc_solid* const* solids, uint32_t n, uint32_t threads);
/* Rayleigh coefficients c = inv(Omega) (2 zeta, 2 zeta), Omega = [[1/f1, f1], [1/f2, f2]]
   (LeafDSH_Solid3D_FEM.h:347-354, LeafDSH_Solid2D_FEM.h:333-341). */
void orc_rayleigh(float f1, float f2, float zeta, float c[2]);
void orc_solid_get_state(const orc_solid* s, float* x, float* v);
void orc_solid_set_state(orc_solid* s, const float* x, const float* v);
float orc_solid_node_mass(const orc_solid* s);
void orc_solid_rayleigh(const orc_solid* s, float c[2]);
/* Multi-solid universe step (SS_AggregateDSH_Basic::MultiDSH_Solid3D_Update):
   fixed dt = min over the solids, Newton passes = max of their nr_max_iter,
   one accumulator (*acc_time, strict '>'); each fixed step runs the leaves'
   NR_Init / NR_PreStep / NR_MidStep / NR_PostStep / NR_End hooks with the
   frame dt. stats: NULL or one per solid (its last fixed step). Returns the
   number of fixed steps taken. */
uint32_t orc_multi_update(orc_solid* const* solids, uint32_t n, float dt, float* acc_time, orc_step_stats* stats);

/* ---- 2D leaf solid (QIE-v, TriangleElementSet2 + LeafDSH_Solid2D_FEM) ----
   Same parameters (gravity[0..1] used); lumped masses Sum(area/3) * density *
   thickness per node, summed over the node's triangles in ascending triangle
   order. Rotations: closed-form 2x2 polar factor for det F > threshold (both
   MSVD and DAPD); below it MSVD takes the 2x2 SVD branch of Compute_R_SVD and
   DAPD R = Id (its invalid-NoC branch); elements with det F <= 0 are counted in
   num_degenerate. */
typedef struct orc_solid2 orc_solid2;
orc_solid2* orc_solid2_create(uint32_t n_nodes, const float* ref_pos2, uint32_t n_tris, const uint32_t* tris,
                              const orc_solid_params* p, float thickness);
void orc_solid2_destroy(orc_solid2* s);
void orc_solid2_set_knc(orc_solid2* s, uint32_t n, const uint32_t* nids);
void orc_solid2_fixed_step(orc_solid2* s, float dt, orc_step_stats* st);
/* Step(dt) of the 2D leaf (LeafDSH_Solid2D_FEM.cpp:264-294), as orc_solid_step. */
void orc_solid2_step(orc_solid2* s, float dt, orc_step_stats* st);
float orc_solid2_total_time(const orc_solid2* s);
void orc_solids2_fixed_step_parallel(orc_solid2* const* solids, uint32_t n, uint32_t threads);
void orc_solid2_get_state(const orc_solid2* s, float* x, float* v);
void orc_solid2_set_state(orc_solid2* s, const float* x, const float* v);
void orc_solid2_node_masses(const orc_solid2* s, float* m);
float orc_solid2_total_area(const orc_solid2* s);
void orc_fan_weighted_areas(uint32_t n_nodes, uint32_t n_tris, const uint32_t* tris, const float* area, float* wa);
/* element operators on the current rotations (set by the last step or begin_evaluation) */
void orc_solid2_begin_evaluation(orc_solid2* s, const float* prev_x /* may be NULL */, const float* x,
                                 float* R_out, float* detF_out);
/* 2D degenerate triangles (DAPD): node of collapse per triangle (3 = invalid) */
void orc_solid2_get_noc(const orc_solid2* s, uint8_t* noc);
void orc_solid2_set_dapd(orc_solid2* s, float l_factor, float nl_factor, float nl_exponent);
uint32_t orc_solid2_num_fallback(const orc_solid2* s);
uint8_t orc_compute_noc2(const float a[6], const float b[6], float area, float thr, float* toc);
void orc_r_dapd_degenerate2(float thr, uint8_t noc, const float p[6], float area, const float inv_dm[4],
                            float l_factor, float nl_factor, float nl_exponent, float R[4]);
void orc_solid2_f_e(const orc_solid2* s, const float* x, float* f);
void orc_solid2_df_x(const orc_solid2* s, const float* dx, float* df);
float orc_solid2_V(const orc_solid2* s, const float* x);

#ifdef __cplusplus
}
#endif
#endif
