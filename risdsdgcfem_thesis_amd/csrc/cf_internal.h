// Internal declarations shared by the corofem host and device sources.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <vector>
#include "../../include/corofem.h"

namespace cf {

void set_error(const char* fmt, ...);

#define CF_CHECK_HIP(expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            cf::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
            return CF_ERR_HIP;                                                               \
        }                                                                                    \
    } while (0)

#define CF_TRY(expr)               \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != CF_OK) return rc_; \
    } while (0)

// Threads per block of the element and node kernels (4 wave64s).
constexpr int kBlock = 256;
// Upper bound on the blocks of the grid-stride node/vector/element kernels,
// and so on the FP64 per-block partial sums they write. A one-block kernel
// sums the partials in a fixed order, so reductions are bit-reproducible.
constexpr int kMaxPartials = 8192;

// Elastic constants of the reference's isotropic E matrix
// (ms/fem/TetrahedronElement3.h:141-156): a = E/((1+nu)(1-2nu)).
struct Material {
    float a_k1;  // a(1-nu)   = lambda + 2 mu
    float a_nu;  // a nu      = lambda
    float a_k2;  // a(0.5-nu) = mu
    int32_t mm = CF_MM_C_WRP;           // material model
    float mu = 0.f, lambda = 0.f;       // Lame parameters (TetrahedronElementSet3.cpp:143-148)
    const float* F = nullptr;           // LCM or DM = Exact: 9 planes of the evaluation's F
    const float* w = nullptr;           // LCMH: per-element blend weight (df_x), null = LCM
    int32_t dr = 0;                     // DM = Exact: 0 none, 1 every element (MSVD), 2 det F > thr (DAPD)
    float thr = 0.f;                    // degeneration threshold (dr == 2)
    const float* detF = nullptr;        // per-element det F of the evaluation (dr == 2)
};
Material material_from(float young, float poisson, int32_t mm = CF_MM_C_WRP);
// Rayleigh coefficients c = inv(Omega) (2 zeta, 2 zeta) (LeafDSH_Solid3D_FEM.h:347-354)
void rayleigh_coeffs(float f1, float f2, float zeta, float c[2]);

// Device layout of an element set, elements in tile order (see cf_tiles.cpp).
struct TetDevice {
    uint32_t N = 0, T = 0;
    float* ref = nullptr;          // 3N AoS rest positions
    uint32_t* conn = nullptr;      // 4 planes of T global node ids
    float* dinv = nullptr;         // 9 planes of T (row-major Dm^-1)
    float* vol = nullptr;          // T
    // LDS-tile node gather
    uint32_t num_tiles = 0;
    uint32_t max_tile_elems = 0;
    uint32_t max_tile_nodes = 0;
    uint32_t num_parts = 0;        // sum over tiles of tile node counts
    uint32_t uniform_te = 0;       // every tile but the last has exactly this many elements (0: not so)
    uint32_t tile_threads = 256;   // workgroup size of k_tile_forces (256, 512 or 1024)
    uint32_t tile_grid = 8;        // grid of k_tile_forces (persistent, or one tile per workgroup)
    int tile_persistent = 0;       // 0: one tile per workgroup, 1: persistent pipelined, 2: pipelined kernel, one tile per workgroup
    uint32_t* tile_elem_ptr = nullptr;  // num_tiles+1
    uint32_t* tile_node_ptr = nullptr;  // num_tiles+1
    uint32_t* tile_nodes = nullptr;     // num_parts, global node ids
    float* ref_tile = nullptr;          // 3 num_parts: rest position of each tile node, tile-major
    uint32_t* slot_ptr = nullptr;       // num_parts+1, slot runs (tile slots start at 4*e0)
    uint32_t* lv = nullptr;             // T: the element's 4 tile-local node ids, one byte each
    uint32_t* rk = nullptr;             // T: per vertex, the element's rank in that node's slot run (byte each)
    uint32_t* node_part_ptr = nullptr;  // N+1: node n's partials are [ptr[n], ptr[n+1])
    uint32_t* part_dst = nullptr;       // num_parts: partial entry of (tile, tile node), tile-major
    float* partials = nullptr;          // 3 * num_parts
    double* tile_part = nullptr;        // num_tiles: per-tile FP64 sums (elastic energy)
};

// Reference-order FP32 evaluation of one solid (cf_exact.hip): stored Ke blocks
// in caller element order, node -> (element, vertex) runs in ascending element
// order for the serial scatter, and the scratch of the element and dot kernels.
struct ExactDevice {
    uint32_t N = 0, T = 0;
    float* kb = nullptr;        // 144 planes of T: block (bi, bj) entry (r, c) at plane (4 bi + bj) 9 + 3 r + c
    uint32_t* tets = nullptr;   // 4T node ids, caller order
    uint32_t* tile_of = nullptr;  // T: caller element -> index of its R in the cache planes
    uint32_t* ptr = nullptr;    // N+1: node n's incident (element, vertex) codes are [ptr[n], ptr[n+1])
    uint32_t* code = nullptr;   // 4T codes 4 e + vertex, ascending e per node
    float* w = nullptr;         // 12T vertex terms (also 4T energy terms)
    float* scal = nullptr;      // the dot kernel's result
};
enum ExactElemMode { kExactFe = 0, kExactDfx = 1, kExactV = 2 };
enum ExactNodeMode { kExactRhs = 0, kExactAv = 1 };
struct ExactNodeTerms {
    const uint8_t* knc = nullptr;
    const float* knc_target = nullptr;
    const float *x = nullptr, *v = nullptr, *d = nullptr;
    float* out = nullptr;
    float* y = nullptr;
    float mass = 0.f, dt = 0.f, inv_dt = 0.f, g[3] = {0, 0, 0}, neg_fk = 0.f, m_fm = 0.f;
    int warm_start = 1;
};
int exact_build(uint32_t N, const float* ref, uint32_t T, const uint32_t* tets, const uint32_t* perm, float E,
                float nu, hipStream_t st, ExactDevice& out);
void exact_free(ExactDevice& d);
// Element terms into `out` (null: d.w); R: the cache's rotation planes, r_stride apart.
int launch_exact_elem(const ExactDevice& d, int mode, const float* R, uint32_t r_stride, const float* ref,
                      const float* in, float* out, hipStream_t s);
int launch_exact_node(const ExactDevice& d, int mode, const ExactNodeTerms& t, hipStream_t s);
// mode 0: ns::real_array::dot(a, b, n); mode 1: 0.5 * the ordered sum of a[0..n).
// Synchronizes the stream and returns the value on the host.
int exact_dot(const ExactDevice& d, int mode, uint32_t n, const float* a, const float* b, hipStream_t s, float* h_out);
int launch_exact_vec(int mode, uint32_t n3, float scalar, const uint8_t* knc, float* x, float* r, float* dvec,
                     const float* q, hipStream_t s);
// TetrahedronElement3::Compute_K in the reference's order (host); returns the volume.
float exact_element_stiffness(const float* r0, const float* r1, const float* r2, const float* r3, float E, float nu,
                              float K[144]);

// Host-side tile construction (cf_tiles.cpp).
struct TileHost {
    std::vector<uint32_t> perm;          // internal element -> caller element
    std::vector<uint32_t> conn_planes;   // 4T
    std::vector<uint32_t> lv, rk;        // T each: 4 bytes per element
    std::vector<uint32_t> tile_elem_ptr, tile_node_ptr, tile_nodes, slot_ptr;
    std::vector<uint32_t> node_part_ptr, part_dst;
    uint32_t max_tile_elems = 0, max_tile_nodes = 0;
};
// Tiles hold at most min(tile_elems, max_per_tile) elements and max_per_tile - 1 nodes.
int build_tiles(uint32_t N, uint32_t T, const float* ref_pos, const uint32_t* tets, uint32_t tile_elems,
                uint32_t max_per_tile, TileHost& out);
int tile_grid_size(const TetDevice& t, int device, int persistent);
uint32_t default_tile_elems();

// CG scalar state in device memory, indexed by iteration so that no kernel
// reads a slot that another block of the same launch writes.
struct CgScalars {
    double delta[1024 + 2];  // delta[k]: r^T r entering iteration k
    double dq[1024 + 2];
    int32_t done[1024 + 2];  // done[k]: iteration k does not run
    int32_t reason[1024 + 2];
    float alpha[1024 + 2];
    float beta[1024 + 2];
    double delta0;
    double target;
    double abs_sq;
    uint32_t max_iter;
};
constexpr uint32_t kMaxCgIter = 1024;

// Work buffers of one CG solve: r, d, q (n floats each), the scalar record,
// and FP64 partials (kMaxPartials each) for d.q and r.r.
struct CgWork {
    float *r, *d, *q;
    CgScalars* st;
    double *part_a, *part_b;
};
int cg_solve_on(const CgWork& w, uint32_t n, hipStream_t st, cf_av_fn av, cf_constraints_fn cons, void* user,
                float* x, const float* b, float rel, float abs_eps, uint32_t max_it, int32_t* h_flag,
                cf_cg_result* res);


// ---------------------------------------------------------------- launchers
enum ForceMode { kModeDfx = 0, kModeFe = 1 };
// kGatherCgInit: the matvec q = S A d followed by the CG start r = S(b - q),
// d = r (ILSS_CG.h:55-63) with per-chunk r.r partials; q is not stored.
enum GatherMode { kGatherAccumulate = 0, kGatherMatvec = 1, kGatherRhs = 2, kGatherRhsDx = 3, kGatherCgInit = 4 };

struct NodeTerms {  // per-node terms fused into the node gather
    const float* d = nullptr;       // matvec: direction d (3N)
    float* out = nullptr;           // accumulate: += ; matvec: q ; rhs: b
    float* y = nullptr;             // rhs: warm start y
    const float* x = nullptr;       // rhs: positions (for KNC velocity)
    const float* xk = nullptr;      // FIE-dx rhs: the Newton iterate x_k
    const float* v = nullptr;       // rhs: velocities
    const uint8_t* knc_mask = nullptr;
    const float* knc_target = nullptr;
    float mass = 0.f;               // uniform node mass (3D)
    const float* mass_array = nullptr;  // per-node masses (2D) or null
    float fM = 0.f, fK = 0.f;
    float dt = 0.f, inv_dt = 0.f;
    float g[3] = {0, 0, 0};
    int warm_start = 1;
    double* partials = nullptr;     // matvec: per-block d.q ; FIE-dx rhs: per-block b.b ; CG start: r.r
    const float* b = nullptr;       // CG start: right-hand side
    float* r_out = nullptr;         // CG start: residual r
    float* d_out = nullptr;         // CG start: direction d = r
};

int launch_precompute(const TetDevice& t, hipStream_t s);
// Per-evaluation outputs of the LCM materials: F (9 planes) and, for LCMH,
// the blend weight w = clamp01(tr|R^T F - I| inv_k / 3) (1 when !hybrid).
struct LcmOut {
    float* F = nullptr;
    float* w = nullptr;
    float inv_k = 1.f;
    bool hybrid = false;
};
// Degenerate-element state of BeginEvaluation: the per-element direction of
// collapse (one byte, 4*vit0 + vit1; kDocInvalid = (3,0)), kept across
// evaluations and updated from prev_x -> x, and the DAPD projection factors
// (m_DAPD_L_Factor, m_DAPD_NL_Factor, m_DAPD_NL_Exponent). doc == nullptr: no
// tracking, DAPD elements at or below the threshold take R = Id.
constexpr uint8_t kDocInvalid = 12;
struct DegenArgs {
    const float* prev_x = nullptr;  // nullptr: x
    uint8_t* doc = nullptr;
    const float* vol = nullptr;
    float l_factor = 1.f, nl_factor = 0.f, nl_exponent = 2.f;
};
// blk_count holds 2*nblk counters: [0,nblk) det F <= 0, [nblk,2nblk) DAPD R=Id fallbacks
int launch_begin_eval(const TetDevice& t, const float* x, int rm, float thr, float* R, float* detF,
                      uint32_t* blk_count, double* blk_sum, float* blk_min, int nblk, hipStream_t s,
                      const LcmOut& lcm = LcmOut(), const DegenArgs& dg = DegenArgs());
int launch_forces(const TetDevice& t, int mode, const float* in, const float* R, Material m, hipStream_t s);
// BeginEvaluation fused with the front of a QIE-v step, for C_WRP without the
// exact rotation differential: per tile, x and the rest positions staged once
// give each element's F, det F and rotation (stored as launch_begin_eval
// stores them, with the same degeneration update), then f_e(x) into
// t.partials and K y into partials2, y being the CG start of the right-hand
// side gather. Statistics per tile: st_count [0, num_tiles) det F <= 0,
// [num_tiles, 2 num_tiles) DAPD fallbacks; st_sum, st_min.
struct FrontOut {
    int rm = 0;
    float thr = 0.f;
    float* R = nullptr;
    float* detF = nullptr;
    uint32_t* st_count = nullptr;
    double* st_sum = nullptr;
    float* st_min = nullptr;
    DegenArgs dg;
    float* partials2 = nullptr;
};
int launch_qie_front(const TetDevice& t, const float* x, Material m, const FrontOut& fo, const NodeTerms& yt,
                     hipStream_t s);
// After the fused front: b and y (right-hand side gather) and the CG start
// r = S(b - A y), d = r with per-chunk r.r partials, in one pass; nt carries
// both gathers' terms (out = b, y, r_out, d_out, partials = the r.r partials)
// and ay the K y partials.
int launch_rhs_start(const TetDevice& t, const float* ay, const NodeTerms& nt, hipStream_t s);
// C_f_d element stage (Rayleigh damping of the velocities v) into the tile partials
struct DampArgs {
    const float* inv_mass;  // N
    float c0, c1;
};
int launch_damping(const TetDevice& t, const float* v, const float* R, Material m, const DampArgs& d, hipStream_t s);
int launch_forces_skip(const TetDevice& t, int mode, const float* in, const float* R, Material m,
                       const int32_t* skip, hipStream_t s, bool reverse = false);
int launch_gather(const TetDevice& t, int mode, const NodeTerms& nt, int nblk, hipStream_t s);
// kGatherMatvec launches its own grid of matvec_chunks(N) workgroups (nblk is
// ignored) and writes that many d.q partials.
int launch_gather_skip(const TetDevice& t, int mode, const NodeTerms& nt, int nblk, const int32_t* skip,
                       hipStream_t s);
// Chunk layouts of the leaf CG, shared by the plain and batched paths so that
// both sum the same partials in the same order. The matvec gather (and the
// CG start fused into it) runs in chunks of kGatherNpt kBlock nodes. The CG
// vector kernels run in chunks of kBlock float4 rows of kCgVpt rows each where
// that still makes at least kCgMinChunks chunks (large meshes: fewer partials
// to sum, more loads in flight per thread), else of one row (small meshes:
// more, shorter workgroups; their kernels are latency-bound). Where a layout
// would make more than kMaxPartials chunks, a chunk takes more nodes (rows).
constexpr uint32_t kGatherNpt = 1;
constexpr uint32_t kCgVpt = 4;
constexpr uint32_t kCgMinChunks = 256;
inline uint32_t matvec_npt(uint32_t N) {
    const uint32_t cap = (uint32_t)kMaxPartials * kBlock;
    const uint32_t need = (N + cap - 1) / cap;
    return need > kGatherNpt ? need : kGatherNpt;
}
inline uint32_t matvec_chunks_of(uint32_t N) {
    const uint32_t per = matvec_npt(N) * kBlock;
    const uint32_t c = (N + per - 1) / per;
    return c ? c : 1u;
}
inline uint32_t cg_rows(uint32_t n) {
    const uint32_t n4 = (n + 3) / 4;
    const uint32_t cap = (uint32_t)kMaxPartials * kBlock;
    const uint32_t need = (n4 + cap - 1) / cap;
    const uint32_t base = (n4 >= kCgVpt * kBlock * kCgMinChunks) ? kCgVpt : 1u;
    return need > base ? need : base;
}
inline uint32_t cg_chunks(uint32_t n) {
    const uint32_t per = cg_rows(n) * kBlock;
    const uint32_t c = ((n + 3) / 4 + per - 1) / per;
    return c ? c : 1u;
}
int matvec_chunks(uint32_t N);
int launch_energy(const TetDevice& t, const float* x, const float* R, Material m, double* blk, int nblk,
                  double* total, hipStream_t s);
int vec_blocks(uint32_t n);    // vector kernels over n floats (float4 per thread)
int node_blocks(uint32_t n);   // one node (or element) per thread

// CG kernels over n = 3N floats
int launch_cg_init(uint32_t n, const float* q, const float* b, float* r, float* d, const uint8_t* mask,
                   double* rr_part, int nblk, hipStream_t s);
int launch_cg_start(CgScalars* st, const double* rr_part, int nblk_prev, float rel, float abs_eps,
                    uint32_t max_iter, hipStream_t s);
int launch_cg_dot(uint32_t n, const float* a, const float* b, double* part, int nblk, hipStream_t s);
// One CG iteration after q = S A d (alpha, x and r, then beta, the stop
// test and d); r.r partials per vec_blocks(n) chunk.
// alpha ; r -= alpha q ; r.r partials (x is updated by the d kernel)
int launch_cg_step_xr(uint32_t n, uint32_t k, CgScalars* st, float* r, const float* q, const double* dq_part,
                      int ndq, double* rr_part, hipStream_t s);
// beta and the stop test ; x += alpha d ; d = S(beta d + r) unless stopped.
// host_done (pinned host memory, or null) receives the stop flag done[k + 1].
int launch_cg_step_d(uint32_t n, uint32_t k, CgScalars* st, float* x, float* d, const float* r,
                     const uint8_t* mask, const double* rr_part, hipStream_t s, int32_t* host_done = nullptr);
// Workgroups of the two CG step kernels (0 = default 1024; at most kMaxPartials).
int set_cg_workgroups(uint32_t n);
uint32_t get_cg_workgroups();
// FIE-dx: x_k += y ; v_k = (x_k - x) / dt ; per-block y.y into part (nblk blocks)
int launch_nr_update(uint32_t n, float* xk, float* vk, const float* x, const float* y, float inv_dt, double* part,
                     int nblk, hipStream_t s);
// x = x_k ; v = v_k, scaled by (1 - dt drag) when drag > 0
int launch_nr_finish(uint32_t n, float* x, float* v, const float* xk, const float* vk, float dt, float drag,
                     hipStream_t s);
int launch_integrate(uint32_t n, float* x, float* v, const float* y, float dt, float drag, hipStream_t s);
int launch_zero_mask(uint32_t n_nodes, float* w, const uint8_t* mask, hipStream_t s);


// ---------------------------------------------------------------- batches
// A batch steps many simulations together. Every simulation keeps the arrays
// of the single-mesh path (tiles, partials, CG vectors) and its own FP64
// partial buffers; the batched kernels find their simulation and their block
// within that simulation's grid through block tables, and run the same block
// bodies, so a simulation's arithmetic does not depend on the batch around it.
struct SimDev {
    // sizes and per-simulation grids (same grids as the single-mesh path)
    uint32_t N, T, num_tiles;
    uint32_t nb_elem;  // node_blocks(T): BeginEvaluation, energy
    uint32_t nb_node;  // node_blocks(N): gathers
    uint32_t nb_init;  // node_blocks(3N): CG init, integration
    uint32_t nb_vec;   // cg_chunks(3N): CG vector updates (the leaf's chunk layout)
    uint32_t nb_mv;    // matvec_chunks_of(N): matvec gather and CG start chunks
    uint32_t npt_mv;   // matvec_npt(N): nodes per thread of one matvec chunk
    uint32_t rows_cg;  // cg_rows(3N): float4s per thread of one CG chunk
    // topology (may be shared by several simulations)
    const uint32_t* conn;
    const uint32_t* tile_elem_ptr;
    const uint32_t* tile_node_ptr;
    const uint32_t* tile_nodes;
    const uint32_t* slot_ptr;
    const uint32_t* node_part_ptr;
    const uint32_t* part_dst;
    const uint32_t* lv;
    const uint32_t* rk;
    const float* ref;
    const float* ref_tile;  // rest positions per tile node, tile-major
    const float* dinv;
    const float* vol;
    // state of this simulation
    float *x, *v, *b, *y, *r, *d, *q, *R, *detF, *partials, *knc_target;
    const uint8_t* knc_mask;  // null without KNC
    double* part_a;           // nb_mv d.q partials
    double* part_b;           // max(nb_mv, nb_vec) r.r partials
    double* part_e;           // num_tiles energy partials
    uint32_t* eval_count;     // 2 nb_elem
    double* eval_sum;         // nb_elem
    float* eval_min;          // nb_elem
    // parameters
    Material m;
    float mass, fM, fK, dt, inv_dt, drag, thr, rel, abs_eps;
    float g[3];
    int32_t rm, warm_start, compute_energy;
    uint32_t max_iter;
    // DAPD degenerate elements: previous positions (m_vecPos0, written by the
    // integration) and per-element DoC; null unless rm is DAPD
    float* x0;
    uint8_t* doc;
    float dapd_l, dapd_nl, dapd_e;
};

// CG scalars of one simulation for the current solve, its step results, and
// where it is in its step: idle, starting (prologue due), iterating (CG), or
// ending (epilogue due). Every batched kernel only touches simulations in its
// own phase, so steps of different simulations may overlap in one round.
enum SimPhase { kPhaseIdle = 0, kPhaseStart = 1, kPhaseIter = 2, kPhaseEnd = 3 };
struct SimCg {
    double delta, delta0, target, abs_sq, dq;
    float alpha, beta;
    uint32_t iters, max_iter;
    int32_t done, reason;  // done: 1 stopped by the tests, 2 breakdown
    double energy, sum_degenerate;
    uint32_t num_degenerate, num_fallback;
    float min_det;
    int32_t phase;
    uint32_t steps_left;   // steps to start after the current one
    uint32_t steps_done;   // steps completed since the simulation was added
};

struct BlockRef { uint32_t slot, blk; };  // slot 0xffffffff: idle block
constexpr uint32_t kIdleSlot = 0xffffffffu;

struct BatchTables {
    const SimDev* sims;
    SimCg* cgs;
    const BlockRef* tile; uint32_t n_tile;
    const BlockRef* elem; uint32_t n_elem;
    const BlockRef* node; uint32_t n_node;
    const BlockRef* init; uint32_t n_init;
    const BlockRef* vec; uint32_t n_vec;
    const BlockRef* mv; uint32_t n_mv;
    const uint32_t* active; uint32_t n_active;  // slots of the stepped simulations
    uint32_t* counter;                          // simulations still iterating
    uint32_t* busy;                             // simulations not idle after a round
    uint32_t max_tn, max_te, tile_threads;
    // Lists compacted on the device (kb_compact_*): when set, the CG-loop
    // kernels take their entry counts from dyn (kDyn*) and walk the entries
    // with a stride of their grid, whose size is then the n_* above.
    const uint32_t* dyn = nullptr;
};

// dyn[]: entries of the tile list (padded to a multiple of 8), of the matvec
// and CG vector lists, the simulations still iterating, the tile list's
// entries per XCD, and its unpadded length.
enum DynCount { kDynTile = 0, kDynMv = 1, kDynVec = 2, kDynActive = 3, kDynTilePer = 4, kDynTiles = 5, kDynCount = 8 };

// Where the device compaction writes (capacities of the full tables).
struct CompactOut {
    BlockRef* tile;
    BlockRef* mv;
    BlockRef* vec;
    uint32_t* active;
    uint32_t* off;  // 4 per simulation of the full list: tile, mv, vec and active offsets
    uint32_t* dyn;
};

// element-stage input: y in the prologue, d in the CG loop, y for every
// simulation regardless of phase (K*x timing)
enum BatchVec { kSelY = 0, kSelD = 1, kSelYAll = 2 };
int launch_batch_begin_eval(const BatchTables& t, hipStream_t s);
int launch_batch_forces(const BatchTables& t, int mode, int sel, hipStream_t s);
int launch_batch_gather_rhs(const BatchTables& t, hipStream_t s);
int launch_batch_gather_matvec(const BatchTables& t, int sel, hipStream_t s);
int launch_batch_cg_start(const BatchTables& t, hipStream_t s);
int launch_batch_cg_alpha(const BatchTables& t, hipStream_t s);
int launch_batch_update_xr(const BatchTables& t, hipStream_t s);
int launch_batch_cg_beta(const BatchTables& t, hipStream_t s);
int launch_batch_update_d(const BatchTables& t, hipStream_t s);
int launch_batch_integrate(const BatchTables& t, hipStream_t s);
int launch_batch_energy(const BatchTables& t, hipStream_t s);
int launch_batch_finish(const BatchTables& t, hipStream_t s);
int launch_batch_fill_u32(uint32_t* p, uint32_t value, uint32_t n, hipStream_t s);
// every active simulation: start one step now (phase start, no further steps)
int launch_batch_arm_all(const BatchTables& t, hipStream_t s);
// (slot, steps) pairs: queue steps; an idle simulation starts at once
int launch_batch_arm(const BatchTables& t, const uint2* list, uint32_t n, hipStream_t s);
int launch_batch_count_busy(const BatchTables& t, hipStream_t s);
// The CG-loop lists of the simulations of t (full tables) still iterating,
// built on the device: offsets by a one-workgroup scan, then the entries.
// Also clears the busy counter.
int launch_batch_compact(const BatchTables& t, const CompactOut& o, uint32_t fill_grid, hipStream_t s);
// LDS bytes of one batched tile workgroup of the given threads
size_t batch_tile_lds_bytes(uint32_t tile_threads);

// ---------------------------------------------------------------- 2D resident step
// A 2D mesh small enough to live in one workgroup is stepped entirely inside
// it (BeginEvaluation, f_e, RHS, the whole CG loop, integration, energy).
// Workgroups have 512 threads (8 wavefronts) split into groups of 1, 2, 4 or
// 8 wavefronts, one mesh per group; each thread keeps up to kTriRegs
// triangles and kNodeRegs nodes in registers, node vectors shared between
// threads live in LDS. Groups of several wavefronts share the workgroup's
// barriers: they step in lockstep (a finished group waits through the
// barriers) but never share data, so a mesh's arithmetic does not depend on
// its neighbours. Triangles are sorted by a greedy colouring (no two
// triangles of one colour share a node), so nodal forces are accumulated in
// LDS colour by colour without atomics, in a fixed order.
constexpr int kTriRegs = 4, kNodeRegs = 3, kMaxColors = 16, k2Threads = 512;
constexpr int kGroupTris = 64 * kTriRegs, kGroupNodes = 64 * kNodeRegs;  // per wavefront of a group

struct Mesh2Dev {
    uint32_t N, T, C, gw;          // gw: wavefronts of the mesh's group (1, 2, 4 or 8)
    uint32_t node_off, tri_off, color_off, pad;
    Material m;                    // plane strain: a_k1 = a - a nu, a_nu = a nu, a_k2 = a/2 - a nu
    float fM, fK, dt, inv_dt, g0, g1, rel, abs_eps, thr, drag;
    int32_t rm, warm_start, compute_energy;
    uint32_t max_iter;
    float dapd_l, dapd_nl, dapd_e;  // DAPD degenerate projection factors
};

struct Sim2Out {
    int32_t iters, reason;
    float rel, abs_res;
    double energy, sum_degenerate;
    uint32_t num_degenerate, num_fallback;
    float min_det, pad;
};

struct Work2 {  // one workgroup: 8/gw meshes of group size gw (kIdleSlot: empty group)
    uint32_t gw, cmax;  // cmax: most colours among the workgroup's meshes
    uint32_t sim[8];
};

struct Batch2Args {
    const Work2* work;
    const Mesh2Dev* meshes;
    const float2* ref;       // per node (node_off)
    const float* mass;
    const uint8_t* knc;
    float2* x;
    float2* v;
    const uint32_t* tri_ids; // per triangle: 3 local node ids, colour-sorted (tri_off)
    const float4* dinv;      // Dm^-1 row-major
    const float* area;
    const uint16_t* color_start;  // per mesh C+1 entries (color_off)
    Sim2Out* out;
    // DAPD degenerate triangles: previous positions (m_vecPos0, node_off) and
    // node of collapse per triangle (tri_off, 3 = none); used by DAPD meshes
    float2* x0;
    uint8_t* noc;
    int32_t* nonfinite;  // set to 1 by any mesh whose CG met a NaN or Inf r.r
};

int launch_batch2_step(const Batch2Args& a, uint32_t n_work, hipStream_t s);

// A 2D mesh beyond one workgroup (cf_solid2 above the resident kernel's
// triangle, node or colour limits), stepped by multi-workgroup kernels on its
// own arrays: triangles in the caller's order, per-(triangle, vertex) forces
// tf summed per node over its incident entries in ascending triangle order
// (the reference's serial scatter order), node vectors as float2.
struct Large2Dev {
    uint32_t N, T, nblk_tri, nblk_node;
    Mesh2Dev c;               // constants (material, dt, CG settings); its N, T and offsets are unused
    const float2* ref;
    const float* mass;
    const uint8_t* knc;
    float2 *x, *v, *x0;       // x0: previous positions (DAPD)
    const uint32_t* tri;      // 3 node ids per triangle
    const float4* dinv;       // Dm^-1 row-major
    const float* area;
    uint8_t* noc;             // node of collapse per triangle (3 = none)
    float4* R;                // the step's rotation per triangle
    float2* tf;               // 3 vertex forces per triangle
    const uint32_t* ptr;      // N + 1: node n's entries are code[ptr[n] .. ptr[n+1])
    const uint32_t* code;     // 3 T entries 3 e + vertex, ascending e per node
    float2 *b, *y, *r, *d, *q;
    CgScalars* st;
    double *part_a, *part_b;  // d.q and r.r partials, one per node block (nblk_node)
    double* stat;             // 4 nblk_tri: degenerate count, degenerate det F sum, fallbacks, energy
    float* stat_min;          // nblk_tri: minimum det F
    Sim2Out* out;             // energy and the BeginEvaluation statistics of the step
};
int launch_l2_begin(const Large2Dev& a, hipStream_t s);
int launch_l2_elem(const Large2Dev& a, bool fe, const float2* in, const int32_t* skip, hipStream_t s);
int launch_l2_rhs(const Large2Dev& a, hipStream_t s);
int launch_l2_matvec(const Large2Dev& a, bool init, const float2* in, const int32_t* skip, hipStream_t s);
int launch_l2_xr(const Large2Dev& a, uint32_t k, hipStream_t s);
int launch_l2_d(const Large2Dev& a, uint32_t k, int32_t* host_done, hipStream_t s);
int launch_l2_integrate(const Large2Dev& a, hipStream_t s);
int launch_l2_energy(const Large2Dev& a, hipStream_t s);
int launch_l2_finish(const Large2Dev& a, hipStream_t s);
// A solve's result from its scalar record (the iteration that stopped, its
// reason and residuals; CF_CG_NONFINITE when a reduction was NaN or Inf).
void cg_result_from(const CgScalars& hs, uint32_t max_it, cf_cg_result* res);
int launch_rotations2(const float* F, const float* det, uint32_t n, int rm, float thr, float* R, uint32_t* fallback,
                      hipStream_t s);
}  // namespace cf
