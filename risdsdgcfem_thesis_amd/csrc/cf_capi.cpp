// corofem C-ABI: handles, device residency and the host-side orchestration of
// the reference's element-set / solver / leaf-step interfaces (include/corofem.h).
#include "cf_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <limits>
#include <vector>

namespace cf {

static thread_local char g_err[512] = "";
void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

Material material_from(float E, float nu, int32_t mm) {
    // TetrahedronElement3::Compute_E (ms/fem/TetrahedronElement3.h:141-156)
    Material m;
    const float a = E / ((1.0f + nu) * (1.0f - 2.0f * nu));
    m.a_nu = a * nu;
    m.a_k1 = a * (1.0f - nu);
    m.a_k2 = a * (0.5f - nu);
    // LameParameters_From_YoungAndPoisson (TetrahedronElementSet3.cpp:143-148)
    m.mu = 0.5f * E / (1 + nu);
    m.lambda = E * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
    m.mm = mm;
    return m;
}

static uint32_t g_tile_elems = 0;
static uint32_t g_tile_threads = 256;
static int g_tile_persistent = 0;
static int g_tile_alternate = 0;  // 1: odd CG iterations walk the tiles backwards
static int g_fused_front = 1;     // 1: QIE-v steps fuse BeginEvaluation, f_e and the CG start's K y

template <class T>
static int dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu bytes) failed: %s", count * sizeof(T), hipGetErrorString(e));
        return (e == hipErrorOutOfMemory) ? CF_ERR_OUT_OF_MEMORY : CF_ERR_HIP;
    }
    return CF_OK;
}
template <class T>
static void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}
template <class T>
static int upload(T** p, const std::vector<T>& v, hipStream_t s) {
    CF_TRY(dalloc(p, v.size()));
    if (!v.empty()) CF_CHECK_HIP(hipMemcpyAsync(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return CF_OK;
}

// Rayleigh coefficients c = inv(Omega) (2 zeta, 2 zeta) (LeafDSH_Solid3D_FEM.h:347-354)
void rayleigh_coeffs(float f1, float f2, float zeta, float c[2]) {
    const float o00 = 1.0f / f1, o01 = f1, o10 = 1.0f / f2, o11 = f2;
    const float det = o00 * o11 - o10 * o01;
    const float id = 1.0f / det;
    const float i00 = id * o11, i01 = -id * o01, i10 = -id * o10, i11 = id * o00;
    const float z = 2 * zeta;
    c[0] = i00 * z + i01 * z;
    c[1] = i10 * z + i11 * z;
}

}  // namespace cf

using namespace cf;

// ================================================================ handles
struct cf_tetset {
    int device = 0;
    hipStream_t stream = nullptr;
    cf_fem_params params{};
    Material mat{};
    TetDevice dev;
    std::vector<uint32_t> perm;   // internal -> caller element
    float total_volume = 0.f;
    // edition
    bool editing = false;
    uint32_t expected_T = 0;
    std::vector<float> ref_host;
    std::vector<uint32_t> tets_host;
};

struct cf_cache {
    cf_tetset* set = nullptr;
    float* R = nullptr;
    float* detF = nullptr;
    float* inv_mass = nullptr;
    uint32_t* blk_count = nullptr;
    double* blk_sum = nullptr;
    float* blk_min = nullptr;
    double* blk_energy = nullptr;
    int nblk = 1;
    const float* x = nullptr;
    const float* prev_x = nullptr;
    const float* v = nullptr;
    float time = 0.f, timestep = 0.f;
    uint32_t num_degenerate = 0, num_fallback = 0;
    float sum_degenerate = 0.f, min_det = 0.f;
    bool stats_valid = false;
    // pinned host copies of the BeginEvaluation block statistics
    uint32_t* h_count = nullptr;
    double* h_sum = nullptr;
    float* h_min = nullptr;
    // entries of the last evaluation's statistics: one per BeginEvaluation
    // block (nblk), or one per tile after the fused QIE-v front
    int stat_n = 1;
    float* F = nullptr;   // LCM / LCMH: 9 planes of F at the evaluation
    float* w = nullptr;   // LCMH: per-element blend weight
    uint8_t* doc = nullptr;  // per-element direction of collapse (ElementCache3::m_DoC)
};

// The set's material with the cache's per-evaluation arrays (LCM: F for f_e;
// LCMH: the blend weights for df_x).
static Material cache_mat(const cf_cache* c) {
    const cf_fem_params& p = c->set->params;
    Material m = c->set->mat;
    if (m.mm != CF_MM_C_WRP || p.dm == CF_DM_EXACT) m.F = c->F;
    if (m.mm == CF_MM_C_LCMH) m.w = c->w;
    if (p.dm == CF_DM_EXACT) {  // Compute_dR_From_RM (TetrahedronElementSet3.cpp:894-953)
        m.dr = (p.rm == CF_RM_MSVD) ? 1 : (p.rm == CF_RM_DAPD ? 2 : 0);
        m.thr = p.degenerate_threshold_det_f;
        m.detF = c->detF;
    }
    return m;
}

struct CgBuffers {
    uint32_t n = 0;
    float *r = nullptr, *d = nullptr, *q = nullptr;
    CgScalars* st = nullptr;
    double* part_a = nullptr;  // d.q partials
    double* part_b = nullptr;  // r.r partials
    int32_t* h_flag = nullptr; // pinned
    int32_t* h_done = nullptr; // pinned, kMaxCgIter + 2: done[k + 1], written by the iteration-k d kernel
    int32_t* d_done = nullptr; // h_done's device address
    hipEvent_t ev_chunk[2] = {nullptr, nullptr};  // ends of the two chunks in flight
};

struct cf_cg {
    hipStream_t stream = nullptr;
    CgBuffers b;
};

// Recorded CG chunks of one leaf: graph c runs iterations 8c .. 8c+7 (clipped
// to the iteration cap). Every kernel argument is fixed for the leaf's
// lifetime except those in the key, so the graphs are re-recorded only when
// the key changes.
struct CgGraphs {
    std::vector<hipGraphExec_t> chunk;
    float fM = 0.f, fK = 0.f;
    uint32_t max_it = 0;
    const uint8_t* knc_mask = nullptr;
    int alternate = -1;           // the tile-alternation setting recorded
    uint32_t cg_workgroups = 0;   // the CG step kernels' grid recorded
    void clear() {
        for (hipGraphExec_t g : chunk)
            if (g) (void)hipGraphExecDestroy(g);
        chunk.clear();
    }
};

struct cf_solid {
    cf_solid_params p{};
    cf_tetset* set = nullptr;
    cf_cache* cache = nullptr;
    hipStream_t stream = nullptr;
    uint32_t N = 0;
    float *x = nullptr, *v = nullptr, *x0 = nullptr, *bvec = nullptr, *y = nullptr;
    uint8_t* knc_mask = nullptr;
    float* knc_target = nullptr;
    uint32_t num_knc = 0;
    float node_mass = 0.f;
    float rayleigh[2] = {0, 0};
    float acc_time = 0.f, total_time = 0.f;
    float air_drag_per_second = 0.f;
    CgBuffers cg;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float* scratch_vec = nullptr;
    float *xk = nullptr, *vk = nullptr;  // FIE-dx iterates (m_IE_vec_x_k, m_IE_vec_v_k)
    double* h_part = nullptr;             // pinned: block partials of |b|^2, |y|^2
    CgScalars* h_cgs = nullptr;           // pinned: the CG record, read once the step has run
    double* h_energy = nullptr;           // pinned: V() of the step
    float* ay_part = nullptr;             // K y tile partials of the fused QIE-v front (3 num_parts)
    cf_cg_result last_ls{};               // last linear solve
    int32_t ls_total = 0;                 // linear-solve iterations of the last fixed step
    bool nonfinite = false;               // a reduction of the last fixed step was NaN or Inf
    ExactDevice ex;                       // reference_order: the reference-order FP32 step's data
    CgGraphs graphs;                      // recorded CG chunks
    bool use_graphs = true;               // the leaf solve replays its CG chunks as recorded graphs
};

struct cf_multi {
    std::vector<cf_solid*> solids;
    float acc_time = 0.f;
    bool use_graphs = true;
};

// ================================================================ helpers
static int cg_alloc(CgBuffers& b, uint32_t n) {
    b.n = n;
    CF_TRY(dalloc(&b.r, n));
    CF_TRY(dalloc(&b.d, n));
    CF_TRY(dalloc(&b.q, n));
    CF_TRY(dalloc(&b.st, 1));
    CF_TRY(dalloc(&b.part_a, kMaxPartials));
    CF_TRY(dalloc(&b.part_b, kMaxPartials));
    CF_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&b.h_flag), sizeof(int32_t)));
    CF_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&b.h_done), (kMaxCgIter + 2) * sizeof(int32_t)));
    std::memset(b.h_done, 0, (kMaxCgIter + 2) * sizeof(int32_t));
    CF_CHECK_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&b.d_done), b.h_done, 0));
    for (hipEvent_t& e : b.ev_chunk) CF_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return CF_OK;
}
static void cg_free(CgBuffers& b) {
    dfree(b.r); dfree(b.d); dfree(b.q); dfree(b.st); dfree(b.part_a); dfree(b.part_b);
    if (b.h_flag) (void)hipHostFree(b.h_flag);
    if (b.h_done) (void)hipHostFree(b.h_done);
    for (hipEvent_t& e : b.ev_chunk) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    b.h_flag = nullptr;
    b.h_done = nullptr;
    b.d_done = nullptr;
}

static void tetset_free(cf_tetset* s) {
    TetDevice& t = s->dev;
    dfree(t.ref); dfree(t.conn); dfree(t.dinv); dfree(t.vol);
    dfree(t.tile_elem_ptr); dfree(t.tile_node_ptr); dfree(t.tile_nodes); dfree(t.slot_ptr); dfree(t.ref_tile);
    dfree(t.lv); dfree(t.rk); dfree(t.node_part_ptr); dfree(t.part_dst); dfree(t.partials); dfree(t.tile_part);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    s->stream = nullptr;
}

static bool valid_params(const cf_fem_params* p) {
    if (!p) return false;
    if (!(p->young_modulus > 0) || !(p->poisson_ratio > -1.0f) || !(p->poisson_ratio < 0.5f)) {
        set_error("invalid material: E=%g nu=%g", p->young_modulus, p->poisson_ratio);
        return false;
    }
    if (p->mm != CF_MM_C_WRP && p->mm != CF_MM_C_LCM && p->mm != CF_MM_C_LCMH) {
        set_error("material method %d not implemented (C_WRP, C_LCM, C_LCMH)", p->mm);
        return false;
    }
    if (p->dm != CF_DM_TRUNCATED && p->dm != CF_DM_EXACT) {
        set_error("differentials method %d not implemented (Truncated, Exact)", p->dm);
        return false;
    }
    if (p->rm != CF_RM_ID && p->rm != CF_RM_MSVD && p->rm != CF_RM_DAPD) {
        set_error("rotation method %d not implemented (Id, MSVD, DAPD)", p->rm);
        return false;
    }
    if (!std::isfinite(p->dapd_l_factor) || !std::isfinite(p->dapd_nl_factor) ||
        !std::isfinite(p->dapd_nl_exponent)) {
        set_error("DAPD factors must be finite");
        return false;
    }
    return true;
}

// Builds device arrays from ref_host/tets_host (EndEdition).
static int tetset_bake(cf_tetset* s) {
    const uint32_t N = uint32_t(s->ref_host.size() / 3), T = uint32_t(s->tets_host.size() / 4);
    for (size_t i = 0; i < s->tets_host.size(); i++)
        if (s->tets_host[i] >= N) {
            set_error("element %zu references node %u >= %u", i / 4, s->tets_host[i], N);
            return CF_ERR_INVALID_ARG;
        }
    TileHost th;
    CF_TRY(build_tiles(N, T, s->ref_host.data(), s->tets_host.data(), g_tile_elems, g_tile_threads, th));
    TetDevice& t = s->dev;
    t.N = N; t.T = T;
    hipStream_t st = s->stream;
    CF_TRY(upload(&t.ref, s->ref_host, st));
    CF_TRY(upload(&t.conn, th.conn_planes, st));
    CF_TRY(dalloc(&t.dinv, 9 * (size_t)T));
    CF_TRY(dalloc(&t.vol, T));
    CF_TRY(upload(&t.tile_elem_ptr, th.tile_elem_ptr, st));
    CF_TRY(upload(&t.tile_node_ptr, th.tile_node_ptr, st));
    CF_TRY(upload(&t.tile_nodes, th.tile_nodes, st));
    CF_TRY(upload(&t.slot_ptr, th.slot_ptr, st));
    CF_TRY(upload(&t.lv, th.lv, st));
    CF_TRY(upload(&t.rk, th.rk, st));
    CF_TRY(upload(&t.node_part_ptr, th.node_part_ptr, st));
    CF_TRY(upload(&t.part_dst, th.part_dst, st));
    {
        // rest positions per tile node, tile-major: the tile stage streams them
        std::vector<float> rt(3 * th.tile_nodes.size());
        for (size_t q = 0; q < th.tile_nodes.size(); q++)
            for (int c = 0; c < 3; c++) rt[3 * q + c] = s->ref_host[3 * (size_t)th.tile_nodes[q] + c];
        CF_TRY(upload(&t.ref_tile, rt, st));
    }
    t.num_tiles = uint32_t(th.tile_elem_ptr.size() - 1);
    t.max_tile_elems = th.max_tile_elems;
    t.max_tile_nodes = th.max_tile_nodes;
    t.num_parts = uint32_t(th.tile_nodes.size());
    t.uniform_te = 0;
    if (t.num_tiles > 0) {
        const uint32_t te = th.tile_elem_ptr[1] - th.tile_elem_ptr[0];
        bool uniform = te > 0 && th.tile_elem_ptr[t.num_tiles] - th.tile_elem_ptr[t.num_tiles - 1] <= te;
        for (uint32_t i = 0; uniform && i + 1 < t.num_tiles; i++)
            uniform = th.tile_elem_ptr[i] == i * te && th.tile_elem_ptr[i + 1] - th.tile_elem_ptr[i] == te;
        if (uniform) t.uniform_te = te;
    }
    t.tile_threads = g_tile_threads;
    t.tile_persistent = g_tile_persistent;
    t.tile_grid = uint32_t(tile_grid_size(t, s->device, g_tile_persistent));
    CF_TRY(dalloc(&t.partials, 3 * (size_t)t.num_parts));
    CF_TRY(dalloc(&t.tile_part, std::max<size_t>(1, t.num_tiles)));
    CF_TRY(launch_precompute(t, st));
    s->perm.swap(th.perm);
    // total volume: FP32 serial sum in the caller's element order, like
    // EndEdition's accumulation (TetrahedronElementSet3.cpp:2118-2162)
    std::vector<float> vol(T);
    if (T) CF_CHECK_HIP(hipMemcpyAsync(vol.data(), t.vol, T * sizeof(float), hipMemcpyDeviceToHost, st));
    CF_CHECK_HIP(hipStreamSynchronize(st));
    std::vector<float> vol_user(T);
    for (uint32_t i = 0; i < T; i++) vol_user[s->perm[i]] = vol[i];
    float acc = 0.f;
    for (uint32_t i = 0; i < T; i++) acc += vol_user[i];
    s->total_volume = acc;
    s->mat = material_from(s->params.young_modulus, s->params.poisson_ratio, s->params.mm);
    s->ref_host.clear(); s->ref_host.shrink_to_fit();
    s->tets_host.clear(); s->tets_host.shrink_to_fit();
    s->editing = false;
    return CF_OK;
}

// The BeginEvaluation block statistics: copies into pinned memory enqueued on
// the set's stream, and their sums once the stream has passed them.
static int cache_stats_enqueue(cf_cache* c) {
    hipStream_t st = c->set->stream;
    const size_t n = size_t(c->stat_n);
    CF_CHECK_HIP(hipMemcpyAsync(c->h_count, c->blk_count, 2 * n * 4, hipMemcpyDeviceToHost, st));
    CF_CHECK_HIP(hipMemcpyAsync(c->h_sum, c->blk_sum, n * 8, hipMemcpyDeviceToHost, st));
    CF_CHECK_HIP(hipMemcpyAsync(c->h_min, c->blk_min, n * 4, hipMemcpyDeviceToHost, st));
    return CF_OK;
}

static void cache_stats_reduce(cf_cache* c) {
    uint32_t a = 0, f = 0;
    double s = 0;
    float m = 3e38f;
    for (int i = 0; i < c->stat_n; i++) {
        a += c->h_count[i]; f += c->h_count[c->stat_n + i]; s += c->h_sum[i]; m = std::min(m, c->h_min[i]);
    }
    c->num_degenerate = a;
    c->num_fallback = f;
    c->sum_degenerate = float(s);
    c->min_det = m;
    c->stats_valid = true;
}

static int cache_stats_fetch(cf_cache* c) {
    if (c->stats_valid) return CF_OK;
    CF_TRY(cache_stats_enqueue(c));
    CF_CHECK_HIP(hipStreamSynchronize(c->set->stream));
    cache_stats_reduce(c);
    return CF_OK;
}

// Device CG loop of ILSS_CG::Solve for the leaf operator A = fM M - fK K.
// The loop runs on the device; the host enqueues chunks of iterations and
// checks the device's done flag between chunks.
//
// One leaf's CG in resumable phases: cg_begin enqueues r = S(b - A y), d = r
// and the start test; cg_launch enqueues up to kCgChunk iterations (each
// skipped on the device once the solve has stopped), whose last d kernel
// writes the stop flag into pinned host memory; cg_poll waits for the stream
// and reads it; cg_finish reads the scalars. The multi-solid step interleaves
// several leaves' phases on their own streams; the single-leaf solve keeps
// one chunk queued behind the one it waits for.
constexpr uint32_t kCgChunk = 8;

struct CgRun {
    NodeTerms nt;
    uint32_t k = 0, max_it = 0;
    bool done = false;
};

// ay: the K y tile partials when the fused front computed them (else null:
// the tile pass runs here); rhs: the right-hand side gather's terms when it
// is still to run (after the fused front), fused here with the CG start.
static int cg_begin(cf_solid* s, float fM, float fK, CgRun* run, const float* ay = nullptr,
                    const NodeTerms* rhs = nullptr) {
    const TetDevice& t = s->set->dev;
    hipStream_t st = s->stream;
    CgBuffers& b = s->cg;
    run->max_it = std::min<uint32_t>(s->p.cg_max_iter, kMaxCgIter);
    run->k = 0;
    run->done = false;
    NodeTerms& nt = run->nt;
    nt = NodeTerms{};
    nt.knc_mask = s->num_knc ? s->knc_mask : nullptr;
    nt.mass = s->node_mass;
    nt.fM = fM; nt.fK = fK;
    // r = S(b - A y), d = r and the r.r partials in the gather of A y
    NodeTerms ni = nt;
    ni.d = s->y; ni.b = s->bvec; ni.r_out = b.r; ni.d_out = b.d; ni.partials = b.part_b;
    if (ay && rhs) {
        NodeTerms nr = *rhs;
        nr.fM = fM; nr.fK = fK; nr.r_out = b.r; nr.d_out = b.d; nr.partials = b.part_b;
        CF_TRY(launch_rhs_start(t, ay, nr, st));
    } else if (ay) {
        TetDevice ta = t;
        ta.partials = const_cast<float*>(ay);
        CF_TRY(launch_gather(ta, kGatherCgInit, ni, 0, st));
    } else {
        CF_TRY(launch_forces(t, kModeDfx, s->y, s->cache->R, cache_mat(s->cache), st));
        CF_TRY(launch_gather(t, kGatherCgInit, ni, 0, st));
    }
    CF_TRY(launch_cg_start(b.st, b.part_b, matvec_chunks(s->N), s->p.cg_rel_epsilon, s->p.cg_abs_epsilon,
                           run->max_it, st));
    nt.partials = b.part_a;
    nt.d = b.d; nt.out = b.q;
    return CF_OK;
}

static int cg_enqueue(cf_solid* s, const CgRun& run, uint32_t k0, uint32_t kend) {
    const TetDevice& t = s->set->dev;
    hipStream_t st = s->stream;
    CgBuffers& b = s->cg;
    const uint32_t n = 3 * s->N;
    const int nnode = node_blocks(s->N);
    for (uint32_t k = k0; k < kend; k++) {
        const int32_t* skip = &b.st->done[k];
        CF_TRY(launch_forces_skip(t, kModeDfx, b.d, s->cache->R, cache_mat(s->cache), skip, st,
                                  g_tile_alternate && (k & 1)));
        CF_TRY(launch_gather_skip(t, kGatherMatvec, run.nt, nnode, skip, st));
        CF_TRY(launch_cg_step_xr(n, k, b.st, b.r, b.q, b.part_a, matvec_chunks(s->N), b.part_b, st));
        CF_TRY(launch_cg_step_d(n, k, b.st, s->y, b.d, b.r, run.nt.knc_mask, b.part_b, st,
                                k + 1 == kend ? b.d_done : nullptr));
    }
    return CF_OK;
}

// The recorded chunk starting at iteration k0 (recorded on first use).
static int cg_graph(cf_solid* s, const CgRun& run, uint32_t k0, uint32_t kend, hipGraphExec_t* out) {
    CgGraphs& g = s->graphs;
    const uint32_t wg = get_cg_workgroups();
    if (g.fM != run.nt.fM || g.fK != run.nt.fK || g.max_it != run.max_it || g.knc_mask != run.nt.knc_mask ||
        g.alternate != g_tile_alternate || g.cg_workgroups != wg) {
        g.clear();
        g.fM = run.nt.fM; g.fK = run.nt.fK; g.max_it = run.max_it; g.knc_mask = run.nt.knc_mask;
        g.alternate = g_tile_alternate; g.cg_workgroups = wg;
    }
    const size_t c = k0 / kCgChunk;
    if (g.chunk.size() <= c) g.chunk.resize(c + 1, nullptr);
    if (!g.chunk[c]) {
        CF_CHECK_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        const int rc = cg_enqueue(s, run, k0, kend);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(s->stream, &graph);
        if (rc != CF_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) { set_error("CG graph capture failed: %s", hipGetErrorString(e)); return CF_ERR_HIP; }
        e = hipGraphInstantiate(&g.chunk[c], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g.chunk[c] = nullptr;
            set_error("CG graph instantiate failed: %s", hipGetErrorString(e));
            return CF_ERR_HIP;
        }
    }
    *out = g.chunk[c];
    return CF_OK;
}

static int cg_launch(cf_solid* s, CgRun* run, bool graphs = false) {
    const uint32_t kend = std::min(run->max_it, run->k + kCgChunk);
    if (graphs && kend > run->k) {
        hipGraphExec_t g = nullptr;
        CF_TRY(cg_graph(s, *run, run->k, kend, &g));
        CF_CHECK_HIP(hipGraphLaunch(g, s->stream));
    } else {
        CF_TRY(cg_enqueue(s, *run, run->k, kend));
    }
    run->k = kend;
    return CF_OK;
}

// Reads the stop flag of the chunk ending at iteration kend once the device
// has finished that chunk (host_done[kend] was written by its last d kernel).
static bool cg_chunk_stopped(const cf_solid* s, const CgRun& run, uint32_t kend) {
    return kend >= run.max_it || reinterpret_cast<volatile const int32_t*>(s->cg.h_done)[kend] != 0;
}

static int cg_poll(cf_solid* s, CgRun* run) {
    CF_CHECK_HIP(hipStreamSynchronize(s->stream));
    if (cg_chunk_stopped(s, *run, run->k)) run->done = true;
    return CF_OK;
}

// The solve's result from its scalar record: the iteration that stopped, its
// reason and residuals. A NaN or Inf in r.r or d.q (a non-finite state or
// right-hand side, which the reference only asserts against,
// TetrahedronElementSet3.cpp:253-255) stops the tests at once; it is
// reported as CF_CG_NONFINITE.
void cf::cg_result_from(const CgScalars& hs, uint32_t max_it, cf_cg_result* res) {
    uint32_t j = 0;
    while (j <= max_it && !hs.done[j]) j++;
    if (j > max_it) j = max_it;
    int reason = hs.reason[j];
    uint32_t iters = j;
    double delta = hs.delta[j];
    if (reason == CF_CG_BREAKDOWN) { iters = j - 1; delta = hs.delta[j - 1]; }
    // r.r entering iterations 0..iters and d.q of iterations 0..j-1 were computed
    bool finite = std::isfinite(hs.delta0);
    for (uint32_t i = 0; i <= iters && finite; i++) finite = std::isfinite(hs.delta[i]);
    for (uint32_t i = 0; i < j && finite; i++) finite = std::isfinite(hs.dq[i]);
    res->iterations = int32_t(iters);
    res->reason = finite ? reason : CF_CG_NONFINITE;
    res->rel_residual = (hs.delta0 > 0) ? float(std::sqrt(delta / hs.delta0)) : 0.f;
    res->abs_residual = float(std::sqrt(delta));
}

static int cg_finish(cf_solid* s, const CgRun& run, cf_cg_result* res) {
    static thread_local CgScalars hs;
    CF_CHECK_HIP(hipMemcpyAsync(&hs, s->cg.st, sizeof(CgScalars), hipMemcpyDeviceToHost, s->stream));
    CF_CHECK_HIP(hipStreamSynchronize(s->stream));
    cg_result_from(hs, run.max_it, res);
    if (res->reason == CF_CG_NONFINITE) s->nonfinite = true;
    return CF_OK;
}

// The single-leaf solve: chunk c + 1 is queued before the host waits for
// chunk c, so the device does not idle while the host reads a stop flag and
// enqueues; a solve that stops inside chunk c leaves chunk c + 1 to run as
// skipped launches. Chunks are replayed as recorded graphs (one launch instead
// of 32) unless the solid's graphs are switched off. The host spins on the
// chunk's event: a blocking wait added 1-2 us per CG iteration on small meshes.
//
// defer: return as soon as every chunk that can run is queued, with the copy
// of the scalar record into pinned memory (s->h_cgs) queued behind them; the
// caller reads the result (cg_result_from) once the stream has passed it.
static int solid_cg(cf_solid* s, float fM, float fK, cf_cg_result* res, bool defer = false,
                    const float* ay = nullptr, const NodeTerms* rhs = nullptr) {
    CgRun run;
    CF_TRY(cg_begin(s, fM, fK, &run, ay, rhs));
    CgBuffers& b = s->cg;
    uint32_t kend[2] = {0, 0};
    CF_TRY(cg_launch(s, &run, s->use_graphs));
    kend[0] = run.k;
    CF_CHECK_HIP(hipEventRecord(b.ev_chunk[0], s->stream));
    for (uint32_t c = 0;; c++) {
        const uint32_t cur = c & 1, nxt = cur ^ 1;
        if (defer && run.k >= run.max_it) break;  // the last chunk is queued
        if (run.k < run.max_it) {
            CF_TRY(cg_launch(s, &run, s->use_graphs));
            kend[nxt] = run.k;
            CF_CHECK_HIP(hipEventRecord(b.ev_chunk[nxt], s->stream));
        }
        hipError_t q;
        while ((q = hipEventQuery(b.ev_chunk[cur])) == hipErrorNotReady) {
        }
        CF_CHECK_HIP(q);
        if (cg_chunk_stopped(s, run, kend[cur])) break;
    }
    run.done = true;
    if (defer) {
        CF_CHECK_HIP(hipMemcpyAsync(s->h_cgs, b.st, sizeof(CgScalars), hipMemcpyDeviceToHost, s->stream));
        return CF_OK;
    }
    return cg_finish(s, run, res);
}

// ================================================================ FIE-dx
// Sum of nblk FP64 block partials in block order, read back to the host.
static int host_sum(cf_solid* s, const double* d_part, int nblk, double* out) {
    CF_CHECK_HIP(hipMemcpyAsync(s->h_part, d_part, size_t(nblk) * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    CF_CHECK_HIP(hipStreamSynchronize(s->stream));
    double acc = 0.0;
    for (int i = 0; i < nblk; i++) acc += s->h_part[i];
    *out = acc;
    return CF_OK;
}

// Right-hand side of one Newton iteration at the iterate x_k (f_e evaluated at
// x_k under the current rotations), warm start on the first iteration only,
// KNC rows in increment form; |b| when bnorm is not null.
static int fie_rhs(cf_solid* s, float dt, float inv_dt, bool first, float* bnorm) {
    cf_tetset* set = s->set;
    const TetDevice& t = set->dev;
    hipStream_t st = s->stream;
    CF_TRY(launch_forces(t, kModeFe, s->xk, s->cache->R, cache_mat(s->cache), st));
    NodeTerms nt;
    nt.out = s->bvec; nt.y = s->y; nt.x = s->x; nt.xk = s->xk; nt.v = s->v;
    nt.knc_mask = s->num_knc ? s->knc_mask : nullptr;
    nt.knc_target = s->knc_target;
    nt.mass = s->node_mass;
    nt.dt = dt; nt.inv_dt = inv_dt;
    nt.g[0] = s->p.gravity[0]; nt.g[1] = s->p.gravity[1]; nt.g[2] = s->p.gravity[2];
    nt.warm_start = (first && s->p.warm_start) ? 1 : 0;
    nt.partials = s->cg.part_b;
    const int nb = node_blocks(s->N);
    CF_TRY(launch_gather(t, kGatherRhsDx, nt, nb, st));
    if (bnorm) {
        double bb = 0.0;
        CF_TRY(host_sum(s, s->cg.part_b, nb, &bb));
        *bnorm = float(std::sqrt(bb));
        if (!std::isfinite(bb)) s->nonfinite = true;
    }
    return CF_OK;
}

// Solve, x_k += y, v_k = (x_k - x)/dt, re-open the evaluation at x_k
// (:1479-1530, NR_MidStep/NR_PostStep :2967-3037); |y| when ynorm is not null.
static int fie_solve_update(cf_solid* s, float inv_dt, float* ynorm) {
    const float fM = inv_dt * inv_dt + inv_dt * s->rayleigh[0];
    const float fK = 1.0f + inv_dt * s->rayleigh[1];
    CF_TRY(solid_cg(s, fM, fK, &s->last_ls));
    s->ls_total += s->last_ls.iterations;
    const uint32_t n3 = 3 * s->N;
    const int nb = node_blocks(n3);
    CF_TRY(launch_nr_update(n3, s->xk, s->vk, s->x, s->y, inv_dt, s->cg.part_b, nb, s->stream));
    if (ynorm) {
        double yy = 0.0;
        CF_TRY(host_sum(s, s->cg.part_b, nb, &yy));
        *ynorm = float(std::sqrt(yy));
        if (!std::isfinite(yy)) s->nonfinite = true;
    }
    return cf_begin_evaluation(s->cache, s->x, s->xk, s->vk, s->total_time);
}

// Step_FullyImplicitEuler_dx (ds/dsh/LeafDSH_Solid3D_FEM.cpp:1357-1552); dt = 0
// is the quasi-static variant (:957-959). The evaluation is open at x on entry.
static int step_fie_dx(cf_solid* s, float dt, int32_t* nr_iters, float* nr_rel, float* nr_abs) {
    const float inv_dt = (dt == 0.0f) ? 0.0f : 1.0f / dt;
    const size_t bytes = 3 * (size_t)s->N * sizeof(float);
    CF_CHECK_HIP(hipMemcpyAsync(s->xk, s->x, bytes, hipMemcpyDeviceToDevice, s->stream));
    CF_CHECK_HIP(hipMemcpyAsync(s->vk, s->v, bytes, hipMemcpyDeviceToDevice, s->stream));
    uint32_t nr_k = 0;
    float delta_new = 0.0f, delta0 = -1.0f;
    bool conv = false;
    s->ls_total = 0;
    s->last_ls = cf_cg_result{};
    do {
        CF_TRY(fie_rhs(s, dt, inv_dt, nr_k == 0, &delta_new));  // |b| after KNC (:1468-1474)
        if (delta0 < 0) delta0 = delta_new;
        conv = delta_new < s->p.nr_abs_epsilon;
        if (!conv) {
            CF_TRY(fie_solve_update(s, inv_dt, &delta_new));
            nr_k++;
        }
    } while (!conv && nr_k < s->p.nr_max_iter);
    *nr_iters = int32_t(nr_k);
    *nr_rel = (delta0 > 0) ? delta_new / delta0 : 0.0f;
    *nr_abs = delta_new;
    return CF_OK;
}

static int solid_fill_stats(cf_solid* s, double energy, float ms, cf_step_stats* stats) {
    CF_TRY(cache_stats_fetch(s->cache));
    stats->num_fixed_steps = 1;
    stats->cg_iters = s->ls_total;
    stats->cg_rel_residual = s->last_ls.rel_residual;
    stats->cg_abs_residual = s->last_ls.abs_residual;
    stats->cg_reason = s->last_ls.reason;
    stats->num_degenerate = s->cache->num_degenerate;
    stats->sum_degenerate_det_f = s->cache->sum_degenerate;
    stats->elastic_energy = float(energy);
    stats->min_det_f = s->cache->min_det;
    stats->step_ms = ms;
    return CF_OK;
}

// ================================================================ C API
extern "C" {

const char* cf_last_error(void) { return g_err; }
int cf_version(void) { return 10000; }

void cf_default_fem_params(cf_fem_params* p) {
    if (!p) return;
    // ms::fem::Params() (ms/fem/Params.h:74-98)
    p->young_modulus = 1000.0f;
    p->poisson_ratio = 0.25f;
    p->damping_ratio = 1.0f;
    p->rayleigh_freq1 = 0.1f;
    p->rayleigh_freq2 = 10000.0f;
    p->degenerate_threshold_det_f = 0.1f;
    p->mm = CF_MM_C_WRP;
    p->ecie_k_factor = 1.0f;
    p->rm = CF_RM_MSVD;
    p->dm = CF_DM_TRUNCATED;
    p->dapd_l_factor = 1.0f;   // Params.h:88
    p->dapd_nl_factor = 0.0f;
    p->dapd_nl_exponent = 2.0f;
}

void cf_rayleigh_coeffs(float f1, float f2, float zeta, float c[2]) {
    if (c) rayleigh_coeffs(f1, f2, zeta, c);
}

void cf_default_solid_params(cf_solid_params* p) {
    if (!p) return;
    // LeafDSH_Solid3D_FEM::Params() (ds/dsh/LeafDSH_Solid3D_FEM.h:203-274)
    cf_default_fem_params(&p->fem);
    p->fixed_dt = 0.001f;
    p->time_scale = 1.0f;
    p->density = 1000.0f;
    p->gravity[0] = p->gravity[1] = p->gravity[2] = 0.0f;
    p->air_drag_coeff = 0.0f;
    p->cg_rel_epsilon = 1e-3f;
    p->cg_abs_epsilon = 1e-6f;
    p->cg_max_iter = 20;
    p->warm_start = 1;
    p->compute_energy = 1;
    p->thickness = 1.0f;
    p->integrator = CF_IT_QIE_V;
    p->nr_max_iter = 3;
    p->nr_abs_epsilon = 1e-6f;
    p->reference_order = 0;
}

int cf_device_count(int* n) {
    if (!n) return CF_ERR_INVALID_ARG;
    hipError_t e = hipGetDeviceCount(n);
    if (e != hipSuccess) { *n = 0; set_error("hipGetDeviceCount: %s", hipGetErrorString(e)); return CF_ERR_NO_DEVICE; }
    return CF_OK;
}
int cf_set_device(int d) { CF_CHECK_HIP(hipSetDevice(d)); return CF_OK; }
int cf_device_alloc(size_t bytes, void** p) {
    if (!p) return CF_ERR_INVALID_ARG;
    return dalloc(reinterpret_cast<uint8_t**>(p), bytes);
}
int cf_device_free(void* p) { if (p) CF_CHECK_HIP(hipFree(p)); return CF_OK; }
// The library's handles run on non-blocking streams, which the legacy default
// stream does not wait for: these convenience copies synchronize the device
// first so that they observe every enqueued operation.
int cf_memcpy_htod(void* d, const void* h, size_t n) {
    CF_CHECK_HIP(hipDeviceSynchronize());
    CF_CHECK_HIP(hipMemcpy(d, h, n, hipMemcpyHostToDevice));
    return CF_OK;
}
int cf_memcpy_dtoh(void* h, const void* d, size_t n) {
    CF_CHECK_HIP(hipDeviceSynchronize());
    CF_CHECK_HIP(hipMemcpy(h, d, n, hipMemcpyDeviceToHost));
    return CF_OK;
}
int cf_memcpy_dtod(void* d, const void* s, size_t n) {
    CF_CHECK_HIP(hipDeviceSynchronize());
    CF_CHECK_HIP(hipMemcpy(d, s, n, hipMemcpyDeviceToDevice));
    CF_CHECK_HIP(hipDeviceSynchronize());
    return CF_OK;
}
int cf_memset_device(void* d, int v, size_t n) {
    CF_CHECK_HIP(hipDeviceSynchronize());
    CF_CHECK_HIP(hipMemset(d, v, n));
    CF_CHECK_HIP(hipDeviceSynchronize());
    return CF_OK;
}
int cf_device_synchronize(void) { CF_CHECK_HIP(hipDeviceSynchronize()); return CF_OK; }

// ---------------------------------------------------------------- meshes
int cf_mesh_box_counts(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t* nn, uint32_t* nt) {
    if (nx < 2 || ny < 2 || nz < 2 || !nn || !nt) { set_error("box needs >= 2 vertices per axis"); return CF_ERR_INVALID_ARG; }
    const uint64_t n = uint64_t(nx) * ny * nz, t = 5ull * (nx - 1) * (ny - 1) * (nz - 1);
    if (n > 0xffffffffull || 4 * t > 0xffffffffull) { set_error("box too large for uint32 indices"); return CF_ERR_CAPACITY; }
    *nn = uint32_t(n);
    *nt = uint32_t(t);
    return CF_OK;
}

// Make_TetSolidShape3_Box (Geo/shape/TetSolidShape3.cpp:921-999)
int cf_mesh_box(const float half[3], uint32_t nx, uint32_t ny, uint32_t nz, float* pos, uint32_t* tets) {
    uint32_t nn, nt;
    CF_TRY(cf_mesh_box_counts(nx, ny, nz, &nn, &nt));
    if (!half || !pos || !tets) return CF_ERR_INVALID_ARG;
    const uint32_t cnt[3] = {nx, ny, nz};
    float size[3];
    for (int c = 0; c < 3; c++) size[c] = 2 * half[c];
    size_t v = 0;
    for (uint32_t i = 0; i < nx; i++)
        for (uint32_t j = 0; j < ny; j++)
            for (uint32_t k = 0; k < nz; k++, v++) {
                const uint32_t ijk[3] = {i, j, k};
                for (int c = 0; c < 3; c++) pos[3 * v + c] = (-half[c]) + (size[c] * float(ijk[c])) / float(cnt[c] - 1);
            }
    size_t t = 0;
    auto emit = [&](uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
        tets[4 * t] = a; tets[4 * t + 1] = b; tets[4 * t + 2] = c; tets[4 * t + 3] = d; t++;
    };
    const uint32_t sx = ny * nz, sy = nz;
    for (uint32_t i = 0; i + 1 < nx; i++)
        for (uint32_t j = 0; j + 1 < ny; j++)
            for (uint32_t k = 0; k + 1 < nz; k++) {
                // cube corners: bit0 = +x, bit1 = +y, bit2 = +z, mirrored on odd i/j/k
                uint32_t c[8];
                for (uint32_t b = 0; b < 8; b++) {
                    uint32_t di = (b & 1) ^ (i & 1), dj = ((b >> 1) & 1) ^ (j & 1), dk = ((b >> 2) & 1) ^ (k & 1);
                    c[b] = (i + di) * sx + (j + dj) * sy + (k + dk);
                }
                // reference labels: v0=---, v1=-+-, v2=++-, v3=+--, v4=--+, v5=-++, v6=+++, v7=+-+
                const uint32_t v0 = c[0], v1 = c[2], v2 = c[3], v3 = c[1], v4 = c[4], v5 = c[6], v6 = c[7], v7 = c[5];
                const uint32_t mirrors = (i & 1) + (j & 1) + (k & 1);
                if (mirrors % 2 == 0) {
                    emit(v0, v1, v4, v3); emit(v2, v1, v3, v6); emit(v5, v1, v6, v4); emit(v6, v1, v3, v4); emit(v7, v3, v4, v6);
                } else {
                    emit(v0, v1, v3, v4); emit(v2, v1, v6, v3); emit(v5, v1, v4, v6); emit(v6, v1, v4, v3); emit(v7, v3, v6, v4);
                }
            }
    return CF_OK;
}

int cf_mesh_box2_counts(uint32_t nx, uint32_t ny, uint32_t* nn, uint32_t* nt) {
    if (nx < 2 || ny < 2 || !nn || !nt) { set_error("grid needs >= 2 vertices per axis"); return CF_ERR_INVALID_ARG; }
    *nn = nx * ny;
    *nt = 2 * (nx - 1) * (ny - 1);
    return CF_OK;
}

// Make_MeshSolidShape2_Box (Geo/shape/MeshSolidShape2.cpp:1651-1675)
int cf_mesh_box2(const float half[2], uint32_t nx, uint32_t ny, float* pos, uint32_t* tris) {
    uint32_t nn, nt;
    CF_TRY(cf_mesh_box2_counts(nx, ny, &nn, &nt));
    if (!half || !pos || !tris) return CF_ERR_INVALID_ARG;
    const float step_x = 2 * half[0] / float(nx - 1), step_y = 2 * half[1] / float(ny - 1);
    size_t v = 0;
    for (uint32_t iy = 0; iy < ny; iy++)
        for (uint32_t ix = 0; ix < nx; ix++, v++) {
            pos[2 * v] = (-half[0]) + float(ix) * step_x;
            pos[2 * v + 1] = (-half[1]) + float(iy) * step_y;
        }
    size_t t = 0;
    for (uint32_t iy = 0; iy + 1 < ny; iy++)
        for (uint32_t ix = 0; ix + 1 < nx; ix++) {
            const uint32_t a = iy * nx + ix, b = a + 1, c = (iy + 1) * nx + ix + 1, d = (iy + 1) * nx + ix;
            tris[3 * t] = a; tris[3 * t + 1] = b; tris[3 * t + 2] = c; t++;
            tris[3 * t] = c; tris[3 * t + 1] = d; tris[3 * t + 2] = a; t++;
        }
    return CF_OK;
}

// ---------------------------------------------------------------- element set
int cf_tetset_begin_edition(uint32_t n_nodes, const float* ref, uint32_t n_tets, const cf_fem_params* params,
                            cf_tetset** out) {
    if (!out || !ref || n_nodes == 0 || !valid_params(params)) {
        if (!out || !ref || n_nodes == 0) set_error("cf_tetset_begin_edition: invalid arguments");
        return CF_ERR_INVALID_ARG;
    }
    *out = nullptr;
    cf_tetset* s = new (std::nothrow) cf_tetset;
    if (!s) return CF_ERR_OUT_OF_MEMORY;
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_error("no HIP device: %s", hipGetErrorString(e));
        delete s;
        return CF_ERR_NO_DEVICE;
    }
    s->params = *params;
    s->editing = true;
    s->expected_T = n_tets;
    s->ref_host.assign(ref, ref + 3 * (size_t)n_nodes);
    s->tets_host.reserve(4 * (size_t)n_tets);
    *out = s;
    return CF_OK;
}

int cf_tetset_add_element(cf_tetset* s, uint32_t i, uint32_t j, uint32_t k, uint32_t l) {
    if (!s || !s->editing) { set_error("set is not being edited"); return CF_ERR_INVALID_ARG; }
    if (s->tets_host.size() / 4 >= s->expected_T) { set_error("more elements than announced"); return CF_ERR_CAPACITY; }
    s->tets_host.push_back(i); s->tets_host.push_back(j); s->tets_host.push_back(k); s->tets_host.push_back(l);
    return CF_OK;
}

int cf_tetset_end_edition(cf_tetset* s) {
    if (!s || !s->editing) { set_error("set is not being edited"); return CF_ERR_INVALID_ARG; }
    if (s->tets_host.size() / 4 != s->expected_T) {
        set_error("EndEdition with %zu of %u elements", s->tets_host.size() / 4, s->expected_T);
        return CF_ERR_INVALID_ARG;
    }
    return tetset_bake(s);
}

int cf_tetset_create(uint32_t n_nodes, const float* ref, uint32_t n_tets, const uint32_t* tets,
                     const cf_fem_params* params, cf_tetset** out) {
    if (!tets && n_tets) { set_error("null tets"); return CF_ERR_INVALID_ARG; }
    CF_TRY(cf_tetset_begin_edition(n_nodes, ref, n_tets, params, out));
    (*out)->tets_host.assign(tets, tets + 4 * (size_t)n_tets);
    int rc = tetset_bake(*out);
    if (rc != CF_OK) { cf_tetset_destroy(*out); *out = nullptr; }
    return rc;
}

int cf_tetset_set_params(cf_tetset* s, const cf_fem_params* params) {
    if (!s || !valid_params(params)) return CF_ERR_INVALID_ARG;
    s->params = *params;
    s->mat = material_from(params->young_modulus, params->poisson_ratio, params->mm);
    return CF_OK;
}

int cf_tetset_destroy(cf_tetset* s) {
    if (!s) return CF_OK;
    tetset_free(s);
    delete s;
    return CF_OK;
}

int cf_tet_compute_k(const float* r, float E, float nu, float* K, float* volume) {
    if (!r || !K) { set_error("cf_tet_compute_k: null argument"); return CF_ERR_INVALID_ARG; }
    const float v = exact_element_stiffness(r, r + 3, r + 6, r + 9, E, nu, K);
    if (volume) *volume = v;
    return CF_OK;
}

int cf_tetset_info(const cf_tetset* s, uint32_t* nn, uint32_t* nt, float* vol) {
    if (!s) return CF_ERR_INVALID_ARG;
    if (nn) *nn = s->dev.N;
    if (nt) *nt = s->dev.T;
    if (vol) *vol = s->total_volume;
    return CF_OK;
}

int cf_tetset_get_elements(const cf_tetset* s, float* h_inv_dm, float* h_vol) {
    if (!s || s->editing) return CF_ERR_INVALID_ARG;
    const uint32_t T = s->dev.T;
    std::vector<float> planes(9 * (size_t)T), vol(T);
    CF_CHECK_HIP(hipMemcpyAsync(planes.data(), s->dev.dinv, planes.size() * 4, hipMemcpyDeviceToHost, s->stream));
    CF_CHECK_HIP(hipMemcpyAsync(vol.data(), s->dev.vol, vol.size() * 4, hipMemcpyDeviceToHost, s->stream));
    CF_CHECK_HIP(hipStreamSynchronize(s->stream));
    for (uint32_t i = 0; i < T; i++) {
        const uint32_t e = s->perm[i];
        if (h_inv_dm)
            for (int k = 0; k < 9; k++) h_inv_dm[9 * (size_t)e + k] = planes[k * (size_t)T + i];
        if (h_vol) h_vol[e] = vol[i];
    }
    return CF_OK;
}

// ---------------------------------------------------------------- cache
int cf_cache_create(cf_tetset* set, const float* inv_mass, cf_cache** out) {
    if (!set || set->editing || !out) { set_error("cf_cache_create: invalid set"); return CF_ERR_INVALID_ARG; }
    cf_cache* c = new (std::nothrow) cf_cache;
    if (!c) return CF_ERR_OUT_OF_MEMORY;
    c->set = set;
    const uint32_t T = set->dev.T;
    c->nblk = node_blocks(T);
    c->stat_n = c->nblk;
    // statistics entries: per BeginEvaluation block, or per tile (fused front)
    const size_t ns = std::max<size_t>(kMaxPartials, set->dev.num_tiles);
    const size_t nh = std::max<size_t>(c->nblk, set->dev.num_tiles);
    int rc = dalloc(&c->R, 9 * (size_t)T);
    if (rc == CF_OK) rc = dalloc(&c->detF, T);
    if (rc == CF_OK) rc = dalloc(&c->blk_count, 2 * ns);
    if (rc == CF_OK) rc = dalloc(&c->blk_sum, ns);
    if (rc == CF_OK) rc = dalloc(&c->blk_min, ns);
    if (rc == CF_OK) rc = dalloc(&c->blk_energy, kMaxPartials + 1);  // block partials, then their total
    if (rc == CF_OK) rc = dalloc(&c->doc, std::max<size_t>(1, T));
    if (rc == CF_OK && (hipHostMalloc(reinterpret_cast<void**>(&c->h_count), 2 * nh * 4) != hipSuccess ||
                        hipHostMalloc(reinterpret_cast<void**>(&c->h_sum), nh * 8) != hipSuccess ||
                        hipHostMalloc(reinterpret_cast<void**>(&c->h_min), nh * 4) != hipSuccess)) {
        set_error("pinned allocation failed");
        rc = CF_ERR_HIP;
    }
    if (rc == CF_OK && inv_mass) {
        rc = dalloc(&c->inv_mass, set->dev.N);
        if (rc == CF_OK && hipMemcpy(c->inv_mass, inv_mass, set->dev.N * 4, hipMemcpyHostToDevice) != hipSuccess)
            rc = CF_ERR_HIP;
    }
    if (rc == CF_OK) {
        // CreateCache: R = Id, det F = 1 (TetrahedronElementSet3.cpp:190-209)
        std::vector<float> planes(9 * (size_t)T, 0.f), ones(T, 1.f);
        for (size_t e = 0; e < T; e++) { planes[e] = 1.f; planes[4 * (size_t)T + e] = 1.f; planes[8 * (size_t)T + e] = 1.f; }
        // and every DoC invalid (TetrahedronElement3.h:30)
        if (hipMemcpy(c->R, planes.data(), planes.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->detF, ones.data(), ones.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(c->doc, kDocInvalid, std::max<size_t>(1, T)) != hipSuccess)
            rc = CF_ERR_HIP;
    }
    if (rc != CF_OK) { cf_cache_destroy(c); return rc; }
    *out = c;
    return CF_OK;
}

int cf_cache_destroy(cf_cache* c) {
    if (!c) return CF_OK;
    dfree(c->R); dfree(c->detF); dfree(c->inv_mass); dfree(c->blk_count); dfree(c->blk_sum);
    dfree(c->blk_min); dfree(c->blk_energy); dfree(c->F); dfree(c->w); dfree(c->doc);
    if (c->h_count) (void)hipHostFree(c->h_count);
    if (c->h_sum) (void)hipHostFree(c->h_sum);
    if (c->h_min) (void)hipHostFree(c->h_min);
    delete c;
    return CF_OK;
}

int cf_begin_evaluation(cf_cache* c, const float* prev_x, const float* x, const float* v, float t) {
    if (!c || !x) { set_error("cf_begin_evaluation: null cache or x"); return CF_ERR_INVALID_ARG; }
    c->prev_x = prev_x; c->x = x; c->v = v;
    c->timestep = t - c->time;
    c->time = t;
    const cf_tetset* s = c->set;
    const bool lcm = s->params.mm != CF_MM_C_WRP;
    const bool keep_F = lcm || s->params.dm == CF_DM_EXACT;
    if (keep_F && !c->F) CF_TRY(dalloc(&c->F, 9 * (size_t)s->dev.T));
    if (lcm && !c->w) CF_TRY(dalloc(&c->w, (size_t)s->dev.T));
    LcmOut lo;
    if (keep_F) lo.F = c->F;
    if (lcm) {
        lo.w = (s->params.mm == CF_MM_C_LCMH) ? c->w : nullptr;
        lo.inv_k = 1.0f / s->params.ecie_k_factor;  // mal::Rcp (:1898)
        lo.hybrid = s->params.ecie_k_factor > 0.1f;   // LCMH blend branch (:1893-1897)
    }
    // UpdateDegeneration runs under every rotation method (:261); its DoC
    // only enters the DAPD rotations
    DegenArgs dg;
    dg.prev_x = prev_x ? prev_x : x;
    dg.doc = c->doc;
    dg.vol = s->dev.vol;
    dg.l_factor = s->params.dapd_l_factor;
    dg.nl_factor = s->params.dapd_nl_factor;
    dg.nl_exponent = s->params.dapd_nl_exponent;
    CF_TRY(launch_begin_eval(s->dev, x, s->params.rm, s->params.degenerate_threshold_det_f, c->R, c->detF,
                             c->blk_count, c->blk_sum, c->blk_min, c->nblk, s->stream, lo, dg));
    c->stat_n = c->nblk;
    c->stats_valid = false;
    return CF_OK;
}

// DoC bytes in the caller's element order (the device keeps them in tile order;
// vertex order within an element is the caller's).
int cf_cache_get_doc(cf_cache* c, uint8_t* h_doc) {
    if (!c || !h_doc) return CF_ERR_INVALID_ARG;
    const cf_tetset* s = c->set;
    std::vector<uint8_t> dev(s->dev.T);
    CF_CHECK_HIP(hipStreamSynchronize(s->stream));
    CF_CHECK_HIP(hipMemcpy(dev.data(), c->doc, s->dev.T, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < s->dev.T; i++) h_doc[s->perm[i]] = dev[i];
    return CF_OK;
}

int cf_cache_set_doc(cf_cache* c, const uint8_t* h_doc) {
    if (!c || !h_doc) return CF_ERR_INVALID_ARG;
    const cf_tetset* s = c->set;
    for (uint32_t e = 0; e < s->dev.T; e++)
        if ((h_doc[e] >> 2) > 3 || (h_doc[e] != kDocInvalid && (h_doc[e] >> 2) > (h_doc[e] & 3))) {
            set_error("cf_cache_set_doc: element %u has DoC byte %u (4*vit0 + vit1 with vit0 <= vit1, or 12)", e,
                      unsigned(h_doc[e]));
            return CF_ERR_INVALID_ARG;
        }
    std::vector<uint8_t> dev(s->dev.T);
    for (uint32_t i = 0; i < s->dev.T; i++) dev[i] = h_doc[s->perm[i]];
    CF_CHECK_HIP(hipStreamSynchronize(s->stream));
    CF_CHECK_HIP(hipMemcpy(c->doc, dev.data(), s->dev.T, hipMemcpyHostToDevice));
    return CF_OK;
}

int cf_f_e(cf_cache* c, float* f) {
    if (!c || !c->x || !f) { set_error("cf_f_e outside Begin/EndEvaluation or null f"); return CF_ERR_INVALID_ARG; }
    const cf_tetset* s = c->set;
    CF_TRY(launch_forces(s->dev, kModeFe, c->x, c->R, cache_mat(c), s->stream));
    NodeTerms nt;
    nt.out = f;
    return launch_gather(s->dev, kGatherAccumulate, nt, node_blocks(s->dev.N), s->stream);
}

int cf_f(cf_cache* c, float* f) { return cf_f_e(c, f); }

// C_f_d (TetrahedronElementSet3.cpp:1098-1150) with the set's Rayleigh coefficients
int cf_f_d(cf_cache* c, float* f) {
    if (!c || !f || !c->x) { set_error("cf_f_d outside Begin/EndEvaluation"); return CF_ERR_INVALID_ARG; }
    if (!c->v || !c->inv_mass) {
        set_error("cf_f_d needs the evaluation's velocities and the cache's inverse masses");
        return CF_ERR_INVALID_ARG;
    }
    const cf_tetset* s = c->set;
    DampArgs d;
    float coef[2];
    rayleigh_coeffs(s->params.rayleigh_freq1, s->params.rayleigh_freq2, s->params.damping_ratio, coef);
    d.inv_mass = c->inv_mass;
    d.c0 = coef[0];
    d.c1 = coef[1];
    Material m = s->mat;  // the stored-Ke (C_WRP) stiffness for every corotational material
    m.mm = CF_MM_C_WRP;
    m.F = m.w = nullptr;
    CF_TRY(launch_damping(s->dev, c->v, c->R, m, d, s->stream));
    NodeTerms nt;
    nt.out = f;
    return launch_gather(s->dev, kGatherAccumulate, nt, node_blocks(s->dev.N), s->stream);
}  // f = f_e + f_p, f_p a no-op (:294-300,350-353)

int cf_df_x(cf_cache* c, const float* dx, float* df) {
    if (!c || !dx || !df) { set_error("cf_df_x: null argument"); return CF_ERR_INVALID_ARG; }
    const cf_tetset* s = c->set;
    CF_TRY(launch_forces(s->dev, kModeDfx, dx, c->R, cache_mat(c), s->stream));
    NodeTerms nt;
    nt.out = df;
    return launch_gather(s->dev, kGatherAccumulate, nt, node_blocks(s->dev.N), s->stream);
}

// V() on the set's stream, its total copied to h_total (pinned or not) behind it.
static int cache_energy_enqueue(cf_cache* c, double* h_total) {
    const cf_tetset* s = c->set;
    double* total = c->blk_energy + kMaxPartials;
    CF_TRY(launch_energy(s->dev, c->x, c->R, cache_mat(c), c->blk_energy, c->nblk, total, s->stream));
    CF_CHECK_HIP(hipMemcpyAsync(h_total, total, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    return CF_OK;
}

int cf_V(cf_cache* c, double* energy) {
    if (!c || !c->x || !energy) { set_error("cf_V outside Begin/EndEvaluation"); return CF_ERR_INVALID_ARG; }
    double h = 0.0;
    CF_TRY(cache_energy_enqueue(c, &h));
    CF_CHECK_HIP(hipStreamSynchronize(c->set->stream));
    *energy = h;
    return CF_OK;
}

int cf_end_evaluation(cf_cache* c) {
    if (!c) return CF_ERR_INVALID_ARG;
    c->x = c->prev_x = c->v = nullptr;
    return CF_OK;
}

int cf_cache_stats(const cf_cache* c, uint32_t* nd, float* sd, float* md) {
    if (!c) return CF_ERR_INVALID_ARG;
    CF_TRY(cache_stats_fetch(const_cast<cf_cache*>(c)));
    if (nd) *nd = c->num_degenerate;
    if (sd) *sd = c->sum_degenerate;
    if (md) *md = c->min_det;
    return CF_OK;
}

int cf_cache_num_fallback(const cf_cache* c, uint32_t* n) {
    if (!c || !n) return CF_ERR_INVALID_ARG;
    CF_TRY(cache_stats_fetch(const_cast<cf_cache*>(c)));
    *n = c->num_fallback;
    return CF_OK;
}

int cf_cache_get_rotations(const cf_cache* c, float* h_R, float* h_det) {
    if (!c) return CF_ERR_INVALID_ARG;
    const cf_tetset* s = c->set;
    const uint32_t T = s->dev.T;
    std::vector<float> planes(9 * (size_t)T), det(T);
    CF_CHECK_HIP(hipMemcpyAsync(planes.data(), c->R, planes.size() * 4, hipMemcpyDeviceToHost, s->stream));
    CF_CHECK_HIP(hipMemcpyAsync(det.data(), c->detF, det.size() * 4, hipMemcpyDeviceToHost, s->stream));
    CF_CHECK_HIP(hipStreamSynchronize(s->stream));
    for (uint32_t i = 0; i < T; i++) {
        const uint32_t e = s->perm[i];
        if (h_R)
            for (int k = 0; k < 9; k++) h_R[9 * (size_t)e + k] = planes[k * (size_t)T + i];
        if (h_det) h_det[e] = det[i];
    }
    return CF_OK;
}

int cf_cache_synchronize(cf_cache* c) {
    if (!c) return CF_ERR_INVALID_ARG;
    CF_CHECK_HIP(hipStreamSynchronize(c->set->stream));
    return CF_OK;
}

// ---------------------------------------------------------------- generic CG
int cf_cg_create(uint32_t n, cf_cg** out) {
    if (!out || n == 0) return CF_ERR_INVALID_ARG;
    cf_cg* c = new (std::nothrow) cf_cg;
    if (!c) return CF_ERR_OUT_OF_MEMORY;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        set_error("hipStreamCreate failed");
        return CF_ERR_NO_DEVICE;
    }
    int rc = cg_alloc(c->b, n);
    if (rc != CF_OK) { cf_cg_destroy(c); return rc; }
    *out = c;
    return CF_OK;
}

int cf_cg_destroy(cf_cg* c) {
    if (!c) return CF_OK;
    cg_free(c->b);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CF_OK;
}

void* cf_cg_stream(cf_cg* c) { return c ? (void*)c->stream : nullptr; }

int cf_cg_solve(cf_cg* c, cf_av_fn av, cf_constraints_fn cons, void* user, float* x, const float* bvec, float rel,
                float abs_eps, uint32_t max_it, cf_cg_result* res) {
    if (!c || !av || !x || !bvec || !res || max_it == 0) { set_error("cf_cg_solve: invalid arguments"); return CF_ERR_INVALID_ARG; }
    const CgBuffers& b = c->b;
    const CgWork w{b.r, b.d, b.q, b.st, b.part_a, b.part_b};
    return cg_solve_on(w, b.n, c->stream, av, cons, user, x, bvec, rel, abs_eps, max_it, b.h_flag, res);
}

}  // extern "C"

// ILSS_CG::Solve (ns/ILSS_CG.h:38-151) with the caller's operator and
// constraint callbacks, on the given work buffers and stream: the body of
// cf_cg_solve (buffers of a cf_cg handle) and of cf_cg_solve_static (the
// library's static scratch). h_flag receives each iteration's stop flag.
int cf::cg_solve_on(const CgWork& b, uint32_t n, hipStream_t st, cf_av_fn av, cf_constraints_fn cons, void* user,
                    float* x, const float* bvec, float rel, float abs_eps, uint32_t max_it, int32_t* h_flag,
                    cf_cg_result* res) {
    max_it = std::min(max_it, kMaxCgIter);
    const int nvec = node_blocks(n);
    // r = S(b - A x); d = r; delta0 = r.r   (ILSS_CG.h:53-65)
    av(user, b.q, x);
    CF_TRY(launch_cg_init(n, b.q, bvec, b.r, b.d, nullptr, b.part_b, nvec, st));
    if (cons) {
        cons(user, b.r);
        CF_CHECK_HIP(hipMemcpyAsync(b.d, b.r, n * 4, hipMemcpyDeviceToDevice, st));
        CF_TRY(launch_cg_dot(n, b.r, b.r, b.part_b, nvec, st));
    }
    CF_TRY(launch_cg_start(b.st, b.part_b, nvec, rel, abs_eps, max_it, st));
    CF_CHECK_HIP(hipMemcpyAsync(h_flag, &b.st->done[0], 4, hipMemcpyDeviceToHost, st));
    CF_CHECK_HIP(hipStreamSynchronize(st));
    uint32_t k = 0;
    while (!*h_flag && k < max_it) {
        av(user, b.q, b.d);
        if (cons) cons(user, b.q);
        CF_TRY(launch_cg_dot(n, b.d, b.q, b.part_a, nvec, st));
        CF_TRY(launch_cg_step_xr(n, k, b.st, b.r, b.q, b.part_a, nvec, b.part_b, st));
        CF_TRY(launch_cg_step_d(n, k, b.st, x, b.d, b.r, nullptr, b.part_b, st));
        if (cons) cons(user, b.d);
        k++;
        CF_CHECK_HIP(hipMemcpyAsync(h_flag, &b.st->done[k], 4, hipMemcpyDeviceToHost, st));
        CF_CHECK_HIP(hipStreamSynchronize(st));
    }
    static thread_local CgScalars hs;
    CF_CHECK_HIP(hipMemcpyAsync(&hs, b.st, sizeof(CgScalars), hipMemcpyDeviceToHost, st));
    CF_CHECK_HIP(hipStreamSynchronize(st));
    cg_result_from(hs, max_it, res);
    if (res->reason == CF_CG_NONFINITE) {
        set_error("cf_cg_solve: r.r or d.q is not finite (NaN or Inf in x, b or A)");
        return CF_ERR_NAN;
    }
    return CF_OK;
}

extern "C" {

// ---------------------------------------------------------------- leaf solid
int cf_solid_destroy(cf_solid* s) {
    if (!s) return CF_OK;
    dfree(s->x); dfree(s->v); dfree(s->x0); dfree(s->bvec); dfree(s->y);
    dfree(s->knc_mask); dfree(s->knc_target); dfree(s->scratch_vec);
    dfree(s->xk); dfree(s->vk); dfree(s->ay_part);
    if (s->h_part) (void)hipHostFree(s->h_part);
    if (s->h_cgs) (void)hipHostFree(s->h_cgs);
    if (s->h_energy) (void)hipHostFree(s->h_energy);
    cg_free(s->cg);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    s->graphs.clear();
    exact_free(s->ex);
    cf_cache_destroy(s->cache);
    cf_tetset_destroy(s->set);
    delete s;
    return CF_OK;
}

int cf_solid_create(uint32_t N, const float* ref, uint32_t T, const uint32_t* tets, const cf_solid_params* p,
                    cf_solid** out) {
    if (!out || !ref || !tets || !p || N == 0 || T == 0) { set_error("cf_solid_create: invalid arguments"); return CF_ERR_INVALID_ARG; }
    if (!(p->fixed_dt > 0) || !(p->density > 0)) { set_error("fixed_dt and density must be > 0"); return CF_ERR_INVALID_ARG; }
    if (p->integrator != CF_IT_QIE_V && p->integrator != CF_IT_FIE_DX) {
        set_error("integrator %d is not supported (CF_IT_QIE_V or CF_IT_FIE_DX)", p->integrator);
        return CF_ERR_UNSUPPORTED;
    }
    if (p->reference_order &&
        (p->integrator != CF_IT_QIE_V || p->fem.mm != CF_MM_C_WRP || p->fem.dm != CF_DM_TRUNCATED)) {
        set_error("reference_order takes QIE-v, C_WRP and DM Truncated");
        return CF_ERR_UNSUPPORTED;
    }
    *out = nullptr;
    cf_solid* s = new (std::nothrow) cf_solid;
    if (!s) return CF_ERR_OUT_OF_MEMORY;
    s->p = *p;
    s->N = N;
    int rc = cf_tetset_create(N, ref, T, tets, &p->fem, &s->set);
    if (rc == CF_OK) rc = cf_cache_create(s->set, nullptr, &s->cache);
    if (rc != CF_OK) { cf_solid_destroy(s); return rc; }
    s->stream = s->set->stream;
    const size_t n3 = 3 * (size_t)N;
    if (rc == CF_OK) rc = dalloc(&s->x, n3);
    if (rc == CF_OK) rc = dalloc(&s->v, n3);
    if (rc == CF_OK) rc = dalloc(&s->x0, n3);
    if (rc == CF_OK) rc = dalloc(&s->bvec, n3);
    if (rc == CF_OK) rc = dalloc(&s->y, n3);
    if (rc == CF_OK) rc = dalloc(&s->knc_mask, N);
    if (rc == CF_OK) rc = dalloc(&s->knc_target, n3);
    if (rc == CF_OK) rc = cg_alloc(s->cg, uint32_t(n3));
    if (rc == CF_OK) rc = dalloc(&s->xk, n3);
    if (rc == CF_OK) rc = dalloc(&s->vk, n3);
    if (rc == CF_OK && (hipHostMalloc(reinterpret_cast<void**>(&s->h_part), kMaxPartials * sizeof(double)) != hipSuccess ||
                        hipHostMalloc(reinterpret_cast<void**>(&s->h_cgs), sizeof(CgScalars)) != hipSuccess ||
                        hipHostMalloc(reinterpret_cast<void**>(&s->h_energy), sizeof(double)) != hipSuccess)) {
        set_error("pinned allocation failed");
        rc = CF_ERR_HIP;
    }
    if (rc != CF_OK) { cf_solid_destroy(s); return rc; }
    if (hipMemcpy(s->x, ref, n3 * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->x0, ref, n3 * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(s->v, 0, n3 * 4) != hipSuccess || hipMemset(s->knc_mask, 0, N) != hipSuccess ||
        hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess) {
        set_error("device initialisation of the solid failed");
        cf_solid_destroy(s);
        return CF_ERR_HIP;
    }
    // RebuildMass: uniform node mass rho*V/N stored as an inverse (:1591-1611,
    // ms/ParticleSystem.h:74,80); GetMass returns 1/(1/m).
    const float total_mass = p->density * s->set->total_volume;
    const float mass_per_node = total_mass / float(N);
    const float inv_mass = 1.0f / mass_per_node;
    s->node_mass = 1.0f / inv_mass;
    rayleigh_coeffs(p->fem.rayleigh_freq1, p->fem.rayleigh_freq2, p->fem.damping_ratio, s->rayleigh);
    s->air_drag_per_second = -std::log(10e-6f) * p->air_drag_coeff;  // RebuildAirDrag (:1586-1589)
    if (p->reference_order) {
        rc = exact_build(N, ref, T, tets, s->set->perm.data(), p->fem.young_modulus, p->fem.poisson_ratio, s->stream,
                         s->ex);
        if (rc != CF_OK) { cf_solid_destroy(s); return rc; }
    }
    *out = s;
    return CF_OK;
}

int cf_solid_set_knc(cf_solid* s, uint32_t n, const uint32_t* nids, const float* targets) {
    if (!s || (n && !nids)) return CF_ERR_INVALID_ARG;
    std::vector<uint8_t> mask(s->N, 0);
    std::vector<float> tgt(3 * (size_t)s->N, 0.f);
    std::vector<float> ref;
    if (!targets) {
        ref.resize(3 * (size_t)s->N);
        CF_CHECK_HIP(hipMemcpy(ref.data(), s->set->dev.ref, ref.size() * 4, hipMemcpyDeviceToHost));
    }
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t nid = nids[i];
        if (nid >= s->N) { set_error("KNC node %u out of range", nid); return CF_ERR_INVALID_ARG; }
        mask[nid] = 1;
        for (int c = 0; c < 3; c++) tgt[3 * (size_t)nid + c] = targets ? targets[3 * (size_t)i + c] : ref[3 * (size_t)nid + c];
    }
    CF_CHECK_HIP(hipMemcpy(s->knc_mask, mask.data(), mask.size(), hipMemcpyHostToDevice));
    CF_CHECK_HIP(hipMemcpy(s->knc_target, tgt.data(), tgt.size() * 4, hipMemcpyHostToDevice));
    s->num_knc = n;
    return CF_OK;
}

// FixedStep with QIE-v in the reference's FP32 order (reference_order = 1):
// BeginEvaluation as the fast path (bit-exact F, det F and DAPD rotations), then
// b, y, the CG and the integration term by term as Step_QuasiImplicitEuler_v
// (:1209-1327), LS_Solve (:1081-1155) and ILSS_CG::Solve (ns/ILSS_CG.h:38-151)
// compute them, with float scalars on the host.
static int solid_fixed_step_exact(cf_solid* s, float dt, cf_step_stats* stats) {
#pragma clang fp contract(off)
    if (!(dt > 0)) return CF_ERR_INVALID_ARG;
    s->nonfinite = false;
    const TetDevice& t = s->set->dev;
    const ExactDevice& ex = s->ex;
    hipStream_t st = s->stream;
    const uint32_t n3 = 3 * s->N;
    CgBuffers& b = s->cg;
    CF_CHECK_HIP(hipEventRecord(s->ev0, st));
    CF_TRY(cf_begin_evaluation(s->cache, s->x0, s->x, s->v, s->total_time));
    CF_CHECK_HIP(hipMemcpyAsync(s->x0, s->x, n3 * 4, hipMemcpyDeviceToDevice, st));
    const float* R = s->cache->R;
    const uint8_t* knc = s->num_knc ? s->knc_mask : nullptr;
    // b = dt (M g + f_e) + M v, y = v, KNC rows (:1209-1295)
    CF_TRY(launch_exact_elem(ex, kExactFe, R, t.T, t.ref, s->x, nullptr, st));
    ExactNodeTerms rhs;
    rhs.knc = knc; rhs.knc_target = s->knc_target; rhs.x = s->x; rhs.v = s->v; rhs.out = s->bvec; rhs.y = s->y;
    rhs.mass = s->node_mass; rhs.dt = dt; rhs.inv_dt = 1.0f / dt; rhs.warm_start = s->p.warm_start;
    for (int c = 0; c < 3; c++) rhs.g[c] = s->p.gravity[c];
    CF_TRY(launch_exact_node(ex, kExactRhs, rhs, st));
    // A = (1 + dt c0) M + (dt^2 + dt c1) K (:1300-1301); LS_A_times_d scales K d
    // by -factor_K and adds (m factor_M) d (:1056-1079)
    const float fM = 1.0f + dt * s->rayleigh[0];
    const float fK = dt * dt + dt * s->rayleigh[1];
    ExactNodeTerms av;
    av.knc = knc; av.neg_fk = -fK; av.m_fm = s->node_mass * fM;
    const float rel = s->p.cg_rel_epsilon, abs_sq = s->p.cg_abs_epsilon * s->p.cg_abs_epsilon;
    // r = S(b - A y); d = r; delta0 = r.r (ILSS_CG.h:53-65)
    CF_TRY(launch_exact_elem(ex, kExactDfx, R, t.T, t.ref, s->y, nullptr, st));
    av.d = s->y; av.out = b.r;
    CF_TRY(launch_exact_node(ex, kExactAv, av, st));
    CF_TRY(launch_exact_vec(0, n3, 0.0f, knc, nullptr, b.r, b.d, s->bvec, st));
    float delta0 = 0.0f;
    CF_TRY(exact_dot(ex, 0, n3, b.r, b.r, st, &delta0));
    float dnew = delta0;
    const float target = (rel * rel) * delta0;
    bool finite = std::isfinite(delta0);
    uint32_t k = 0;
    int reason = 0;
    av.d = b.d; av.out = b.q;
    while (dnew > abs_sq && dnew > target && k < s->p.cg_max_iter) {
        CF_TRY(launch_exact_elem(ex, kExactDfx, R, t.T, t.ref, b.d, nullptr, st));
        CF_TRY(launch_exact_node(ex, kExactAv, av, st));
        float dq = 0.0f;
        CF_TRY(exact_dot(ex, 0, n3, b.d, b.q, st, &dq));
        finite = finite && std::isfinite(dq);
        if (std::fabs(dq) < abs_sq) { reason = CF_CG_BREAKDOWN; break; }  // ILSS_CG.h:97
        const float alpha = dnew / dq;
        CF_TRY(launch_exact_vec(1, n3, alpha, knc, s->y, b.r, b.d, b.q, st));
        const float dold = dnew;
        CF_TRY(exact_dot(ex, 0, n3, b.r, b.r, st, &dnew));
        finite = finite && std::isfinite(dnew);
        const float beta = dnew / dold;
        CF_TRY(launch_exact_vec(2, n3, beta, knc, nullptr, b.r, b.d, nullptr, st));
        k++;
    }
    if (!reason) reason = !(dnew > abs_sq) ? CF_CG_CONVERGED_ABS : !(dnew > target) ? CF_CG_CONVERGED_REL : CF_CG_MAX_ITER;
    if (!finite) { reason = CF_CG_NONFINITE; s->nonfinite = true; }
    // v = y, x = x + dt y (:1322-1327)
    CF_TRY(launch_exact_vec(3, n3, dt, nullptr, s->x, s->v, nullptr, s->y, st));
    // V() with the new x (:977; C_V :1456-1488): 0.5 * the element-ordered sum
    float energy = 0.0f;
    if (s->p.compute_energy) {
        CF_TRY(launch_exact_elem(ex, kExactV, R, t.T, t.ref, s->x, nullptr, st));
        CF_TRY(exact_dot(ex, 1, 4 * t.T, ex.w, nullptr, st, &energy));
    }
    CF_TRY(cf_end_evaluation(s->cache));
    CF_CHECK_HIP(hipEventRecord(s->ev1, st));
    CF_CHECK_HIP(hipEventSynchronize(s->ev1));
    s->last_ls.iterations = int32_t(k);
    s->last_ls.reason = reason;
    s->last_ls.rel_residual = (delta0 > 0) ? std::sqrt(dnew / delta0) : 0.0f;
    s->last_ls.abs_residual = std::sqrt(dnew);
    s->ls_total = int32_t(k);
    if (stats) {
        float ms = 0.f;
        CF_CHECK_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        CF_TRY(solid_fill_stats(s, energy, ms, stats));
        stats->nr_iters = 0;
        stats->nr_rel_residual = 0.f;
        stats->nr_abs_residual = 0.f;
    }
    if (!s->nonfinite) return CF_OK;
    set_error("fixed step: a CG reduction is NaN or Inf (non-finite state or forces)");
    return CF_ERR_NAN;
}

static int nonfinite_status(cf_solid* s) {
    if (!s->nonfinite) return CF_OK;
    set_error("fixed step: a CG or Newton reduction is NaN or Inf (non-finite state or forces)");
    return CF_ERR_NAN;
}

// FixedStep (:898-990). The total time is the caller's: Step advances it once
// per call (:229-230), a direct fixed step by its own dt.
// Whether a QIE-v step may open its evaluation with the fused front: C_WRP
// without the exact rotation differential (nothing else is kept per element).
static bool front_fusable(const cf_solid* s) {
    const cf_tetset* set = s->set;
    const TetDevice& t = set->dev;
    const uint64_t lim = uint64_t(1) << 32;  // its buffer accesses take 32-bit byte offsets
    return g_fused_front && s->p.integrator != CF_IT_FIE_DX && t.num_tiles > 0 && set->params.mm == CF_MM_C_WRP &&
           set->params.dm != CF_DM_EXACT && 36 * uint64_t(t.T) < lim && 12 * uint64_t(t.N) < lim &&
           12 * uint64_t(t.num_parts) < lim && 8 * uint64_t(t.num_tiles) < lim;
}

// BeginEvaluation(prev_x = x0, x, v, t) (FixedStep :927) fused with f_e(x)
// into the tile partials and the CG start's K y into s->ay_part: the same R,
// det F, DoC and partials as cf_begin_evaluation followed by the two tile
// passes; y is the right-hand side gather's (yt).
static int begin_front(cf_solid* s, const NodeTerms& yt) {
    cf_cache* c = s->cache;
    const cf_tetset* set = s->set;
    if (!s->ay_part) CF_TRY(dalloc(&s->ay_part, 3 * (size_t)set->dev.num_parts));
    c->prev_x = s->x0; c->x = s->x; c->v = s->v;
    c->timestep = s->total_time - c->time;
    c->time = s->total_time;
    FrontOut fo;
    fo.rm = set->params.rm;
    fo.thr = set->params.degenerate_threshold_det_f;
    fo.R = c->R;
    fo.detF = c->detF;
    fo.st_count = c->blk_count; fo.st_sum = c->blk_sum; fo.st_min = c->blk_min;
    fo.dg.prev_x = s->x0;
    fo.dg.doc = c->doc;
    fo.dg.vol = set->dev.vol;
    fo.dg.l_factor = set->params.dapd_l_factor;
    fo.dg.nl_factor = set->params.dapd_nl_factor;
    fo.dg.nl_exponent = set->params.dapd_nl_exponent;
    fo.partials2 = s->ay_part;
    CF_TRY(launch_qie_front(set->dev, s->x, cache_mat(c), fo, yt, s->stream));
    c->stat_n = int(set->dev.num_tiles);
    c->stats_valid = false;
    return CF_OK;
}

static int solid_fixed_step(cf_solid* s, float dt, cf_step_stats* stats) {
    if (!s || !(dt >= 0) || (dt == 0 && s->p.integrator != CF_IT_FIE_DX)) return CF_ERR_INVALID_ARG;
    if (s->p.reference_order) return solid_fixed_step_exact(s, dt, stats);
    s->nonfinite = false;
    cf_tetset* set = s->set;
    const TetDevice& t = set->dev;
    hipStream_t st = s->stream;
    const uint32_t n3 = 3 * s->N;
    CF_CHECK_HIP(hipEventRecord(s->ev0, st));
    // b = dt (M g + f_e(x)) + M v ; y = v ; KNC rows (Step_QuasiImplicitEuler_v :1209-1295)
    NodeTerms nt;
    nt.out = s->bvec; nt.y = s->y; nt.x = s->x; nt.v = s->v;
    nt.knc_mask = s->num_knc ? s->knc_mask : nullptr;
    nt.knc_target = s->knc_target;
    nt.mass = s->node_mass;
    nt.dt = dt; nt.inv_dt = 1.0f / dt;
    nt.g[0] = s->p.gravity[0]; nt.g[1] = s->p.gravity[1]; nt.g[2] = s->p.gravity[2];
    nt.warm_start = s->p.warm_start;
    // BeginEvaluation(prev_x, x, v, t) (FixedStep :927), fused with f_e and
    // the CG start's K y where the material allows
    const bool fused = front_fusable(s);
    if (fused) CF_TRY(begin_front(s, nt));
    else CF_TRY(cf_begin_evaluation(s->cache, s->x0, s->x, s->v, s->total_time));
    // m_vecPos0 <- x (:936-937)
    CF_CHECK_HIP(hipMemcpyAsync(s->x0, s->x, n3 * 4, hipMemcpyDeviceToDevice, st));
    if (s->p.integrator == CF_IT_FIE_DX) {
        int32_t nr = 0;
        float nr_rel = 0.f, nr_abs = 0.f;
        CF_TRY(step_fie_dx(s, dt, &nr, &nr_rel, &nr_abs));
        // pos = x_k, vel = v_k (:1546-1551), air drag (:982-984)
        CF_TRY(launch_nr_finish(n3, s->x, s->v, s->xk, s->vk, dt, s->air_drag_per_second, st));
        double energy = 0.0;
        if (s->p.compute_energy) CF_TRY(cf_V(s->cache, &energy));  // V() at the final iterate (:977)
        CF_TRY(cf_end_evaluation(s->cache));
        CF_CHECK_HIP(hipEventRecord(s->ev1, st));
        CF_CHECK_HIP(hipEventSynchronize(s->ev1));
        if (stats) {
            float ms = 0.f;
            CF_CHECK_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
            CF_TRY(solid_fill_stats(s, energy, ms, stats));
            stats->nr_iters = nr;
            stats->nr_rel_residual = nr_rel;
            stats->nr_abs_residual = nr_abs;
            if (s->nonfinite) stats->cg_reason = CF_CG_NONFINITE;
        }
        return nonfinite_status(s);
    }
    if (!fused) {
        CF_TRY(launch_forces(t, kModeFe, s->x, s->cache->R, cache_mat(s->cache), st));
        CF_TRY(launch_gather(t, kGatherRhs, nt, node_blocks(s->N), st));
    }  // fused: the right-hand side gather runs with the CG start
    // A = (1 + dt c0) M + (dt^2 + dt c1) K  (:1300-1302)
    const float fM = 1.0f + dt * s->rayleigh[0];
    const float fK = dt * dt + dt * s->rayleigh[1];
    // The rest of the step is queued behind the CG without waiting for it;
    // the CG record, V() and the BeginEvaluation statistics come back through
    // pinned memory with the step's one synchronisation.
    cf_cg_result res{};
    CF_TRY(solid_cg(s, fM, fK, &res, true, fused ? s->ay_part : nullptr, fused ? &nt : nullptr));
    // v = y ; x = x + dt y (:1322-1327) ; air drag (:982-984)
    CF_TRY(launch_integrate(n3, s->x, s->v, s->y, dt, s->air_drag_per_second, st));
    *s->h_energy = 0.0;
    if (s->p.compute_energy) CF_TRY(cache_energy_enqueue(s->cache, s->h_energy));  // V() with the new x (:977)
    if (stats) CF_TRY(cache_stats_enqueue(s->cache));
    CF_CHECK_HIP(hipEventRecord(s->ev1, st));
    CF_CHECK_HIP(hipEventSynchronize(s->ev1));
    CF_TRY(cf_end_evaluation(s->cache));
    cg_result_from(*s->h_cgs, std::min<uint32_t>(s->p.cg_max_iter, kMaxCgIter), &res);
    if (res.reason == CF_CG_NONFINITE) s->nonfinite = true;
    s->last_ls = res;
    const double energy = *s->h_energy;
    if (stats) {
        float ms = 0.f;
        CF_CHECK_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        cache_stats_reduce(s->cache);
        stats->num_fixed_steps = 1;
        stats->cg_iters = res.iterations;
        stats->cg_rel_residual = res.rel_residual;
        stats->cg_abs_residual = res.abs_residual;
        stats->cg_reason = res.reason;
        stats->num_degenerate = s->cache->num_degenerate;
        stats->sum_degenerate_det_f = s->cache->sum_degenerate;
        stats->elastic_energy = float(energy);
        stats->min_det_f = s->cache->min_det;
        stats->step_ms = ms;
        stats->nr_iters = 0;
        stats->nr_rel_residual = 0.f;
        stats->nr_abs_residual = 0.f;
    }
    return nonfinite_status(s);
}

int cf_solid_fixed_step(cf_solid* s, float dt, cf_step_stats* stats) {
    const int rc = solid_fixed_step(s, dt, stats);
    if (s && (rc == CF_OK || rc == CF_ERR_NAN)) s->total_time += dt;
    return rc;
}

int cf_solid_total_time(const cf_solid* s, float* t) {
    if (!s || !t) return CF_ERR_INVALID_ARG;
    *t = s->total_time;
    return CF_OK;
}

// ---------------------------------------------------------------- multi-solid
int cf_multi_create(cf_solid* const* solids, uint32_t n, cf_multi** out) {
    if (!out || !solids || n == 0) { set_error("cf_multi_create: invalid arguments"); return CF_ERR_INVALID_ARG; }
    *out = nullptr;
    cf_multi* m = new (std::nothrow) cf_multi;
    if (!m) return CF_ERR_OUT_OF_MEMORY;
    for (uint32_t i = 0; i < n; i++) {
        if (!solids[i]) { delete m; set_error("cf_multi_create: solid %u is null", i); return CF_ERR_INVALID_ARG; }
        if (solids[i]->p.reference_order) {
            delete m;
            set_error("cf_multi_create: solid %u uses reference_order (leaf steps only)", i);
            return CF_ERR_UNSUPPORTED;
        }
        m->solids.push_back(solids[i]);
    }
    *out = m;
    return CF_OK;
}

int cf_multi_destroy(cf_multi* m) {
    delete m;
    return CF_OK;
}

int cf_multi_set_graphs(cf_multi* m, int use_graphs) {
    if (!m) return CF_ERR_INVALID_ARG;
    m->use_graphs = use_graphs != 0;
    return CF_OK;
}

// SS_AggregateDSH_Basic::MultiDSH_Solid3D_Update (ds/ss/SS_AggregateDSH_Basic.cpp:229-300)
// without contacts: fixed dt = min of the solids' fixed_dt, Newton passes = max
// of their nr_max_iter, one accumulator with strict '>'; each fixed step runs
// the leaves' NR hooks (LeafDSH_Solid3D_FEM.cpp:2824-3039) with the frame dt.
// NR_PreStep resets x_k = x, v_k = v and the pass count every pass, so each
// pass evaluates f_e at x under the rotations of the previous pass's end
// positions and warm-starts y = dt v.
int cf_multi_update(cf_multi* m, float dt, uint32_t* num_fixed_steps, cf_step_stats* stats) {
    if (!m || !(dt >= 0)) return CF_ERR_INVALID_ARG;
    float fixed_dt = std::numeric_limits<float>::infinity();
    uint32_t max_nr = 0;
    for (cf_solid* s : m->solids) {
        fixed_dt = std::min(fixed_dt, s->p.fixed_dt);
        max_nr = std::max(max_nr, s->p.nr_max_iter);
        s->total_time += dt;
    }
    m->acc_time += dt;
    const float inv_dt = (dt == 0.0f) ? 0.0f : 1.0f / dt;
    const size_t n = m->solids.size();
    uint32_t steps = 0;
    bool nonfinite = false;
    std::vector<CgRun> runs(n);
    while (m->acc_time > fixed_dt) {
        for (cf_solid* s : m->solids) {  // NR_Init (:2824-2836)
            s->nonfinite = false;
            CF_TRY(cf_begin_evaluation(s->cache, s->x0, s->x, s->v, s->total_time));
            CF_CHECK_HIP(hipMemcpyAsync(s->x0, s->x, 3 * (size_t)s->N * 4, hipMemcpyDeviceToDevice, s->stream));
            s->ls_total = 0;
        }
        uint32_t nr = 0;
        do {
            // NR_PreStep + NR_MidStep (:2857-2985): every leaf's solve is
            // enqueued on its own stream, eight iterations at a time (one
            // recorded graph per chunk unless graphs are off), before the
            // host waits on any of them
            for (size_t i = 0; i < n; i++) {
                cf_solid* s = m->solids[i];
                const size_t bytes = 3 * (size_t)s->N * 4;
                CF_CHECK_HIP(hipMemcpyAsync(s->xk, s->x, bytes, hipMemcpyDeviceToDevice, s->stream));
                CF_CHECK_HIP(hipMemcpyAsync(s->vk, s->v, bytes, hipMemcpyDeviceToDevice, s->stream));
                CF_TRY(fie_rhs(s, dt, inv_dt, true, nullptr));
                s->ls_total = 0;
                const float fM = inv_dt * inv_dt + inv_dt * s->rayleigh[0];
                const float fK = 1.0f + inv_dt * s->rayleigh[1];
                CF_TRY(cg_begin(s, fM, fK, &runs[i]));
            }
            for (size_t active = n; active > 0;) {
                for (size_t i = 0; i < n; i++)
                    if (!runs[i].done) CF_TRY(cg_launch(m->solids[i], &runs[i], m->use_graphs));
                active = 0;
                for (size_t i = 0; i < n; i++) {
                    if (runs[i].done) continue;
                    CF_TRY(cg_poll(m->solids[i], &runs[i]));
                    active += runs[i].done ? 0 : 1;
                }
            }
            for (size_t i = 0; i < n; i++) {
                cf_solid* s = m->solids[i];
                CF_TRY(cg_finish(s, runs[i], &s->last_ls));
                s->ls_total = s->last_ls.iterations;
            }
            for (cf_solid* s : m->solids) {  // NR_PostStep (:2987-3037)
                const uint32_t n3 = 3 * s->N;
                CF_TRY(launch_nr_update(n3, s->xk, s->vk, s->x, s->y, inv_dt, s->cg.part_b, node_blocks(n3), s->stream));
                CF_TRY(cf_begin_evaluation(s->cache, s->x, s->xk, s->vk, s->total_time));
            }
        } while (++nr < max_nr);
        for (size_t i = 0; i < n; i++) {  // NR_End (:2838-2852)
            cf_solid* s = m->solids[i];
            CF_TRY(launch_nr_finish(3 * s->N, s->x, s->v, s->xk, s->vk, dt, s->air_drag_per_second, s->stream));
            CF_TRY(cf_end_evaluation(s->cache));
            CF_CHECK_HIP(hipStreamSynchronize(s->stream));
            nonfinite = nonfinite || s->nonfinite;
            if (stats) {
                CF_TRY(solid_fill_stats(s, 0.0, 0.f, &stats[i]));
                stats[i].nr_iters = int32_t(nr);
                stats[i].nr_rel_residual = 0.f;
                stats[i].nr_abs_residual = 0.f;
            }
        }
        m->acc_time -= fixed_dt;
        steps++;
    }
    if (num_fixed_steps) *num_fixed_steps = steps;
    if (nonfinite) {
        set_error("multi-solid step: a CG reduction is NaN or Inf (non-finite state or forces)");
        return CF_ERR_NAN;
    }
    return CF_OK;
}

// Step (:225-256): fixed-dt accumulator with strict '>'
// Step (:225-256): scaled_dt = time_scale dt goes to the total time once per
// call (:229-230) and to the accumulator; fixed steps of time_scale fixed_dt
// run while the accumulator is strictly greater (:239)
int cf_solid_step(cf_solid* s, float dt, cf_step_stats* stats) {
    if (!s || !(dt >= 0)) return CF_ERR_INVALID_ARG;
    const float scaled_dt = s->p.time_scale * dt;
    s->total_time += scaled_dt;
    s->acc_time += scaled_dt;
    const float fixed = s->p.time_scale * s->p.fixed_dt;
    cf_step_stats last{};
    uint32_t n = 0;
    int rc = CF_OK;
    while (s->acc_time > fixed) {
        rc = solid_fixed_step(s, fixed, &last);
        if (rc != CF_OK && rc != CF_ERR_NAN) return rc;
        s->acc_time -= fixed;
        n++;
        if (rc == CF_ERR_NAN) break;
    }
    last.num_fixed_steps = n;
    if (stats) *stats = last;
    return rc;
}

int cf_solid_get_state(const cf_solid* s, float* pos, float* vel) {
    if (!s) return CF_ERR_INVALID_ARG;
    CF_CHECK_HIP(hipStreamSynchronize(s->stream));
    if (pos) CF_CHECK_HIP(hipMemcpy(pos, s->x, 3 * (size_t)s->N * 4, hipMemcpyDeviceToHost));
    if (vel) CF_CHECK_HIP(hipMemcpy(vel, s->v, 3 * (size_t)s->N * 4, hipMemcpyDeviceToHost));
    return CF_OK;
}

int cf_solid_set_state(cf_solid* s, const float* pos, const float* vel) {
    if (!s) return CF_ERR_INVALID_ARG;
    CF_CHECK_HIP(hipStreamSynchronize(s->stream));
    if (pos) CF_CHECK_HIP(hipMemcpy(s->x, pos, 3 * (size_t)s->N * 4, hipMemcpyHostToDevice));
    if (vel) CF_CHECK_HIP(hipMemcpy(s->v, vel, 3 * (size_t)s->N * 4, hipMemcpyHostToDevice));
    return CF_OK;
}

int cf_solid_device_state(cf_solid* s, float** pos, float** vel) {
    if (!s) return CF_ERR_INVALID_ARG;
    if (pos) *pos = s->x;
    if (vel) *vel = s->v;
    return CF_OK;
}

int cf_solid_info(const cf_solid* s, uint32_t* nn, uint32_t* nt, float* mass, float rayleigh[2]) {
    if (!s) return CF_ERR_INVALID_ARG;
    if (nn) *nn = s->N;
    if (nt) *nt = s->set->dev.T;
    if (mass) *mass = s->node_mass;
    if (rayleigh) { rayleigh[0] = s->rayleigh[0]; rayleigh[1] = s->rayleigh[1]; }
    return CF_OK;
}

int cf_solid_assemble_rhs(cf_solid* s, float dt, float* h_b, float* h_y0) {
    if (!s || !(dt > 0)) return CF_ERR_INVALID_ARG;
    cf_tetset* set = s->set;
    hipStream_t st = s->stream;
    CF_TRY(cf_begin_evaluation(s->cache, s->x0, s->x, s->v, s->total_time));
    CF_TRY(launch_forces(set->dev, kModeFe, s->x, s->cache->R, cache_mat(s->cache), st));
    NodeTerms nt;
    nt.out = s->bvec; nt.y = s->y; nt.x = s->x; nt.v = s->v;
    nt.knc_mask = s->num_knc ? s->knc_mask : nullptr;
    nt.knc_target = s->knc_target;
    nt.mass = s->node_mass;
    nt.dt = dt; nt.inv_dt = 1.0f / dt;
    nt.g[0] = s->p.gravity[0]; nt.g[1] = s->p.gravity[1]; nt.g[2] = s->p.gravity[2];
    nt.warm_start = s->p.warm_start;
    CF_TRY(launch_gather(set->dev, kGatherRhs, nt, node_blocks(s->N), st));
    CF_TRY(cf_end_evaluation(s->cache));
    CF_CHECK_HIP(hipStreamSynchronize(st));
    if (h_b) CF_CHECK_HIP(hipMemcpy(h_b, s->bvec, 3 * (size_t)s->N * 4, hipMemcpyDeviceToHost));
    if (h_y0) CF_CHECK_HIP(hipMemcpy(h_y0, s->y, 3 * (size_t)s->N * 4, hipMemcpyDeviceToHost));
    return CF_OK;
}

int cf_solid_apply_A(cf_solid* s, float fM, float fK, const float* h_d, float* h_q) {
    if (!s || !h_d || !h_q) return CF_ERR_INVALID_ARG;
    cf_tetset* set = s->set;
    hipStream_t st = s->stream;
    CgBuffers& b = s->cg;
    const size_t n3 = 3 * (size_t)s->N;
    CF_CHECK_HIP(hipStreamSynchronize(st));
    CF_CHECK_HIP(hipMemcpy(b.d, h_d, n3 * 4, hipMemcpyHostToDevice));
    NodeTerms nt;
    nt.knc_mask = s->num_knc ? s->knc_mask : nullptr;
    nt.mass = s->node_mass;
    nt.fM = fM; nt.fK = fK;
    nt.partials = b.part_a;
    nt.d = b.d; nt.out = b.q;
    CF_TRY(launch_forces(set->dev, kModeDfx, b.d, s->cache->R, cache_mat(s->cache), st));
    CF_TRY(launch_gather(set->dev, kGatherMatvec, nt, node_blocks(s->N), st));
    CF_CHECK_HIP(hipStreamSynchronize(st));
    CF_CHECK_HIP(hipMemcpy(h_q, b.q, n3 * 4, hipMemcpyDeviceToHost));
    return CF_OK;
}

int cf_solid_time_df_x(cf_solid* s, uint32_t reps, double* ms) {
    if (!s || reps == 0 || !ms) return CF_ERR_INVALID_ARG;
    cf_tetset* set = s->set;
    hipStream_t st = s->stream;
    if (!s->scratch_vec) CF_TRY(dalloc(&s->scratch_vec, 3 * (size_t)s->N));
    NodeTerms nt;
    nt.out = s->scratch_vec;
    CF_CHECK_HIP(hipEventRecord(s->ev0, st));
    for (uint32_t r = 0; r < reps; r++) {
        CF_TRY(launch_forces(set->dev, kModeDfx, s->v, s->cache->R, cache_mat(s->cache), st));
        CF_TRY(launch_gather(set->dev, kGatherAccumulate, nt, node_blocks(s->N), st));
    }
    CF_CHECK_HIP(hipEventRecord(s->ev1, st));
    CF_CHECK_HIP(hipEventSynchronize(s->ev1));
    float t = 0.f;
    CF_CHECK_HIP(hipEventElapsedTime(&t, s->ev0, s->ev1));
    *ms = double(t) / reps;
    return CF_OK;
}

int cf_solid_time_kx(cf_solid* s, uint32_t reps, double* ms_tile, double* ms_gather) {
    if (!s || reps == 0 || !ms_tile || !ms_gather) return CF_ERR_INVALID_ARG;
    cf_tetset* set = s->set;
    hipStream_t st = s->stream;
    CgBuffers& b = s->cg;
    CF_CHECK_HIP(hipMemcpyAsync(b.d, s->v, 3 * (size_t)s->N * 4, hipMemcpyDeviceToDevice, st));
    NodeTerms nt;
    nt.knc_mask = s->num_knc ? s->knc_mask : nullptr;
    nt.mass = s->node_mass;
    nt.fM = 1.0f; nt.fK = s->p.fixed_dt * s->p.fixed_dt;
    nt.partials = b.part_a;
    nt.d = b.d; nt.out = b.q;
    const int nnode = node_blocks(s->N);
    float t = 0.f;
    CF_CHECK_HIP(hipEventRecord(s->ev0, st));
    for (uint32_t r = 0; r < reps; r++) CF_TRY(launch_forces(set->dev, kModeDfx, b.d, s->cache->R, cache_mat(s->cache), st));
    CF_CHECK_HIP(hipEventRecord(s->ev1, st));
    CF_CHECK_HIP(hipEventSynchronize(s->ev1));
    CF_CHECK_HIP(hipEventElapsedTime(&t, s->ev0, s->ev1));
    *ms_tile = double(t) / reps;
    CF_CHECK_HIP(hipEventRecord(s->ev0, st));
    for (uint32_t r = 0; r < reps; r++) CF_TRY(launch_gather(set->dev, kGatherMatvec, nt, nnode, st));
    CF_CHECK_HIP(hipEventRecord(s->ev1, st));
    CF_CHECK_HIP(hipEventSynchronize(s->ev1));
    CF_CHECK_HIP(hipEventElapsedTime(&t, s->ev0, s->ev1));
    *ms_gather = double(t) / reps;
    return CF_OK;
}

int cf_set_tile_elems(uint32_t te) {
    if (te != 0 && (te < 16 || te > 16383)) { set_error("tile_elems out of [16, 16383]"); return CF_ERR_INVALID_ARG; }
    g_tile_elems = te;
    return CF_OK;
}
int cf_set_tile_threads(uint32_t n) {
    if (n != 0 && n != 128 && n != 256 && n != 512 && n != 1024) {
        set_error("tile threads must be 128, 256, 512 or 1024");
        return CF_ERR_INVALID_ARG;
    }
    g_tile_threads = n ? n : 256;
    return CF_OK;
}

int cf_set_tile_persistent(int persistent) {
    if (persistent < 0 || persistent > 2) { set_error("tile schedule must be 0, 1 or 2"); return CF_ERR_INVALID_ARG; }
    g_tile_persistent = persistent;
    return CF_OK;
}

int cf_solid_set_graphs(cf_solid* s, int use_graphs) {
    if (!s) return CF_ERR_INVALID_ARG;
    s->use_graphs = use_graphs != 0;
    return CF_OK;
}

int cf_set_fused_front(int enable) {
    if (enable != 0 && enable != 1) { set_error("fused front must be 0 or 1"); return CF_ERR_INVALID_ARG; }
    g_fused_front = enable;
    return CF_OK;
}

int cf_set_tile_alternate(int alternate) {
    if (alternate != 0 && alternate != 1) { set_error("tile alternation must be 0 or 1"); return CF_ERR_INVALID_ARG; }
    g_tile_alternate = alternate;
    return CF_OK;
}

int cf_get_tile_elems(uint32_t* te) {
    if (!te) return CF_ERR_INVALID_ARG;
    *te = g_tile_elems ? g_tile_elems : default_tile_elems();
    return CF_OK;
}

int cf_set_cg_workgroups(uint32_t n) {
    if (set_cg_workgroups(n) != CF_OK) {
        set_error("CG workgroups must be at most %d", kMaxPartials);
        return CF_ERR_INVALID_ARG;
    }
    return CF_OK;
}

int cf_get_cg_workgroups(uint32_t* n) {
    if (!n) return CF_ERR_INVALID_ARG;
    *n = get_cg_workgroups();
    return CF_OK;
}

}  // extern "C"
