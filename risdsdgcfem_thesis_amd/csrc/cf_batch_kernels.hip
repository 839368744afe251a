// Batched kernels: one launch steps the same stage of every simulation of a
// batch. A block reads its (simulation slot, block within that simulation's
// grid) from a block table and runs the single-mesh block body with that
// simulation's arrays; per-simulation scalars (the CG's delta, alpha, beta and
// stop tests) are finalised by one wavefront per simulation that reproduces
// the single-mesh one-block reduction. Simulations whose solve has stopped
// skip the iteration kernels, so a batch iterates until its slowest member is
// done while converged members cost one table and one flag read per block.
#include "cf_device.h"

namespace cf {

namespace {

__device__ inline TileArgs tile_args(const SimDev& s, const float* in, uint32_t max_tn, uint32_t max_te) {
    TileArgs a;
    a.T = s.T;
    a.num_tiles = s.num_tiles;
    a.max_te = max_te;
    a.max_tn = max_tn;
    a.tile_elem_ptr = s.tile_elem_ptr;
    a.tile_node_ptr = s.tile_node_ptr;
    a.tile_nodes = s.tile_nodes;
    a.part_dst = s.part_dst;
    a.slot_ptr = s.slot_ptr;
    a.lv = s.lv;
    a.rk = s.rk;
    a.R = s.R;
    a.dinv = s.dinv;
    a.vol = s.vol;
    a.ref = s.ref;
    a.ref_tile = s.ref_tile;
    a.in = in;
    a.partials = s.partials;
    a.m = s.m;
    a.skip = nullptr;
    return a;
}

__device__ inline NodeTerms node_terms(const SimDev& s) {
    NodeTerms t;
    t.knc_mask = s.knc_mask;
    t.knc_target = s.knc_target;
    t.mass = s.mass;
    t.mass_array = nullptr;
    t.fM = s.fM;
    t.fK = s.fK;
    t.dt = s.dt;
    t.inv_dt = s.inv_dt;
    t.g[0] = s.g[0]; t.g[1] = s.g[1]; t.g[2] = s.g[2];
    t.warm_start = s.warm_start;
    return t;
}

// The slot of the wavefront's simulation in a per-simulation kernel
// (4 wavefronts per block), or kIdleSlot. The grid covers n_active
// simulations; a list compacted on the device (t.dyn) holds at most that many.
__device__ inline uint32_t wave_slot(const BatchTables& t) {
    const uint32_t w = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    // the slot is loaded together with the device length (a position below
    // n_active is inside the list's capacity; past the length it is stale)
    const uint32_t slot = (w < t.n_active) ? t.active[w] : kIdleSlot;
    const uint32_t n = t.dyn ? t.dyn[kDynActive] : t.n_active;
    return (w < n) ? slot : kIdleSlot;
}

}  // namespace

// ------------------------------------------------------------ step prologue
__global__ __launch_bounds__(kBlock, 8) void kb_begin_eval(BatchTables t) {
    __shared__ double sh[kBlock / 64];
    __shared__ float shm[kBlock / 64];
    __shared__ uint32_t shc[kBlock / 64], shf[kBlock / 64];
    if (blockIdx.x == 0 && threadIdx.x == 0) { *t.counter = 0; *t.busy = 0; }
    const BlockRef br = t.elem[blockIdx.x];
    if (br.slot == kIdleSlot || t.cgs[br.slot].phase != kPhaseStart) return;
    const SimDev& s = t.sims[br.slot];
    DegenArgs dg;
    dg.prev_x = s.x0;
    dg.doc = s.doc;
    dg.vol = s.vol;
    dg.l_factor = s.dapd_l;
    dg.nl_factor = s.dapd_nl;
    dg.nl_exponent = s.dapd_e;
    begin_eval_block(s.T, s.conn, s.x, s.dinv, s.rm, s.thr, s.R, s.detF, s.eval_count, s.eval_sum, s.eval_min,
                     br.blk, s.nb_elem, sh, shm, shc, shf, LcmOut(), dg);
}

// Element stage over every tile of the batch: f_e on x, or df_x on y (the
// warm start, before the solve) for starting simulations, or df_x on d for
// iterating ones; kSelYAll: df_x on y for every simulation.
template <int MODE, int TB>
__device__ inline void tiles_entry(const BatchTables& t, int sel, BlockRef br, const TileLds<TB>& L) {
    if (br.slot == kIdleSlot) return;
    if (sel != kSelYAll && t.cgs[br.slot].phase != (sel == kSelD ? kPhaseIter : kPhaseStart)) return;
    const SimDev& s = t.sims[br.slot];
    const float* in = (MODE == kModeFe) ? s.x : (sel == kSelD ? s.d : s.y);
    const TileArgs a = tile_args(s, in, t.max_tn, t.max_te);
    tile_block<MODE>(a, br.blk, L.v, L.r, L.F, L.sp);
}

// LOOP: a list compacted on the device (t.dyn), walked with a stride of the
// grid. A workgroup loads its first entry together with the list length (the
// grid never exceeds the list's capacity, so the load is in bounds; an entry
// past the length is stale and not used), so the length costs no extra
// latency. Otherwise: one table entry per workgroup.
template <int MODE, int TB, bool LOOP>
__global__ __launch_bounds__(TB) void kb_tiles(BatchTables t, int sel) {
    extern __shared__ float smem_f[];
    const TileLds<TB> L(smem_f);
    const BlockRef br = t.tile[blockIdx.x];
    if (!LOOP) {
        tiles_entry<MODE, TB>(t, sel, br, L);
        return;
    }
    const uint32_t n = t.dyn[kDynTile];
    if (blockIdx.x >= n) return;
    tiles_entry<MODE, TB>(t, sel, br, L);
    for (uint32_t i = blockIdx.x + gridDim.x; i < n; i += gridDim.x) {
        __syncthreads();  // the next entry restages the LDS planes
        tiles_entry<MODE, TB>(t, sel, t.tile[i], L);
    }
}

// b = dt (M g + f_e) + M v, y = v, KNC rows (Step_QuasiImplicitEuler_v,
// ds/dsh/LeafDSH_Solid3D_FEM.cpp:1209-1295)
__global__ __launch_bounds__(kBlock) void kb_gather_rhs(BatchTables t) {
    __shared__ double sh[kBlock / 64];
    const BlockRef br = t.node[blockIdx.x];
    if (br.slot == kIdleSlot || t.cgs[br.slot].phase != kPhaseStart) return;
    const SimDev& s = t.sims[br.slot];
    NodeTerms nt = node_terms(s);
    nt.out = s.b;
    nt.y = s.y;
    nt.x = s.x;
    nt.v = s.v;
    gather_block<kGatherRhs>(s.N, s.node_part_ptr, s.partials, nt, br.blk, s.nb_node, sh);
}

// Over the leaf's matvec chunks (one chunk per block): q = S(fM M d - fK K d)
// and per-chunk d.q (LS_A_times_d :1056-1079) for iterating simulations
// (kSelD) or, on y, for every simulation (kSelYAll); for starting simulations
// (kSelY) the CG start r = S(b - A y), d = r with per-chunk r.r instead
// (ILSS_CG.h:55-63), as the leaf's fused start gather computes it.
__device__ inline void gather_matvec_entry(const BatchTables& t, int sel, BlockRef br, double* sh) {
    if (br.slot == kIdleSlot) return;
    if (sel != kSelYAll && t.cgs[br.slot].phase != (sel == kSelD ? kPhaseIter : kPhaseStart)) return;
    const SimDev& s = t.sims[br.slot];
    NodeTerms nt = node_terms(s);
    double part;
    if (sel == kSelY) {
        nt.d = s.y;
        nt.b = s.b; nt.r_out = s.r; nt.d_out = s.d;
        part = block_sum(gather_matvec_chunk<true>(s.N, br.blk, s.npt_mv, s.node_part_ptr, s.partials, nt), sh);
        if (threadIdx.x == 0) s.part_b[br.blk] = part;
    } else {
        nt.d = (sel == kSelD) ? s.d : s.y;
        nt.out = s.q;
        part = block_sum(gather_matvec_chunk<false>(s.N, br.blk, s.npt_mv, s.node_part_ptr, s.partials, nt), sh);
        if (threadIdx.x == 0) s.part_a[br.blk] = part;
    }
}

template <bool LOOP>
__global__ __launch_bounds__(kBlock) void kb_gather_matvec(BatchTables t, int sel) {
    __shared__ double sh[kBlock / 64];
    const BlockRef br = t.mv[blockIdx.x];
    if (!LOOP) {
        gather_matvec_entry(t, sel, br, sh);
        return;
    }
    const uint32_t n = t.dyn[kDynMv];
    if (blockIdx.x >= n) return;
    gather_matvec_entry(t, sel, br, sh);
    for (uint32_t i = blockIdx.x + gridDim.x; i < n; i += gridDim.x) gather_matvec_entry(t, sel, t.mv[i], sh);
}

// ------------------------------------------------------------ CG (ns/ILSS_CG.h:38-151)

__global__ __launch_bounds__(kBlock) void kb_cg_start(BatchTables t) {
    const uint32_t slot = wave_slot(t);
    if (slot == kIdleSlot || t.cgs[slot].phase != kPhaseStart) return;
    const SimDev& s = t.sims[slot];
    const double d0 = wave_reduce_partials(s.part_b, s.nb_mv);
    if ((threadIdx.x & 63) == 0) {
        SimCg& c = t.cgs[slot];
        const float abs_sq = s.abs_eps * s.abs_eps;
        const float rel_sq = s.rel * s.rel;
        c.delta0 = d0;
        c.delta = d0;
        c.target = (double)rel_sq * d0;
        c.abs_sq = (double)abs_sq;
        c.max_iter = s.max_iter;
        c.iters = 0;
        c.dq = 0.0;
        c.alpha = 0.f;
        c.beta = 0.f;
        int reason = 0;
        if (!(d0 > (double)abs_sq)) reason = CF_CG_CONVERGED_ABS;
        else if (!(d0 > c.target)) reason = CF_CG_CONVERGED_REL;
        else if (!(0u < s.max_iter)) reason = CF_CG_MAX_ITER;
        c.done = reason != 0;
        c.reason = reason;
        c.phase = reason ? kPhaseEnd : kPhaseIter;
        if (!reason) atomicAdd(t.counter, 1u);
    }
}

// alpha = delta / d.q, or breakdown when |d.q| < abs^2 (ILSS_CG.h:91-99)
__global__ __launch_bounds__(kBlock) void kb_cg_alpha(BatchTables t) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *t.counter = 0;
    const uint32_t slot = wave_slot(t);
    if (slot == kIdleSlot) return;
    SimCg& c = t.cgs[slot];
    if (c.phase != kPhaseIter) return;
    const SimDev& s = t.sims[slot];
    const double dq = wave_reduce_partials(s.part_a, s.nb_mv);
    if ((threadIdx.x & 63) == 0) {
        c.dq = dq;
        if (fabs(dq) < c.abs_sq) {
            c.reason = CF_CG_BREAKDOWN;
            c.done = 2;
            c.phase = kPhaseEnd;
        } else {
            c.alpha = (float)(c.delta / dq);
        }
    }
}

// x += alpha d ; r -= alpha q over the leaf's CG chunks, with per-chunk r.r
__device__ inline void update_xr_entry(const BatchTables& t, BlockRef br, double* sh) {
    if (br.slot == kIdleSlot) return;
    const SimCg& c = t.cgs[br.slot];
    if (c.phase != kPhaseIter) return;
    const SimDev& s = t.sims[br.slot];
    const double part = block_sum(xr_rows_chunk(3 * s.N, br.blk, s.rows_cg, c.alpha, s.y, s.r, s.d, s.q), sh);
    if (threadIdx.x == 0) s.part_b[br.blk] = part;
}

template <bool LOOP>
__global__ __launch_bounds__(kBlock) void kb_update_xr(BatchTables t) {
    __shared__ double sh[kBlock / 64];
    const BlockRef br = t.vec[blockIdx.x];
    if (!LOOP) {
        update_xr_entry(t, br, sh);
        return;
    }
    const uint32_t n = t.dyn[kDynVec];
    if (blockIdx.x >= n) return;
    update_xr_entry(t, br, sh);
    for (uint32_t i = blockIdx.x + gridDim.x; i < n; i += gridDim.x) update_xr_entry(t, t.vec[i], sh);
}

// delta' = r.r ; beta = delta'/delta ; stop tests (ILSS_CG.h:86-88,115-133)
__global__ __launch_bounds__(kBlock) void kb_cg_beta(BatchTables t) {
    const uint32_t slot = wave_slot(t);
    if (slot == kIdleSlot) return;
    SimCg& c = t.cgs[slot];
    if (c.phase != kPhaseIter) return;
    const SimDev& s = t.sims[slot];
    const double dn = wave_reduce_partials(s.part_b, s.nb_vec);
    if ((threadIdx.x & 63) == 0) {
        c.beta = (float)(dn / c.delta);
        c.delta = dn;
        const uint32_t k1 = c.iters + 1;
        c.iters = k1;
        int reason = 0;
        if (!(dn > c.abs_sq)) reason = CF_CG_CONVERGED_ABS;
        else if (!(dn > c.target)) reason = CF_CG_CONVERGED_REL;
        else if (!(k1 < c.max_iter)) reason = CF_CG_MAX_ITER;
        c.reason = reason;
        c.done = reason != 0;
        if (reason) c.phase = kPhaseEnd;
        else atomicAdd(t.counter, 1u);
    }
}

__device__ inline void update_d_entry(const BatchTables& t, BlockRef br) {
    if (br.slot == kIdleSlot) return;
    const SimCg& c = t.cgs[br.slot];
    if (c.phase != kPhaseIter) return;
    const SimDev& s = t.sims[br.slot];
    update_d_block<true>(3 * s.N, c.beta, s.d, s.r, s.knc_mask, br.blk, s.nb_vec);
}

template <bool LOOP>
__global__ __launch_bounds__(kBlock) void kb_update_d(BatchTables t) {
    const BlockRef br = t.vec[blockIdx.x];
    if (!LOOP) {
        update_d_entry(t, br);
        return;
    }
    const uint32_t n = t.dyn[kDynVec];
    if (blockIdx.x >= n) return;
    update_d_entry(t, br);
    for (uint32_t i = blockIdx.x + gridDim.x; i < n; i += gridDim.x) update_d_entry(t, t.vec[i]);
}

// ------------------------------------------------------------ step epilogue
__global__ __launch_bounds__(kBlock) void kb_integrate(BatchTables t) {
    const BlockRef br = t.init[blockIdx.x];
    if (br.slot == kIdleSlot || t.cgs[br.slot].phase != kPhaseEnd) return;
    const SimDev& s = t.sims[br.slot];
    integrate_block(3 * s.N, s.x, s.v, s.y, s.dt, s.drag, br.blk, s.nb_init, s.x0);
}

// V() with the new positions (FixedStep :977), C_WRP per tile as the leaf's
// tile energy computes it (part_e[tile]); batches hold C_WRP simulations only.
template <int TB>
__global__ __launch_bounds__(TB) void kb_tile_energy(BatchTables t) {
    extern __shared__ float smem_f[];
    __shared__ double sh[TB / 64];
    const BlockRef br = t.tile[blockIdx.x];
    if (br.slot == kIdleSlot || t.cgs[br.slot].phase != kPhaseEnd) return;
    const SimDev& s = t.sims[br.slot];
    if (!s.compute_energy) return;
    const TileArgs a = tile_args(s, s.x, t.max_tn, t.max_te);
    const double e = tile_energy_block<TB>(a, br.blk, smem_f, sh);
    if (threadIdx.x == 0) s.part_e[br.blk] = e;
}

// Per-simulation step results, one wavefront per simulation: the
// BeginEvaluation counts and minimum over the element blocks (any order gives
// the same values), the sum of degenerate det F in block order (as the leaf's
// host sum; only when there are degenerate elements, else it is 0), and the
// per-tile energies in the leaf's one-block order.
__global__ __launch_bounds__(kBlock) void kb_finish(BatchTables t) {
    const uint32_t slot = wave_slot(t);
    if (slot == kIdleSlot) return;
    const SimDev& s = t.sims[slot];
    SimCg& c = t.cgs[slot];
    if (c.phase != kPhaseEnd) return;
    const uint32_t lane = threadIdx.x & 63;
    uint32_t nd = 0, nf = 0;
    float mn = 3e38f;
    for (uint32_t i = lane; i < s.nb_elem; i += 64) {
        nd += s.eval_count[i];
        nf += s.eval_count[s.nb_elem + i];
        mn = fminf(mn, s.eval_min[i]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        nd += __shfl_xor(nd, o);
        nf += __shfl_xor(nf, o);
        mn = fminf(mn, __shfl_xor(mn, o));
    }
    double en = 0.0;
    if (s.compute_energy)
        en = wave_reduce_partials(s.part_e, (int)s.num_tiles);
    if (lane != 0) return;
    double sd = 0.0;
    if (nd)
        for (uint32_t i = 0; i < s.nb_elem; i++) sd += s.eval_sum[i];
    c.num_degenerate = nd;
    c.num_fallback = nf;
    c.sum_degenerate = sd;
    c.min_det = mn;
    c.energy = en;
    c.steps_done++;
    if (c.steps_left) {
        c.steps_left--;
        c.phase = kPhaseStart;
    } else {
        c.phase = kPhaseIdle;
    }
}

// Every active simulation starts one step (cf_batch_fixed_step).
__global__ void kb_arm_all(BatchTables t) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t.n_active) return;
    SimCg& c = t.cgs[t.active[i]];
    c.phase = kPhaseStart;
    c.steps_left = 0;
}

// Queued steps: an idle simulation starts its first one at once.
__global__ void kb_arm(BatchTables t, const uint2* list, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || list[i].y == 0) return;
    SimCg& c = t.cgs[list[i].x];
    if (c.phase == kPhaseIdle) {
        c.phase = kPhaseStart;
        c.steps_left += list[i].y - 1;
    } else {
        c.steps_left += list[i].y;
    }
}

// Simulations with work left after a round.
__global__ void kb_count_busy(BatchTables t) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < t.n_active && t.cgs[t.active[i]].phase != kPhaseIdle) atomicAdd(t.busy, 1u);
}

// ------------------------------------------------------------ device compaction
// The CG-loop lists of the simulations still iterating, in the order of the
// full active list (the order the host compaction uses): per simulation its
// tiles (0 .. num_tiles), matvec chunks and CG vector chunks, and its slot.
// The tile list is dealt to the XCDs as the host tables are: natural entry
// m goes to position (m % per) * 8 + m / per, per = ceil(tiles / 8), so that
// a workgroup striding by a multiple of 8 stays on one XCD and each XCD walks
// a contiguous range of tiles; the positions of m in [tiles, 8 per) hold
// idle entries.
constexpr uint32_t kScanThreads = 1024;

__device__ inline bool iterating(const BatchTables& t, uint32_t slot) { return t.cgs[slot].phase == kPhaseIter; }

// One workgroup: per-simulation offsets (exclusive prefix sums over the full
// active list) and the list lengths. Thread k owns a contiguous run of
// simulations, so the offsets follow the list order.
__global__ __launch_bounds__(kScanThreads) void kb_compact_scan(BatchTables t, CompactOut o) {
    __shared__ uint32_t sc[4][kScanThreads];
    const uint32_t tid = threadIdx.x, n = t.n_active;
    const uint32_t per = (n + kScanThreads - 1) / kScanThreads;
    const uint32_t i0 = min(n, tid * per), i1 = min(n, i0 + per);
    uint32_t c[4] = {0, 0, 0, 0};
    for (uint32_t i = i0; i < i1; i++) {
        const uint32_t slot = t.active[i];
        if (!iterating(t, slot)) continue;
        const SimDev& s = t.sims[slot];
        c[0] += s.num_tiles; c[1] += s.nb_mv; c[2] += s.nb_vec; c[3] += 1;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) sc[k][tid] = c[k];
    __syncthreads();
    // inclusive Hillis-Steele scan of the four counts
    for (uint32_t w = 1; w < kScanThreads; w <<= 1) {
        uint32_t v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = (tid >= w) ? sc[k][tid - w] : 0u;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; k++) sc[k][tid] += v[k];
        __syncthreads();
    }
    uint32_t run[4];
#pragma unroll
    for (int k = 0; k < 4; k++) run[k] = sc[k][tid] - c[k];
    for (uint32_t i = i0; i < i1; i++) {
        const uint32_t slot = t.active[i];
#pragma unroll
        for (int k = 0; k < 4; k++) o.off[4 * i + k] = run[k];
        if (!iterating(t, slot)) continue;
        const SimDev& s = t.sims[slot];
        run[0] += s.num_tiles; run[1] += s.nb_mv; run[2] += s.nb_vec; run[3] += 1;
    }
    if (tid == kScanThreads - 1) {
        const uint32_t tiles = sc[0][tid];
        const uint32_t tper = (tiles + 7) / 8;
        o.dyn[kDynTile] = 8 * tper;
        o.dyn[kDynMv] = sc[1][tid];
        o.dyn[kDynVec] = sc[2][tid];
        o.dyn[kDynActive] = sc[3][tid];
        o.dyn[kDynTilePer] = tper;
        o.dyn[kDynTiles] = tiles;
        *t.busy = 0;
    }
}

// The entries, one simulation per workgroup at a time (coalesced writes).
__global__ __launch_bounds__(kBlock) void kb_compact_fill(BatchTables t, CompactOut o) {
    const uint32_t tper = o.dyn[kDynTilePer], tiles = o.dyn[kDynTiles];
    for (uint32_t i = blockIdx.x; i < t.n_active; i += gridDim.x) {
        const uint32_t slot = t.active[i];
        if (!iterating(t, slot)) continue;
        const SimDev& s = t.sims[slot];
        const uint32_t* off = o.off + 4 * i;
        for (uint32_t j = threadIdx.x; j < s.num_tiles; j += kBlock) {
            const uint32_t m = off[0] + j;
            o.tile[(m % tper) * 8 + m / tper] = BlockRef{slot, j};
        }
        for (uint32_t j = threadIdx.x; j < s.nb_mv; j += kBlock) o.mv[off[1] + j] = BlockRef{slot, j};
        for (uint32_t j = threadIdx.x; j < s.nb_vec; j += kBlock) o.vec[off[2] + j] = BlockRef{slot, j};
        if (threadIdx.x == 0) o.active[off[3]] = slot;
    }
    if (blockIdx.x == 0 && threadIdx.x < 8 * tper - tiles) {
        const uint32_t m = tiles + threadIdx.x;
        o.tile[(m % tper) * 8 + m / tper] = BlockRef{kIdleSlot, 0};
    }
}

__global__ void kb_fill_u32(uint32_t* p, uint32_t value, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

// ------------------------------------------------------------ launchers
namespace {
int check_launch_b() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("batched kernel launch failed: %s", hipGetErrorString(e));
        return CF_ERR_HIP;
    }
    return CF_OK;
}
inline uint32_t mesh_grid(const BatchTables& t) { return (t.n_active + 3) / 4; }
}  // namespace

int launch_batch_begin_eval(const BatchTables& t, hipStream_t s) {
    if (!t.n_elem) return CF_OK;
    hipLaunchKernelGGL(kb_begin_eval, dim3(t.n_elem), dim3(kBlock), 0, s, t);
    return check_launch_b();
}

int launch_batch_forces(const BatchTables& t, int mode, int sel, hipStream_t s) {
    if (!t.n_tile) return CF_OK;
#define CF_BTILE(TBV)                                                                                     \
    do {                                                                                                  \
        const size_t sm = TileLds<TBV>::kBytes;                                                           \
        if (mode == kModeFe) hipLaunchKernelGGL((kb_tiles<kModeFe, TBV, false>), dim3(t.n_tile), dim3(TBV), sm, s, t, sel); \
        else if (t.dyn) hipLaunchKernelGGL((kb_tiles<kModeDfx, TBV, true>), dim3(t.n_tile), dim3(TBV), sm, s, t, sel); \
        else hipLaunchKernelGGL((kb_tiles<kModeDfx, TBV, false>), dim3(t.n_tile), dim3(TBV), sm, s, t, sel); \
    } while (0)
    switch (t.tile_threads) {
        case 1024: CF_BTILE(1024); break;
        case 512: CF_BTILE(512); break;
        case 128: CF_BTILE(128); break;
        default: CF_BTILE(256); break;
    }
#undef CF_BTILE
    return check_launch_b();
}

int launch_batch_gather_rhs(const BatchTables& t, hipStream_t s) {
    if (!t.n_node) return CF_OK;
    hipLaunchKernelGGL(kb_gather_rhs, dim3(t.n_node), dim3(kBlock), 0, s, t);
    return check_launch_b();
}

int launch_batch_gather_matvec(const BatchTables& t, int sel, hipStream_t s) {
    if (!t.n_mv) return CF_OK;
    if (t.dyn) hipLaunchKernelGGL(kb_gather_matvec<true>, dim3(t.n_mv), dim3(kBlock), 0, s, t, sel);
    else hipLaunchKernelGGL(kb_gather_matvec<false>, dim3(t.n_mv), dim3(kBlock), 0, s, t, sel);
    return check_launch_b();
}

int launch_batch_cg_start(const BatchTables& t, hipStream_t s) {
    if (!t.n_active) return CF_OK;
    hipLaunchKernelGGL(kb_cg_start, dim3(mesh_grid(t)), dim3(kBlock), 0, s, t);
    return check_launch_b();
}

int launch_batch_cg_alpha(const BatchTables& t, hipStream_t s) {
    if (!t.n_active) return CF_OK;
    hipLaunchKernelGGL(kb_cg_alpha, dim3(mesh_grid(t)), dim3(kBlock), 0, s, t);
    return check_launch_b();
}

int launch_batch_update_xr(const BatchTables& t, hipStream_t s) {
    if (!t.n_vec) return CF_OK;
    if (t.dyn) hipLaunchKernelGGL(kb_update_xr<true>, dim3(t.n_vec), dim3(kBlock), 0, s, t);
    else hipLaunchKernelGGL(kb_update_xr<false>, dim3(t.n_vec), dim3(kBlock), 0, s, t);
    return check_launch_b();
}

int launch_batch_cg_beta(const BatchTables& t, hipStream_t s) {
    if (!t.n_active) return CF_OK;
    hipLaunchKernelGGL(kb_cg_beta, dim3(mesh_grid(t)), dim3(kBlock), 0, s, t);
    return check_launch_b();
}

int launch_batch_update_d(const BatchTables& t, hipStream_t s) {
    if (!t.n_vec) return CF_OK;
    if (t.dyn) hipLaunchKernelGGL(kb_update_d<true>, dim3(t.n_vec), dim3(kBlock), 0, s, t);
    else hipLaunchKernelGGL(kb_update_d<false>, dim3(t.n_vec), dim3(kBlock), 0, s, t);
    return check_launch_b();
}

int launch_batch_integrate(const BatchTables& t, hipStream_t s) {
    if (!t.n_init) return CF_OK;
    hipLaunchKernelGGL(kb_integrate, dim3(t.n_init), dim3(kBlock), 0, s, t);
    return check_launch_b();
}

int launch_batch_energy(const BatchTables& t, hipStream_t s) {
    if (!t.n_tile) return CF_OK;
    switch (t.tile_threads) {
        case 1024: hipLaunchKernelGGL(kb_tile_energy<1024>, dim3(t.n_tile), dim3(1024), 6 * 1024 * 4, s, t); break;
        case 512: hipLaunchKernelGGL(kb_tile_energy<512>, dim3(t.n_tile), dim3(512), 6 * 512 * 4, s, t); break;
        case 128: hipLaunchKernelGGL(kb_tile_energy<128>, dim3(t.n_tile), dim3(128), 6 * 128 * 4, s, t); break;
        default: hipLaunchKernelGGL(kb_tile_energy<256>, dim3(t.n_tile), dim3(256), 6 * 256 * 4, s, t); break;
    }
    return check_launch_b();
}

int launch_batch_finish(const BatchTables& t, hipStream_t s) {
    if (!t.n_active) return CF_OK;
    hipLaunchKernelGGL(kb_finish, dim3(mesh_grid(t)), dim3(kBlock), 0, s, t);
    return check_launch_b();
}

int launch_batch_arm_all(const BatchTables& t, hipStream_t s) {
    if (!t.n_active) return CF_OK;
    hipLaunchKernelGGL(kb_arm_all, dim3((t.n_active + 255) / 256), dim3(256), 0, s, t);
    return check_launch_b();
}

int launch_batch_arm(const BatchTables& t, const uint2* list, uint32_t n, hipStream_t s) {
    if (!n) return CF_OK;
    hipLaunchKernelGGL(kb_arm, dim3((n + 255) / 256), dim3(256), 0, s, t, list, n);
    return check_launch_b();
}

int launch_batch_count_busy(const BatchTables& t, hipStream_t s) {
    if (!t.n_active) return CF_OK;
    hipLaunchKernelGGL(kb_count_busy, dim3((t.n_active + 255) / 256), dim3(256), 0, s, t);
    return check_launch_b();
}

size_t batch_tile_lds_bytes(uint32_t tile_threads) {
    switch (tile_threads) {
        case 1024: return TileLds<1024>::kBytes;
        case 512: return TileLds<512>::kBytes;
        case 128: return TileLds<128>::kBytes;
        default: return TileLds<256>::kBytes;
    }
}

int launch_batch_compact(const BatchTables& t, const CompactOut& o, uint32_t fill_grid, hipStream_t s) {
    if (!t.n_active) return CF_OK;
    hipLaunchKernelGGL(kb_compact_scan, dim3(1), dim3(kScanThreads), 0, s, t, o);
    CF_TRY(check_launch_b());
    const uint32_t g = fill_grid < t.n_active ? fill_grid : t.n_active;
    hipLaunchKernelGGL(kb_compact_fill, dim3(g ? g : 1), dim3(kBlock), 0, s, t, o);
    return check_launch_b();
}

int launch_batch_fill_u32(uint32_t* p, uint32_t value, uint32_t n, hipStream_t s) {
    if (!n) return CF_OK;
    hipLaunchKernelGGL(kb_fill_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, value, n);
    return check_launch_b();
}

}  // namespace cf
