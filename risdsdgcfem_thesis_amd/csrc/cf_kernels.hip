// corofem device kernels for gfx950 (MI355X).
//
// Element data are SoA planes in tile order; node vectors are flat FP32
// (node i at [3i,3i+2]). The scatter-add of the reference's element loops
// (C_f_e / C_df_x, ms/fem/TetrahedronElementSet3.cpp:1053-1096,1761-1836) is
// replaced by an atomics-free two-stage gather:
//   k_tile_forces : one workgroup per tile of elements; tile nodes staged in
//                   LDS, per-element nodal forces staged in LDS, summed per
//                   tile node in a fixed slot order -> one partial per (tile,node)
//   k_gather      : per node, sums its tile partials in tile order and fuses the
//                   per-node terms of the caller (mass term + KNC projection +
//                   d.q for CG, or RHS assembly for QIE-v).
// Reductions accumulate in FP64 with a fixed tree, so every result is
// bit-reproducible run to run.
#include "cf_device.h"
#include <cmath>

namespace cf {

// ============================================================ precompute
// EndEdition per element (ms/fem/TetrahedronElementSet3.cpp:2120-2161):
// V = det(Dm)/6, Dm^-1 = Inverse(Dm), Dm = cols(r1-r0, r2-r0, r3-r0).
__global__ __launch_bounds__(kBlock) void k_precompute(uint32_t T, const uint32_t* __restrict__ conn,
                                                       const float* __restrict__ ref, float* __restrict__ dinv,
                                                       float* __restrict__ vol) {
#pragma clang fp contract(off)
    const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= T) return;
    float r[4][3];
    for (int j = 0; j < 4; j++) {
        const uint32_t g = conn[j * (size_t)T + e];
        r[j][0] = ref[3 * (size_t)g]; r[j][1] = ref[3 * (size_t)g + 1]; r[j][2] = ref[3 * (size_t)g + 2];
    }
    float D[9];
    for (int i = 0; i < 3; i++) {
        D[3 * i + 0] = r[1][i] - r[0][i];
        D[3 * i + 1] = r[2][i] - r[0][i];
        D[3 * i + 2] = r[3][i] - r[0][i];
    }
    const float det = refmath::det3(D);
    vol[e] = det / 6.0f;
    float inv[9];
    refmath::inv3(D, det, inv);
    for (int k = 0; k < 9; k++) dinv[k * (size_t)T + e] = inv[k];
}

// ============================================================ BeginEvaluation
__global__ __launch_bounds__(kBlock, 8) void k_begin_eval(uint32_t T, const uint32_t* __restrict__ conn,
                                                       const float* __restrict__ x, const float* __restrict__ dinv,
                                                       int rm, float thr, float* __restrict__ R,
                                                       float* __restrict__ detF, uint32_t* blk_count,
                                                       double* blk_sum, float* blk_min, LcmOut lcm,
                                                       DegenArgs dg) {
    __shared__ double sh[kBlock / 64];
    __shared__ float shm[kBlock / 64];
    __shared__ uint32_t shc[kBlock / 64], shf[kBlock / 64];
    begin_eval_block(T, conn, x, dinv, rm, thr, R, detF, blk_count, blk_sum, blk_min, blockIdx.x, gridDim.x, sh,
                     shm, shc, shf, lcm, dg);
}

// Tile kernel: one element and at most one tile node per thread. A workgroup
// walks its tiles; while tile t is computed from LDS, the gathers of tile t+1
// (node vectors, rest positions, element data) and the node ids of tile t+2
// are in flight (LDS-only barriers do not wait for them). With a grid of one
// workgroup per tile each workgroup handles a single tile. Workgroups are
// dealt round-robin to the 8 XCDs; each XCD owns a contiguous range of tiles,
// so tiles that share halo nodes share an L2.
template <int MODE, int TB, bool PIPE>
__global__ __launch_bounds__(TB) void k_tile_forces(TileArgs a) {
    extern __shared__ float smem_f[];
    const TileLds<TB> L(smem_f);
    const Slots& s_v = L.v;  // tile node vectors
    const Slots& s_r = L.r;  // tile node rest positions
    const Slots& s_F = L.F;  // 4 force slots per element
    uint32_t* s_sp = L.sp;   // slot-run starts

    const uint32_t tid = threadIdx.x;
    // Every argument the prologue and the first loads use, read in one batch
    // of scalar loads (the compiler would otherwise fetch them in three
    // dependent rounds, as each branch first needs them).
    asm volatile("" ::"s"(a.T), "s"(a.num_tiles), "s"(a.reverse), "s"(a.uniform_te), "s"(a.skip),
                 "s"(a.tile_node_ptr), "s"(a.tile_elem_ptr), "s"(a.tile_nodes), "s"(a.lv), "s"(a.rk), "s"(a.R),
                 "s"(a.slot_ptr), "s"(a.part_dst), "s"(a.ref_tile), "s"(a.in));
    const uint32_t gx = gridDim.x / 8, xcd = blockIdx.x % 8, j = blockIdx.x / 8;
    const uint32_t per = (a.num_tiles + 7) / 8;
    const uint32_t lo = xcd * per;
    const uint32_t hi = min(a.num_tiles, lo + per);
    uint32_t t = lo + j;
    if (t >= hi) return;
    if (!PIPE && a.reverse) t = hi - 1 - j;

    // The skip flag and the tile's descriptor are loaded together: the empty
    // asm consumes the descriptor before the branch, so the compiler cannot
    // sink its loads behind the skip test (one scalar round trip, not two).
    TileDesc dt = tile_desc(a, t);
    const int32_t sk = a.skip ? *a.skip : 0;
    asm volatile("" ::"s"(dt.n0), "s"(dt.nn), "s"(dt.e0), "s"(dt.ne));
    if (sk) return;
    TileRegs g;
    load_tile_regs_cold(a, dt, tid, g);
    uint32_t t1 = t + gx;
    TileDesc d1 = dt;
    uint32_t gid1 = 0;
    if (PIPE && t1 < hi) {
        d1 = tile_desc(a, t1);
        gid1 = (tid < d1.nn) ? ld32(a.tile_nodes, d1.n0 + tid) : 0u;
    }
    while (true) {
        // stage tile t into LDS
        if (tid < dt.nn) {
            s_v.x[tid] = g.v[0]; s_v.y[tid] = g.v[1]; s_v.z[tid] = g.v[2];
            s_r.x[tid] = g.r[0]; s_r.y[tid] = g.r[1]; s_r.z[tid] = g.r[2];
        }
        if (tid <= dt.nn) s_sp[tid] = g.sp;
        const uint32_t cpd = g.pd;
        const uint32_t cl[4] = {g.l[0], g.l[1], g.l[2], g.l[3]};
        const uint32_t cf[4] = {g.f[0], g.f[1], g.f[2], g.f[3]};
        float cR[9];
#pragma unroll
        for (int k = 0; k < 9; k++) cR[k] = g.R[k];
        lds_barrier();
        // prefetch: gathers of tile t1, node ids of tile t2
        const uint32_t t2 = t1 + gx;
        TileDesc d2 = d1;
        uint32_t gid2 = 0;
        if (PIPE && t1 < hi) load_tile_regs(a, d1, gid1, tid, g);
        if (PIPE && t2 < hi) {
            d2 = tile_desc(a, t2);
            gid2 = (tid < d2.nn) ? ld32(a.tile_nodes, d2.n0 + tid) : 0u;
        }
        // element forces of tile t into the force slots
        if (tid < dt.ne) {
            uint32_t fs[4];
#pragma unroll
            for (int q = 0; q < 4; q++) fs[q] = s_sp[cl[q]] + cf[q];
            element_force<MODE>(s_v, s_r, s_F, cl, fs, cR, a.m);
        }
        lds_barrier();
        // per tile node: its contiguous slot run (ascending element order) -> one partial
        if (tid < dt.nn) {
            float fx, fy, fz;
            sum_slot_run(s_F, s_sp[tid], s_sp[tid + 1], fx, fy, fz);
            float* o = a.partials + 3 * (size_t)cpd;
            o[0] = fx; o[1] = fy; o[2] = fz;
        }
        if (!PIPE || t1 >= hi) break;
        lds_barrier();
        t = t1; dt = d1;
        t1 = t2; d1 = d2; gid1 = gid2;
    }
}

// LCM / LCMH element stage, and C_WRP's K*x with the exact rotation
// differential: one tile per workgroup, same tile mapping.
template <int MODE, int TB, int KIND>
__global__ __launch_bounds__(TB) void k_tile_forces_kind(TileArgs a) {
    if (a.skip && *a.skip) return;
    extern __shared__ float smem_f[];
    const TileLds<TB> L(smem_f);
    const uint32_t xcd = blockIdx.x % 8, j = blockIdx.x / 8;
    const uint32_t per = (a.num_tiles + 7) / 8;
    const uint32_t lo = xcd * per;
    const uint32_t hi = min(a.num_tiles, lo + per);
    if (lo + j >= hi) return;
    const uint32_t t = a.reverse ? hi - 1 - j : lo + j;
    tile_block<MODE, KIND>(a, t, L.v, L.r, L.F, L.sp);
}

// BeginEvaluation fused with the front of a QIE-v step (launch_qie_front):
// one tile per workgroup, same tile mapping as the tile kernels. The tile's
// positions and rest positions, staged once, give each element's F (Dm^-1
// rebuilt from the rest positions as k_precompute builds it), det F and
// rotation, which stays in registers for f_e(x) and, after y replaces x in
// LDS, for K y. R, det F, the DoC and the two partial sets carry the same
// bits as BeginEvaluation, the f_e tile pass and the CG start's K y tile pass
// run one after another. Every global store goes through a uniform base and
// a 32-bit offset (st32), so that no per-lane 64-bit address has to live
// across the rotation's branches and the degenerate branch's call.
struct FrontArgs {
    TileArgs a;  // a.in = x, a.partials = f_e
    NodeTerms yt;
    const uint32_t* conn;
    const float* dinv;
    FrontOut fo;
    uint32_t num_nodes, num_parts;  // byte sizes below 4 GiB (checked on the host)
};

// Buffer accesses for the fused front: a resource over n bytes from a uniform
// base, a 32-bit byte offset per lane, and accesses past n bytes dropped by
// the hardware. No per-lane 64-bit address lives across the rotation's
// branches and the degenerate branch's call (a 64-bit store address kept
// there was once compiled into a register pair that held other values on
// one path). aux 2: non-temporal.
using BufRsrc = __amdgpu_buffer_rsrc_t;
typedef uint32_t Bu2 __attribute__((ext_vector_type(2)));
__device__ inline BufRsrc buf(const void* p, uint32_t nbytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, int(nbytes), 0x00020000);
}
__device__ inline float bld_f(BufRsrc r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ inline uint32_t bld_u8(BufRsrc r, uint32_t off) { return __builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0); }
__device__ inline void bst_f(BufRsrc r, uint32_t off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, off, 0, 0);
}
__device__ inline void bst_f_nt(BufRsrc r, uint32_t off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, off, 0, 2);
}
__device__ inline void bst_u(BufRsrc r, uint32_t off, uint32_t v) { __builtin_amdgcn_raw_buffer_store_b32(v, r, off, 0, 0); }
__device__ inline void bst_u8(BufRsrc r, uint32_t off, uint32_t v) {
    __builtin_amdgcn_raw_buffer_store_b8(uint8_t(v), r, off, 0, 0);
}
__device__ inline void bst_d(BufRsrc r, uint32_t off, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(Bu2, v), r, off, 0, 0);
}

// The CG start y of Step_QuasiImplicitEuler_v (:1209-1295) at node n, the
// right-hand side gather's y (gather_node<kGatherRhs>) in the same
// operations: v with the warm start, else 0; a KNC node takes the constraint
// velocity (target - x) / dt, rescaled to at most 50 and zero below 0.001.
__device__ inline void front_start_y(uint32_t n, const FrontArgs& f, float y[3]) {
#pragma clang fp contract(off)
    const NodeTerms& t = f.yt;
    const uint32_t N = f.num_nodes;
    const uint32_t o = 12 * n;
    if (t.knc_mask && bld_u8(buf(t.knc_mask, N), n)) {
        const BufRsrc rt = buf(t.knc_target, 12 * N), rx = buf(t.x, 12 * N);
        float vk[3];
#pragma unroll
        for (int c = 0; c < 3; c++) vk[c] = (bld_f(rt, o + 4 * c) - bld_f(rx, o + 4 * c)) * t.inv_dt;
        const float nk = sqrtf(vk[0] * vk[0] + vk[1] * vk[1] + vk[2] * vk[2]);
        if (nk > 0.001f) {
            const float inv = 1.0f / nk, s = fminf(nk, 50.0f);
#pragma unroll
            for (int c = 0; c < 3; c++) vk[c] = s * (vk[c] * inv);
        } else {
            vk[0] = vk[1] = vk[2] = 0.f;
        }
#pragma unroll
        for (int c = 0; c < 3; c++) y[c] = vk[c];
    } else {
        const BufRsrc rv = buf(t.v, 12 * N);
#pragma unroll
        for (int c = 0; c < 3; c++) y[c] = t.warm_start ? bld_f(rv, o + 4 * c) : 0.0f;
    }
}

// BeginEvaluation's rotation and degeneration update of element e, the
// sequence of begin_eval_block: the rotation of the method, then
// UpdateDegeneration (:477-532), which clears an undegenerate element's DoC
// and, for a degenerate one, updates the DoC and under DAPD writes R itself.
// R is stored non-temporally (the tile stage reads it in the next pass) and
// returned in Rm; fb counts DAPD's R = Id fallbacks.
template <int TAG>
__device__ inline void front_rotation(const FrontArgs& f, uint32_t e, const float F[9], float d, float Rm[9],
                                      uint32_t& fb) {
#pragma clang fp contract(off)
    const FrontOut& fo = f.fo;
    const uint32_t T = f.a.T;
    const BufRsrc rR = buf(fo.R, 36 * T);
#pragma unroll
    for (int k = 0; k < 9; k++) Rm[k] = (k % 4 == 0) ? 1.f : 0.f;
    if (fo.rm == CF_RM_MSVD) {
        refmath::svd_rotation(F, Rm);
    } else if (fo.rm == CF_RM_DAPD) {
        if (d > fo.thr) refmath::polar_newton(F, Rm);
    }
    bool r_stored = false;
    if (fo.dg.doc) {
        const BufRsrc rd = buf(fo.dg.doc, T);
        if (d > fo.thr) {
            if (bld_u8(rd, e) != kDocInvalid) bst_u8(rd, e, kDocInvalid);
        } else {
            const float* x = f.a.in;
            const uint32_t fl = degenerate_element<TAG>(e, T, f.conn, x, fo.dg.prev_x ? fo.dg.prev_x : x, fo.dg.doc,
                                                        fo.dg.vol, f.dinv, fo.rm, fo.thr, d, fo.dg.l_factor,
                                                        fo.dg.nl_factor, fo.dg.nl_exponent, fo.R);
            if (fo.rm == CF_RM_DAPD) {
                fb += fl;
                r_stored = true;
                for (int k = 0; k < 9; k++) Rm[k] = bld_f(rR, 4 * (k * T + e));
            }
        }
    } else if (fo.rm == CF_RM_DAPD && !(d > fo.thr)) {
        fb++;  // no DoC tracking: R = Id
    }
    if (!r_stored)
        for (int k = 0; k < 9; k++) bst_f_nt(rR, 4 * (k * T + e), Rm[k]);
}

template <int TB, int W>
__global__ __launch_bounds__(TB, W) void k_qie_front(FrontArgs f) {
#pragma clang fp contract(off)
    extern __shared__ float smem_f[];
    __shared__ double s_sum[TB / 64];
    __shared__ float s_min[TB / 64];
    __shared__ uint32_t s_cnt[TB / 64], s_fb[TB / 64];
    const TileLds<TB> L(smem_f);
    const TileArgs& a = f.a;
    const FrontOut& fo = f.fo;
    const uint32_t tid = threadIdx.x;
    const uint32_t xcd = blockIdx.x % 8, j = blockIdx.x / 8;
    const uint32_t per = (a.num_tiles + 7) / 8;
    const uint32_t lo = xcd * per;
    const uint32_t hi = min(a.num_tiles, lo + per);
    const uint32_t t = lo + j;
    if (t >= hi) return;
    const TileDesc dt = tile_desc(a, t);
    // the cold tile loads without R; lanes past the tile load valid duplicates
    const uint32_t tn = min(tid, dt.nn - 1);
    const uint32_t gid = ld32(a.tile_nodes, dt.n0 + tn);
    const uint32_t e = dt.e0 + min(tid, dt.ne - 1);
    const uint32_t lv = ld32_nt(a.lv, e), rk = ld32_nt(a.rk, e);
    const uint32_t sp = ld32(a.slot_ptr, dt.n0 + min(tid, dt.nn)) - 4 * dt.e0;
    const uint32_t pd = ld32(a.part_dst, dt.n0 + tn);
    float rp[3] = {0.f, 0.f, 0.f}, xp[3] = {0.f, 0.f, 0.f};
    if (a.ref_tile) {
        const uint32_t q = 3 * (dt.n0 + tn);
        rp[0] = ld32(a.ref_tile, q + 0); rp[1] = ld32(a.ref_tile, q + 1); rp[2] = ld32(a.ref_tile, q + 2);
    }
    if (tid < dt.nn) {
        xp[0] = ld32(a.in, 3 * gid + 0); xp[1] = ld32(a.in, 3 * gid + 1); xp[2] = ld32(a.in, 3 * gid + 2);
        if (!a.ref_tile) {
            const BufRsrc rr = buf(a.ref, 12 * f.num_nodes);
            rp[0] = bld_f(rr, 12 * gid); rp[1] = bld_f(rr, 12 * gid + 4); rp[2] = bld_f(rr, 12 * gid + 8);
        }
        L.v.x[tid] = xp[0]; L.v.y[tid] = xp[1]; L.v.z[tid] = xp[2];
        L.r.x[tid] = rp[0]; L.r.y[tid] = rp[1]; L.r.z[tid] = rp[2];
    }
    if (tid <= dt.nn) L.sp[tid] = sp;
    uint32_t cl[4], fs[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int q = 0; q < 4; q++) cl[q] = (lv >> (8 * q)) & 255u;
    lds_barrier();
    // BeginEvaluation of the thread's element
    float Rm[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
    uint32_t cnt = 0, fb = 0;
    double sdet = 0.0;
    float mdet = 3.0e38f;
    if (tid < dt.ne) {
#pragma unroll
        for (int q = 0; q < 4; q++) fs[q] = L.sp[cl[q]] + ((rk >> (8 * q)) & 255u);
        float xs[4][3], rs[4][3];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            xs[q][0] = L.v.x[cl[q]]; xs[q][1] = L.v.y[cl[q]]; xs[q][2] = L.v.z[cl[q]];
            rs[q][0] = L.r.x[cl[q]]; rs[q][1] = L.r.y[cl[q]]; rs[q][2] = L.r.z[cl[q]];
        }
        float Ds[9], Dm[9], G[9], F[9];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            Ds[3 * i + 0] = xs[1][i] - xs[0][i];
            Ds[3 * i + 1] = xs[2][i] - xs[0][i];
            Ds[3 * i + 2] = xs[3][i] - xs[0][i];
            Dm[3 * i + 0] = rs[1][i] - rs[0][i];
            Dm[3 * i + 1] = rs[2][i] - rs[0][i];
            Dm[3 * i + 2] = rs[3][i] - rs[0][i];
        }
        refmath::inv3(Dm, refmath::det3(Dm), G);
        refmath::mul3(Ds, G, F);
        const float d = refmath::det3(F);
        bst_f(buf(fo.detF, 4 * a.T), 4 * e, d);
        if (d <= 0.0f) { cnt++; sdet += d; }
        mdet = fminf(mdet, d);
        front_rotation<16 * TB + W>(f, e, F, d, Rm, fb);  // a callee copy per instantiation
        element_force<kModeFe>(L.v, L.r, L.F, cl, fs, Rm, a.m);
    }
    // the tile's statistics: wave sums, then the waves in order
    sdet = wave_sum(sdet);
    for (int o = 32; o > 0; o >>= 1) {
        mdet = fminf(mdet, __shfl_xor(mdet, o, 64));
        cnt += __shfl_xor(cnt, o, 64);
        fb += __shfl_xor(fb, o, 64);
    }
    if ((tid & 63) == 0) {
        const uint32_t w = tid >> 6;
        s_sum[w] = sdet; s_min[w] = mdet; s_cnt[w] = cnt; s_fb[w] = fb;
    }
    lds_barrier();
    if (tid == 0) {
        double ss = s_sum[0];
        float mm = s_min[0];
        uint32_t cc = s_cnt[0], ff = s_fb[0];
        for (int w = 1; w < TB / 64; w++) { ss += s_sum[w]; mm = fminf(mm, s_min[w]); cc += s_cnt[w]; ff += s_fb[w]; }
        const BufRsrc rc = buf(fo.st_count, 8 * a.num_tiles);
        bst_u(rc, 4 * t, cc);
        bst_u(rc, 4 * (a.num_tiles + t), ff);
        bst_d(buf(fo.st_sum, 8 * a.num_tiles), 8 * t, ss);
        bst_f(buf(fo.st_min, 4 * a.num_tiles), 4 * t, mm);
    }
    // f_e per tile node
    if (tid < dt.nn) {
        float fx, fy, fz;
        sum_slot_run(L.F, L.sp[tid], L.sp[tid + 1], fx, fy, fz);
        const BufRsrc rp = buf(a.partials, 12 * f.num_parts);
        bst_f(rp, 12 * pd, fx); bst_f(rp, 12 * pd + 4, fy); bst_f(rp, 12 * pd + 8, fz);
    }
    lds_barrier();
    // K y with the same rotations (y read now: nothing held across the rotation)
    if (tid < dt.nn) {
        float y[3];
        front_start_y(gid, f, y);
        L.v.x[tid] = y[0]; L.v.y[tid] = y[1]; L.v.z[tid] = y[2];
    }
    lds_barrier();
    if (tid < dt.ne) element_force<kModeDfx>(L.v, L.r, L.F, cl, fs, Rm, a.m);
    lds_barrier();
    if (tid < dt.nn) {
        float fx, fy, fz;
        sum_slot_run(L.F, L.sp[tid], L.sp[tid + 1], fx, fy, fz);
        const BufRsrc rp = buf(fo.partials2, 12 * f.num_parts);
        bst_f(rp, 12 * pd, fx); bst_f(rp, 12 * pd + 4, fy); bst_f(rp, 12 * pd + 8, fz);
    }
}

// C_f_d element stage (TetrahedronElementSet3.cpp:1098-1150): the C_WRP
// stiffness force of the velocities (a.in) scaled by c1, minus the
// per-vertex mass term R (c0 m_j R^T v_j); tile node masses staged in LDS.
template <int TB>
__global__ __launch_bounds__(TB) void k_tile_damping(TileArgs a, DampArgs d) {
    extern __shared__ float smem_f[];
    const TileLds<TB, 1> L(smem_f);
    const uint32_t xcd = blockIdx.x % 8, j = blockIdx.x / 8;
    const uint32_t per = (a.num_tiles + 7) / 8;
    const uint32_t lo = xcd * per;
    const uint32_t hi = min(a.num_tiles, lo + per);
    const uint32_t t = lo + j;
    if (t >= hi) return;
    const uint32_t tid = threadIdx.x;
    const TileDesc dt = tile_desc(a, t);
    TileRegs g;
    load_tile_regs_cold(a, dt, tid, g);
    if (tid < dt.nn) {
        const uint32_t gid = ld32(a.tile_nodes, dt.n0 + tid);
        L.v.x[tid] = g.v[0]; L.v.y[tid] = g.v[1]; L.v.z[tid] = g.v[2];
        L.r.x[tid] = g.r[0]; L.r.y[tid] = g.r[1]; L.r.z[tid] = g.r[2];
        L.extra[tid] = d.c0 * (1.0f / ld32(d.inv_mass, gid));  // c0 mal::Rcp(inv_mass) (:1126)
    }
    if (tid <= dt.nn) L.sp[tid] = g.sp;
    lds_barrier();
    if (tid < dt.ne) {
        uint32_t fs[4];
#pragma unroll
        for (int q = 0; q < 4; q++) fs[q] = L.sp[g.l[q]] + g.f[q];
        element_force_fd(L.v, L.r, L.F, L.extra, g.l, fs, g.R, a.m, d.c1);
    }
    lds_barrier();
    if (tid < dt.nn) {
        float fx, fy, fz;
        sum_slot_run(L.F, L.sp[tid], L.sp[tid + 1], fx, fy, fz);
        float* o = a.partials + 3 * (size_t)g.pd;
        o[0] = fx; o[1] = fy; o[2] = fz;
    }
}

// ============================================================ node gather
struct GatherArgs {
    uint32_t N;
    const uint32_t* node_part_ptr;
    const float* partials;
    NodeTerms t;
    const int32_t* skip;
};

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_gather(GatherArgs a) {
    __shared__ double sh[kBlock / 64];
    if (a.skip && *a.skip) return;
    gather_block<MODE>(a.N, a.node_part_ptr, a.partials, a.t, blockIdx.x, gridDim.x, sh);
}

// The matvec gather in chunks of npt kBlock nodes, one chunk per workgroup;
// chunk c's d.q partial in t.partials[c].
// The QIE-v right-hand side and the CG start in one pass, after the fused
// step front: over the matvec chunks (node (c npt + j) kBlock + thread), b and
// y of each node from the f_e partials in the operations of
// gather_node<kGatherRhs>, then q = S(fM m y - fK K y) from the K y partials
// (ay) and r = S(b - q), d = r in those of gather_matvec_chunk<true>, with the
// chunk's r.r partial summed in the same order. The outputs carry the bits of
// the two gathers run one after the other.
__global__ __launch_bounds__(kBlock) void k_rhs_start(GatherArgs a, const float* __restrict__ ay, uint32_t npt) {
#pragma clang fp contract(off)
    __shared__ double sh[kBlock / 64];
    const NodeTerms& t = a.t;
    double acc = 0.0;
    for (uint32_t j = 0; j < npt; j++) {
        const uint32_t n = (blockIdx.x * npt + j) * kBlock + threadIdx.x;
        if (n >= a.N) continue;
        const size_t i = 3 * (size_t)n;
        const uint32_t p0 = a.node_part_ptr[n], p1 = a.node_part_ptr[n + 1];
        // right-hand side (gather_node<kGatherRhs>)
        float fx = 0.f, fy = 0.f, fz = 0.f;
        for (uint32_t p = p0; p < p1; p++) {
            const float* pp = a.partials + 3 * (size_t)p;
            fx += pp[0]; fy += pp[1]; fz += pp[2];
        }
        const float mass = t.mass_array ? t.mass_array[n] : t.mass;
        const float f[3] = {fx, fy, fz};
        float b[3], y[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float vc = t.v[i + c];
            b[c] = (mass * t.g[c] + f[c]) * t.dt + mass * vc;
            y[c] = t.warm_start ? vc : 0.0f;
        }
        const bool knc = t.knc_mask && t.knc_mask[n];
        if (knc) {
            float vk[3];
#pragma unroll
            for (int c = 0; c < 3; c++) vk[c] = (t.knc_target[i + c] - t.x[i + c]) * t.inv_dt;
            const float nk = sqrtf(vk[0] * vk[0] + vk[1] * vk[1] + vk[2] * vk[2]);
            if (nk > 0.001f) {
                const float inv = 1.0f / nk, sc = fminf(nk, 50.0f);
#pragma unroll
                for (int c = 0; c < 3; c++) vk[c] = sc * (vk[c] * inv);
            } else {
                vk[0] = vk[1] = vk[2] = 0.f;
            }
#pragma unroll
            for (int c = 0; c < 3; c++) { b[c] = vk[c]; y[c] = vk[c]; }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) { t.out[i + c] = b[c]; t.y[i + c] = y[c]; }
        // CG start (gather_matvec_chunk<true>, cg_start_node)
        float kx = 0.f, ky = 0.f, kz = 0.f;
        for (uint32_t p = p0; p < p1; p++) {
            const float* pp = ay + 3 * (size_t)p;
            kx += pp[0]; ky += pp[1]; kz += pp[2];
        }
        const float fac = mass * t.fM;
        float q[3] = {-t.fK * kx + fac * y[0], -t.fK * ky + fac * y[1], -t.fK * kz + fac * y[2]};
        if (knc) q[0] = q[1] = q[2] = 0.f;
        double node = 0.0;  // the node's r.r, then the thread's sum (cg_start_node)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float ri = -1.0f * q[k] + b[k];
            if (knc) ri = 0.f;
            t.r_out[i + k] = ri;
            t.d_out[i + k] = ri;
            node += (double)ri * ri;
        }
        acc += node;
    }
    const double sum = block_sum(acc, sh);
    if (threadIdx.x == 0) t.partials[blockIdx.x] = sum;
}

template <bool INIT>
__global__ __launch_bounds__(kBlock) void k_gather_matvec(GatherArgs a, uint32_t npt) {
    __shared__ double sh[kBlock / 64];
    // the arguments in one batch of scalar loads, before the skip test needs
    // the first of them (as in the tile kernel)
    asm volatile("" ::"s"(a.N), "s"(a.node_part_ptr), "s"(a.partials), "s"(a.t.d), "s"(a.t.out), "s"(a.t.mass),
                 "s"(a.t.fM), "s"(a.t.fK), "s"(a.t.knc_mask), "s"(a.t.mass_array), "s"(a.t.partials), "s"(a.skip),
                 "s"(npt));
    if (a.skip && *a.skip) return;
    const double s =
        block_sum(gather_matvec_chunk<INIT>(a.N, blockIdx.x, npt, a.node_part_ptr, a.partials, a.t), sh);
    if (threadIdx.x == 0) a.t.partials[blockIdx.x] = s;
}

// ============================================================ energy
__global__ __launch_bounds__(kBlock) void k_energy(uint32_t T, const uint32_t* __restrict__ conn,
                                                   const float* __restrict__ ref,
                                                   const float* __restrict__ x, const float* __restrict__ R,
                                                   const float* __restrict__ dinv, const float* __restrict__ vol,
                                                   Material m, double* blk) {
    __shared__ double sh[kBlock / 64];
    energy_block(T, conn, ref, x, R, dinv, vol, m, blk, blockIdx.x, gridDim.x, sh);
}

// C_V of C_WRP per tile (tile_energy_block), one tile per workgroup with the
// K*x tiles' XCD mapping; part[t] is tile t's FP64 sum.
template <int TB>
__global__ __launch_bounds__(TB) void k_tile_energy(TileArgs a, double* part) {
    extern __shared__ float smem_f[];
    __shared__ double sh[TB / 64];
    const uint32_t xcd = blockIdx.x % 8, j = blockIdx.x / 8;
    const uint32_t per = (a.num_tiles + 7) / 8;
    const uint32_t lo = xcd * per;
    const uint32_t hi = min(a.num_tiles, lo + per);
    if (lo + j >= hi) return;
    const uint32_t t = lo + j;
    const double e = tile_energy_block<TB>(a, t, smem_f, sh);
    if (threadIdx.x == 0) part[t] = e;
}

// *out = the n partials summed in the one-block order (one workgroup).
__global__ __launch_bounds__(kBlock) void k_sum_partials(const double* part, int n, double* out) {
    __shared__ double sh[kBlock / 64];
    const double s = reduce_partials(part, n, sh);
    if (threadIdx.x == 0) *out = s;
}

// ============================================================ CG vector kernels
// ILSS_CG::Solve (ns/ILSS_CG.h:38-151); the projection S zeroes KNC nodes.
// Vector kernels walk n floats as float4 when every pointer is 16-byte
// aligned (V4), with the n % 4 tail done by the first threads of the grid.
__global__ __launch_bounds__(kBlock) void k_cg_init(uint32_t n, const float* __restrict__ q,
                                                    const float* __restrict__ b, float* __restrict__ r,
                                                    float* __restrict__ d, const uint8_t* __restrict__ mask,
                                                    double* rr_part) {
    __shared__ double sh[kBlock / 64];
    cg_init_block(n, q, b, r, d, mask, rr_part, blockIdx.x, gridDim.x, sh);
}

__global__ __launch_bounds__(kBlock) void k_cg_start(CgScalars* st, const double* rr_part, int nprev, float rel,
                                                     float abs_eps, uint32_t max_iter) {
    __shared__ double sh[kBlock / 64];
    const double d0 = reduce_partials(rr_part, nprev, sh);
    for (uint32_t k = threadIdx.x; k < kMaxCgIter + 2; k += kBlock) {
        st->done[k] = 0;
        st->reason[k] = 0;
        st->delta[k] = 0.0;
        st->dq[k] = 0.0;
        st->alpha[k] = 0.f;
        st->beta[k] = 0.f;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const float abs_sq = abs_eps * abs_eps;
        const float rel_sq = rel * rel;
        st->delta0 = d0;
        st->delta[0] = d0;
        st->target = (double)rel_sq * d0;
        st->abs_sq = (double)abs_sq;
        st->max_iter = max_iter;
        int reason = 0;
        if (!(d0 > (double)abs_sq)) reason = CF_CG_CONVERGED_ABS;
        else if (!(d0 > st->target)) reason = CF_CG_CONVERGED_REL;
        else if (!(0u < max_iter)) reason = CF_CG_MAX_ITER;
        st->done[0] = reason != 0;
        st->reason[0] = reason;
    }
}

__global__ __launch_bounds__(kBlock) void k_dot(uint32_t n, const float* __restrict__ a,
                                                const float* __restrict__ b, double* part) {
    __shared__ double sh[kBlock / 64];
    double acc = 0.0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) acc += (double)a[i] * b[i];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// One CG iteration's scalar and vector work after q = S A d, in two kernels
// (ILSS_CG.h:86-133). Every workgroup sums the same FP64 partials in the same
// fixed order as a one-block reduction (so all agree on alpha / beta to the
// bit), workgroup 0 records the scalars and stop tests, and the vector work is
// split into nchunk chunks laid out as a grid of nchunk blocks would be: chunk
// c's r.r partial is the same value whichever workgroup computes it.

// Default grid of the two CG step kernels: 1,024 workgroups, four per CU
// (swept 256-8,192). Each workgroup loads its first chunk's operands before
// the partial sum; loading more chunks early cost more in held registers and
// serial barriers than the latency they hid on the 5.24M-tet mesh's
// 3.26M-float vectors.
constexpr uint32_t kCgWorkgroups = 1024;

// One float4 of the two CG step kernels. The x update of iteration k moves
// from the first kernel (which then streams r and q only) to the second, which
// loads d_k anyway: x += alpha_k d_k there, before d is overwritten. The
// arithmetic of every element is the reference's (ILSS_CG.h:100-133).
struct R4 { float4 r, q; };
__device__ inline R4 load_r4(const float* r, const float* q, uint32_t i) {
    R4 v;
    v.q = reinterpret_cast<const float4*>(q)[i];
    v.r = reinterpret_cast<const float4*>(r)[i];
    return v;
}
// r -= alpha q, stored; returns the float4's r.r terms
__device__ inline double apply_r4(R4 v, float alpha, float* r, uint32_t i) {
#pragma clang fp contract(off)
    v.r.x = v.r.x - v.q.x * alpha; v.r.y = v.r.y - v.q.y * alpha;
    v.r.z = v.r.z - v.q.z * alpha; v.r.w = v.r.w - v.q.w * alpha;
    reinterpret_cast<float4*>(r)[i] = v.r;
    return (double)v.r.x * v.r.x + (double)v.r.y * v.r.y + (double)v.r.z * v.r.z + (double)v.r.w * v.r.w;
}
// scalar elements first, first + stride, ... below n
__device__ inline double tail_r(uint32_t n, float alpha, float* r, const float* q, uint32_t first, uint32_t stride,
                                double acc) {
#pragma clang fp contract(off)
    for (uint32_t i = first; i < n; i += stride) {
        const float ri = r[i] - q[i] * alpha;
        r[i] = ri;
        acc += (double)ri * ri;
    }
    return acc;
}

struct Dx4 { float4 d, r, x; };
__device__ inline Dx4 load_dx4(const float* d, const float* r, const float* x, uint32_t i) {
    Dx4 v;
    v.d = reinterpret_cast<const float4*>(d)[i];
    v.x = reinterpret_cast<const float4*>(x)[i];
    v.r = reinterpret_cast<const float4*>(r)[i];
    return v;
}
// x += alpha d ; then, unless the solve stopped, d = S(beta d + r)
__device__ inline void apply_dx4(Dx4 v, float alpha, float beta, bool upd_d, float* x, float* d,
                                 const uint8_t* mask, uint32_t i) {
#pragma clang fp contract(off)
    float4 xi = v.x;
    xi.x = xi.x + v.d.x * alpha; xi.y = xi.y + v.d.y * alpha;
    xi.z = xi.z + v.d.z * alpha; xi.w = xi.w + v.d.w * alpha;
    reinterpret_cast<float4*>(x)[i] = xi;
    if (upd_d) apply_d4(D4{v.d, v.r}, beta, d, mask, i);
}
__device__ inline void tail_dx(uint32_t n, float alpha, float beta, bool upd_d, float* x, float* d, const float* r,
                               const uint8_t* mask, uint32_t first, uint32_t stride) {
#pragma clang fp contract(off)
    for (uint32_t i = first; i < n; i += stride) {
        const float di = d[i];
        x[i] = x[i] + di * alpha;
        if (upd_d) {
            float dn = beta * di + r[i];
            if (mask && mask[i / 3]) dn = 0.f;
            d[i] = dn;
        }
    }
}

// Chunk c of the two CG step kernels: `rows` float4s per thread (cg_rows),
// float4 (c rows + j) kBlock + thread for j < rows (each j a coalesced row),
// plus the n % 4 tail in chunk 0; without float4 access, the floats
// c kBlock + thread + m nchunk kBlock. Chunk c's r.r partial is the same value
// whichever workgroup computes it, so results do not depend on the grid. A
// workgroup loads the first min(rows, kCgVpt) rows of its first chunk before
// it sums the previous kernel's partials (PRE: chunk c's operands are in v).
struct R4Rows { R4 v[kCgVpt]; };
struct Dx4Rows { Dx4 v[kCgVpt]; };

template <bool V4, bool PRE>
__device__ inline double r_chunk(uint32_t n, uint32_t c, uint32_t nchunk, uint32_t rows, float alpha, float* r,
                                 const float* q, const R4Rows& v) {
    const uint32_t tid = threadIdx.x;
    double acc = 0.0;
    if (V4) {
        const uint32_t n4 = n / 4;
#pragma unroll
        for (uint32_t j = 0; j < kCgVpt; j++) {
            const uint32_t i = (c * rows + j) * kBlock + tid;
            if (j < rows && i < n4) {
                R4 w;
                if (PRE) w = v.v[j];
                else w = load_r4(r, q, i);
                acc += apply_r4(w, alpha, r, i);
            }
        }
        for (uint32_t j = kCgVpt; j < rows; j++) {
            const uint32_t i = (c * rows + j) * kBlock + tid;
            if (i < n4) acc += apply_r4(load_r4(r, q, i), alpha, r, i);
        }
        if (c == 0) acc = tail_r(n, alpha, r, q, 4 * n4 + tid, kBlock, acc);
    } else {
        acc = tail_r(n, alpha, r, q, c * kBlock + tid, nchunk * kBlock, acc);
    }
    return acc;
}
template <bool V4, bool PRE>
__device__ inline void dx_chunk(uint32_t n, uint32_t c, uint32_t nchunk, uint32_t rows, float alpha, float beta,
                                bool upd_d, float* x, float* d, const float* r, const uint8_t* mask,
                                const Dx4Rows& v) {
    const uint32_t tid = threadIdx.x;
    if (V4) {
        const uint32_t n4 = n / 4;
#pragma unroll
        for (uint32_t j = 0; j < kCgVpt; j++) {
            const uint32_t i = (c * rows + j) * kBlock + tid;
            if (j < rows && i < n4) {
                Dx4 w;
                if (PRE) w = v.v[j];
                else w = load_dx4(d, r, x, i);
                apply_dx4(w, alpha, beta, upd_d, x, d, mask, i);
            }
        }
        for (uint32_t j = kCgVpt; j < rows; j++) {
            const uint32_t i = (c * rows + j) * kBlock + tid;
            if (i < n4) apply_dx4(load_dx4(d, r, x, i), alpha, beta, upd_d, x, d, mask, i);
        }
        if (c == 0) tail_dx(n, alpha, beta, upd_d, x, d, r, mask, 4 * n4 + tid, kBlock);
    } else {
        tail_dx(n, alpha, beta, upd_d, x, d, r, mask, c * kBlock + tid, nchunk * kBlock);
    }
}

// alpha = delta_k / d.q (break down if |d.q| < abs^2) ; r -= alpha q ;
// per-chunk partials of r.r. Workgroup g walks chunks g, g + grid, ...
template <bool V4>
__global__ __launch_bounds__(kBlock) void k_cg_step_xr(uint32_t n, uint32_t k, CgScalars* st, float* __restrict__ r,
                                                       const float* __restrict__ q, const double* dq_part, int ndq,
                                                       double* rr_part, uint32_t nchunk, uint32_t rows) {
    __shared__ double sh[kBlock / 64];
    if (st->done[k]) return;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const uint32_t c0 = blockIdx.x;
    R4Rows v;
#pragma unroll
    for (uint32_t j = 0; j < kCgVpt; j++) {
        const uint32_t i = (c0 * rows + j) * kBlock + threadIdx.x;
        v.v[j] = R4{};
        if (V4 && j < rows && blockIdx.x < nchunk && i < n / 4) v.v[j] = load_r4(r, q, i);
    }
    const double dq = reduce_partials_wide(dq_part, ndq, sh);
    if (fabs(dq) < st->abs_sq) {
        if (lead) {
            st->dq[k] = dq;
            st->reason[k + 1] = CF_CG_BREAKDOWN;
            st->done[k + 1] = 2;
        }
        return;
    }
    const float alpha = (float)(st->delta[k] / dq);
    if (lead) {
        st->dq[k] = dq;
        st->alpha[k] = alpha;
    }
    if (c0 < nchunk) {
        const double s = block_sum(r_chunk<V4, true>(n, c0, nchunk, rows, alpha, r, q, v), sh);
        if (threadIdx.x == 0) rr_part[c0] = s;
    }
    for (uint32_t c = c0 + gridDim.x; c < nchunk; c += gridDim.x) {
        const double s = block_sum(r_chunk<V4, false>(n, c, nchunk, rows, alpha, r, q, v), sh);
        if (threadIdx.x == 0) rr_part[c] = s;
    }
}

// delta_{k+1} = r.r ; beta ; stop test ; x += alpha_k d ; d = S(beta d + r)
// unless the solve stopped. host_done, when not null, receives done[k + 1].
template <bool V4>
__global__ __launch_bounds__(kBlock) void k_cg_step_d(uint32_t n, uint32_t k, CgScalars* st, float* __restrict__ x,
                                                      float* __restrict__ d, const float* __restrict__ r,
                                                      const uint8_t* __restrict__ mask, const double* rr_part,
                                                      uint32_t nchunk, uint32_t rows, int32_t* host_done) {
    __shared__ double sh[kBlock / 64];
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (st->done[k]) {
        // keep every later iteration switched off once the solve has ended
        if (lead) {
            st->done[k + 1] = 1;
            if (host_done) host_done[k + 1] = 1;
        }
        return;
    }
    if (st->reason[k + 1] == CF_CG_BREAKDOWN) {  // done[k + 1] is set; x keeps its value
        if (lead && host_done) host_done[k + 1] = 2;
        return;
    }
    const uint32_t c0 = blockIdx.x;
    Dx4Rows v;
#pragma unroll
    for (uint32_t j = 0; j < kCgVpt; j++) {
        const uint32_t i = (c0 * rows + j) * kBlock + threadIdx.x;
        v.v[j] = Dx4{};
        if (V4 && j < rows && blockIdx.x < nchunk && i < n / 4) v.v[j] = load_dx4(d, r, x, i);
    }
    const double dn = reduce_partials_wide(rr_part, (int)nchunk, sh);
    const uint32_t k1 = k + 1;
    int reason = 0;
    if (!(dn > st->abs_sq)) reason = CF_CG_CONVERGED_ABS;
    else if (!(dn > st->target)) reason = CF_CG_CONVERGED_REL;
    else if (!(k1 < st->max_iter)) reason = CF_CG_MAX_ITER;
    const float beta = (float)(dn / st->delta[k]);
    const float alpha = st->alpha[k];
    if (lead) {
        st->delta[k1] = dn;
        st->beta[k] = beta;
        st->reason[k1] = reason;
        st->done[k1] = reason != 0;
        if (host_done) host_done[k1] = reason != 0;
    }
    if (c0 < nchunk) dx_chunk<V4, true>(n, c0, nchunk, rows, alpha, beta, reason == 0, x, d, r, mask, v);
    for (uint32_t c = c0 + gridDim.x; c < nchunk; c += gridDim.x)
        dx_chunk<V4, false>(n, c, nchunk, rows, alpha, beta, reason == 0, x, d, r, mask, v);
}

__global__ __launch_bounds__(kBlock) void k_integrate(uint32_t n, float* __restrict__ x, float* __restrict__ v,
                                                      const float* __restrict__ y, float dt, float drag) {
    integrate_block(n, x, v, y, dt, drag, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(kBlock) void k_nr_update(uint32_t n, float* __restrict__ xk, float* __restrict__ vk,
                                                      const float* __restrict__ x, const float* __restrict__ y,
                                                      float inv_dt, double* part) {
    __shared__ double sh[kBlock / 64];
    nr_update_block(n, xk, vk, x, y, inv_dt, part, blockIdx.x, gridDim.x, sh);
}

__global__ __launch_bounds__(kBlock) void k_nr_finish(uint32_t n, float* __restrict__ x, float* __restrict__ v,
                                                      const float* __restrict__ xk, const float* __restrict__ vk,
                                                      float dt, float drag) {
    nr_finish_block(n, x, v, xk, vk, dt, drag, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(kBlock) void k_zero_mask(uint32_t nn, float* __restrict__ w,
                                                      const uint8_t* __restrict__ mask) {
    for (uint32_t n = blockIdx.x * kBlock + threadIdx.x; n < nn; n += gridDim.x * kBlock)
        if (mask[n]) { w[3 * (size_t)n] = 0.f; w[3 * (size_t)n + 1] = 0.f; w[3 * (size_t)n + 2] = 0.f; }
}

// ============================================================ launchers
int vec_blocks(uint32_t n) {
    uint32_t b = (n + 4 * kBlock - 1) / (4 * kBlock);
    if (b < 1) b = 1;
    if (b > (uint32_t)kMaxPartials) b = kMaxPartials;
    return (int)b;
}

int node_blocks(uint32_t n) {
    uint32_t b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > (uint32_t)kMaxPartials) b = kMaxPartials;
    return (int)b;
}

static bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

static int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("kernel launch failed: %s", hipGetErrorString(e));
        return CF_ERR_HIP;
    }
    return CF_OK;
}

int launch_precompute(const TetDevice& t, hipStream_t s) {
    if (t.T == 0) return CF_OK;
    hipLaunchKernelGGL(k_precompute, dim3((t.T + kBlock - 1) / kBlock), dim3(kBlock), 0, s, t.T, t.conn, t.ref,
                       t.dinv, t.vol);
    return check_launch();
}

int launch_begin_eval(const TetDevice& t, const float* x, int rm, float thr, float* R, float* detF,
                      uint32_t* blk_count, double* blk_sum, float* blk_min, int nblk, hipStream_t s,
                      const LcmOut& lcm, const DegenArgs& dg) {
    hipLaunchKernelGGL(k_begin_eval, dim3(nblk), dim3(kBlock), 0, s, t.T, t.conn, x, t.dinv, rm, thr, R, detF,
                       blk_count, blk_sum, blk_min, lcm, dg);
    return check_launch();
}


int launch_qie_front(const TetDevice& t, const float* x, Material m, const FrontOut& fo, const NodeTerms& yt,
                     hipStream_t s) {
    if (t.num_tiles == 0) return CF_OK;
    FrontArgs f;
    TileArgs& a = f.a;
    a.T = t.T; a.num_tiles = t.num_tiles; a.max_te = t.max_tile_elems; a.max_tn = t.max_tile_nodes;
    a.tile_elem_ptr = t.tile_elem_ptr; a.tile_node_ptr = t.tile_node_ptr; a.tile_nodes = t.tile_nodes;
    a.part_dst = t.part_dst;
    a.slot_ptr = t.slot_ptr; a.lv = t.lv; a.rk = t.rk; a.R = fo.R; a.dinv = t.dinv; a.vol = t.vol;
    a.ref = t.ref;
    a.ref_tile = t.ref_tile;
    a.uniform_te = t.uniform_te;
    a.in = x; a.partials = t.partials; a.m = m; a.skip = nullptr;
    f.yt = yt;
    f.conn = t.conn;
    f.dinv = t.dinv;
    f.fo = fo;
    f.num_nodes = t.N;
    f.num_parts = t.num_parts;
    const uint32_t g1 = 8 * ((t.num_tiles + 7) / 8);  // one workgroup per tile
// 6 waves per SIMD (80 VGPRs): the element stage's rotation and forces are
// ALU- and latency-bound, and 3 workgroups of 512 threads per CU measured
// 183 us on the 5.24M-tet mesh against 220 (2 per CU, 96 VGPRs), 227 (4, 64
// VGPRs with spills) and 327 (1, 164 VGPRs).
#define CF_FRONT(TBV, WV) \
    hipLaunchKernelGGL((k_qie_front<TBV, WV>), dim3(g1), dim3(TBV), TileLds<TBV>::kBytes, s, f)
    switch (t.tile_threads) {
        case 1024: CF_FRONT(1024, 4); break;
        case 512: CF_FRONT(512, 6); break;
        case 128: CF_FRONT(128, 6); break;
        default: CF_FRONT(256, 6); break;
    }
#undef CF_FRONT
    return check_launch();
}

int launch_forces_skip(const TetDevice& t, int mode, const float* in, const float* R, Material m,
                       const int32_t* skip, hipStream_t s, bool reverse) {
    if (t.num_tiles == 0) return CF_OK;
    TileArgs a;
    a.reverse = reverse ? 1u : 0u;
    a.T = t.T; a.num_tiles = t.num_tiles; a.max_te = t.max_tile_elems; a.max_tn = t.max_tile_nodes;
    a.tile_elem_ptr = t.tile_elem_ptr; a.tile_node_ptr = t.tile_node_ptr; a.tile_nodes = t.tile_nodes;
    a.part_dst = t.part_dst;
    a.slot_ptr = t.slot_ptr; a.lv = t.lv; a.rk = t.rk; a.R = R; a.dinv = t.dinv; a.vol = t.vol;
    a.ref = t.ref;
    a.ref_tile = t.ref_tile;
    a.uniform_te = t.uniform_te;
    a.in = in; a.partials = t.partials; a.m = m; a.skip = skip;
    const uint32_t grid = t.tile_grid;
    const bool fe = mode == kModeFe;
    if (m.mm != CF_MM_C_WRP || (m.dr && !fe)) {
        const uint32_t g1 = 8 * ((t.num_tiles + 7) / 8);  // one workgroup per tile
        const bool lcm = m.mm != CF_MM_C_WRP;
#define CF_TILE_LCM(TBV)                                                                                           \
    do {                                                                                                           \
        const size_t sm = TileLds<TBV>::kBytes;                                                                    \
        if (fe) hipLaunchKernelGGL((k_tile_forces_kind<kModeFe, TBV, kTileLcm>), dim3(g1), dim3(TBV), sm, s, a);   \
        else if (lcm) hipLaunchKernelGGL((k_tile_forces_kind<kModeDfx, TBV, kTileLcm>), dim3(g1), dim3(TBV), sm, s, a); \
        else hipLaunchKernelGGL((k_tile_forces_kind<kModeDfx, TBV, kTileWrpExact>), dim3(g1), dim3(TBV), sm, s, a); \
    } while (0)
        switch (t.tile_threads) {
            case 1024: CF_TILE_LCM(1024); break;
            case 512: CF_TILE_LCM(512); break;
            case 128: CF_TILE_LCM(128); break;
            default: CF_TILE_LCM(256); break;
        }
#undef CF_TILE_LCM
        return check_launch();
    }
#define CF_TILE_LAUNCH(TBV, PV)                                                                          \
    do {                                                                                                \
        const size_t sm = TileLds<TBV>::kBytes;                                                         \
        if (fe) hipLaunchKernelGGL((k_tile_forces<kModeFe, TBV, PV>), dim3(grid), dim3(TBV), sm, s, a); \
        else hipLaunchKernelGGL((k_tile_forces<kModeDfx, TBV, PV>), dim3(grid), dim3(TBV), sm, s, a);   \
    } while (0)
    if (t.tile_persistent) {
        switch (t.tile_threads) {
            case 1024: CF_TILE_LAUNCH(1024, true); break;
            case 512: CF_TILE_LAUNCH(512, true); break;
            case 128: CF_TILE_LAUNCH(128, true); break;
            default: CF_TILE_LAUNCH(256, true); break;
        }
    } else {
        switch (t.tile_threads) {
            case 1024: CF_TILE_LAUNCH(1024, false); break;
            case 512: CF_TILE_LAUNCH(512, false); break;
            case 128: CF_TILE_LAUNCH(128, false); break;
            default: CF_TILE_LAUNCH(256, false); break;
        }
    }
#undef CF_TILE_LAUNCH
    return check_launch();
}

// Grid of the tile kernel, a multiple of 8 (one share per XCD): one workgroup
// per tile, or (persistent) the resident workgroups per CU from the occupancy
// calculator for this LDS size times the CUs, at most one per tile.
int tile_grid_size(const TetDevice& t, int device, int persistent) {
    int cus = 0, per_cu = 0;
    CF_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    hipError_t e;
    switch (t.tile_threads) {
        case 1024: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tile_forces<kModeDfx, 1024, true>, 1024, TileLds<1024>::kBytes); break;
        case 512: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tile_forces<kModeDfx, 512, true>, 512, TileLds<512>::kBytes); break;
        case 128: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tile_forces<kModeDfx, 128, true>, 128, TileLds<128>::kBytes); break;
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tile_forces<kModeDfx, 256, true>, 256, TileLds<256>::kBytes); break;
    }
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    const uint32_t per_xcd_tiles = (t.num_tiles + 7) / 8;
    uint32_t wg_per_xcd = (persistent == 1) ? uint32_t(per_cu) * uint32_t(cus) / 8 : per_xcd_tiles;
    if (wg_per_xcd < 1) wg_per_xcd = 1;
    if (wg_per_xcd > per_xcd_tiles) wg_per_xcd = per_xcd_tiles;
    return int(8 * wg_per_xcd);
}

int launch_damping(const TetDevice& t, const float* v, const float* R, Material m, const DampArgs& d,
                   hipStream_t s) {
    if (t.num_tiles == 0) return CF_OK;
    TileArgs a;
    a.T = t.T; a.num_tiles = t.num_tiles; a.max_te = t.max_tile_elems; a.max_tn = t.max_tile_nodes;
    a.tile_elem_ptr = t.tile_elem_ptr; a.tile_node_ptr = t.tile_node_ptr; a.tile_nodes = t.tile_nodes;
    a.part_dst = t.part_dst;
    a.slot_ptr = t.slot_ptr; a.lv = t.lv; a.rk = t.rk; a.R = R; a.dinv = t.dinv; a.vol = t.vol;
    a.ref = t.ref;
    a.ref_tile = t.ref_tile;
    a.uniform_te = t.uniform_te;
    a.in = v; a.partials = t.partials; a.m = m; a.skip = nullptr;
    const uint32_t g1 = 8 * ((t.num_tiles + 7) / 8);
    switch (t.tile_threads) {
        case 1024: hipLaunchKernelGGL(k_tile_damping<1024>, dim3(g1), dim3(1024), (TileLds<1024, 1>::kBytes), s, a, d); break;
        case 512: hipLaunchKernelGGL(k_tile_damping<512>, dim3(g1), dim3(512), (TileLds<512, 1>::kBytes), s, a, d); break;
        case 128: hipLaunchKernelGGL(k_tile_damping<128>, dim3(g1), dim3(128), (TileLds<128, 1>::kBytes), s, a, d); break;
        default: hipLaunchKernelGGL(k_tile_damping<256>, dim3(g1), dim3(256), (TileLds<256, 1>::kBytes), s, a, d); break;
    }
    return check_launch();
}

int launch_forces(const TetDevice& t, int mode, const float* in, const float* R, Material m, hipStream_t s) {
    return launch_forces_skip(t, mode, in, R, m, nullptr, s);
}

int matvec_chunks(uint32_t N) { return (int)matvec_chunks_of(N); }

int launch_gather_skip(const TetDevice& t, int mode, const NodeTerms& nt, int nblk, const int32_t* skip,
                       hipStream_t s) {
    GatherArgs a;
    a.N = t.N; a.node_part_ptr = t.node_part_ptr; a.partials = t.partials;
    a.t = nt; a.skip = skip;
    if (mode == kGatherAccumulate)
        hipLaunchKernelGGL(k_gather<kGatherAccumulate>, dim3(nblk), dim3(kBlock), 0, s, a);
    else if (mode == kGatherMatvec)
        hipLaunchKernelGGL(k_gather_matvec<false>, dim3(matvec_chunks(t.N)), dim3(kBlock), 0, s, a, matvec_npt(t.N));
    else if (mode == kGatherCgInit)
        hipLaunchKernelGGL(k_gather_matvec<true>, dim3(matvec_chunks(t.N)), dim3(kBlock), 0, s, a, matvec_npt(t.N));
    else if (mode == kGatherRhsDx)
        hipLaunchKernelGGL(k_gather<kGatherRhsDx>, dim3(nblk), dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL(k_gather<kGatherRhs>, dim3(nblk), dim3(kBlock), 0, s, a);
    return check_launch();
}

int launch_rhs_start(const TetDevice& t, const float* ay, const NodeTerms& nt, hipStream_t s) {
    if (t.N == 0) return CF_OK;
    GatherArgs a;
    a.N = t.N; a.node_part_ptr = t.node_part_ptr; a.partials = t.partials; a.t = nt; a.skip = nullptr;
    hipLaunchKernelGGL(k_rhs_start, dim3(matvec_chunks(t.N)), dim3(kBlock), 0, s, a, ay, matvec_npt(t.N));
    return check_launch();
}

int launch_gather(const TetDevice& t, int mode, const NodeTerms& nt, int nblk, hipStream_t s) {
    return launch_gather_skip(t, mode, nt, nblk, nullptr, s);
}

int launch_energy(const TetDevice& t, const float* x, const float* R, Material m, double* blk, int nblk,
                  double* total, hipStream_t s) {
    if (m.mm == CF_MM_C_WRP && t.num_tiles > 0 && t.tile_part) {
        TileArgs a;
        a.T = t.T; a.num_tiles = t.num_tiles; a.max_te = t.max_tile_elems; a.max_tn = t.max_tile_nodes;
        a.tile_elem_ptr = t.tile_elem_ptr; a.tile_node_ptr = t.tile_node_ptr; a.tile_nodes = t.tile_nodes;
        a.part_dst = t.part_dst;
        a.slot_ptr = t.slot_ptr; a.lv = t.lv; a.rk = t.rk; a.R = R; a.dinv = t.dinv; a.vol = t.vol;
        a.ref = t.ref;
        a.ref_tile = t.ref_tile;
        a.uniform_te = t.uniform_te;
        a.in = x; a.partials = t.partials; a.m = m; a.skip = nullptr;
        const uint32_t g1 = 8 * ((t.num_tiles + 7) / 8);  // one workgroup per tile
        switch (t.tile_threads) {
            case 1024: hipLaunchKernelGGL(k_tile_energy<1024>, dim3(g1), dim3(1024), 6 * 1024 * 4, s, a, t.tile_part); break;
            case 512: hipLaunchKernelGGL(k_tile_energy<512>, dim3(g1), dim3(512), 6 * 512 * 4, s, a, t.tile_part); break;
            case 128: hipLaunchKernelGGL(k_tile_energy<128>, dim3(g1), dim3(128), 6 * 128 * 4, s, a, t.tile_part); break;
            default: hipLaunchKernelGGL(k_tile_energy<256>, dim3(g1), dim3(256), 6 * 256 * 4, s, a, t.tile_part); break;
        }
        CF_TRY(check_launch());
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, s, t.tile_part, int(t.num_tiles), total);
        return check_launch();
    }
    hipLaunchKernelGGL(k_energy, dim3(nblk), dim3(kBlock), 0, s, t.T, t.conn, t.ref, x, R, t.dinv, t.vol, m, blk);
    CF_TRY(check_launch());
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, s, blk, nblk, total);
    return check_launch();
}

int launch_cg_init(uint32_t n, const float* q, const float* b, float* r, float* d, const uint8_t* mask,
                   double* rr_part, int nblk, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_init, dim3(nblk), dim3(kBlock), 0, s, n, q, b, r, d, mask, rr_part);
    return check_launch();
}

int launch_cg_start(CgScalars* st, const double* rr_part, int nblk_prev, float rel, float abs_eps,
                    uint32_t max_iter, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(kBlock), 0, s, st, rr_part, nblk_prev, rel, abs_eps, max_iter);
    return check_launch();
}

int launch_cg_dot(uint32_t n, const float* a, const float* b, double* part, int nblk, hipStream_t s) {
    hipLaunchKernelGGL(k_dot, dim3(nblk), dim3(kBlock), 0, s, n, a, b, part);
    return check_launch();
}

static uint32_t g_cg_workgroups = 0;  // 0: kCgWorkgroups

int set_cg_workgroups(uint32_t n) {
    if (n > (uint32_t)kMaxPartials) return CF_ERR_INVALID_ARG;
    g_cg_workgroups = n;
    return CF_OK;
}
uint32_t get_cg_workgroups() { return g_cg_workgroups ? g_cg_workgroups : kCgWorkgroups; }

static dim3 cg_step_grid(uint32_t nchunk) {
    const uint32_t g = g_cg_workgroups ? g_cg_workgroups : kCgWorkgroups;
    return dim3(std::max(1u, std::min(nchunk, g)));
}

int launch_cg_step_xr(uint32_t n, uint32_t k, CgScalars* st, float* r, const float* q, const double* dq_part,
                      int ndq, double* rr_part, hipStream_t s) {
    const uint32_t nchunk = cg_chunks(n);  // the d kernel sums this many r.r partials
    const dim3 grid = cg_step_grid(nchunk);
    if (aligned16(r) && aligned16(q))
        hipLaunchKernelGGL(k_cg_step_xr<true>, grid, dim3(kBlock), 0, s, n, k, st, r, q, dq_part, ndq, rr_part,
                           nchunk, cg_rows(n));
    else
        hipLaunchKernelGGL(k_cg_step_xr<false>, grid, dim3(kBlock), 0, s, n, k, st, r, q, dq_part, ndq, rr_part,
                           nchunk, cg_rows(n));
    return check_launch();
}

int launch_cg_step_d(uint32_t n, uint32_t k, CgScalars* st, float* x, float* d, const float* r,
                     const uint8_t* mask, const double* rr_part, hipStream_t s, int32_t* host_done) {
    const uint32_t nchunk = cg_chunks(n);
    const dim3 grid = cg_step_grid(nchunk);
    if (aligned16(x) && aligned16(d) && aligned16(r))
        hipLaunchKernelGGL(k_cg_step_d<true>, grid, dim3(kBlock), 0, s, n, k, st, x, d, r, mask, rr_part, nchunk,
                           cg_rows(n), host_done);
    else
        hipLaunchKernelGGL(k_cg_step_d<false>, grid, dim3(kBlock), 0, s, n, k, st, x, d, r, mask, rr_part, nchunk,
                           cg_rows(n), host_done);
    return check_launch();
}

int launch_nr_update(uint32_t n, float* xk, float* vk, const float* x, const float* y, float inv_dt, double* part,
                     int nblk, hipStream_t s) {
    hipLaunchKernelGGL(k_nr_update, dim3(nblk), dim3(kBlock), 0, s, n, xk, vk, x, y, inv_dt, part);
    return check_launch();
}

int launch_nr_finish(uint32_t n, float* x, float* v, const float* xk, const float* vk, float dt, float drag,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_nr_finish, dim3(node_blocks(n)), dim3(kBlock), 0, s, n, x, v, xk, vk, dt, drag);
    return check_launch();
}

int launch_integrate(uint32_t n, float* x, float* v, const float* y, float dt, float drag, hipStream_t s) {
    hipLaunchKernelGGL(k_integrate, dim3(node_blocks(n)), dim3(kBlock), 0, s, n, x, v, y, dt, drag);
    return check_launch();
}

int launch_zero_mask(uint32_t n_nodes, float* w, const uint8_t* mask, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_mask, dim3(node_blocks(n_nodes)), dim3(kBlock), 0, s, n_nodes, w, mask);
    return check_launch();
}

}  // namespace cf
