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
#include "cf_internal.h"
#include <hip/hip_runtime.h>
#include <cmath>

namespace cf {

// ============================================================ FP32 math in the
// reference's evaluation order. Contraction is disabled in these functions so
// that, like the reference's SSE4.1 build (no FMA), each product is rounded.
namespace refmath {

__device__ inline float det3(const float* m) {  // Mal/GMat.h:236-244
#pragma clang fp contract(off)
    return m[0] * m[4] * m[8] - m[0] * m[7] * m[5] + m[3] * m[7] * m[2] - m[3] * m[1] * m[8] +
           m[6] * m[1] * m[5] - m[6] * m[4] * m[2];
}

__device__ inline void inv3(const float* m, float det, float* r) {  // Mal/GMat.h:291-307
#pragma clang fp contract(off)
    float id = 1.0f / det;
    r[0] = id * (m[4] * m[8] - m[5] * m[7]);
    r[1] = id * (m[2] * m[7] - m[1] * m[8]);
    r[2] = id * (m[1] * m[5] - m[2] * m[4]);
    r[3] = id * (m[5] * m[6] - m[3] * m[8]);
    r[4] = id * (m[0] * m[8] - m[2] * m[6]);
    r[5] = id * (m[2] * m[3] - m[0] * m[5]);
    r[6] = id * (m[3] * m[7] - m[4] * m[6]);
    r[7] = id * (m[1] * m[6] - m[0] * m[7]);
    r[8] = id * (m[0] * m[4] - m[1] * m[3]);
}

__device__ inline void mul3(const float* a, const float* b, float* p) {  // Mal/GMat.h:135-145
#pragma clang fp contract(off)
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            float acc = 0.0f;
            acc += a[3 * i + 0] * b[0 + j];
            acc += a[3 * i + 1] * b[3 + j];
            acc += a[3 * i + 2] * b[6 + j];
            p[3 * i + j] = acc;
        }
}

// Newton polar decomposition with the reference's loose stop rule
// (Mal/GMatDecomposition.h:411-432): error = |H1-H0|_F^2 <= 0.01 or 10 steps.
__device__ inline void polar_newton(const float* H, float* R) {
#pragma clang fp contract(off)
    float H0[9];
    for (int i = 0; i < 9; i++) H0[i] = H[i];
    unsigned n = 0;
    float err;
    do {
        float Ht[9] = {H0[0], H0[3], H0[6], H0[1], H0[4], H0[7], H0[2], H0[5], H0[8]};
        float it[9];
        inv3(Ht, det3(Ht), it);
        float H1[9];
        for (int i = 0; i < 9; i++) H1[i] = 0.5f * (H0[i] + it[i]);
        err = 0.0f;
        for (int i = 0; i < 9; i++) {
            float d = H1[i] - H0[i];
            err += d * d;
        }
        for (int i = 0; i < 9; i++) H0[i] = H1[i];
        n++;
    } while (n < 10u && err > 0.01f);
    for (int i = 0; i < 9; i++) R[i] = H0[i];
}

// ---- PhysBAM-style FP32 SVD as adapted by the reference
// (Mal/GMatDecomposition_SVD3x3_PhysBAM.h:35-280).
struct v3 { float x, y, z; };

__device__ inline v3 cross(v3 a, v3 b) {  // Mal/GVec.h:89-91
#pragma clang fp contract(off)
    return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline float dot(v3 a, v3 b) {
#pragma clang fp contract(off)
    float acc = 0.0f;
    acc += a.x * b.x;
    acc += a.y * b.y;
    acc += a.z * b.z;
    return acc;
}
__device__ inline v3 normalized(v3 a) {  // v * (1/|v|), Mal/GVec.h:85,97
#pragma clang fp contract(off)
    float inv = 1.0f / sqrtf(dot(a, a));
    return v3{inv * a.x, inv * a.y, inv * a.z};
}
__device__ inline v3 orthogonal_vector(v3 v) {  // :35-49
    float ax = fabsf(v.x), ay = fabsf(v.y), az = fabsf(v.z);
    if (ax < ay) return (ax < az) ? v3{0.f, v.z, -v.y} : v3{v.y, -v.x, 0.f};
    return (ay < az) ? v3{-v.z, 0.f, v.x} : v3{v.y, -v.x, 0.f};
}
struct sym3 { float x11, x21, x31, x22, x32, x33; };

__device__ inline v3 largest_column_normalized(const sym3& s) {  // :149-156
#pragma clang fp contract(off)
    float s11 = s.x11 * s.x11, s12 = s.x21 * s.x21, s13 = s.x31 * s.x31;
    float s22 = s.x22 * s.x22, s23 = s.x32 * s.x32, s33 = s.x33 * s.x33;
    float c1 = s11 + s12 + s13, c2 = s12 + s22 + s23, c3 = s13 + s23 + s33;
    if (c1 > c2) {
        if (c1 > c3) { float inv = 1.0f / sqrtf(c1); return v3{inv * s.x11, inv * s.x21, inv * s.x31}; }
    } else if (c2 > c3) { float inv = 1.0f / sqrtf(c2); return v3{inv * s.x21, inv * s.x22, inv * s.x32}; }
    if (c3 > 0) { float inv = 1.0f / sqrtf(c3); return v3{inv * s.x31, inv * s.x32, inv * s.x33}; }
    return v3{1.f, 0.f, 0.f};
}

__device__ inline void svd_rotation(const float* F, float* R) {
#pragma clang fp contract(off)
    // normal equations F^T F (:129-137)
    sym3 A;
    A.x11 = F[0] * F[0] + F[3] * F[3] + F[6] * F[6];
    A.x21 = F[1] * F[0] + F[4] * F[3] + F[7] * F[6];
    A.x31 = F[2] * F[0] + F[5] * F[3] + F[8] * F[6];
    A.x22 = F[1] * F[1] + F[4] * F[4] + F[7] * F[7];
    A.x32 = F[2] * F[1] + F[5] * F[4] + F[8] * F[7];
    A.x33 = F[2] * F[2] + F[5] * F[5] + F[8] * F[8];
    // Smith (1961) eigenvalues (:161-197). sqrt(3) is written Sqrt(3) with an
    // int argument in the reference, which deduces Sqrt<int> and yields 1.
    const float one_third = 1.0f / 3.0f;
    const float one_sixth = 1.0f / 6.0f;
    const float root_three = 1.0f;
    float m = one_third * (A.x11 + A.x22 + A.x33);
    float a11 = A.x11 - m, a22 = A.x22 - m, a33 = A.x33 - m;
    float a12s = A.x21 * A.x21, a13s = A.x31 * A.x31, a23s = A.x32 * A.x32;
    float p = one_sixth * (a11 * a11 + a22 * a22 + a33 * a33 + 2 * (a12s + a13s + a23s));
    float q = 0.5f * (a11 * (a22 * a33 - a23s) - a22 * a13s - a33 * a12s) + A.x21 * A.x31 * A.x32;
    float sqrt_p = sqrtf(p), disc = p * p * p - q * q;
    float phi = one_third * atan2f(sqrtf(fmaxf(0.0f, disc)), q);
    float c = cosf(phi), sn = sinf(phi);
    float spc = sqrt_p * c, r3sps = root_three * sqrt_p * sn;
    float l0 = m + 2 * spc, l1 = m - spc - r3sps, l2 = m - spc + r3sps;
    float L0, L1, L2;
    if (l0 < l1) {
        if (l1 < l2) { L0 = l2; L1 = l1; L2 = l0; }
        else if (l0 < l2) { L0 = l1; L1 = l2; L2 = l0; }
        else { L0 = l1; L1 = l0; L2 = l2; }
    } else {
        if (l1 > l2) { L0 = l0; L1 = l1; L2 = l2; }
        else if (l0 > l2) { L0 = l0; L1 = l2; L2 = l1; }
        else { L0 = l2; L1 = l0; L2 = l1; }
    }
    // eigenvectors (:199-236)
    bool flipped = false;
    float f0 = L0, f2 = L2;
    if (L0 - L1 < L1 - L2) { f0 = L2; f2 = L0; flipped = true; }
    sym3 am = A;
    am.x11 -= f0; am.x22 -= f0; am.x33 -= f0;
    sym3 cf_;
    cf_.x11 = am.x22 * am.x33 - am.x32 * am.x32;
    cf_.x21 = am.x32 * am.x31 - am.x21 * am.x33;
    cf_.x31 = am.x21 * am.x32 - am.x22 * am.x31;
    cf_.x22 = am.x11 * am.x33 - am.x31 * am.x31;
    cf_.x32 = am.x21 * am.x31 - am.x11 * am.x32;
    cf_.x33 = am.x11 * am.x22 - am.x21 * am.x21;
    v3 v1 = largest_column_normalized(cf_);
    v3 o0 = normalized(orthogonal_vector(v1));
    v3 o1 = cross(v1, o0);
    // reduced 2x2: (full_A * other)^T other, accumulate from 0
    auto Amul = [&](v3 u) {
        v3 r;
        float acc;
        acc = 0.0f; acc += A.x11 * u.x; acc += A.x21 * u.y; acc += A.x31 * u.z; r.x = acc;
        acc = 0.0f; acc += A.x21 * u.x; acc += A.x22 * u.y; acc += A.x32 * u.z; r.y = acc;
        acc = 0.0f; acc += A.x31 * u.x; acc += A.x32 * u.y; acc += A.x33 * u.z; r.z = acc;
        return r;
    };
    v3 b0 = Amul(o0), b1 = Amul(o1);
    float r11 = b0.x * o0.x + b0.y * o0.y + b0.z * o0.z;
    float r21 = b1.x * o0.x + b1.y * o0.y + b1.z * o0.z;
    float r22 = b1.x * o1.x + b1.y * o1.y + b1.z * o1.z;
    r11 -= f2; r22 -= f2;
    // 2x2 cofactor (x22, -x21, x11) and its largest column (:70-82)
    float c11 = r22, c21 = -r21, c22 = r11;
    float w0, w1;
    {
        float s11 = c11 * c11, s12 = c21 * c21, s22 = c22 * c22;
        float k1 = s11 + s12, k2 = s12 + s22;
        if (k1 > k2) { float inv = 1.0f / sqrtf(k1); w0 = inv * c11; w1 = inv * c21; }
        else if (k2 > 0) { float inv = 1.0f / sqrtf(k2); w0 = inv * c21; w1 = inv * c22; }
        else { w0 = 1.f; w1 = 0.f; }
    }
    v3 v3v;
    {
        float acc;
        acc = 0.0f; acc += o0.x * w0; acc += o1.x * w1; v3v.x = acc;
        acc = 0.0f; acc += o0.y * w0; acc += o1.y * w1; v3v.y = acc;
        acc = 0.0f; acc += o0.z * w0; acc += o1.z * w1; v3v.z = acc;
    }
    v3 v2 = cross(v3v, v1);
    v3 V0, V1, V2;  // columns of V
    if (flipped) { V0 = v3v; V1 = v2; V2 = v3{-v1.x, -v1.y, -v1.z}; }
    else { V0 = v1; V1 = v2; V2 = v3v; }
    // singular vectors (:259-279); the singular values only enter through
    // the sign of det F, which does not change U or V here.
    auto Fmul = [&](v3 u) {
        v3 r;
        float acc;
        acc = 0.0f; acc += F[0] * u.x; acc += F[1] * u.y; acc += F[2] * u.z; r.x = acc;
        acc = 0.0f; acc += F[3] * u.x; acc += F[4] * u.y; acc += F[5] * u.z; r.y = acc;
        acc = 0.0f; acc += F[6] * u.x; acc += F[7] * u.y; acc += F[8] * u.z; r.z = acc;
        return r;
    };
    v3 U0 = normalized(Fmul(V0));
    v3 p0 = normalized(orthogonal_vector(U0));
    v3 p1 = cross(U0, p0);
    v3 fv1 = normalized(Fmul(V1));
    float t0 = dot(p0, fv1), t1 = dot(p1, fv1);
    v3 U1;
    {
        float acc;
        acc = 0.0f; acc += p0.x * t0; acc += p1.x * t1; U1.x = acc;
        acc = 0.0f; acc += p0.y * t0; acc += p1.y * t1; U1.y = acc;
        acc = 0.0f; acc += p0.z * t0; acc += p1.z * t1; U1.z = acc;
    }
    v3 U2 = cross(U0, U1);
    // R = U * V^T  (Rotation.cpp:55)
    const float U[9] = {U0.x, U1.x, U2.x, U0.y, U1.y, U2.y, U0.z, U1.z, U2.z};
    const float Vt[9] = {V0.x, V0.y, V0.z, V1.x, V1.y, V1.z, V2.x, V2.y, V2.z};
    mul3(U, Vt, R);
}

}  // namespace refmath

// ============================================================ reductions
__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// Sum over the block; result valid in every thread. sh: kBlock/64 doubles.
__device__ inline double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    for (int i = 0; i < kBlock / 64; i++) r += sh[i];
    return r;
}
// Deterministic sum of n partials by one block (fixed order and tree).
__device__ inline double reduce_partials(const double* part, int n, double* sh) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock) v += part[i];
    return block_sum(v, sh);
}

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
// (ms/fem/TetrahedronElementSet3.cpp:219-281 + ComputeRotations_*, :748-892)
__global__ __launch_bounds__(kBlock) void k_begin_eval(uint32_t T, const uint32_t* __restrict__ conn,
                                                       const float* __restrict__ x, const float* __restrict__ dinv,
                                                       int rm, float thr, float* __restrict__ R,
                                                       float* __restrict__ detF, uint32_t* blk_count,
                                                       double* blk_sum, float* blk_min) {
#pragma clang fp contract(off)
    __shared__ double sh[kBlock / 64];
    __shared__ float shm[kBlock / 64];
    __shared__ uint32_t shc[kBlock / 64];
    __shared__ uint32_t shf[kBlock / 64];
    uint32_t cnt = 0, fb = 0;
    double sdet = 0.0;
    float mdet = 3.0e38f;
    for (uint32_t e = blockIdx.x * kBlock + threadIdx.x; e < T; e += gridDim.x * kBlock) {
        float xs[4][3];
        for (int j = 0; j < 4; j++) {
            const uint32_t g = conn[j * (size_t)T + e];
            xs[j][0] = x[3 * (size_t)g]; xs[j][1] = x[3 * (size_t)g + 1]; xs[j][2] = x[3 * (size_t)g + 2];
        }
        float Ds[9], G[9], F[9];
        for (int i = 0; i < 3; i++) {
            Ds[3 * i + 0] = xs[1][i] - xs[0][i];
            Ds[3 * i + 1] = xs[2][i] - xs[0][i];
            Ds[3 * i + 2] = xs[3][i] - xs[0][i];
        }
        for (int k = 0; k < 9; k++) G[k] = dinv[k * (size_t)T + e];
        refmath::mul3(Ds, G, F);
        const float d = refmath::det3(F);
        float Rm[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
        if (rm == CF_RM_MSVD) {
            refmath::svd_rotation(F, Rm);
        } else if (rm == CF_RM_DAPD) {
            if (d > thr) refmath::polar_newton(F, Rm);
        }
        for (int k = 0; k < 9; k++) R[k * (size_t)T + e] = Rm[k];
        detF[e] = d;
        if (d <= 0.0f) { cnt++; sdet += d; }
        // DAPD elements at or below the threshold take R = Id: counted too.
        if (rm == CF_RM_DAPD && !(d > thr)) fb++;
        mdet = fminf(mdet, d);
    }
    // block reductions (fixed order)
    double s = block_sum(sdet, sh);
    for (int o = 32; o > 0; o >>= 1) {
        mdet = fminf(mdet, __shfl_xor(mdet, o, 64));
        cnt += __shfl_xor(cnt, o, 64);
        fb += __shfl_xor(fb, o, 64);
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { shm[w] = mdet; shc[w] = cnt; shf[w] = fb; }
    __syncthreads();
    if (threadIdx.x == 0) {
        float mm = shm[0];
        uint32_t cc = shc[0], ff = shf[0];
        for (int i = 1; i < kBlock / 64; i++) { mm = fminf(mm, shm[i]); cc += shc[i]; ff += shf[i]; }
        blk_count[blockIdx.x] = cc;
        blk_count[gridDim.x + blockIdx.x] = ff;
        blk_sum[blockIdx.x] = s;
        blk_min[blockIdx.x] = mm;
    }
}

// ============================================================ element forces
// Corotated element force differential via the equivalent Piola form of
// Ke = V B^T E B (TetrahedronElement3.h:157-173):
//   H = R^T [v1-v0, v2-v0, v3-v0] Dm^-1            (df_x: v = dx)
//   H = (R^T [x1-x0, x2-x0, x3-x0] - Dm) Dm^-1     (f_e:  v = x, u = R^T x - r)
//   sigma = E(sym H) with the reference's a_k1, a_nu, a_k2
//   P = -V R sigma Dm^-T ; node c+1 gets column c, node 0 gets -(sum of columns)
// which equals -R * sum_j K_ij * (local vector j) of C_f_e / C_df_x.
struct TileArgs {
    uint32_t T;
    uint32_t num_tiles;
    uint32_t max_te, max_tn;
    const uint32_t* tile_elem_ptr;
    const uint32_t* tile_node_ptr;
    const uint32_t* tile_nodes;
    const uint32_t* part_dst;
    const uint32_t* slot_ptr;
    const uint16_t* dst;
    const uint16_t* lvid;
    const float* R;
    const float* dinv;
    const float* vol;
    const float* ref;
    const float* in;
    float* partials;
    Material m;
    const int32_t* skip;
};

// LDS-only workgroup barrier: orders LDS accesses across the barrier without
// waiting for outstanding global loads, so that prefetches stay in flight.
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

struct TileDesc { uint32_t e0, ne, n0, nn; };

__device__ inline TileDesc tile_desc(const TileArgs& a, uint32_t t) {
    TileDesc d;
    d.e0 = a.tile_elem_ptr[t];
    d.ne = a.tile_elem_ptr[t + 1] - d.e0;
    d.n0 = a.tile_node_ptr[t];
    d.nn = a.tile_node_ptr[t + 1] - d.n0;
    return d;
}

// Per-thread registers of one tile: this thread's tile node (vector, rest
// position, slot offset) and element (local vertex ids, force slots, R).
struct TileRegs {
    float v[3], r[3], R[9];
    uint32_t sp, pd;
    uint32_t l[4];
    uint32_t f[4];
};

__device__ inline void load_tile_regs(const TileArgs& a, const TileDesc& d, uint32_t gid, uint32_t tid,
                                      TileRegs& g) {
    const size_t T = a.T;
    if (tid < d.nn) {
        g.v[0] = a.in[3 * (size_t)gid + 0]; g.v[1] = a.in[3 * (size_t)gid + 1]; g.v[2] = a.in[3 * (size_t)gid + 2];
        g.r[0] = a.ref[3 * (size_t)gid + 0]; g.r[1] = a.ref[3 * (size_t)gid + 1]; g.r[2] = a.ref[3 * (size_t)gid + 2];
    }
    if (tid <= d.nn) g.sp = a.slot_ptr[d.n0 + tid] - 4 * d.e0;
    if (tid < d.nn) g.pd = a.part_dst[d.n0 + tid];
    if (tid < d.ne) {
        const size_t e = d.e0 + tid;
        g.l[0] = a.lvid[e]; g.l[1] = a.lvid[T + e]; g.l[2] = a.lvid[2 * T + e]; g.l[3] = a.lvid[3 * T + e];
        g.f[0] = a.dst[e]; g.f[1] = a.dst[T + e]; g.f[2] = a.dst[2 * T + e]; g.f[3] = a.dst[3 * T + e];
#pragma unroll
        for (int k = 0; k < 9; k++) g.R[k] = a.R[k * T + e];
    }
}

// Nodal forces of one element into its four tile force slots.
template <int MODE>
__device__ inline void element_force(const float4* s_v, const float4* s_r, float4* s_F, const uint32_t* lv,
                                     const uint32_t* fs, const float* R, const Material& m) {
    const float4 x0 = s_v[lv[0]], x1 = s_v[lv[1]], x2 = s_v[lv[2]], x3 = s_v[lv[3]];
    const float4 r0 = s_r[lv[0]], r1 = s_r[lv[1]], r2 = s_r[lv[2]], r3 = s_r[lv[3]];
    // Dm^-1 and V rebuilt from the rest positions staged in LDS with the same
    // operations as k_precompute (bit-identical to the stored values): 12 B
    // per tile node instead of 40 B per element from HBM.
    const float Dm[9] = {r1.x - r0.x, r2.x - r0.x, r3.x - r0.x,
                         r1.y - r0.y, r2.y - r0.y, r3.y - r0.y,
                         r1.z - r0.z, r2.z - r0.z, r3.z - r0.z};
    const float detDm = refmath::det3(Dm);
    const float V = detDm / 6.0f;
    float G[9];
    refmath::inv3(Dm, detDm, G);
    const float D[9] = {x1.x - x0.x, x2.x - x0.x, x3.x - x0.x,   // columns v1-v0, v2-v0, v3-v0
                        x1.y - x0.y, x2.y - x0.y, x3.y - x0.y,
                        x1.z - x0.z, x2.z - x0.z, x3.z - x0.z};
    float RD[9];  // R^T D  (f_e: minus the rest edges Dm, the local displacement differences)
#pragma unroll
    for (int i3 = 0; i3 < 3; i3++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            RD[3 * i3 + j] = R[0 + i3] * D[0 + j] + R[3 + i3] * D[3 + j] + R[6 + i3] * D[6 + j];
    if (MODE == kModeFe) {
#pragma unroll
        for (int k = 0; k < 9; k++) RD[k] -= Dm[k];
    }
    float H[9];  // RD G
#pragma unroll
    for (int i3 = 0; i3 < 3; i3++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            H[3 * i3 + j] = RD[3 * i3 + 0] * G[0 + j] + RD[3 * i3 + 1] * G[3 + j] + RD[3 * i3 + 2] * G[6 + j];
    const float sxx = m.a_k1 * H[0] + m.a_nu * (H[4] + H[8]);
    const float syy = m.a_k1 * H[4] + m.a_nu * (H[0] + H[8]);
    const float szz = m.a_k1 * H[8] + m.a_nu * (H[0] + H[4]);
    const float sxy = m.a_k2 * (H[1] + H[3]);
    const float syz = m.a_k2 * (H[5] + H[7]);
    const float szx = m.a_k2 * (H[2] + H[6]);
    const float S[9] = {sxx, sxy, szx, sxy, syy, syz, szx, syz, szz};
    float W[9];  // -V sigma G^T
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            W[3 * r + c] = -V * (S[3 * r + 0] * G[3 * c + 0] + S[3 * r + 1] * G[3 * c + 1] + S[3 * r + 2] * G[3 * c + 2]);
    float P[9];  // R W: column c is the force on vertex c+1
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            P[3 * r + c] = R[3 * r + 0] * W[0 + c] + R[3 * r + 1] * W[3 + c] + R[3 * r + 2] * W[6 + c];
    s_F[fs[1]] = make_float4(P[0], P[3], P[6], 0.f);
    s_F[fs[2]] = make_float4(P[1], P[4], P[7], 0.f);
    s_F[fs[3]] = make_float4(P[2], P[5], P[8], 0.f);
    s_F[fs[0]] = make_float4(-(P[0] + P[1] + P[2]), -(P[3] + P[4] + P[5]), -(P[6] + P[7] + P[8]), 0.f);
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
    if (a.skip && *a.skip) return;
    extern __shared__ float4 smem4[];
    float4* s_v = smem4;                                 // max_tn
    float4* s_r = s_v + a.max_tn;                        // max_tn
    float4* s_F = s_r + a.max_tn;                        // 4*max_te force slots
    uint32_t* s_sp = reinterpret_cast<uint32_t*>(s_F + 4 * a.max_te);  // max_tn+1

    const uint32_t tid = threadIdx.x;
    const uint32_t gx = gridDim.x / 8, xcd = blockIdx.x % 8, j = blockIdx.x / 8;
    const uint32_t per = (a.num_tiles + 7) / 8;
    const uint32_t lo = xcd * per;
    const uint32_t hi = min(a.num_tiles, lo + per);
    uint32_t t = lo + j;
    if (t >= hi) return;

    TileDesc dt = tile_desc(a, t);
    TileRegs g;
    g.sp = 0;
    g.pd = 0;
    {
        const uint32_t gid = (tid < dt.nn) ? a.tile_nodes[dt.n0 + tid] : 0u;
        load_tile_regs(a, dt, gid, tid, g);
    }
    uint32_t t1 = t + gx;
    TileDesc d1 = dt;
    uint32_t gid1 = 0;
    if (PIPE && t1 < hi) {
        d1 = tile_desc(a, t1);
        gid1 = (tid < d1.nn) ? a.tile_nodes[d1.n0 + tid] : 0u;
    }
    while (true) {
        // stage tile t into LDS
        if (tid < dt.nn) {
            s_v[tid] = make_float4(g.v[0], g.v[1], g.v[2], 0.f);
            s_r[tid] = make_float4(g.r[0], g.r[1], g.r[2], 0.f);
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
            gid2 = (tid < d2.nn) ? a.tile_nodes[d2.n0 + tid] : 0u;
        }
        // element forces of tile t into the force slots
        if (tid < dt.ne) element_force<MODE>(s_v, s_r, s_F, cl, cf, cR, a.m);
        lds_barrier();
        // per tile node: its contiguous slot run (ascending element order) -> one partial
        if (tid < dt.nn) {
            float fx = 0.f, fy = 0.f, fz = 0.f;
            const uint32_t s1 = s_sp[tid + 1];
            for (uint32_t sidx = s_sp[tid]; sidx < s1; sidx++) {
                const float4 f = s_F[sidx];
                fx += f.x; fy += f.y; fz += f.z;
            }
            float* o = a.partials + 3 * (size_t)cpd;
            o[0] = fx; o[1] = fy; o[2] = fz;
        }
        if (!PIPE || t1 >= hi) break;
        lds_barrier();
        t = t1; dt = d1;
        t1 = t2; d1 = d2; gid1 = gid2;
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
    double local = 0.0;
    for (uint32_t n = blockIdx.x * kBlock + threadIdx.x; n < a.N; n += gridDim.x * kBlock) {
        float fx = 0.f, fy = 0.f, fz = 0.f;
        const uint32_t p1 = a.node_part_ptr[n + 1];
        for (uint32_t p = a.node_part_ptr[n]; p < p1; p++) {
            const float* pp = a.partials + 3 * (size_t)p;
            fx += pp[0]; fy += pp[1]; fz += pp[2];
        }
        const size_t i = 3 * (size_t)n;
        if (MODE == kGatherAccumulate) {
            a.t.out[i] += fx; a.t.out[i + 1] += fy; a.t.out[i + 2] += fz;
        } else if (MODE == kGatherMatvec) {
            // LS_A_times_d (LeafDSH_Solid3D_FEM.cpp:1056-1079) + LS_ZeroKNC (:1045-1053)
            const float mass = a.t.mass_array ? a.t.mass_array[n] : a.t.mass;
            const float fac = mass * a.t.fM;
            const float d0 = a.t.d[i], d1 = a.t.d[i + 1], d2 = a.t.d[i + 2];
            float q0 = -a.t.fK * fx + fac * d0;
            float q1 = -a.t.fK * fy + fac * d1;
            float q2 = -a.t.fK * fz + fac * d2;
            if (a.t.knc_mask && a.t.knc_mask[n]) { q0 = 0.f; q1 = 0.f; q2 = 0.f; }
            a.t.out[i] = q0; a.t.out[i + 1] = q1; a.t.out[i + 2] = q2;
            local += (double)d0 * q0 + (double)d1 * q1 + (double)d2 * q2;
        } else {
            // b = dt (M g + f_e) + M v ; y = v ; KNC rows take the constraint
            // velocity (Step_QuasiImplicitEuler_v :1209-1295).
            const float mass = a.t.mass_array ? a.t.mass_array[n] : a.t.mass;
            const float f[3] = {fx, fy, fz};
            float b[3], y[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float vc = a.t.v[i + c];
                b[c] = (mass * a.t.g[c] + f[c]) * a.t.dt + mass * vc;
                y[c] = a.t.warm_start ? vc : 0.0f;
            }
            if (a.t.knc_mask && a.t.knc_mask[n]) {
                float vk[3];
#pragma unroll
                for (int c = 0; c < 3; c++) vk[c] = (a.t.knc_target[i + c] - a.t.x[i + c]) * a.t.inv_dt;
                const float nk = sqrtf(vk[0] * vk[0] + vk[1] * vk[1] + vk[2] * vk[2]);
                if (nk > 0.001f) {
                    const float inv = 1.0f / nk, s = fminf(nk, 50.0f);
#pragma unroll
                    for (int c = 0; c < 3; c++) vk[c] = s * (vk[c] * inv);
                } else {
                    vk[0] = vk[1] = vk[2] = 0.f;
                }
#pragma unroll
                for (int c = 0; c < 3; c++) { b[c] = vk[c]; y[c] = vk[c]; }
            }
#pragma unroll
            for (int c = 0; c < 3; c++) { a.t.out[i + c] = b[c]; a.t.y[i + c] = y[c]; }
        }
    }
    if (MODE == kGatherMatvec) {
        const double s = block_sum(local, sh);
        if (threadIdx.x == 0) a.t.partials[blockIdx.x] = s;
    }
}

// ============================================================ energy
// C_V (ms/fem/TetrahedronElementSet3.cpp:1456-1488): 1/2 u^T Ke u = 1/2 V sigma:H
__global__ __launch_bounds__(kBlock) void k_energy(uint32_t T, const uint32_t* __restrict__ conn,
                                                   const float* __restrict__ ref,
                                                   const float* __restrict__ x, const float* __restrict__ R,
                                                   const float* __restrict__ dinv, const float* __restrict__ vol,
                                                   Material m, double* blk) {
    __shared__ double sh[kBlock / 64];
    double acc = 0.0;
    for (uint32_t e = blockIdx.x * kBlock + threadIdx.x; e < T; e += gridDim.x * kBlock) {
        float xs[4][3];
        for (int j = 0; j < 4; j++) {
            const uint32_t g = conn[j * (size_t)T + e];
            xs[j][0] = x[3 * (size_t)g]; xs[j][1] = x[3 * (size_t)g + 1]; xs[j][2] = x[3 * (size_t)g + 2];
        }
        float rs[4][3];
        for (int j = 0; j < 4; j++) {
            const uint32_t g = conn[j * (size_t)T + e];
            rs[j][0] = ref[3 * (size_t)g]; rs[j][1] = ref[3 * (size_t)g + 1]; rs[j][2] = ref[3 * (size_t)g + 2];
        }
        float D[9], Rm[9], G[9], RD[9], H[9];
        for (int c = 0; c < 3; c++) {
            D[3 * c + 0] = xs[1][c] - xs[0][c];
            D[3 * c + 1] = xs[2][c] - xs[0][c];
            D[3 * c + 2] = xs[3][c] - xs[0][c];
        }
        for (int k = 0; k < 9; k++) { Rm[k] = R[k * (size_t)T + e]; G[k] = dinv[k * (size_t)T + e]; }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                RD[3 * i + j] = Rm[i] * D[j] + Rm[3 + i] * D[3 + j] + Rm[6 + i] * D[6 + j] - (rs[j + 1][i] - rs[0][i]);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) H[3 * i + j] = RD[3 * i] * G[j] + RD[3 * i + 1] * G[3 + j] + RD[3 * i + 2] * G[6 + j];
        const float sxx = m.a_k1 * H[0] + m.a_nu * (H[4] + H[8]);
        const float syy = m.a_k1 * H[4] + m.a_nu * (H[0] + H[8]);
        const float szz = m.a_k1 * H[8] + m.a_nu * (H[0] + H[4]);
        const float sxy = m.a_k2 * (H[1] + H[3]);
        const float syz = m.a_k2 * (H[5] + H[7]);
        const float szx = m.a_k2 * (H[2] + H[6]);
        const float sH = sxx * H[0] + syy * H[4] + szz * H[8] + sxy * (H[1] + H[3]) + syz * (H[5] + H[7]) +
                         szx * (H[2] + H[6]);
        acc += 0.5 * (double)vol[e] * (double)sH;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) blk[blockIdx.x] = s;
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
    double acc = 0.0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        float ri = -1.0f * q[i] + b[i];  // muleq_addeq(r, -1, b)
        if (mask && mask[i / 3]) ri = 0.f;
        r[i] = ri;
        d[i] = ri;
        acc += (double)ri * ri;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) rr_part[blockIdx.x] = s;
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

// alpha for iteration k from per-block d.q partials (generic solver path)
__global__ __launch_bounds__(kBlock) void k_cg_alpha(CgScalars* st, uint32_t k, const double* part, int n) {
    __shared__ double sh[kBlock / 64];
    if (st->done[k]) return;
    const double dq = reduce_partials(part, n, sh);
    if (threadIdx.x == 0) {
        st->dq[k] = dq;
        if (fabs(dq) < st->abs_sq) {
            st->reason[k + 1] = CF_CG_BREAKDOWN;
            st->done[k + 1] = 2;
        } else {
            st->alpha[k] = (float)(st->delta[k] / dq);
        }
    }
}

// x += alpha d ; r -= alpha q ; per-block partials of r^T r (ILSS_CG.h:100-117)
template <bool V4>
__global__ __launch_bounds__(kBlock) void k_cg_update_xr(uint32_t n, uint32_t k, CgScalars* st,
                                                         float* __restrict__ x, float* __restrict__ r,
                                                         const float* __restrict__ d, const float* __restrict__ q,
                                                         double* rr_part) {
    __shared__ double sh[kBlock / 64];
    if (st->done[k] || st->reason[k + 1] == CF_CG_BREAKDOWN) return;
    const float alpha = st->alpha[k];
    double acc = 0.0;
    const uint32_t gid = blockIdx.x * kBlock + threadIdx.x, stride = gridDim.x * kBlock;
    uint32_t tail0 = 0;
    if (V4) {
        const uint32_t n4 = n / 4;
        tail0 = 4 * n4;
        float4* x4 = reinterpret_cast<float4*>(x);
        float4* r4 = reinterpret_cast<float4*>(r);
        const float4* d4 = reinterpret_cast<const float4*>(d);
        const float4* q4 = reinterpret_cast<const float4*>(q);
        for (uint32_t i = gid; i < n4; i += stride) {
            const float4 di = d4[i], qi = q4[i];
            float4 xi = x4[i], ri = r4[i];
            xi.x = xi.x + di.x * alpha; xi.y = xi.y + di.y * alpha;
            xi.z = xi.z + di.z * alpha; xi.w = xi.w + di.w * alpha;
            ri.x = ri.x - qi.x * alpha; ri.y = ri.y - qi.y * alpha;
            ri.z = ri.z - qi.z * alpha; ri.w = ri.w - qi.w * alpha;
            x4[i] = xi;
            r4[i] = ri;
            acc += (double)ri.x * ri.x + (double)ri.y * ri.y + (double)ri.z * ri.z + (double)ri.w * ri.w;
        }
    }
    for (uint32_t i = tail0 + gid; i < n; i += stride) {
        x[i] = x[i] + d[i] * alpha;
        const float ri = r[i] - q[i] * alpha;
        r[i] = ri;
        acc += (double)ri * ri;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) rr_part[blockIdx.x] = s;
}

// delta_{k+1} = r^T r from per-block partials ; beta ; stop test (ILSS_CG.h:86-88,115-133)
__global__ __launch_bounds__(kBlock) void k_cg_beta(CgScalars* st, uint32_t k, const double* part, int n) {
    __shared__ double sh[kBlock / 64];
    if (st->done[k] || st->reason[k + 1] == CF_CG_BREAKDOWN) return;
    const double dn = reduce_partials(part, n, sh);
    if (threadIdx.x == 0) {
        st->delta[k + 1] = dn;
        st->beta[k] = (float)(dn / st->delta[k]);
        const uint32_t k1 = k + 1;
        int reason = 0;
        if (!(dn > st->abs_sq)) reason = CF_CG_CONVERGED_ABS;
        else if (!(dn > st->target)) reason = CF_CG_CONVERGED_REL;
        else if (!(k1 < st->max_iter)) reason = CF_CG_MAX_ITER;
        st->reason[k1] = reason;
        st->done[k1] = reason != 0;
    }
}

// d = S(beta d + r)
template <bool V4>
__global__ __launch_bounds__(kBlock) void k_cg_update_d(uint32_t n, uint32_t k, CgScalars* st,
                                                        float* __restrict__ d, const float* __restrict__ r,
                                                        const uint8_t* __restrict__ mask) {
    if (st->done[k]) {
        // keep every later iteration switched off once the solve has ended
        if (blockIdx.x == 0 && threadIdx.x == 0) st->done[k + 1] = 1;
        return;
    }
    if (st->done[k + 1]) return;  // converged or broke down at iteration k
    const float beta = st->beta[k];
    const uint32_t gid = blockIdx.x * kBlock + threadIdx.x, stride = gridDim.x * kBlock;
    uint32_t tail0 = 0;
    if (V4) {
        const uint32_t n4 = n / 4;
        tail0 = 4 * n4;
        float4* d4 = reinterpret_cast<float4*>(d);
        const float4* r4 = reinterpret_cast<const float4*>(r);
        for (uint32_t i = gid; i < n4; i += stride) {
            const float4 ri = r4[i];
            float4 di = d4[i];
            di.x = beta * di.x + ri.x; di.y = beta * di.y + ri.y;
            di.z = beta * di.z + ri.z; di.w = beta * di.w + ri.w;
            if (mask) {
                const uint32_t e = 4 * i;
                if (mask[e / 3]) di.x = 0.f;
                if (mask[(e + 1) / 3]) di.y = 0.f;
                if (mask[(e + 2) / 3]) di.z = 0.f;
                if (mask[(e + 3) / 3]) di.w = 0.f;
            }
            d4[i] = di;
        }
    }
    for (uint32_t i = tail0 + gid; i < n; i += stride) {
        float di = beta * d[i] + r[i];
        if (mask && mask[i / 3]) di = 0.f;
        d[i] = di;
    }
}

__global__ __launch_bounds__(kBlock) void k_integrate(uint32_t n, float* __restrict__ x, float* __restrict__ v,
                                                      const float* __restrict__ y, float dt, float drag) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const float yi = y[i];
        x[i] = x[i] + dt * yi;
        v[i] = (drag > 0.f) ? (1.0f - dt * drag) * yi : yi;
    }
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
                      uint32_t* blk_count, double* blk_sum, float* blk_min, int nblk, hipStream_t s) {
    hipLaunchKernelGGL(k_begin_eval, dim3(nblk), dim3(kBlock), 0, s, t.T, t.conn, x, t.dinv, rm, thr, R, detF,
                       blk_count, blk_sum, blk_min);
    return check_launch();
}

static size_t tile_smem(const TetDevice& t) {
    size_t b = sizeof(float4) * (2 * (size_t)t.max_tile_nodes + 4 * (size_t)t.max_tile_elems);
    b += sizeof(uint32_t) * ((size_t)t.max_tile_nodes + 1);
    return (b + 63) & ~size_t(63);
}

int launch_forces_skip(const TetDevice& t, int mode, const float* in, const float* R, Material m,
                       const int32_t* skip, hipStream_t s) {
    if (t.num_tiles == 0) return CF_OK;
    TileArgs a;
    a.T = t.T; a.num_tiles = t.num_tiles; a.max_te = t.max_tile_elems; a.max_tn = t.max_tile_nodes;
    a.tile_elem_ptr = t.tile_elem_ptr; a.tile_node_ptr = t.tile_node_ptr; a.tile_nodes = t.tile_nodes;
    a.part_dst = t.part_dst;
    a.slot_ptr = t.slot_ptr; a.dst = t.dst; a.lvid = t.lvid; a.R = R; a.dinv = t.dinv; a.vol = t.vol;
    a.ref = t.ref;
    a.in = in; a.partials = t.partials; a.m = m; a.skip = skip;
    const size_t sm = tile_smem(t);
    const uint32_t grid = t.tile_grid;
    const bool fe = mode == kModeFe;
#define CF_TILE_LAUNCH(TBV, PV)                                                                        \
    do {                                                                                              \
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
    const size_t sm = tile_smem(t);
    hipError_t e;
    switch (t.tile_threads) {
        case 1024: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tile_forces<kModeDfx, 1024, true>, 1024, sm); break;
        case 512: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tile_forces<kModeDfx, 512, true>, 512, sm); break;
        case 128: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tile_forces<kModeDfx, 128, true>, 128, sm); break;
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tile_forces<kModeDfx, 256, true>, 256, sm); break;
    }
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    const uint32_t per_xcd_tiles = (t.num_tiles + 7) / 8;
    uint32_t wg_per_xcd = (persistent == 1) ? uint32_t(per_cu) * uint32_t(cus) / 8 : per_xcd_tiles;
    if (wg_per_xcd < 1) wg_per_xcd = 1;
    if (wg_per_xcd > per_xcd_tiles) wg_per_xcd = per_xcd_tiles;
    return int(8 * wg_per_xcd);
}

int launch_forces(const TetDevice& t, int mode, const float* in, const float* R, Material m, hipStream_t s) {
    return launch_forces_skip(t, mode, in, R, m, nullptr, s);
}

int launch_gather_skip(const TetDevice& t, int mode, const NodeTerms& nt, int nblk, const int32_t* skip,
                       hipStream_t s) {
    GatherArgs a;
    a.N = t.N; a.node_part_ptr = t.node_part_ptr; a.partials = t.partials;
    a.t = nt; a.skip = skip;
    if (mode == kGatherAccumulate)
        hipLaunchKernelGGL(k_gather<kGatherAccumulate>, dim3(nblk), dim3(kBlock), 0, s, a);
    else if (mode == kGatherMatvec)
        hipLaunchKernelGGL(k_gather<kGatherMatvec>, dim3(nblk), dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL(k_gather<kGatherRhs>, dim3(nblk), dim3(kBlock), 0, s, a);
    return check_launch();
}

int launch_gather(const TetDevice& t, int mode, const NodeTerms& nt, int nblk, hipStream_t s) {
    return launch_gather_skip(t, mode, nt, nblk, nullptr, s);
}

int launch_energy(const TetDevice& t, const float* x, const float* R, Material m, double* blk, int nblk,
                  hipStream_t s) {
    hipLaunchKernelGGL(k_energy, dim3(nblk), dim3(kBlock), 0, s, t.T, t.conn, t.ref, x, R, t.dinv, t.vol, m, blk);
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

int launch_cg_alpha(CgScalars* st, uint32_t k, const double* dq_part, int nblk_dq, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(kBlock), 0, s, st, k, dq_part, nblk_dq);
    return check_launch();
}

int launch_cg_update_xr(uint32_t n, uint32_t k, CgScalars* st, float* x, float* r, const float* d, const float* q,
                        double* rr_part, hipStream_t s) {
    const int nb = vec_blocks(n);
    if (aligned16(x) && aligned16(r) && aligned16(d) && aligned16(q))
        hipLaunchKernelGGL(k_cg_update_xr<true>, dim3(nb), dim3(kBlock), 0, s, n, k, st, x, r, d, q, rr_part);
    else
        hipLaunchKernelGGL(k_cg_update_xr<false>, dim3(nb), dim3(kBlock), 0, s, n, k, st, x, r, d, q, rr_part);
    CF_TRY(check_launch());
    hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(kBlock), 0, s, st, k, rr_part, nb);
    return check_launch();
}

int launch_cg_update_d(uint32_t n, uint32_t k, CgScalars* st, float* d, const float* r, const uint8_t* mask,
                       hipStream_t s) {
    const int nb = vec_blocks(n);
    if (aligned16(d) && aligned16(r))
        hipLaunchKernelGGL(k_cg_update_d<true>, dim3(nb), dim3(kBlock), 0, s, n, k, st, d, r, mask);
    else
        hipLaunchKernelGGL(k_cg_update_d<false>, dim3(nb), dim3(kBlock), 0, s, n, k, st, d, r, mask);
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
