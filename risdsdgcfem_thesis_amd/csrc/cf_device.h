// Device code shared by the single-mesh kernels (cf_kernels.hip) and the
// batched kernels (cf_batch_kernels.hip): the reference-order FP32 math, the
// fixed-order FP64 reductions, the tile staging and the element force.
#pragma once
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

// Right singular vectors V0, V1, V2 of F (descending singular values): the
// eigenvectors of the normal equations F^T F.
__device__ inline void svd_right_vectors(const float* F, v3& V0, v3& V1, v3& V2) {
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
    if (flipped) { V0 = v3v; V1 = v2; V2 = v3{-v1.x, -v1.y, -v1.z}; }
    else { V0 = v1; V1 = v2; V2 = v3v; }
}

__device__ inline void svd_rotation(const float* F, float* R) {
#pragma clang fp contract(off)
    v3 V0, V1, V2;  // columns of V
    svd_right_vectors(F, V0, V1, V2);
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

// ---- degenerate elements (DAPD) --------------------------------------------

// gsl_poly_solve_cubic (GNU Scientific Library, poly/solve_cubic.c, the same
// in GSL 1.x and 2.x; the reference links the system GSL): real roots of
// x^3 + a x^2 + b x + c in double, Cardano for one, Viete's trigonometric form
// for three (ascending), the exact double and triple roots by the
// discriminant tests. Returns the number of roots.
__device__ inline int gsl_cubic(double a, double b, double c, double* x) {
#pragma clang fp contract(off)
    const double q = a * a - 3 * b;
    const double r = 2 * a * a * a - 9 * a * b + 27 * c;
    const double Q = q / 9, R = r / 54;
    const double Q3 = Q * Q * Q, R2 = R * R;
    const double CR2 = 729 * r * r, CQ3 = 2916 * q * q * q;
    if (R == 0 && Q == 0) {
        x[0] = x[1] = x[2] = -a / 3;
        return 3;
    }
    if (CR2 == CQ3) {
        const double sq = sqrt(Q);
        if (R > 0) { x[0] = -2 * sq - a / 3; x[1] = sq - a / 3; x[2] = sq - a / 3; }
        else { x[0] = -sq - a / 3; x[1] = -sq - a / 3; x[2] = 2 * sq - a / 3; }
        return 3;
    }
    if (R2 < Q3) {
        const double sgn = (R >= 0) ? 1.0 : -1.0;
        const double theta = acos(sgn * sqrt(R2 / Q3));
        const double norm = -2 * sqrt(Q);
        double r0 = norm * cos(theta / 3) - a / 3;
        double r1 = norm * cos((theta + 2.0 * M_PI) / 3) - a / 3;
        double r2 = norm * cos((theta - 2.0 * M_PI) / 3) - a / 3;
        double t;
        if (r0 > r1) { t = r0; r0 = r1; r1 = t; }
        if (r1 > r2) { t = r1; r1 = r2; r2 = t; if (r0 > r1) { t = r0; r0 = r1; r1 = t; } }
        x[0] = r0; x[1] = r1; x[2] = r2;
        return 3;
    }
    const double sgn = (R >= 0) ? 1.0 : -1.0;
    const double A = -sgn * pow(fabs(R) + sqrt(R2 - Q3), 1.0 / 3.0);
    x[0] = A + Q / A - a / 3;
    return 1;
}

// mal::GSolvePolynomialEq2 (Mal/GSolvePolynomialEq.h:11-38) in float;
// IsZero takes an absolute epsilon of 1e-4 (Mal/RealUtils.h:110-114).
__device__ inline int solve_eq2(float a, float b, float c, float& x1, float& x2) {
#pragma clang fp contract(off)
    if (fabsf(a) <= 0.0001f) {
        if (fabsf(b) <= 0.0001f) return 0;
        x1 = x2 = -c / b;
        return 1;
    }
    const float disc = b * b - 4 * a * c;
    if (disc < 0) return 0;
    const float root = sqrtf(disc);
    const float inv = 1.0f / (2.0f * a);
    x1 = (-b - root) * inv;
    x2 = (-b + root) * inv;
    if (x1 > x2) { const float t = x1; x1 = x2; x2 = t; }
    return (x1 == x2) ? 1 : 2;
}

// mal::GSolvePolynomialEq3 (ms/fem/TetrahedronElementSet3.cpp:50-69):
// coefficients normalised in float, the cubic in double, roots to float.
__device__ inline int solve_eq3(float a, float b, float c, float d, float* x) {
#pragma clang fp contract(off)
    if (a != 0.0f) {
        double r[3] = {0.0, 0.0, 0.0};
        const int n = gsl_cubic(double(b / a), double(c / a), double(d / a), r);
        x[0] = float(r[0]); x[1] = float(r[1]); x[2] = float(r[2]);
        return n;
    }
    x[2] = 0.0f;
    return solve_eq2(b, c, d, x[0], x[1]);
}

// One permutation term u_x v_y w_z of a 3x3 determinant whose columns p, q, r
// are taken from the edge-difference sets u, v, w.
__device__ inline float perm_term(const float (*u)[3], const float (*v)[3], const float (*w)[3], int p, int q,
                                  int r) {
#pragma clang fp contract(off)
    return u[p][0] * v[q][1] * w[r][2];
}
// det Ds(t) on the linear path Ds(t) = Da + t Dd as A t^3 + B t^2 + C t + D:
// each coefficient sums the six permutation terms + + + - - - (column
// triples 012, 120, 201, 210, 102, 021), a term taking its factors from Da
// (a) or Dd (d); in B (C) the single Da (Dd) factor moves from x to y to z.
__device__ inline float det_coef(const float (*a)[3], const float (*d)[3], int k, int p, int q, int r) {
#pragma clang fp contract(off)
    switch (k) {
        case 3: return perm_term(d, d, d, p, q, r);
        case 2: return perm_term(a, d, d, p, q, r) + perm_term(d, a, d, p, q, r) + perm_term(d, d, a, p, q, r);
        case 1: return perm_term(a, a, d, p, q, r) + perm_term(a, d, a, p, q, r) + perm_term(d, a, a, p, q, r);
        default: return perm_term(a, a, a, p, q, r);
    }
}
__device__ inline float det_poly(const float (*a)[3], const float (*d)[3], int k) {
#pragma clang fp contract(off)
    return det_coef(a, d, k, 0, 1, 2) + det_coef(a, d, k, 1, 2, 0) + det_coef(a, d, k, 2, 0, 1) -
           det_coef(a, d, k, 2, 1, 0) - det_coef(a, d, k, 1, 0, 2) - det_coef(a, d, k, 0, 2, 1);
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
// v += part[j], j = i, i + stride, ... < n, in that order. The loads
// of eight terms are issued before their adds: the partials come from blocks
// on every XCD, so most are misses in this XCD's L2, and one latency per
// eight terms replaces one per term. The additions keep their order.
__device__ inline double strided_sum_from(double v, const double* part, int i, int n, int stride) {
    for (; i + 7 * stride < n; i += 8 * stride) {
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; u++) t[u] = part[i + u * stride];
#pragma unroll
        for (int u = 0; u < 8; u++) v += t[u];
    }
    for (; i < n; i += stride) v += part[i];
    return v;
}
__device__ inline double strided_sum(const double* part, int first, int n, int stride) {
    return strided_sum_from(0.0, part, first, n, stride);
}
// The same sum with the terms of the first 24 strides all loaded before the
// first add (one latency for up to 24 terms; more registers).
__device__ inline double strided_sum_wide(const double* part, int first, int n, int stride) {
    constexpr int M = 24;
    double t[M];
#pragma unroll
    for (int u = 0; u < M; u++) {
        const int j = first + u * stride;
        t[u] = j < n ? part[j] : 0.0;
    }
    double v = 0.0;
#pragma unroll
    for (int u = 0; u < M; u++)
        if (first + u * stride < n) v += t[u];
    return strided_sum_from(v, part, first + M * stride, n, stride);
}
// Deterministic sum of n partials by one block (fixed order and tree).
__device__ inline double reduce_partials(const double* part, int n, double* sh) {
    return block_sum(strided_sum(part, threadIdx.x, n, kBlock), sh);
}
__device__ inline double reduce_partials_wide(const double* part, int n, double* sh) {
    return block_sum(strided_sum_wide(part, threadIdx.x, n, kBlock), sh);
}
// reduce_partials computed by one wavefront: lane l plays the threads l,
// l+64, l+128, l+192 of the block and the four wave sums are added in the
// block's order, so one wavefront per mesh reproduces the one-block sum bit
// for bit. All 64 lanes must be active.
__device__ inline double wave_reduce_partials(const double* part, int n) {
    const int l = threadIdx.x & 63;
    double r = 0.0;
    for (int w = 0; w < kBlock / 64; w++) r += wave_sum(strided_sum(part, l + 64 * w, n, kBlock));
    return r;
}

// a0 b0 + a1 b1 + a2 b2 as fma(a2, b2, fma(a1, b1, a0 b0)): a fixed rounding
// sequence, the same in every kernel.
__device__ inline float fma3(float a0, float b0, float a1, float b1, float a2, float b2) {
#pragma clang fp contract(off)
    return fmaf(a2, b2, fmaf(a1, b1, a0 * b0));
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
    const uint32_t* lv;
    const uint32_t* rk;
    const float* R;
    const float* dinv;
    const float* vol;
    const float* ref;
    const float* ref_tile;  // rest positions per tile node (tile-major), or null: gather from ref
    const float* in;
    float* partials;
    Material m;
    const int32_t* skip;
    // Walk each XCD's tile range from its end (one tile per workgroup only).
    // Consecutive K*x passes in opposite directions start on the tiles the
    // previous pass touched last, which the memory-side cache may still hold;
    // every tile writes the same partials either way.
    uint32_t reverse = 0;
    // When every tile but the last holds exactly uniform_te elements, tile t's
    // elements are [t uniform_te, ...) and their loads need not wait for the
    // tile's descriptor (0: read tile_elem_ptr).
    uint32_t uniform_te = 0;
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
    if (a.uniform_te) {
        d.e0 = t * a.uniform_te;
        d.ne = min(a.uniform_te, a.T - d.e0);
    } else {
        d.e0 = a.tile_elem_ptr[t];
        d.ne = a.tile_elem_ptr[t + 1] - d.e0;
    }
    d.n0 = a.tile_node_ptr[t];
    d.nn = a.tile_node_ptr[t + 1] - d.n0;
    return d;
}

// Per-thread registers of one tile: this thread's tile node (vector, rest
// position, slot offset) and element (local vertex ids, its rank in each
// vertex's slot run, R). The element's force slot at vertex j is the run start
// of node l[j] (staged in LDS) plus f[j].
struct TileRegs {
    float v[3], r[3], R[9];
    uint32_t sp, pd;
    uint32_t l[4];
    uint32_t f[4];
};

// Element i of an array addressed by a 32-bit byte offset from a uniform
// base: the load takes the base in scalar registers and one 32-bit vector
// offset, with no 64-bit address arithmetic per lane. Every per-mesh array
// the tile stage reads is below 4 GiB (T, N < 2^28 are checked on the host).
// The arrays are device allocations, so the load says global: where the base
// comes from memory (the batch's per-simulation descriptors) rather than from
// a kernel argument, the compiler would otherwise issue flat loads.
#define CF_GLOBAL __attribute__((address_space(1)))
template <typename V>
__device__ inline V ld32(const V* base, uint32_t i) {
    return *(const CF_GLOBAL V*)((const CF_GLOBAL char*)base + i * uint32_t(sizeof(V)));
}
// The same load marked non-temporal. The tile stage reads each element's R
// and packed ids (44 B) once per K*x pass, 231 MB on the 5.24M-tet mesh; read
// non-temporally they leave the caches to the node vectors (gathered by ~2.5
// tiles each), the tile-order rest positions and the partials the gather
// reads next: tile kernel 71.7 -> 65.5 us per K*x, QIE-v step 5.64 -> 5.57 ms.
// The small static per-tile-node streams and the gather's partials read this
// way were slower (step 6.1 ms): their 12-byte records share cache lines; so
// were the partials stored non-temporally (6.9 ms: the gather reads them next).
template <typename V>
__device__ inline V ld32_nt(const V* base, uint32_t i) {
    return __builtin_nontemporal_load((const CF_GLOBAL V*)((const CF_GLOBAL char*)base + i * uint32_t(sizeof(V))));
}

__device__ inline void load_tile_regs(const TileArgs& a, const TileDesc& d, uint32_t gid, uint32_t tid,
                                      TileRegs& g) {
    if (tid < d.nn) {
        g.v[0] = ld32(a.in, 3 * gid + 0); g.v[1] = ld32(a.in, 3 * gid + 1); g.v[2] = ld32(a.in, 3 * gid + 2);
        g.r[0] = ld32(a.ref, 3 * gid + 0); g.r[1] = ld32(a.ref, 3 * gid + 1); g.r[2] = ld32(a.ref, 3 * gid + 2);
    }
    if (tid <= d.nn) g.sp = ld32(a.slot_ptr, d.n0 + tid) - 4 * d.e0;
    if (tid < d.nn) g.pd = ld32(a.part_dst, d.n0 + tid);
    if (tid < d.ne) {
        const uint32_t e = d.e0 + tid;
        const uint32_t lv = ld32(a.lv, e), rk = ld32(a.rk, e);
#pragma unroll
        for (int j = 0; j < 4; j++) { g.l[j] = (lv >> (8 * j)) & 255u; g.f[j] = (rk >> (8 * j)) & 255u; }
#pragma unroll
        for (int k = 0; k < 9; k++) g.R[k] = ld32(a.R + k * (size_t)a.T, e);
    }
}

// The same registers for a tile whose loads start cold, in an order that
// keeps bytes in flight: the tile's node ids first, then the element stream
// (R and the packed ids/ranks) and the slot offsets while the ids are in
// flight, then the node gathers, which need the ids. Lanes past the tile's
// elements or nodes load a valid duplicate instead of branching around the
// loads (their values are not used).
__device__ inline void load_tile_regs_cold(const TileArgs& a, const TileDesc& d, uint32_t tid, TileRegs& g) {
    const uint32_t tn = min(tid, d.nn - 1);
    const uint32_t gid = ld32(a.tile_nodes, d.n0 + tn);
    const uint32_t e = d.e0 + min(tid, d.ne - 1);
    const uint32_t lv = ld32_nt(a.lv, e), rk = ld32_nt(a.rk, e);
#pragma unroll
    for (int k = 0; k < 9; k++) g.R[k] = ld32_nt(a.R + k * (size_t)a.T, e);
    g.sp = ld32(a.slot_ptr, d.n0 + min(tid, d.nn)) - 4 * d.e0;
    g.pd = ld32(a.part_dst, d.n0 + tn);
    if (a.ref_tile) {  // static: streamed in tile order, not gathered behind the ids
        const uint32_t q = 3 * (d.n0 + tn);
        g.r[0] = ld32(a.ref_tile, q + 0); g.r[1] = ld32(a.ref_tile, q + 1); g.r[2] = ld32(a.ref_tile, q + 2);
    }
    if (tid < d.nn) {
        g.v[0] = ld32(a.in, 3 * gid + 0); g.v[1] = ld32(a.in, 3 * gid + 1); g.v[2] = ld32(a.in, 3 * gid + 2);
        if (!a.ref_tile) {
            g.r[0] = ld32(a.ref, 3 * gid + 0); g.r[1] = ld32(a.ref, 3 * gid + 1); g.r[2] = ld32(a.ref, 3 * gid + 2);
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) { g.l[j] = (lv >> (8 * j)) & 255u; g.f[j] = (rk >> (8 * j)) & 255u; }
}

// Nodal forces of one element into its four tile force slots.
// Tile data in LDS as float planes (x, y, z): the tile's node vectors and rest
// positions (one entry per tile node) and its force slots (4 per element),
// 12 bytes per entry.
struct Slots {
    float* x;
    float* y;
    float* z;
};
__device__ inline Slots make_slots(float* base, uint32_t n) { return Slots{base, base + n, base + 2 * n}; }

// Physical LDS word of force slot s: bits 0-4 XORed with bits 5-9. Tile
// nodes are sorted by valence, so the slot runs summed by neighbouring lanes
// often start 32, 16 or 8 slots apart; unswizzled, those lanes would hit the
// same few LDS banks on every step of the run. The XOR stays inside each
// 32-slot block (the slot planes are multiples of 32 long).
__device__ inline uint32_t slot_swizzle(uint32_t s) { return s ^ ((s >> 5) & 31u); }


// LDS layout of a tile workgroup of TB threads (at most TB - 1 tile nodes and
// TB elements): node vectors and rest positions as planes of TB floats, the
// 4 TB force slots as planes of 4 TB floats (swizzled, slot_swizzle), then
// TB + 1 slot-run starts.
// Compile-time plane strides let one LDS address serve the x, y and z planes
// (immediate offsets).
template <int TB, int EXTRA = 0>
struct TileLds {
    static_assert(TB % 8 == 0, "slot planes must be whole 32-slot swizzle blocks");
    static constexpr uint32_t kFloats = (18 + EXTRA) * TB;
    static constexpr size_t kBytes = (sizeof(float) * kFloats + sizeof(uint32_t) * (TB + 1) + 63) & ~size_t(63);
    Slots v, r, F;
    float* extra;  // EXTRA planes of TB floats (per tile node)
    uint32_t* sp;
    __device__ explicit TileLds(float* base)
        : v(make_slots(base, TB)), r(make_slots(base + 3 * TB, TB)), F(make_slots(base + 6 * TB, 4 * TB)),
          extra(base + 18 * TB), sp(reinterpret_cast<uint32_t*>(base + kFloats)) {}
};

// Sum of one node's slot run [s0, s1) in ascending slot order; the loads of
// four slots are issued together, the adds stay sequential.
__device__ inline void sum_slot_run(const Slots& s_F, uint32_t s0, uint32_t s1, float& fx, float& fy, float& fz) {
#pragma clang fp contract(off)
    fx = 0.f; fy = 0.f; fz = 0.f;
    uint32_t i = s0;
    for (; i + 4 <= s1; i += 4) {
        float ax[4], ay[4], az[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t p = slot_swizzle(i + q);
            ax[q] = s_F.x[p]; ay[q] = s_F.y[p]; az[q] = s_F.z[p];
        }
#pragma unroll
        for (int q = 0; q < 4; q++) { fx += ax[q]; fy += ay[q]; fz += az[q]; }
    }
    for (; i < s1; i++) {
        const uint32_t p = slot_swizzle(i);
        fx += s_F.x[p]; fy += s_F.y[p]; fz += s_F.z[p];
    }
}

// C_WRP element force in Piola form: P (column c is the force on vertex c+1).
// The C_WRP forces of one element from its edge matrix D (columns v1-v0,
// v2-v0, v3-v0 of the input vectors), rest edges Dm, Dm^-1 (G) and volume V.
template <int MODE>
__device__ inline void wrp_forces_dg(const float D[9], const float Dm[9], const float G[9], float V, const float* R,
                                     const Material& m, float P[9]) {
    // Every fused multiply-add is written out (fma3); nothing else may be
    // contracted, so the result is fixed by this source in every kernel that
    // inlines it.
#pragma clang fp contract(off)
    float RD[9];  // R^T D  (f_e: minus the rest edges Dm, the local displacement differences)
#pragma unroll
    for (int i3 = 0; i3 < 3; i3++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            RD[3 * i3 + j] = fma3(R[0 + i3], D[0 + j], R[3 + i3], D[3 + j], R[6 + i3], D[6 + j]);
    if (MODE == kModeFe) {
#pragma unroll
        for (int k = 0; k < 9; k++) RD[k] -= Dm[k];
    }
    float H[9];  // RD G
#pragma unroll
    for (int i3 = 0; i3 < 3; i3++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            H[3 * i3 + j] = fma3(RD[3 * i3 + 0], G[0 + j], RD[3 * i3 + 1], G[3 + j], RD[3 * i3 + 2], G[6 + j]);
    const float sxx = fmaf(m.a_k1, H[0], m.a_nu * (H[4] + H[8]));
    const float syy = fmaf(m.a_k1, H[4], m.a_nu * (H[0] + H[8]));
    const float szz = fmaf(m.a_k1, H[8], m.a_nu * (H[0] + H[4]));
    const float sxy = m.a_k2 * (H[1] + H[3]);
    const float syz = m.a_k2 * (H[5] + H[7]);
    const float szx = m.a_k2 * (H[2] + H[6]);
    const float S[9] = {sxx, sxy, szx, sxy, syy, syz, szx, syz, szz};
    float W[9];  // -V sigma G^T
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            W[3 * r + c] = -V * fma3(S[3 * r + 0], G[3 * c + 0], S[3 * r + 1], G[3 * c + 1], S[3 * r + 2], G[3 * c + 2]);
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            P[3 * r + c] = fma3(R[3 * r + 0], W[0 + c], R[3 * r + 1], W[3 + c], R[3 * r + 2], W[6 + c]);
}

template <int MODE>
__device__ inline void wrp_forces(const Slots& s_v, const Slots& s_r, const uint32_t* lv, const float* R,
                                  const Material& m, float P[9]) {
#pragma clang fp contract(off)
    const float3 x0 = make_float3(s_v.x[lv[0]], s_v.y[lv[0]], s_v.z[lv[0]]);
    const float3 x1 = make_float3(s_v.x[lv[1]], s_v.y[lv[1]], s_v.z[lv[1]]);
    const float3 x2 = make_float3(s_v.x[lv[2]], s_v.y[lv[2]], s_v.z[lv[2]]);
    const float3 x3 = make_float3(s_v.x[lv[3]], s_v.y[lv[3]], s_v.z[lv[3]]);
    const float3 r0 = make_float3(s_r.x[lv[0]], s_r.y[lv[0]], s_r.z[lv[0]]);
    const float3 r1 = make_float3(s_r.x[lv[1]], s_r.y[lv[1]], s_r.z[lv[1]]);
    const float3 r2 = make_float3(s_r.x[lv[2]], s_r.y[lv[2]], s_r.z[lv[2]]);
    const float3 r3 = make_float3(s_r.x[lv[3]], s_r.y[lv[3]], s_r.z[lv[3]]);
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
    wrp_forces_dg<MODE>(D, Dm, G, V, R, m, P);
}

template <int MODE>
__device__ inline void element_force(const Slots& s_v, const Slots& s_r, const Slots& s_F, const uint32_t* lv,
                                     const uint32_t* fs, const float* R, const Material& m) {
#pragma clang fp contract(off)
    float P[9];  // R W: column c is the force on vertex c+1
    wrp_forces<MODE>(s_v, s_r, lv, R, m, P);
    const uint32_t f0 = slot_swizzle(fs[0]), f1 = slot_swizzle(fs[1]), f2 = slot_swizzle(fs[2]),
                   f3 = slot_swizzle(fs[3]);
    s_F.x[f1] = P[0]; s_F.y[f1] = P[3]; s_F.z[f1] = P[6];
    s_F.x[f2] = P[1]; s_F.y[f2] = P[4]; s_F.z[f2] = P[7];
    s_F.x[f3] = P[2]; s_F.y[f3] = P[5]; s_F.z[f3] = P[8];
    s_F.x[f0] = -(P[0] + P[1] + P[2]); s_F.y[f0] = -(P[3] + P[4] + P[5]); s_F.z[f0] = -(P[6] + P[7] + P[8]);
}


// C_f_d (TetrahedronElementSet3.cpp:1098-1150) of one element into its force
// slots: f_j = c1 (C_WRP stiffness force of the velocities)_j - R (s_m[j] R^T v_j),
// with s_m holding c0 m per tile node and s_v the velocities.
__device__ inline void element_force_fd(const Slots& s_v, const Slots& s_r, const Slots& s_F, const float* s_m,
                                        const uint32_t* lv, const uint32_t* fs, const float* R, const Material& m,
                                        float c1) {
#pragma clang fp contract(off)
    float P[9];
    wrp_forces<kModeDfx>(s_v, s_r, lv, R, m, P);
    float fk[4][3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        fk[1][i] = c1 * P[3 * i];
        fk[2][i] = c1 * P[3 * i + 1];
        fk[3][i] = c1 * P[3 * i + 2];
        fk[0][i] = c1 * (-(P[3 * i] + P[3 * i + 1] + P[3 * i + 2]));
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t l = lv[j];
        const float v0 = s_v.x[l], v1 = s_v.y[l], v2 = s_v.z[l];
        const float sm = s_m[l];
        float mv[3];
#pragma unroll
        for (int i = 0; i < 3; i++) mv[i] = sm * fma3(R[i], v0, R[3 + i], v1, R[6 + i], v2);
        const uint32_t f = slot_swizzle(fs[j]);
        s_F.x[f] = fk[j][0] - fma3(R[0], mv[0], R[1], mv[1], R[2], mv[2]);
        s_F.y[f] = fk[j][1] - fma3(R[3], mv[0], R[4], mv[1], R[5], mv[2]);
        s_F.z[f] = fk[j][2] - fma3(R[6], mv[0], R[7], mv[1], R[8], mv[2]);
    }
}

// Compute_dR (ms/fem/Rotation.cpp:622-644, McAdams 2011) in the reference's
// order: S = R^T F, W = R^T dF, G = S - tr(S) I, w = (W12 - W21, W20 - W02,
// W01 - W10); dR = R skew(G^-1 w) when |det G| > 1e-4, else no dR.
__device__ inline bool compute_dR(const float dF[9], const float F[9], const float R[9], float dR[9]) {
#pragma clang fp contract(off)
    float S[9], W[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float a = 0.0f, b = 0.0f;
            a += R[i] * F[j]; a += R[3 + i] * F[3 + j]; a += R[6 + i] * F[6 + j];
            b += R[i] * dF[j]; b += R[3 + i] * dF[3 + j]; b += R[6 + i] * dF[6 + j];
            S[3 * i + j] = a;
            W[3 * i + j] = b;
        }
    float tr = 0.0f;
    tr += S[0]; tr += S[4]; tr += S[8];
    float G[9];
#pragma unroll
    for (int k = 0; k < 9; k++) G[k] = S[k] - tr * ((k % 4 == 0) ? 1.0f : 0.0f);
    const float w0 = W[5] - W[7], w1 = W[6] - W[2], w2 = W[1] - W[3];
    const float d = refmath::det3(G);
    if (!((d > 0.0f ? d : -d) > 0.0001f)) return false;
    float Gi[9];
    refmath::inv3(G, d, Gi);
    float r[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float a = 0.0f;
        a += Gi[3 * i] * w0; a += Gi[3 * i + 1] * w1; a += Gi[3 * i + 2] * w2;
        r[i] = a;
    }
    const float K[9] = {0.0f, -r[2], r[1], r[2], 0.0f, -r[0], -r[1], r[0], 0.0f};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float a = 0.0f;
            a += R[3 * i] * K[j]; a += R[3 * i + 1] * K[3 + j]; a += R[3 * i + 2] * K[6 + j];
            dR[3 * i + j] = a;
        }
    return true;
}

// Isotropic stress of the Ke-based C_WRP material for a (non-symmetric)
// displacement gradient H (TetrahedronElement3::Compute_E, engineering shear).
__device__ inline void wrp_stress(const Material& m, const float H[9], float S[9]) {
#pragma clang fp contract(off)
    const float sxx = fmaf(m.a_k1, H[0], m.a_nu * (H[4] + H[8]));
    const float syy = fmaf(m.a_k1, H[4], m.a_nu * (H[0] + H[8]));
    const float szz = fmaf(m.a_k1, H[8], m.a_nu * (H[0] + H[4]));
    const float sxy = m.a_k2 * (H[1] + H[3]);
    const float syz = m.a_k2 * (H[5] + H[7]);
    const float szx = m.a_k2 * (H[2] + H[6]);
    S[0] = sxx; S[1] = sxy; S[2] = szx; S[3] = sxy; S[4] = syy; S[5] = syz; S[6] = szx; S[7] = syz; S[8] = szz;
}

// C_df_x with DM = Exact (TetrahedronElementSet3.cpp:1761-1836) in Piola form:
// df = R W(R^T dF + dR^T F) + dR W(R^T F - I), W(H) = -V sigma(H) Dm^-T, where
// the first term is the truncated K*x plus -R K dR^T x and the second is
// -dR K u; F is the evaluation's (cached), dR from Compute_dR.
__device__ inline void element_force_exact(const Slots& s_v, const Slots& s_r, const Slots& s_F, const uint32_t* lv,
                                           const uint32_t* fs, const float* R, const Material& m, const float* Fe,
                                           bool allow_dR) {
#pragma clang fp contract(off)
    const float3 r0 = make_float3(s_r.x[lv[0]], s_r.y[lv[0]], s_r.z[lv[0]]);
    const float3 r1 = make_float3(s_r.x[lv[1]], s_r.y[lv[1]], s_r.z[lv[1]]);
    const float3 r2 = make_float3(s_r.x[lv[2]], s_r.y[lv[2]], s_r.z[lv[2]]);
    const float3 r3 = make_float3(s_r.x[lv[3]], s_r.y[lv[3]], s_r.z[lv[3]]);
    const float Dm[9] = {r1.x - r0.x, r2.x - r0.x, r3.x - r0.x,
                         r1.y - r0.y, r2.y - r0.y, r3.y - r0.y,
                         r1.z - r0.z, r2.z - r0.z, r3.z - r0.z};
    const float detDm = refmath::det3(Dm);
    const float V = detDm / 6.0f;
    float Gi[9];
    refmath::inv3(Dm, detDm, Gi);
    const float3 x0 = make_float3(s_v.x[lv[0]], s_v.y[lv[0]], s_v.z[lv[0]]);
    const float3 x1 = make_float3(s_v.x[lv[1]], s_v.y[lv[1]], s_v.z[lv[1]]);
    const float3 x2 = make_float3(s_v.x[lv[2]], s_v.y[lv[2]], s_v.z[lv[2]]);
    const float3 x3 = make_float3(s_v.x[lv[3]], s_v.y[lv[3]], s_v.z[lv[3]]);
    const float D[9] = {x1.x - x0.x, x2.x - x0.x, x3.x - x0.x,
                        x1.y - x0.y, x2.y - x0.y, x3.y - x0.y,
                        x1.z - x0.z, x2.z - x0.z, x3.z - x0.z};
    float dF[9], H1[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float a = 0.0f;
            a += D[3 * i] * Gi[j]; a += D[3 * i + 1] * Gi[3 + j]; a += D[3 * i + 2] * Gi[6 + j];
            dF[3 * i + j] = a;
        }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) H1[3 * i + j] = fma3(R[i], dF[j], R[3 + i], dF[3 + j], R[6 + i], dF[6 + j]);
    float dR[9];
    const bool ok = allow_dR && compute_dR(dF, Fe, R, dR);
    float H2[9];
    if (ok) {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                H1[3 * i + j] += fma3(dR[i], Fe[j], dR[3 + i], Fe[3 + j], dR[6 + i], Fe[6 + j]);
                H2[3 * i + j] = fma3(R[i], Fe[j], R[3 + i], Fe[3 + j], R[6 + i], Fe[6 + j]) - ((i == j) ? 1.0f : 0.0f);
            }
    }
    float S[9], W[9], P[9];
    wrp_stress(m, H1, S);
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            W[3 * r + c] = -V * fma3(S[3 * r + 0], Gi[3 * c + 0], S[3 * r + 1], Gi[3 * c + 1], S[3 * r + 2], Gi[3 * c + 2]);
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            P[3 * r + c] = fma3(R[3 * r + 0], W[0 + c], R[3 * r + 1], W[3 + c], R[3 * r + 2], W[6 + c]);
    if (ok) {
        wrp_stress(m, H2, S);
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++)
                W[3 * r + c] = -V * fma3(S[3 * r + 0], Gi[3 * c + 0], S[3 * r + 1], Gi[3 * c + 1], S[3 * r + 2], Gi[3 * c + 2]);
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++)
                P[3 * r + c] += fma3(dR[3 * r + 0], W[0 + c], dR[3 * r + 1], W[3 + c], dR[3 * r + 2], W[6 + c]);
    }
    const uint32_t f0 = slot_swizzle(fs[0]), f1 = slot_swizzle(fs[1]), f2 = slot_swizzle(fs[2]),
                   f3 = slot_swizzle(fs[3]);
    s_F.x[f1] = P[0]; s_F.y[f1] = P[3]; s_F.z[f1] = P[6];
    s_F.x[f2] = P[1]; s_F.y[f2] = P[4]; s_F.z[f2] = P[7];
    s_F.x[f3] = P[2]; s_F.y[f3] = P[5]; s_F.z[f3] = P[8];
    s_F.x[f0] = -(P[0] + P[1] + P[2]); s_F.y[f0] = -(P[3] + P[4] + P[5]); s_F.z[f0] = -(P[6] + P[7] + P[8]);
}

// LCM / LCMH nodal forces of one element into its four tile force slots
// (C_LCM_f_e :1155-1178, C_LCM_df_x :1838-1872, C_LCMH_df_x :1874-1929, with
// DM = Truncated), in the reference's operation order: f_e from the F of the
// evaluation, P = R (2 mu (S - I) + lambda tr(S - I) I), S = R^T F; df_x with
// dF = dDs Dm^-1, dP = 2 mu dF + lambda tr(R^T dF) R, or for LCMH
// 2 mu (w dF + (1 - w) R sym(R^T dF)) + lambda tr(R^T dF) R; then
// H = -V P Dm^-T, column c to vertex c+1 and -(sum) to vertex 0. Dm^-1 and V
// are rebuilt from the rest positions in LDS as for C_WRP.
template <int MODE>
__device__ inline void element_force_lcm(const Slots& s_v, const Slots& s_r, const Slots& s_F, const uint32_t* lv,
                                         const uint32_t* fs, const float* R, const Material& m, const float* Fe,
                                         float w, bool allow_dR = false) {
#pragma clang fp contract(off)
    const float3 r0 = make_float3(s_r.x[lv[0]], s_r.y[lv[0]], s_r.z[lv[0]]);
    const float3 r1 = make_float3(s_r.x[lv[1]], s_r.y[lv[1]], s_r.z[lv[1]]);
    const float3 r2 = make_float3(s_r.x[lv[2]], s_r.y[lv[2]], s_r.z[lv[2]]);
    const float3 r3 = make_float3(s_r.x[lv[3]], s_r.y[lv[3]], s_r.z[lv[3]]);
    const float Dm[9] = {r1.x - r0.x, r2.x - r0.x, r3.x - r0.x,
                         r1.y - r0.y, r2.y - r0.y, r3.y - r0.y,
                         r1.z - r0.z, r2.z - r0.z, r3.z - r0.z};
    const float detDm = refmath::det3(Dm);
    const float V = detDm / 6.0f;
    float G[9];
    refmath::inv3(Dm, detDm, G);
    float P[9];
    if (MODE == kModeFe) {
        float SmI[9];
        float tr = 0.0f;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                float a = 0.0f;
                a += R[i] * Fe[j]; a += R[3 + i] * Fe[3 + j]; a += R[6 + i] * Fe[6 + j];
                SmI[3 * i + j] = a - ((i == j) ? 1.0f : 0.0f);
            }
        tr += SmI[0]; tr += SmI[4]; tr += SmI[8];
        const float c = 2 * m.mu, lt = m.lambda * tr;
        float Q[9];
#pragma unroll
        for (int k = 0; k < 9; k++) Q[k] = c * SmI[k] + lt * ((k % 4 == 0) ? 1.0f : 0.0f);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                float a = 0.0f;
                a += R[3 * i] * Q[j]; a += R[3 * i + 1] * Q[3 + j]; a += R[3 * i + 2] * Q[6 + j];
                P[3 * i + j] = a;
            }
    } else {
        const float3 x0 = make_float3(s_v.x[lv[0]], s_v.y[lv[0]], s_v.z[lv[0]]);
        const float3 x1 = make_float3(s_v.x[lv[1]], s_v.y[lv[1]], s_v.z[lv[1]]);
        const float3 x2 = make_float3(s_v.x[lv[2]], s_v.y[lv[2]], s_v.z[lv[2]]);
        const float3 x3 = make_float3(s_v.x[lv[3]], s_v.y[lv[3]], s_v.z[lv[3]]);
        const float D[9] = {x1.x - x0.x, x2.x - x0.x, x3.x - x0.x,
                            x1.y - x0.y, x2.y - x0.y, x3.y - x0.y,
                            x1.z - x0.z, x2.z - x0.z, x3.z - x0.z};
        float dF[9], RtdF[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                float a = 0.0f;
                a += D[3 * i] * G[j]; a += D[3 * i + 1] * G[3 + j]; a += D[3 * i + 2] * G[6 + j];
                dF[3 * i + j] = a;
            }
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                float a = 0.0f;
                a += R[i] * dF[j]; a += R[3 + i] * dF[3 + j]; a += R[6 + i] * dF[6 + j];
                RtdF[3 * i + j] = a;
            }
        float tr = 0.0f;
        tr += RtdF[0]; tr += RtdF[4]; tr += RtdF[8];
        const float c = 2 * m.mu, lt = m.lambda * tr;
        if (m.w) {
            float Sy[9], B[9];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) Sy[3 * i + j] = 0.5f * (RtdF[3 * i + j] + RtdF[3 * j + i]);
            const float ow = 1 - w;
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    float a = 0.0f;
                    a += (ow * R[3 * i]) * Sy[j]; a += (ow * R[3 * i + 1]) * Sy[3 + j]; a += (ow * R[3 * i + 2]) * Sy[6 + j];
                    B[3 * i + j] = w * dF[3 * i + j] + a;
                }
#pragma unroll
            for (int k = 0; k < 9; k++) P[k] = c * B[k] + lt * R[k];
        } else {
#pragma unroll
            for (int k = 0; k < 9; k++) P[k] = c * dF[k] + lt * R[k];
        }
        float dR[9];
        if (allow_dR && compute_dR(dF, Fe, R, dR)) {
            // dP += (lambda tr(S - I) - 2 mu) dR (C_LCM_df_x :1862-1864), S = R^T F
            float trs = 0.0f;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                float a = 0.0f;
                a += R[i] * Fe[i]; a += R[3 + i] * Fe[3 + i]; a += R[6 + i] * Fe[6 + i];
                trs += a - 1.0f;
            }
            const float coef = m.lambda * trs - 2 * m.mu;
#pragma unroll
            for (int k = 0; k < 9; k++) P[k] = P[k] + coef * dR[k];
        }
    }
    float H[9];  // (-V P) Dm^-T
    const float nv = -V;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float a = 0.0f;
            a += (nv * P[3 * i]) * G[3 * j]; a += (nv * P[3 * i + 1]) * G[3 * j + 1]; a += (nv * P[3 * i + 2]) * G[3 * j + 2];
            H[3 * i + j] = a;
        }
    const uint32_t f0 = slot_swizzle(fs[0]), f1 = slot_swizzle(fs[1]), f2 = slot_swizzle(fs[2]),
                   f3 = slot_swizzle(fs[3]);
    s_F.x[f1] = H[0]; s_F.y[f1] = H[3]; s_F.z[f1] = H[6];
    s_F.x[f2] = H[1]; s_F.y[f2] = H[4]; s_F.z[f2] = H[7];
    s_F.x[f3] = H[2]; s_F.y[f3] = H[5]; s_F.z[f3] = H[8];
    s_F.x[f0] = (-H[0] - H[1]) - H[2]; s_F.y[f0] = (-H[3] - H[4]) - H[5]; s_F.z[f0] = (-H[6] - H[7]) - H[8];
}

// ============================================================ block bodies
// Each body is the work of block `blk` of a grid of `nblk` blocks over one
// mesh (grid-stride loops over that grid). The single-mesh kernels call them
// with (blockIdx.x, gridDim.x); the batched kernels with the block's place in
// its own mesh's grid, so a mesh's arithmetic, and the order of every sum, are
// the same alone and inside a batch.

// BeginEvaluation (ms/fem/TetrahedronElementSet3.cpp:219-281, rotations
// :748-892): F = Ds Dm^-1, det F, R by the rotation method; per-block counts
// of det F <= 0 (and their sum), of DAPD R = Id fallbacks, and min det F.
// blk_count holds [0,nblk) degenerate counts and [nblk,2nblk) fallbacks.
// ComputeDoC (ms/fem/TetrahedronElementSet3.cpp:589-743) for the path a -> b
// of one element: the time of collapse toc in (0,1] where det Ds(t) reaches
// thr * 6 V, the reference's collapse direction N at toc (row 2 of V^T,
// Ds = U S V^T), then vertex-face when a face's N-projected area is half the
// sum of all four (absolute epsilon 1e-4, the last such face wins), else the
// first edge pair (0j, kl) whose triangle normals around edge 0j disagree.
// Returns the DoC byte (4*vit0 + vit1, kDocInvalid without a crossing).
__device__ inline uint8_t compute_doc(const float (*a)[3], const float (*b)[3], float volume, float thr) {
#pragma clang fp contract(off)
    float d[4][3], dd[3][3], da[3][3];
    for (int i = 0; i < 4; i++)
        for (int c = 0; c < 3; c++) d[i][c] = b[i][c] - a[i][c];
    for (int i = 0; i < 3; i++)
        for (int c = 0; c < 3; c++) { dd[i][c] = d[i + 1][c] - d[0][c]; da[i][c] = a[i + 1][c] - a[0][c]; }
    const float K = thr * 6.0f * volume;
    float roots[3] = {0.0f, 0.0f, 0.0f};
    const int n = refmath::solve_eq3(refmath::det_poly(da, dd, 3), refmath::det_poly(da, dd, 2),
                                     refmath::det_poly(da, dd, 1), refmath::det_poly(da, dd, 0) - K, roots);
    if (n <= 0) return kDocInvalid;
    float toc = 2.0f;
    int n01 = 0;
    for (int i = 0; i < n; i++)
        if (roots[i] > 0.0f && roots[i] <= 1.0f && roots[i] < toc) { toc = roots[i]; n01++; }
    if (n01 != 1 && n01 != 3) return kDocInvalid;
    refmath::v3 q[4];
    for (int i = 0; i < 4; i++)
        q[i] = refmath::v3{a[i][0] + toc * d[i][0], a[i][1] + toc * d[i][1], a[i][2] + toc * d[i][2]};
    auto sub = [](refmath::v3 u, refmath::v3 v) { return refmath::v3{u.x - v.x, u.y - v.y, u.z - v.z}; };
    const refmath::v3 q10 = sub(q[1], q[0]), q20 = sub(q[2], q[0]), q30 = sub(q[3], q[0]);
    const float Q[9] = {q10.x, q20.x, q30.x, q10.y, q20.y, q30.y, q10.z, q20.z, q30.z};
    refmath::v3 V0, V1, N;
    refmath::svd_right_vectors(Q, V0, V1, N);
    float area[4];
    area[0] = fabsf(refmath::dot(N, refmath::cross(sub(q[2], q[1]), sub(q[3], q[1]))));
    area[1] = fabsf(refmath::dot(N, refmath::cross(q20, q30)));
    area[2] = fabsf(refmath::dot(N, refmath::cross(q10, q30)));
    area[3] = fabsf(refmath::dot(N, refmath::cross(q10, q20)));
    const float sum = area[0] + area[1] + area[2] + area[3];
    uint8_t doc = kDocInvalid;
    for (int i = 0; i < 4; i++)
        if (fabsf(area[i] - 0.5f * sum) <= 0.0001f) doc = uint8_t(5 * i);
    if (doc != kDocInvalid) return doc;
    if (refmath::dot(refmath::cross(q10, q20), refmath::cross(q10, q30)) < 0.0f) return 1;
    if (refmath::dot(refmath::cross(q20, q10), refmath::cross(q20, q30)) < 0.0f) return 2;
    if (refmath::dot(refmath::cross(q30, q10), refmath::cross(q30, q20)) < 0.0f) return 3;
    return 0;
}

// Compute_R_DAPD_Degenerate (ms/fem/Rotation.cpp:364-545): the collapsing
// vertex (vertex-face DoC) or edge (edge-edge DoC) moves along the unit
// collapse axis by lambda_u, a cubic Hermite in w / dh for 0 < w and linear
// (plus the optional power term) for w <= 0, dh = thr 6 V / |axis|; R is the
// Newton polar factor of the projected F. Returns false (R untouched) when the
// axis is shorter than 1e-3.
__device__ inline bool r_dapd_degenerate(float (*x)[3], float detF, float thr, uint8_t doc, float volume,
                                         const float* inv_dm, float fL, float fNL, float eNL, float* R) {
#pragma clang fp contract(off)
    using refmath::v3;
    auto P = [&](int i) { return v3{x[i][0], x[i][1], x[i][2]}; };
    auto sub = [](v3 u, v3 v) { return v3{u.x - v.x, u.y - v.y, u.z - v.z}; };
    const int v0 = doc >> 2, v1 = doc & 3;
    int i1, i2;  // the axis' reference vertex (vertex-face: x1; edge-edge: x2)
    v3 axis;
    if (v0 == v1) {
        const int a1 = (v0 + 1) & 3, a2 = (v0 + 2) & 3, a3 = (v0 + 3) & 3;
        axis = refmath::cross(sub(P(a2), P(a1)), sub(P(a3), P(a1)));
        i1 = a1; i2 = a1;
    } else {
        const int k = (v1 == 1) ? 2 : 1, l = (v1 == 3) ? 2 : 3;
        axis = refmath::cross(sub(P(v1), P(v0)), sub(P(l), P(k)));
        i1 = v1; i2 = k;
    }
    const float len_sq = refmath::dot(axis, axis);
    if (!(len_sq > 0.000001f)) return false;
    const float len = sqrtf(len_sq);
    const float inv = 1.0f / len;
    axis = v3{inv * axis.x, inv * axis.y, inv * axis.z};
    const float dh = thr * 6.0f * volume / len;
    float w = refmath::dot(sub(P(v0), P(i2)), axis);
    if (w * detF < 0.0f) { w = -w; axis = v3{-axis.x, -axis.y, -axis.z}; }
    const float s = w / dh;
    const float t = 1.0f - s;
    const float lambda = (w > 0.0f) ? fL * dh * ((2.0f - t) * t * t) : fL * dh * (t + fNL * powf(-s, eNL));
    const v3 mv{lambda * axis.x, lambda * axis.y, lambda * axis.z};
    x[v0][0] = x[v0][0] + mv.x; x[v0][1] = x[v0][1] + mv.y; x[v0][2] = x[v0][2] + mv.z;
    if (v0 != v1) { x[i1][0] = x[i1][0] + mv.x; x[i1][1] = x[i1][1] + mv.y; x[i1][2] = x[i1][2] + mv.z; }
    float Ds[9], F[9];
    for (int i = 0; i < 3; i++) {
        Ds[3 * i + 0] = x[1][i] - x[0][i];
        Ds[3 * i + 1] = x[2][i] - x[0][i];
        Ds[3 * i + 2] = x[3][i] - x[0][i];
    }
    refmath::mul3(Ds, inv_dm, F);
    refmath::polar_newton(F, R);
    return true;
}

// The degenerate branch of BeginEvaluation for one element (det F <= thr),
// kept out of line so that the common path's registers are not sized for it:
// UpdateDegeneration (TetrahedronElementSet3.cpp:477-532) — a new DoC when the
// element was undegenerate at prev_x, else the DoC is kept — and, under DAPD,
// the rotation (ComputeRotations_DAPD :868-892), written to R. Returns 1 when
// DAPD fell back to R = Id (no DoC, or a collapsed axis).
// TAG gives each calling kernel its own copy: a shared non-inlined function
// is allocated for the most registers any caller allows.
template <int TAG = 0>
__device__ __noinline__ uint32_t degenerate_element(uint32_t e, uint32_t T, const uint32_t* conn, const float* x,
                                                    const float* prev_x, uint8_t* doc, const float* vol,
                                                    const float* dinv, int rm, float thr, float detF, float fL,
                                                    float fNL, float eNL, float* R) {
#pragma clang fp contract(off)
    uint32_t g[4];
    for (int j = 0; j < 4; j++) g[j] = conn[j * (size_t)T + e];
    float G[9];
    for (int k = 0; k < 9; k++) G[k] = dinv[k * (size_t)T + e];
    float xs[4][3], ps[4][3];
    for (int j = 0; j < 4; j++)
        for (int c = 0; c < 3; c++) {
            xs[j][c] = x[3 * (size_t)g[j] + c];
            ps[j][c] = prev_x[3 * (size_t)g[j] + c];
        }
    const float V = vol[e];
    uint8_t dc = doc[e];
    {
        float Ds[9], F[9];
        for (int i = 0; i < 3; i++) {
            Ds[3 * i + 0] = ps[1][i] - ps[0][i];
            Ds[3 * i + 1] = ps[2][i] - ps[0][i];
            Ds[3 * i + 2] = ps[3][i] - ps[0][i];
        }
        refmath::mul3(Ds, G, F);
        if (refmath::det3(F) > thr) {
            const uint8_t nd = compute_doc(ps, xs, V, thr);
            if (nd != dc) { doc[e] = nd; dc = nd; }
        }
    }
    if (rm != CF_RM_DAPD) return 0;
    float Rm[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
    bool ok = false;
    if ((dc >> 2) <= (dc & 3)) ok = r_dapd_degenerate(xs, detF, thr, dc, V, G, fL, fNL, eNL, Rm);
    if (!ok) {
        for (int k = 0; k < 9; k++) Rm[k] = (k % 4 == 0) ? 1.f : 0.f;
    }
    for (int k = 0; k < 9; k++) R[k * (size_t)T + e] = Rm[k];
    return ok ? 0u : 1u;
}

// The element streams (connectivity, Dm^-1 in, R out: 88 B per tet) are read
// and written non-temporally, leaving the caches to the gathered positions
// (step 5.57 -> 5.53 ms on the 5.24M-tet mesh).
__device__ inline void begin_eval_block(uint32_t T, const uint32_t* __restrict__ conn, const float* __restrict__ x,
                                        const float* __restrict__ dinv, int rm, float thr, float* __restrict__ R,
                                        float* __restrict__ detF, uint32_t* blk_count, double* blk_sum,
                                        float* blk_min, uint32_t blk, uint32_t nblk, double* sh, float* shm,
                                        uint32_t* shc, uint32_t* shf, const LcmOut lcm = LcmOut(),
                                        const DegenArgs dg = DegenArgs()) {
#pragma clang fp contract(off)
    uint32_t cnt = 0, fb = 0;
    double sdet = 0.0;
    float mdet = 3.0e38f;
    for (uint32_t e = blk * kBlock + threadIdx.x; e < T; e += nblk * kBlock) {
        float xs[4][3];
        for (int j = 0; j < 4; j++) {
            const uint32_t g = __builtin_nontemporal_load(conn + j * (size_t)T + e);
            xs[j][0] = x[3 * (size_t)g]; xs[j][1] = x[3 * (size_t)g + 1]; xs[j][2] = x[3 * (size_t)g + 2];
        }
        float Ds[9], G[9], F[9];
        for (int i = 0; i < 3; i++) {
            Ds[3 * i + 0] = xs[1][i] - xs[0][i];
            Ds[3 * i + 1] = xs[2][i] - xs[0][i];
            Ds[3 * i + 2] = xs[3][i] - xs[0][i];
        }
        for (int k = 0; k < 9; k++) G[k] = __builtin_nontemporal_load(dinv + k * (size_t)T + e);
        refmath::mul3(Ds, G, F);
        const float d = refmath::det3(F);
        float Rm[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
        if (rm == CF_RM_MSVD) {
            refmath::svd_rotation(F, Rm);
        } else if (rm == CF_RM_DAPD) {
            if (d > thr) refmath::polar_newton(F, Rm);
        }
        // UpdateDegeneration (:477-532): undegenerate clears the DoC; the
        // degenerate branch (rare) updates it and, under DAPD, writes R itself
        bool r_stored = false;
        if (dg.doc) {
            if (d > thr) {
                if (dg.doc[e] != kDocInvalid) dg.doc[e] = kDocInvalid;
            } else {
                const uint32_t f = degenerate_element(e, T, conn, x, dg.prev_x ? dg.prev_x : x, dg.doc, dg.vol, dinv,
                                                      rm, thr, d, dg.l_factor, dg.nl_factor, dg.nl_exponent, R);
                if (rm == CF_RM_DAPD) {
                    fb += f;
                    r_stored = true;
                    for (int k = 0; k < 9; k++) Rm[k] = R[k * (size_t)T + e];
                }
            }
        } else if (rm == CF_RM_DAPD && !(d > thr)) {
            fb++;  // no DoC tracking: R = Id
        }
        if (!r_stored)
            for (int k = 0; k < 9; k++) __builtin_nontemporal_store(Rm[k], R + k * (size_t)T + e);
        detF[e] = d;
        if (lcm.F) {
            for (int k = 0; k < 9; k++) lcm.F[k * (size_t)T + e] = F[k];
        }
        if (lcm.w) {
            // LCMH weight (C_LCMH_df_x :1893-1899): S = R^T F, only its diagonal enters
            float wt = 1.0f;
            if (lcm.hybrid) {
                float tr = 0.0f;
                for (int i = 0; i < 3; i++) {
                    float sii = 0.0f;
                    for (int k = 0; k < 3; k++) sii += Rm[3 * k + i] * F[3 * k + i];
                    const float dii = sii - 1.0f;
                    tr += (dii > 0.0f) ? dii : -dii;
                }
                wt = lcm.inv_k * tr / 3.0f;
                wt = fmaxf(0.0f, fminf(wt, 1.0f));
            }
            lcm.w[e] = wt;
        }
        if (d <= 0.0f) { cnt++; sdet += d; }
        mdet = fminf(mdet, d);
    }
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
        blk_count[blk] = cc;
        blk_count[nblk + blk] = ff;
        blk_sum[blk] = s;
        blk_min[blk] = mm;
    }
}

// One tile of the element stage without prefetching: stage the tile's node
// vectors and rest positions in LDS, write each element's nodal forces to its
// force slots, sum each tile node's contiguous slot run into its partial.
// Element kinds of the tile stage: the stored-Ke C_WRP forces, the LCM / LCMH
// Piola forces, and C_WRP's K*x with the exact rotation differential.
enum TileKind { kTileWrp = 0, kTileLcm = 1, kTileWrpExact = 2 };

template <int MODE, int KIND = kTileWrp>
__device__ inline void tile_block(const TileArgs& a, uint32_t t, const Slots& s_v, const Slots& s_r, const Slots& s_F,
                                  uint32_t* s_sp) {
#pragma clang fp contract(off)
    const uint32_t tid = threadIdx.x;
    const TileDesc dt = tile_desc(a, t);
    TileRegs g;
    load_tile_regs_cold(a, dt, tid, g);
    float Fe[9];
    float w = 1.0f;
    bool allow_dR = false;
    if (KIND != kTileWrp) {
        const uint32_t e = dt.e0 + min(tid, dt.ne - 1);
        if (MODE == kModeFe || a.m.dr) {
#pragma unroll
            for (int k = 0; k < 9; k++) Fe[k] = ld32(a.m.F + k * (size_t)a.T, e);
        }
        if (MODE == kModeDfx && a.m.w) w = ld32(a.m.w, e);
        if (MODE == kModeDfx && a.m.dr) allow_dR = (a.m.dr == 1) || (ld32(a.m.detF, e) > a.m.thr);
    }
    if (tid < dt.nn) {
        s_v.x[tid] = g.v[0]; s_v.y[tid] = g.v[1]; s_v.z[tid] = g.v[2];
        s_r.x[tid] = g.r[0]; s_r.y[tid] = g.r[1]; s_r.z[tid] = g.r[2];
    }
    if (tid <= dt.nn) s_sp[tid] = g.sp;
    lds_barrier();
    if (tid < dt.ne) {
        uint32_t fs[4];
#pragma unroll
        for (int j = 0; j < 4; j++) fs[j] = s_sp[g.l[j]] + g.f[j];
        if (KIND == kTileLcm) element_force_lcm<MODE>(s_v, s_r, s_F, g.l, fs, g.R, a.m, Fe, w, allow_dR);
        else if (KIND == kTileWrpExact) element_force_exact(s_v, s_r, s_F, g.l, fs, g.R, a.m, Fe, allow_dR);
        else element_force<MODE>(s_v, s_r, s_F, g.l, fs, g.R, a.m);
    }
    lds_barrier();
    if (tid < dt.nn) {
        float fx, fy, fz;
        sum_slot_run(s_F, s_sp[tid], s_sp[tid + 1], fx, fy, fz);
        float* o = a.partials + 3 * (size_t)g.pd;
        o[0] = fx; o[1] = fy; o[2] = fz;
    }
}

// Node gather: node n sums its tile partials in tile order, then the caller's
// per-node terms. Returns the node's d.q contribution (matvec), else 0.
template <int MODE>
__device__ inline double gather_node(uint32_t n, const uint32_t* __restrict__ node_part_ptr,
                                     const float* __restrict__ partials, const NodeTerms& t) {
#pragma clang fp contract(off)
    const size_t i = 3 * (size_t)n;
    // the matvec's node data does not depend on the partials: load it first
    float d0 = 0.f, d1 = 0.f, d2 = 0.f, mass = 0.f;
    bool knc = false;
    if (MODE == kGatherMatvec) {
        mass = t.mass_array ? t.mass_array[n] : t.mass;
        d0 = t.d[i]; d1 = t.d[i + 1]; d2 = t.d[i + 2];
        knc = t.knc_mask && t.knc_mask[n];
    }
    // partials in ascending order, loaded two at a time (most nodes have one
    // or two), added one after the other as before
    float fx = 0.f, fy = 0.f, fz = 0.f;
    const uint32_t p1 = node_part_ptr[n + 1];
    uint32_t p = node_part_ptr[n];
    for (; p + 1 < p1; p += 2) {
        const float* pp = partials + 3 * (size_t)p;
        const float ax = pp[0], ay = pp[1], az = pp[2], bx = pp[3], by = pp[4], bz = pp[5];
        fx += ax; fy += ay; fz += az;
        fx += bx; fy += by; fz += bz;
    }
    if (p < p1) {
        const float* pp = partials + 3 * (size_t)p;
        fx += pp[0]; fy += pp[1]; fz += pp[2];
    }
    if (MODE == kGatherAccumulate) {
        t.out[i] += fx; t.out[i + 1] += fy; t.out[i + 2] += fz;
        return 0.0;
    } else if (MODE == kGatherMatvec) {
        // LS_A_times_d (LeafDSH_Solid3D_FEM.cpp:1056-1079) + LS_ZeroKNC (:1045-1053)
        const float fac = mass * t.fM;
        float q0 = -t.fK * fx + fac * d0;
        float q1 = -t.fK * fy + fac * d1;
        float q2 = -t.fK * fz + fac * d2;
        if (knc) { q0 = 0.f; q1 = 0.f; q2 = 0.f; }
        t.out[i] = q0; t.out[i + 1] = q1; t.out[i + 2] = q2;
        return (double)d0 * q0 + (double)d1 * q1 + (double)d2 * q2;
    } else if (MODE == kGatherRhsDx) {
        // One Newton iteration of Step_FullyImplicitEuler_dx (:1378-1465):
        // b = M g + f_e + (1/dt) M (v - (x_k - x)/dt); y = dt v when warm_start
        // (set by the caller for the first iteration only) else 0; KNC rows
        // b = y = target - x_k. Returns the node's b.b.
        const float mass = t.mass;
        const float sm = t.inv_dt * mass;
        const float f[3] = {fx, fy, fz};
        float b[3], y[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float vc = t.v[i + c];
            b[c] = (mass * t.g[c] + f[c]) + sm * (vc - t.inv_dt * (t.xk[i + c] - t.x[i + c]));
            y[c] = t.warm_start ? t.dt * vc : 0.0f;
        }
        if (t.knc_mask && t.knc_mask[n]) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                b[c] = t.knc_target[i + c] - t.xk[i + c];
                y[c] = b[c];
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) { t.out[i + c] = b[c]; t.y[i + c] = y[c]; }
        return (double)b[0] * b[0] + (double)b[1] * b[1] + (double)b[2] * b[2];
    } else {
        // b = dt (M g + f_e) + M v ; y = v ; KNC rows take the constraint
        // velocity (Step_QuasiImplicitEuler_v :1209-1295).
        const float mass = t.mass_array ? t.mass_array[n] : t.mass;
        const float f[3] = {fx, fy, fz};
        float b[3], y[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float vc = t.v[i + c];
            b[c] = (mass * t.g[c] + f[c]) * t.dt + mass * vc;
            y[c] = t.warm_start ? vc : 0.0f;
        }
        if (t.knc_mask && t.knc_mask[n]) {
            float vk[3];
#pragma unroll
            for (int c = 0; c < 3; c++) vk[c] = (t.knc_target[i + c] - t.x[i + c]) * t.inv_dt;
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
        for (int c = 0; c < 3; c++) { t.out[i + c] = b[c]; t.y[i + c] = y[c]; }
        return 0.0;
    }
}

template <int MODE>
__device__ inline void gather_block(uint32_t N, const uint32_t* __restrict__ node_part_ptr,
                                    const float* __restrict__ partials, const NodeTerms& t, uint32_t blk,
                                    uint32_t nblk, double* sh) {
#pragma clang fp contract(off)
    double local = 0.0;
    for (uint32_t n = blk * kBlock + threadIdx.x; n < N; n += nblk * kBlock)
        local += gather_node<MODE>(n, node_part_ptr, partials, t);
    if (MODE == kGatherMatvec || MODE == kGatherRhsDx) {
        const double s = block_sum(local, sh);
        if (threadIdx.x == 0) t.partials[blk] = s;
    }
}

// The matvec gather (LS_A_times_d, LeafDSH_Solid3D_FEM.cpp:1056-1079, and
// LS_ZeroKNC, :1045-1053) over chunk c of `npt` kBlock nodes: node
// (c npt + j) kBlock + thread for j < npt. For the first kGatherNpt nodes of a
// thread the node terms and the first two partials are loaded before any sum,
// so their latencies overlap; each node still adds its partials in ascending
// order. Returns the thread's d.q terms, summed in j order. INIT: the CG start
// instead, r = S(b - q) and d = r written for each node (q is not stored),
// returning the r.r terms.
__device__ inline double cg_start_node(uint32_t n, const float q[3], const NodeTerms& t) {
#pragma clang fp contract(off)
    const size_t i = 3 * (size_t)n;
    const bool knc = t.knc_mask && t.knc_mask[n];
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float ri = -1.0f * q[k] + t.b[i + k];  // muleq_addeq(r, -1, b) (ILSS_CG.h:55-63)
        if (knc) ri = 0.f;
        t.r_out[i + k] = ri;
        t.d_out[i + k] = ri;
        acc += (double)ri * ri;
    }
    return acc;
}
template <bool INIT>
__device__ inline double gather_matvec_chunk(uint32_t N, uint32_t c, uint32_t npt,
                                             const uint32_t* __restrict__ node_part_ptr,
                                             const float* __restrict__ partials, const NodeTerms& t) {
#pragma clang fp contract(off)
    const float* __restrict__ dv = t.d;
    float* __restrict__ out = t.out;
    const uint32_t tid = threadIdx.x;
    uint32_t nn[kGatherNpt], p0[kGatherNpt], p1[kGatherNpt];
    float dd[kGatherNpt][3], ms[kGatherNpt], a[kGatherNpt][3], b[kGatherNpt][3];
    bool kn[kGatherNpt];
#pragma unroll
    for (uint32_t j = 0; j < kGatherNpt; j++) {
        const uint32_t n = (c * npt + j) * kBlock + tid;
        nn[j] = n;
        p0[j] = p1[j] = 0;
        dd[j][0] = dd[j][1] = dd[j][2] = 0.f;
        ms[j] = 0.f;
        kn[j] = false;
        if (n < N) {
            p0[j] = node_part_ptr[n];
            p1[j] = node_part_ptr[n + 1];
            ms[j] = t.mass_array ? t.mass_array[n] : t.mass;
            dd[j][0] = dv[3 * (size_t)n]; dd[j][1] = dv[3 * (size_t)n + 1]; dd[j][2] = dv[3 * (size_t)n + 2];
            kn[j] = t.knc_mask && t.knc_mask[n];
        }
    }
#pragma unroll
    for (uint32_t j = 0; j < kGatherNpt; j++) {
        a[j][0] = a[j][1] = a[j][2] = 0.f;
        b[j][0] = b[j][1] = b[j][2] = 0.f;
        if (p0[j] < p1[j]) {
            const float* pp = partials + 3 * (size_t)p0[j];
            a[j][0] = pp[0]; a[j][1] = pp[1]; a[j][2] = pp[2];
        }
        if (p0[j] + 1 < p1[j]) {
            const float* pp = partials + 3 * (size_t)p0[j] + 3;
            b[j][0] = pp[0]; b[j][1] = pp[1]; b[j][2] = pp[2];
        }
    }
    double acc = 0.0;
#pragma unroll
    for (uint32_t j = 0; j < kGatherNpt; j++) {
        if (nn[j] >= N) continue;
        float fx = 0.f, fy = 0.f, fz = 0.f;
        if (p0[j] < p1[j]) { fx += a[j][0]; fy += a[j][1]; fz += a[j][2]; }
        if (p0[j] + 1 < p1[j]) { fx += b[j][0]; fy += b[j][1]; fz += b[j][2]; }
        for (uint32_t p = p0[j] + 2; p < p1[j]; p++) {
            const float* pp = partials + 3 * (size_t)p;
            fx += pp[0]; fy += pp[1]; fz += pp[2];
        }
        const float fac = ms[j] * t.fM;
        float q0 = -t.fK * fx + fac * dd[j][0];
        float q1 = -t.fK * fy + fac * dd[j][1];
        float q2 = -t.fK * fz + fac * dd[j][2];
        if (kn[j]) { q0 = 0.f; q1 = 0.f; q2 = 0.f; }
        if (INIT) {
            const float q[3] = {q0, q1, q2};
            acc += cg_start_node(nn[j], q, t);
            continue;
        }
        const size_t i = 3 * (size_t)nn[j];
        out[i] = q0; out[i + 1] = q1; out[i + 2] = q2;
        acc += (double)dd[j][0] * q0 + (double)dd[j][1] * q1 + (double)dd[j][2] * q2;
    }
    for (uint32_t j = kGatherNpt; j < npt; j++) {
        const uint32_t n = (c * npt + j) * kBlock + tid;
        if (n >= N) continue;
        if (INIT) {
            // q of node n alone, then the CG start terms
            float q[3];
            const float* pd = dv + 3 * (size_t)n;
            float fx = 0.f, fy = 0.f, fz = 0.f;
            for (uint32_t p = node_part_ptr[n]; p < node_part_ptr[n + 1]; p++) {
                const float* pp = partials + 3 * (size_t)p;
                fx += pp[0]; fy += pp[1]; fz += pp[2];
            }
            const float fac = (t.mass_array ? t.mass_array[n] : t.mass) * t.fM;
            q[0] = -t.fK * fx + fac * pd[0];
            q[1] = -t.fK * fy + fac * pd[1];
            q[2] = -t.fK * fz + fac * pd[2];
            if (t.knc_mask && t.knc_mask[n]) q[0] = q[1] = q[2] = 0.f;
            acc += cg_start_node(n, q, t);
        } else {
            acc += gather_node<kGatherMatvec>(n, node_part_ptr, partials, t);
        }
    }
    return acc;
}

// One element's C_V term (ms/fem/TetrahedronElementSet3.cpp:1456-1488),
// 1/2 u^T Ke u = 1/2 V sigma:H with H = (R^T D - Dm) Dm^-1: D and Dm hold the
// current and rest edges (entry 3 c + j: component c of vertex j+1 minus
// vertex 0), G = Dm^-1, V the rest volume.
__device__ inline double wrp_energy(const float D[9], const float Dm[9], const float Rm[9], const float G[9],
                                    float V, const Material& m) {
#pragma clang fp contract(off)
    float RD[9], H[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            RD[3 * i + j] = Rm[i] * D[j] + Rm[3 + i] * D[3 + j] + Rm[6 + i] * D[6 + j] - Dm[3 * i + j];
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
    return 0.5 * (double)V * (double)sH;
}

// C_V of C_WRP over tile t (ms/fem/TetrahedronElementSet3.cpp:1456-1488): the
// tile's positions (a.in) and rest positions staged in LDS planes (smem: 6 TB
// floats) as for K*x, Dm^-1 and V rebuilt from them as k_precompute computes
// them, one element per thread. Returns the tile's FP64 sum in thread 0 (each
// wave's terms summed by wave_sum, the waves in order).
template <int TB>
__device__ inline double tile_energy_block(const TileArgs& a, uint32_t t, float* smem, double* sh) {
#pragma clang fp contract(off)
    const Slots s_x = make_slots(smem, TB), s_r = make_slots(smem + 3 * TB, TB);
    const uint32_t tid = threadIdx.x;
    const TileDesc dt = tile_desc(a, t);
    TileRegs g;
    load_tile_regs_cold(a, dt, tid, g);
    if (tid < dt.nn) {
        s_x.x[tid] = g.v[0]; s_x.y[tid] = g.v[1]; s_x.z[tid] = g.v[2];
        s_r.x[tid] = g.r[0]; s_r.y[tid] = g.r[1]; s_r.z[tid] = g.r[2];
    }
    __syncthreads();
    double e = 0.0;
    if (tid < dt.ne) {
        const uint32_t l0 = g.l[0];
        float D[9], Dm[9], G[9];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const uint32_t l = g.l[q + 1];
            D[q] = s_x.x[l] - s_x.x[l0]; D[3 + q] = s_x.y[l] - s_x.y[l0]; D[6 + q] = s_x.z[l] - s_x.z[l0];
            Dm[q] = s_r.x[l] - s_r.x[l0]; Dm[3 + q] = s_r.y[l] - s_r.y[l0]; Dm[6 + q] = s_r.z[l] - s_r.z[l0];
        }
        const float detDm = refmath::det3(Dm);
        const float V = detDm / 6.0f;
        refmath::inv3(Dm, detDm, G);
        e = wrp_energy(D, Dm, g.R, G, V, a.m);
    }
    e = wave_sum(e);
    if ((tid & 63) == 0) sh[tid >> 6] = e;
    __syncthreads();
    double sum = 0.0;
    if (tid == 0)
        for (int w = 0; w < TB / 64; w++) sum += sh[w];
    return sum;
}

// C_V (ms/fem/TetrahedronElementSet3.cpp:1456-1488): 1/2 u^T Ke u = 1/2 V sigma:H
__device__ inline void energy_block(uint32_t T, const uint32_t* __restrict__ conn, const float* __restrict__ ref,
                                    const float* __restrict__ x, const float* __restrict__ R,
                                    const float* __restrict__ dinv, const float* __restrict__ vol, Material m,
                                    double* blk_part, uint32_t blk, uint32_t nblk, double* sh) {
#pragma clang fp contract(off)
    double acc = 0.0;
    if (m.mm != CF_MM_C_WRP && m.F) {
        // C_LCM_V (TetrahedronElementSet3.cpp:1489-1508): V (mu |S - I|_F^2 + lambda/2 tr(S - I)^2), S = R^T F
        for (uint32_t e = blk * kBlock + threadIdx.x; e < T; e += nblk * kBlock) {
            float Rm[9], F[9];
            for (int k = 0; k < 9; k++) { Rm[k] = R[k * (size_t)T + e]; F[k] = m.F[k * (size_t)T + e]; }
            float nsq = 0.0f, tr = 0.0f;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    float sij = 0.0f;
                    for (int k = 0; k < 3; k++) sij += Rm[3 * k + i] * F[3 * k + j];
                    const float d = sij - ((i == j) ? 1.0f : 0.0f);
                    nsq += d * d;
                    if (i == j) tr += d;
                }
            acc += (double)(vol[e] * (m.mu * nsq + 0.5f * m.lambda * (tr * tr)));
        }
        const double s = block_sum(acc, sh);
        if (threadIdx.x == 0) blk_part[blk] = s;
        return;
    }
    for (uint32_t e = blk * kBlock + threadIdx.x; e < T; e += nblk * kBlock) {
        float xs[4][3], rs[4][3];
        for (int j = 0; j < 4; j++) {
            const uint32_t g = conn[j * (size_t)T + e];
            xs[j][0] = x[3 * (size_t)g]; xs[j][1] = x[3 * (size_t)g + 1]; xs[j][2] = x[3 * (size_t)g + 2];
            rs[j][0] = ref[3 * (size_t)g]; rs[j][1] = ref[3 * (size_t)g + 1]; rs[j][2] = ref[3 * (size_t)g + 2];
        }
        float D[9], Dm[9], Rm[9], G[9];
        for (int c = 0; c < 3; c++)
            for (int j = 0; j < 3; j++) {
                D[3 * c + j] = xs[j + 1][c] - xs[0][c];
                Dm[3 * c + j] = rs[j + 1][c] - rs[0][c];
            }
        for (int k = 0; k < 9; k++) { Rm[k] = R[k * (size_t)T + e]; G[k] = dinv[k * (size_t)T + e]; }
        acc += wrp_energy(D, Dm, Rm, G, vol[e], m);
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) blk_part[blk] = s;
}

// ------------------------------------------------------------ CG bodies
// ILSS_CG::Solve (ns/ILSS_CG.h:38-151) over n = 3N floats; the projection S
// zeroes the rows of KNC nodes. Vector bodies walk n floats as float4 when
// every pointer is 16-byte aligned (V4), the n % 4 tail by block 0.

// r = S(b - q) ; d = r ; per-block r.r (one float per thread)
__device__ inline void cg_init_block(uint32_t n, const float* __restrict__ q, const float* __restrict__ b,
                                     float* __restrict__ r, float* __restrict__ d, const uint8_t* __restrict__ mask,
                                     double* rr_part, uint32_t blk, uint32_t nblk, double* sh) {
#pragma clang fp contract(off)
    double acc = 0.0;
    for (uint32_t i = blk * kBlock + threadIdx.x; i < n; i += nblk * kBlock) {
        float ri = -1.0f * q[i] + b[i];  // muleq_addeq(r, -1, b)
        if (mask && mask[i / 3]) ri = 0.f;
        r[i] = ri;
        d[i] = ri;
        acc += (double)ri * ri;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) rr_part[blk] = s;
}

// One float4 of the CG vector updates: the operands, then the arithmetic
// (shared by the block helpers below and the CG step kernels, which may load
// the operands early).
struct Xr4 { float4 x, r, d, q; };
__device__ inline Xr4 load_xr4(const float* x, const float* r, const float* d, const float* q, uint32_t i) {
    Xr4 v;
    v.d = reinterpret_cast<const float4*>(d)[i];
    v.q = reinterpret_cast<const float4*>(q)[i];
    v.x = reinterpret_cast<const float4*>(x)[i];
    v.r = reinterpret_cast<const float4*>(r)[i];
    return v;
}
// x += alpha d ; r -= alpha q for one float4, stored; returns its r.r terms
__device__ inline double apply_xr4(Xr4 v, float alpha, float* x, float* r, uint32_t i) {
#pragma clang fp contract(off)
    v.x.x = v.x.x + v.d.x * alpha; v.x.y = v.x.y + v.d.y * alpha;
    v.x.z = v.x.z + v.d.z * alpha; v.x.w = v.x.w + v.d.w * alpha;
    v.r.x = v.r.x - v.q.x * alpha; v.r.y = v.r.y - v.q.y * alpha;
    v.r.z = v.r.z - v.q.z * alpha; v.r.w = v.r.w - v.q.w * alpha;
    reinterpret_cast<float4*>(x)[i] = v.x;
    reinterpret_cast<float4*>(r)[i] = v.r;
    return (double)v.r.x * v.r.x + (double)v.r.y * v.r.y + (double)v.r.z * v.r.z + (double)v.r.w * v.r.w;
}
// the n % 4 floats after the float4s, for the threads gid of a stride
__device__ inline double tail_xr(uint32_t n, float alpha, float* x, float* r, const float* d, const float* q,
                                 uint32_t gid, uint32_t stride, double acc) {
#pragma clang fp contract(off)
    for (uint32_t i = 4 * (n / 4) + gid; i < n; i += stride) {
        x[i] = x[i] + d[i] * alpha;
        const float ri = r[i] - q[i] * alpha;
        r[i] = ri;
        acc += (double)ri * ri;
    }
    return acc;
}

// x += alpha d ; r -= alpha q ; per-block r.r (ILSS_CG.h:100-117)
template <bool V4>
__device__ inline void update_xr_block(uint32_t n, float alpha, float* __restrict__ x, float* __restrict__ r,
                                       const float* __restrict__ d, const float* __restrict__ q, double* rr_part,
                                       uint32_t blk, uint32_t nblk, double* sh) {
#pragma clang fp contract(off)
    double acc = 0.0;
    const uint32_t gid = blk * kBlock + threadIdx.x, stride = nblk * kBlock;
    if (V4) {
        for (uint32_t i = gid; i < n / 4; i += stride) acc += apply_xr4(load_xr4(x, r, d, q, i), alpha, x, r, i);
        acc = tail_xr(n, alpha, x, r, d, q, gid, stride, acc);
    } else {
        for (uint32_t i = gid; i < n; i += stride) {
            x[i] = x[i] + d[i] * alpha;
            const float ri = r[i] - q[i] * alpha;
            r[i] = ri;
            acc += (double)ri * ri;
        }
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) rr_part[blk] = s;
}

// x += alpha d ; r -= alpha q over chunk c of the leaf CG's chunk layout
// (rows float4s per thread: float4 (c rows + j) kBlock + thread, j ascending,
// then the n % 4 tail in chunk 0); returns the thread's r.r terms in the order
// the leaf's CG step kernel adds them, so the chunk partial is the same.
__device__ inline double xr_rows_chunk(uint32_t n, uint32_t c, uint32_t rows, float alpha, float* __restrict__ x,
                                       float* __restrict__ r, const float* __restrict__ d,
                                       const float* __restrict__ q) {
    const uint32_t n4 = n / 4, tid = threadIdx.x;
    double acc = 0.0;
    for (uint32_t j = 0; j < rows; j++) {
        const uint32_t i = (c * rows + j) * kBlock + tid;
        if (i < n4) acc += apply_xr4(load_xr4(x, r, d, q, i), alpha, x, r, i);
    }
    if (c == 0) acc = tail_xr(n, alpha, x, r, d, q, tid, kBlock, acc);
    return acc;
}

struct D4 { float4 d, r; };
__device__ inline D4 load_d4(const float* d, const float* r, uint32_t i) {
    D4 v;
    v.r = reinterpret_cast<const float4*>(r)[i];
    v.d = reinterpret_cast<const float4*>(d)[i];
    return v;
}
// d = S(beta d + r) for one float4, stored
__device__ inline void apply_d4(D4 v, float beta, float* d, const uint8_t* mask, uint32_t i) {
#pragma clang fp contract(off)
    float4 di = v.d;
    di.x = beta * di.x + v.r.x; di.y = beta * di.y + v.r.y;
    di.z = beta * di.z + v.r.z; di.w = beta * di.w + v.r.w;
    if (mask) {
        const uint32_t e = 4 * i;
        if (mask[e / 3]) di.x = 0.f;
        if (mask[(e + 1) / 3]) di.y = 0.f;
        if (mask[(e + 2) / 3]) di.z = 0.f;
        if (mask[(e + 3) / 3]) di.w = 0.f;
    }
    reinterpret_cast<float4*>(d)[i] = di;
}
__device__ inline void tail_d(uint32_t n, float beta, float* d, const float* r, const uint8_t* mask, uint32_t first,
                              uint32_t stride) {
#pragma clang fp contract(off)
    for (uint32_t i = first; i < n; i += stride) {
        float di = beta * d[i] + r[i];
        if (mask && mask[i / 3]) di = 0.f;
        d[i] = di;
    }
}

// d = S(beta d + r)
template <bool V4>
__device__ inline void update_d_block(uint32_t n, float beta, float* __restrict__ d, const float* __restrict__ r,
                                      const uint8_t* __restrict__ mask, uint32_t blk, uint32_t nblk) {
    const uint32_t gid = blk * kBlock + threadIdx.x, stride = nblk * kBlock;
    if (V4) {
        for (uint32_t i = gid; i < n / 4; i += stride) apply_d4(load_d4(d, r, i), beta, d, mask, i);
        tail_d(n, beta, d, r, mask, 4 * (n / 4) + gid, stride);
    } else {
        tail_d(n, beta, d, r, mask, gid, stride);
    }
}

// v = y ; x = x + dt y (Step_QuasiImplicitEuler_v :1322-1327), air drag (FixedStep :982-984)
// FIE-dx iterate update (:1500-1511): x_k += y ; v_k = (x_k - x) / dt ;
// per-block sum of y.y (FP64) into part[blk].
__device__ inline void nr_update_block(uint32_t n, float* __restrict__ xk, float* __restrict__ vk,
                                       const float* __restrict__ x, const float* __restrict__ y, float inv_dt,
                                       double* part, uint32_t blk, uint32_t nblk, double* sh) {
#pragma clang fp contract(off)
    double acc = 0.0;
    for (uint32_t i = blk * kBlock + threadIdx.x; i < n; i += nblk * kBlock) {
        const float yi = y[i];
        const float xn = xk[i] + yi;
        xk[i] = xn;
        vk[i] = inv_dt * (xn - x[i]);
        acc += (double)yi * yi;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blk] = s;
}

// End of a FIE-dx step (:1546-1551, air drag :982-984): x = x_k ; v = v_k.
__device__ inline void nr_finish_block(uint32_t n, float* __restrict__ x, float* __restrict__ v,
                                       const float* __restrict__ xk, const float* __restrict__ vk, float dt,
                                       float drag, uint32_t blk, uint32_t nblk) {
#pragma clang fp contract(off)
    for (uint32_t i = blk * kBlock + threadIdx.x; i < n; i += nblk * kBlock) {
        x[i] = xk[i];
        v[i] = (drag > 0.f) ? (1.0f - dt * drag) * vk[i] : vk[i];
    }
}

// x0 (may be null) receives x before the update: the leaf's m_vecPos0 = x,
// taken after BeginEvaluation (LeafDSH_Solid3D_FEM.cpp:936-937) and not read
// again before the next one.
__device__ inline void integrate_block(uint32_t n, float* __restrict__ x, float* __restrict__ v,
                                       const float* __restrict__ y, float dt, float drag, uint32_t blk,
                                       uint32_t nblk, float* __restrict__ x0 = nullptr) {
#pragma clang fp contract(off)
    for (uint32_t i = blk * kBlock + threadIdx.x; i < n; i += nblk * kBlock) {
        const float yi = y[i];
        const float xi = x[i];
        if (x0) x0[i] = xi;
        x[i] = xi + dt * yi;
        v[i] = (drag > 0.f) ? (1.0f - dt * drag) * yi : yi;
    }
}

}  // namespace cf
