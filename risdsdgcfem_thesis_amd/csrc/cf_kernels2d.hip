// 2D corotational triangle FEM: the whole QIE-v fixed step of a mesh inside
// one wavefront or one workgroup (TriangleElementSet2 + LeafDSH_Solid2D_FEM).
//
// Reference path: BeginEvaluation (ms/fem/TriangleElementSet2.cpp:183-262),
// rotations (:1007-1018,1078-1110 -> Mal/GMatDecomposition.h:17-30 for
// det F > threshold, ms/fem/Rotation.cpp:19-37 for MSVD below it),
// C_f_e (:1253-1286), C_df_x (:1859-1915), C_V
// (:1620-1645), LS_A_times_d (ds/dsh/LeafDSH_Solid2D_FEM.cpp:1101-1132),
// Step_QuasiImplicitEuler_v (:1287-1403) with the CG of ns/ILSS_CG.h:38-151.
//
// Ke = A B^T E B (ms/fem/TriangleElement2.h:98-139) is applied in the
// equivalent Piola form from Dm^-1 and the area:
//   H = R^T [v1-v0, v2-v0] Dm^-1         (df_x: v = d)
//   H = (R^T [x1-x0, x2-x0] - Dm) Dm^-1  (f_e)
//   sigma = E(sym H), P = -A R sigma Dm^-T ; vertex c+1 gets column c,
//   vertex 0 minus their sum.
// No memory traffic inside the CG loop: triangle data in registers, the CG
// direction and the nodal forces in LDS, per-node CG vectors in registers.
#include "cf_device.h"

namespace cf {

namespace {

// A group of GW wavefronts stepping one mesh inside a 512-thread workgroup.
// GW == 1: wave-level barriers and reductions, groups fully independent.
// GW > 1: workgroup barriers shared by all groups (same barrier sequence in
// every group), reductions over the group's own wavefronts in fixed order.
template <int GW>
struct Grp {
    static constexpr int G = 64 * GW;
    __device__ static uint32_t lane() { return threadIdx.x % G; }
    __device__ static uint32_t group() { return threadIdx.x / G; }
    __device__ static void bar() {
        if (GW == 1) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
            __syncthreads();
        }
    }
    // red holds two buffers of 8 wave sums, used alternately (par flips on
    // every call, the same in every thread): a buffer is written again only
    // two sums later, after every thread has passed the other sum's barrier
    // and so has finished reading it. One barrier per sum.
    __device__ static double sum(double v, double* red, uint32_t& par) {
        v = wave_sum(v);
        if (GW == 1) return v;
        double* rb = red + 8 * par;
        par ^= 1u;
        if ((threadIdx.x & 63) == 0) rb[threadIdx.x >> 6] = v;
        __syncthreads();
        const uint32_t w0 = group() * GW;
        double r = 0.0;
        for (int i = 0; i < GW; i++) r += rb[w0 + i];
        return r;
    }
    __device__ static float min(float v, float* red) {
        for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
        if (GW == 1) return v;
        __syncthreads();
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
        const uint32_t w0 = group() * GW;
        float r = red[w0];
        for (int i = 1; i < GW; i++) r = fminf(r, red[w0 + i]);
        return r;
    }
    // whether any group of the workgroup is still active: its own flag when
    // the group is alone in its barrier sequence (GW == 1: a wavefront; GW ==
    // 8: the whole workgroup, where `active` is uniform); else through two
    // alternating flag buffers, one barrier (as sum). The barrier also orders
    // the group's LDS writes before the call against the reads after it.
    static constexpr bool kAlone = GW == 1 || GW == 8;
    __device__ static bool any(bool active, uint32_t* flags, uint32_t& par) {
        if (kAlone) return active;
        uint32_t* fb = flags + 8 * par;
        par ^= 1u;
        if (lane() == 0) fb[group()] = active ? 1u : 0u;
        __syncthreads();
        uint32_t f = 0;
        for (int i = 0; i < 8 / GW; i++) f |= fb[i];
        return f != 0;
    }
};

// A triangle's constant data, kept in registers by the thread that owns it;
// its rotation of the step lives in LDS (one float4 per owned triangle).
struct Tri {
    uint32_t i01, i2;          // local node ids: i0 | i1 << 16, i2 (meshes hold at most 1,536 nodes)
    float g00, g01, g10, g11;  // Dm^-1
    float A;
    __device__ uint32_t n0() const { return i01 & 0xffffu; }
    __device__ uint32_t n1() const { return i01 >> 16; }
    __device__ uint32_t n2() const { return i2; }
};

// The thread's k-th triangle for a k known only at run time. The fields are
// combined by masks on their bits: a register array indexed at run time (or
// a select chain the compiler turns into one) would live in scratch memory.
__device__ inline uint32_t pick_bits(const uint32_t (&v)[kTriRegs], uint32_t k) {
    uint32_t r = 0;
#pragma unroll
    for (uint32_t j = 0; j < uint32_t(kTriRegs); j++) r |= v[j] & (0u - uint32_t(k == j));
    return r;
}
__device__ inline Tri pick(const Tri (&tr)[kTriRegs], uint32_t k) {
    uint32_t i01[kTriRegs], i2[kTriRegs], g00[kTriRegs], g01[kTriRegs], g10[kTriRegs], g11[kTriRegs], A[kTriRegs];
#pragma unroll
    for (int j = 0; j < kTriRegs; j++) {
        i01[j] = tr[j].i01; i2[j] = tr[j].i2;
        g00[j] = __float_as_uint(tr[j].g00); g01[j] = __float_as_uint(tr[j].g01);
        g10[j] = __float_as_uint(tr[j].g10); g11[j] = __float_as_uint(tr[j].g11);
        A[j] = __float_as_uint(tr[j].A);
    }
    Tri t;
    t.i01 = pick_bits(i01, k); t.i2 = pick_bits(i2, k);
    t.g00 = __uint_as_float(pick_bits(g00, k)); t.g01 = __uint_as_float(pick_bits(g01, k));
    t.g10 = __uint_as_float(pick_bits(g10, k)); t.g11 = __uint_as_float(pick_bits(g11, k));
    t.A = __uint_as_float(pick_bits(A, k));
    return t;
}

struct Rot {
    float r00, r01, r10, r11;  // R, row-major
};
__device__ inline Rot rot_of(float4 v) { return Rot{v.x, v.y, v.z, v.w}; }

// Closed-form 2x2 polar factor (Mal/GMatDecomposition.h:17-30)
__device__ inline void polar2(float f00, float f01, float f10, float f11, float det, Rot& t) {
#pragma clang fp contract(off)
    const float sg = (det > 0.f) ? 1.0f : ((det < 0.f) ? -1.0f : 0.0f);
    float r00 = f00 + sg * f11, r01 = f01 + sg * (-f10);
    float r10 = f10 + sg * (-f01), r11 = f11 + sg * f00;
    const float i0 = 1.0f / sqrtf(r00 * r00 + r10 * r10);
    r00 *= i0; r10 *= i0;
    const float i1 = 1.0f / sqrtf(r01 * r01 + r11 * r11);
    r01 *= i1; r11 *= i1;
    t.r00 = r00; t.r01 = r01; t.r10 = r10; t.r11 = r11;
}

// Degenerate branch of the 2D Compute_R_SVD (ms/fem/Rotation.cpp:27-36), in
// the reference's FP32 operation order: eigenvalues of F^T F by the quadratic
// formula (Mal/GMatEigen.h:151-164), eigenvectors from the first row of
// F^T F - lambda Id whose squared norm is not below 1e-8 (:172-184,
// Mal/GVec.h:193-199), U = F V diag(s)^+ with singular values <= 1e-4 taken
// as 0 and their directions completed by the CCW perpendicular
// (Mal/GMatDecomposition.h:255-345), reflections removed from U and V^T,
// R = U V^T. Only the second eigenvalue and singular value are NaN-tested
// (Mal/GVec.h:201-208 starts at index 1); a NaN pair with |det F - 1| >= 0.05
// flows into R, as in the reference's release build.
struct Mat2 {
    float m00, m01, m10, m11;
};

__device__ inline Mat2 mul2(const Mat2& x, const Mat2& y) {  // Mal/GMat.h:135-145, accumulated from 0
#pragma clang fp contract(off)
    Mat2 r;
    float a;
    a = 0.0f; a += x.m00 * y.m00; a += x.m01 * y.m10; r.m00 = a;
    a = 0.0f; a += x.m00 * y.m01; a += x.m01 * y.m11; r.m01 = a;
    a = 0.0f; a += x.m10 * y.m00; a += x.m11 * y.m10; r.m10 = a;
    a = 0.0f; a += x.m10 * y.m01; a += x.m11 * y.m11; r.m11 = a;
    return r;
}

__device__ inline float det2m(const Mat2& m) {
#pragma clang fp contract(off)
    return m.m00 * m.m11 - m.m10 * m.m01;
}

__device__ inline bool zero2(float a, float b) {
#pragma clang fp contract(off)
    float acc = 0.0f;
    acc += a * a;
    acc += b * b;
    return acc < 0.0001f * 0.0001f;
}

__device__ inline void eigvec2(const Mat2& m, float lam, float& x, float& y) {
#pragma clang fp contract(off)
    const float e00 = m.m00 - lam, e01 = m.m01, e10 = m.m10, e11 = m.m11 - lam;
    if (!zero2(e00, e01)) { x = e01; y = -e00; }
    else if (!zero2(e10, e11)) { x = -e11; y = e10; }
    else { x = 0.0f; y = 0.0f; return; }
    float nsq = 0.0f;
    nsq += x * x;
    nsq += y * y;
    const float inv = 1.0f / sqrtf(nsq);
    x = inv * x;
    y = inv * y;
}

// Not inlined: the branch is rare, and inlining it into the step kernel's
// unrolled triangle loop would cost the kernel registers it needs elsewhere.
// fix: remove reflections from U and V^T as Compute_R_SVD does (Rotation.cpp:31-33);
// GRotation2x2_PolarDecomposition_From_SVD (GMatDecomposition.h:365-379) does not.
__device__ __noinline__ float4 svd2_degenerate(float f00, float f01, float f10, float f11, bool fix = true) {
#pragma clang fp contract(off)
    const Mat2 F{f00, f01, f10, f11};
    const Mat2 A = mul2(Mat2{f00, f10, f01, f11}, F);
    float tr = 0.0f;
    tr += A.m00;
    tr += A.m11;
    const float b = -tr, c = det2m(A);
    const float pm = sqrtf(b * b - 4.0f * c);
    const float ev0 = 0.5f * (-b + pm), ev1 = 0.5f * (-b - pm);
    bool identity = false;
    if (ev1 != ev1) {
        const float d = det2m(F) - 1.0f;
        identity = double(d > 0.0f ? d : -d) < 0.05;
    }
    Mat2 U{1.0f, 0.0f, 0.0f, 1.0f}, Vt{1.0f, 0.0f, 0.0f, 1.0f};
    if (!identity) {
        float s0 = sqrtf(ev0), s1 = sqrtf(ev1);
        if (s1 != s1) {
            if (ev0 < 0.0f) s0 = 0.0f;
            if (ev1 < 0.0f) s1 = 0.0f;
        }
        const bool z0 = (s0 > 0.0f ? s0 : -s0) <= 0.0001f;  // Mal/RealUtils.h:114
        const bool z1 = (s1 > 0.0f ? s1 : -s1) <= 0.0001f;
        Mat2 V;
        eigvec2(A, ev0, V.m00, V.m10);
        eigvec2(A, ev1, V.m01, V.m11);
        const Mat2 P{z0 ? 0.0f : 1.0f / s0, 0.0f, 0.0f, z1 ? 0.0f : 1.0f / s1};
        if (!(z0 && z1)) {
            U = mul2(mul2(F, V), P);
            Vt = Mat2{V.m00, V.m10, V.m01, V.m11};
            if (z0) {
                Vt.m00 = Vt.m11; Vt.m01 = -Vt.m10;
                U.m00 = U.m11; U.m10 = -U.m01;
            } else if (z1) {
                Vt.m10 = Vt.m01; Vt.m11 = -Vt.m00;
                U.m01 = U.m10; U.m11 = -U.m00;
            }
        }
    }
    if (fix) {
        if (det2m(U) < 0.0f) { U.m01 = -U.m01; U.m11 = -U.m11; }
        if (det2m(Vt) < 0.0f) { Vt.m10 = -Vt.m10; Vt.m11 = -Vt.m11; }
    }
    const Mat2 R = mul2(U, Vt);
    return make_float4(R.m00, R.m01, R.m10, R.m11);
}

// ---- DAPD degenerate triangles (rare path, out of line)
constexpr uint32_t kNocInvalid = 3;  // cInvalidNoC (TriangleElementSet2.h:114)

// UpdateDegeneration for one triangle at or below the threshold
// (TriangleElementSet2.cpp:583-640): a new node of collapse when the triangle
// was undegenerate at prev_x (ComputeNoC, :739-835: the time of collapse as a
// root of the quadratic det Ds(t) = thr 2 A in float, the first positive one
// else the second, then the vertex opposite the longest edge), else the old.
__device__ __noinline__ uint32_t update_noc2(float2 a0, float2 a1, float2 a2, float2 b0, float2 b1, float2 b2,
                                             float g00, float g01, float g10, float g11, float area, float thr,
                                             uint32_t old) {
#pragma clang fp contract(off)
    {
        const float D00 = a1.x - a0.x, D01 = a2.x - a0.x, D10 = a1.y - a0.y, D11 = a2.y - a0.y;
        float f00 = 0.0f; f00 += D00 * g00; f00 += D01 * g10;
        float f01 = 0.0f; f01 += D00 * g01; f01 += D01 * g11;
        float f10 = 0.0f; f10 += D10 * g00; f10 += D11 * g10;
        float f11 = 0.0f; f11 += D10 * g01; f11 += D11 * g11;
        if (!(f00 * f11 - f10 * f01 > thr)) return old;  // degenerate already: keep
    }
    const float a[3][2] = {{a0.x, a0.y}, {a1.x, a1.y}, {a2.x, a2.y}};
    float d[3][2] = {{b0.x - a0.x, b0.y - a0.y}, {b1.x - a1.x, b1.y - a1.y}, {b2.x - a2.x, b2.y - a2.y}};
    const float K = thr * 2.0f * area;
    const float A = d[1][0] * d[2][1] + d[2][0] * d[0][1] + d[0][0] * d[1][1] - d[1][0] * d[0][1] -
                    d[0][0] * d[2][1] - d[2][0] * d[1][1];
    const float B = (d[1][0] * a[2][1] + a[1][0] * d[2][1]) + (d[2][0] * a[0][1] + a[2][0] * d[0][1]) +
                    (d[0][0] * a[1][1] + a[0][0] * d[1][1]) - (d[1][0] * a[0][1] + a[1][0] * d[0][1]) -
                    (d[0][0] * a[2][1] + a[0][0] * d[2][1]) - (d[2][0] * a[1][1] + a[2][0] * d[1][1]);
    const float C = a[1][0] * a[2][1] + a[2][0] * a[0][1] + a[0][0] * a[1][1] - a[1][0] * a[0][1] -
                    a[0][0] * a[2][1] - a[2][0] * a[1][1] - K;
    float t1 = 0.0f, t2 = 1.0f;
    if (refmath::solve_eq2(A, B, C, t1, t2) <= 0) return kNocInvalid;
    const float toc = (t1 > 0.0f) ? t1 : t2;
    float q[3][2];
    for (int i = 0; i < 3; i++) { q[i][0] = a[i][0] + toc * d[i][0]; q[i][1] = a[i][1] + toc * d[i][1]; }
    float l[3];  // squared lengths of edges 01, 12, 20
    for (int i = 0; i < 3; i++) {
        const int j = (i + 1) % 3;
        const float ex = q[j][0] - q[i][0], ey = q[j][1] - q[i][1];
        float acc = 0.0f;
        acc += ex * ex;
        acc += ey * ey;
        l[i] = acc;
    }
    return (l[0] >= l[1]) ? ((l[0] >= l[2]) ? 2u : 1u) : ((l[1] >= l[2]) ? 0u : 1u);
}

// Compute_R_DAPD_Degenerate, 2D (Rotation.cpp:300-362, __ENABLE_REGULARIZE_DAPD
// :6-8): the node of collapse moves along the perpendicular of its opposite
// edge (squared length + 1e-6) by lambda_u; R is the closed-form polar factor
// of the projected F, or its SVD-based form when det <= 0.001.
__device__ __noinline__ float4 r_dapd_degenerate2(float2 p0, float2 p1, float2 p2, uint32_t noc, float area,
                                                  float g00, float g01, float g10, float g11, float thr, float fL,
                                                  float fNL, float eNL) {
#pragma clang fp contract(off)
    float x[3][2] = {{p0.x, p0.y}, {p1.x, p1.y}, {p2.x, p2.y}};
    const uint32_t i0 = noc, i1 = (noc + 1) % 3, i2 = (noc + 2) % 3;
    const float ex = x[i2][0] - x[i1][0], ey = x[i2][1] - x[i1][1];
    float dsq = 0.0f;
    dsq += ex * ex;
    dsq += ey * ey;
    dsq += 0.001f * 0.001f;
    const float dist = sqrtf(dsq);
    const float inv = 1.0f / dist;
    const float ax = inv * -ey, ay = inv * ex;  // PerpendicularCW (Mal/GVec.h:223-226) / dist
    const float dh = thr * 2.0f * area / dist;
    const float mx = 0.5f * (x[i1][0] + x[i2][0]), my = 0.5f * (x[i1][1] + x[i2][1]);
    float w = 0.0f;
    w += (x[i0][0] - mx) * ax;
    w += (x[i0][1] - my) * ay;
    const float s = w / dh;
    const float t = 1.0f - s;
    const float lambda = (w >= 0.0f) ? fL * dh * ((2.0f - t) * t * t) : fL * dh * (t + fNL * powf(-s, eNL));
    x[i0][0] = x[i0][0] + lambda * ax;
    x[i0][1] = x[i0][1] + lambda * ay;
    const float D00 = x[1][0] - x[0][0], D01 = x[2][0] - x[0][0], D10 = x[1][1] - x[0][1], D11 = x[2][1] - x[0][1];
    float f00 = 0.0f; f00 += D00 * g00; f00 += D01 * g10;
    float f01 = 0.0f; f01 += D00 * g01; f01 += D01 * g11;
    float f10 = 0.0f; f10 += D10 * g00; f10 += D11 * g10;
    float f11 = 0.0f; f11 += D10 * g01; f11 += D11 * g11;
    const float dp = f00 * f11 - f10 * f01;
    const float dd = dp - 1.0f;
    if (dp > 0.001f || (dd > 0.0f ? dd : -dd) <= 0.0001f) {
        Rot t_;
        polar2(f00, f01, f10, f11, dp, t_);
        return make_float4(t_.r00, t_.r01, t_.r10, t_.r11);
    }
    return svd2_degenerate(f00, f01, f10, f11, false);
}

// Rotation of one triangle (ComputeRotations_MSVD/DAPD,
// TriangleElementSet2.cpp:1007-1018,1078-1110): polar factor above the
// threshold; below it MSVD's 2x2 SVD, DAPD R = Id (its invalid-NoC branch).
// Returns 1 for a DAPD fallback.
__device__ inline float rotation2(int rm, float thr, float f00, float f01, float f10, float f11, float det, Rot& t) {
    t.r00 = 1.f; t.r01 = 0.f; t.r10 = 0.f; t.r11 = 1.f;
    if (rm == CF_RM_ID) return 0.0f;
    if (det > thr) {
        polar2(f00, f01, f10, f11, det, t);
        return 0.0f;
    }
    if (rm == CF_RM_MSVD) {
        const float4 r = svd2_degenerate(f00, f01, f10, f11);
        t.r00 = r.x; t.r01 = r.y; t.r10 = r.z; t.r11 = r.w;
        return 0.0f;
    }
    return 1.0f;
}

// Displacement gradient H of a triangle: R^T D Dm^-1, or for f_e
// (R^T D - Dm) Dm^-1 so that a rest state gives exactly 0.
template <bool FE>
__device__ inline void tri_gradient(const Tri& t, const Rot& R, float2 p0, float2 p1, float2 p2, float2 q0, float2 q1,
                                    float2 q2, float H[4]) {
#pragma clang fp contract(off)
    const float D00 = p1.x - p0.x, D01 = p2.x - p0.x, D10 = p1.y - p0.y, D11 = p2.y - p0.y;
    float RD00 = fmaf(R.r10, D10, R.r00 * D00), RD01 = fmaf(R.r10, D11, R.r00 * D01);
    float RD10 = fmaf(R.r11, D10, R.r01 * D00), RD11 = fmaf(R.r11, D11, R.r01 * D01);
    if (FE) {
        RD00 -= q1.x - q0.x; RD01 -= q2.x - q0.x;
        RD10 -= q1.y - q0.y; RD11 -= q2.y - q0.y;
    }
    H[0] = fmaf(RD01, t.g10, RD00 * t.g00);
    H[1] = fmaf(RD01, t.g11, RD00 * t.g01);
    H[2] = fmaf(RD11, t.g10, RD10 * t.g00);
    H[3] = fmaf(RD11, t.g11, RD10 * t.g01);
}

// Nodal forces of a triangle from its displacement gradient.
__device__ inline void tri_forces(const Tri& t, const Rot& R, const float H[4], const Material& m, float2& f0,
                                  float2& f1, float2& f2) {
#pragma clang fp contract(off)
    const float sxx = fmaf(m.a_k1, H[0], m.a_nu * H[3]);
    const float syy = fmaf(m.a_k1, H[3], m.a_nu * H[0]);
    const float sxy = m.a_k2 * (H[1] + H[2]);
    const float W00 = -t.A * fmaf(sxy, t.g01, sxx * t.g00), W01 = -t.A * fmaf(sxy, t.g11, sxx * t.g10);
    const float W10 = -t.A * fmaf(syy, t.g01, sxy * t.g00), W11 = -t.A * fmaf(syy, t.g11, sxy * t.g10);
    const float P00 = fmaf(R.r01, W10, R.r00 * W00), P01 = fmaf(R.r01, W11, R.r00 * W01);
    const float P10 = fmaf(R.r11, W10, R.r10 * W00), P11 = fmaf(R.r11, W11, R.r10 * W01);
    f1 = make_float2(P00, P10);
    f2 = make_float2(P01, P11);
    f0 = make_float2(-(P00 + P01), -(P10 + P11));
}

// Element pass: s_f = sum of the group's triangle forces on the vectors in
// s_v (FE: f_e on positions, with the rest positions), colour by colour.
// Triangles are sorted by colour and thread `lane` owns triangles lane + k G,
// so its triangles of one colour are a run of consecutive k (on a grid mesh
// at most one); no two triangles of a colour share a node, so the order
// within a colour does not change a node's sum. `run` false: the group only
// takes part in the barriers.
// ZERO: clear the node forces first (behind a barrier); else the caller has
// cleared them (each thread its own nodes, as it consumed them) and ordered
// its LDS writes before the call with a barrier of its own.
template <int GW, bool FE, bool ZERO = true>
__device__ inline void accumulate_forces(const Tri (&tr)[kTriRegs], const float4* s_R, uint32_t lane, uint32_t N,
                                         uint32_t T, uint32_t C, uint32_t CL, bool run, const uint16_t* cs,
                                         const float2* ref, const Material& m, const float2* s_v, float2* s_f) {
#pragma clang fp contract(off)
    constexpr int G = Grp<GW>::G;
    if (ZERO) {
        if (run) {
#pragma unroll
            for (int k = 0; k < kNodeRegs; k++) {
                const uint32_t n = lane + k * G;
                if (n < N) s_f[n] = make_float2(0.f, 0.f);
            }
        }
        Grp<GW>::bar();
    }
    for (uint32_t c = 0; c < CL; c++) {
        if (run && c < C) {
            const uint32_t lo = cs[c], hi = min(uint32_t(cs[c + 1]), T);
            uint32_t k = (lane >= lo) ? 0u : (lo - lane + G - 1) / G;  // first k with lane + k G >= lo
#pragma unroll 1
            for (; k < uint32_t(kTriRegs) && lane + k * G < hi; k++) {
                const Tri t = pick(tr, k);
                const Rot R = rot_of(s_R[k * G + lane]);
                float H[4];
                const float2 z = make_float2(0.f, 0.f);
                if (FE) tri_gradient<true>(t, R, s_v[t.n0()], s_v[t.n1()], s_v[t.n2()], ref[t.n0()], ref[t.n1()], ref[t.n2()], H);
                else tri_gradient<false>(t, R, s_v[t.n0()], s_v[t.n1()], s_v[t.n2()], z, z, z, H);
                float2 f0, f1, f2;
                tri_forces(t, R, H, m, f0, f1, f2);
                float2 a0 = s_f[t.n0()], a1 = s_f[t.n1()], a2 = s_f[t.n2()];
                a0.x += f0.x; a0.y += f0.y;
                a1.x += f1.x; a1.y += f1.y;
                a2.x += f2.x; a2.y += f2.y;
                s_f[t.n0()] = a0; s_f[t.n1()] = a1; s_f[t.n2()] = a2;
            }
        }
        Grp<GW>::bar();
    }
}

template <int GW>
__device__ void mesh_step(const Batch2Args& a, uint32_t cmax, float2* s_v_all, float2* s_f_all, float2* s_y_all,
                          float4* s_R_all, double* red, float* redm, uint32_t* flags) {
#pragma clang fp contract(off)
    using Gp = Grp<GW>;
    constexpr int G = Gp::G;
    // the group index is the same in every lane of a wavefront: make it (and
    // so the mesh descriptor) scalar
    const uint32_t lane = Gp::lane();
    const uint32_t grp = __builtin_amdgcn_readfirstlane(Gp::group());
    const uint32_t sim = __builtin_amdgcn_readfirstlane(a.work[blockIdx.x].sim[grp]);
    const bool real = sim != kIdleSlot;
    if (GW == 1 && !real) return;  // independent wavefronts: nothing to wait for
    Mesh2Dev M;
    if (real) {
        M = a.meshes[sim];
    } else {
        M.N = M.T = M.C = 0;
        M.node_off = M.tri_off = M.color_off = 0;
        M.m = Material{0.f, 0.f, 0.f};
        M.fM = M.fK = M.dt = M.inv_dt = M.g0 = M.g1 = M.rel = M.abs_eps = M.thr = M.drag = 0.f;
        M.rm = M.warm_start = M.compute_energy = 0;
        M.max_iter = 0;
        M.dapd_l = M.dapd_nl = M.dapd_e = 0.f;
    }
    const uint32_t N = M.N, T = M.T, C = M.C;
    const uint32_t CL = (GW == 1) ? C : cmax;  // colour passes: the same count in every group
    float2* s_v = s_v_all + grp * (G * kNodeRegs);
    float2* s_f = s_f_all + grp * (G * kNodeRegs);
    float2* s_y = s_y_all + grp * (G * kNodeRegs);  // the solution y; s_v holds the CG direction d
    float4* s_R = s_R_all + grp * (G * kTriRegs);   // the rotation of triangle lane + k G at k G + lane
    const float2* ref = a.ref + M.node_off;
    float2* xg = a.x + M.node_off;
    float2* vg = a.v + M.node_off;
    // DoC-style state only matters to DAPD: other meshes neither read nor write it
    const bool track = real && a.noc && M.rm == CF_RM_DAPD;
    float2* x0g = track ? a.x0 + M.node_off : nullptr;
    const uint16_t* cs = a.color_start + M.color_off;

    // ---- own triangles and nodes
    Tri tr[kTriRegs];
#pragma unroll
    for (int k = 0; k < kTriRegs; k++) {
        const uint32_t e = lane + k * G;
        if (e < T) {
            const size_t ge = M.tri_off + (size_t)e;
            tr[k].i01 = a.tri_ids[3 * ge] | (a.tri_ids[3 * ge + 1] << 16);
            tr[k].i2 = a.tri_ids[3 * ge + 2];
            const float4 g = a.dinv[ge];
            tr[k].g00 = g.x; tr[k].g01 = g.y; tr[k].g10 = g.z; tr[k].g11 = g.w;
            tr[k].A = a.area[ge];
        }
    }
    float2 r[kNodeRegs];
    float mass[kNodeRegs];
    bool kn[kNodeRegs];
#pragma unroll
    for (int k = 0; k < kNodeRegs; k++) {
        const uint32_t n = lane + k * G;
        if (n < N) {
            mass[k] = a.mass[M.node_off + n];
            kn[k] = a.knc[M.node_off + n] != 0;
            s_v[n] = xg[n];
        }
    }
    Gp::bar();

    // ---- BeginEvaluation: F = Ds Dm^-1, det F, R
    double ndeg = 0.0, sdeg = 0.0, nfb = 0.0;
    float mind = 3.0e38f;
    // one triangle at a time (k at run time): the rotation's temporaries are
    // live for one triangle only
#pragma unroll 1
    for (uint32_t k = 0; k < uint32_t(kTriRegs); k++) {
        const uint32_t e = lane + k * G;
        if (e >= T) break;
        const Tri t = pick(tr, k);
        const float2 p0 = s_v[t.n0()], p1 = s_v[t.n1()], p2 = s_v[t.n2()];
        const float D00 = p1.x - p0.x, D01 = p2.x - p0.x, D10 = p1.y - p0.y, D11 = p2.y - p0.y;
        float f00 = 0.0f; f00 += D00 * t.g00; f00 += D01 * t.g10;  // Mal GMat product, accumulated from 0
        float f01 = 0.0f; f01 += D00 * t.g01; f01 += D01 * t.g11;
        float f10 = 0.0f; f10 += D10 * t.g00; f10 += D11 * t.g10;
        float f11 = 0.0f; f11 += D10 * t.g01; f11 += D11 * t.g11;
        const float det = f00 * f11 - f10 * f01;  // Mal/GMat.h:230-233
        // UpdateDegeneration (TriangleElementSet2.cpp:583-640) and the DAPD
        // rotation of a degenerate triangle with a node of collapse
        uint32_t nc = kNocInvalid;
        if (track) {
            const size_t ge = M.tri_off + (size_t)e;
            const uint32_t old = a.noc[ge];
            nc = (det > M.thr) ? kNocInvalid
                               : update_noc2(x0g[t.n0()], x0g[t.n1()], x0g[t.n2()], p0, p1, p2, t.g00, t.g01, t.g10, t.g11,
                                             t.A, M.thr, old);
            if (nc != old) a.noc[ge] = uint8_t(nc);
        }
        if (track && !(det > M.thr) && nc != kNocInvalid) {
            s_R[k * G + lane] = r_dapd_degenerate2(p0, p1, p2, nc, t.A, t.g00, t.g01, t.g10, t.g11, M.thr, M.dapd_l,
                                                   M.dapd_nl, M.dapd_e);
        } else {
            Rot rt;
            nfb += rotation2(M.rm, M.thr, f00, f01, f10, f11, det, rt);
            s_R[k * G + lane] = make_float4(rt.r00, rt.r01, rt.r10, rt.r11);
        }
        if (det <= 0.f) { ndeg += 1.0; sdeg += det; }
        mind = fminf(mind, det);
    }

    // ---- b = dt (M g + f_e) + M v ; y = v ; KNC rows take the constraint velocity
    accumulate_forces<GW, true>(tr, s_R, lane, N, T, C, CL, real, cs, ref, M.m, s_v, s_f);
    float2 b[kNodeRegs];
#pragma unroll
    for (int k = 0; k < kNodeRegs; k++) {
        const uint32_t n = lane + k * G;
        if (n >= N) continue;
        const float2 f = s_f[n], vv = vg[n];
        const float m = mass[k];
        b[k].x = (m * M.g0 + f.x) * M.dt + m * vv.x;
        b[k].y = (m * M.g1 + f.y) * M.dt + m * vv.y;
        float2 yk = M.warm_start ? vv : make_float2(0.f, 0.f);
        if (kn[k]) {
            const float2 q = ref[n], xx = s_v[n];
            float vx = (q.x - xx.x) * M.inv_dt, vy = (q.y - xx.y) * M.inv_dt;
            const float nk = sqrtf(vx * vx + vy * vy);
            if (nk > 0.001f) {
                const float inv = 1.0f / nk, s = fminf(nk, 50.0f);
                vx = s * (vx * inv); vy = s * (vy * inv);
            } else {
                vx = 0.f; vy = 0.f;
            }
            b[k] = make_float2(vx, vy);
            yk = b[k];
        }
        s_y[n] = yk;
    }
    Gp::bar();  // every read of s_v (positions) is done
#pragma unroll
    for (int k = 0; k < kNodeRegs; k++) {
        const uint32_t n = lane + k * G;
        if (n < N) s_v[n] = s_y[n];
    }
    Gp::bar();

    // ---- CG (ns/ILSS_CG.h:38-151): r = S(b - A y), d = r
    accumulate_forces<GW, false>(tr, s_R, lane, N, T, C, CL, real, cs, ref, M.m, s_v, s_f);
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kNodeRegs; k++) {
        const uint32_t n = lane + k * G;
        if (n >= N) continue;
        const float2 f = s_f[n], yk = s_y[n];
        s_f[n] = make_float2(0.f, 0.f);  // cleared for the first iteration's forces
        const float fac = mass[k] * M.fM;
        float q0 = -M.fK * f.x + fac * yk.x, q1 = -M.fK * f.y + fac * yk.y;
        if (kn[k]) { q0 = 0.f; q1 = 0.f; }
        float r0 = -1.0f * q0 + b[k].x, r1 = -1.0f * q1 + b[k].y;
        if (kn[k]) { r0 = 0.f; r1 = 0.f; }
        r[k] = make_float2(r0, r1);
        acc += (double)r0 * r0 + (double)r1 * r1;
    }
    uint32_t rpar = 0, fpar = 0;  // alternating reduction and flag buffers
    const double delta0 = Gp::sum(acc, red, rpar);
    double delta = delta0;
    const float abs_sq = M.abs_eps * M.abs_eps;
    const float rel_sq = M.rel * M.rel;
    const double target = (double)rel_sq * delta0;
    const double abs2 = (double)abs_sq;
    int reason = 0;
    uint32_t it = 0;
    if (!real) reason = -1;
    else if (!(delta0 > abs2)) reason = CF_CG_CONVERGED_ABS;
    else if (!(delta0 > target)) reason = CF_CG_CONVERGED_REL;
    else if (!(0u < M.max_iter)) reason = CF_CG_MAX_ITER;
    Gp::bar();  // every read of s_v (y) is done
#pragma unroll
    for (int k = 0; k < kNodeRegs; k++) {
        const uint32_t n = lane + k * G;
        if (n < N) s_v[n] = r[k];  // d = r
    }
    while (Gp::any(reason == 0, flags, fpar)) {
        const bool run = reason == 0;
        if (Gp::kAlone) Gp::bar();  // else any() has just passed a barrier
        accumulate_forces<GW, false, false>(tr, s_R, lane, N, T, C, CL, run, cs, ref, M.m, s_v, s_f);
        float2 q[kNodeRegs];
        acc = 0.0;
        if (run) {
#pragma unroll
            for (int k = 0; k < kNodeRegs; k++) {
                const uint32_t n = lane + k * G;
                if (n >= N) continue;
                const float2 f = s_f[n], dk = s_v[n];
                s_f[n] = make_float2(0.f, 0.f);  // cleared for the next iteration's forces
                const float fac = mass[k] * M.fM;
                float q0 = -M.fK * f.x + fac * dk.x, q1 = -M.fK * f.y + fac * dk.y;
                if (kn[k]) { q0 = 0.f; q1 = 0.f; }
                q[k] = make_float2(q0, q1);
                acc += (double)dk.x * q0 + (double)dk.y * q1;
            }
        }
        const double dq = Gp::sum(acc, red, rpar);
        if (run && fabs(dq) < abs2) reason = CF_CG_BREAKDOWN;
        const bool upd = run && reason == 0;
        const float alpha = upd ? (float)(delta / dq) : 0.f;
        acc = 0.0;
        if (upd) {
#pragma unroll
            for (int k = 0; k < kNodeRegs; k++) {
                const uint32_t n = lane + k * G;
                if (n >= N) continue;
                const float2 dk = s_v[n];
                float2 yk = s_y[n];
                yk.x = yk.x + dk.x * alpha; yk.y = yk.y + dk.y * alpha;
                s_y[n] = yk;
                r[k].x = r[k].x - q[k].x * alpha; r[k].y = r[k].y - q[k].y * alpha;
                acc += (double)r[k].x * r[k].x + (double)r[k].y * r[k].y;
            }
        }
        const double dn = Gp::sum(acc, red, rpar);
        float beta = 0.f;
        if (upd) {
            beta = (float)(dn / delta);
            delta = dn;
            it++;
            if (!(dn > abs2)) reason = CF_CG_CONVERGED_ABS;
            else if (!(dn > target)) reason = CF_CG_CONVERGED_REL;
            else if (!(it < M.max_iter)) reason = CF_CG_MAX_ITER;
        }
        // Every read of s_v (d) by another thread was in the colour passes,
        // ordered before this point by their last barrier and the two sums'
        // barriers; a lone wavefront keeps its wave barrier.
        if (GW == 1) Gp::bar();
        if (upd && reason == 0) {
#pragma unroll
            for (int k = 0; k < kNodeRegs; k++) {
                const uint32_t n = lane + k * G;
                if (n >= N) continue;
                const float2 dk = s_v[n];
                float d0 = beta * dk.x + r[k].x, d1 = beta * dk.y + r[k].y;
                if (kn[k]) { d0 = 0.f; d1 = 0.f; }
                s_v[n] = make_float2(d0, d1);
            }
        }
    }

    // ---- integrate: v = y, x = x + dt y (LeafDSH_Solid2D_FEM.cpp:1394-1398)
    Gp::bar();
#pragma unroll
    for (int k = 0; k < kNodeRegs; k++) {
        const uint32_t n = lane + k * G;
        if (n >= N) continue;
        const float2 yy = s_y[n], xx = xg[n];
        if (track) x0g[n] = xx;  // m_vecPos0 = x after BeginEvaluation (LeafDSH_Solid2D_FEM.cpp:974-985)
        const float2 xn = make_float2(xx.x + M.dt * yy.x, xx.y + M.dt * yy.y);
        const float2 vn = (M.drag > 0.f) ? make_float2((1.0f - M.dt * M.drag) * yy.x, (1.0f - M.dt * M.drag) * yy.y)
                                         : yy;
        xg[n] = xn;
        vg[n] = vn;
        s_v[n] = xn;
    }
    Gp::bar();

    // ---- elastic energy on the new positions with the step's rotations (C_V)
    double en = 0.0;
    if (M.compute_energy) {
#pragma unroll 1
        for (uint32_t k = 0; k < uint32_t(kTriRegs); k++) {
            const uint32_t e = lane + k * G;
            if (e >= T) break;
            const Tri t = pick(tr, k);
            const Rot R = rot_of(s_R[k * G + lane]);
            float H[4];
            tri_gradient<true>(t, R, s_v[t.n0()], s_v[t.n1()], s_v[t.n2()], ref[t.n0()], ref[t.n1()], ref[t.n2()], H);
            const float sxx = fmaf(M.m.a_k1, H[0], M.m.a_nu * H[3]);
            const float syy = fmaf(M.m.a_k1, H[3], M.m.a_nu * H[0]);
            const float sxy = M.m.a_k2 * (H[1] + H[2]);
            const float sH = sxx * H[0] + syy * H[3] + sxy * (H[1] + H[2]);
            en += 0.5 * (double)t.A * (double)sH;
        }
    }
    const double energy = Gp::sum(en, red, rpar);
    const double nd = Gp::sum(ndeg, red, rpar);
    const double sd = Gp::sum(sdeg, red, rpar);
    const double nf = Gp::sum(nfb, red, rpar);
    const float md = Gp::min(mind, redm);
    if (real && lane == 0) {
        Sim2Out& o = a.out[sim];
        o.iters = int32_t(it);
        o.reason = reason;
        o.rel = (delta0 > 0) ? float(sqrt(delta / delta0)) : 0.f;
        o.abs_res = float(sqrt(delta));
        o.energy = energy;
        o.sum_degenerate = sd;
        o.num_degenerate = uint32_t(nd);
        o.num_fallback = uint32_t(nf);
        o.min_det = md;
        if (!(isfinite(delta0) && isfinite(delta))) *a.nonfinite = 1;
    }
}

}  // namespace

// One workgroup per work item: 8/gw meshes, one per group of gw wavefronts.
// LDS per group: three float2 node arrays of 64 gw kNodeRegs entries (d or
// positions, forces, the solution y) and one float4 rotation per owned
// triangle, 68.7 KB per workgroup. Bounded to 4 waves per SIMD (at most 128
// VGPRs), so that two workgroups share a CU and hide each other's barrier
// and LDS latency.
__global__ __launch_bounds__(k2Threads, 4) void k2_step(Batch2Args a) {
    __shared__ float2 s_v[k2Threads * kNodeRegs];
    __shared__ float2 s_f[k2Threads * kNodeRegs];
    __shared__ float2 s_y[k2Threads * kNodeRegs];
    __shared__ float4 s_R[k2Threads * kTriRegs];
    __shared__ double red[16];
    __shared__ float redm[8];
    __shared__ uint32_t flags[16];
    const uint32_t gw = a.work[blockIdx.x].gw, cmax = a.work[blockIdx.x].cmax;
    switch (gw) {
        case 1: mesh_step<1>(a, cmax, s_v, s_f, s_y, s_R, red, redm, flags); break;
        case 2: mesh_step<2>(a, cmax, s_v, s_f, s_y, s_R, red, redm, flags); break;
        case 4: mesh_step<4>(a, cmax, s_v, s_f, s_y, s_R, red, redm, flags); break;
        default: mesh_step<8>(a, cmax, s_v, s_f, s_y, s_R, red, redm, flags); break;
    }
}

namespace {
// One thread per deformation gradient: the triangle rotation of the step kernel.
__global__ void __launch_bounds__(256) k2_rotations(const float4* __restrict__ F, const float* __restrict__ det,
                                                    uint32_t n, int rm, float thr, float4* __restrict__ R,
                                                    uint32_t* __restrict__ fallback) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float4 f = F[i];
    Rot t;
    const float fb = rotation2(rm, thr, f.x, f.y, f.z, f.w, det[i], t);
    R[i] = make_float4(t.r00, t.r01, t.r10, t.r11);
    fallback[i] = fb != 0.0f;
}
}  // namespace

int launch_rotations2(const float* F, const float* det, uint32_t n, int rm, float thr, float* R, uint32_t* fallback,
                      hipStream_t s) {
    if (!n) return CF_OK;
    hipLaunchKernelGGL(k2_rotations, dim3((n + 255) / 256), dim3(256), 0, s, reinterpret_cast<const float4*>(F),
                       det, n, rm, thr, reinterpret_cast<float4*>(R), fallback);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("2D rotation launch failed: %s", hipGetErrorString(e));
        return CF_ERR_HIP;
    }
    return CF_OK;
}

int launch_batch2_step(const Batch2Args& a, uint32_t n_work, hipStream_t s) {
    if (!n_work) return CF_OK;
    hipLaunchKernelGGL(k2_step, dim3(n_work), dim3(k2Threads), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("2D step launch failed: %s", hipGetErrorString(e));
        return CF_ERR_HIP;
    }
    return CF_OK;
}

// ============================================================ large 2D meshes
// A mesh beyond one workgroup, one thread per triangle or node per pass, with
// the resident step's arithmetic: the same rotation, node-of-collapse,
// gradient and force functions, and the same b / y / q / CG expressions. A
// node's force is its incident (triangle, vertex) entries summed from 0 in
// ascending triangle order. FP64 reductions: per block, then the blocks'
// partials in the fixed one-block order.
namespace {

__device__ inline Tri l2_tri(const Large2Dev& a, uint32_t e, uint32_t& i0, uint32_t& i1, uint32_t& i2) {
    i0 = a.tri[3 * (size_t)e]; i1 = a.tri[3 * (size_t)e + 1]; i2 = a.tri[3 * (size_t)e + 2];
    const float4 g = a.dinv[e];
    Tri t;
    t.i01 = 0; t.i2 = 0;  // node ids are passed separately (meshes here may exceed 16-bit ids)
    t.g00 = g.x; t.g01 = g.y; t.g10 = g.z; t.g11 = g.w;
    t.A = a.area[e];
    return t;
}

__device__ inline float2 l2_node_force(const Large2Dev& a, uint32_t n) {
#pragma clang fp contract(off)
    float2 f = make_float2(0.f, 0.f);
    for (uint32_t c = a.ptr[n]; c < a.ptr[n + 1]; c++) {
        const float2 t = a.tf[a.code[c]];
        f.x += t.x; f.y += t.y;
    }
    return f;
}

__device__ inline float block_min(float v, float* sh) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = sh[0];
    for (int i = 1; i < kBlock / 64; i++) r = fminf(r, sh[i]);
    return r;
}

// BeginEvaluation: F, det F, node of collapse, R (as the resident step)
__global__ __launch_bounds__(kBlock) void kl2_begin(Large2Dev a) {
#pragma clang fp contract(off)
    __shared__ double sh[kBlock / 64];
    __shared__ float shm[kBlock / 64];
    const Mesh2Dev& M = a.c;
    const bool track = M.rm == CF_RM_DAPD;
    double ndeg = 0.0, sdeg = 0.0, nfb = 0.0;
    float mind = 3.0e38f;
    for (uint32_t e = blockIdx.x * kBlock + threadIdx.x; e < a.T; e += gridDim.x * kBlock) {
        uint32_t i0, i1, i2;
        const Tri t = l2_tri(a, e, i0, i1, i2);
        const float2 p0 = a.x[i0], p1 = a.x[i1], p2 = a.x[i2];
        const float D00 = p1.x - p0.x, D01 = p2.x - p0.x, D10 = p1.y - p0.y, D11 = p2.y - p0.y;
        float f00 = 0.0f; f00 += D00 * t.g00; f00 += D01 * t.g10;
        float f01 = 0.0f; f01 += D00 * t.g01; f01 += D01 * t.g11;
        float f10 = 0.0f; f10 += D10 * t.g00; f10 += D11 * t.g10;
        float f11 = 0.0f; f11 += D10 * t.g01; f11 += D11 * t.g11;
        const float det = f00 * f11 - f10 * f01;
        uint32_t nc = kNocInvalid;
        if (track) {
            const uint32_t old = a.noc[e];
            nc = (det > M.thr) ? kNocInvalid
                               : update_noc2(a.x0[i0], a.x0[i1], a.x0[i2], p0, p1, p2, t.g00, t.g01, t.g10, t.g11,
                                             t.A, M.thr, old);
            if (nc != old) a.noc[e] = uint8_t(nc);
        }
        if (track && !(det > M.thr) && nc != kNocInvalid) {
            a.R[e] = r_dapd_degenerate2(p0, p1, p2, nc, t.A, t.g00, t.g01, t.g10, t.g11, M.thr, M.dapd_l, M.dapd_nl,
                                        M.dapd_e);
        } else {
            Rot rt;
            nfb += rotation2(M.rm, M.thr, f00, f01, f10, f11, det, rt);
            a.R[e] = make_float4(rt.r00, rt.r01, rt.r10, rt.r11);
        }
        if (det <= 0.f) { ndeg += 1.0; sdeg += det; }
        mind = fminf(mind, det);
    }
    const double s0 = block_sum(ndeg, sh);
    const double s1 = block_sum(sdeg, sh);
    const double s2 = block_sum(nfb, sh);
    const float mn = block_min(mind, shm);
    if (threadIdx.x == 0) {
        a.stat[blockIdx.x] = s0;
        a.stat[a.nblk_tri + blockIdx.x] = s1;
        a.stat[2 * a.nblk_tri + blockIdx.x] = s2;
        a.stat_min[blockIdx.x] = mn;
    }
}

// Triangle forces on the vectors `in` (FE: f_e on positions, with the rest
// positions) into tf
template <bool FE>
__global__ __launch_bounds__(kBlock) void kl2_elem(Large2Dev a, const float2* __restrict__ in, const int32_t* skip) {
    if (skip && *skip) return;
    for (uint32_t e = blockIdx.x * kBlock + threadIdx.x; e < a.T; e += gridDim.x * kBlock) {
        uint32_t i0, i1, i2;
        const Tri t = l2_tri(a, e, i0, i1, i2);
        const Rot R = rot_of(a.R[e]);
        float H[4];
        const float2 z = make_float2(0.f, 0.f);
        if (FE) tri_gradient<true>(t, R, in[i0], in[i1], in[i2], a.ref[i0], a.ref[i1], a.ref[i2], H);
        else tri_gradient<false>(t, R, in[i0], in[i1], in[i2], z, z, z, H);
        float2 f0, f1, f2;
        tri_forces(t, R, H, a.c.m, f0, f1, f2);
        a.tf[3 * (size_t)e] = f0;
        a.tf[3 * (size_t)e + 1] = f1;
        a.tf[3 * (size_t)e + 2] = f2;
    }
}

// b = dt (M g + f_e) + M v ; y = v (warm start) ; KNC rows take the
// constraint velocity (Step_QuasiImplicitEuler_v, LeafDSH_Solid2D_FEM.cpp:1287-1403)
__global__ __launch_bounds__(kBlock) void kl2_rhs(Large2Dev a) {
#pragma clang fp contract(off)
    const Mesh2Dev& M = a.c;
    for (uint32_t n = blockIdx.x * kBlock + threadIdx.x; n < a.N; n += gridDim.x * kBlock) {
        const float2 f = l2_node_force(a, n), vv = a.v[n];
        const float m = a.mass[n];
        float2 b;
        b.x = (m * M.g0 + f.x) * M.dt + m * vv.x;
        b.y = (m * M.g1 + f.y) * M.dt + m * vv.y;
        float2 yk = M.warm_start ? vv : make_float2(0.f, 0.f);
        if (a.knc[n]) {
            const float2 q = a.ref[n], xx = a.x[n];
            float vx = (q.x - xx.x) * M.inv_dt, vy = (q.y - xx.y) * M.inv_dt;
            const float nk = sqrtf(vx * vx + vy * vy);
            if (nk > 0.001f) {
                const float inv = 1.0f / nk, sc = fminf(nk, 50.0f);
                vx = sc * (vx * inv); vy = sc * (vy * inv);
            } else {
                vx = 0.f; vy = 0.f;
            }
            b = make_float2(vx, vy);
            yk = b;
        }
        a.b[n] = b;
        a.y[n] = yk;
    }
}

// q = S(fM M d - fK f) with per-block d.q; INIT: the CG start r = S(b - q),
// d = r with per-block r.r instead (q not stored)
template <bool INIT>
__global__ __launch_bounds__(kBlock) void kl2_matvec(Large2Dev a, const float2* __restrict__ in, const int32_t* skip) {
#pragma clang fp contract(off)
    __shared__ double sh[kBlock / 64];
    if (skip && *skip) return;
    const Mesh2Dev& M = a.c;
    double acc = 0.0;
    for (uint32_t n = blockIdx.x * kBlock + threadIdx.x; n < a.N; n += gridDim.x * kBlock) {
        const float2 f = l2_node_force(a, n), dk = in[n];
        const float fac = a.mass[n] * M.fM;
        const bool kn = a.knc[n] != 0;
        float q0 = -M.fK * f.x + fac * dk.x, q1 = -M.fK * f.y + fac * dk.y;
        if (kn) { q0 = 0.f; q1 = 0.f; }
        if (INIT) {
            const float2 bb = a.b[n];
            float r0 = -1.0f * q0 + bb.x, r1 = -1.0f * q1 + bb.y;
            if (kn) { r0 = 0.f; r1 = 0.f; }
            a.r[n] = make_float2(r0, r1);
            a.d[n] = make_float2(r0, r1);
            acc += (double)r0 * r0 + (double)r1 * r1;
        } else {
            a.q[n] = make_float2(q0, q1);
            acc += (double)dk.x * q0 + (double)dk.y * q1;
        }
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) (INIT ? a.part_b : a.part_a)[blockIdx.x] = s;
}

// alpha = delta_k / d.q (break down if |d.q| < abs^2) ; y += alpha d ;
// r -= alpha q ; per-block r.r
__global__ __launch_bounds__(kBlock) void kl2_xr(Large2Dev a, uint32_t k) {
#pragma clang fp contract(off)
    __shared__ double sh[kBlock / 64];
    CgScalars* st = a.st;
    if (st->done[k]) return;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const double dq = reduce_partials(a.part_a, (int)a.nblk_node, sh);
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
    double acc = 0.0;
    for (uint32_t n = blockIdx.x * kBlock + threadIdx.x; n < a.N; n += gridDim.x * kBlock) {
        const float2 dk = a.d[n], qk = a.q[n];
        float2 yk = a.y[n], rk = a.r[n];
        yk.x = yk.x + dk.x * alpha; yk.y = yk.y + dk.y * alpha;
        rk.x = rk.x - qk.x * alpha; rk.y = rk.y - qk.y * alpha;
        a.y[n] = yk;
        a.r[n] = rk;
        acc += (double)rk.x * rk.x + (double)rk.y * rk.y;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) a.part_b[blockIdx.x] = s;
}

// delta_{k+1} = r.r ; beta ; stop test ; d = S(beta d + r). host_done, when
// not null, receives done[k + 1].
__global__ __launch_bounds__(kBlock) void kl2_d(Large2Dev a, uint32_t k, int32_t* host_done) {
#pragma clang fp contract(off)
    __shared__ double sh[kBlock / 64];
    CgScalars* st = a.st;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (st->done[k]) {
        if (lead) {
            st->done[k + 1] = 1;
            if (host_done) host_done[k + 1] = 1;
        }
        return;
    }
    if (st->reason[k + 1] == CF_CG_BREAKDOWN) {
        if (lead && host_done) host_done[k + 1] = 2;
        return;
    }
    const double dn = reduce_partials(a.part_b, (int)a.nblk_node, sh);
    const uint32_t k1 = k + 1;
    int reason = 0;
    if (!(dn > st->abs_sq)) reason = CF_CG_CONVERGED_ABS;
    else if (!(dn > st->target)) reason = CF_CG_CONVERGED_REL;
    else if (!(k1 < st->max_iter)) reason = CF_CG_MAX_ITER;
    const float beta = (float)(dn / st->delta[k]);
    if (lead) {
        st->delta[k1] = dn;
        st->beta[k] = beta;
        st->reason[k1] = reason;
        st->done[k1] = reason != 0;
        if (host_done) host_done[k1] = reason != 0;
    }
    if (reason) return;
    for (uint32_t n = blockIdx.x * kBlock + threadIdx.x; n < a.N; n += gridDim.x * kBlock) {
        const float2 dk = a.d[n], rk = a.r[n];
        float d0 = beta * dk.x + rk.x, d1 = beta * dk.y + rk.y;
        if (a.knc[n]) { d0 = 0.f; d1 = 0.f; }
        a.d[n] = make_float2(d0, d1);
    }
}

// v = y, x = x + dt y (LeafDSH_Solid2D_FEM.cpp:1394-1398); DAPD keeps the
// previous positions (m_vecPos0)
__global__ __launch_bounds__(kBlock) void kl2_integrate(Large2Dev a) {
#pragma clang fp contract(off)
    const Mesh2Dev& M = a.c;
    const bool track = M.rm == CF_RM_DAPD;
    for (uint32_t n = blockIdx.x * kBlock + threadIdx.x; n < a.N; n += gridDim.x * kBlock) {
        const float2 yy = a.y[n], xx = a.x[n];
        if (track) a.x0[n] = xx;
        a.x[n] = make_float2(xx.x + M.dt * yy.x, xx.y + M.dt * yy.y);
        a.v[n] = (M.drag > 0.f) ? make_float2((1.0f - M.dt * M.drag) * yy.x, (1.0f - M.dt * M.drag) * yy.y) : yy;
    }
}

// C_V on the new positions with the step's rotations, per-block sums
__global__ __launch_bounds__(kBlock) void kl2_energy(Large2Dev a) {
#pragma clang fp contract(off)
    __shared__ double sh[kBlock / 64];
    const Mesh2Dev& M = a.c;
    double en = 0.0;
    for (uint32_t e = blockIdx.x * kBlock + threadIdx.x; e < a.T; e += gridDim.x * kBlock) {
        uint32_t i0, i1, i2;
        const Tri t = l2_tri(a, e, i0, i1, i2);
        const Rot R = rot_of(a.R[e]);
        float H[4];
        tri_gradient<true>(t, R, a.x[i0], a.x[i1], a.x[i2], a.ref[i0], a.ref[i1], a.ref[i2], H);
        const float sxx = fmaf(M.m.a_k1, H[0], M.m.a_nu * H[3]);
        const float syy = fmaf(M.m.a_k1, H[3], M.m.a_nu * H[0]);
        const float sxy = M.m.a_k2 * (H[1] + H[2]);
        const float sH = sxx * H[0] + syy * H[3] + sxy * (H[1] + H[2]);
        en += 0.5 * (double)t.A * (double)sH;
    }
    const double s = block_sum(en, sh);
    if (threadIdx.x == 0) a.stat[3 * a.nblk_tri + blockIdx.x] = s;
}

// The step's BeginEvaluation statistics and energy (one workgroup)
__global__ __launch_bounds__(kBlock) void kl2_finish(Large2Dev a, int with_energy) {
    __shared__ double sh[kBlock / 64];
    __shared__ float shm[kBlock / 64];
    const int nb = (int)a.nblk_tri;
    const double nd = reduce_partials(a.stat, nb, sh);
    const double sd = reduce_partials(a.stat + nb, nb, sh);
    const double nf = reduce_partials(a.stat + 2 * nb, nb, sh);
    const double en = with_energy ? reduce_partials(a.stat + 3 * nb, nb, sh) : 0.0;
    float mn = 3.0e38f;
    for (int i = threadIdx.x; i < nb; i += kBlock) mn = fminf(mn, a.stat_min[i]);
    mn = block_min(mn, shm);
    if (threadIdx.x == 0) {
        Sim2Out& o = *a.out;
        o.energy = en;
        o.sum_degenerate = sd;
        o.num_degenerate = uint32_t(nd);
        o.num_fallback = uint32_t(nf);
        o.min_det = mn;
    }
}

int check_l2() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("large 2D kernel launch failed: %s", hipGetErrorString(e));
        return CF_ERR_HIP;
    }
    return CF_OK;
}

}  // namespace

int launch_l2_begin(const Large2Dev& a, hipStream_t s) {
    hipLaunchKernelGGL(kl2_begin, dim3(a.nblk_tri), dim3(kBlock), 0, s, a);
    return check_l2();
}
int launch_l2_elem(const Large2Dev& a, bool fe, const float2* in, const int32_t* skip, hipStream_t s) {
    if (fe) hipLaunchKernelGGL(kl2_elem<true>, dim3(a.nblk_tri), dim3(kBlock), 0, s, a, in, skip);
    else hipLaunchKernelGGL(kl2_elem<false>, dim3(a.nblk_tri), dim3(kBlock), 0, s, a, in, skip);
    return check_l2();
}
int launch_l2_rhs(const Large2Dev& a, hipStream_t s) {
    hipLaunchKernelGGL(kl2_rhs, dim3(a.nblk_node), dim3(kBlock), 0, s, a);
    return check_l2();
}
int launch_l2_matvec(const Large2Dev& a, bool init, const float2* in, const int32_t* skip, hipStream_t s) {
    if (init) hipLaunchKernelGGL(kl2_matvec<true>, dim3(a.nblk_node), dim3(kBlock), 0, s, a, in, skip);
    else hipLaunchKernelGGL(kl2_matvec<false>, dim3(a.nblk_node), dim3(kBlock), 0, s, a, in, skip);
    return check_l2();
}
int launch_l2_xr(const Large2Dev& a, uint32_t k, hipStream_t s) {
    hipLaunchKernelGGL(kl2_xr, dim3(a.nblk_node), dim3(kBlock), 0, s, a, k);
    return check_l2();
}
int launch_l2_d(const Large2Dev& a, uint32_t k, int32_t* host_done, hipStream_t s) {
    hipLaunchKernelGGL(kl2_d, dim3(a.nblk_node), dim3(kBlock), 0, s, a, k, host_done);
    return check_l2();
}
int launch_l2_integrate(const Large2Dev& a, hipStream_t s) {
    hipLaunchKernelGGL(kl2_integrate, dim3(a.nblk_node), dim3(kBlock), 0, s, a);
    return check_l2();
}
int launch_l2_energy(const Large2Dev& a, hipStream_t s) {
    hipLaunchKernelGGL(kl2_energy, dim3(a.nblk_tri), dim3(kBlock), 0, s, a);
    return check_l2();
}
int launch_l2_finish(const Large2Dev& a, hipStream_t s) {
    hipLaunchKernelGGL(kl2_finish, dim3(1), dim3(kBlock), 0, s, a, (int)a.c.compute_energy);
    return check_l2();
}

}  // namespace cf
