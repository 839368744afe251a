// Reference-order FP32 evaluation of a 3D leaf: the solid's step computed with
// the reference's own arithmetic and summation order, so that its results are
// bit-identical to the reference's FP32 solver on the same inputs.
//
//   - Ke stored per element as 16 3x3 blocks from TetrahedronElement3::Compute_K
//     (ms/fem/TetrahedronElement3.h:85-173), computed on the host;
//   - C_f_e / C_df_x (ms/fem/TetrahedronElementSet3.cpp:1053-1096,1761-1836):
//     per element and vertex i, R * (K_i0 l_0 + K_i1 l_1 + K_i2 l_2 + K_i3 l_3)
//     with Mal's row-by-row matrix-vector products (Mal/GMatUtils.h:13-24),
//     subtracted from the node's value in element order (the reference's serial
//     scatter), here by one thread per node walking its incident (element,
//     vertex) pairs in ascending element order;
//   - ns::real_array::dot (ns/real_array.h:34-49, SSE4.1): each 4-float chunk
//     reduced as (p0 + p1) + (p2 + p3) by _mm_dp_ps and added serially, the lanes
//     past n taken as zero;
//   - ILSS_CG (ns/ILSS_CG.h:38-151) with float scalars, and the QIE-v step's
//     vector updates (ds/dsh/LeafDSH_Solid3D_FEM.cpp:1205-1328) term by term.
// Every function here is compiled without FP contraction.
#include "cf_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace cf {
namespace {

// One element's stiffness matrix (TetrahedronElement3::Compute_K): B from the
// AFEM-16 inverse-Jacobian formula, the isotropic E, K = V B^T (E B), products
// accumulated from 0 in index order. Returns the volume.
float ke_reference_order(const float* r0, const float* r1, const float* r2, const float* r3, float E, float nu,
                         float K[144]) {
#pragma clang fp contract(off)
    // Compute_D: columns r1 - r0, r2 - r0, r3 - r0; Det (Mal/GMat.h:236-244)
    float D[3][3];
    for (int i = 0; i < 3; i++) { D[i][0] = r1[i] - r0[i]; D[i][1] = r2[i] - r0[i]; D[i][2] = r3[i] - r0[i]; }
    const float det = D[0][0] * D[1][1] * D[2][2] - D[0][0] * D[2][1] * D[1][2] + D[1][0] * D[2][1] * D[0][2] -
                      D[1][0] * D[0][1] * D[2][2] + D[2][0] * D[0][1] * D[1][2] - D[2][0] * D[1][1] * D[0][2];
    const float volume = det / 6.0f;
    const float x1 = r0[0], x2 = r1[0], x3 = r2[0], x4 = r3[0];
    const float y1 = r0[1], y2 = r1[1], y3 = r2[1], y4 = r3[1];
    const float z1 = r0[2], z2 = r1[2], z3 = r2[2], z4 = r3[2];
    // derivative columns of the inverse Jacobian (rows n = 1..4, columns x, y, z)
    float J[4][3];
    J[0][0] = (y4 - y2) * (z3 - z2) - (y3 - y2) * (z4 - z2);
    J[0][1] = (x3 - x2) * (z4 - z2) - (x4 - x2) * (z3 - z2);
    J[0][2] = (x4 - x2) * (y3 - y2) - (x3 - x2) * (y4 - y2);
    J[1][0] = (y3 - y1) * (z4 - z3) - (y3 - y4) * (z1 - z3);
    J[1][1] = (x4 - x3) * (z3 - z1) - (x1 - x3) * (z3 - z4);
    J[1][2] = (x3 - x1) * (y4 - y3) - (x3 - x4) * (y1 - y3);
    J[2][0] = (y2 - y4) * (z1 - z4) - (y1 - y4) * (z2 - z4);
    J[2][1] = (x1 - x4) * (z2 - z4) - (x2 - x4) * (z1 - z4);
    J[2][2] = (x2 - x4) * (y1 - y4) - (x1 - x4) * (y2 - y4);
    J[3][0] = (y1 - y3) * (z2 - z1) - (y1 - y2) * (z3 - z1);
    J[3][1] = (x2 - x1) * (z1 - z3) - (x3 - x1) * (z1 - z2);
    J[3][2] = (x1 - x3) * (y2 - y1) - (x1 - x2) * (y3 - y1);
    const float s = 1.0f / (6.0f * volume);
    float B[6][12];
    std::memset(B, 0, sizeof(B));
    for (int n = 0; n < 4; n++) {
        const float a = J[n][0] * s, b = J[n][1] * s, c = J[n][2] * s;
        B[0][3 * n + 0] = a;
        B[1][3 * n + 1] = b;
        B[2][3 * n + 2] = c;
        B[3][3 * n + 0] = b; B[3][3 * n + 1] = a;
        B[4][3 * n + 1] = c; B[4][3 * n + 2] = b;
        B[5][3 * n + 0] = c; B[5][3 * n + 2] = a;
    }
    float Em[6][6];
    std::memset(Em, 0, sizeof(Em));
    const float aa = E / ((1.0f + nu) * (1.0f - 2.0f * nu));
    const float a_nu = aa * nu, a_k1 = aa * (1.0f - nu), a_k2 = aa * (0.5f - nu);
    Em[0][0] = a_k1; Em[0][1] = a_nu; Em[0][2] = a_nu;
    Em[1][0] = a_nu; Em[1][1] = a_k1; Em[1][2] = a_nu;
    Em[2][0] = a_nu; Em[2][1] = a_nu; Em[2][2] = a_k1;
    Em[3][3] = a_k2; Em[4][4] = a_k2; Em[5][5] = a_k2;
    float EB[6][12];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 12; j++) {
            float acc = 0.0f;
            for (int k = 0; k < 6; k++) acc += Em[i][k] * B[k][j];
            EB[i][j] = acc;
        }
    for (int i = 0; i < 12; i++)
        for (int j = 0; j < 12; j++) {
            float acc = 0.0f;
            for (int k = 0; k < 6; k++) acc += B[k][i] * EB[k][j];
            K[i * 12 + j] = acc * volume;
        }
    return volume;
}

// Mal's matrix-vector product, row by row from 0 (Mal/GMatUtils.h:13-24).
__device__ inline void mvec3(const float m[9], const float v[3], float out[3]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float acc = 0.0f;
        acc += m[3 * i + 0] * v[0];
        acc += m[3 * i + 1] * v[1];
        acc += m[3 * i + 2] * v[2];
        out[i] = acc;
    }
}

// Per caller element e: the four vertex terms of C_f_e / C_df_x / C_V.
//   mode kExactFe:  l_j = R^T x_j - r_j, W[e][i] = R sum_j K_ij l_j
//   mode kExactDfx: l_j = R^T dx_j,      W[e][i] = R sum_j K_ij l_j
//   mode kExactV:   l_j = R^T x_j - r_j, E[e][i] = l_i . sum_j K_ij l_j
__global__ __launch_bounds__(kBlock) void k_exact_elem(int mode, uint32_t T, const uint32_t* __restrict__ tets,
                                                       const uint32_t* __restrict__ tile_of,
                                                       const float* __restrict__ Rp, uint32_t r_stride,
                                                       const float* __restrict__ kb, const float* __restrict__ ref,
                                                       const float* __restrict__ in, float* __restrict__ out) {
#pragma clang fp contract(off)
    const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= T) return;
    const uint32_t ti = tile_of[e];
    float R[9], Rt[9];
#pragma unroll
    for (int k = 0; k < 9; k++) R[k] = Rp[(size_t)k * r_stride + ti];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Rt[3 * i + j] = R[3 * j + i];
    float l[4][3];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t n = tets[4 * (size_t)e + j];
        const float v[3] = {in[3 * (size_t)n], in[3 * (size_t)n + 1], in[3 * (size_t)n + 2]};
        mvec3(Rt, v, l[j]);
        if (mode != kExactDfx) {
            l[j][0] = l[j][0] - ref[3 * (size_t)n];
            l[j][1] = l[j][1] - ref[3 * (size_t)n + 1];
            l[j][2] = l[j][2] - ref[3 * (size_t)n + 2];
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        float acc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            float K[9], t[3];
#pragma unroll
            for (int k = 0; k < 9; k++) K[k] = kb[(size_t)((4 * i + j) * 9 + k) * T + e];
            mvec3(K, l[j], t);
            if (j == 0) {
                acc[0] = t[0]; acc[1] = t[1]; acc[2] = t[2];
            } else {
                acc[0] = acc[0] + t[0]; acc[1] = acc[1] + t[1]; acc[2] = acc[2] + t[2];
            }
        }
        if (mode == kExactV) {
            float d = 0.0f;
            d += l[i][0] * acc[0];
            d += l[i][1] * acc[1];
            d += l[i][2] * acc[2];
            out[4 * (size_t)e + i] = d;
        } else {
            float w[3];
            mvec3(R, acc, w);
            float* o = out + 12 * (size_t)e + 3 * i;
            o[0] = w[0]; o[1] = w[1]; o[2] = w[2];
        }
    }
}

// Per node n: its value minus the vertex terms of its elements in ascending
// element order (the reference's serial scatter, o[c] -= w[c]), then the
// caller's node terms:
//   kExactRhs: b = ((0 + m g) - terms) * dt + m v; y = v (warm start) or 0;
//              KNC rows b = y = the constraint velocity (:1209-1295)
//   kExactAv:  q = (0 - terms) * (-fK) + (m fM) d; KNC rows q = 0
//              (LS_A_times_d :1056-1079 + LS_ZeroKNC :1045-1053)
struct ExactNodeArgs {
    const uint32_t* ptr;
    const uint32_t* code;
    const float* w;
    const uint8_t* knc;     // may be null
    const float* knc_target;
    const float* x;
    const float* v;
    const float* d;
    float* out;
    float* y;
    float mass, dt, inv_dt, g0, g1, g2, neg_fk, m_fm;
    int warm_start;
};

__global__ __launch_bounds__(kBlock) void k_exact_node(int mode, uint32_t N, ExactNodeArgs a) {
#pragma clang fp contract(off)
    const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= N) return;
    float s0, s1, s2;
    if (mode == kExactRhs) {
        s0 = 0.0f + a.mass * a.g0;
        s1 = 0.0f + a.mass * a.g1;
        s2 = 0.0f + a.mass * a.g2;
    } else {
        s0 = s1 = s2 = 0.0f;
    }
    const uint32_t k1 = a.ptr[n + 1];
    for (uint32_t k = a.ptr[n]; k < k1; k++) {
        const float* w = a.w + 3 * (size_t)a.code[k];
        s0 = s0 - w[0];
        s1 = s1 - w[1];
        s2 = s2 - w[2];
    }
    const size_t i = 3 * (size_t)n;
    const bool fixed = a.knc && a.knc[n];
    if (mode == kExactRhs) {
        float b0 = s0 * a.dt, b1 = s1 * a.dt, b2 = s2 * a.dt;
        b0 = b0 + a.mass * a.v[i]; b1 = b1 + a.mass * a.v[i + 1]; b2 = b2 + a.mass * a.v[i + 2];
        float y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
        if (a.warm_start) { y0 = a.v[i]; y1 = a.v[i + 1]; y2 = a.v[i + 2]; }
        if (fixed) {
            // vk = (target - x) / dt; |vk| > 0.001: min(|vk|, 50) vk / |vk| (Vec3::Normalized
            // scales by 1/|vk|, Mal/GVec.h:85), else 0
            const float v0 = (a.knc_target[i] - a.x[i]) * a.inv_dt;
            const float v1 = (a.knc_target[i + 1] - a.x[i + 1]) * a.inv_dt;
            const float v2 = (a.knc_target[i + 2] - a.x[i + 2]) * a.inv_dt;
            float sq = 0.0f;
            sq += v0 * v0; sq += v1 * v1; sq += v2 * v2;
            const float nk = sqrtf(sq);
            float k0 = 0.0f, k1v = 0.0f, k2 = 0.0f;
            if (nk > 0.001f) {
                const float inv = 1.0f / nk;
                const float m = nk < 50.0f ? nk : 50.0f;
                k0 = m * (inv * v0); k1v = m * (inv * v1); k2 = m * (inv * v2);
            }
            b0 = y0 = k0; b1 = y1 = k1v; b2 = y2 = k2;
        }
        a.out[i] = b0; a.out[i + 1] = b1; a.out[i + 2] = b2;
        a.y[i] = y0; a.y[i + 1] = y1; a.y[i + 2] = y2;
    } else {
        float q0 = s0 * a.neg_fk, q1 = s1 * a.neg_fk, q2 = s2 * a.neg_fk;
        q0 = q0 + a.m_fm * a.d[i]; q1 = q1 + a.m_fm * a.d[i + 1]; q2 = q2 + a.m_fm * a.d[i + 2];
        if (fixed) q0 = q1 = q2 = 0.0f;
        a.out[i] = q0; a.out[i + 1] = q1; a.out[i + 2] = q2;
    }
}

// ns::real_array::dot over n floats in one workgroup: each 4-float chunk
// (p0 + p1) + (p2 + p3), zero past n, added in chunk order to an accumulator
// that starts at 0. mode 1: the sum of n floats in order instead (C_V's
// serial energy sum, acc += term), then halved.
__global__ __launch_bounds__(kBlock) void k_exact_dot(int mode, uint32_t n, const float* __restrict__ a,
                                                      const float* __restrict__ b, float* __restrict__ out) {
#pragma clang fp contract(off)
    __shared__ float4 chunk4[kBlock / 4];
    float* chunk = reinterpret_cast<float*>(chunk4);
    const uint32_t nchunk = mode == 1 ? n : (n + 3) / 4;
    float acc = 0.0f;
    for (uint32_t base = 0; base < nchunk; base += kBlock) {
        const uint32_t k = base + threadIdx.x;
        if (k < nchunk) {
            if (mode == 1) {
                chunk[threadIdx.x] = a[k];
            } else {
                float p[4];
#pragma unroll
                for (int l = 0; l < 4; l++) {
                    const uint32_t i = 4 * k + l;
                    p[l] = (i < n) ? a[i] * b[i] : 0.0f;
                }
                chunk[threadIdx.x] = (p[0] + p[1]) + (p[2] + p[3]);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            // in chunk order; the float4 reads only let the loads run ahead of the adds
            const uint32_t m = std::min<uint32_t>(kBlock, nchunk - base);
            uint32_t j = 0;
#pragma unroll 8
            for (; j + 4 <= m; j += 4) {
                const float4 c = chunk4[j / 4];
                acc = acc + c.x;
                acc = acc + c.y;
                acc = acc + c.z;
                acc = acc + c.w;
            }
            for (; j < m; j++) acc = acc + chunk[j];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = mode == 1 ? 0.5f * acc : acc;
}

// ILSS_CG's vector updates (ns/ILSS_CG.h:53-118, real_array.h:66-78), with
// S (zero the KNC rows) where the algorithm applies it.
//   0: r = -1 r + b, S(r), d = r                      (start)
//   1: x = x + d alpha, r = r - q alpha               (addeq_scaled / subeq_scaled)
//   2: d = beta d + r, S(d)                           (muleq_addeq)
//   3: v = y, x = x + dt y                            (QIE-v integration :1322-1327)
__global__ __launch_bounds__(kBlock) void k_exact_vec(int mode, uint32_t n3, float s, const uint8_t* __restrict__ knc,
                                                      float* __restrict__ x, float* __restrict__ r,
                                                      float* __restrict__ d, const float* __restrict__ q) {
#pragma clang fp contract(off)
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n3; i += gridDim.x * kBlock) {
        const bool fixed = knc && knc[i / 3];
        if (mode == 0) {
            float v = -1.0f * r[i] + q[i];
            if (fixed) v = 0.0f;
            r[i] = v;
            d[i] = v;
        } else if (mode == 1) {
            x[i] = x[i] + d[i] * s;
            r[i] = r[i] - q[i] * s;
        } else if (mode == 2) {
            float v = s * d[i] + r[i];
            if (fixed) v = 0.0f;
            d[i] = v;
        } else {
            const float y = q[i];
            r[i] = y;              // velocity
            x[i] = x[i] + s * y;   // position
        }
    }
}

}  // namespace

float exact_element_stiffness(const float* r0, const float* r1, const float* r2, const float* r3, float E, float nu,
                              float K[144]) {
    return ke_reference_order(r0, r1, r2, r3, E, nu, K);
}

int exact_build(uint32_t N, const float* ref, uint32_t T, const uint32_t* tets, const uint32_t* perm, float E,
                float nu, hipStream_t st, ExactDevice& out) {
    out = ExactDevice{};
    out.N = N;
    out.T = T;
    std::vector<float> kb(144 * (size_t)T);
    float K[144];
    for (uint32_t e = 0; e < T; e++) {
        const uint32_t* t = tets + 4 * (size_t)e;
        ke_reference_order(ref + 3 * (size_t)t[0], ref + 3 * (size_t)t[1], ref + 3 * (size_t)t[2],
                           ref + 3 * (size_t)t[3], E, nu, K);
        for (int bi = 0; bi < 4; bi++)
            for (int bj = 0; bj < 4; bj++)
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++)
                        kb[(size_t)((4 * bi + bj) * 9 + 3 * r + c) * T + e] = K[(3 * bi + r) * 12 + 3 * bj + c];
    }
    std::vector<uint32_t> tile_of(T), ptr(N + 1, 0), code(4 * (size_t)T);
    for (uint32_t i = 0; i < T; i++) tile_of[perm[i]] = i;
    for (size_t k = 0; k < 4 * (size_t)T; k++) ptr[tets[k] + 1]++;
    for (uint32_t n = 0; n < N; n++) ptr[n + 1] += ptr[n];
    std::vector<uint32_t> fill(ptr.begin(), ptr.end() - 1);
    for (uint32_t e = 0; e < T; e++)  // ascending element order within each node's run
        for (int c = 0; c < 4; c++) code[fill[tets[4 * (size_t)e + c]]++] = 4 * e + uint32_t(c);
    auto up = [&](auto** d, const auto& h) -> int {
        const size_t bytes = std::max<size_t>(1, h.size()) * sizeof(h[0]);
        CF_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(d), bytes));
        CF_CHECK_HIP(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice, st));
        return CF_OK;
    };
    std::vector<uint32_t> tv(tets, tets + 4 * (size_t)T);
    int rc = up(&out.kb, kb);
    if (rc == CF_OK) rc = up(&out.tets, tv);
    if (rc == CF_OK) rc = up(&out.tile_of, tile_of);
    if (rc == CF_OK) rc = up(&out.ptr, ptr);
    if (rc == CF_OK) rc = up(&out.code, code);
    if (rc == CF_OK && hipMalloc(reinterpret_cast<void**>(&out.w), 12 * std::max<size_t>(1, T) * sizeof(float)) != hipSuccess)
        rc = CF_ERR_OUT_OF_MEMORY;
    if (rc == CF_OK && hipMalloc(reinterpret_cast<void**>(&out.scal), 4 * sizeof(float)) != hipSuccess)
        rc = CF_ERR_OUT_OF_MEMORY;
    if (rc == CF_OK && hipStreamSynchronize(st) != hipSuccess) rc = CF_ERR_HIP;
    if (rc != CF_OK) {
        exact_free(out);
        if (rc == CF_ERR_OUT_OF_MEMORY) set_error("reference-order mode: device allocation failed");
    }
    return rc;
}

void exact_free(ExactDevice& d) {
    for (void* p : {(void*)d.kb, (void*)d.tets, (void*)d.tile_of, (void*)d.ptr, (void*)d.code, (void*)d.w,
                    (void*)d.scal})
        if (p) (void)hipFree(p);
    d = ExactDevice{};
}

int launch_exact_elem(const ExactDevice& d, int mode, const float* R, uint32_t r_stride, const float* ref,
                      const float* in, float* out, hipStream_t s) {
    if (!d.T) return CF_OK;
    const uint32_t grid = (d.T + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_exact_elem, dim3(grid), dim3(kBlock), 0, s, mode, d.T, d.tets, d.tile_of, R, r_stride, d.kb,
                       ref, in, out ? out : d.w);
    CF_CHECK_HIP(hipGetLastError());
    return CF_OK;
}

int launch_exact_node(const ExactDevice& d, int mode, const ExactNodeTerms& t, hipStream_t s) {
    ExactNodeArgs a;
    a.ptr = d.ptr; a.code = d.code; a.w = d.w;
    a.knc = t.knc; a.knc_target = t.knc_target; a.x = t.x; a.v = t.v; a.d = t.d; a.out = t.out; a.y = t.y;
    a.mass = t.mass; a.dt = t.dt; a.inv_dt = t.inv_dt;
    a.g0 = t.g[0]; a.g1 = t.g[1]; a.g2 = t.g[2];
    a.neg_fk = t.neg_fk; a.m_fm = t.m_fm; a.warm_start = t.warm_start;
    const uint32_t grid = (d.N + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_exact_node, dim3(grid), dim3(kBlock), 0, s, mode, d.N, a);
    CF_CHECK_HIP(hipGetLastError());
    return CF_OK;
}

int exact_dot(const ExactDevice& d, int mode, uint32_t n, const float* a, const float* b, hipStream_t s, float* h_out) {
    hipLaunchKernelGGL(k_exact_dot, dim3(1), dim3(kBlock), 0, s, mode, n, a, b, d.scal);
    CF_CHECK_HIP(hipGetLastError());
    CF_CHECK_HIP(hipMemcpyAsync(h_out, d.scal, sizeof(float), hipMemcpyDeviceToHost, s));
    CF_CHECK_HIP(hipStreamSynchronize(s));
    return CF_OK;
}

int launch_exact_vec(int mode, uint32_t n3, float scalar, const uint8_t* knc, float* x, float* r, float* dvec,
                     const float* q, hipStream_t s) {
    if (!n3) return CF_OK;
    const uint32_t grid = std::min<uint32_t>((n3 + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(k_exact_vec, dim3(grid), dim3(kBlock), 0, s, mode, n3, scalar, knc, x, r, dvec, q);
    CF_CHECK_HIP(hipGetLastError());
    return CF_OK;
}

}  // namespace cf
