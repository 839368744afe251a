// corofem oracle -- CPU restatement of the reference's FP32 corotational FEM
// step. Test infrastructure only (see corofem_oracle.h). Every function below
// names the reference file:line whose arithmetic it restates; the evaluation
// order of each FP32 expression follows the reference so that, compiled
// without FMA contraction, results are reproducible bit for bit.
#include "corofem_oracle.h"

#include <cmath>
#include <cstring>
#include <cstdlib>
#include <limits>
#include <map>
#include <vector>
#include <thread>
#include <atomic>
#include <algorithm>
#include <utility>

namespace {

// ---------------------------------------------------------------- small math
// Row-major fixed-size matrices, products accumulate from 0 in k order
// (Mal/GMat.h:135-145, Mal/GMatUtils.h:13-24).
struct M3 { float a[3][3]; };
struct V3 { float v[3]; };

inline M3 mzero() { M3 m; std::memset(&m, 0, sizeof(m)); return m; }
inline float mtrace_(const M3& m) { float t = 0.0f; for (int i = 0; i < 3; i++) t += m.a[i][i]; return t; }  // Mal/GMat.h:221-227
inline M3 mid() { M3 m = mzero(); m.a[0][0] = m.a[1][1] = m.a[2][2] = 1.0f; return m; }
inline M3 mload(const float* p) { M3 m; std::memcpy(m.a, p, 9 * sizeof(float)); return m; }
inline void mstore(const M3& m, float* p) { std::memcpy(p, m.a, 9 * sizeof(float)); }
inline M3 mtr(const M3& m) { M3 t; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) t.a[i][j] = m.a[j][i]; return t; }
inline M3 mmul(const M3& x, const M3& y) {
    M3 p;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            float acc = 0.0f;
            for (int k = 0; k < 3; k++) acc += x.a[i][k] * y.a[k][j];
            p.a[i][j] = acc;
        }
    return p;
}
inline V3 mvec(const M3& m, const V3& v) {
    V3 r;
    for (int i = 0; i < 3; i++) {
        float acc = 0.0f;
        for (int j = 0; j < 3; j++) acc += m.a[i][j] * v.v[j];
        r.v[i] = acc;
    }
    return r;
}
inline V3 vadd(const V3& a, const V3& b) { V3 r; for (int i = 0; i < 3; i++) r.v[i] = a.v[i] + b.v[i]; return r; }
inline V3 vsub(const V3& a, const V3& b) { V3 r; for (int i = 0; i < 3; i++) r.v[i] = a.v[i] - b.v[i]; return r; }
inline V3 vscale(const V3& a, float s) { V3 r; for (int i = 0; i < 3; i++) r.v[i] = s * a.v[i]; return r; }
inline float vdot(const V3& a, const V3& b) { float acc = 0.0f; for (int i = 0; i < 3; i++) acc += a.v[i] * b.v[i]; return acc; }
inline V3 vcross(const V3& a, const V3& b) {  // Mal/GVec.h:89-91
    V3 r;
    r.v[0] = a.v[1] * b.v[2] - a.v[2] * b.v[1];
    r.v[1] = a.v[2] * b.v[0] - a.v[0] * b.v[2];
    r.v[2] = a.v[0] * b.v[1] - a.v[1] * b.v[0];
    return r;
}
inline float vnorm(const V3& a) { return std::sqrt(vdot(a, a)); }
// v / s == v * (1/s)  (Mal/GVec.h:85)
inline V3 vnormalized(const V3& a) { float inv = 1.0f / vnorm(a); return vscale(a, inv); }
inline V3 vload(const float* p) { V3 r; r.v[0] = p[0]; r.v[1] = p[1]; r.v[2] = p[2]; return r; }
inline M3 from_cols(const V3& c0, const V3& c1, const V3& c2) {
    M3 m;
    for (int i = 0; i < 3; i++) { m.a[i][0] = c0.v[i]; m.a[i][1] = c1.v[i]; m.a[i][2] = c2.v[i]; }
    return m;
}
inline V3 col(const M3& m, int j) { V3 r; for (int i = 0; i < 3; i++) r.v[i] = m.a[i][j]; return r; }
inline void setcol(M3& m, int j, const V3& c) { for (int i = 0; i < 3; i++) m.a[i][j] = c.v[i]; }

// Mal/GMat.h:236-244
inline float det3(const M3& m) {
    return m.a[0][0] * m.a[1][1] * m.a[2][2]
         - m.a[0][0] * m.a[2][1] * m.a[1][2]
         + m.a[1][0] * m.a[2][1] * m.a[0][2]
         - m.a[1][0] * m.a[0][1] * m.a[2][2]
         + m.a[2][0] * m.a[0][1] * m.a[1][2]
         - m.a[2][0] * m.a[1][1] * m.a[0][2];
}
// Mal/GMat.h:291-313
inline M3 inv3(const M3& m, float det) {
    float id = 1.0f / det;
    M3 r;
    r.a[0][0] = id * (m.a[1][1] * m.a[2][2] - m.a[1][2] * m.a[2][1]);
    r.a[0][1] = id * (m.a[0][2] * m.a[2][1] - m.a[0][1] * m.a[2][2]);
    r.a[0][2] = id * (m.a[0][1] * m.a[1][2] - m.a[0][2] * m.a[1][1]);
    r.a[1][0] = id * (m.a[1][2] * m.a[2][0] - m.a[1][0] * m.a[2][2]);
    r.a[1][1] = id * (m.a[0][0] * m.a[2][2] - m.a[0][2] * m.a[2][0]);
    r.a[1][2] = id * (m.a[0][2] * m.a[1][0] - m.a[0][0] * m.a[1][2]);
    r.a[2][0] = id * (m.a[1][0] * m.a[2][1] - m.a[1][1] * m.a[2][0]);
    r.a[2][1] = id * (m.a[0][1] * m.a[2][0] - m.a[0][0] * m.a[2][1]);
    r.a[2][2] = id * (m.a[0][0] * m.a[1][1] - m.a[0][1] * m.a[1][0]);
    return r;
}
inline M3 inv3(const M3& m) { return inv3(m, det3(m)); }

// Newton polar decomposition, Mal/GMatDecomposition.h:411-432:
// H <- 0.5*(H + inv(H^T)), error = |H1-H0|_F^2 (sum in storage order),
// stop when error <= 0.01 or after 10 iterations.
inline M3 polar_newton(const M3& H, uint32_t* iters) {
    M3 H0 = H;
    float err = 0.0f;
    unsigned n = 0;
    do {
        M3 it = inv3(mtr(H0));
        M3 H1;
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) H1.a[i][j] = 0.5f * (H0.a[i][j] + it.a[i][j]);
        err = 0.0f;
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) { float d = H1.a[i][j] - H0.a[i][j]; err += d * d; }
        H0 = H1;
        n++;
    } while (n < 10u && err > 0.01f);
    if (iters) *iters = n;
    return H0;
}

// ---- PhysBAM-style 3x3 SVD, FP32, as adapted in Mal/GMatDecomposition_SVD3x3_PhysBAM.h
struct Sym3 { float x11, x21, x31, x22, x32, x33; };
struct Sym2 { float x11, x21, x22; };

inline V3 orthogonal_vector(const V3& v) {  // :35-49
    float ax = std::fabs(v.v[0]), ay = std::fabs(v.v[1]), az = std::fabs(v.v[2]);
    V3 r;
    if (ax < ay) {
        if (ax < az) { r.v[0] = 0; r.v[1] = v.v[2]; r.v[2] = -v.v[1]; }
        else { r.v[0] = v.v[1]; r.v[1] = -v.v[0]; r.v[2] = 0; }
    } else {
        if (ay < az) { r.v[0] = -v.v[2]; r.v[1] = 0; r.v[2] = v.v[0]; }
        else { r.v[0] = v.v[1]; r.v[1] = -v.v[0]; r.v[2] = 0; }
    }
    return r;
}
inline Sym3 normal_equations(const M3& F) {  // :129-137
    Sym3 s;
    s.x11 = F.a[0][0] * F.a[0][0] + F.a[1][0] * F.a[1][0] + F.a[2][0] * F.a[2][0];
    s.x21 = F.a[0][1] * F.a[0][0] + F.a[1][1] * F.a[1][0] + F.a[2][1] * F.a[2][0];
    s.x31 = F.a[0][2] * F.a[0][0] + F.a[1][2] * F.a[1][0] + F.a[2][2] * F.a[2][0];
    s.x22 = F.a[0][1] * F.a[0][1] + F.a[1][1] * F.a[1][1] + F.a[2][1] * F.a[2][1];
    s.x32 = F.a[0][2] * F.a[0][1] + F.a[1][2] * F.a[1][1] + F.a[2][2] * F.a[2][1];
    s.x33 = F.a[0][2] * F.a[0][2] + F.a[1][2] * F.a[1][2] + F.a[2][2] * F.a[2][2];
    return s;
}
inline Sym3 cofactor(const Sym3& s) {  // :139-147
    Sym3 c;
    c.x11 = s.x22 * s.x33 - s.x32 * s.x32;
    c.x21 = s.x32 * s.x31 - s.x21 * s.x33;
    c.x31 = s.x21 * s.x32 - s.x22 * s.x31;
    c.x22 = s.x11 * s.x33 - s.x31 * s.x31;
    c.x32 = s.x21 * s.x31 - s.x11 * s.x32;
    c.x33 = s.x11 * s.x22 - s.x21 * s.x21;
    return c;
}
inline V3 largest_column_normalized(const Sym3& s) {  // :149-156
    float s11 = s.x11 * s.x11, s12 = s.x21 * s.x21, s13 = s.x31 * s.x31;
    float s22 = s.x22 * s.x22, s23 = s.x32 * s.x32, s33 = s.x33 * s.x33;
    float c1 = s11 + s12 + s13, c2 = s12 + s22 + s23, c3 = s13 + s23 + s33;
    V3 r;
    if (c1 > c2) {
        if (c1 > c3) { float inv = 1.0f / std::sqrt(c1); r.v[0] = s.x11 * inv; r.v[1] = s.x21 * inv; r.v[2] = s.x31 * inv; return r; }
    } else if (c2 > c3) { float inv = 1.0f / std::sqrt(c2); r.v[0] = s.x21 * inv; r.v[1] = s.x22 * inv; r.v[2] = s.x32 * inv; return r; }
    if (c3 > 0) { float inv = 1.0f / std::sqrt(c3); r.v[0] = s.x31 * inv; r.v[1] = s.x32 * inv; r.v[2] = s.x33 * inv; return r; }
    r.v[0] = 1; r.v[1] = 0; r.v[2] = 0;
    return r;
}
inline void largest_column_normalized2(const Sym2& s, float out[2]) {  // :75-82
    float s11 = s.x11 * s.x11, s12 = s.x21 * s.x21, s22 = s.x22 * s.x22;
    float c1 = s11 + s12, c2 = s12 + s22;
    if (c1 > c2) { float inv = 1.0f / std::sqrt(c1); out[0] = s.x11 * inv; out[1] = s.x21 * inv; }
    else if (c2 > 0) { float inv = 1.0f / std::sqrt(c2); out[0] = s.x21 * inv; out[1] = s.x22 * inv; }
    else { out[0] = 1; out[1] = 0; }
}
// Smith (1961) trigonometric eigenvalues, sorted decreasingly (:161-197)
// The reference writes Sqrt(3) with an int argument (:165), which deduces
// mal::Sqrt<int> (Mal/RealUtils.h:26-27) and truncates sqrt(3) to 1.
const float kRefRootThree = float(int(std::sqrt(3.0f)));
inline V3 fast_eigenvalues(const Sym3& s, float root_three) {
    const float one_third = 1.0f / 3.0f;
    const float one_sixth = 1.0f / 6.0f;
    float m = one_third * (s.x11 + s.x22 + s.x33);
    float a11 = s.x11 - m, a22 = s.x22 - m, a33 = s.x33 - m;
    float a12s = s.x21 * s.x21, a13s = s.x31 * s.x31, a23s = s.x32 * s.x32;
    float p = one_sixth * (a11 * a11 + a22 * a22 + a33 * a33 + 2 * (a12s + a13s + a23s));
    float q = 0.5f * (a11 * (a22 * a33 - a23s) - a22 * a13s - a33 * a12s) + s.x21 * s.x31 * s.x32;
    float sqrt_p = std::sqrt(p), disc = p * p * p - q * q;
    float phi = one_third * std::atan2(std::sqrt(std::max(0.0f, disc)), q);
    float c = std::cos(phi), sn = std::sin(phi);
    float spc = sqrt_p * c, r3sps = root_three * sqrt_p * sn;
    float l0 = m + 2 * spc, l1 = m - spc - r3sps, l2 = m - spc + r3sps;
    V3 r;
    if (l0 < l1) {
        if (l1 < l2) { r.v[0] = l2; r.v[1] = l1; r.v[2] = l0; }
        else if (l0 < l2) { r.v[0] = l1; r.v[1] = l2; r.v[2] = l0; }
        else { r.v[0] = l1; r.v[1] = l0; r.v[2] = l2; }
    } else {
        if (l1 > l2) { r.v[0] = l0; r.v[1] = l1; r.v[2] = l2; }
        else if (l0 > l2) { r.v[0] = l0; r.v[1] = l2; r.v[2] = l1; }
        else { r.v[0] = l2; r.v[1] = l0; r.v[2] = l1; }
    }
    return r;
}
// 3x2 matrices as two columns
struct M32 { V3 c0, c1; };
inline V3 m32_vec(const M32& m, const float v[2]) {  // GMat<3,2> * GVec<2>
    V3 r;
    for (int i = 0; i < 3; i++) { float acc = 0.0f; acc += m.c0.v[i] * v[0]; acc += m.c1.v[i] * v[1]; r.v[i] = acc; }
    return r;
}
inline Sym2 conjugate_with_transpose(const M32& A, const Sym3& B) {  // :100-107
    M3 fb;
    fb.a[0][0] = B.x11; fb.a[0][1] = B.x21; fb.a[0][2] = B.x31;
    fb.a[1][0] = B.x21; fb.a[1][1] = B.x22; fb.a[1][2] = B.x32;
    fb.a[2][0] = B.x31; fb.a[2][1] = B.x32; fb.a[2][2] = B.x33;
    // full_B * A (3x3 * 3x2), accumulate from 0
    M32 BA;
    BA.c0 = mvec(fb, A.c0);
    BA.c1 = mvec(fb, A.c1);
    // Transpose_Times_With_Symmetric_Result(BA, A)  (:86-96)
    Sym2 r;
    r.x11 = BA.c0.v[0] * A.c0.v[0] + BA.c0.v[1] * A.c0.v[1] + BA.c0.v[2] * A.c0.v[2];
    r.x21 = BA.c1.v[0] * A.c0.v[0] + BA.c1.v[1] * A.c0.v[1] + BA.c1.v[2] * A.c0.v[2];
    r.x22 = BA.c1.v[0] * A.c1.v[0] + BA.c1.v[1] * A.c1.v[1] + BA.c1.v[2] * A.c1.v[2];
    return r;
}
inline M3 fast_eigenvectors(const Sym3& A, const V3& lambda) {  // :199-236
    bool flipped = false;
    V3 lf = lambda;
    if (lambda.v[0] - lambda.v[1] < lambda.v[1] - lambda.v[2]) {
        lf.v[0] = lambda.v[2];
        lf.v[2] = lambda.v[0];
        flipped = true;
    }
    Sym3 am = A;
    am.x11 -= lf.v[0]; am.x22 -= lf.v[0]; am.x33 -= lf.v[0];
    V3 v1 = largest_column_normalized(cofactor(am));
    V3 v1o = vnormalized(orthogonal_vector(v1));
    M32 other;
    other.c0 = v1o;
    other.c1 = vcross(v1, v1o);
    Sym2 ar = conjugate_with_transpose(other, A);
    ar.x11 -= lf.v[2]; ar.x22 -= lf.v[2];
    Sym2 arc; arc.x11 = ar.x22; arc.x21 = -ar.x21; arc.x22 = ar.x11;  // GSymmetricMat2x2::Cofactor_Matrix
    float w[2];
    largest_column_normalized2(arc, w);
    V3 v3 = m32_vec(other, w);
    V3 v2 = vcross(v3, v1);
    if (flipped) return from_cols(v3, v2, vscale(v1, -1.0f));
    return from_cols(v1, v2, v3);
}
// GFastSVD3x3 (:246-280): F = U diag(s) V^T
inline void fast_svd3(const M3& F, M3& U, V3& s, M3& Vt, float root_three = kRefRootThree) {
    Sym3 ne = normal_equations(F);
    V3 lambda = fast_eigenvalues(ne, root_three);
    M3 V = fast_eigenvectors(ne, lambda);
    Vt = mtr(V);
    if (lambda.v[2] < 0) {
        lambda.v[0] = std::max(lambda.v[0], 0.0f);
        lambda.v[1] = std::max(lambda.v[1], 0.0f);
        lambda.v[2] = std::max(lambda.v[2], 0.0f);
    }
    s.v[0] = std::sqrt(lambda.v[0]);
    s.v[1] = std::sqrt(lambda.v[1]);
    s.v[2] = std::sqrt(lambda.v[2]);
    if (det3(F) < 0) s.v[2] = -s.v[2];
    setcol(U, 0, vnormalized(mvec(F, col(V, 0))));
    V3 v1o = vnormalized(orthogonal_vector(col(U, 0)));
    M32 other;
    other.c0 = v1o;
    other.c1 = vcross(col(U, 0), v1o);
    // other_v * ( other_v^T * Normalized(F * V.col1) )
    V3 fv1 = vnormalized(mvec(F, col(V, 1)));
    float t[2];
    { float acc = 0.0f; for (int i = 0; i < 3; i++) acc += other.c0.v[i] * fv1.v[i]; t[0] = acc; }
    { float acc = 0.0f; for (int i = 0; i < 3; i++) acc += other.c1.v[i] * fv1.v[i]; t[1] = acc; }
    setcol(U, 1, m32_vec(other, t));
    setcol(U, 2, vcross(col(U, 0), col(U, 1)));
}

// ---- element precompute (TetrahedronElement3.h:85-173) ----
inline M3 compute_D(const V3& r0, const V3& r1, const V3& r2, const V3& r3) {
    return from_cols(vsub(r1, r0), vsub(r2, r0), vsub(r3, r0));
}
// ---- degenerate elements: direction of collapse and the DAPD rotation ----
// DoC (ms/fem/TetrahedronElement3.h:28-38) is kept as one byte, 4*vit0 + vit1:
// vertex-face when vit0 == vit1, edge-edge when vit0 < vit1, invalid (3,0) = 12.
constexpr uint8_t kDocInvalid = 12;
inline bool doc_valid(uint8_t c) { return (c >> 2) <= (c & 3); }

// gsl_poly_solve_cubic (GNU Scientific Library, poly/solve_cubic.c; the
// reference links the system GSL and pins no version, the routine is the same
// in GSL 1.x and 2.x): x^3 + a x^2 + b x + c = 0 in double. One real root by
// Cardano's formula, three by Viete's trigonometric form, sorted ascending;
// the exact double root and the triple root are caught by the discriminant
// tests. Returns the number of roots written.
int gsl_cubic(double a, double b, double c, double* x0, double* x1, double* x2) {
    const double q = a * a - 3 * b;
    const double r = 2 * a * a * a - 9 * a * b + 27 * c;
    const double Q = q / 9, R = r / 54;
    const double Q3 = Q * Q * Q, R2 = R * R;
    const double CR2 = 729 * r * r, CQ3 = 2916 * q * q * q;
    if (R == 0 && Q == 0) {
        *x0 = *x1 = *x2 = -a / 3;
        return 3;
    }
    if (CR2 == CQ3) {  // R^2 == Q^3 in integer-exact form: a double root
        const double sq = std::sqrt(Q);
        if (R > 0) { *x0 = -2 * sq - a / 3; *x1 = sq - a / 3; *x2 = sq - a / 3; }
        else { *x0 = -sq - a / 3; *x1 = -sq - a / 3; *x2 = 2 * sq - a / 3; }
        return 3;
    }
    if (R2 < Q3) {
        const double sgn = (R >= 0) ? 1.0 : -1.0;
        const double theta = std::acos(sgn * std::sqrt(R2 / Q3));
        const double norm = -2 * std::sqrt(Q);
        double r0 = norm * std::cos(theta / 3) - a / 3;
        double r1 = norm * std::cos((theta + 2.0 * M_PI) / 3) - a / 3;
        double r2 = norm * std::cos((theta - 2.0 * M_PI) / 3) - a / 3;
        if (r0 > r1) std::swap(r0, r1);
        if (r1 > r2) { std::swap(r1, r2); if (r0 > r1) std::swap(r0, r1); }
        *x0 = r0; *x1 = r1; *x2 = r2;
        return 3;
    }
    const double sgn = (R >= 0) ? 1.0 : -1.0;
    const double A = -sgn * std::pow(std::fabs(R) + std::sqrt(R2 - Q3), 1.0 / 3.0);
    *x0 = A + Q / A - a / 3;
    return 1;
}

// mal::GSolvePolynomialEq1/2 (Mal/GSolvePolynomialEq.h:11-38) in float, with
// IsZero's absolute epsilon 1e-4 (Mal/RealUtils.h:110-114).
inline int solve_eq1(float a, float b, float& x) {
    if (std::fabs(a) <= 0.0001f) return 0;
    x = -b / a;
    return 1;
}
inline int solve_eq2(float a, float b, float c, float& x1, float& x2) {
    if (std::fabs(a) <= 0.0001f) {
        const int n = solve_eq1(b, c, x1);
        x2 = x1;
        return n;
    }
    const float disc = b * b - 4 * a * c;
    if (disc < 0) return 0;
    const float root = std::sqrt(disc);
    const float inv = 1.0f / (2.0f * a);
    x1 = (-b - root) * inv;
    x2 = (-b + root) * inv;
    if (x1 > x2) std::swap(x1, x2);
    return (x1 == x2) ? 1 : 2;
}
// mal::GSolvePolynomialEq3 (ms/fem/TetrahedronElementSet3.cpp:50-69):
// coefficients normalised in float, GSL's cubic in double, roots to float.
inline int solve_eq3(float a, float b, float c, float d, float x[3]) {
    if (a != 0.0f) {
        double r0 = 0, r1 = 0, r2 = 0;
        const int n = gsl_cubic(double(b / a), double(c / a), double(d / a), &r0, &r1, &r2);
        x[0] = float(r0); x[1] = float(r1); x[2] = float(r2);
        return n;
    }
    x[2] = 0.0f;
    return solve_eq2(b, c, d, x[0], x[1]);
}

// det Ds(t) of the linear path Ds(t) = Da + t Dd is the cubic
// A t^3 + B t^2 + C t + D'. Its coefficients are sums over the six permutation
// terms u_x v_y w_z of the determinant, (u, v, w) running over the edge
// columns 0, 1, 2; a term takes k of its factors from Da and 3-k from Dd.
// The sums run in the reference's order (:607-636): terms + + + - - -, and in
// B and C the Da factor (resp. the Dd factor) moving from x to y to z.
template <class Term>
inline float perm6(Term f) {
    return f(0, 1, 2) + f(1, 2, 0) + f(2, 0, 1) - f(2, 1, 0) - f(1, 0, 2) - f(0, 2, 1);
}

// ComputeDoC (ms/fem/TetrahedronElementSet3.cpp:589-743): the time of
// collapse toc in (0,1] at which det Ds on the path a -> b reaches
// thr * 6 V, the collapse "normal" N (row 2 of V^T in Ds(toc) = U S V^T: the
// right singular vector of the smallest singular value, as the reference
// takes it, not the left one normal to the flattened element), then the
// crossing: vertex-face when one face's
// N-projected area is half the sum of all four (absolute epsilon 1e-4), else
// the first edge pair (0j, kl) whose triangle normals around edge 0j disagree.
uint8_t compute_doc(const V3 a[4], const V3 b[4], float volume, float thr, float* toc_out = nullptr) {
    const V3 d0 = vsub(b[0], a[0]), d1 = vsub(b[1], a[1]), d2 = vsub(b[2], a[2]), d3 = vsub(b[3], a[3]);
    const V3 dd[3] = {vsub(d1, d0), vsub(d2, d0), vsub(d3, d0)};
    const V3 da[3] = {vsub(a[1], a[0]), vsub(a[2], a[0]), vsub(a[3], a[0])};
    const float K = thr * 6.0f * volume;
    const float A = perm6([&](int p, int q, int r) { return dd[p].v[0] * dd[q].v[1] * dd[r].v[2]; });
    const float B = perm6([&](int p, int q, int r) {
        return da[p].v[0] * dd[q].v[1] * dd[r].v[2] + dd[p].v[0] * da[q].v[1] * dd[r].v[2] +
               dd[p].v[0] * dd[q].v[1] * da[r].v[2];
    });
    const float C = perm6([&](int p, int q, int r) {
        return da[p].v[0] * da[q].v[1] * dd[r].v[2] + da[p].v[0] * dd[q].v[1] * da[r].v[2] +
               dd[p].v[0] * da[q].v[1] * da[r].v[2];
    });
    const float D = perm6([&](int p, int q, int r) { return da[p].v[0] * da[q].v[1] * da[r].v[2]; }) - K;
    float roots[3] = {0.0f, 0.0f, 0.0f};
    const int n = solve_eq3(A, B, C, D, roots);
    if (n <= 0) return kDocInvalid;
    float toc = 2.0f;
    int n01 = 0;
    for (int i = 0; i < n; i++)
        if (roots[i] > 0.0f && roots[i] <= 1.0f && roots[i] < toc) { toc = roots[i]; n01++; }
    if (n01 != 1 && n01 != 3) return kDocInvalid;
    if (toc_out) *toc_out = toc;
    V3 q[4];
    q[0] = vadd(a[0], vscale(d0, toc)); q[1] = vadd(a[1], vscale(d1, toc));
    q[2] = vadd(a[2], vscale(d2, toc)); q[3] = vadd(a[3], vscale(d3, toc));
    const V3 q10 = vsub(q[1], q[0]), q20 = vsub(q[2], q[0]), q30 = vsub(q[3], q[0]);
    M3 U, Vt;
    V3 sv;
    fast_svd3(from_cols(q10, q20, q30), U, sv, Vt);
    V3 N;
    for (int i = 0; i < 3; i++) N.v[i] = Vt.a[2][i];
    float area[4];
    area[0] = std::fabs(vdot(N, vcross(vsub(q[2], q[1]), vsub(q[3], q[1]))));
    area[1] = std::fabs(vdot(N, vcross(q20, q30)));
    area[2] = std::fabs(vdot(N, vcross(q10, q30)));
    area[3] = std::fabs(vdot(N, vcross(q10, q20)));
    const float sum = area[0] + area[1] + area[2] + area[3];
    uint8_t doc = kDocInvalid;
    for (int i = 0; i < 4; i++)
        if (std::fabs(area[i] - 0.5f * sum) <= 0.0001f) doc = uint8_t(5 * i);
    if (doc_valid(doc)) return doc;
    if (vdot(vcross(q10, q20), vcross(q10, q30)) < 0.0f) return 1;  // (0,1)
    if (vdot(vcross(q20, q10), vcross(q20, q30)) < 0.0f) return 2;  // (0,2)
    if (vdot(vcross(q30, q10), vcross(q30, q20)) < 0.0f) return 3;  // (0,3)
    return 0;                                                        // forced vertex-face 0
}

// Compute_R_DAPD_Degenerate (ms/fem/Rotation.cpp:364-545): move the collapsing
// vertex (vertex-face) or edge (edge-edge) along the unit collapse axis past the
// collapse by lambda_u (cubic Hermite for 0 < w <= dh, linear plus optional
// power term for w <= 0), then R = Newton polar factor of the projected F.
// Axis shorter than 1e-3 (squared length 1e-6): R = Id.
M3 r_dapd_degenerate(float detF, float thr, uint8_t doc, const V3 p[4], float volume, const M3& invDm,
                     float fL, float fNL, float eNL, bool* ok) {
    V3 x[4] = {p[0], p[1], p[2], p[3]};
    const int v0 = doc >> 2, v1 = doc & 3;
    int i0, i1, i2, i3;  // moved vertex/vertices, then the ones defining the axis
    V3 axis;
    if (v0 == v1) {  // vertex-face: vertex v0 against the face of the next three
        i0 = v0; i1 = (v0 + 1) % 4; i2 = (v0 + 2) % 4; i3 = (v0 + 3) % 4;
        axis = vcross(vsub(x[i2], x[i1]), vsub(x[i3], x[i1]));
    } else {         // edge-edge: edge (0, v1) against the opposite edge
        i0 = v0; i1 = v1;
        i2 = (v1 == 1) ? 2 : 1;
        i3 = (v1 == 3) ? 2 : 3;
        axis = vcross(vsub(x[i1], x[i0]), vsub(x[i3], x[i2]));
    }
    const float len_sq = vdot(axis, axis);
    if (!(len_sq > 0.000001f)) { *ok = false; return mid(); }
    const float len = std::sqrt(len_sq);
    axis = vscale(axis, 1.0f / len);
    const float dh = thr * 6.0f * volume / len;
    // the moved vertex's height over the axis' plane
    float w = (v0 == v1) ? vdot(vsub(x[i0], x[i1]), axis) : vdot(vsub(x[i0], x[i2]), axis);
    if (w * detF < 0.0f) { w = -w; axis = vscale(axis, -1.0f); }
    const float s = w / dh;
    const float t = 1.0f - s;
    const float lambda = (w > 0.0f) ? fL * dh * ((2.0f - t) * t * t) : fL * dh * (t + fNL * std::pow(-s, eNL));
    x[i0] = vadd(x[i0], vscale(axis, lambda));
    if (v0 != v1) x[i1] = vadd(x[i1], vscale(axis, lambda));
    const M3 projF = mmul(compute_D(x[0], x[1], x[2], x[3]), invDm);
    *ok = true;
    return polar_newton(projF, nullptr);
}

float compute_K(const V3& r0, const V3& r1, const V3& r2, const V3& r3, float E, float nu, float K[144]) {
    float volume = det3(compute_D(r0, r1, r2, r3)) / 6.0f;
    float x1 = r0.v[0], x2 = r1.v[0], x3 = r2.v[0], x4 = r3.v[0];
    float y1 = r0.v[1], y2 = r1.v[1], y3 = r2.v[1], y4 = r3.v[1];
    float z1 = r0.v[2], z2 = r1.v[2], z3 = r2.v[2], z4 = r3.v[2];
    float J[4][4];
    // inverse Jacobian, AFEM-16 p.7 (TetrahedronElement3.h:116-123); column 0 unused by B
    J[0][0] = x2 * (y3 * z4 - y4 * z3) + x3 * (y4 * z2 - y2 * z4) + x4 * (y2 * z3 - y3 * z2);
    J[1][0] = x1 * (y4 * z3 - y3 * z4) + x3 * (y1 * z4 - y4 * z1) + x4 * (y3 * z1 - y1 * z3);
    J[2][0] = x1 * (y2 * z4 - y4 * z2) + x2 * (y4 * z1 - y1 * z4) + x4 * (y1 * z2 - y2 * z1);
    J[3][0] = x1 * (y3 * z2 - y2 * z3) + x2 * (y1 * z3 - y3 * z1) + x3 * (y2 * z1 - y1 * z2);
    J[0][1] = (y4 - y2) * (z3 - z2) - (y3 - y2) * (z4 - z2); J[0][2] = (x3 - x2) * (z4 - z2) - (x4 - x2) * (z3 - z2); J[0][3] = (x4 - x2) * (y3 - y2) - (x3 - x2) * (y4 - y2);
    J[1][1] = (y3 - y1) * (z4 - z3) - (y3 - y4) * (z1 - z3); J[1][2] = (x4 - x3) * (z3 - z1) - (x1 - x3) * (z3 - z4); J[1][3] = (x3 - x1) * (y4 - y3) - (x3 - x4) * (y1 - y3);
    J[2][1] = (y2 - y4) * (z1 - z4) - (y1 - y4) * (z2 - z4); J[2][2] = (x1 - x4) * (z2 - z4) - (x2 - x4) * (z1 - z4); J[2][3] = (x2 - x4) * (y1 - y4) - (x1 - x4) * (y2 - y4);
    J[3][1] = (y1 - y3) * (z2 - z1) - (y1 - y2) * (z3 - z1); J[3][2] = (x2 - x1) * (z1 - z3) - (x3 - x1) * (z1 - z2); J[3][3] = (x1 - x3) * (y2 - y1) - (x1 - x2) * (y3 - y1);
    float s = 1.0f / (6.0f * volume);
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) J[i][j] = J[i][j] * s;
    float a[4], b[4], c[4];
    for (int n = 0; n < 4; n++) { a[n] = J[n][1]; b[n] = J[n][2]; c[n] = J[n][3]; }
    float B[6][12];
    std::memset(B, 0, sizeof(B));
    for (int n = 0; n < 4; n++) {  // AFEM-16 p.13 (TetrahedronElement3.h:132-137)
        B[0][3 * n + 0] = a[n];
        B[1][3 * n + 1] = b[n];
        B[2][3 * n + 2] = c[n];
        B[3][3 * n + 0] = b[n]; B[3][3 * n + 1] = a[n];
        B[4][3 * n + 1] = c[n]; B[4][3 * n + 2] = b[n];
        B[5][3 * n + 0] = c[n]; B[5][3 * n + 2] = a[n];
    }
    // E (TetrahedronElement3.h:141-156)
    float Em[6][6];
    std::memset(Em, 0, sizeof(Em));
    float aa = E / ((1.0f + nu) * (1.0f - 2.0f * nu));
    float a_nu = aa * nu, a_k1 = aa * (1.0f - nu), a_k2 = aa * (0.5f - nu);
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

// 2D (TriangleElement2.h:71-139)
float compute_K2(const float r0[2], const float r1[2], const float r2[2], float E, float nu, float K[36]) {
    float area = 0.5f * ((r1[0] - r0[0]) * (r2[1] - r0[1]) - (r2[0] - r0[0]) * (r1[1] - r0[1]));
    float rcp = 1.0f / (2.0f * area);
    float x02 = r0[0] - r2[0]; x02 *= rcp;
    float x10 = r1[0] - r0[0]; x10 *= rcp;
    float x21 = r2[0] - r1[0]; x21 *= rcp;
    float y01 = r0[1] - r1[1]; y01 *= rcp;
    float y12 = r1[1] - r2[1]; y12 *= rcp;
    float y20 = r2[1] - r0[1]; y20 *= rcp;
    float B[3][6] = {{y12, 0, y20, 0, y01, 0}, {0, x21, 0, x02, 0, x10}, {x21, y12, x02, y20, x10, y01}};
    float a = E / ((1 + nu) * (1 - 2 * nu));
    float anu = a * nu;
    float Em[3][3] = {{a - anu, anu, 0}, {anu, a - anu, 0}, {0, 0, 0.5f * a - anu}};
    float EB[3][6];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 6; j++) { float acc = 0.0f; for (int k = 0; k < 3; k++) acc += Em[i][k] * B[k][j]; EB[i][j] = acc; }
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) { float acc = 0.0f; for (int k = 0; k < 3; k++) acc += B[k][i] * EB[k][j]; K[i * 6 + j] = acc * area; }
    return area;
}

// ---- SSE real_array reductions (ns/real_array.h:34-49): each 4-lane chunk
// is reduced by _mm_dp_ps as (p0+p1)+(p2+p3) and added serially to lane 0 of
// the accumulator; lanes past n are taken as zero.
inline float dp4(const float* a, const float* b, uint32_t i, uint32_t n) {
    float p[4];
    for (int l = 0; l < 4; l++) p[l] = (i + l < n) ? a[i + l] * b[i + l] : 0.0f;
    return (p[0] + p[1]) + (p[2] + p[3]);
}
float sse_dot(const float* a, const float* b, uint32_t n) {
    float acc = 0.0f;
    uint32_t an = 4 * ((n + 3) / 4);
    for (uint32_t i = 0; i < an; i += 4) acc = acc + dp4(a, b, i, n);
    return acc;
}

// Same products accumulated in FP64 in index order, rounded once to FP32: the
// reduction precision the GPU path uses (not the reference's; used to measure
// how much of a GPU/oracle difference the reduction precision alone explains).
float fp64_dot(const float* a, const float* b, uint32_t n) {
    double acc = 0.0;
    for (uint32_t i = 0; i < n; i++) acc += double(a[i]) * double(b[i]);
    return float(acc);
}

}  // namespace

// ================================================================ C API
extern "C" {

void orc_box_counts(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t* n_nodes, uint32_t* n_tets) {
    *n_nodes = nx * ny * nz;
    *n_tets = 5u * (nx - 1) * (ny - 1) * (nz - 1);
}

// Geo/shape/TetSolidShape3.cpp:921-999 (counts are vertices per axis)
void orc_mesh_box(float hx, float hy, float hz, uint32_t nx, uint32_t ny, uint32_t nz, float* pos, uint32_t* tets) {
    float h[3] = {hx, hy, hz};
    float sz[3] = {2 * hx, 2 * hy, 2 * hz};
    uint32_t n[3] = {nx, ny, nz};
    uint32_t v = 0;
    for (uint32_t i = 0; i < nx; i++)
        for (uint32_t j = 0; j < ny; j++)
            for (uint32_t k = 0; k < nz; k++) {
                uint32_t ijk[3] = {i, j, k};
                for (int c = 0; c < 3; c++) pos[3 * v + c] = (-h[c]) + (sz[c] * float(ijk[c])) / float(n[c] - 1);
                v++;
            }
    uint32_t t = 0;
    auto add = [&](uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
        tets[4 * t + 0] = a; tets[4 * t + 1] = b; tets[4 * t + 2] = c; tets[4 * t + 3] = d; t++;
    };
    for (uint32_t i = 0; i + 1 < nx; i++)
        for (uint32_t j = 0; j + 1 < ny; j++)
            for (uint32_t k = 0; k + 1 < nz; k++) {
                uint32_t v0 = i * ny * nz + j * nz + k;
                uint32_t v1 = i * ny * nz + (j + 1) * nz + k;
                uint32_t v2 = (i + 1) * ny * nz + (j + 1) * nz + k;
                uint32_t v3 = (i + 1) * ny * nz + j * nz + k;
                uint32_t v4 = i * ny * nz + j * nz + (k + 1);
                uint32_t v5 = i * ny * nz + (j + 1) * nz + (k + 1);
                uint32_t v6 = (i + 1) * ny * nz + (j + 1) * nz + (k + 1);
                uint32_t v7 = (i + 1) * ny * nz + j * nz + (k + 1);
                int mirrors = 0;
                if (i % 2 == 1) { mirrors++; std::swap(v0, v3); std::swap(v1, v2); std::swap(v5, v6); std::swap(v4, v7); }
                if (j % 2 == 1) { mirrors++; std::swap(v0, v1); std::swap(v3, v2); std::swap(v4, v5); std::swap(v7, v6); }
                if (k % 2 == 1) { mirrors++; std::swap(v0, v4); std::swap(v1, v5); std::swap(v2, v6); std::swap(v3, v7); }
                if (mirrors % 2 == 0) {
                    add(v0, v1, v4, v3); add(v2, v1, v3, v6); add(v5, v1, v6, v4); add(v6, v1, v3, v4); add(v7, v3, v4, v6);
                } else {
                    add(v0, v1, v3, v4); add(v2, v1, v6, v3); add(v5, v1, v4, v6); add(v6, v1, v4, v3); add(v7, v3, v6, v4);
                }
            }
}

void orc_box2_counts(uint32_t nx, uint32_t ny, uint32_t* n_nodes, uint32_t* n_tris) {
    *n_nodes = nx * ny;
    *n_tris = 2u * (nx - 1) * (ny - 1);
}

// Geo/shape/MeshSolidShape2.cpp:1651-1675
void orc_mesh_box2(float hx, float hy, uint32_t nx, uint32_t ny, float* pos, uint32_t* tris) {
    float step_x = 2 * hx / float(nx - 1);
    float step_y = 2 * hy / float(ny - 1);
    uint32_t v = 0;
    for (uint32_t y = 0; y < ny; y++)
        for (uint32_t x = 0; x < nx; x++) {
            pos[2 * v + 0] = (-hx) + float(x) * step_x;
            pos[2 * v + 1] = (-hy) + float(y) * step_y;
            v++;
        }
    uint32_t t = 0;
    for (uint32_t y = 0; y + 1 < ny; y++)
        for (uint32_t x = 0; x + 1 < nx; x++) {
            uint32_t v0 = y * nx + x, v1 = y * nx + x + 1, v2 = (y + 1) * nx + x + 1, v3 = (y + 1) * nx + x;
            tris[3 * t + 0] = v0; tris[3 * t + 1] = v1; tris[3 * t + 2] = v2; t++;
            tris[3 * t + 0] = v2; tris[3 * t + 1] = v3; tris[3 * t + 2] = v0; t++;
        }
}

float orc_det3(const float m[9]) { return det3(mload(m)); }
void orc_inverse3(const float m[9], float out[9]) { mstore(inv3(mload(m)), out); }
void orc_pd3(const float F[9], float detF, float R[9], uint32_t* iters) {
    (void)detF;
    mstore(polar_newton(mload(F), iters), R);
}
void orc_svd3(const float F[9], float U[9], float s[3], float Vt[9]) {
    M3 u, vt; V3 sv;
    fast_svd3(mload(F), u, sv, vt);
    mstore(u, U); mstore(vt, Vt);
    s[0] = sv.v[0]; s[1] = sv.v[1]; s[2] = sv.v[2];
}
// The same SVD with the true sqrt(3): a textbook Smith/cofactor SVD, used only
// to pin the rest of the restated algorithm against an independent SVD.
void orc_svd3_sqrt3(const float F[9], float U[9], float s[3], float Vt[9]) {
    M3 u, vt; V3 sv;
    fast_svd3(mload(F), u, sv, vt, std::sqrt(3.0f));
    mstore(u, U); mstore(vt, Vt);
    s[0] = sv.v[0]; s[1] = sv.v[1]; s[2] = sv.v[2];
}
// ms/fem/Rotation.cpp:39-57: R = U * V^T of the PhysBAM SVD
void orc_svd_rotation3(const float F[9], float R[9]) {
    M3 u, vt; V3 sv;
    fast_svd3(mload(F), u, sv, vt);
    mstore(mmul(u, vt), R);
}
// Mal/GMatDecomposition.h:17-30
void orc_pd2(const float F[4], float detF, float R[4]) {
    float sg = (detF > 0) ? 1.0f : ((detF < 0) ? -1.0f : 0.0f);
    float r00 = F[0] + sg * F[3], r01 = F[1] + sg * (-F[2]);
    float r10 = F[2] + sg * (-F[1]), r11 = F[3] + sg * F[0];
    float i0 = 1.0f / std::sqrt(r00 * r00 + r10 * r10);
    r00 *= i0; r10 *= i0;
    float i1 = 1.0f / std::sqrt(r01 * r01 + r11 * r11);
    r01 *= i1; r11 *= i1;
    R[0] = r00; R[1] = r01; R[2] = r10; R[3] = r11;
}
float orc_compute_K(const float r[12], float E, float nu, float K[144]) {
    return compute_K(vload(r), vload(r + 3), vload(r + 6), vload(r + 9), E, nu, K);
}
float orc_compute_K2(const float r[6], float E, float nu, float K[36]) {
    return compute_K2(r, r + 2, r + 4, E, nu, K);
}
float orc_sse_dot(const float* a, const float* b, uint32_t n) { return sse_dot(a, b, n); }

}  // extern "C"

// ================================================================ element set
struct orc_tetset {
    uint32_t N = 0, T = 0;
    std::vector<float> ref;          // 3N
    std::vector<uint32_t> tet;       // 4T
    std::vector<float> Kb;           // T * 16 blocks * 9, block (i,j) at (4*i+j)*9
    std::vector<float> Dinv;         // 9T
    std::vector<float> vol;          // T
    float total_volume = 0.0f;
    std::vector<float> R;            // 9T cache
    std::vector<float> detF;         // T
    std::vector<float> F;            // 9T cache (LCM / LCMH forces)
    int mm = ORC_MM_C_WRP;           // material model (ms/fem/Params.h:51-61)
    int dm = ORC_DM_TRUNCATED;       // differentials method (Params.h:68-72)
    int rm_eval = ORC_RM_MSVD;       // rotation method and threshold of the last BeginEvaluation
    float thr_eval = 0.1f;
    std::vector<float> X;            // 3N positions of the last BeginEvaluation (cache.m_vecX)
    float ecie_k = 1.0f;             // m_ECIE_k_factor (Params.h:89)
    float mu = 0.0f, lambda = 0.0f;  // Lame parameters (TetrahedronElementSet3.cpp:143-148)
    std::vector<uint8_t> doc;        // T directions of collapse (ElementCache3::m_DoC), kept across evaluations
    float dapd_l = 1.0f, dapd_nl = 0.0f, dapd_e = 2.0f;  // m_DAPD_L/NL_Factor, m_DAPD_NL_Exponent (Params.h:88)
    uint32_t n_fallback = 0;         // DAPD elements at or below the threshold left with R = Id
};

extern "C" {

// EditableTetrahedronElementSet3::BeginEdition/AddElement/EndEdition
// (ms/fem/TetrahedronElementSet3.cpp:2084-2166)
orc_tetset* orc_tetset_create(uint32_t n_nodes, const float* ref_pos, uint32_t n_tets, const uint32_t* tets,
                              float young, float poisson) {
    orc_tetset* s = new orc_tetset;
    s->N = n_nodes; s->T = n_tets;
    // LameParameters_From_YoungAndPoisson (TetrahedronElementSet3.cpp:143-148)
    s->mu = 0.5f * young / (1 + poisson);
    s->lambda = young * poisson / ((1.0f + poisson) * (1.0f - 2.0f * poisson));
    s->F.assign(9 * size_t(n_tets), 0.0f);
    s->ref.assign(ref_pos, ref_pos + 3 * size_t(n_nodes));
    s->tet.assign(tets, tets + 4 * size_t(n_tets));
    s->Kb.resize(size_t(n_tets) * 144);
    s->Dinv.resize(size_t(n_tets) * 9);
    s->vol.resize(n_tets);
    s->R.resize(size_t(n_tets) * 9);
    s->detF.resize(n_tets);
    s->doc.assign(n_tets, kDocInvalid);  // CreateCache: DoC() is invalid (TetrahedronElement3.h:30)
    float acc = 0.0f;
    float K[144];
    for (uint32_t e = 0; e < n_tets; e++) {
        V3 r0 = vload(&s->ref[3 * size_t(tets[4 * e + 0])]);
        V3 r1 = vload(&s->ref[3 * size_t(tets[4 * e + 1])]);
        V3 r2 = vload(&s->ref[3 * size_t(tets[4 * e + 2])]);
        V3 r3 = vload(&s->ref[3 * size_t(tets[4 * e + 3])]);
        M3 Dm = compute_D(r0, r1, r2, r3);
        float v = det3(Dm) / 6.0f;
        s->vol[e] = v;
        acc += v;
        compute_K(r0, r1, r2, r3, young, poisson, K);
        float* kb = &s->Kb[size_t(e) * 144];
        for (int bi = 0; bi < 4; bi++)
            for (int bj = 0; bj < 4; bj++)
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++) kb[(4 * bi + bj) * 9 + 3 * r + c] = K[(3 * bi + r) * 12 + 3 * bj + c];
        mstore(inv3(Dm), &s->Dinv[size_t(e) * 9]);
        mstore(mid(), &s->R[size_t(e) * 9]);  // CreateCache: R = Id (:190-209)
        s->detF[e] = 1.0f;
    }
    s->total_volume = acc;
    return s;
}

void orc_tetset_destroy(orc_tetset* s) { delete s; }
float orc_tetset_total_volume(const orc_tetset* s) { return s->total_volume; }

void orc_tetset_element(const orc_tetset* s, uint32_t e, float K144[144], float Dinv[9], float* vol) {
    if (K144) {
        const float* kb = &s->Kb[size_t(e) * 144];
        for (int bi = 0; bi < 4; bi++)
            for (int bj = 0; bj < 4; bj++)
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++) K144[(3 * bi + r) * 12 + 3 * bj + c] = kb[(4 * bi + bj) * 9 + 3 * r + c];
    }
    if (Dinv) std::memcpy(Dinv, &s->Dinv[size_t(e) * 9], 9 * sizeof(float));
    if (vol) *vol = s->vol[e];
}

// BeginEvaluation (ms/fem/TetrahedronElementSet3.cpp:219-281), its
// UpdateDegeneration (:477-532, the __NEW_CROSSING_DETECTION variant) and
// ComputeRotations_* (:748-892). prev_x may be NULL (then x). Under DAPD an
// element at or below the threshold with a valid DoC takes
// Compute_R_DAPD_Degenerate; one without (or with a collapsed axis) takes
// R = Id (:885-890) and is counted in n_fallback.
void orc_begin_evaluation(orc_tetset* s, const float* prev_x, const float* x, int rm, float thr, float* R_out,
                          float* detF_out, uint32_t* num_degenerate, float* sum_degenerate_detF) {
    if (!prev_x) prev_x = x;
    s->rm_eval = rm;
    s->thr_eval = thr;
    s->X.assign(x, x + 3 * size_t(s->N));
    uint32_t ndeg = 0;
    float sdeg = 0.0f;
    s->n_fallback = 0;
    for (uint32_t e = 0; e < s->T; e++) {
        const uint32_t* t = &s->tet[4 * size_t(e)];
        const V3 xs[4] = {vload(&x[3 * size_t(t[0])]), vload(&x[3 * size_t(t[1])]), vload(&x[3 * size_t(t[2])]),
                          vload(&x[3 * size_t(t[3])])};
        const M3 Dinv = mload(&s->Dinv[9 * size_t(e)]);
        M3 Ds = compute_D(xs[0], xs[1], xs[2], xs[3]);
        M3 F = mmul(Ds, Dinv);
        float dF = det3(F);
        s->detF[e] = dF;
        mstore(F, &s->F[9 * size_t(e)]);
        ndeg += (dF <= 0) ? 1u : 0u;
        sdeg += (dF <= 0) ? dF : 0.0f;
        // UpdateDegeneration: undegenerate clears the DoC; an element that was
        // undegenerate at prev_x and is degenerate now gets a new one; one that
        // was degenerate already keeps its DoC
        if (dF > thr) {
            s->doc[e] = kDocInvalid;
        } else {
            const V3 ps[4] = {vload(&prev_x[3 * size_t(t[0])]), vload(&prev_x[3 * size_t(t[1])]),
                              vload(&prev_x[3 * size_t(t[2])]), vload(&prev_x[3 * size_t(t[3])])};
            if (det3(mmul(compute_D(ps[0], ps[1], ps[2], ps[3]), Dinv)) > thr)
                s->doc[e] = compute_doc(ps, xs, s->vol[e], thr);
        }
        M3 R;
        switch (rm) {
            case ORC_RM_MSVD: { M3 u, vt; V3 sv; fast_svd3(F, u, sv, vt); R = mmul(u, vt); } break;
            case ORC_RM_DAPD:
                if (dF > thr) {
                    R = polar_newton(F, nullptr);
                } else {
                    bool ok = false;
                    if (doc_valid(s->doc[e]))
                        R = r_dapd_degenerate(dF, thr, s->doc[e], xs, s->vol[e], Dinv, s->dapd_l, s->dapd_nl,
                                              s->dapd_e, &ok);
                    if (!ok) { R = mid(); s->n_fallback++; }
                }
                break;
            case ORC_RM_ID:
            default: R = mid(); break;
        }
        mstore(R, &s->R[9 * size_t(e)]);
    }
    if (R_out) std::memcpy(R_out, s->R.data(), s->R.size() * sizeof(float));
    if (detF_out) std::memcpy(detF_out, s->detF.data(), s->detF.size() * sizeof(float));
    if (num_degenerate) *num_degenerate = ndeg;
    if (sum_degenerate_detF) *sum_degenerate_detF = sdeg;
}

void orc_set_rotations(orc_tetset* s, const float* R) { std::memcpy(s->R.data(), R, s->R.size() * sizeof(float)); }

void orc_tetset_get_doc(const orc_tetset* s, uint8_t* doc) { std::memcpy(doc, s->doc.data(), s->doc.size()); }
void orc_tetset_set_doc(orc_tetset* s, const uint8_t* doc) { std::memcpy(s->doc.data(), doc, s->doc.size()); }
void orc_tetset_set_dapd(orc_tetset* s, float l_factor, float nl_factor, float nl_exponent) {
    s->dapd_l = l_factor; s->dapd_nl = nl_factor; s->dapd_e = nl_exponent;
}
uint32_t orc_tetset_num_fallback(const orc_tetset* s) { return s->n_fallback; }

int orc_gsl_cubic(double a, double b, double c, double roots[3]) {
    roots[0] = roots[1] = roots[2] = 0.0;
    return gsl_cubic(a, b, c, &roots[0], &roots[1], &roots[2]);
}
int orc_solve_eq3(float a, float b, float c, float d, float roots[3]) { return solve_eq3(a, b, c, d, roots); }
uint8_t orc_compute_doc(const float a[12], const float b[12], float volume, float thr, float* toc) {
    const V3 av[4] = {vload(a), vload(a + 3), vload(a + 6), vload(a + 9)};
    const V3 bv[4] = {vload(b), vload(b + 3), vload(b + 6), vload(b + 9)};
    return compute_doc(av, bv, volume, thr, toc);
}
int orc_r_dapd_degenerate(float detF, float thr, uint8_t doc, const float p[12], float volume, const float inv_dm[9],
                          float l_factor, float nl_factor, float nl_exponent, float R[9]) {
    const V3 pv[4] = {vload(p), vload(p + 3), vload(p + 6), vload(p + 9)};
    bool ok = false;
    mstore(r_dapd_degenerate(detF, thr, doc, pv, volume, mload(inv_dm), l_factor, nl_factor, nl_exponent, &ok), R);
    return ok ? 1 : 0;
}

namespace {
// Shared body of C_f_e / C_df_x: v_i -= R * sum_j K_ij * l_j, with l_j the
// local-frame vectors, accumulated per element in element order.
inline void corot_apply(const orc_tetset* s, uint32_t e, const V3 l[4], float* out) {
    const uint32_t* t = &s->tet[4 * size_t(e)];
    const float* kb = &s->Kb[size_t(e) * 144];
    M3 R = mload(&s->R[9 * size_t(e)]);
    for (int i = 0; i < 4; i++) {
        V3 acc = mvec(mload(kb + (4 * i + 0) * 9), l[0]);
        acc = vadd(acc, mvec(mload(kb + (4 * i + 1) * 9), l[1]));
        acc = vadd(acc, mvec(mload(kb + (4 * i + 2) * 9), l[2]));
        acc = vadd(acc, mvec(mload(kb + (4 * i + 3) * 9), l[3]));
        V3 w = mvec(R, acc);
        float* o = &out[3 * size_t(t[i])];
        o[0] -= w.v[0]; o[1] -= w.v[1]; o[2] -= w.v[2];
    }
}
}  // namespace

void orc_tetset_set_material(orc_tetset* s, int mm, float ecie_k_factor) {
    s->mm = mm;
    s->ecie_k = ecie_k_factor;
}
void orc_tetset_set_dm(orc_tetset* s, int dm) { s->dm = dm; }
}  // extern "C"
namespace {
bool compute_dR(const M3& dDs, const M3& invDm, const M3& F, const M3& R, const M3& Rt, M3& dR);
}
extern "C" int orc_dR(const float F[9], const float dF[9], const float R[9], float dR[9]) {
    M3 out = mzero();
    const M3 Rm = mload(R);
    const bool ok = compute_dR(mload(dF), mid(), mload(F), Rm, mtr(Rm), out);
    mstore(out, dR);
    return ok ? 1 : 0;
}

namespace {
// Compute_dR (ms/fem/Rotation.cpp:622-644, McAdams 2011): dF = dDs Dm^-1,
// S = R^T F, W = R^T dF, G = S - tr(S) I, w = axial(W - W^T);
// dR = R skew(G^-1 w) when |det G| > 1e-4 (mal::Epsilon), else no dR.
bool compute_dR(const M3& dDs, const M3& invDm, const M3& F, const M3& R, const M3& Rt, M3& dR) {
    const M3 dF = mmul(dDs, invDm);
    const M3 S = mmul(Rt, F);
    const M3 W = mmul(Rt, dF);
    const float tr = mtrace_(S);
    M3 G = S;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) G.a[i][j] = S.a[i][j] - tr * ((i == j) ? 1.0f : 0.0f);
    V3 w;
    w.v[0] = W.a[1][2] - W.a[2][1];
    w.v[1] = W.a[2][0] - W.a[0][2];
    w.v[2] = W.a[0][1] - W.a[1][0];
    const float d = det3(G);
    if (!((d > 0.0f ? d : -d) > 0.0001f)) return false;
    const V3 r = mvec(inv3(G, d), w);
    M3 K = mzero();
    K.a[0][1] = -r.v[2]; K.a[0][2] = r.v[1];
    K.a[1][0] = r.v[2];  K.a[1][2] = -r.v[0];
    K.a[2][0] = -r.v[1]; K.a[2][1] = r.v[0];
    dR = mmul(R, K);
    return true;
}

// Compute_dR_From_RM (TetrahedronElementSet3.cpp:894-953) for DM = Exact:
// MSVD always, DAPD above the degeneration threshold.
bool element_dR(const orc_tetset* s, uint32_t e, const float* dx, M3& dR) {
    if (s->dm != ORC_DM_EXACT) return false;
    if (s->rm_eval == ORC_RM_DAPD && !(s->detF[e] > s->thr_eval)) return false;
    if (s->rm_eval != ORC_RM_DAPD && s->rm_eval != ORC_RM_MSVD) return false;
    const uint32_t* t = &s->tet[4 * size_t(e)];
    const M3 R = mload(&s->R[9 * size_t(e)]);
    const M3 dDs = compute_D(vload(&dx[3 * size_t(t[0])]), vload(&dx[3 * size_t(t[1])]),
                             vload(&dx[3 * size_t(t[2])]), vload(&dx[3 * size_t(t[3])]));
    return compute_dR(dDs, mload(&s->Dinv[9 * size_t(e)]), mload(&s->F[9 * size_t(e)]), R, mtr(R), dR);
}
inline M3 mlin(float a, const M3& x, float b, const M3& y) {  // a x + b y, element-wise
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.a[i][j] = a * x.a[i][j] + b * y.a[i][j];
    return r;
}
inline M3 madd(const M3& x, const M3& y) { M3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.a[i][j] = x.a[i][j] + y.a[i][j]; return r; }
inline M3 mscale(float a, const M3& x) { M3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.a[i][j] = a * x.a[i][j]; return r; }
inline float mtrace(const M3& m) { float t = 0.0f; for (int i = 0; i < 3; i++) t += m.a[i][i]; return t; }

// Piola force of one element, H = -V P Dm^-T, scattered as the reference
// does: h0 = -h1 - h2 - h3 to vertex 0, column c to vertex c+1
// (TetrahedronElementSet3.cpp:1171-1176).
inline void scatter_piola(const orc_tetset* s, uint32_t e, const M3& P, float* out) {
    const uint32_t* t = &s->tet[4 * size_t(e)];
    const M3 H = mmul(mscale(-s->vol[e], P), mtr(mload(&s->Dinv[9 * size_t(e)])));
    const V3 c0 = col(H, 0), c1 = col(H, 1), c2 = col(H, 2);
    const V3 h0 = vsub(vsub(vscale(c0, -1.0f), c1), c2);
    float* o0 = &out[3 * size_t(t[0])];
    for (int i = 0; i < 3; i++) o0[i] += h0.v[i];
    for (int k = 0; k < 3; k++) {
        float* o = &out[3 * size_t(t[k + 1])];
        const V3 c = col(H, k);
        for (int i = 0; i < 3; i++) o[i] += c.v[i];
    }
}

// C_LCM_f_e (:1155-1178): S = R^T F, P = R (2 mu (S - I) + lambda tr(S - I) I)
void lcm_f_e(const orc_tetset* s, float* f) {
    for (uint32_t e = 0; e < s->T; e++) {
        const M3 R = mload(&s->R[9 * size_t(e)]);
        const M3 S = mmul(mtr(R), mload(&s->F[9 * size_t(e)]));
        M3 SmI = S;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) SmI.a[i][j] = S.a[i][j] - mid().a[i][j];
        const M3 P = mmul(R, mlin(2 * s->mu, SmI, s->lambda * mtrace(SmI), mid()));
        scatter_piola(s, e, P, f);
    }
}

// C_LCM_df_x (:1838-1872) and C_LCMH_df_x (:1874-1929), DM = Truncated:
// dF = dDs Dm^-1; LCM dP = 2 mu dF + lambda tr(R^T dF) R; LCMH blends 2 mu dF
// with 2 mu R sym(R^T dF) by w = clamp01(tr|S - I| / (3 k)) when k > 0.1.
void lcm_df_x(const orc_tetset* s, const float* dx, float* df, bool hybrid) {
    for (uint32_t e = 0; e < s->T; e++) {
        const uint32_t* t = &s->tet[4 * size_t(e)];
        const M3 R = mload(&s->R[9 * size_t(e)]);
        const M3 Rt = mtr(R);
        const M3 dDs = compute_D(vload(&dx[3 * size_t(t[0])]), vload(&dx[3 * size_t(t[1])]),
                                 vload(&dx[3 * size_t(t[2])]), vload(&dx[3 * size_t(t[3])]));
        const M3 dF = mmul(dDs, mload(&s->Dinv[9 * size_t(e)]));
        const M3 RtdF = mmul(Rt, dF);
        M3 dP;
        if (hybrid && s->ecie_k > 0.1f) {
            const M3 S = mmul(Rt, mload(&s->F[9 * size_t(e)]));
            float tr_abs = 0.0f;
            for (int i = 0; i < 3; i++) tr_abs += std::fabs(S.a[i][i] - 1.0f);
            float w = (1.0f / s->ecie_k) * tr_abs / 3.0f;
            w = std::max(0.0f, std::min(w, 1.0f));
            M3 sym;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) sym.a[i][j] = 0.5f * (RtdF.a[i][j] + RtdF.a[j][i]);
            const M3 blend = madd(mscale(w, dF), mmul(mscale(1.0f - w, R), sym));
            dP = madd(mscale(2 * s->mu, blend), mscale(s->lambda * mtrace(RtdF), R));
        } else {
            dP = madd(mscale(2 * s->mu, dF), mscale(s->lambda * mtrace(RtdF), R));
        }
        M3 dR;
        if (element_dR(s, e, dx, dR)) {  // dP += (lambda tr(S - I) - 2 mu) dR (:1862-1864)
            const M3 S = mmul(Rt, mload(&s->F[9 * size_t(e)]));
            M3 SmI = S;
            for (int i = 0; i < 3; i++) SmI.a[i][i] = S.a[i][i] - 1.0f;
            dP = madd(dP, mscale(s->lambda * mtrace(SmI) - 2 * s->mu, dR));
        }
        scatter_piola(s, e, dP, df);
    }
}
}  // namespace

extern "C" {

// C_f_e (ms/fem/TetrahedronElementSet3.cpp:1053-1096): u_j = R^T x_j - r_j
// (LCM / LCMH: C_LCM_f_e from the F cached at BeginEvaluation)
void orc_f_e(const orc_tetset* s, const float* x, float* f) {
    if (s->mm == ORC_MM_C_LCM || s->mm == ORC_MM_C_LCMH) {
        lcm_f_e(s, f);
        return;
    }
    for (uint32_t e = 0; e < s->T; e++) {
        const uint32_t* t = &s->tet[4 * size_t(e)];
        M3 Rt = mtr(mload(&s->R[9 * size_t(e)]));
        V3 u[4];
        for (int j = 0; j < 4; j++) u[j] = vsub(mvec(Rt, vload(&x[3 * size_t(t[j])])), vload(&s->ref[3 * size_t(t[j])]));
        corot_apply(s, e, u, f);
    }
}

// C_df_x with DM = Truncated (:1761-1836; the dR terms are skipped, :901)
void orc_df_x(const orc_tetset* s, const float* dx, float* df) {
    if (s->mm == ORC_MM_C_LCM || s->mm == ORC_MM_C_LCMH) {
        lcm_df_x(s, dx, df, s->mm == ORC_MM_C_LCMH);
        return;
    }
    for (uint32_t e = 0; e < s->T; e++) {
        const uint32_t* t = &s->tet[4 * size_t(e)];
        M3 Rt = mtr(mload(&s->R[9 * size_t(e)]));
        V3 l[4];
        for (int j = 0; j < 4; j++) l[j] = mvec(Rt, vload(&dx[3 * size_t(t[j])]));
        corot_apply(s, e, l, df);
        M3 dR;
        if (element_dR(s, e, dx, dR)) {  // dR terms of C_df_x (:1804-1834)
            const float* kb = &s->Kb[size_t(e) * 144];
            V3 u[4], q[4];
            const M3 dRt = mtr(dR);
            for (int j = 0; j < 4; j++) {
                const V3 xj = vload(&s->X[3 * size_t(t[j])]);
                u[j] = vsub(mvec(Rt, xj), vload(&s->ref[3 * size_t(t[j])]));
                q[j] = mvec(dRt, xj);
            }
            for (int i = 0; i < 4; i++) {  // df -= dR K u
                V3 acc = mvec(mload(kb + (4 * i + 0) * 9), u[0]);
                acc = vadd(acc, mvec(mload(kb + (4 * i + 1) * 9), u[1]));
                acc = vadd(acc, mvec(mload(kb + (4 * i + 2) * 9), u[2]));
                acc = vadd(acc, mvec(mload(kb + (4 * i + 3) * 9), u[3]));
                const V3 w = mvec(dR, acc);
                float* o = &df[3 * size_t(t[i])];
                o[0] -= w.v[0]; o[1] -= w.v[1]; o[2] -= w.v[2];
            }
            for (int i = 0; i < 4; i++) {  // df -= R K dR^T x
                V3 acc = mvec(mload(kb + (4 * i + 0) * 9), q[0]);
                acc = vadd(acc, mvec(mload(kb + (4 * i + 1) * 9), q[1]));
                acc = vadd(acc, mvec(mload(kb + (4 * i + 2) * 9), q[2]));
                acc = vadd(acc, mvec(mload(kb + (4 * i + 3) * 9), q[3]));
                const V3 w = mvec(mtr(Rt), acc);
                float* o = &df[3 * size_t(t[i])];
                o[0] -= w.v[0]; o[1] -= w.v[1]; o[2] -= w.v[2];
            }
        }
    }
}

// C_f_d (:1098-1150): corotational Rayleigh damping from the stored Ke (for
// every corotational material), f_i -= R ((c0 m_i) R^T v_i + c1 sum_j K_ij R^T v_j)
// per element, m_i = 1 / inv_mass_i.
void orc_f_d(const orc_tetset* s, const float* v, const float* inv_mass, float c0, float c1, float* f) {
    for (uint32_t e = 0; e < s->T; e++) {
        const uint32_t* t = &s->tet[4 * size_t(e)];
        const float* kb = &s->Kb[size_t(e) * 144];
        const M3 R = mload(&s->R[9 * size_t(e)]);
        const M3 Rt = mtr(R);
        V3 l[4];
        for (int j = 0; j < 4; j++) l[j] = mvec(Rt, vload(&v[3 * size_t(t[j])]));
        for (int i = 0; i < 4; i++) {
            V3 acc = mvec(mload(kb + (4 * i + 0) * 9), l[0]);
            acc = vadd(acc, mvec(mload(kb + (4 * i + 1) * 9), l[1]));
            acc = vadd(acc, mvec(mload(kb + (4 * i + 2) * 9), l[2]));
            acc = vadd(acc, mvec(mload(kb + (4 * i + 3) * 9), l[3]));
            const float sm = c0 * (1.0f / inv_mass[t[i]]);
            const V3 w = mvec(R, vadd(vscale(l[i], sm), vscale(acc, c1)));
            float* o = &f[3 * size_t(t[i])];
            o[0] -= w.v[0]; o[1] -= w.v[1]; o[2] -= w.v[2];
        }
    }
}

// C_V (:1456-1488); LCM / LCMH: C_LCM_V (:1489-1508),
// sum of V (mu |S - I|_F^2 + lambda/2 tr(S - I)^2) with S = R^T F
float orc_V(const orc_tetset* s, const float* x) {
    if (s->mm == ORC_MM_C_LCM || s->mm == ORC_MM_C_LCMH) {
        float acc = 0.0f;
        for (uint32_t e = 0; e < s->T; e++) {
            const M3 S = mmul(mtr(mload(&s->R[9 * size_t(e)])), mload(&s->F[9 * size_t(e)]));
            float nsq = 0.0f, tr = 0.0f;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    const float d = S.a[i][j] - ((i == j) ? 1.0f : 0.0f);
                    nsq += d * d;
                }
            for (int i = 0; i < 3; i++) tr += S.a[i][i] - 1.0f;
            acc += s->vol[e] * (s->mu * nsq + 0.5f * s->lambda * (tr * tr));
        }
        return acc;
    }
    float acc2 = 0.0f;
    for (uint32_t e = 0; e < s->T; e++) {
        const uint32_t* t = &s->tet[4 * size_t(e)];
        const float* kb = &s->Kb[size_t(e) * 144];
        M3 Rt = mtr(mload(&s->R[9 * size_t(e)]));
        V3 u[4];
        for (int j = 0; j < 4; j++) u[j] = vsub(mvec(Rt, vload(&x[3 * size_t(t[j])])), vload(&s->ref[3 * size_t(t[j])]));
        for (int i = 0; i < 4; i++) {
            V3 acc = mvec(mload(kb + (4 * i + 0) * 9), u[0]);
            acc = vadd(acc, mvec(mload(kb + (4 * i + 1) * 9), u[1]));
            acc = vadd(acc, mvec(mload(kb + (4 * i + 2) * 9), u[2]));
            acc = vadd(acc, mvec(mload(kb + (4 * i + 3) * 9), u[3]));
            acc2 += vdot(u[i], acc);
        }
    }
    return 0.5f * acc2;
}

}  // extern "C"

// ================================================================ CG + leaf
namespace {

// Rayleigh coefficients c = inv(Omega) * 2*zeta (LeafDSH_Solid3D_FEM.h:347-354)
void rayleigh(float f1, float f2, float zeta, float c[2]) {
    float o00 = 1.0f / f1, o01 = f1, o10 = 1.0f / f2, o11 = f2;
    float det = o00 * o11 - o10 * o01;
    float id = 1.0f / det;
    float i00 = id * o11, i01 = -id * o01, i10 = -id * o10, i11 = id * o00;
    float z0 = 2 * zeta, z1 = 2 * zeta;
    float acc = 0.0f; acc += i00 * z0; acc += i01 * z1; c[0] = acc;
    acc = 0.0f; acc += i10 * z0; acc += i11 * z1; c[1] = acc;
}

// Unpreconditioned modified CG with projection S (ns/ILSS_CG.h:38-151).
template <class AV, class SP>
void cg_solve(uint32_t n, AV&& Av, SP&& S, float* x, const float* b, float rel, float abs_eps, uint32_t max_it,
              std::vector<float>& r, std::vector<float>& d, std::vector<float>& q,
              int32_t* iters, float* rel_res, float* abs_res, bool red64) {
    auto sse_dot = [red64](const float* a, const float* c, uint32_t m) { return red64 ? fp64_dot(a, c, m) : ::sse_dot(a, c, m); };
    float abs_sq = abs_eps * abs_eps;
    Av(r.data(), x);
    for (uint32_t i = 0; i < n; i++) r[i] = -1.0f * r[i] + b[i];  // muleq_addeq (real_array.h:78)
    S(r.data());
    std::memcpy(d.data(), r.data(), n * sizeof(float));
    float delta0 = sse_dot(r.data(), r.data(), n);
    float dnew = delta0, dold = delta0;
    uint32_t k = 0;
    float target = (rel * rel) * delta0;
    while (dnew > abs_sq && dnew > target && k < max_it) {
        Av(q.data(), d.data());
        S(q.data());
        float dq = sse_dot(d.data(), q.data(), n);
        if (std::fabs(dq) < abs_sq) break;
        float alpha = dnew / dq;
        for (uint32_t i = 0; i < n; i++) x[i] = x[i] + d[i] * alpha;   // addeq_scaled (:66-67)
        for (uint32_t i = 0; i < n; i++) r[i] = r[i] - q[i] * alpha;   // subeq_scaled (:69-70)
        dold = dnew;
        dnew = sse_dot(r.data(), r.data(), n);
        float beta = dnew / dold;
        for (uint32_t i = 0; i < n; i++) d[i] = beta * d[i] + r[i];    // muleq_addeq (:78)
        S(d.data());
        k++;
    }
    *iters = int32_t(k);
    *rel_res = (delta0 > 0) ? std::sqrt(dnew / delta0) : 0.0f;
    *abs_res = std::sqrt(dnew);
}

}  // namespace

struct orc_solid {
    orc_solid_params p;
    orc_tetset* tes = nullptr;
    uint32_t N = 0;
    std::vector<float> x, v, x0;   // positions, velocities, previous positions (m_vecPos0)
    float inv_mass = 0.0f;         // uniform
    float c[2] = {0, 0};           // Rayleigh
    std::vector<uint32_t> knc;
    std::vector<float> knc_pos;    // 3N: each KNC node's target position (KNC::m_Pos), the rest pose by default
    std::vector<float> b, y, r, d, q;
    std::vector<float> xk, vk;     // FIE-dx iterates (m_IE_vec_x_k, m_IE_vec_v_k)
    float acc_time = 0.0f;
    float total_time = 0.0f;
    // last BeginEvaluation's counts
    uint32_t ndeg = 0;
    float sdeg = 0.0f, min_det = 0.0f;
    int32_t ls_iters = 0;
    float ls_rel = 0.0f, ls_abs = 0.0f;
};

namespace {

void solid_begin_evaluation(orc_solid* s, const float* prev_x, const float* x) {
    orc_begin_evaluation(s->tes, prev_x, x, s->p.rm, s->p.degenerate_threshold, nullptr, nullptr, &s->ndeg, &s->sdeg);
    s->min_det = 1e30f;
    for (uint32_t e = 0; e < s->tes->T; e++) s->min_det = std::min(s->min_det, s->tes->detF[e]);
}

// LS_Solve (ds/dsh/LeafDSH_Solid3D_FEM.cpp:1081-1155): CG on A y = b with
// A = fM M - fK df_x (LS_A_times_d :1056-1079) and KNC rows zeroed (:1045-1053).
void solid_solve(orc_solid* s, float fM, float fK) {
    const uint32_t N = s->N;
    const size_t n3 = 3 * size_t(N);
    const float mass = 1.0f / s->inv_mass;
    auto Av = [&](float* q, const float* d) {
        std::memset(q, 0, n3 * sizeof(float));
        orc_df_x(s->tes, d, q);
        for (size_t i = 0; i < n3; i++) q[i] *= -fK;
        for (uint32_t i = 0; i < N; i++) {
            float factor = mass * fM;
            q[3 * i] += factor * d[3 * i];
            q[3 * i + 1] += factor * d[3 * i + 1];
            q[3 * i + 2] += factor * d[3 * i + 2];
        }
    };
    auto S = [&](float* w) {
        for (uint32_t nid : s->knc) { w[3 * size_t(nid)] = 0; w[3 * size_t(nid) + 1] = 0; w[3 * size_t(nid) + 2] = 0; }
    };
    cg_solve(uint32_t(n3), Av, S, s->y.data(), s->b.data(), s->p.cg_rel_epsilon, s->p.cg_abs_epsilon, s->p.cg_max_iter,
             s->r, s->d, s->q, &s->ls_iters, &s->ls_rel, &s->ls_abs, s->p.reduce_fp64 != 0);
}

float solid_norm(const orc_solid* s, const std::vector<float>& w) {  // ns::real_array::norm_2 (real_array.h:42-50)
    const float sq = s->p.reduce_fp64 ? fp64_dot(w.data(), w.data(), uint32_t(w.size()))
                                      : sse_dot(w.data(), w.data(), uint32_t(w.size()));
    return std::sqrt(sq);
}

// Right-hand side of one FIE-dx Newton iteration at the iterate x_k, with the
// elastic force evaluated at x_eval under the current rotations
// (Step_FullyImplicitEuler_dx :1378-1465): b = M g + f_e + (1/dt) M (v - (x_k - x)/dt),
// y = dt v on the first iteration (warm start) else 0, KNC rows b = y = target - x_k.
void fie_rhs(orc_solid* s, float dt, float inv_dt, const float* x_eval, bool first) {
    const uint32_t N = s->N;
    const float mass = 1.0f / s->inv_mass;
    for (uint32_t i = 0; i < N; i++)
        for (int c = 0; c < 3; c++) s->b[3 * i + c] = 0.0f + mass * s->p.gravity[c];
    orc_f_e(s->tes, x_eval, s->b.data());
    const float sm = inv_dt * mass;
    for (size_t i = 0; i < s->b.size(); i++) s->b[i] += sm * (s->v[i] - inv_dt * (s->xk[i] - s->x[i]));
    if (first && s->p.warm_start) {
        for (size_t i = 0; i < s->y.size(); i++) s->y[i] = dt * s->v[i];
    } else {
        std::fill(s->y.begin(), s->y.end(), 0.0f);
    }
    for (uint32_t nid : s->knc)
        for (int c = 0; c < 3; c++) {
            const size_t i = 3 * size_t(nid) + c;
            s->b[i] = s->knc_pos[i] - s->xk[i];
            s->y[i] = s->knc_pos[i] - s->xk[i];
        }
}

// x_k += y ; v_k = (x_k - x) / dt (:1500-1511)
void fie_update(orc_solid* s, float inv_dt) {
    for (size_t i = 0; i < s->xk.size(); i++) s->xk[i] = s->xk[i] + s->y[i];
    for (size_t i = 0; i < s->vk.size(); i++) s->vk[i] = inv_dt * (s->xk[i] - s->x[i]);
}

}  // namespace

extern "C" {

// LeafDSH_Solid3D_FEM::Create (ds/dsh/LeafDSH_Solid3D_FEM.cpp:125-223) + RebuildMass (:1591-1611)
orc_solid* orc_solid_create(uint32_t n_nodes, const float* ref_pos, uint32_t n_tets, const uint32_t* tets,
                            const orc_solid_params* p) {
    orc_solid* s = new orc_solid;
    s->p = *p;
    s->N = n_nodes;
    s->tes = orc_tetset_create(n_nodes, ref_pos, n_tets, tets, p->young_modulus, p->poisson_ratio);
    orc_tetset_set_material(s->tes, p->mm ? p->mm : ORC_MM_C_WRP, p->ecie_k_factor);
    orc_tetset_set_dm(s->tes, p->dm);
    s->x.assign(ref_pos, ref_pos + 3 * size_t(n_nodes));
    s->x0 = s->x;
    s->v.assign(3 * size_t(n_nodes), 0.0f);
    float total_mass = p->density * s->tes->total_volume;
    float mass_per_node = total_mass / float(n_nodes);
    s->inv_mass = 1.0f / mass_per_node;  // GParticleSystem::SetMass (ms/ParticleSystem.h:80)
    rayleigh(p->rayleigh_freq1, p->rayleigh_freq2, p->damping_ratio, s->c);
    size_t n3 = 3 * size_t(n_nodes);
    s->b.resize(n3); s->y.resize(n3); s->r.resize(n3); s->d.resize(n3); s->q.resize(n3);
    return s;
}

void orc_solid_destroy(orc_solid* s) { orc_tetset_destroy(s->tes); delete s; }
void orc_solid_set_knc(orc_solid* s, uint32_t n, const uint32_t* nids) {
    s->knc.assign(nids, nids + n);
    s->knc_pos = s->tes->ref;
}
void orc_solid_set_knc_targets(orc_solid* s, uint32_t n, const uint32_t* nids, const float* targets) {
    orc_solid_set_knc(s, n, nids);
    for (uint32_t i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) s->knc_pos[3 * size_t(nids[i]) + c] = targets[3 * size_t(i) + c];
}
float orc_solid_node_mass(const orc_solid* s) { return 1.0f / s->inv_mass; }
void orc_solid_rayleigh(const orc_solid* s, float c[2]) { c[0] = s->c[0]; c[1] = s->c[1]; }

// Step_FullyImplicitEuler_dx (:1357-1552): Newton iterations on the implicit
// Euler equations in position increments; dt = 0 is the quasi-static variant
// (:957-959). Returns the iterations taken and the last |b| or |y|.
static void step_fie_dx(orc_solid* s, float dt, int32_t* nr_iters, float* nr_rel, float* nr_abs, int32_t* ls_total) {
    const float inv_dt = (dt == 0.0f) ? 0.0f : 1.0f / dt;
    s->xk = s->x;
    s->vk = s->v;
    uint32_t nr_k = 0;
    float delta_new = 0.0f, delta0 = -1.0f;
    bool conv = false;
    *ls_total = 0;
    do {
        fie_rhs(s, dt, inv_dt, s->xk.data(), nr_k == 0);
        delta_new = solid_norm(s, s->b);  // after KNC (:1468-1474)
        if (delta0 < 0) delta0 = delta_new;
        conv = delta_new < s->p.nr_abs_epsilon;
        if (!conv) {
            const float fM = inv_dt * inv_dt + inv_dt * s->c[0];
            const float fK = 1.0f + inv_dt * s->c[1];
            solid_solve(s, fM, fK);
            *ls_total += s->ls_iters;
            fie_update(s, inv_dt);
            delta_new = solid_norm(s, s->y);
            nr_k++;
            solid_begin_evaluation(s, s->x.data(), s->xk.data());  // End/BeginEvaluation at x_k (:1526-1530)
        }
    } while (!conv && nr_k < s->p.nr_max_iter);
    *nr_iters = int32_t(nr_k);
    *nr_rel = (delta0 > 0) ? delta_new / delta0 : 0.0f;
    *nr_abs = delta_new;
    s->x = s->xk;
    s->v = s->vk;
}

// FixedStep (:898-990) with Step_QuasiImplicitEuler_v (:1205-1328) or
// Step_FullyImplicitEuler_dx (:1357-1552). FixedStep itself leaves the total
// time alone: Step advances it once per call (:229-230).
static void solid_fixed_step(orc_solid* s, float dt, orc_step_stats* st) {
    if (s->p.integrator == ORC_IT_FIE_DX) {
        solid_begin_evaluation(s, s->x0.data(), s->x.data());
        s->x0 = s->x;
        int32_t nr = 0, ls = 0;
        float nr_rel = 0, nr_abs = 0;
        step_fie_dx(s, dt, &nr, &nr_rel, &nr_abs, &ls);
        const float V = orc_V(s->tes, s->x.data());  // V() at the final iterate (:977)
        if (st) {
            st->num_fixed_steps = 1;
            st->cg_iters = ls; st->cg_rel = s->ls_rel; st->cg_abs = s->ls_abs;
            st->num_degenerate = s->ndeg; st->sum_degenerate_detF = s->sdeg;
            st->elastic_energy = V; st->min_detF = s->min_det;
            st->nr_iters = nr; st->nr_rel = nr_rel; st->nr_abs = nr_abs;
        }
        return;
    }
    const uint32_t N = s->N;
    const size_t n3 = 3 * size_t(N);
    const float mass = 1.0f / s->inv_mass;  // GetMass (ms/ParticleSystem.h:74)
    // UpdateKNC, static mode: KNC target = rest position (:877-896)
    uint32_t ndeg = 0; float sdeg = 0.0f;
    orc_begin_evaluation(s->tes, s->x0.data(), s->x.data(), s->p.rm, s->p.degenerate_threshold, nullptr, nullptr, &ndeg,
                         &sdeg);
    float min_det = 1e30f;
    for (uint32_t e = 0; e < s->tes->T; e++) min_det = std::min(min_det, s->tes->detF[e]);
    s->x0 = s->x;
    float inv_dt = 1.0f / dt;
    // b = M g
    for (uint32_t i = 0; i < N; i++)
        for (int c = 0; c < 3; c++) s->b[3 * i + c] = 0.0f + mass * s->p.gravity[c];
    orc_f_e(s->tes, s->x.data(), s->b.data());
    for (size_t i = 0; i < n3; i++) s->b[i] *= dt;
    for (uint32_t i = 0; i < N; i++)
        for (int c = 0; c < 3; c++) s->b[3 * i + c] += mass * s->v[3 * i + c];
    if (s->p.warm_start) s->y = s->v; else std::fill(s->y.begin(), s->y.end(), 0.0f);
    // KNC (:1268-1295): position constraints at their targets
    for (uint32_t nid : s->knc) {
        V3 vk;
        for (int c = 0; c < 3; c++) vk.v[c] = (s->knc_pos[3 * size_t(nid) + c] - s->x[3 * size_t(nid) + c]) * inv_dt;
        float nk = vnorm(vk);
        V3 kv;
        if (nk > 0.001f) kv = vscale(vnormalized(vk), std::min(nk, 50.0f));
        else { kv.v[0] = kv.v[1] = kv.v[2] = 0.0f; }
        for (int c = 0; c < 3; c++) { s->b[3 * size_t(nid) + c] = kv.v[c]; s->y[3 * size_t(nid) + c] = kv.v[c]; }
    }
    const float fM = 1.0f + dt * s->c[0];
    const float fK = dt * dt + dt * s->c[1];
    auto Av = [&](float* q, const float* d) {  // LS_A_times_d (:1056-1079)
        std::memset(q, 0, n3 * sizeof(float));
        orc_df_x(s->tes, d, q);
        for (size_t i = 0; i < n3; i++) q[i] *= -fK;
        for (uint32_t i = 0; i < N; i++) {
            float factor = mass * fM;
            q[3 * i] += factor * d[3 * i];
            q[3 * i + 1] += factor * d[3 * i + 1];
            q[3 * i + 2] += factor * d[3 * i + 2];
        }
    };
    auto S = [&](float* w) {  // LS_ZeroKNC (:1045-1053)
        for (uint32_t nid : s->knc) { w[3 * size_t(nid)] = 0; w[3 * size_t(nid) + 1] = 0; w[3 * size_t(nid) + 2] = 0; }
    };
    int32_t it = 0; float rr = 0, ar = 0;
    cg_solve(uint32_t(n3), Av, S, s->y.data(), s->b.data(), s->p.cg_rel_epsilon, s->p.cg_abs_epsilon, s->p.cg_max_iter,
             s->r, s->d, s->q, &it, &rr, &ar, s->p.reduce_fp64 != 0);
    for (uint32_t i = 0; i < N; i++)
        for (int c = 0; c < 3; c++) {
            s->v[3 * i + c] = s->y[3 * i + c];
            s->x[3 * i + c] = s->x[3 * i + c] + dt * s->y[3 * i + c];
        }
    float V = orc_V(s->tes, s->x.data());
    if (st) {
        st->num_fixed_steps = 1;
        st->cg_iters = it; st->cg_rel = rr; st->cg_abs = ar;
        st->num_degenerate = ndeg; st->sum_degenerate_detF = sdeg;
        st->elastic_energy = V; st->min_detF = min_det;
        st->nr_iters = 0; st->nr_rel = 0; st->nr_abs = 0;
    }
}

void orc_solid_fixed_step(orc_solid* s, float dt, orc_step_stats* st) {
    solid_fixed_step(s, dt, st);
    s->total_time += dt;
}

// SS_AggregateDSH_Basic::MultiDSH_Solid3D_Update (ds/ss/SS_AggregateDSH_Basic.cpp:229-300)
// without contacts, with the leaves' NR hooks (LeafDSH_Solid3D_FEM.cpp:2824-3039).
// NR_PreStep resets x_k = x, v_k = v and the iteration count on every pass, and
// the cache keeps pointing at x_k, so each pass evaluates f_e at x under the
// rotations of the previous pass's end positions and warm-starts y = dt v.
uint32_t orc_multi_update(orc_solid* const* solids, uint32_t n, float dt, float* acc_time, orc_step_stats* stats) {
    float fixed_dt = std::numeric_limits<float>::infinity();
    uint32_t max_nr = 0;
    for (uint32_t i = 0; i < n; i++) {
        fixed_dt = std::min(fixed_dt, solids[i]->p.fixed_dt);
        max_nr = std::max(max_nr, solids[i]->p.nr_max_iter);
        solids[i]->total_time += dt;
    }
    *acc_time += dt;
    const float inv_dt = (dt == 0.0f) ? 0.0f : 1.0f / dt;
    uint32_t steps = 0;
    while (*acc_time > fixed_dt) {
        for (uint32_t i = 0; i < n; i++) {  // NR_Init (:2824-2836)
            orc_solid* s = solids[i];
            solid_begin_evaluation(s, s->x0.data(), s->x.data());
            s->x0 = s->x;
        }
        uint32_t nr = 0;
        do {
            for (uint32_t i = 0; i < n; i++) {  // NR_PreStep + NR_MidStep (:2857-2985)
                orc_solid* s = solids[i];
                s->xk = s->x;
                s->vk = s->v;
                fie_rhs(s, dt, inv_dt, s->xk.data(), true);
                solid_solve(s, inv_dt * inv_dt + inv_dt * s->c[0], 1.0f + inv_dt * s->c[1]);
            }
            for (uint32_t i = 0; i < n; i++) {  // NR_PostStep (:2987-3037)
                orc_solid* s = solids[i];
                fie_update(s, inv_dt);
                solid_begin_evaluation(s, s->x.data(), s->xk.data());
            }
        } while (++nr < max_nr);
        for (uint32_t i = 0; i < n; i++) {  // NR_End (:2838-2852)
            orc_solid* s = solids[i];
            s->x = s->xk;
            s->v = s->vk;
            if (stats) {
                orc_step_stats& st = stats[i];
                st.num_fixed_steps = 1;
                st.cg_iters = s->ls_iters; st.cg_rel = s->ls_rel; st.cg_abs = s->ls_abs;
                st.num_degenerate = s->ndeg; st.sum_degenerate_detF = s->sdeg;
                st.elastic_energy = 0.0f; st.min_detF = s->min_det;
                st.nr_iters = int32_t(nr); st.nr_rel = 0.0f; st.nr_abs = 0.0f;
            }
        }
        *acc_time -= fixed_dt;
        steps++;
    }
    return steps;
}

// Step (:225-256): scaled_dt = time_scale dt is added to the total time and
// to the accumulator; fixed steps of time_scale fixed_dt run while the
// accumulator is strictly greater (:239)
void orc_solid_step(orc_solid* s, float dt, orc_step_stats* st) {
    const float scaled_dt = s->p.time_scale * dt;
    s->total_time += scaled_dt;
    s->acc_time += scaled_dt;
    const float scaled_fixed_dt = s->p.time_scale * s->p.fixed_dt;
    orc_step_stats last;
    std::memset(&last, 0, sizeof(last));
    uint32_t n = 0;
    while (s->acc_time > scaled_fixed_dt) {
        solid_fixed_step(s, scaled_fixed_dt, &last);
        s->acc_time -= scaled_fixed_dt;
        n++;
    }
    last.num_fixed_steps = n;
    if (st) *st = last;
}

float orc_solid_total_time(const orc_solid* s) { return s->total_time; }

// Many independent solids, one fixed step each, spread over `threads` host
// threads (each solid stepped whole by one thread): the mesh-parallel CPU
// baseline of the batched workloads.
void orc_solids_fixed_step_parallel(orc_solid* const* solids, uint32_t n, uint32_t threads) {
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        for (uint32_t i = next++; i < n; i = next++) orc_solid_fixed_step(solids[i], solids[i]->p.fixed_dt, nullptr);
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < std::max(1u, threads); t++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
}

void orc_rayleigh(float f1, float f2, float zeta, float c[2]) { rayleigh(f1, f2, zeta, c); }

void orc_solid_get_state(const orc_solid* s, float* x, float* v) {
    if (x) std::memcpy(x, s->x.data(), s->x.size() * sizeof(float));
    if (v) std::memcpy(v, s->v.data(), s->v.size() * sizeof(float));
}
void orc_solid_set_state(orc_solid* s, const float* x, const float* v) {
    if (x) std::memcpy(s->x.data(), x, s->x.size() * sizeof(float));
    if (v) std::memcpy(s->v.data(), v, s->v.size() * sizeof(float));
}

}  // extern "C"

// ================================================================ 2D twin
struct orc_solid2 {
    orc_solid_params p;
    float thickness = 1.0f;
    uint32_t N = 0, T = 0;
    std::vector<float> ref;       // 2N
    std::vector<uint32_t> tri;    // 3T
    std::vector<float> Kb;        // T * 9 blocks * 4, block (i,j) at (3*i+j)*4, row-major 2x2
    std::vector<float> Dinv;      // 4T
    std::vector<float> area;      // T
    float total_area = 0.0f;
    std::vector<float> R, detF;   // 4T, T
    std::vector<float> inv_mass;  // N
    float c[2] = {0, 0};
    std::vector<uint32_t> knc;
    std::vector<float> x, v, b, y, r, d, q;
    std::vector<float> x0;           // 2N previous positions (m_vecPos0)
    std::vector<uint8_t> noc;        // T nodes of collapse (ElementCache::m_NoC), 3 = invalid
    float dapd_l = 1.0f, dapd_nl = 0.0f, dapd_e = 2.0f;
    float acc_time = 0.0f, total_time = 0.0f;
    uint32_t n_fallback = 0;         // last BeginEvaluation's DAPD R = Id count
};

namespace {
struct M2 { float a[2][2]; };
inline M2 m2load(const float* p) { M2 m; m.a[0][0] = p[0]; m.a[0][1] = p[1]; m.a[1][0] = p[2]; m.a[1][1] = p[3]; return m; }
inline void m2store(const M2& m, float* p) { p[0] = m.a[0][0]; p[1] = m.a[0][1]; p[2] = m.a[1][0]; p[3] = m.a[1][1]; }
inline M2 m2mul(const M2& x, const M2& y) {  // Mal/GMat.h:135-145 (accumulate from 0)
    M2 r;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) { float acc = 0.0f; acc += x.a[i][0] * y.a[0][j]; acc += x.a[i][1] * y.a[1][j]; r.a[i][j] = acc; }
    return r;
}
inline void m2vec(const M2& m, const float v[2], float out[2]) {
    for (int i = 0; i < 2; i++) { float acc = 0.0f; acc += m.a[i][0] * v[0]; acc += m.a[i][1] * v[1]; out[i] = acc; }
}
inline float det2(const M2& m) { return m.a[0][0] * m.a[1][1] - m.a[1][0] * m.a[0][1]; }  // Mal/GMat.h:230-233
inline M2 inv2(const M2& m) {  // Mal/GMat.h:260-288
    float id = 1.0f / det2(m);
    M2 r;
    r.a[0][0] = id * m.a[1][1]; r.a[0][1] = -id * m.a[0][1];
    r.a[1][0] = -id * m.a[1][0]; r.a[1][1] = id * m.a[0][0];
    return r;
}
inline M2 d2(const float* r0, const float* r1, const float* r2) {  // TriangleElement2::Compute_D
    M2 D;
    D.a[0][0] = r1[0] - r0[0]; D.a[0][1] = r2[0] - r0[0];
    D.a[1][0] = r1[1] - r0[1]; D.a[1][1] = r2[1] - r0[1];
    return D;
}
// v_i -= R * sum_j K_ij l_j (C_f_e / C_df_x, TriangleElementSet2.cpp:1253-1286,1859-1915)
inline void corot_apply2(const orc_solid2* s, uint32_t e, const float l[3][2], float* out) {
    const uint32_t* t = &s->tri[3 * size_t(e)];
    const float* kb = &s->Kb[size_t(e) * 36];
    M2 R = m2load(&s->R[4 * size_t(e)]);
    for (int i = 0; i < 3; i++) {
        float acc[2], w[2], tmp[2];
        m2vec(m2load(kb + (3 * i + 0) * 4), l[0], acc);
        m2vec(m2load(kb + (3 * i + 1) * 4), l[1], tmp); acc[0] += tmp[0]; acc[1] += tmp[1];
        m2vec(m2load(kb + (3 * i + 2) * 4), l[2], tmp); acc[0] += tmp[0]; acc[1] += tmp[1];
        m2vec(R, acc, w);
        float* o = &out[2 * size_t(t[i])];
        o[0] -= w[0]; o[1] -= w[1];
    }
}
inline M2 m2tr(const M2& m) { M2 t; t.a[0][0] = m.a[0][0]; t.a[0][1] = m.a[1][0]; t.a[1][0] = m.a[0][1]; t.a[1][1] = m.a[1][1]; return t; }

// Mal/GVec.h:193-199: squared norm accumulated from 0, strictly below
// epsilon^2 with epsilon = 1e-4 (Mal/RealUtils.h:110)
inline bool zero2(float a, float b) {
    float acc = 0.0f;
    acc += a * a;
    acc += b * b;
    return acc < 0.0001f * 0.0001f;
}
// Mal/GMatEigen.h:172-184: unit eigenvector of the symmetric m for lambda
// from the first non-negligible row of m - lambda*Id, normalised as
// v * (1 / sqrt(v.v)) (Mal/GVec.h:85,97,103)
inline void eigvec2(const M2& m, float lam, float& x, float& y) {
    const float e00 = m.a[0][0] - lam, e01 = m.a[0][1], e10 = m.a[1][0], e11 = m.a[1][1] - lam;
    if (!zero2(e00, e01)) { x = e01; y = -e00; }
    else if (!zero2(e10, e11)) { x = -e11; y = e10; }
    else { x = 0.0f; y = 0.0f; return; }
    float nsq = 0.0f;
    nsq += x * x;
    nsq += y * y;
    const float inv = 1.0f / std::sqrt(nsq);
    x = inv * x;
    y = inv * y;
}
// Degenerate branch of the 2D Compute_R_SVD (ms/fem/Rotation.cpp:27-36):
// SVD F = U diag(s) V^T (Mal/GMatDecomposition.h:255-345) from the
// eigenvalues of F^T F by the quadratic formula (Mal/GMatEigen.h:151-164),
// U = F V diag(s)^+, null singular directions completed by the CCW
// perpendicular; then U and V^T are made proper rotations and R = U V^T.
// Release-build behaviour: a NaN eigenvalue pair with |det F - 1| >= 0.05 is
// not caught and flows into R. Only the second eigenvalue and the second
// singular value are NaN-tested (Mal/GVec.h:201-208 starts at index 1).
// fix_reflection: Compute_R_SVD's removal of reflections from U and V^T
// (Rotation.cpp:31-33); GRotation2x2_PolarDecomposition_From_SVD returns U V^T
// as it is (Mal/GMatDecomposition.h:365-379).
M2 svd_rotation2_degenerate(const M2& F, bool fix_reflection = true) {
    M2 U, Vt;
    const M2 A = m2mul(m2tr(F), F);
    float tr = 0.0f;
    tr += A.a[0][0];
    tr += A.a[1][1];
    const float b = -tr, c = det2(A);
    const float pm = std::sqrt(b * b - 4.0f * c);
    const float ev0 = 0.5f * (-b + pm), ev1 = 0.5f * (-b - pm);
    bool identity = false;
    if (ev1 != ev1) {
        const float d = det2(F) - 1.0f;
        identity = double(d > 0.0f ? d : -d) < 0.05;
    }
    if (identity) {
        U = M2{{{1.0f, 0.0f}, {0.0f, 1.0f}}};
        Vt = U;
    } else {
        float s0 = std::sqrt(ev0), s1 = std::sqrt(ev1);
        if (s1 != s1) {
            if (ev0 < 0.0f) s0 = 0.0f;
            if (ev1 < 0.0f) s1 = 0.0f;
        }
        const bool z0 = (s0 > 0.0f ? s0 : -s0) <= 0.0001f;  // Mal/RealUtils.h:114
        const bool z1 = (s1 > 0.0f ? s1 : -s1) <= 0.0001f;
        M2 V;
        eigvec2(A, ev0, V.a[0][0], V.a[1][0]);
        eigvec2(A, ev1, V.a[0][1], V.a[1][1]);
        M2 P{{{z0 ? 0.0f : 1.0f / s0, 0.0f}, {0.0f, z1 ? 0.0f : 1.0f / s1}}};
        U = m2mul(m2mul(F, V), P);
        Vt = m2tr(V);
        if (z0 && z1) {
            U = M2{{{1.0f, 0.0f}, {0.0f, 1.0f}}};
            Vt = U;
        } else if (z0) {
            Vt.a[0][0] = Vt.a[1][1]; Vt.a[0][1] = -Vt.a[1][0];
            U.a[0][0] = U.a[1][1]; U.a[1][0] = -U.a[0][1];
        } else if (z1) {
            Vt.a[1][0] = Vt.a[0][1]; Vt.a[1][1] = -Vt.a[0][0];
            U.a[0][1] = U.a[1][0]; U.a[1][1] = -U.a[0][0];
        }
    }
    if (fix_reflection) {
        if (det2(U) < 0.0f) { U.a[0][1] = -U.a[0][1]; U.a[1][1] = -U.a[1][1]; }
        if (det2(Vt) < 0.0f) { Vt.a[1][0] = -Vt.a[1][0]; Vt.a[1][1] = -Vt.a[1][1]; }
    }
    return m2mul(U, Vt);
}

// ---- 2D degenerate triangles (DAPD)
constexpr uint8_t kNocInvalid = 3;  // cInvalidNoC (TriangleElementSet2.h:114)

// ComputeNoC (ms/fem/TriangleElementSet2.cpp:739-835): the time of collapse as
// a root of the quadratic det Ds(t) = thr 2 A (mal::GSolvePolynomialEq2 in
// float, the first positive root, else the second, even above 1), then the
// node of collapse: the vertex opposite the longest edge at that time.
uint8_t compute_noc2(const float a[3][2], const float b[3][2], float area, float thr, float* toc_out = nullptr) {
    float d[3][2];
    for (int i = 0; i < 3; i++) { d[i][0] = b[i][0] - a[i][0]; d[i][1] = b[i][1] - a[i][1]; }
    const float K = thr * 2.0f * area;
    const float A = d[1][0] * d[2][1] + d[2][0] * d[0][1] + d[0][0] * d[1][1] - d[1][0] * d[0][1] -
                    d[0][0] * d[2][1] - d[2][0] * d[1][1];
    const float B = (d[1][0] * a[2][1] + a[1][0] * d[2][1]) + (d[2][0] * a[0][1] + a[2][0] * d[0][1]) +
                    (d[0][0] * a[1][1] + a[0][0] * d[1][1]) - (d[1][0] * a[0][1] + a[1][0] * d[0][1]) -
                    (d[0][0] * a[2][1] + a[0][0] * d[2][1]) - (d[2][0] * a[1][1] + a[2][0] * d[1][1]);
    const float C = a[1][0] * a[2][1] + a[2][0] * a[0][1] + a[0][0] * a[1][1] - a[1][0] * a[0][1] -
                    a[0][0] * a[2][1] - a[2][0] * a[1][1] - K;
    float t1 = 0.0f, t2 = 1.0f;
    if (solve_eq2(A, B, C, t1, t2) <= 0) return kNocInvalid;
    const float toc = (t1 > 0.0f) ? t1 : t2;
    if (toc_out) *toc_out = toc;
    float q[3][2];
    for (int i = 0; i < 3; i++) { q[i][0] = a[i][0] + toc * d[i][0]; q[i][1] = a[i][1] + toc * d[i][1]; }
    auto nsq = [&](int i, int j) {
        const float ex = q[i][0] - q[j][0], ey = q[i][1] - q[j][1];
        float acc = 0.0f;
        acc += ex * ex;
        acc += ey * ey;
        return acc;
    };
    const float l01 = nsq(1, 0), l12 = nsq(2, 1), l20 = nsq(0, 2);
    return (l01 >= l12) ? ((l01 >= l20) ? 2 : 1) : ((l12 >= l20) ? 0 : 1);
}

// Compute_R_DAPD_Degenerate, 2D (ms/fem/Rotation.cpp:300-362, with
// __ENABLE_REGULARIZE_DAPD, :6-8): the collapsing node moves along the unit
// perpendicular of its opposite edge (squared length regularised by 1e-6) by
// lambda_u; R is the closed-form polar factor of the projected F, or
// GRotation2x2_PolarDecomposition_From_SVD when its det is <= 0.001.
M2 r_dapd_degenerate2(float thr, uint8_t noc, const float p[3][2], float area, const M2& invDm, float fL, float fNL,
                      float eNL) {
    float x[3][2];
    std::memcpy(x, p, sizeof(x));
    const int i0 = noc, i1 = (noc + 1) % 3, i2 = (noc + 2) % 3;
    const float ex = x[i2][0] - x[i1][0], ey = x[i2][1] - x[i1][1];
    float dsq = 0.0f;
    dsq += ex * ex;
    dsq += ey * ey;
    dsq += 0.001f * 0.001f;
    const float dist = std::sqrt(dsq);
    const float inv = 1.0f / dist;
    const float ax = inv * -ey, ay = inv * ex;  // PerpendicularCW (Mal/GVec.h:223-226) / dist
    const float dh = thr * 2.0f * area / dist;
    const float mx = 0.5f * (x[i1][0] + x[i2][0]), my = 0.5f * (x[i1][1] + x[i2][1]);
    float w = 0.0f;
    w += (x[i0][0] - mx) * ax;
    w += (x[i0][1] - my) * ay;
    const float s = w / dh;
    const float t = 1.0f - s;
    const float lambda = (w >= 0.0f) ? fL * dh * ((2.0f - t) * t * t) : fL * dh * (t + fNL * std::pow(-s, eNL));
    x[i0][0] = x[i0][0] + lambda * ax;
    x[i0][1] = x[i0][1] + lambda * ay;
    const M2 projF = m2mul(d2(x[0], x[1], x[2]), invDm);
    const float dp = det2(projF);
    float Fa[4], R[4];
    m2store(projF, Fa);
    if (dp > 0.001f || std::fabs(dp - 1.0f) <= 0.0001f) {
        orc_pd2(Fa, dp, R);
        return m2load(R);
    }
    return svd_rotation2_degenerate(projF, false);
}
}  // namespace

extern "C" {

// ms/fem/Rotation.cpp:19-37 (2D Compute_R_SVD)
void orc_svd_rotation2(const float F[4], float detF, float thr, float R[4]) {
    if (detF > thr) {
        orc_pd2(F, detF, R);
        return;
    }
    m2store(svd_rotation2_degenerate(m2load(F)), R);
}

// EditableTriangleElementSet2::EndEdition (TriangleElementSet2.cpp:2187-2239) +
// LeafDSH_Solid2D_FEM::RebuildMass (LeafDSH_Solid2D_FEM.cpp:1842-1863)

// RebuildMass's weighted areas (ds/dsh/LeafDSH_Solid2D_FEM.cpp:1842-1863):
// weighted_area += Rcp(3) * Area(pid) over iterator_polygons_around_vertex_ccw
// (Geo/shape/MeshSolidShape2.h:587-619). Half-edge 3p+k of triangle p runs
// from its vertex k to vertex k+1 (AddPolygonN, in the order the triangles
// were added; baking remaps polygon ids, not half-edges); a vertex's first
// half-edge is its lowest-numbered outgoing one, or its outgoing boundary
// half-edge (RebuildTopology, MeshSolidShape2.cpp:1330-1356); the fan advances
// h <- Sym(Prev(h)) until it reaches the boundary or comes back.
static std::vector<float> fan_weighted_areas(uint32_t n_nodes, uint32_t n_tris, const uint32_t* tris,
                                             const float* area) {
    const uint32_t none = 0xffffffffu;
    const uint32_t nhe = 3 * n_tris;
    auto origin = [&](uint32_t h) { return tris[h]; };
    auto next = [&](uint32_t h) { return 3 * (h / 3) + (h % 3 + 1) % 3; };
    auto prev = [&](uint32_t h) { return 3 * (h / 3) + (h % 3 + 2) % 3; };
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> by_ends;
    for (uint32_t h = 0; h < nhe; h++) by_ends[{origin(h), origin(next(h))}] = h;
    std::vector<uint32_t> sym(nhe, none), out(n_nodes, none);
    for (uint32_t h = 0; h < nhe; h++) {
        auto it = by_ends.find({origin(next(h)), origin(h)});
        if (it != by_ends.end()) sym[h] = it->second;
    }
    std::vector<uint32_t> deg(n_nodes, 0);
    for (uint32_t h = 0; h < nhe; h++) {
        if (out[origin(h)] == none || sym[h] == none) out[origin(h)] = h;
        deg[origin(h)]++;
    }
    const float third = 1.0f / 3.0f;
    std::vector<float> wa(n_nodes, 0.0f);
    for (uint32_t v = 0; v < n_nodes; v++) {
        if (out[v] == none) continue;
        float w = 0.0f;
        uint32_t h = out[v];
        for (uint32_t steps = deg[v]; steps > 0; steps--) {  // a manifold fan ends within its degree
            w += third * area[h / 3];
            const uint32_t nh = sym[prev(h)];
            if (nh == none || nh == out[v]) break;
            h = nh;
        }
        wa[v] = w;
    }
    return wa;
}

orc_solid2* orc_solid2_create(uint32_t n_nodes, const float* ref2, uint32_t n_tris, const uint32_t* tris,
                              const orc_solid_params* p, float thickness) {
    orc_solid2* s = new orc_solid2;
    s->p = *p;
    s->thickness = thickness;
    s->N = n_nodes; s->T = n_tris;
    s->ref.assign(ref2, ref2 + 2 * size_t(n_nodes));
    s->tri.assign(tris, tris + 3 * size_t(n_tris));
    s->Kb.resize(size_t(n_tris) * 36);
    s->Dinv.resize(4 * size_t(n_tris));
    s->area.resize(n_tris);
    s->R.assign(4 * size_t(n_tris), 0.0f);
    s->detF.assign(n_tris, 1.0f);
    float acc = 0.0f, K[36];
    for (uint32_t e = 0; e < n_tris; e++) {
        const float* r0 = &s->ref[2 * size_t(tris[3 * e])];
        const float* r1 = &s->ref[2 * size_t(tris[3 * e + 1])];
        const float* r2 = &s->ref[2 * size_t(tris[3 * e + 2])];
        float a = 0.5f * ((r1[0] - r0[0]) * (r2[1] - r0[1]) - (r2[0] - r0[0]) * (r1[1] - r0[1]));  // Compute_Area
        s->area[e] = a;
        acc += a;
        compute_K2(r0, r1, r2, p->young_modulus, p->poisson_ratio, K);
        float* kb = &s->Kb[size_t(e) * 36];
        for (int bi = 0; bi < 3; bi++)
            for (int bj = 0; bj < 3; bj++)
                for (int rr = 0; rr < 2; rr++)
                    for (int cc = 0; cc < 2; cc++) kb[(3 * bi + bj) * 4 + 2 * rr + cc] = K[(2 * bi + rr) * 6 + 2 * bj + cc];
        m2store(inv2(d2(r0, r1, r2)), &s->Dinv[4 * size_t(e)]);
        s->R[4 * size_t(e)] = 1.0f; s->R[4 * size_t(e) + 3] = 1.0f;  // CreateCache: R = Id
    }
    s->total_area = acc;
    // lumped masses: weighted area Sum(Rcp(3) * area) over the node's triangles
    // in the mesh's CCW fan order
    const float density_per_area = p->density * thickness;
    const std::vector<float> wa = fan_weighted_areas(n_nodes, n_tris, tris, s->area.data());
    s->inv_mass.resize(n_nodes);
    for (uint32_t n = 0; n < n_nodes; n++) s->inv_mass[n] = 1.0f / (wa[n] * density_per_area);  // SetMass stores 1/m
    rayleigh(p->rayleigh_freq1, p->rayleigh_freq2, p->damping_ratio, s->c);
    const size_t n2 = 2 * size_t(n_nodes);
    s->x.assign(ref2, ref2 + n2);
    s->v.assign(n2, 0.0f);
    s->x0 = s->x;                       // m_vecPos0 starts at the rest pose
    s->noc.assign(n_tris, kNocInvalid);  // CreateCache: m_NoC = cInvalidNoC
    s->b.resize(n2); s->y.resize(n2); s->r.resize(n2); s->d.resize(n2); s->q.resize(n2);
    return s;
}

void orc_solid2_destroy(orc_solid2* s) { delete s; }
void orc_solid2_set_knc(orc_solid2* s, uint32_t n, const uint32_t* nids) { s->knc.assign(nids, nids + n); }
float orc_solid2_total_area(const orc_solid2* s) { return s->total_area; }
// The fan-ordered weighted areas for caller-given triangle areas.
void orc_fan_weighted_areas(uint32_t n_nodes, uint32_t n_tris, const uint32_t* tris, const float* area, float* wa) {
    const std::vector<float> w = fan_weighted_areas(n_nodes, n_tris, tris, area);
    std::copy(w.begin(), w.end(), wa);
}
void orc_solid2_node_masses(const orc_solid2* s, float* m) {
    for (uint32_t n = 0; n < s->N; n++) m[n] = 1.0f / s->inv_mass[n];
}

// BeginEvaluation (TriangleElementSet2.cpp:183-262), its UpdateDegeneration
// (:583-640) and ComputeRotations_MSVD/DAPD (:1007-1018,1078-1110):
// Compute_R_PD for det F > threshold (GMatDecomposition.h:17-30); below it
// MSVD takes the 2x2 SVD (Rotation.cpp:19-37) and DAPD the projected rotation
// of the triangle's node of collapse, or R = Id without one (:1100-1105).
// prev_x may be NULL (then x).
void orc_solid2_begin_evaluation(orc_solid2* s, const float* prev_x, const float* x, float* R_out, float* detF_out) {
    if (!prev_x) prev_x = x;
    const float thr = s->p.degenerate_threshold;
    s->n_fallback = 0;
    for (uint32_t e = 0; e < s->T; e++) {
        const uint32_t* t = &s->tri[3 * size_t(e)];
        const M2 Dinv = m2load(&s->Dinv[4 * size_t(e)]);
        M2 F = m2mul(d2(&x[2 * size_t(t[0])], &x[2 * size_t(t[1])], &x[2 * size_t(t[2])]), Dinv);
        float dF = det2(F);
        s->detF[e] = dF;
        const float xs[3][2] = {{x[2 * size_t(t[0])], x[2 * size_t(t[0]) + 1]},
                                {x[2 * size_t(t[1])], x[2 * size_t(t[1]) + 1]},
                                {x[2 * size_t(t[2])], x[2 * size_t(t[2]) + 1]}};
        if (dF > thr) {
            s->noc[e] = kNocInvalid;
        } else {
            const float ps[3][2] = {{prev_x[2 * size_t(t[0])], prev_x[2 * size_t(t[0]) + 1]},
                                    {prev_x[2 * size_t(t[1])], prev_x[2 * size_t(t[1]) + 1]},
                                    {prev_x[2 * size_t(t[2])], prev_x[2 * size_t(t[2]) + 1]}};
            if (det2(m2mul(d2(ps[0], ps[1], ps[2]), Dinv)) > thr) s->noc[e] = compute_noc2(ps, xs, s->area[e], thr);
        }
        float R[4] = {1.0f, 0.0f, 0.0f, 1.0f};
        if (s->p.rm != ORC_RM_ID && dF > thr) {
            float Fa[4];
            m2store(F, Fa);
            orc_pd2(Fa, dF, R);
        } else if (s->p.rm == ORC_RM_MSVD) {
            m2store(svd_rotation2_degenerate(F), R);
        } else if (s->p.rm == ORC_RM_DAPD) {
            if (s->noc[e] != kNocInvalid)
                m2store(r_dapd_degenerate2(thr, s->noc[e], xs, s->area[e], Dinv, s->dapd_l, s->dapd_nl, s->dapd_e), R);
            else
                s->n_fallback++;
        }
        std::memcpy(&s->R[4 * size_t(e)], R, sizeof(R));
    }
    if (R_out) std::memcpy(R_out, s->R.data(), s->R.size() * sizeof(float));
    if (detF_out) std::memcpy(detF_out, s->detF.data(), s->detF.size() * sizeof(float));
}

void orc_solid2_f_e(const orc_solid2* s, const float* x, float* f) {
    for (uint32_t e = 0; e < s->T; e++) {
        const uint32_t* t = &s->tri[3 * size_t(e)];
        M2 Rt = m2tr(m2load(&s->R[4 * size_t(e)]));
        float u[3][2];
        for (int j = 0; j < 3; j++) {
            m2vec(Rt, &x[2 * size_t(t[j])], u[j]);
            u[j][0] -= s->ref[2 * size_t(t[j])]; u[j][1] -= s->ref[2 * size_t(t[j]) + 1];
        }
        corot_apply2(s, e, u, f);
    }
}

void orc_solid2_df_x(const orc_solid2* s, const float* dx, float* df) {
    for (uint32_t e = 0; e < s->T; e++) {
        const uint32_t* t = &s->tri[3 * size_t(e)];
        M2 Rt = m2tr(m2load(&s->R[4 * size_t(e)]));
        float l[3][2];
        for (int j = 0; j < 3; j++) m2vec(Rt, &dx[2 * size_t(t[j])], l[j]);
        corot_apply2(s, e, l, df);
    }
}

// C_V (TriangleElementSet2.cpp:1620-1645)
float orc_solid2_V(const orc_solid2* s, const float* x) {
    float acc2 = 0.0f;
    for (uint32_t e = 0; e < s->T; e++) {
        const uint32_t* t = &s->tri[3 * size_t(e)];
        const float* kb = &s->Kb[size_t(e) * 36];
        M2 Rt = m2tr(m2load(&s->R[4 * size_t(e)]));
        float u[3][2];
        for (int j = 0; j < 3; j++) {
            m2vec(Rt, &x[2 * size_t(t[j])], u[j]);
            u[j][0] -= s->ref[2 * size_t(t[j])]; u[j][1] -= s->ref[2 * size_t(t[j]) + 1];
        }
        for (int i = 0; i < 3; i++) {
            float acc[2], tmp[2];
            m2vec(m2load(kb + (3 * i + 0) * 4), u[0], acc);
            m2vec(m2load(kb + (3 * i + 1) * 4), u[1], tmp); acc[0] += tmp[0]; acc[1] += tmp[1];
            m2vec(m2load(kb + (3 * i + 2) * 4), u[2], tmp); acc[0] += tmp[0]; acc[1] += tmp[1];
            float dd = 0.0f; dd += u[i][0] * acc[0]; dd += u[i][1] * acc[1];
            acc2 += dd;
        }
    }
    return 0.5f * acc2;
}

// FixedStep with Step_QuasiImplicitEuler_v (LeafDSH_Solid2D_FEM.cpp:1287-1403)
void orc_solid2_fixed_step(orc_solid2* s, float dt, orc_step_stats* st) {
    const uint32_t N = s->N;
    const size_t n2 = 2 * size_t(N);
    orc_solid2_begin_evaluation(s, s->x0.data(), s->x.data(), nullptr, nullptr);  // FixedStep :974
    s->x0 = s->x;
    uint32_t ndeg = 0; float sdeg = 0.0f, min_det = 1e30f;
    for (uint32_t e = 0; e < s->T; e++) {
        if (s->detF[e] <= 0) { ndeg++; sdeg += s->detF[e]; }
        min_det = std::min(min_det, s->detF[e]);
    }
    const float inv_dt = 1.0f / dt;
    for (uint32_t i = 0; i < N; i++) {
        const float m = 1.0f / s->inv_mass[i];  // GetMass
        s->b[2 * i] = 0.0f + m * s->p.gravity[0];
        s->b[2 * i + 1] = 0.0f + m * s->p.gravity[1];
    }
    orc_solid2_f_e(s, s->x.data(), s->b.data());
    for (size_t i = 0; i < n2; i++) s->b[i] *= dt;
    for (uint32_t i = 0; i < N; i++) {
        const float m = 1.0f / s->inv_mass[i];
        s->b[2 * i] += m * s->v[2 * i];
        s->b[2 * i + 1] += m * s->v[2 * i + 1];
    }
    if (s->p.warm_start) s->y = s->v; else std::fill(s->y.begin(), s->y.end(), 0.0f);
    for (uint32_t nid : s->knc) {  // KNC (:1352-1372): position constraints at the rest pose
        float vk[2] = {(s->ref[2 * size_t(nid)] - s->x[2 * size_t(nid)]) * inv_dt,
                       (s->ref[2 * size_t(nid) + 1] - s->x[2 * size_t(nid) + 1]) * inv_dt};
        float nk = std::sqrt(vk[0] * vk[0] + vk[1] * vk[1]);
        float kv[2] = {0.0f, 0.0f};
        if (nk > 0.001f) {
            float inv = 1.0f / nk, sc = std::min(nk, 50.0f);
            kv[0] = sc * (vk[0] * inv); kv[1] = sc * (vk[1] * inv);
        }
        s->b[2 * size_t(nid)] = kv[0]; s->b[2 * size_t(nid) + 1] = kv[1];
        s->y[2 * size_t(nid)] = kv[0]; s->y[2 * size_t(nid) + 1] = kv[1];
    }
    const float fM = 1.0f + dt * s->c[0];
    const float fK = dt * dt + dt * s->c[1];
    auto Av = [&](float* q, const float* d) {  // LS_A_times_d (:1101-1132)
        std::memset(q, 0, n2 * sizeof(float));
        orc_solid2_df_x(s, d, q);
        for (size_t i = 0; i < n2; i++) q[i] *= -fK;
        for (uint32_t i = 0; i < N; i++) {
            float factor = (1.0f / s->inv_mass[i]) * fM;
            q[2 * i] += factor * d[2 * i];
            q[2 * i + 1] += factor * d[2 * i + 1];
        }
    };
    auto S = [&](float* w) {
        for (uint32_t nid : s->knc) { w[2 * size_t(nid)] = 0; w[2 * size_t(nid) + 1] = 0; }
    };
    int32_t it = 0; float rr = 0, ar = 0;
    cg_solve(uint32_t(n2), Av, S, s->y.data(), s->b.data(), s->p.cg_rel_epsilon, s->p.cg_abs_epsilon,
             s->p.cg_max_iter, s->r, s->d, s->q, &it, &rr, &ar, s->p.reduce_fp64 != 0);
    for (size_t i = 0; i < n2; i++) {
        s->v[i] = s->y[i];
        s->x[i] = s->x[i] + dt * s->y[i];
    }
    float V = orc_solid2_V(s, s->x.data());
    if (st) {
        st->num_fixed_steps = 1;
        st->cg_iters = it; st->cg_rel = rr; st->cg_abs = ar;
        st->num_degenerate = ndeg; st->sum_degenerate_detF = sdeg;
        st->elastic_energy = V; st->min_detF = min_det;
    }
}

// Step (LeafDSH_Solid2D_FEM.cpp:264-294): the 3D leaf's accumulator
void orc_solid2_step(orc_solid2* s, float dt, orc_step_stats* st) {
    const float scaled_dt = s->p.time_scale * dt;
    s->total_time += scaled_dt;
    s->acc_time += scaled_dt;
    const float scaled_fixed_dt = s->p.time_scale * s->p.fixed_dt;
    orc_step_stats last;
    std::memset(&last, 0, sizeof(last));
    uint32_t n = 0;
    while (s->acc_time > scaled_fixed_dt) {
        orc_solid2_fixed_step(s, scaled_fixed_dt, &last);
        s->acc_time -= scaled_fixed_dt;
        n++;
    }
    last.num_fixed_steps = n;
    if (st) *st = last;
}

float orc_solid2_total_time(const orc_solid2* s) { return s->total_time; }

void orc_solids2_fixed_step_parallel(orc_solid2* const* solids, uint32_t n, uint32_t threads) {
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        for (uint32_t i = next++; i < n; i = next++) orc_solid2_fixed_step(solids[i], solids[i]->p.fixed_dt, nullptr);
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < std::max(1u, threads); t++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
}

void orc_solid2_get_state(const orc_solid2* s, float* x, float* v) {
    if (x) std::memcpy(x, s->x.data(), s->x.size() * sizeof(float));
    if (v) std::memcpy(v, s->v.data(), s->v.size() * sizeof(float));
}
void orc_solid2_set_state(orc_solid2* s, const float* x, const float* v) {
    if (x) std::memcpy(s->x.data(), x, s->x.size() * sizeof(float));
    if (v) std::memcpy(s->v.data(), v, s->v.size() * sizeof(float));
}

void orc_solid2_get_noc(const orc_solid2* s, uint8_t* noc) { std::memcpy(noc, s->noc.data(), s->noc.size()); }
void orc_solid2_set_dapd(orc_solid2* s, float l_factor, float nl_factor, float nl_exponent) {
    s->dapd_l = l_factor; s->dapd_nl = nl_factor; s->dapd_e = nl_exponent;
}
uint32_t orc_solid2_num_fallback(const orc_solid2* s) { return s->n_fallback; }
uint8_t orc_compute_noc2(const float a[6], const float b[6], float area, float thr, float* toc) {
    const float av[3][2] = {{a[0], a[1]}, {a[2], a[3]}, {a[4], a[5]}};
    const float bv[3][2] = {{b[0], b[1]}, {b[2], b[3]}, {b[4], b[5]}};
    return compute_noc2(av, bv, area, thr, toc);
}
void orc_r_dapd_degenerate2(float thr, uint8_t noc, const float p[6], float area, const float inv_dm[4],
                            float l_factor, float nl_factor, float nl_exponent, float R[4]) {
    const float pv[3][2] = {{p[0], p[1]}, {p[2], p[3]}, {p[4], p[5]}};
    m2store(r_dapd_degenerate2(thr, noc, pv, area, m2load(inv_dm), l_factor, nl_factor, nl_exponent), R);
}

}  // extern "C"
