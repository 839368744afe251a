// Thin extern "C" drivers over the reference's own header-only math and
// ms/fem/Rotation.cpp, compiled in place from /root/reference/src by
// oracle/Makefile into oracle/_ref/. Used only to pin the oracle (tests and
// tests/golden/gen_golden.py); never linked into the shipped library.
//
// Reference entry points driven here:
//   S2::ms::fem::TetrahedronElement3::Compute_K/Compute_Volume  (ms/fem/TetrahedronElement3.h:85-173)
//   S2::ms::fem::TriangleElement2::Compute_K                    (ms/fem/TriangleElement2.h:123-139)
//   mal::Det / mal::Inverse                                     (Mal/GMat.h:236-313)
//   mal::GRotation3x3_PolarDecomposition                        (Mal/GMatDecomposition.h:411-432)
//   mal::GRotation2x2_PolarDecomposition                        (Mal/GMatDecomposition.h:17-30)
//   mal::GSingularValueDecomposition_USVt (PhysBAM 3x3)         (Mal/GMatDecomposition.h:351-361)
//   S2::ms::fem::Compute_R_SVD (2D and 3D)                      (ms/fem/Rotation.cpp:19-57)
//   S2::ns::real_array::dot / sq_norm_2 (SSE4.1)                (ns/real_array.h:34-49)
//   the leaves' Rayleigh coefficients, mal::Inverse(Omega) * (2*zeta) on Mal's
//   Mat2x2 / Vec2 (LeafDSH_Solid3D_FEM.h:347-354; the leaf header itself needs
//   boost, so the four statements are driven here on the reference's types)
#include <cstring>
#include <cstdint>
#include <cstdlib>
#include <Saphyre2/ms/fem/TetrahedronElement3.h>
#include <Saphyre2/ms/fem/TriangleElement2.h>
#include <Saphyre2/ms/fem/Rotation.h>
#include <Mal/GMatDecomposition.h>
#include <Saphyre2/ns/real_array.h>

using namespace S2;

namespace {
Mat3x3 load3(const float* p) {
    Mat3x3 m;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) m(i, j) = p[3 * i + j];
    return m;
}
void store3(const Mat3x3& m, float* p) {
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) p[3 * i + j] = m(i, j);
}
Mat2x2 load2(const float* p) {
    Mat2x2 m;
    m(0, 0) = p[0]; m(0, 1) = p[1]; m(1, 0) = p[2]; m(1, 1) = p[3];
    return m;
}
void store2(const Mat2x2& m, float* p) { p[0] = m(0, 0); p[1] = m(0, 1); p[2] = m(1, 0); p[3] = m(1, 1); }
}  // namespace

extern "C" {

float ref_compute_K(const float r[12], float E, float nu, float K[144]) {
    Vec3 r0(r[0], r[1], r[2]), r1(r[3], r[4], r[5]), r2(r[6], r[7], r[8]), r3(r[9], r[10], r[11]);
    Mat12x12 Km;
    float v = ms::fem::TetrahedronElement3::Compute_K(r0, r1, r2, r3, E, nu, Km);
    for (int i = 0; i < 12; i++) for (int j = 0; j < 12; j++) K[12 * i + j] = Km(i, j);
    return v;
}

float ref_compute_volume(const float r[12]) {
    Vec3 r0(r[0], r[1], r[2]), r1(r[3], r[4], r[5]), r2(r[6], r[7], r[8]), r3(r[9], r[10], r[11]);
    return ms::fem::TetrahedronElement3::Compute_Volume(r0, r1, r2, r3);
}

void ref_inv_dm(const float r[12], float out[9]) {
    Vec3 r0(r[0], r[1], r[2]), r1(r[3], r[4], r[5]), r2(r[6], r[7], r[8]), r3(r[9], r[10], r[11]);
    Mat3x3 Dm;
    ms::fem::TetrahedronElement3::Compute_D(r0, r1, r2, r3, Dm);
    store3(mal::Inverse(Dm), out);
}

float ref_compute_K2(const float r[6], float E, float nu, float K[36]) {
    Vec2 r0(r[0], r[1]), r1(r[2], r[3]), r2(r[4], r[5]);
    Mat6x6 Km;
    float a = ms::fem::TriangleElement2::Compute_K(r0, r1, r2, E, nu, Km);
    for (int i = 0; i < 6; i++) for (int j = 0; j < 6; j++) K[6 * i + j] = Km(i, j);
    return a;
}

float ref_det3(const float m[9]) { return mal::Det(load3(m)); }
void ref_inverse3(const float m[9], float out[9]) { store3(mal::Inverse(load3(m)), out); }

void ref_pd3(const float F[9], float detF, float R[9]) {
    store3(mal::GRotation3x3_PolarDecomposition(load3(F), detF), R);
}
void ref_svd3(const float F[9], float U[9], float s[3], float Vt[9]) {
    Mat3x3 u, vt;
    Vec3 sv;
    mal::GSingularValueDecomposition_USVt(load3(F), u, sv, vt);
    store3(u, U);
    store3(vt, Vt);
    s[0] = sv[0]; s[1] = sv[1]; s[2] = sv[2];
}
void ref_R_svd3(const float F[9], float detF, float thr, float R[9]) {
    store3(ms::fem::Compute_R_SVD(load3(F), detF, thr), R);
}
void ref_pd2(const float F[4], float detF, float R[4]) {
    store2(mal::GRotation2x2_PolarDecomposition(load2(F), detF), R);
}
void ref_R_svd2(const float F[4], float detF, float thr, float R[4]) {
    store2(ms::fem::Compute_R_SVD(load2(F), detF, thr), R);
}

// SSE reductions over a 16-byte aligned copy whose padding lanes are zero.
float ref_dot(const float* a, const float* b, uint32_t n) {
    size_t an = 4 * ((n + 3) / 4);
    float* pa = ns::real_array::alloc(an);
    float* pb = ns::real_array::alloc(an);
    std::memset(pa, 0, an * sizeof(float));
    std::memset(pb, 0, an * sizeof(float));
    std::memcpy(pa, a, n * sizeof(float));
    std::memcpy(pb, b, n * sizeof(float));
    float r = ns::real_array::dot(pa, pb, n);
    ns::real_array::dealloc(pa);
    ns::real_array::dealloc(pb);
    return r;
}

void ref_rayleigh(float f1, float f2, float zeta, float c[2]) {
    Mat2x2 Omega;
    Omega(0, 0) = mal::Rcp(f1); Omega(0, 1) = f1;
    Omega(1, 0) = mal::Rcp(f2); Omega(1, 1) = f2;
    Vec2 z(zeta, zeta);
    Vec2 r = mal::Inverse(Omega) * (2 * z);
    c[0] = r[0];
    c[1] = r[1];
}

int ref_has_sse_path(void) {
#ifdef __S2_NS_ENABLE_SIMD
    return 1;
#else
    return 0;
#endif
}

}  // extern "C"
