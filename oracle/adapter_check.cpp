// Runs the reference-side IForce3 adapter of INTEGRATION.md section 1 (written
// to _ref/TetrahedronElementSet3_MI355X.h by extract_adapter.py and compiled
// against the reference's own ms/fem/TetrahedronElementSet3.h and Params.h) the
// way LeafDSH_Solid3D_FEM drives EditableTetrahedronElementSet3 (BeginEdition,
// AddElement per tet, EndEdition, CreateCache(inv_mass), BeginEvaluation with
// host arrays, f_e / df_x accumulating into host arrays, V, f_d), and compares
// every result with the same operators called through the bulk C-ABI on device
// arrays. Prints "adapter ok" and returns 0 when all are bit-identical; returns
// 2 without a device. Test infrastructure: built by oracle/Makefile.
#include "TetrahedronElementSet3_MI355X.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

using S2::Vec3;
using S2::ms::fem::TetrahedronElementSet3_MI355X;

namespace {
uint32_t g_state = 12345u;
float jitter(float scale) {  // deterministic LCG in [-scale, scale)
    g_state = g_state * 1664525u + 1013904223u;
    return scale * (float(g_state >> 8) / float(1u << 24) * 2.0f - 1.0f);
}
bool same(const std::vector<float>& a, const std::vector<float>& b, const char* what) {
    if (a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(float)) == 0) return true;
    std::printf("adapter mismatch: %s\n", what);
    return false;
}
int ok(int rc, const char* what) {
    if (rc != CF_OK) std::printf("C-ABI call failed: %s (%s)\n", what, cf_last_error());
    return rc == CF_OK;
}
}  // namespace

int main() {
    int ndev = 0;
    if (cf_device_count(&ndev) != CF_OK || ndev < 1) {
        std::printf("no device\n");
        return 2;
    }
    uint32_t N = 0, T = 0;
    cf_mesh_box_counts(5, 4, 6, &N, &T);
    std::vector<float> ref(3 * N);
    std::vector<uint32_t> tets(4 * T);
    const float half[3] = {0.02f, 0.015f, 0.025f};
    cf_mesh_box(half, 5, 4, 6, ref.data(), tets.data());
    std::vector<float> x(ref), prev(ref), v(3 * N), dx(3 * N), inv_mass(N);
    for (uint32_t i = 0; i < 3 * N; i++) {
        x[i] += jitter(2e-3f);
        prev[i] += jitter(1e-3f);
        v[i] = jitter(0.1f);
        dx[i] = jitter(1.0f);
    }
    for (uint32_t i = 0; i < N; i++) inv_mass[i] = 1.0f / (0.01f + 0.001f * float(i % 7));

    // the reference-side adapter, as the leaf uses EditableTetrahedronElementSet3
    S2::ms::fem::Params params;
    params.m_YoungModulus = 1e5f;
    params.m_PoissonRatio = 0.3f;
    params.m_RM = S2::ms::fem::Params::eRM_DAPD;
    TetrahedronElementSet3_MI355X tes;
    tes.BeginEdition(N, reinterpret_cast<const Vec3*>(ref.data()), T, params);
    for (uint32_t e = 0; e < T; e++) tes.AddElement(tets[4 * e], tets[4 * e + 1], tets[4 * e + 2], tets[4 * e + 3]);
    if (!tes.EndEdition()) { std::printf("EndEdition failed\n"); return 1; }
    S2::ms::IForce3::ICache* cache = tes.CreateCache(inv_mass.data());
    tes.BeginEvaluation(cache, reinterpret_cast<const Vec3*>(prev.data()), reinterpret_cast<const Vec3*>(x.data()),
                        reinterpret_cast<const Vec3*>(v.data()), 0.0f);
    std::vector<float> fa(3 * N, 0.25f), dfa(3 * N, -0.5f), fda(3 * N, 0.0f);
    tes.f_e(cache, reinterpret_cast<Vec3*>(fa.data()));
    tes.df_x(cache, reinterpret_cast<const Vec3*>(dx.data()), reinterpret_cast<Vec3*>(dfa.data()));
    tes.df_x(cache, reinterpret_cast<const Vec3*>(dx.data()), reinterpret_cast<Vec3*>(dfa.data()));
    tes.f_d(cache, reinterpret_cast<Vec3*>(fda.data()));
    const float Va = tes.V(cache);
    tes.EndEvaluation(cache);
    tes.DestroyCache(cache);

    // the same through the bulk C-ABI on device arrays
    cf_fem_params q;
    cf_default_fem_params(&q);
    q.young_modulus = 1e5f;
    q.poisson_ratio = 0.3f;
    q.rm = CF_RM_DAPD;
    cf_tetset* set = nullptr;
    cf_cache* c = nullptr;
    const size_t bytes = 3 * sizeof(float) * N;
    float *d_x = nullptr, *d_prev = nullptr, *d_v = nullptr, *d_dx = nullptr, *d_f = nullptr, *d_df = nullptr,
          *d_fd = nullptr;
    std::vector<float> fb(3 * N, 0.25f), dfb(3 * N, -0.5f), fdb(3 * N, 0.0f);
    double Vb = 0.0;
    int good = ok(cf_tetset_create(N, ref.data(), T, tets.data(), &q, &set), "cf_tetset_create") &&
               ok(cf_cache_create(set, inv_mass.data(), &c), "cf_cache_create");
    for (float** p : {&d_x, &d_prev, &d_v, &d_dx, &d_f, &d_df, &d_fd})
        good = good && ok(cf_device_alloc(bytes, reinterpret_cast<void**>(p)), "cf_device_alloc");
    good = good && ok(cf_memcpy_htod(d_x, x.data(), bytes), "htod") && ok(cf_memcpy_htod(d_prev, prev.data(), bytes), "htod") &&
           ok(cf_memcpy_htod(d_v, v.data(), bytes), "htod") && ok(cf_memcpy_htod(d_dx, dx.data(), bytes), "htod") &&
           ok(cf_memcpy_htod(d_f, fb.data(), bytes), "htod") && ok(cf_memcpy_htod(d_df, dfb.data(), bytes), "htod") &&
           ok(cf_memcpy_htod(d_fd, fdb.data(), bytes), "htod") &&
           ok(cf_begin_evaluation(c, d_prev, d_x, d_v, 0.0f), "cf_begin_evaluation") &&
           ok(cf_f_e(c, d_f), "cf_f_e") && ok(cf_df_x(c, d_dx, d_df), "cf_df_x") && ok(cf_df_x(c, d_dx, d_df), "cf_df_x") &&
           ok(cf_f_d(c, d_fd), "cf_f_d") && ok(cf_V(c, &Vb), "cf_V") && ok(cf_end_evaluation(c), "cf_end_evaluation") &&
           ok(cf_memcpy_dtoh(fb.data(), d_f, bytes), "dtoh") && ok(cf_memcpy_dtoh(dfb.data(), d_df, bytes), "dtoh") &&
           ok(cf_memcpy_dtoh(fdb.data(), d_fd, bytes), "dtoh");
    for (float* p : {d_x, d_prev, d_v, d_dx, d_f, d_df, d_fd})
        if (p) cf_device_free(p);
    if (c) cf_cache_destroy(c);
    if (set) cf_tetset_destroy(set);
    if (!good) return 1;
    good = same(fa, fb, "f_e") && same(dfa, dfb, "df_x") && same(fda, fdb, "f_d");
    if (good && Va != float(Vb)) { std::printf("adapter mismatch: V\n"); good = false; }
    if (good) std::printf("adapter ok: %u nodes, %u tets, V %.9g, total volume %.9g\n", N, T, double(Va),
                          double(tes.TotalVolume()));
    return good ? 0 : 1;
}
