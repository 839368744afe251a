// 2D solids (cf_batch2_*, cf_solid2_* in include/corofem.h): host-side mesh
// preparation for the workgroup-resident step kernel (cf_kernels2d.hip).
//
// Per mesh, once: element precompute in the reference's FP32 order (area,
// Dm^-1: ms/fem/TriangleElement2.h:71-96, Mal/GMat.h:260-288), lumped masses
// (LeafDSH_Solid2D_FEM::RebuildMass, ds/dsh/LeafDSH_Solid2D_FEM.cpp:1842-1863,
// summed over the node's triangles in ascending triangle order), a greedy
// triangle colouring, and the triangles sorted by colour.
#include "cf_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

using namespace cf;

struct cf_batch2 {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<Mesh2Dev> meshes;
    std::vector<float2> ref, x, v;
    std::vector<float> mass, area;
    std::vector<uint8_t> knc;
    std::vector<uint32_t> tri_ids;
    std::vector<float4> dinv;
    std::vector<uint16_t> cstart;
    std::vector<float2> x0;          // previous positions (DAPD node-of-collapse tracking)
    std::vector<uint8_t> noc;        // per triangle (colour order), 3 = none
    std::vector<uint32_t> tri_src;   // colour order -> the caller's triangle index
    std::vector<Work2> work;
    uint64_t tris = 0;
    bool dirty = true;        // static data or state must be uploaded
    bool on_device = false;   // the device holds the newest state
    // device copies
    Mesh2Dev* d_meshes = nullptr;
    float2 *d_ref = nullptr, *d_x = nullptr, *d_v = nullptr;
    float *d_mass = nullptr, *d_area = nullptr;
    uint8_t* d_knc = nullptr;
    uint32_t* d_tri = nullptr;
    float4* d_dinv = nullptr;
    uint16_t* d_cs = nullptr;
    float2* d_x0 = nullptr;
    uint8_t* d_noc = nullptr;
    Work2* d_work = nullptr;
    std::vector<int32_t> last_iters;  // each mesh's CG iterations in the last step (empty before the first)
    Sim2Out* d_out = nullptr;
    std::vector<Sim2Out> h_out;
    int32_t* d_flag = nullptr;  // a mesh's CG met a NaN or Inf
};

namespace {

constexpr uint32_t kNocNone = 3;  // cInvalidNoC (TriangleElementSet2.h:114)

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void free_device(cf_batch2* b) {
    dfree(b->d_meshes); dfree(b->d_ref); dfree(b->d_x); dfree(b->d_v); dfree(b->d_mass); dfree(b->d_area);
    dfree(b->d_knc); dfree(b->d_tri); dfree(b->d_dinv); dfree(b->d_cs); dfree(b->d_work); dfree(b->d_out);
    dfree(b->d_x0); dfree(b->d_noc); dfree(b->d_flag);
}

template <class T>
int put(T** d, const std::vector<T>& h, hipStream_t s) {
    CF_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(d), std::max<size_t>(1, h.size()) * sizeof(T)));
    if (!h.empty()) CF_CHECK_HIP(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return CF_OK;
}

// Pulls the device state into the host mirrors (before a re-upload).
int pull_state(cf_batch2* b) {
    if (!b->on_device) return CF_OK;
    CF_CHECK_HIP(hipMemcpyAsync(b->x.data(), b->d_x, b->x.size() * sizeof(float2), hipMemcpyDeviceToHost, b->stream));
    CF_CHECK_HIP(hipMemcpyAsync(b->v.data(), b->d_v, b->v.size() * sizeof(float2), hipMemcpyDeviceToHost, b->stream));
    CF_CHECK_HIP(hipMemcpyAsync(b->x0.data(), b->d_x0, b->x0.size() * sizeof(float2), hipMemcpyDeviceToHost, b->stream));
    CF_CHECK_HIP(hipMemcpyAsync(b->noc.data(), b->d_noc, b->noc.size(), hipMemcpyDeviceToHost, b->stream));
    CF_CHECK_HIP(hipStreamSynchronize(b->stream));
    b->on_device = false;
    return CF_OK;
}

// Work list: meshes of one group size share workgroups (8/gw meshes each).
// A workgroup holds its LDS until its slowest mesh's CG has stopped, so once a
// step has run, meshes are grouped by the iteration count of their last step
// (most first; the counts change little from one step to the next), then by
// size; before that by size alone. Largest first, so that the tail is short.
// A mesh's result does not depend on its group.
void build_work(cf_batch2* b) {
    std::vector<uint32_t> cls[4];
    auto slot = [](uint32_t gw) { return gw == 1 ? 0 : gw == 2 ? 1 : gw == 4 ? 2 : 3; };
    for (uint32_t i = 0; i < b->meshes.size(); i++) cls[slot(b->meshes[i].gw)].push_back(i);
    const bool by_iters = b->last_iters.size() == b->meshes.size();
    auto by_size = [&](uint32_t p, uint32_t q) {
        return b->meshes[p].T != b->meshes[q].T ? b->meshes[p].T > b->meshes[q].T : p < q;
    };
    // counting sort on the iteration count (most first), stable: O(meshes)
    std::vector<uint32_t> count, sorted;
    auto sort_by_iters = [&](std::vector<uint32_t>& v) {
        count.assign(kMaxCgIter + 2, 0);
        for (uint32_t i : v) count[kMaxCgIter + 1 - std::min<uint32_t>(uint32_t(b->last_iters[i]), kMaxCgIter + 1)]++;
        uint32_t acc = 0;
        for (uint32_t& c : count) { const uint32_t t = c; c = acc; acc += t; }
        sorted.resize(v.size());
        for (uint32_t i : v) sorted[count[kMaxCgIter + 1 - std::min<uint32_t>(uint32_t(b->last_iters[i]), kMaxCgIter + 1)]++] = i;
        v.swap(sorted);
    };
    b->work.clear();
    for (int c = 3; c >= 0; c--) {
        if (by_iters) sort_by_iters(cls[c]);
        else std::stable_sort(cls[c].begin(), cls[c].end(), by_size);
        const uint32_t gw = 1u << c, per = 8 / gw;
        for (size_t i = 0; i < cls[c].size(); i += per) {
            Work2 w{};
            w.gw = gw;
            w.cmax = 0;
            for (uint32_t j = 0; j < 8; j++) w.sim[j] = kIdleSlot;
            for (uint32_t j = 0; j < per && i + j < cls[c].size(); j++) {
                w.sim[j] = cls[c][i + j];
                w.cmax = std::max(w.cmax, b->meshes[cls[c][i + j]].C);
            }
            b->work.push_back(w);
        }
    }
}

int upload(cf_batch2* b) {
    CF_TRY(pull_state(b));
    free_device(b);
    build_work(b);
    hipStream_t s = b->stream;
    CF_TRY(put(&b->d_meshes, b->meshes, s));
    CF_TRY(put(&b->d_ref, b->ref, s));
    CF_TRY(put(&b->d_x, b->x, s));
    CF_TRY(put(&b->d_v, b->v, s));
    CF_TRY(put(&b->d_mass, b->mass, s));
    CF_TRY(put(&b->d_area, b->area, s));
    CF_TRY(put(&b->d_knc, b->knc, s));
    CF_TRY(put(&b->d_tri, b->tri_ids, s));
    CF_TRY(put(&b->d_dinv, b->dinv, s));
    CF_TRY(put(&b->d_cs, b->cstart, s));
    CF_TRY(put(&b->d_x0, b->x0, s));
    CF_TRY(put(&b->d_noc, b->noc, s));
    CF_TRY(put(&b->d_work, b->work, s));
    b->h_out.assign(b->meshes.size(), Sim2Out{});
    CF_TRY(put(&b->d_out, b->h_out, s));
    CF_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(&b->d_flag), sizeof(int32_t)));
    CF_CHECK_HIP(hipStreamSynchronize(s));
    b->dirty = false;
    b->on_device = true;
    return CF_OK;
}

Batch2Args args(const cf_batch2* b) {
    Batch2Args a;
    a.work = b->d_work;
    a.meshes = b->d_meshes;
    a.ref = b->d_ref;
    a.mass = b->d_mass;
    a.knc = b->d_knc;
    a.x = b->d_x;
    a.v = b->d_v;
    a.tri_ids = b->d_tri;
    a.dinv = b->d_dinv;
    a.area = b->d_area;
    a.color_start = b->d_cs;
    a.out = b->d_out;
    a.x0 = b->d_x0;
    a.noc = b->d_noc;
    a.nonfinite = b->d_flag;
    return a;
}

Material material2(float E, float nu) {
    // TriangleElement2::Compute_E (ms/fem/TriangleElement2.h:121-129), plane strain
    const float a = E / ((1 + nu) * (1 - 2 * nu));
    const float anu = a * nu;
    Material m;
    m.a_k1 = a - anu;
    m.a_nu = anu;
    m.a_k2 = 0.5f * a - anu;
    return m;
}

}  // namespace

extern "C" {

int cf_rotations2(const float* h_F, const float* h_detF, uint32_t n, int32_t rm, float threshold, float* h_R,
                  uint32_t* h_fallback) {
    using namespace cf;
    if (n && (!h_F || !h_detF || !h_R)) {
        set_error("cf_rotations2: null array");
        return CF_ERR_INVALID_ARG;
    }
    if (rm != CF_RM_ID && rm != CF_RM_MSVD && rm != CF_RM_DAPD) {
        set_error("cf_rotations2: rotation method %d is not supported", rm);
        return CF_ERR_UNSUPPORTED;
    }
    if (!n) return CF_OK;
    float *d_F = nullptr, *d_det = nullptr, *d_R = nullptr;
    uint32_t* d_fb = nullptr;
    struct Free {
        void* p[4];
        ~Free() {
            for (void* q : p)
                if (q) (void)hipFree(q);
        }
    } guard{{nullptr, nullptr, nullptr, nullptr}};
    CF_CHECK_HIP(hipMalloc(&d_F, size_t(n) * 16));
    guard.p[0] = d_F;
    CF_CHECK_HIP(hipMalloc(&d_det, size_t(n) * 4));
    guard.p[1] = d_det;
    CF_CHECK_HIP(hipMalloc(&d_R, size_t(n) * 16));
    guard.p[2] = d_R;
    CF_CHECK_HIP(hipMalloc(&d_fb, size_t(n) * 4));
    guard.p[3] = d_fb;
    CF_CHECK_HIP(hipMemcpy(d_F, h_F, size_t(n) * 16, hipMemcpyHostToDevice));
    CF_CHECK_HIP(hipMemcpy(d_det, h_detF, size_t(n) * 4, hipMemcpyHostToDevice));
    CF_TRY(launch_rotations2(d_F, d_det, n, rm, threshold, d_R, d_fb, nullptr));
    CF_CHECK_HIP(hipDeviceSynchronize());
    CF_CHECK_HIP(hipMemcpy(h_R, d_R, size_t(n) * 16, hipMemcpyDeviceToHost));
    if (h_fallback) CF_CHECK_HIP(hipMemcpy(h_fallback, d_fb, size_t(n) * 4, hipMemcpyDeviceToHost));
    return CF_OK;
}

int cf_batch2_create(cf_batch2** out) {
    if (!out) return CF_ERR_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        (void)hipGetLastError();
        set_error("no HIP device");
        return CF_ERR_NO_DEVICE;
    }
    cf_batch2* b = new (std::nothrow) cf_batch2;
    if (!b) return CF_ERR_OUT_OF_MEMORY;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        set_error("2D batch: stream/event creation failed");
        cf_batch2_destroy(b);
        return CF_ERR_HIP;
    }
    *out = b;
    return CF_OK;
}

int cf_batch2_destroy(cf_batch2* b) {
    if (!b) return CF_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    free_device(b);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return CF_OK;
}


// RebuildMass's weighted areas (ds/dsh/LeafDSH_Solid2D_FEM.cpp:1842-1863):
// Rcp(3) * area summed over iterator_polygons_around_vertex_ccw
// (Geo/shape/MeshSolidShape2.h:587-619). Half-edge 3p+k of triangle p runs
// from its vertex k to vertex k+1, numbered in the order the triangles are
// given (the reference's AddPolygonN order; its baking remaps polygon ids but
// keeps half-edges); a vertex's fan starts at its lowest-numbered outgoing
// half-edge, or at its outgoing boundary half-edge (RebuildTopology,
// MeshSolidShape2.cpp:1330-1356), and advances h <- Sym(Prev(h)) until the
// boundary or the start, at most once per outgoing half-edge.
static void fan_weighted_areas(uint32_t N, uint32_t T, const uint32_t* tris, const float* area,
                               std::vector<float>& wa) {
    const uint32_t none = 0xffffffffu, nhe = 3 * T;
    auto next = [](uint32_t h) { return h - h % 3 + (h % 3 + 1) % 3; };
    auto prev = [](uint32_t h) { return h - h % 3 + (h % 3 + 2) % 3; };
    std::vector<uint32_t> first(N + 1, 0), outs(nhe);  // outgoing half-edges per vertex, ascending
    for (uint32_t h = 0; h < nhe; h++) first[tris[h] + 1]++;
    for (uint32_t v = 0; v < N; v++) first[v + 1] += first[v];
    std::vector<uint32_t> fill(first.begin(), first.end() - 1);
    for (uint32_t h = 0; h < nhe; h++) outs[fill[tris[h]]++] = h;
    std::vector<uint32_t> sym(nhe, none);
    for (uint32_t h = 0; h < nhe; h++) {
        const uint32_t o = tris[h], d = tris[next(h)];
        for (uint32_t i = first[d]; i < first[d + 1]; i++)
            if (tris[next(outs[i])] == o) { sym[h] = outs[i]; break; }
    }
    const float third = 1.0f / 3.0f;
    wa.assign(N, 0.0f);
    for (uint32_t v = 0; v < N; v++) {
        uint32_t start = none;
        for (uint32_t i = first[v]; i < first[v + 1]; i++)
            if (start == none || sym[outs[i]] == none) start = outs[i];
        if (start == none) continue;
        float w = 0.0f;
        uint32_t h = start;
        for (uint32_t steps = first[v + 1] - first[v]; steps > 0; steps--) {
            w += third * area[h / 3];
            const uint32_t nh = sym[prev(h)];
            if (nh == none || nh == start) break;
            h = nh;
        }
        wa[v] = w;
    }
}

// Arguments and parameters a 2D solid accepts (C_WRP, QIE-v, truncated
// differential, ID / MSVD / DAPD rotations; ids in range).
static int check2(const char* fn, uint32_t N, const float* ref, uint32_t T, const uint32_t* tris,
                  const cf_solid_params* p, uint32_t n_knc, const uint32_t* knc) {
    if (!ref || !tris || !p || N == 0 || T == 0 || (n_knc && !knc)) {
        set_error("%s: invalid arguments", fn);
        return CF_ERR_INVALID_ARG;
    }
    if (p->fem.mm != CF_MM_C_WRP) {
        set_error("%s: only the C_WRP material is batched", fn);
        return CF_ERR_UNSUPPORTED;
    }
    if (p->integrator != CF_IT_QIE_V) {
        set_error("%s: only the QIE-v integrator is batched", fn);
        return CF_ERR_UNSUPPORTED;
    }
    if (p->reference_order) {
        set_error("%s: 2D solids take the resident step kernel (reference_order = 0)", fn);
        return CF_ERR_UNSUPPORTED;
    }
    if (!(p->fixed_dt > 0) || !(p->density > 0) || !(p->time_scale > 0) || !(p->thickness > 0) ||
        !(p->fem.young_modulus > 0) || !(p->fem.poisson_ratio > -1.0f) || !(p->fem.poisson_ratio < 0.5f) ||
        p->fem.mm != CF_MM_C_WRP || p->fem.dm != CF_DM_TRUNCATED ||
        (p->fem.rm != CF_RM_ID && p->fem.rm != CF_RM_MSVD && p->fem.rm != CF_RM_DAPD)) {
        set_error("%s: unsupported parameters", fn);
        return CF_ERR_INVALID_ARG;
    }
    for (size_t i = 0; i < 3 * (size_t)T; i++)
        if (tris[i] >= N) { set_error("triangle %zu references node %u >= %u", i / 3, tris[i], N); return CF_ERR_INVALID_ARG; }
    for (uint32_t i = 0; i < n_knc; i++)
        if (knc[i] >= N) { set_error("KNC node %u out of range", knc[i]); return CF_ERR_INVALID_ARG; }
    return CF_OK;
}

// Smallest group of wavefronts that holds the mesh in registers (above 8: the
// mesh does not fit one workgroup).
static uint32_t resident_group(uint32_t N, uint32_t T) {
    uint32_t gw = 1;
    while (gw <= 8 && (T > gw * uint32_t(kGroupTris) || N > gw * uint32_t(kGroupNodes))) gw *= 2;
    return gw;
}

// Element precompute (EndEdition, TriangleElementSet2.cpp:2187-2239) and the
// lumped masses: Sum(Rcp(3) area) * density * thickness in the mesh's CCW fan
// order, stored as 1/m (GetMass = 1/(1/m)).
static int precompute2(uint32_t N, const float* ref, uint32_t T, const uint32_t* tris, const cf_solid_params* p,
                       std::vector<float>& area, std::vector<float4>& dinv, std::vector<float>& mass) {
    area.resize(T);
    dinv.resize(T);
    for (uint32_t e = 0; e < T; e++) {
        const float* r0 = &ref[2 * (size_t)tris[3 * e]];
        const float* r1 = &ref[2 * (size_t)tris[3 * e + 1]];
        const float* r2 = &ref[2 * (size_t)tris[3 * e + 2]];
        area[e] = 0.5f * ((r1[0] - r0[0]) * (r2[1] - r0[1]) - (r2[0] - r0[0]) * (r1[1] - r0[1]));
        const float D00 = r1[0] - r0[0], D01 = r2[0] - r0[0], D10 = r1[1] - r0[1], D11 = r2[1] - r0[1];
        const float det = D00 * D11 - D10 * D01;
        if (!(det > 0)) { set_error("triangle %u is degenerate or inverted at rest", e); return CF_ERR_INVALID_ARG; }
        const float id = 1.0f / det;
        dinv[e] = make_float4(id * D11, -id * D01, -id * D10, id * D00);
    }
    const float density_per_area = p->density * p->thickness;
    std::vector<float> wa;
    fan_weighted_areas(N, T, tris, area.data(), wa);
    mass.resize(N);
    for (uint32_t n = 0; n < N; n++) {
        const float inv_mass = 1.0f / (wa[n] * density_per_area);
        mass[n] = 1.0f / inv_mass;
    }
    return CF_OK;
}

// Greedy colouring, no two triangles of one colour sharing a node; the number
// of colours, or 33 when 32 do not suffice.
static uint32_t colour2(uint32_t N, uint32_t T, const uint32_t* tris, std::vector<uint32_t>& color) {
    std::vector<uint32_t> used(N, 0);
    color.assign(T, 0);
    uint32_t C = 0;
    for (uint32_t e = 0; e < T; e++) {
        uint32_t mask = used[tris[3 * e]] | used[tris[3 * e + 1]] | used[tris[3 * e + 2]];
        uint32_t c = 0;
        while (c < 32 && (mask >> c & 1u)) c++;
        if (c >= 32) return 33;
        color[e] = c;
        for (int j = 0; j < 3; j++) used[tris[3 * e + j]] |= 1u << c;
        C = std::max(C, c + 1);
    }
    return C;
}

// The solid's constants of the step (material, time step, Rayleigh factors,
// CG settings, DAPD factors).
static void consts2(const cf_solid_params* p, Mesh2Dev& m) {
    m.m = material2(p->fem.young_modulus, p->fem.poisson_ratio);
    float c[2];
    rayleigh_coeffs(p->fem.rayleigh_freq1, p->fem.rayleigh_freq2, p->fem.damping_ratio, c);
    const float dt = p->time_scale * p->fixed_dt;
    m.dt = dt;
    m.inv_dt = 1.0f / dt;
    m.fM = 1.0f + dt * c[0];
    m.fK = dt * dt + dt * c[1];
    m.g0 = p->gravity[0];
    m.g1 = p->gravity[1];
    m.rel = p->cg_rel_epsilon;
    m.abs_eps = p->cg_abs_epsilon;
    m.thr = p->fem.degenerate_threshold_det_f;
    m.drag = -std::log(10e-6f) * p->air_drag_coeff;
    m.rm = p->fem.rm;
    m.warm_start = p->warm_start;
    m.compute_energy = p->compute_energy;
    m.max_iter = std::min<uint32_t>(p->cg_max_iter, kMaxCgIter);
    m.dapd_l = p->fem.dapd_l_factor;
    m.dapd_nl = p->fem.dapd_nl_factor;
    m.dapd_e = p->fem.dapd_nl_exponent;
}

int cf_batch2_add_sim(cf_batch2* b, uint32_t N, const float* ref, uint32_t T, const uint32_t* tris,
                      const cf_solid_params* p, uint32_t n_knc, const uint32_t* knc, uint32_t* sim_id) {
    if (!b || !sim_id) {
        set_error("cf_batch2_add_sim: invalid arguments");
        return CF_ERR_INVALID_ARG;
    }
    CF_TRY(check2("cf_batch2_add_sim", N, ref, T, tris, p, n_knc, knc));
    const uint32_t gw = resident_group(N, T);
    if (gw > 8) {
        set_error("2D mesh of %u triangles / %u nodes exceeds one workgroup (%d / %d)", T, N, 8 * kGroupTris,
                  8 * kGroupNodes);
        return CF_ERR_UNSUPPORTED;
    }
    std::vector<float> area, mass;
    std::vector<float4> dinv;
    CF_TRY(precompute2(N, ref, T, tris, p, area, dinv, mass));
    std::vector<uint32_t> color;
    const uint32_t C = colour2(N, T, tris, color);
    if (C > uint32_t(kMaxColors)) { set_error("2D mesh needs more than %d colours", kMaxColors); return CF_ERR_UNSUPPORTED; }
    std::vector<uint32_t> order(T);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t q) { return color[a] < color[q]; });
    CF_TRY(pull_state(b));
    Mesh2Dev m{};
    m.N = N; m.T = T; m.C = C; m.gw = gw;
    m.node_off = uint32_t(b->ref.size());
    m.tri_off = uint32_t(b->area.size());
    m.color_off = uint32_t(b->cstart.size());
    consts2(p, m);
    for (uint32_t n = 0; n < N; n++) {
        b->ref.push_back(make_float2(ref[2 * (size_t)n], ref[2 * (size_t)n + 1]));
        b->x.push_back(b->ref.back());
        b->x0.push_back(b->ref.back());  // m_vecPos0 starts at the rest pose
        b->v.push_back(make_float2(0.f, 0.f));
        b->mass.push_back(mass[n]);
        b->knc.push_back(0);
    }
    for (uint32_t i = 0; i < n_knc; i++) b->knc[m.node_off + knc[i]] = 1;
    std::vector<uint32_t> cstart(C + 1, 0);
    for (uint32_t e = 0; e < T; e++) cstart[color[e] + 1]++;
    for (uint32_t k = 0; k < C; k++) cstart[k + 1] += cstart[k];
    for (uint32_t k = 0; k <= C; k++) b->cstart.push_back(uint16_t(cstart[k]));
    for (uint32_t i = 0; i < T; i++) {
        const uint32_t e = order[i];
        for (int j = 0; j < 3; j++) b->tri_ids.push_back(tris[3 * e + j]);
        b->dinv.push_back(dinv[e]);
        b->area.push_back(area[e]);
        b->noc.push_back(uint8_t(kNocNone));
        b->tri_src.push_back(e);
    }
    b->meshes.push_back(m);
    b->tris += T;
    b->dirty = true;
    *sim_id = uint32_t(b->meshes.size() - 1);
    return CF_OK;
}

int cf_batch2_fixed_step(cf_batch2* b, cf_step_stats* stats) {
    if (!b) return CF_ERR_INVALID_ARG;
    if (b->meshes.empty()) return CF_OK;
    if (b->dirty) CF_TRY(upload(b));
    CF_CHECK_HIP(hipMemsetAsync(b->d_flag, 0, sizeof(int32_t), b->stream));
    CF_CHECK_HIP(hipEventRecord(b->ev0, b->stream));
    CF_TRY(launch_batch2_step(args(b), uint32_t(b->work.size()), b->stream));
    CF_CHECK_HIP(hipEventRecord(b->ev1, b->stream));
    CF_CHECK_HIP(hipEventSynchronize(b->ev1));
    int32_t nonfinite = 0;
    CF_CHECK_HIP(hipMemcpy(&nonfinite, b->d_flag, sizeof(int32_t), hipMemcpyDeviceToHost));
    CF_CHECK_HIP(hipMemcpy(b->h_out.data(), b->d_out, b->h_out.size() * sizeof(Sim2Out), hipMemcpyDeviceToHost));
    // the next step's workgroups group meshes of similar iteration counts
    b->last_iters.resize(b->meshes.size());
    for (size_t i = 0; i < b->h_out.size(); i++) b->last_iters[i] = b->h_out[i].iters;
    build_work(b);
    CF_CHECK_HIP(hipMemcpyAsync(b->d_work, b->work.data(), b->work.size() * sizeof(Work2), hipMemcpyHostToDevice,
                                b->stream));
    if (stats) {
        float ms = 0.f;
        CF_CHECK_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        for (size_t i = 0; i < b->h_out.size(); i++) {
            const Sim2Out& o = b->h_out[i];
            cf_step_stats& s = stats[i];
            s.num_fixed_steps = 1;
            s.cg_iters = o.iters;
            s.cg_reason = (std::isfinite(o.rel) && std::isfinite(o.abs_res)) ? o.reason : CF_CG_NONFINITE;
            s.cg_rel_residual = o.rel;
            s.cg_abs_residual = o.abs_res;
            s.num_degenerate = o.num_degenerate;
            s.sum_degenerate_det_f = float(o.sum_degenerate);
            s.elastic_energy = float(o.energy);
            s.min_det_f = o.min_det;
            s.step_ms = ms;
        }
    }
    if (nonfinite) {
        set_error("2D step: a CG reduction is NaN or Inf (non-finite state)");
        return CF_ERR_NAN;
    }
    return CF_OK;
}

int cf_batch2_get_state(cf_batch2* b, uint32_t id, float* pos, float* vel) {
    if (!b || id >= b->meshes.size()) { set_error("no 2D simulation %u", id); return CF_ERR_INVALID_ARG; }
    const Mesh2Dev& m = b->meshes[id];
    if (b->on_device) {
        CF_CHECK_HIP(hipStreamSynchronize(b->stream));
        if (pos) CF_CHECK_HIP(hipMemcpy(pos, b->d_x + m.node_off, m.N * sizeof(float2), hipMemcpyDeviceToHost));
        if (vel) CF_CHECK_HIP(hipMemcpy(vel, b->d_v + m.node_off, m.N * sizeof(float2), hipMemcpyDeviceToHost));
    } else {
        if (pos) std::memcpy(pos, &b->x[m.node_off], m.N * sizeof(float2));
        if (vel) std::memcpy(vel, &b->v[m.node_off], m.N * sizeof(float2));
    }
    return CF_OK;
}

int cf_batch2_set_state(cf_batch2* b, uint32_t id, const float* pos, const float* vel) {
    if (!b || id >= b->meshes.size()) { set_error("no 2D simulation %u", id); return CF_ERR_INVALID_ARG; }
    const Mesh2Dev& m = b->meshes[id];
    if (b->on_device) {
        CF_CHECK_HIP(hipStreamSynchronize(b->stream));
        if (pos) CF_CHECK_HIP(hipMemcpy(b->d_x + m.node_off, pos, m.N * sizeof(float2), hipMemcpyHostToDevice));
        if (vel) CF_CHECK_HIP(hipMemcpy(b->d_v + m.node_off, vel, m.N * sizeof(float2), hipMemcpyHostToDevice));
    } else {
        if (pos) std::memcpy(&b->x[m.node_off], pos, m.N * sizeof(float2));
        if (vel) std::memcpy(&b->v[m.node_off], vel, m.N * sizeof(float2));
    }
    return CF_OK;
}

int cf_batch2_get_noc(cf_batch2* b, uint32_t id, uint8_t* h_noc) {
    if (!b || !h_noc || id >= b->meshes.size()) {
        set_error("no 2D simulation %u", id);
        return CF_ERR_INVALID_ARG;
    }
    const Mesh2Dev& m = b->meshes[id];
    std::vector<uint8_t> sorted(m.T);
    if (b->on_device) {
        CF_CHECK_HIP(hipStreamSynchronize(b->stream));
        CF_CHECK_HIP(hipMemcpy(sorted.data(), b->d_noc + m.tri_off, m.T, hipMemcpyDeviceToHost));
    } else {
        std::memcpy(sorted.data(), &b->noc[m.tri_off], m.T);
    }
    for (uint32_t i = 0; i < m.T; i++) h_noc[b->tri_src[m.tri_off + i]] = sorted[i];
    return CF_OK;
}

int cf_batch2_fallbacks(cf_batch2* b, uint32_t id, uint32_t* num_fallback) {
    if (!b || !num_fallback || id >= b->meshes.size()) {
        set_error("no 2D simulation %u", id);
        return CF_ERR_INVALID_ARG;
    }
    *num_fallback = 0;
    if (!b->on_device) return CF_OK;  // not stepped since it was added
    Sim2Out o;
    CF_CHECK_HIP(hipStreamSynchronize(b->stream));
    CF_CHECK_HIP(hipMemcpy(&o, b->d_out + id, sizeof(Sim2Out), hipMemcpyDeviceToHost));
    *num_fallback = o.num_fallback;
    return CF_OK;
}

int cf_batch2_node_masses(const cf_batch2* b, uint32_t id, float* mass) {
    if (!b || !mass || id >= b->meshes.size()) { set_error("no 2D simulation %u", id); return CF_ERR_INVALID_ARG; }
    const Mesh2Dev& m = b->meshes[id];
    std::memcpy(mass, &b->mass[m.node_off], m.N * sizeof(float));
    return CF_OK;
}

int cf_batch2_info(const cf_batch2* b, uint32_t* num_sims, uint64_t* tris, uint64_t* nodes, uint32_t* workgroups) {
    if (!b) return CF_ERR_INVALID_ARG;
    if (num_sims) *num_sims = uint32_t(b->meshes.size());
    if (tris) *tris = b->tris;
    if (nodes) *nodes = b->ref.size();
    if (workgroups) {
        uint32_t n[9] = {0};
        for (const auto& m : b->meshes) n[m.gw]++;
        *workgroups = (n[1] + 7) / 8 + (n[2] + 3) / 4 + (n[4] + 1) / 2 + n[8];
    }
    return CF_OK;
}

// ---------------------------------------------------------------- single 2D solid
// A mesh that fits one workgroup is a one-simulation batch (the resident step
// kernel); a larger one (more triangles, nodes or colours than the resident
// kernel holds) is stepped by the multi-workgroup kernels on its own arrays.
}  // extern "C"

namespace {

struct Large2 {
    Large2Dev d{};
    uint32_t max_it = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evc[2] = {nullptr, nullptr};
    int32_t* h_done = nullptr;  // pinned: done[k + 1] from the last d kernel of each chunk
    int32_t* d_done = nullptr;
    std::vector<void*> mem;
    bool stepped = false;
    template <class T>
    int alloc(T** p, size_t n) {
        void* q = nullptr;
        CF_CHECK_HIP(hipMalloc(&q, std::max<size_t>(1, n) * sizeof(T)));
        mem.push_back(q);
        *p = static_cast<T*>(q);
        return CF_OK;
    }
    template <class T>
    int upload(const T** p, const std::vector<T>& h) {
        T* q = nullptr;
        CF_TRY(alloc(&q, h.size()));
        if (!h.empty()) CF_CHECK_HIP(hipMemcpy(q, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        *p = q;
        return CF_OK;
    }
    ~Large2() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* q : mem) (void)hipFree(q);
        if (h_done) (void)hipHostFree(h_done);
        for (hipEvent_t e : {ev0, ev1, evc[0], evc[1]})
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

int large2_create(uint32_t N, const float* ref, uint32_t T, const uint32_t* tris, const cf_solid_params* p,
                  uint32_t n_knc, const uint32_t* knc, Large2** out) {
    std::vector<float> area, mass;
    std::vector<float4> dinv;
    CF_TRY(precompute2(N, ref, T, tris, p, area, dinv, mass));
    Large2* L = new (std::nothrow) Large2;
    if (!L) return CF_ERR_OUT_OF_MEMORY;
    struct Guard {
        Large2*& L;
        ~Guard() { delete L; }
    } guard{L};
    Large2Dev& a = L->d;
    a.N = N;
    a.T = T;
    a.nblk_tri = std::min<uint32_t>((T + kBlock - 1) / kBlock, kMaxPartials);
    a.nblk_node = std::min<uint32_t>((N + kBlock - 1) / kBlock, kMaxPartials);
    a.c = Mesh2Dev{};
    consts2(p, a.c);
    L->max_it = a.c.max_iter;
    if (hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&L->ev0) != hipSuccess || hipEventCreate(&L->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&L->evc[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&L->evc[1], hipEventDisableTiming) != hipSuccess) {
        set_error("large 2D solid: stream/event creation failed");
        return CF_ERR_HIP;
    }
    std::vector<float2> r2(N);
    for (uint32_t n = 0; n < N; n++) r2[n] = make_float2(ref[2 * (size_t)n], ref[2 * (size_t)n + 1]);
    std::vector<uint8_t> kn(N, 0), noc(T, uint8_t(kNocNone));
    for (uint32_t i = 0; i < n_knc; i++) kn[knc[i]] = 1;
    std::vector<uint32_t> tri(tris, tris + 3 * (size_t)T);
    // node -> (triangle, vertex) codes 3 e + v, ascending e (counting sort)
    std::vector<uint32_t> ptr(N + 1, 0), code(3 * (size_t)T);
    for (size_t i = 0; i < 3 * (size_t)T; i++) ptr[tris[i] + 1]++;
    for (uint32_t n = 0; n < N; n++) ptr[n + 1] += ptr[n];
    std::vector<uint32_t> fill(ptr.begin(), ptr.end() - 1);
    for (size_t i = 0; i < 3 * (size_t)T; i++) code[fill[tris[i]]++] = uint32_t(i);
    CF_TRY(L->upload(&a.ref, r2));
    CF_TRY(L->upload(&a.mass, mass));
    CF_TRY(L->upload(&a.knc, kn));
    CF_TRY(L->upload(&a.tri, tri));
    CF_TRY(L->upload(&a.dinv, dinv));
    CF_TRY(L->upload(&a.area, area));
    CF_TRY(L->upload(&a.ptr, ptr));
    CF_TRY(L->upload(&a.code, code));
    const float2* cx = nullptr;
    CF_TRY(L->upload(&cx, r2));
    a.x = const_cast<float2*>(cx);
    CF_TRY(L->upload(&cx, r2));  // m_vecPos0 starts at the rest pose
    a.x0 = const_cast<float2*>(cx);
    const uint8_t* cn = nullptr;
    CF_TRY(L->upload(&cn, noc));
    a.noc = const_cast<uint8_t*>(cn);
    CF_TRY(L->alloc(&a.v, N));
    CF_CHECK_HIP(hipMemset(a.v, 0, N * sizeof(float2)));
    CF_TRY(L->alloc(&a.R, T));
    CF_TRY(L->alloc(&a.tf, 3 * (size_t)T));
    for (float2** v : {&a.b, &a.y, &a.r, &a.d, &a.q}) CF_TRY(L->alloc(v, N));
    CF_TRY(L->alloc(&a.st, 1));
    CF_TRY(L->alloc(&a.part_a, a.nblk_node));
    CF_TRY(L->alloc(&a.part_b, a.nblk_node));
    CF_TRY(L->alloc(&a.stat, 4 * (size_t)a.nblk_tri));
    CF_TRY(L->alloc(&a.stat_min, a.nblk_tri));
    CF_TRY(L->alloc(&a.out, 1));
    CF_CHECK_HIP(hipMemset(a.out, 0, sizeof(Sim2Out)));
    CF_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&L->h_done), (kMaxCgIter + 2) * sizeof(int32_t)));
    std::memset(L->h_done, 0, (kMaxCgIter + 2) * sizeof(int32_t));
    CF_CHECK_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&L->d_done), L->h_done, 0));
    *out = L;
    L = nullptr;  // owned by the caller now
    return CF_OK;
}

// Iterations [k0, kend) of the CG, each skipped on the device once the solve
// has stopped; the last d kernel writes the stop flag to the host.
int large2_chunk(Large2* L, uint32_t k0, uint32_t kend) {
    const Large2Dev& a = L->d;
    for (uint32_t k = k0; k < kend; k++) {
        const int32_t* skip = &a.st->done[k];
        CF_TRY(launch_l2_elem(a, false, a.d, skip, L->stream));
        CF_TRY(launch_l2_matvec(a, false, a.d, skip, L->stream));
        CF_TRY(launch_l2_xr(a, k, L->stream));
        CF_TRY(launch_l2_d(a, k, k + 1 == kend ? L->d_done : nullptr, L->stream));
    }
    return CF_OK;
}

// One QIE-v fixed step (Step_QuasiImplicitEuler_v, LeafDSH_Solid2D_FEM.cpp:1287-1403)
int large2_step(Large2* L, cf_step_stats* stats) {
    const Large2Dev& a = L->d;
    hipStream_t st = L->stream;
    CF_CHECK_HIP(hipEventRecord(L->ev0, st));
    CF_TRY(launch_l2_begin(a, st));                       // BeginEvaluation
    CF_TRY(launch_l2_elem(a, true, a.x, nullptr, st));    // f_e(x)
    CF_TRY(launch_l2_rhs(a, st));                         // b, y, KNC rows
    CF_TRY(launch_l2_elem(a, false, a.y, nullptr, st));   // A y
    CF_TRY(launch_l2_matvec(a, true, a.y, nullptr, st));  // r = S(b - A y), d = r
    CF_TRY(launch_cg_start(a.st, a.part_b, int(a.nblk_node), a.c.rel, a.c.abs_eps, L->max_it, st));
    // chunks of eight iterations, one queued behind the one the host waits for
    constexpr uint32_t kChunk = 8;
    uint32_t k = 0, kend[2] = {0, 0};
    auto launch = [&](uint32_t slot) -> int {
        const uint32_t e = std::min(L->max_it, k + kChunk);
        CF_TRY(large2_chunk(L, k, e));
        k = e;
        kend[slot] = e;
        CF_CHECK_HIP(hipEventRecord(L->evc[slot], st));
        return CF_OK;
    };
    if (L->max_it > 0) {
        CF_TRY(launch(0));
        for (uint32_t c = 0;; c++) {
            const uint32_t cur = c & 1, nxt = cur ^ 1;
            if (k < L->max_it) CF_TRY(launch(nxt));
            hipError_t q;
            while ((q = hipEventQuery(L->evc[cur])) == hipErrorNotReady) {
            }
            CF_CHECK_HIP(q);
            if (kend[cur] >= L->max_it || reinterpret_cast<volatile const int32_t*>(L->h_done)[kend[cur]] != 0) break;
        }
    }
    CF_TRY(launch_l2_integrate(a, st));
    if (a.c.compute_energy) CF_TRY(launch_l2_energy(a, st));
    CF_TRY(launch_l2_finish(a, st));
    CF_CHECK_HIP(hipEventRecord(L->ev1, st));
    CF_CHECK_HIP(hipEventSynchronize(L->ev1));
    L->stepped = true;
    static thread_local CgScalars hs;
    CF_CHECK_HIP(hipMemcpy(&hs, a.st, sizeof(CgScalars), hipMemcpyDeviceToHost));
    cf_cg_result res{};
    cg_result_from(hs, L->max_it, &res);
    if (stats) {
        Sim2Out o;
        CF_CHECK_HIP(hipMemcpy(&o, a.out, sizeof(Sim2Out), hipMemcpyDeviceToHost));
        float ms = 0.f;
        CF_CHECK_HIP(hipEventElapsedTime(&ms, L->ev0, L->ev1));
        *stats = cf_step_stats{};
        stats->num_fixed_steps = 1;
        stats->cg_iters = res.iterations;
        stats->cg_reason = res.reason;
        stats->cg_rel_residual = res.rel_residual;
        stats->cg_abs_residual = res.abs_residual;
        stats->num_degenerate = o.num_degenerate;
        stats->sum_degenerate_det_f = float(o.sum_degenerate);
        stats->elastic_energy = float(o.energy);
        stats->min_det_f = o.min_det;
        stats->step_ms = ms;
    }
    if (res.reason == CF_CG_NONFINITE) {
        set_error("2D step: a CG reduction is NaN or Inf (non-finite state)");
        return CF_ERR_NAN;
    }
    return CF_OK;
}

}  // namespace

extern "C" {

struct cf_solid2 {
    cf_batch2* b = nullptr;
    Large2* large = nullptr;
    float acc_time = 0.f;
    cf_solid_params p{};
};

int cf_solid2_create(uint32_t N, const float* ref, uint32_t T, const uint32_t* tris, const cf_solid_params* p,
                     uint32_t n_knc, const uint32_t* knc, cf_solid2** out) {
    if (!out || !p) return CF_ERR_INVALID_ARG;
    *out = nullptr;
    CF_TRY(check2("cf_solid2_create", N, ref, T, tris, p, n_knc, knc));
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        (void)hipGetLastError();
        set_error("no HIP device");
        return CF_ERR_NO_DEVICE;
    }
    cf_solid2* s = new (std::nothrow) cf_solid2;
    if (!s) return CF_ERR_OUT_OF_MEMORY;
    s->p = *p;
    std::vector<uint32_t> color;
    const bool resident = resident_group(N, T) <= 8 && colour2(N, T, tris, color) <= uint32_t(kMaxColors);
    int rc;
    if (resident) {
        rc = cf_batch2_create(&s->b);
        uint32_t id = 0;
        if (rc == CF_OK) rc = cf_batch2_add_sim(s->b, N, ref, T, tris, p, n_knc, knc, &id);
    } else {
        rc = large2_create(N, ref, T, tris, p, n_knc, knc, &s->large);
    }
    if (rc != CF_OK) { cf_solid2_destroy(s); return rc; }
    *out = s;
    return CF_OK;
}

int cf_solid2_destroy(cf_solid2* s) {
    if (!s) return CF_OK;
    if (s->b) cf_batch2_destroy(s->b);
    delete s->large;
    delete s;
    return CF_OK;
}

static int solid2_fixed(cf_solid2* s, cf_step_stats* stats) {
    return s->large ? large2_step(s->large, stats) : cf_batch2_fixed_step(s->b, stats);
}

int cf_solid2_fixed_step(cf_solid2* s, cf_step_stats* stats) {
    if (!s) return CF_ERR_INVALID_ARG;
    return solid2_fixed(s, stats);
}

// Step (LeafDSH_Solid2D_FEM: fixed-dt accumulator with strict '>')
int cf_solid2_step(cf_solid2* s, float dt, cf_step_stats* stats) {
    if (!s) return CF_ERR_INVALID_ARG;
    s->acc_time += s->p.time_scale * dt;
    const float fixed = s->p.time_scale * s->p.fixed_dt;
    cf_step_stats last{};
    uint32_t n = 0;
    int rc = CF_OK;
    while (s->acc_time > fixed) {
        rc = solid2_fixed(s, &last);
        if (rc != CF_OK && rc != CF_ERR_NAN) return rc;
        s->acc_time -= fixed;
        n++;
        if (rc == CF_ERR_NAN) break;
    }
    last.num_fixed_steps = n;
    if (stats) *stats = last;
    return rc;
}

int cf_solid2_get_noc(cf_solid2* s, uint8_t* h_noc) {
    if (!s || !h_noc) return CF_ERR_INVALID_ARG;
    if (s->large) {
        CF_CHECK_HIP(hipStreamSynchronize(s->large->stream));
        CF_CHECK_HIP(hipMemcpy(h_noc, s->large->d.noc, s->large->d.T, hipMemcpyDeviceToHost));
        return CF_OK;
    }
    return cf_batch2_get_noc(s->b, 0, h_noc);
}

int cf_solid2_fallbacks(cf_solid2* s, uint32_t* num_fallback) {
    if (!s || !num_fallback) return CF_ERR_INVALID_ARG;
    if (s->large) {
        *num_fallback = 0;
        if (!s->large->stepped) return CF_OK;
        Sim2Out o;
        CF_CHECK_HIP(hipStreamSynchronize(s->large->stream));
        CF_CHECK_HIP(hipMemcpy(&o, s->large->d.out, sizeof(Sim2Out), hipMemcpyDeviceToHost));
        *num_fallback = o.num_fallback;
        return CF_OK;
    }
    return cf_batch2_fallbacks(s->b, 0, num_fallback);
}

int cf_solid2_get_state(const cf_solid2* s, float* pos, float* vel) {
    if (!s) return CF_ERR_INVALID_ARG;
    if (s->large) {
        const Large2Dev& a = s->large->d;
        CF_CHECK_HIP(hipStreamSynchronize(s->large->stream));
        if (pos) CF_CHECK_HIP(hipMemcpy(pos, a.x, a.N * sizeof(float2), hipMemcpyDeviceToHost));
        if (vel) CF_CHECK_HIP(hipMemcpy(vel, a.v, a.N * sizeof(float2), hipMemcpyDeviceToHost));
        return CF_OK;
    }
    return cf_batch2_get_state(s->b, 0, pos, vel);
}

int cf_solid2_set_state(cf_solid2* s, const float* pos, const float* vel) {
    if (!s) return CF_ERR_INVALID_ARG;
    if (s->large) {
        const Large2Dev& a = s->large->d;
        CF_CHECK_HIP(hipStreamSynchronize(s->large->stream));
        if (pos) CF_CHECK_HIP(hipMemcpy(a.x, pos, a.N * sizeof(float2), hipMemcpyHostToDevice));
        if (vel) CF_CHECK_HIP(hipMemcpy(a.v, vel, a.N * sizeof(float2), hipMemcpyHostToDevice));
        return CF_OK;
    }
    return cf_batch2_set_state(s->b, 0, pos, vel);
}

}  // extern "C"
