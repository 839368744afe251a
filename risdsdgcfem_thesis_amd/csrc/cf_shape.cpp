// Tet shapes as the reference bakes them (host code): the edit-then-bake
// pipeline of EditableTetSolidShape3 (Geo/shape/TetSolidShape3.cpp) and the
// TetGen importer of examples/tetgen2geo/main.cpp.
//
// EndEdition (:262-268) -> RebuildAllFromEditData (:558-566):
//   FixDegenerateVertices (:809-845)  drop vertices no tetrahedron uses, remap
//   RebuildTopology       (:619-749)  face neighbours (NTID) and oriented
//                                     boundary faces
//   RebuildLayers         (:752-807)  tetrahedra stably sorted by layer, the
//                                     layer being the least graph distance
//                                     of their vertices from the boundary
// AddTetrahedron (:278-307) refuses tetrahedra with det Ds <= 0.
//
// The baked tetrahedron order is the order the reference's element sets (and
// therefore its serial force scatter) see, so a shape baked here reproduces a
// reference scene element for element. Indices are uint32 here (the reference
// stores uint16, so its shapes stop at 65,535 vertices and tetrahedra).
#include "cf_internal.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

struct cf_tetshape {
    std::vector<float> pos;        // 3V
    std::vector<uint32_t> vid;     // 4T
    std::vector<uint32_t> ntid;    // 4T, kNone without a neighbour
    std::vector<uint32_t> bf;      // 3BF
    std::vector<uint32_t> layers;  // 2L: first tet, count
    std::vector<uint32_t> src;     // T: index of the tetrahedron in the edit input
};

namespace cf {
namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;  // cTetSolid3_InvalidFeatureIndex

struct V3f { float x, y, z; };

inline V3f sub(V3f a, V3f b) {
#pragma clang fp contract(off)
    return V3f{a.x - b.x, a.y - b.y, a.z - b.z};
}
inline V3f cross(V3f a, V3f b) {  // Mal/GVec.h:89-91
#pragma clang fp contract(off)
    return V3f{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float dot(V3f a, V3f b) {
#pragma clang fp contract(off)
    float acc = 0.0f;
    acc += a.x * b.x;
    acc += a.y * b.y;
    acc += a.z * b.z;
    return acc;
}
// mal::Det of GMat3x3_From_Columns(c0, c1, c2) (Mal/GMat.h:236-244)
inline float det_cols(V3f c0, V3f c1, V3f c2) {
#pragma clang fp contract(off)
    const float m00 = c0.x, m01 = c1.x, m02 = c2.x;
    const float m10 = c0.y, m11 = c1.y, m12 = c2.y;
    const float m20 = c0.z, m21 = c1.z, m22 = c2.z;
    return m00 * m11 * m22 - m00 * m21 * m12 + m10 * m21 * m02 - m10 * m01 * m22 + m20 * m01 * m12 -
           m20 * m11 * m02;
}

struct Edit {
    std::vector<V3f> v;
    std::vector<uint32_t> t;    // 4 per tet
    std::vector<uint32_t> src;  // input index per accepted tet
};

// AddTetrahedron: positively oriented tetrahedra only.
int add_tets(Edit& ed, uint32_t n_tets, const uint32_t* tets) {
    const uint32_t V = uint32_t(ed.v.size());
    for (uint32_t e = 0; e < n_tets; e++) {
        const uint32_t* q = tets + 4 * (size_t)e;
        for (int j = 0; j < 4; j++)
            if (q[j] >= V) {
                set_error("tetrahedron %u references vertex %u >= %u", e, q[j], V);
                return CF_ERR_INVALID_ARG;
            }
        const V3f p0 = ed.v[q[0]];
        if (det_cols(sub(ed.v[q[1]], p0), sub(ed.v[q[2]], p0), sub(ed.v[q[3]], p0)) <= 0.0f) continue;
        ed.t.insert(ed.t.end(), q, q + 4);
        ed.src.push_back(e);
    }
    return CF_OK;
}

cf_tetshape* bake(Edit& ed) {
    const uint32_t T = uint32_t(ed.src.size());
    // FixDegenerateVertices: keep the vertices some tetrahedron uses, in order
    std::vector<uint32_t> degree(ed.v.size(), 0), remap(ed.v.size(), kNone);
    for (uint32_t x : ed.t) degree[x]++;
    std::vector<V3f> v;
    v.reserve(ed.v.size());
    for (size_t i = 0; i < ed.v.size(); i++)
        if (degree[i]) { remap[i] = uint32_t(v.size()); v.push_back(ed.v[i]); }
    std::vector<uint32_t> t(ed.t.size());
    for (size_t i = 0; i < t.size(); i++) t[i] = remap[ed.t[i]];
    const uint32_t V = uint32_t(v.size());

    // RebuildTopology: the face opposite vertex n of tet e is
    // (vid[(n+1)%4], vid[(n+2)%4], vid[(n+3)%4]); faces with the same vertex
    // set are neighbours, matched in the reference's scan order (each face
    // against the earlier ones)
    std::vector<uint32_t> ntid(4 * (size_t)T, kNone);
    {
        struct Key {
            uint32_t a, b, c;
            bool operator==(const Key& o) const { return a == o.a && b == o.b && c == o.c; }
        };
        struct Hash {
            size_t operator()(const Key& k) const {
                uint64_t h = k.a * 0x9E3779B97F4A7C15ull;
                h ^= (k.b + 0x7F4A7C159E3779B9ull) + (h << 6) + (h >> 2);
                h ^= (k.c + 0x94D049BB133111EBull) + (h << 6) + (h >> 2);
                return size_t(h);
            }
        };
        std::unordered_map<Key, std::vector<uint32_t>, Hash> faces;  // key -> earlier faces (4 e + n)
        faces.reserve(4 * (size_t)T);
        for (uint32_t e = 0; e < T; e++)
            for (uint32_t n = 0; n < 4; n++) {
                uint32_t f[3] = {t[4 * e + (n + 1) % 4], t[4 * e + (n + 2) % 4], t[4 * e + (n + 3) % 4]};
                std::sort(f, f + 3);
                auto& list = faces[Key{f[0], f[1], f[2]}];
                for (uint32_t g : list) {
                    ntid[4 * e + n] = g / 4;
                    ntid[g] = e;
                }
                list.push_back(4 * e + n);
            }
    }
    // boundary faces, oriented outwards (:732-744)
    std::vector<uint32_t> bf;
    for (uint32_t e = 0; e < T; e++)
        for (uint32_t n = 0; n < 4; n++) {
            if (ntid[4 * e + n] != kNone) continue;
            const uint32_t a = t[4 * e + (n + 1) % 4], b = t[4 * e + (n + 2) % 4], c = t[4 * e + (n + 3) % 4];
            const float s = dot(sub(v[t[4 * e + n]], v[a]), cross(sub(v[b], v[a]), sub(v[c], v[a])));
            if (s <= 0.0f) bf.insert(bf.end(), {a, b, c});
            else bf.insert(bf.end(), {a, c, b});
        }

    // RebuildLayers: vertex layer = least number of tetrahedron hops from a
    // boundary vertex (the fixed point the reference reaches by relaxing
    // layer[v1] = min(layer[v1], layer[v2] + 1) over all tetrahedra T times);
    // unreachable vertices keep T. Breadth-first from the boundary vertices.
    const uint32_t cmax = T;
    std::vector<uint32_t> vl(V, cmax);
    {
        std::vector<uint32_t> vt_ptr(V + 1, 0), vt(4 * (size_t)T);  // vertex -> tets
        for (uint32_t x : t) vt_ptr[x + 1]++;
        for (uint32_t i = 0; i < V; i++) vt_ptr[i + 1] += vt_ptr[i];
        std::vector<uint32_t> fill(vt_ptr.begin(), vt_ptr.end() - 1);
        for (uint32_t e = 0; e < T; e++)
            for (int j = 0; j < 4; j++) vt[fill[t[4 * e + j]]++] = e;
        std::vector<uint32_t> frontier, next;
        for (uint32_t x : bf)
            if (vl[x] != 0) { vl[x] = 0; frontier.push_back(x); }
        uint32_t level = 0;
        while (!frontier.empty() && level + 1 < cmax) {
            next.clear();
            for (uint32_t x : frontier)
                for (uint32_t k = vt_ptr[x]; k < vt_ptr[x + 1]; k++)
                    for (int j = 0; j < 4; j++) {
                        const uint32_t y = t[4 * vt[k] + j];
                        if (vl[y] > level + 1) { vl[y] = level + 1; next.push_back(y); }
                    }
            frontier.swap(next);
            level++;
        }
    }
    std::vector<uint32_t> tl(T, cmax);
    uint32_t max_layer = 0;
    for (uint32_t e = 0; e < T; e++) {
        for (int j = 0; j < 4; j++) tl[e] = std::min(tl[e], vl[t[4 * e + j]]);
        max_layer = std::max(max_layer, tl[e]);
    }
    const uint32_t L = T ? max_layer + 1 : 0;
    std::vector<uint32_t> count(L + 1, 0);
    for (uint32_t e = 0; e < T; e++) count[tl[e] + 1]++;
    for (uint32_t l = 0; l < L; l++) count[l + 1] += count[l];
    cf_tetshape* s = new cf_tetshape;
    s->layers.resize(2 * (size_t)L);
    for (uint32_t l = 0; l < L; l++) { s->layers[2 * l] = count[l]; s->layers[2 * l + 1] = count[l + 1] - count[l]; }
    std::vector<uint32_t> old2new(T);
    {
        std::vector<uint32_t> at(count.begin(), count.end() - 1);
        for (uint32_t e = 0; e < T; e++) old2new[e] = at[tl[e]]++;  // stable within a layer
    }
    s->pos.resize(3 * (size_t)V);
    for (uint32_t i = 0; i < V; i++) { s->pos[3 * i] = v[i].x; s->pos[3 * i + 1] = v[i].y; s->pos[3 * i + 2] = v[i].z; }
    s->vid.resize(4 * (size_t)T);
    s->ntid.resize(4 * (size_t)T);
    s->src.resize(T);
    for (uint32_t e = 0; e < T; e++) {
        const uint32_t d = old2new[e];
        for (int j = 0; j < 4; j++) {
            s->vid[4 * d + j] = t[4 * e + j];
            const uint32_t nb = ntid[4 * e + j];
            s->ntid[4 * d + j] = (nb == kNone) ? kNone : old2new[nb];
        }
        s->src[d] = ed.src[e];
    }
    s->bf.swap(bf);
    return s;
}

// Whitespace-separated tokens with '#' comments skipped (TetGen writes a
// trailing comment line; the reference's stream reader takes none before the
// data either).
struct Tokens {
    std::string text;
    size_t at = 0;
    bool open(const std::string& name) {
        std::ifstream f(name, std::ios::binary);
        if (!f) return false;
        std::ostringstream ss;
        ss << f.rdbuf();
        text = ss.str();
        return true;
    }
    bool next(std::string& tok) {
        for (;;) {
            while (at < text.size() && std::isspace((unsigned char)text[at])) at++;
            if (at < text.size() && text[at] == '#') {
                while (at < text.size() && text[at] != '\n') at++;
                continue;
            }
            break;
        }
        if (at >= text.size()) return false;
        const size_t b = at;
        while (at < text.size() && !std::isspace((unsigned char)text[at]) && text[at] != '#') at++;
        tok.assign(text, b, at - b);
        return true;
    }
    bool integer(long long& out) {
        std::string tok;
        if (!next(tok)) return false;
        char* end = nullptr;
        out = std::strtoll(tok.c_str(), &end, 10);
        return end && *end == '\0';
    }
    bool real(float& out) {  // decimal text straight to float, as the reference's >> into float
        std::string tok;
        if (!next(tok)) return false;
        char* end = nullptr;
        out = std::strtof(tok.c_str(), &end);
        return end && *end == '\0';
    }
    bool skip(long long n) {
        std::string tok;
        for (long long i = 0; i < n; i++)
            if (!next(tok)) return false;
        return true;
    }
};

}  // namespace
}  // namespace cf

using namespace cf;

extern "C" {

int cf_tetshape_bake(uint32_t n_vertices, const float* h_pos, uint32_t n_tets, const uint32_t* h_tets,
                     cf_tetshape** out) {
    if (!out || (n_vertices && !h_pos) || (n_tets && !h_tets)) {
        set_error("cf_tetshape_bake: null argument");
        return CF_ERR_INVALID_ARG;
    }
    Edit ed;
    ed.v.resize(n_vertices);
    for (uint32_t i = 0; i < n_vertices; i++) ed.v[i] = V3f{h_pos[3 * i], h_pos[3 * i + 1], h_pos[3 * i + 2]};
    CF_TRY(add_tets(ed, n_tets, h_tets));
    *out = bake(ed);
    return CF_OK;
}

// examples/tetgen2geo/main.cpp:80-180: <base>.node (count, dimension 3,
// attributes, boundary markers; per node id x y z ...) and <base>.ele (count,
// 4 nodes per tet, attributes; per tet id v0..v3 ...), ids relative to the
// first node's id; <base>.face is read but unused by the reference and is not
// read here. natural_scale > 0 applies the importer's -nsac option: centre on
// the bounding box and scale its largest extent to natural_scale.
int cf_tetshape_load_tetgen(const char* base_name, float natural_scale, cf_tetshape** out) {
#pragma clang fp contract(off)
    if (!base_name || !out) return CF_ERR_INVALID_ARG;
    const std::string base(base_name);
    Tokens nf, ef;
    if (!nf.open(base + ".node")) { set_error("cannot open %s.node", base_name); return CF_ERR_INVALID_ARG; }
    if (!ef.open(base + ".ele")) { set_error("cannot open %s.ele", base_name); return CF_ERR_INVALID_ARG; }
    long long nn, dim, nattr, nbm;
    if (!nf.integer(nn) || !nf.integer(dim) || !nf.integer(nattr) || !nf.integer(nbm) || nn < 0 || dim != 3 ||
        nattr < 0 || nbm < 0 || nn > 0xFFFFFFFFll) {
        set_error("%s.node: bad header (need: count 3 attributes markers)", base_name);
        return CF_ERR_INVALID_ARG;
    }
    Edit ed;
    ed.v.resize(size_t(nn));
    long long base_vid = 0;
    // bounding box the way GAABB::Merge keeps it: centre and half sizes
    float pos_c[3] = {0, 0, 0}, half[3] = {0, 0, 0};
    for (long long i = 0; i < nn; i++) {
        long long id;
        float p[3];
        if (!nf.integer(id) || !nf.real(p[0]) || !nf.real(p[1]) || !nf.real(p[2]) || !nf.skip(nattr + nbm)) {
            set_error("%s.node: bad node record %lld", base_name, i);
            return CF_ERR_INVALID_ARG;
        }
        if (i == 0) base_vid = id;
        ed.v[size_t(i)] = V3f{p[0], p[1], p[2]};
        if (i == 0) {
            for (int c = 0; c < 3; c++) { pos_c[c] = 0.5f * (p[c] + p[c]); half[c] = 0.5f * (p[c] - p[c]); }
        } else {
            for (int c = 0; c < 3; c++) {
                float lo = pos_c[c] - half[c], hi = pos_c[c] + half[c];
                lo = std::min(lo, p[c]);
                hi = std::max(hi, p[c]);
                pos_c[c] = 0.5f * (lo + hi);
                half[c] = 0.5f * (hi - lo);
            }
        }
    }
    long long nt, npt, tattr;
    if (!ef.integer(nt) || !ef.integer(npt) || !ef.integer(tattr) || nt < 0 || npt != 4 || tattr < 0 ||
        nt > 0xFFFFFFFFll) {
        set_error("%s.ele: bad header (need: count 4 attributes)", base_name);
        return CF_ERR_INVALID_ARG;
    }
    std::vector<uint32_t> tets(4 * size_t(nt));
    for (long long e = 0; e < nt; e++) {
        long long id, q[4];
        if (!ef.integer(id) || !ef.integer(q[0]) || !ef.integer(q[1]) || !ef.integer(q[2]) || !ef.integer(q[3]) ||
            !ef.skip(tattr)) {
            set_error("%s.ele: bad tetrahedron record %lld", base_name, e);
            return CF_ERR_INVALID_ARG;
        }
        for (int j = 0; j < 4; j++) {
            const long long v = q[j] - base_vid;
            if (v < 0 || v >= nn) {
                set_error("%s.ele: tetrahedron %lld references node %lld", base_name, e, q[j]);
                return CF_ERR_INVALID_ARG;
            }
            tets[4 * size_t(e) + j] = uint32_t(v);
        }
    }
    if (natural_scale > 0.0f && nn > 0) {
        float ext = 0.0f;
        for (int c = 0; c < 3; c++) ext = (c == 0) ? 2.0f * half[0] : std::max(ext, 2.0f * half[c]);
        const float sf = natural_scale / ext;
        for (V3f& p : ed.v) {  // Transform(-centre, Id) then Transform(0, diag(sf))
            p = V3f{p.x + -pos_c[0], p.y + -pos_c[1], p.z + -pos_c[2]};
            p = V3f{sf * p.x, sf * p.y, sf * p.z};
        }
    }
    CF_TRY(add_tets(ed, uint32_t(nt), tets.data()));
    *out = bake(ed);
    return CF_OK;
}

int cf_tetshape_info(const cf_tetshape* s, uint32_t* n_vertices, uint32_t* n_tets, uint32_t* n_boundary_faces,
                     uint32_t* n_layers) {
    if (!s) return CF_ERR_INVALID_ARG;
    if (n_vertices) *n_vertices = uint32_t(s->pos.size() / 3);
    if (n_tets) *n_tets = uint32_t(s->src.size());
    if (n_boundary_faces) *n_boundary_faces = uint32_t(s->bf.size() / 3);
    if (n_layers) *n_layers = uint32_t(s->layers.size() / 2);
    return CF_OK;
}

int cf_tetshape_get(const cf_tetshape* s, float* h_pos, uint32_t* h_tets, uint32_t* h_ntid, uint32_t* h_bf,
                    uint32_t* h_layers, uint32_t* h_src) {
    if (!s) return CF_ERR_INVALID_ARG;
    if (h_pos) std::memcpy(h_pos, s->pos.data(), s->pos.size() * 4);
    if (h_tets) std::memcpy(h_tets, s->vid.data(), s->vid.size() * 4);
    if (h_ntid) std::memcpy(h_ntid, s->ntid.data(), s->ntid.size() * 4);
    if (h_bf) std::memcpy(h_bf, s->bf.data(), s->bf.size() * 4);
    if (h_layers) std::memcpy(h_layers, s->layers.data(), s->layers.size() * 4);
    if (h_src) std::memcpy(h_src, s->src.data(), s->src.size() * 4);
    return CF_OK;
}

int cf_tetshape_destroy(cf_tetshape* s) {
    delete s;
    return CF_OK;
}

}  // extern "C"
