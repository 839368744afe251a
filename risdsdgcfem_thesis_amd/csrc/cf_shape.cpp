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
#include <iterator>
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

// ---- ItemStream shape files (util/ItemStream.h, util/ItemStream.cpp,
// util/ItemStreamSerialization.cpp; the shape item of
// Geo/shape/ShapeLibrary.cpp:332-359).
// Binary: uint32 magic 'UISM', version 0, begin, end, string_end, data size
// and string size (both rounded up to 4), the data block, the string block.
// Items: int32 id (> 0 numeric, < 0 minus the offset of the name in the
// string block, whose byte 0 is unused), uint16 type (low 11 bits the value
// type, 0x1000 simple, 0x4000 array, 0x8000 complex), two pad bytes, uint32
// size including the header; arrays add a uint32 count; payloads are padded
// to 4 bytes; a complex item's size covers its sub-items.
// Text: one item per line, "name value" with value 21u / -3i / 1.5f / 2.0d,
// ^HEX for byte arrays, and "name" followed by { ... } for complex items,
// indented 4 spaces per level; numeric ids are written as #id.
constexpr uint32_t kIsMagic = (uint32_t('U') << 24) | (uint32_t('I') << 16) | (uint32_t('S') << 8) | uint32_t('M');
constexpr uint16_t kIsSimple = 0x1000, kIsArray = 0x4000, kIsComplex = 0x8000, kIsTypeMask = 0x07FF;
constexpr uint16_t kPlaUInt32 = 5, kPlaUInt8 = 9;          // pla_types.h enum order
constexpr uint16_t kTypeShapeDef = 40;                      // cNumPlaTypes + 1 (ShapeLibrary.h:9)
constexpr uint32_t kShapeTetSolid3 = 21;                    // EShapeType (Geo/ShapeTypes.h)
constexpr uint32_t kTetSolidVersion = 4;                    // TetSolidShape3::cVersion
constexpr uint16_t kNone16 = 0xFFFF;

struct IsItem {
    std::string name;  // empty for numeric ids
    int32_t id = 0;
    uint16_t kind = 0, type = 0;
    double number = 0.0;  // simple numeric value
    std::vector<uint8_t> bytes;
    std::vector<IsItem> sub;
    const IsItem* find(const char* n) const {
        for (const IsItem& it : sub)
            if (it.name == n) return &it;
        return nullptr;
    }
};

bool is_parse_bin(const std::vector<uint8_t>& data, const std::vector<char>& strings, uint32_t b, uint32_t e,
                  std::vector<IsItem>& out, int depth) {
    if (depth > 32) return false;
    while (b < e) {
        if (e - b < 12) return false;
        IsItem it;
        int32_t id;
        uint16_t ty;
        uint32_t size;
        std::memcpy(&id, &data[b], 4);
        std::memcpy(&ty, &data[b + 4], 2);
        std::memcpy(&size, &data[b + 8], 4);
        if (size < 12 || size > e - b) return false;
        it.id = id;
        if (id < 0) {
            const size_t off = size_t(-(int64_t)id);
            if (off >= strings.size()) return false;
            const char* str = &strings[off];
            it.name.assign(str, strnlen(str, strings.size() - off));
        }
        it.type = ty & kIsTypeMask;
        if (ty & kIsComplex) {
            it.kind = kIsComplex;
            if (!is_parse_bin(data, strings, b + 12, b + size, it.sub, depth + 1)) return false;
        } else if (ty & kIsArray) {
            it.kind = kIsArray;
            if (size < 16) return false;
            uint32_t count;
            std::memcpy(&count, &data[b + 12], 4);
            if (it.type == kPlaUInt8 || it.type == 8) {  // uint8 or int8 (strings) arrays
                if (count > size - 16) return false;
                it.bytes.assign(&data[b + 16], &data[b + 16] + count);
            }
        } else if (ty & kIsSimple) {
            it.kind = kIsSimple;
            const uint8_t* p = &data[b + 12];
            switch (it.type) {
                case 4: { int32_t v; std::memcpy(&v, p, 4); it.number = v; } break;
                case kPlaUInt32: { uint32_t v; std::memcpy(&v, p, 4); it.number = v; } break;
                case 11: { float v; std::memcpy(&v, p, 4); it.number = v; } break;
                case 12: if (size >= 20) { double v; std::memcpy(&v, p, 8); it.number = v; } break;
                default: break;
            }
        }
        out.push_back(std::move(it));
        b += size;
    }
    return true;
}

// ItemStream::Parse subset (util/ItemStream.cpp:305-624): names, numbers with
// an i/u/f/d suffix (none: double), ^hex byte arrays, nested { } items, '/'
// comment lines, optional '=' and ',' separators.
struct IsText {
    const std::string& s;
    size_t at = 0;
    explicit IsText(const std::string& str) : s(str) {}
    void skip(const char* set) {
        while (at < s.size() && std::strchr(set, s[at]) && s[at]) at++;
    }
    bool items(std::vector<IsItem>& out, bool nested, int depth) {
        if (depth > 32) return false;
        for (;;) {
            skip(" \t\r\n,");
            if (at < s.size() && s[at] == '/') {
                while (at < s.size() && s[at] != '\n') at++;
                continue;
            }
            if (at >= s.size()) return !nested;
            if (s[at] == '}') {
                if (!nested) return false;
                at++;
                return true;
            }
            IsItem it;
            const size_t nb = at;
            while (at < s.size() && !std::strchr(" \t\r\n=", s[at])) at++;
            std::string name = s.substr(nb, at - nb);
            if (name.empty()) return false;
            if (name[0] == '#') {
                char* end = nullptr;
                it.id = int32_t(std::strtol(name.c_str() + 1, &end, 10));
                if (!end || *end) it.name = name;
            } else {
                it.name = name;
            }
            skip(" \t\r\n=");
            if (at >= s.size()) return false;
            const char c = s[at];
            if (c == '{') {
                at++;
                it.kind = kIsComplex;
                if (!items(it.sub, true, depth + 1)) return false;
            } else if (c == '^') {
                at++;
                skip(" ");
                it.kind = kIsArray;
                it.type = kPlaUInt8;
                auto hex = [](char h) -> int {
                    return (h >= '0' && h <= '9') ? h - '0' : (h >= 'A' && h <= 'F') ? h - 'A' + 10 : -1;
                };
                while (at + 1 < s.size() && hex(s[at]) >= 0 && hex(s[at + 1]) >= 0) {
                    it.bytes.push_back(uint8_t(hex(s[at]) * 16 + hex(s[at + 1])));
                    at += 2;
                }
            } else if (std::isdigit((unsigned char)c) || c == '+' || c == '-' || c == '.') {
                const size_t b = at;
                while (at < s.size() && std::strchr("0123456789+-.", s[at]) && s[at]) at++;
                const std::string num = s.substr(b, at - b);
                it.kind = kIsSimple;
                const char suf = (at < s.size()) ? s[at] : '\0';
                if (suf == 'u') { it.type = kPlaUInt32; it.number = double(uint32_t(std::atoi(num.c_str()))); at++; }
                else if (suf == 'i') { it.type = 4; it.number = std::atoi(num.c_str()); at++; }
                else if (suf == 'f') { it.type = 11; it.number = float(std::atof(num.c_str())); at++; }
                else if (suf == 'd') { it.type = 12; it.number = std::atof(num.c_str()); at++; }
                else { it.type = 12; it.number = std::atof(num.c_str()); }
            } else {
                return false;  // keywords, strings, arrays, NiR: not part of a shape item
            }
            out.push_back(std::move(it));
        }
    }
};

int is_read_file(const char* path, std::vector<IsItem>& items) {
    std::ifstream f(path, std::ios::binary);
    if (!f) { set_error("cannot open %s", path); return CF_ERR_INVALID_ARG; }
    std::vector<uint8_t> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    uint32_t magic = 0;
    if (raw.size() >= 4) std::memcpy(&magic, raw.data(), 4);
    if (magic == kIsMagic) {
        if (raw.size() < 28) { set_error("%s: truncated item stream", path); return CF_ERR_INVALID_ARG; }
        uint32_t h[7];
        std::memcpy(h, raw.data(), 28);  // magic, version, begin, end, string_end, data size, string size
        if (h[1] != 0 || h[2] > h[3] || h[3] > h[5] || h[4] > h[6] || 28ull + h[5] + h[6] > raw.size()) {
            set_error("%s: bad item stream header", path);
            return CF_ERR_INVALID_ARG;
        }
        std::vector<uint8_t> data(raw.begin() + 28, raw.begin() + 28 + h[5]);
        std::vector<char> strings(raw.begin() + 28 + h[5], raw.begin() + 28 + h[5] + h[6]);
        strings.push_back('\0');
        if (!is_parse_bin(data, strings, h[2], h[3], items, 0)) {
            set_error("%s: malformed binary item stream", path);
            return CF_ERR_INVALID_ARG;
        }
        return CF_OK;
    }
    const std::string text(raw.begin(), raw.end());
    IsText p(text);
    if (!p.items(items, false, 0)) {
        set_error("%s: malformed text item stream (at byte %zu)", path, p.at);
        return CF_ERR_INVALID_ARG;
    }
    return CF_OK;
}

// the arrays of a baked shape in the reference's 16-bit layout
struct Shape16 {
    std::vector<uint8_t> points, v, t, bf, l;
};

int shape16(const cf_tetshape* s, Shape16& o) {
    const size_t V = s->pos.size() / 3, T = s->src.size(), BF = s->bf.size() / 3, L = s->layers.size() / 2;
    if (V > 0xFFFF || T > 0xFFFF || L > 0xFFFF) {
        set_error("the reference's shape files hold 16-bit indices: %zu vertices, %zu tetrahedra", V, T);
        return CF_ERR_UNSUPPORTED;
    }
    o.points.resize(12 * V);
    std::memcpy(o.points.data(), s->pos.data(), 12 * V);
    o.v.assign(V, 0);
    std::vector<uint16_t> t(8 * T), bf(3 * BF), l(2 * L);
    for (size_t e = 0; e < T; e++)
        for (int j = 0; j < 4; j++) {
            t[8 * e + j] = uint16_t(s->vid[4 * e + j]);
            t[8 * e + 4 + j] = (s->ntid[4 * e + j] == kNone) ? kNone16 : uint16_t(s->ntid[4 * e + j]);
        }
    for (size_t i = 0; i < 3 * BF; i++) bf[i] = uint16_t(s->bf[i]);
    for (size_t i = 0; i < 2 * L; i++) l[i] = uint16_t(s->layers[i]);
    auto bytes = [](const std::vector<uint16_t>& x, std::vector<uint8_t>& out) {
        out.resize(2 * x.size());
        std::memcpy(out.data(), x.data(), out.size());
    };
    bytes(t, o.t);
    bytes(bf, o.bf);
    bytes(l, o.l);
    return CF_OK;
}

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

// ShapeLibrary::Load (Geo/shape/ShapeLibrary.cpp:460-468) of a library file
// holding a TetSolidShape3 (binary or text, told apart by the magic number):
// the first top-level complex item with shape_type == eShape_TetSolid3.
int cf_tetshape_load(const char* path, cf_tetshape** out) {
    if (!path || !out) return CF_ERR_INVALID_ARG;
    std::vector<IsItem> items;
    CF_TRY(is_read_file(path, items));
    for (const IsItem& it : items) {
        if (it.kind != kIsComplex) continue;
        const IsItem* st = it.find("shape_type");
        if (!st || st->kind != kIsSimple || st->number != double(kShapeTetSolid3)) continue;
        const IsItem* ver = it.find("version");
        if (!ver || ver->number != double(kTetSolidVersion)) {
            set_error("%s: TetSolidShape3 version %g (only version %u is read)", path, ver ? ver->number : -1.0,
                      kTetSolidVersion);
            return CF_ERR_UNSUPPORTED;
        }
        const char* names[5] = {"vec_points", "vec_v", "vec_t", "vec_bf", "vec_l"};
        const char* counts[4] = {"num_v", "num_t", "num_bf", "num_l"};
        double n[4];
        for (int k = 0; k < 4; k++) {
            const IsItem* c = it.find(counts[k]);
            if (!c || c->kind != kIsSimple || !(c->number >= 0) || c->number > 4294967295.0) {
                set_error("%s: missing or bad %s", path, counts[k]);
                return CF_ERR_INVALID_ARG;
            }
            n[k] = c->number;
        }
        const size_t V = size_t(n[0]), T = size_t(n[1]), BF = size_t(n[2]), L = size_t(n[3]);
        const size_t want[5] = {12 * V, V, 16 * T, 6 * BF, 4 * L};
        const IsItem* arr[5];
        for (int k = 0; k < 5; k++) {
            arr[k] = it.find(names[k]);
            if (!arr[k] || arr[k]->kind != kIsArray || arr[k]->bytes.size() != want[k]) {
                set_error("%s: missing or bad %s", path, names[k]);
                return CF_ERR_INVALID_ARG;
            }
        }
        cf_tetshape* s = new cf_tetshape;
        s->pos.resize(3 * V);
        std::memcpy(s->pos.data(), arr[0]->bytes.data(), 12 * V);
        s->vid.resize(4 * T);
        s->ntid.resize(4 * T);
        s->src.resize(T);
        for (size_t e = 0; e < T; e++) {
            uint16_t q[8];
            std::memcpy(q, &arr[2]->bytes[16 * e], 16);
            for (int j = 0; j < 4; j++) {
                s->vid[4 * e + j] = q[j];
                s->ntid[4 * e + j] = (q[4 + j] == kNone16) ? kNone : q[4 + j];
            }
            s->src[e] = uint32_t(e);
        }
        s->bf.resize(3 * BF);
        for (size_t i = 0; i < 3 * BF; i++) {
            uint16_t x;
            std::memcpy(&x, &arr[3]->bytes[2 * i], 2);
            s->bf[i] = x;
        }
        s->layers.resize(2 * L);
        for (size_t i = 0; i < 2 * L; i++) {
            uint16_t x;
            std::memcpy(&x, &arr[4]->bytes[2 * i], 2);
            s->layers[i] = x;
        }
        for (size_t i = 0; i < 4 * T; i++)
            if (s->vid[i] >= V || (s->ntid[i] != kNone && s->ntid[i] >= T)) {
                delete s;
                set_error("%s: tetrahedron index out of range", path);
                return CF_ERR_INVALID_ARG;
            }
        *out = s;
        return CF_OK;
    }
    set_error("%s: no TetSolidShape3 item", path);
    return CF_ERR_INVALID_ARG;
}

// ShapeLibrary::Register_TetSolidShape3 + Save (ShapeLibrary.cpp:332-359,
// :446-450): a one-shape library, the shape as item #1.
int cf_tetshape_save(const cf_tetshape* s, const char* path, int binary) {
    if (!s || !path) return CF_ERR_INVALID_ARG;
    Shape16 a;
    CF_TRY(shape16(s, a));
    const uint32_t counts[4] = {uint32_t(s->pos.size() / 3), uint32_t(s->src.size()), uint32_t(s->bf.size() / 3),
                                uint32_t(s->layers.size() / 2)};
    const char* cnames[6] = {"shape_type", "version", "num_v", "num_t", "num_bf", "num_l"};
    const uint32_t cvals[6] = {kShapeTetSolid3, kTetSolidVersion, counts[0], counts[1], counts[2], counts[3]};
    const char* anames[5] = {"vec_points", "vec_v", "vec_t", "vec_bf", "vec_l"};
    const std::vector<uint8_t>* arrs[5] = {&a.points, &a.v, &a.t, &a.bf, &a.l};
    std::string out;
    if (binary) {
        std::vector<char> strings(1, '\0');  // offset 0 is never a name
        std::vector<uint32_t> off;
        for (const char* nm : cnames) { off.push_back(uint32_t(strings.size())); strings.insert(strings.end(), nm, nm + std::strlen(nm) + 1); }
        for (const char* nm : anames) { off.push_back(uint32_t(strings.size())); strings.insert(strings.end(), nm, nm + std::strlen(nm) + 1); }
        std::vector<uint8_t> data;
        auto put32 = [&](uint32_t x) { const uint8_t* p = reinterpret_cast<const uint8_t*>(&x); data.insert(data.end(), p, p + 4); };
        auto header = [&](int32_t id, uint16_t type, uint32_t size) {
            put32(uint32_t(id));
            data.push_back(uint8_t(type & 0xFF));
            data.push_back(uint8_t(type >> 8));
            data.push_back(0);
            data.push_back(0);
            put32(size);
        };
        header(1, uint16_t(kTypeShapeDef | kIsComplex), 0);
        for (int k = 0; k < 6; k++) {
            header(-int32_t(off[k]), uint16_t(kPlaUInt32 | kIsSimple), 16);
            put32(cvals[k]);
        }
        for (int k = 0; k < 5; k++) {
            const std::vector<uint8_t>& b = *arrs[k];
            const uint32_t padded = uint32_t(4 * ((b.size() + 3) / 4));
            header(-int32_t(off[6 + k]), uint16_t(kPlaUInt8 | kIsArray), 16 + padded);
            put32(uint32_t(b.size()));
            data.insert(data.end(), b.begin(), b.end());
            data.resize(data.size() + (padded - b.size()), 0);
        }
        const uint32_t total = uint32_t(data.size());
        std::memcpy(&data[8], &total, 4);
        const uint32_t send = uint32_t(strings.size());
        strings.resize(4 * ((strings.size() + 3) / 4), '\0');
        const uint32_t hdr[7] = {kIsMagic, 0u, 0u, total, send, total, uint32_t(strings.size())};
        out.append(reinterpret_cast<const char*>(hdr), 28);
        out.append(reinterpret_cast<const char*>(data.data()), data.size());
        out.append(strings.data(), strings.size());
    } else {  // ItemToFileTxt (ItemStreamSerialization.cpp:14-173), 4-space indentation
        static const char hexd[] = "0123456789ABCDEF";
        out += "#1\n{\n";
        char line[64];
        for (int k = 0; k < 6; k++) {
            std::snprintf(line, sizeof(line), "    %s %du\n", cnames[k], int(cvals[k]));
            out += line;
        }
        for (int k = 0; k < 5; k++) {
            out += "    ";
            out += anames[k];
            out += " ^";
            for (uint8_t b : *arrs[k]) { out += hexd[b >> 4]; out += hexd[b & 15]; }
            out += "\n";
        }
        out += "}\n";
    }
    std::ofstream f(path, std::ios::binary);
    if (!f || !f.write(out.data(), std::streamsize(out.size()))) {
        set_error("cannot write %s", path);
        return CF_ERR_INVALID_ARG;
    }
    return CF_OK;
}

}  // extern "C"
