// Host construction of the LDS tiles used by the atomics-free node gather.
//
// Elements are sorted along a Morton (Z-order) curve of their rest centroids
// and cut into tiles of at most `tile_elems` elements and fewer distinct nodes
// than the tile kernel has threads (one element and one node per thread, plus
// the end of the last node's slot run). For every tile we record
//   - its distinct nodes (ascending global id) -> one partial per (tile, node)
//   - per element, the tile-local index of each of its 4 vertices (uint16)
//   - per (element, vertex) its force slot: the slots of one tile node form a
//     contiguous run in ascending element order (the order the reference's
//     serial scatter accumulates them), and tile nodes are ordered by valence
// and, per global node, the contiguous run of its partials in ascending tile order.
// All of this is integer work done once per mesh.
#include "cf_internal.h"
#include <algorithm>
#include <numeric>
#include <cstring>

namespace cf {

namespace {
constexpr uint32_t kMaxTileNodes = 256;  // tile-local node ids are one byte
constexpr uint32_t kMaxRank = 63;        // elements per node within one tile
// LDS bytes of one tile in k_tile_forces: node vectors and rest positions
// (2 x 3 floats), 4 force slots per element (3 floats each) and slot offsets
// (4 B per node).
constexpr uint32_t kMaxTileLds = 64 * 1024 - 512;
inline uint32_t tile_lds(uint32_t ne, uint32_t nn) { return 28 * nn + 48 * ne + 64; }

inline uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}
}  // namespace

uint32_t default_tile_elems() { return 256; }

int build_tiles(uint32_t N, uint32_t T, const float* ref, const uint32_t* tets, uint32_t tile_elems,
                uint32_t max_per_tile, TileHost& out) {
    // the tile stage addresses per-mesh arrays with 32-bit byte offsets
    if (N >= (1u << 28) || T >= (1u << 28)) {
        set_error("mesh too large for one element set: %u nodes, %u elements (limit 2^28 each)", N, T);
        return CF_ERR_UNSUPPORTED;
    }
    if (tile_elems == 0) tile_elems = default_tile_elems();
    if (tile_elems > max_per_tile) tile_elems = max_per_tile;  // one element per thread
    // one node per thread, and one more thread stages the end of the last
    // node's slot run: at most max_per_tile - 1 nodes
    const uint32_t max_nodes = std::min(kMaxTileNodes, max_per_tile) - 1;
    // ---- Morton order of centroids
    float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
    for (uint32_t n = 0; n < N; n++)
        for (int c = 0; c < 3; c++) {
            lo[c] = std::min(lo[c], ref[3 * (size_t)n + c]);
            hi[c] = std::max(hi[c], ref[3 * (size_t)n + c]);
        }
    float ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    if (!(ext > 0)) ext = 1.0f;
    const double scale = double((1u << 21) - 1) / double(ext);
    std::vector<uint64_t> key(T);
    for (uint32_t e = 0; e < T; e++) {
        double c[3] = {0, 0, 0};
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 3; k++) c[k] += ref[3 * (size_t)tets[4 * (size_t)e + j] + k];
        uint64_t q[3];
        for (int k = 0; k < 3; k++) {
            double v = (c[k] * 0.25 - lo[k]) * scale;
            if (v < 0) v = 0;
            if (v > double((1u << 21) - 1)) v = double((1u << 21) - 1);
            q[k] = uint64_t(v);
        }
        key[e] = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
    }
    out.perm.resize(T);
    std::iota(out.perm.begin(), out.perm.end(), 0u);
    std::stable_sort(out.perm.begin(), out.perm.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });

    // ---- cut into tiles
    out.tile_elem_ptr.assign(1, 0);
    out.tile_node_ptr.assign(1, 0);
    out.tile_nodes.clear();
    out.slot_ptr.assign(1, 0);
    out.conn_planes.assign(4 * (size_t)T, 0);
    out.lv.assign(T, 0);
    out.rk.assign(T, 0);
    out.max_tile_elems = 0;
    out.max_tile_nodes = 0;
    std::vector<int32_t> local(N, -1);
    std::vector<uint32_t> in_tile(N, 0);  // elements of the current tile at each node
    std::vector<uint32_t> nodes;  // distinct nodes of the current tile
    uint32_t e = 0;
    while (e < T) {
        const uint32_t e0 = e;
        nodes.clear();
        while (e < T && e - e0 < tile_elems) {
            const uint32_t* t = &tets[4 * (size_t)out.perm[e]];
            uint32_t fresh = 0;
            for (int j = 0; j < 4; j++) fresh += (local[t[j]] < 0) ? 1u : 0u;
            const uint32_t nn1 = uint32_t(nodes.size()) + fresh;
            if (nn1 > max_nodes || tile_lds(e - e0 + 1, nn1) > kMaxTileLds) break;
            // an element's rank in a node's slot run is stored in 6 bits
            bool full = false;
            for (int j = 0; j < 4; j++) full |= in_tile[t[j]] >= kMaxRank;
            if (full) break;
            for (int j = 0; j < 4; j++) {
                if (local[t[j]] < 0) { local[t[j]] = 0; nodes.push_back(t[j]); }
                in_tile[t[j]]++;
            }
            e++;
        }
        const uint32_t ne = e - e0;
        // valence (slots) of each tile node; nodes ordered by descending valence
        // (ties by id) so that the lanes of a wave sum runs of similar length
        std::sort(nodes.begin(), nodes.end());
        for (uint32_t l = 0; l < nodes.size(); l++) local[nodes[l]] = int32_t(l);
        std::vector<uint32_t> val(nodes.size(), 0);
        for (uint32_t i = 0; i < ne; i++) {
            const uint32_t* t = &tets[4 * (size_t)out.perm[e0 + i]];
            for (int j = 0; j < 4; j++) val[local[t[j]]]++;
        }
        std::vector<uint32_t> order(nodes.size());
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return val[a] > val[b]; });
        std::vector<uint32_t> sorted_nodes(nodes.size());
        for (uint32_t r = 0; r < order.size(); r++) sorted_nodes[r] = nodes[order[r]];
        nodes.swap(sorted_nodes);
        for (uint32_t l = 0; l < nodes.size(); l++) local[nodes[l]] = int32_t(l);
        // slot ranges per tile node; within a node, slots in ascending element order
        std::vector<uint32_t> cnt(nodes.size() + 1, 0);
        for (uint32_t i = 0; i < ne; i++) {
            const uint32_t* t = &tets[4 * (size_t)out.perm[e0 + i]];
            for (int j = 0; j < 4; j++) cnt[local[t[j]] + 1]++;
        }
        for (size_t l = 0; l < nodes.size(); l++) cnt[l + 1] += cnt[l];
        const size_t sbase = 4 * (size_t)e0;
        std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
        for (uint32_t i = 0; i < ne; i++) {
            const size_t ie = e0 + i;
            const uint32_t* t = &tets[4 * (size_t)out.perm[ie]];
            uint32_t lv = 0, rk = 0;
            for (int j = 0; j < 4; j++) {
                const uint32_t l = uint32_t(local[t[j]]);
                const uint32_t rank = fill[l]++ - cnt[l];  // slot = run start + rank
                out.conn_planes[j * (size_t)T + ie] = t[j];
                lv |= l << (8 * j);
                rk |= rank << (8 * j);
            }
            out.lv[ie] = lv;
            out.rk[ie] = rk;
        }
        for (size_t l = 0; l < nodes.size(); l++) {
            out.tile_nodes.push_back(nodes[l]);
            out.slot_ptr.push_back(uint32_t(sbase + cnt[l + 1]));
            local[nodes[l]] = -1;
            in_tile[nodes[l]] = 0;
        }
        out.tile_elem_ptr.push_back(e);
        out.tile_node_ptr.push_back(uint32_t(out.tile_nodes.size()));
        out.max_tile_elems = std::max(out.max_tile_elems, ne);
        out.max_tile_nodes = std::max(out.max_tile_nodes, uint32_t(nodes.size()));
    }
    // ---- partials stored node-major: node n owns [node_part_ptr[n], node_part_ptr[n+1]),
    // one entry per tile touching it in ascending tile order; part_dst maps each
    // (tile, tile node) to its entry
    const uint32_t P = uint32_t(out.tile_nodes.size());
    out.node_part_ptr.assign((size_t)N + 1, 0);
    for (uint32_t p = 0; p < P; p++) out.node_part_ptr[out.tile_nodes[p] + 1]++;
    for (uint32_t n = 0; n < N; n++) out.node_part_ptr[n + 1] += out.node_part_ptr[n];
    out.part_dst.assign(P, 0);
    std::vector<uint32_t> fill(out.node_part_ptr.begin(), out.node_part_ptr.end() - 1);
    for (uint32_t p = 0; p < P; p++) out.part_dst[p] = fill[out.tile_nodes[p]]++;
    return CF_OK;
}

}  // namespace cf
