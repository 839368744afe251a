// Batched stepping of many solids (cf_batch_* in include/corofem.h).
//
// Host side of the batch: a device arena that sub-allocates meshes and
// simulations (so admission and removal do not call hipMalloc/hipFree, which
// synchronise the device), shared mesh topologies, simulations in slots of a
// device descriptor array, the block tables of the batched kernels, and the
// step recorded as HIP graphs.
#include "cf_internal.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

namespace cf {

// ---------------------------------------------------------------- arena
// First-fit sub-allocator over device chunks of at least kChunk bytes, 256 B
// alignment, neighbours coalesced on free. Chunks are returned to HIP only
// when the batch is destroyed.
class Arena {
public:
    static constexpr size_t kAlign = 256;
    static constexpr size_t kChunk = size_t(1) << 30;
    ~Arena() { release(); }
    int alloc(size_t bytes, char** out) {
        bytes = round(bytes ? bytes : 1);
        for (size_t c = 0; c < chunks_.size(); c++) {
            auto& fr = chunks_[c].free;
            for (auto it = fr.begin(); it != fr.end(); ++it) {
                if (it->second < bytes) continue;
                const size_t off = it->first, sz = it->second;
                fr.erase(it);
                if (sz > bytes) fr[off + bytes] = sz - bytes;
                char* p = chunks_[c].base + off;
                live_[p] = {c, bytes};
                in_use_ += bytes;
                *out = p;
                return CF_OK;
            }
        }
        const size_t csz = std::max(kChunk, bytes);
        char* base = nullptr;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&base), csz);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("batch arena: hipMalloc(%zu) failed: %s", csz, hipGetErrorString(e));
            return e == hipErrorOutOfMemory ? CF_ERR_OUT_OF_MEMORY : CF_ERR_HIP;
        }
        chunks_.push_back(Chunk{base, csz, {}});
        reserved_ += csz;
        if (csz > bytes) chunks_.back().free[bytes] = csz - bytes;
        live_[base] = {chunks_.size() - 1, bytes};
        in_use_ += bytes;
        *out = base;
        return CF_OK;
    }
    void free(char* p) {
        if (!p) return;
        auto it = live_.find(p);
        if (it == live_.end()) return;
        const size_t c = it->second.first, sz = it->second.second;
        live_.erase(it);
        in_use_ -= sz;
        auto& fr = chunks_[c].free;
        size_t off = size_t(p - chunks_[c].base), len = sz;
        auto nx = fr.lower_bound(off);
        if (nx != fr.end() && nx->first == off + len) { len += nx->second; nx = fr.erase(nx); }
        if (nx != fr.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == off) { off = pv->first; len += pv->second; fr.erase(pv); }
        }
        fr[off] = len;
    }
    void release() {
        for (auto& c : chunks_) (void)hipFree(c.base);
        chunks_.clear();
        live_.clear();
        in_use_ = reserved_ = 0;
    }
    static size_t round(size_t n) { return (n + kAlign - 1) & ~(kAlign - 1); }
    size_t reserved() const { return reserved_; }

private:
    struct Chunk {
        char* base;
        size_t size;
        std::map<size_t, size_t> free;  // offset -> length
    };
    std::vector<Chunk> chunks_;
    std::unordered_map<char*, std::pair<size_t, size_t>> live_;
    size_t in_use_ = 0, reserved_ = 0;
};

// Offsets of sub-arrays inside one arena block.
struct Layout {
    size_t off = 0;
    template <class T>
    size_t add(size_t count) {
        off = Arena::round(off);
        const size_t o = off;
        off += (count ? count : 1) * sizeof(T);
        return o;
    }
};

}  // namespace cf

using namespace cf;

namespace {

struct BMesh {
    uint32_t N = 0, T = 0, num_tiles = 0, num_parts = 0, max_te = 0, max_tn = 0, users = 0;
    float total_volume = 0.f;
    TetDevice dev;               // pointers into `block` (partials unused)
    char* block = nullptr;
    size_t bytes = 0;
    std::vector<float> ref_host; // rest positions (default KNC targets)
};

struct BSim {
    uint32_t mesh = 0, slot = 0;
    bool active = true;
    cf_solid_params p{};
    char* block = nullptr;
    size_t bytes = 0;
    SimDev d{};
};

// Grid sizes of one chunk's kernels on device-compacted lists (tile, matvec
// and CG vector kernels). A grid at least as long as its list runs one entry
// per workgroup, as the full tables do; a shorter one strides.
struct TailGrid {
    uint32_t tile = 8, mv = 1, vec = 1;
    std::tuple<uint32_t, uint32_t, uint32_t> key() const { return std::make_tuple(tile, mv, vec); }
};

struct TileBuild {
    TileHost th;
    std::vector<float> ref;
    float total_volume = 0.f;
    int rc = CF_OK;
};

}  // namespace

struct cf_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0, used = 0;
    Arena arena;
    uint32_t tile_threads = 256, tile_elems = 256;
    uint32_t cg_chunk = 8;
    int use_graphs = 1;
    std::map<uint32_t, BMesh> meshes;
    std::map<uint32_t, BSim> sims;
    uint32_t next_mesh = 1, next_sim = 1;
    // slots
    uint32_t cap = 0;
    SimDev* d_sims = nullptr;
    SimCg* d_cgs = nullptr;
    std::vector<uint32_t> free_slots;
    uint32_t slots_used = 0;
    // tables
    bool dirty = true;
    std::vector<uint32_t> active_ids;  // sim ids in table order
    BatchTables tab{};
    BlockRef* d_tab = nullptr;
    size_t tab_cap = 0;                // BlockRef entries
    uint32_t* d_active = nullptr;
    size_t active_cap = 0;
    uint32_t* d_counter = nullptr;
    uint32_t* h_counter = nullptr;     // pinned, 4 entries
    // compacted tables of the simulations still iterating (CG loop only)
    BlockRef* d_tab_cg = nullptr;
    size_t tab_cg_cap = 0;
    uint32_t* d_active_cg = nullptr;
    size_t active_cg_cap = 0;
    std::vector<BlockRef> h_tab_cg;
    std::vector<uint32_t> h_active_cg;
    SimCg* h_cgs_pin[2] = {nullptr, nullptr};  // pinned copies of the CG scalars, per chunk parity
    uint32_t h_cgs_cap = 0;
    uint32_t max_iter = 0;
    uint64_t active_tets = 0, active_nodes = 0;
    // lists compacted on the device (CF_BATCH_STEP_DEVICE*): list lengths,
    // per-simulation offsets, and the tables the tail kernels walk
    uint32_t* d_dyn = nullptr;
    uint32_t* d_comp_off = nullptr;
    size_t comp_off_cap = 0;
    BatchTables tab_dyn{};
    CompactOut comp{};
    uint32_t fill_grid = 1;
    // step schedule and, for the device schedule, the last step's chunk
    // count (chunks until no simulation iterated)
    int step_mode = CF_BATCH_STEP_DEVICE;
    bool hist_valid = false;
    uint32_t hist_chunks = 0;
    uint32_t last_continuations = 0;
    // grids of each chunk of the next step on device-compacted lists, from
    // the previous step's list lengths at that chunk (power-of-two classes)
    std::vector<TailGrid> hist_grids;
    // graphs; tail chunks recorded per grid class
    hipGraphExec_t g_pro = nullptr, g_chunk = nullptr, g_epi = nullptr, g_round = nullptr;
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, hipGraphExec_t> g_tails;
    uint32_t g_chunk_len = 0, g_round_len = 0;
    // steps queued for the next round: (slot, steps)
    std::vector<uint2> pending;
    uint2* d_arm = nullptr;
    size_t arm_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evc[2] = {nullptr, nullptr};
    std::vector<SimCg> h_cgs;
};

namespace {

void free_graphs(cf_batch* b) {
    if (b->g_pro) (void)hipGraphExecDestroy(b->g_pro);
    if (b->g_chunk) (void)hipGraphExecDestroy(b->g_chunk);
    if (b->g_epi) (void)hipGraphExecDestroy(b->g_epi);
    if (b->g_round) (void)hipGraphExecDestroy(b->g_round);
    for (auto& kv : b->g_tails) (void)hipGraphExecDestroy(kv.second);
    b->g_tails.clear();
    b->g_pro = b->g_chunk = b->g_epi = b->g_round = nullptr;
}

template <class T>
T* at(char* base, size_t off) { return reinterpret_cast<T*>(base + off); }

// Device layout of a mesh's topology (one arena block).
struct MeshLayout {
    size_t ref, conn, dinv, vol, tep, tnp, tn, sp, lv, rk, npp, pd, rt, total;
    MeshLayout(uint32_t N, uint32_t T, uint32_t nt, uint32_t P) {
        Layout l;
        ref = l.add<float>(3 * (size_t)N);
        conn = l.add<uint32_t>(4 * (size_t)T);
        dinv = l.add<float>(9 * (size_t)T);
        vol = l.add<float>(T);
        tep = l.add<uint32_t>(nt + 1);
        tnp = l.add<uint32_t>(nt + 1);
        tn = l.add<uint32_t>(P);
        sp = l.add<uint32_t>(P + 1);
        lv = l.add<uint32_t>(T);
        rk = l.add<uint32_t>(T);
        npp = l.add<uint32_t>((size_t)N + 1);
        pd = l.add<uint32_t>(P);
        rt = l.add<float>(3 * (size_t)P);  // rest positions per tile node, tile-major
        total = Arena::round(l.off);
    }
};

// Device layout of a simulation's state (one arena block).
struct SimLayout {
    size_t vec[8], R, detF, part, mask, pa, pb, pe, ec, es, em, x0, doc, total;
    uint32_t nb_elem, nb_node, nb_init, nb_vec, nb_mv;
    SimLayout(uint32_t N, uint32_t T, uint32_t P, uint32_t num_tiles) {
        nb_elem = uint32_t(node_blocks(T));
        nb_node = uint32_t(node_blocks(N));
        nb_init = uint32_t(node_blocks(3 * N));
        nb_vec = cg_chunks(3 * N);
        nb_mv = matvec_chunks_of(N);
        Layout l;
        for (int i = 0; i < 8; i++) vec[i] = l.add<float>(3 * (size_t)N);
        R = l.add<float>(9 * (size_t)T);
        detF = l.add<float>(T);
        part = l.add<float>(3 * (size_t)P);
        mask = l.add<uint8_t>(N);
        pa = l.add<double>(nb_mv);
        pb = l.add<double>(std::max(nb_mv, nb_vec));
        pe = l.add<double>(std::max(num_tiles, 1u));
        ec = l.add<uint32_t>(2 * (size_t)nb_elem);
        es = l.add<double>(nb_elem);
        em = l.add<float>(nb_elem);
        x0 = l.add<float>(3 * (size_t)N);  // DAPD: previous positions and DoC
        doc = l.add<uint8_t>(T);
        total = Arena::round(l.off);
    }
};

int build_mesh_host(uint32_t N, const float* ref, uint32_t T, const uint32_t* tets, uint32_t tile_elems,
                    uint32_t tile_threads, TileBuild& out) {
    for (size_t i = 0; i < 4 * (size_t)T; i++)
        if (tets[i] >= N) {
            set_error("element %zu references node %u >= %u", i / 4, tets[i], N);
            return CF_ERR_INVALID_ARG;
        }
    CF_TRY(build_tiles(N, T, ref, tets, tile_elems, tile_threads, out.th));
    out.ref.assign(ref, ref + 3 * (size_t)N);
    return CF_OK;
}

int upload_mesh(cf_batch* b, TileBuild& tb, uint32_t* mesh_id) {
    const TileHost& th = tb.th;
    const uint32_t N = uint32_t(tb.ref.size() / 3), T = uint32_t(th.perm.size());
    const uint32_t nt = uint32_t(th.tile_elem_ptr.size() - 1), P = uint32_t(th.tile_nodes.size());
    const MeshLayout L(N, T, nt, P);
    if (b->budget && b->used + L.total > b->budget) {
        set_error("mesh of %zu bytes exceeds the batch's HBM budget", L.total);
        return CF_ERR_CAPACITY;
    }
    BMesh m;
    m.N = N; m.T = T; m.num_tiles = nt; m.num_parts = P;
    m.max_te = th.max_tile_elems; m.max_tn = th.max_tile_nodes;
    CF_TRY(b->arena.alloc(L.total, &m.block));
    m.bytes = L.total;
    TetDevice& t = m.dev;
    t.N = N; t.T = T;
    t.ref = at<float>(m.block, L.ref);
    t.conn = at<uint32_t>(m.block, L.conn);
    t.dinv = at<float>(m.block, L.dinv);
    t.vol = at<float>(m.block, L.vol);
    t.tile_elem_ptr = at<uint32_t>(m.block, L.tep);
    t.tile_node_ptr = at<uint32_t>(m.block, L.tnp);
    t.tile_nodes = at<uint32_t>(m.block, L.tn);
    t.slot_ptr = at<uint32_t>(m.block, L.sp);
    t.lv = at<uint32_t>(m.block, L.lv);
    t.rk = at<uint32_t>(m.block, L.rk);
    t.node_part_ptr = at<uint32_t>(m.block, L.npp);
    t.part_dst = at<uint32_t>(m.block, L.pd);
    t.ref_tile = at<float>(m.block, L.rt);
    t.num_tiles = nt; t.max_tile_elems = m.max_te; t.max_tile_nodes = m.max_tn; t.num_parts = P;
    hipStream_t s = b->stream;
    auto put = [&](void* dst, const void* src, size_t bytes) -> int {
        if (bytes) CF_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
        return CF_OK;
    };
    int rc = put(t.ref, tb.ref.data(), tb.ref.size() * 4);
    if (rc == CF_OK) rc = put(t.conn, th.conn_planes.data(), th.conn_planes.size() * 4);
    if (rc == CF_OK) rc = put(t.tile_elem_ptr, th.tile_elem_ptr.data(), th.tile_elem_ptr.size() * 4);
    if (rc == CF_OK) rc = put(t.tile_node_ptr, th.tile_node_ptr.data(), th.tile_node_ptr.size() * 4);
    if (rc == CF_OK) rc = put(t.tile_nodes, th.tile_nodes.data(), th.tile_nodes.size() * 4);
    if (rc == CF_OK) rc = put(t.slot_ptr, th.slot_ptr.data(), th.slot_ptr.size() * 4);
    if (rc == CF_OK) rc = put(t.lv, th.lv.data(), th.lv.size() * 4);
    if (rc == CF_OK) rc = put(t.rk, th.rk.data(), th.rk.size() * 4);
    if (rc == CF_OK) rc = put(t.node_part_ptr, th.node_part_ptr.data(), th.node_part_ptr.size() * 4);
    if (rc == CF_OK) rc = put(t.part_dst, th.part_dst.data(), th.part_dst.size() * 4);
    std::vector<float> rt(3 * (size_t)P);
    for (size_t q = 0; q < P; q++)
        for (int c = 0; c < 3; c++) rt[3 * q + c] = tb.ref[3 * (size_t)th.tile_nodes[q] + c];
    if (rc == CF_OK) rc = put(t.ref_tile, rt.data(), rt.size() * 4);
    if (rc == CF_OK) rc = launch_precompute(t, s);
    // total volume: FP32 serial sum in the caller's element order (as cf_tetset)
    std::vector<float> vol(T), vol_user(T);
    if (rc == CF_OK && T && hipMemcpyAsync(vol.data(), t.vol, T * 4, hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = CF_ERR_HIP;
    if (rc == CF_OK && hipStreamSynchronize(s) != hipSuccess) rc = CF_ERR_HIP;
    if (rc != CF_OK) {
        if (rc == CF_ERR_HIP) set_error("batch mesh upload failed");
        b->arena.free(m.block);
        return rc;
    }
    for (uint32_t i = 0; i < T; i++) vol_user[th.perm[i]] = vol[i];
    float acc = 0.f;
    for (uint32_t i = 0; i < T; i++) acc += vol_user[i];
    m.total_volume = acc;
    m.ref_host.swap(tb.ref);
    b->used += m.bytes;
    const uint32_t id = b->next_mesh++;
    b->meshes.emplace(id, std::move(m));
    *mesh_id = id;
    return CF_OK;
}

int ensure_slots(cf_batch* b, uint32_t need) {
    if (need <= b->cap) return CF_OK;
    uint32_t cap = std::max<uint32_t>(64, b->cap);
    while (cap < need) cap *= 2;
    SimDev* ns = nullptr;
    SimCg* nc = nullptr;
    CF_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(&ns), cap * sizeof(SimDev)));
    if (hipMalloc(reinterpret_cast<void**>(&nc), cap * sizeof(SimCg)) != hipSuccess) {
        (void)hipFree(ns);
        set_error("batch slot arrays: hipMalloc failed");
        return CF_ERR_OUT_OF_MEMORY;
    }
    CF_CHECK_HIP(hipMemsetAsync(nc, 0, cap * sizeof(SimCg), b->stream));
    if (b->cap) {
        CF_CHECK_HIP(hipMemcpyAsync(ns, b->d_sims, b->cap * sizeof(SimDev), hipMemcpyDeviceToDevice, b->stream));
        CF_CHECK_HIP(hipMemcpyAsync(nc, b->d_cgs, b->cap * sizeof(SimCg), hipMemcpyDeviceToDevice, b->stream));
    }
    CF_CHECK_HIP(hipStreamSynchronize(b->stream));
    if (b->d_sims) (void)hipFree(b->d_sims);
    if (b->d_cgs) (void)hipFree(b->d_cgs);
    b->d_sims = ns;
    b->d_cgs = nc;
    b->cap = cap;
    b->dirty = true;
    return CF_OK;
}

int ensure_device(void** p, size_t* cap, size_t need, size_t elem) {
    if (need <= *cap && *p) return CF_OK;
    size_t c = std::max<size_t>(1024, *cap);
    while (c < need) c *= 2;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    CF_CHECK_HIP(hipMalloc(p, c * elem));
    *cap = c;
    return CF_OK;
}

// Block tables of the given simulations, in the given order. The tile table
// is laid out so that the workgroups the hardware deals to one XCD
// (blockIdx % 8) take a contiguous range of tiles. cg_only: only the tables
// of the CG-loop kernels (tile, node, vec) and the slot list.
struct TableLists {
    std::vector<BlockRef> tile, elem, node, init, vec, mv;
    std::vector<uint32_t> slots;
    uint32_t max_tn = 1, max_te = 1, max_it = 0;
    uint64_t tets = 0, nodes = 0;
};

void table_lists(cf_batch* b, const std::vector<uint32_t>& ids, bool cg_only, TableLists& L) {
    std::vector<BlockRef> tile;
    for (uint32_t id : ids) {
        const BSim& s = b->sims[id];
        const BMesh& m = b->meshes[s.mesh];
        const uint32_t sl = s.slot;
        L.slots.push_back(sl);
        for (uint32_t i = 0; i < m.num_tiles; i++) tile.push_back({sl, i});
        for (uint32_t i = 0; i < s.d.nb_node; i++) L.node.push_back({sl, i});
        for (uint32_t i = 0; i < s.d.nb_vec; i++) L.vec.push_back({sl, i});
        for (uint32_t i = 0; i < s.d.nb_mv; i++) L.mv.push_back({sl, i});
        if (!cg_only) {
            for (uint32_t i = 0; i < s.d.nb_elem; i++) L.elem.push_back({sl, i});
            for (uint32_t i = 0; i < s.d.nb_init; i++) L.init.push_back({sl, i});
        }
        L.max_tn = std::max(L.max_tn, m.max_tn);
        L.max_te = std::max(L.max_te, m.max_te);
        L.max_it = std::max(L.max_it, s.d.max_iter);
        L.tets += m.T;
        L.nodes += m.N;
    }
    if (!tile.empty()) {
        const size_t per = (tile.size() + 7) / 8;
        L.tile.assign(8 * per, BlockRef{kIdleSlot, 0});
        for (size_t k = 0; k < L.tile.size(); k++) {
            const size_t nat = (k % 8) * per + k / 8;
            if (nat < tile.size()) L.tile[k] = tile[nat];
        }
    }
}

// Lays the lists out in one buffer (tile, elem, node, init, vec) and points t at it.
void table_views(const TableLists& L, BlockRef* base, uint32_t* active, BatchTables& t) {
    size_t o = 0;
    t.tile = base + o; t.n_tile = uint32_t(L.tile.size()); o += L.tile.size();
    t.elem = base + o; t.n_elem = uint32_t(L.elem.size()); o += L.elem.size();
    t.node = base + o; t.n_node = uint32_t(L.node.size()); o += L.node.size();
    t.init = base + o; t.n_init = uint32_t(L.init.size()); o += L.init.size();
    t.vec = base + o; t.n_vec = uint32_t(L.vec.size()); o += L.vec.size();
    t.mv = base + o; t.n_mv = uint32_t(L.mv.size());
    t.active = active;
    t.n_active = uint32_t(L.slots.size());
}

void table_concat(const TableLists& L, std::vector<BlockRef>& all) {
    all.clear();
    all.reserve(L.tile.size() + L.elem.size() + L.node.size() + L.init.size() + L.vec.size() + L.mv.size());
    all.insert(all.end(), L.tile.begin(), L.tile.end());
    all.insert(all.end(), L.elem.begin(), L.elem.end());
    all.insert(all.end(), L.node.begin(), L.node.end());
    all.insert(all.end(), L.init.begin(), L.init.end());
    all.insert(all.end(), L.vec.begin(), L.vec.end());
    all.insert(all.end(), L.mv.begin(), L.mv.end());
}

// Tables of the tail chunks (lists compacted on the device, in the
// compacted-table buffers at fixed offsets: tiles, matvec chunks, CG vector
// chunks) and their persistent grids: as many workgroups as the CUs hold at
// once (tiles: by the tile kernel's LDS, a multiple of 8 so that a workgroup
// stays on one XCD while it strides), never more than the full tables need.
int dyn_tables(cf_batch* b) {
    const BatchTables& t = b->tab;
    int cus = 0;
    CF_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
    const size_t lds = batch_tile_lds_bytes(b->tile_threads);
    const uint32_t per_cu_tiles = std::max<uint32_t>(1, std::min<uint32_t>(uint32_t(160 * 1024 / lds),
                                                                             2048 / b->tile_threads));
    const uint32_t g_tile = std::max<uint32_t>(8, (uint32_t(cus) * per_cu_tiles) & ~7u);
    const uint32_t g_vec = uint32_t(cus) * 8;
    BatchTables d = t;
    d.tile = b->d_tab_cg;
    d.mv = b->d_tab_cg + t.n_tile;
    d.vec = b->d_tab_cg + t.n_tile + t.n_mv;
    d.n_tile = std::min(t.n_tile, g_tile);
    d.n_mv = std::min(t.n_mv, g_vec);
    d.n_vec = std::min(t.n_vec, g_vec);
    d.elem = d.node = d.init = nullptr;
    d.n_elem = d.n_node = d.n_init = 0;
    d.active = b->d_active_cg;
    d.dyn = b->d_dyn;
    b->tab_dyn = d;
    b->comp = CompactOut{b->d_tab_cg, b->d_tab_cg + t.n_tile, b->d_tab_cg + t.n_tile + t.n_mv, b->d_active_cg,
                         b->d_comp_off, b->d_dyn};
    b->fill_grid = g_vec;
    return CF_OK;
}

int rebuild_tables(cf_batch* b) {
    std::vector<uint32_t> ids;
    for (auto& kv : b->sims)
        if (kv.second.active) ids.push_back(kv.first);
    TableLists L;
    table_lists(b, ids, false, L);
    std::vector<BlockRef> all;
    table_concat(L, all);
    CF_TRY(ensure_device(reinterpret_cast<void**>(&b->d_tab), &b->tab_cap, all.size(), sizeof(BlockRef)));
    CF_TRY(ensure_device(reinterpret_cast<void**>(&b->d_active), &b->active_cap, L.slots.size(), sizeof(uint32_t)));
    // room for the compacted tables (never larger than the full ones)
    CF_TRY(ensure_device(reinterpret_cast<void**>(&b->d_tab_cg), &b->tab_cg_cap, all.size(), sizeof(BlockRef)));
    CF_TRY(ensure_device(reinterpret_cast<void**>(&b->d_active_cg), &b->active_cg_cap, L.slots.size(),
                         sizeof(uint32_t)));
    if (!all.empty())
        CF_CHECK_HIP(hipMemcpyAsync(b->d_tab, all.data(), all.size() * sizeof(BlockRef), hipMemcpyHostToDevice,
                                    b->stream));
    if (!L.slots.empty())
        CF_CHECK_HIP(hipMemcpyAsync(b->d_active, L.slots.data(), L.slots.size() * 4, hipMemcpyHostToDevice,
                                    b->stream));
    CF_CHECK_HIP(hipStreamSynchronize(b->stream));
    if (b->h_cgs_cap < b->cap) {
        for (auto& p : b->h_cgs_pin)
            if (p) (void)hipHostFree(p);
        for (auto& p : b->h_cgs_pin) {
            p = nullptr;
            CF_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(1, b->cap) * sizeof(SimCg)));
        }
        b->h_cgs_cap = b->cap;
    }
    CF_TRY(ensure_device(reinterpret_cast<void**>(&b->d_comp_off), &b->comp_off_cap, 4 * L.slots.size(),
                         sizeof(uint32_t)));
    BatchTables& t = b->tab;
    t.sims = b->d_sims;
    t.cgs = b->d_cgs;
    table_views(L, b->d_tab, b->d_active, t);
    t.counter = b->d_counter;
    t.busy = b->d_counter + 1;
    t.max_tn = L.max_tn;
    t.max_te = L.max_te;
    t.tile_threads = b->tile_threads;
    t.dyn = nullptr;
    CF_TRY(dyn_tables(b));
    b->active_ids.swap(ids);
    b->max_iter = L.max_it;
    b->active_tets = L.tets;
    b->active_nodes = L.nodes;
    free_graphs(b);
    b->dirty = false;
    return CF_OK;
}

// Tables of the CG loop over the simulations still iterating, read from the
// pinned copy of the CG scalars; uploaded on the stream (after the chunk in
// flight) into the compacted buffers.
// A simulation whose last solve met a NaN or Inf in r.r or d.q (its state,
// forces or operator hold non-finite values).
static bool sim_cg_nonfinite(const SimCg& c) {
    return !std::isfinite(c.delta0) || !std::isfinite(c.delta) || !std::isfinite(c.dq);
}

int compact_tables(cf_batch* b, const SimCg* cgs, BatchTables& out, uint32_t* n_sims) {
    std::vector<uint32_t> ids;
    for (uint32_t id : b->active_ids)
        if (!cgs[b->sims[id].slot].done) ids.push_back(id);
    TableLists L;
    table_lists(b, ids, true, L);
    table_concat(L, b->h_tab_cg);
    b->h_active_cg = L.slots;
    if (!b->h_tab_cg.empty())
        CF_CHECK_HIP(hipMemcpyAsync(b->d_tab_cg, b->h_tab_cg.data(), b->h_tab_cg.size() * sizeof(BlockRef),
                                    hipMemcpyHostToDevice, b->stream));
    if (!b->h_active_cg.empty())
        CF_CHECK_HIP(hipMemcpyAsync(b->d_active_cg, b->h_active_cg.data(), b->h_active_cg.size() * 4,
                                    hipMemcpyHostToDevice, b->stream));
    out = b->tab;
    table_views(L, b->d_tab_cg, b->d_active_cg, out);
    *n_sims = uint32_t(ids.size());
    return CF_OK;
}

int enqueue_prologue(cf_batch* b, bool arm_all) {
    const BatchTables& t = b->tab;
    hipStream_t s = b->stream;
    if (arm_all) CF_TRY(launch_batch_arm_all(t, s));      // every active simulation starts a step
    CF_TRY(launch_batch_begin_eval(t, s));               // BeginEvaluation (FixedStep :927)
    CF_TRY(launch_batch_forces(t, kModeFe, kSelY, s));    // f_e(x)
    CF_TRY(launch_batch_gather_rhs(t, s));                // b, y0, KNC (:1209-1295)
    CF_TRY(launch_batch_forces(t, kModeDfx, kSelY, s));   // r = S(b - A y0)
    CF_TRY(launch_batch_gather_matvec(t, kSelY, s));  // fused: r = S(b - A y0), d = r, r.r partials
    return launch_batch_cg_start(t, s);
}

int enqueue_chunk(cf_batch* b, uint32_t iters, const BatchTables& t) {
    hipStream_t s = b->stream;
    for (uint32_t k = 0; k < iters; k++) {
        CF_TRY(launch_batch_forces(t, kModeDfx, kSelD, s));
        CF_TRY(launch_batch_gather_matvec(t, kSelD, s));
        CF_TRY(launch_batch_cg_alpha(t, s));
        CF_TRY(launch_batch_update_xr(t, s));
        CF_TRY(launch_batch_cg_beta(t, s));
        CF_TRY(launch_batch_update_d(t, s));
    }
    return CF_OK;
}

// A chunk of CG iterations of the simulations still iterating, on lists the
// device compacts first: no host work between chunks.
int enqueue_tail_chunk(cf_batch* b, uint32_t iters, const TailGrid& g) {
    CF_TRY(launch_batch_compact(b->tab, b->comp, b->fill_grid, b->stream));
    BatchTables d = b->tab_dyn;
    d.n_tile = std::min(g.tile, b->tab.n_tile);
    d.n_mv = std::min(g.mv, b->tab.n_mv);
    d.n_vec = std::min(g.vec, b->tab.n_vec);
    return enqueue_chunk(b, iters, d);
}

int enqueue_epilogue(cf_batch* b) {
    const BatchTables& t = b->tab;
    hipStream_t s = b->stream;
    CF_TRY(launch_batch_integrate(t, s));  // v = y, x += dt y (:1322-1327)
    CF_TRY(launch_batch_energy(t, s));     // V() (:977)
    return launch_batch_finish(t, s);
}

// One round of iteration-level batching: starting simulations run their
// prologue, iterating ones `iters` CG iterations, those whose solve stopped
// their epilogue (and start their next queued step in the following round).
int enqueue_round(cf_batch* b, uint32_t iters) {
    CF_TRY(enqueue_prologue(b, false));
    CF_TRY(enqueue_chunk(b, iters, b->tab));
    CF_TRY(enqueue_epilogue(b));
    return launch_batch_count_busy(b->tab, b->stream);
}

int record(cf_batch* b, int which, uint32_t iters, hipGraphExec_t* exec, const TailGrid& grid = TailGrid()) {
    CF_CHECK_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = which == 0   ? enqueue_prologue(b, true)
             : which == 1 ? enqueue_chunk(b, iters, b->tab)
             : which == 2 ? enqueue_epilogue(b)
             : which == 4 ? enqueue_tail_chunk(b, iters, grid)
                          : enqueue_round(b, iters);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc != CF_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        set_error("batch graph capture failed: %s", hipGetErrorString(e));
        return CF_ERR_HIP;
    }
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        set_error("batch graph instantiate failed: %s", hipGetErrorString(e));
        return CF_ERR_HIP;
    }
    return CF_OK;
}

int prepare(cf_batch* b, bool step_graphs = false) {
    if (b->dirty) CF_TRY(rebuild_tables(b));
    if (step_graphs && b->use_graphs && b->tab.n_active) {
        const uint32_t chunk = std::min(b->cg_chunk, std::max<uint32_t>(1, b->max_iter));
        if (b->g_chunk && b->g_chunk_len != chunk) free_graphs(b);
        if (!b->g_pro) CF_TRY(record(b, 0, 0, &b->g_pro));
        if (!b->g_chunk) CF_TRY(record(b, 1, chunk, &b->g_chunk));
        if (!b->g_epi) CF_TRY(record(b, 2, 0, &b->g_epi));
        b->g_chunk_len = chunk;
    }
    return CF_OK;
}

void sim_params(BSim& s, const BMesh& m) {
    const cf_solid_params& p = s.p;
    SimDev& d = s.d;
    d.m = material_from(p.fem.young_modulus, p.fem.poisson_ratio);
    // RebuildMass: uniform rho V / N stored as an inverse (LeafDSH_Solid3D_FEM.cpp:1591-1611)
    const float inv_mass = 1.0f / (p.density * m.total_volume / float(m.N));
    d.mass = 1.0f / inv_mass;
    float c[2];
    rayleigh_coeffs(p.fem.rayleigh_freq1, p.fem.rayleigh_freq2, p.fem.damping_ratio, c);
    const float dt = p.time_scale * p.fixed_dt;
    d.dt = dt;
    d.inv_dt = 1.0f / dt;
    d.fM = 1.0f + dt * c[0];
    d.fK = dt * dt + dt * c[1];
    d.drag = -std::log(10e-6f) * p.air_drag_coeff;
    d.thr = p.fem.degenerate_threshold_det_f;
    d.rel = p.cg_rel_epsilon;
    d.abs_eps = p.cg_abs_epsilon;
    d.g[0] = p.gravity[0]; d.g[1] = p.gravity[1]; d.g[2] = p.gravity[2];
    d.rm = p.fem.rm;
    d.dapd_l = p.fem.dapd_l_factor;
    d.dapd_nl = p.fem.dapd_nl_factor;
    d.dapd_e = p.fem.dapd_nl_exponent;
    d.warm_start = p.warm_start;
    d.compute_energy = p.compute_energy;
    d.max_iter = std::min<uint32_t>(p.cg_max_iter, kMaxCgIter);
}

// CG chunks of a step under CF_BATCH_STEP_HOST: the counter and the CG
// scalars written by each chunk are read while the next one runs. Once the
// simulations still iterating are at most half of those in the tables, the
// remaining chunks run on tables of just those, compacted on the host and
// launched directly.
int step_chunks_host(cf_batch* b, uint32_t chunk, uint32_t nchunks) {
    hipStream_t s = b->stream;
    BatchTables cur = b->tab;
    bool compacted = false;
    uint32_t cur_sims = b->tab.n_active;
    const size_t cg_bytes = b->slots_used * sizeof(SimCg);
    for (uint32_t i = 0; i < nchunks; i++) {
        if (b->use_graphs && !compacted) CF_CHECK_HIP(hipGraphLaunch(b->g_chunk, s));
        else CF_TRY(enqueue_chunk(b, chunk, cur));
        CF_CHECK_HIP(hipMemcpyAsync(&b->h_counter[i & 1], b->d_counter, 4, hipMemcpyDeviceToHost, s));
        CF_CHECK_HIP(hipMemcpyAsync(b->h_cgs_pin[i & 1], b->d_cgs, cg_bytes, hipMemcpyDeviceToHost, s));
        CF_CHECK_HIP(hipEventRecord(b->evc[i & 1], s));
        if (i >= 1) {
            CF_CHECK_HIP(hipEventSynchronize(b->evc[(i - 1) & 1]));
            const uint32_t left = b->h_counter[(i - 1) & 1];
            if (left == 0) break;
            if (2 * left <= cur_sims && i + 1 < nchunks) {
                CF_TRY(compact_tables(b, b->h_cgs_pin[(i - 1) & 1], cur, &cur_sims));
                compacted = true;
            }
        }
    }
    return CF_OK;
}

// The whole step submitted at once under the device schedule: prologue,
// chunks on lists the device compacts before each chunk, and the epilogue
// (which only touches simulations whose solve has ended). The chunk count
// and each chunk's grids come from the previous step; the host reads nothing
// until the step has run. Simulations still iterating then (a step that
// needs more chunks than the last one) continue with chunks on grids of the
// persistent sizes, one host read of the iterating count per chunk, and a
// second epilogue ends them.
int launch_tail(cf_batch* b, uint32_t chunk, const TailGrid& g) {
    if (!b->use_graphs) return enqueue_tail_chunk(b, chunk, g);
    auto it = b->g_tails.find(g.key());
    if (it == b->g_tails.end()) {
        hipGraphExec_t e = nullptr;
        CF_TRY(record(b, 4, chunk, &e, g));
        it = b->g_tails.emplace(g.key(), e).first;
    }
    CF_CHECK_HIP(hipGraphLaunch(it->second, b->stream));
    return CF_OK;
}

TailGrid persistent_grid(const cf_batch* b) {
    TailGrid g;
    g.tile = b->tab_dyn.n_tile;
    g.mv = b->tab_dyn.n_mv;
    g.vec = b->tab_dyn.n_vec;
    return g;
}

int step_device(cf_batch* b, uint32_t chunk, uint32_t nchunks) {
    hipStream_t s = b->stream;
    const uint32_t total = b->hist_valid ? std::max(1u, std::min(b->hist_chunks, nchunks)) : 1u;
    for (uint32_t i = 0; i < total; i++)
        CF_TRY(launch_tail(b, chunk, (b->hist_valid && i < b->hist_grids.size()) ? b->hist_grids[i]
                                                                                  : persistent_grid(b)));
    CF_CHECK_HIP(hipMemcpyAsync(&b->h_counter[0], b->d_counter, 4, hipMemcpyDeviceToHost, s));
    if (b->use_graphs) CF_CHECK_HIP(hipGraphLaunch(b->g_epi, s));
    else CF_TRY(enqueue_epilogue(b));
    CF_CHECK_HIP(hipStreamSynchronize(s));
    uint32_t done_chunks = total;
    b->last_continuations = 0;
    if (b->h_counter[0]) {
        while (done_chunks < nchunks) {
            CF_TRY(launch_tail(b, chunk, persistent_grid(b)));
            done_chunks++;
            b->last_continuations++;
            CF_CHECK_HIP(hipMemcpyAsync(&b->h_counter[0], b->d_counter, 4, hipMemcpyDeviceToHost, s));
            CF_CHECK_HIP(hipStreamSynchronize(s));
            if (!b->h_counter[0]) break;
        }
        if (b->use_graphs) CF_CHECK_HIP(hipGraphLaunch(b->g_epi, s));
        else CF_TRY(enqueue_epilogue(b));
    }
    return CF_OK;
}

// The chunk count for the next step from this step's CG iterations (chunks
// until no simulation iterated), and each chunk's grids: the lengths of the lists compacted at the start of
// chunk i are those of the simulations that ran more than i chunks of
// iterations, each rounded up to the next m 2^k with 16 <= m < 32 (at most
// 1/16 above the length, while the recorded tail graphs repeat from step to
// step).
uint32_t grid_class(uint32_t v) {
    if (v <= 32) return std::max(v, 1u);
    uint32_t k = 0;
    while ((v >> k) >= 32) k++;
    const uint32_t m = (v + (1u << k) - 1) >> k;  // in [16, 32]
    return m << k;
}

void update_history(cf_batch* b, uint32_t chunk) {
    std::vector<uint32_t> it;
    it.reserve(b->active_ids.size());
    for (uint32_t id : b->active_ids) it.push_back(b->h_cgs[b->sims[id].slot].iters);
    if (it.empty()) return;
    const uint32_t most = *std::max_element(it.begin(), it.end());
    b->hist_chunks = std::max<uint32_t>(1, (most + chunk - 1) / chunk);
    const std::vector<TailGrid> prev = b->hist_valid ? b->hist_grids : std::vector<TailGrid>();
    b->hist_grids.assign(b->hist_chunks, TailGrid());
    // a chunk keeps its previous grid while the new length fits it with at
    // most a quarter to spare, so that steps replay graphs already recorded
    auto keep = [](uint32_t old_g, uint32_t len, uint32_t g) { return (len <= old_g && 4 * len >= 3 * old_g) ? old_g : g; };
    for (uint32_t i = 0; i < b->hist_chunks; i++) {
        uint64_t tiles = 0, mv = 0, vec = 0;
        for (size_t k = 0; k < it.size(); k++) {
            if (it[k] <= i * chunk) continue;
            const SimDev& d = b->sims[b->active_ids[k]].d;
            tiles += d.num_tiles;
            mv += d.nb_mv;
            vec += d.nb_vec;
        }
        TailGrid& g = b->hist_grids[i];
        // tile grids stay multiples of 8 (a workgroup keeps its XCD while it strides)
        const uint32_t lt = uint32_t(8 * ((tiles + 7) / 8)), lm = uint32_t(mv), lv = uint32_t(vec);
        g.tile = std::min(b->tab.n_tile, 8 * grid_class(std::max(1u, lt / 8)));
        g.mv = std::min(b->tab.n_mv, grid_class(std::max(1u, lm)));
        g.vec = std::min(b->tab.n_vec, grid_class(std::max(1u, lv)));
        if (i < prev.size()) {
            g.tile = std::min(b->tab.n_tile, keep(prev[i].tile, lt, g.tile));
            g.mv = std::min(b->tab.n_mv, keep(prev[i].mv, lm, g.mv));
            g.vec = std::min(b->tab.n_vec, keep(prev[i].vec, lv, g.vec));
        }
    }
    b->hist_valid = true;
}

}  // namespace

extern "C" {

int cf_batch_create(uint64_t budget, cf_batch** out) {
    if (!out) return CF_ERR_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        (void)hipGetLastError();
        set_error("no HIP device");
        return CF_ERR_NO_DEVICE;
    }
    cf_batch* b = new (std::nothrow) cf_batch;
    if (!b) return CF_ERR_OUT_OF_MEMORY;
    b->budget = budget;
    (void)hipGetDevice(&b->device);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&b->d_counter), 16) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&b->d_dyn), kDynCount * sizeof(uint32_t)) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&b->h_counter), 16) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&b->evc[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->evc[1], hipEventDisableTiming) != hipSuccess) {
        set_error("batch: HIP stream/event/memory setup failed");
        cf_batch_destroy(b);
        return CF_ERR_HIP;
    }
    *out = b;
    return CF_OK;
}

int cf_batch_destroy(cf_batch* b) {
    if (!b) return CF_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    free_graphs(b);
    b->arena.release();
    if (b->d_sims) (void)hipFree(b->d_sims);
    if (b->d_cgs) (void)hipFree(b->d_cgs);
    if (b->d_tab) (void)hipFree(b->d_tab);
    if (b->d_active) (void)hipFree(b->d_active);
    if (b->d_counter) (void)hipFree(b->d_counter);
    if (b->d_dyn) (void)hipFree(b->d_dyn);
    if (b->d_comp_off) (void)hipFree(b->d_comp_off);
    if (b->h_counter) (void)hipHostFree(b->h_counter);
    for (auto& p : b->h_cgs_pin)
        if (p) (void)hipHostFree(p);
    if (b->d_tab_cg) (void)hipFree(b->d_tab_cg);
    if (b->d_arm) (void)hipFree(b->d_arm);
    if (b->d_active_cg) (void)hipFree(b->d_active_cg);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    for (auto& e : b->evc)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return CF_OK;
}

int cf_batch_set_tiles(cf_batch* b, uint32_t threads, uint32_t elems) {
    if (!b || (threads != 128 && threads != 256 && threads != 512 && threads != 1024) || elems > threads) {
        set_error("cf_batch_set_tiles: threads must be 128/256/512/1024 and elems <= threads");
        return CF_ERR_INVALID_ARG;
    }
    if (!b->meshes.empty() && threads != b->tile_threads) {
        set_error("cf_batch_set_tiles: the tile threads are fixed once meshes exist");
        return CF_ERR_INVALID_ARG;
    }
    b->tile_threads = threads;
    b->tile_elems = elems ? elems : threads;
    return CF_OK;
}

int cf_batch_set_schedule(cf_batch* b, uint32_t chunk, int use_graphs) {
    if (!b || chunk < 1 || chunk > kMaxCgIter) return CF_ERR_INVALID_ARG;
    b->cg_chunk = chunk;
    b->use_graphs = use_graphs ? 1 : 0;
    b->hist_valid = false;  // chunk counts are in chunks of the old length
    free_graphs(b);
    return CF_OK;
}

int cf_batch_set_step_mode(cf_batch* b, int mode) {
    if (!b || (mode != CF_BATCH_STEP_HOST && mode != CF_BATCH_STEP_DEVICE)) {
        set_error("cf_batch_set_step_mode: mode must be CF_BATCH_STEP_HOST or CF_BATCH_STEP_DEVICE");
        return CF_ERR_INVALID_ARG;
    }
    b->step_mode = mode;
    b->hist_valid = false;
    free_graphs(b);
    return CF_OK;
}

int cf_batch_add_mesh(cf_batch* b, uint32_t N, const float* ref, uint32_t T, const uint32_t* tets, uint32_t* id) {
    if (!b || !ref || !tets || !id || N == 0 || T == 0) { set_error("cf_batch_add_mesh: invalid arguments"); return CF_ERR_INVALID_ARG; }
    TileBuild tb;
    CF_TRY(build_mesh_host(N, ref, T, tets, b->tile_elems, b->tile_threads, tb));
    return upload_mesh(b, tb, id);
}

int cf_batch_add_meshes(cf_batch* b, uint32_t count, const uint32_t* N, const float* const* ref, const uint32_t* T,
                        const uint32_t* const* tets, uint32_t threads, uint32_t* ids) {
    if (!b || !N || !ref || !T || !tets || !ids) { set_error("cf_batch_add_meshes: invalid arguments"); return CF_ERR_INVALID_ARG; }
    for (uint32_t i = 0; i < count; i++)
        if (!ref[i] || !tets[i] || N[i] == 0 || T[i] == 0) { set_error("cf_batch_add_meshes: mesh %u is empty", i); return CF_ERR_INVALID_ARG; }
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    threads = std::min(threads, std::max(1u, count));
    std::vector<TileBuild> builds(count);
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        for (uint32_t i = next++; i < count; i = next++)
            builds[i].rc = build_mesh_host(N[i], ref[i], T[i], tets[i], b->tile_elems, b->tile_threads, builds[i]);
    };
    std::vector<std::thread> pool;
    for (uint32_t k = 1; k < threads; k++) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    for (uint32_t i = 0; i < count; i++) {
        if (builds[i].rc != CF_OK) { set_error("cf_batch_add_meshes: mesh %u is invalid", i); return builds[i].rc; }
        CF_TRY(upload_mesh(b, builds[i], &ids[i]));
        builds[i] = TileBuild();
    }
    return CF_OK;
}

int cf_batch_remove_mesh(cf_batch* b, uint32_t id) {
    if (!b) return CF_ERR_INVALID_ARG;
    auto it = b->meshes.find(id);
    if (it == b->meshes.end()) { set_error("no mesh %u", id); return CF_ERR_INVALID_ARG; }
    if (it->second.users) { set_error("mesh %u is used by %u simulations", id, it->second.users); return CF_ERR_INVALID_ARG; }
    CF_CHECK_HIP(hipStreamSynchronize(b->stream));
    b->arena.free(it->second.block);
    b->used -= it->second.bytes;
    b->meshes.erase(it);
    return CF_OK;
}

int cf_batch_sim_bytes(const cf_batch* b, uint32_t mesh_id, uint64_t* bytes) {
    if (!b || !bytes) return CF_ERR_INVALID_ARG;
    auto it = b->meshes.find(mesh_id);
    if (it == b->meshes.end()) { set_error("no mesh %u", mesh_id); return CF_ERR_INVALID_ARG; }
    *bytes = SimLayout(it->second.N, it->second.T, it->second.num_parts, it->second.num_tiles).total;
    return CF_OK;
}

int cf_batch_add_sim(cf_batch* b, uint32_t mesh_id, const cf_solid_params* p, uint32_t n_knc, const uint32_t* knc,
                     const float* knc_tgt, const float* pos, const float* vel, uint32_t* sim_id) {
    if (!b || !p || !sim_id || (n_knc && !knc)) { set_error("cf_batch_add_sim: invalid arguments"); return CF_ERR_INVALID_ARG; }
    auto mit = b->meshes.find(mesh_id);
    if (mit == b->meshes.end()) { set_error("no mesh %u", mesh_id); return CF_ERR_INVALID_ARG; }
    BMesh& m = mit->second;
    if (p->fem.mm != CF_MM_C_WRP) {
        set_error("%s: only the C_WRP material is batched", "cf_batch_add_sim");
        return CF_ERR_UNSUPPORTED;
    }
    if (p->integrator != CF_IT_QIE_V) {
        set_error("%s: only the QIE-v integrator is batched", "cf_batch_add_sim");
        return CF_ERR_UNSUPPORTED;
    }
    if (!(p->fixed_dt > 0) || !(p->density > 0) || !(p->time_scale > 0)) { set_error("fixed_dt, time_scale and density must be > 0"); return CF_ERR_INVALID_ARG; }
    if (p->reference_order) { set_error("batches take the fast path (reference_order = 0)"); return CF_ERR_UNSUPPORTED; }
    if (!(p->fem.young_modulus > 0) || !(p->fem.poisson_ratio > -1.0f) || !(p->fem.poisson_ratio < 0.5f) ||
        p->fem.mm != CF_MM_C_WRP || p->fem.dm != CF_DM_TRUNCATED ||
        (p->fem.rm != CF_RM_ID && p->fem.rm != CF_RM_MSVD && p->fem.rm != CF_RM_DAPD)) {
        set_error("cf_batch_add_sim: unsupported material or method");
        return CF_ERR_INVALID_ARG;
    }
    for (uint32_t i = 0; i < n_knc; i++)
        if (knc[i] >= m.N) { set_error("KNC node %u out of range", knc[i]); return CF_ERR_INVALID_ARG; }
    const SimLayout L(m.N, m.T, m.num_parts, m.num_tiles);
    if (b->budget && b->used + L.total > b->budget) {
        set_error("simulation of %zu bytes exceeds the HBM budget (%llu of %llu in use)", L.total,
                  (unsigned long long)b->used, (unsigned long long)b->budget);
        return CF_ERR_CAPACITY;
    }
    uint32_t slot;
    if (!b->free_slots.empty()) {
        slot = b->free_slots.back();
    } else {
        CF_TRY(ensure_slots(b, b->slots_used + 1));
        slot = b->slots_used;
    }
    BSim s;
    s.mesh = mesh_id;
    s.slot = slot;
    s.p = *p;
    CF_TRY(b->arena.alloc(L.total, &s.block));
    s.bytes = L.total;
    SimDev& d = s.d;
    std::memset(&d, 0, sizeof(d));
    const TetDevice& t = m.dev;
    d.N = m.N; d.T = m.T; d.num_tiles = m.num_tiles;
    d.nb_elem = L.nb_elem; d.nb_node = L.nb_node; d.nb_init = L.nb_init; d.nb_vec = L.nb_vec;
    d.nb_mv = L.nb_mv; d.npt_mv = matvec_npt(m.N); d.rows_cg = cg_rows(3 * m.N);
    d.conn = t.conn; d.tile_elem_ptr = t.tile_elem_ptr; d.tile_node_ptr = t.tile_node_ptr;
    d.tile_nodes = t.tile_nodes; d.slot_ptr = t.slot_ptr; d.node_part_ptr = t.node_part_ptr;
    d.part_dst = t.part_dst; d.lv = t.lv; d.rk = t.rk; d.ref = t.ref; d.dinv = t.dinv; d.vol = t.vol;
    d.ref_tile = t.ref_tile;
    float* vecs[8];
    for (int i = 0; i < 8; i++) vecs[i] = at<float>(s.block, L.vec[i]);
    d.x = vecs[0]; d.v = vecs[1]; d.b = vecs[2]; d.y = vecs[3];
    d.r = vecs[4]; d.d = vecs[5]; d.q = vecs[6]; d.knc_target = vecs[7];
    d.R = at<float>(s.block, L.R);
    d.detF = at<float>(s.block, L.detF);
    d.partials = at<float>(s.block, L.part);
    uint8_t* mask = at<uint8_t>(s.block, L.mask);
    d.knc_mask = n_knc ? mask : nullptr;
    d.part_a = at<double>(s.block, L.pa);
    d.part_b = at<double>(s.block, L.pb);
    d.part_e = at<double>(s.block, L.pe);
    d.eval_count = at<uint32_t>(s.block, L.ec);
    d.eval_sum = at<double>(s.block, L.es);
    d.eval_min = at<float>(s.block, L.em);
    if (p->fem.rm == CF_RM_DAPD) {
        d.x0 = at<float>(s.block, L.x0);
        d.doc = at<uint8_t>(s.block, L.doc);
    }
    sim_params(s, m);
    // initial state, KNC, and R = Id (CreateCache, TetrahedronElementSet3.cpp:190-209)
    hipStream_t st = b->stream;
    const size_t n3 = 3 * (size_t)m.N;
    std::vector<uint8_t> hmask(m.N, 0);
    std::vector<float> tgt(n3, 0.f);
    for (uint32_t i = 0; i < n_knc; i++) {
        hmask[knc[i]] = 1;
        for (int c = 0; c < 3; c++) tgt[3 * (size_t)knc[i] + c] = knc_tgt ? knc_tgt[3 * (size_t)i + c] : m.ref_host[3 * (size_t)knc[i] + c];
    }
    std::vector<float> Rid(9 * (size_t)m.T, 0.f);
    for (size_t e = 0; e < m.T; e++) { Rid[e] = 1.f; Rid[4 * (size_t)m.T + e] = 1.f; Rid[8 * (size_t)m.T + e] = 1.f; }
    int rc = CF_OK;
    auto chk = [&](hipError_t e) { if (rc == CF_OK && e != hipSuccess) { set_error("batch sim upload: %s", hipGetErrorString(e)); rc = CF_ERR_HIP; } };
    chk(hipMemsetAsync(s.block, 0, L.total, st));
    chk(hipMemcpyAsync(d.x, pos ? pos : m.ref_host.data(), n3 * 4, hipMemcpyHostToDevice, st));
    if (vel) chk(hipMemcpyAsync(d.v, vel, n3 * 4, hipMemcpyHostToDevice, st));
    chk(hipMemcpyAsync(mask, hmask.data(), m.N, hipMemcpyHostToDevice, st));
    chk(hipMemcpyAsync(d.knc_target, tgt.data(), n3 * 4, hipMemcpyHostToDevice, st));
    chk(hipMemcpyAsync(d.R, Rid.data(), Rid.size() * 4, hipMemcpyHostToDevice, st));
    if (d.x0) {  // m_vecPos0 starts at the rest pose like a leaf's (a given initial
                 // state is a set_state, which leaves it), every DoC invalid
        chk(hipMemcpyAsync(d.x0, m.ref_host.data(), n3 * 4, hipMemcpyHostToDevice, st));
        chk(hipMemsetAsync(d.doc, kDocInvalid, m.T, st));
    }
    chk(hipMemcpyAsync(b->d_sims + slot, &d, sizeof(SimDev), hipMemcpyHostToDevice, st));
    chk(hipMemsetAsync(b->d_cgs + slot, 0, sizeof(SimCg), st));  // idle, no steps done
    chk(hipStreamSynchronize(st));
    if (rc != CF_OK) {
        b->arena.free(s.block);
        return rc;
    }
    if (!b->free_slots.empty() && b->free_slots.back() == slot) b->free_slots.pop_back();
    else b->slots_used++;
    b->used += s.bytes;
    m.users++;
    const uint32_t id = b->next_sim++;
    b->sims.emplace(id, s);
    b->dirty = true;
    *sim_id = id;
    return CF_OK;
}

int cf_batch_remove_sim(cf_batch* b, uint32_t id) {
    if (!b) return CF_ERR_INVALID_ARG;
    auto it = b->sims.find(id);
    if (it == b->sims.end()) { set_error("no simulation %u", id); return CF_ERR_INVALID_ARG; }
    CF_CHECK_HIP(hipStreamSynchronize(b->stream));
    b->arena.free(it->second.block);
    b->used -= it->second.bytes;
    b->meshes[it->second.mesh].users--;
    b->free_slots.push_back(it->second.slot);
    const uint32_t slot = it->second.slot;
    b->pending.erase(std::remove_if(b->pending.begin(), b->pending.end(),
                                    [slot](const uint2& p) { return p.x == slot; }),
                     b->pending.end());
    if (it->second.active) b->dirty = true;
    b->sims.erase(it);
    return CF_OK;
}

int cf_batch_set_active(cf_batch* b, uint32_t id, int active) {
    if (!b) return CF_ERR_INVALID_ARG;
    auto it = b->sims.find(id);
    if (it == b->sims.end()) { set_error("no simulation %u", id); return CF_ERR_INVALID_ARG; }
    const bool a = active != 0;
    if (it->second.active != a) { it->second.active = a; b->dirty = true; }
    return CF_OK;
}

int cf_batch_get_state(cf_batch* b, uint32_t id, float* pos, float* vel) {
    if (!b) return CF_ERR_INVALID_ARG;
    auto it = b->sims.find(id);
    if (it == b->sims.end()) { set_error("no simulation %u", id); return CF_ERR_INVALID_ARG; }
    const size_t n3 = 3 * (size_t)it->second.d.N;
    if (pos) CF_CHECK_HIP(hipMemcpyAsync(pos, it->second.d.x, n3 * 4, hipMemcpyDeviceToHost, b->stream));
    if (vel) CF_CHECK_HIP(hipMemcpyAsync(vel, it->second.d.v, n3 * 4, hipMemcpyDeviceToHost, b->stream));
    CF_CHECK_HIP(hipStreamSynchronize(b->stream));
    return CF_OK;
}

int cf_batch_set_state(cf_batch* b, uint32_t id, const float* pos, const float* vel) {
    if (!b) return CF_ERR_INVALID_ARG;
    auto it = b->sims.find(id);
    if (it == b->sims.end()) { set_error("no simulation %u", id); return CF_ERR_INVALID_ARG; }
    const size_t n3 = 3 * (size_t)it->second.d.N;
    if (pos) CF_CHECK_HIP(hipMemcpyAsync(it->second.d.x, pos, n3 * 4, hipMemcpyHostToDevice, b->stream));
    if (vel) CF_CHECK_HIP(hipMemcpyAsync(it->second.d.v, vel, n3 * 4, hipMemcpyHostToDevice, b->stream));
    CF_CHECK_HIP(hipStreamSynchronize(b->stream));
    return CF_OK;
}

int cf_batch_fixed_step(cf_batch* b, cf_step_stats* stats) {
    if (!b) return CF_ERR_INVALID_ARG;
    CF_TRY(prepare(b, true));
    const BatchTables& t = b->tab;
    if (!t.n_active) return CF_OK;
    hipStream_t s = b->stream;
    CF_CHECK_HIP(hipEventRecord(b->ev0, s));
    if (b->use_graphs) CF_CHECK_HIP(hipGraphLaunch(b->g_pro, s));
    else CF_TRY(enqueue_prologue(b, true));
    const uint32_t chunk = b->use_graphs ? b->g_chunk_len : std::min(b->cg_chunk, std::max<uint32_t>(1, b->max_iter));
    const uint32_t nchunks = (b->max_iter + chunk - 1) / chunk;
    if (b->step_mode == CF_BATCH_STEP_HOST) {
        CF_TRY(step_chunks_host(b, chunk, nchunks));
        if (b->use_graphs) CF_CHECK_HIP(hipGraphLaunch(b->g_epi, s));
        else CF_TRY(enqueue_epilogue(b));
    } else {
        CF_TRY(step_device(b, chunk, nchunks));
    }
    CF_CHECK_HIP(hipEventRecord(b->ev1, s));
    CF_CHECK_HIP(hipEventSynchronize(b->ev1));
    // every step's CG records: the stats, and the non-finite test
    b->h_cgs.resize(b->slots_used);
    CF_CHECK_HIP(hipMemcpy(b->h_cgs.data(), b->d_cgs, b->slots_used * sizeof(SimCg), hipMemcpyDeviceToHost));
    if (b->step_mode != CF_BATCH_STEP_HOST) update_history(b, chunk);
    uint32_t nonfinite = 0;
    for (uint32_t id : b->active_ids) nonfinite += sim_cg_nonfinite(b->h_cgs[b->sims[id].slot]) ? 1 : 0;
    if (stats) {
        float ms = 0.f;
        CF_CHECK_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        for (size_t i = 0; i < b->active_ids.size(); i++) {
            const BSim& sim = b->sims[b->active_ids[i]];
            const SimCg& c = b->h_cgs[sim.slot];
            cf_step_stats& o = stats[i];
            o.num_fixed_steps = 1;
            o.cg_iters = int32_t(c.iters);
            o.cg_reason = sim_cg_nonfinite(c) ? CF_CG_NONFINITE : c.reason;
            o.cg_rel_residual = (c.delta0 > 0) ? float(std::sqrt(c.delta / c.delta0)) : 0.f;
            o.cg_abs_residual = float(std::sqrt(c.delta));
            o.num_degenerate = c.num_degenerate;
            o.sum_degenerate_det_f = float(c.sum_degenerate);
            o.elastic_energy = float(c.energy);
            o.min_det_f = c.min_det;
            o.step_ms = ms;
        }
    }
    if (nonfinite) {
        set_error("batched step: %u simulation(s) with a NaN or Inf CG reduction", nonfinite);
        return CF_ERR_NAN;
    }
    return CF_OK;
}

int cf_batch_request_steps(cf_batch* b, uint32_t id, uint32_t steps) {
    if (!b) return CF_ERR_INVALID_ARG;
    auto it = b->sims.find(id);
    if (it == b->sims.end()) { set_error("no simulation %u", id); return CF_ERR_INVALID_ARG; }
    if (steps) b->pending.push_back(make_uint2(it->second.slot, steps));
    return CF_OK;
}

int cf_batch_round(cf_batch* b, uint32_t iters, uint32_t* steps_done, uint32_t* busy) {
    if (!b || iters == 0 || iters > kMaxCgIter) return CF_ERR_INVALID_ARG;
    CF_TRY(prepare(b));
    const BatchTables& t = b->tab;
    hipStream_t s = b->stream;
    if (!b->pending.empty()) {
        CF_TRY(ensure_device(reinterpret_cast<void**>(&b->d_arm), &b->arm_cap, b->pending.size(), sizeof(uint2)));
        CF_CHECK_HIP(hipMemcpyAsync(b->d_arm, b->pending.data(), b->pending.size() * sizeof(uint2),
                                    hipMemcpyHostToDevice, s));
        CF_TRY(launch_batch_arm(t, b->d_arm, uint32_t(b->pending.size()), s));
        CF_CHECK_HIP(hipStreamSynchronize(s));  // the host list is reused
        b->pending.clear();
    }
    if (t.n_active) {
        if (b->use_graphs) {
            if (b->g_round && b->g_round_len != iters) {
                (void)hipGraphExecDestroy(b->g_round);
                b->g_round = nullptr;
            }
            if (!b->g_round) CF_TRY(record(b, 3, iters, &b->g_round));
            b->g_round_len = iters;
            CF_CHECK_HIP(hipGraphLaunch(b->g_round, s));
        } else {
            CF_TRY(enqueue_round(b, iters));
        }
    }
    const size_t cg_bytes = b->slots_used * sizeof(SimCg);
    CF_CHECK_HIP(hipMemcpyAsync(&b->h_counter[2], b->d_counter + 1, 4, hipMemcpyDeviceToHost, s));
    CF_CHECK_HIP(hipMemcpyAsync(b->h_cgs_pin[0], b->d_cgs, cg_bytes, hipMemcpyDeviceToHost, s));
    CF_CHECK_HIP(hipStreamSynchronize(s));
    if (steps_done)
        for (size_t i = 0; i < b->active_ids.size(); i++)
            steps_done[i] = b->h_cgs_pin[0][b->sims[b->active_ids[i]].slot].steps_done;
    if (busy) *busy = t.n_active ? b->h_counter[2] : 0;
    uint32_t nonfinite = 0;
    for (uint32_t id : b->active_ids) nonfinite += sim_cg_nonfinite(b->h_cgs_pin[0][b->sims[id].slot]) ? 1 : 0;
    if (nonfinite) {
        set_error("batch round: %u simulation(s) with a NaN or Inf CG reduction", nonfinite);
        return CF_ERR_NAN;
    }
    return CF_OK;
}

int cf_batch_active_ids(const cf_batch* b, uint32_t* ids) {
    if (!b || !ids) return CF_ERR_INVALID_ARG;
    CF_TRY(prepare(const_cast<cf_batch*>(b)));
    std::copy(b->active_ids.begin(), b->active_ids.end(), ids);
    return CF_OK;
}

int cf_batch_info_get(const cf_batch* b, cf_batch_info* info) {
    if (!b || !info) return CF_ERR_INVALID_ARG;
    CF_TRY(prepare(const_cast<cf_batch*>(b)));
    info->num_meshes = uint32_t(b->meshes.size());
    info->num_sims = uint32_t(b->sims.size());
    info->num_active = uint32_t(b->active_ids.size());
    info->bytes_in_use = b->used;
    info->bytes_reserved = b->arena.reserved();
    info->budget_bytes = b->budget;
    info->active_tets = b->active_tets;
    info->active_nodes = b->active_nodes;
    info->tile_threads = b->tile_threads;
    info->tile_elems = b->tile_elems;
    info->cg_chunk = b->cg_chunk;
    info->use_graphs = uint32_t(b->use_graphs);
    info->step_mode = uint32_t(b->step_mode);
    info->next_chunks = b->hist_valid ? b->hist_chunks : 1;
    info->last_continuations = b->last_continuations;
    info->chunk_graphs = uint32_t(b->g_tails.size());
    return CF_OK;
}

int cf_batch_time_kx(cf_batch* b, uint32_t reps, double* ms) {
    if (!b || !ms || reps == 0) return CF_ERR_INVALID_ARG;
    CF_TRY(prepare(b));
    const BatchTables& t = b->tab;
    hipStream_t s = b->stream;
    CF_TRY(launch_batch_forces(t, kModeDfx, kSelYAll, s));
    CF_TRY(launch_batch_gather_matvec(t, kSelYAll, s));
    CF_CHECK_HIP(hipEventRecord(b->ev0, s));
    for (uint32_t i = 0; i < reps; i++) {
        CF_TRY(launch_batch_forces(t, kModeDfx, kSelYAll, s));
        CF_TRY(launch_batch_gather_matvec(t, kSelYAll, s));
    }
    CF_CHECK_HIP(hipEventRecord(b->ev1, s));
    CF_CHECK_HIP(hipEventSynchronize(b->ev1));
    float tot = 0.f;
    CF_CHECK_HIP(hipEventElapsedTime(&tot, b->ev0, b->ev1));
    *ms = double(tot) / reps;
    return CF_OK;
}

int cf_batch_synchronize(cf_batch* b) {
    if (!b) return CF_ERR_INVALID_ARG;
    CF_CHECK_HIP(hipStreamSynchronize(b->stream));
    return CF_OK;
}

}  // extern "C"
