// Real-time serving loop over a batch (cf_serve_trace in include/corofem.h):
// arrivals, FIFO admission by HBM budget and a cap on the tets in flight
// (fixed, or adapted to a target round time),
// iteration-level batching (every round runs, in the same launches, the step
// prologues, a few CG iterations and the step epilogues of all admitted
// requests, so each step completes in the round its own solve stops),
// departures that free their memory, lazily uploaded and LRU-evicted meshes,
// per-step latencies.
#include "cf_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

using clk = std::chrono::steady_clock;

struct CachedMesh {
    uint32_t id = 0, users = 0;
    double last_used = 0.0;
};

struct Live {
    uint32_t req, sim, done;
    double enqueued;
};

double percentile(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    size_t k = size_t(std::ceil(p * double(v.size())));
    k = k ? k - 1 : 0;
    return v[std::min(k, v.size() - 1)];
}

// Evicts the least recently used mesh that no live request uses (not `keep`).
bool evict_one(cf_batch* b, std::map<uint32_t, CachedMesh>& cache, uint32_t keep) {
    auto victim = cache.end();
    for (auto it = cache.begin(); it != cache.end(); ++it)
        if (it->first != keep && it->second.users == 0 &&
            (victim == cache.end() || it->second.last_used < victim->second.last_used))
            victim = it;
    if (victim == cache.end()) return false;
    if (cf_batch_remove_mesh(b, victim->second.id) != CF_OK) return false;
    cache.erase(victim);
    return true;
}

}  // namespace

extern "C" int cf_serve_trace(const cf_serve_mesh* meshes, uint32_t n_meshes, const cf_serve_request* reqs,
                              uint32_t n_reqs, const cf_serve_config* cfg, float* const* final_pos,
                              double* lat_out, cf_serve_result* out) {
    using namespace cf;
    if (!meshes || !reqs || !cfg || !out || n_meshes == 0) {
        set_error("cf_serve_trace: invalid arguments");
        return CF_ERR_INVALID_ARG;
    }
    for (uint32_t i = 0; i < n_reqs; i++) {
        if (reqs[i].mesh >= n_meshes || reqs[i].steps == 0 || (reqs[i].n_knc && !reqs[i].knc_nids) ||
            (i && reqs[i].arrival_s < reqs[i - 1].arrival_s)) {
            set_error("cf_serve_trace: request %u is invalid (mesh, steps, KNC, or arrivals out of order)", i);
            return CF_ERR_INVALID_ARG;
        }
    }
    cf_batch* b = nullptr;
    CF_TRY(cf_batch_create(cfg->hbm_budget_bytes, &b));
    struct Guard {
        cf_batch* b;
        ~Guard() { cf_batch_destroy(b); }
    } guard{b};
    CF_TRY(cf_batch_set_tiles(b, cfg->tile_threads ? cfg->tile_threads : 256, 0));
    CF_TRY(cf_batch_set_schedule(b, cfg->cg_chunk ? cfg->cg_chunk : 8, cfg->use_graphs));

    cf_serve_result r{};
    std::map<uint32_t, CachedMesh> cache;  // mesh index -> batch mesh
    std::deque<uint32_t> waiting;
    std::vector<Live> live;
    std::vector<double> lat, first;
    std::vector<uint32_t> ids, done;
    lat.reserve(1024);
    uint64_t live_tets = 0;
    uint32_t next = 0;
    uint32_t eligible = 0;
    while (eligible < n_reqs && reqs[eligible].arrival_s <= cfg->duration_s) eligible++;
    r.requests = eligible;
    const clk::time_point t0 = clk::now();
    auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
    double tick_total = 0.0;
    // the tets cap on admission, adapted to the measured round time when a
    // target is set
    uint64_t cap = cfg->max_tick_tets;
    const uint64_t cap_min = std::max<uint64_t>(1, cap / 64);
    const double target = cfg->target_round_ms;

    while (true) {
        double t = now();
        while (next < eligible && reqs[next].arrival_s <= t) waiting.push_back(next++);
        // admission, in arrival order
        bool cap_blocked = false;
        while (!waiting.empty()) {
            const cf_serve_request& q = reqs[waiting.front()];
            const cf_serve_mesh& m = meshes[q.mesh];
            if (cap && !live.empty() && live_tets + m.n_tets > cap) {
                cap_blocked = true;
                break;
            }
            auto it = cache.find(q.mesh);
            if (it == cache.end()) {
                uint32_t mid = 0;
                int rc;
                while ((rc = cf_batch_add_mesh(b, m.n_nodes, m.ref_pos, m.n_tets, m.tets, &mid)) == CF_ERR_CAPACITY)
                    if (!evict_one(b, cache, q.mesh)) break;
                if (rc == CF_ERR_CAPACITY) break;
                if (rc != CF_OK) return rc;
                it = cache.emplace(q.mesh, CachedMesh{mid, 0, t}).first;
                r.mesh_uploads++;
            }
            cf_solid_params p = cfg->base;
            p.fem.young_modulus = q.young_modulus;
            p.fem.poisson_ratio = q.poisson_ratio;
            uint32_t sim = 0;
            int rc;
            while ((rc = cf_batch_add_sim(b, it->second.id, &p, q.n_knc, q.knc_nids, nullptr, nullptr, nullptr,
                                          &sim)) == CF_ERR_CAPACITY)
                if (!evict_one(b, cache, q.mesh)) break;
            if (rc == CF_ERR_CAPACITY) break;
            if (rc != CF_OK) return rc;
            CF_TRY(cf_batch_request_steps(b, sim, q.steps));
            it->second.users++;
            live.push_back(Live{waiting.front(), sim, 0, q.arrival_s});
            live_tets += m.n_tets;
            waiting.pop_front();
            r.admitted++;
        }
        r.max_active = std::max<uint32_t>(r.max_active, uint32_t(live.size()));
        if (live.empty()) {
            if (!waiting.empty()) {
                // nothing runs and the head still does not fit: it never will
                set_error("request %u does not fit the HBM budget on its own", waiting.front());
                return CF_ERR_CAPACITY;
            }
            if (next >= eligible) break;
            if (!cfg->drain && t >= cfg->duration_s) break;
            const double wait = reqs[next].arrival_s - now();
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
            continue;
        }
        if (!cfg->drain && t >= cfg->duration_s) break;
        // one round: prologues, cfg->cg_chunk CG iterations and epilogues of
        // every admitted request; a request's step completes in the round its
        // own solve stops
        cf_batch_info info{};
        CF_TRY(cf_batch_info_get(b, &info));
        r.peak_bytes = std::max<uint64_t>(r.peak_bytes, info.bytes_in_use);
        const double ts = now();
        uint32_t busy = 0;
        ids.resize(info.num_active);
        done.resize(info.num_active);
        CF_TRY(cf_batch_round(b, cfg->cg_chunk ? cfg->cg_chunk : 8, done.data(), &busy));
        CF_TRY(cf_batch_active_ids(b, ids.data()));
        const double te = now();
        tick_total += te - ts;
        r.ticks++;
        if (target > 0 && cap) {
            const double ms = (te - ts) * 1e3;
            if (cap_blocked && ms < 0.8 * target) cap += cap / 4;
            else if (ms > 1.25 * target) cap = std::max(cap_min, cap - cap / 5);
        }
        std::unordered_map<uint32_t, uint32_t> done_of;
        for (size_t i = 0; i < ids.size(); i++) done_of[ids[i]] = done[i];
        std::vector<Live> keep;
        keep.reserve(live.size());
        for (Live& l : live) {
            const cf_serve_request& q = reqs[l.req];
            const uint32_t now_done = done_of[l.sim];
            for (; l.done < now_done; l.done++) {
                lat.push_back((te - l.enqueued) * 1e3);
                if (l.done == 0) first.push_back((te - q.arrival_s) * 1e3);
                l.enqueued = te;
                r.steps++;
                r.tet_steps += meshes[q.mesh].n_tets;
            }
            if (l.done < q.steps) {
                keep.push_back(l);
                continue;
            }
            if (final_pos && final_pos[l.req]) CF_TRY(cf_batch_get_state(b, l.sim, final_pos[l.req], nullptr));
            CF_TRY(cf_batch_remove_sim(b, l.sim));
            CachedMesh& cm = cache[q.mesh];
            cm.users--;
            cm.last_used = te;
            live_tets -= meshes[q.mesh].n_tets;
            r.completed++;
        }
        live.swap(keep);
    }
    r.wall_s = now();
    r.steps_per_s = r.wall_s > 0 ? double(r.steps) / r.wall_s : 0.0;
    r.tick_mean_ms = r.ticks ? tick_total / r.ticks * 1e3 : 0.0;
    r.step_p50_ms = percentile(lat, 0.50);
    r.step_p99_ms = percentile(lat, 0.99);
    r.step_max_ms = lat.empty() ? 0.0 : *std::max_element(lat.begin(), lat.end());
    double s = 0.0;
    for (double v : lat) s += v;
    r.step_mean_ms = lat.empty() ? 0.0 : s / double(lat.size());
    r.first_p50_ms = percentile(first, 0.50);
    r.first_p99_ms = percentile(first, 0.99);
    r.final_tick_tets = cap;
    if (lat_out) std::copy(lat.begin(), lat.end(), lat_out);
    *out = r;
    return CF_OK;
}
