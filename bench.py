"""Benchmark: the corotational FEM implicit step on one MI355X.

Default run (one JSON line): the four workloads of the hot path.
  single      the large mesh, plain GPU drop-in -- a 64x64x256 cube grid
              (65x65x257 vertices, 5,242,880 tets / 1,085,825 nodes), edge
              0.01 m, E=1e5 Pa, nu=0.3, rho=1000, dt=1/60 s, gravity -9.81 y, CG
              rel 1e-4 / abs 1e-6 / max 50, polar-decomposition rotations, FP32
              state with FP64 CG reductions; initial pose one seeded rigid
              rotation times a stretch in [0.7, 1.3] plus 1% jitter, z-min face
              held. One "step" = one QIE-v fixed step. Top-level `value` =
              tet-matvecs per second over the timed steps; `roofline` = one K*x
              (element tile kernel + node gather) timed with HIP events on the
              solid's stream; `cpu_baseline` = the oracle's K*x (stored Ke,
              serial scatter, one core) on a bounded shard of the same mesh,
              with the host's CPU model and, for config 1 (the 20,000-tet
              beam), per-step medians of the oracle on one core and of the
              GPU leaf (fast path and reference order).
  batched_2d  65,536 2D grids (n x n, n in [4,32]) stepped by ONE launch per
              step, against the plain single-mesh path on every 64th mesh.
  batched_3d  1,024 uneven 3D cube grids (4-40 cubes per axis) in one
              graph-recorded batched step, against the plain path on every 8th.
              Both batches run under two CG settings (the reference's abs
              floor 1e-6, and the relative test alone, where every solve
              stops at rel 1e-4: the headline), each with its achieved
              relative residuals; `cpu_baseline` = the oracle stepping a
              sample of the meshes on all host threads (up to 16).
  serving     Poisson arrivals at 100 / 1k / 10k requests/s, HBM budget 200 GB,
              continuous batching; per-step latency p50/p99.
`--workload single|batch2d|batch3d|serve` runs one of them alone.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from risdsdgcfem_thesis_amd.workloads import large_mesh_state  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "K·x Gtets/s (%HBM peak); steps/s batched vs plain at 1e-4 residual; p99 ms"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_traffic():
    p = os.path.join(ROOT, "profiles", "kx_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def host_cpu():
    """The host's CPU model and the cores this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return {"model": model, "cores_available": cores}


def cpu_threads():
    """Host threads for the mesh-parallel CPU baselines: the cores this
    process may use, at most 16 (one GPU's share of a box's CPUs)."""
    return max(1, min(16, host_cpu()["cores_available"]))


def config1_steps(cf, steps=10, warmup=2):
    """Config 1 (the 10x10x40-cube beam in the reference's baked element
    order, CG rel 1e-4 max 50, undamped): per-step wall time, median over
    `steps` after `warmup`, of the oracle's FixedStep on one core (the
    reference's own per-step statistic, LeafDSH_Solid3D_FEM.h:599-621) and of
    the GPU leaf, fast path and reference order."""
    from oracle import orc
    sh = cf.mesh_box_baked((0.05, 0.05, 0.2), (11, 11, 41))
    pos, tets = sh.pos.copy(), sh.tets.copy()
    knc = np.nonzero(np.isclose(pos[2::3], -0.2))[0].astype(np.uint32)

    def timed(step):
        for _ in range(warmup):
            step()
        ts = []
        for _ in range(steps):
            t = time.perf_counter()
            step()
            ts.append(time.perf_counter() - t)
        return round(float(np.median(ts)) * 1e3, 3)

    o = orc.Solid(pos, tets, orc.default_params(rm=orc.RM_DAPD, cg_max_iter=50), knc)
    out = {"workload": "config 1: 10x10x40-cube beam, 20,000 tets, baked order, CG rel 1e-4 max 50",
           "cpu_oracle_ms_median": timed(o.fixed_step), "cpu_cores": 1}
    for name, ro in (("gpu_fast_ms_median", 0), ("gpu_reference_order_ms_median", 1)):
        p = cf.default_solid_params(fixed_dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), young_modulus=1e5,
                                    poisson_ratio=0.3, damping_ratio=0.0, rm=cf.RM_DAPD, cg_rel_epsilon=1e-4,
                                    cg_abs_epsilon=1e-6, cg_max_iter=50, density=1000.0, reference_order=ro)
        g = cf.LeafSolid3D(pos, tets, p, knc=knc)
        out[name] = timed(g.fixed_step)
        g.close()
    return out


def cpu_batch_baseline(meshes, make_solid, unit, what, steps=1):
    """The oracle stepping a sample of a batch's meshes, the meshes spread over
    the host threads (each stepped whole by one thread)."""
    from oracle import orc
    solids = [make_solid(m) for m in meshes]
    threads = cpu_threads()
    t = time.perf_counter()
    for _ in range(steps):
        orc.step_parallel(solids, threads)
    dt = time.perf_counter() - t
    return {"value": round(len(meshes) * steps / dt, 2), "unit": unit, "cores": threads, "kind": "port",
            "sample": f"{what}: {len(meshes)} meshes x {steps} step(s) in {dt:.1f} s on {threads} threads"}


def cpu_baseline(budget_s=12.0):
    """The oracle's K*x (reference algorithm: stored 3x3 Ke blocks, serial
    scatter-add, FP32, one core) on a 64x64x64-cube shard of the mesh."""
    from oracle import orc
    pos, tets = orc.mesh_box((0.32, 0.32, 0.32), (65, 65, 65))
    T = tets.size // 4
    ts = orc.TetSet(pos, tets, 1e5, 0.3)
    x = large_mesh_state(pos)
    ts.begin_evaluation(x, rm=orc.RM_DAPD)
    d = np.random.default_rng(1).standard_normal(pos.size).astype(np.float32)
    df = np.zeros_like(d)
    ts.df_x(d, df)  # warm
    n = 0
    t0 = time.perf_counter()
    while True:
        df[:] = 0
        ts.df_x(d, df)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 200:
            break
    dt = time.perf_counter() - t0
    return {"value": T * n / dt / 1e9, "unit": "Gtet-matvec/s", "cores": 1, "kind": "port",
            "sample": f"{n} K*x of the oracle on a 64x64x64-cube shard ({T} tets) of the same mesh, "
                      f"{dt:.1f} s"}


def residual_stats(stats):
    """Achieved relative CG residuals of a step's solves (sqrt(delta/delta0),
    ILSS_CG.h:148-150) and why they stopped."""
    rel = np.array([s.cg_rel_residual for s in stats], np.float64)
    reasons = np.array([s.cg_reason for s in stats])
    return {"rel_p50": float(np.percentile(rel, 50)), "rel_p99": float(np.percentile(rel, 99)),
            "rel_max": float(rel.max()), "solves": int(rel.size), "at_or_below_1e-4": int((rel <= 1e-4).sum()),
            "cg_reasons": {str(k): int((reasons == k).sum()) for k in (1, 2, 3, 4)}}


def with_abs(params, abs_eps):
    p = type(params).from_buffer_copy(params)
    p.cg_abs_epsilon = abs_eps
    return p


# The two CG settings of the batched workloads: the reference's defaults (rel
# 1e-4 and the absolute floor 1e-6, below which a solve also stops, as does a
# |d.q| < abs^2 breakdown), and the relative test alone (abs 0), where every
# solve stops at rel 1e-4 (or the iteration cap).
CG_VARIANTS = [("reference_abs_1e-6", 1e-6), ("rel_1e-4_only", 0.0)]


def bench_batch3d(cf, args):
    """Uneven 3D batch: 1,024 cube grids (cubes per axis uniform in [4,40],
    seed 4), E log-uniform [1e4,1e7], nu [0.2,0.45], a random face held. The
    batched step (all meshes in the same launches, graph-recorded) against the
    plain single-mesh path run mesh by mesh on every 8th mesh, under both CG
    settings."""
    from risdsdgcfem_thesis_amd import workloads, autotune
    t0 = time.perf_counter()
    meshes = workloads.uneven_batch(count=args.batch_count, cg_max_iter=args.batch_max_iter)
    T = sum(m["tets"].size // 4 for m in meshes)
    N = sum(m["pos"].size // 3 for m in meshes)
    if args.no_autotune:
        tuned = {"tile_threads": args.tile_threads, "tile_elems": args.tile_elems, "cg_chunk": args.cg_chunk}
    else:
        tuned = autotune.tune(meshes, use_graphs=not args.no_graphs)
        log(f"[batch3d] autotuned in {tuned['seconds']} s: tiles {tuned['tile_threads']}/{tuned['tile_elems']} "
            f"(ns/tet {tuned['kx_ns_per_tet']}), chunk {tuned['cg_chunk']} (step ms {tuned['step_ms']})")
    b = cf.Batch(tile_threads=tuned["tile_threads"], tile_elems=tuned["tile_elems"], cg_chunk=tuned["cg_chunk"],
                 use_graphs=not args.no_graphs, step_mode=args.step_mode)
    mids = b.add_meshes([(m["pos"], m["tets"]) for m in meshes], threads=16)
    log(f"[batch3d] {len(meshes)} meshes, {T} tets, {N} nodes; setup {time.perf_counter() - t0:.1f} s")
    b_alg = 92 * T + 24 * N
    out = {
        "workload": "uneven 3D batch: 1,024 cube grids, 4-40 cubes per axis (seed 4), QIE-v, CG rel 1e-4",
        "cg_reason_codes": "1 rel, 2 abs, 3 max_iter, 4 breakdown (|d.q| < abs^2, ILSS_CG.h:97)",
        "meshes": len(meshes), "tets": T, "nodes": N, "steps": args.batch_steps, "warmup": args.warmup,
        "cg_max_iter": args.batch_max_iter,
        "tile": {"threads": tuned["tile_threads"], "elems": tuned["tile_elems"], "cg_chunk": tuned["cg_chunk"],
                 "graphs": not args.no_graphs,
                 "step_mode": {0: "host", 1: "device"}[args.step_mode]},
        "autotune": None if args.no_autotune else {k: tuned[k] for k in ("kx_ns_per_tet", "step_ms", "sample",
                                                                          "seconds")},
        "variants": {},
    }
    for name, abs_eps in CG_VARIANTS:
        sids = [b.add_sim(mid, with_abs(m["params"], abs_eps), knc=m["knc"]) for mid, m in zip(mids, meshes)]
        for _ in range(args.warmup):
            b.fixed_step(stats=False)
        b.synchronize()
        iters = []
        t1 = time.perf_counter()
        for _ in range(args.batch_steps):
            st = b.fixed_step()
            iters.append([st[s].cg_iters for s in sids])
        b.synchronize()
        wall = time.perf_counter() - t1
        it = np.array(iters)
        res = residual_stats([st[s] for s in sids])
        batched_sps = len(meshes) * args.batch_steps / wall
        if name == CG_VARIANTS[0][0]:
            kx_ms = b.time_kx(args.kx_reps)
            out["kx"] = {"ms": round(kx_ms, 4), "gtets": round(T / (kx_ms * 1e-3) / 1e9, 2),
                         "achieved_gbs": round(b_alg / (kx_ms * 1e-3) / 1e9, 1),
                         "frac": round(b_alg / (kx_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        for s in sids:
            b.remove_sim(s)
        # plain path, one mesh at a time
        sample = meshes[::8]
        plain_wall = 0.0
        for m in sample:
            s = cf.LeafSolid3D(m["pos"], m["tets"], with_abs(m["params"], abs_eps), knc=m["knc"])
            s.set_graphs(not args.no_graphs)
            for _ in range(args.warmup):
                s.fixed_step()
            cf.synchronize()
            t2 = time.perf_counter()
            for _ in range(args.batch_steps):
                s.fixed_step()
            cf.synchronize()
            plain_wall += time.perf_counter() - t2
            s.close()
        plain_sps = len(sample) * args.batch_steps / plain_wall
        log(f"[batch3d] {name}: batched {wall / args.batch_steps * 1e3:.1f} ms/step, iters mean {it.mean():.1f} "
            f"max {it.max()}, rel p99 {res['rel_p99']:.2e}; plain "
            f"{plain_wall / (len(sample) * args.batch_steps) * 1e3:.2f} ms per mesh-step")
        out["variants"][name] = {
            "cg_abs_epsilon": abs_eps, "cg_iters_mean": round(float(it.mean()), 2), "cg_iters_max": int(it.max()),
            "residual": res,
            "batched_steps_per_s": round(batched_sps, 1), "batched_ms_per_step": round(wall / args.batch_steps * 1e3, 2),
            "plain_steps_per_s": round(plain_sps, 1), "plain_sample": f"{len(sample)} of {len(meshes)} meshes (every 8th)",
            "speedup": round(batched_sps / plain_sps, 2)}
    b.close()
    head = out["variants"][CG_VARIANTS[1][0]]
    out.update({k: head[k] for k in ("batched_steps_per_s", "plain_steps_per_s", "speedup")})
    out["headline_variant"] = CG_VARIANTS[1][0]
    if not args.no_cpu_baseline:
        from oracle import orc
        abs_eps = CG_VARIANTS[1][1]
        out["cpu_baseline"] = cpu_batch_baseline(
            meshes[::8], lambda m: orc.Solid(m["pos"], m["tets"], orc.default_params(
                young_modulus=m["E"], poisson_ratio=m["nu"], rm=orc.RM_DAPD, cg_max_iter=args.batch_max_iter,
                cg_abs_epsilon=abs_eps), m["knc"]),
            "steps/s", "the oracle (stored Ke, FP32) on every 8th mesh, CG rel 1e-4 only")
    return out


def bench_batch2d(cf, args):
    """2D batch: 65,536 n x n grids (n uniform in [4,32], seed 3), E
    log-uniform [1e4,1e7], nu [0.2,0.45], bottom edge held. The batched step is
    ONE kernel launch for all meshes; the plain single-mesh path (the same
    kernel, one mesh per launch) runs on every 64th mesh; under both CG
    settings."""
    from risdsdgcfem_thesis_amd import workloads
    meshes = workloads.batch2d(count=args.batch2d_count, cg_max_iter=args.batch_max_iter)
    T = np.array([m["tris"].size // 3 for m in meshes])
    NN = np.array([m["pos"].size // 2 for m in meshes])
    out = {"workload": "2D batch: 65,536 n x n grids, n in [4,32] (seed 3), QIE-v, CG rel 1e-4, one launch per step",
           "meshes": len(meshes), "steps": args.batch_steps, "warmup": args.warmup,
           "cg_max_iter": args.batch_max_iter, "variants": {}}
    for name, abs_eps in CG_VARIANTS:
        t0 = time.perf_counter()
        b = cf.Batch2()
        for m in meshes:
            b.add_sim(m["pos"], m["tris"], with_abs(m["params"], abs_eps), knc=m["knc"])
        info = b.info()
        b.fixed_step(stats=False)  # uploads; first step
        log(f"[batch2d] {name}: {info['sims']} meshes, {info['tris']} tris, {info['nodes']} nodes, "
            f"{info['workgroups']} workgroups; setup {time.perf_counter() - t0:.1f} s")
        out.update({"tris": int(info["tris"]), "nodes": int(info["nodes"]), "workgroups": int(info["workgroups"])})
        for _ in range(args.warmup):
            b.fixed_step(stats=False)
        cf.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.batch_steps):
            b.fixed_step(stats=False)
        wall = time.perf_counter() - t1
        st = b.fixed_step()  # one more step, untimed, for the CG statistics
        it = np.array([x.cg_iters for x in st])
        res = residual_stats(st)
        dev_ms = st[0].step_ms
        tri_matvecs = float(((it + 1) * T).sum())
        # the 2D K*x convention 48 Ne + 16 Nn bytes per matvec (the data sit in
        # LDS inside the resident kernel: an HBM-equivalent rate, not HBM traffic)
        b_alg2 = float(((it + 1) * (48 * T + 16 * NN)).sum())
        batched_sps = len(meshes) * args.batch_steps / wall
        b.close()
        sample = meshes[::64]
        plain_wall = 0.0
        for m in sample:
            s = cf.LeafSolid2D(m["pos"], m["tris"], with_abs(m["params"], abs_eps), knc=m["knc"])
            s.fixed_step()
            for _ in range(args.warmup):
                s.fixed_step()
            t2 = time.perf_counter()
            for _ in range(args.batch_steps):
                s.fixed_step()
            plain_wall += time.perf_counter() - t2
            s.close()
        plain_sps = len(sample) * args.batch_steps / plain_wall
        log(f"[batch2d] {name}: batched {wall / args.batch_steps * 1e3:.2f} ms/step (device {dev_ms:.2f} ms), "
            f"iters mean {it.mean():.1f}, rel p99 {res['rel_p99']:.2e}; plain "
            f"{plain_wall / (len(sample) * args.batch_steps) * 1e3:.3f} ms per mesh-step")
        out["variants"][name] = {
            "cg_abs_epsilon": abs_eps, "cg_iters_mean": round(float(it.mean()), 2), "cg_iters_max": int(it.max()),
            "residual": res,
            "batched_steps_per_s": round(batched_sps, 1),
            "batched_ms_per_step": round(wall / args.batch_steps * 1e3, 3),
            "batched_device_ms_per_step": round(float(dev_ms), 3),
            "gtri_matvecs_per_s": round(tri_matvecs / (dev_ms * 1e-3) / 1e9, 2),
            "kx_hbm_equiv_frac": round(b_alg2 / (dev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "plain_steps_per_s": round(plain_sps, 1),
            "plain_sample": f"{len(sample)} of {len(meshes)} meshes (every 64th)",
            "speedup": round(batched_sps / plain_sps, 1)}
    head = out["variants"][CG_VARIANTS[1][0]]
    out.update({k: head[k] for k in ("batched_steps_per_s", "plain_steps_per_s", "speedup",
                                     "batched_device_ms_per_step", "gtri_matvecs_per_s", "kx_hbm_equiv_frac")})
    out["headline_variant"] = CG_VARIANTS[1][0]
    if not args.no_cpu_baseline:
        from oracle import orc
        abs_eps = CG_VARIANTS[1][1]
        out["cpu_baseline"] = cpu_batch_baseline(
            meshes[::64], lambda m: orc.Solid2(m["pos"], m["tris"], orc.default_params(
                young_modulus=m["E"], poisson_ratio=m["nu"], rm=orc.RM_DAPD, cg_max_iter=args.batch_max_iter,
                cg_abs_epsilon=abs_eps, gravity=(0.0, -9.81, 0.0)), knc=m["knc"]),
            "steps/s", "the oracle (stored Ke, FP32) on every 64th mesh, CG rel 1e-4 only", steps=3)
    return out


def bench_serve(cf, args):
    """Serving: Poisson arrivals at each rate for a fixed real-time window;
    requests are seeded cube grids of 500-200,000 tets (log-uniform) with 1-60
    steps and config-3 materials, admitted by a 200 GB HBM budget and a cap on
    the tets in flight that adapts to a target round time."""
    from risdsdgcfem_thesis_amd import workloads
    out = []
    for rate in args.serve_rates:
        meshes, reqs = workloads.serving_trace(rate=rate, duration_s=args.serve_seconds, seed=5)
        base = workloads.solid_params(1e5, 0.3, args.batch_max_iter)
        res, lat = cf.serve(meshes, reqs, base, hbm_budget_bytes=int(args.serve_budget_gb * 1e9),
                            duration_s=args.serve_seconds, drain=False, max_tick_tets=args.serve_tick_tets,
                            cg_chunk=args.cg_chunk, use_graphs=not args.no_graphs,
                            target_round_ms=args.serve_target_round_ms)
        log(f"[serve] rate {rate}/s: {res['requests']} requests, {res['admitted']} admitted, {res['completed']} done, "
            f"{res['steps']} steps in {res['wall_s']:.2f} s, p50 {res['step_p50_ms']:.2f} ms, "
            f"p99 {res['step_p99_ms']:.2f} ms, max active {res['max_active']}, round {res['tick_mean_ms']:.2f} ms, "
            f"cap {res['final_tick_tets'] / 1e6:.1f}M tets")
        out.append({"rate_per_s": rate, "requests": res["requests"], "admitted": res["admitted"],
                    "completed": res["completed"], "steps": res["steps"],
                    "steps_per_s": round(res["steps_per_s"], 1),
                    "gtet_steps_per_s": round(res["tet_steps"] / res["wall_s"] / 1e9, 4) if res["wall_s"] else 0,
                    "step_p50_ms": round(res["step_p50_ms"], 3), "step_p99_ms": round(res["step_p99_ms"], 3),
                    "first_step_p50_ms": round(res["first_p50_ms"], 3),
                    "first_step_p99_ms": round(res["first_p99_ms"], 3),
                    "tick_mean_ms": round(res["tick_mean_ms"], 3), "max_active": res["max_active"],
                    "final_tick_tets": res["final_tick_tets"],
                    "peak_gb": round(res["peak_bytes"] / 1e9, 3), "window_s": args.serve_seconds})
    return {"workload": "serving: Poisson arrivals, cube grids 500-200,000 tets, 1-60 steps, config-3 materials",
            "hbm_budget_gb": args.serve_budget_gb, "max_tick_tets": args.serve_tick_tets,
            "target_round_ms": args.serve_target_round_ms,
            "cg_max_iter": args.batch_max_iter, "rates": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--kx-reps", type=int, default=50)
    ap.add_argument("--tile-elems", type=int, default=512,
                    help="elements per tile of the 5.24M-tet mesh's element kernel (swept: 512 x 512 best)")
    ap.add_argument("--tile-threads", type=int, default=512)
    ap.add_argument("--tile-schedule", type=int, default=0)
    ap.add_argument("--cg-workgroups", type=int, default=0,
                    help="workgroups of the leaf CG iteration kernels (0 = library default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="all", choices=["all", "single", "batch3d", "batch2d", "serve", "fie"])
    ap.add_argument("--fie-steps", type=int, default=5)
    ap.add_argument("--batch-count", type=int, default=1024)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--batch-max-iter", type=int, default=1000,
                    help="CG iteration cap of the batched and served workloads (high enough that every solve "
                         "reaches the 1e-4 residual)")
    ap.add_argument("--cg-chunk", type=int, default=8)
    ap.add_argument("--step-mode", type=int, default=1, choices=[0, 1],
                    help="batched 3D: 0 host-polled chunks, 1 the step submitted at once with device-compacted "
                         "lists (default)")
    ap.add_argument("--tile-alternate", type=int, default=0, choices=[0, 1],
                    help="single mesh: 0 every CG iteration walks the tiles forwards; 1 odd iterations walk them "
                         "backwards, starting on the tiles the previous K*x touched last")
    ap.add_argument("--no-autotune", action="store_true", help="batched 3D: use --tile-threads/--tile-elems/--cg-chunk")
    ap.add_argument("--no-graphs", action="store_true",
                    help="single mesh, batched 3D and serving: launch the step's kernels directly instead of "
                         "recorded HIP graphs (for kernel-trace profiling)")
    ap.add_argument("--batch2d-count", type=int, default=65536)
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[100.0, 300.0, 1000.0, 10000.0],
                    help="serving: Poisson arrival rates (per second); 100 and 300 are sustainable, 1k and 10k "
                         "overload the GPU")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-tick-tets", type=int, default=16_000_000,
                    help="serving: tets cap on admission (the starting cap when --serve-target-round-ms > 0)")
    ap.add_argument("--serve-target-round-ms", type=float, default=10.0,
                    help="serving: adapt the tets cap so that a round takes about this long (0: fixed cap; "
                         "swept 5, 10, 20: 10 gave the most steps at 1k arrivals/s)")
    ap.add_argument("--kx-only", type=int, default=0, metavar="REPS",
                    help="profiling mode: one step, then REPS K*x launches and one 256 MiB device copy "
                         "(the copy calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE); prints timings only")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the workload is one GPU's mesh; --gpus {args.gpus} is not "
              "supported (see the sharding design in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import risdsdgcfem_thesis_amd as cf
    if cf.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)

    if args.workload == "batch3d":
        print(json.dumps(bench_batch3d(cf, args)), flush=True)
        return
    if args.workload == "batch2d":
        print(json.dumps(bench_batch2d(cf, args)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(bench_serve(cf, args)), flush=True)
        return
    if args.workload == "fie":
        print(json.dumps(bench_fie(cf, args)), flush=True)
        return
    if args.workload == "single" or args.kx_only:
        print(json.dumps(bench_single(cf, args)), flush=True)
        return
    # default: every workload, one JSON line
    out = bench_single(cf, args)
    out["batched_2d"] = bench_batch2d(cf, args)
    out["batched_3d"] = bench_batch3d(cf, args)
    out["serving"] = bench_serve(cf, args)
    out["fie_dx"] = bench_fie(cf, args)
    out["headline"] = {
        "kx_gtets": out["roofline"]["kx_gtets"], "kx_frac_hbm_peak": out["roofline"]["frac"],
        "batched_vs_plain_2d": out["batched_2d"]["speedup"], "batched_vs_plain_3d": out["batched_3d"]["speedup"],
        "serve_p99_ms": {str(int(r["rate_per_s"])): r["step_p99_ms"] for r in out["serving"]["rates"]},
    }
    print(json.dumps(out), flush=True)


def reset_tiles(cf):
    cf.set_tile_elems(0)
    cf.set_tile_threads(0)
    cf.set_tile_persistent(0)


def bench_single(cf, args):
    """The single large mesh (64x64x256 cube grid), plain GPU drop-in."""
    cf.set_tile_elems(args.tile_elems)
    cf.set_tile_threads(args.tile_threads)
    cf.set_tile_persistent(args.tile_schedule)
    cf.set_cg_workgroups(args.cg_workgroups)
    cf.set_tile_alternate(args.tile_alternate)
    counts = (65, 65, 257)
    half = (0.32, 0.32, 1.28)
    pos, tets = cf.mesh_box(half, counts)
    N, T = pos.size // 3, tets.size // 4
    x0 = large_mesh_state(pos)
    knc = np.nonzero(np.isclose(pos[2::3], -half[2]))[0].astype(np.uint32)
    params = cf.default_solid_params(fixed_dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), young_modulus=1e5,
                                     poisson_ratio=0.3, damping_ratio=0.0, rm=cf.RM_DAPD, cg_rel_epsilon=1e-4,
                                     cg_abs_epsilon=1e-6, cg_max_iter=50, density=1000.0)
    log(f"[bench] mesh {N} nodes {T} tets; building solid")
    t_build = time.perf_counter()
    solid = cf.LeafSolid3D(pos, tets, params)
    reset_tiles(cf)  # the solid keeps its tiles; later workloads build theirs with the library default
    solid.set_state(x0, np.zeros_like(x0))
    solid.set_knc(knc, x0.reshape(-1, 3)[knc].ravel())
    solid.set_graphs(not args.no_graphs)
    log(f"[bench] setup {time.perf_counter() - t_build:.2f} s; warmup {args.warmup} steps")
    if args.kx_only:
        solid.fixed_step()
        ms_tile, ms_gather = solid.time_kx(args.kx_reps if args.kx_reps else args.kx_only)
        nbytes = 256 << 20
        a, b = cf.DeviceArray(nbytes // 4), cf.DeviceArray(nbytes // 4)
        cf.lib.cf_memcpy_dtod(b.ptr, a.ptr, nbytes)
        cf.synchronize()
        return {"mode": "kx-only", "tile_ms": ms_tile, "gather_ms": ms_gather, "copy_bytes": nbytes,
                "tets": T, "nodes": N}
    for _ in range(args.warmup):
        solid.fixed_step()
    cf.synchronize()
    log(f"[bench] timing {args.steps} steps")
    iters = []
    step_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = solid.fixed_step()
        iters.append(st.cg_iters)
        step_ms.append(st.step_ms)
    cf.synchronize()
    wall = time.perf_counter() - t0
    matvecs = sum(i + 1 for i in iters)
    value = T * matvecs / wall / 1e9
    log(f"[bench] {wall * 1e3 / args.steps:.2f} ms/step, cg iters {iters}")

    ms_tile, ms_gather = solid.time_kx(args.kx_reps)
    ms_kx = ms_tile + ms_gather
    b_alg = 92 * T + 24 * N
    achieved = b_alg / (ms_kx * 1e-3) / 1e9
    traffic = load_traffic()
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic.get("bytes_per_kx") if traffic else None,
        "kernels": {"k_tile_forces_ms": round(ms_tile, 5), "k_gather_ms": round(ms_gather, 5)},
        "b_alg_bytes": b_alg, "kx_gtets": round(T / (ms_kx * 1e-3) / 1e9, 2),
        "tile": {"elems": args.tile_elems, "threads": args.tile_threads, "schedule": args.tile_schedule,
                 "alternate": args.tile_alternate},
        "cg_workgroups": cf.get_cg_workgroups(),
    }
    pcie = pcie_inclusive(cf, solid, pos, tets, x0, args)
    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle K*x, one core)")
        cpu = cpu_baseline()
        cpu["host"] = host_cpu()
        log("[bench] config 1: per-step medians, oracle and GPU")
        cpu["config1_step"] = config1_steps(cf)
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gtet-matvec/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded cube-grid mesh, seeded rotation+stretch initial pose)",
        "config": {"workload": "single 64x64x256 cube-grid mesh, plain GPU drop-in (QIE-v, CG rel 1e-4 max 50)",
                   "tets": T, "nodes": N, "rotations": "polar (DAPD)", "parallelism": "single GPU"},
        "steps_per_s": round(args.steps / wall, 3), "cg_iters_mean": round(float(np.mean(iters)), 2),
        "roofline": roofline, "cpu_baseline": cpu, "pcie_inclusive": pcie,
    }
    solid.close()
    return out


def pcie_inclusive(cf, solid, pos, tets, x0, args, reps=5):
    """Rates when the caller's arrays live in host memory (not `value`): a K*x
    through the element-set boundary with the input vector uploaded and the
    accumulated output copied in and back per call (what the IForce3 adapter
    does with the leaf's ParticleSystem arrays), and a leaf step followed by a
    read-back of positions and velocities. Pageable numpy arrays."""
    N, T = pos.size // 3, tets.size // 4
    cf.set_tile_elems(args.tile_elems)
    cf.set_tile_threads(args.tile_threads)
    ts = cf.TetrahedronElementSet3(pos, tets, cf.default_fem_params(young_modulus=1e5, poisson_ratio=0.3,
                                                                   rm=cf.RM_DAPD))
    reset_tiles(cf)
    dx = cf.DeviceArray.from_host(x0)
    ts.begin_evaluation(dx)
    v = np.random.default_rng(1).standard_normal(3 * N).astype(np.float32)
    out = np.zeros(3 * N, np.float32)
    dv, dout = cf.DeviceArray(3 * N), cf.DeviceArray(3 * N)
    ts.df_x(dv, dout)
    cf.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        dv.upload(v)
        dout.upload(out)
        ts.df_x(dv, dout)
        out = dout.numpy()
    kx_ms = (time.perf_counter() - t) / reps * 1e3
    ts.end_evaluation()
    ts.close()
    cf.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        solid.fixed_step()
        solid.state()
    step_ms = (time.perf_counter() - t) / reps * 1e3
    log(f"[bench] host arrays: K*x {kx_ms:.2f} ms per call, step + state read-back {step_ms:.2f} ms")
    return {"kx_ms": round(kx_ms, 3), "kx_gtets": round(T / (kx_ms * 1e-3) / 1e9, 2),
            "step_with_state_readback_ms": round(step_ms, 3),
            "what": "K*x with the input uploaded and the output copied in and back per call (3 x 13 MB over "
                    "PCIe, pageable host arrays); a QIE-v step followed by a read-back of x and v (26 MB)"}


def bench_fie(cf, args):
    """FIE-dx on the single large mesh (reference defaults: 3 Newton iterations,
    CG rel 1e-4 max 50), and the multi-solid universe step over 16 config-1
    beams (20,000 tets each)."""
    counts = (65, 65, 257)
    half = (0.32, 0.32, 1.28)
    pos, tets = cf.mesh_box(half, counts)
    N, T = pos.size // 3, tets.size // 4
    x0 = large_mesh_state(pos)
    knc = np.nonzero(np.isclose(pos[2::3], -half[2]))[0].astype(np.uint32)
    params = cf.default_solid_params(fixed_dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), young_modulus=1e5,
                                     poisson_ratio=0.3, damping_ratio=0.0, rm=cf.RM_DAPD, cg_rel_epsilon=1e-4,
                                     cg_abs_epsilon=1e-6, cg_max_iter=50, density=1000.0,
                                     integrator=cf.IT_FIE_DX, nr_max_iter=3)
    log(f"[bench] FIE-dx on {T} tets")
    cf.set_tile_elems(args.tile_elems)
    cf.set_tile_threads(args.tile_threads)
    solid = cf.LeafSolid3D(pos, tets, params)
    reset_tiles(cf)
    solid.set_graphs(not args.no_graphs)
    solid.set_state(x0, np.zeros_like(x0))
    solid.set_knc(knc, x0.reshape(-1, 3)[knc].ravel())
    solid.fixed_step()
    cf.synchronize()
    nr, ls = [], []
    t0 = time.perf_counter()
    for _ in range(args.fie_steps):
        st = solid.fixed_step()
        nr.append(st.nr_iters)
        ls.append(st.cg_iters)
    cf.synchronize()
    wall = time.perf_counter() - t0
    solid.close()
    matvecs = sum(ls) + sum(nr)  # CG iterations + one residual matvec per solve
    single = {"workload": "FIE-dx, 64x64x256 cube grid, 3 Newton iterations, CG rel 1e-4 max 50",
              "steps": args.fie_steps, "ms_per_step": round(wall * 1e3 / args.fie_steps, 3),
              "steps_per_s": round(args.fie_steps / wall, 3), "nr_iters": nr, "cg_iters": ls,
              "gtet_matvecs_per_s": round(T * matvecs / wall / 1e9, 3)}
    bpos, btets = cf.mesh_box((0.05, 0.05, 0.2), (11, 11, 41))
    bknc = np.nonzero(np.isclose(bpos[2::3], -0.2))[0].astype(np.uint32)
    bp = cf.default_solid_params(fixed_dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), young_modulus=1e5,
                                 poisson_ratio=0.3, damping_ratio=0.0, rm=cf.RM_DAPD, cg_rel_epsilon=1e-4,
                                 cg_abs_epsilon=1e-6, cg_max_iter=50, density=1000.0, nr_max_iter=3)
    solids = [cf.LeafSolid3D(bpos, btets, bp, knc=bknc) for _ in range(16)]
    multi = cf.MultiSolid3D(solids, use_graphs=not args.no_graphs)
    multi.update(1.0 / 60.0)
    multi.update(1.0 / 60.0)
    cf.synchronize()
    frames, steps = 5, 0
    t0 = time.perf_counter()
    for _ in range(frames):
        n, _ = multi.update(1.0 / 60.0)
        steps += n
    cf.synchronize()
    mwall = time.perf_counter() - t0
    multi.close()
    for sd in solids:
        sd.close()
    log(f"[bench] FIE-dx {single['ms_per_step']} ms/step; multi-solid {mwall * 1e3 / max(steps, 1):.2f} ms/fixed step")
    return {"single": single,
            "multi": {"workload": "multi-solid universe step, 16 config-1 beams (20,000 tets each), 3 Newton passes",
                      "solids": 16, "fixed_steps": steps, "ms_per_fixed_step": round(mwall * 1e3 / max(steps, 1), 3),
                      "solid_steps_per_s": round(16 * steps / mwall, 2)}}


if __name__ == "__main__":
    main()
