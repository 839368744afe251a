"""GPU: the serving loop (cf_serve_trace). Requests that arrive while others
run are admitted into the running batch; each request's final state must be
bit-identical to the same request stepped alone on the single-mesh path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone(gpu):
    from risdsdgcfem_thesis_amd import workloads
    meshes, reqs = workloads.serving_trace(rate=400.0, duration_s=0.05, seed=11, tets_hi=20_000, steps_hi=12)
    assert len(reqs) >= 8
    base = workloads.solid_params(1e5, 0.3, 50)
    finals = [np.zeros(meshes[q["mesh"]][0].size, np.float32) for q in reqs]
    res, lat = gpu.serve(meshes, reqs, base, hbm_budget_bytes=8 << 30, duration_s=1.0, drain=True,
                         final_positions=finals)
    assert res["completed"] == len(reqs) and res["steps"] == sum(q["steps"] for q in reqs)
    assert res["max_active"] >= 2  # requests overlapped on the device
    assert len(lat) == res["steps"] and (lat > 0).all()
    assert res["step_p50_ms"] <= res["step_p99_ms"] <= res["step_max_ms"]
    for i in range(0, len(reqs), max(1, len(reqs) // 6)):
        q = reqs[i]
        pos, tets = meshes[q["mesh"]]
        p = workloads.solid_params(q["E"], q["nu"], 50)
        s = gpu.LeafSolid3D(pos, tets, p, knc=q["knc"])
        for _ in range(q["steps"]):
            s.fixed_step()
        x, _ = s.state()
        s.close()
        assert np.array_equal(finals[i], x), f"request {i} differs from its run alone"


def test_budget_limits_residency(gpu):
    from risdsdgcfem_thesis_amd import workloads
    meshes, reqs = workloads.serving_trace(rate=2000.0, duration_s=0.02, seed=3, tets_lo=2000, tets_hi=2001,
                                           steps_lo=3, steps_hi=3)
    base = workloads.solid_params(1e5, 0.3, 20)
    b = gpu.Batch()
    m = b.add_mesh(*meshes[0])
    per_sim = b.sim_bytes(m)
    mesh_bytes = b.info()["bytes_in_use"]
    b.close()
    budget = mesh_bytes + 3 * per_sim + per_sim // 2  # room for three simulations at once
    res, _ = gpu.serve(meshes, reqs, base, hbm_budget_bytes=budget, duration_s=1.0, drain=True)
    assert res["completed"] == len(reqs)
    assert res["max_active"] <= 3 and res["peak_bytes"] <= budget


def test_adaptive_cap_moves_and_results_hold(gpu):
    """target_round_ms: with an unreachable target (1 ns) every round is over
    it and the cap shrinks to its floor, 1/64 of the start; with a target far
    above any round, every round whose admission hit the cap grows it. Either
    way every request completes and its final state equals its run alone."""
    from risdsdgcfem_thesis_amd import workloads
    meshes, reqs = workloads.serving_trace(rate=600.0, duration_s=0.04, seed=12, tets_hi=12_000, steps_hi=6)
    assert len(reqs) >= 8
    base = workloads.solid_params(1e5, 0.3, 40)
    start = 20_000
    runs = {}
    for name, target in (("shrink", 1e-6), ("grow", 1e6)):
        finals = [np.zeros(meshes[q["mesh"]][0].size, np.float32) for q in reqs]
        res, _ = gpu.serve(meshes, reqs, base, hbm_budget_bytes=8 << 30, duration_s=1.0, drain=True,
                           max_tick_tets=start, target_round_ms=target, final_positions=finals)
        assert res["completed"] == len(reqs) and res["steps"] == sum(q["steps"] for q in reqs)
        runs[name] = (res, finals)
        print(f"\n{name}: final cap {res['final_tick_tets']}, max active {res['max_active']}, rounds {res['ticks']}")
    assert runs["shrink"][0]["final_tick_tets"] == start // 64
    assert runs["grow"][0]["final_tick_tets"] > start
    for i in (0, len(reqs) // 2, len(reqs) - 1):
        q = reqs[i]
        pos, tets = meshes[q["mesh"]]
        s = gpu.LeafSolid3D(pos, tets, workloads.solid_params(q["E"], q["nu"], 40), knc=q["knc"])
        for _ in range(q["steps"]):
            s.fixed_step()
        x, _ = s.state()
        s.close()
        for name in runs:
            assert np.array_equal(runs[name][1][i], x), f"request {i} ({name}) differs from its run alone"
