"""GPU: batched stepping (cf_batch_*) against the single-mesh path.

A simulation's step inside a batch runs the single-mesh block bodies on its
own arrays with its own per-mesh reduction trees, so its positions,
velocities and CG statistics must be bit-identical to LeafSolid3D.fixed_step
on the same mesh, parameters and state: alone, in any mix, with or without
HIP graphs, for any CG chunk length, and across admissions and removals.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EDGE = 0.01


def box_case(cf, counts, E, nu, face, rm=None, max_iter=50):
    """Cube-grid mesh with `counts` vertices per axis, edge 0.01 m, one face
    (axis, side) held at rest."""
    half = tuple(0.5 * EDGE * (n - 1) for n in counts)
    pos, tets = cf.mesh_box(half, counts)
    axis, side = face
    c = pos[axis::3]
    knc = np.nonzero(np.isclose(c, side * half[axis]))[0].astype(np.uint32)
    params = cf.default_solid_params(fixed_dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), young_modulus=E,
                                     poisson_ratio=nu, damping_ratio=0.0, rm=cf.RM_DAPD if rm is None else rm,
                                     cg_rel_epsilon=1e-4, cg_abs_epsilon=1e-6, cg_max_iter=max_iter, density=1000.0)
    return {"pos": pos, "tets": tets, "params": params, "knc": knc, "N": pos.size // 3}


def cases(cf):
    return [
        box_case(cf, (5, 5, 9), 1e5, 0.30, (2, -1)),
        box_case(cf, (7, 4, 6), 3e6, 0.45, (0, 1)),
        box_case(cf, (9, 9, 9), 2e4, 0.20, (1, -1), rm=cf.RM_MSVD),
        box_case(cf, (4, 12, 5), 1e7, 0.35, (2, 1)),
        box_case(cf, (6, 6, 6), 5e4, 0.25, (1, -1), max_iter=7),
    ]


def run_plain(cf, case, steps):
    s = cf.LeafSolid3D(case["pos"], case["tets"], case["params"], knc=case["knc"])
    st = [s.fixed_step() for _ in range(steps)]
    x, v = s.state()
    s.close()
    return x, v, [(t.cg_iters, t.cg_reason, t.cg_rel_residual, t.elastic_energy, t.min_det_f) for t in st]


def run_batch(cf, cs, steps, **kw):
    b = cf.Batch(**kw)
    mids = b.add_meshes([(c["pos"], c["tets"]) for c in cs], threads=4)
    sids = [b.add_sim(m, c["params"], knc=c["knc"]) for m, c in zip(mids, cs)]
    hist = {s: [] for s in sids}
    for _ in range(steps):
        st = b.fixed_step()
        for s in sids:
            t = st[s]
            hist[s].append((t.cg_iters, t.cg_reason, t.cg_rel_residual, t.elastic_energy, t.min_det_f))
    out = []
    for s, c in zip(sids, cs):
        x, v = b.state(s, c["N"])
        out.append((x, v, hist[s]))
    b.close()
    return out


def same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


@pytest.fixture(scope="module")
def plain_results(gpu):
    cs = cases(gpu)
    return cs, [run_plain(gpu, c, 3) for c in cs]


def test_batch_equals_plain_bitexact(gpu, plain_results):
    cs, plain = plain_results
    got = run_batch(gpu, cs, 3)
    for i, (g, p) in enumerate(zip(got, plain)):
        assert same(g, p), f"case {i}: batched step differs from the single-mesh step"
    # the capped case stops at its own max_iter while the others go on
    assert plain[4][2][0][0] == 7 and plain[4][2][0][1] == gpu.CG_MAX_ITER


def test_alone_equals_in_mix(gpu, plain_results):
    cs, plain = plain_results
    for i in (0, 3):
        alone = run_batch(gpu, [cs[i]], 3)[0]
        assert same(alone, plain[i])
    mix = run_batch(gpu, [cs[3], cs[1], cs[0]], 3)
    assert same(mix[0], plain[3]) and same(mix[1], plain[1]) and same(mix[2], plain[0])


@pytest.mark.parametrize("kw", [dict(use_graphs=False), dict(cg_chunk=1), dict(cg_chunk=64),
                                dict(cg_chunk=3, use_graphs=False)])
def test_schedules_equal(gpu, plain_results, kw):
    cs, plain = plain_results
    got = run_batch(gpu, cs, 3, **kw)
    for i, (g, p) in enumerate(zip(got, plain)):
        assert same(g, p), f"case {i} with {kw}"


@pytest.mark.parametrize("mode", ["BATCH_STEP_HOST", "BATCH_STEP_DEVICE"])
@pytest.mark.parametrize("kw", [dict(), dict(use_graphs=False), dict(cg_chunk=1), dict(cg_chunk=5)])
def test_step_modes_equal(gpu, plain_results, mode, kw):
    """Host-polled chunks, and the step submitted at once with the lists
    compacted on the device before every chunk: the same results."""
    cs, plain = plain_results
    got = run_batch(gpu, cs, 3, step_mode=getattr(gpu, mode), **kw)
    for i, (g, p) in enumerate(zip(got, plain)):
        assert same(g, p), f"case {i} under {mode} with {kw}"


def test_device_schedule_outrun_and_overshoot(gpu):
    """The device schedule submits a step with the previous step's chunk
    counts. Steps from rest and from a strongly disturbed state need
    different CG iteration counts (here the step from rest needs more): the
    step that needs more chunks than the one before it gets chunks added after
    the submitted ones (one host read each), the one that needs fewer runs
    trailing chunks that find no simulation iterating. Every step stays
    bit-identical to the single-mesh path given the same states."""
    cs = [box_case(gpu, (5, 5, 9), 1e5, 0.30, (2, -1), max_iter=200),
          box_case(gpu, (7, 4, 6), 3e6, 0.45, (0, 1), max_iter=200),
          box_case(gpu, (6, 6, 6), 5e4, 0.25, (1, -1), max_iter=200)]
    rng = np.random.default_rng(11)
    kicks = [(c["pos"] + rng.uniform(-2e-3, 2e-3, c["pos"].shape).astype(np.float32),
              rng.uniform(-0.5, 0.5, c["pos"].shape).astype(np.float32)) for c in cs]
    plains = []
    for c, (xk, vk) in zip(cs, kicks):
        s = gpu.LeafSolid3D(c["pos"], c["tets"], c["params"], knc=c["knc"])
        st = [s.fixed_step()]
        s.set_state(xk, vk)
        st.append(s.fixed_step())
        s.set_state(c["pos"], np.zeros_like(c["pos"]))
        st.append(s.fixed_step())
        x, v = s.state()
        s.close()
        plains.append((x, v, [(t.cg_iters, t.cg_reason, t.cg_rel_residual, t.elastic_energy, t.min_det_f)
                              for t in st]))
    b = gpu.Batch(cg_chunk=2, step_mode=gpu.BATCH_STEP_DEVICE)
    mids = b.add_meshes([(c["pos"], c["tets"]) for c in cs], threads=2)
    sids = [b.add_sim(m, c["params"], knc=c["knc"]) for m, c in zip(mids, cs)]
    hist = {s: [] for s in sids}
    conts, chunks = [], []
    for step in range(3):
        if step == 1:
            for sid, (xk, vk) in zip(sids, kicks):
                b.set_state(sid, xk, vk)
        if step == 2:
            for sid, c in zip(sids, cs):
                b.set_state(sid, c["pos"], np.zeros_like(c["pos"]))
        submitted = b.info()["next_chunks"]
        st = b.fixed_step()
        info = b.info()
        conts.append(info["last_continuations"])
        chunks.append((submitted, info["next_chunks"]))
        for sid in sids:
            t = st[sid]
            hist[sid].append((t.cg_iters, t.cg_reason, t.cg_rel_residual, t.elastic_energy, t.min_det_f))
    print(f"\n(submitted, needed) chunks per step {chunks}, chunks added {conts}")
    for sid, c, p in zip(sids, cs, plains):
        x, v = b.state(sid, c["N"])
        assert same((x, v, hist[sid]), p)
    b.close()
    assert any(c > 0 for c in conts[1:]), "no step needed chunks beyond the submitted ones"
    assert any(sub > need for sub, need in chunks[1:]), "no step needed fewer chunks than submitted"


def test_admission_removal_and_inactive(gpu, plain_results):
    cs, plain = plain_results
    b = gpu.Batch()
    m0 = b.add_mesh(cs[0]["pos"], cs[0]["tets"])
    m1 = b.add_mesh(cs[1]["pos"], cs[1]["tets"])
    s0 = b.add_sim(m0, cs[0]["params"], knc=cs[0]["knc"])
    extra = b.add_sim(m1, cs[1]["params"], knc=cs[1]["knc"])
    b.fixed_step()
    b.remove_sim(extra)  # frees a slot and its memory
    s1 = b.add_sim(m1, cs[1]["params"], knc=cs[1]["knc"])  # reuses them, starts at rest
    s2 = b.add_sim(m0, cs[0]["params"], knc=cs[0]["knc"])  # second simulation on a shared mesh
    b.set_active(s2, False)
    x_rest = b.state(s2, cs[0]["N"])[0]
    b.fixed_step()
    b.fixed_step()
    assert np.array_equal(b.state(s2, cs[0]["N"])[0], x_rest)  # inactive: untouched
    b.set_active(s2, True)
    b.fixed_step()
    # s0 took 4 steps, s1 3 from rest, s2 1
    p0 = run_plain(gpu, cs[0], 4)
    x0, v0 = b.state(s0, cs[0]["N"])
    assert np.array_equal(x0, p0[0]) and np.array_equal(v0, p0[1])
    p1 = run_plain(gpu, cs[1], 3)
    x1, v1 = b.state(s1, cs[1]["N"])
    assert np.array_equal(x1, p1[0]) and np.array_equal(v1, p1[1])
    p2 = run_plain(gpu, cs[0], 1)
    x2, v2 = b.state(s2, cs[0]["N"])
    assert np.array_equal(x2, p2[0]) and np.array_equal(v2, p2[1])
    with pytest.raises(gpu.CorofemError):
        b.remove_mesh(m0)  # still used
    info = b.info()
    assert info["num_sims"] == 3 and info["num_active"] == 3 and info["bytes_in_use"] > 0
    b.close()


def test_hbm_budget_admission(gpu):
    c = box_case(gpu, (6, 6, 6), 1e5, 0.3, (1, -1))
    probe = gpu.Batch()
    m = probe.add_mesh(c["pos"], c["tets"])
    mesh_bytes = probe.info()["bytes_in_use"]
    sim_bytes = probe.sim_bytes(m)
    probe.close()
    b = gpu.Batch(hbm_budget_bytes=mesh_bytes + sim_bytes + sim_bytes // 2)
    m = b.add_mesh(c["pos"], c["tets"])
    s = b.add_sim(m, c["params"], knc=c["knc"])
    with pytest.raises(gpu.CorofemError) as e:
        b.add_sim(m, c["params"], knc=c["knc"])
    assert "status -8" in str(e.value)
    b.remove_sim(s)
    b.add_sim(m, c["params"], knc=c["knc"])  # fits again once memory is freed
    assert b.info()["bytes_in_use"] == mesh_bytes + sim_bytes
    b.close()


def test_batch_kx_timing_runs(gpu):
    cs = cases(gpu)[:3]
    b = gpu.Batch()
    mids = b.add_meshes([(c["pos"], c["tets"]) for c in cs])
    for m, c in zip(mids, cs):
        b.add_sim(m, c["params"], knc=c["knc"])
    b.fixed_step(stats=False)
    ms = b.time_kx(5)
    assert ms > 0
    b.close()


def test_rounds_equal_plain_bitexact(gpu, plain_results):
    """Iteration-level batching: steps of different simulations start and end
    in different rounds; each simulation's state equals its plain run."""
    cs, _ = plain_results
    want_steps = [2, 3, 1, 2, 4]
    b = gpu.Batch()
    mids = b.add_meshes([(c["pos"], c["tets"]) for c in cs])
    sids = [b.add_sim(m, c["params"], knc=c["knc"]) for m, c in zip(mids, cs)]
    for s, n in zip(sids, want_steps):
        b.request_steps(s, n)
    rounds, busy = 0, 1
    seen = {s: 0 for s in sids}
    while busy:
        done, busy = b.round(iters=4)
        for s in sids:
            assert done[s] >= seen[s]
            seen[s] = done[s]
        rounds += 1
        assert rounds < 1000
    assert [seen[s] for s in sids] == want_steps
    # a later request continues from the current state
    b.request_steps(sids[2], 1)
    while b.round(iters=4)[1]:
        pass
    want_steps[2] += 1
    for s, c, n in zip(sids, cs, want_steps):
        p = run_plain(gpu, c, n)
        x, v = b.state(s, c["N"])
        assert np.array_equal(x, p[0]) and np.array_equal(v, p[1]), f"simulation {s} differs after {n} steps"
    b.close()


@pytest.mark.parametrize("max_iter", [50, 500])
def test_multichunk_meshes_equal_plain_bitexact(gpu, max_iter):
    """Meshes with several matvec-gather chunks (1,024 nodes), CG vector
    chunks (4,096 floats) and tiles: the 11x11x41-vertex beam (4,961 nodes,
    20,000 tets) and a 17^3-vertex cube (4,913 nodes), 10 steps, batched
    together against each alone in the single-mesh path. Every CG iteration
    sums the same partials in the same order on both paths, so positions,
    velocities and statistics stay bit-identical over the 10 steps."""
    cs = [box_case(gpu, (11, 11, 41), 1e5, 0.30, (2, -1), max_iter=max_iter),
          box_case(gpu, (17, 17, 17), 3e5, 0.40, (1, -1), max_iter=max_iter)]
    for c in cs:
        assert c["N"] > 4 * 1024  # several chunks of each kind
    plain = [run_plain(gpu, c, 10) for c in cs]
    batched = run_batch(gpu, cs, 10)
    for i, (g, p) in enumerate(zip(batched, plain)):
        print(f"\ncase {i}: CG iterations per step {[h[0] for h in p[2]]}")
        assert same(g, p), f"case {i}: batched steps differ from the single-mesh steps"


@pytest.mark.slow
def test_four_row_cg_chunks_equal_plain_bitexact(gpu):
    """A 70^3-cube mesh (357,911 nodes, 1,715,000 tets): large enough that the
    CG vector chunks take four float4 rows per thread (263 chunks) - the
    layout of the large single-mesh workload - batched against alone, 2 steps."""
    c = box_case(gpu, (71, 71, 71), 2e5, 0.35, (2, -1), max_iter=60)
    assert 3 * c["N"] >= 4 * 4 * 256 * 256
    p = run_plain(gpu, c, 2)
    (g,) = run_batch(gpu, [c], 2)
    print(f"\nCG iterations per step {[h[0] for h in p[2]]}")
    assert same(g, p), "batched steps differ from the single-mesh steps"
