"""GPU: edge cases of the leaf, the operator and the batch against the oracle.

- One tetrahedron (the smallest mesh: one tile of one element, one matvec and
  CG chunk) stepped by the leaf and inside a batch next to a 4,961-node beam.
- Nodes no element references (TetSolidShape3 keeps unused vertices; the
  leaf's uniform node mass counts them, LeafDSH_Solid3D_FEM.cpp:1591-1611):
  they take no partial and fall freely under gravity.
- Every node held (LS_ZeroKNC, LeafDSH_Solid3D_FEM.cpp:1045-1053): the CG
  start residual is zero, the solve stops on the absolute test before its
  first iteration and the solid stays at its targets.
- A node shared by 144 elements: the tile builder cuts a tile before a node's
  in-tile valence passes 63 (the rank byte), so that node's forces come from
  several tiles and several partials; K*x and a step against the oracle.
- A batch with no simulation, or with every simulation inactive, steps as a
  no-op.

Tolerances as in the other leaf tests: K*x 1e-6 normwise (the GPU evaluates
Ke u in the Piola form with FP32 arithmetic), steps 1e-4 of the displacement
from rest with a converging CG (rel 1e-6), CG iterations within one.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def gpu_params(cf, **kw):
    base = dict(fixed_dt=DT, gravity=(0.0, -9.81, 0.0), young_modulus=1e5, poisson_ratio=0.3, damping_ratio=0.0,
                rm=cf.RM_DAPD, cg_rel_epsilon=1e-6, cg_abs_epsilon=1e-12, cg_max_iter=2000, density=1000.0)
    base.update(kw)
    return cf.default_solid_params(**base)


def orc_params(orc, **kw):
    base = dict(young_modulus=1e5, poisson_ratio=0.3, rm=orc.RM_DAPD, cg_rel_epsilon=1e-6, cg_abs_epsilon=1e-12,
                cg_max_iter=2000, gravity=(0.0, -9.81, 0.0), reduce_fp64=1)
    base.update(kw)
    return orc.default_params(**base)


def step_pair(cf, orc, pos, tets, knc, steps=3, **kw):
    """Steps the leaf and the oracle side by side; returns the largest error
    relative to the displacement from rest (a step of a solid near rest moves
    it by a few thousand ulps of its coordinates, so one ulp of difference
    would be 1e-3 of that step), and the CG iteration pairs."""
    s = cf.LeafSolid3D(pos, tets, gpu_params(cf, **kw), knc=knc)
    o = orc.Solid(pos, tets, orc_params(orc, **kw), knc)
    worst, iters = 0.0, []
    for _ in range(steps):
        sg, so = s.fixed_step(), o.fixed_step()
        iters.append((sg.cg_iters, so.cg_iters))
        xg, _ = s.state()
        xo, _ = o.state()
        worst = max(worst, np.abs(xg - xo).max() / max(np.abs(xo - pos).max(), 1e-30))
    s.close()
    return worst, iters, xg, xo


def orient(pos, tets):
    """Swaps two vertices of every element of negative volume."""
    x = pos.reshape(-1, 3)
    t = tets.reshape(-1, 4).copy()
    d = np.einsum("ij,ij->i", np.cross(x[t[:, 1]] - x[t[:, 0]], x[t[:, 2]] - x[t[:, 0]]), x[t[:, 3]] - x[t[:, 0]])
    neg = d < 0
    t[neg, 2], t[neg, 3] = t[neg, 3].copy(), t[neg, 2].copy()
    return t.ravel().astype(np.uint32)


def fan_mesh(sectors=48, radius=0.05, height=0.02):
    """A disc of prisms around the z axis: every prism (c0, p_i, p_i+1 / c1,
    q_i, q_i+1) is split into three tetrahedra that all hold the bottom centre
    c0, so c0 has 3 * sectors elements (144 for 48 sectors)."""
    ang = 2 * np.pi * np.arange(sectors) / sectors
    ring = np.stack([radius * np.cos(ang), radius * np.sin(ang)], axis=1)
    pts = [(0.0, 0.0, 0.0), (0.0, 0.0, height)]
    pts += [(px, py, 0.0) for px, py in ring]
    pts += [(px, py, height) for px, py in ring]
    pos = np.array(pts, np.float32).ravel()
    c0, c1 = 0, 1
    tets = []
    for i in range(sectors):
        j = (i + 1) % sectors
        a0, a1, a2 = c0, 2 + i, 2 + j
        b0, b1, b2 = c1, 2 + sectors + i, 2 + sectors + j
        tets += [(a0, a1, a2, b2), (a0, a1, b2, b1), (a0, b1, b2, b0)]
    return pos, orient(pos, np.array(tets, np.uint32))


def test_single_tet_leaf_vs_oracle(gpu, orc):
    pos = np.array([0, 0, 0, 0.01, 0, 0, 0, 0.01, 0, 0, 0, 0.01], np.float32)
    tets = np.array([0, 1, 2, 3], np.uint32)
    knc = np.array([0], np.uint32)
    worst, iters, xg, xo = step_pair(gpu, orc, pos, tets, knc, steps=4)
    assert worst <= 1e-4, worst
    assert all(abs(a - b) <= 1 for a, b in iters), iters
    assert np.array_equal(xg[:3], pos[:3])  # the held node stays


def test_single_tet_in_a_batch_equals_leaf(gpu):
    """The one-element mesh batched with the 10x10x40-cube beam: both equal
    their single-mesh steps bit for bit."""
    small = (np.array([0, 0, 0, 0.01, 0, 0, 0, 0.01, 0, 0, 0, 0.01], np.float32), np.array([0, 1, 2, 3], np.uint32),
             np.array([0], np.uint32))
    bpos, btets = gpu.mesh_box((0.05, 0.05, 0.2), (11, 11, 41))
    big = (bpos, btets, np.nonzero(np.isclose(bpos[2::3], -0.2))[0].astype(np.uint32))
    p = gpu_params(gpu, cg_rel_epsilon=1e-4, cg_abs_epsilon=1e-6, cg_max_iter=200)
    plain = []
    for pos, tets, knc in (small, big):
        s = gpu.LeafSolid3D(pos, tets, p, knc=knc)
        st = [s.fixed_step() for _ in range(3)]
        plain.append((s.state(), [(t.cg_iters, t.cg_reason, t.elastic_energy) for t in st]))
        s.close()
    for mode in (gpu.BATCH_STEP_DEVICE, gpu.BATCH_STEP_HOST):
        b = gpu.Batch(step_mode=mode)
        sids = [b.add_sim(b.add_mesh(pos, tets), p, knc=knc) for pos, tets, knc in (small, big)]
        hist = {sid: [] for sid in sids}
        for _ in range(3):
            st = b.fixed_step()
            for sid in sids:
                hist[sid].append((st[sid].cg_iters, st[sid].cg_reason, st[sid].elastic_energy))
        for sid, (pos, _, _), (state, h) in zip(sids, (small, big), plain):
            x, v = b.state(sid, pos.size // 3)
            assert np.array_equal(x, state[0]) and np.array_equal(v, state[1]) and hist[sid] == h, mode
        b.close()


def test_unreferenced_nodes_fall_freely(gpu, orc):
    """A 4x4x6-cube beam with five vertices no element uses, two of them
    between the beam's node ids."""
    pos, tets = gpu.mesh_box((0.02, 0.02, 0.03), (5, 5, 7))
    n = pos.size // 3
    extra = np.array([[0.1, 0.1, 0.1], [0.2, 0.0, 0.0], [0.0, 0.3, 0.0], [0.1, 0.2, 0.3], [0.3, 0.3, 0.3]], np.float32)
    # renumber so that two unused vertices sit at ids 3 and 40 of the new order
    order = list(range(n))
    order.insert(3, n)
    order.insert(40, n + 1)
    order += [n + 2, n + 3, n + 4]
    allpos = np.concatenate([pos.reshape(-1, 3), extra])
    new_of_old = np.empty(len(order), np.uint32)
    new_of_old[np.array(order)] = np.arange(len(order), dtype=np.uint32)
    pos2 = allpos[np.array(order)].ravel().astype(np.float32)
    tets2 = new_of_old[tets].astype(np.uint32)
    knc = new_of_old[np.nonzero(np.isclose(pos[2::3], -0.03))[0]].astype(np.uint32)
    worst, iters, xg, xo = step_pair(gpu, orc, pos2, tets2, knc, steps=3)
    assert worst <= 1e-4, worst
    assert all(abs(a - b) <= 1 for a, b in iters), iters
    free = np.array([3, 40, n + 2, n + 3, n + 4])
    x0 = pos2.reshape(-1, 3)[free]
    # three backward-Euler steps of free fall: y_k = y_{k-1} + g dt, x += dt y
    drop = -9.81 * DT * DT * (1 + 2 + 3)
    assert np.allclose(xg.reshape(-1, 3)[free][:, 1] - x0[:, 1], drop, rtol=1e-4)
    assert np.array_equal(xg.reshape(-1, 3)[free][:, [0, 2]], x0[:, [0, 2]])


def test_every_node_held(gpu, orc):
    pos, tets = gpu.mesh_box((0.02, 0.02, 0.03), (4, 4, 5))
    knc = np.arange(pos.size // 3, dtype=np.uint32)
    s = gpu.LeafSolid3D(pos, tets, gpu_params(gpu), knc=knc)
    o = orc.Solid(pos, tets, orc_params(orc), knc)
    for _ in range(2):
        sg, so = s.fixed_step(), o.fixed_step()
        assert sg.cg_iters == so.cg_iters == 0
        assert sg.cg_reason == gpu.CG_CONVERGED_ABS and so.cg_abs == 0.0
    xg, vg = s.state()
    xo, vo = o.state()
    assert np.array_equal(xg, pos) and np.array_equal(xo, pos)
    assert not vg.any() and not vo.any()
    s.close()


def test_high_valence_node_kx_and_step_vs_oracle(gpu, orc):
    pos, tets = fan_mesh()
    N = pos.size // 3
    assert np.count_nonzero(tets == 0) == 144  # above the 63 elements a tile may give one node
    p = gpu.default_fem_params(young_modulus=1e5, poisson_ratio=0.3, rm=gpu.RM_DAPD)
    rng = np.random.default_rng(5)
    x = pos + rng.uniform(-5e-4, 5e-4, pos.size).astype(np.float32)
    for threads in (128, 256):
        gpu.set_tile_threads(threads)
        try:
            s = gpu.TetrahedronElementSet3(pos, tets, p)
        finally:
            gpu.set_tile_threads(256)
        o = orc.TetSet(pos, tets, 1e5, 0.3)
        s.begin_evaluation(gpu.DeviceArray.from_host(x))
        Ro, _, _, _ = o.begin_evaluation(x, rm=gpu.RM_DAPD)
        R, _ = s.rotations()
        assert np.array_equal(R.reshape(-1), Ro)
        v = rng.standard_normal(3 * N).astype(np.float32)
        df = gpu.DeviceArray.from_host(np.zeros(3 * N, np.float32))
        s.df_x(gpu.DeviceArray.from_host(v), df)
        dg, do = df.numpy(), o.df_x(v)
        assert np.linalg.norm(dg - do) <= 1e-6 * np.linalg.norm(do), threads
        # the shared node's row alone, which sums several tiles' partials
        assert np.linalg.norm(dg[:3] - do[:3]) <= 1e-6 * np.linalg.norm(do), threads
        s.end_evaluation()
        s.close()
    knc = np.nonzero(np.isclose(pos[2::3], 0.0))[0].astype(np.uint32)[1::2]  # every other bottom node
    worst, iters, _, _ = step_pair(gpu, orc, pos, tets, knc, steps=3)
    assert worst <= 1e-4, worst
    assert all(abs(a - b) <= 1 for a, b in iters), iters


def test_empty_and_inactive_batches(gpu):
    b = gpu.Batch()
    assert b.fixed_step() == {}
    pos, tets = gpu.mesh_box((0.02, 0.02, 0.03), (3, 3, 4))
    sid = b.add_sim(b.add_mesh(pos, tets), gpu_params(gpu))
    b.set_active(sid, False)
    assert b.fixed_step() == {}
    x, v = b.state(sid, pos.size // 3)
    assert np.array_equal(x, pos) and not v.any()
    b.set_active(sid, True)
    st = b.fixed_step()
    assert list(st) == [sid] and st[sid].cg_iters > 0
    b.close()
