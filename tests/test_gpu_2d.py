"""GPU: the 2D twin (workgroup-resident step kernel; the multi-workgroup path
for meshes beyond one workgroup) against the oracle, and batched against alone.

Tolerances: lumped masses bit-exact (same FP32 host arithmetic); the GPU
applies Ke in the equivalent Piola form, so steps agree to rounding
amplified by the CG: positions within 1e-4 relative after one step (5e-5 with
a converged CG), CG iterations within +-1. A simulation's result in a batch is
bit-identical to the same simulation alone (same kernel, its own group).
"""
import numpy as np
import pytest

from conftest import golden, rel_node_err

pytestmark = pytest.mark.gpu

EDGE = 0.01


def grid(cf, n):
    half = (0.5 * EDGE * n, 0.5 * EDGE * n)
    pos, tris = cf.mesh_box2(half, (n + 1, n + 1))
    knc = np.nonzero(np.isclose(pos[1::2], -half[1]))[0].astype(np.uint32)  # bottom edge
    return pos, tris, knc


def params(cf, E=1e5, nu=0.3, max_iter=50, rel=1e-4, rm=None):
    return cf.default_solid_params(fixed_dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), young_modulus=E,
                                   poisson_ratio=nu, damping_ratio=0.0, rm=cf.RM_DAPD if rm is None else rm,
                                   cg_rel_epsilon=rel, cg_abs_epsilon=1e-6, cg_max_iter=max_iter, density=1000.0)


def oracle_params(orc, E=1e5, nu=0.3, max_iter=50, rel=1e-4, rm=None):
    return orc.default_params(young_modulus=E, poisson_ratio=nu, rm=orc.RM_DAPD if rm is None else rm,
                              cg_rel_epsilon=rel, cg_abs_epsilon=1e-6, cg_max_iter=max_iter,
                              gravity=(0.0, -9.81, 0.0), reduce_fp64=1)


def test_lumped_masses_bitexact(gpu, orc):
    for n in (4, 13):
        pos, tris, knc = grid(gpu, n)
        b = gpu.Batch2()
        sid = b.add_sim(pos, tris, params(gpu), knc=knc)
        o = orc.Solid2(pos, tris, oracle_params(orc), knc=knc)
        assert np.array_equal(b.node_masses(sid), o.node_masses())
        b.close()


@pytest.mark.parametrize("n,E,nu", [(4, 1e5, 0.3), (12, 3e5, 0.25), (20, 5e4, 0.4), (30, 2e6, 0.35)])
def test_steps_vs_oracle(gpu, orc, n, E, nu):
    """n = 4, 12, 20, 30 use groups of 1, 2, 4 and 8 wavefronts."""
    pos, tris, knc = grid(gpu, n)
    s = gpu.LeafSolid2D(pos, tris, params(gpu, E, nu), knc=knc)
    o = orc.Solid2(pos, tris, oracle_params(orc, E, nu), knc=knc)
    for step in range(3):
        sg, so = s.fixed_step(), o.fixed_step()
        assert abs(sg.cg_iters - so.cg_iters) <= 1, (step, sg.cg_iters, so.cg_iters)
        xg, vg = s.state()
        xo, vo = o.state()
        err = rel_node_err(np.c_[xg.reshape(-1, 2), np.zeros(xg.size // 2)],
                           np.c_[xo.reshape(-1, 2), np.zeros(xo.size // 2)])
        assert err <= 1e-4 * (step + 1), (step, err)
        assert abs(sg.elastic_energy - so.elastic_energy) <= 1e-3 * abs(so.elastic_energy) + 1e-12
    assert np.abs(vg.reshape(-1, 2)[knc]).max() == 0


def test_converged_step_vs_oracle(gpu, orc):
    pos, tris, knc = grid(gpu, 10)
    s = gpu.LeafSolid2D(pos, tris, params(gpu, max_iter=1000, rel=1e-6), knc=knc)
    o = orc.Solid2(pos, tris, oracle_params(orc, max_iter=1000, rel=1e-6), knc=knc)
    sg, so = s.fixed_step(), o.fixed_step()
    assert sg.cg_reason in (gpu.CG_CONVERGED_REL, gpu.CG_CONVERGED_ABS)
    xg, _ = s.state()
    xo, _ = o.state()
    d = np.abs(xg - xo).max() / np.abs(xo - pos).max()
    assert d <= 5e-5, d


def test_batch_equals_alone_bitexact(gpu):
    rng = np.random.default_rng(3)
    sizes = [4, 9, 12, 15, 18, 22, 27, 32, 5, 16, 24, 31, 11, 6]
    cases = []
    for n in sizes:
        pos, tris, knc = grid(gpu, n)
        E = float(np.exp(rng.uniform(np.log(1e4), np.log(1e7))))
        nu = float(rng.uniform(0.2, 0.45))
        cases.append((pos, tris, knc, params(gpu, E, nu)))
    b = gpu.Batch2()
    ids = [b.add_sim(p, t, pr, knc=k) for p, t, k, pr in cases]
    info = b.info()
    assert info["sims"] == len(sizes)
    hist = [b.fixed_step() for _ in range(3)]
    for i, (p, t, k, pr) in enumerate(cases):
        s = gpu.LeafSolid2D(p, t, pr, knc=k)
        st = [s.fixed_step() for _ in range(3)]
        x, v = s.state()
        xb, vb = b.state(ids[i])
        assert np.array_equal(x, xb) and np.array_equal(v, vb), f"mesh {i} (n={sizes[i]}) differs in the batch"
        for j in range(3):
            assert (st[j].cg_iters, st[j].cg_reason, st[j].cg_rel_residual) == \
                   (hist[j][i].cg_iters, hist[j][i].cg_reason, hist[j][i].cg_rel_residual)
        s.close()
    b.close()


def test_rotations2_against_reference_golden(gpu):
    """1024 golden F (56 with det F <= 0.1, 29 inverted): MSVD bit-exact with
    the reference's Compute_R_SVD including its 2x2 SVD branch; DAPD polar
    above the threshold and R = Id (flagged) below."""
    g = golden("tris.npz")
    R, fb = gpu.rotations2(g["F"], g["det"], gpu.RM_MSVD)
    assert np.array_equal(R, g["R_svd"]) and not fb.any()
    R, fb = gpu.rotations2(g["F"], g["det"], gpu.RM_DAPD)
    ok = g["det"] > 0.1
    assert np.array_equal(R[ok], g["R_pd"][ok])
    assert np.array_equal(R[~ok], np.tile(np.eye(2, dtype=np.float32).ravel(), ((~ok).sum(), 1)))
    assert np.array_equal(fb.astype(bool), ~ok)


def test_rotations2_degenerate_vs_oracle(gpu, orc):
    """The 2x2 SVD branch on generic, small, inverted, rank-1, near rank-1 and
    zero F: bit-exact with the oracle (NaN where the reference yields NaN)."""
    rng = np.random.default_rng(12)
    F = [rng.uniform(-1, 1, (4000, 4)), rng.uniform(-0.3, 0.3, (4000, 4))]
    u, v = rng.normal(size=(1000, 2)), rng.normal(size=(1000, 2))
    r1 = np.einsum("ni,nj->nij", u, v).reshape(-1, 4)
    F += [r1, r1 + 1e-4 * rng.normal(size=r1.shape), np.zeros((1, 4)), [[1e-5, 0, 0, 1e-5]], [[0, 0, 0, 1.0]]]
    F = np.ascontiguousarray(np.concatenate([np.asarray(f, np.float64) for f in F]).astype(np.float32))
    det = (F[:, 0] * F[:, 3] - F[:, 2] * F[:, 1]).astype(np.float32)
    assert (det <= 0.1).sum() > 5000
    R, _ = gpu.rotations2(F, det, gpu.RM_MSVD)
    want = np.zeros_like(F)
    for i in range(len(F)):
        orc.lib().orc_svd_rotation2(orc.fp(F[i]), float(det[i]), 0.1, orc.fp(want[i]))
    np.testing.assert_array_equal(R, want)


def _squashed_state(pos):
    """Upper half compressed to 4% of its height (det F = 0.04 <= 0.1), its top
    quarter also mirrored in x (det F = -0.04, inverted)."""
    x = pos.reshape(-1, 2).copy()
    up = x[:, 1] > 0
    x[up, 1] *= 0.04
    top = pos.reshape(-1, 2)[:, 1] > 0.5 * pos.reshape(-1, 2)[:, 1].max()
    x[top, 0] *= -1.0
    return np.ascontiguousarray(x.ravel())


@pytest.mark.parametrize("rm", ["MSVD", "DAPD"])
def test_degenerate_step_vs_oracle(gpu, orc, rm):
    """A converged step from a state with degenerate and inverted triangles:
    MSVD takes the 2x2 SVD rotation there; DAPD the projected rotation of the
    triangles that get a node of collapse on the way from rest and R = Id for
    the others (at this cm scale the reference's absolute epsilon in the
    time-of-collapse solve rejects many; counted as fallbacks, as the oracle
    counts them); positions within 1e-4 of the oracle relative to the
    displacement."""
    pos, tris, knc = grid(gpu, 12)
    x0 = _squashed_state(pos)
    s = gpu.LeafSolid2D(pos, tris, params(gpu, max_iter=1000, rel=1e-6, rm=getattr(gpu, "RM_" + rm)), knc=knc)
    o = orc.Solid2(pos, tris, oracle_params(orc, max_iter=1000, rel=1e-6, rm=getattr(orc, "RM_" + rm)), knc=knc)
    v0 = np.zeros_like(x0)
    s.set_state(x0, v0)
    o.set_state(x0, v0)
    sg, so = s.fixed_step(), o.fixed_step()
    assert sg.num_degenerate == so.num_degenerate > 0
    X = x0.reshape(-1, 2)
    t = tris.reshape(-1, 3)
    P = pos.reshape(-1, 2)
    def cross(a, b):
        return a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]

    dF = cross(X[t[:, 1]] - X[t[:, 0]], X[t[:, 2]] - X[t[:, 0]]) / cross(P[t[:, 1]] - P[t[:, 0]], P[t[:, 2]] - P[t[:, 0]])
    n_low = int((dF <= 0.1).sum())
    assert n_low > 20
    if rm == "DAPD":
        assert np.array_equal(s.noc(), o.noc())
        assert s.num_fallback() == o.num_fallback() and 0 < o.num_fallback() < n_low
    else:
        assert s.num_fallback() == 0
    xg, _ = s.state()
    xo, _ = o.state()
    d = np.abs(xg - xo).max() / np.abs(xo - x0).max()
    assert d <= 1e-4, d
    s.close()


def test_msvd_and_dapd_differ_on_degenerate_elements(gpu):
    """The SVD branch is exercised: the same degenerate state steps differently
    under MSVD (SVD rotation) and DAPD (R = Id)."""
    pos, tris, knc = grid(gpu, 12)
    x0 = _squashed_state(pos)
    out = []
    for rm in (gpu.RM_MSVD, gpu.RM_DAPD):
        s = gpu.LeafSolid2D(pos, tris, params(gpu, rm=rm), knc=knc)
        s.set_state(x0, np.zeros_like(x0))
        s.fixed_step()
        out.append(s.state()[0])
        s.close()
    assert np.abs(out[0] - out[1]).max() > 1e-3 * np.abs(out[1] - x0).max()


def test_batch_refuses_mesh_beyond_one_workgroup(gpu):
    """Batches hold meshes that fit the resident kernel; a larger one is a
    single solid's (the multi-workgroup path below)."""
    pos, tris, knc = grid(gpu, 40)
    b = gpu.Batch2()
    with pytest.raises(gpu.CorofemError) as e:
        b.add_sim(pos, tris, params(gpu), knc=knc)
    assert "status -6" in str(e.value)
    b.close()


@pytest.mark.parametrize("n,E,nu,rm", [(40, 1e5, 0.3, "DAPD"), (64, 3e5, 0.4, "MSVD"), (90, 2e6, 0.25, "DAPD")])
def test_large_mesh_steps_vs_oracle(gpu, orc, n, E, nu, rm):
    """Meshes beyond one workgroup (3,200 / 8,192 / 16,200 triangles) take the
    multi-workgroup path, whose nodal forces sum in the reference's serial
    order: 3 steps against the oracle with the tolerances of the resident
    path (positions 1e-4 relative per step, CG iterations +-1, energy 1e-3)."""
    pos, tris, knc = grid(gpu, n)
    assert tris.size // 3 > 2048
    s = gpu.LeafSolid2D(pos, tris, params(gpu, E, nu, rm=getattr(gpu, "RM_" + rm)), knc=knc)
    o = orc.Solid2(pos, tris, oracle_params(orc, E, nu, rm=getattr(orc, "RM_" + rm)), knc=knc)
    for step in range(3):
        sg, so = s.fixed_step(), o.fixed_step()
        assert abs(sg.cg_iters - so.cg_iters) <= 1, (step, sg.cg_iters, so.cg_iters)
        xg, vg = s.state()
        xo, vo = o.state()
        err = rel_node_err(np.c_[xg.reshape(-1, 2), np.zeros(xg.size // 2)],
                           np.c_[xo.reshape(-1, 2), np.zeros(xo.size // 2)])
        print(f"\nn={n} step {step}: CG {sg.cg_iters}/{so.cg_iters}, rel node err {err:.2e}")
        assert err <= 1e-4 * (step + 1), (step, err)
        assert abs(sg.elastic_energy - so.elastic_energy) <= 1e-3 * abs(so.elastic_energy) + 1e-12
        assert sg.num_degenerate == so.num_degenerate
    assert np.abs(vg.reshape(-1, 2)[knc]).max() == 0
    s.close()


@pytest.mark.parametrize("rm", ["MSVD", "DAPD"])
def test_large_mesh_degenerate_step_vs_oracle(gpu, orc, rm):
    """The squashed and partly inverted state on a 40x40 grid (3,200
    triangles): degenerate counts, nodes of collapse and fallbacks equal to
    the oracle's, positions within 1e-4 of the displacement (converged CG)."""
    pos, tris, knc = grid(gpu, 40)
    x0 = _squashed_state(pos)
    s = gpu.LeafSolid2D(pos, tris, params(gpu, max_iter=1000, rel=1e-6, rm=getattr(gpu, "RM_" + rm)), knc=knc)
    o = orc.Solid2(pos, tris, oracle_params(orc, max_iter=1000, rel=1e-6, rm=getattr(orc, "RM_" + rm)), knc=knc)
    v0 = np.zeros_like(x0)
    s.set_state(x0, v0)
    o.set_state(x0, v0)
    for step in range(2):
        sg, so = s.fixed_step(), o.fixed_step()
        assert sg.num_degenerate == so.num_degenerate
        if step == 0:
            assert so.num_degenerate > 0
        if rm == "DAPD":
            assert np.array_equal(s.noc(), o.noc())
            assert s.num_fallback() == o.num_fallback()
        else:
            assert s.num_fallback() == 0
    xg, _ = s.state()
    xo, _ = o.state()
    d = np.abs(xg - xo).max() / np.abs(xo - x0).max()
    print(f"\n{rm}: {so.num_degenerate} degenerate, {s.num_fallback()} fallbacks, displacement error {d:.2e}")
    assert d <= 1e-4, d
    s.close()


def test_large_path_switches_at_the_resident_limits(gpu, orc):
    """The largest grid that fits one workgroup (2,048 triangles at 32x32)
    and the next (33x33, 2,178 triangles, multi-workgroup) both follow the
    oracle; the Step accumulator and set_state work on the large path."""
    for n in (32, 33):
        pos, tris, knc = grid(gpu, n)
        s = gpu.LeafSolid2D(pos, tris, params(gpu), knc=knc)
        o = orc.Solid2(pos, tris, oracle_params(orc), knc=knc)
        st = s.step(2.5 / 60.0)  # two fixed steps, the half carried over
        o.fixed_step()
        o.fixed_step()
        assert st.num_fixed_steps == 2
        xg, _ = s.state()
        xo, _ = o.state()
        err = rel_node_err(np.c_[xg.reshape(-1, 2), np.zeros(xg.size // 2)],
                           np.c_[xo.reshape(-1, 2), np.zeros(xo.size // 2)])
        assert err <= 2e-4, (n, err)
        s.set_state(pos, np.zeros_like(pos))
        assert np.array_equal(s.state()[0], pos) and not s.state()[1].any()
        s.close()


def _pushed_metre_grid(cf):
    """A 12x12 grid of 1 m cells with three interior nodes pushed through their
    neighbours (six inverted triangles, all with a node of collapse)."""
    pos, tris = cf.mesh_box2((6.0, 6.0), (13, 13))
    knc = np.nonzero(np.isclose(pos[1::2], -6.0))[0].astype(np.uint32)
    x = pos.reshape(-1, 2).copy()
    inner = np.nonzero((np.abs(x[:, 0]) < 5.5) & (x[:, 1] > -4.5) & (x[:, 1] < 5.5))[0]
    for n, d in zip(inner[[20, 45, 77]], ([1.4, 0.3], [-0.2, -1.5], [1.2, 1.1])):
        x[n] += d
    return pos, tris, knc, np.ascontiguousarray(x.astype(np.float32).ravel())


def test_noc_metre_grid_steps_vs_oracle(gpu, orc):
    pos, tris, knc, x0 = _pushed_metre_grid(gpu)
    s = gpu.LeafSolid2D(pos, tris, params(gpu, E=1e7, max_iter=1000, rel=1e-6), knc=knc)
    o = orc.Solid2(pos, tris, oracle_params(orc, E=1e7, max_iter=1000, rel=1e-6), knc=knc)
    v0 = np.zeros_like(x0)
    s.set_state(x0, v0)
    o.set_state(x0, v0)
    prev = x0
    for step in range(3):
        sg, so = s.fixed_step(), o.fixed_step()
        assert sg.num_degenerate == so.num_degenerate
        if step == 0:
            assert so.num_degenerate >= 6 and o.num_fallback() == 0 and (o.noc() != 3).sum() >= 6
        assert np.array_equal(s.noc(), o.noc()) and s.num_fallback() == o.num_fallback()
        xg, _ = s.state()
        xo, _ = o.state()
        d = np.abs(xg - xo).max() / np.abs(xo - prev).max()
        assert d <= 1e-4, (step, d)
        prev = xo
    s.close()


def test_noc_batch_equals_leaf_bitexact(gpu):
    """Degenerate DAPD states (cm grid, metre grid) and an MSVD one stepped in
    one batch: bit-identical to each alone, node-of-collapse bytes included."""
    pos_c, tris_c, knc_c = grid(gpu, 12)
    x_c = _squashed_state(pos_c)
    pos_m, tris_m, knc_m, x_m = _pushed_metre_grid(gpu)
    cases = [(pos_c, tris_c, knc_c, x_c, params(gpu)), (pos_m, tris_m, knc_m, x_m, params(gpu, E=1e7)),
             (pos_c, tris_c, knc_c, x_c, params(gpu, rm=gpu.RM_MSVD))]
    alone = []
    for p, t, k, x, pr in cases:
        s = gpu.LeafSolid2D(p, t, pr, knc=k)
        s.set_state(x, np.zeros_like(x))
        for _ in range(3):
            s.fixed_step()
        alone.append((s.state(), s.noc(), s.num_fallback()))
        s.close()
    b = gpu.Batch2()
    ids = [b.add_sim(p, t, pr, knc=k) for p, t, k, _, pr in cases]
    for i, (_, _, _, x, _) in zip(ids, cases):
        b.set_state(i, x, np.zeros_like(x))
    for _ in range(3):
        b.fixed_step()
    for i, ((xa, va), noc, fb) in zip(ids, alone):
        xb, vb = b.state(i)
        assert np.array_equal(xb, xa) and np.array_equal(vb, va)
        assert np.array_equal(b.noc(i), noc) and b.num_fallback(i) == fb
    b.close()
