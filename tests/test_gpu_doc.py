"""GPU: DAPD's degenerate branch through the C-ABI against the oracle:
UpdateDegeneration / ComputeDoC (ms/fem/TetrahedronElementSet3.cpp:477-743)
and Compute_R_DAPD_Degenerate (ms/fem/Rotation.cpp:364-545).

- BeginEvaluation(prev_x, x) on thousands of disjoint elements moving through
  vertex-face and edge-edge collapses and random motions: det F bit-exact,
  the same DoC bytes (the GPU's SVD and double-precision libm may differ from
  the host's in the last ulp, which can flip a borderline crossing class, so
  at most 0.5% may differ), R within 1e-5 where the DoC agrees, the DoC state
  kept and cleared across evaluations, fallbacks counted alike;
- non-default projection factors;
- QIE-v steps from a state with inverted elements, leaf vs oracle;
- batched steps of such a state bit-identical to the leaf's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E, NU, THR = 1e5, 0.3, 0.1
E_BOX = 1e7  # stiff enough for the 8 m box under gravity


def rest_tet(rng):
    while True:
        a = rng.normal(size=(4, 3))
        Dm = (a[1:] - a[0]).T
        if np.linalg.det(Dm) > 0.3 and np.linalg.cond(Dm) < 6:
            return a


def random_rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def motions(rng, T, kinds=tuple(range(12))):
    """Rest and current corners of T disjoint elements: of the kinds 0-11, 0-2
    are vertex-face collapses, 3-5 edge-edge, 6-9 random motions, 10-11 mild."""
    A = np.zeros((T, 4, 3))
    B = np.zeros((T, 4, 3))
    for e in range(T):
        kind = kinds[e % len(kinds)]
        if kind < 3:  # vertex through the opposite face
            a = rest_tet(rng)
            i = int(rng.integers(4))
            c = np.delete(a, i, axis=0).mean(axis=0)
            b = a.copy()
            b[i] = c - rng.uniform(0.0, 0.8) * (a[i] - c) + 0.05 * rng.normal(size=3)
        elif kind < 6:  # edge (0, j) through edge (k, l), rotated and shifted
            j = int(rng.integers(1, 4))
            k, l = [m for m in (1, 2, 3) if m != j]
            a = np.zeros((4, 3))
            a[0], a[j] = (0.0, 0.0, 0.8), (1.0, 0.1, 0.8)
            a[k], a[l] = (0.5, -0.6, 0.0), (0.45, 0.6, 0.0)
            if np.linalg.det((a[1:] - a[0]).T) < 0:
                a[[k, l]] = a[[l, k]]
            b = a.copy()
            b[0, 2] = b[j, 2] = rng.uniform(-0.7, 0.05)
            Q, t = random_rotation(rng), rng.normal(size=3)
            a, b = a @ Q.T + t, b @ Q.T + t
        elif kind < 10:
            a = rest_tet(rng)
            b = a + rng.normal(size=(4, 3)) * 1.2
        else:
            a = rest_tet(rng)
            b = a + rng.normal(size=(4, 3)) * 0.05
        A[e], B[e] = a, b
    return A.astype(np.float32).reshape(-1), B.astype(np.float32).reshape(-1)


def compare(gpu, orc, s, o, x, prev, rm, allowed):
    s.begin_evaluation(gpu.DeviceArray.from_host(x), prev_x=gpu.DeviceArray.from_host(prev))
    Ro, deto, _, _ = o.begin_evaluation(x, rm=rm, thr=THR, prev_x=prev)
    Rg, detg = s.rotations()
    assert np.array_equal(detg, deto)
    dg, do = s.doc(), o.doc()
    same = dg == do
    assert (~same).sum() <= allowed, f"{(~same).sum()} DoC bytes differ"
    err = np.abs(Rg.reshape(-1, 9) - Ro.reshape(-1, 9)).max(axis=1)
    assert err[same].max() <= 1e-5, err[same].max()
    if rm == orc.RM_DAPD and same.all():
        assert s.num_fallback() == o.num_fallback()
    s.end_evaluation()
    return dg, do, deto


@pytest.mark.parametrize("rm", ["RM_DAPD", "RM_MSVD"])
def test_doc_and_rotations_vs_oracle(gpu, orc, rm):
    rng = np.random.default_rng(21)
    T = 6000
    A, B = motions(rng, T)
    tets = np.arange(4 * T, dtype=np.uint32)
    s = gpu.TetrahedronElementSet3(A, tets, gpu.default_fem_params(young_modulus=E, poisson_ratio=NU,
                                                                    rm=getattr(gpu, rm)))
    o = orc.TetSet(A, tets, E, NU)
    allowed = int(0.005 * T)
    # rest -> current: new DoCs for every element that collapsed on the way
    dg, do, det1 = compare(gpu, orc, s, o, B, A, getattr(orc, rm), allowed)
    degenerate = det1 <= THR
    valid = (do >> 2) <= (do & 3)
    assert degenerate.sum() > 0.4 * T and (valid & degenerate).sum() > 0.3 * T
    assert np.all(do[~degenerate] == orc.DOC_INVALID)
    vf = valid & ((do >> 2) == (do & 3))
    assert vf.sum() > 0.1 * T and (valid & ~vf).sum() > 0.05 * T  # both crossing classes occur
    # degenerate at prev and now: the DoC is kept (continue from the GPU's own)
    o.set_doc(dg)
    B2 = (B + rng.normal(size=B.size).astype(np.float32) * 0.01).astype(np.float32)
    dg2, do2, det2 = compare(gpu, orc, s, o, B2, B, getattr(orc, rm), allowed)
    kept = (det1 <= THR) & (det2 <= THR)
    assert np.array_equal(dg2[kept], dg[kept])
    # back at rest: every DoC cleared
    dg3, _, _ = compare(gpu, orc, s, o, A, B2, getattr(orc, rm), 0)
    assert np.all(dg3 == orc.DOC_INVALID)
    # degenerate without a DoC: DAPD falls back to R = Id
    if rm == "RM_DAPD":
        s.set_doc(np.full(T, orc.DOC_INVALID, np.uint8))
        o.set_doc(np.full(T, orc.DOC_INVALID, np.uint8))
        compare(gpu, orc, s, o, B, B, orc.RM_DAPD, 0)
        s.begin_evaluation(gpu.DeviceArray.from_host(B), prev_x=gpu.DeviceArray.from_host(B))
        assert s.num_fallback() == int(degenerate.sum())
        s.end_evaluation()
    s.close()


def test_projection_factors_vs_oracle(gpu, orc):
    """Non-default L / NL factors and exponent on the collapse motions (with
    random deep inversions the power term makes the projected F so
    ill-conditioned that the 10-step float32 Newton polar does not converge and
    a last-ulp difference in powf decides between finite and NaN on either
    side)."""
    rng = np.random.default_rng(22)
    T = 3000
    A, B = motions(rng, T, kinds=(0, 1, 2, 3, 4, 5))
    tets = np.arange(4 * T, dtype=np.uint32)
    fem = gpu.default_fem_params(young_modulus=E, poisson_ratio=NU, rm=gpu.RM_DAPD)
    fem.dapd_l_factor, fem.dapd_nl_factor, fem.dapd_nl_exponent = 1.5, 0.2, 2.0
    s = gpu.TetrahedronElementSet3(A, tets, fem)
    o = orc.TetSet(A, tets, E, NU)
    o.set_dapd(1.5, 0.2, 2.0)
    compare(gpu, orc, s, o, B, A, orc.RM_DAPD, int(0.005 * T))
    s.close()


def inverted_box(cf):
    """A 4x4x8-cube box of 1 m cubes, z-min face held, with three interior
    nodes pushed through their neighbours so that a dozen tets start inverted.
    Metre-sized cubes: for cm-sized ones the reference's absolute epsilon in
    the time-of-collapse solve (IsZero, 1e-4) rejects a single moving vertex's
    (linear) collapse and no DoC is found."""
    pos, tets = cf.mesh_box((2.0, 2.0, 4.0), (5, 5, 9))
    knc = np.nonzero(np.isclose(pos[2::3], -4.0))[0].astype(np.uint32)
    x = pos.reshape(-1, 3).copy()
    interior = np.nonzero((np.abs(x[:, 0]) < 1.5) & (np.abs(x[:, 1]) < 1.5) & (x[:, 2] > -3.0) &
                          (x[:, 2] < 3.0))[0]
    for n, d in zip(interior[[3, 10, 17]], ([1.6, 0.0, 0.0], [0.0, -1.7, 0.4], [0.0, 0.0, 1.8])):
        x[n] += d
    return pos, tets, knc, x.astype(np.float32).reshape(-1)


def test_inverted_leaf_steps_vs_oracle(gpu, orc):
    pos, tets, knc, x0 = inverted_box(gpu)
    o_set = orc.TetSet(pos, tets, E_BOX, NU)
    _, det, _, _ = o_set.begin_evaluation(x0, rm=orc.RM_DAPD, prev_x=pos)
    doc = o_set.doc()
    assert (det <= 0).sum() >= 6 and ((doc >> 2) <= (doc & 3)).sum() >= 6  # inverted, with DoCs
    p = gpu.default_solid_params(fixed_dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), young_modulus=E_BOX,
                                 poisson_ratio=NU, damping_ratio=0.0, rm=gpu.RM_DAPD, cg_rel_epsilon=1e-6,
                                 cg_abs_epsilon=1e-12, cg_max_iter=3000, density=1000.0)
    s = gpu.LeafSolid3D(pos, tets, p, knc=knc)
    s.set_state(x0, np.zeros_like(x0))
    op = orc.default_params(young_modulus=E_BOX, poisson_ratio=NU, rm=orc.RM_DAPD, cg_rel_epsilon=1e-6,
                            cg_abs_epsilon=1e-12, cg_max_iter=3000, gravity=(0.0, -9.81, 0.0), reduce_fp64=1)
    o = orc.Solid(pos, tets, op, knc=knc)
    o.set_state(x0, np.zeros_like(x0))
    prev = x0
    for step in range(3):
        print(f"[doc] leaf step {step}", flush=True)
        s.fixed_step()
        o.fixed_step()
        xg, _ = s.state()
        xo, _ = o.state()
        d = np.abs(xg - xo).max() / np.abs(xo - prev).max()
        assert d <= 1e-4, (step, d)
        prev = xo
    s.close()


def test_inverted_batch_equals_leaf(gpu):
    pos, tets, knc, x0 = inverted_box(gpu)
    v0 = np.zeros_like(x0)
    runs = []
    for rm in (gpu.RM_DAPD, gpu.RM_MSVD):
        p = gpu.default_solid_params(fixed_dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), young_modulus=E_BOX,
                                     poisson_ratio=NU, damping_ratio=0.0, rm=rm, cg_rel_epsilon=1e-4,
                                     cg_abs_epsilon=1e-6, cg_max_iter=200, density=1000.0)
        s = gpu.LeafSolid3D(pos, tets, p, knc=knc)
        s.set_state(x0, v0)
        for _ in range(3):
            s.fixed_step()
        runs.append((p, s.state()))
        s.close()
    b = gpu.Batch()
    mid = b.add_meshes([(pos, tets)], threads=1)[0]
    sids = [b.add_sim(mid, p, knc=knc, x=x0, v=v0) for p, _ in runs]
    for _ in range(3):
        b.fixed_step()
    for sid, (_, (xl, vl)) in zip(sids, runs):
        xb, vb = b.state(sid, pos.size // 3)
        assert np.array_equal(xb, xl) and np.array_equal(vb, vl)
    b.close()
