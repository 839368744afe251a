"""CPU: pins the oracle (the CPU restatement of the reference) against
(1) the reference's own compiled math, via the committed golden vectors and,
    where it is built, oracle/_ref directly;
(2) what the mathematics fixes: rigid-body null space and symmetry of Ke,
    an independent float64 construction of Ke, SVD/polar closed forms, a
    dense solve of the CG system, free fall, mesh conformity.
"""
import numpy as np
import pytest

from conftest import golden, rel_node_err, beam_config1


# ------------------------------------------------------------ golden vectors
def test_ke_matches_reference_golden(orc):
    g = golden("ke_tets.npz")
    for i in range(len(g["V"])):
        r = np.ascontiguousarray(g["r"][i])
        K = np.zeros(144, np.float32)
        v = orc.lib().orc_compute_K(orc.fp(r), float(g["E"][i]), float(g["nu"][i]), orc.fp(K))
        assert v == g["V"][i]
        assert np.array_equal(K, g["K"][i]), i
        D = np.zeros(9, np.float32)
        m = (r[3:].reshape(3, 3) - r[:3]).T.copy()  # columns r1-r0, r2-r0, r3-r0
        orc.lib().orc_inverse3(orc.fp(np.ascontiguousarray(m.ravel())), orc.fp(D))
        assert np.array_equal(D, g["Dinv"][i]), i


def test_rotations_match_reference_golden(orc):
    g = golden("rotations.npz")
    L = orc.lib()
    for i in range(len(g["det"])):
        F = np.ascontiguousarray(g["F"][i])
        assert L.orc_det3(orc.fp(F)) == g["det"][i]
        R = np.zeros(9, np.float32)
        it = np.zeros(1, np.uint32)
        L.orc_pd3(orc.fp(F), float(g["det"][i]), orc.fp(R), orc.up(it))
        assert np.array_equal(R, g["R_pd"][i]), ("polar", i)
        L.orc_svd_rotation3(orc.fp(F), orc.fp(R))
        assert np.array_equal(R, g["R_svd"][i]), ("svd", i)
        U, S, Vt = np.zeros(9, np.float32), np.zeros(3, np.float32), np.zeros(9, np.float32)
        L.orc_svd3(orc.fp(F), orc.fp(U), orc.fp(S), orc.fp(Vt))
        assert np.array_equal(U, g["U"][i]) and np.array_equal(S, g["S"][i]) and np.array_equal(Vt, g["Vt"][i])


def test_sse_dot_matches_reference_golden(orc):
    g = golden("dots.npz")
    o = 0
    for L, want in zip(g["lens"], g["dot"]):
        a = np.ascontiguousarray(g["a"][o:o + L])
        b = np.ascontiguousarray(g["b"][o:o + L])
        assert orc.lib().orc_sse_dot(orc.fp(a), orc.fp(b), int(L)) == want
        o += L


def test_2d_matches_reference_golden(orc):
    g = golden("tris.npz")
    L = orc.lib()
    for i in range(len(g["A"])):
        r = np.ascontiguousarray(g["r"][i])
        K = np.zeros(36, np.float32)
        a = L.orc_compute_K2(orc.fp(r), float(g["E"][i]), float(g["nu"][i]), orc.fp(K))
        assert a == g["A"][i] and np.array_equal(K, g["K"][i]), i
    assert (g["det"] <= 0.1).sum() >= 50  # the fixture exercises the 2x2 SVD branch
    for i in range(len(g["det"])):
        F = np.ascontiguousarray(g["F"][i])
        R = np.zeros(4, np.float32)
        if g["det"][i] > 0.1:
            L.orc_pd2(orc.fp(F), float(g["det"][i]), orc.fp(R))
            assert np.array_equal(R, g["R_pd"][i]), i
        L.orc_svd_rotation2(orc.fp(F), float(g["det"][i]), 0.1, orc.fp(R))
        assert np.array_equal(R, g["R_svd"][i]), i  # polar above the threshold, 2x2 SVD below


def test_oracle_against_compiled_reference_live(orc):
    """Where oracle/_ref is built (reference sources present), compare on fresh seeds."""
    R = orc.ref()
    if R is None:
        pytest.skip("oracle/_ref not built here")
    rng = np.random.default_rng(7)
    L = orc.lib()
    for _ in range(500):
        r = rng.uniform(-1, 1, 12).astype(np.float32)
        K1, K2 = np.zeros(144, np.float32), np.zeros(144, np.float32)
        assert L.orc_compute_K(orc.fp(r), 3e4, 0.35, orc.fp(K1)) == R.ref_compute_K(orc.fp(r), 3e4, 0.35, orc.fp(K2))
        assert np.array_equal(K1, K2)
        F = (np.eye(3) + rng.uniform(-0.7, 0.7, (3, 3))).astype(np.float32).ravel()
        d = L.orc_det3(orc.fp(F))
        a, b = np.zeros(9, np.float32), np.zeros(9, np.float32)
        L.orc_pd3(orc.fp(F), d, orc.fp(a), None)
        R.ref_pd3(orc.fp(F), d, orc.fp(b))
        assert np.array_equal(a, b)
        L.orc_svd_rotation3(orc.fp(F), orc.fp(a))
        R.ref_R_svd3(orc.fp(F), d, 0.1, orc.fp(b))
        assert np.array_equal(a, b)
    # 2x2 SVD branch: generic, small, inverted, rank-1 and near rank-1 F
    F2 = [rng.uniform(-1, 1, 4) for _ in range(2000)] + [rng.uniform(-0.3, 0.3, 4) for _ in range(2000)]
    for _ in range(500):
        u, v = rng.normal(size=2), rng.normal(size=2)
        F2.append(np.outer(u, v).ravel())
        F2.append((np.outer(u, v) + 1e-4 * rng.normal(size=(2, 2))).ravel())
    F2 += [np.zeros(4), np.array([1e-5, 0, 0, 1e-5]), np.array([0, 0, 0, 1.0]), np.array([-1, 0, 0, 1.0])]
    for f in np.array(F2, np.float32):
        f = np.ascontiguousarray(f)
        d = np.float32(f[0] * f[3] - f[2] * f[1])
        a, b = np.zeros(4, np.float32), np.zeros(4, np.float32)
        L.orc_svd_rotation2(orc.fp(f), float(d), 0.1, orc.fp(a))
        R.ref_R_svd2(orc.fp(f), float(d), 0.1, orc.fp(b))
        assert np.array_equal(a, b), f


# ------------------------------------------------------------ closed forms
def _ke_float64(r, E, nu):
    """Independent Ke = V B^T E B from shape-function gradients (float64)."""
    r = r.reshape(4, 3).astype(np.float64)
    Dm = (r[1:] - r[0]).T
    V = np.linalg.det(Dm) / 6.0
    G = np.linalg.inv(Dm)                       # rows: grad N1..N3
    grads = np.vstack([-G.sum(0), G])           # grad N0..N3
    B = np.zeros((6, 12))
    for n, (a, b, c) in enumerate(grads):
        B[0, 3 * n], B[1, 3 * n + 1], B[2, 3 * n + 2] = a, b, c
        B[3, 3 * n], B[3, 3 * n + 1] = b, a
        B[4, 3 * n + 1], B[4, 3 * n + 2] = c, b
        B[5, 3 * n], B[5, 3 * n + 2] = c, a
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    Em = np.zeros((6, 6))
    Em[:3, :3] = lam
    Em[np.arange(3), np.arange(3)] += 2 * mu
    Em[3, 3] = Em[4, 4] = Em[5, 5] = mu
    return V * B.T @ Em @ B, V


def test_ke_independent_construction_and_null_space(orc):
    g = golden("ke_tets.npz")
    for i in range(0, len(g["V"]), 8):
        K = g["K"][i].reshape(12, 12).astype(np.float64)
        K64, V64 = _ke_float64(g["r"][i], float(g["E"][i]), float(g["nu"][i]))
        scale = np.abs(K64).max()
        assert np.abs(K - K64).max() <= 2e-4 * scale, i
        assert abs(g["V"][i] - V64) <= 1e-5 * abs(V64)
        assert np.abs(K - K.T).max() <= 1e-5 * scale
        # rigid modes: 3 translations + 3 infinitesimal rotations
        r = g["r"][i].reshape(4, 3).astype(np.float64)
        modes = [np.tile(np.eye(3)[k], 4) for k in range(3)]
        for k in range(3):
            w = np.eye(3)[k]
            modes.append(np.cross(w, r).ravel())
        for u in modes:
            assert np.abs(K @ u).max() <= 1e-4 * scale * np.abs(u).max()


def test_svd_structure_with_true_sqrt3_matches_numpy(orc):
    """With the true sqrt(3), the restated Smith/cofactor SVD is an SVD: pins
    every step except the reference's Sqrt<int>(3) truncation."""
    rng = np.random.default_rng(3)
    L = orc.lib()
    for _ in range(300):
        F = (np.eye(3) + rng.uniform(-0.5, 0.5, (3, 3))).astype(np.float32)
        U, S, Vt = np.zeros(9, np.float32), np.zeros(3, np.float32), np.zeros(9, np.float32)
        L.orc_svd3_sqrt3(orc.fp(np.ascontiguousarray(F.ravel())), orc.fp(U), orc.fp(S), orc.fp(Vt))
        s_np = np.linalg.svd(F.astype(np.float64), compute_uv=False)
        if np.linalg.det(F) < 0:
            s_np[2] = -s_np[2]
        if min(abs(s_np[0] - s_np[1]), abs(s_np[1] - s_np[2])) < 1e-2:
            continue
        assert np.allclose(S, s_np, rtol=2e-4, atol=2e-5)
        rec = U.reshape(3, 3) @ np.diag(S) @ Vt.reshape(3, 3)
        assert np.abs(rec - F).max() < 2e-3


def test_polar_newton_closed_form(orc):
    """F = Q diag(s), s in [0.7, 1.3]: the loose Newton polar factor is within
    the stop rule's reach of Q, and det F > 0.1 (undegenerate branch)."""
    g = golden("rotations.npz")
    n = len(g["s"])
    for i in range(n):
        F = g["F"][i]
        Q = g["Q"][i].reshape(3, 3)
        R = np.zeros(9, np.float32)
        it = np.zeros(1, np.uint32)
        orc.lib().orc_pd3(orc.fp(np.ascontiguousarray(F)), float(g["det"][i]), orc.fp(R), orc.up(it))
        assert g["det"][i] > 0.1
        assert 1 <= it[0] <= 10
        assert np.abs(R.reshape(3, 3) - Q).max() < 0.06


def test_svd2_degenerate_closed_forms(orc):
    """The 2x2 SVD branch (det F <= 0.1) against what the mathematics fixes:
    F = Q1 diag(s0, +-s1) Q2^T with s0 > s1 >= 0 and Q1, Q2 proper rotations
    gives R = Q1 Q2^T (the polar factor when det F > 0; with the smallest
    singular direction flipped when F is inverted); a rank-1 F = s u v^T maps
    v onto u for 0.05 <= s <= 0.1 (the branch tests singular values and rows
    of F^T F - lambda Id against absolute epsilons: above s = 0.1 the float32
    noise of the null singular value, about 3.5e-4 s, passes its 1e-4 zero
    test, below s = 0.05 whole rows fall under it; outside the window the
    reference itself returns no rotation, and only the bit-exact comparison
    with the compiled reference covers it); F = 0 gives the identity. The branch works on the eigenvectors of
    F^T F in float32, so its error grows with the squared condition number
    kappa = s0 / s1: the tolerance is 2e-4 + 60 eps kappa^2 (kappa <= 50),
    far below the O(1) error of a flipped sign or a transposed factor."""
    rng = np.random.default_rng(21)
    L = orc.lib()

    def rot(t):
        return np.array([[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]])

    def R_of(F):
        F = np.ascontiguousarray(F, np.float32).ravel()
        d = np.float32(F[0] * F[3] - F[2] * F[1])
        assert d <= 0.1
        R = np.zeros(4, np.float32)
        L.orc_svd_rotation2(orc.fp(F), float(d), 0.1, orc.fp(R))
        return R.reshape(2, 2).astype(np.float64)

    for _ in range(2000):
        Q1, Q2 = rot(rng.uniform(0, 2 * np.pi)), rot(rng.uniform(0, 2 * np.pi))
        s0 = rng.uniform(0.3, 1.5)
        s1 = rng.uniform(0.02 * s0, min(0.09 / s0, 0.5 * s0))
        tol = 2e-4 + 60 * np.finfo(np.float32).eps * (s0 / s1) ** 2
        for sign in (1.0, -1.0):
            R = R_of(Q1 @ np.diag([s0, sign * s1]) @ Q2.T)
            assert np.abs(R - Q1 @ Q2.T).max() <= tol, (s0, sign * s1)
            assert abs(np.linalg.det(R) - 1.0) <= tol and np.abs(R @ R.T - np.eye(2)).max() <= tol
        u, v = Q1[:, 0], Q2[:, 0]
        R = R_of(rng.uniform(0.05, 0.1) * np.outer(u, v))
        assert np.abs(R @ v - u).max() <= 2e-4
        assert abs(np.linalg.det(R) - 1.0) <= 1e-5
    assert np.array_equal(R_of(np.zeros((2, 2))), np.eye(2))


def test_mesh_box_conformity(orc, cf):
    pos, tets = orc.mesh_box((0.05, 0.05, 0.2), (11, 11, 41))
    assert pos.size // 3 == 4961 and tets.size // 4 == 20000
    p2, t2 = cf.mesh_box((0.05, 0.05, 0.2), (11, 11, 41))
    assert np.array_equal(pos, p2) and np.array_equal(tets, t2)
    P = pos.reshape(-1, 3).astype(np.float64)
    T = tets.reshape(-1, 4)
    D = np.stack([P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]], P[T[:, 3]] - P[T[:, 0]]], 2)
    vol = np.linalg.det(D) / 6
    assert (vol > 0).all()
    assert abs(vol.sum() - 0.1 * 0.1 * 0.4) < 1e-9
    faces = np.sort(np.concatenate([T[:, [0, 1, 2]], T[:, [0, 1, 3]], T[:, [0, 2, 3]], T[:, [1, 2, 3]]]), 1)
    _, counts = np.unique(faces, axis=0, return_counts=True)
    assert counts.max() == 2
    boundary = (counts == 1).sum()
    assert boundary == 2 * 2 * (10 * 10 + 10 * 40 + 10 * 40)


def test_mesh_box2(orc, cf):
    pos, tris = orc.mesh_box2((0.5, 0.25), (5, 3))
    p2, t2 = cf.mesh_box2((0.5, 0.25), (5, 3))
    assert np.array_equal(pos, p2) and np.array_equal(tris, t2)
    P = pos.reshape(-1, 2)
    T = tris.reshape(-1, 3)
    a = 0.5 * ((P[T[:, 1], 0] - P[T[:, 0], 0]) * (P[T[:, 2], 1] - P[T[:, 0], 1]) -
               (P[T[:, 2], 0] - P[T[:, 0], 0]) * (P[T[:, 1], 1] - P[T[:, 0], 1]))
    assert (a > 0).all() and abs(a.sum() - 0.5) < 1e-6


# ------------------------------------------------------------ operators
def _dense_K(orc, pos, tets, R, E, nu):
    """Assembled sum_e R_e Ke R_e^T from the float64 construction (independent)."""
    N = pos.size // 3
    T = tets.reshape(-1, 4)
    Kg = np.zeros((3 * N, 3 * N))
    for e in range(T.shape[0]):
        ke, _ = _ke_float64(pos.reshape(-1, 3)[T[e]].ravel(), E, nu)
        Re = np.kron(np.eye(4), R.reshape(-1, 9)[e].reshape(3, 3).astype(np.float64))
        kr = Re @ ke @ Re.T
        idx = np.concatenate([[3 * n, 3 * n + 1, 3 * n + 2] for n in T[e]])
        Kg[np.ix_(idx, idx)] += kr
    return Kg


def test_df_x_and_f_e_against_dense_assembly(orc):
    pos, tets = orc.mesh_box((0.02, 0.02, 0.03), (3, 3, 4))
    E, nu = 1e5, 0.3
    s = orc.TetSet(pos, tets, E, nu)
    rng = np.random.default_rng(11)
    x = pos + rng.uniform(-2e-3, 2e-3, pos.size).astype(np.float32)
    R, det, nd, _ = s.begin_evaluation(x, rm=orc.RM_DAPD)
    assert nd == 0
    Kg = _dense_K(orc, pos, tets, R, E, nu)
    dx = rng.standard_normal(pos.size).astype(np.float32)
    df = s.df_x(dx)
    want = -Kg @ dx.astype(np.float64)
    assert np.linalg.norm(df - want) <= 1e-5 * np.linalg.norm(want)
    # f_e: -sum R Ke (R^T x - r)
    T = tets.reshape(-1, 4)
    want = np.zeros(pos.size)
    for e in range(T.shape[0]):
        ke, _ = _ke_float64(pos.reshape(-1, 3)[T[e]].ravel(), E, nu)
        Re = R[9 * e:9 * e + 9].reshape(3, 3).astype(np.float64)
        u = np.concatenate([Re.T @ x.reshape(-1, 3)[n] - pos.reshape(-1, 3)[n] for n in T[e]])
        f = -np.kron(np.eye(4), Re) @ (ke @ u)
        for j, n in enumerate(T[e]):
            want[3 * n:3 * n + 3] += f[3 * j:3 * j + 3]
    fe = s.f_e(x)
    assert np.linalg.norm(fe - want) <= 2e-4 * np.linalg.norm(want)
    # energy: 1/2 u^T K u summed
    V = s.V(x)
    assert V > 0


# ------------------------------------------------------------ solver / step
def test_cg_step_against_dense_solve(orc):
    """One QIE-v step with a tight CG: y solves (fM M + fK K) y = b on the free
    rows, with KNC rows pinned, like a dense float64 solve."""
    pos, tets = orc.mesh_box((0.02, 0.02, 0.03), (3, 3, 4))
    knc = np.nonzero(np.isclose(pos[2::3], -0.03))[0].astype(np.uint32)
    p = orc.default_params(rm=orc.RM_DAPD, cg_rel_epsilon=1e-7, cg_abs_epsilon=1e-12, cg_max_iter=500,
                           reduce_fp64=1)
    s = orc.Solid(pos, tets, p, knc)
    # perturb so that rotations and elastic forces are non-trivial
    rng = np.random.default_rng(5)
    x0 = pos + rng.uniform(-1e-3, 1e-3, pos.size).astype(np.float32)
    x0[np.repeat(knc, 3) * 3 + np.tile([0, 1, 2], knc.size)] = pos[np.repeat(knc, 3) * 3 + np.tile([0, 1, 2], knc.size)]
    v0 = rng.uniform(-0.1, 0.1, pos.size).astype(np.float32)
    v0.reshape(-1, 3)[knc] = 0
    s.set_state(x0, v0)
    ts = orc.TetSet(pos, tets, 1e5, 0.3)
    R, _, _, _ = ts.begin_evaluation(x0, rm=orc.RM_DAPD)
    Kg = _dense_K(orc, pos, tets, R, 1e5, 0.3)
    m = s.node_mass()
    dt = 1.0 / 60
    N = pos.size // 3
    fe = ts.f_e(x0).astype(np.float64)
    g = np.tile([0, -9.81, 0], N)
    b = dt * (m * g + fe) + m * v0
    A = m * np.eye(3 * N) + dt * dt * Kg
    free = np.ones(N, bool)
    free[knc] = False
    fr = np.repeat(free, 3)
    y = np.zeros(3 * N)
    y[fr] = np.linalg.solve(A[np.ix_(fr, fr)], b[fr])
    st = s.fixed_step(dt)
    x1, v1 = s.state()
    assert st.cg_iters > 5
    assert np.abs(v1 - y).max() <= 1e-4 * np.abs(y).max()
    assert np.abs(v1.reshape(-1, 3)[knc]).max() == 0


def test_free_fall_invariant(orc):
    """From rest with no KNC, elastic forces vanish and K annihilates the
    uniform velocity: one step gives v = dt g, x = r + dt^2 g exactly-ish."""
    pos, tets = orc.mesh_box((0.02, 0.02, 0.02), (3, 3, 3))
    p = orc.default_params(rm=orc.RM_MSVD)
    s = orc.Solid(pos, tets, p)
    st = s.fixed_step(1.0 / 60)
    x, v = s.state()
    dt = np.float32(1.0 / 60)
    assert np.allclose(v.reshape(-1, 3), [0, -9.81 * dt, 0], atol=1e-6)
    assert np.allclose(x, pos + np.tile([0, -9.81 * dt * dt, 0], pos.size // 3).astype(np.float32), atol=1e-7)
    assert st.num_degenerate == 0


def test_step_accumulator_strict_greater(orc):
    pos, tets = orc.mesh_box((0.02, 0.02, 0.02), (3, 3, 3))
    s = orc.Solid(pos, tets, orc.default_params(fixed_dt=1.0 / 60))
    assert s.step(1.0 / 60).num_fixed_steps == 0   # 1/60 > 1/60 is false
    assert s.step(1.0 / 60).num_fixed_steps == 1


def test_config1_oracle_noise_floor(orc):
    """Documents the sensitivity of config 1 (undamped, CG capped at 50): the
    reduction precision alone moves positions by O(1e-4) relative; this is
    the floor any non-bit-identical implementation is measured against."""
    pos, tets, knc = beam_config1(orc)
    a = orc.Solid(pos, tets, orc.default_params(rm=orc.RM_DAPD), knc)
    b = orc.Solid(pos, tets, orc.default_params(rm=orc.RM_DAPD, reduce_fp64=1), knc)
    for _ in range(10):
        sa, sb = a.fixed_step(), b.fixed_step()
        assert abs(sa.cg_iters - sb.cg_iters) <= 1
    err = rel_node_err(a.state()[0], b.state()[0])
    assert 1e-7 < err < 1e-3
