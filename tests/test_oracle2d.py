"""CPU: the oracle's 2D twin (TriangleElementSet2 + LeafDSH_Solid2D_FEM),
pinned against an independent float64 construction of the plane-strain
triangle stiffness, dense assembly, a dense solve of one QIE-v step, the
lumped-mass closed form and free fall. (The 2D Ke and polar factor themselves
are pinned bit-exactly to the reference in test_oracle.py.)"""
import numpy as np
import pytest


def ke2_float64(r, E, nu):
    """Plane-strain linear triangle from barycentric gradients (independent of
    the reference's B formula): K = A B^T C B, engineering shear."""
    r = np.asarray(r, np.float64).reshape(3, 2)
    M = np.column_stack([np.ones(3), r])
    G = np.linalg.inv(M)[1:, :]  # d(lambda_i)/d(x,y), 2x3
    area = 0.5 * np.linalg.det(np.column_stack([r[1] - r[0], r[2] - r[0]]))
    B = np.zeros((3, 6))
    for i in range(3):
        B[0, 2 * i] = G[0, i]
        B[1, 2 * i + 1] = G[1, i]
        B[2, 2 * i] = G[1, i]
        B[2, 2 * i + 1] = G[0, i]
    a = E / ((1 + nu) * (1 - 2 * nu))
    C = np.array([[a * (1 - nu), a * nu, 0], [a * nu, a * (1 - nu), 0], [0, 0, a * (0.5 - nu)]])
    return area * B.T @ C @ B, area


def dense_K2(pos, tris, R, E, nu):
    N = pos.size // 2
    K = np.zeros((2 * N, 2 * N))
    T = tris.reshape(-1, 3)
    P = pos.reshape(-1, 2)
    for e in range(T.shape[0]):
        ke, _ = ke2_float64(P[T[e]], E, nu)
        Re = R[4 * e:4 * e + 4].reshape(2, 2).astype(np.float64)
        Rb = np.kron(np.eye(3), Re)
        ke_w = Rb @ ke @ Rb.T
        idx = np.concatenate([[2 * n, 2 * n + 1] for n in T[e]])
        K[np.ix_(idx, idx)] += ke_w
    return K


def test_ke2_independent_construction(orc):
    rng = np.random.default_rng(2)
    for _ in range(50):
        r = rng.uniform(-1, 1, 6).astype(np.float32)
        if 0.5 * ((r[2] - r[0]) * (r[5] - r[1]) - (r[4] - r[0]) * (r[3] - r[1])) < 0.05:
            continue
        K = np.zeros(36, np.float32)
        a = orc.lib().orc_compute_K2(orc.fp(r), 1e5, 0.3, orc.fp(K))
        want, area = ke2_float64(r, 1e5, 0.3)
        assert abs(a - area) <= 1e-6 * abs(area)
        assert np.abs(K.reshape(6, 6) - want).max() <= 2e-5 * np.abs(want).max()
        # rigid translations in the null space, symmetric
        t = np.tile([1.0, 0.0], 3)
        assert np.abs(want @ t).max() <= 1e-9 * np.abs(want).max()
        assert np.abs(K.reshape(6, 6) - K.reshape(6, 6).T).max() <= 1e-5 * np.abs(K).max()


def test_lumped_masses_closed_form(orc):
    pos, tris = orc.mesh_box2((0.04, 0.03), (9, 7))
    p = orc.default_params(density=1000.0)
    s = orc.Solid2(pos, tris, p, thickness=1.0)
    m = s.node_masses()
    area = s.total_area()
    assert abs(area - 0.08 * 0.06) <= 1e-6
    assert abs(m.sum() - 1000.0 * area) <= 1e-4 * 1000.0 * area
    # each node gets a third of each incident triangle's area
    inc = np.bincount(tris, minlength=pos.size // 2)
    tri_area = area / (tris.size // 3)
    assert np.allclose(m, inc * tri_area / 3 * 1000.0, rtol=1e-5)


def fan_sums_geometric(pos, tris, area):
    """RebuildMass's weighted areas, Sum Rcp(3) * area in float32, with each
    vertex's triangles ordered by the angle of their centroids around it
    (counter-clockwise): an interior vertex's fan starts at its lowest-numbered
    triangle (which holds its lowest-numbered outgoing half-edge), a boundary
    vertex's just after the empty sector."""
    P = pos.reshape(-1, 2).astype(np.float64)
    T3 = tris.reshape(-1, 3)
    third = np.float32(1.0) / np.float32(3.0)
    inc = [[] for _ in range(len(P))]
    edges = {}
    for t, tri in enumerate(T3):
        for k in range(3):
            inc[tri[k]].append(t)
            e = tuple(sorted((tri[k], tri[(k + 1) % 3])))
            edges[e] = edges.get(e, 0) + 1
    wa = np.zeros(len(P), np.float32)
    for v, ts in enumerate(inc):
        if not ts:
            continue
        c = P[T3[ts]].mean(axis=1) - P[v]
        ang = np.arctan2(c[:, 1], c[:, 0])
        order = np.argsort(ang)
        boundary = any(n == 1 for e, n in edges.items() if v in e)
        if boundary:
            a = ang[order]
            gaps = np.diff(np.r_[a, a[0] + 2 * np.pi])
            start = (int(np.argmax(gaps)) + 1) % len(order)
        else:
            start = int(np.nonzero(np.array(ts)[order] == min(ts))[0][0])
        w = np.float32(0.0)
        for i in np.roll(order, -start):
            w = np.float32(w + np.float32(third * area[ts[i]]))
        wa[v] = w
    return wa


@pytest.mark.parametrize("counts,hole", [((9, 7), None), ((6, 6), (2, 3)), ((12, 5), (7, 2))])
def test_lumped_mass_fan_order(orc, counts, hole):
    """RebuildMass sums over iterator_polygons_around_vertex_ccw
    (LeafDSH_Solid2D_FEM.cpp:1842-1863, MeshSolidShape2.h:587-619): with
    random areas of several magnitudes, where any other order would round
    differently, the oracle's sums equal the geometric CCW fan sums bit for
    bit, on boxes and on boxes with a one-quad hole."""
    pos, tris = orc.mesh_box2((0.4, 0.3), counts)
    T3 = tris.reshape(-1, 3)
    if hole is not None:  # the two triangles of the quad at (i, j): quads are added row-major
        q = hole[0] * (counts[1] - 1) + hole[1]
        T3 = np.delete(T3, [2 * q, 2 * q + 1], axis=0)
    tris = np.ascontiguousarray(T3.ravel(), np.uint32)
    rng = np.random.default_rng(6)
    area = (10.0 ** rng.uniform(-3, 0, len(T3))).astype(np.float32)
    wa = np.zeros(pos.size // 2, np.float32)
    orc.lib().orc_fan_weighted_areas(pos.size // 2, len(T3), orc.up(tris), orc.fp(area), orc.fp(wa))
    expect = fan_sums_geometric(pos, tris, area)
    assert np.array_equal(wa, expect)
    # the order matters: summing in ascending triangle order differs somewhere
    asc = np.zeros_like(wa)
    for t, tri in enumerate(T3):
        for v in tri:
            asc[v] = np.float32(asc[v] + np.float32(np.float32(1.0) / np.float32(3.0) * area[t]))
    assert not np.array_equal(asc, expect)


def test_operators_against_dense_assembly(orc):
    pos, tris = orc.mesh_box2((0.03, 0.02), (7, 5))
    E, nu = 2e5, 0.35
    p = orc.default_params(young_modulus=E, poisson_ratio=nu, rm=orc.RM_DAPD)
    s = orc.Solid2(pos, tris, p)
    rng = np.random.default_rng(4)
    x = pos + rng.uniform(-2e-3, 2e-3, pos.size).astype(np.float32)
    R, det = s.begin_evaluation(x)
    assert (det > 0.1).all()
    # R is the polar factor: orthonormal with det +1
    Rm = R.reshape(-1, 2, 2).astype(np.float64)
    assert np.abs(np.einsum("eij,eik->ejk", Rm, Rm) - np.eye(2)).max() <= 1e-6
    Kg = dense_K2(pos, tris, R, E, nu)
    dx = rng.standard_normal(pos.size).astype(np.float32)
    df = s.df_x(dx)
    want = -Kg @ dx.astype(np.float64)
    assert np.linalg.norm(df - want) <= 1e-5 * np.linalg.norm(want)
    T = tris.reshape(-1, 3)
    P = pos.reshape(-1, 2)
    X = x.reshape(-1, 2)
    want = np.zeros(pos.size)
    for e in range(T.shape[0]):
        ke, _ = ke2_float64(P[T[e]], E, nu)
        Re = R[4 * e:4 * e + 4].reshape(2, 2).astype(np.float64)
        u = np.concatenate([Re.T @ X[n] - P[n] for n in T[e]])
        f = -np.kron(np.eye(3), Re) @ (ke @ u)
        for j, n in enumerate(T[e]):
            want[2 * n:2 * n + 2] += f[2 * j:2 * j + 2]
    fe = s.f_e(x)
    assert np.linalg.norm(fe - want) <= 2e-4 * np.linalg.norm(want)
    assert s.V(x) > 0


def test_step_against_dense_solve(orc):
    pos, tris = orc.mesh_box2((0.03, 0.03), (7, 7))
    knc = np.nonzero(np.isclose(pos[1::2], -0.03))[0].astype(np.uint32)
    p = orc.default_params(rm=orc.RM_DAPD, cg_rel_epsilon=1e-7, cg_abs_epsilon=1e-12, cg_max_iter=500,
                           reduce_fp64=1, gravity=(0.0, -9.81, 0.0))
    s = orc.Solid2(pos, tris, p, knc=knc)
    rng = np.random.default_rng(6)
    x0 = pos + rng.uniform(-1e-3, 1e-3, pos.size).astype(np.float32)
    x0.reshape(-1, 2)[knc] = pos.reshape(-1, 2)[knc]
    v0 = rng.uniform(-0.1, 0.1, pos.size).astype(np.float32)
    v0.reshape(-1, 2)[knc] = 0
    s.set_state(x0, v0)
    probe = orc.Solid2(pos, tris, p)
    R, _ = probe.begin_evaluation(x0)
    fe = probe.f_e(x0).astype(np.float64)
    Kg = dense_K2(pos, tris, R, 1e5, 0.3)
    m = np.repeat(s.node_masses().astype(np.float64), 2)
    dt = 1.0 / 60
    N = pos.size // 2
    g = np.tile([0.0, -9.81], N)
    b = dt * (m * g + fe) + m * v0
    A = np.diag(m) + dt * dt * Kg
    free = np.ones(N, bool)
    free[knc] = False
    fr = np.repeat(free, 2)
    y = np.zeros(2 * N)
    y[fr] = np.linalg.solve(A[np.ix_(fr, fr)], b[fr])
    st = s.fixed_step(dt)
    x1, v1 = s.state()
    assert st.cg_iters > 5
    assert np.abs(v1 - y).max() <= 1e-4 * np.abs(y).max()
    assert np.abs(v1.reshape(-1, 2)[knc]).max() == 0
    assert np.allclose(x1, x0 + np.float32(dt) * v1, atol=1e-7)


def test_free_fall(orc):
    pos, tris = orc.mesh_box2((0.02, 0.02), (5, 5))
    # lumped masses are not uniform, so b = M g dt is not a multiple of a
    # rigid mode and CG needs several iterations: solve tightly
    p = orc.default_params(rm=orc.RM_MSVD, gravity=(0.0, -9.81, 0.0), cg_rel_epsilon=1e-7,
                           cg_abs_epsilon=1e-12, cg_max_iter=500, reduce_fp64=1)
    s = orc.Solid2(pos, tris, p)
    st = s.fixed_step(1.0 / 60)
    x, v = s.state()
    dt = np.float32(1.0 / 60)
    # FP32 rounding leaves the rest state not exactly force-free (F = I + O(eps)),
    # and dt^2 K / M is about 560 here: v matches g dt to ~1e-5 relative
    assert np.allclose(v.reshape(-1, 2), [0, -9.81 * dt], atol=5e-6)
    assert np.allclose(x, pos + np.tile([0, -9.81 * dt * dt], pos.size // 2).astype(np.float32), atol=1e-7)
    assert st.num_degenerate == 0
