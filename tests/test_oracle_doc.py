"""CPU: pins the oracle's degenerate-element branch of DAPD:
- gsl_poly_solve_cubic as the reference links it (GSL, no version pinned;
  TetrahedronElementSet3.cpp:36-69) and mal::GSolvePolynomialEq2's float
  fallback (Mal/GSolvePolynomialEq.h:11-38): against known roots;
- ComputeDoC (TetrahedronElementSet3.cpp:589-743): the time of collapse is
  the first t in (0,1] where det Ds(t) = thr * 6 V on the linear path (brute
  force in float64), and the crossing class is right on motions whose
  crossing is known by construction (a vertex through its opposite face, an
  edge through its opposite edge);
- UpdateDegeneration (TetrahedronElementSet3.cpp:477-532): the DoC state
  across evaluations;
- Compute_R_DAPD_Degenerate (Rotation.cpp:364-545): inverted elements are
  projected to det F = thr along the collapse axis (V-F and E-E), and R is
  continuous where the branches meet (det F = thr and w = 0).
The GSL library is not in this image, so nothing here calls it.
"""
import ctypes as C

import numpy as np
import pytest

THR = 0.1


def pd3(orc, F):
    """The oracle's Newton polar factor (bit-exact with the reference's)."""
    F = np.ascontiguousarray(F, np.float32).ravel()
    R = np.zeros(9, np.float32)
    it = np.zeros(1, np.uint32)
    orc.lib().orc_pd3(orc.fp(F), C.c_float(float(np.linalg.det(F.reshape(3, 3)))), orc.fp(R), orc.up(it))
    return R.reshape(3, 3)


# ------------------------------------------------------------------ cubic
def test_gsl_cubic_known_roots(orc):
    rng = np.random.default_rng(1)
    for _ in range(2000):
        r = np.sort(rng.uniform(-3, 3, 3))
        if np.diff(r).min() < 0.05:
            continue
        a, b, c = -r.sum(), r[0] * r[1] + r[0] * r[2] + r[1] * r[2], -r.prod()
        got = orc.gsl_cubic(a, b, c)
        assert got.size == 3
        assert np.abs(got - r).max() <= 1e-9 * (1 + np.abs(r).max())
    for _ in range(2000):  # one real root and a complex pair u +- iv
        x, u, v = rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(0.2, 3)
        q = u * u + v * v
        got = orc.gsl_cubic(-(x + 2 * u), 2 * u * x + q, -x * q)
        assert got.size == 1 and abs(got[0] - x) <= 1e-9 * (1 + abs(x))
    # exact double roots (caught by the integer-form discriminant) and a triple root
    assert np.array_equal(orc.gsl_cubic(0.0, -3.0, 2.0), [-2.0, 1.0, 1.0])     # (x-1)^2 (x+2)
    assert np.array_equal(orc.gsl_cubic(0.0, -3.0, -2.0), [-1.0, -1.0, 2.0])   # (x+1)^2 (x-2)
    assert np.array_equal(orc.gsl_cubic(-1.5, 0.75, -0.125), [0.5, 0.5, 0.5])  # (x-1/2)^3


def test_solve_eq3_float_paths(orc):
    # a != 0: the normalised cubic, roots to float
    n, r = orc.solve_eq3(2.0, -3.0, -3.0, 2.0)  # 2 (x+1)(x-2)(x-1/2)
    assert n == 3 and np.allclose(r, [-1.0, 0.5, 2.0], atol=1e-6)
    # a == 0: quadratic in float, ascending
    n, r = orc.solve_eq3(0.0, 2.0, -3.0, 1.0)
    assert n == 2 and np.allclose(r[:2], [0.5, 1.0], atol=1e-7)
    n, r = orc.solve_eq3(0.0, -1.0, 0.0, 4.0)  # -x^2 + 4: 1/(2a) < 0 swaps the roots
    assert n == 2 and np.allclose(r[:2], [-2.0, 2.0])
    n, _ = orc.solve_eq3(0.0, 1.0, 0.0, 1.0)   # x^2 + 1: no real root
    assert n == 0
    n, r = orc.solve_eq3(0.0, 1.0, -2.0, 1.0)  # (x-1)^2: one root
    assert n == 1 and r[0] == 1.0
    # IsZero uses an absolute epsilon of 1e-4: a leading 5e-5 is taken as zero
    n, r = orc.solve_eq3(0.0, 5e-5, 2.0, -1.0)
    assert n == 1 and r[0] == 0.5 and r[1] == 0.5
    n, _ = orc.solve_eq3(0.0, 0.0, 5e-5, 1.0)
    assert n == 0


# ------------------------------------------------------------------ ComputeDoC
def rest_tet(rng, scale=1.0):
    while True:
        a = rng.normal(size=(4, 3)) * scale
        Dm = (a[1:] - a[0]).T
        if np.linalg.det(Dm) > 0.3 * scale ** 3 and np.linalg.cond(Dm) < 6:
            return a


def det_path(a, b, t):
    p = a + t * (b - a)
    return np.linalg.det((p[1:] - p[0]).T)


def test_doc_time_of_collapse(orc):
    """toc is the first t in (0,1] with det Ds(t) = thr * 6 V (float64 brute force)."""
    rng = np.random.default_rng(2)
    found = 0
    for _ in range(600):
        a = rest_tet(rng)
        V = np.linalg.det((a[1:] - a[0]).T) / 6.0
        b = a + rng.normal(size=(4, 3)) * 1.2
        if det_path(a, b, 1.0) / (6 * V) > THR:  # must end degenerate
            continue
        a32, b32 = a.astype(np.float32), b.astype(np.float32)
        V32 = float(np.float32(np.linalg.det((a32[1:] - a32[0]).T.astype(np.float64)) / 6.0))
        doc, toc = orc.compute_doc(a32, b32, V32, THR, with_toc=True)
        # det Ds(t) - K on the float64 path, a cubic fitted exactly from 4 samples
        ts = np.array([0.0, 1 / 3, 2 / 3, 1.0])
        K = THR * 6 * V32
        coef = np.polyfit(ts, [det_path(a32.astype(np.float64), b32.astype(np.float64), t) - K for t in ts], 3)
        roots = np.roots(coef)
        real = np.sort(roots[np.abs(roots.imag) < 1e-9].real)
        first = real[(real > 0) & (real <= 1)]
        assert first.size > 0  # det goes from 6V > K to <= K
        if doc == orc.DOC_INVALID:
            continue
        found += 1
        assert abs(toc - first[0]) <= 2e-5 * max(1.0, 1.0 / abs(np.polyval(np.polyder(coef), first[0])))
    assert found > 150


def vf_motion(rng, i):
    """Vertex i moves through the centroid of its opposite face to the far side."""
    a = rest_tet(rng)
    others = [j for j in range(4) if j != i]
    c = a[others].mean(axis=0)
    b = a.copy()
    b[i] = c - 0.6 * (a[i] - c) + 0.05 * rng.normal(size=3)
    return a, b


def ee_motion(j):
    """Edge (0, j) moves through the opposite edge (k, l)."""
    k, l = [m for m in (1, 2, 3) if m != j]
    h = 0.8
    a = np.zeros((4, 3))
    a[0], a[j] = (0.0, 0.0, h), (1.0, 0.1, h)
    a[k], a[l] = (0.5, -0.6, 0.0), (0.45, 0.6, 0.0)
    if np.linalg.det((a[1:] - a[0]).T) < 0:
        a[[k, l]] = a[[l, k]]
    b = a.copy()
    b[0, 2] = b[j, 2] = -0.7
    return a, b


def test_doc_classifies_known_crossings(orc):
    """The reference's collapse normal is the right singular vector of Ds(toc)
    (coefficient space), not the physical normal of the flattened element, so
    the class follows the geometry only when the element is nearly flat at toc.
    At thr = 1e-3 a vertex through its opposite face must come out vertex-face
    on that vertex in at least 99% of cases (a wrong face index or pair fails
    a whole class), and an edge through its opposite edge edge-edge on that
    pair."""
    rng = np.random.default_rng(3)
    for thr, need in ((1e-3, 0.99), (1e-2, 0.95)):
        for i in range(4):
            agree = sum(orc.compute_doc(a, b, np.linalg.det((a[1:] - a[0]).T) / 6.0, thr) == 5 * i
                        for a, b in (vf_motion(rng, i) for _ in range(200)))
            assert agree >= need * 200, (thr, i, agree)
    for j in (1, 2, 3):
        a, b = ee_motion(j)
        V = np.linalg.det((a[1:] - a[0]).T) / 6.0
        for thr in (1e-3, 1e-2, THR):
            assert orc.compute_doc(a, b, V, thr) == j  # edge-edge (0, j)


def test_doc_state_across_evaluations(orc):
    """UpdateDegeneration: new DoC on an undegenerate -> degenerate step, kept
    while degenerate, cleared when undegenerate; without a DoC DAPD falls back."""
    rng = np.random.default_rng(4)
    a, b = vf_motion(rng, 2)
    tets = np.array([0, 1, 2, 3], np.uint32)
    ts = orc.TetSet(a.astype(np.float32).ravel(), tets, 1e5, 0.3)
    x0 = a.astype(np.float32).ravel()
    x1 = b.astype(np.float32).ravel()
    ts.begin_evaluation(x0, rm=orc.RM_DAPD, prev_x=x0)
    assert ts.doc()[0] == orc.DOC_INVALID
    ts.begin_evaluation(x1, rm=orc.RM_DAPD, prev_x=x0)
    assert ts.doc()[0] == 10 and ts.num_fallback() == 0
    x2 = (b + 0.02 * rng.normal(size=(4, 3))).astype(np.float32).ravel()
    ts.set_doc(np.array([5], np.uint8))  # a stale DoC is kept while prev_x is degenerate
    R, _, _, _ = ts.begin_evaluation(x2, rm=orc.RM_DAPD, prev_x=x1)
    assert ts.doc()[0] == 5 and ts.num_fallback() == 0
    ts.set_doc(np.array([orc.DOC_INVALID], np.uint8))
    R, _, _, _ = ts.begin_evaluation(x2, rm=orc.RM_DAPD, prev_x=x1)
    assert ts.doc()[0] == orc.DOC_INVALID and ts.num_fallback() == 1
    assert np.array_equal(R.reshape(3, 3), np.eye(3, dtype=np.float32))
    ts.begin_evaluation(x0, rm=orc.RM_DAPD, prev_x=x2)
    assert ts.doc()[0] == orc.DOC_INVALID
    # the state is tracked under every rotation method
    ts.begin_evaluation(x1, rm=orc.RM_MSVD, prev_x=x0)
    assert ts.doc()[0] == 10


# ------------------------------------------------------------------ R_DAPD_Degenerate
def projected_vf(p, i, V, thr=THR):
    """Float64: vertex i moved along its opposite face's normal to height
    thr 6V / |n| on the side of positive det, its foot point kept."""
    idx = [(i + 1) % 4, (i + 2) % 4, (i + 3) % 4]
    n = np.cross(p[idx[1]] - p[idx[0]], p[idx[2]] - p[idx[0]])
    ln = np.linalg.norm(n)
    nh = n / ln
    dh = thr * 6 * V / ln
    foot = p[i] - np.dot(p[i] - p[idx[0]], nh) * nh
    out = []
    for sgn in (1.0, -1.0):
        q = p.copy()
        q[i] = foot + sgn * dh * nh
        out.append(q)
    return max(out, key=lambda q: np.linalg.det((q[1:] - q[0]).T))


def projected_ee(p, j, V, thr=THR):
    k, l = [m for m in (1, 2, 3) if m != j]
    n = np.cross(p[j] - p[0], p[l] - p[k])
    ln = np.linalg.norm(n)
    nh = n / ln
    dh = thr * 6 * V / ln
    d = np.dot(p[0] - p[k], nh)
    out = []
    for sgn in (1.0, -1.0):
        q = p.copy()
        q[0] += (sgn * dh - d) * nh
        q[j] += (sgn * dh - d) * nh
        out.append(q)
    return max(out, key=lambda q: np.linalg.det((q[1:] - q[0]).T))


def test_dapd_degenerate_projects_inverted_to_threshold(orc):
    rng = np.random.default_rng(5)
    cases = 0
    for i in range(4):
        for _ in range(20):
            a, b = vf_motion(rng, i)
            a32, b32 = a.astype(np.float32), b.astype(np.float32)
            Dm = (a32[1:] - a32[0]).T.astype(np.float64)
            V = float(np.float32(np.linalg.det(Dm) / 6.0))
            Dinv = np.linalg.inv(Dm).astype(np.float32)
            F = (b32[1:] - b32[0]).T.astype(np.float64) @ Dinv.astype(np.float64)
            detF = float(np.linalg.det(F))
            assert detF < 0
            ok, R = orc.r_dapd_degenerate(detF, THR, 5 * i, b32, V, Dinv)
            assert ok
            q = projected_vf(b32.astype(np.float64), i, V)
            projF = (q[1:] - q[0]).T @ Dinv.astype(np.float64)
            assert abs(np.linalg.det(projF) - THR) < 1e-4
            assert np.abs(R - pd3(orc, projF)).max() < 2e-5
            cases += 1
    for j in (1, 2, 3):
        a, b = ee_motion(j)
        b[0, 2] = b[j, 2] = -0.3
        a32, b32 = a.astype(np.float32), b.astype(np.float32)
        Dm = (a32[1:] - a32[0]).T.astype(np.float64)
        V = float(np.float32(np.linalg.det(Dm) / 6.0))
        Dinv = np.linalg.inv(Dm).astype(np.float32)
        F = (b32[1:] - b32[0]).T.astype(np.float64) @ Dinv.astype(np.float64)
        assert np.linalg.det(F) < 0
        ok, R = orc.r_dapd_degenerate(float(np.linalg.det(F)), THR, j, b32, V, Dinv)
        assert ok
        q = projected_ee(b32.astype(np.float64), j, V)
        projF = (q[1:] - q[0]).T @ Dinv.astype(np.float64)
        assert abs(np.linalg.det(projF) - THR) < 1e-4
        assert np.abs(R - pd3(orc, projF)).max() < 2e-5
        cases += 1
    assert cases == 83


def test_dapd_degenerate_continuity(orc):
    """R is continuous in the vertex height w: at det F = thr it meets the
    undegenerate branch (polar of F) and at w = 0 the two lambda_u branches meet."""
    rng = np.random.default_rng(6)
    for i in range(4):
        a = rest_tet(rng)
        Dm = (a[1:] - a[0]).T
        V = np.linalg.det(Dm) / 6.0
        Dinv = np.linalg.inv(Dm).astype(np.float32)
        idx = [(i + 1) % 4, (i + 2) % 4, (i + 3) % 4]
        n = np.cross(a[idx[1]] - a[idx[0]], a[idx[2]] - a[idx[0]])
        nh = n / np.linalg.norm(n)
        h0 = np.dot(a[i] - a[idx[0]], nh)  # signed height of the rest pose
        tilt = 0.3 * rng.normal(size=3)
        tilt -= np.dot(tilt, nh) * nh

        def pose(frac):  # vertex i at frac * (rest height), shifted sideways
            p = a.copy()
            p[i] = a[i] - (1 - frac) * h0 * nh + tilt
            return p.astype(np.float32)

        def rot(frac):
            p = pose(frac)
            F = (p[1:] - p[0]).T.astype(np.float64) @ Dinv.astype(np.float64)
            d = float(np.linalg.det(F))
            if d > THR:
                return pd3(orc, F)
            ok, R = orc.r_dapd_degenerate(d, THR, 5 * i, p, np.float32(V), Dinv)
            assert ok
            return R

        # det F is linear in frac: det F = frac (the tilt keeps the height)
        for edge in (THR, 0.0):
            lo, hi = rot(edge - 1e-4), rot(edge + 1e-4)
            assert np.abs(lo - hi).max() < 2e-3, (i, edge)
        # no jumps along the whole degenerate range; the exactly flat pose is
        # left out: there w and det F round to opposite signs and the axis flip
        # (Rotation.cpp:399-403) gives a reflection, a measure-zero case
        fr = np.linspace(-0.5, 0.2, 141) + 0.0025
        Rs = np.array([rot(f) for f in fr])
        assert np.abs(np.diff(Rs, axis=0)).max() < 0.05


# ------------------------------------------------------------------ 2D (TriangleElementSet2 / Rotation.cpp:300-362)
def rest_tri(rng):
    while True:
        a = rng.normal(size=(3, 2))
        D = (a[1:] - a[0]).T
        if np.linalg.det(D) > 0.4 and np.linalg.cond(D) < 5:
            return a


def pd2(F):
    """Exact 2x2 polar factor in float64."""
    U, _, Vt = np.linalg.svd(F)
    R = U @ Vt
    if np.linalg.det(R) < 0:
        U[:, -1] *= -1
        R = U @ Vt
    return R


def test_noc2_time_and_node(orc):
    """ComputeNoC: toc is the first positive root of det Ds(t) = thr 2 A on the
    linear path (brute force in float64) and a vertex moved through its
    opposite edge is the node of collapse (its opposite edge is the longest
    once the triangle is nearly flat)."""
    rng = np.random.default_rng(7)
    for i in range(3):
        for _ in range(100):
            a = rest_tri(rng)
            j, k = [m for m in range(3) if m != i]
            c = 0.5 * (a[j] + a[k]) + rng.uniform(-0.2, 0.2) * (a[k] - a[j])
            b = a.copy()
            b[i] = c - rng.uniform(0.05, 0.8) * (a[i] - c)
            A = np.linalg.det((a[1:] - a[0]).T) / 2
            a32, b32 = a.astype(np.float32), b.astype(np.float32)
            noc, toc = orc.compute_noc2(a32, b32, np.float32(A), THR, with_toc=True)
            assert noc == i
            aa, bb = a32.astype(np.float64), b32.astype(np.float64)
            ts = np.array([0.0, 0.5, 1.0])
            vals = [np.linalg.det(((aa + t * (bb - aa))[1:] - (aa + t * (bb - aa))[0]).T) - THR * 2 * np.float32(A)
                    for t in ts]
            r = np.sort(np.roots(np.polyfit(ts, vals, 2)).real)
            first = r[r > 0][0]
            assert abs(toc - first) <= 1e-5 * (1 + abs(first))


def test_dapd_degenerate2_projects_inverted_to_threshold(orc):
    rng = np.random.default_rng(8)
    cases = 0
    for i in range(3):
        for _ in range(50):
            a = rest_tri(rng)
            j, k = [m for m in range(3) if m != i]
            c = 0.5 * (a[j] + a[k])
            b = a.copy()
            b[i] = c - rng.uniform(0.1, 0.9) * (a[i] - c) + 0.05 * rng.normal(size=2)
            Dm = (a[1:] - a[0]).T
            A = np.float32(np.linalg.det(Dm) / 2)
            Dinv = np.linalg.inv(Dm).astype(np.float32)
            b32 = b.astype(np.float32)
            if np.linalg.det((b32[1:] - b32[0]).T.astype(np.float64)) >= 0:
                continue  # not inverted: the Hermite branch, not a projection to thr
            cases += 1
            R = orc.r_dapd_degenerate2(THR, i, b32, A, Dinv)
            # float64: the node moved perpendicular to its opposite edge to height
            # thr 2 A / |edge| on the triangle's side, its foot point kept
            q = b32.astype(np.float64)
            e = q[(i + 2) % 3] - q[(i + 1) % 3]
            n = np.array([-e[1], e[0]]) / np.linalg.norm(e)
            h = THR * 2 * A / np.linalg.norm(e)
            q[i] = q[i] - np.dot(q[i] - q[(i + 1) % 3], n) * n + h * n
            projF = (q[1:] - q[0]).T @ Dinv.astype(np.float64)
            assert abs(np.linalg.det(projF) - THR) < 1e-4
            assert np.abs(R - pd2(projF)).max() < 2e-5
    assert cases > 130


def test_dapd_degenerate2_continuity(orc):
    rng = np.random.default_rng(9)
    for i in range(3):
        a = rest_tri(rng)
        Dm = (a[1:] - a[0]).T
        A = np.float32(np.linalg.det(Dm) / 2)
        Dinv = np.linalg.inv(Dm).astype(np.float32)
        j, k = (i + 1) % 3, (i + 2) % 3
        e = a[k] - a[j]
        n = np.array([-e[1], e[0]]) / np.linalg.norm(e)
        h0 = np.dot(a[i] - a[j], n)
        slide = 0.3 * rng.normal() * e / np.linalg.norm(e)

        def rot(frac):
            p = a.copy()
            p[i] = a[i] - (1 - frac) * h0 * n + slide
            p = p.astype(np.float32)
            F = (p[1:] - p[0]).T.astype(np.float64) @ Dinv.astype(np.float64)
            if np.linalg.det(F) > THR:
                return pd2(F)
            return orc.r_dapd_degenerate2(THR, i, p, A, Dinv)

        for edge in (THR, 0.0):
            assert np.abs(rot(edge - 1e-4) - rot(edge + 1e-4)).max() < 2e-3, (i, edge)
        Rs = np.array([rot(f) for f in np.linspace(-0.5, 0.2, 141) + 0.0025])
        assert np.abs(np.diff(Rs, axis=0)).max() < 0.05
