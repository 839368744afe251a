"""CPU: pins the oracle's exact rotation differential (DM = Exact):
Compute_dR (ms/fem/Rotation.cpp:622-644, McAdams 2011) and the dR terms of
C_df_x / C_LCM_df_x (ms/fem/TetrahedronElementSet3.cpp:1804-1834,1862-1864).
Neither compiles here, so the checks are mathematical:
- dR is the directional derivative of the polar rotation: central
  differences of the float64 polar factor of F +- eps dF;
- with exact polar rotations, C_df_x with dR is the consistent tangent of
  C_f_e: central differences of the float64 corotational force
  -sum R K (R^T x - r) with R the exact polar factor at x +- eps dx;
- the truncated K*x (no dR) is not, so the dR terms are exercised.
"""
import numpy as np
import pytest


def polar(F):
    U, _, Vt = np.linalg.svd(F)
    R = U @ Vt
    if np.linalg.det(R) < 0:
        U[:, -1] *= -1
        R = U @ Vt
    return R


def test_dR_is_derivative_of_polar_rotation(orc):
    rng = np.random.default_rng(11)
    L = orc.lib()
    checked = 0
    for _ in range(500):
        F = np.eye(3) + rng.uniform(-0.4, 0.4, (3, 3))
        if np.linalg.det(F) < 0.3:
            continue
        dF = rng.uniform(-1, 1, (3, 3))
        R = polar(F)
        out = np.zeros(9, np.float32)
        ok = L.orc_dR(orc.fp(np.ascontiguousarray(F.astype(np.float32).ravel())),
                      orc.fp(np.ascontiguousarray(dF.astype(np.float32).ravel())),
                      orc.fp(np.ascontiguousarray(R.astype(np.float32).ravel())), orc.fp(out))
        assert ok
        eps = 1e-6
        fd = (polar(F + eps * dF) - polar(F - eps * dF)) / (2 * eps)
        assert np.abs(out.reshape(3, 3) - fd).max() <= 2e-4 * max(1.0, np.abs(fd).max())
        checked += 1
    assert checked > 300


def random_rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def test_dR_singular_branch(orc):
    """The branch |det G| <= 1e-4 -> no dR. With R given and S = R^T F =
    P diag(s) P^T symmetric, G = S - tr(S) I has the eigenvalues -(s_j + s_k),
    so det G = -(s0 + s1)(s1 + s2)(s0 + s2): F is built for a chosen det G
    (s = (a, b, -b + e), e solving (a + b) e (a - b + e) = -det G). Below the
    threshold the oracle reports no dR and leaves it zero; above it, dR is
    finite and R^T dR is skew (dR = R skew(.), a rotation's derivative)."""
    rng = np.random.default_rng(12)
    L = orc.lib()
    for _ in range(40):
        Q, P = random_rotation(rng), random_rotation(rng)
        a, b = rng.uniform(1.2, 2.0), rng.uniform(0.3, 0.6)
        dF = rng.uniform(-1, 1, (3, 3))
        for det_g, expect in ((0.0, 0), (5e-5, 0), (-5e-5, 0), (5e-4, 1), (-5e-4, 1), (1e-2, 1)):
            e = 0.5 * (-(a - b) + np.sqrt((a - b) ** 2 - 4 * det_g / (a + b)))
            sv = np.array([a, b, -b + e])
            assert abs(-(sv[0] + sv[1]) * (sv[1] + sv[2]) * (sv[0] + sv[2]) - det_g) < 1e-12
            F = Q @ P @ np.diag(sv) @ P.T
            out = np.full(9, np.nan, np.float32)
            ok = L.orc_dR(orc.fp(np.ascontiguousarray(F.astype(np.float32).ravel())),
                          orc.fp(np.ascontiguousarray(dF.astype(np.float32).ravel())),
                          orc.fp(np.ascontiguousarray(Q.astype(np.float32).ravel())), orc.fp(out))
            assert ok == expect, (det_g, ok)
            if not expect:
                assert np.all(out == 0.0)
                continue
            dR = out.reshape(3, 3).astype(np.float64)
            assert np.all(np.isfinite(dR)) and np.abs(dR).max() > 0
            K = Q.T @ dR
            assert np.abs(K + K.T).max() <= 1e-5 * np.abs(K).max()


@pytest.mark.parametrize("rm", ["RM_MSVD", "RM_DAPD"])
def test_exact_df_x_without_dR_branches(orc, rm):
    """The oracle's K*x with DM = Exact drops exactly the dR terms where the
    reference computes none: under MSVD on inverted elements with singular
    values (2, 1, 1) (G singular, see test_dR_singular_branch), under DAPD on
    elements with det F = 0.05 <= 0.1; there it equals the truncated K*x bit
    for bit, and on the regular elements beside them it does not."""
    from conftest import special_elements
    rng = np.random.default_rng(14)
    T = 256
    A, B, special = special_elements(rng, T, rm)
    tets = np.arange(4 * T, dtype=np.uint32)
    v = rng.standard_normal(A.size).astype(np.float32)
    o = orc.TetSet(A, tets, E, NU)
    o.begin_evaluation(B, rm=getattr(orc, rm))
    o.set_dm(orc.DM_EXACT)
    ex = o.df_x(v).reshape(T, 12)
    o.set_dm(orc.DM_TRUNCATED)
    tr = o.df_x(v).reshape(T, 12)
    assert np.array_equal(ex[special], tr[special])
    scale = np.abs(tr).max(axis=1)
    assert np.mean(np.abs(ex - tr)[~special].max(axis=1) > 1e-3 * scale[~special]) > 0.95


E, NU = 1e5, 0.3
MU = 0.5 * E / (1 + NU)
LAM = E * NU / ((1 + NU) * (1 - 2 * NU))


def f_e64(pos, tets, Ks, x, lcm=False):
    """float64 corotational force with the exact polar rotation: stored Ke
    (C_WRP) or the LCM Piola form P = R (2 mu (S - I) + lambda tr(S - I) I)."""
    P = pos.reshape(-1, 3).astype(np.float64)
    X = x.reshape(-1, 3).astype(np.float64)
    f = np.zeros_like(X)
    for e, t in enumerate(tets.reshape(-1, 4)):
        Dm = (P[t[1:]] - P[t[0]]).T
        F = (X[t[1:]] - X[t[0]]).T @ np.linalg.inv(Dm)
        R = polar(F)
        if lcm:
            SmI = R.T @ F - np.eye(3)
            Pk = R @ (2 * MU * SmI + LAM * np.trace(SmI) * np.eye(3))
            H = -(np.linalg.det(Dm) / 6.0) * Pk @ np.linalg.inv(Dm).T
            fe = np.vstack([-H.sum(1), H.T])
        else:
            u = (R.T @ X[t].T).T - P[t]
            fe = -(R @ (Ks[e] @ u.ravel()).reshape(4, 3).T).T
        f[t] += fe
    return f.ravel()


@pytest.mark.parametrize("mm", ["MM_C_WRP", "MM_C_LCM"])
def test_exact_df_x_is_consistent_tangent(orc, cf, mm):
    pos, tets = cf.mesh_box((0.05, 0.05, 0.05), (3, 3, 3))
    rng = np.random.default_rng(4)
    x = (pos + 0.004 * rng.normal(size=pos.size)).astype(np.float32)
    T = tets.size // 4
    ts = orc.TetSet(pos, tets, E, NU)
    lcm = mm != "MM_C_WRP"
    ts.set_material(getattr(orc, mm))
    Ks = np.array([ts.element(e)[0].astype(np.float64) for e in range(T)])
    ts.begin_evaluation(x, rm=orc.RM_MSVD)
    P = pos.reshape(-1, 3).astype(np.float64)
    X = x.reshape(-1, 3).astype(np.float64)
    Rex = np.array([polar((X[t[1:]] - X[t[0]]).T @ np.linalg.inv((P[t[1:]] - P[t[0]]).T)).ravel()
                    for t in tets.reshape(-1, 4)], np.float32).ravel()
    ts.set_rotations(Rex)
    dx = rng.normal(size=pos.size)
    eps = 1e-5
    tangent = (f_e64(pos, tets, Ks, x + eps * dx, lcm) - f_e64(pos, tets, Ks, x - eps * dx, lcm)) / (2 * eps)
    ts.set_dm(orc.DM_EXACT)
    exact = ts.df_x(dx.astype(np.float32)).astype(np.float64)
    ts.set_dm(orc.DM_TRUNCATED)
    trunc = ts.df_x(dx.astype(np.float32)).astype(np.float64)
    scale = np.abs(tangent).max()
    assert np.abs(exact - tangent).max() <= 1e-5 * scale
    assert np.abs(trunc - tangent).max() > 10 * np.abs(exact - tangent).max()
