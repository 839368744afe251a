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
