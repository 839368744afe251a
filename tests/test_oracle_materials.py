"""CPU: pins the oracle's LCM / LCMH material forces (C_LCM_f_e, C_LCM_df_x,
C_LCMH_df_x, C_LCM_V: ms/fem/TetrahedronElementSet3.cpp:1155-1178,1489-1508,
1838-1929) against what the mathematics fixes. That file does not compile
here (GSL), so there are no reference vectors for these:
- with the exact polar factor as rotation, S = R^T F is symmetric and the
  LCM force and energy equal the stored-Ke (WRP) ones, which the compiled
  reference pins (Ke, golden vectors);
- with R = Id, LCM's K*x equals WRP's on symmetric affine fields, and on an
  infinitesimal rotation (where WRP gives 0) its work is -sum V 2 mu |W|^2;
  the isotropic dilation pins lambda;
- LCMH blends by w = clamp01(tr|S - I| / (3k)): on a rotation field under a
  uniaxial stretch s it is w times LCM; k <= 0.1 is plain LCM, bit for bit;
- rigid translations are in every null space.
"""
import numpy as np
import pytest

E, NU = 1e5, 0.3
MU = 0.5 * E / (1 + NU)
LAM = E * NU / ((1 + NU) * (1 - 2 * NU))


def mesh(cf):
    pos, tets = cf.mesh_box((0.05, 0.05, 0.2), (4, 4, 6))
    return pos, tets


def tetset(orc, pos, tets, mm, k=1.0, x=None, R=None):
    ts = orc.TetSet(pos, tets, E, NU)
    ts.set_material(mm, k)
    ts.begin_evaluation(pos if x is None else x, rm=orc.RM_DAPD)
    if R is not None:
        ts.set_rotations(R)
    return ts


def exact_polar(pos, tets, x):
    P = pos.reshape(-1, 3).astype(np.float64)
    X = x.reshape(-1, 3).astype(np.float64)
    t = tets.reshape(-1, 4)
    out = []
    for e in range(len(t)):
        F = (X[t[e, 1:]] - X[t[e, 0]]).T @ np.linalg.inv((P[t[e, 1:]] - P[t[e, 0]]).T)
        U, _, Vt = np.linalg.svd(F)
        out.append((U @ Vt).ravel())
    return np.asarray(out, np.float32).ravel()


def identity_R(tets):
    return np.tile(np.eye(3, dtype=np.float32).ravel(), tets.size // 4)


def volumes(orc, pos, tets):
    ts = orc.TetSet(pos, tets, E, NU)
    return np.array([ts.element(e)[2] for e in range(tets.size // 4)], np.float64)


def test_lcm_force_and_energy_equal_wrp_with_exact_polar(orc, cf):
    pos, tets = mesh(cf)
    rng = np.random.default_rng(1)
    x = (pos + 0.002 * rng.normal(size=pos.size)).astype(np.float32)
    R = exact_polar(pos, tets, x)
    w = tetset(orc, pos, tets, orc.MM_C_WRP, x=x, R=R)
    fw, vw = w.f_e(x), w.V(x)
    for mm in (orc.MM_C_LCM, orc.MM_C_LCMH):
        l = tetset(orc, pos, tets, mm, x=x, R=R)
        fl, vl = l.f_e(x), l.V(x)
        assert np.abs(fl - fw).max() <= 2e-6 * np.abs(fw).max()
        assert abs(vl - vw) <= 1e-5 * abs(vw)


def test_lcm_stiffness_on_affine_fields(orc, cf):
    pos, tets = mesh(cf)
    R = identity_R(tets)
    lcm = tetset(orc, pos, tets, orc.MM_C_LCM, R=R)
    wrp = tetset(orc, pos, tets, orc.MM_C_WRP, R=R)
    X = pos.reshape(-1, 3).astype(np.float64)
    vol = volumes(orc, pos, tets).sum()
    A = np.array([[1.0, 0.3, -0.2], [0.3, -0.5, 0.1], [-0.2, 0.1, 0.4]]) * 1e-3  # symmetric
    dx = (X @ A.T).astype(np.float32).ravel()
    fl, fw = lcm.df_x(dx), wrp.df_x(dx)
    assert np.abs(fl - fw).max() <= 1e-5 * np.abs(fw).max()
    W = np.array([[0.0, 0.2, -0.1], [-0.2, 0.0, 0.3], [0.1, -0.3, 0.0]]) * 1e-3  # infinitesimal rotation
    dx = (X @ W.T).astype(np.float32).ravel()
    fl, fw = lcm.df_x(dx).astype(np.float64), wrp.df_x(dx).astype(np.float64)
    assert np.abs(fw).max() <= 1e-4 * np.abs(fl).max()
    work = dx.astype(np.float64) @ fl
    assert abs(work - (-vol * 2 * MU * np.sum(W * W))) <= 1e-4 * abs(work)
    a = 1e-3  # isotropic dilation: work = -V (2 mu + 3 lambda) 3 a^2
    dx = (a * X).astype(np.float32).ravel()
    work = dx.astype(np.float64) @ lcm.df_x(dx).astype(np.float64)
    assert abs(work - (-vol * (2 * MU + 3 * LAM) * 3 * a * a)) <= 1e-4 * abs(work)


def test_lcmh_blend_weight(orc, cf):
    pos, tets = mesh(cf)
    R = identity_R(tets)
    s = 0.3
    x = (pos.reshape(-1, 3) * np.array([1 + s, 1.0, 1.0])).astype(np.float32).ravel()  # F = diag(1+s, 1, 1)
    X = pos.reshape(-1, 3).astype(np.float64)
    W = np.array([[0.0, 0.2, -0.1], [-0.2, 0.0, 0.3], [0.1, -0.3, 0.0]]) * 1e-3
    dx = (X @ W.T).astype(np.float32).ravel()
    lcm = tetset(orc, pos, tets, orc.MM_C_LCM, x=x, R=R).df_x(dx).astype(np.float64)
    for k in (0.5, 2.0):
        w = min(1.0, s / (3 * k))
        h = tetset(orc, pos, tets, orc.MM_C_LCMH, k=k, x=x, R=R).df_x(dx).astype(np.float64)
        assert np.abs(h - w * lcm).max() <= 1e-4 * np.abs(lcm).max(), k
    plain = tetset(orc, pos, tets, orc.MM_C_LCMH, k=0.05, x=x, R=R).df_x(dx)
    assert np.array_equal(plain, lcm.astype(np.float32))


@pytest.mark.parametrize("mm", ["MM_C_LCM", "MM_C_LCMH"])
def test_translations_in_null_space(orc, cf, mm):
    pos, tets = mesh(cf)
    rng = np.random.default_rng(2)
    x = (pos + 0.003 * rng.normal(size=pos.size)).astype(np.float32)
    ts = tetset(orc, pos, tets, getattr(orc, mm), x=x)
    dx = np.tile(np.array([0.3, -0.2, 0.5], np.float32), pos.size // 3)
    f = ts.df_x(dx)
    scale = np.abs(ts.df_x(rng.normal(size=pos.size).astype(np.float32))).max()
    assert np.abs(f).max() <= 1e-5 * scale


def test_damping_force_identities(orc, cf):
    """C_f_d (TetrahedronElementSet3.cpp:1098-1150): with c0 = 0 it is c1 times
    the K*x of the velocities; with c1 = 0 and orthonormal rotations each node
    gets -c0 m_i v_i once per incident element."""
    pos, tets = mesh(cf)
    rng = np.random.default_rng(3)
    x = (pos + 0.002 * rng.normal(size=pos.size)).astype(np.float32)
    R = exact_polar(pos, tets, x)
    ts = tetset(orc, pos, tets, orc.MM_C_WRP, x=x, R=R)
    N = pos.size // 3
    inv_mass = rng.uniform(50.0, 200.0, N).astype(np.float32)
    v = rng.normal(size=pos.size).astype(np.float32)
    c1 = 0.02
    fd = ts.f_d(v, inv_mass, 0.0, c1).astype(np.float64)
    kv = ts.df_x(v).astype(np.float64)
    assert np.abs(fd - c1 * kv).max() <= 1e-6 * np.abs(c1 * kv).max()
    c0 = 0.7
    fd = ts.f_d(v, inv_mass, c0, 0.0).astype(np.float64).reshape(-1, 3)
    valence = np.bincount(tets.astype(np.int64), minlength=N)
    want = -c0 * (1.0 / inv_mass.astype(np.float64))[:, None] * valence[:, None] * v.reshape(-1, 3)
    assert np.abs(fd - want).max() <= 1e-5 * np.abs(want).max()
