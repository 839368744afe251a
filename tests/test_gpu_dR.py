"""GPU: K*x with the exact rotation differential (DM = Exact, Compute_dR of
ms/fem/Rotation.cpp:622-644 and the dR terms of C_df_x / C_LCM_df_x,
TetrahedronElementSet3.cpp:1804-1834,1862-1864) through the C-ABI against
the oracle on the same rotations and F. The GPU evaluates the dR terms in
Piola form (stored-Ke products on the oracle side): K*x agrees to 1e-5
normwise; a QIE-v step with exact differentials agrees to 1e-4 of its
displacement with a converging CG.
"""
import numpy as np
import pytest

from conftest import beam_config1, special_elements

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rm,mm", [("RM_DAPD", "MM_C_WRP"), ("RM_MSVD", "MM_C_WRP"), ("RM_DAPD", "MM_C_LCM"),
                                   ("RM_DAPD", "MM_C_LCMH")])
def test_exact_kx_vs_oracle(gpu, orc, rm, mm):
    pos, tets, _ = beam_config1(gpu)
    N = pos.size // 3
    rng = np.random.default_rng(9)
    x = (pos + rng.standard_normal(pos.size).astype(np.float32) * 3e-3).astype(np.float32)
    s = gpu.TetrahedronElementSet3(pos, tets, gpu.default_fem_params(young_modulus=1e5, poisson_ratio=0.3,
                                                                    rm=getattr(gpu, rm), mm=getattr(gpu, mm),
                                                                    dm=gpu.DM_EXACT))
    s.begin_evaluation(gpu.DeviceArray.from_host(x))
    o = orc.TetSet(pos, tets, 1e5, 0.3)
    o.set_material(getattr(orc, mm))
    o.set_dm(orc.DM_EXACT)
    o.begin_evaluation(x, rm=getattr(orc, rm))
    R, _ = s.rotations()
    o.set_rotations(R.reshape(-1))
    v = rng.standard_normal(3 * N).astype(np.float32)
    df = gpu.DeviceArray.from_host(np.zeros(3 * N, np.float32))
    s.df_x(gpu.DeviceArray.from_host(v), df)
    dg, do = df.numpy(), o.df_x(v)
    assert np.linalg.norm(dg - do) <= 1e-5 * np.linalg.norm(do)
    o.set_dm(orc.DM_TRUNCATED)
    assert np.linalg.norm(dg - o.df_x(v)) > 100 * np.linalg.norm(dg - do)  # the dR terms are there
    s.end_evaluation()
    s.close()


def test_exact_leaf_step_vs_oracle(gpu, orc):
    pos, tets, knc = beam_config1(gpu)
    p = gpu.default_solid_params(fixed_dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), young_modulus=1e5,
                                 poisson_ratio=0.3, damping_ratio=0.0, rm=gpu.RM_DAPD, dm=gpu.DM_EXACT,
                                 cg_rel_epsilon=1e-6, cg_abs_epsilon=1e-12, cg_max_iter=2000, density=1000.0)
    s = gpu.LeafSolid3D(pos, tets, p, knc=knc)
    op = orc.default_params(young_modulus=1e5, poisson_ratio=0.3, rm=orc.RM_DAPD, cg_rel_epsilon=1e-6,
                            cg_abs_epsilon=1e-12, cg_max_iter=2000, gravity=(0.0, -9.81, 0.0), reduce_fp64=1,
                            dm=orc.DM_EXACT)
    o = orc.Solid(pos, tets, op, knc=knc)
    prev = pos
    for step in range(3):
        s.fixed_step()
        o.fixed_step()
        xg, _ = s.state()
        xo, _ = o.state()
        d = np.abs(xg - xo).max() / np.abs(xo - prev).max()
        assert d <= 1e-4, (step, d)
        prev = xo
    s.close()



@pytest.mark.parametrize("rm", ["RM_MSVD", "RM_DAPD"])
def test_exact_kx_without_dR_branches(gpu, orc, rm):
    """Compute_dR's singular branch (|det G| <= 1e-4, Rotation.cpp:622-644)
    under MSVD and Compute_dR_From_RM's threshold switch under DAPD (no dR at
    det F <= 0.1, TetrahedronElementSet3.cpp:894-953): on those elements K*x
    with DM = Exact equals K*x with DM = Truncated, while on the regular
    elements beside them the dR terms are there; against the oracle on the
    GPU's rotations."""
    rng = np.random.default_rng(13)
    T = 512
    A, B, special = special_elements(rng, T, rm)
    tets = np.arange(4 * T, dtype=np.uint32)
    v = rng.standard_normal(A.size).astype(np.float32)
    out = {}
    for dm in ("DM_EXACT", "DM_TRUNCATED"):
        s = gpu.TetrahedronElementSet3(A, tets, gpu.default_fem_params(young_modulus=1e5, poisson_ratio=0.3,
                                                                        rm=getattr(gpu, rm), dm=getattr(gpu, dm)))
        s.begin_evaluation(gpu.DeviceArray.from_host(B))
        df = gpu.DeviceArray.from_host(np.zeros(A.size, np.float32))
        s.df_x(gpu.DeviceArray.from_host(v), df)
        out[dm] = (df.numpy().reshape(T, 12), s.rotations())
        s.end_evaluation()
        s.close()
    (ex, (R, det)), (tr, _) = out["DM_EXACT"], out["DM_TRUNCATED"]
    if rm == "RM_MSVD":
        Rs = R.reshape(T, 3, 3).astype(np.float64)
        Ds = (B.reshape(T, 4, 3)[:, 1:] - B.reshape(T, 4, 3)[:, :1]).transpose(0, 2, 1).astype(np.float64)
        Dm = (A.reshape(T, 4, 3)[:, 1:] - A.reshape(T, 4, 3)[:, :1]).transpose(0, 2, 1).astype(np.float64)
        S = Rs.transpose(0, 2, 1) @ Ds @ np.linalg.inv(Dm)
        G = S - np.trace(S, axis1=1, axis2=2)[:, None, None] * np.eye(3)
        assert np.abs(np.linalg.det(G[special])).max() < 1e-5
    else:
        assert np.all(det[special] < 0.06) and np.all(det[~special] > 0.2)
    scale = np.abs(tr).max(axis=1)
    assert np.all(np.abs(ex - tr)[special].max(axis=1) <= 1e-6 * scale[special])
    assert np.mean(np.abs(ex - tr)[~special].max(axis=1) > 1e-3 * scale[~special]) > 0.95
    o = orc.TetSet(A, tets, 1e5, 0.3)
    o.set_dm(orc.DM_EXACT)
    o.begin_evaluation(B, rm=getattr(orc, rm))
    o.set_rotations(R.reshape(-1))
    do = o.df_x(v).reshape(T, 12)
    assert np.abs(ex - do).max() <= 1e-5 * np.abs(do).max()
