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

from conftest import beam_config1

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
