"""GPU: the LCM and LCMH materials (C_LCM_f_e, C_LCM_df_x, C_LCMH_df_x,
C_LCM_V) through the C-ABI against the oracle, on the same rotations (DAPD
rotations are bit-exact on both sides) and the same evaluation F.

Tolerances: the GPU sums element forces per node in tile order instead of
element order, so f_e and the energy agree to 1e-5 and K*x to 1e-6 normwise
(as for C_WRP); a QIE-v step of a leaf with the scene-default LCMH material
agrees to 1e-4 of the step's displacement with a converging CG.
"""
import numpy as np
import pytest

from conftest import beam_config1

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mm,k", [("MM_C_LCM", 1.0), ("MM_C_LCMH", 1.0), ("MM_C_LCMH", 0.3), ("MM_C_LCMH", 0.05)])
def test_lcm_forces_vs_oracle(gpu, orc, mm, k):
    pos, tets, _ = beam_config1(gpu)
    N = pos.size // 3
    rng = np.random.default_rng(5)
    x = (pos + rng.standard_normal(pos.size).astype(np.float32) * 2e-3).astype(np.float32)
    o = orc.TetSet(pos, tets, 1e5, 0.3)
    o.set_material(getattr(orc, mm), k)
    s = gpu.TetrahedronElementSet3(pos, tets, gpu.default_fem_params(young_modulus=1e5, poisson_ratio=0.3,
                                                                    rm=gpu.RM_DAPD, mm=getattr(gpu, mm),
                                                                    ecie_k_factor=k))
    s.begin_evaluation(gpu.DeviceArray.from_host(x))
    Ro, _, _, _ = o.begin_evaluation(x, rm=orc.RM_DAPD)
    R, _ = s.rotations()
    assert np.array_equal(R.reshape(-1), Ro)
    f = gpu.DeviceArray.from_host(np.zeros(3 * N, np.float32))
    s.f_e(f)
    fg, fo = f.numpy(), o.f_e(x)
    assert np.linalg.norm(fg - fo) <= 1e-5 * np.linalg.norm(fo)
    v = rng.standard_normal(3 * N).astype(np.float32)
    df = gpu.DeviceArray.from_host(np.zeros(3 * N, np.float32))
    s.df_x(gpu.DeviceArray.from_host(v), df)
    dg, do = df.numpy(), o.df_x(v)
    assert np.linalg.norm(dg - do) <= 1e-6 * np.linalg.norm(do)
    assert abs(s.V() - o.V(x)) <= 1e-5 * abs(o.V(x))
    s.end_evaluation()
    s.close()


def test_lcmh_leaf_steps_vs_oracle(gpu, orc):
    pos, tets, knc = beam_config1(gpu)
    p = gpu.default_solid_params(fixed_dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), young_modulus=1e5,
                                 poisson_ratio=0.3, damping_ratio=0.0, rm=gpu.RM_DAPD, mm=gpu.MM_C_LCMH,
                                 cg_rel_epsilon=1e-6, cg_abs_epsilon=1e-12, cg_max_iter=2000, density=1000.0)
    s = gpu.LeafSolid3D(pos, tets, p, knc=knc)
    op = orc.default_params(young_modulus=1e5, poisson_ratio=0.3, rm=orc.RM_DAPD, cg_rel_epsilon=1e-6,
                            cg_abs_epsilon=1e-12, cg_max_iter=2000, gravity=(0.0, -9.81, 0.0), reduce_fp64=1,
                            mm=orc.MM_C_LCMH)
    o = orc.Solid(pos, tets, op, knc=knc)
    prev = pos
    for step in range(3):
        sg, so = s.fixed_step(), o.fixed_step()
        xg, _ = s.state()
        xo, _ = o.state()
        d = np.abs(xg - xo).max() / np.abs(xo - prev).max()
        assert d <= 1e-4, (step, d)
        assert abs(sg.elastic_energy - so.elastic_energy) <= 1e-3 * abs(so.elastic_energy)
        prev = xo
    s.close()


def test_batches_refuse_lcm(gpu):
    pos, tets, knc = beam_config1(gpu)
    b = gpu.Batch()
    mid = b.add_mesh(pos, tets)
    with pytest.raises(gpu.CorofemError) as e:
        b.add_sim(mid, gpu.default_solid_params(mm=gpu.MM_C_LCMH), knc=knc)
    assert "status -6" in str(e.value)
    b.close()


def test_damping_force_vs_oracle(gpu, orc):
    pos, tets, _ = beam_config1(gpu)
    N = pos.size // 3
    rng = np.random.default_rng(7)
    x = (pos + rng.standard_normal(pos.size).astype(np.float32) * 2e-3).astype(np.float32)
    v = rng.standard_normal(pos.size).astype(np.float32)
    inv_mass = rng.uniform(50.0, 200.0, N).astype(np.float32)
    fem = gpu.default_fem_params(young_modulus=1e5, poisson_ratio=0.3, rm=gpu.RM_DAPD, damping_ratio=0.5)
    s = gpu.TetrahedronElementSet3(pos, tets, fem, inv_mass=inv_mass)
    dxa, dva = gpu.DeviceArray.from_host(x), gpu.DeviceArray.from_host(v)
    s.begin_evaluation(dxa, v=dva)
    o = orc.TetSet(pos, tets, 1e5, 0.3)
    o.begin_evaluation(x, rm=orc.RM_DAPD)
    c = gpu.rayleigh_coefficients(0.1, 10000.0, 0.5)
    f = gpu.DeviceArray.from_host(np.zeros(3 * N, np.float32))
    s.f_d(f)
    fo = o.f_d(v, inv_mass, c[0], c[1])
    assert np.linalg.norm(f.numpy() - fo) <= 1e-5 * np.linalg.norm(fo)
    s.end_evaluation()
    s.close()
