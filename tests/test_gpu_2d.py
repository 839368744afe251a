"""GPU: the 2D twin (workgroup-resident step kernel) against the oracle, and
batched against alone.

Tolerances: lumped masses bit-exact (same FP32 host arithmetic); the GPU
applies Ke in the equivalent Piola form, so steps agree to rounding
amplified by the CG: positions within 1e-4 relative after one step (5e-5 with
a converged CG), CG iterations within +-1. A simulation's result in a batch is
bit-identical to the same simulation alone (same kernel, its own group).
"""
import numpy as np
import pytest

from conftest import rel_node_err

pytestmark = pytest.mark.gpu

EDGE = 0.01


def grid(cf, n):
    half = (0.5 * EDGE * n, 0.5 * EDGE * n)
    pos, tris = cf.mesh_box2(half, (n + 1, n + 1))
    knc = np.nonzero(np.isclose(pos[1::2], -half[1]))[0].astype(np.uint32)  # bottom edge
    return pos, tris, knc


def params(cf, E=1e5, nu=0.3, max_iter=50, rel=1e-4):
    return cf.default_solid_params(fixed_dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), young_modulus=E,
                                   poisson_ratio=nu, damping_ratio=0.0, rm=cf.RM_DAPD, cg_rel_epsilon=rel,
                                   cg_abs_epsilon=1e-6, cg_max_iter=max_iter, density=1000.0)


def oracle_params(orc, E=1e5, nu=0.3, max_iter=50, rel=1e-4):
    return orc.default_params(young_modulus=E, poisson_ratio=nu, rm=orc.RM_DAPD, cg_rel_epsilon=rel,
                              cg_abs_epsilon=1e-6, cg_max_iter=max_iter, gravity=(0.0, -9.81, 0.0), reduce_fp64=1)


def test_lumped_masses_bitexact(gpu, orc):
    for n in (4, 13):
        pos, tris, knc = grid(gpu, n)
        b = gpu.Batch2()
        sid = b.add_sim(pos, tris, params(gpu), knc=knc)
        o = orc.Solid2(pos, tris, oracle_params(orc), knc=knc)
        assert np.array_equal(b.node_masses(sid), o.node_masses())
        b.close()


@pytest.mark.parametrize("n,E,nu", [(4, 1e5, 0.3), (12, 3e5, 0.25), (20, 5e4, 0.4), (30, 2e6, 0.35)])
def test_steps_vs_oracle(gpu, orc, n, E, nu):
    """n = 4, 12, 20, 30 use groups of 1, 2, 4 and 8 wavefronts."""
    pos, tris, knc = grid(gpu, n)
    s = gpu.LeafSolid2D(pos, tris, params(gpu, E, nu), knc=knc)
    o = orc.Solid2(pos, tris, oracle_params(orc, E, nu), knc=knc)
    for step in range(3):
        sg, so = s.fixed_step(), o.fixed_step()
        assert abs(sg.cg_iters - so.cg_iters) <= 1, (step, sg.cg_iters, so.cg_iters)
        xg, vg = s.state()
        xo, vo = o.state()
        err = rel_node_err(np.c_[xg.reshape(-1, 2), np.zeros(xg.size // 2)],
                           np.c_[xo.reshape(-1, 2), np.zeros(xo.size // 2)])
        assert err <= 1e-4 * (step + 1), (step, err)
        assert abs(sg.elastic_energy - so.elastic_energy) <= 1e-3 * abs(so.elastic_energy) + 1e-12
    assert np.abs(vg.reshape(-1, 2)[knc]).max() == 0


def test_converged_step_vs_oracle(gpu, orc):
    pos, tris, knc = grid(gpu, 10)
    s = gpu.LeafSolid2D(pos, tris, params(gpu, max_iter=1000, rel=1e-6), knc=knc)
    o = orc.Solid2(pos, tris, oracle_params(orc, max_iter=1000, rel=1e-6), knc=knc)
    sg, so = s.fixed_step(), o.fixed_step()
    assert sg.cg_reason in (gpu.CG_CONVERGED_REL, gpu.CG_CONVERGED_ABS)
    xg, _ = s.state()
    xo, _ = o.state()
    d = np.abs(xg - xo).max() / np.abs(xo - pos).max()
    assert d <= 5e-5, d


def test_batch_equals_alone_bitexact(gpu):
    rng = np.random.default_rng(3)
    sizes = [4, 9, 12, 15, 18, 22, 27, 32, 5, 16, 24, 31, 11, 6]
    cases = []
    for n in sizes:
        pos, tris, knc = grid(gpu, n)
        E = float(np.exp(rng.uniform(np.log(1e4), np.log(1e7))))
        nu = float(rng.uniform(0.2, 0.45))
        cases.append((pos, tris, knc, params(gpu, E, nu)))
    b = gpu.Batch2()
    ids = [b.add_sim(p, t, pr, knc=k) for p, t, k, pr in cases]
    info = b.info()
    assert info["sims"] == len(sizes)
    hist = [b.fixed_step() for _ in range(3)]
    for i, (p, t, k, pr) in enumerate(cases):
        s = gpu.LeafSolid2D(p, t, pr, knc=k)
        st = [s.fixed_step() for _ in range(3)]
        x, v = s.state()
        xb, vb = b.state(ids[i])
        assert np.array_equal(x, xb) and np.array_equal(v, vb), f"mesh {i} (n={sizes[i]}) differs in the batch"
        for j in range(3):
            assert (st[j].cg_iters, st[j].cg_reason, st[j].cg_rel_residual) == \
                   (hist[j][i].cg_iters, hist[j][i].cg_reason, hist[j][i].cg_rel_residual)
        s.close()
    b.close()


def test_too_large_mesh_is_refused(gpu):
    pos, tris, knc = grid(gpu, 40)
    with pytest.raises(gpu.CorofemError) as e:
        gpu.LeafSolid2D(pos, tris, params(gpu), knc=knc)
    assert "status -6" in str(e.value)
