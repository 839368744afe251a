"""GPU: config 1, the parity anchor, on the reference's own element order.

The 10x10x40-cube beam (11x11x41 vertices, 20,000 tets, 4,961 nodes), half
sizes (0.05, 0.05, 0.2) m, E = 1e5, nu = 0.3, rho = 1000, uniform node mass,
z-min face fixed, g = -9.81 y, dt = 1/60, QIE-v, DAPD rotations, CG rel 1e-4 /
abs 1e-6, 100 steps. The element order is the baked shape's
(Make_TetSolidShape3_Box then EndEdition, which sorts tetrahedra by layer,
Geo/shape/TetSolidShape3.cpp:786-801), the order LeafDSH_Solid3D_FEM::Create
adds them in (ds/dsh/LeafDSH_Solid3D_FEM.cpp:158-167).

- Reference-order mode (cf_solid_params.reference_order = 1): the GPU computes
  the reference's FP32 arithmetic in its order, so positions and velocities
  are bit-identical to the oracle (FP32 SSE-order reductions) after every one
  of the 100 steps, and CG iteration counts are equal on every step, with the
  CG capped at 50 as specified and with a converging CG (cap 500), undamped
  and at the Solid3D default damping.
- Fast path (FP64 reductions, Piola-form element forces): this problem moves
  by more than 1e-4 under rounding alone - the oracle changes by 1.2e-4 (cap
  50) and 2.1e-4 (cap 500) relative per node after 100 steps when only its
  reductions switch from FP32 to FP64 - and amplifies any per-step difference
  about a hundredfold over 100 steps. The fast path is held to 1e-4 after one
  step (measured 1.0e-6 and 2.2e-6 against the FP64-reduction oracle) and to
  about three times its measured deviation at 10 and 100 steps (measured:
  cap 50, 1.1e-4 and 1.0e-3; cap 500, 1.6e-4 and 2.6e-4); the printed values
  are recorded in the design notes.
"""
import numpy as np
import pytest

from conftest import rel_node_err

pytestmark = pytest.mark.gpu

HALF = (0.05, 0.05, 0.2)
COUNTS = (11, 11, 41)


def baked_beam(cf):
    sh = cf.mesh_box_baked(HALF, COUNTS)
    pos, tets = sh.pos.copy(), sh.tets.copy()
    knc = np.nonzero(np.isclose(pos[2::3], -HALF[2]))[0].astype(np.uint32)
    return pos, tets, knc


def pair(gpu, orc, max_iter, reference_order, zeta=0.0, reduce_fp64=0):
    pos, tets, knc = baked_beam(gpu)
    assert pos.size // 3 == 4961 and tets.size // 4 == 20000 and knc.size == 121
    p = gpu.default_solid_params(fixed_dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), young_modulus=1e5,
                                 poisson_ratio=0.3, damping_ratio=zeta, rm=gpu.RM_DAPD, cg_rel_epsilon=1e-4,
                                 cg_abs_epsilon=1e-6, cg_max_iter=max_iter, density=1000.0,
                                 reference_order=reference_order)
    g = gpu.LeafSolid3D(pos, tets, p, knc=knc)
    o = orc.Solid(pos, tets, orc.default_params(rm=orc.RM_DAPD, cg_max_iter=max_iter, damping_ratio=zeta,
                                                reduce_fp64=reduce_fp64), knc)
    return g, o, pos


@pytest.mark.slow
@pytest.mark.parametrize("max_iter,zeta,steps", [(50, 0.0, 100), (500, 0.0, 100), (50, 1.0, 40)])
def test_reference_order_bitexact(gpu, orc, max_iter, zeta, steps):
    g, o, _ = pair(gpu, orc, max_iter, 1, zeta)
    for k in range(1, steps + 1):
        a, b = g.fixed_step(), o.fixed_step()
        assert a.cg_iters == b.cg_iters, (k, a.cg_iters, b.cg_iters)
        xg, vg = g.state()
        xo, vo = o.state()
        if not (np.array_equal(xg, xo) and np.array_equal(vg, vo)):
            bad = np.nonzero(xg != xo)[0]
            pytest.fail(f"step {k}: {bad.size} position floats differ (first {bad[:5]}), "
                        f"max rel {rel_node_err(xg, xo):.3e}")
        assert a.num_degenerate == b.num_degenerate == 0
        assert np.float32(a.cg_rel_residual) == np.float32(b.cg_rel)
        assert np.float32(a.elastic_energy) == np.float32(b.elastic_energy)
    print(f"\nreference order, cap {max_iter}, zeta {zeta}: {steps} steps bit-identical, last CG {a.cg_iters} "
          f"iterations (rel {a.cg_rel_residual:.2e})")


@pytest.mark.slow
@pytest.mark.parametrize("max_iter,bound10,bound100", [(50, 3e-4, 3e-3), (500, 5e-4, 1e-3)])
def test_fast_path_config1_baked(gpu, orc, max_iter, bound10, bound100):
    """The fast path against the oracle with FP64 reductions (e64) and with the
    reference's FP32 reductions (e32); the oracle's own FP32-vs-FP64 change is
    the floor. Bounds: 1e-4 after one step; at 10 and 100 steps about three
    times the deviation measured on MI355X."""
    g, o64, pos = pair(gpu, orc, max_iter, 0, reduce_fp64=1)
    o32 = orc.Solid(pos, o64.tets, orc.default_params(rm=orc.RM_DAPD, cg_max_iter=max_iter), o64_knc(pos))
    out = {}
    for k in range(1, 101):
        a, b, c = g.fixed_step(), o64.fixed_step(), o32.fixed_step()
        if max_iter == 50:
            assert a.cg_iters == b.cg_iters == c.cg_iters == 50
        if k in (1, 10, 100):
            xg, x64, x32 = g.state()[0], o64.state()[0], o32.state()[0]
            out[k] = (rel_node_err(xg, x64), rel_node_err(xg, x32), rel_node_err(x64, x32))
    for k, (e64, e32, floor) in out.items():
        print(f"\ncap {max_iter} step {k}: GPU-vs-FP64 oracle {e64:.2e}, GPU-vs-FP32 oracle {e32:.2e}, "
              f"oracle FP32-vs-FP64 {floor:.2e}")
    assert out[1][0] <= 1e-4 and out[1][1] <= 1e-4
    assert out[10][0] <= bound10 and out[100][0] <= bound100


def o64_knc(pos):
    return np.nonzero(np.isclose(pos[2::3], -HALF[2]))[0].astype(np.uint32)
