"""GPU: the fused QIE-v step front (cf_set_fused_front) against the separate
passes it replaces: BeginEvaluation (TetrahedronElementSet3.cpp:219-275 with
UpdateDegeneration :477-532), the f_e tile pass of the right-hand side
(LeafDSH_Solid3D_FEM.cpp:1209-1295) and the K y tile pass of the CG start
(ILSS_CG.h:55-63).

Both run the same steps from the same state; positions, velocities, the CG
record (iterations, residuals, reason), the elastic energy and the
BeginEvaluation statistics (count of det F <= 0, min det F) must be equal bit
for bit. The sum of the degenerate elements' det F is reduced per tile by the
fused front and per block by BeginEvaluation, so only it is compared to a
relative 1e-6. Cases:
- the config-1 beam (z-min face held), with and without the warm start, at
  the library's default tiles and at 512 elements per 512-thread workgroup;
- the same beam under MSVD and with the constraint targets moved, so that
  the KNC rows take the clamped constraint velocity;
- the inverted box (a dozen tets start inverted): DAPD with the DoC state and
  its degenerate rotations, and MSVD;
- 128-element tiles on 128-thread workgroups.
"""
import numpy as np
import pytest

from conftest import beam_config1
from test_gpu_doc import E_BOX, inverted_box

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def run(cf, fused, pos, tets, p, knc, x0, v0, steps, tiles=None, targets=None):
    """Positions, velocities and per-step records after `steps` fixed steps."""
    if tiles:
        cf.set_tile_elems(tiles)
        cf.set_tile_threads(tiles)
    try:
        s = cf.LeafSolid3D(pos, tets, p, knc=knc, knc_targets=targets)
    finally:
        cf.set_tile_elems(0)
        cf.set_tile_threads(256)
    cf.set_fused_front(fused)
    try:
        s.set_state(x0, v0)
        hist = []
        for k in range(steps):
            print(f"[front] fused={fused} tiles={tiles} step {k}", flush=True)
            st = s.fixed_step()
            hist.append(st)
        x, v = s.state()
    finally:
        cf.set_fused_front(1)
        s.close()
    return x, v, hist


def same_runs(a, b):
    xa, va, ha = a
    xb, vb, hb = b
    assert np.array_equal(xa, xb)
    assert np.array_equal(va, vb)
    for k, (sa, sb) in enumerate(zip(ha, hb)):
        rec_a = (sa.cg_iters, sa.cg_reason, sa.cg_rel_residual, sa.cg_abs_residual, sa.elastic_energy,
                 sa.num_degenerate, sa.min_det_f)
        rec_b = (sb.cg_iters, sb.cg_reason, sb.cg_rel_residual, sb.cg_abs_residual, sb.elastic_energy,
                 sb.num_degenerate, sb.min_det_f)
        assert rec_a == rec_b, (k, rec_a, rec_b)
        assert abs(sa.sum_degenerate_det_f - sb.sum_degenerate_det_f) <= 1e-6 * abs(sb.sum_degenerate_det_f), k


def beam_params(cf, **kw):
    base = dict(fixed_dt=DT, gravity=(0.0, -9.81, 0.0), young_modulus=1e5, poisson_ratio=0.3, damping_ratio=0.0,
                rm=cf.RM_DAPD, cg_rel_epsilon=1e-4, cg_abs_epsilon=1e-6, cg_max_iter=200, density=1000.0)
    base.update(kw)
    return cf.default_solid_params(**base)


def bent_state(pos, seed=1):
    rng = np.random.default_rng(seed)
    x = pos.reshape(-1, 3).astype(np.float64)
    z = x[:, 2] + 0.2
    x[:, 1] += 0.04 * (z / 0.4) ** 2  # bent beam, fixed end at rest
    v = rng.normal(0.0, 0.05, x.shape) * (z / 0.4)[:, None]
    return x.astype(np.float32).reshape(-1), v.astype(np.float32).reshape(-1)


@pytest.mark.parametrize("warm", [1, 0])
@pytest.mark.parametrize("tiles", [None, 512])
def test_front_beam(gpu, warm, tiles):
    pos, tets, knc = beam_config1(gpu)
    x0, v0 = bent_state(pos)
    p = beam_params(gpu, warm_start=warm)
    a = run(gpu, 1, pos, tets, p, knc, x0, v0, 4, tiles)
    b = run(gpu, 0, pos, tets, p, knc, x0, v0, 4, tiles)
    assert a[2][0].cg_iters > 5
    same_runs(a, b)


def test_front_msvd_and_moved_targets(gpu):
    pos, tets, knc = beam_config1(gpu)
    x0, v0 = bent_state(pos, 2)
    targets = pos.reshape(-1, 3)[knc].copy()
    targets[: len(knc) // 2, 0] += 0.02  # half the held face pulled sideways
    targets[len(knc) // 2:, 1] -= 2.0   # the other half far away: speed clamped to 50
    p = beam_params(gpu, rm=gpu.RM_MSVD)
    a = run(gpu, 1, pos, tets, p, knc, x0, v0, 3, targets=targets.reshape(-1))
    b = run(gpu, 0, pos, tets, p, knc, x0, v0, 3, targets=targets.reshape(-1))
    same_runs(a, b)


@pytest.mark.parametrize("rm", ["DAPD", "MSVD"])
def test_front_inverted_box(gpu, rm):
    pos, tets, knc, x0 = inverted_box(gpu)
    p = gpu.default_solid_params(fixed_dt=DT, gravity=(0.0, -9.81, 0.0), young_modulus=E_BOX, poisson_ratio=0.3,
                                 damping_ratio=0.0, rm=getattr(gpu, "RM_" + rm), cg_rel_epsilon=1e-4,
                                 cg_abs_epsilon=1e-6, cg_max_iter=200, density=1000.0)
    v0 = np.zeros_like(x0)
    a = run(gpu, 1, pos, tets, p, knc, x0, v0, 3)
    b = run(gpu, 0, pos, tets, p, knc, x0, v0, 3)
    assert a[2][0].num_degenerate >= 6
    same_runs(a, b)


def test_front_small_tiles(gpu):
    pos, tets, knc = beam_config1(gpu)
    x0, v0 = bent_state(pos, 3)
    p = beam_params(gpu)
    a = run(gpu, 1, pos, tets, p, knc, x0, v0, 3, 128)
    b = run(gpu, 0, pos, tets, p, knc, x0, v0, 3, 128)
    same_runs(a, b)
