"""GPU: the leaf boundary through the C-ABI against the oracle.

- Rayleigh damping (LeafDSH_Solid3D_FEM.h:347-354): the library's coefficients
  equal the oracle's bit for bit (both are pinned to the reference's compiled
  Mal arithmetic and the closed form in tests/test_oracle_damping.py); damped
  QIE-v and FIE-dx steps (3D) and QIE-v steps (2D) follow the oracle with a
  converging CG, at the Solid3D default zeta = 1 and at a strong setting
  (zeta = 2, f = (1, 100)) where undamped results are far away.
- Step(dt) (LeafDSH_Solid3D_FEM.cpp:225-256, 2D LeafDSH_Solid2D_FEM.cpp:264-294):
  strict '>' accumulator, time_scale, zero or several fixed steps per call,
  total time once per call; the number of fixed steps per call equals the
  oracle's and positions follow it.
- Incremental edition (BeginEdition / AddElement / EndEdition,
  TetrahedronElementSet3.cpp:2084-2166) builds the same set as the bulk call:
  precompute and operator results bit for bit.
- A non-finite state (NaN positions set from the host; NaN arithmetic only)
  returns CF_ERR_NAN from the 3D leaf (QIE-v and FIE-dx), a batch (the other
  simulations' results unchanged), the 2D leaf and the CG solver.

Tolerances: FP32 state with FP64 reductions and the Piola form of Ke on the
GPU; with CG rel 1e-6 positions agree to 1e-4 of the step's displacement (3D
leaf steps) or of the displacement from rest (2D, and the Step sequences,
whose steps move the solid by only a few hundred ulps).
"""
import numpy as np
import pytest

from conftest import rel_node_err

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
DAMPINGS = [(0.1, 1e4, 1.0), (1.0, 100.0, 2.0)]


def beam(cf, counts=(5, 5, 11), half=(0.05, 0.05, 0.2)):
    pos, tets = cf.mesh_box(half, counts)
    knc = np.nonzero(np.isclose(pos[2::3], -half[2]))[0].astype(np.uint32)
    return pos, tets, knc


def gpu_params(cf, f1=0.1, f2=1e4, zeta=0.0, integrator=None, **kw):
    base = dict(fixed_dt=DT, gravity=(0.0, -9.81, 0.0), young_modulus=1e5, poisson_ratio=0.3, damping_ratio=zeta,
                rayleigh_freq1=f1, rayleigh_freq2=f2, rm=cf.RM_DAPD, cg_rel_epsilon=1e-6, cg_abs_epsilon=1e-12,
                cg_max_iter=2000, density=1000.0)
    base.update(kw)
    p = cf.default_solid_params(**base)
    if integrator is not None:
        p.integrator = integrator
    return p


def orc_params(orc, f1=0.1, f2=1e4, zeta=0.0, integrator=None, **kw):
    base = dict(young_modulus=1e5, poisson_ratio=0.3, rm=orc.RM_DAPD, cg_rel_epsilon=1e-6, cg_abs_epsilon=1e-12,
                cg_max_iter=2000, gravity=(0.0, -9.81, 0.0), reduce_fp64=1, damping_ratio=zeta, rayleigh_freq1=f1,
                rayleigh_freq2=f2)
    base.update(kw)
    p = orc.default_params(**base)
    if integrator is not None:
        p.integrator = integrator
    return p


def displacement_err(xg, xo, x_prev):
    return np.abs(xg - xo).max() / max(np.abs(xo - x_prev).max(), 1e-30)


@pytest.mark.parametrize("f1,f2,zeta", [(0.1, 1e4, 0.0), (0.1, 1e4, 1.0), (1.0, 100.0, 2.0), (2.5, 40.0, 0.7)])
def test_rayleigh_coefficients_bitexact(gpu, orc, f1, f2, zeta):
    assert np.array_equal(np.float32(gpu.rayleigh_coefficients(f1, f2, zeta)), orc.rayleigh(f1, f2, zeta))
    pos, tets, knc = beam(gpu, (3, 3, 4))
    s = gpu.LeafSolid3D(pos, tets, gpu_params(gpu, f1, f2, zeta), knc=knc)
    o = orc.Solid(pos, tets, orc_params(orc, f1, f2, zeta), knc)
    assert np.array_equal(s.info()["rayleigh"], o.rayleigh())


@pytest.mark.parametrize("integrator", ["QIE_V", "FIE_DX"])
@pytest.mark.parametrize("f1,f2,zeta", DAMPINGS)
def test_damped_leaf_steps_vs_oracle(gpu, orc, integrator, f1, f2, zeta):
    pos, tets, knc = beam(gpu)
    gi = gpu.IT_QIE_V if integrator == "QIE_V" else gpu.IT_FIE_DX
    oi = orc.IT_QIE_V if integrator == "QIE_V" else orc.IT_FIE_DX
    s = gpu.LeafSolid3D(pos, tets, gpu_params(gpu, f1, f2, zeta, gi), knc=knc)
    o = orc.Solid(pos, tets, orc_params(orc, f1, f2, zeta, oi), knc)
    u = orc.Solid(pos, tets, orc_params(orc, f1, f2, 0.0, oi), knc)
    prev = pos
    for step in range(3):
        sg, so = s.fixed_step(), o.fixed_step()
        u.fixed_step()
        if integrator == "FIE_DX":
            assert sg.nr_iters == so.nr_iters
        xg, _ = s.state()
        xo, _ = o.state()
        assert displacement_err(xg, xo, prev) <= 1e-4, (step, displacement_err(xg, xo, prev))
        prev = xo
    assert abs(sg.elastic_energy - so.elastic_energy) <= 1e-3 * abs(so.elastic_energy)
    if zeta > 1.5:  # the strong damping is visible far above the tolerance
        assert np.abs(u.state()[0] - xo).max() > 1e-2 * np.abs(xo - pos).max()


@pytest.mark.parametrize("f1,f2,zeta", DAMPINGS)
def test_damped_2d_steps_vs_oracle(gpu, orc, f1, f2, zeta):
    n = 12
    half = (0.005 * n, 0.005 * n)
    pos, tris = gpu.mesh_box2(half, (n + 1, n + 1))
    knc = np.nonzero(np.isclose(pos[1::2], -half[1]))[0].astype(np.uint32)
    s = gpu.LeafSolid2D(pos, tris, gpu_params(gpu, f1, f2, zeta, cg_max_iter=1000), knc=knc)
    o = orc.Solid2(pos, tris, orc_params(orc, f1, f2, zeta, cg_max_iter=1000), knc=knc)
    for step in range(3):
        sg, so = s.fixed_step(), o.fixed_step()
        assert abs(sg.cg_iters - so.cg_iters) <= 1, (step, sg.cg_iters, so.cg_iters)
        xg, _ = s.state()
        xo, _ = o.state()
        # a 2D step moves this block by ~1e-5 m, a few hundred ulps of its
        # coordinates: the error is taken against the displacement from rest
        assert displacement_err(xg, xo, pos) <= 1e-4, (step, displacement_err(xg, xo, pos))


def test_solid_step_accumulator_vs_oracle(gpu, orc):
    """Step with time_scale 0.5: zero, one and several fixed steps per call, the
    strict '>' test on the float32 accumulator, and the total time advanced by
    time_scale * dt once per call."""
    pos, tets, knc = beam(gpu, (3, 3, 5), (0.02, 0.02, 0.04))
    kw = dict(fixed_dt=0.01, time_scale=0.5, cg_max_iter=60, cg_rel_epsilon=1e-6)
    s = gpu.LeafSolid3D(pos, tets, gpu_params(gpu, **kw), knc=knc)
    o = orc.Solid(pos, tets, orc_params(orc, fixed_dt=0.01, time_scale=0.5, cg_max_iter=60), knc)
    for dt in (0.01, 0.01, 0.035, 0.0, 0.02, 0.004, 0.013):
        sg, so = s.step(dt), o.step(dt)
        assert sg.num_fixed_steps == so.num_fixed_steps, (dt, sg.num_fixed_steps, so.num_fixed_steps)
        assert s.total_time() == o.total_time()
        xg, _ = s.state()
        xo, _ = o.state()
        if sg.num_fixed_steps:
            assert displacement_err(xg, xo, pos) <= 1e-4
    # the first 0.01 of scaled 0.005 does not exceed the fixed 0.005: no step
    t = gpu.LeafSolid3D(pos, tets, gpu_params(gpu, **kw), knc=knc)
    assert t.step(0.01).num_fixed_steps == 0 and t.step(0.01).num_fixed_steps == 1
    # a direct fixed step advances the total time by its own dt
    t.fixed_step(0.004)
    assert t.total_time() == np.float32(np.float32(0.005) + np.float32(0.005)) + np.float32(0.004)


def test_solid2_step_accumulator_vs_oracle(gpu, orc):
    n = 6
    half = (0.005 * n, 0.005 * n)
    pos, tris = gpu.mesh_box2(half, (n + 1, n + 1))
    knc = np.nonzero(np.isclose(pos[1::2], -half[1]))[0].astype(np.uint32)
    s = gpu.LeafSolid2D(pos, tris, gpu_params(gpu, fixed_dt=0.01, time_scale=2.0, cg_max_iter=200), knc=knc)
    o = orc.Solid2(pos, tris, orc_params(orc, fixed_dt=0.01, time_scale=2.0, cg_max_iter=200), knc=knc)
    for dt in (0.025, 0.005, 0.0, 0.004, 0.03):
        sg, so = s.step(dt), o.step(dt)
        assert sg.num_fixed_steps == so.num_fixed_steps, (dt, sg.num_fixed_steps, so.num_fixed_steps)
        xg, _ = s.state()
        xo, _ = o.state()
        if sg.num_fixed_steps:
            assert displacement_err(xg, xo, pos) <= 1e-4


def test_incremental_edition_equals_bulk(gpu):
    pos, tets, _ = beam(gpu, (6, 5, 9))
    p = gpu.default_fem_params(young_modulus=1e5, poisson_ratio=0.3, rm=gpu.RM_DAPD)
    a = gpu.TetrahedronElementSet3(pos, tets, p)
    b = gpu.TetrahedronElementSet3(pos, tets, p, incremental=True)
    da, va = a.elements()
    db, vb = b.elements()
    assert np.array_equal(da, db) and np.array_equal(va, vb)
    assert a.total_volume() == b.total_volume()
    rng = np.random.default_rng(1)
    x = (pos + rng.uniform(-2e-3, 2e-3, pos.size)).astype(np.float32)
    dx = rng.standard_normal(pos.size).astype(np.float32)
    out = []
    for s in (a, b):
        s.begin_evaluation(gpu.DeviceArray.from_host(x))
        f = gpu.DeviceArray.from_host(np.zeros(pos.size, np.float32))
        s.f_e(f)
        df = gpu.DeviceArray.from_host(np.zeros(pos.size, np.float32))
        s.df_x(gpu.DeviceArray.from_host(dx), df)
        out.append((s.rotations()[0], f.numpy(), df.numpy(), s.V()))
        s.end_evaluation()
    for u, v in zip(out[0], out[1]):
        assert np.array_equal(u, v)


def test_incremental_edition_errors(gpu):
    import ctypes as C
    pos, tets, _ = beam(gpu, (3, 3, 3))
    h = C.c_void_p()
    p = gpu.default_fem_params()
    lib = gpu.lib
    gpu.check(lib.cf_tetset_begin_edition(pos.size // 3, pos.ctypes.data_as(C.POINTER(C.c_float)), 2, C.byref(p),
                                          C.byref(h)))
    assert lib.cf_tetset_add_element(h, *tets[:4].tolist()) == 0
    assert lib.cf_tetset_end_edition(h) == gpu.ERR_INVALID_ARG       # 1 of 2 announced elements
    assert lib.cf_tetset_add_element(h, *tets[4:8].tolist()) == 0
    assert lib.cf_tetset_add_element(h, *tets[8:12].tolist()) == gpu.ERR_CAPACITY  # a third one
    assert lib.cf_tetset_end_edition(h) == 0
    lib.cf_tetset_destroy(h)


def _nan_state(pos, node):
    x = pos.copy()
    x[3 * node + 1] = np.nan
    return x, np.zeros_like(x)


@pytest.mark.parametrize("integrator", ["QIE_V", "FIE_DX"])
def test_nonfinite_state_reports_nan(gpu, integrator):
    pos, tets, knc = beam(gpu, (4, 4, 6), (0.02, 0.02, 0.04))
    gi = gpu.IT_QIE_V if integrator == "QIE_V" else gpu.IT_FIE_DX
    s = gpu.LeafSolid3D(pos, tets, gpu_params(gpu, integrator=gi, cg_max_iter=50), knc=knc)
    s.fixed_step()
    x, v = s.state()
    assert np.isfinite(x).all()
    s.set_state(*_nan_state(pos, 40))
    with pytest.raises(gpu.CorofemError) as e:
        s.fixed_step()
    assert e.value.status == gpu.ERR_NAN
    with pytest.raises(gpu.CorofemError) as e:
        s.step(DT * 1.5)
    assert e.value.status == gpu.ERR_NAN
    # a finite state steps normally again
    s.set_state(pos, np.zeros_like(pos))
    assert s.fixed_step().cg_reason != gpu.CG_NONFINITE


def test_nonfinite_batch_and_2d_and_cg(gpu):
    pos, tets, knc = beam(gpu, (4, 4, 6), (0.02, 0.02, 0.04))
    p = gpu_params(gpu, cg_max_iter=50, cg_rel_epsilon=1e-4)
    b = gpu.Batch()
    mid = b.add_mesh(pos, tets)
    ok, bad = b.add_sim(mid, p, knc=knc), b.add_sim(mid, p, knc=knc)
    b.set_state(bad, *_nan_state(pos, 40))
    with pytest.raises(gpu.CorofemError) as e:
        b.fixed_step()
    assert e.value.status == gpu.ERR_NAN
    alone = gpu.LeafSolid3D(pos, tets, p, knc=knc)
    alone.fixed_step()
    assert np.array_equal(b.state(ok, pos.size // 3)[0], alone.state()[0])
    b.close()

    n = 6
    half = (0.005 * n, 0.005 * n)
    pos2, tris = gpu.mesh_box2(half, (n + 1, n + 1))
    s2 = gpu.LeafSolid2D(pos2, tris, gpu_params(gpu, cg_max_iter=50))
    s2.fixed_step()
    x2 = pos2.copy()
    x2[5] = np.nan
    s2.set_state(x2, np.zeros_like(x2))
    with pytest.raises(gpu.CorofemError) as e:
        s2.fixed_step()
    assert e.value.status == gpu.ERR_NAN

    N = 8
    cg = gpu.ILSS_CG(N)

    def av(q_ptr, v_ptr):
        import ctypes as C
        gpu.lib.cf_memcpy_dtod(C.cast(q_ptr, C.c_void_p), C.cast(v_ptr, C.c_void_p), 4 * N)

    b_nan = np.ones(N, np.float32)
    b_nan[3] = np.nan
    x = gpu.DeviceArray.from_host(np.zeros(N, np.float32))
    with pytest.raises(gpu.CorofemError) as e:
        cg.solve(av, None, x, gpu.DeviceArray.from_host(b_nan), rel=1e-6, abs_eps=1e-12, max_iterations=20)
    assert e.value.status == gpu.ERR_NAN
    res = cg.solve(av, None, x, gpu.DeviceArray.from_host(np.ones(N, np.float32)), rel=1e-6, abs_eps=1e-12,
                   max_iterations=20)
    assert res.reason in (gpu.CG_CONVERGED_REL, gpu.CG_CONVERGED_ABS)


def test_leaf_graphs_equal_direct_launches(gpu):
    """The leaf CG's chunks replayed as recorded graphs (default) and launched
    directly give bit-identical steps, including after the tile-alternation
    setting changes (the recorded chunks are re-recorded) on a mesh with
    several CG chunks of each kind and solves that stop inside a chunk."""
    half, counts = (0.05, 0.05, 0.2), (11, 11, 41)
    pos, tets = gpu.mesh_box(half, counts)
    knc = np.nonzero(np.isclose(pos[2::3], -half[2]))[0].astype(np.uint32)
    p = gpu.default_solid_params(fixed_dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), young_modulus=1e5,
                                 poisson_ratio=0.3, damping_ratio=0.0, rm=gpu.RM_DAPD, cg_rel_epsilon=1e-4,
                                 cg_abs_epsilon=0.0, cg_max_iter=300, density=1000.0)
    a = gpu.LeafSolid3D(pos, tets, p, knc=knc)
    b = gpu.LeafSolid3D(pos, tets, p, knc=knc)
    b.set_graphs(False)
    gpu.set_tile_alternate(1)
    try:
        iters = []
        for k in range(6):
            if k == 3:
                gpu.set_tile_alternate(0)
            sa, sb = a.fixed_step(), b.fixed_step()
            iters.append(sa.cg_iters)
            assert sa.cg_iters == sb.cg_iters and sa.cg_reason == sb.cg_reason
            xa, va = a.state()
            xb, vb = b.state()
            assert np.array_equal(xa, xb) and np.array_equal(va, vb), k
        assert any(i % 8 for i in iters)  # some solve stops inside a chunk
        print(f"\nCG iterations {iters}")
    finally:
        gpu.set_tile_alternate(0)  # the default
        a.close()
        b.close()
