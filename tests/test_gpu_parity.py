"""GPU parity: the HIP path (through the C-ABI) against the oracle and the
reference's golden vectors.

Tolerances (FP32 state, FP64 reductions on the GPU):
  - element precompute (Dm^-1, V) and polar-decomposition rotations: bit-exact
  - PhysBAM-SVD rotations: 1e-5 absolute (the GPU's atan2f/cosf/sinf may differ
    from the host libm by an ulp)
  - f_e, K*x: 1e-5 relative (norm), the GPU evaluates Ke*u in the equivalent
    Piola form from Dm^-1 and V instead of stored Ke blocks
  - steps: CG iterations within +-1 per step; positions within 1e-4 relative
    after one step. The 100-step config-1 checks, on the reference's baked
    element order, are in tests/test_gpu_config1.py.
"""
import numpy as np
import pytest

from conftest import golden, rel_node_err, beam_config1

pytestmark = pytest.mark.gpu


def unit_tets_from_F(F):
    """One disjoint unit right tet per F (rest Dm = I), deformed corners x_j = F e_j,
    so that the kernel's F = Ds Dm^-1 equals F exactly."""
    n = F.shape[0]
    ref = np.tile(np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32), n)
    x = np.zeros((n, 4, 3), np.float32)
    Fm = F.reshape(n, 3, 3)
    x[:, 1] = Fm[:, :, 0]
    x[:, 2] = Fm[:, :, 1]
    x[:, 3] = Fm[:, :, 2]
    tets = np.arange(4 * n, dtype=np.uint32)
    return ref, tets, x.ravel()


def test_precompute_bitexact(gpu, orc):
    pos, tets, _ = beam_config1(gpu)
    s = gpu.TetrahedronElementSet3(pos, tets, gpu.default_fem_params(young_modulus=1e5, poisson_ratio=0.3))
    o = orc.TetSet(pos, tets, 1e5, 0.3)
    dinv, vol = s.elements()
    T = tets.size // 4
    od = np.stack([o.element(e)[1] for e in range(T)])
    ov = np.array([o.element(e)[2] for e in range(T)], np.float32)
    assert np.array_equal(dinv, od)
    assert np.array_equal(vol, ov)
    assert s.total_volume() == o.total_volume()


def test_rotations_against_reference_golden(gpu):
    g = golden("rotations.npz")
    ref, tets, x = unit_tets_from_F(g["F"])
    for rm, key in ((gpu.RM_DAPD, "R_pd"), (gpu.RM_MSVD, "R_svd")):
        s = gpu.TetrahedronElementSet3(ref, tets, gpu.default_fem_params(rm=rm))
        dx = gpu.DeviceArray.from_host(x)
        s.begin_evaluation(dx)
        R, det = s.rotations()
        assert np.array_equal(det, g["det"])
        R = R.reshape(-1, 9)
        want = g[key]
        ok = g["det"] > 0.1
        if rm == gpu.RM_DAPD:
            assert np.array_equal(R[ok], want[ok])
            assert np.array_equal(R[~ok], np.tile(np.eye(3, dtype=np.float32).ravel(), ((~ok).sum(), 1)))
        else:
            assert np.abs(R - want).max() <= 1e-5
        s.close()


def test_rotation_stress_config2_style(gpu, orc):
    """Per-tet F = Q diag(s), Q uniform on SO(3), s in [0.7,1.3]^3 (Philox seed 2):
    polar rotations bit-exact with the oracle, close to Q."""
    rng = np.random.Generator(np.random.Philox(2))
    n = 20000
    u = rng.random((n, 3))
    q = np.stack([np.sqrt(1 - u[:, 0]) * np.sin(2 * np.pi * u[:, 1]), np.sqrt(1 - u[:, 0]) * np.cos(2 * np.pi * u[:, 1]),
                  np.sqrt(u[:, 0]) * np.sin(2 * np.pi * u[:, 2]), np.sqrt(u[:, 0]) * np.cos(2 * np.pi * u[:, 2])], 1)
    x_, y_, z_, w_ = q.T
    Q = np.stack([1 - 2 * (y_ * y_ + z_ * z_), 2 * (x_ * y_ - z_ * w_), 2 * (x_ * z_ + y_ * w_),
                  2 * (x_ * y_ + z_ * w_), 1 - 2 * (x_ * x_ + z_ * z_), 2 * (y_ * z_ - x_ * w_),
                  2 * (x_ * z_ - y_ * w_), 2 * (y_ * z_ + x_ * w_), 1 - 2 * (x_ * x_ + y_ * y_)], 1).reshape(n, 3, 3)
    s = rng.uniform(0.7, 1.3, (n, 3))
    F = np.einsum("nij,nj->nij", Q, s).astype(np.float32).reshape(n, 9)
    ref, tets, x = unit_tets_from_F(F)
    es = gpu.TetrahedronElementSet3(ref, tets, gpu.default_fem_params(rm=gpu.RM_DAPD))
    es.begin_evaluation(gpu.DeviceArray.from_host(x))
    R, det = es.rotations()
    o = orc.TetSet(ref, tets, 1000.0, 0.25)
    Ro, deto, nd, _ = o.begin_evaluation(x, rm=orc.RM_DAPD)
    assert nd == 0 and (det > 0.1).all()
    assert np.array_equal(R.reshape(-1), Ro)
    assert np.abs(R - Q).max() < 0.06


def test_f_e_df_x_energy_vs_oracle(gpu, orc):
    pos, tets, _ = beam_config1(gpu)
    N = pos.size // 3
    rng = np.random.default_rng(0)
    x = pos + rng.standard_normal(pos.size).astype(np.float32) * 1e-3
    o = orc.TetSet(pos, tets, 1e5, 0.3)
    for rm in (gpu.RM_DAPD, gpu.RM_MSVD):
        s = gpu.TetrahedronElementSet3(pos, tets, gpu.default_fem_params(young_modulus=1e5, poisson_ratio=0.3, rm=rm))
        dx = gpu.DeviceArray.from_host(x)
        s.begin_evaluation(dx)
        Ro, _, _, _ = o.begin_evaluation(x, rm=rm)
        R, _ = s.rotations()
        if rm == gpu.RM_DAPD:
            assert np.array_equal(R.reshape(-1), Ro)
        o.set_rotations(R.reshape(-1))  # same rotations on both sides for the operator checks
        f = gpu.DeviceArray.from_host(np.zeros(3 * N, np.float32))
        s.f_e(f)
        fg, fo = f.numpy(), o.f_e(x)
        assert np.linalg.norm(fg - fo) <= 1e-5 * np.linalg.norm(fo)
        # accumulate semantics: a second call adds the same forces again
        s.f_e(f)
        assert np.linalg.norm(f.numpy() - 2 * fo) <= 2e-5 * np.linalg.norm(fo)
        v = rng.standard_normal(3 * N).astype(np.float32)
        df = gpu.DeviceArray.from_host(np.zeros(3 * N, np.float32))
        s.df_x(gpu.DeviceArray.from_host(v), df)
        dg, do = df.numpy(), o.df_x(v)
        assert np.linalg.norm(dg - do) <= 1e-6 * np.linalg.norm(do)
        assert abs(s.V() - o.V(x)) <= 1e-5 * abs(o.V(x))
        s.end_evaluation()
        s.close()


def test_kx_properties_config2_shard(gpu, orc):
    """64x64x32-cube shard of the config-2 mesh in a rotated+stretched pose:
    K x symmetric, rigid translations in the null space, matches the oracle."""
    pos, tets = gpu.mesh_box((0.32, 0.32, 0.16), (65, 65, 33))
    N = pos.size // 3
    rng = np.random.Generator(np.random.Philox(2))
    A = np.linalg.qr(rng.standard_normal((3, 3)))[0] @ np.diag(rng.uniform(0.7, 1.3, 3))
    x = (pos.reshape(-1, 3) @ A.T).astype(np.float32).ravel()
    x += rng.uniform(-1e-4, 1e-4, x.size).astype(np.float32)
    s = gpu.TetrahedronElementSet3(pos, tets, gpu.default_fem_params(young_modulus=1e5, poisson_ratio=0.3,
                                                                    rm=gpu.RM_DAPD))
    s.begin_evaluation(gpu.DeviceArray.from_host(x))
    R, _ = s.rotations()

    def K(v):
        out = gpu.DeviceArray.from_host(np.zeros(3 * N, np.float32))
        s.df_x(gpu.DeviceArray.from_host(v), out)
        return out.numpy().astype(np.float64)

    a = rng.standard_normal(3 * N).astype(np.float32)
    b = rng.standard_normal(3 * N).astype(np.float32)
    Ka, Kb = K(a), K(b)
    assert abs(a @ Kb - b @ Ka) <= 1e-5 * abs(a @ Ka)
    assert np.abs(K(np.tile(np.float32([1.0, -2.0, 0.5]), N))).max() == 0.0
    o = orc.TetSet(pos, tets, 1e5, 0.3)
    o.set_rotations(R.reshape(-1))
    want = o.df_x(a)
    assert np.linalg.norm(Ka - want) <= 1e-6 * np.linalg.norm(want)


def _run_pair(gpu, orc, rm, max_iter, steps, check_at):
    pos, tets, knc = beam_config1(gpu)
    p = gpu.default_solid_params(fixed_dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), young_modulus=1e5,
                                 poisson_ratio=0.3, damping_ratio=0.0, rm=rm, cg_rel_epsilon=1e-4,
                                 cg_abs_epsilon=1e-6, cg_max_iter=max_iter)
    g = gpu.LeafSolid3D(pos, tets, p, knc=knc)
    o64 = orc.Solid(pos, tets, orc.default_params(rm=rm, cg_max_iter=max_iter, reduce_fp64=1), knc)
    o32 = orc.Solid(pos, tets, orc.default_params(rm=rm, cg_max_iter=max_iter), knc)
    out = {}
    for k in range(1, steps + 1):
        a, b, c = g.fixed_step(), o64.fixed_step(), o32.fixed_step()
        assert abs(a.cg_iters - b.cg_iters) <= 1, (k, a.cg_iters, b.cg_iters)
        assert abs(a.cg_iters - c.cg_iters) <= 1, (k, a.cg_iters, c.cg_iters)
        assert a.num_degenerate == b.num_degenerate == 0
        if k in check_at:
            xg, vg = g.state()
            x64, _ = o64.state()
            x32, _ = o32.state()
            out[k] = (rel_node_err(xg, x64), rel_node_err(xg, x32), rel_node_err(x64, x32), a, b)
    return out


@pytest.mark.parametrize("rm", ["DAPD", "MSVD"])
def test_step_parity_converged(gpu, orc, rm):
    """CG allowed to converge (max 500): the solve reaches rel 1e-4 and the
    GPU follows the oracle step by step."""
    r = gpu.RM_DAPD if rm == "DAPD" else gpu.RM_MSVD
    res = _run_pair(gpu, orc, r, 500, 3, (1, 3))
    e64, e32, floor, a, b = res[1]
    assert e64 <= 1e-4 and e32 <= 1e-4
    # the solve ends on rel 1e-4, on the absolute floor 1e-6, or on the |d.q| <
    # abs^2 break that a nearly converged direction triggers (ILSS_CG.h:86-97)
    assert a.cg_reason in (gpu.CG_CONVERGED_REL, gpu.CG_CONVERGED_ABS, gpu.CG_BREAKDOWN)
    assert a.cg_rel_residual <= 2e-4
    assert abs(a.elastic_energy - b.elastic_energy) <= 1e-3 * abs(b.elastic_energy)
    assert res[3][0] <= 1e-3


def test_stage_inspection_matches_oracle(gpu, orc):
    pos, tets, knc = beam_config1(gpu)
    p = gpu.default_solid_params(fixed_dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), young_modulus=1e5,
                                 poisson_ratio=0.3, damping_ratio=0.0, rm=gpu.RM_DAPD)
    g = gpu.LeafSolid3D(pos, tets, p, knc=knc)
    m = g.info()["node_mass"]
    o = orc.Solid(pos, tets, orc.default_params(rm=orc.RM_DAPD), knc)
    assert m == o.node_mass()
    b, y0 = g.assemble_rhs()
    N = pos.size // 3
    want = (np.tile(np.float32([0, -9.81, 0]), N) * np.float32(m)) * np.float32(1 / 60)
    want.reshape(-1, 3)[knc] = 0
    assert np.array_equal(b, want.astype(np.float32)) or np.abs(b - want).max() <= 1e-12
    assert np.abs(y0).max() == 0
    ts = orc.TetSet(pos, tets, 1e5, 0.3)
    ts.begin_evaluation(pos, rm=orc.RM_DAPD)
    d = np.random.default_rng(2).standard_normal(3 * N).astype(np.float32)
    fK = np.float32(1 / 60) ** 2
    q = g.apply_A(d, 1.0, fK)
    qo = -fK * ts.df_x(d) + np.float32(m) * d
    qo.reshape(-1, 3)[knc] = 0
    assert np.linalg.norm(q - qo) <= 1e-6 * np.linalg.norm(qo)


def test_generic_cg_solver_against_numpy(gpu):
    """ILSS_CG through callbacks on an SPD system built from the element set:
    A v = v + c K v on a small free mesh, solved to 1e-6."""
    pos, tets = gpu.mesh_box((0.01, 0.01, 0.01), (3, 3, 3))
    N = pos.size // 3
    es = gpu.TetrahedronElementSet3(pos, tets, gpu.default_fem_params(young_modulus=1e3, poisson_ratio=0.3,
                                                                     rm=gpu.RM_DAPD))
    x0 = gpu.DeviceArray.from_host(pos)
    es.begin_evaluation(x0)
    # dense A from unit vectors, through df_x
    Kd = np.zeros((3 * N, 3 * N))
    for i in range(3 * N):
        e = np.zeros(3 * N, np.float32)
        e[i] = 1
        out = gpu.DeviceArray.from_host(np.zeros(3 * N, np.float32))
        es.df_x(gpu.DeviceArray.from_host(e), out)
        Kd[:, i] = -out.numpy()
    c = 1e-3
    A = np.eye(3 * N) + c * Kd
    cg = gpu.ILSS_CG(3 * N)
    tmp_in = gpu.DeviceArray(3 * N)
    tmp_out = gpu.DeviceArray(3 * N)
    import ctypes as C
    nbytes = 3 * N * 4

    def av(q_ptr, v_ptr):
        vals = np.empty(3 * N, np.float32)
        gpu.lib.cf_memcpy_dtoh(vals.ctypes.data_as(C.c_void_p), C.cast(v_ptr, C.c_void_p), nbytes)
        res = (A @ vals.astype(np.float64)).astype(np.float32)
        gpu.lib.cf_memcpy_htod(C.cast(q_ptr, C.c_void_p), res.ctypes.data_as(C.c_void_p), nbytes)

    b = np.random.default_rng(4).standard_normal(3 * N).astype(np.float32)
    x = gpu.DeviceArray.from_host(np.zeros(3 * N, np.float32))
    res = cg.solve(av, None, x, gpu.DeviceArray.from_host(b), rel=1e-6, abs_eps=1e-12, max_iterations=500)
    assert res.reason == gpu.CG_CONVERGED_REL
    want = np.linalg.solve(A, b.astype(np.float64))
    assert np.abs(x.numpy() - want).max() <= 1e-4 * np.abs(want).max()
    del tmp_in, tmp_out


def test_df_x_disjoint_tets_vs_oracle(gpu, orc):
    """Disjoint tets fill a tile's node budget (4 nodes per element): the tile
    builder must leave one thread to stage the end of the last slot run."""
    rng = np.random.default_rng(7)
    n = 1000
    F = np.eye(3, dtype=np.float32)[None] + 0.1 * rng.standard_normal((n, 3, 3)).astype(np.float32)
    ref, tets, x = unit_tets_from_F(F.reshape(n, 9))
    s = gpu.TetrahedronElementSet3(ref, tets, gpu.default_fem_params(young_modulus=1e5, poisson_ratio=0.3,
                                                                     rm=gpu.RM_DAPD))
    o = orc.TetSet(ref, tets, 1e5, 0.3)
    s.begin_evaluation(gpu.DeviceArray.from_host(x))
    R, _ = s.rotations()
    o.set_rotations(R.reshape(-1))
    v = rng.standard_normal(x.size).astype(np.float32)
    df = gpu.DeviceArray.from_host(np.zeros(x.size, np.float32))
    s.df_x(gpu.DeviceArray.from_host(v), df)
    dg, do = df.numpy(), o.df_x(v)
    assert np.linalg.norm(dg - do) <= 1e-6 * np.linalg.norm(do)
    s.close()
