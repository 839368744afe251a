"""CPU: pins the oracle's Rayleigh damping, the damped integrators, the
elastic energy C_V and the leaves' Step accumulator against what the
reference's own compiled math and the mathematics fix.

- Rayleigh coefficients c = inv(Omega) (2 zeta, 2 zeta) with
  Omega = [[1/f1, f1], [1/f2, f2]] (LeafDSH_Solid3D_FEM.h:347-354): bit-exact
  with Mal's Inverse and product compiled from the reference (oracle/_ref), and
  equal to the closed form c0 = 2 zeta f1 f2 / (f1 + f2), c1 = 2 zeta / (f1 + f2).
- A damped QIE-v step solves ((1 + dt c0) M + (dt^2 + dt c1) K) y = b
  (LeafDSH_Solid3D_FEM.cpp:1300-1301; 2D LeafDSH_Solid2D_FEM.cpp:1287-1403): a
  dense float64 solve; the frequencies f1 = 1, f2 = 100 make c0 a hundred
  times c1, so swapping the coefficients or dropping either term moves the
  answer by far more than the tolerance.
- A damped FIE-dx Newton iteration scales that system by 1/dt^2
  (factor_M = 1/dt^2 + c0/dt, factor_K = 1 + c1/dt, :1481-1482): one
  iteration is the damped QIE-v step.
- C_V (TetrahedronElementSet3.cpp:1456-1488) with R = Id is 1/2 u^T K u; on a
  linear displacement field every element carries the same strain, so the
  energy is the volume times the strain energy density: uniaxial
  1/2 (lambda + 2 mu) s^2, simple shear mu g^2 / 2 (engineering shear in E),
  dilation 3 a^2 (mu + 3 lambda / 2), with mu and lambda from E and nu
  (TetrahedronElementSet3.cpp:143-148).
- Step (LeafDSH_Solid3D_FEM.cpp:225-256, 2D :264-294): the total time grows by
  time_scale dt per call, fixed steps of time_scale fixed_dt run while the
  accumulator is strictly greater, and each equals a FixedStep of that length.
"""
import numpy as np
import pytest

from test_oracle import _dense_K

DT = 1.0 / 60.0


def closed_form_rayleigh(f1, f2, zeta):
    return 2 * zeta * f1 * f2 / (f1 + f2), 2 * zeta / (f1 + f2)


@pytest.mark.parametrize("f1,f2,zeta", [(0.1, 1e4, 1.0), (0.1, 1e4, 0.0), (0.1, 1e4, 0.35), (1.0, 100.0, 2.0),
                                        (2.5, 40.0, 0.7)])
def test_rayleigh_coefficients(orc, f1, f2, zeta):
    c = orc.rayleigh(f1, f2, zeta)
    ref = orc.ref()
    if ref is not None:
        want = np.zeros(2, np.float32)
        ref.ref_rayleigh(f1, f2, zeta, orc.fp(want))
        assert np.array_equal(c, want)
    c0, c1 = closed_form_rayleigh(f1, f2, zeta)
    assert abs(c[0] - c0) <= 1e-5 * max(abs(c0), 1e-30)
    assert abs(c[1] - c1) <= 1e-5 * max(abs(c1), 1e-30)


def test_rayleigh_default_zeta_one(orc):
    """The Solid3D default damping ratio 1 (bs/Solid3D.h:49) with the library
    frequencies 0.1 and 1e4 (ms/fem/Params.h:79)."""
    c = orc.rayleigh(0.1, 1e4, 1.0)
    assert abs(c[0] - 0.199998) <= 1e-6 and abs(c[1] - 1.99998e-4) <= 1e-9
    pos, tets = orc.mesh_box((0.01, 0.01, 0.01), (2, 2, 2))
    s = orc.Solid(pos, tets, orc.default_params(damping_ratio=1.0))
    assert np.array_equal(s.rayleigh(), c)


def _perturbed(pos, knc, dim, seed):
    rng = np.random.default_rng(seed)
    x0 = pos + rng.uniform(-1e-3, 1e-3, pos.size).astype(np.float32)
    x0.reshape(-1, dim)[knc] = pos.reshape(-1, dim)[knc]
    v0 = rng.uniform(-0.1, 0.1, pos.size).astype(np.float32)
    v0.reshape(-1, dim)[knc] = 0
    return x0, v0


@pytest.mark.parametrize("f1,f2,zeta", [(0.1, 1e4, 1.0), (1.0, 100.0, 2.0)])
def test_damped_qiev_step_against_dense_solve(orc, f1, f2, zeta):
    pos, tets = orc.mesh_box((0.02, 0.02, 0.03), (3, 3, 4))
    knc = np.nonzero(np.isclose(pos[2::3], -0.03))[0].astype(np.uint32)
    p = orc.default_params(rm=orc.RM_DAPD, cg_rel_epsilon=1e-7, cg_abs_epsilon=1e-12, cg_max_iter=500,
                           reduce_fp64=1, damping_ratio=zeta, rayleigh_freq1=f1, rayleigh_freq2=f2)
    s = orc.Solid(pos, tets, p, knc)
    x0, v0 = _perturbed(pos, knc, 3, 5)
    s.set_state(x0, v0)
    ts = orc.TetSet(pos, tets, 1e5, 0.3)
    R, _, _, _ = ts.begin_evaluation(x0, rm=orc.RM_DAPD)
    Kg = _dense_K(orc, pos, tets, R, 1e5, 0.3)
    m = float(s.node_mass())
    c0, c1 = closed_form_rayleigh(f1, f2, zeta)
    N = pos.size // 3
    fe = ts.f_e(x0).astype(np.float64)
    b = DT * (m * np.tile([0, -9.81, 0], N) + fe) + m * v0
    free = np.repeat(~np.isin(np.arange(N), knc), 3)

    def solve(fM, fK):
        A = fM * m * np.eye(3 * N) + fK * Kg
        y = np.zeros(3 * N)
        y[free] = np.linalg.solve(A[np.ix_(free, free)], b[free])
        return y

    y = solve(1 + DT * c0, DT * DT + DT * c1)
    st = s.fixed_step(DT)
    x1, v1 = s.state()
    assert st.cg_iters > 5
    assert np.abs(v1 - y).max() <= 1e-4 * np.abs(y).max()
    assert np.allclose(x1, x0 + np.float32(DT) * v1, atol=1e-7)
    # the damping terms matter at this tolerance: undamped or swapped coefficients miss it
    if zeta > 1.5:
        for wrong in (solve(1.0, DT * DT), solve(1 + DT * c1, DT * DT + DT * c0)):
            assert np.abs(wrong - y).max() > 100 * 1e-4 * np.abs(y).max()


def test_damped_qiev_2d_step_against_dense_solve(orc):
    from test_oracle2d import dense_K2
    f1, f2, zeta = 1.0, 100.0, 2.0
    pos, tris = orc.mesh_box2((0.03, 0.03), (7, 7))
    knc = np.nonzero(np.isclose(pos[1::2], -0.03))[0].astype(np.uint32)
    p = orc.default_params(rm=orc.RM_DAPD, cg_rel_epsilon=1e-7, cg_abs_epsilon=1e-12, cg_max_iter=500,
                           reduce_fp64=1, gravity=(0.0, -9.81, 0.0), damping_ratio=zeta, rayleigh_freq1=f1,
                           rayleigh_freq2=f2)
    s = orc.Solid2(pos, tris, p, knc=knc)
    x0, v0 = _perturbed(pos, knc, 2, 6)
    s.set_state(x0, v0)
    probe = orc.Solid2(pos, tris, p)
    R, _ = probe.begin_evaluation(x0)
    fe = probe.f_e(x0).astype(np.float64)
    Kg = dense_K2(pos, tris, R, 1e5, 0.3)
    m = np.repeat(s.node_masses().astype(np.float64), 2)
    c0, c1 = closed_form_rayleigh(f1, f2, zeta)
    N = pos.size // 2
    b = DT * (m * np.tile([0.0, -9.81], N) + fe) + m * v0
    A = (1 + DT * c0) * np.diag(m) + (DT * DT + DT * c1) * Kg
    free = np.repeat(~np.isin(np.arange(N), knc), 2)
    y = np.zeros(2 * N)
    y[free] = np.linalg.solve(A[np.ix_(free, free)], b[free])
    st = s.fixed_step(DT)
    x1, v1 = s.state()
    assert st.cg_iters > 5
    assert np.abs(v1 - y).max() <= 1e-4 * np.abs(y).max()
    A0 = np.diag(m) + DT * DT * Kg
    y0 = np.zeros(2 * N)
    y0[free] = np.linalg.solve(A0[np.ix_(free, free)], b[free])
    assert np.abs(y0 - y).max() > 100 * 1e-4 * np.abs(y).max()


def test_damped_fie_one_iteration_is_damped_qiev(orc):
    pos, tets = orc.mesh_box((0.05, 0.05, 0.2), (5, 5, 11))
    knc = np.nonzero(np.isclose(pos[2::3], -0.2))[0].astype(np.uint32)
    kw = dict(cg_rel_epsilon=1e-7, cg_abs_epsilon=1e-12, cg_max_iter=3000, reduce_fp64=1, rm=orc.RM_DAPD,
              damping_ratio=2.0, rayleigh_freq1=1.0, rayleigh_freq2=100.0)
    q = orc.Solid(pos, tets, orc.default_params(**kw), knc=knc)
    f = orc.Solid(pos, tets, orc.default_params(integrator=orc.IT_FIE_DX, nr_max_iter=1, **kw), knc=knc)
    u = orc.Solid(pos, tets, orc.default_params(**dict(kw, damping_ratio=0.0)), knc=knc)
    for _ in range(3):
        q.fixed_step(DT), f.fixed_step(DT), u.fixed_step(DT)
    xq, vq = q.state()
    xf, vf = f.state()
    xu, _ = u.state()
    assert np.abs(xf - xq).max() <= 1e-5 * np.abs(xq - pos).max()
    assert np.abs(vf - vq).max() <= 1e-4 * np.abs(vq).max()
    assert np.abs(xu - xq).max() > 1e-2 * np.abs(xq - pos).max()


def _jittered_box(orc, seed):
    pos, tets = orc.mesh_box((0.02, 0.015, 0.03), (4, 3, 5))
    rng = np.random.default_rng(seed)
    pos = (pos + rng.uniform(-1.5e-3, 1.5e-3, pos.size)).astype(np.float32)
    return pos, tets


@pytest.mark.parametrize("E,nu", [(1e5, 0.3), (3e4, 0.45)])
def test_energy_closed_form_linear_fields(orc, E, nu):
    pos, tets = _jittered_box(orc, 3)
    s = orc.TetSet(pos, tets, E, nu)
    Vtot = float(s.total_volume())
    T = tets.size // 4
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    X = pos.reshape(-1, 3).astype(np.float64)
    s.set_rotations(np.tile(np.eye(3, dtype=np.float32).ravel(), T))

    def energy(G):
        x = (X + X @ G.T).astype(np.float32).ravel()
        return float(s.V(x))

    a = 1e-3
    cases = [
        (np.diag([a, 0.0, 0.0]), 0.5 * (lam + 2 * mu) * a * a),              # uniaxial strain
        (np.array([[0.0, a, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]]), mu * a * a / 2),  # simple shear
        (a * np.eye(3), 3 * a * a * (mu + 1.5 * lam)),                        # dilation
    ]
    for G, w in cases:
        assert abs(energy(G) - Vtot * w) <= 2e-3 * Vtot * w, (G, energy(G), Vtot * w)
    # a factor-2 slip in the 1/2, in mu or in the engineering shear moves one of these by 2x
    assert abs(energy(cases[1][0]) / energy(cases[0][0]) - cases[1][1] / cases[0][1]) <= 5e-3 * cases[1][1] / cases[0][1]


def test_step_accumulator_time_scale(orc):
    """Step with time_scale 0.5 and several fixed steps per call: the same
    states as direct fixed steps of 0.5 fixed_dt, the strict '>' test, and the
    total time advanced once per call."""
    pos, tets = orc.mesh_box((0.02, 0.02, 0.04), (3, 3, 5))
    knc = np.nonzero(np.isclose(pos[2::3], -0.04))[0].astype(np.uint32)
    p = orc.default_params(fixed_dt=0.01, time_scale=0.5, rm=orc.RM_DAPD, cg_max_iter=30)
    a = orc.Solid(pos, tets, p, knc)
    b = orc.Solid(pos, tets, p, knc)
    dts = [0.01, 0.035, 0.0, 0.02, 0.004]
    acc = np.float32(0)
    total = np.float32(0)
    fixed = np.float32(0.5) * np.float32(0.01)
    for dt in dts:
        st = a.step(dt)
        sdt = np.float32(0.5) * np.float32(dt)
        total = np.float32(total + sdt)
        acc = np.float32(acc + sdt)
        n = 0
        while acc > fixed:
            b.fixed_step(float(fixed))
            acc = np.float32(acc - fixed)
            n += 1
        assert st.num_fixed_steps == n
        assert a.total_time() == total
        xa, va = a.state()
        xb, vb = b.state()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    # 0.01 * 0.5 accumulated twice from 0.01: exactly one fixed step, not two
    c = orc.Solid(pos, tets, orc.default_params(fixed_dt=0.01, time_scale=0.5), knc)
    assert c.step(0.01).num_fixed_steps == 0
    assert c.step(0.01).num_fixed_steps == 1


def test_step_accumulator_2d(orc):
    pos, tris = orc.mesh_box2((0.03, 0.03), (5, 5))
    knc = np.nonzero(np.isclose(pos[1::2], -0.03))[0].astype(np.uint32)
    p = orc.default_params(fixed_dt=0.01, time_scale=2.0, gravity=(0.0, -9.81, 0.0), cg_max_iter=40)
    a = orc.Solid2(pos, tris, p, knc=knc)
    b = orc.Solid2(pos, tris, p, knc=knc)
    st = a.step(0.025)          # 0.05 of scaled time: 2 fixed steps of 0.02 (0.05 > 0.04 > 0.02)
    assert st.num_fixed_steps == 2
    for _ in range(2):
        b.fixed_step(0.02)
    assert np.array_equal(a.state()[0], b.state()[0])
    assert a.total_time() == np.float32(np.float32(2.0) * np.float32(0.025))
    # the float32 accumulator decides the next call: 0.05 - 0.02 - 0.02 + 0.01
    acc = np.float32(np.float32(np.float32(2.0) * np.float32(0.025)) - np.float32(0.02))
    acc = np.float32(np.float32(acc - np.float32(0.02)) + np.float32(2.0) * np.float32(0.005))
    assert a.step(0.005).num_fixed_steps == int(acc > np.float32(2.0) * np.float32(0.01))
