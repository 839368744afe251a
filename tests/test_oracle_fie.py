"""CPU: pins the oracle's FIE-dx integrator (Step_FullyImplicitEuler_dx,
ds/dsh/LeafDSH_Solid3D_FEM.cpp:1357-1552) and the multi-solid universe step
(SS_AggregateDSH_Basic::MultiDSH_Solid3D_Update, ds/ss/SS_AggregateDSH_Basic.cpp:229-300,
with the NR hooks of LeafDSH_Solid3D_FEM.cpp:2824-3039) against what the
mathematics fixes. These parts of the reference do not compile here (boost),
so there are no reference vectors for them:
- one Newton iteration of implicit Euler in position increments is the
  QIE-v step (same linear system scaled by 1/dt^2, y = dt v);
- converged Newton iterates solve the backward-Euler equations
  M (x1 - x0 - dt v0) / dt^2 = M g + f_e(x1) at the free nodes, and the
  quasi-static variant (dt = 0) solves f_e(x) + M g = 0; the floor of both
  residuals is the float32 rounding of f_e (about 5e-4 and 7e-3 of a node's
  weight here), far below the O(1) error of a wrong term or sign;
- a multi-solid pass after the first solves the QIE-like system at x with
  the rotations of the previous pass's end positions (NR_PreStep resets the
  iterate), checked against a dense float64 solve built from the oracle's
  own f_e and df_x at those rotations;
- the multi-solid step takes the smallest fixed dt, the largest Newton pass
  count and the strict '>' accumulator.
"""
import numpy as np
import pytest

DT = 1.0 / 60.0


def small_beam(cf):
    pos, tets = cf.mesh_box((0.05, 0.05, 0.2), (5, 5, 11))
    knc = np.nonzero(np.isclose(pos[2::3], -0.2))[0].astype(np.uint32)
    free = np.ones(pos.size // 3, bool)
    free[knc] = False
    return pos, tets, knc, free


def params(orc, **kw):
    base = dict(cg_rel_epsilon=1e-7, cg_abs_epsilon=1e-12, cg_max_iter=3000, reduce_fp64=1, rm=orc.RM_DAPD)
    base.update(kw)
    return orc.default_params(**base)


def f_e_at(orc, pos, tets, x, E=1e5, rot_x=None):
    ts = orc.TetSet(pos, tets, E, 0.3)
    ts.begin_evaluation(np.ascontiguousarray(x if rot_x is None else rot_x, np.float32), rm=orc.RM_DAPD)
    return ts, ts.f_e(np.ascontiguousarray(x, np.float32)).astype(np.float64)


def test_single_newton_iteration_is_qie_v(orc, cf):
    pos, tets, knc, _ = small_beam(cf)
    q = orc.Solid(pos, tets, params(orc), knc=knc)
    f = orc.Solid(pos, tets, params(orc, integrator=orc.IT_FIE_DX, nr_max_iter=1), knc=knc)
    for _ in range(3):
        sq, sf = q.fixed_step(DT), f.fixed_step(DT)
        assert sf.nr_iters == 1
    xq, vq = q.state()
    xf, vf = f.state()
    assert np.abs(xf - xq).max() <= 1e-5 * np.abs(xq - pos).max()
    assert np.abs(vf - vq).max() <= 1e-4 * np.abs(vq).max()


def test_converged_newton_solves_backward_euler(orc, cf):
    pos, tets, knc, free = small_beam(cf)
    s = orc.Solid(pos, tets, params(orc, integrator=orc.IT_FIE_DX, nr_max_iter=12, nr_abs_epsilon=1e-12), knc=knc)
    q = orc.Solid(pos, tets, params(orc), knc=knc)
    m = s.node_mass()
    g = np.tile([0.0, -9.81, 0.0], pos.size // 3)
    for _ in range(3):
        x0, v0 = s.state()
        st = s.fixed_step(DT)
        q.fixed_step(DT)
        assert st.nr_iters == 12
        x1, v1 = s.state()
        _, fe = f_e_at(orc, pos, tets, x1)
        r = m * (x1.astype(np.float64) - x0 - DT * v0) / DT ** 2 - m * g - fe
        assert np.abs(r.reshape(-1, 3)[free]).max() <= 2e-3 * m * 9.81
        assert np.abs(v1 - (x1 - x0) / DT).max() <= 1e-6 * np.abs(v1).max() + 1e-7
    # the nonlinear solve is not the linearised one
    assert np.abs(s.state()[0] - q.state()[0]).max() > 1e-3 * np.abs(q.state()[0] - pos).max()


def test_quasistatic_newton_reaches_equilibrium(orc, cf):
    pos, tets, knc, free = small_beam(cf)
    E = 1e6
    s = orc.Solid(pos, tets, params(orc, integrator=orc.IT_FIE_DX, young_modulus=E, nr_max_iter=20,
                                    nr_abs_epsilon=1e-12), knc=knc)
    st = s.fixed_step(0.0)
    x, v = s.state()
    m = s.node_mass()
    _, fe = f_e_at(orc, pos, tets, x, E=E)
    r = fe + m * np.tile([0.0, -9.81, 0.0], pos.size // 3)
    assert np.abs(r.reshape(-1, 3)[free]).max() <= 2e-2 * m * 9.81
    sag = (x - pos).reshape(-1, 3)[:, 1]
    assert sag.min() < -1e-2 and np.abs(sag[~free]).max() == 0.0
    assert st.nr_iters == 20 and np.all(v == 0.0)  # inv_dt = 0: v_k = 0


def test_multi_one_pass_matches_leaf_and_accumulator(orc, cf):
    pos, tets, knc, _ = small_beam(cf)
    leaf = orc.Solid(pos, tets, params(orc, integrator=orc.IT_FIE_DX, nr_max_iter=1), knc=knc)
    solo = orc.Solid(pos, tets, params(orc, nr_max_iter=1), knc=knc)
    multi = orc.Multi([solo])
    n, _ = multi.update(DT)
    assert n == 0  # acc = dt is not > fixed dt
    for _ in range(3):
        leaf.fixed_step(DT)
        n, st = multi.update(DT)
        assert n == 1 and st[0].nr_iters == 1
    xl, vl = leaf.state()
    xm, vm = solo.state()
    assert np.array_equal(xl, xm) and np.array_equal(vl, vm)


def test_multi_later_passes_use_lagged_rotations(orc, cf):
    """Pass 2 solves (M/dt^2 + K_R1) y = M g + f_e(x; R1) + M v / dt from x,
    with R1 the rotations at the end positions of pass 1."""
    pos, tets, knc, free = small_beam(cf)
    one = orc.Solid(pos, tets, params(orc, nr_max_iter=1), knc=knc)
    two = orc.Solid(pos, tets, params(orc, nr_max_iter=2), knc=knc)
    m1, m2 = orc.Multi([one]), orc.Multi([two])
    m1.update(DT)
    m2.update(DT)
    m1.update(DT)
    m2.update(DT)
    x1, _ = one.state()  # end of pass 1 from the rest state
    x2, _ = two.state()
    ts, fe = f_e_at(orc, pos, tets, pos, rot_x=x1)
    n3 = pos.size
    mass = two.node_mass()
    A = np.eye(n3) * mass / DT ** 2
    for j in range(n3):
        e = np.zeros(n3, np.float32)
        e[j] = 1.0
        A[:, j] -= ts.df_x(e)
    b = mass * np.tile([0.0, -9.81, 0.0], n3 // 3) + fe
    fm = np.repeat(free, 3)
    y = np.zeros(n3)
    y[fm] = np.linalg.solve(A[np.ix_(fm, fm)], b[fm])
    want = pos + y
    assert np.abs(x2 - want).max() <= 1e-4 * np.abs(want - pos).max()
    assert np.abs(x2 - x1).max() > 1e-3 * np.abs(x1 - pos).max()  # the second pass matters


def test_multi_takes_smallest_dt_and_most_passes(orc, cf):
    pos, tets, knc, _ = small_beam(cf)
    a = orc.Solid(pos, tets, params(orc, fixed_dt=1.0 / 60.0, nr_max_iter=2), knc=knc)
    b = orc.Solid(pos, tets, params(orc, fixed_dt=1.0 / 120.0, nr_max_iter=3), knc=knc)
    multi = orc.Multi([a, b])
    n, st = multi.update(1.0 / 60.0)  # acc 1/60 > 1/120 once
    assert n == 1 and [s.nr_iters for s in st] == [3, 3]
    # float32 accumulator with strict '>' (1/120 + 1/30 leaves a rounding excess: five steps)
    acc, fixed, want = np.float32(1.0 / 60.0) - np.float32(1.0 / 120.0), np.float32(1.0 / 120.0), 0
    acc = np.float32(acc + np.float32(1.0 / 30.0))
    while acc > fixed:
        acc = np.float32(acc - fixed)
        want += 1
    n, _ = multi.update(1.0 / 30.0)
    assert n == want
