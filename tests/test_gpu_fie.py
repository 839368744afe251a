"""GPU: the FIE-dx leaf integrator and the multi-solid universe step through
the C-ABI, against the oracle on the same inputs.

Tolerances: the GPU applies Ke in the Piola form and reduces in FP64, so with
a converging CG (rel 1e-6) positions agree to 1e-4 of the step's displacement
per step; Newton iteration counts are equal (they are fixed by nr_max_iter
here: the FP32 floor of |b| is above the default 1e-6 threshold); one Newton
iteration reproduces QIE-v to 1e-5. Batches and 2D solids refuse FIE-dx.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def beam(cf, counts=(5, 5, 11), half=(0.05, 0.05, 0.2)):
    pos, tets = cf.mesh_box(half, counts)
    knc = np.nonzero(np.isclose(pos[2::3], -half[2]))[0].astype(np.uint32)
    return pos, tets, knc


def gpu_params(cf, integrator=None, **kw):
    base = dict(fixed_dt=DT, gravity=(0.0, -9.81, 0.0), young_modulus=1e5, poisson_ratio=0.3, damping_ratio=0.0,
                rm=cf.RM_DAPD, cg_rel_epsilon=1e-6, cg_abs_epsilon=1e-12, cg_max_iter=2000, density=1000.0)
    nr = {k: kw.pop(k) for k in ("nr_max_iter", "nr_abs_epsilon") if k in kw}
    base.update(kw)
    p = cf.default_solid_params(**base)
    p.integrator = cf.IT_FIE_DX if integrator is None else integrator
    for k, v in nr.items():
        setattr(p, k, v)
    return p


def orc_params(orc, integrator=None, **kw):
    base = dict(young_modulus=1e5, poisson_ratio=0.3, rm=orc.RM_DAPD, cg_rel_epsilon=1e-6, cg_abs_epsilon=1e-12,
                cg_max_iter=2000, gravity=(0.0, -9.81, 0.0), reduce_fp64=1)
    base.update(kw)
    p = orc.default_params(**base)
    p.integrator = orc.IT_FIE_DX if integrator is None else integrator
    return p


def close(xg, xo, pos, tol):
    d = np.abs(xg - xo).max() / np.abs(xo - pos).max()
    assert d <= tol, d


@pytest.mark.parametrize("counts", [(5, 5, 11), (11, 11, 41)])
def test_fie_dx_steps_vs_oracle(gpu, orc, counts):
    pos, tets, knc = beam(gpu, counts)
    s = gpu.LeafSolid3D(pos, tets, gpu_params(gpu), knc=knc)
    o = orc.Solid(pos, tets, orc_params(orc), knc=knc)
    prev = pos
    for step in range(3):
        sg, so = s.fixed_step(), o.fixed_step()
        assert sg.nr_iters == so.nr_iters == 3, (step, sg.nr_iters, so.nr_iters)
        xg, _ = s.state()
        xo, _ = o.state()
        close(xg, xo, prev, 1e-4)
        prev = xo
        assert abs(sg.elastic_energy - so.elastic_energy) <= 1e-3 * abs(so.elastic_energy)
    s.close()


def test_fie_dx_one_iteration_is_qie_v(gpu):
    pos, tets, knc = beam(gpu)
    f = gpu.LeafSolid3D(pos, tets, gpu_params(gpu, nr_max_iter=1), knc=knc)
    q = gpu.LeafSolid3D(pos, tets, gpu_params(gpu, integrator=gpu.IT_QIE_V), knc=knc)
    for _ in range(3):
        assert f.fixed_step().nr_iters == 1
        q.fixed_step()
    close(f.state()[0], q.state()[0], pos, 1e-5)
    f.close()
    q.close()


def test_fie_dx_quasistatic_vs_oracle(gpu, orc):
    pos, tets, knc = beam(gpu)
    s = gpu.LeafSolid3D(pos, tets, gpu_params(gpu, young_modulus=1e6, nr_max_iter=20), knc=knc)
    o = orc.Solid(pos, tets, orc_params(orc, young_modulus=1e6, nr_max_iter=20), knc=knc)
    sg, so = s.fixed_step(0.0), o.fixed_step(0.0)
    assert sg.nr_iters == so.nr_iters == 20
    xg, vg = s.state()
    close(xg, o.state()[0], pos, 1e-3)
    assert np.all(vg == 0.0)
    s.close()


def test_multi_vs_oracle(gpu, orc):
    meshes = [beam(gpu), beam(gpu, (6, 4, 9), (0.05, 0.03, 0.16))]
    nr = [2, 3]
    dts = [1.0 / 60.0, 1.0 / 120.0]
    gs = [gpu.LeafSolid3D(p, t, gpu_params(gpu, integrator=gpu.IT_QIE_V, fixed_dt=d, nr_max_iter=k), knc=c)
          for (p, t, c), k, d in zip(meshes, nr, dts)]
    os_ = [orc.Solid(p, t, orc_params(orc, integrator=orc.IT_QIE_V, fixed_dt=d, nr_max_iter=k), knc=c)
           for (p, t, c), k, d in zip(meshes, nr, dts)]
    mg, mo = gpu.MultiSolid3D(gs), orc.Multi(os_)
    prev = [m[0] for m in meshes]
    for frame in range(3):
        ng, stg = mg.update(DT)
        no, sto = mo.update(DT)
        assert ng == no
        for i in range(2):
            assert stg[i].nr_iters == sto[i].nr_iters == 3
            xg, _ = gs[i].state()
            xo, _ = os_[i].state()
            close(xg, xo, prev[i], 1e-4)
            prev[i] = xo
    mg.close()
    for g in gs:
        g.close()


def test_multi_interleaved_equals_one_leaf_universes(gpu):
    """The universe step enqueues every leaf's CG on its own stream before it
    waits on any, as recorded graphs: three leaves of different sizes and
    iteration caps stepped together end bit-identical to each leaf stepped by a
    universe of its own, and to the same step with the kernels launched
    directly, also after the frame dt and the fixed nodes change."""
    meshes = [beam(gpu), beam(gpu, (6, 4, 9), (0.05, 0.03, 0.16)), beam(gpu, (3, 3, 12), (0.02, 0.02, 0.3))]
    params = [gpu_params(gpu, fixed_dt=1.0 / 60.0, nr_max_iter=3, cg_max_iter=m) for m in (200, 7, 40)]
    runs = []
    for together, graphs in ((True, True), (False, True), (True, False)):
        leaves = [gpu.LeafSolid3D(p, t, q, knc=c) for (p, t, c), q in zip(meshes, params)]
        multis = ([gpu.MultiSolid3D(leaves, use_graphs=graphs)] if together else
                  [gpu.MultiSolid3D([s], use_graphs=graphs) for s in leaves])
        for frame, dt in enumerate((DT, 2 * DT, DT, DT)):
            # a new frame dt (new fM, fK) and new fixed nodes: the recorded
            # chunks are re-recorded
            if frame == 3:
                for s, (p, _, c) in zip(leaves, meshes):
                    s.set_knc(c[: len(c) // 2])
            for m in multis:
                m.update(dt)
        runs.append([s.state() for s in leaves])
        for m in multis:
            m.close()
        for s in leaves:
            s.close()
    for other in runs[1:]:
        for (xa, va), (xb, vb) in zip(runs[0], other):
            assert np.array_equal(xa, xb) and np.array_equal(va, vb)


def test_batches_and_2d_refuse_fie_dx(gpu):
    pos, tets, knc = beam(gpu)
    b = gpu.Batch()
    mid = b.add_mesh(pos, tets)
    with pytest.raises(gpu.CorofemError) as e:
        b.add_sim(mid, gpu_params(gpu), knc=knc)
    assert "status -6" in str(e.value)
    b.close()
    pos2, tris = gpu.mesh_box2((0.05, 0.05), (5, 5))
    with pytest.raises(gpu.CorofemError) as e:
        gpu.LeafSolid2D(pos2, tris, gpu_params(gpu))
    assert "status -6" in str(e.value)
