"""GPU: cf_cg_solve_static, the CG solve that allocates nothing (its work
buffers are the library's static device scratch), against cf_cg_solve on the
same systems through the C-ABI: solutions and results bit-identical, on two
calls in a row with the second system smaller than the first (the scratch
keeps the first call's contents), with and without a constraint callback; a
system beyond the scratch's capacity is refused before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def spd_system(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n))
    A = M @ M.T / n + np.eye(n)
    b = rng.standard_normal(n).astype(np.float32)
    return A, b


def callbacks(gpu, A, fixed):
    n = A.shape[0]
    nbytes = 4 * n

    def av(q_ptr, v_ptr):  # q = A v on host copies of the device vectors
        v = np.empty(n, np.float32)
        gpu.lib.cf_memcpy_dtoh(v.ctypes.data_as(C.c_void_p), C.cast(v_ptr, C.c_void_p), nbytes)
        q = (A @ v.astype(np.float64)).astype(np.float32)
        gpu.lib.cf_memcpy_htod(C.cast(q_ptr, C.c_void_p), q.ctypes.data_as(C.c_void_p), nbytes)

    def cons(v_ptr):  # S: zero the fixed rows
        v = np.empty(n, np.float32)
        gpu.lib.cf_memcpy_dtoh(v.ctypes.data_as(C.c_void_p), C.cast(v_ptr, C.c_void_p), nbytes)
        v[fixed] = 0.0
        gpu.lib.cf_memcpy_htod(C.cast(v_ptr, C.c_void_p), v.ctypes.data_as(C.c_void_p), nbytes)

    return av, (cons if fixed is not None else None)


@pytest.mark.parametrize("constrained", [False, True])
def test_static_scratch_equals_handle(gpu, constrained):
    for n, seed in ((700, 1), (130, 2)):  # the second call is smaller than the first
        A, b = spd_system(n, seed)
        fixed = np.arange(0, n, 7) if constrained else None
        av, cons = callbacks(gpu, A, fixed)
        kw = dict(rel=1e-6, abs_eps=1e-10, max_iterations=400)
        x_h = gpu.DeviceArray.from_host(np.zeros(n, np.float32))
        cg = gpu.ILSS_CG(n)
        r_h = cg.solve(av, cons, x_h, gpu.DeviceArray.from_host(b), **kw)
        cg.close()
        x_s = gpu.DeviceArray.from_host(np.zeros(n, np.float32))
        r_s = gpu.cg_solve_static(av, cons, x_s, gpu.DeviceArray.from_host(b), **kw)
        xs, xh = x_s.numpy(), x_h.numpy()
        assert np.array_equal(xs, xh), (n, np.abs(xs - xh).max())
        assert (r_s.iterations, r_s.reason) == (r_h.iterations, r_h.reason)
        assert r_s.rel_residual == r_h.rel_residual and r_s.abs_residual == r_h.abs_residual
        assert r_s.reason in (gpu.CG_CONVERGED_REL, gpu.CG_CONVERGED_ABS)
        free = np.ones(n, bool)
        if fixed is not None:
            free[fixed] = False
        want = np.zeros(n)
        want[free] = np.linalg.solve(A[np.ix_(free, free)], b[free].astype(np.float64))
        assert np.abs(xs - want).max() <= 1e-4 * np.abs(want).max()


def test_static_scratch_capacity(gpu):
    called = []
    x = gpu.DeviceArray(4)
    b = gpu.DeviceArray(4)
    cav = gpu._capi.AV_FN(lambda user, q, v: called.append(1))
    res = gpu.CgResult()
    n = gpu._capi.CF_CG_STATIC_MAX_N + 1
    rc = gpu.lib.cf_cg_solve_static(None, n, cav, gpu._capi.CONS_FN(0), None, x.ptr, b.ptr, 1e-6, 1e-10, 10,
                                    C.byref(res))
    assert rc == gpu.ERR_CAPACITY and not called
    assert b"static scratch" in gpu.lib.cf_last_error()
