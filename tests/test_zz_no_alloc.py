"""The allocation-free entry point (aphpois_solve_device_noalloc) against the
handle path it is a variant of (aphpois_create + aphpois_set_system_device +
aphpois_solve_device): the same kernels and tilings, so solutions and info
agree bit for bit; and against the oracle's restatement of the reference's
SolverConjugate / SolverJacobi (src/linear/linear.ipp:24-129, 158-217).

Two calls in a row on the same device, the second with a smaller system, so a
stale table, partial or flag of the first call would show in the second.
"""
import ctypes as C

import numpy as np
import pytest

from aphros_amd import _lib as L
from aphros_amd.devmem import DeviceBuffer, synchronize
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _desc(shape, method, resmode, tol, miniter, maxiter, h3):
    d = L.Desc()
    L.lib().aphpois_desc_default(C.byref(d))
    d.nx, d.ny, d.nz = shape[2], shape[1], shape[0]
    d.method = method
    d.resmode = resmode
    d.tol = tol
    d.miniter = miniter
    d.maxiter = maxiter
    d.h3 = h3
    return d


def _handle_solve(d, e_dev, x0_dev, x_dev):
    h = C.c_void_p()
    L.check(L.lib().aphpois_create(C.byref(d), C.byref(h)))
    try:
        L.check(L.lib().aphpois_set_system_device(h, e_dev.ptr))
        info = L.Info()
        L.check(L.lib().aphpois_solve_device(h, x0_dev.ptr if x0_dev else None, x_dev.ptr,
                                             C.byref(info)))
    finally:
        L.lib().aphpois_destroy(h)
    return info


def _noalloc_solve(d, e_dev, x0_dev, x_dev):
    info = L.Info()
    L.check(L.lib().aphpois_solve_device_noalloc(C.byref(d), e_dev.ptr,
                                                 x0_dev.ptr if x0_dev else None, x_dev.ptr,
                                                 None, C.byref(info)))
    return info


def _cases():
    # (name, Expr [nz, ny, nx, 8], h3): the reference's linear-test system at
    # 32^3, then at 16^3 (the second call is smaller)
    E1, _, h1 = O.tlinear_system(32)
    E2, _, h2 = O.tlinear_system(16)
    return [("tlinear32", E1, h1 ** 3), ("tlinear16", E2, h2 ** 3)]


@pytest.mark.parametrize("method", [L.CONJUGATE, L.JACOBI])
def test_noalloc_equals_handle_path_two_calls(method):
    # Jacobi: the 32^3 system converges (613 sweeps), the 16^3 one stops at
    # maxiter (4001 sweeps): both ends of the stop rule
    tol = 1e-6 if method == L.CONJUGATE else 1e-5
    for name, E, h3 in _cases():
        E = np.ascontiguousarray(E, dtype=np.float64)
        shape = E.shape[:3]
        n = E.shape[0] * E.shape[1] * E.shape[2]
        _say(f"noalloc method={method} case={name} shape={shape}")
        d = _desc(shape, method, L.RES_L2, tol, 0, 4000, h3)
        e_dev = DeviceBuffer(E.nbytes)
        e_dev.from_host(E)
        xa = DeviceBuffer(n * 8)
        xb = DeviceBuffer(n * 8)
        ia = _handle_solve(d, e_dev, None, xa)
        ib = _noalloc_solve(d, e_dev, None, xb)
        synchronize()
        a = xa.to_host(np.empty(shape))
        b = xb.to_host(np.empty(shape))
        assert ib.iter == ia.iter
        assert np.array_equal(a, b)
        for k in ("residual", "res_l2", "res_max", "true_relres", "converged", "status"):
            assert getattr(ib, k) == getattr(ia, k), k
        assert ib.max_diff == -1.0 and ib.mean == 0.0
        # the oracle's restatement of the reference solver
        if method == L.CONJUGATE:
            xo, ito, _, _ = O.conjugate(E, h3, tol=tol, miniter=0, maxiter=4000)
            assert ito == ib.iter
            assert np.abs(b - xo).max() < 1e-9
        else:
            xo, ito, _, _ = O.jacobi(E, tol=tol, miniter=0, maxiter=4000)
            assert ito == ib.iter
            assert np.array_equal(b, xo)
        for buf in (e_dev, xa, xb):
            buf.free()


def test_noalloc_warm_start_and_maxnorm():
    E, _, h = O.tlinear_system(16)
    E = np.ascontiguousarray(E)
    shape = E.shape[:3]
    n = int(np.prod(shape))
    h3 = h ** 3
    x0 = np.random.default_rng(3).standard_normal(shape)
    d = _desc(shape, L.CONJUGATE, L.RES_MAX, 1e-3, 0, 1000, h3)
    e_dev = DeviceBuffer(E.nbytes)
    e_dev.from_host(E)
    x0_dev = DeviceBuffer(n * 8)
    x0_dev.from_host(x0)
    xa = DeviceBuffer(n * 8)
    xb = DeviceBuffer(n * 8)
    _say("noalloc warm start, separate x0, max-norm")
    ia = _handle_solve(d, e_dev, x0_dev, xa)
    ib = _noalloc_solve(d, e_dev, x0_dev, xb)
    a = xa.to_host(np.empty(shape))
    b = xb.to_host(np.empty(shape))
    assert ib.iter == ia.iter and np.array_equal(a, b)
    # x0 == x: the solve starts from what x holds
    _say("noalloc warm start in place")
    xb.from_host(x0)
    ic = _noalloc_solve(d, e_dev, xb, xb)
    c = xb.to_host(np.empty(shape))
    assert ic.iter == ia.iter and np.array_equal(a, c)
    xo, ito, _, _ = O.conjugate(E, h3, tol=1e-3, miniter=0, maxiter=1000, maxnorm=True, x0=x0)
    assert ito == ic.iter and np.abs(c - xo).max() < 1e-9
    for buf in (e_dev, x0_dev, xa, xb):
        buf.free()


def test_noalloc_rejects():
    E, _, h = O.tlinear_system(16)
    shape = E.shape[:3]
    lib = L.lib()
    # beyond the static work space: refused before any device work
    assert lib.aphpois_noalloc_max_cells() == 1 << 21
    d = _desc((128, 128, 129), L.CONJUGATE, L.RES_L2, 1e-6, 0, 10, 1.0)
    info = L.Info()
    rc = lib.aphpois_solve_device_noalloc(C.byref(d), C.c_void_p(16), None, C.c_void_p(16),
                                          None, C.byref(info))
    assert rc == L.EINVAL and b"2^21" in lib.aphpois_last_error()
    d = _desc(shape, L.PCG, L.RES_L2, 1e-6, 0, 10, 1.0)
    rc = lib.aphpois_solve_device_noalloc(C.byref(d), C.c_void_p(16), None, C.c_void_p(16),
                                          None, C.byref(info))
    assert rc == L.EINVAL
    # a NaN in the system is found by the device check
    E = np.ascontiguousarray(E).copy()
    E[1, 2, 3, 0] = np.nan
    n = int(np.prod(shape))
    e_dev = DeviceBuffer(E.nbytes)
    e_dev.from_host(E)
    x = DeviceBuffer(n * 8)
    d = _desc(shape, L.CONJUGATE, L.RES_L2, 1e-6, 0, 10, h ** 3)
    rc = lib.aphpois_solve_device_noalloc(C.byref(d), e_dev.ptr, None, x.ptr, None,
                                          C.byref(info))
    assert rc == L.ENUMERIC
    e_dev.free()
    x.free()
