"""CPU oracle for the pressure-Poisson path: TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker.  The product (libaphpois) never calls it.

Contents:
  * ctypes binding of oracle/liboracle.so (oracle.c: restatements of
    SolverConjugate, SolverJacobi, the pressure assembly, FP64 Jacobi-PCG);
  * numpy twin of the synthetic-system generator (independent of the
    library's C++ generator, used to check it);
  * the reference's linear-test system (src/test/linear/main.cpp:45-85);
  * a runner for oracle/_ref/ref_driver (the reference's own solver templates,
    built from its sources where they exist).
"""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")

_dp = C.POINTER(C.c_double)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `python -m aphros_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.orc_assemble.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, C.c_int, _dp]
        L.orc_assemble.restype = None
        L.orc_apply.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int]
        L.orc_apply.restype = None
        L.orc_conjugate.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp,
                                    C.c_double, C.c_int, C.c_int, C.c_int, _dp, C.c_int, _dp]
        L.orc_conjugate.restype = C.c_int
        L.orc_jacobi.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_int,
                                 C.c_int, _dp, C.c_int, _dp]
        L.orc_jacobi.restype = C.c_int
        L.orc_true_relres.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_true_relres.restype = C.c_double
        L.orc_pcg_jacobi.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_double,
                                     C.c_int, C.c_int, _dp]
        L.orc_pcg_jacobi.restype = C.c_int
        L.orc_assemble_flux.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _dp,
                                         _dp, _dp, _dp, C.c_int, _dp]
        L.orc_assemble_flux.restype = None
        L.orc_flux_correct.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _dp,
                                       _dp, _dp, _dp, _dp]
        L.orc_flux_correct.restype = None
        L.orc_mg_apply.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   _dp, _dp]
        L.orc_mg_apply.restype = C.c_int
        L.orc_mgpcg.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_int,
                                C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_mgpcg.restype = C.c_int
        L.orc_num_threads.argtypes = []
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _dims(shape):
    nz, ny, nx = shape
    return nx, ny, nz


# ------------------------------------------------------------------ solvers

def assemble(rho, f, sign=1, extent=1.0):
    """Expr (nz,ny,nx,8) from cell density and source, reference op order."""
    rho = np.ascontiguousarray(rho, np.float64)
    f = np.ascontiguousarray(f, np.float64)
    nx, ny, nz = _dims(rho.shape)
    e = np.empty(rho.shape + (8,))
    lib().orc_assemble(nx, ny, nz, extent, _p(rho), _p(f), sign, _p(e))
    return e


def assemble_flux(rho, ffv, src=None, cell=None, dt=1.0, sign=1, fk=None):
    """Expr (nz,ny,nx,8) of the projection with face volume fluxes ffv
    (3, nz, ny, nx: x, y, z lower faces of each cell) and volume source; with
    fk (rho None) the face denominators are given (SIMPLE)."""
    shape = ffv.shape[1:]
    nx, ny, nz = _dims(shape)
    rho = None if rho is None else np.ascontiguousarray(rho, np.float64)
    fk = None if fk is None else np.ascontiguousarray(fk, np.float64)
    ffv = np.ascontiguousarray(ffv, np.float64)
    src = None if src is None else np.ascontiguousarray(src, np.float64)
    cell = 1.0 / nx if cell is None else cell
    e = np.empty(tuple(shape) + (8,))
    lib().orc_assemble_flux(nx, ny, nz, cell, dt, _p(rho), _p(fk), _p(ffv), _p(src), sign,
                            _p(e))
    return e


def flux_correct(rho, ffv, p, cell=None, dt=1.0, fk=None):
    """Corrected face fluxes (3, nz, ny, nx) after the pressure solve."""
    shape = ffv.shape[1:]
    nx, ny, nz = _dims(shape)
    rho = None if rho is None else np.ascontiguousarray(rho, np.float64)
    fk = None if fk is None else np.ascontiguousarray(fk, np.float64)
    ffv = np.ascontiguousarray(ffv, np.float64)
    p = np.ascontiguousarray(p, np.float64)
    cell = 1.0 / nx if cell is None else cell
    out = np.empty((3,) + tuple(shape))
    lib().orc_flux_correct(nx, ny, nz, cell, dt, _p(rho), _p(fk), _p(ffv), _p(p), _p(out))
    return out


def apply(expr, x, with_const=False):
    expr = np.ascontiguousarray(expr, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    nx, ny, nz = _dims(expr.shape[:3])
    y = np.empty(expr.shape[:3])
    lib().orc_apply(_p(expr), nx, ny, nz, _p(x), _p(y), int(with_const))
    return y


def conjugate(expr, h3, x0=None, tol=0.0, miniter=1, maxiter=100, maxnorm=False, hist=0):
    """SolverConjugate; returns (x, iter, residual, history)."""
    expr = np.ascontiguousarray(expr, np.float64)
    nx, ny, nz = _dims(expr.shape[:3])
    x = np.empty(expr.shape[:3])
    x0 = None if x0 is None else np.ascontiguousarray(x0, np.float64)
    res = C.c_double()
    hb = np.zeros(max(hist, 1))
    it = lib().orc_conjugate(_p(expr), nx, ny, nz, h3, _p(x0), _p(x), tol, miniter, maxiter,
                             int(maxnorm), _p(hb) if hist else None, hist, C.byref(res))
    if it < 0:
        raise MemoryError("oracle allocation failed")
    return x, it, res.value, hb[:min(it, hist)]


def jacobi(expr, x0=None, tol=0.0, miniter=1, maxiter=100, hist=0):
    """SolverJacobi; returns (x, iter, residual, history)."""
    expr = np.ascontiguousarray(expr, np.float64)
    nx, ny, nz = _dims(expr.shape[:3])
    x = np.empty(expr.shape[:3])
    x0 = None if x0 is None else np.ascontiguousarray(x0, np.float64)
    res = C.c_double()
    hb = np.zeros(max(hist, 1))
    it = lib().orc_jacobi(_p(expr), nx, ny, nz, _p(x0), _p(x), tol, miniter, maxiter,
                          _p(hb) if hist else None, hist, C.byref(res))
    if it < 0:
        raise MemoryError("oracle allocation failed")
    return x, it, res.value, hb[:min(it, hist)]


def true_relres(expr, x):
    expr = np.ascontiguousarray(expr, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    nx, ny, nz = _dims(expr.shape[:3])
    return lib().orc_true_relres(_p(expr), nx, ny, nz, _p(x))


def pcg_jacobi(expr, tol=1e-8, maxiter=100000, remove_mean=True, x0=None):
    """FP64 Jacobi-PCG to a true relative residual; returns (x, iter, relres).
    maxiter < 0 runs exactly -maxiter iterations (timing sample)."""
    expr = np.ascontiguousarray(expr, np.float64)
    nx, ny, nz = _dims(expr.shape[:3])
    x = np.empty(expr.shape[:3])
    x0 = None if x0 is None else np.ascontiguousarray(x0, np.float64)
    rr = C.c_double()
    it = lib().orc_pcg_jacobi(_p(expr), nx, ny, nz, _p(x0), _p(x), tol, maxiter,
                              int(remove_mean), C.byref(rr))
    if it < 0:
        raise MemoryError("oracle allocation failed")
    return x, it, rr.value


MG_SWEEPS = 20  # damped Jacobi sweeps on the coarsest level (the engine's kCoarseSweeps)


def mg_apply(expr, r, sweeps=MG_SWEEPS, bf16=False, w=0):
    """One cycle z = M r of the multigrid preconditioner; returns (z, nlev).
    bf16: coefficients stored as the GPU's FP32 engines store them; w: levels
    with two coarse corrections."""
    expr = np.ascontiguousarray(expr, np.float64)
    nx, ny, nz = _dims(expr.shape[:3])
    r = np.ascontiguousarray(r, np.float64)
    z = np.empty(expr.shape[:3])
    nlev = lib().orc_mg_apply(_p(expr), nx, ny, nz, sweeps, int(bf16), w, _p(r), _p(z))
    if nlev == -2:
        raise ValueError("grid cannot be coarsened for the multigrid preconditioner")
    if nlev < 0:
        raise MemoryError("oracle allocation failed")
    return z, nlev


def mgpcg(expr, tol=1e-8, maxiter=100000, remove_mean=True, x0=None, sweeps=MG_SWEEPS, w=0):
    """FP64 multigrid-preconditioned CG to a true relative residual; returns
    (x, iter, relres, setup_seconds).  maxiter < 0 runs exactly -maxiter
    iterations (timing sample)."""
    expr = np.ascontiguousarray(expr, np.float64)
    nx, ny, nz = _dims(expr.shape[:3])
    x = np.empty(expr.shape[:3])
    x0 = None if x0 is None else np.ascontiguousarray(x0, np.float64)
    rr = C.c_double()
    setup = C.c_double()
    it = lib().orc_mgpcg(_p(expr), nx, ny, nz, _p(x0), _p(x), tol, maxiter, sweeps, w,
                         int(remove_mean), C.byref(rr), C.byref(setup))
    if it == -2:
        raise ValueError("grid cannot be coarsened for the multigrid preconditioner")
    if it < 0:
        raise MemoryError("oracle allocation failed")
    return x, it, rr.value, setup.value


def num_threads():
    return lib().orc_num_threads()


# ------------------------------------------------- numpy twin of the generator

_M64 = (1 << 64) - 1


def _sm64_int(x):
    x = (x + 0x9E3779B97F4A7C15) & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def _sm64_np(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def _key(seed, stream):
    return _sm64_int((_sm64_int(seed) + stream) & _M64)


def _u01(k, i):
    with np.errstate(over="ignore"):
        v = _sm64_np(np.uint64(k) + np.asarray(i, dtype=np.uint64))
    return (v >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def twin_density(n, nspheres, rmin, rmax, ratio, seed):
    """Brute-force sphere density on (n,n,n) (z,y,x)."""
    h = 1.0 / n
    ks = _key(seed, 1)
    s = np.arange(nspheres, dtype=np.uint64)
    cx, cy, cz = _u01(ks, 4 * s), _u01(ks, 4 * s + 1), _u01(ks, 4 * s + 2)
    r = rmin + (rmax - rmin) * _u01(ks, 4 * s + 3)
    c = (np.arange(n) + 0.5) * h
    rho = np.full((n, n, n), float(ratio))
    for i in range(nspheres):
        dx = c - cx[i]
        dx -= np.rint(dx)
        dy = c - cy[i]
        dy -= np.rint(dy)
        dz = c - cz[i]
        dz -= np.rint(dz)
        d2 = (dx * dx)[None, None, :] + (dy * dy)[None, :, None] + (dz * dz)[:, None, None]
        rho[d2 < r[i] * r[i]] = 1.0
    return rho


def twin_rhs(n, seed):
    kf = _key(seed, 2)
    f = 2.0 * _u01(kf, np.arange(n ** 3, dtype=np.uint64)) - 1.0
    f = f.reshape(n, n * n)
    psum = np.cumsum(f, axis=1)[:, -1]          # per-plane sequential sums
    tot = np.cumsum(psum)[-1]                   # planes in order
    f = f - tot / float(n ** 3)
    return f.reshape(n, n, n)


# --------------------------------------- the reference's linear-test system

def tlinear_system(n=32):
    """System of src/test/linear/main.cpp:45-85 on an n^3 periodic grid:
    u* = sin(2 pi x) sin(2 pi y^2) sin(2 pi z^3) at cell centres; face
    resistivity 10 within 0.2 of the centre, else 1; flux = grad/rho*area
    (GradientImplicit, AppendExpr), so e0 < 0; e7 = -(A u*).
    Returns (expr (n,n,n,8), u_exact (n,n,n), h)."""
    h = 1.0 / n
    c = (np.arange(n) + 0.5) * h
    f = np.arange(n) * h
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    ue = np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y ** 2) * np.sin(2 * np.pi * Z ** 3)

    def rho(fx, fy, fz):
        d = np.sqrt((fx - 0.5) ** 2 + (fy - 0.5) ** 2 + (fz - 0.5) ** 2)
        return np.where(d < 0.2, 10.0, 1.0)

    vol = h * h * h
    area = vol / h
    a = 1.0 / h
    # minus faces of each cell (face x = i*h for cell i)
    Zf, Yf, Xf = np.meshgrid(c, c, f, indexing="ij")
    kx = (a / rho(Xf, Yf, Zf)) * area
    Zf, Yf, Xf = np.meshgrid(c, f, c, indexing="ij")
    ky = (a / rho(Xf, Yf, Zf)) * area
    Zf, Yf, Xf = np.meshgrid(f, c, c, indexing="ij")
    kz = (a / rho(Xf, Yf, Zf)) * area
    E = np.zeros((n, n, n, 8))
    E[..., 1] = kx
    E[..., 2] = np.roll(kx, -1, axis=2)
    E[..., 3] = ky
    E[..., 4] = np.roll(ky, -1, axis=1)
    E[..., 5] = kz
    E[..., 6] = np.roll(kz, -1, axis=0)
    # AppendExpr accumulates the diagonal over q = 0..5 (mesh.h:575-579)
    e0 = np.zeros((n, n, n))
    for q in range(6):
        e0 = e0 - E[..., 1 + q]
    E[..., 0] = e0
    E[..., 7] = -apply(E, ue)
    return E, ue, h


def max_diff_mean_removed(x, ue):
    """max|d - mean(d)|, d = x - u* (src/test/linear/main.cpp:107-128)."""
    d = x - ue
    d = d - d.mean()
    return np.abs(d).max()


# ------------------------------------------------ the reference's own solver

def ref_available():
    return os.path.exists(REF_DRIVER)


def ref_solve(name, expr, tol, miniter, maxiter, maxnorm=0, x0=None, timed=False):
    """Runs the reference's SolverConjugate/SolverJacobi (oracle/_ref);
    returns (x, iter, residual) and, if timed, the solve's wall seconds."""
    expr = np.ascontiguousarray(expr, np.float64)
    nz, ny, nx = expr.shape[:3]
    with tempfile.TemporaryDirectory() as d:
        pe = os.path.join(d, "e.bin")
        px = os.path.join(d, "x.bin")
        po = os.path.join(d, "o.bin")
        expr.tofile(pe)
        if x0 is not None:
            np.ascontiguousarray(x0, np.float64).tofile(px)
        out = subprocess.run(
            [REF_DRIVER, name, str(nx), str(ny), str(nz), repr(float(tol)), str(miniter),
             str(maxiter), str(int(maxnorm)), pe, px if x0 is not None else "-", po],
            check=True, stdout=subprocess.PIPE, text=True).stdout
        x = np.fromfile(po).reshape(nz, ny, nx)
    kv = dict(t.split("=") for t in out.split())
    if timed:
        return x, int(kv["iter"]), float(kv["res"]), float(kv["sec"])
    return x, int(kv["iter"]), float(kv["res"])
