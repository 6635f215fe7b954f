"""Dense-array front end of the C ABI: one resident system per handle."""
import ctypes as C

import numpy as np

from . import _lib as L


class Poisson:
    """aphpois handle for a dense (nz, ny, nx) periodic system.

    Poisson(shape, method=PCG, precision=MIXED, resmode=RES_REL_TRUE, tol=1e-8,
            maxiter=..., h3=...) ; set_system(expr) ; solve(x0) -> (x, info)
    """

    def __init__(self, shape, method=L.PCG, precision=L.FP64, resmode=L.RES_REL_TRUE,
                 tol=1e-8, miniter=0, maxiter=100000, h3=None, nullspace=L.NULL_AUTO,
                 delta=0.0, check_every=0, device=0, mem_budget=0, precond=L.PRECOND_JACOBI,
                 mg_w=L.MG_W_AUTO):
        nz, ny, nx = shape
        self.shape = (nz, ny, nx)
        d = L.Desc()
        L.lib().aphpois_desc_default(C.byref(d))
        d.nx, d.ny, d.nz = nx, ny, nz
        d.h3 = h3 if h3 is not None else 1.0 / (nx * ny * nz)
        d.method, d.precision, d.resmode = method, precision, resmode
        d.tol, d.miniter, d.maxiter = tol, miniter, maxiter
        d.nullspace, d.delta, d.check_every = nullspace, delta, check_every
        d.device, d.mem_budget, d.precond = device, mem_budget, precond
        d.mg_w = mg_w
        self.desc = d
        self.h = C.c_void_p()
        L.check(L.lib().aphpois_create(C.byref(d), C.byref(self.h)))

    def set_system(self, expr):
        e = np.ascontiguousarray(expr, np.float64).reshape(-1)
        nz, ny, nx = self.shape
        if e.size != 8 * nx * ny * nz:
            raise ValueError("expr must hold 8 doubles per cell")
        L.check(L.lib().aphpois_set_system(self.h, L.dptr(e), 0, nx, nx * ny))

    def set_rhs(self, e7):
        e = np.ascontiguousarray(e7, np.float64).reshape(-1)
        nz, ny, nx = self.shape
        L.check(L.lib().aphpois_set_rhs(self.h, L.dptr(e), 0, nx, nx * ny))

    def set_solution_ref(self, xref):
        """Reference solution (nz, ny, nx) for info["max_diff"] of later
        solves; None removes it."""
        nz, ny, nx = self.shape
        x = None if xref is None else np.ascontiguousarray(xref, np.float64).reshape(-1)
        if x is not None and x.size != nx * ny * nz:
            raise ValueError("xref has the wrong size")
        L.check(L.lib().aphpois_set_solution_ref(self.h, L.dptr(x), 0, nx, nx * ny))

    def set_conf(self, tol, miniter, maxiter):
        L.check(L.lib().aphpois_set_conf(self.h, tol, miniter, maxiter))

    def solve(self, x0=None, out=None):
        nz, ny, nx = self.shape
        x = out if out is not None else np.empty(self.shape)
        x0p = None
        if x0 is not None:
            x0p = x0 if x0 is x else np.ascontiguousarray(x0, np.float64)
        info = L.Info()
        L.check(L.lib().aphpois_solve(self.h, L.dptr(x0p), L.dptr(x), 0, nx, nx * ny,
                                      C.byref(info)))
        return x, info.as_dict()

    def solve_device(self, x_dev_ptr, x0_dev_ptr=None):
        info = L.Info()
        L.check(L.lib().aphpois_solve_device(self.h, x0_dev_ptr, x_dev_ptr, C.byref(info)))
        return info.as_dict()

    def set_system_device(self, expr_dev_ptr):
        L.check(L.lib().aphpois_set_system_device(self.h, expr_dev_ptr))

    def set_profiling(self, on=True):
        L.check(L.lib().aphpois_set_profiling(self.h, int(on)))

    def kernel_times(self):
        avg = (C.c_double * 16)()
        cnt = (C.c_longlong * 16)()
        nk = C.c_int()
        L.check(L.lib().aphpois_kernel_times(self.h, avg, cnt, C.byref(nk)))
        return [(avg[j], cnt[j]) for j in range(nk.value)]

    def launch_info(self):
        a, b, v = C.c_int(), C.c_int(), C.c_int()
        L.check(L.lib().aphpois_launch_info(self.h, C.byref(a), C.byref(b), C.byref(v)))
        return {"stencil_blocks": a.value, "stream_blocks": b.value, "vx": v.value}

    def set_tile(self, lx=0, bz=0, vx=0, sm=0):
        L.check(L.lib().aphpois_set_tile(self.h, lx, bz, vx, sm))

    def assemble(self, rho, ffv, src=None, cell=None, dt=1.0, sign=1):
        """The projection's pressure system from density (nz, ny, nx), face
        volume fluxes (3, nz, ny, nx; lower faces) and volume source,
        assembled on the device and made resident."""
        nz, ny, nx = self.shape
        rho = np.ascontiguousarray(rho, np.float64)
        ffv = np.ascontiguousarray(ffv, np.float64)
        if rho.size != nx * ny * nz or ffv.size != 3 * rho.size:
            raise ValueError("rho needs nx*ny*nz values and ffv 3*nx*ny*nz")
        src = None if src is None else np.ascontiguousarray(src, np.float64)
        cell = 1.0 / nx if cell is None else cell
        L.check(L.lib().aphpois_assemble(self.h, L.dptr(rho), L.dptr(ffv), L.dptr(src), cell,
                                         dt, sign))

    def assemble_faces(self, fk, ffv, src=None, cell=None, sign=1):
        """SIMPLE's pressure-correction system: face denominators fk (3, nz,
        ny, nx) instead of the density."""
        nz, ny, nx = self.shape
        fk = np.ascontiguousarray(fk, np.float64)
        ffv = np.ascontiguousarray(ffv, np.float64)
        if fk.size != 3 * nx * ny * nz or ffv.size != fk.size:
            raise ValueError("fk and ffv need 3*nx*ny*nz values")
        src = None if src is None else np.ascontiguousarray(src, np.float64)
        cell = 1.0 / nx if cell is None else cell
        L.check(L.lib().aphpois_assemble_faces(self.h, L.dptr(fk), L.dptr(ffv), L.dptr(src),
                                               cell, sign))

    def flux_correct_faces(self, p, fk, ffv, cell=None):
        nz, ny, nx = self.shape
        p = np.ascontiguousarray(p, np.float64)
        fk = np.ascontiguousarray(fk, np.float64)
        ffv = np.ascontiguousarray(ffv, np.float64)
        out = np.empty((3, nz, ny, nx))
        cell = 1.0 / nx if cell is None else cell
        L.check(L.lib().aphpois_flux_correct_faces(self.h, L.dptr(p), L.dptr(fk), L.dptr(ffv),
                                                   L.dptr(out), cell))
        return out

    def flux_correct(self, p, rho, ffv, cell=None, dt=1.0):
        """Corrected face fluxes (3, nz, ny, nx) for the pressure p."""
        nz, ny, nx = self.shape
        p = np.ascontiguousarray(p, np.float64)
        rho = np.ascontiguousarray(rho, np.float64)
        ffv = np.ascontiguousarray(ffv, np.float64)
        out = np.empty((3, nz, ny, nx))
        cell = 1.0 / nx if cell is None else cell
        L.check(L.lib().aphpois_flux_correct(self.h, L.dptr(p), L.dptr(rho), L.dptr(ffv),
                                             L.dptr(out), cell, dt))
        return out

    def precond_apply(self, r):
        """z = M r with the multigrid preconditioner (PCG with precond MG);
        r in the positive definite sign convention, shape (nz, ny, nx)."""
        r = np.ascontiguousarray(r, np.float64)
        if r.size != int(np.prod(self.shape)):
            raise ValueError("r has the wrong size")
        z = np.empty_like(r)
        L.check(L.lib().aphpois_precond_apply(self.h, L.dptr(r), L.dptr(z)))
        return z

    def device_bytes(self):
        return L.lib().aphpois_device_bytes(self.h)

    def close(self):
        if self.h:
            L.lib().aphpois_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
