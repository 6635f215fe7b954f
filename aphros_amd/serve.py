"""Request server front end (libaphpois aphpois_server_*): continuous
batching of warm-started correlated solve sequences on one GPU."""
import ctypes as C

import numpy as np

from . import _lib as L


class Server:
    def __init__(self, precision=L.MIXED, tol=1e-8, maxiter=200000, delta=0.0, check_every=0,
                 device=0, mem_budget=0, max_active=0, precond=L.PRECOND_JACOBI,
                 mg_w=L.MG_W_AUTO):
        d = L.ServerDesc()
        d.precond = precond
        d.mg_w = mg_w
        d.precision, d.tol, d.maxiter, d.delta = precision, tol, maxiter, delta
        d.check_every, d.device, d.mem_budget, d.max_active = (check_every, device,
                                                                int(mem_budget), max_active)
        self.h = C.c_void_p()
        rc = L.lib().aphpois_server_create(C.byref(d), C.byref(self.h))
        if rc != L.OK:
            raise L.AphpoisError(rc, L.lib().aphpois_server_last_error().decode())
        self._keep = {}

    def submit(self, expr, nsteps=1, perturb=0.01, seed=0, e7_steps=None, want_x=False,
               h3=None):
        """expr: (nz, ny, nx, 8).  e7_steps: optional list of nsteps arrays (entry 0
        ignored, None entries use the perturbation)."""
        e = np.ascontiguousarray(expr, np.float64)
        nz, ny, nx = e.shape[:3]
        r = L.Request()
        r.nx, r.ny, r.nz = nx, ny, nz
        r.h3 = h3 if h3 is not None else 1.0 / (nx * ny * nz)
        r.expr = L.dptr(e.reshape(-1))
        r.nsteps = nsteps
        keep = [e, r]
        if e7_steps is not None:
            arrs = [None if a is None else np.ascontiguousarray(a, np.float64).reshape(-1)
                    for a in e7_steps]
            ptrs = (C.POINTER(C.c_double) * nsteps)(
                *[L.dptr(a) if a is not None else C.POINTER(C.c_double)() for a in arrs])
            r.e7 = ptrs
            keep += [arrs, ptrs]
        r.perturb = perturb
        r.seed = seed
        x = np.empty((nz, ny, nx)) if want_x else None
        r.x_out = L.dptr(x)
        keep.append(x)
        rid = C.c_longlong()
        rc = L.lib().aphpois_server_submit(self.h, C.byref(r), C.byref(rid))
        if rc != L.OK:
            raise L.AphpoisError(rc, L.lib().aphpois_server_last_error().decode())
        self._keep[rid.value] = keep
        return rid.value

    def wait(self, rid, timeout=-1.0):
        res = L.Result()
        rc = L.lib().aphpois_server_wait(self.h, rid, float(timeout), C.byref(res))
        if rc != L.OK:
            raise L.AphpoisError(rc, L.lib().aphpois_server_last_error().decode())
        keep = self._keep.pop(rid)
        out = res.as_dict()
        out["x"] = keep[-1]
        return out

    def cancel(self, rid):
        L.lib().aphpois_server_cancel(self.h, rid)

    def stats(self):
        st = L.ServerStats()
        L.lib().aphpois_server_get_stats(self.h, C.byref(st))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self.h:
            L.lib().aphpois_server_destroy(self.h)
            self.h = C.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
