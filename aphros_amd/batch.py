"""Batched solves: many independent periodic systems advanced by the same
kernel launches (libaphpois aphpois_batch_*)."""
import ctypes as C

import numpy as np

from . import _lib as L


class Batch:
    def __init__(self, precision=L.MIXED, resmode=L.RES_REL_TRUE, tol=1e-8, maxiter=200000,
                 delta=0.0, check_every=0, device=0, mem_budget=0, max_systems=1024,
                 precond=L.PRECOND_JACOBI, mg_w=L.MG_W_BATCH):
        d = L.BatchDesc()
        d.precond = precond
        d.mg_w = mg_w
        d.precision, d.resmode, d.tol, d.maxiter = precision, resmode, tol, maxiter
        d.delta, d.check_every, d.device = delta, check_every, device
        d.mem_budget, d.max_systems = int(mem_budget), max_systems
        self.h = C.c_void_p()
        L.check(L.lib().aphpois_batch_create(C.byref(d), C.byref(self.h)))
        self.shapes = {}

    def add(self, expr, x0=None, h3=None):
        e = np.ascontiguousarray(expr, np.float64)
        nz, ny, nx = e.shape[:3]
        h3 = h3 if h3 is not None else 1.0 / (nx * ny * nz)
        x0p = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        slot = C.c_int()
        L.check(L.lib().aphpois_batch_add(self.h, nx, ny, nz, h3, L.dptr(e.reshape(-1)),
                                          L.dptr(x0p), C.byref(slot)))
        self.shapes[slot.value] = (nz, ny, nx)
        return slot.value

    def set_rhs(self, slot, e7=None, x0=None):
        e = None if e7 is None else np.ascontiguousarray(e7, np.float64).reshape(-1)
        x = None if x0 is None else np.ascontiguousarray(x0, np.float64).reshape(-1)
        L.check(L.lib().aphpois_batch_set_rhs(self.h, slot, L.dptr(e), L.dptr(x)))

    def set_system_tile(self, slot, lx, bz):
        L.check(L.lib().aphpois_batch_set_system_tile(self.h, slot, lx, bz))

    def solve(self):
        L.check(L.lib().aphpois_batch_solve(self.h))
        return L.lib().aphpois_batch_last_seconds(self.h)

    def get(self, slot, want_x=True):
        info = L.Info()
        x = np.empty(self.shapes[slot]) if want_x else None
        L.check(L.lib().aphpois_batch_get(self.h, slot, L.dptr(x), C.byref(info)))
        return x, info.as_dict()

    def remove(self, slot):
        L.check(L.lib().aphpois_batch_remove(self.h, slot))
        self.shapes.pop(slot, None)

    def device_bytes(self):
        return L.lib().aphpois_batch_device_bytes(self.h)

    def close(self):
        if self.h:
            L.lib().aphpois_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
