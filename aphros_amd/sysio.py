"""System files: the linear system as an array (nz, ny, nx, 8) of float64,
the dataset of the reference's --system_in / --system_out
(src/test/linear/main.cpp:181-189), stored as .npy through libaphpois
(aphpois_system_file_*)."""
import ctypes as C
import os

import numpy as np

from . import _lib as L


def _check(rc):
    if rc != L.OK:
        raise L.AphpoisError(rc, L.lib().aphpois_system_file_error().decode())


def load_system(path):
    """Expr array (nz, ny, nx, 8) from a system file."""
    nx, ny, nz = C.c_int(), C.c_int(), C.c_int()
    p = os.fsencode(path)
    _check(L.lib().aphpois_system_file_info(p, C.byref(nx), C.byref(ny), C.byref(nz)))
    e = np.empty((nz.value, ny.value, nx.value, 8))
    _check(L.lib().aphpois_system_file_read(p, L.dptr(e.reshape(-1)), e.size))
    return e


def save_system(path, expr):
    """Writes an Expr array (nz, ny, nx, 8) as a system file."""
    e = np.ascontiguousarray(expr, np.float64)
    if e.ndim != 4 or e.shape[3] != 8:
        raise ValueError("expr must have shape (nz, ny, nx, 8)")
    nz, ny, nx = e.shape[:3]
    _check(L.lib().aphpois_system_file_write(os.fsencode(path), nx, ny, nz,
                                             L.dptr(e.reshape(-1))))
