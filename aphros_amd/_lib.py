"""ctypes binding of libaphpois (include/aphpois.h).

The library is the product: HIP kernels for gfx950 behind a C ABI.  Loading
fails loudly when it has not been built; there is no CPU fallback.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaphpois.so")

OK, EINVAL, ENOMEM, EDEVICE, ESTATE, EBUDGET, ENUMERIC, ECANCEL = range(8)
CONJUGATE, JACOBI, PCG = 0, 1, 2
FP64, MIXED = 0, 1
RES_L2, RES_MAX, RES_REL_TRUE = 0, 1, 2
NULL_AUTO, NULL_OFF, NULL_ON = 0, 1, 2
PRECOND_JACOBI, PRECOND_MG = 0, 1

ERROR_NAMES = {OK: "OK", EINVAL: "EINVAL", ENOMEM: "ENOMEM", EDEVICE: "EDEVICE",
               ESTATE: "ESTATE", EBUDGET: "EBUDGET", ENUMERIC: "ENUMERIC",
               ECANCEL: "ECANCEL"}


class Desc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("h3", C.c_double),
                ("method", C.c_int), ("precision", C.c_int), ("resmode", C.c_int),
                ("tol", C.c_double), ("miniter", C.c_int), ("maxiter", C.c_int),
                ("nullspace", C.c_int), ("delta", C.c_double), ("check_every", C.c_int),
                ("device", C.c_int), ("mem_budget", C.c_longlong), ("precond", C.c_int),
                ("mg_w", C.c_int)]


class Info(C.Structure):
    _fields_ = [("iter", C.c_int), ("residual", C.c_double), ("res_l2", C.c_double),
                ("res_max", C.c_double), ("true_relres", C.c_double), ("mean", C.c_double),
                ("converged", C.c_int), ("refinements", C.c_int), ("status", C.c_int),
                ("seconds", C.c_double), ("max_diff", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchDesc(C.Structure):
    _fields_ = [("precision", C.c_int), ("resmode", C.c_int), ("tol", C.c_double),
                ("maxiter", C.c_int), ("delta", C.c_double), ("check_every", C.c_int),
                ("device", C.c_int), ("mem_budget", C.c_longlong), ("max_systems", C.c_int),
                ("precond", C.c_int), ("mg_w", C.c_int)]


class ServerDesc(C.Structure):
    _fields_ = [("precision", C.c_int), ("tol", C.c_double), ("maxiter", C.c_int),
                ("delta", C.c_double), ("check_every", C.c_int), ("device", C.c_int),
                ("mem_budget", C.c_longlong), ("max_active", C.c_int), ("precond", C.c_int),
                ("mg_w", C.c_int)]


class Request(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("h3", C.c_double),
                ("expr", C.POINTER(C.c_double)), ("nsteps", C.c_int),
                ("e7", C.POINTER(C.POINTER(C.c_double))), ("perturb", C.c_double),
                ("seed", C.c_ulonglong), ("x_out", C.POINTER(C.c_double))]


class Result(C.Structure):
    _fields_ = [("id", C.c_longlong), ("status", C.c_int), ("steps_done", C.c_int),
                ("iters_total", C.c_longlong), ("refinements_total", C.c_int),
                ("last_relres", C.c_double), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [("queued", C.c_int), ("active", C.c_int), ("completed", C.c_longlong),
                ("device_bytes", C.c_longlong), ("time", C.c_double)]


class AphpoisError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


_dp = C.POINTER(C.c_double)
_vp = C.c_void_p
_ll = C.c_longlong

# name -> (restype, argtypes); the exported surface of include/aphpois.h
MG_W_DEFAULT = 3  # APHPOIS_MG_W_DEFAULT
# Batch default: a batch shares its launches among its systems, so W levels
# cost little latency per system (96 systems of 32^3 .. 160^3 to 1e-8: 0.084 s
# with 2, 0.088 s with 3, 0.091 s with 1, 0.106 s with 4, 0.108 s automatic)
MG_W_BATCH = 2
MG_W_AUTO = -1  # APHPOIS_MG_W_AUTO


def mg_w_auto(n):
    """The W levels APHPOIS_MG_W_AUTO gives a system of n cells (the engine's
    rule, before clamping to the system's depth)."""
    lg = math.log2(n)
    return 0 if lg < 22.5 else (1 if lg < 24.5 else (2 if lg < 25.5 else 3))

SIGNATURES = {
    "aphpois_desc_default": (None, [C.POINTER(Desc)]),
    "aphpois_create": (C.c_int, [C.POINTER(Desc), C.POINTER(_vp)]),
    "aphpois_set_system": (C.c_int, [_vp, _dp, _ll, _ll, _ll]),
    "aphpois_set_system_device": (C.c_int, [_vp, _vp]),
    "aphpois_set_rhs": (C.c_int, [_vp, _dp, _ll, _ll, _ll]),
    "aphpois_solve": (C.c_int, [_vp, _dp, _dp, _ll, _ll, _ll, C.POINTER(Info)]),
    "aphpois_solve_device": (C.c_int, [_vp, _vp, _vp, C.POINTER(Info)]),
    "aphpois_solve_device_noalloc": (C.c_int, [C.POINTER(Desc), _vp, _vp, _vp, _vp,
                                               C.POINTER(Info)]),
    "aphpois_noalloc_max_cells": (C.c_longlong, []),
    "aphpois_set_conf": (C.c_int, [_vp, C.c_double, C.c_int, C.c_int]),
    "aphpois_set_solution_ref": (C.c_int, [_vp, _dp, _ll, _ll, _ll]),
    "aphpois_device_bytes": (C.c_longlong, [_vp]),
    "aphpois_destroy": (C.c_int, [_vp]),
    "aphpois_last_error": (C.c_char_p, []),
    "aphpois_set_profiling": (C.c_int, [_vp, C.c_int]),
    "aphpois_kernel_times": (C.c_int, [_vp, _dp, C.POINTER(C.c_longlong), C.POINTER(C.c_int)]),
    "aphpois_launch_info": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    "aphpois_set_tile": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "aphpois_precond_apply": (C.c_int, [_vp, _dp, _dp]),
    "aphpois_copy_bandwidth": (C.c_int, [C.c_int, _ll, C.c_int, _dp]),
    "aphpois_system_file_info": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                           C.POINTER(C.c_int)]),
    "aphpois_system_file_read": (C.c_int, [C.c_char_p, _dp, _ll]),
    "aphpois_system_file_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, _dp]),
    "aphpois_system_file_error": (C.c_char_p, []),
    "aphpois_assemble": (C.c_int, [_vp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int]),
    "aphpois_assemble_device": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_double,
                                          C.c_int]),
    "aphpois_flux_correct": (C.c_int, [_vp, _dp, _dp, _dp, _dp, C.c_double, C.c_double]),
    "aphpois_assemble_faces": (C.c_int, [_vp, _dp, _dp, _dp, C.c_double, C.c_int]),
    "aphpois_flux_correct_faces": (C.c_int, [_vp, _dp, _dp, _dp, _dp, C.c_double]),
    "aphpois_flux_correct_device": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_double,
                                              C.c_double]),
    "aphpois_batch_set_tile": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "aphpois_batch_set_system_tile": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "aphpois_batch_last_seconds": (C.c_double, [_vp]),
    "aphpois_server_create": (C.c_int, [C.POINTER(ServerDesc), C.POINTER(_vp)]),
    "aphpois_server_submit": (C.c_int, [_vp, C.POINTER(Request), C.POINTER(C.c_longlong)]),
    "aphpois_server_wait": (C.c_int, [_vp, C.c_longlong, C.c_double, C.POINTER(Result)]),
    "aphpois_server_cancel": (C.c_int, [_vp, C.c_longlong]),
    "aphpois_server_get_stats": (C.c_int, [_vp, C.POINTER(ServerStats)]),
    "aphpois_server_destroy": (C.c_int, [_vp]),
    "aphpois_server_last_error": (C.c_char_p, []),
    "aphpois_batch_create": (C.c_int, [C.POINTER(BatchDesc), C.POINTER(_vp)]),
    "aphpois_batch_add": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp,
                                    C.POINTER(C.c_int)]),
    "aphpois_batch_set_rhs": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "aphpois_batch_solve": (C.c_int, [_vp]),
    "aphpois_batch_get": (C.c_int, [_vp, C.c_int, _dp, C.POINTER(Info)]),
    "aphpois_batch_remove": (C.c_int, [_vp, C.c_int]),
    "aphpois_batch_device_bytes": (C.c_longlong, [_vp]),
    "aphpois_batch_destroy": (C.c_int, [_vp]),
    "aphpois_synth_density": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.c_double, C.c_ulonglong, _dp]),
    "aphpois_synth_rhs": (C.c_int, [C.c_int, C.c_ulonglong, _dp]),
    "aphpois_synth_assemble": (C.c_int, [C.c_int, _dp, _dp, _dp]),
}

_lib = None


def lib():
    """Loads libaphpois.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m aphros_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != OK:
        raise AphpoisError(code, lib().aphpois_last_error().decode())
    return code


def dptr(a):
    """double* of a contiguous float64 numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)
