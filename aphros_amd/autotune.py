"""Tile autotuner for the PCG stencil kernel.

For a grid edge and precision it times the SpMV kernel (device events, a few
hundred iterations on a synthetic system of that size) for each candidate
launch shape (threads per row lx, planes per workgroup bz) and keeps the
fastest.  Results are cached in a JSON file keyed by the device (its name
and compute-unit count), the precision, the preconditioner and the edge
($APHPOIS_TUNE_CACHE, default ~/.cache/aphpois/tiles.json), so a tile tuned on
one GPU model is not reused on another.
"""
import json
import os

from . import _lib as L
from . import synth
from .devmem import device_name
from .solver import Poisson

CANDIDATES = [(64, 0), (32, 0), (16, 0), (8, 0), (64, 8), (64, 32), (64, 64), (32, 16),
              (128, 0)]


def cache_path():
    return os.environ.get("APHPOIS_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "aphpois", "tiles.json"))


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(d):
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        with open(p + ".tmp", "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(p + ".tmp", p)
    except OSError:
        pass


def tune_edge(n, precision=L.MIXED, iters=150, candidates=None, precond=L.PRECOND_JACOBI,
              device=0):
    """Returns ((lx, bz), {candidate: spmv ms})."""
    e, _, _ = synth.system(n, 2, 0.1, 0.2, 100.0, 12345)
    s = Poisson((n, n, n), method=L.PCG, precision=precision, resmode=L.RES_REL_TRUE,
                tol=1e-30, maxiter=iters, precond=precond, device=device)
    s.set_system(e)
    times = {}
    for lx, bz in candidates or CANDIDATES:
        try:
            s.set_tile(lx, bz)
        except L.AphpoisError:
            continue
        s.set_profiling(True)
        s.solve(None)
        times[(lx, bz)] = s.kernel_times()[0][0]
    s.close()
    best = min(times, key=times.get)
    return best, times


def cache_key(n, precision, precond, device=0):
    return f"{device_name(device)}|p{precision}_c{precond}_n{n}"


def tile_for(n, precision=L.MIXED, retune=False, precond=L.PRECOND_JACOBI, device=0):
    key = cache_key(n, precision, precond, device)
    d = _load()
    if key in d and not retune:
        return tuple(d[key]["tile"])
    best, times = tune_edge(n, precision, precond=precond, device=device)
    d[key] = {"tile": list(best), "ms": {f"{a}x{b}": t for (a, b), t in times.items()}}
    _save(d)
    return best
