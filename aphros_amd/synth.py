"""Synthetic benchmark systems (product generator, host side, libaphpois).

Unit periodic cube, n^3 cells, h = 1/n.  Spheres with splitmix64-drawn
centres/radii, rho = 1 inside and `ratio` outside; harmonic face density;
Proj sign (e0 > 0); RHS f uniform in [-1, 1] with the mean removed and
e7 = -f*h^3 (pressure system of src/solver/proj.ipp:343-383).
"""
import numpy as np

from . import _lib as L

# The benchmark configurations (grid edge, sphere count, radius range, density
# ratio, seed) of the measured workloads.
CONFIGS = {
    "c1_64": dict(n=64, nspheres=8, rmin=0.08, rmax=0.15, ratio=1000.0, seed=1),
    "c2_256": dict(n=256, nspheres=32, rmin=0.08, rmax=0.15, ratio=1000.0, seed=2),
    "c3_512": dict(n=512, nspheres=128, rmin=0.02, rmax=0.06, ratio=1000.0, seed=3),
}


def density(n, nspheres, rmin, rmax, ratio, seed):
    rho = np.empty(n * n * n)
    L.check(L.lib().aphpois_synth_density(n, nspheres, rmin, rmax, ratio, seed, L.dptr(rho)))
    return rho


def rhs(n, seed):
    f = np.empty(n * n * n)
    L.check(L.lib().aphpois_synth_rhs(n, seed, L.dptr(f)))
    return f


def assemble(n, rho, f):
    """Expr AoS, shape (n^3, 8), dense x-fastest."""
    rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    e = np.empty((n * n * n, 8))
    L.check(L.lib().aphpois_synth_assemble(n, L.dptr(rho), L.dptr(f), L.dptr(e)))
    return e


def system(n, nspheres, rmin, rmax, ratio, seed):
    """(expr, rho, f) of one synthetic pressure system."""
    rho = density(n, nspheres, rmin, rmax, ratio, seed)
    f = rhs(n, seed)
    return assemble(n, rho, f), rho, f


def config(name):
    c = CONFIGS[name]
    return system(c["n"], c["nspheres"], c["rmin"], c["rmax"], c["ratio"], c["seed"])
