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


# ---------------------------------------------------------------- workloads
# Parameters of the batched and served workloads are drawn with the same
# splitmix64 counter generator as the fields: u(seed, stream, i) =
# (splitmix64(key(seed, stream) + i) >> 11) * 2^-53, stream 3 for the
# per-system parameters and stream 4 for the arrival times.

_M64 = (1 << 64) - 1


def _sm64(x):
    x = (x + 0x9E3779B97F4A7C15) & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def _u(seed, stream, i):
    k = _sm64((_sm64(seed) + stream) & _M64)
    return (_sm64((k + i) & _M64) >> 11) * 2.0 ** -53


BATCH_EDGES = (32, 48, 64, 96, 128, 160)
SERVE_EDGES = (48, 64, 96, 128, 192, 256)


def batch_specs(count=96, seed=4):
    """The batched workload: `count` periodic systems, system i drawn from
    seed + i: edge uniform in BATCH_EDGES, 1..16 spheres of radius in
    [0.05, 0.15], density ratio log-uniform in [10, 1000]."""
    out = []
    for i in range(count):
        sd = seed + i
        n = BATCH_EDGES[min(5, int(_u(sd, 3, 0) * 6))]
        ns = 1 + min(15, int(_u(sd, 3, 1) * 16))
        ratio = 10.0 ** (1.0 + 2.0 * _u(sd, 3, 2))
        out.append(dict(n=n, nspheres=ns, rmin=0.05, rmax=0.15, ratio=ratio, seed=sd))
    return out


def serve_specs(count, rate, seed=5):
    """The served workload: request j (seed + j) is a periodic system with an
    edge uniform in SERVE_EDGES, 8 spheres of radius in [0.05, 0.15], density
    ratio 1000; arrival times of a Poisson process of `rate` requests/s."""
    out, t = [], 0.0
    for j in range(count):
        sd = seed + j
        t += -np.log1p(-_u(seed, 4, j)) / rate
        n = SERVE_EDGES[min(5, int(_u(sd, 3, 0) * 6))]
        out.append(dict(n=n, nspheres=8, rmin=0.05, rmax=0.15, ratio=1000.0, seed=sd,
                        arrival=t))
    return out


def spec_system(spec):
    """(expr (n, n, n, 8), rho, f) of a workload entry."""
    e, rho, f = system(spec["n"], spec["nspheres"], spec["rmin"], spec["rmax"], spec["ratio"],
                       spec["seed"])
    n = spec["n"]
    return e.reshape(n, n, n, 8), rho, f
