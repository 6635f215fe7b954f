"""The drop-in through the reference's own plugin interface.

oracle/_ref/ref_hip_driver is the reference's stage loop (one block covering
the periodic grid) with the solver created through the reference's module
registry, ModuleLinear<M>::GetInstance(name)->Make(var, "symm", m): "conjugate"
and "jacobi" are the reference's own modules, "hip" is the adapter of
aphros_amd/adapter (linear::SolverHip over the aphpois C ABI).  It is built
from the reference's sources by oracle/build_ref.sh where they are present;
the tests skip without it.

CPU: the registry route reproduces the direct construction bit for bit.
GPU: the "hip" module against the reference's modules on the same system:
Jacobi bit for bit, CG with equal iteration counts and solutions within 1e-9
(only the order of the dot-product sums differs), PCG to the true relative
residual.
"""
import os
import subprocess

import numpy as np
import pytest

from aphros_amd import synth
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
HIP_DRV = os.path.join(ROOT, "oracle", "_ref", "ref_hip_driver")

needs_hip_driver = pytest.mark.skipif(not os.path.exists(HIP_DRV),
                                      reason="oracle/_ref/ref_hip_driver not built")


def _run(drv, name, E, tmp, tol, miniter, maxiter, maxnorm=0, x0=None, extra=()):
    nz, ny, nx = E.shape[:3]
    e = os.path.join(tmp, "e.bin")
    np.ascontiguousarray(E, np.float64).tofile(e)
    x0p = "-"
    if x0 is not None:
        x0p = os.path.join(tmp, "x0.bin")
        np.ascontiguousarray(x0, np.float64).tofile(x0p)
    out = os.path.join(tmp, f"x_{name}.bin")
    cmd = [drv, name, str(nx), str(ny), str(nz), repr(tol), str(miniter), str(maxiter),
           str(maxnorm), e, x0p, out, *extra]
    print("run:", " ".join(cmd[:9]), *extra, flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    kv = dict(t.split("=") for t in r.stdout.split())
    return np.fromfile(out).reshape(nz, ny, nx), int(kv["iter"]), float(kv["res"])


@pytest.fixture(scope="module")
def tlin16():
    return O.tlinear_system(16)[0]


@needs_hip_driver
@pytest.mark.parametrize("name", ["conjugate", "jacobi"])
def test_registry_route_matches_direct(tlin16, tmp_path, name):
    x1, it1, r1 = _run(DRV, name, tlin16, str(tmp_path), 1e-10, 1, 150)
    x2, it2, r2 = _run(HIP_DRV, name, tlin16, str(tmp_path), 1e-10, 1, 150)
    assert it1 == it2 and r1 == r2 and np.array_equal(x1, x2)


@pytest.fixture(scope="module")
def c1():
    e, _, _ = synth.config("c1_64")
    return e.reshape(64, 64, 64, 8)


@needs_hip_driver
@pytest.mark.gpu
@pytest.mark.parametrize("case", ["tlinear32", "c1"])
def test_hip_module_conjugate(case, c1, tmp_path):
    # t.linear at its own tolerance; the sphere system capped at 51 iterations
    E = O.tlinear_system(32)[0] if case == "tlinear32" else c1
    tol, maxiter = (1e-5, 1000) if case == "tlinear32" else (0.0, 50)
    xr, itr, rr = _run(HIP_DRV, "conjugate", E, str(tmp_path), tol, 0, maxiter)
    xh, ith, rh = _run(HIP_DRV, "hip", E, str(tmp_path), tol, 0, maxiter)
    print(f"  reference iter={itr} res={rr:.6e}  hip iter={ith} res={rh:.6e}", flush=True)
    assert ith == itr
    assert abs(rh - rr) <= 1e-6 * rr
    assert np.abs(xh - xr).max() <= 1e-9 * max(1.0, np.abs(xr).max())


@needs_hip_driver
@pytest.mark.gpu
def test_hip_module_jacobi_bitwise(tmp_path):
    E = O.tlinear_system(32)[0]
    xr, itr, rr = _run(HIP_DRV, "jacobi", E, str(tmp_path), 1e-8, 1, 300)
    xh, ith, rh = _run(HIP_DRV, "hip", E, str(tmp_path), 1e-8, 1, 300,
                       extra=("linsolver_hip_method=jacobi",))
    assert ith == itr and rh == rr and np.array_equal(xh, xr)


@needs_hip_driver
@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hip_module_pcg_true_residual(c1, tmp_path, precision):
    """Multigrid PCG selected by configuration, stopping on the FP64 true
    relative residual; warm start from the CG solution (fc_init)."""
    x, it, res = _run(HIP_DRV, "hip", c1, str(tmp_path), 1e-8, 0, 1000,
                      extra=("linsolver_hip_method=pcg", f"linsolver_hip_precision={precision}",
                             "linsolver_hip_true_relres=1"))
    rel = O.true_relres(c1, x)
    print(f"  pcg {precision}: iter={it} true_relres={rel:.3e}", flush=True)
    assert rel <= 1e-8 * (1 + 1e-6)
    x2, it2, _ = _run(HIP_DRV, "hip", c1, str(tmp_path), 1e-8, 0, 1000, x0=x,
                      extra=("linsolver_hip_method=pcg", f"linsolver_hip_precision={precision}",
                             "linsolver_hip_true_relres=1"))
    assert it2 <= 2 and O.true_relres(c1, x2) <= 1e-8 * (1 + 1e-6)
