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

oracle/_ref/ref_blocks_driver runs the same modules on a rank split into
several blocks that share one lead (t.linear's --mesh 32 --block 16: 8
blocks), serving halos, block-order reductions and lead broadcasts between
stages as the reference's local backend does.  On the CPU it is checked
against the one-block harness; on the GPU the "hip" module with 8 blocks
must give the one-block result bit for bit (the lead solves the gathered
rank-wide system either way), keep its coefficients resident across Solve
calls when only e7 changes, and solve wall-bounded systems.
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
BLK_DRV = os.path.join(ROOT, "oracle", "_ref", "ref_blocks_driver")

needs_hip_driver = pytest.mark.skipif(not os.path.exists(HIP_DRV),
                                      reason="oracle/_ref/ref_hip_driver not built")
needs_blocks_driver = pytest.mark.skipif(not os.path.exists(BLK_DRV),
                                         reason="oracle/_ref/ref_blocks_driver not built")


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


def _run_blocks(name, E, tmp, bs, tol, miniter, maxiter, x0=None, extra=()):
    nz, ny, nx = E.shape[:3]
    e = os.path.join(tmp, "e.bin")
    np.ascontiguousarray(E, np.float64).tofile(e)
    x0p = "-"
    if x0 is not None:
        x0p = os.path.join(tmp, "x0.bin")
        np.ascontiguousarray(x0, np.float64).tofile(x0p)
    out = os.path.join(tmp, f"xb_{name}.bin")
    cmd = [BLK_DRV, name, str(nx), str(ny), str(nz), str(bs), str(bs), str(bs), repr(tol),
           str(miniter), str(maxiter), "0", e, x0p, out, *extra]
    print("run:", " ".join(cmd[:11]), *extra, flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        return None, r.stderr
    kv = dict(t.split("=") for t in r.stdout.split())
    return np.fromfile(out).reshape(nz, ny, nx), {k: float(v) for k, v in kv.items()}


def _walled(n=32):
    """t.linear's system with every link across the domain boundary removed
    (zero coefficient, diagonal without it), as the projection assembles
    walls; e7 keeps u* an exact solution."""
    E, ue, h = O.tlinear_system(n)
    for q, (ax, side) in enumerate([(2, 0), (2, -1), (1, 0), (1, -1), (0, 0), (0, -1)]):
        sl = [slice(None)] * 3
        sl[ax] = side
        sl = tuple(sl)
        E[sl + (0,)] += E[sl + (1 + q,)]
        E[sl + (1 + q,)] = 0.0
    E[..., 7] = 0.0
    E[..., 7] = -O.apply(E, ue)
    return E, ue, h


@needs_blocks_driver
@pytest.mark.parametrize("name", ["conjugate", "jacobi"])
def test_blocks_harness_against_one_block(tmp_path, name):
    """CPU: the harness with one block reproduces the one-block harness bit
    for bit; with 8 blocks Jacobi (max-reductions only) is bit-identical and
    CG sums its dot products per block, in block order (same iteration count,
    solutions within round-off)."""
    E = O.tlinear_system(32)[0]
    x1, it1, r1 = _run(DRV, name, E, str(tmp_path), 1e-5, 0, 1000)
    xb1, kv1 = _run_blocks(name, E, str(tmp_path), 32, 1e-5, 0, 1000)
    assert kv1["blocks"] == 1 and kv1["iter"] == it1 and kv1["res"] == r1
    assert np.array_equal(xb1, x1)
    xb8, kv8 = _run_blocks(name, E, str(tmp_path), 16, 1e-5, 0, 1000)
    assert kv8["blocks"] == 8 and kv8["iter"] == it1
    if name == "jacobi":
        assert kv8["res"] == r1 and np.array_equal(xb8, x1)
    else:
        assert abs(kv8["res"] - r1) <= 1e-9 * r1
        assert np.abs(xb8 - x1).max() <= 1e-12


@needs_blocks_driver
def test_blocks_harness_walls_and_repeat(tmp_path):
    """CPU: a wall-bounded system (no link across the boundary) solved with
    non-periodic blocks equals the periodic solve (the wrapped neighbours
    carry zero coefficients); a repeated solve with e7 doubled equals the
    one-block solve of that system."""
    E, ue, h = _walled()
    xp, kp = _run_blocks("conjugate", E, str(tmp_path), 16, 1e-7, 0, 2000)
    xw, kw = _run_blocks("conjugate", E, str(tmp_path), 16, 1e-7, 0, 2000,
                         extra=("harness_periodic=0",))
    assert kp["iter"] == kw["iter"] and np.array_equal(xp, xw)
    assert np.abs(xw - ue).max() < 1e-4
    x2, k2 = _run_blocks("conjugate", E, str(tmp_path), 16, 1e-7, 0, 2000,
                         extra=("harness_repeat=2",))
    E2 = E.copy()
    E2[..., 7] *= 2.0
    x1, it1, r1 = _run_blocks("conjugate", E2, str(tmp_path), 16, 1e-7, 0, 2000)[0], None, None
    assert np.array_equal(x2, x1)


@needs_blocks_driver
@pytest.mark.gpu
@pytest.mark.parametrize("method", ["conjugate", "pcg"])
def test_hip_module_eight_blocks_equal_one(tmp_path, method):
    """The "hip" module on 8 blocks (t.linear's --mesh 32 --block 16) gives
    the one-block result bit for bit: each block's gather into the rank-wide
    field, the lead's single solve and the scatter back are exact copies."""
    E = O.tlinear_system(32)[0]
    extra = (f"linsolver_hip_method={method}",)
    if method == "pcg":
        extra += ("linsolver_hip_precision=mixed", "linsolver_hip_true_relres=1")
    tol = 1e-5 if method == "conjugate" else 1e-9
    x8, k8 = _run_blocks("hip", E, str(tmp_path), 16, tol, 0, 1000, extra=extra)
    assert x8 is not None, k8
    x1, k1 = _run_blocks("hip", E, str(tmp_path), 32, tol, 0, 1000, extra=extra)
    assert x1 is not None, k1
    print(f"  hip {method}: 8 blocks iter={k8['iter']:.0f} res={k8['res']:.6e}; "
          f"1 block iter={k1['iter']:.0f}", flush=True)
    assert k8["blocks"] == 8 and k1["blocks"] == 1
    assert k8["iter"] == k1["iter"] and k8["res"] == k1["res"]
    assert np.array_equal(x8, x1)
    if method == "conjugate":
        # the reference's own module on the same 8 blocks
        xr, kr = _run_blocks("conjugate", E, str(tmp_path), 16, tol, 0, 1000)
        assert kr["iter"] == k8["iter"]
        assert np.abs(xr - x8).max() <= 1e-9
    else:
        assert O.true_relres(E, x8) <= 1e-9 * (1 + 1e-6)


@needs_blocks_driver
@pytest.mark.gpu
def test_hip_module_keeps_coefficients_resident(tmp_path):
    """Two Solve calls on 8 blocks: e7 alone changes -> one system upload and
    one constant-term upload, and the result equals a fresh solve of the new
    system; coefficients change -> two system uploads; reuse off -> two."""
    E = O.tlinear_system(32)[0]
    base = ("linsolver_hip_method=pcg", "linsolver_hip_precision=mixed",
            "linsolver_hip_true_relres=1")
    x2, k2 = _run_blocks("hip", E, str(tmp_path), 16, 1e-9, 0, 1000,
                         extra=base + ("harness_repeat=2",))
    assert x2 is not None, k2
    assert (k2["uploads"], k2["rhs_uploads"]) == (1, 1)
    E2 = E.copy()
    E2[..., 7] *= 2.0
    xf, kf = _run_blocks("hip", E2, str(tmp_path), 16, 1e-9, 0, 1000, extra=base)
    assert kf["iter"] == k2["iter"] and np.array_equal(x2, xf)
    xc, kc = _run_blocks("hip", E, str(tmp_path), 16, 1e-9, 0, 1000,
                         extra=base + ("harness_repeat=2", "harness_coef_scale=3.0"))
    assert (kc["uploads"], kc["rhs_uploads"]) == (2, 0)
    E3 = E.copy()
    E3[..., :7] *= 3.0
    E3[..., 7] *= 2.0
    assert O.true_relres(E3, xc) <= 1e-9 * (1 + 1e-6)
    _, ko = _run_blocks("hip", E, str(tmp_path), 16, 1e-9, 0, 1000,
                        extra=base + ("harness_repeat=2", "linsolver_hip_reuse=0"))
    assert (ko["uploads"], ko["rhs_uploads"]) == (2, 0)


@needs_blocks_driver
@pytest.mark.gpu
@pytest.mark.parametrize("method", ["conjugate", "pcg"])
def test_hip_module_wall_bounded(tmp_path, method):
    """Non-periodic blocks: a system without links across the domain
    boundary (walls) is solved, matching the reference's CG; a system that
    keeps such a link is refused."""
    E, ue, h = _walled()
    extra = ("harness_periodic=0", f"linsolver_hip_method={method}")
    if method == "pcg":
        extra += ("linsolver_hip_precision=fp64", "linsolver_hip_true_relres=1")
    xh, kh = _run_blocks("hip", E, str(tmp_path), 16, 1e-9 if method == "pcg" else 1e-7, 0,
                         3000, extra=extra)
    assert xh is not None, kh
    if method == "conjugate":
        xr, kr = _run_blocks("conjugate", E, str(tmp_path), 16, 1e-7, 0, 3000,
                             extra=("harness_periodic=0",))
        assert kh["iter"] == kr["iter"]
        assert np.abs(xh - xr).max() <= 1e-9
    else:
        assert O.true_relres(E, xh) <= 1e-9 * (1 + 1e-6)
    P = O.tlinear_system(32)[0]  # periodic links kept
    xbad, err = _run_blocks("hip", P, str(tmp_path), 16, 1e-7, 0, 3000, extra=extra)
    assert xbad is None and "non-periodic" in err
