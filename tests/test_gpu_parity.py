"""GPU parity of the HIP path against the CPU oracle (which restates the
reference's SolverConjugate / SolverJacobi bit for bit; see test_oracle.py).

Tolerances: CONJUGATE and JACOBI keep the reference's per-cell operation order
without FMA, so Jacobi (max-reductions only) is bit-exact and CG differs only
through the order of the dot-product sums (iteration counts equal, solutions
within 1e-9 absolute).  PCG is checked on the FP64 true relative residual
(<= 1e-8, recomputed by the oracle) and against an FP64 oracle solution.
"""
import numpy as np
import pytest

from aphros_amd import _lib as L
from aphros_amd import synth
from aphros_amd.solver import Poisson
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def tlin32():
    return O.tlinear_system(32)


@pytest.fixture(scope="module")
def c1():
    e, rho, f = synth.config("c1_64")
    return e.reshape(64, 64, 64, 8)


def test_conjugate_tlinear_matches_oracle(tlin32):
    E, ue, h = tlin32
    _say("conjugate t.linear 32^3")
    xo, ito, reso, _ = O.conjugate(E, h ** 3, tol=1e-5, miniter=0, maxiter=1000)
    s = Poisson(E.shape[:3], method=L.CONJUGATE, precision=L.FP64, resmode=L.RES_L2,
                tol=1e-5, miniter=0, maxiter=1000, h3=h ** 3)
    s.set_system(E)
    x, info = s.solve(None)
    assert info["iter"] == ito
    assert abs(info["residual"] - reso) <= 1e-6 * reso
    assert np.abs(x - xo).max() < 1e-9
    # the reference test's own rule (src/test/linear/test:29)
    assert O.max_diff_mean_removed(x, ue) < 2 * 6.926950e-03 + 1e-12


def test_jacobi_tlinear_bitwise(tlin32):
    E, ue, h = tlin32
    _say("jacobi t.linear 32^3")
    xo, ito, reso, _ = O.jacobi(E, tol=1e-5, miniter=0, maxiter=1000)
    s = Poisson(E.shape[:3], method=L.JACOBI, precision=L.FP64, resmode=L.RES_L2,
                tol=1e-5, miniter=0, maxiter=1000, h3=h ** 3)
    s.set_system(E)
    x, info = s.solve(None)
    assert info["iter"] == ito
    assert info["residual"] == reso
    assert np.array_equal(x, xo)
    assert O.max_diff_mean_removed(x, ue) < 2 * 8.284385e-03 + 1e-12


def test_conjugate_sphere_history(c1):
    """Fixed iteration counts (tol 0) on the 64^3 sphere system: every capped
    run ends where the oracle's does."""
    h3 = (1 / 64) ** 3
    for maxiter in (0, 1, 7, 50):
        _say(f"conjugate c1 maxiter={maxiter}")
        xo, ito, reso, _ = O.conjugate(c1, h3, tol=0, miniter=0, maxiter=maxiter)
        s = Poisson(c1.shape[:3], method=L.CONJUGATE, resmode=L.RES_L2, tol=0, miniter=0,
                    maxiter=maxiter, h3=h3)
        s.set_system(c1)
        x, info = s.solve(None)
        assert info["iter"] == ito == maxiter + 1
        assert abs(info["residual"] - reso) <= 1e-9 * reso
        assert np.abs(x - xo).max() <= 1e-10 * max(1.0, np.abs(xo).max())


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
@pytest.mark.parametrize("precond", [L.PRECOND_JACOBI, L.PRECOND_MG])
def test_pcg_c1_true_residual(c1, precision, precond):
    _say(f"pcg c1 precision={precision} precond={precond}")
    s = Poisson(c1.shape[:3], method=L.PCG, precision=precision, resmode=L.RES_REL_TRUE,
                tol=1e-8, maxiter=20000, precond=precond)
    s.set_system(c1)
    x, info = s.solve(None)
    assert info["converged"] == 1 and info["status"] == 0
    assert info["true_relres"] <= 1e-8
    # independent FP64 recomputation on the CPU
    assert O.true_relres(c1, x) <= 1.0e-8 * (1 + 1e-6)
    assert abs(x.mean()) < 1e-12
    xo, ito, rro = O.pcg_jacobi(c1, tol=1e-11)
    err = np.abs(x - xo).max() / np.abs(xo).max()
    _say(f"  iter={info['iter']} oracle_iter={ito} rel_err={err:.3e} ru={info['refinements']}")
    assert err < 1e-5


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
@pytest.mark.parametrize("precond", [L.PRECOND_JACOBI, L.PRECOND_MG])
def test_pcg_zero_start_equals_explicit_zero_guess(c1, precision, precond):
    """Without an initial guess the solve starts from zero-filled vectors with
    the residual b itself (no pass over x or the FP64 coefficients); an
    explicit all-zero guess is uploaded and takes the full start (x64
    accumulated, b - A x in FP64).  A x is exactly zero, so both runs have the
    same iterates: iteration counts, reliable updates, residuals and the
    solution, bit for bit."""
    _say(f"pcg zero start precision={precision} precond={precond}")
    s = Poisson(c1.shape[:3], method=L.PCG, precision=precision, resmode=L.RES_REL_TRUE,
                tol=1e-8, maxiter=20000, precond=precond)
    s.set_system(c1)
    xa, ia = s.solve(None)
    xb, ib = s.solve(np.zeros(c1.shape[:3]))
    s.close()
    _say(f"  iter={ia['iter']} / {ib['iter']} relres={ia['true_relres']:.6e} / "
         f"{ib['true_relres']:.6e}")
    for k in ("iter", "refinements", "true_relres", "residual", "converged", "status"):
        assert ia[k] == ib[k], k
    assert ia["converged"] == 1
    assert np.array_equal(xa, xb)


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
@pytest.mark.parametrize("case", ["c1", "c1_mass", "c1_ulp"])
def test_pcg_residual_diagonal_formed_or_loaded(c1, precision, case):
    """The FP64 true residual of the reliable updates forms e0 from the six
    neighbour coefficients when every e0 equals -(e1 + ... + e6), summed in
    neighbour order, bit for bit (the assembled c1 system), and loads e0
    otherwise: c1 with a mass term on the diagonal, and c1 with one diagonal
    entry one ulp larger (still singular, no mass term).  Either way the
    residual the solver reports is the oracle's recomputation of b - A x."""
    e = c1
    if case == "c1_mass":
        e = _with_mass(c1)
    elif case == "c1_ulp":
        e = c1.copy()
        e[3, 5, 7, 0] = np.nextafter(e[3, 5, 7, 0], np.inf)
    d = e[..., 0]
    formed = -(((((e[..., 1] + e[..., 2]) + e[..., 3]) + e[..., 4]) + e[..., 5]) + e[..., 6])
    assert np.array_equal(d, formed) == (case == "c1")
    _say(f"pcg residual diagonal {case} precision={precision}")
    s = Poisson(e.shape[:3], method=L.PCG, precision=precision, resmode=L.RES_REL_TRUE,
                tol=1e-8, maxiter=20000, precond=L.PRECOND_MG)
    s.set_system(e)
    x, info = s.solve(None)
    s.close()
    rel = O.true_relres(e, x)
    _say(f"  iter={info['iter']} ru={info['refinements']} relres={rel:.6e} "
         f"(gpu {info['true_relres']:.6e})")
    assert info["converged"] == 1 and info["status"] == 0
    assert rel <= 1e-8 * (1 + 1e-6)
    assert abs(rel - info["true_relres"]) <= 1e-3 * rel


@pytest.mark.parametrize("precond", [L.PRECOND_JACOBI, L.PRECOND_MG])
def test_pcg_frequent_reliable_updates(c1, precond):
    """delta = 0.999: a reliable update whenever the recursive residual falls
    below its maximum since the last one (a third or more of the iterations),
    so updates often follow each other with no x^ write in between (x^ stale,
    counted as zero, in the accumulation of the second).  The solve still
    reaches the oracle's 1e-8 true residual and the oracle's solution."""
    s = Poisson(c1.shape[:3], method=L.PCG, precision=L.MIXED, resmode=L.RES_REL_TRUE,
                tol=1e-8, maxiter=20000, precond=precond, delta=0.999)
    s.set_system(c1)
    x, info = s.solve(None)
    s.close()
    rel = O.true_relres(c1, x)
    _say(f"  precond={precond} iter={info['iter']} ru={info['refinements']} relres={rel:.3e}")
    assert info["converged"] == 1
    assert info["refinements"] >= info["iter"] // 3
    assert rel <= 1e-8 * (1 + 1e-6)
    xo, _, _ = O.pcg_jacobi(c1, tol=1e-11)
    assert np.abs(x - xo).max() / np.abs(xo).max() < 1e-5


def test_rhs_norm_follows_set_rhs(c1):
    """||b|| of the relative stop test is kept between solves and recomputed
    after set_rhs.  With b scaled by 1024 (exact in binary floating point)
    every iterate scales exactly: the same iterations and relative residual,
    and 1024 times the solution, bit for bit; a stale ||b|| would stop the
    solve 1024 times too late.  Back to the first b, the first solve again."""
    s = Poisson(c1.shape[:3], method=L.PCG, precision=L.MIXED, resmode=L.RES_REL_TRUE,
                tol=1e-8, maxiter=20000, precond=L.PRECOND_MG)
    s.set_system(c1)
    x1, i1 = s.solve(None)
    x1b, i1b = s.solve(None)
    e2 = c1.copy()
    e2[..., 7] *= 1024.0
    s.set_rhs(e2[..., 7])
    x2, i2 = s.solve(None)
    s.set_rhs(c1[..., 7])
    x3, i3 = s.solve(None)
    s.close()
    _say(f"  iter {i1['iter']} {i2['iter']} {i3['iter']} relres {i1['true_relres']:.6e} "
         f"{i2['true_relres']:.6e}")
    assert i1["converged"] == 1 and O.true_relres(c1, x1) <= 1e-8 * (1 + 1e-6)
    assert np.array_equal(x1, x1b) and np.array_equal(x1, x3)
    for k in ("iter", "refinements", "true_relres"):
        assert i1[k] == i2[k] == i3[k], k
    assert np.array_equal(x2, 1024.0 * x1)


@pytest.mark.parametrize("precond", [L.PRECOND_JACOBI, L.PRECOND_MG])
def test_pcg_tlinear_negative_diagonal(tlin32, precond):
    """e0 < 0 convention (t.linear / SolvePoisson)."""
    E, ue, h = tlin32
    s = Poisson(E.shape[:3], method=L.PCG, precision=L.MIXED, resmode=L.RES_REL_TRUE,
                tol=1e-10, maxiter=5000, h3=h ** 3, precond=precond)
    s.set_system(E)
    x, info = s.solve(None)
    assert info["converged"] == 1
    assert O.true_relres(E, x) <= 1e-10 * (1 + 1e-6)
    assert O.max_diff_mean_removed(x, ue) < 1e-7


def _with_mass(e, frac=0.05, seed=4):
    e = e.copy()
    rng = np.random.default_rng(seed)
    e[..., 0] += np.sign(e[..., 0]) * frac * np.abs(e[..., 0]).mean() * rng.random(e.shape[:3])
    return e


@pytest.mark.parametrize("precision,tol", [(L.FP64, 1e-12), (L.MIXED, 1e-4)])
@pytest.mark.parametrize("case,w", [("c1", 0), ("tlin_negative", 0), ("c1_mass", 0), ("c1", 2),
                                    ("c1_mass", 9)])
def test_mg_vcycle_matches_oracle(c1, tlin32, precision, tol, case, w):
    """One cycle on the GPU against the oracle's (FP64: rounding only; MIXED:
    FP32 arithmetic and bfloat16 coefficients restated by the oracle), on the
    same residual; w levels with two coarse corrections (9: every level that
    can, i.e. 4 of the 6 levels of 64^3)."""
    e = {"c1": c1, "tlin_negative": tlin32[0], "c1_mass": _with_mass(c1)}[case]
    _say(f"vcycle {case} precision={precision} w={w}")
    s = Poisson(e.shape[:3], method=L.PCG, precision=precision, resmode=L.RES_REL_TRUE,
                tol=1e-8, maxiter=20000, precond=L.PRECOND_MG, mg_w=w)
    s.set_system(e)
    rng = np.random.default_rng(12)
    r = rng.standard_normal(e.shape[:3])
    r -= r.mean()
    z = s.precond_apply(r)
    # MIXED stores the V-cycle coefficients in bfloat16
    zo, nlev = O.mg_apply(e, r, bf16=precision == L.MIXED, w=w)
    err = np.abs(z - zo).max() / np.abs(zo).max()
    _say(f"  nlev={nlev} rel_err={err:.3e}")
    assert err <= tol
    # the solve state is untouched: a solve after it matches a fresh handle
    x, info = s.solve(None)
    s2 = Poisson(e.shape[:3], method=L.PCG, precision=precision, resmode=L.RES_REL_TRUE,
                 tol=1e-8, maxiter=20000, precond=L.PRECOND_MG, mg_w=w)
    s2.set_system(e)
    x2, info2 = s2.solve(None)
    assert info["iter"] == info2["iter"] and np.array_equal(x, x2)


@pytest.fixture(scope="module")
def c2():
    e, rho, f = synth.config("c2_256")
    return e.reshape(256, 256, 256, 8)


@pytest.mark.parametrize("precision,tol", [(L.FP64, 1e-12), (L.MIXED, 1e-4)])
@pytest.mark.parametrize("w", [0, 3])
def test_mg_vcycle_large_fused_post(c2, precision, tol, w):
    """256^3 (2^24 cells): the level-0 post-smoothing runs as one fused kernel
    (whole-row tiles, halo rows and planes recomputed, z[0] -> z[1]); one
    cycle against the oracle's, as for the small systems."""
    _say(f"vcycle c2_256 precision={precision} w={w}")
    s = Poisson(c2.shape[:3], method=L.PCG, precision=precision, resmode=L.RES_REL_TRUE,
                tol=1e-8, maxiter=20000, precond=L.PRECOND_MG, mg_w=w)
    s.set_system(c2)
    rng = np.random.default_rng(13)
    r = rng.standard_normal(c2.shape[:3])
    r -= r.mean()
    z = s.precond_apply(r)
    zo, nlev = O.mg_apply(c2, r, bf16=precision == L.MIXED, w=w)
    err = np.abs(z - zo).max() / np.abs(zo).max()
    _say(f"  nlev={nlev} rel_err={err:.3e}")
    assert err <= tol


def test_mg_vcycle_fused_noncubic_mass():
    """A 256 x 256 x 320 (z, y, x) system with a mass term: fused level-0
    kernels with rows of 80 threads (12 thread rows per tile, 10 output rows,
    a partial last tile in y) and the mass term in both fused kernels."""
    shape = (256, 256, 320)
    rng = np.random.default_rng(21)
    rho = np.where(rng.random(shape) < 0.2, 100.0, 1.0)
    f = rng.standard_normal(shape)
    f -= f.mean()
    e = _with_mass(O.assemble(rho, f))
    _say("vcycle 320x256x256 mass, mixed, w=3")
    s = Poisson(shape, method=L.PCG, precision=L.MIXED, resmode=L.RES_REL_TRUE, tol=1e-8,
                maxiter=20000, precond=L.PRECOND_MG, mg_w=3)
    s.set_system(e)
    r = rng.standard_normal(shape)
    z = s.precond_apply(r)
    zo, nlev = O.mg_apply(e, r, bf16=True, w=3)
    err = np.abs(z - zo).max() / np.abs(zo).max()
    _say(f"  nlev={nlev} rel_err={err:.3e}")
    assert err <= 1e-4


def test_mg_vcycle_512_fused_w_levels():
    """The benchmark system at 512^3 with 3 W levels: fused kernels on level 0
    and on level 1 (256^3), whose second coarse correction ends with the
    fused post-smoothing adding the first correction."""
    e, _, _ = synth.config("c3_512")
    e = e.reshape(512, 512, 512, 8)
    _say("vcycle c3_512 mixed w=3")
    s = Poisson(e.shape[:3], method=L.PCG, precision=L.MIXED, resmode=L.RES_REL_TRUE,
                tol=1e-8, maxiter=20000, precond=L.PRECOND_MG, mg_w=3)
    s.set_system(e)
    rng = np.random.default_rng(14)
    r = rng.standard_normal(e.shape[:3])
    r -= r.mean()
    z = s.precond_apply(r)
    s.close()
    zo, nlev = O.mg_apply(e, r, bf16=True, w=3)
    err = np.abs(z - zo).max() / np.abs(zo).max()
    _say(f"  nlev={nlev} rel_err={err:.3e}")
    assert err <= 1e-4


def test_pcg_large_fused_post_true_residual(c2):
    """The fused level-0 post-smoothing also forms (r, z) for beta: a mixed
    multigrid PCG solve at 256^3 reaches the FP64 true residual (recomputed on
    the CPU) within the oracle's FP64 multigrid PCG iteration count + 2."""
    s = Poisson(c2.shape[:3], method=L.PCG, precision=L.MIXED, resmode=L.RES_REL_TRUE,
                tol=1e-8, maxiter=20000, precond=L.PRECOND_MG, mg_w=L.MG_W_DEFAULT)
    s.set_system(c2)
    x, info = s.solve(None)
    assert info["converged"] == 1 and info["status"] == 0
    rel = O.true_relres(c2, x)
    _, ito, _, _ = O.mgpcg(c2, tol=1e-8, w=L.MG_W_DEFAULT)
    _say(f"  iter={info['iter']} oracle_iter={ito} relres={rel:.3e}")
    assert rel <= 1e-8 * (1 + 1e-6)
    assert info["iter"] <= ito + 2


@pytest.mark.parametrize("name,w", [("c1", 0), ("c2", 1), ("s384", 3)])
def test_mg_w_auto_picks_by_size(request, name, w):
    """The default W levels (APHPOIS_MG_W_AUTO) are the engine's choice by
    cell count: 0 at 64^3, 1 at 256^3, 3 at 384^3 (2^25.75 cells); the
    default solve equals the solve with that mg_w bit for bit."""
    if name == "s384":
        e = synth.system(384, 32, 0.08, 0.15, 1000.0, 2)[0].reshape(384, 384, 384, 8)
    else:
        e = request.getfixturevalue(name)
    assert L.mg_w_auto(e.shape[0] * e.shape[1] * e.shape[2]) == w
    out = []
    for mw in (L.MG_W_AUTO, w):
        s = Poisson(e.shape[:3], method=L.PCG, precision=L.MIXED, resmode=L.RES_REL_TRUE,
                    tol=1e-8, maxiter=20000, precond=L.PRECOND_MG, mg_w=mw)
        s.set_system(e)
        x, info = s.solve(None)
        s.close()
        assert info["converged"] == 1
        out.append((x, info["iter"]))
    _say(f"  {name}: auto iter={out[0][1]} w={w} iter={out[1][1]}")
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])


@pytest.mark.parametrize("w", [0, 2])
def test_mg_fp64_iterations_match_oracle(c1, w):
    """FP64 multigrid PCG on the GPU and the oracle's FP64 multigrid PCG (same
    preconditioner, sums in a different order): equal iteration counts up to
    the stop test's granularity."""
    s = Poisson(c1.shape[:3], method=L.PCG, precision=L.FP64, resmode=L.RES_REL_TRUE,
                tol=1e-8, maxiter=20000, precond=L.PRECOND_MG, mg_w=w)
    s.set_system(c1)
    x, info = s.solve(None)
    xo, ito, rro, _ = O.mgpcg(c1, tol=1e-8, w=w)
    _say(f"  gpu iter={info['iter']} oracle iter={ito}")
    assert abs(info["iter"] - ito) <= 2
    err = np.abs(x - xo).max() / np.abs(xo).max()
    assert err < 1e-6


def test_warm_start_alias_and_setconf(c1):
    s = Poisson(c1.shape[:3], method=L.PCG, precision=L.MIXED, resmode=L.RES_REL_TRUE,
                tol=1e-6, maxiter=20000)
    s.set_system(c1)
    x, info1 = s.solve(None)
    s.set_conf(1e-8, 0, 20000)
    x2, info2 = s.solve(x, out=x)  # init aliases sol
    assert x2 is x
    assert info2["converged"] == 1 and info2["true_relres"] <= 1e-8
    assert info2["iter"] < info1["iter"] * 2


def test_errors(c1):
    s = Poisson(c1.shape[:3], method=L.PCG)
    with pytest.raises(L.AphpoisError) as ei:
        s.solve(None)
    assert ei.value.code == L.ESTATE
    bad = c1.copy()
    bad[3, 4, 5, 2] *= 1.5  # breaks symmetry
    with pytest.raises(L.AphpoisError) as ei:
        s.set_system(bad)
    assert ei.value.code == L.EINVAL
    bad = c1.copy()
    bad[1, 2, 3, 7] = np.nan
    with pytest.raises(L.AphpoisError) as ei:
        s.set_system(bad)
    assert ei.value.code == L.ENUMERIC
    with pytest.raises(L.AphpoisError) as ei:
        Poisson(c1.shape[:3], method=L.PCG, mem_budget=1 << 20)
    assert ei.value.code == L.EBUDGET
