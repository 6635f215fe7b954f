"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

1. Against the reference itself: vectors written by tests/golden/make_golden.py
   from the reference's own SolverConjugate / SolverJacobi templates
   (src/linear/linear.ipp, compiled from the reference sources into
   oracle/_ref/ref_driver).  The oracle reproduces them bit for bit: iteration
   counts, residuals, per-iteration residual history and solution vectors.
2. Against the reference test's own rule on its own system
   (src/test/linear/test:29, src/test/linear/main.cpp:45-128).
3. Against mathematics that fixes the answer independently: FFT closed form
   for constant coefficients, a dense least-squares solve on a tiny sphere
   system, CG finite termination, the Jacobi fixed point and its Fourier
   amplification factor, and closed forms of the assembly.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_fixtures.npz")


@pytest.fixture(scope="module")
def G():
    return np.load(GOLD, allow_pickle=False)


# ----------------------------------------------------- reference fixtures


@pytest.mark.parametrize("name", ["conjugate", "jacobi"])
def test_tlin16_bitwise(G, name):
    E = G["tlin16_expr"]
    h3 = (1.0 / 16) ** 3
    if name == "conjugate":
        x, it, res, _ = O.conjugate(E, h3, tol=1e-5, miniter=0, maxiter=1000)
    else:
        x, it, res, _ = O.jacobi(E, tol=1e-5, miniter=0, maxiter=1000)
    assert it == int(G[f"tlin16_{name}_iter"])
    assert res == float(G[f"tlin16_{name}_res"])
    assert np.array_equal(x, G[f"tlin16_{name}_x"])


@pytest.mark.parametrize("name", ["conjugate", "jacobi"])
def test_tlin32_matches_reference(G, name):
    E, ue, h = O.tlinear_system(32)
    if name == "conjugate":
        x, it, res, _ = O.conjugate(E, h ** 3, tol=1e-5, miniter=0, maxiter=1000)
    else:
        x, it, res, _ = O.jacobi(E, tol=1e-5, miniter=0, maxiter=1000)
    assert it == int(G[f"tlin32_{name}_iter"])
    assert res == float(G[f"tlin32_{name}_res"])
    flat = x.reshape(-1)
    s = np.array([flat.sum(), (flat * flat).sum(), flat.max(), flat.min()])
    assert np.array_equal(s, G[f"tlin32_{name}_sum"])
    assert np.array_equal(flat[::97], G[f"tlin32_{name}_sample"])


def test_sphere16_history_and_solution(G):
    E = G["sph16_expr"]
    h3 = (1.0 / 16) ** 3
    _, it, _, hist = O.conjugate(E, h3, tol=0.0, miniter=0, maxiter=39, hist=40)
    assert it == 40
    assert np.array_equal(hist, G["sph16_conjugate_hist"])
    x, it, res, _ = O.conjugate(E, h3, tol=1e-10, miniter=0, maxiter=5000)
    assert it == int(G["sph16_conjugate_iter"]) and res == float(G["sph16_conjugate_res"])
    assert np.array_equal(x, G["sph16_conjugate_x"])
    x, it, res, _ = O.jacobi(E, tol=1e-9, miniter=0, maxiter=100000)
    assert it == int(G["sph16_jacobi_iter"]) == 100001  # maxiter rule: iter > maxiter
    assert res == float(G["sph16_jacobi_res"])
    assert np.array_equal(x, G["sph16_jacobi_x"])


def test_warm_start_and_maxnorm(G):
    E = G["tlin16_expr"]
    h3 = (1.0 / 16) ** 3
    x, it, res, _ = O.conjugate(E, h3, x0=G["tlin16_x0"], tol=1e-6, miniter=0, maxiter=1000)
    assert it == int(G["tlin16_warm_iter"]) and res == float(G["tlin16_warm_res"])
    assert np.array_equal(x, G["tlin16_warm_x"])
    _, it, res, _ = O.conjugate(E, h3, tol=1e-4, miniter=0, maxiter=1000, maxnorm=True)
    assert it == int(G["tlin16_maxnorm_iter"]) and res == float(G["tlin16_maxnorm_res"])


def test_c1_64_matches_reference(G):
    from aphros_amd import synth
    e, _, _ = synth.config("c1_64")
    E = e.reshape(64, 64, 64, 8)
    x, it, res, _ = O.conjugate(E, (1 / 64) ** 3, tol=1e-7, miniter=0, maxiter=100000)
    assert it == int(G["c1_64_conjugate_iter"]) and res == float(G["c1_64_conjugate_res"])
    flat = x.reshape(-1)
    s = np.array([flat.sum(), (flat * flat).sum(), flat.max(), flat.min()])
    assert np.array_equal(s, G["c1_64_conjugate_sum"])


@pytest.mark.skipif(not O.ref_available(), reason="reference harness not built here")
@pytest.mark.parametrize("seed", [3, 4])
def test_live_reference_random_systems(seed):
    """Random 8^3 systems with a random initial guess, run live through the
    reference's own solvers."""
    n = 8
    rng = np.random.default_rng(seed)
    rho = np.where(rng.random((n, n, n)) < 0.3, 1.0, 50.0)
    E = O.assemble(rho, rng.uniform(-1, 1, (n, n, n)))
    x0 = rng.uniform(-1, 1, (n, n, n))
    for name in ("conjugate", "jacobi"):
        xr, itr, resr = O.ref_solve(name, E, 1e-7, 3, 400, x0=x0)
        if name == "conjugate":
            x, it, res, _ = O.conjugate(E, (1 / n) ** 3, x0=x0, tol=1e-7, miniter=3, maxiter=400)
        else:
            x, it, res, _ = O.jacobi(E, x0=x0, tol=1e-7, miniter=3, maxiter=400)
        assert (it, res) == (itr, resr)
        assert np.array_equal(x, xr)


# ----------------------------------------------------- the reference test's rule


def test_reference_linear_test_rule():
    """src/test/linear/test:29: max_diff_exact < 2 * ref + 1e-12 with the
    values held in src/test/linear/ref/{conjugate,jacobi}/outdiff."""
    E, ue, h = O.tlinear_system(32)
    x, _, _, _ = O.conjugate(E, h ** 3, tol=1e-5, miniter=0, maxiter=1000)
    assert O.max_diff_mean_removed(x, ue) < 2 * 6.926950e-03 + 1e-12
    x, _, _, _ = O.jacobi(E, tol=1e-5, miniter=0, maxiter=1000)
    assert O.max_diff_mean_removed(x, ue) < 2 * 8.284385e-03 + 1e-12


# ----------------------------------------------------- closed forms


def _laplacian_symbol(n):
    k = 2 * np.pi * np.arange(n) / n
    lam1 = 2 - 2 * np.cos(k)
    return lam1[:, None, None] + lam1[None, :, None] + lam1[None, None, :]


def test_constant_density_fft_closed_form():
    n = 24
    rho = np.full((n, n, n), 7.0)
    rng = np.random.default_rng(1)
    f = rng.uniform(-1, 1, (n, n, n))
    f -= f.mean()
    E = O.assemble(rho, f)
    h = 1.0 / n
    kf = h / 7.0
    assert np.allclose(E[..., 2], -kf, rtol=1e-14, atol=0)
    b = f * h ** 3
    lam = _laplacian_symbol(n) * kf
    lam[0, 0, 0] = 1.0
    xh = np.fft.fftn(b) / lam
    xh[0, 0, 0] = 0.0
    x_exact = np.real(np.fft.ifftn(xh))
    x, it, res, _ = O.conjugate(E, h ** 3, tol=1e-14, miniter=0, maxiter=2000)
    x = x - x.mean()
    assert np.abs(x - x_exact).max() <= 1e-9 * np.abs(x_exact).max()
    xp, itp, rr = O.pcg_jacobi(E, tol=1e-12)
    assert np.abs(xp - x_exact).max() <= 1e-9 * np.abs(x_exact).max()
    assert rr <= 1e-12


def _dense(E):
    n = E.shape[0]
    N = n ** 3
    A = np.empty((N, N))
    e = np.zeros(N)
    for j in range(N):
        e[:] = 0
        e[j] = 1
        A[:, j] = O.apply(E, e.reshape(n, n, n)).reshape(-1)
    return A


def test_dense_solve_tiny_sphere_system():
    n = 6
    rho = O.twin_density(n, 2, 0.15, 0.3, 1000.0, 5)
    assert 0 < (rho == 1).sum() < rho.size
    f = O.twin_rhs(n, 5)
    E = O.assemble(rho, f)
    A = _dense(E)
    assert np.array_equal(A, A.T)                     # symmetric
    assert np.abs(A.sum(axis=1)).max() < 1e-15 * np.abs(A).max() * 8  # zero row sums
    b = -E[..., 7].reshape(-1)
    xs = np.linalg.lstsq(A, b, rcond=None)[0]
    xs -= xs.mean()
    x, it, res, _ = O.conjugate(E, (1 / n) ** 3, tol=1e-13, miniter=0, maxiter=5000)
    x = x.reshape(-1) - x.mean()
    assert np.abs(x - xs).max() <= 1e-9 * np.abs(xs).max()
    xp, _, _ = O.pcg_jacobi(E, tol=1e-12)
    assert np.abs(xp.reshape(-1) - xs).max() <= 1e-9 * np.abs(xs).max()
    # Jacobi: the solution is a fixed point of one sweep
    xj, itj, resj, _ = O.jacobi(E, x0=(xs + 0.25).reshape(n, n, n), tol=0, miniter=0, maxiter=0)
    assert itj == 1 and resj < 1e-10 * (np.abs(xs).max() + 0.25)


def test_cg_finite_termination():
    """Constant coefficients on 4^3: the periodic Laplacian has 6 distinct
    non-zero eigenvalues, so exact CG converges in at most 6 iterations."""
    n = 4
    rng = np.random.default_rng(7)
    f = rng.uniform(-1, 1, (n, n, n))
    f -= f.mean()
    E = O.assemble(np.ones((n, n, n)), f)
    _, it, res0, _ = O.conjugate(E, 1.0, tol=0, miniter=0, maxiter=0)
    _, it, res, _ = O.conjugate(E, 1.0, tol=0, miniter=0, maxiter=5)
    assert it == 6
    assert res < 1e-12 * res0


def test_jacobi_fourier_amplification():
    """Constant coefficients, b = 0: a Fourier mode is multiplied by
    g = (cos kx + cos ky + cos kz) / 3 per sweep."""
    n = 16
    E = O.assemble(np.ones((n, n, n)), np.zeros((n, n, n)))
    c = np.arange(n)
    kx, ky, kz = 2 * np.pi * 3 / n, 2 * np.pi * 1 / n, 2 * np.pi * 5 / n
    u0 = np.cos(kx * c)[None, None, :] * np.cos(ky * c)[None, :, None] * \
        np.cos(kz * c)[:, None, None]
    u1, it, _, _ = O.jacobi(E, x0=u0, tol=0, miniter=0, maxiter=0)
    g = (np.cos(kx) + np.cos(ky) + np.cos(kz)) / 3
    assert it == 1
    assert np.abs(u1 - g * u0).max() < 1e-13


def test_assemble_harmonic_face_density():
    n = 8
    rho = np.full((n, n, n), 4.0)
    rho[:, :, 3] = 1000.0
    E = O.assemble(rho, np.zeros((n, n, n)))
    h = 1.0 / n
    # face between x=2 (rho 4) and x=3 (rho 1000)
    k = -E[0, 0, 2, 2]
    assert np.isclose(k, h * 0.5 * (1 / 4.0 + 1 / 1000.0), rtol=1e-14)
    assert E[0, 0, 3, 1] == E[0, 0, 2, 2]             # symmetric pair
    assert np.isclose(E[0, 0, 3, 0], -E[0, 0, 3, 1:7].sum(), rtol=1e-14)
    # SolvePoisson / t.linear sign: everything negated
    En = O.assemble(rho, np.ones((n, n, n)), sign=-1)
    assert np.array_equal(En[..., :7], -E[..., :7])


def test_planted_mistake_is_caught():
    """A dropped neighbour term in the operator breaks the dense symmetry and
    the FFT closed form: the checks above would fail on such an oracle."""
    n = 6
    E = O.assemble(O.twin_density(n, 2, 0.15, 0.3, 1000.0, 5), O.twin_rhs(n, 5))
    bad = E.copy()
    bad[..., 6] = 0.0
    A = _dense(bad)
    assert not np.array_equal(A, A.T)


# ----------------------------------------------------- projection with face fluxes


def _flux_case(n=12, seed=8):
    rng = np.random.default_rng(seed)
    rho = np.where(rng.random((n, n, n)) < 0.4, 1000.0, 1.0)
    ffv = rng.standard_normal((3, n, n, n)) * 1e-3
    src = rng.standard_normal((n, n, n))
    src -= src.mean()  # solvable periodic system
    return rho, ffv, src


def test_assemble_flux_without_fluxes_is_assemble():
    """Zero fluxes, dt = 1: the pinned assembly bit for bit (both signs)."""
    rho, _, src = _flux_case()
    n = rho.shape[0]
    for sign in (1, -1):
        e1 = O.assemble(rho, src, sign=sign)
        e2 = O.assemble_flux(rho, np.zeros((3, n, n, n)), src, cell=1.0 / n, dt=1.0, sign=sign)
        assert np.array_equal(e1, e2)


def test_assemble_flux_constant_term_is_divergence():
    """e7 = net outward flux - src * vol (the divergence of a uniform flux
    field vanishes exactly); coefficients scale with dt."""
    rho, ffv, src = _flux_case()
    n = rho.shape[0]
    h = 1.0 / n
    e = O.assemble_flux(rho, ffv, src, cell=h, dt=0.5)
    div = (np.roll(ffv[0], -1, axis=2) - ffv[0] + np.roll(ffv[1], -1, axis=1) - ffv[1]
           + np.roll(ffv[2], -1, axis=0) - ffv[2])
    assert np.allclose(e[..., 7], div - src * h ** 3, rtol=0, atol=1e-15)
    uni = np.ones((3, n, n, n)) * np.array([0.3, -0.2, 0.7])[:, None, None, None]
    e_u = O.assemble_flux(rho, uni, None, cell=h, dt=0.5)
    assert np.all(np.abs(e_u[..., 7]) <= 1e-16)
    e1 = O.assemble_flux(rho, ffv, src, cell=h, dt=1.0)
    assert np.allclose(e[..., :7], 0.5 * e1[..., :7], rtol=1e-15, atol=0)


def test_flux_correction_is_the_projection():
    """The linear system at p is the divergence of the corrected fluxes minus
    the source: sum_q outward(q) * F_q(p) - src * vol = A p + e7 for any p.
    A sign or index mistake in either function breaks the identity."""
    rho, ffv, src = _flux_case()
    n = rho.shape[0]
    h = 1.0 / n
    e = O.assemble_flux(rho, ffv, src, cell=h, dt=0.7)
    p = np.random.default_rng(2).standard_normal((n, n, n))
    F = O.flux_correct(rho, ffv, p, cell=h, dt=0.7)
    div = (np.roll(F[0], -1, axis=2) - F[0] + np.roll(F[1], -1, axis=1) - F[1]
           + np.roll(F[2], -1, axis=0) - F[2])
    lhs = div - src * h ** 3
    rhs = O.apply(e, p, with_const=True)
    assert np.max(np.abs(lhs - rhs)) <= 1e-12 * np.max(np.abs(rhs))
    # after a solve the corrected fluxes are divergence free up to the source
    x, it, rel = O.pcg_jacobi(e, tol=1e-12)
    F = O.flux_correct(rho, ffv, x, cell=h, dt=0.7)
    div = (np.roll(F[0], -1, axis=2) - F[0] + np.roll(F[1], -1, axis=1) - F[1]
           + np.roll(F[2], -1, axis=0) - F[2])
    assert rel <= 1e-12
    assert np.max(np.abs(div - src * h ** 3)) <= 1e-9 * np.max(np.abs(src * h ** 3))
    # planted mistake (the oracle's convention with the faces' roles swapped)
    Fw = O.flux_correct(rho, ffv, -p, cell=h, dt=0.7)
    divw = (np.roll(Fw[0], -1, axis=2) - Fw[0] + np.roll(Fw[1], -1, axis=1) - Fw[1]
            + np.roll(Fw[2], -1, axis=0) - Fw[2])
    assert np.max(np.abs(divw - src * h ** 3 - O.apply(e, p, with_const=True))) > 1e-6


def test_assemble_faces_is_the_density_form():
    """SIMPLE's form with the harmonic face density passed as fk (and dt = 1)
    is the density form bit for bit; the flux correction likewise."""
    rho, ffv, src = _flux_case()
    n = rho.shape[0]
    inv = 1.0 / rho
    fk = np.empty((3, n, n, n))
    for d, ax in enumerate((2, 1, 0)):
        fk[d] = 1.0 / ((inv + np.roll(inv, 1, axis=ax)) * 0.5)
    e1 = O.assemble_flux(rho, ffv, src, cell=1.0 / n, dt=1.0)
    e2 = O.assemble_flux(None, ffv, src, cell=1.0 / n, dt=1.0, fk=fk)
    assert np.array_equal(e1, e2)
    p = np.random.default_rng(4).standard_normal((n, n, n))
    assert np.array_equal(O.flux_correct(rho, ffv, p, cell=1.0 / n),
                          O.flux_correct(None, ffv, p, cell=1.0 / n, fk=fk))
