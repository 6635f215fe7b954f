"""Pins the oracle's true relative residual (orc_true_relres), the checker that
every preconditioned, mixed-precision, batched and served GPU test accepts a
solution by.

orc_true_relres(E, x) = ||A x + e7||_2 / ||e7||_2, i.e. ||b - A x|| / ||b||
with b = -e7 (the reference's system reads e0*x + sum_q e[1+q]*x[nb(q)] + e7
= 0, src/linear/linear.h:37-39).  It is checked against what fixes it
independently of the oracle's code:

  * a dense matrix built here with numpy index arithmetic from the Expr
    coefficients (not through the oracle's operator), with numpy's norms;
  * closed forms: x = 0 gives exactly 1; the exact solution of the
    reference's linear-test system gives ~0; scaling the whole system leaves
    it unchanged; scaling x by c on a system with e7 = 0 scales ||A x|| by c;
  * mutants of the formula (a dropped square root, the opposite sign of e7,
    a dropped neighbour term, the norm of x instead of b) fail these checks.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _dense_np(E):
    """A (N x N) from the Expr by index arithmetic: A[c, c] = e0(c),
    A[c, nb_q(c)] += e_{1+q}(c), q = 2*d + side, periodic wrap."""
    nz, ny, nx = E.shape[:3]
    N = nx * ny * nz
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    c = (x + nx * (y + ny * z)).reshape(-1)
    nbs = [((x - 1) % nx, y, z), ((x + 1) % nx, y, z), (x, (y - 1) % ny, z),
           (x, (y + 1) % ny, z), (x, y, (z - 1) % nz), (x, y, (z + 1) % nz)]
    A = np.zeros((N, N))
    e = E.reshape(N, 8)
    A[c, c] += e[c, 0]
    for q, (a, b, d) in enumerate(nbs):
        j = (a + nx * (b + ny * d)).reshape(-1)
        np.add.at(A, (c, j), e[c, 1 + q])
    return A


def _relres_np(E, x):
    A = _dense_np(E)
    b = -E[..., 7].reshape(-1)
    return np.linalg.norm(b - A @ x.reshape(-1)) / np.linalg.norm(b)


def _system(n=6, seed=5, shape=None):
    shape = shape or (n, n, n)
    rng = np.random.default_rng(seed)
    rho = np.where(rng.random(shape) < 0.3, 1.0, 1000.0)
    f = rng.uniform(-1, 1, shape)
    f -= f.mean()
    return O.assemble(rho, f)


# --- mutants of the formula: each must fail at least one check below ----

def _m_no_sqrt(E, x):
    y = O.apply(E, x, with_const=True)
    return float((y * y).sum() / (E[..., 7] ** 2).sum())


def _m_sign(E, x):
    y = O.apply(E, x, with_const=False) - E[..., 7]  # b = +e7
    return float(np.linalg.norm(y) / np.linalg.norm(E[..., 7]))


def _m_drop_zp(E, x):
    Eb = E.copy()
    Eb[..., 6] = 0.0
    return O.true_relres(Eb, x)


def _m_norm_x(E, x):
    y = O.apply(E, x, with_const=True)
    nx = np.linalg.norm(x)
    return float(np.linalg.norm(y) / (nx if nx > 0 else 1.0))


def _checks(fn):
    """The checks of this file as one predicate over a candidate function."""
    ok = True
    for seed, shape in ((5, None), (9, (4, 6, 8))):
        E = _system(seed=seed, shape=shape)
        rng = np.random.default_rng(seed + 100)
        for _ in range(3):
            x = rng.standard_normal(E.shape[:3])
            ok &= bool(np.isclose(fn(E, x), _relres_np(E, x), rtol=1e-12, atol=0))
        ok &= fn(E, np.zeros(E.shape[:3])) == 1.0
    Et, ue, h = O.tlinear_system(8)
    ok &= fn(Et, ue) < 1e-13
    E = _system(seed=11)
    x = np.random.default_rng(3).standard_normal(E.shape[:3])
    ok &= bool(np.isclose(fn(3.5 * E, x), fn(E, x), rtol=1e-13))
    return ok


def test_true_relres_matches_dense_numpy():
    for seed, shape in ((5, None), (9, (4, 6, 8)), (13, (8, 4, 6))):
        E = _system(seed=seed, shape=shape)
        rng = np.random.default_rng(seed)
        for _ in range(4):
            x = rng.standard_normal(E.shape[:3])
            a, b = O.true_relres(E, x), _relres_np(E, x)
            assert np.isclose(a, b, rtol=1e-12, atol=0), (a, b)


def test_true_relres_closed_forms():
    E = _system()
    assert O.true_relres(E, np.zeros(E.shape[:3])) == 1.0
    # the reference's linear-test system: e7 = -A u*, so u* solves it
    Et, ue, h = O.tlinear_system(16)
    assert O.true_relres(Et, ue) < 1e-13
    # a scaled system has the same relative residual
    x = np.random.default_rng(2).standard_normal(E.shape[:3])
    assert np.isclose(O.true_relres(7.25 * E, x), O.true_relres(E, x), rtol=1e-13)
    # a solution of the system has a small residual, a perturbed one not
    xs = np.linalg.lstsq(_dense_np(E), -E[..., 7].reshape(-1), rcond=None)[0]
    xs = xs.reshape(E.shape[:3])
    assert O.true_relres(E, xs) < 1e-10
    assert O.true_relres(E, xs + 1e-3 * x) > 1e-6


def test_true_relres_linear_in_x_without_constant():
    """With e7 = 0 except one cell (b = e_k), ||A (c x)|| = |c| ||A x||."""
    E = _system(seed=17)
    E[..., 7] = 0.0
    E[0, 0, 0, 7] = 1.0
    x = np.random.default_rng(4).standard_normal(E.shape[:3])
    A = _dense_np(E)
    for c in (0.5, -2.0, 10.0):
        y = A @ (c * x).reshape(-1) + E[..., 7].reshape(-1)
        assert np.isclose(O.true_relres(E, c * x), np.linalg.norm(y), rtol=1e-12)


def test_oracle_true_relres_passes_checks():
    assert _checks(O.true_relres)


@pytest.mark.parametrize("mutant", [_m_no_sqrt, _m_sign, _m_drop_zp, _m_norm_x])
def test_planted_mistakes_fail_checks(mutant):
    assert not _checks(mutant)
