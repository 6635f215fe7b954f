"""Pins the oracle's multigrid preconditioner and multigrid PCG (CPU only).

The reference's symmetric plugin has no preconditioner, so nothing of the
reference fixes these functions directly.  They are pinned by what the
mathematics fixes instead:
  * a dense matrix restatement of the V-cycle, built from the system matrix
    itself (not from face conductances): the coarse operator of aggregation
    coarsening is 0.5 * P^T A P with P the injection to the 2x2x2 children
    (internal faces cancel in P^T A P), so an index or transposition mistake
    in the stencil coarsening, the restriction or the prolongation fails it;
  * symmetry and positive definiteness of M, which CG relies on;
  * MG-PCG reaching the tolerance on the true residual (orc_true_relres is
    pinned against a dense numpy residual, closed forms and planted mutants
    in test_oracle_relres.py) and agreeing with the Jacobi-PCG solution.
"""
import numpy as np
import pytest

from oracle import oracle as O

OMEGA = 0.9  # coarsest level
# the degree-2 Chebyshev polynomial on [0.5, 1.8] for D^-1 A: its two roots
_C, _D = (0.5 + 1.8) / 2, (1.8 - 0.5) / 2
OMEGA_A = 1 / (_C + _D * np.cos(np.pi / 4))
OMEGA_B = 1 / (_C - _D * np.cos(np.pi / 4))
PROLONG = 1.1  # scale of the prolonged coarse correction


def _system(n, seed, ratio=1000.0, mass=0.0, sign=1, shape=None):
    rng = np.random.default_rng(seed)
    shape = shape or (n, n, n)
    rho = np.where(rng.random(shape) < 0.3, ratio, 1.0)
    f = rng.standard_normal(shape)
    f -= f.mean()
    e = O.assemble(rho, f, sign=sign)
    if mass:
        e[..., 0] += sign * mass * np.abs(e[..., 0]).mean() * rng.random(shape)
    return e


def _dense(e):
    """Dense s*A (s = sign of the diagonal) from the Expr by applying it to the
    unit vectors."""
    n = e[..., 0].size
    A = np.empty((n, n))
    for j in range(n):
        u = np.zeros(n)
        u[j] = 1
        A[:, j] = O.apply(e, u.reshape(e.shape[:3])).ravel()
    return np.sign(e.flat[0]) * A


def _injection(shape, shift=0):
    nz, ny, nx = shape
    cz, cy, cx = nz // 2, ny // 2, nx // 2
    P = np.zeros((nx * ny * nz, cx * cy * cz))
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                X = ((x + shift) % nx) // 2
                P[(z * ny + y) * nx + x, ((z // 2) * cy + y // 2) * cx + X] = 1
    return P, (cz, cy, cx)


def _dense_vcycle(e, sweeps, scale=0.5, shift=0, w=0):
    """M as a dense matrix, from the cycle's definition in matrix form
    (scale, shift: planted mistakes for the check's own test; w: levels with
    two coarse corrections)."""
    A = [_dense(e)]
    shapes = [e.shape[:3]]
    Ps = []
    while True:
        nz, ny, nx = shapes[-1]
        if (nx | ny | nz) & 1 or (nx // 2) * (ny // 2) * (nz // 2) < 8:
            break
        P, cs = _injection(shapes[-1], shift)
        Ps.append(P)
        A.append(scale * P.T @ A[-1] @ P)
        shapes.append(cs)

    def smooth(l, u, r, om):
        d = np.diag(A[l])
        dinv = np.where(d != 0, 1 / np.where(d != 0, d, 1), 0)
        return u + om * dinv * (r - A[l] @ u)

    def vc(l, r):
        d = np.diag(A[l])
        dinv = np.where(d != 0, 1 / np.where(d != 0, d, 1), 0)
        if l == len(A) - 1:
            u = OMEGA * dinv * r
            for _ in range(sweeps - 1):
                u = smooth(l, u, r, OMEGA)
            return u
        z = smooth(l, OMEGA_A * dinv * r, r, OMEGA_B)
        rc = Ps[l].T @ (r - A[l] @ z)
        zc = vc(l + 1, rc)
        if l < w and l + 1 < len(A) - 1:
            zs = zc
            zc = vc(l + 1, rc - A[l + 1] @ zs)
            zc = zc + zs
        z = smooth(l, z + PROLONG * (Ps[l] @ zc), r, OMEGA_B)
        return smooth(l, z, r, OMEGA_A)

    n = A[0].shape[0]
    return np.stack([vc(0, np.eye(n)[:, j]) for j in range(n)], axis=1), len(A)


@pytest.mark.parametrize("shape,mass,sign", [((8, 8, 8), 0.0, 1), ((4, 4, 4), 0.0, -1),
                                             ((4, 8, 8), 0.05, 1), ((8, 8, 4), 0.05, -1)])
def test_vcycle_matches_dense_definition(shape, mass, sign):
    e = _system(None, 7, mass=mass, sign=sign, shape=shape)
    for sweeps in (1, 20, 40):
        M, nlev = _dense_vcycle(e, sweeps)
        rng = np.random.default_rng(3)
        r = rng.standard_normal(shape)
        z, nl = O.mg_apply(e, r, sweeps=sweeps)
        assert nl == nlev
        ze = (M @ r.ravel()).reshape(shape)
        assert np.max(np.abs(z - ze)) <= 1e-11 * np.max(np.abs(ze))


@pytest.mark.parametrize("w", [1, 2])
def test_w_levels_match_dense_definition(w):
    """Two coarse corrections on the top w levels (8^3: 3 levels, so w = 2
    acts like w = 1: level 1's coarse level is the coarsest)."""
    e = _system(None, 7, shape=(8, 8, 8), mass=0.05)
    M, nlev = _dense_vcycle(e, O.MG_SWEEPS, w=w)
    r = np.random.default_rng(3).standard_normal((8, 8, 8))
    z, nl = O.mg_apply(e, r, w=w)
    ze = (M @ r.ravel()).reshape(8, 8, 8)
    assert nl == nlev == 3
    assert np.max(np.abs(z - ze)) <= 1e-11 * np.max(np.abs(ze))
    z0, _ = O.mg_apply(e, r, w=0)
    assert np.max(np.abs(z - z0)) > 1e-4 * np.max(np.abs(ze))  # it does something


def test_w_levels_symmetric_and_fewer_iterations():
    e = _system(32, 11)
    rng = np.random.default_rng(5)
    r1, r2 = rng.standard_normal((2, 32, 32, 32))
    z1, _ = O.mg_apply(e, r1, w=2)
    z2, _ = O.mg_apply(e, r2, w=2)
    a, b = np.vdot(z1, r2), np.vdot(r1, z2)
    assert abs(a - b) <= 1e-12 * np.sqrt(np.vdot(z1, r1) * np.vdot(z2, r2))
    # heavy spheres on a 64^3 grid (6 levels): the second corrections pay off
    e = _spheres(64, 12, 1)
    _, it0, rel0, _ = O.mgpcg(e, tol=1e-10, w=0)
    _, it2, rel2, _ = O.mgpcg(e, tol=1e-10, w=2)
    assert rel0 <= 1e-10 and rel2 <= 1e-10 and it2 < it0


def _spheres(n, k, seed, ratio=1000.0):
    """Density `ratio` inside k random periodic spheres (radius 0.05-0.15)."""
    rng = np.random.default_rng(seed)
    g = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    rho = np.ones((n, n, n))
    for _ in range(k):
        c = rng.random(3)
        r = 0.05 + 0.1 * rng.random()
        d2 = sum(np.minimum(np.abs(a - ci), 1 - np.abs(a - ci)) ** 2 for a, ci in zip((X, Y, Z), c))
        rho[d2 < r * r] = ratio
    f = rng.standard_normal((n, n, n))
    f -= f.mean()
    return O.assemble(rho, f)


def test_vcycle_symmetric_positive():
    e = _system(16, 11)
    rng = np.random.default_rng(5)
    r1, r2 = rng.standard_normal((2, 16, 16, 16))
    r1 -= r1.mean()
    r2 -= r2.mean()
    z1, nlev = O.mg_apply(e, r1)
    z2, _ = O.mg_apply(e, r2)
    assert nlev == 4  # 16, 8, 4, 2 cells per edge
    a, b = np.vdot(z1, r2), np.vdot(r1, z2)
    assert abs(a - b) <= 1e-12 * np.sqrt(np.vdot(z1, r1) * np.vdot(z2, r2))
    assert np.vdot(z1, r1) > 0 and np.vdot(z2, r2) > 0


@pytest.mark.parametrize("scale,shift", [(1.0, 0), (0.5, 1)])
def test_vcycle_check_catches_planted_mistake(scale, shift):
    """A Galerkin coarse operator without the 0.5, or aggregates offset by one
    cell, gives a different M: the dense-definition check has teeth."""
    e = _system(None, 7, shape=(4, 4, 4))
    M, _ = _dense_vcycle(e, O.MG_SWEEPS, scale=scale, shift=shift)
    r = np.random.default_rng(3).standard_normal((4, 4, 4))
    z, _ = O.mg_apply(e, r)
    ze = (M @ r.ravel()).reshape(4, 4, 4)
    assert np.max(np.abs(z - ze)) > 1e-3 * np.max(np.abs(ze))


@pytest.mark.parametrize("n,mass", [(32, 0.0), (32, 0.05)])
def test_mgpcg_reaches_tolerance(n, mass):
    e = _system(n, 21, mass=mass)
    # the mean is a free constant only without the mass term
    x, it, rel, _ = O.mgpcg(e, tol=1e-10, remove_mean=not mass)
    assert rel <= 1e-10
    assert O.true_relres(e, x) <= 1e-10
    xj, itj, relj = O.pcg_jacobi(e, tol=1e-10, remove_mean=not mass)
    assert it < itj / 3  # the V-cycle is a far better preconditioner
    d = x - xj
    if not mass:
        d -= d.mean()
    assert np.max(np.abs(d)) <= 1e-7 * np.max(np.abs(xj))


def test_mgpcg_rejects_odd_grid():
    e = _system(None, 1, shape=(6, 6, 5))
    with pytest.raises(ValueError):
        O.mgpcg(e)


def test_vcycle_bf16_coefficients():
    """The FP32 engines' storage (bfloat16 coefficients): still a symmetric
    positive definite M, within bfloat16 accuracy of the FP64 V-cycle, and
    different from it (the rounding is applied)."""
    e = _system(16, 11)
    rng = np.random.default_rng(5)
    r1, r2 = rng.standard_normal((2, 16, 16, 16))
    z1, _ = O.mg_apply(e, r1, bf16=True)
    z2, _ = O.mg_apply(e, r2, bf16=True)
    a, b = np.vdot(z1, r2), np.vdot(r1, z2)
    assert abs(a - b) <= 1e-12 * np.sqrt(np.vdot(z1, r1) * np.vdot(z2, r2))
    assert np.vdot(z1, r1) > 0
    zd, _ = O.mg_apply(e, r1)
    rel = np.abs(z1 - zd).max() / np.abs(zd).max()
    assert 1e-5 < rel < 3e-2
