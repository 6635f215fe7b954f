"""The benchmark configurations as stated, solved on the GPU and checked
against the CPU oracle:

  C2  256^3, 32 spheres, density ratio 1000: FP64 Jacobi-preconditioned CG
      (the drop-in's plain form) to a 1e-10 true relative residual, against
      the oracle's FP64 multigrid-PCG solution;
  C3  512^3, 128 spheres: the mixed-precision multigrid solve to 1e-8 with the
      true residual recomputed by the oracle, and its error against the FP64
      GPU solution reported by the solver itself (info.max_diff);
  C5  a served request with explicit right-hand sides per time step, whose
      last solution is compared with the oracle's solution of the last step.

Solutions of a periodic system are defined up to a constant: they are
compared after removing their means.  The true relative residual is the
oracle's orc_true_relres (pinned in test_oracle_relres.py).
"""
import numpy as np
import pytest

from aphros_amd import _lib as L
from aphros_amd import synth
from aphros_amd.serve import Server
from aphros_amd.solver import Poisson
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _rel_diff(x, y):
    d = (x - x.mean()) - (y - y.mean())
    return np.abs(d).max() / np.abs(y - y.mean()).max()


def test_c2_fp64_jacobi_pcg_against_oracle():
    e, _, _ = synth.config("c2_256")
    E = e.reshape(256, 256, 256, 8)
    _say("c2_256 fp64 jacobi-pcg to 1e-10")
    s = Poisson(E.shape[:3], method=L.PCG, precision=L.FP64, resmode=L.RES_REL_TRUE,
                tol=1e-10, maxiter=200000, precond=L.PRECOND_JACOBI)
    s.set_system(E)
    x, info = s.solve(None)
    s.close()
    assert info["converged"] == 1 and info["status"] == 0
    rel = O.true_relres(E, x)
    xo, ito, rro, _ = O.mgpcg(E, tol=1e-10, w=1)
    err = _rel_diff(x, xo)
    _say(f"  gpu iter={info['iter']} ({info['seconds']:.3f}s) relres={rel:.3e}; "
         f"oracle mgpcg iter={ito} relres={rro:.3e}; rel diff {err:.3e}")
    assert rel <= 1e-10 * (1 + 1e-6)
    assert err <= 1e-6


def test_c3_mixed_solve_512():
    e, _, _ = synth.config("c3_512")
    E = e.reshape(512, 512, 512, 8)
    del e
    _say("c3_512 fp64 multigrid solve")
    s64 = Poisson(E.shape[:3], method=L.PCG, precision=L.FP64, resmode=L.RES_REL_TRUE,
                  tol=1e-8, precond=L.PRECOND_MG)
    s64.set_system(E)
    x64, i64 = s64.solve(None)
    s64.close()
    assert i64["converged"] == 1
    _say(f"  fp64 iter={i64['iter']} relres={i64['true_relres']:.3e}")
    _say("c3_512 mixed multigrid solve")
    s = Poisson(E.shape[:3], method=L.PCG, precision=L.MIXED, resmode=L.RES_REL_TRUE,
                tol=1e-8, precond=L.PRECOND_MG)
    s.set_system(E)
    s.set_solution_ref(x64)
    x, info = s.solve(None)
    s.close()
    assert info["converged"] == 1 and info["status"] == 0
    rel = O.true_relres(E, x)
    d = x - x64
    want = np.abs(d - d.mean()).max()
    _say(f"  mixed iter={info['iter']} relres={rel:.3e} (gpu {info['true_relres']:.3e}) "
         f"max_diff={info['max_diff']:.3e} numpy {want:.3e} max|x|={np.abs(x64).max():.3e}")
    assert rel <= 1e-8 * (1 + 1e-6)
    assert abs(rel - info["true_relres"]) <= 1e-3 * rel
    assert np.isclose(info["max_diff"], want, rtol=1e-9)
    assert info["max_diff"] <= 1e-5 * np.abs(x64).max()


@pytest.mark.parametrize("precond", [L.PRECOND_JACOBI, L.PRECOND_MG])
def test_served_result_matches_oracle(precond):
    """Three time steps with explicit constant terms (each warm-started from
    the previous step's solution): the served solution of the last step equals
    the oracle's FP64 solution of that system."""
    n = 64
    e, _, _ = synth.system(n, 8, 0.08, 0.15, 1000.0, 41)
    E = e.reshape(n, n, n, 8)
    rng = np.random.default_rng(42)
    steps = [None]
    b = E[..., 7].copy()
    for k in range(1, 3):
        eta = rng.uniform(-1, 1, b.shape)
        eta -= eta.mean()
        b = b + 0.01 * np.abs(b).max() * eta
        steps.append(b.copy())
    srv = Server(precision=L.MIXED, tol=1e-8, precond=precond)
    rid = srv.submit(E, nsteps=3, e7_steps=steps, want_x=True)
    r = srv.wait(rid, timeout=120)
    srv.close()
    assert r["status"] == 0 and r["steps_done"] == 3
    last = E.copy()
    last[..., 7] = steps[-1]
    x = r["x"]
    rel = O.true_relres(last, x)
    xo, ito, rro = O.pcg_jacobi(last, tol=1e-11)
    err = _rel_diff(x, xo)
    _say(f"  served iters={r['iters_total']} relres={rel:.3e} rel diff vs oracle {err:.3e}")
    assert rel <= 1e-8 * (1 + 1e-6)
    assert err <= 1e-5
