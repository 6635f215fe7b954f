"""Batched and served solves on the GPU.

Each system's launch geometry (tiles, workgroups, reduction order) depends on
that system only, so a system's result is bit-identical alone, in a batch,
and in a served mix; the tests assert exact equality.  Convergence is checked
on the FP64 true relative residual recomputed by the oracle.
"""
import numpy as np
import pytest

from aphros_amd import _lib as L
from aphros_amd import synth
from aphros_amd.batch import Batch
from aphros_amd.serve import Server
from aphros_amd.solver import Poisson
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SYSTEMS = [(32, 3, 0.1, 0.2, 1000.0, 21), (48, 5, 0.05, 0.15, 10.0, 22),
           (64, 8, 0.08, 0.15, 1000.0, 23), (32, 1, 0.2, 0.3, 100.0, 24)]


def _sys(n, ns, rmin, rmax, ratio, seed):
    e, _, _ = synth.system(n, ns, rmin, rmax, ratio, seed)
    return e.reshape(n, n, n, 8)


def _single(E, precision=L.MIXED, x0=None, tol=1e-8, precond=L.PRECOND_JACOBI,
            mg_w=L.MG_W_AUTO):
    s = Poisson(E.shape[:3], method=L.PCG, precision=precision, resmode=L.RES_REL_TRUE,
                tol=tol, maxiter=50000, precond=precond, mg_w=mg_w)
    s.set_system(E)
    x, info = s.solve(x0)
    s.close()
    return x, info


@pytest.mark.parametrize("precision", [L.MIXED, L.FP64])
@pytest.mark.parametrize("precond,w", [(L.PRECOND_JACOBI, 0), (L.PRECOND_MG, 0),
                                       (L.PRECOND_MG, 3), (L.PRECOND_MG, 6)])
def test_batch_equals_single(precision, precond, w):
    """Systems of different depths (32^3 and 48^3: 5 levels, 64^3: 6): with
    W levels each system keeps its own number (mg_w clamped to its depth)."""
    systems = [_sys(*p) for p in SYSTEMS]
    b = Batch(precision=precision, tol=1e-8, maxiter=50000, precond=precond, mg_w=w)
    slots = [b.add(E) for E in systems]
    b.solve()
    for E, sl in zip(systems, slots):
        xb, ib = b.get(sl)
        xs, is_ = _single(E, precision, precond=precond, mg_w=w)
        print(f"n={E.shape[0]} iter={ib['iter']} relres={ib['true_relres']:.2e}", flush=True)
        assert ib["converged"] == 1 and ib["iter"] == is_["iter"]
        assert np.array_equal(xb, xs)
        assert O.true_relres(E, xb) <= 1e-8 * (1 + 1e-6)
    b.close()


@pytest.mark.parametrize("precision", [L.MIXED, L.FP64])
def test_batch_odd_row_length(precision):
    """An odd row length (33^3) takes the one-cell-per-thread CG kernels and
    FP64 residual.  Cells per thread are chosen for the whole launch, so a
    batch that holds a 33^3 system runs its 32^3 neighbour at one cell per
    thread too: the 33^3 result equals its lone solve bit for bit, the 32^3
    one meets the same tolerance with another summation order."""
    e33 = _sys(33, 4, 0.1, 0.2, 1000.0, 31)
    e32 = _sys(32, 3, 0.1, 0.2, 1000.0, 21)
    b = Batch(precision=precision, tol=1e-8, maxiter=50000, precond=L.PRECOND_JACOBI)
    s33, s32 = b.add(e33), b.add(e32)
    b.solve()
    for E, sl in ((e33, s33), (e32, s32)):
        xb, ib = b.get(sl)
        xs, is_ = _single(E, precision, precond=L.PRECOND_JACOBI)
        print(f"n={E.shape[0]} batch iter={ib['iter']} alone iter={is_['iter']}", flush=True)
        assert ib["converged"] == 1 and is_["converged"] == 1
        assert O.true_relres(E, xb) <= 1e-8 * (1 + 1e-6)
        assert O.true_relres(E, xs) <= 1e-8 * (1 + 1e-6)
        if E.shape[0] == 33:
            assert ib["iter"] == is_["iter"] and np.array_equal(xb, xs)
        else:
            assert abs(ib["iter"] - is_["iter"]) <= 2
            scale = np.abs(xs - xs.mean()).max()
            assert np.abs((xb - xb.mean()) - (xs - xs.mean())).max() <= 1e-5 * scale
    b.close()


@pytest.mark.parametrize("precision", [L.MIXED, L.FP64])
def test_batch_fused_and_plain_levels(precision):
    """A 256^3 system (fused pre/post kernels on its level 0) in one batch with
    32^3 and 48^3 systems (separate sweeps): every level launch then runs both
    kinds, each skipping the other's systems; results equal the single solves
    bit for bit."""
    systems = [_sys(256, 16, 0.05, 0.15, 1000.0, 31), _sys(*SYSTEMS[0]), _sys(*SYSTEMS[1])]
    b = Batch(precision=precision, tol=1e-8, maxiter=50000, precond=L.PRECOND_MG)
    slots = [b.add(E) for E in systems]
    b.solve()
    for E, sl in zip(systems, slots):
        xb, ib = b.get(sl)
        xs, is_ = _single(E, precision, precond=L.PRECOND_MG, mg_w=L.MG_W_BATCH)
        print(f"n={E.shape[0]} iter={ib['iter']} relres={ib['true_relres']:.2e}", flush=True)
        assert ib["converged"] == 1 and ib["iter"] == is_["iter"]
        assert np.array_equal(xb, xs)
        assert O.true_relres(E, xb) <= 1e-8 * (1 + 1e-6)
    b.close()


def test_batch_warm_start_new_rhs():
    E = _sys(*SYSTEMS[2])
    b = Batch(precision=L.MIXED, tol=1e-8)
    sl = b.add(E)
    b.solve()
    x1, i1 = b.get(sl)
    E2 = E.copy()
    rng = np.random.default_rng(3)
    pert = rng.uniform(-1, 1, E.shape[:3])
    pert -= pert.mean()
    E2[..., 7] += 0.01 * np.abs(E[..., 7]).max() * pert
    b.set_rhs(sl, E2[..., 7])           # warm start from x1
    b.solve()
    x2, i2 = b.get(sl)
    assert i2["converged"] == 1 and i2["iter"] < i1["iter"]
    assert O.true_relres(E2, x2) <= 1e-8 * (1 + 1e-6)
    b.close()


@pytest.mark.parametrize("precond", [L.PRECOND_JACOBI, L.PRECOND_MG])
def test_server_alone_equals_mix(precond):
    A = _sys(*SYSTEMS[2])
    srv = Server(precision=L.MIXED, tol=1e-8, mem_budget=0, precond=precond)
    ra = srv.submit(A, nsteps=3, perturb=0.01, seed=5, want_x=True)
    alone = srv.wait(ra, timeout=120)
    assert alone["status"] == 0 and alone["steps_done"] == 3
    ids = [srv.submit(_sys(*SYSTEMS[i]), nsteps=2, perturb=0.01, seed=10 + i)
           for i in (0, 1, 3)]
    rb = srv.submit(A, nsteps=3, perturb=0.01, seed=5, want_x=True)
    ids.append(srv.submit(_sys(*SYSTEMS[1]), nsteps=1, seed=30))
    mixed = srv.wait(rb, timeout=120)
    others = [srv.wait(i, timeout=120) for i in ids]
    srv.close()
    assert mixed["status"] == 0
    assert all(o["status"] == 0 for o in others)
    assert mixed["iters_total"] == alone["iters_total"]
    assert np.array_equal(mixed["x"], alone["x"])
    assert mixed["last_relres"] <= 1e-8


def test_server_admission_by_budget_and_cancel():
    E = _sys(*SYSTEMS[2])
    # one 64^3 mixed system needs about 25 MB; budget for one at a time
    srv = Server(precision=L.MIXED, tol=1e-8, mem_budget=40 << 20)
    r1 = srv.submit(E, nsteps=2, seed=1)
    r2 = srv.submit(E, nsteps=2, seed=2)
    r3 = srv.submit(E, nsteps=2, seed=3)
    srv.cancel(r3)
    a = srv.wait(r1, timeout=120)
    b = srv.wait(r2, timeout=120)
    c = srv.wait(r3, timeout=120)
    assert a["status"] == 0 and b["status"] == 0
    assert b["t_admit"] >= a["t_done"] - 1e-6   # admitted once memory freed
    assert c["status"] == L.ECANCEL
    big = _sys(96, 2, 0.1, 0.2, 10.0, 4)        # never fits
    r4 = srv.submit(big, nsteps=1)
    d = srv.wait(r4, timeout=120)
    assert d["status"] == L.EBUDGET
    srv.close()


def test_server_stats_while_running_and_destroy_with_waiter():
    """stats() polled from another thread while requests are admitted,
    stepped and completed; destroy() with a thread blocked in an unbounded
    wait on an admitted request releases that thread."""
    import ctypes as C
    import threading
    import time
    srv = Server(precision=L.MIXED, tol=1e-8, precond=L.PRECOND_MG)
    systems = [_sys(*p) for p in SYSTEMS]
    polled = []
    stop = threading.Event()

    def poll():
        while not stop.is_set():
            polled.append(srv.stats())

    t = threading.Thread(target=poll)
    t.start()
    ids = [srv.submit(E, nsteps=3, perturb=0.01, seed=60 + k) for k, E in enumerate(systems)]
    ids.append(srv.submit(_sys(256, 16, 0.05, 0.15, 1000.0, 61), nsteps=2, seed=62))
    res = [srv.wait(i, timeout=120) for i in ids]
    stop.set()
    t.join(timeout=30)
    assert not t.is_alive()
    assert all(r["status"] == 0 for r in res)
    assert len(polled) > 0 and max(p["device_bytes"] for p in polled) > 0
    assert polled[-1]["completed"] <= len(ids)
    # a long request and a thread waiting on it without a time limit
    rid = srv.submit(_sys(256, 16, 0.05, 0.15, 1000.0, 63), nsteps=200, perturb=0.01, seed=64)
    out = {}

    def waiter():
        res = L.Result()
        out["rc"] = L.lib().aphpois_server_wait(srv.h, rid, -1.0, C.byref(res))
        out["status"] = res.status

    w = threading.Thread(target=waiter)
    w.start()
    time.sleep(1.0)
    srv.close()
    w.join(timeout=60)
    assert not w.is_alive()
    print(f"destroy released the waiter: rc={out.get('rc')} status={out.get('status')}",
          flush=True)
    assert out["rc"] == L.OK and out["status"] in (L.OK, L.ECANCEL)
