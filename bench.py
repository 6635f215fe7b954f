#!/usr/bin/env python3
"""Benchmark: Poisson solves/s at 512^3 to a 1e-8 true relative residual,
mixed precision versus FP64, on one MI355X.

A step is one complete solve of the resident 512^3 pressure system (128
spheres, density ratio 1000, zero initial guess) to ||b - A x|| / ||b|| <= 1e-8,
with the system already in HBM and the solution left in HBM.  The FP64 form of
the same solver runs on the same system for the mixed/FP64 ratio.

Prints one JSON line (see the fields below).  --gpus N > 1 is refused: this
build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Poisson solves/s @512^3 to 1e-8 (mixed vs FP64 GPU); p50/p99 served latency"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM traffic per cell of each iteration kernel: `vec` values of
# the engine's element type T (4 bytes MIXED, 8 FP64) and `coef` values of
# the V-cycle coefficient type (bfloat16, 2 bytes, in MIXED; T in FP64), each
# array read or written once per launch; +1 T value for the mass term of a
# non-singular system.  slot: index in Poisson.kernel_times(); name: the
# kernel instance in profiles.
KERNELS = {
    # p = z + beta p: read z, p_in, kx, ky, kz (the CG's FP32 operator); write
    # p_out, q; and the x^ update with alpha p_in of the previous iteration
    # (xh: x^ vectors per launch, averaged by xh_vectors)
    "spmv": {"slot": 0, "vec": 7.0, "xh": True, "coef": 0.0, "m": 1,
             "name": "k_pcg_spmv<T, VX, false, true>"},
    # r -= alpha q: read r, q; write r
    "update": {"slot": 2, "vec": 3.0, "coef": 0.0, "m": 0, "name": "k_pcg_update<T, VX>"},
    # u = w D^-1 r fused: read r, kx, ky, kz, dinv; write z
    "mg_presweep_l0": {"slot": 13, "vec": 2.0, "coef": 4.0, "m": 1,
                       "name": "k_mg_sweep<T, VX, 1, false, true>"},
    # read z_in, z_coarse (1/8), r, kx, ky, kz, dinv; write z
    "mg_postsweep1_l0": {"slot": 14, "vec": 3.125, "coef": 4.0, "m": 1,
                         "name": "k_mg_sweep<T, VX, 2, false, true>"},
    # read z_in, r, kx, ky, kz, dinv; write z; (r, z) partials
    "mg_postsweep2_l0": {"slot": 15, "vec": 3.0, "coef": 4.0, "m": 1,
                         "name": "k_mg_sweep<T, VX, 0, true, true>"},
    # read r, z, kx, ky, kz; write r_coarse (1/8)
    "mg_restrict_l0": {"slot": 10, "vec": 2.125, "coef": 3.0, "m": 1,
                       "name": "k_mg_sweep<T, VX, 3, false, true>"},
    # both post-sweeps fused (slot 15 then stays empty): read z_in, z_coarse
    # (1/8), r, kx, ky, kz, dinv; write z; (r, z) partials
    "mg_post_l0": {"slot": 14, "vec": 3.125, "coef": 4.0, "m": 1,
                   "name": "k_mg_post<T, VX, true, true>"},
    # pre-sweep and restriction fused (slot 10 then stays empty): read r, kx,
    # ky, kz, dinv; write z, r_coarse (1/8)
    "mg_pre_l0": {"slot": 13, "vec": 2.125, "coef": 4.0, "m": 1,
                  "name": "k_mg_pre<T, VX, true>"},
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo) and logical CPU count."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return f"{line.split(':', 1)[1].strip()} ({os.cpu_count()} logical CPUs)"
    except OSError:
        pass
    return None


# the Jacobi-scaled form of the operator: p = S r + beta p, read r, p_in, S,
# kx, ky, kz; write p_out, q; and x^ as in the multigrid form
SPMV_JACOBI = {"slot": 0, "vec": 8.0, "xh": True, "coef": 0.0, "m": 1,
               "name": "k_pcg_spmv<T, VX, true, true>"}


def launched_name(name, elem, vx):
    """The kernel instance the engine launches for a table entry: the
    template arguments of the precision and cells per thread filled in; FP32
    engines with 4 cells per thread run the software-pipelined fused level
    kernels k_mg_pre4 / k_mg_post4."""
    if elem == 4 and vx == 4:
        if name == "k_mg_post<T, VX, true, true>":
            return "k_mg_post4<true, true>"
        if name == "k_mg_pre<T, VX, true>":
            return "k_mg_pre4<true>"
    return name.replace("T, VX", f"{'float' if elem == 4 else 'double'}, {vx}")


def xh_vectors(solves):
    """x^ vectors the CG operator moves per launch, averaged over the
    launches of `solves` ((iterations, reliable updates including the start)
    pairs): 0 in the launch after each reliable update but the last (which
    folds alpha p into x64 itself), 1 (write x^ = alpha p_in) in the launch
    after that, 2 (read and write x^ += alpha p_in) in the others."""
    moved = launches = 0
    for n_iter, n_ru in solves:
        skip = min(n_iter, max(0, n_ru - 1))
        fresh = min(n_iter - skip, skip)
        moved += 2 * (n_iter - skip - fresh) + fresh
        launches += n_iter
    return moved / max(1, launches)


def kernel_table(kt, ncell, elem, coef, has_m, per_iter_launches, jacobi=False, vx=1,
                 xh_vecs=2.0):
    """Per-kernel average duration and achieved algorithmic bandwidth.
    xh_vecs: x^ vectors per CG-operator launch (xh_vectors)."""
    out = {}
    table = dict(KERNELS)
    if jacobi:
        table = {"spmv": SPMV_JACOBI, "update": KERNELS["update"]}
    fused = len(kt) > 15 and kt[15][1] == 0
    fused_pre = len(kt) > 10 and kt[10][1] == 0
    skip = {"mg_postsweep1_l0" if fused else "mg_post_l0"}
    skip |= {"mg_presweep_l0", "mg_restrict_l0"} if fused_pre else {"mg_pre_l0"}
    for key, k in table.items():
        if k["slot"] >= len(kt):
            continue
        if key in skip:
            continue
        ms, cnt = kt[k["slot"]]
        if cnt == 0 or ms <= 0:
            continue
        per_cell = (k["vec"] + (xh_vecs if k.get("xh") else 0.0) + (k["m"] if has_m else 0)) * elem \
            + k["coef"] * coef
        b = per_cell * ncell
        out[key] = {"kernel": launched_name(k["name"], elem, vx), "avg_ms": round(ms, 5),
                    "launches": cnt,
                    "alg_bytes_per_cell": round(per_cell, 4), "alg_bytes_per_launch": int(b),
                    "GBps": round(b / (ms * 1e-3) / 1e9, 1),
                    "ms_per_iter": round(ms * per_iter_launches.get(key, 1), 5)}
        if k.get("xh"):
            out[key]["xh_vectors_per_launch"] = round(xh_vecs, 4)
    return out


def measured_traffic(config, prec, precond, kernel, li, L):
    """HBM bytes per working launch of `kernel` (the launches that did work,
    as the HIP event averages count them) from the committed FETCH_SIZE /
    WRITE_SIZE counter passes (profiles/traffic_<workload>.json, calibrated
    against a plain copy), when they cover this workload; else None."""
    if prec != L.MIXED or precond != L.PRECOND_MG:
        return None
    path = os.path.join(ROOT, "profiles", f"traffic_{config}_mixed_mg.json")
    try:
        with open(path) as f:
            table = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    rec = table.get(kernel)
    return None if rec is None else int(rec["fetch_bytes"] + rec["write_bytes"])


def run_c2(args):
    """256^3 (32 spheres r in [0.08, 0.15], density ratio 1000, seed 2): FP64
    Jacobi-preconditioned CG, the drop-in's plain form, to a 1e-8 true
    relative residual from a zero guess; HBM GB/s per iteration from the
    event-timed iteration kernels."""
    from aphros_amd import _lib as L
    from aphros_amd import synth
    from aphros_amd.devmem import DeviceBuffer
    from aphros_amd.solver import Poisson
    cfg = synth.CONFIGS["c2_256"]
    n = cfg["n"]
    ncell = n ** 3
    e, _, _ = synth.config("c2_256")
    s = Poisson((n, n, n), method=L.PCG, precision=L.FP64, resmode=L.RES_REL_TRUE, tol=1e-8,
                maxiter=200000, precond=L.PRECOND_JACOBI)
    s.set_system(e)
    del e
    xb = DeviceBuffer(ncell * 8)
    t = time.perf_counter()
    info0 = s.solve_device(xb.ptr)  # graph replays
    plain = time.perf_counter() - t
    s.set_profiling(True)
    t = time.perf_counter()
    info = s.solve_device(xb.ptr)  # kernels launched singly between events
    prof = time.perf_counter() - t
    kt = s.kernel_times()
    li = s.launch_info()
    s.close()
    xb.free()
    if not (info0["converged"] and info["converged"]):
        raise RuntimeError("c2 solve did not converge")
    kernels = kernel_table(kt, ncell, 8, 8, False, {}, jacobi=True, vx=li["vx"],
                           xh_vecs=xh_vectors([(info["iter"], info["refinements"])]))
    b_it = sum(k["alg_bytes_per_launch"] for k in kernels.values())
    ms_it = sum(k["avg_ms"] for k in kernels.values())
    log(f"[bench] c2: {info0['iter']} iterations, {plain:.3f}s; iteration kernels "
        f"{b_it / (ms_it * 1e-3) / 1e9:.0f} GB/s")
    return {"workload": f"c2_256: {n}^3 periodic, {cfg['nspheres']} spheres r in "
                        f"[{cfg['rmin']},{cfg['rmax']}], density ratio {cfg['ratio']:g}, seed "
                        f"{cfg['seed']}; FP64 Jacobi-preconditioned CG to a 1e-8 true "
                        f"relative residual, zero guess",
            "iters": info0["iter"], "refinements": info0["refinements"],
            "true_relres": info0["true_relres"], "s_per_solve": round(plain, 4),
            "s_per_solve_event_timed": round(prof, 4),
            "alg_bytes_per_iter": int(b_it), "iter_kernels_ms": round(ms_it, 5),
            "GBps_per_iter": round(b_it / (ms_it * 1e-3) / 1e9, 1),
            "frac_of_peak": round(b_it / (ms_it * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "GBps_whole_solve": round(b_it * info0["iter"] / plain / 1e9, 1),
            "kernels": kernels}


def run_c1(args):
    """64^3 (8 spheres, density ratio 1000, seed 1) to a 1e-8 true relative
    residual: the reference's CG restated on the GPU (CONJUGATE, FP64, its
    stop rule with tol = 1e-8 ||b|| / sqrt(h^3)), the multigrid PCG forms on
    the GPU, and the reference's CG on the host's cores (cpu_baseline)."""
    from aphros_amd import _lib as L
    from aphros_amd import synth
    from aphros_amd.devmem import DeviceBuffer
    from aphros_amd.solver import Poisson
    from oracle import oracle as O
    n = 64
    ncell = n ** 3
    h3 = (1.0 / n) ** 3
    e, _, _ = synth.config("c1_64")
    E = e.reshape(n, n, n, 8)
    # the reference's residual is sqrt(sum r^2 / h^3): this tol is a relative
    # residual of 1e-8 on the recursive residual
    tol_ref = 1e-8 * float(np.linalg.norm(E[..., 7])) / np.sqrt(h3)
    xb = DeviceBuffer(ncell * 8)
    out = {"workload": "c1_64: 64^3 periodic, 8 spheres r in [0.08,0.15], density ratio 1000, "
                       "seed 1, zero guess, 1e-8 true relative residual"}

    def gpu(name, **kw):
        s = Poisson((n, n, n), **kw)
        s.set_system(E)
        s.solve_device(xb.ptr)
        t = time.perf_counter()
        reps = 3
        for _ in range(reps):
            info = s.solve_device(xb.ptr)
        dt = (time.perf_counter() - t) / reps
        s.close()
        out[name] = {"s_per_solve": round(dt, 5), "iters": info["iter"],
                     "true_relres": info["true_relres"]}
        log(f"[bench] c1 {name}: {dt * 1e3:.2f} ms, {info['iter']} iterations, "
            f"relres {info['true_relres']:.2e}")
        return info

    ic = gpu("gpu_conjugate", method=L.CONJUGATE, resmode=L.RES_L2, tol=tol_ref, miniter=0,
             maxiter=100000, h3=h3)
    gpu("gpu_mg_pcg_mixed", method=L.PCG, precision=L.MIXED, tol=1e-8,
        precond=L.PRECOND_MG)
    gpu("gpu_mg_pcg_fp64", method=L.PCG, precision=L.FP64, tol=1e-8, precond=L.PRECOND_MG)
    xb.free()
    iters = ic["iter"]
    cpu = None
    if args.cpu_iters > 0:
        k = max(1, min(args.cpu_iters, iters))
        t = time.perf_counter()
        _, it, _, _ = O.conjugate(E, h3, tol=0.0, miniter=0, maxiter=k - 1)
        dt = time.perf_counter() - t
        spi = dt / it
        cpu = {"value": round(1.0 / (spi * iters), 5),
               "unit": "solves/s @64^3 (SolverConjugate to a 1e-8 relative residual)",
               "cores": O.num_threads(), "kind": "port", "cpu_model": cpu_model(),
               "iters": iters, "s_per_solve": round(spi * iters, 3),
               "sample": f"{it} iterations of the oracle's SolverConjugate restatement "
                         f"(bit-identical iterates, OpenMP) on the 64^3 system in "
                         f"{dt:.2f}s, scaled to the {iters} iterations of the solve "
                         f"(the GPU CONJUGATE count, the same stop rule)"}
        if args.ref_iters > 0 and O.ref_available():
            kr = max(1, min(args.ref_iters, iters))
            _, itr, _, sec = O.ref_solve("conjugate", E, 0.0, 0, kr - 1, timed=True)
            cpu["reference_compiled"] = {
                "s_per_solve": round(sec / itr * iters, 3), "cores": 1,
                "sample": f"{itr} iterations of the reference's own SolverConjugate "
                          f"(oracle/_ref/ref_driver: the reference's sources, one block, "
                          f"one thread) in {sec:.2f}s, scaled to {iters} iterations"}
        for key in ("gpu_conjugate", "gpu_mg_pcg_mixed", "gpu_mg_pcg_fp64"):
            out[key]["speedup_vs_cpu"] = round(cpu["s_per_solve"] / out[key]["s_per_solve"], 1)
        log(f"[bench] c1 cpu: {cpu['s_per_solve']:.2f}s per solve ({cpu['cores']} threads)")
    return out, cpu


def run_batched(args, precond):
    """Batched workload: independent periodic systems of uneven edges solved
    together (one set of launches per kernel for all of them), mixed
    precision to a 1e-8 true relative residual from a zero guess."""
    from aphros_amd import _lib as L
    from aphros_amd import synth
    from aphros_amd.batch import Batch
    from aphros_amd import autotune
    specs = synth.batch_specs(args.batch_systems)
    # tile shapes of the CG stencil per grid edge from the autotuner (timed
    # once per edge, cached)
    t = time.time()
    tiles = {n: autotune.tile_for(n, L.MIXED, precond=precond)
             for n in sorted({sp["n"] for sp in specs})}
    log(f"[bench] batched: tiles {tiles} ({time.time() - t:.1f}s)")
    b = Batch(precision=L.MIXED, tol=1e-8, precond=precond)
    t = time.time()
    slots = []
    for sp in specs:
        e, _, _ = synth.spec_system(sp)
        sl = b.add(e)
        b.set_system_tile(sl, *tiles[sp["n"]])
        slots.append((sl, sp["n"]))
    log(f"[bench] batched: {len(slots)} systems resident in {time.time() - t:.1f}s, "
        f"device_bytes={b.device_bytes() / 1e9:.2f} GB")
    b.solve()  # first solve: launch tables and graph capture
    for sl, n in slots:
        b.set_rhs(sl, None, np.zeros(n ** 3))
    t = time.perf_counter()
    b.solve()
    dt = time.perf_counter() - t
    infos = [b.get(sl, want_x=False)[1] for sl, _ in slots]
    b.close()
    if not all(i["converged"] for i in infos):
        raise RuntimeError("batched solve did not converge")
    cells = sum(n ** 3 for _, n in slots)
    log(f"[bench] batched: {dt:.3f}s for {len(slots)} solves")
    return {"systems": len(slots), "cells": cells, "seconds": round(dt, 4),
            "tiles": {str(n): list(tl) for n, tl in tiles.items()},
            "solves_per_s": round(len(slots) / dt, 3),
            "iters_min": min(i["iter"] for i in infos),
            "iters_max": max(i["iter"] for i in infos),
            "max_true_relres": max(i["true_relres"] for i in infos),
            "workload": f"{len(slots)} periodic systems, edges from {synth.BATCH_EDGES}, "
                        f"1-16 spheres, density ratio log-uniform in [10, 1000], seeds 4.."
                        f"{3 + len(slots)}; mixed precision to 1e-8, zero guess"}


def run_served(args, precond):
    """Served workload: Poisson arrivals of requests, each a sequence of 10
    correlated right-hand sides (1% perturbation per step, warm started),
    continuous batching under a 200 GB memory budget."""
    from aphros_amd import _lib as L
    from aphros_amd import synth
    from aphros_amd.serve import Server
    specs = synth.serve_specs(args.serve_requests, args.serve_rate)
    systems = [synth.spec_system(sp)[0] for sp in specs]
    srv = Server(precision=L.MIXED, tol=1e-8, mem_budget=200e9, precond=precond)
    # reference point: a request of the largest edge alone on the idle server
    big = max(range(len(specs)), key=lambda i: specs[i]["n"])
    ra = srv.submit(systems[big], nsteps=10, perturb=0.01, seed=specs[big]["seed"])
    r_alone = srv.wait(ra, timeout=120)
    alone_s = r_alone["t_done"] - r_alone["t_submit"]
    log(f"[bench]   served: a {specs[big]['n']}^3 request alone: {alone_s:.3f}s")
    t0 = time.perf_counter()
    ids = []
    for sp, e in zip(specs, systems):
        delay = sp["arrival"] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(e, nsteps=10, perturb=0.01, seed=sp["seed"]))
    res = []
    deadline = time.perf_counter() + 120.0
    for k, i in enumerate(ids):
        while True:
            try:
                r = srv.wait(i, timeout=5.0)
                break
            except L.AphpoisError:
                if time.perf_counter() > deadline:
                    srv.close()
                    raise RuntimeError(f"served request {k} not finished within the deadline")
                log(f"[bench]   served: waiting for request {k}, {srv.stats()}")
        res.append(r)
        log(f"[bench]   served: request {k} ({specs[k]['n']}^3) done: "
            f"{r['t_done'] - r['t_submit']:.3f}s, {r['iters_total']} iterations")
    wall = time.perf_counter() - t0
    srv.close()
    bad = [r for r in res if r["status"] != 0 or r["steps_done"] != 10]
    if bad:
        raise RuntimeError(f"served requests failed: {bad[:2]}")
    lat = np.array([r["t_done"] - r["t_submit"] for r in res])
    per_solve = np.array([(r["t_done"] - r["t_admit"]) / r["steps_done"] for r in res])
    t_first = min(r["t_submit"] for r in res)
    t_last = max(r["t_done"] for r in res)
    solves = sum(r["steps_done"] for r in res)
    log(f"[bench] served: {len(res)} requests, {solves} solves in {t_last - t_first:.2f}s, "
        f"p50 {np.percentile(lat, 50):.3f}s p99 {np.percentile(lat, 99):.3f}s")
    # latency per grid edge (size class)
    by_edge = {}
    for sp, r in zip(specs, res):
        by_edge.setdefault(sp["n"], []).append(r["t_done"] - r["t_submit"])
    per_size = {str(n): {"requests": len(v), "p50_request_s": round(float(np.percentile(v, 50)), 4),
                         "p99_request_s": round(float(np.percentile(v, 99)), 4)}
                for n, v in sorted(by_edge.items())}
    return {"requests": len(res), "solves": solves,
            "solves_per_s": round(solves / (t_last - t_first), 3),
            "p50_request_s": round(float(np.percentile(lat, 50)), 4),
            "p99_request_s": round(float(np.percentile(lat, 99)), 4),
            "p50_solve_s": round(float(np.percentile(per_solve, 50)), 5),
            "p99_solve_s": round(float(np.percentile(per_solve, 99)), 5),
            "max_last_relres": max(r["last_relres"] for r in res),
            "by_edge": per_size,
            "alone_largest_request_s": round(alone_s, 4),
            "p99_over_alone_largest": round(float(np.percentile(lat, 99)) / alone_s, 3),
            "wall_s": round(wall, 3),
            "workload": f"{len(res)} requests, Poisson arrivals at {args.serve_rate}/s, edges "
                        f"from {synth.SERVE_EDGES}, 8 spheres, density ratio 1000, 10 "
                        f"correlated RHS per request (1% perturbation, warm start), 200 GB "
                        f"budget, mixed precision to 1e-8"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3_512")
    ap.add_argument("--precond", choices=["mg", "jacobi"], default="mg")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed",
                    help="form of the timed solves (the other one is the comparison)")
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--cpu-iters", type=int, default=2000,
                    help="iterations of the oracle's SolverConjugate on the 64^3 system "
                         "timed for cpu_baseline (0 = skip)")
    ap.add_argument("--ref-iters", type=int, default=300,
                    help="iterations of the compiled reference (oracle/_ref/ref_driver, if "
                         "built) timed beside it (0 = skip)")
    ap.add_argument("--cpu-mg-iters", type=int, default=15,
                    help="iterations of the oracle's FP64 multigrid PCG at the main config "
                         "(0 = skip)")
    ap.add_argument("--c2", type=int, default=1,
                    help="run the 256^3 FP64 Jacobi-PCG leg (GB/s per iteration)")
    ap.add_argument("--c1", type=int, default=1,
                    help="run the 64^3 reference-CG leg and the CPU baseline")
    ap.add_argument("--no-profile", action="store_true",
                    help="no per-kernel events in the timed region")
    ap.add_argument("--batch-systems", type=int, default=96,
                    help="systems of the batched workload (0 = skip)")
    ap.add_argument("--serve-requests", type=int, default=32,
                    help="requests of the served workload (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=4.0, help="arrivals per second")
    ap.add_argument("--copy-gib", type=float, default=2.0,
                    help="buffer size of the plain-copy bandwidth measurement (0 = skip)")
    ap.add_argument("--only-copy", action="store_true",
                    help="run only the plain-copy measurement (profiler calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a batch of systems would shard by whole system across GPUs)", file=sys.stderr)
        return 2

    from aphros_amd import _lib as L
    from aphros_amd import synth
    from aphros_amd.devmem import DeviceBuffer
    from aphros_amd.solver import Poisson

    copy_gbps = None
    if args.copy_gib > 0:
        import ctypes as C
        g = C.c_double()
        nbytes = int(args.copy_gib * 2 ** 30)
        L.check(L.lib().aphpois_copy_bandwidth(0, nbytes, 10, C.byref(g)))
        copy_gbps = round(g.value, 1)
        log(f"[bench] plain copy of {nbytes} bytes: {copy_gbps} GB/s (read + write)")
    if args.only_copy:
        print(json.dumps({"copy_bytes": int(args.copy_gib * 2 ** 30), "copy_GBps": copy_gbps}))
        return 0

    precond = L.PRECOND_MG if args.precond == "mg" else L.PRECOND_JACOBI
    cfg = synth.CONFIGS[args.config]
    n = cfg["n"]
    t0 = time.time()
    expr, rho, f = synth.config(args.config)
    log(f"[bench] generated {args.config} ({n}^3) in {time.time() - t0:.1f}s")
    ncell = n ** 3
    xbuf = DeviceBuffer(ncell * 8)
    has_m = False  # the periodic pressure system: row sums vanish

    def run_form(precision, steps, warmup, profile, xref=None):
        s = Poisson((n, n, n), method=L.PCG, precision=precision, resmode=L.RES_REL_TRUE,
                    tol=1e-8, maxiter=200000, precond=precond)
        t = time.time()
        s.set_system(expr)
        log(f"[bench] precision={precision} set_system {time.time() - t:.2f}s "
            f"device_bytes={s.device_bytes() / 1e9:.2f} GB")
        if xref is not None:
            s.set_solution_ref(xref.reshape(n, n, n))
        plain = []
        for i in range(warmup):
            t = time.perf_counter()
            info = s.solve_device(xbuf.ptr)
            plain.append(time.perf_counter() - t)
            log(f"[bench]   warmup {i}: {plain[-1]:.3f}s iter={info['iter']} "
                f"relres={info['true_relres']:.3e}")
        if profile:
            s.set_profiling(True)
        lat, infos = [], []
        for i in range(steps):
            t = time.perf_counter()
            info = s.solve_device(xbuf.ptr)
            lat.append(time.perf_counter() - t)
            infos.append(info)
            log(f"[bench]   step {i}: {lat[-1]:.3f}s iter={info['iter']} "
                f"ru={info['refinements']} relres={info['true_relres']:.3e} "
                f"converged={info['converged']} max_diff={info['max_diff']:.3e}")
            if not info["converged"] or info["true_relres"] > 1e-8:
                raise RuntimeError("solve did not converge to 1e-8")
        kt = s.kernel_times() if profile else None
        li = s.launch_info()
        s.close()
        return lat, infos, kt, li, plain

    main_prec = L.MIXED if args.precision == "mixed" else L.FP64
    # the other form first: with the FP64 solution on the host, every mixed
    # solve reports its error against it (info.max_diff)
    fp64, x_other = None, None
    if args.fp64_steps > 0:
        lat64, inf64, _, _, _ = run_form(L.FP64 if main_prec == L.MIXED else L.MIXED,
                                         args.fp64_steps, 0, False)
        fp64 = {"solves_per_s": round(args.fp64_steps / sum(lat64), 4),
                "iters": [i["iter"] for i in inf64],
                "s_per_solve": round(sum(lat64) / args.fp64_steps, 4),
                "true_relres": [i["true_relres"] for i in inf64]}
        if main_prec == L.MIXED:
            x_other = xbuf.to_host(np.empty(ncell))
    lat, infos, kt, li, plain = run_form(main_prec, args.steps, args.warmup,
                                         not args.no_profile, xref=x_other)
    total = sum(lat)
    value = args.steps / total
    iters = [i["iter"] for i in infos]
    error_vs_fp64 = None
    if x_other is not None:
        scale = float(np.abs(x_other - x_other.mean()).max())
        md = max(i["max_diff"] for i in infos)
        error_vs_fp64 = {"max_abs_diff": md, "max_abs_fp64_solution": scale,
                         "relative": md / scale,
                         "what": "max |x_mixed - x_fp64| over the timed solves, means "
                                 "removed (info.max_diff, computed on the GPU)"}
        del x_other

    roofline, kernels = None, None
    if kt:
        per_iter = {"mg_presweep_l0": 1, "mg_postsweep1_l0": 1, "mg_postsweep2_l0": 1,
                    "mg_post_l0": 1, "mg_pre_l0": 1}
        elem = 4 if main_prec == L.MIXED else 8
        coef = 2 if precond == L.PRECOND_MG and main_prec == L.MIXED else elem
        kernels = kernel_table(kt, ncell, elem, coef, has_m, per_iter,
                               jacobi=precond != L.PRECOND_MG, vx=li["vx"],
                               xh_vecs=xh_vectors([(i["iter"], i["refinements"]) for i in infos]))
        if kernels:
            # the dominant kernel: the most device time per iteration
            key = max(kernels, key=lambda k: kernels[k]["ms_per_iter"])
            k = kernels[key]
            achieved = k["GBps"]
            traffic = measured_traffic(args.config, main_prec, precond, k["kernel"], li, L)
            roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": traffic, "kernel": k["kernel"],
                        "alg_bytes_per_launch": k["alg_bytes_per_launch"],
                        "element": "float" if main_prec == L.MIXED else "double",
                        "avg_ms": k["avg_ms"], "launches": k["launches"],
                        "plain_copy_GBps": copy_gbps}
        # level-0 kernels of one iteration together: algorithmic bytes over
        # their summed average durations
        l0 = [v for v in kernels.values() if isinstance(v, dict)]
        if l0:
            kernels["level0_alg_GBps"] = round(
                sum(v["alg_bytes_per_launch"] for v in l0) /
                (sum(v["avg_ms"] for v in l0) * 1e-3) / 1e9, 1)
        if precond == L.PRECOND_MG and len(kt) > 12:
            kernels["vcycle_total_ms"] = round(kt[8][0], 5)
            kernels["mg_coarse_levels_ms_per_vcycle"] = round(kt[11][0], 5)
            if kt[12][1] > 0:
                kernels["mg_bottom_levels_ms"] = round(kt[12][0], 5)
        ms_it = 1e3 * total / max(1, sum(iters))
        kernels["ms_per_iter_wall"] = round(ms_it, 5)

    cpu_mg = None
    if args.cpu_mg_iters > 0:
        from oracle import oracle as O
        e4 = expr.reshape(n, n, n, 8)
        # iterations of the FP64 GPU solve
        ref_iters = fp64["iters"][0] if (fp64 and main_prec == L.MIXED) else iters[0]
        t = time.perf_counter()
        if precond == L.PRECOND_MG:
            _, it, _, setup = O.mgpcg(e4, tol=0.0, maxiter=-args.cpu_mg_iters,
                                      remove_mean=False, w=L.mg_w_auto(ncell))
            name = "multigrid PCG"
        else:
            _, it, _ = O.pcg_jacobi(e4, tol=0.0, maxiter=-args.cpu_mg_iters, remove_mean=False)
            setup, name = 0.0, "Jacobi-PCG"
        dt = time.perf_counter() - t
        s_per_iter = (dt - setup) / it
        cpu_mg = {"value": round(1.0 / (setup + s_per_iter * ref_iters), 6),
                  "unit": f"solves/s @{n}^3", "cores": O.num_threads(), "kind": "port",
                  "sample": f"{it} FP64 {name} iterations of the oracle (OpenMP) on the "
                            f"{n}^3 system ({dt:.1f}s: setup {setup:.1f}s, {s_per_iter:.3f} "
                            f"s/iter), scaled to the {ref_iters} iterations the FP64 GPU "
                            f"solve needs; the reference has no preconditioned CPU solver, "
                            f"so this is this build's algorithm on the CPU"}
    del expr, rho, f

    c2 = run_c2(args) if args.c2 else None
    c1, cpu = run_c1(args) if args.c1 else (None, None)
    batched = run_batched(args, precond) if args.batch_systems > 0 else None
    served = run_served(args, precond) if args.serve_requests > 0 else None

    pc_name = (f"multigrid (V(2,2), W on the top {L.mg_w_auto(ncell)} levels, "
               f"the automatic choice at {ncell} cells)"
               if precond == L.PRECOND_MG else "Jacobi")
    form = ("mixed-precision (FP32 inner, FP64 reliable updates)" if main_prec == L.MIXED
            else "FP64")
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if main_prec == L.MIXED else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {n}^3 periodic, {cfg['nspheres']} spheres "
                               f"r in [{cfg['rmin']},{cfg['rmax']}], density ratio "
                               f"{cfg['ratio']:g}, seed {cfg['seed']}; {form} "
                               f"{pc_name}-preconditioned CG to a 1e-8 true relative "
                               f"residual, zero initial guess",
                   "cells": ncell, "iters": iters,
                   "p50_s": round(float(np.percentile(lat, 50)), 4),
                   "p99_s": round(float(np.percentile(lat, 99)), 4),
                   "unprofiled_warmup_s": [round(v, 4) for v in plain],
                   "launch": li},
        "fp64" if main_prec == L.MIXED else "mixed": fp64,
        "mixed_speedup_vs_fp64": (None if not fp64 else
                                  round(value / fp64["solves_per_s"], 3) if main_prec == L.MIXED
                                  else round(fp64["solves_per_s"] / value, 3)),
        "error_vs_fp64": error_vs_fp64,
        "roofline": roofline,
        "kernels": kernels,
        "c2_fp64_jacobi_pcg": c2,
        "c1_reference_cg": c1,
        "batched": batched,
        "served": served,
        "cpu_baseline": cpu,
        "cpu_512_multigrid_port": cpu_mg,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
