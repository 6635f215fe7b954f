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

# algorithmic bytes per cell of the dominant kernel (PCG spmv with the fused
# p update): read r, p_in, S, kx, ky, kz; write p_out, q  -> 8 words per cell
SPMV_WORDS = 8


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3_512")
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--cpu-iters", type=int, default=8,
                    help="iterations of the CPU oracle timed for cpu_baseline (0 = skip)")
    ap.add_argument("--no-profile", action="store_true",
                    help="no per-kernel events in the timed region")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the systems of a batch shard by whole system, see DESIGN.md)", file=sys.stderr)
        return 2

    from aphros_amd import _lib as L
    from aphros_amd import synth
    from aphros_amd.devmem import DeviceBuffer
    from aphros_amd.solver import Poisson

    cfg = synth.CONFIGS[args.config]
    n = cfg["n"]
    t0 = time.time()
    expr, rho, f = synth.config(args.config)
    log(f"[bench] generated {args.config} ({n}^3) in {time.time() - t0:.1f}s")
    ncell = n ** 3
    xbuf = DeviceBuffer(ncell * 8)

    def run_form(precision, steps, warmup, profile):
        s = Poisson((n, n, n), method=L.PCG, precision=precision, resmode=L.RES_REL_TRUE,
                    tol=1e-8, maxiter=200000)
        t = time.time()
        s.set_system(expr)
        log(f"[bench] precision={precision} set_system {time.time() - t:.2f}s "
            f"device_bytes={s.device_bytes() / 1e9:.2f} GB")
        for i in range(warmup):
            info = s.solve_device(xbuf.ptr)
            log(f"[bench]   warmup {i}: iter={info['iter']} relres={info['true_relres']:.3e}")
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
                f"converged={info['converged']}")
            if not info["converged"]:
                raise RuntimeError("solve did not converge")
        kt = s.kernel_times() if profile else None
        li = s.launch_info()
        s.close()
        return lat, infos, kt, li

    lat, infos, kt, li = run_form(L.MIXED, args.steps, args.warmup, not args.no_profile)
    total = sum(lat)
    value = args.steps / total
    iters = [i["iter"] for i in infos]

    fp64 = None
    if args.fp64_steps > 0:
        lat64, inf64, kt64, _ = run_form(L.FP64, args.fp64_steps, 0, False)
        fp64 = {"solves_per_s": args.fp64_steps / sum(lat64),
                "iters": [i["iter"] for i in inf64],
                "s_per_solve": sum(lat64) / args.fp64_steps}

    roofline = None
    if kt:
        spmv_ms, spmv_n = kt[0]
        upd_ms, upd_n = kt[2]
        bytes_per_launch = SPMV_WORDS * 4 * ncell
        achieved = bytes_per_launch / (spmv_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": None,
                    "kernel": "k_pcg_spmv<float,4>",
                    "alg_bytes_per_launch": bytes_per_launch,
                    "avg_ms": round(spmv_ms, 5), "launches": spmv_n,
                    "update_kernel_avg_ms": round(upd_ms, 5),
                    "update_kernel_GBps": round(6 * 4 * ncell / (upd_ms * 1e-3) / 1e9, 1)}

    cpu = None
    if args.cpu_iters > 0:
        from oracle import oracle as O
        e4 = expr.reshape(n, n, n, 8)
        ref_iters = fp64["iters"][0] if fp64 else iters[0]
        t = time.perf_counter()
        _, it, _ = O.pcg_jacobi(e4, tol=0.0, maxiter=-args.cpu_iters, remove_mean=False)
        dt = time.perf_counter() - t
        s_per_iter = dt / it
        cpu = {"value": 1.0 / (s_per_iter * ref_iters), "unit": "solves/s",
               "cores": O.num_threads(), "kind": "port",
               "sample": f"{it} FP64 Jacobi-PCG iterations of the oracle on the {n}^3 system "
                         f"({dt:.1f}s, {s_per_iter:.3f} s/iter), scaled by the {ref_iters} "
                         f"iterations the FP64 GPU solve needs"}

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
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {n}^3 periodic, {cfg['nspheres']} spheres "
                               f"r in [{cfg['rmin']},{cfg['rmax']}], density ratio "
                               f"{cfg['ratio']:g}, seed {cfg['seed']}; mixed-precision "
                               f"Jacobi-PCG (FP32 inner, FP64 reliable updates) to a 1e-8 "
                               f"true relative residual, zero initial guess",
                   "cells": ncell, "iters": iters,
                   "p50_s": float(np.percentile(lat, 50)),
                   "p99_s": float(np.percentile(lat, 99)),
                   "launch": li},
        "fp64": fp64,
        "mixed_speedup_vs_fp64": round(value / fp64["solves_per_s"], 3) if fp64 else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
