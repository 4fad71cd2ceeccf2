#!/usr/bin/env python3
"""Throughput of the FUSS explicit solver step on one MI355X.

Workload: the steady 3-D cube at 256^3 cells (0.1 m, seeded smooth per-axis
stretching, 32-knot k(T) table resampled to 1 K rows, qvol = 1e6 W/m^3,
convective 303 walls h = 50 W/m^2K, T_inf = 300 K), SSP-RK3 with local time
stepping, implicit residual smoothing beta = 0.5 (2 Jacobi sweeps per
direction), 3 multigrid levels allocated.  One step = one Explicit_Step on
the finest level: local dt + global min, 3 x (ghost fill, fused
flux/BC/newstate stage, 3 IRS line passes), residual sum -- inputs resident
in HBM.  value = interior cells x steps / wall time (cell-updates/s).

Also reported: the fused stage kernel's achieved bandwidth (algorithmic
bytes per launch / average launch time from HIP events on the library's
stream) against the MI355X peak of 8 TB/s, and the oracle (scalar C port of
the reference, FP64) timed on the host on a bounded sample of the same
workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the stage kernel measured with rocprofv3 PMC
# (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, calibrated on a 1 GiB device copy);
# source: profiles/explicit_step_cube256_fp64.md (256^3 cube, FP64).
STAGE_TRAFFIC_MEASURED = {"k_stage3d<false, true, true, double>": 1023e6 + 142e6}


def device_sync():
    hip = ctypes.CDLL("libamdhip64.so")
    if hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def calibration_copy(nbytes=1 << 30, reps=3):
    """Device-to-device copies of a known size: the reference kernel for
    scaling the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # device to device
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def cpu_baseline(n, iters):
    """Oracle (scalar FP64 C port) on the host: `iters` finest-level
    iterations of the same cube at n^3 (single level, same numerics)."""
    from fuss_amd import cases
    from oracle import pyoracle

    c = cases.cube3d(n=n, mgl=1)
    o = pyoracle.Oracle(c)
    t0 = time.perf_counter()
    for _ in range(iters):
        o.step(1)
    dt = time.perf_counter() - t0
    return {"value": c.ncell * iters / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle (scalar C, FP64), {iters} iteration(s) of the {n}^3 cube, RK3+IRS, "
                      f"{dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-iters", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calib-copy", action="store_true", help="also run 1 GiB D2D copies (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); "
              "multi-GPU sharding of the path is design-only", file=sys.stderr)
        sys.exit(2)

    from fuss_amd import cases, gpu, perfmodel

    print(f"[bench] building the {args.n}^3 cube", file=sys.stderr, flush=True)
    case = cases.cube3d(n=args.n, mgl=3)
    t0 = time.perf_counter()
    s = gpu.Solver(case)
    print(f"[bench] setup {time.perf_counter()-t0:.1f} s", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        s.explicit_step(1)
    device_sync()
    s.L.fuss_profile(s.h, 1)
    device_sync()
    t0 = time.perf_counter()
    res = None
    for _ in range(args.steps):
        res = s.explicit_step(1)
    device_sync()
    wall = time.perf_counter() - t0
    ncl = gpu.C.c_int * 8
    ms = (ctypes.c_double * 8)()
    cnt = ncl()
    s.L.fuss_profile_read(s.h, 8, ms, cnt)
    s.L.fuss_profile(s.h, 0)
    ncell = case.ncell
    value = ncell * args.steps / wall
    print(f"[bench] {args.steps} steps in {wall:.3f} s, residual {np.sqrt(res.res_sq):.4e}", file=sys.stderr,
          flush=True)
    names = ["dt", "ghost", "stage", "irs", "irs_update", "reduce", "bc"]
    kern = {names[q]: {"ms_avg": ms[q] / cnt[q], "launches": cnt[q]} for q in range(7) if cnt[q]}
    stage_b = perfmodel.stage_bytes(ndir=3, tang=False, irs=True, prim=True) * ncell
    stage_ms = ms[2] / cnt[2]
    achieved = stage_b / (stage_ms * 1e-3) / 1e9
    out = {
        "metric": "cell-updates/s/GPU + time-to-1e-10 residual, FP32-refined vs plain FP64",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"steady cube {args.n}^3, plain FP64, RK3 + LTS + IRS(jacobi2, 0.5), "
                               "qvol 1e6, 303 h=50, k(T) 32-knot table, 3 MG levels; step = one "
                               "finest-level Explicit_Step",
                   "cells": ncell, "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": STAGE_TRAFFIC_MEASURED["k_stage3d<false, true, true, double>"] if args.n == 256 else None,
                     "kernel": "k_stage (fused interior+BC fluxes, scaled residual)",
                     "bytes_per_cell": perfmodel.stage_bytes(ndir=3, tang=False, irs=True, prim=True),
                     "ms_avg": stage_ms},
        "kernels": kern,
        "iteration_bytes_per_cell": perfmodel.iteration_bytes(3, 3, False, True, True),
    }
    s.close()
    if args.calib_copy:
        calibration_copy()
    if not args.no_cpu_baseline:
        print(f"[bench] cpu baseline: oracle {args.cpu_n}^3 x {args.cpu_iters}", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_iters)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
