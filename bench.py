#!/usr/bin/env python3
"""Throughput of the FUSS explicit solver step on one MI355X.

Workload: the steady 3-D cube at 256^3 cells (0.1 m, seeded smooth per-axis
stretching, 32-knot k(T) table resampled to 1 K rows, qvol = 1e6 W/m^3,
convective 303 walls h = 50 W/m^2K, T_inf = 300 K), four-stage RK with local
time stepping, implicit residual smoothing beta = 0.5 (2 Jacobi sweeps per
direction), 3 multigrid levels allocated.  One step = one Explicit_Step on
the finest level: local dt + global min, 4 x (ghost fill, fused
flux/BC/newstate stage, 3 IRS line passes), residual sum -- inputs resident
in HBM.  value = interior cells x steps / wall time (cell-updates/s).

Time to residual 1e-10 (||T^{n+1}-T^n||_2 of an FP64 Explicit_Step, the
reference's stopping test), from T = 300 K: measured for the FP64 defect
correction with FP32 inner sweeps (fuss_refine, 3-level correction V-cycle as
the configuration states, and full depth) and with FP64 inner sweeps at
256^3; measured for both the plain FP64 march and the FP32-refined solve at
64^3 (--plain-conv-n), whose ratio is the measured speedup.  At 256^3 the
plain march needs ~3e6 iterations (about an hour): its time there is
reported as an estimate (measured time per iteration x an iteration count
extrapolated from measured plain runs at 32^3, 48^3, 64^3).

Also reported: per-kernel achieved bandwidth of the stage kernels and the
fused IRS pass (algorithmic bytes per launch / average launch time from HIP
events on the library's stream) against the MI355X peak of 8 TB/s; the
batched sweep and the transient lattice in FP64 and FP32 bulk; the adjoint;
and the oracle's OpenMP build on the host cores on a bounded sample of the
same workload.

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
# (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, calibrated on a 1 GiB device copy) on
# the 256^3 cube, FP64: the LDS-ring stage kernel (stages 2-4 of a steady
# step: no dt) on 64 x 4 or 64 x 8 columns of kc planes, one material,
# primitive variables, IRS on; rect = per-axis metric factors (rectilinear
# level).


def stage_kernel(rect, kc, rows):
    """The stage kernel of stages 2-4 on the cube (one material, uniform qvol,
    primitive, IRS): the LDS ring with 4 or 8 rows, or register columns."""
    if rows == 1:
        return f"fuss::k_stage3r<true, true, true, true, false, false, double, {kc}>"
    return f"k_stage3p<true, true, true, {'true' if rect else 'false'}, false, double, {kc}, {rows}>"


STAGE_TRAFFIC_MEASURED = {stage_kernel(True, kc, rows): fetch * 1024 * 2.0 + 131072 * 1024 * 1.0
                          for kc, rows, fetch in ((4, 4, 153140), (8, 4, 154981), (16, 4, 149399), (32, 4, 143788),
                                                  (64, 4, 140721), (32, 8, 143789), (64, 8, 140803))}
# Plain FP64 path (four-stage RK + LTS + IRS 0.5, 3-level sequencing) run to
# residual 1e-10 on the GPU: iterations per cube size (bench runs with
# --plain-conv-n 32, 48, 64 and, with --n 128, 128).  Least-squares fit of
# log(iters) on log(n) through the four: slope PLAIN_EXP; the 256^3 count is
# extrapolated from the largest measured cube: 826104 (n/128)^PLAIN_EXP.
PLAIN_ITERS_MEASURED = {32: 64506, 48: 135672, 64: 230218, 128: 826104}
PLAIN_EXP = 1.840


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


def stream_copy_gbs(nbytes=1 << 30, reps=5):
    """Achievable HBM bandwidth: the best of `reps` device-to-device copies of
    nbytes (read + write), timed with HIP events; GB/s of read + written bytes."""
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipEventCreate(ctypes.byref(e0))
    hip.hipEventCreate(ctypes.byref(e1))
    best = 0.0
    for _ in range(reps):
        hip.hipEventRecord(e0, None)
        assert hip.hipMemcpyAsync(b, a, ctypes.c_size_t(nbytes), 3, None) == 0  # device to device
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float(0.0)
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        best = max(best, 2.0 * nbytes / (ms.value * 1e-3) / 1e9)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return best


def cpu_baseline(n, iters):
    """The oracle's OpenMP build (the C restatement of the reference, FP64,
    its per-cell and per-line loops shared over the host's cores as the
    reference's OpenMP loops are) on the host: `iters` finest-level
    iterations of the same cube at n^3 (single level, same numerics)."""
    from fuss_amd import cases
    from oracle import pyoracle

    c = cases.cube3d(n=n, mgl=1, n_rk=4)
    o = pyoracle.Oracle(c, omp=True)
    threads = int(o.L.ob_threads())
    t0 = time.perf_counter()
    for _ in range(iters):
        o.step(1)
    dt = time.perf_counter() - t0
    return {"value": c.ncell * iters / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle, OpenMP build ({threads} threads; {os.cpu_count()} host CPUs visible), FP64, "
                      f"{iters} iteration(s) of the {n}^3 cube, RK4 + LTS + IRS 0.5, {dt:.1f} s wall"}


def refined_run(s, T0, prec, max_levels, label, max_outer=300, coarse_cg=0):
    """FP64 defect correction with `prec` inner sweeps from T0; wall time of
    the second of two identical runs (the first loads the modules)."""
    from fuss_amd import gpu

    rp = gpu.refine_params(precision=prec, max_outer=max_outer, res_threshold=1e-10, max_levels=max_levels,
                           coarse_cg=coarse_cg)
    gpu._chk(s.L.fuss_refine_setup(s.h, gpu.C.byref(rp)), "fuss_refine_setup")
    s.set_T(T0)
    s.refine(rp, setup=False)
    s.set_T(T0)
    device_sync()
    t0 = time.perf_counter()
    r, hist = s.refine(rp, setup=False)
    device_sync()
    dt = time.perf_counter() - t0
    print(f"[bench] {label}: {dt*1e3:.1f} ms, {r.outer} outer, {r.levels} levels, residual {r.res:.3e}",
          file=sys.stderr, flush=True)
    return dt, r, hist, s.get("T")


def time_to_converge(s, case, args, ms_step):
    """Wall time to the reference's residual 1e-10 (||T^{n+1} - T^n||_2 of an
    FP64 Explicit_Step) from T = 300 K.

    Measured: the FP64 defect correction with FP32 and with FP64 inner
    sweeps at the benchmark size (correction V-cycle to full depth, and a
    3-level V-cycle -- the configuration's multigrid depth -- whose coarsest
    64^3 level is solved by 80 conjugate-gradient iterations; the same cut
    with that level only smoothed is reported too: it cannot carry the global
    heat-up mode, and 30 outer iterations leave the residual far above the
    threshold), and at a smaller cube both the plain FP64 march (the
    reference's method, 3-level coarse-to-fine sequencing) and the
    FP32-refined solve, so the speedup there is a ratio of two measurements.  The plain march at the benchmark
    size takes ~3e6 iterations; its time there is an estimate (measured time
    per iteration x an iteration count extrapolated from measured runs)."""
    from fuss_amd import cases, gpu

    T0 = case.blocks[0].T
    out = {"threshold": 1e-10, "from": "initial state T = 300 K"}
    finals = {}
    for prec, name, lv, mo, cg in ((gpu.FUSS_FP32, "fp32_refined", 0, 300, 0),
                                   (gpu.FUSS_FP64, "fp64_refined", 0, 300, 0),
                                   (gpu.FUSS_FP32, "fp32_refined_3_levels", 3, 300, 80),
                                   (gpu.FUSS_FP32, "fp32_refined_3_levels_smoothed", 3, 30, 0)):
        dt, r, hist, T = refined_run(s, T0, prec, lv, name, mo, cg)
        finals[name] = T
        out[name] = {"seconds": dt, "converged": bool(r.converged), "outer": r.outer, "vcycles": r.vcycles,
                     "levels": r.levels, "coarsest": f"{cg} CG iterations" if cg else "direct or smoothed",
                     "final_residual": r.res, "defect_W": r.defect, "measured": True,
                     "residual_history": [float(x) for x in hist]}
    a, b = finals["fp32_refined"], finals["fp64_refined"]
    out["fp32_vs_fp64_refined_max_rel"] = float(np.max(np.abs(a - b) / np.abs(b)))
    out["speedup_fp32_vs_fp64_refined"] = out["fp64_refined"]["seconds"] / out["fp32_refined"]["seconds"]
    m = args.plain_conv_n
    if m:
        # measured at m^3: the plain FP64 march to 1e-10 and the FP32-refined
        # solve of the same cube
        c = cases.cube3d(n=m, mgl=3, n_rk=args.rk)
        sp = gpu.Solver(c)
        sp.autotune(1, steps=3)
        device_sync()
        t0 = time.perf_counter()

        def progress(done, lv, res):  # (printed between fuss_run calls of 50 000 steps)
            print(f"[bench] plain FP64 {m}^3: {done} iterations, level {lv}, residual {res:.3e}",
                  file=sys.stderr, flush=True)

        nit, hist = sp.solve(chunk=50000, progress=progress)
        device_sync()
        t_plain = time.perf_counter() - t0
        Tp = sp.get("T")
        sp.close()
        sr = gpu.Solver(cases.cube3d(n=m, mgl=1, n_rk=args.rk))
        t_ref, r, _, Tr = refined_run(sr, cases.cube3d(n=m, mgl=1, n_rk=args.rk).blocks[0].T, gpu.FUSS_FP32, 0,
                                      f"fp32_refined {m}^3")
        sr.close()
        out["measured_at"] = {"n": m, "plain_fp64": {"seconds": t_plain, "iterations": int(nit),
                                                      "final_residual": float(hist[-1, 2])},
                              "fp32_refined": {"seconds": t_ref, "outer": r.outer, "final_residual": r.res},
                              "speedup_fp32_refined_vs_plain": t_plain / t_ref,
                              "fp32_refined_vs_plain_max_rel": float(np.max(np.abs(Tr - Tp) / np.abs(Tp)))}
        print(f"[bench] plain FP64 {m}^3 to 1e-10: {nit} iterations in {t_plain:.1f} s; FP32-refined "
              f"{t_ref*1e3:.1f} ms; speedup {t_plain / t_ref:.0f}", file=sys.stderr, flush=True)
    iters = PLAIN_ITERS_MEASURED[128] * (args.n / 128.0) ** PLAIN_EXP
    plain_s = iters * ms_step * 1e-3
    out["fp64_plain"] = {"measured": False, "seconds_est": plain_s, "iterations_est": iters,
                         "ms_per_iteration": ms_step, "iterations_measured": PLAIN_ITERS_MEASURED,
                         "fit": f"826104 (n/128)^{PLAIN_EXP}",
                         "measured_128": {"iterations": 826104, "seconds": 151.9,
                                          "fp32_refined_seconds": 0.0115, "speedup_measured": 13224,
                                          "source": "bench.py --n 128 --plain-conv-n 128"},
                         "note": "estimate: measured ms per iteration x extrapolated iteration count"}
    out["speedup_fp32_refined_vs_plain_est"] = plain_s / out["fp32_refined"]["seconds"]
    return out


def adjoint_check(args):
    """Heat-flux inverse problem on the 512^2 transient plate: forward and
    reverse-sweep wall times, the adjoint gradient against central finite
    differences (h = 1e-6) on 8 seeded directions, the FP32-field adjoint
    against the FP64 one, and a short L-BFGS recovery of q and c."""
    import torch

    from fuss_amd import adjoint, cases, gpu

    case, info = cases.flux_plate(n=args.adj_n, nseg=64, nprobe=32)
    nseg = info["q_true"].size
    prob = adjoint.FluxProblem(case, info["seg_entries"], info["probes"], args.adj_steps, 20, gpu.FUSS_FP64)
    obs = prob.forward(info["q_true"], 1.0)
    q0, c0 = np.full(nseg, 3.0e4), 1.15
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = prob.forward(q0, c0)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    g = 2.0 * (hist - obs)
    t0 = time.perf_counter()
    gq, gk = prob.backward(g)
    t_bwd = time.perf_counter() - t0
    grad = np.concatenate([gq, [gk]])
    loss0 = float(((hist - obs) ** 2).sum())
    rng = np.random.default_rng(2024)
    scale = np.concatenate([np.full(nseg, 1.0e4), [1.0]])
    h = 1e-6
    errs = []
    for _ in range(8):
        v = rng.normal(size=nseg + 1) * scale
        Lp = float(((prob.forward(q0 + h * v[:-1], c0 + h * v[-1]) - obs) ** 2).sum())
        Lm = float(((prob.forward(q0 - h * v[:-1], c0 - h * v[-1]) - obs) ** 2).sum())
        fd = (Lp - Lm) / (2 * h)
        ad = float(grad @ v)
        errs.append(abs(fd - ad) / abs(ad))
    p32 = adjoint.FluxProblem(case, info["seg_entries"], info["probes"], args.adj_steps, 20, gpu.FUSS_FP32)
    h32 = p32.forward(q0, c0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gq32, gk32 = p32.backward(2.0 * (h32 - obs))
    t_bwd32 = time.perf_counter() - t0
    p32.close()
    grad32 = np.concatenate([gq32, [gk32]])
    x = torch.tensor(np.concatenate([q0 / 1e4, [c0]]), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.LBFGS([x], lr=1.0, max_iter=args.adj_lbfgs, line_search_fn="strong_wolfe")

    def closure():
        opt.zero_grad()
        hh = adjoint.ProbeHistory.apply(x[:nseg] * 1e4, x[nseg], prob)
        loss = ((hh - obs) ** 2).sum()
        loss.backward()
        return loss

    t0 = time.perf_counter()
    opt.step(closure)
    t_inv = time.perf_counter() - t0
    loss1 = float(closure().detach())
    xq = x.detach().numpy()
    prob.close()
    out = {"workload": f"transient plate {args.adj_n}^2, 64 flux segments + k scale, 32 probes, "
                       f"{args.adj_steps} SSP-RK3 steps (global dt), checkpoints every 20",
           "forward_s": t_fwd, "backward_s": t_bwd, "backward_fp32_fields_s": t_bwd32,
           "fd_directions": 8, "fd_h": h, "fd_max_rel_err": max(errs), "fd_rel_errs": errs,
           "fp32_vs_fp64_grad_rel": float(np.linalg.norm(grad32 - grad) / np.linalg.norm(grad)),
           "inverse": {"lbfgs_max_iter": args.adj_lbfgs, "seconds": t_inv, "loss_start": loss0, "loss_end": loss1,
                       "q_rel_err": float(np.linalg.norm(xq[:nseg] * 1e4 - info["q_true"]) /
                                          np.linalg.norm(info["q_true"])),
                       "k_scale": float(xq[nseg]), "k_scale_true": 1.0}}
    print(f"[bench] adjoint: fwd {t_fwd:.3f} s, bwd {t_bwd:.3f} s, FD max rel err {max(errs):.2e}, "
          f"inverse loss {loss0:.3e} -> {loss1:.3e}", file=sys.stderr, flush=True)
    return out


def sweep_check(args):
    """Parameter sweep of independent steady cubes (k, h, qvol drawn per case,
    uneven 32^3-96^3 blocks) as one batched context: the plain FP64
    Explicit_Step over all cases in one launch per pass, and the steady solve
    of every case to its own residual 1e-10 by FP64 defect correction with
    FP32 (and, as the FP64 solution, FP64) inner sweeps."""
    from fuss_amd import cases, gpu

    t0 = time.perf_counter()
    case = cases.sweep(ncase=args.sweep_cases, lo=32, hi=96, mgl=3, n_rk=args.rk)
    gen_s = time.perf_counter() - t0
    s = gpu.Solver(case)
    setup_s = time.perf_counter() - t0
    ncell = case.ncell
    tuned = s.autotune(1, steps=3)
    for _ in range(2):
        s.explicit_step(1)
    device_sync()
    t0 = time.perf_counter()
    for _ in range(args.sweep_steps):
        s.explicit_step(1)
    device_sync()
    step_s = (time.perf_counter() - t0) / args.sweep_steps
    T0 = [b.T for b in case.blocks]
    out = {"workload": f"{args.sweep_cases} independent steady cubes, dims 32..96 (multiples of 4), "
                       f"k 5-200, h 10-500, qvol 0-1e6, RK{args.rk} + IRS 0.5, one context",
           "cells": ncell, "setup_s": setup_s,
           # (numpy generation of the synthetic cases, then the C-ABI set-up:
           # block and BC uploads, device metrics, BC records, levels)
           "case_generation_s": gen_s, "solver_setup_s": setup_s - gen_s,
           "plain_fp64_ms_per_step": step_s * 1e3,
           "tiles": {"column_planes": tuned[0], "stage_kernel": "row march" if tuned[1] else "LDS ring",
                     "stage_rows": s.tiles(1)[2], "irs_rows": s.tiles(1)[3]},
           "plain_fp64_cell_updates_per_s": ncell / step_s}
    finals = {}
    for prec, name in ((gpu.FUSS_FP32, "fp32_refined"), (gpu.FUSS_FP64, "fp64_refined")):
        rp = gpu.refine_params(precision=prec, max_outer=80, res_threshold=1e-10, per_block=True)
        gpu._chk(s.L.fuss_refine_setup(s.h, gpu.C.byref(rp)), "fuss_refine_setup")
        for b, T in enumerate(T0):
            s.set_T(T, 1, b + 1)
        device_sync()
        t0 = time.perf_counter()
        r, _ = s.refine(rp, setup=False)
        device_sync()
        dt = time.perf_counter() - t0
        conv, res = s.refine_blocks()
        finals[name] = s.T()
        out[name] = {"seconds": dt, "all_converged": bool(r.converged), "outer": r.outer, "vcycles": r.vcycles,
                     "levels": r.levels, "max_case_residual": float(res.max()),
                     "case_outer_at_convergence": {"min": int(conv.min()), "median": float(np.median(conv)),
                                                   "max": int(conv.max())}}
        print(f"[bench] sweep {name}: {dt*1e3:.1f} ms, {r.outer} outer, max case residual {res.max():.2e}",
              file=sys.stderr, flush=True)
    errs = np.array([float(np.max(np.abs(a - b) / np.abs(b))) for a, b in zip(finals["fp32_refined"],
                                                                               finals["fp64_refined"])])
    out["per_case_max_rel_err_fp32_vs_fp64"] = {"max": float(errs.max()), "median": float(np.median(errs))}
    s.close()
    return out


def lattice_check(args):
    """Transient 4x4x4 block lattice (uneven per-slab dims 16-128, 1-to-1
    connections, curvilinear, 3 tabulated materials, 302/301/304 walls),
    global-dt RK4 for `lattice_steps` steps through fuss_run (captured
    graphs): plain FP64, and FP32 bulk (FUSS_FP32: state, fluxes and update in
    FP32; dt and the residual reductions in FP64) with its error against the
    FP64 state after the same steps."""
    from fuss_amd import cases, gpu

    out = {"workload": "transient 4x4x4 lattice of 64 connected curvilinear blocks (slab dims 16-128), "
                       "3 materials, 302/301/304 walls, RK4 global dt"}
    states = {}
    for prec, name in ((gpu.FUSS_FP64, "fp64"), (gpu.FUSS_FP32, "fp32_bulk")):
        t0 = time.perf_counter()
        case = cases.lattice(nl=4, lo=16, hi=128, n_rk=4, time_accurate=True)
        s = gpu.Solver(case, precision=prec)
        setup_s = time.perf_counter() - t0
        if args.no_graphs:
            s.graphs(False)
        tuned = s.autotune(1, steps=3)
        s.run_level(1, 30, 30, -1.0, 1e300, hist=False)   # warm-up (graph capture, module load)
        s.set_T_all([b.T for b in case.blocks])
        s.set_clock(1, 0.0, 0)
        device_sync()
        t0 = time.perf_counter()
        k, _, _, r = s.run_level(1, args.lattice_steps, args.lattice_steps, -1.0, 1e300, hist=False)
        device_sync()
        dt = time.perf_counter() - t0
        states[name] = [s.get("T", 1, b + 1) for b in range(case.nb)]
        out["cells"], out["blocks"], out["steps"] = case.ncell, case.nb, k
        out[name] = {"seconds": dt, "setup_s": setup_s, "cell_updates_per_s": case.ncell * k / dt,
                     "simulated_time_s": r.time,
                     "tiles": {"column_planes": tuned[0], "stage_kernel": "row march" if tuned[1] else "LDS ring"}}
        print(f"[bench] lattice {name}: {case.ncell} cells x {k} steps in {dt:.2f} s", file=sys.stderr, flush=True)
        s.close()
    err = max(float(np.max(np.abs(a - b) / np.abs(b))) for a, b in zip(states["fp32_bulk"], states["fp64"]))
    out["fp32_bulk"]["max_rel_err_vs_fp64"] = err
    out["fp32_bulk"]["speedup_vs_fp64"] = out["fp64"]["seconds"] / out["fp32_bulk"]["seconds"]
    print(f"[bench] lattice FP32 bulk vs FP64 after {out['steps']} steps: max rel err {err:.2e}", file=sys.stderr,
          flush=True)
    return out


def fp32_step_check(args, case, kc, stage_rows, irs_rows, s64):
    """The same cube steps in an FP32 context (FUSS_FP32: state, fluxes,
    update and IRS in FP32; dt formed in FP64; FP64 reductions), on the FP64
    run's tiles: ms per step, per-kernel bandwidth at FP32 bytes, and the
    state's difference from the FP64 steps after the same count."""
    from fuss_amd import gpu, perfmodel

    sd = gpu.Solver(case)
    sf = gpu.Solver(case, precision=gpu.FUSS_FP32)
    for s in (sd, sf):
        s.set_tiles(1, kc=kc, row_march=0, stage_rows=stage_rows, irs_rows=irs_rows)
        s.run_level(1, args.steps, args.steps, -1.0, 1e300, hist=False)
    device_sync()
    t0 = time.perf_counter()
    sf.run_level(1, args.steps, args.steps, -1.0, 1e300, hist=False)
    device_sync()
    wall = time.perf_counter() - t0
    sd.run_level(1, args.steps, args.steps, -1.0, 1e300, hist=False)
    Td, Tf = sd.get("T"), sf.get("T")
    sf.L.fuss_profile(sf.h, 1)
    for _ in range(args.steps):
        sf.explicit_step(1)
    device_sync()
    ms = (ctypes.c_double * 8)()
    cnt = (ctypes.c_int * 8)()
    sf.L.fuss_profile_read(sf.h, 8, ms, cnt)
    sf.L.fuss_profile(sf.h, 0)
    names = ["dt", "ghost", "stage", "irs", "irs_update", "reduce", "bc", "stage_dt"]
    kern = {names[q]: {"ms_avg": ms[q] / cnt[q], "launches": cnt[q]} for q in range(8) if cnt[q]}
    rect = sf.rect_info(1)[0]
    kb = {"stage": perfmodel.stage_bytes(ndir=3, tang=False, irs=True, prim=True, fp=4, mat=False, qvol=False,
                                         rect=rect),
          "stage_dt": perfmodel.stage_bytes(ndir=3, tang=False, irs=True, prim=True, fp=4, mat=False, qvol=False,
                                            rect=rect),
          "irs_update": perfmodel.irs_fused_bytes(4, mat=False, stage_state=False)}
    for kname, bpc in kb.items():
        if kname in kern:
            gbs = bpc * case.ncell / (kern[kname]["ms_avg"] * 1e-3) / 1e9
            kern[kname].update({"bytes_per_cell": bpc, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS})
    sd.close()
    sf.close()
    out = {"ms_per_step": wall / args.steps * 1e3, "cell_updates_per_s": case.ncell * args.steps / wall,
           "speedup_vs_fp64_step": s64 / (wall / args.steps), "kernels": kern,
           "max_rel_diff_vs_fp64_after_steps": float(np.max(np.abs(Tf - Td) / np.abs(Td))),
           "steps_compared": 2 * args.steps}
    print(f"[bench] FP32 step: {out['ms_per_step']:.3f} ms ({out['speedup_vs_fp64_step']:.2f}x FP64), "
          f"max rel diff {out['max_rel_diff_vs_fp64_after_steps']:.2e}", file=sys.stderr, flush=True)
    return out


def c1_check(args):
    """The 2-D parity plate (128 x 128, RK3, local dt, vnn 0.8, no IRS) from
    its seeded initial state to residual 1e-10: fuss_run on captured graphs
    (one launch per kernel) and as persistent launches (one cooperative
    launch per 32768 steps), stop iteration and wall time of each, and the
    field against the oracle's run to the same stop (a committed fixture)."""
    from fuss_amd import cases, gpu

    out = {"workload": "2-D plate 128^2, k=50, rho=7800, cp=460, 302 500/300 K, 301 q=0, RK3 LTS vnn 0.8, "
                       "to residual 1e-10"}
    fields = {}
    for persistent, name in ((False, "graphs"), (True, "persistent")):
        c = cases.plate2d(128)
        s = gpu.Solver(c)
        s.persistent(persistent)
        s.run_level(1, 60, 60, -1.0, 1e300, hist=False)   # warm-up: modules, graph capture
        s.set_T(c.blocks[0].T)
        s.set_clock(1, -1.0, 0)
        device_sync()
        t0 = time.perf_counter()
        n, hist = s.solve()
        device_sync()
        dt = time.perf_counter() - t0
        fields[name] = s.get("T")
        out[name] = {"iterations": int(n), "seconds": dt, "us_per_step": dt / n * 1e6,
                     "cell_updates_per_s": c.ncell * n / dt, "final_residual": float(hist[-1, 2])}
        print(f"[bench] C1 {name}: {n} iterations in {dt:.2f} s ({dt / n * 1e6:.1f} us/step)", file=sys.stderr,
              flush=True)
        s.close()
    out["persistent_vs_graphs_max_rel"] = float(np.max(np.abs(fields["persistent"] - fields["graphs"]) /
                                                      np.abs(fields["graphs"])))
    out["speedup_persistent_vs_graphs"] = out["graphs"]["seconds"] / out["persistent"]["seconds"]
    fx = os.path.join(ROOT, "tests", "golden", "plate2d_128_oracle.npz")
    if os.path.isfile(fx):
        z = np.load(fx, allow_pickle=False)
        out["oracle_stop_iteration"] = int(z["stop_iter"])
        out["max_rel_err_vs_oracle"] = float(np.max(np.abs(fields["persistent"] - z["T"]) / np.abs(z["T"])))
    return out


def summarize_profile(db_path, top=16, only=None):
    """Markdown summary of a rocprofv3 results database (kernel-trace
    statistics, or per-kernel PMC counter values when the run collected
    counters): calls, mean / min / max duration and share per kernel; for
    counters the median and max value per dispatch.  top: rows kept; only:
    kernels whose name contains this text (counters)."""
    import sqlite3

    db = sqlite3.connect(db_path)
    cur = db.cursor()
    lines = []
    npmc = cur.execute("select count(*) from pmc_events").fetchone()[0]
    if npmc:
        lines.append("| kernel | counter | dispatches | median | max |")
        lines.append("|---|---|---:|---:|---:|")
        rows = cur.execute("select name, counter_name, counter_value from pmc_events").fetchall()
        by = {}
        for name, ctr, v in rows:
            if only is None or only in name:
                by.setdefault((name, ctr), []).append(v)
        for (name, ctr), vs in sorted(by.items(), key=lambda kv: -max(kv[1]))[:top]:
            lines.append(f"| `{name[:100]}` | {ctr} | {len(vs)} | {float(np.median(vs)):.1f} | {max(vs):.1f} |")
    else:
        tot = cur.execute("select sum(duration) from kernels").fetchone()[0] or 1
        lines.append("| kernel | calls | avg us | min us | max us | % |")
        lines.append("|---|---:|---:|---:|---:|---:|")
        q = ("select name, count(*), avg(duration), min(duration), max(duration), sum(duration) from kernels "
             "group by name order by sum(duration) desc limit ?")
        for name, n, avg, mn, mx, sm in cur.execute(q, (top,)):
            lines.append(f"| `{name[:100]}` | {n} | {avg / 1e3:.1f} | {mn / 1e3:.1f} | {mx / 1e3:.1f} | "
                         f"{100.0 * sm / tot:.1f} |")
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calib-copy", action="store_true", help="also run 1 GiB D2D copies (counter calibration)")
    ap.add_argument("--rk", type=int, default=4)
    ap.add_argument("--kc", type=int, default=0, help="column height (4..64); with --irs-rows: tiles without autotune")
    ap.add_argument("--stage-rows", type=int, default=4,
                    help="rows of the LDS-ring stage kernel's columns (4 or 8), or 1: register columns")
    ap.add_argument("--irs-rows", type=int, default=4, help="rows of the fused IRS pass's columns (4 or 8)")
    ap.add_argument("--row-march", action="store_true", help="with --kc: the row-march stage kernel")
    ap.add_argument("--no-rect", action="store_true",
                    help="stream the stored 3-D metric arrays instead of the per-axis factors")
    ap.add_argument("--no-refine", action="store_true", help="skip the time-to-1e-10 runs")
    ap.add_argument("--plain-conv-n", type=int, default=64,
                    help="cube size of the measured plain FP64 and FP32-refined runs to 1e-10 (0: skip)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint (inverse problem) measurement")
    ap.add_argument("--no-batched", action="store_true", help="skip the sweep and lattice measurements")
    ap.add_argument("--no-c1", action="store_true", help="skip the 2-D plate run to 1e-10")
    ap.add_argument("--no-fp32-step", action="store_true", help="skip the FP32 explicit step on the cube")
    ap.add_argument("--only", choices=["lattice", "sweep", "c1", "adjoint"],
                    help="run this measurement alone (profiling) and print its JSON")
    ap.add_argument("--no-graphs", action="store_true", help="lattice: fuss_run step by step (no captured graphs)")
    ap.add_argument("--sweep-cases", type=int, default=256)
    ap.add_argument("--sweep-steps", type=int, default=5)
    ap.add_argument("--lattice-steps", type=int, default=1000)
    ap.add_argument("--adj-n", type=int, default=512)
    ap.add_argument("--adj-steps", type=int, default=400)
    ap.add_argument("--adj-lbfgs", type=int, default=20)
    ap.add_argument("--summarize-profile", metavar="DB", help="print a markdown summary of a rocprofv3 database")
    ap.add_argument("--profile-rows", type=int, default=16, help="with --summarize-profile: rows kept")
    ap.add_argument("--profile-only", metavar="TEXT", help="with --summarize-profile: kernels whose name contains TEXT")
    args = ap.parse_args()
    if args.summarize_profile:
        summarize_profile(args.summarize_profile, args.profile_rows, args.profile_only)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); "
              "multi-GPU sharding of the path is design-only", file=sys.stderr)
        sys.exit(2)

    from fuss_amd import cases, gpu, perfmodel

    if args.only:
        fn = {"lattice": lattice_check, "sweep": sweep_check, "c1": c1_check, "adjoint": adjoint_check}[args.only]
        print(json.dumps({args.only: fn(args)}), flush=True)
        return
    print(f"[bench] building the {args.n}^3 cube", file=sys.stderr, flush=True)
    case = cases.cube3d(n=args.n, mgl=3, n_rk=args.rk)
    t0 = time.perf_counter()
    s = gpu.Solver(case)
    print(f"[bench] setup {time.perf_counter()-t0:.1f} s", file=sys.stderr, flush=True)
    rect = s.rect_info(1)
    print(f"[bench] per-axis metric factors: weights {rect[0]}, dt {rect[1]}, "
          f"max relative deviation {rect[2]:.3e}", file=sys.stderr, flush=True)
    if args.no_rect:
        s.rect(False)
        rect = (False, False, rect[2])
    if args.kc:
        s.set_tiles(1, kc=args.kc, row_march=int(args.row_march), stage_rows=args.stage_rows,
                    irs_rows=args.irs_rows)
        tuned = (args.kc, int(args.row_march), 0.0)
        chosen = "fixed (--kc)"
    else:
        tuned = s.autotune(1, steps=4)
        chosen = "fuss_autotune"
    _, _, stage_rows, irs_rows = s.tiles(1)
    skind = "row march" if tuned[1] else ("register columns" if stage_rows == 1 else "LDS ring")
    print(f"[bench] tiles ({chosen}): column height {tuned[0]}, {skind}, "
          f"stage columns of {stage_rows} rows, IRS columns of {irs_rows} rows, {tuned[2]:.3f} ms/step",
          file=sys.stderr, flush=True)
    # warm-up: module load, then a call of the timed size, which captures the
    # graph of fuss_run's steps (chunks of min(steps, 30) rounded up to a
    # multiple of 3) that the timed call replays
    s.run_level(1, max(args.warmup, 1), max(args.warmup, 1), -1.0, 1e300, hist=False)
    s.run_level(1, args.steps, args.steps, -1.0, 1e300, hist=False)
    device_sync()
    # timed region: `steps` Explicit_Steps through fuss_run (the library's
    # step loop: captured graph replay, stop tests on the device)
    t0 = time.perf_counter()
    k, _, _, res = s.run_level(1, args.steps, args.steps, -1.0, 1e300, hist=False)
    device_sync()
    wall = time.perf_counter() - t0
    assert k == args.steps, (k, args.steps)
    # per-kernel durations: the same steps again, one at a time, with HIP
    # events around every launch on the library's stream
    s.L.fuss_profile(s.h, 1)
    for _ in range(args.steps):
        s.explicit_step(1)
    device_sync()
    ncl = gpu.C.c_int * 8
    ms = (ctypes.c_double * 8)()
    cnt = ncl()
    s.L.fuss_profile_read(s.h, 8, ms, cnt)
    s.L.fuss_profile(s.h, 0)
    ncell = case.ncell
    value = ncell * args.steps / wall
    print(f"[bench] {args.steps} steps in {wall:.3f} s, residual {np.sqrt(res.res_sq):.4e}", file=sys.stderr,
          flush=True)
    # "stage": the stage kernel without dt (stages 2-4 of a steady step on a
    # rectilinear level, every stage otherwise); "stage_dt": stage 1 with dt
    names = ["dt", "ghost", "stage", "irs", "irs_update", "reduce", "bc", "stage_dt"]
    kern = {names[q]: {"ms_avg": ms[q] / cnt[q], "launches": cnt[q]} for q in range(8) if cnt[q]}
    # algorithmic bytes per cell of the streamed kernels on this level (one
    # material, uniform qvol): the stage kernel reads T and dt and writes R
    # (the first stage writes dt instead of reading it); the fused x-y-z IRS
    # pass reads R and T^n and writes T (the four-stage primitive update needs
    # no stage state), the last stage also writing the next step's dt
    kb = {"stage": perfmodel.stage_bytes(ndir=3, tang=False, irs=True, prim=True, mat=False, qvol=False,
                                         rect=rect[0]),
          "stage_dt": perfmodel.stage_bytes(ndir=3, tang=False, irs=True, prim=True, mat=False, qvol=False,
                                            rect=rect[0]),
          "irs_update": perfmodel.irs_fused_bytes(8, mat=False, stage_state=False)}
    for kname, bpc in kb.items():
        if kname in kern:
            gbs = bpc * case.ncell / (kern[kname]["ms_avg"] * 1e-3) / 1e9
            kern[kname].update({"bytes_per_cell": bpc, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS})
    # one material and a uniform qvol: the stage kernel reads neither stream;
    # on a rectilinear level it reads no per-cell geometry either
    stage_bpc = perfmodel.stage_bytes(ndir=3, tang=False, irs=True, prim=True, mat=False, qvol=False, rect=rect[0])
    stage_b = stage_bpc * ncell
    skern = stage_kernel(rect[0], tuned[0], stage_rows)
    ldsring = args.n == 256 and not tuned[1]
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
        "config": {"workload": f"steady cube {args.n}^3, plain FP64, RK{args.rk} + LTS + IRS(jacobi2, 0.5), "
                               "qvol 1e6, 303 h=50, k(T) 32-knot table, 3 MG levels; step = one "
                               "finest-level Explicit_Step",
                   "cells": ncell, "parallelism": "single-gpu",
                   "tiles": {"column_planes": tuned[0],
                             "stage_kernel": "row march" if tuned[1] else ("register columns" if stage_rows == 1
                                                                           else "LDS ring"),
                             "stage_rows": stage_rows, "irs_rows": irs_rows, "chosen_by": chosen}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": STAGE_TRAFFIC_MEASURED.get(skern) if ldsring else None,
                     "kernel": (skern if ldsring else "stage kernel")
                               + " (fused interior + BC fluxes, scaled residual)",
                     "bytes_per_cell": stage_bpc,
                     "ms_avg": stage_ms},
        "kernels": kern,
        "metric_factors": {"per_axis": rect[0], "dt_per_axis": rect[1], "max_rel_dev": rect[2]},
        "iteration_bytes_per_cell": perfmodel.iteration_bytes(args.rk, 3, False, True, True, mat=False,
                                                             fused_irs=True, qvol=False, rect=rect[0],
                                                             fused_dt=True),
        # the bytes a straightforward implementation streams per cell-update
        # (every operand per stage), and the throughput they amount to here
        "canonical_bytes_per_cell_update": perfmodel.canonical_bytes(args.rk, True, 3),
        "canonical_GBs": perfmodel.canonical_bytes(args.rk, True, 3) * value / 1e9,
    }
    ms_step = wall / args.steps * 1e3
    if not args.no_fp32_step:
        out["fp32_step"] = fp32_step_check(args, case, tuned[0], stage_rows, irs_rows, wall / args.steps)
    if not args.no_refine:
        out["time_to_1e-10"] = time_to_converge(s, case, args, ms_step)
    s.close()
    if not args.no_adjoint:
        out["adjoint"] = adjoint_check(args)
    if not args.no_c1:
        out["c1_plate128"] = c1_check(args)
    if not args.no_batched:
        out["sweep"] = sweep_check(args)
        out["lattice"] = lattice_check(args)
    if args.calib_copy:
        calibration_copy()
    # the achievable bandwidth beside the peak: a 1 GiB device-to-device copy
    copy_gbs = stream_copy_gbs()
    out["roofline"]["achievable_copy_GBs"] = copy_gbs
    out["roofline"]["frac_of_copy"] = out["roofline"]["achieved"] / copy_gbs
    print(f"[bench] device copy {copy_gbs:.0f} GB/s; stage kernel at {out['roofline']['frac_of_copy']:.2f} of it",
          file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        print(f"[bench] cpu baseline: oracle {args.cpu_n}^3 x {args.cpu_iters}", file=sys.stderr, flush=True)
        cb = cpu_baseline(args.cpu_n, args.cpu_iters)
        if "time_to_1e-10" in out and args.cpu_n == args.n:
            # the plain march's iteration count (the same trajectory) at the CPU's
            # iteration rate: an estimate, as the GPU's plain time at this size
            it = out["time_to_1e-10"]["fp64_plain"]["iterations_est"]
            cb["time_to_1e-10_est_s"] = it / (cb["value"] / ncell)
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
