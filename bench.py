#!/usr/bin/env python3
"""Benchmark: solves/s and %HBM roofline of the FVM assembly + BiCGSTAB hot
path on one MI355X.

Workload (default): 3D unit-cube hex mesh 128^3 = 2,097,152 cells, SIMPLE
pressure-correction matrix (7-point, nnz 14,581,760, a_P ~ lognormal(0,1),
Dirichlet p'=0 on the outlet, b ~ U[-1,1], seed 1002), BiCGSTAB + ILU(0) to
relative residual 1e-8 (FP64-verified).  One step = device assembly of the
matrix + preconditioner set-up + solve, with mesh, fields, b and x resident
in HBM.  `value` is the mixed-precision (FP32 SpMV/preconditioner inside FP64
refinement) rate; the plain FP64 rate, per-solve p50/p99 latency and the
dominant kernel's roofline are reported beside it, and the scipy CPU
baseline (oracle ILU(0) + scipy bicgstab) is timed on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128]
                       [--precond ilu0|jacobi] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per k_spmv_p launch on the 128^3 mixed default workload, from the
# FETCH_SIZE / WRITE_SIZE counter passes: the narrow-index FP32 SELL copy
# (6 bytes an entry, the library's default) and the packed one (8 bytes an
# entry, KCFD_SELL16=0)
TRAFFIC_SPMV_P_MIXED_HEX128 = 115.2e6          # 106.7 MB read + 8.5 MB written
TRAFFIC_SPMV_P_MIXED_HEX128_PACKED = 153.9e6   # 145.4 MB read + 8.5 MB written


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(A, b, precond, iters_needed, budget_s=15.0):
    """scipy BiCGSTAB (reference call) + the oracle preconditioner on the host,
    timed on a bounded number of iterations and scaled to a full solve."""
    from oracle import solve as osolve
    from scipy.sparse.linalg import LinearOperator, bicgstab

    os.environ.setdefault("OMP_NUM_THREADS", "1")
    n = A.shape[0]
    t0 = time.perf_counter()
    if precond == "ilu0":
        f = osolve.ILU0(A)
        M = LinearOperator((n, n), matvec=f.solve)
    else:
        M = osolve.make_preconditioner(A, "jacobi")
    t_setup = time.perf_counter() - t0
    # time k iterations (maxiter=k) within the budget
    k = 2
    t_iter = None
    while True:
        t1 = time.perf_counter()
        bicgstab(A, b, x0=np.zeros(n), rtol=1e-8, atol=0.0, maxiter=k, M=M)
        dt = time.perf_counter() - t1
        t_iter = dt / k
        if dt * 3 > budget_s or k >= 64:
            break
        k *= 2
    t_solve = t_setup + t_iter * iters_needed
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": (f"oracle {precond} set-up ({t_setup:.2f} s) + {k} scipy bicgstab iterations "
                       f"({t_iter * 1e3:.1f} ms/iter) on the same {n}-row system, scaled to the "
                       f"{iters_needed} iterations an FP64 solve with this preconditioner takes; 1 thread"),
            "s_per_solve": t_solve}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="pressure", choices=["pressure", "batched", "serve"],
                    help="pressure: one hex n^3 pressure-correction system (default); "
                         "batched: --nsys uneven quad/hex/tet systems in one batch; "
                         "serve: Poisson request stream, continuous batching")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: arrivals per second")
    ap.add_argument("--duration", type=float, default=60.0, help="serve: seconds of arrivals")
    ap.add_argument("--templates", type=int, default=8, help="serve: mesh sizes per kind")
    ap.add_argument("--baseline-duration", type=float, default=10.0,
                    help="serve: seconds of the same trace served one at a time in FP64")
    ap.add_argument("--nsys", type=int, default=512, help="batched: number of systems")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--precond", default="ilu0", choices=["ilu0", "jacobi"])
    ap.add_argument("--method", default="bicgstab", choices=["bicgstab", "cg", "gmres"],
                    help="pressure: Krylov method (cg: solve_pressure_correction's solver; gmres: "
                         "solve_linear_system(method='gmres'), restart 20; both FP64 only)")
    ap.add_argument("--no-tune", action="store_true", help="batched: library default launch shape")
    ap.add_argument("--delta", type=float, default=0.1,
                    help="mixed: reliable-update trigger (FP32 residual reduction)")
    ap.add_argument("--order", default="color", choices=["natural", "color"],
                    help="ILU(0) row ordering (colour: first-fit colouring, 2 levels per solve)")
    ap.add_argument("--grid-per-cu", type=int, default=0, help="launch shape (0: library default)")
    ap.add_argument("--trsv-per-cu", type=int, default=0, help="launch shape (0: library default)")
    ap.add_argument("--rows-per-thread", type=int, default=0, choices=[0, 1, 2, 4],
                    help="launch shape: rows per thread of the tile kernels (0: library default 2)")
    ap.add_argument("--tile-walk", type=int, default=0,
                    help="launch shape: 1 interleaved, 2 contiguous tiles per workgroup (0: default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-calibrate", action="store_true",
                    help="serve: skip the per-template precision calibration")
    ap.add_argument("--no-single", action="store_true",
                    help="batched: skip the FP64 one-system-at-a-time GPU baseline")
    ap.add_argument("--check-every", type=int, default=8,
                    help="iterations per captured graph between host status checks")
    ap.add_argument("--precisions", default="fp64,mixed",
                    help="comma list; a profiler run can take one precision per process")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "only one GPU is measured: --gpus must be 1 (independent systems "
                                   "would shard one batch per GPU; one large system would split into "
                                   "row slabs with halo exchanges and block-Jacobi ILU(0); neither is built)"}))
        sys.exit(2)
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "serve":
        return bench_serve(args)
    return bench_pressure(args)


def bench_pressure(args):
    import __graft_entry__
    __graft_entry__.load_package()
    from kcfd_amd import _lib as L, device, workloads
    from kcfd_amd._lib import DevBuf

    ctx = L.default_context()
    info = ctx.info()
    log("device", info)
    copy_gbs = ctx.stream_copy_gbs(1 << 30, 10)
    log(f"STREAM copy (1 GiB): {copy_gbs:.1f} GB/s")
    t = time.perf_counter()
    case = workloads.hex_pressure(args.n, seed=1002)
    mesh = case.mesh
    log(f"mesh {mesh.n_cells} cells built in {time.perf_counter() - t:.1f}s")
    dm = device.DeviceMesh(mesh, ctx)
    rec = dict(case.recipe)["pressure"]
    pd = np.array([1 if nm in rec["dirichlet"] else 0 for nm in dm.patch_names], np.int32)
    dP = DevBuf.from_array(ctx, rec["dP"])
    A = dm.new_system()
    bdev = DevBuf.from_array(ctx, case.rhs)
    xdev = DevBuf(ctx, mesh.n_cells)
    n, nnz = A.n, A.nnz

    def assemble():
        L.check(L.lib().kcfd_asm_reset(A.h))
        L.check(L.lib().kcfd_asm_pressure(dm.h, A.h, dP.p, float(rec["rho"]), pd.ctypes.data))

    results, sols = {}, {}
    precs = [p for p in args.precisions.split(",") if p]
    if args.method in ("cg", "gmres"):
        precs = ["fp64"]                 # the CG and GMRES paths run in FP64
    for prec in precs:
        opts = device.default_opts(rtol=1e-8, precond=args.precond, precision=prec,
                                   inner_rtol=args.delta, check_every=args.check_every,
                                   grid_per_cu=args.grid_per_cu, trsv_per_cu=args.trsv_per_cu,
                                   tile_walk=args.tile_walk, method=args.method,
                                   rows_per_thread=args.rows_per_thread)
        lat, rs = [], []

        def step():
            assemble()
            A.setup_precond(args.precond, prec, order=args.order)
            xdev.fill(0.0)
            return A.solve_dev(bdev.p, xdev.p, opts, setup=False)

        for w in range(args.warmup):
            r = step()
            log(f"{prec} warmup {w}: iters={r.iters} refinements={r.refinements} ms={r.ms:.1f}")
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ts = time.perf_counter()
            r = step()
            lat.append(time.perf_counter() - ts)
            rs.append(r)
            log(f"{prec} step: iters={r.iters} info={r.info} resid={r.rel_resid:.2e} "
                f"{lat[-1] * 1e3:.1f} ms")
        ctx.sync()
        wall = time.perf_counter() - t0
        r = rs[-1]
        sols[prec] = xdev.download()
        results[prec] = dict(solves_per_s=args.steps / wall, ms_per_step=1e3 * wall / args.steps,
                             solve_ms=float(np.mean([q.ms for q in rs])), iters=r.iters,
                             refinements=r.refinements, info=r.info, rel_resid=r.rel_resid,
                             bytes=r.bytes, p50_ms=1e3 * float(np.percentile(lat, 50)),
                             p99_ms=1e3 * float(np.percentile(lat, 99)))
        log(prec, results[prec])
        if args.method == "gmres":       # host-driven per cycle: no batch workspace to profile
            results[prec]["kernels"] = {}
            continue
        # per-kernel timing of the solve's kernels on the library stream
        bh = L.ctypes.c_void_p()
        L.check(L.lib().kcfd_mat_solver(A.h, L.ctypes.byref(bh)))
        kern = {}
        ilu_name = "ilu_apply" if args.order == "color" else "ilu_lower_solve"
        for phase, name in [(1, "spmv_p"), (2, "xr_update")] + ([(3, ilu_name)] if args.precond == "ilu0" else []):
            ms = L.ctypes.c_double()
            by = L.ctypes.c_double()
            L.check(L.lib().kcfd_batch_profile(bh, phase, 10, L.ctypes.byref(ms)))
            L.check(L.lib().kcfd_batch_kernel_bytes(bh, phase, L.ctypes.byref(by)))
            kern[name] = dict(avg_ms=ms.value, bytes=by.value,
                              gbs=by.value / (ms.value * 1e-3) / 1e9)
        results[prec]["kernels"] = kern
        log(prec, "kernels", kern)

    mixed = results.get("mixed") or results[precs[0]]
    fp64 = results.get("fp64") or results[precs[0]]
    # roofline kernel: the SpMV k_spmv_p (v = A M p with the rt.v partials);
    # with k_spmv_s, the same operation on s, it is the largest part of an
    # iteration.  Timed per launch with HIP events on the library stream
    # inside real iterations; its algorithmic bytes are M_A + 3 n s_v (ILU)
    # or M_A + 4 n s_v (Jacobi).
    kern = mixed["kernels"]
    solve_gbs = mixed["bytes"] / (mixed["solve_ms"] * 1e-3) / 1e9
    if args.method == "gmres":
        # no single dominant kernel (SpMV, level-scheduled ILU(0) and the
        # Gram-Schmidt sweeps): the whole solve's algorithmic bytes over its
        # device time
        roofline = {"bound": "hbm", "achieved": round(solve_gbs, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(solve_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "whole GMRES solve", "stream_copy_gbs": round(copy_gbs, 1)}
    else:
        dk = kern["spmv_p"]
        achieved = dk["gbs"]
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_spmv_p",
                    "avg_us": round(1e3 * dk["avg_ms"], 2), "bytes_per_launch": dk["bytes"],
                    "solve_avg_gbs": round(solve_gbs, 1), "stream_copy_gbs": round(copy_gbs, 1)}
        # HBM bytes per k_spmv_p launch measured by the FETCH_SIZE / WRITE_SIZE
        # counter passes on this exact workload (default shape, mixed,
        # two-colour ILU(0), 128^3), scaled by the copy-kernel calibration
        if (args.method == "bicgstab" and args.n == 128 and args.precond == "ilu0" and args.order == "color"
                and "mixed" in results and not (args.grid_per_cu or args.trsv_per_cu or args.tile_walk
                                                 or args.rows_per_thread not in (0, 2))):
            packed = os.environ.get("KCFD_SELL16", "1").startswith("0")
            traffic = TRAFFIC_SPMV_P_MIXED_HEX128_PACKED if packed else TRAFFIC_SPMV_P_MIXED_HEX128
            roofline["traffic"] = traffic
            # the rate of the bytes the kernel actually moves (the algorithmic
            # model counts 4-byte column indices; the narrow-index copy
            # streams 2-byte offsets)
            roofline["traffic_gbs"] = round(traffic / (dk["avg_ms"] * 1e-3) / 1e9, 1)
            roofline["traffic_frac"] = round(roofline["traffic_gbs"] / HBM_PEAK_GBS, 4)
            roofline["traffic_source"] = ("profiles/hbm_traffic_bicgstab_ilu0_mixed_hex128.md: FETCH_SIZE x 2.000 "
                                          "+ WRITE_SIZE x 1.000 KiB per dispatch, median working dispatch"
                                          + (" (packed SELL copy, KCFD_SELL16=0)" if packed else ""))
    cpu = None
    if not args.no_cpu:
        try:
            Ah = A.to_scipy()
            bh_ = case.rhs
            its = fp64["iters"]
            if args.method != "bicgstab":
                raise RuntimeError(f"no CPU baseline for the {args.method} workload")
            if args.precond == "ilu0" and args.order != "natural":
                # the CPU path factors in the natural order, which needs fewer
                # iterations than the colour ordering: count them on the GPU
                A.setup_precond("ilu0", "fp64", order="natural")
                xdev.fill(0.0)
                its = A.solve_dev(bdev.p, xdev.p, device.default_opts(rtol=1e-8, precond="ilu0"),
                                  setup=False).iters
                log(f"natural-order ILU(0) FP64 iterations (CPU baseline scaling): {its}")
            cpu = cpu_baseline(Ah, bh_, args.precond, its, args.cpu_budget)
            cpu["cores"] = 1
            cpu.pop("s_per_solve", None)
        except Exception as e:  # the baseline is informative; never fail the bench on it
            log("cpu baseline failed:", repr(e))
    line = {
        "metric": "solves/s + %HBM roofline at rel-resid 1e-8; p50/p99 latency, mixed vs FP64",
        "value": round(mixed["solves_per_s"], 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 refinement" if "mixed" in results else "f64",
        "data": "synthetic",
        "config": {"workload": f"hex{args.n}^3 SIMPLE pressure correction, "
                               f"{dict(bicgstab='BiCGSTAB', cg='CG', gmres='GMRES(20)')[args.method]}"
                               f"+{args.precond}, "
                               f"rtol 1e-8", "cells": n, "nnz": nnz,
                   "ilu_order": args.order if args.precond == "ilu0" else None,
                   "step": "device assembly + preconditioner set-up + solve"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mixed": {k: v for k, v in mixed.items() if k != "kernels"},
        "fp64": {k: v for k, v in fp64.items() if k != "kernels"},
        "mixed_vs_fp64": round(mixed["solves_per_s"] / fp64["solves_per_s"], 3) if len(precs) > 1 else None,
        # relative L2 of the mixed solution against the FP64 solution of the
        # same system (the FP64 path is the one held to the scipy oracle)
        "mixed_error_vs_fp64": (float(np.linalg.norm(sols["mixed"] - sols["fp64"]) / np.linalg.norm(sols["fp64"]))
                                if "mixed" in sols and "fp64" in sols else None),
        "value_precision": "mixed" if "mixed" in results else precs[0],
        "kernels_mixed": kern,
        "kernels_fp64": fp64["kernels"],
        "device": info["name"],
    }
    print(json.dumps(line))



def bench_batched(args):
    """Batched workload: --nsys systems, cells log-uniform in [2e3, 2.5e5],
    50% quad convection-diffusion / 30% hex pressure correction / 20% tet
    two-fluid momentum (seeds 4000+i), each assembled on the device, solved
    together to rtol 1e-8 in FP64 and in mixed precision with the autotuned
    launch shape.  One step = one batched solve of every system from x0 = 0."""
    import __graft_entry__
    __graft_entry__.load_package()
    from kcfd_amd import _lib as L, batch as B, device, tune, workloads

    specs = workloads.batch_specs(args.nsys)
    t = time.perf_counter()
    cases = B.build_cases(specs)          # forks builders: before the GPU is touched
    log(f"built {len(cases)} host cases in {time.perf_counter() - t:.1f}s: {B.specs_summary(specs)}")
    ctx = L.default_context()
    info = ctx.info()
    log("device", info)
    copy_gbs = ctx.stream_copy_gbs(1 << 30, 10)
    t = time.perf_counter()
    sb = B.SystemBatch.assemble(ctx, cases)
    del cases
    log(f"assembled {len(sb.mats)} systems ({sb.total_rows} rows) on the device in "
        f"{time.perf_counter() - t:.1f}s")
    order = args.order if args.precond == "ilu0" else None
    results, sols = {}, {}
    for prec in [p for p in args.precisions.split(",") if p]:
        t = time.perf_counter()
        sb.setup(args.precond, prec, order or "natural")
        log(f"{prec}: preconditioners set up in {time.perf_counter() - t:.1f}s")
        if args.no_tune:
            choice = dict(grid_per_cu=args.grid_per_cu, trsv_per_cu=args.trsv_per_cu,
                          tile_walk=args.tile_walk, rows_per_thread=args.rows_per_thread,
                          check_every=args.check_every)
        else:
            choice = tune.autotune(sb, prec, args.precond, order or "natural", delta=args.delta, log=log)
        log(f"{prec}: launch shape {tune.describe(choice)}")
        opts = dict(precision=prec, precond=args.precond, order=order or "natural", delta=args.delta,
                    **choice)
        bobj = sb.make_batch(device.default_opts(
            rtol=1e-8, precond=args.precond, precision=prec, inner_rtol=args.delta,
            check_every=choice["check_every"], grid_per_cu=choice["grid_per_cu"],
            trsv_per_cu=choice["trsv_per_cu"], tile_walk=choice.get("tile_walk", 0),
            rows_per_thread=choice.get("rows_per_thread", 0)))
        for w in range(args.warmup):
            r = sb.solve(batch=bobj, **opts)
            log(f"{prec} warmup {w}: {r.summary()}")
        runs = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            runs.append(sb.solve(batch=bobj, **opts))
            log(f"{prec} step: wall {runs[-1].wall_s * 1e3:.1f} ms")
        wall = time.perf_counter() - t0
        r = runs[-1]
        summ = r.summary()
        summ["solves_per_s"] = round(args.steps * len(sb.mats) / wall, 2)
        summ["ms_per_step"] = round(1e3 * wall / args.steps, 3)
        lat = [q.wall_s for q in runs]
        summ["p50_ms"] = round(1e3 * float(np.percentile(lat, 50)), 3)
        summ["p99_ms"] = round(1e3 * float(np.percentile(lat, 99)), 3)
        summ["achieved_gbs"] = round(float(np.mean([q.bytes / (q.device_ms * 1e-3) / 1e9 for q in runs])), 1)
        # the dominant kernel (SpMV) timed inside real iterations of this batch
        log(f"{prec}: timing the SpMV inside batch iterations")
        ms = L.ctypes.c_double()
        by = L.ctypes.c_double()
        L.check(L.lib().kcfd_batch_profile(bobj.h, 1, 10, L.ctypes.byref(ms)))
        L.check(L.lib().kcfd_batch_kernel_bytes(bobj.h, 1, L.ctypes.byref(by)))
        summ["spmv"] = dict(avg_ms=ms.value, bytes=by.value, gbs=by.value / (ms.value * 1e-3) / 1e9)
        if args.precond == "ilu0":
            L.check(L.lib().kcfd_batch_profile(bobj.h, 3, 10, L.ctypes.byref(ms)))
            L.check(L.lib().kcfd_batch_kernel_bytes(bobj.h, 3, L.ctypes.byref(by)))
            summ["ilu_apply" if order == "color" else "ilu_lower_solve"] = dict(
                avg_ms=ms.value, bytes=by.value, gbs=by.value / (ms.value * 1e-3) / 1e9)
        # the profile overwrote x: solve once more for the solutions
        log(f"{prec}: kernel timing done; solving again for the solutions")
        sb.solve(batch=bobj, **opts)
        sols[prec] = sb.solutions()
        bobj.close()
        if prec == "fp64" and not args.no_single:
            summ["one_at_a_time"] = single_fp64(sb, args, choice, log)
        results[prec] = summ
        log(prec, summ)
    mixed = results.get("mixed") or next(iter(results.values()))
    fp64 = results.get("fp64") or next(iter(results.values()))
    err = None
    if "mixed" in sols and "fp64" in sols:
        e = [float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
             for a, b in zip(sols["mixed"], sols["fp64"])]
        err = dict(max=max(e), median=float(np.median(e)))
    roof = {"bound": "hbm", "achieved": round(mixed["spmv"]["gbs"], 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(mixed["spmv"]["gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "spmv_p (batch)", "solve_avg_gbs": mixed["achieved_gbs"],
            "stream_copy_gbs": round(copy_gbs, 1)}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_batch(sb, specs, args.precond, args.cpu_budget)
    single = fp64.get("one_at_a_time")
    line = {
        "metric": "solves/s + %HBM roofline at rel-resid 1e-8; p50/p99 latency, mixed vs FP64",
        "value": mixed["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": mixed["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 inner / f64 refinement",
        "data": "synthetic",
        "config": {"workload": f"{len(sb.mats)} uneven systems in one batch (50% quad conv-diff, "
                               f"30% hex pressure, 20% tet two-fluid; cells log-uniform 2e3..2.5e5), "
                               f"BiCGSTAB+{args.precond}, rtol 1e-8", "rows": sb.total_rows,
                   "ilu_order": order, "step": "one batched solve of every system"},
        "roofline": roof, "cpu_baseline": cpu, "mixed": mixed, "fp64": fp64,
        "mixed_vs_fp64": round(mixed["solves_per_s"] / fp64["solves_per_s"], 3) if len(results) > 1 else None,
        "mixed_vs_fp64_one_at_a_time": round(mixed["solves_per_s"] / single["solves_per_s"], 3)
        if single else None,
        "mixed_error_vs_fp64": err, "device": info["name"],
    }
    print(json.dumps(line))


def single_fp64(sb, args, choice, log):
    """The plain FP64 GPU path: every system of the batch solved on its own
    (kcfd_solve, one system per launch), in order, from x0 = 0.  A first pass
    builds each matrix's single-system workspace and graph; the second pass
    is timed."""
    from kcfd_amd import device
    opts = device.default_opts(rtol=1e-8, precond=args.precond, precision="fp64",
                               check_every=choice["check_every"], grid_per_cu=choice["grid_per_cu"],
                               trsv_per_cu=choice["trsv_per_cu"], tile_walk=choice.get("tile_walk", 0))
    walls, iters = [], []
    for rep in range(2):
        walls.clear()
        iters.clear()
        sb.ctx.sync()
        t0 = time.perf_counter()
        for m, b, x in zip(sb.mats, sb.bs, sb.xs):
            ts = time.perf_counter()
            x.fill(0.0)
            r = m.solve_dev(b.p, x.p, opts, setup=False)
            walls.append(time.perf_counter() - ts)
            iters.append(r.iters)
        wall = time.perf_counter() - t0
        log(f"fp64 one at a time, pass {rep}: {wall:.2f} s")
    return dict(solves_per_s=round(len(sb.mats) / wall, 2), wall_ms=round(1e3 * wall, 1),
                p50_ms=round(1e3 * float(np.percentile(walls, 50)), 3),
                p99_ms=round(1e3 * float(np.percentile(walls, 99)), 3),
                iters_mean=round(float(np.mean(iters)), 1))


def cpu_baseline_batch(sb, specs, precond, budget_s):
    """The reference CPU path (scipy bicgstab + the oracle preconditioner, 1
    thread) on systems of the batch taken in order until the budget is
    spent; the batch rate is extrapolated by cell count."""
    from oracle import solve as osolve
    t_all, cells, done = 0.0, 0, 0
    for k, m in enumerate(sb.mats):
        if t_all > budget_s:
            break
        A = m.to_scipy()
        b = sb.bs[k].download()
        t0 = time.perf_counter()
        osolve.solve_bicgstab(A, b, rtol=1e-8, precond="ilu0_struct" if precond == "ilu0" else "jacobi")
        t_all += time.perf_counter() - t0
        cells += m.n
        done += 1
    rate = cells / t_all                       # cells per second
    est = sb.total_rows / rate                 # seconds for the whole batch
    return {"value": len(sb.mats) / est, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": (f"{done} of {len(sb.mats)} systems ({cells} cells) solved by scipy bicgstab + "
                       f"oracle {precond} in {t_all:.1f} s; batch rate extrapolated by cell count; 1 thread")}



def bench_serve(args):
    """Serving workload: Poisson arrivals at --rate per second for --duration
    seconds; each request = pressure correction + 3 momentum solves on one
    mesh of 1e3..1e6 cells (log-uniform; quad / hex / tet 50/30/20, snapped
    to --templates sizes per kind); mixed precision to rtol 1e-8 with
    continuous batching and HBM admission.  The first --baseline-duration
    seconds of the same trace are also served one request at a time in FP64."""
    import __graft_entry__
    __graft_entry__.load_package()
    from kcfd_amd import _lib as L, serve as S

    sizes = S.template_sizes(per_kind=args.templates)
    t = time.perf_counter()
    built = S.build_templates(sizes=sizes)          # forks builders: before the GPU is touched
    log(f"built {len(built)} mesh templates in {time.perf_counter() - t:.1f}s (sizes {sizes})")
    trace = S.make_trace(rate=args.rate, duration=args.duration)
    kidx = {k: i for i, k in enumerate(S.KINDS)}
    tr = [(ta, kidx[k] * len(sizes) + S.nearest_template(sizes, n), sd) for ta, k, n, sd in trace]
    ctx = L.default_context()
    info = ctx.info()
    log("device", info)
    t = time.perf_counter()
    temps = [S.Template(ctx, k, n, p, m, args.precond, "mixed", args.order) for k, n, (p, m) in built]
    del built
    log(f"templates on the device in {time.perf_counter() - t:.1f}s")
    if not args.no_calibrate:
        # per-template precision: one mixed solve of each prototype matrix
        t = time.perf_counter()
        for tp in temps:
            tp.calibrate()
        log(f"precision calibration in {time.perf_counter() - t:.1f}s: FP64 for "
            f"{sum(len(tp.fp64_kinds) for tp in temps)} of {2 * len(temps)} (template, kind) pairs: "
            + ", ".join(f"{tp.kind}{tp.n}:{'p' if True in tp.fp64_kinds else ''}{'m' if False in tp.fp64_kinds else ''}"
                        for tp in temps if tp.fp64_kinds))
    log(f"templates ready; "
        f"request footprint {min(x.bytes for x in temps) / 1e6:.1f}..{max(x.bytes for x in temps) / 1e6:.1f} MB")
    cells = np.array([temps[ti].n for _, ti, _ in tr])
    log(f"trace: {len(tr)} requests over {args.duration:.0f}s, {cells.sum() / 1e6:.1f}M cells "
        f"(mean {cells.mean():.0f}), {4 * len(tr)} solves")

    def progress(*a):
        log(*a)

    # warm-up: a few requests (graph capture, first launches)
    srv = S.Server(ctx, temps, precision="mixed", precond=args.precond, order=args.order,
                   batched=True, check_every=args.check_every, log=progress)
    srv.run([(0.0, ti, sd) for _, ti, sd in tr[:8]])
    t = time.perf_counter()
    rep = srv.run(tr)
    srv.close()
    mixed = rep.summary()
    log("batched mixed", mixed, f"(wall {time.perf_counter() - t:.1f}s)")
    nb = sum(1 for x in tr if x[0] < args.baseline_duration)
    base = None
    if nb:
        b = S.Server(ctx, temps, precision="fp64", precond=args.precond, order=args.order, batched=False,
                     log=progress)
        rb = b.run(tr[:nb])
        base = rb.summary()
        base["trace_prefix_s"] = args.baseline_duration
        log("one-at-a-time fp64", base)
    line = {
        "metric": "solves/s + %HBM roofline at rel-resid 1e-8; p50/p99 latency, mixed vs FP64",
        "value": mixed["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(1e3 * rep.elapsed_s, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 inner / f64 refinement", "data": "synthetic",
        "config": {"workload": f"serving: Poisson {args.rate:.0f} req/s for {args.duration:.0f} s, "
                               f"each request pressure correction + 3 momentum solves on a 1e3..1e6-cell "
                               f"mesh (quad/hex/tet 50/30/20), BiCGSTAB+{args.precond}, rtol 1e-8",
                   "requests": len(tr), "templates_per_kind": len(sizes),
                   "step": "the whole trace (continuous batching)"},
        "latency_ms": {"p50": mixed["p50_ms"], "p99": mixed["p99_ms"]},
        "serve_mixed_batched": mixed, "baseline_fp64_one_at_a_time": base,
        "vs_one_at_a_time": round(mixed["solves_per_s"] / base["solves_per_s"], 3) if base else None,
        "device": info["name"],
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
