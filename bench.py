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
baseline (oracle ILU(0) + scipy bicgstab) is timed to completion.

Beside it the same line carries the 512-system batch (config 4), the quad
128^2 Jacobi system (config 1) and the tet two-fluid Jacobi system
(config 3), each FP64 and mixed, with CPU baselines timed to completion on
the host before the GPU is touched.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128]
                       [--precond ilu0|jacobi] [--no-cpu] [--no-extra]
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
# HBM bytes per launch on the 128^3 mixed default workload (folded products,
# narrow-index FP32 copies), from the FETCH_SIZE / WRITE_SIZE counter passes
# scaled by the copy-kernel calibration: one ILU(0) apply (the two k_ilu2
# launches of the p apply) and one k_spmv_p.  None: not measured for the
# current kernels.
TRAFFIC_ILU_APPLY_MIXED_HEX128 = 124.4e6   # 55.9 MB colour-1 launch + 68.5 MB colour-0 launch
TRAFFIC_SPMV_P_MIXED_HEX128 = 70.3e6       # 65.9 MB read + 4.3 MB written


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


# The reference's own CPU path for the 128^3 pressure system -- scipy
# bicgstab with spilu(A.tocsc(), fill_factor=1, drop_tol=0) as
# core/preconditioner.py:49-54 and core/linear_solver.py:59 call them -- run
# once to completion on the MI355X host (1 thread, AMD EPYC): three minutes,
# too long for every bench run.  Measured by `bench.py --workload
# cpu-reference`; profiles/cpu_reference_hex128.md has the run.
CPU_REFERENCE_HEX128 = dict(
    value=0.005611, unit="solves/s", cores=1, kind="reference",
    sample=("the reference's call (spilu(A.tocsc(), fill_factor=1, drop_tol=0) as a LinearOperator + "
            "scipy bicgstab, rtol 1e-8, x0 = 0) on the 2097152-row hex 128^3 pressure system, once, to "
            "completion: 52.5 s spilu + 125.7 s for 788 iterations = 178.2 s; 1 thread, AMD EPYC host of "
            "the MI355X; measured by bench.py --workload cpu-reference"),
    iters=788, s_per_solve=178.23, spilu_s=52.5, rel_resid=5.63e-9)


def cpu_info():
    """(threads this process may use, CPU model)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return len(os.sched_getaffinity(0)), model


def _one_thread():
    """BLAS pinned to one thread for the timed CPU solves (threadpoolctl
    limits the libraries numpy and scipy already loaded)."""
    from threadpoolctl import threadpool_limits
    return threadpool_limits(limits=1)


def cpu_solve(A, b, precond, reps=1, progress=False):
    """scipy bicgstab with the reference's call, rtol 1e-8, x0 = 0, timed to
    completion: set-up + solve, one warm-up when reps > 1, the median of
    reps runs.  precond: 'jacobi' (the reference's LinearOperator), 'spilu'
    (the reference's spilu call) or 'ilu0_struct' (the oracle's structural
    ILU(0), the preconditioner the device runs).  ``progress`` logs the
    set-up time and then the iteration count about every 30 s."""
    from oracle import solve as osolve
    times, its = [], []
    last = [0.0]

    def tick(k):
        now = time.perf_counter()
        if now - last[0] >= 30.0:
            log(f"  iteration {k}, {now - t0:.0f} s")
            last[0] = now

    with _one_thread():
        for r in range(reps + (1 if reps > 1 else 0)):
            t0 = time.perf_counter()
            M = osolve.make_preconditioner(A, "ilu0" if precond == "spilu" else precond)
            if progress:
                last[0] = time.perf_counter()
                log(f"  preconditioner set up in {last[0] - t0:.1f} s")
            x, info, it = osolve.solve_bicgstab(A, b, rtol=1e-8, M=M, on_iter=tick if progress else None)
            dt = time.perf_counter() - t0
            if reps == 1 or r > 0:
                times.append(dt)
                its.append(it)
    return dict(s_per_solve=float(np.median(times)), iters=its[-1], info=int(info),
                rel_resid=osolve.true_rel_residual(A, x, b), runs=len(times))


_POOL_CASES = None


def _pool_solve(k):
    from oracle import fvm as ofvm
    case = _POOL_CASES[k]
    A, b, _ = ofvm.system_of_case(case)
    r = cpu_solve(A, b, "spilu")
    return case.mesh.n_cells, r["s_per_solve"], r["iters"], r["info"]


def cpu_batch_pool(cases, every, workers):
    """The reference CPU path (spilu + bicgstab, one thread per process) on a
    fixed sample of the batch -- every `every`-th system -- solved to
    completion by `workers` processes (forked: before the GPU is touched).
    Returns the N-core throughput and the 1-core rate of the same sample."""
    import multiprocessing as mp
    global _POOL_CASES
    _POOL_CASES = cases
    idx = list(range(0, len(cases), every))
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(workers) as pool:
        out = pool.map(_pool_solve, idx, chunksize=1)
    wall = time.perf_counter() - t0
    _POOL_CASES = None
    cpu_s = sum(o[1] for o in out)
    return dict(systems=len(idx), cells=int(sum(o[0] for o in out)), wall_s=round(wall, 2),
                cpu_s=round(cpu_s, 2), workers=workers, solves_per_s_ncore=round(len(idx) / wall, 3),
                solves_per_s_1core=round(len(idx) / cpu_s, 4),
                not_converged=sum(1 for o in out if o[3] != 0))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="pressure", choices=["pressure", "batched", "serve", "cpu-reference"],
                    help="pressure: the default line -- the hex n^3 pressure-correction system (value), "
                         "with the 512-system batch, the quad 128^2 and the tet two-fluid configurations "
                         "beside it; batched: the batch alone; serve: Poisson request stream, continuous "
                         "batching; cpu-reference: the reference's own CPU path on the n^3 system, to "
                         "completion (minutes)")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: arrivals per second")
    ap.add_argument("--duration", type=float, default=60.0, help="serve: seconds of arrivals")
    ap.add_argument("--templates", type=int, default=8, help="serve: mesh sizes per kind")
    ap.add_argument("--baseline-duration", type=float, default=10.0,
                    help="serve: seconds of the same trace served one at a time in FP64")
    ap.add_argument("--nsys", type=int, default=512, help="batched: number of systems")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch-steps", type=int, default=2, help="default line: timed steps of the batch")
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--precond", default="ilu0", choices=["ilu0", "jacobi"])
    ap.add_argument("--method", default="bicgstab", choices=["bicgstab", "cg", "gmres"],
                    help="pressure: Krylov method (cg: solve_pressure_correction's solver; gmres: "
                         "solve_linear_system(method='gmres'), restart 20; both FP64 only)")
    ap.add_argument("--no-tune", action="store_true", help="batched: library default launch shape")
    ap.add_argument("--delta", type=float, default=0.1,
                    help="mixed: reliable-update trigger (FP32 residual reduction)")
    ap.add_argument("--order", default="color", choices=["natural", "color", "segment"],
                    help="ILU(0) row ordering (colour: first-fit colouring, 2 levels per solve)")
    ap.add_argument("--grid-per-cu", type=int, default=0, help="launch shape (0: library default)")
    ap.add_argument("--trsv-per-cu", type=int, default=0, help="launch shape (0: library default)")
    ap.add_argument("--rows-per-thread", type=int, default=0, choices=[0, 1, 2, 4],
                    help="launch shape: rows per thread of the tile kernels (0: library default 2)")
    ap.add_argument("--tile-walk", type=int, default=0,
                    help="launch shape: 1 interleaved, 2 contiguous tiles per workgroup (0: default)")
    ap.add_argument("--pressure-method", default="cg", choices=["cg", "bicgstab"],
                    help="serve: pressure corrections by FP64 CG (the reference's pressure solver) or by "
                         "mixed BiCGSTAB with FP64 CG for those that stall")
    ap.add_argument("--fold", type=int, default=0, choices=[-1, 0, 1],
                    help="two-colour ILU(0) BiCGSTAB: fold the colour-0 rows of each product into the "
                         "apply (0: library default, on in mixed precision; 1: on; -1: off)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="default line: only the pressure system (no batch, no configs 1 and 3)")
    ap.add_argument("--no-calibrate", action="store_true",
                    help="serve: skip the per-template precision calibration")
    ap.add_argument("--no-single", action="store_true",
                    help="batched: skip the FP64 one-system-at-a-time GPU baseline")
    ap.add_argument("--check-every", type=int, default=None,
                    help="iterations per captured graph between host status checks (default: 32 for the "
                         "single pressure system, where it measured fastest; 8 for the configs 1 / 3 "
                         "solves and the server)")
    ap.add_argument("--precisions", default="fp64,mixed",
                    help="comma list; a profiler run can take one precision per process")
    ap.add_argument("--cpu-workers", type=int, default=16, help="processes of the batch's CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "only one GPU is measured: --gpus must be 1 (independent systems "
                                   "would shard one batch per GPU; one large system would split into "
                                   "row slabs with halo exchanges and block-Jacobi ILU(0); neither is built)"}))
        sys.exit(2)
    if args.workload == "batched":
        return bench_batched_only(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "cpu-reference":
        return bench_cpu_reference(args)
    return bench_default(args)


def _kc():
    import __graft_entry__
    __graft_entry__.load_package()
    from kcfd_amd import _lib as L, batch as B, device, tune, workloads
    return L, B, device, tune, workloads


def bench_cpu_reference(args):
    """The reference's CPU path on the n^3 pressure system, to completion."""
    L, B, device, tune, workloads = _kc()
    from oracle import fvm as ofvm
    cores, model = cpu_info()
    case = workloads.hex_pressure(args.n, seed=1002)
    A, b, _ = ofvm.system_of_case(case)
    log(f"reference CPU path on {A.shape[0]} rows: spilu + bicgstab, 1 thread ({model})")
    r = cpu_solve(A, b, "spilu", progress=True)
    r.update(cells=A.shape[0], nnz=int(A.nnz), cores_available=cores, cpu_model=model,
             solves_per_s=1.0 / r["s_per_solve"])
    print(json.dumps(r))


def bench_default(args):
    """The line the driver records: config 2 (hex n^3 pressure correction,
    BiCGSTAB + ILU(0)) as `value`, with the 512-system batch (config 4), the
    quad 128^2 Jacobi system (config 1) and the tet two-fluid Jacobi system
    (config 3) measured beside it.  Host meshes and every CPU baseline come
    first, before the GPU is touched (the batch's CPU pool forks)."""
    L, B, device, tune, workloads = _kc()
    from oracle import fvm as ofvm
    extra = not args.no_extra and args.method == "bicgstab" and args.n == 128
    t = time.perf_counter()
    case = workloads.hex_pressure(args.n, seed=1002)
    log(f"mesh {case.mesh.n_cells} cells built in {time.perf_counter() - t:.1f}s")
    cases4 = c1 = c3 = None
    if extra:
        t = time.perf_counter()
        specs = workloads.batch_specs(args.nsys)
        cases4 = B.build_cases(specs)
        c1 = workloads.quad_convdiff(128, seed=1001)
        c3 = workloads.tet_twofluid(200_000, seed=1003)
        log(f"batch, quad 128^2 and tet meshes built in {time.perf_counter() - t:.1f}s: "
            f"{B.specs_summary(specs)}; tets {c3.mesh.n_cells}")
    cores, model = cpu_info()
    cpu = {}
    if not args.no_cpu:
        t = time.perf_counter()
        A2, b2, _ = ofvm.system_of_case(case)
        if args.method == "bicgstab" and args.precond == "ilu0":
            # the oracle's structural ILU(0) in the natural order (the device's
            # preconditioner; the colour ordering is a GPU choice) with scipy
            # bicgstab, to completion; one run (about 20 s at 128^3)
            r = cpu_solve(A2, b2, "ilu0_struct")
            cpu["config2"] = dict(value=round(1.0 / r["s_per_solve"], 5), unit="solves/s", cores=1, kind="port",
                                  sample=(f"the whole solve, once: oracle structural ILU(0), natural order, "
                                          f"set-up + scipy bicgstab (the reference's call, rtol 1e-8) on the "
                                          f"same {A2.shape[0]}-row system, {r['iters']} iterations, "
                                          f"{r['s_per_solve']:.1f} s; BLAS pinned to 1 thread"),
                                  iters=r["iters"], s_per_solve=round(r["s_per_solve"], 3))
            if CPU_REFERENCE_HEX128 and args.n == 128:
                cpu["config2_reference_path"] = CPU_REFERENCE_HEX128
        if extra:
            A1, b1, _ = ofvm.system_of_case(c1)
            r = cpu_solve(A1, b1, "jacobi", reps=5)
            cpu["config1"] = dict(value=round(1.0 / r["s_per_solve"], 3), unit="solves/s", cores=1,
                                  kind="reference path", iters=r["iters"],
                                  sample=(f"the reference's call exactly (Jacobi LinearOperator + scipy "
                                          f"bicgstab, rtol 1e-8) on the quad 128^2 system, median of 5 after "
                                          f"a warm-up, {r['iters']} iterations; 1 thread"))
            A3, b3, _ = ofvm.system_of_case(c3)
            r = cpu_solve(A3, b3, "jacobi")
            cpu["config3"] = dict(value=round(1.0 / r["s_per_solve"], 4), unit="solves/s", cores=1,
                                  kind="reference path", iters=r["iters"],
                                  sample=(f"the reference's call exactly (Jacobi + scipy bicgstab, rtol 1e-8) "
                                          f"on the {A3.shape[0]}-tet two-fluid momentum system, once, "
                                          f"{r['s_per_solve']:.1f} s; 1 thread"))
            workers = max(1, min(args.cpu_workers, cores))
            pool = cpu_batch_pool(cases4, 16, workers)
            cpu["config4"] = dict(value=pool["solves_per_s_ncore"], unit="solves/s", cores=workers,
                                  kind="reference path", value_1core=pool["solves_per_s_1core"],
                                  sample=(f"every 16th system of the batch ({pool['systems']} systems, "
                                          f"{pool['cells']} cells) solved to completion by the reference's "
                                          f"call (spilu(fill_factor=1, drop_tol=0) + scipy bicgstab, rtol "
                                          f"1e-8) in {workers} processes of 1 thread: {pool['wall_s']} s wall, "
                                          f"{pool['cpu_s']} s of solver time"), pool=pool)
        log(f"CPU baselines in {time.perf_counter() - t:.1f}s: {cpu}")
    ctx = L.default_context()
    info = ctx.info()
    log("device", info)
    copy_gbs = ctx.stream_copy_gbs(1 << 30, 10)
    read_gbs = ctx.stream_read_gbs(1 << 30, 10)
    log(f"copy ceiling {copy_gbs:.1f} GB/s, read ceiling {read_gbs:.1f} GB/s (1 GiB)")
    p = run_pressure(args, ctx, case)
    mixed, fp64 = p["mixed"], p["fp64"]
    roof = p["roofline"]
    roof["stream_copy_gbs"] = round(copy_gbs, 1)
    roof["stream_read_gbs"] = round(read_gbs, 1)
    cfg = {}
    if extra:
        cfg["config1_quad128_jacobi"] = run_single(args, ctx, c1, "jacobi", steps=20)
        cfg["config3_tet_twofluid_jacobi"] = run_single(args, ctx, c3, "jacobi", steps=5)
        cfg["config4_batch512"] = run_batched(args, ctx, cases4, steps=args.batch_steps)
        del cases4
        for k, key in (("config1", "config1_quad128_jacobi"), ("config3", "config3_tet_twofluid_jacobi"),
                       ("config4", "config4_batch512")):
            if k in cpu:
                cfg[key]["cpu_baseline"] = cpu[k]
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
        "dtype": "f32 inner / f64 refinement" if p["precs"] != ["fp64"] else "f64",
        "data": "synthetic",
        "config": {"workload": f"hex{args.n}^3 SIMPLE pressure correction, "
                               f"{dict(bicgstab='BiCGSTAB', cg='CG', gmres='GMRES(20)')[args.method]}"
                               f"+{args.precond}, rtol 1e-8", "cells": p["n"], "nnz": p["nnz"],
                   "ilu_order": args.order if args.precond == "ilu0" else None,
                   "step": "device assembly + preconditioner set-up + solve"},
        "roofline": roof,
        "cpu_baseline": cpu.get("config2"),
        "cpu_reference_path": cpu.get("config2_reference_path"),
        "host": {"cores": cores, "cpu_model": model},
        "mixed": {k: v for k, v in mixed.items() if k != "kernels"},
        "fp64": {k: v for k, v in fp64.items() if k != "kernels"},
        "mixed_vs_fp64": p["mixed_vs_fp64"],
        "mixed_error_vs_fp64": p["mixed_error_vs_fp64"],
        "value_precision": "mixed" if "mixed" in p["precs"] else p["precs"][0],
        "kernels_mixed": mixed.get("kernels"),
        "kernels_fp64": fp64.get("kernels"),
        "configs": cfg,
        "device": info["name"],
    }
    print(json.dumps(line))


def run_pressure(args, ctx, case):
    """Config 2: one step = device assembly of the pressure matrix +
    preconditioner set-up + solve, FP64 and mixed; per-phase kernel timing
    and the roofline of the SpMV k_spmv_p."""
    L, B, device, tune, workloads = _kc()
    from kcfd_amd._lib import DevBuf
    mesh = case.mesh
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
                                   inner_rtol=args.delta, check_every=args.check_every or 32,
                                   grid_per_cu=args.grid_per_cu, trsv_per_cu=args.trsv_per_cu,
                                   tile_walk=args.tile_walk, method=args.method,
                                   rows_per_thread=args.rows_per_thread, fold_colour0=args.fold)
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
        two_colour = args.precond == "ilu0" and args.order == "color"
        seg = args.precond == "ilu0" and args.order == "segment"
        phases = [(1, "spmv_p"), (2, "xr_update"), (4, "p_update")]
        if args.precond == "ilu0":
            if two_colour and args.method == "bicgstab":
                phases += [(3, "ilu_colour1"), (5, "ilu_colour0"), (6, "ilu_apply")]
            elif seg and args.method == "bicgstab":     # the three segment launches
                phases += [(6, "ilu_apply")]
            else:
                phases += [(3, "ilu_apply" if two_colour or seg else "ilu_lower_solve")]
        for phase, name in phases:
            ms = L.ctypes.c_double()
            by = L.ctypes.c_double()
            L.check(L.lib().kcfd_batch_profile(bh, phase, 10, L.ctypes.byref(ms)))
            L.check(L.lib().kcfd_batch_kernel_bytes(bh, phase, L.ctypes.byref(by)))
            kern[name] = dict(avg_ms=ms.value, bytes=by.value,
                              gbs=by.value / (ms.value * 1e-3) / 1e9)
        # ilu_apply (phase 6): one apply, timed from the start of its first
        # launch to the end of its second; ilu_colour1 / ilu_colour0: each
        # launch timed alone (each then includes its own dispatch)
        results[prec]["kernels"] = kern
        log(prec, "kernels", kern)

    mixed = results.get("mixed") or results[precs[0]]
    fp64 = results.get("fp64") or results[precs[0]]
    # roofline kernel: the dominant kernel of an iteration.  With the
    # two-colour ILU(0) that is k_ilu2: two launches per apply (colour-1
    # rows, then colour-0 rows), two applies per iteration, 40% of the mixed
    # iteration's kernel time.  The unit is one apply (both launches), timed
    # with HIP events in the launches' own dispatch packets inside real
    # iterations; its algorithmic bytes are the factor entries at 4-byte
    # indices plus the vector streams (kcfd_batch_kernel_bytes, phase 3).
    # Without ILU(0) (and beside it) the SpMV k_spmv_p.
    kern = mixed["kernels"]
    solve_gbs = mixed["bytes"] / (mixed["solve_ms"] * 1e-3) / 1e9
    if args.method == "gmres":
        # no single dominant kernel (SpMV, level-scheduled ILU(0) and the
        # Gram-Schmidt sweeps): the whole solve's algorithmic bytes over its
        # device time
        roofline = {"bound": "hbm", "achieved": round(solve_gbs, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(solve_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "whole GMRES solve"}
    else:
        two_colour = args.precond == "ilu0" and args.order == "color"
        seg = args.precond == "ilu0" and args.order == "segment" and args.method == "bicgstab"
        name = "ilu_apply" if two_colour or seg else "spmv_p"
        dk = kern[name]
        achieved = dk["gbs"]
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "k_ilu2" if two_colour else ("k_seg" if seg else "k_spmv_p"),
                    "work_unit": ("one ILU(0) apply: the colour-1 and colour-0 launches of k_ilu2, timed from "
                                  "the first's start to the second's end" if two_colour else
                                  ("one ILU(0) apply: the three k_seg launches, timed from the first's start "
                                   "to the last's end" if seg else "one launch")),
                    "avg_us": round(1e3 * dk["avg_ms"], 2), "bytes_per_unit": dk["bytes"],
                    "solve_avg_gbs": round(solve_gbs, 1)}
        if two_colour or seg:
            sp = kern["spmv_p"]
            roofline["k_spmv_p"] = {"avg_us": round(1e3 * sp["avg_ms"], 2), "bytes": sp["bytes"],
                                    "gbs": round(sp["gbs"], 1), "frac": round(sp["gbs"] / HBM_PEAK_GBS, 4)}
        # HBM bytes per unit measured by the FETCH_SIZE / WRITE_SIZE counter
        # passes on this exact workload (default shape, mixed, folded,
        # two-colour ILU(0), 128^3), scaled by the copy-kernel calibration
        if (args.method == "bicgstab" and args.n == 128 and two_colour and "mixed" in results
                and args.fold == 0 and not os.environ.get("KCFD_SELL16", "1").startswith("0")
                and not (args.grid_per_cu or args.trsv_per_cu or args.tile_walk
                         or args.rows_per_thread not in (0, 2))
                and TRAFFIC_ILU_APPLY_MIXED_HEX128 is not None):
            traffic = TRAFFIC_ILU_APPLY_MIXED_HEX128
            roofline["traffic"] = traffic
            # the rate of the bytes the kernel actually moves (the algorithmic
            # model counts 4-byte column indices; the narrow-index copies
            # stream 2-byte offsets)
            roofline["traffic_gbs"] = round(traffic / (dk["avg_ms"] * 1e-3) / 1e9, 1)
            roofline["traffic_frac"] = round(roofline["traffic_gbs"] / HBM_PEAK_GBS, 4)
            roofline["traffic_source"] = ("profiles/hbm_traffic_bicgstab_ilu0_mixed_hex128.md: FETCH_SIZE x 2.000 "
                                          "+ WRITE_SIZE x 1.000 KiB per dispatch, median working dispatch")
            if TRAFFIC_SPMV_P_MIXED_HEX128 is not None:
                roofline["k_spmv_p"]["traffic"] = TRAFFIC_SPMV_P_MIXED_HEX128
    err = (float(np.linalg.norm(sols["mixed"] - sols["fp64"]) / np.linalg.norm(sols["fp64"]))
           if "mixed" in sols and "fp64" in sols else None)
    return dict(mixed=mixed, fp64=fp64, precs=precs, n=n, nnz=nnz, roofline=roofline,
                mixed_vs_fp64=round(mixed["solves_per_s"] / fp64["solves_per_s"], 3) if len(precs) > 1 else None,
                mixed_error_vs_fp64=err)


def run_single(args, ctx, case, precond, steps):
    """One system solved on its own (device-assembled matrix, x0 = 0, rtol
    1e-8), FP64 and mixed: solves/s, per-solve latency, iterations and the
    SpMV's rate inside real iterations."""
    L, B, device, tune, workloads = _kc()
    from kcfd_amd._lib import DevBuf
    dm = device.DeviceMesh(case.mesh, ctx)
    A = workloads.assemble_device(dm, case)
    b = DevBuf.from_array(ctx, A.get_rhs())
    x = DevBuf(ctx, A.n)
    out = {"workload": case.name, "cells": A.n, "nnz": A.nnz, "precond": precond}
    sols = {}
    for prec in ("fp64", "mixed"):
        A.setup_precond(precond, prec)
        opts = device.default_opts(rtol=1e-8, precond=precond, precision=prec, check_every=args.check_every or 8)
        for _ in range(2):
            x.fill(0.0)
            A.solve_dev(b.p, x.p, opts, setup=False)
        lat, rs = [], []
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            ts = time.perf_counter()
            x.fill(0.0)
            r = A.solve_dev(b.p, x.p, opts, setup=False)
            lat.append(time.perf_counter() - ts)
            rs.append(r)
        wall = time.perf_counter() - t0
        sols[prec] = x.download()
        bh = L.ctypes.c_void_p()
        L.check(L.lib().kcfd_mat_solver(A.h, L.ctypes.byref(bh)))
        ms, by = L.ctypes.c_double(), L.ctypes.c_double()
        L.check(L.lib().kcfd_batch_profile(bh, 1, 10, L.ctypes.byref(ms)))
        L.check(L.lib().kcfd_batch_kernel_bytes(bh, 1, L.ctypes.byref(by)))
        r = rs[-1]
        out[prec] = dict(solves_per_s=round(steps / wall, 3), p50_ms=round(1e3 * float(np.percentile(lat, 50)), 3),
                         p99_ms=round(1e3 * float(np.percentile(lat, 99)), 3), iters=r.iters,
                         refinements=r.refinements, info=r.info, rel_resid=r.rel_resid, fallback=getattr(r, "fallback", 0),
                         spmv=dict(avg_us=round(1e3 * ms.value, 2), bytes=by.value,
                                   gbs=round(by.value / (ms.value * 1e-3) / 1e9, 1),
                                   frac=round(by.value / (ms.value * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)))
        log(case.name, prec, out[prec])
    out["mixed_vs_fp64"] = round(out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"], 3)
    out["mixed_error_vs_fp64"] = float(np.linalg.norm(sols["mixed"] - sols["fp64"]) / np.linalg.norm(sols["fp64"]))
    A.close()
    dm.close()
    return out


def run_batched(args, ctx, cases, steps):
    """Config 4: the systems of `cases` assembled on the device and solved as
    one batch to rtol 1e-8, FP64 and mixed, each with its autotuned launch
    shape; one step = one batched solve of every system from x0 = 0.  The
    SpMV and the ILU(0) apply are timed inside batch iterations; the plain
    FP64 path solving the systems one at a time is timed too."""
    L, B, device, tune, workloads = _kc()
    t = time.perf_counter()
    sb = B.SystemBatch.assemble(ctx, cases)
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
                          check_every=args.check_every or 8)
        else:
            choice = tune.autotune(sb, prec, args.precond, order or "natural", delta=args.delta, log=log)
        log(f"{prec}: launch shape {tune.describe(choice)}")
        opts = dict(precision=prec, precond=args.precond, order=order or "natural", delta=args.delta,
                    fold_colour0=args.fold, **choice)
        bobj = sb.make_batch(device.default_opts(
            rtol=1e-8, precond=args.precond, precision=prec, inner_rtol=args.delta,
            check_every=choice["check_every"], grid_per_cu=choice["grid_per_cu"],
            trsv_per_cu=choice["trsv_per_cu"], tile_walk=choice.get("tile_walk", 0),
            rows_per_thread=choice.get("rows_per_thread", 0), fold_colour0=args.fold))
        for w in range(args.warmup):
            r = sb.solve(batch=bobj, **opts)
            log(f"{prec} warmup {w}: {r.summary()}")
        runs = []
        t0 = time.perf_counter()
        for _ in range(steps):
            runs.append(sb.solve(batch=bobj, **opts))
            log(f"{prec} step: wall {runs[-1].wall_s * 1e3:.1f} ms")
        wall = time.perf_counter() - t0
        r = runs[-1]
        summ = r.summary()
        summ["solves_per_s"] = round(steps * len(sb.mats) / wall, 2)
        summ["ms_per_step"] = round(1e3 * wall / steps, 3)
        lat = [q.wall_s for q in runs]
        summ["p50_ms"] = round(1e3 * float(np.percentile(lat, 50)), 3)
        summ["p99_ms"] = round(1e3 * float(np.percentile(lat, 99)), 3)
        summ["achieved_gbs"] = round(float(np.mean([q.bytes / (q.device_ms * 1e-3) / 1e9 for q in runs])), 1)
        # the dominant kernels timed inside real iterations of this batch
        ms = L.ctypes.c_double()
        by = L.ctypes.c_double()
        L.check(L.lib().kcfd_batch_profile(bobj.h, 1, 10, L.ctypes.byref(ms)))
        L.check(L.lib().kcfd_batch_kernel_bytes(bobj.h, 1, L.ctypes.byref(by)))
        summ["spmv"] = dict(avg_ms=ms.value, bytes=by.value, gbs=by.value / (ms.value * 1e-3) / 1e9)
        if args.precond == "ilu0":
            # the apply to p: phase 3 (the two-colour members' colour-1 rows
            # and the other members' L solve) plus, with two-colour members,
            # phase 5 (their colour-0 rows)
            # (segment order: phase 6, the whole apply of every member)
            tot_ms, tot_by = 0.0, 0.0
            for phase in ((3, 5) if order == "color" else ((6,) if order == "segment" else (3,))):
                if phase == 5 and L.lib().kcfd_batch_kernel_bytes(bobj.h, 5, L.ctypes.byref(by)) == 0 \
                        and by.value == 0.0:
                    continue   # no two-colour member
                L.check(L.lib().kcfd_batch_profile(bobj.h, phase, 10, L.ctypes.byref(ms)))
                L.check(L.lib().kcfd_batch_kernel_bytes(bobj.h, phase, L.ctypes.byref(by)))
                tot_ms += ms.value
                tot_by += by.value
            summ["ilu_apply" if order in ("color", "segment") else "ilu_lower_solve"] = dict(
                avg_ms=tot_ms, bytes=tot_by, gbs=tot_by / (tot_ms * 1e-3) / 1e9)
        # the profile overwrote x: solve once more for the solutions
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
    single = fp64.get("one_at_a_time")
    out = {"systems": len(sb.mats), "rows": sb.total_rows, "ilu_order": order, "mixed": mixed, "fp64": fp64,
           "roofline": {"bound": "hbm", "achieved": round(mixed["spmv"]["gbs"], 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(mixed["spmv"]["gbs"] / HBM_PEAK_GBS, 4),
                        "kernel": "k_spmv_p (batch)", "solve_avg_gbs": mixed["achieved_gbs"]},
           "mixed_vs_fp64": round(mixed["solves_per_s"] / fp64["solves_per_s"], 3) if len(results) > 1 else None,
           "mixed_vs_fp64_one_at_a_time": round(mixed["solves_per_s"] / single["solves_per_s"], 3)
           if single else None,
           "mixed_error_vs_fp64": err}
    sb.close()
    return out


def bench_batched_only(args):
    """The batch alone (config 4), as its own JSON line."""
    L, B, device, tune, workloads = _kc()
    specs = workloads.batch_specs(args.nsys)
    t = time.perf_counter()
    cases = B.build_cases(specs)          # forks builders: before the GPU is touched
    log(f"built {len(cases)} host cases in {time.perf_counter() - t:.1f}s: {B.specs_summary(specs)}")
    cpu = None
    if not args.no_cpu:
        cores, model = cpu_info()
        workers = max(1, min(args.cpu_workers, cores))
        pool = cpu_batch_pool(cases, 16, workers)
        cpu = dict(value=pool["solves_per_s_ncore"], unit="solves/s", cores=workers, kind="reference path",
                   value_1core=pool["solves_per_s_1core"], cpu_model=model, pool=pool,
                   sample=f"every 16th system of the batch, spilu + scipy bicgstab, {workers} processes")
    ctx = L.default_context()
    info = ctx.info()
    copy_gbs = ctx.stream_copy_gbs(1 << 30, 10)
    b = run_batched(args, ctx, cases, steps=args.steps)
    roof = dict(b["roofline"], stream_copy_gbs=round(copy_gbs, 1), traffic=None)
    line = {
        "metric": "solves/s + %HBM roofline at rel-resid 1e-8; p50/p99 latency, mixed vs FP64",
        "value": b["mixed"]["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": b["mixed"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 inner / f64 refinement",
        "data": "synthetic",
        "config": {"workload": f"{b['systems']} uneven systems in one batch (50% quad conv-diff, "
                               f"30% hex pressure, 20% tet two-fluid; cells log-uniform 2e3..2.5e5), "
                               f"BiCGSTAB+{args.precond}, rtol 1e-8", "rows": b["rows"],
                   "ilu_order": b["ilu_order"], "step": "one batched solve of every system"},
        "roofline": roof, "cpu_baseline": cpu, "batch": b, "device": info["name"],
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
            tp.calibrate(pressure=args.pressure_method == "bicgstab")
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
                   batched=True, check_every=args.check_every or 8, log=progress, fold_colour0=args.fold,
                   pressure_method=args.pressure_method)
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
                               f"mesh (quad/hex/tet 50/30/20); pressure: FP64 CG+{args.precond}, momentum: "
                               f"mixed BiCGSTAB+{args.precond} (FP64 on the segment ordering for the templates whose "
                               f"mixed calibration solve stalls); rtol 1e-8, every answer FP64-verified",
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
