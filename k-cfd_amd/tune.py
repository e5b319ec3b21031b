"""Launch-shape autotuner for batched BiCGSTAB solves.

The knobs are those of kcfd_solve_opts:

* ``grid_per_cu``: persistent workgroups (256 lanes, 4 wavefronts) per CU
  of the SpMV and vector kernels -- the block count of the tile walk;
* ``trsv_per_cu``: workgroups per CU of the ILU(0) triangular solves;
* ``tile_walk``: 1 interleaved tiles per workgroup, 2 one contiguous run of
  tiles per workgroup (inside its XCD's range either way);
* ``rows_per_thread``: rows each thread takes per tile, 1, 2 or 4 (a wave
  covers 64, 128 or 256 rows; a tile 256, 512 or 1024);
* ``check_every``: iterations per captured hipGraph between host checks.

Stage 1 times every launch shape (grid, ILU grid, walk) on the batch itself,
with the iteration count capped (``maxiter``) so that each candidate costs a
few milliseconds; stage 2 runs full solves for each (``rows_per_thread``,
``check_every``) pair at that shape.  Rows per thread is judged on full
solves because it changes the summation tree of the dot products, and with
it the iteration count of the hardest systems (a few per cent per iteration
can be lost many times over in iterations).  The winner is cached in a JSON file keyed by a signature of the
workload (precision, preconditioner, system count and size class, device),
at ``$KCFD_TUNE_CACHE`` or ``~/.cache/kcfd_amd/autotune.json``.
"""

from __future__ import annotations

import itertools
import json
import math
import os
import time

from . import _lib as L
from . import device

GRID_PER_CU = (2, 4, 8)
TRSV_PER_CU = (2, 4)
TILE_WALK = (1, 2)
ROWS_PER_THREAD = (2, 1, 4)
CHECK_EVERY = (8, 16, 32)


def cache_path() -> str:
    return os.environ.get("KCFD_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "kcfd_amd", "autotune.json")


def signature(sb, precision: str, precond: str, order: str | None, ctx: L.Context) -> str:
    n = sb.n
    info = ctx.info()
    nnz = sum(m.nnz for m in sb.mats)
    return "|".join([
        "v3", precision, precond, str(order if precond == "ilu0" else None),
        f"nsys2^{int(round(math.log2(max(1, len(n)))))}",
        f"rows2^{int(round(math.log2(max(1, int(n.sum())))))}",
        f"nnzrow{int(round(nnz / max(1, int(n.sum()))))}",
        f"cus{info['cus']}", info["name"].strip()])


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db):
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + f".{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass   # a read-only home only costs a re-tune


def _time_batch(sb, opts, reps=2):
    """Wall ms of a batched solve under ``opts`` (best of ``reps``)."""
    b = sb.make_batch(opts)
    try:
        best = math.inf
        for _ in range(reps + 1):     # first run captures the graph
            for x in sb.xs:
                x.fill(0.0)
            sb.ctx.sync()
            t0 = time.perf_counter()
            b.run()
            best = min(best, time.perf_counter() - t0)
        return 1e3 * best
    finally:
        b.close()


def autotune(sb, precision="mixed", precond="jacobi", order="color", rtol=1e-8, delta=0.1,
             probe_iters=24, use_cache=True, log=None) -> dict:
    """Best {grid_per_cu, trsv_per_cu, tile_walk, rows_per_thread, check_every} for this batch."""
    sb.setup(precond, precision, order)
    key = signature(sb, precision, precond, order, sb.ctx)
    db = _load() if use_cache else {}
    if use_cache and key in db:
        if log:
            log(f"autotune: cached {db[key]['best']} for {key}")
        return dict(db[key]["best"])
    trsv = TRSV_PER_CU if precond == "ilu0" else (0,)
    trials = []
    for g, t, w in itertools.product(GRID_PER_CU, trsv, TILE_WALK):
        o = device.default_opts(rtol=rtol, precond=precond, precision=precision, inner_rtol=delta,
                                maxiter=probe_iters, check_every=8, grid_per_cu=g, trsv_per_cu=t,
                                tile_walk=w)
        ms = _time_batch(sb, o)
        trials.append(dict(grid_per_cu=g, trsv_per_cu=t, tile_walk=w, probe_ms=round(ms, 3)))
        if log:
            log(f"autotune: grid_per_cu={g} trsv_per_cu={t} tile_walk={w}: {ms:.2f} ms "
                f"for <= {probe_iters} iterations")
    shape = min(trials, key=lambda r: r["probe_ms"])
    # rows per thread changes the reduction tree, and with it the iteration
    # counts of the hardest systems: judged on full solves, with check_every
    best = None
    full = []
    for rpt, ce in itertools.product(ROWS_PER_THREAD, CHECK_EVERY):
        o = device.default_opts(rtol=rtol, precond=precond, precision=precision, inner_rtol=delta,
                                check_every=ce, grid_per_cu=shape["grid_per_cu"],
                                trsv_per_cu=shape["trsv_per_cu"], tile_walk=shape["tile_walk"],
                                rows_per_thread=rpt)
        ms = _time_batch(sb, o, reps=1)
        full.append(dict(rows_per_thread=rpt, check_every=ce, full_ms=round(ms, 3)))
        if log:
            log(f"autotune: rows_per_thread={rpt} check_every={ce}: {ms:.2f} ms (full solve)")
        if best is None or ms < best[2]:
            best = (rpt, ce, ms)
    choice = dict(grid_per_cu=shape["grid_per_cu"], trsv_per_cu=shape["trsv_per_cu"],
                  tile_walk=shape["tile_walk"], rows_per_thread=best[0], check_every=best[1])
    if use_cache:
        db = _load()
        db[key] = dict(best=choice, trials=trials, full=full, full_solve_ms=round(best[2], 3))
        _store(db)
    return choice


def describe(choice: dict) -> str:
    return ", ".join(f"{k}={v}" for k, v in sorted(choice.items()))


__all__ = ["autotune", "signature", "cache_path", "describe", "GRID_PER_CU", "TRSV_PER_CU",
           "TILE_WALK", "ROWS_PER_THREAD", "CHECK_EVERY"]
