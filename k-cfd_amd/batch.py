"""Batched BiCGSTAB over many uneven, device-resident FVM systems.

The reference solves one system per call (core/linear_solver.py:14-62).
Here a batch of systems shares every launch of the solver: the persistent
tile walk of each phase kernel covers the rows of all systems, per-system
finalize kernels advance each system's scalar state machine, and one
captured hipGraph replays ``check_every`` iterations between host checks
(kcfd_batch_*).  Systems converge independently; a finished system's tiles
are skipped by the later iterations.

Workflow::

    specs = workloads.batch_specs(512)              # (kind, cells, seed)
    cases = build_cases(specs)                      # host meshes + recipes
    sb = SystemBatch.assemble(ctx, cases)           # device meshes, FVM assembly
    res = sb.solve(precision="mixed", precond="ilu0", order="color")

``build_cases`` uses worker processes: forked before the process first
touches the GPU, spawned after.
"""

from __future__ import annotations

import functools
import math
import multiprocessing as mp
import os
import runpy
import time
from concurrent.futures import ProcessPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from . import device, workloads
from ._lib import DevBuf


def _build_one(spec):
    kind, n, seed = spec
    return workloads.case_of_size(kind, n, seed)


def build_cases(specs, workers: int | None = None):
    """Host-side meshes and assembly recipes of ``specs`` (kind, cells, seed),
    in order.  ``workers`` > 1 runs that many builder processes: forked while
    this process has not touched the GPU, started fresh (``spawn``, which
    costs each worker an interpreter start) once it has, since forking a
    process that holds a HIP context is unsafe."""
    if workers is None:
        workers = min(16, os.cpu_count() or 1)
    if workers <= 1 or len(specs) < 4:
        return [_build_one(s) for s in specs]
    if L._default_ctx is None:
        pool = ProcessPoolExecutor(workers, mp_context=mp.get_context("fork"))
    else:
        boot = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_boot.py")
        pool = ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn"),
                                   initializer=functools.partial(runpy.run_path, boot))
    with pool as ex:
        return list(ex.map(_build_one, specs, chunksize=2))


@dataclass
class BatchResult:
    precision: str
    precond: str
    wall_s: float            # host wall time of the batch solve (set-up excluded)
    device_ms: float         # device time of the solve (events)
    results: list            # kcfd_solve_result per system
    bytes: float             # algorithmic HBM bytes (sum over systems)
    opts: dict = field(default_factory=dict)

    @property
    def nsys(self):
        return len(self.results)

    @property
    def solves_per_s(self):
        return self.nsys / self.wall_s

    @property
    def gbs(self):
        return self.bytes / (self.device_ms * 1e-3) / 1e9

    def summary(self):
        it = np.array([r.iters for r in self.results])
        rr = np.array([r.rel_resid for r in self.results])
        return dict(precision=self.precision, precond=self.precond, nsys=self.nsys,
                    solves_per_s=round(self.solves_per_s, 2), wall_ms=round(1e3 * self.wall_s, 2),
                    device_ms=round(self.device_ms, 2), achieved_gbs=round(self.gbs, 1),
                    converged=int(sum(r.info == 0 for r in self.results)),
                    iters_mean=round(float(it.mean()), 1), iters_max=int(it.max()),
                    refinements_total=int(sum(r.refinements for r in self.results)),
                    fp64_fallbacks=int(sum(getattr(r, "fallback", 0) for r in self.results)),
                    max_rel_resid=float(rr.max()), opts=self.opts,
                    not_converged=[dict(index=k, info=r.info, iters=r.iters,
                                        refinements=r.refinements, rel_resid=r.rel_resid)
                                   for k, r in enumerate(self.results) if r.info != 0][:16])


class SystemBatch:
    """Device-resident systems (matrix, rhs, solution) solved as one batch."""

    def __init__(self, ctx: L.Context, mats, bs, kinds=None):
        self.ctx = ctx
        self.mats = list(mats)
        self.bs = list(bs)
        self.xs = [DevBuf(ctx, m.n) for m in self.mats]
        self.kinds = list(kinds) if kinds is not None else ["?"] * len(self.mats)
        self.n = np.array([m.n for m in self.mats])
        self._pc = None

    @classmethod
    def assemble(cls, ctx: L.Context, cases, keep_meshes: bool = False):
        """Upload each case's mesh and play its recipe on the device; host
        arrays are dropped as soon as a system is assembled."""
        mats, bs, kinds = [], [], []
        for c in cases:
            dm = device.DeviceMesh(c.mesh, ctx)
            A = workloads.assemble_device(dm, c)
            mats.append(A)
            bs.append(DevBuf.from_array(ctx, A.get_rhs()))
            kinds.append(c.name.split("_")[0].rstrip("0123456789"))
            if not keep_meshes:
                dm.close()
        return cls(ctx, mats, bs, kinds)

    @property
    def total_rows(self):
        return int(self.n.sum())

    def setup(self, precond="jacobi", precision="fp64", order="color"):
        key = (precond, precision, order)
        if self._pc != key:
            for m in self.mats:
                m.setup_precond(precond, precision, order=order if precond == "ilu0" else None)
            self.ctx.sync()
            self._pc = key

    def make_batch(self, opts: L.SolveOpts) -> device.Batch:
        return device.Batch(self.ctx, self.mats, self.bs, self.xs, opts)

    def solve(self, precision="mixed", precond="jacobi", order="color", rtol=1e-8,
              check_every=8, grid_per_cu=0, trsv_per_cu=0, delta=0.1, batch=None,
              fallback=True, tile_walk=0, rows_per_thread=0, fold_colour0=0) -> BatchResult:
        """One batched solve from x0 = 0 (preconditioners set up beforehand,
        outside the timed region).  In mixed precision, the systems that stop
        without converging continue in FP64 from their best iterate, as one
        more batch inside the timed region (``fallback``)."""
        self.setup(precond, precision, order)
        opts = device.default_opts(rtol=rtol, precond=precond, precision=precision, inner_rtol=delta,
                                   check_every=check_every, grid_per_cu=grid_per_cu,
                                   trsv_per_cu=trsv_per_cu, tile_walk=tile_walk,
                                   rows_per_thread=rows_per_thread, fold_colour0=fold_colour0)
        own = batch is None
        b = batch if batch is not None else self.make_batch(opts)
        nfb = 0
        try:
            for x in self.xs:
                x.fill(0.0)
            self.ctx.sync()
            t0 = time.perf_counter()
            res = b.run()
            dev_ms = max(r.ms for r in res) if res else 0.0
            for r in res:
                r.fallback = 0
            bad = [k for k, r in enumerate(res) if r.info != 0]
            if fallback and precision == "mixed" and bad:
                o64 = device.default_opts(rtol=rtol, precond=precond, precision="fp64",
                                          check_every=check_every, grid_per_cu=grid_per_cu,
                                          trsv_per_cu=trsv_per_cu, tile_walk=tile_walk,
                                          rows_per_thread=rows_per_thread)
                fb = device.Batch(self.ctx, [self.mats[k] for k in bad], [self.bs[k] for k in bad],
                                  [self.xs[k] for k in bad], o64)
                try:
                    res2 = fb.run()
                finally:
                    fb.close()
                dev_ms += max(r.ms for r in res2)
                for k, r2 in zip(bad, res2):
                    r1 = res[k]
                    r2.iters += r1.iters
                    r2.refinements = r1.refinements
                    r2.bytes += r1.bytes
                    r2.fallback = 1
                    res[k] = r2
                nfb = len(bad)
            wall = time.perf_counter() - t0
        finally:
            if own:
                b.close()
        return BatchResult(precision, precond, wall, dev_ms, res, float(sum(r.bytes for r in res)),
                           dict(check_every=check_every, grid_per_cu=grid_per_cu,
                                trsv_per_cu=trsv_per_cu, tile_walk=tile_walk,
                                rows_per_thread=rows_per_thread,
                                order=order if precond == "ilu0" else None,
                                delta=delta if precision == "mixed" else None, fp64_fallbacks=nfb))

    def solutions(self):
        return [x.download() for x in self.xs]

    def close(self):
        for x in self.xs:
            x.free()
        for b in self.bs:
            b.free()
        for m in self.mats:
            m.close()
        self.xs, self.bs, self.mats = [], [], []


def specs_summary(specs):
    n = np.array([s[1] for s in specs])
    kinds = {k: sum(1 for s in specs if s[0] == k) for k in ("quad", "hex", "tet")}
    return dict(nsys=len(specs), cells_total=int(n.sum()), cells_min=int(n.min()),
                cells_max=int(n.max()), kinds=kinds,
                cells_geomean=int(round(math.exp(np.log(n).mean()))))
