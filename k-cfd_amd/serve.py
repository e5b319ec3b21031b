"""Serving loop: continuous batching of FVM solves on one GPU.

A request is one SIMPLE pressure-correction solve plus three momentum
component solves (one momentum matrix, three right-hand sides) on one mesh.
Requests arrive over time.  Every request that fits is admitted between
solver rounds: its matrices are assembled on the device from the mesh's
uploaded recipe, its preconditioners are set up, and its four systems attach
to free slots of one dynamic batch (kcfd_batch_create_dynamic).  Each round
replays a captured graph of ``check_every`` iterations over all members, so
the work of requests that arrived at different times overlaps on the device.
Systems detach as they finish.  A request completes when its four solves do.

The pressure-correction system is symmetric positive definite (a
symmetric M-matrix, whose ILU(0) is a symmetric positive-definite
preconditioner), and the reference solves it with CG
(solve_pressure_correction, core/linear_solver.py:73-93); by default
(``pressure_method='cg'``) the server does the same, in FP64, in a CG batch
of its own.  BiCGSTAB with the two-colour ILU(0) stagnates on the largest
one-sided-Dirichlet pressure systems, in FP32 and even in FP64, where CG
converges.  The momentum solves run in mixed precision.  A mixed solve that
stalls (FP32 operator error of about eps_32 * cond(A) per cycle) continues
in FP64 from its best iterate, and from then on the same kind of system on
the same mesh template goes to FP64 directly instead of repeating the mixed
attempt.  With ``pressure_method='bicgstab'`` the pressure corrections run
in mixed BiCGSTAB too, and those that stall continue by FP64 CG (the
reference's pressure solver); a stalled momentum solve continues by FP64
BiCGSTAB.  ``Template.calibrate`` makes that choice before serving:
one mixed solve of each prototype matrix (seeded right-hand side) marks the
kinds that stall.  Every result is FP64-verified: an FP64 solve that
meets scipy's stopping test (on the recursive residual) but not the
tolerance in its FP64 true residual continues from its x (at most 8
times).

Admission is by HBM footprint: a request is admitted only while the bytes
of the admitted requests (matrices, preconditioners, vectors, right-hand
sides, solutions) stay within ``hbm_budget`` and four slots are free; the
rest wait in arrival order.  Matrices and Krylov state stay on the device
from admission to completion.  A finished request's matrices and vectors
go to a per-template pool and the template's next request is assembled
into them, so that the steady state allocates and frees no device memory
(each device free synchronises the device); pooled arrays are freed,
largest first, when a new allocation would not fit beside them within
``pool_budget`` (by default 80% of the free HBM at start, against 60% for
the admitted requests) or the device's free memory would fall below a 5%
reserve.

The reference solves one system per call (core/linear_solver.py:14-62);
``Server(batched=False, precision="fp64")`` is that one-at-a-time mode on
the GPU (pressure by CG as well), the baseline the batched server is
compared with.
"""

from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from . import device, workloads
from ._lib import DevBuf, check, lib

KINDS = ("quad", "hex", "tet")


# ----------------------------------------------------------------- workload

def template_sizes(nmin=1_000, nmax=1_000_000, per_kind=8):
    """Cell counts of the mesh templates: log-spaced in [nmin, nmax]."""
    return [int(round(math.exp(x))) for x in np.linspace(math.log(nmin), math.log(nmax), per_kind)]


def make_trace(rate=200.0, duration=60.0, seed=5000, nmin=1_000, nmax=1_000_000,
               mix=(0.5, 0.3, 0.2)):
    """Poisson arrivals at ``rate`` per second over ``duration`` seconds;
    cells log-uniform in [nmin, nmax]; kinds quad / hex / tet by ``mix``.
    Returns [(t_arrival, kind, cells, seed)]."""
    rng = np.random.default_rng(seed)
    out, t, i = [], 0.0, 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        u = rng.uniform()
        kind = KINDS[0] if u < mix[0] else (KINDS[1] if u < mix[0] + mix[1] else KINDS[2])
        out.append((t, kind, n, seed + 1 + i))
        i += 1
    return out


def nearest_template(sizes, n):
    """Index of the template size nearest to n in log scale."""
    ls = np.log(np.asarray(sizes, float))
    return int(np.argmin(np.abs(ls - math.log(max(n, 1)))))


def build_template_cases(kind, cells, seed):
    """(pressure case, momentum case) on one mesh of about ``cells`` cells."""
    if kind == "quad":
        s = max(4, int(round(math.sqrt(cells))))
        mesh = workloads.quad_mesh(s)
        mom = workloads.quad_convdiff(s, seed, mesh=mesh)
        pres = workloads.pressure_case(mesh, seed + 1)
    elif kind == "hex":
        s = max(3, int(round(cells ** (1.0 / 3.0))))
        mesh = workloads.hex_mesh(s)
        pres = workloads.hex_pressure(s, seed + 1, mesh=mesh)
        mom = workloads.hex_momentum(s, seed, mesh=mesh)
    else:
        mom = workloads.tet_twofluid(max(64, int(round(cells / 5.3))), seed)
        pres = workloads.pressure_case(mom.mesh, seed + 1)
    return pres, mom


def _build_template(spec):
    kind, cells, seed = spec
    return kind, cells, build_template_cases(kind, cells, seed)


def build_templates(kinds=KINDS, sizes=None, seed=6000, workers=None):
    """Host meshes + recipes of every (kind, size) template, in worker
    processes (forked: call before the GPU context exists)."""
    import multiprocessing as mp
    import os
    from concurrent.futures import ProcessPoolExecutor

    sizes = template_sizes() if sizes is None else list(sizes)
    specs = [(k, n, seed + 100 * i + j) for i, k in enumerate(kinds) for j, n in enumerate(sizes)]
    workers = min(16, os.cpu_count() or 1) if workers is None else workers
    if workers <= 1 or len(specs) < 4:
        return [_build_template(s) for s in specs]
    if L._default_ctx is not None:
        raise RuntimeError("build_templates with workers must run before the GPU context is created")
    # largest first: the pool finishes sooner
    order = sorted(range(len(specs)), key=lambda i: -specs[i][1])
    with ProcessPoolExecutor(workers, mp_context=mp.get_context("fork")) as ex:
        res = dict(zip(order, ex.map(_build_template, [specs[i] for i in order])))
    return [res[i] for i in range(len(specs))]


# ------------------------------------------------------------------ server

class Template:
    """A mesh resident on the device with its two uploaded recipes and one
    prototype matrix per recipe that holds the ILU(0) ordering."""

    def __init__(self, ctx, kind, cells, pres, mom, precond, precision, order):
        self.kind, self.cells = kind, cells
        self.dmesh = device.DeviceMesh(pres.mesh, ctx)
        self.n = self.dmesh.n
        self.rec = [workloads.DeviceRecipe(self.dmesh, pres), workloads.DeviceRecipe(self.dmesh, mom)]
        self.proto = []
        for r in self.rec:
            A = r.assemble(self.dmesh.new_system())
            A.setup_precond(precond, precision, order=order if precond == "ilu0" else None)
            self.proto.append(A)
        vs = 8 if precision == "fp64" else 4
        nvec = (9 if precond == "ilu0" else 6) + (2 if precision != "fp64" else 0)
        # matrices + preconditioners (two), four systems' vectors, b and x
        self.bytes = int(sum(A.bytes() for A in self.proto) + 4 * nvec * (self.n * vs + 256) + 8 * 8 * self.n)
        self.dmesh.mesh = None            # host geometry no longer needed
        # system kinds (True: pressure, False: momentum) solved in FP64 from
        # the start by a mixed-precision server (see calibrate)
        self.fp64_kinds = set()
        self.precond, self.precision = precond, precision

    def calibrate(self, rtol=1e-8, seed=0, check_every=8, pressure=True, fp64_order="segment"):
        """One mixed solve of each prototype matrix with a seeded right-hand
        side (of the momentum matrix only when ``pressure`` is False: the
        pressure system then goes to CG); a kind whose solve does not
        converge in mixed precision is marked for FP64 (route_fp64).
        Returns the solve results."""
        out = []
        if self.precision != "mixed":
            return out
        rng = np.random.default_rng(seed)
        for kind_p, A in zip((True, False), self.proto):
            if kind_p and not pressure:
                continue
            pressure_ = kind_p
            b = DevBuf.from_array(A.ctx, rng.uniform(-1.0, 1.0, self.n))
            x = DevBuf(A.ctx, self.n)
            x.fill(0.0)
            opts = device.default_opts(rtol=rtol, precond=self.precond, precision="mixed",
                                       check_every=check_every)
            r = A.solve_dev(b.p, x.p, opts, setup=False, fallback=False)
            out.append((pressure_, r.info, r.iters))
            b.free()
            x.free()
            if r.info != 0:
                self.route_fp64(pressure_, fp64_order)
        return out

    def route_fp64(self, pressure, order="segment"):
        """Solve this template's systems of one kind (True: pressure, False:
        momentum) in FP64 from the start.  A momentum kind also takes the
        ILU(0) ordering ``order`` for its matrices (by default the segment
        ordering, whose FP64 solves take 25-30% fewer iterations on grids
        than the colour ordering; None keeps the template's ordering).  The
        pressure kind goes to the CG batch and keeps its ordering."""
        self.fp64_kinds.add(pressure)
        if not pressure and order is not None and self.precond == "ilu0":
            self.proto[1].setup_precond(self.precond, self.precision, order=order)

    def new_matrices(self, precond, precision):
        mats = []
        for r, P in zip(self.rec, self.proto):
            A = self.dmesh.new_system()
            if precond == "ilu0":
                A.copy_order_from(P)
            r.assemble(A)
            A.setup_precond(precond, precision)
            mats.append(A)
        return mats

    def close(self):
        for A in self.proto:
            A.close()
        for r in self.rec:
            r.close()
        self.dmesh.close()


@dataclass
class Request:
    rid: int
    t_arrival: float
    template: Template
    seed: int
    t_admit: float = math.nan
    t_done: float = math.nan
    mats: list = field(default_factory=list)
    bufs: list = field(default_factory=list)
    xs: list = field(default_factory=list)
    slots: list = field(default_factory=list)
    results: list = field(default_factory=lambda: [None] * 4)
    remaining: int = 4

    def rhs(self):
        """Pressure b and the three momentum components' b (seeded)."""
        rng = np.random.default_rng(self.seed)
        n = self.template.n
        return [rng.uniform(-1.0, 1.0, n) for _ in range(4)]

    def systems(self):
        """(matrix, b, x) of the four solves."""
        P, M = self.mats
        return [(P, self.bufs[0], self.xs[0])] + [(M, self.bufs[c], self.xs[c]) for c in (1, 2, 3)]

    def release(self):
        for A in self.mats:
            A.close()
        for b in self.bufs + self.xs:
            b.free()
        self.mats, self.bufs, self.xs = [], [], []


@dataclass
class ServeReport:
    mode: str
    precision: str
    requests: int
    solves: int
    elapsed_s: float
    latencies_ms: np.ndarray
    queue_ms: np.ndarray
    rounds: int
    mean_members: float
    peak_hbm_bytes: int
    peak_requests: int
    not_converged: int
    max_rel_resid: float
    fp64_fallbacks: int = 0
    fp64_direct: int = 0
    fp64_continued: int = 0
    arrays_reused: int = 0
    time_s: dict = field(default_factory=dict)

    @property
    def solves_per_s(self):
        return self.solves / self.elapsed_s

    def summary(self):
        lat = self.latencies_ms
        q = self.queue_ms
        return dict(mode=self.mode, precision=self.precision, requests=self.requests, solves=self.solves,
                    elapsed_s=round(self.elapsed_s, 3), solves_per_s=round(self.solves_per_s, 2),
                    requests_per_s=round(self.requests / self.elapsed_s, 2),
                    p50_ms=round(float(np.percentile(lat, 50)), 2),
                    p99_ms=round(float(np.percentile(lat, 99)), 2),
                    mean_ms=round(float(lat.mean()), 2),
                    queue_p50_ms=round(float(np.percentile(q, 50)), 2),
                    queue_p99_ms=round(float(np.percentile(q, 99)), 2),
                    rounds=self.rounds, mean_members=round(self.mean_members, 1),
                    peak_hbm_gb=round(self.peak_hbm_bytes / 1e9, 2), peak_requests=self.peak_requests,
                    not_converged=self.not_converged, max_rel_resid=self.max_rel_resid,
                    fp64_fallbacks=self.fp64_fallbacks, fp64_direct=self.fp64_direct,
                    fp64_continued=self.fp64_continued, arrays_reused=self.arrays_reused,
                    time_s={k: round(v, 2) for k, v in sorted(self.time_s.items())})


class Server:
    """Continuous-batching solver service on one GPU (see the module doc)."""

    def __init__(self, ctx, templates, precision="mixed", precond="ilu0", order="color",
                 batched=True, capacity=2048, hbm_budget=None, pool_budget=None, rtol=1e-8, check_every=8,
                 grid_per_cu=0, trsv_per_cu=0, delta=0.1, log=None, pressure_method="cg", fold_colour0=0):
        self.ctx = ctx
        self.templates = templates
        self.precision, self.precond, self.order = precision, precond, order
        self.batched = batched
        self.capacity = capacity
        info = ctx.info()
        self.hbm_budget = int(0.6 * info["free_bytes"]) if hbm_budget is None else int(hbm_budget)
        # the admitted requests plus the pooled arrays of finished ones
        self.pool_budget = (max(self.hbm_budget, int(0.8 * info["free_bytes"])) if hbm_budget is None
                            else self.hbm_budget) if pool_budget is None else int(pool_budget)
        self.pool_reserve = int(0.05 * info["total_bytes"])
        self.opts = device.default_opts(rtol=rtol, precond=precond, precision=precision, inner_rtol=delta,
                                        check_every=check_every, grid_per_cu=grid_per_cu,
                                        trsv_per_cu=trsv_per_cu, fold_colour0=fold_colour0)
        self.log = log or (lambda *a: None)
        self.batch = None
        if batched:
            h = L.ctypes.c_void_p()
            check(lib().kcfd_batch_create_dynamic(ctx.h, capacity, L.ctypes.byref(self.opts),
                                                  L.ctypes.byref(h)))
            self.batch = h
        self.batch64 = None
        self.batch_cg = None
        self.pressure_cg = pressure_method == "cg"
        self.opts_cg = device.default_opts(rtol=rtol, precond=precond, precision="fp64", method="cg",
                                           check_every=check_every, grid_per_cu=grid_per_cu,
                                           trsv_per_cu=trsv_per_cu)
        self.fallbacks = 0
        self.direct64 = 0
        self.continued = 0      # FP64 solves continued from x for their true residual
        self.rtol = rtol
        # (template, system kind) whose mixed solve stalled: later systems of
        # that kind on that template are solved in FP64 from the start
        self.force64 = set()
        self.keep_solutions = False
        self.solutions = {}
        # finished requests' device arrays by template, reused by the next
        # request of that template (pool_arrays=False: freed at once)
        self.pool_arrays = True
        self._pool = {}
        self._pool_bytes = {id(T): T.bytes for T in templates}
        self.pooled_bytes = 0
        self.reused = 0
        self._t = collections.defaultdict(float)   # host seconds by part of the loop

    # -- request lifecycle ---------------------------------------------------
    def _admit(self, req: Request, now, in_use=0):
        """Assemble the request's matrices, set up their preconditioners and
        upload its right-hand sides.  A finished request's device arrays of
        the same template are reused when pooled (re-assembled in place:
        no device allocation); otherwise new ones are allocated, after
        freeing pooled arrays until the admitted requests (``in_use``
        bytes), the pool and this request fit the HBM budget."""
        t = time.perf_counter()
        T = req.template
        pool = self._pool.get(id(T))
        reuse = bool(pool)
        if pool:
            req.mats, req.bufs, req.xs = pool.pop()
            self.pooled_bytes -= T.bytes
            for r, A in zip(T.rec, req.mats):
                r.assemble(A)
                A.setup_precond(self.precond, self.precision)
            for buf, b in zip(req.bufs, req.rhs()):
                buf.upload(b)
            self.reused += 1
        else:
            te = time.perf_counter()
            self._evict(in_use + T.bytes, T.bytes)
            self._t["evict"] += time.perf_counter() - te
            req.mats = T.new_matrices(self.precond, self.precision)
            req.bufs = [DevBuf.from_array(self.ctx, b) for b in req.rhs()]
            req.xs = [DevBuf(self.ctx, T.n) for _ in range(4)]
        for x in req.xs:
            x.fill(0.0)
        req.t_admit = now
        self._t["admit_reused" if reuse else "admit_new"] += time.perf_counter() - t

    def _finish(self, req: Request, now):
        req.t_done = now
        if self.keep_solutions:
            self.solutions[req.rid] = ([x.download() for x in req.xs], list(req.results))
        if self.pool_arrays and req.mats:
            # keep the arrays for the template's next request (a device
            # free synchronises the device; an allocation costs as much)
            self._pool.setdefault(id(req.template), []).append((req.mats, req.bufs, req.xs))
            self.pooled_bytes += req.template.bytes
            req.mats, req.bufs, req.xs = [], [], []
        else:
            req.release()

    def _evict(self, need, alloc=0):
        """Free pooled arrays, largest template first, until ``need`` bytes
        plus the pool fit ``pool_budget`` and the device has ``alloc``
        bytes free beside a reserve (the byte estimates do not see the
        allocator's rounding).  The batches' kept vector blocks of detached
        systems (counted in the pooled requests' bytes) are freed first."""
        def over():
            if need + self.pooled_bytes > self.pool_budget:
                return True
            return alloc > 0 and self.ctx.info()["free_bytes"] < alloc + self.pool_reserve
        trimmed = False
        while self.pooled_bytes and over():
            if not trimmed:
                trimmed = True
                for h in (self.batch, self.batch64, self.batch_cg):
                    if h is not None:
                        check(lib().kcfd_batch_trim(h))
                continue
            tid = max((k for k, v in self._pool.items() if v), key=lambda k: self._pool_bytes[k])
            mats, bufs, xs = self._pool[tid].pop()
            self.pooled_bytes -= self._pool_bytes[tid]
            for A in mats:
                A.close()
            for b in bufs + xs:
                b.free()

    def free_pool(self):
        """Free every pooled array."""
        self._evict(self.pool_budget + 1)

    def _fp64_batch(self):
        """The FP64 batch that takes over mixed solves that stalled."""
        if self.batch64 is None:
            o = L.SolveOpts()
            L.ctypes.pointer(o)[0] = self.opts
            o.precision = L.FP64
            h = L.ctypes.c_void_p()
            check(lib().kcfd_batch_create_dynamic(self.ctx.h, self.capacity, L.ctypes.byref(o),
                                                  L.ctypes.byref(h)))
            self.batch64 = h
        return self.batch64

    def _cg_batch(self):
        """The FP64 CG batch of the pressure-correction solves."""
        if self.batch_cg is None:
            h = L.ctypes.c_void_p()
            check(lib().kcfd_batch_create_dynamic(self.ctx.h, self.capacity, L.ctypes.byref(self.opts_cg),
                                                  L.ctypes.byref(h)))
            self.batch_cg = h
        return self.batch_cg

    def close(self):
        self.free_pool()
        for name in ("batch", "batch64", "batch_cg"):
            h = getattr(self, name, None)
            if h is not None:
                lib().kcfd_batch_destroy(h)
                setattr(self, name, None)

    # -- the loop ------------------------------------------------------------
    def run(self, trace, time_scale=1.0):
        """Serve ``trace`` = [(t_arrival, template_index, seed)] in real time
        (arrival times multiplied by ``time_scale``).  Returns a ServeReport."""
        reqs = [Request(i, t * time_scale, self.templates[ti], sd) for i, (t, ti, sd) in enumerate(trace)]
        pending = collections.deque(reqs)
        queue = collections.deque()
        owner, owner64 = {}, {}
        self.fallbacks = 0
        self.continued = 0
        reused0 = self.reused
        self._t = collections.defaultdict(float)
        cont = {}
        in_use = 0
        peak_bytes = peak_reqs = 0
        active = 0
        rounds = 0
        members_sum = 0
        done = []
        fin = (L.ctypes.c_int * max(1, self.capacity))()
        nfin = L.ctypes.c_int()
        res = L.SolveResult()
        t0 = time.perf_counter()
        t_log = 10.0
        while pending or queue or active:
            now = time.perf_counter() - t0
            if now >= t_log:       # progress, every 10 s of serving
                self.log(f"serve t={now:.0f}s: {len(done)} done, {active} on the device "
                         f"({len(owner)} + {len(owner64)} systems), {len(queue)} queued, "
                         f"{len(pending)} not yet arrived")
                t_log = now + 10.0
            while pending and pending[0].t_arrival <= now:
                queue.append(pending.popleft())
            if not self.batched:
                if not queue:
                    time.sleep(max(0.0, min(pending[0].t_arrival - now, 0.05)) if pending else 0.0)
                    continue
                req = queue.popleft()
                self._admit(req, time.perf_counter() - t0, in_use)
                peak_reqs = max(peak_reqs, 1)
                peak_bytes = max(peak_bytes, req.template.bytes)
                ts = time.perf_counter()
                for k, (A, b, x) in enumerate(req.systems()):
                    o = self.opts_cg if (k == 0 and self.pressure_cg) else self.opts
                    r = A.solve_dev(b.p, x.p, o, setup=False)
                    its, tries = r.iters, 0
                    while r.info == 0 and not r.rel_resid < self.rtol and tries < 8:   # continue from x
                        tries += 1
                        r = A.solve_dev(b.p, x.p, o, setup=False)
                        its += r.iters
                        self.continued += 1
                    req.results[k] = (r.info, its, r.refinements, r.rel_resid)
                self._t["solve_one_at_a_time"] += time.perf_counter() - ts
                self._finish(req, time.perf_counter() - t0)
                done.append(req)
                rounds += 1
                members_sum += 4
                continue
            # admit in arrival order while the footprint and the slots allow
            while queue and in_use + queue[0].template.bytes <= self.hbm_budget and \
                    4 * (active + 1) <= self.capacity:
                req = queue.popleft()
                self._admit(req, time.perf_counter() - t0, in_use)
                for k, (A, b, x) in enumerate(req.systems()):
                    slot = L.ctypes.c_int()
                    if k == 0 and self.pressure_cg:
                        check(lib().kcfd_batch_attach(self._cg_batch(), A.h, b.p, x.p, L.ctypes.byref(slot)))
                        owner64[("cg", slot.value)] = (req, k, None)
                    elif self.precision == "mixed" and ((k == 0) in req.template.fp64_kinds or
                                                      (id(req.template), k == 0) in self.force64):
                        if k == 0:   # a pressure correction that needs FP64: the CG batch
                            check(lib().kcfd_batch_attach(self._cg_batch(), A.h, b.p, x.p,
                                                          L.ctypes.byref(slot)))
                            owner64[("cg", slot.value)] = (req, k, None)
                        else:
                            check(lib().kcfd_batch_attach(self._fp64_batch(), A.h, b.p, x.p,
                                                          L.ctypes.byref(slot)))
                            owner64[slot.value] = (req, k, None)
                        self.direct64 += 1
                    else:
                        check(lib().kcfd_batch_attach(self.batch, A.h, b.p, x.p, L.ctypes.byref(slot)))
                        owner[slot.value] = (req, k)
                    req.slots.append(slot.value)
                in_use += req.template.bytes
                active += 1
                peak_bytes = max(peak_bytes, in_use)
                peak_reqs = max(peak_reqs, active)
            if queue and not active:
                raise RuntimeError(f"request {queue[0].rid} ({queue[0].template.bytes} bytes) exceeds "
                                   f"the HBM budget ({self.hbm_budget} bytes)")
            if not active:
                if pending:
                    time.sleep(max(0.0, min(pending[0].t_arrival - (time.perf_counter() - t0), 0.05)))
                continue
            finished = []
            if owner:
                ts = time.perf_counter()
                check(lib().kcfd_batch_step(self.batch, fin, L.ctypes.byref(nfin)))
                self._t[f"step_{self.precision}_bicgstab"] += time.perf_counter() - ts
                rounds += 1
                members_sum += len(owner)
                for j in range(nfin.value):
                    slot = fin[j]
                    req, k = owner.pop(slot)
                    check(lib().kcfd_batch_detach(self.batch, slot, L.ctypes.byref(res)))
                    r = (res.info, res.iters, res.refinements, res.rel_resid)
                    if res.info != 0 and self.precision == "mixed":
                        # FP32 recurrences stalled: continue in FP64 from the
                        # best iterate (the matrix also holds its FP64 set-up)
                        # (a pressure correction continues by CG, the
                        # reference's pressure solver)
                        A, b, x = req.systems()[k]
                        s64 = L.ctypes.c_int()
                        if k == 0:
                            check(lib().kcfd_batch_attach(self._cg_batch(), A.h, b.p, x.p,
                                                          L.ctypes.byref(s64)))
                            owner64[("cg", s64.value)] = (req, k, r)
                        else:
                            check(lib().kcfd_batch_attach(self._fp64_batch(), A.h, b.p, x.p,
                                                          L.ctypes.byref(s64)))
                            owner64[s64.value] = (req, k, r)
                        self.fallbacks += 1
                        self.force64.add((id(req.template), k == 0))
                    else:
                        finished.append((req, k, r))
            # the FP64 batches: pressure by CG, and FP64 BiCGSTAB (stalled or
            # FP64-routed mixed solves, or every solve of an FP64 server)
            for tag, bh in (("cg", self.batch_cg), (None, self.batch64)):
                if bh is None or not any(isinstance(key, tuple) == (tag == "cg") for key in owner64):
                    continue
                ts = time.perf_counter()
                check(lib().kcfd_batch_step(bh, fin, L.ctypes.byref(nfin)))
                self._t["step_fp64_cg" if tag == "cg" else "step_fp64_bicgstab"] += time.perf_counter() - ts
                for j in range(nfin.value):
                    key = (tag, fin[j]) if tag else fin[j]
                    req, k, r1 = owner64.pop(key)
                    check(lib().kcfd_batch_detach(bh, fin[j], L.ctypes.byref(res)))
                    r = ((res.info, res.iters, 0, res.rel_resid) if r1 is None else
                         (res.info, r1[1] + res.iters, r1[2], res.rel_resid))
                    cont[(req.rid, k)] = cont.get((req.rid, k), 0) + 1
                    if res.info == 0 and not res.rel_resid < self.rtol and cont[(req.rid, k)] <= 8:
                        # scipy's stopping test is on the recursive residual;
                        # the FP64 true residual decides: continue from x
                        A, b, x = req.systems()[k]
                        s2 = L.ctypes.c_int()
                        check(lib().kcfd_batch_attach(bh, A.h, b.p, x.p, L.ctypes.byref(s2)))
                        owner64[(tag, s2.value) if tag else s2.value] = (req, k, r)
                        self.continued += 1
                        continue
                    finished.append((req, k, r))
            now = time.perf_counter() - t0
            for req, k, r in finished:
                req.results[k] = r
                req.remaining -= 1
                if req.remaining == 0:
                    self._finish(req, now)
                    done.append(req)
                    in_use -= req.template.bytes
                    active -= 1
        elapsed = (max(r.t_done for r in done) if done else 0.0)
        lat = np.array([1e3 * (r.t_done - r.t_arrival) for r in done])
        qd = np.array([1e3 * (r.t_admit - r.t_arrival) for r in done])
        allres = [x for r in done for x in r.results]
        return ServeReport(mode="batched" if self.batched else "one-at-a-time", precision=self.precision,
                           requests=len(done), solves=4 * len(done), elapsed_s=elapsed,
                           latencies_ms=lat, queue_ms=qd, rounds=rounds,
                           mean_members=members_sum / max(1, rounds), peak_hbm_bytes=peak_bytes,
                           peak_requests=peak_reqs,
                           not_converged=sum(1 for x in allres if x[0] != 0),
                           max_rel_resid=float(max((x[3] for x in allres), default=0.0)),
                           fp64_fallbacks=self.fallbacks, fp64_direct=self.direct64,
                           fp64_continued=self.continued, arrays_reused=self.reused - reused0,
                           time_s=dict(self._t))
