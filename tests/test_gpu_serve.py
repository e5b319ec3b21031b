"""Serving loop on the GPU: continuous batching, HBM admission, and every
result the same alone and in the mix."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def templates(ctx):
    from kcfd_amd import serve as S
    out = []
    for i, (kind, n) in enumerate([("quad", 1000), ("hex", 3375), ("tet", 4000), ("quad", 9000),
                                   ("hex", 8000)]):
        pres, mom = S.build_template_cases(kind, n, 6000 + i)
        out.append(S.Template(ctx, kind, n, pres, mom, "ilu0", "mixed", "color"))
    yield out
    for t in out:
        t.close()


def _trace(n, rate, ntemp, seed=1):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    return [(float(t[i]), int(rng.integers(ntemp)), 7000 + i) for i in range(n)]


def test_serve_mix_equals_alone(ctx, templates):
    """Requests that overlap on the device give exactly the solutions (and
    iteration counts) they give when each system is solved alone: the
    pressure corrections by CG in FP64, the momentum solves by
    mixed-precision BiCGSTAB."""
    from kcfd_amd import serve as S
    trace = _trace(24, 2000.0, len(templates))
    srv = S.Server(ctx, templates, precision="mixed", precond="ilu0", batched=True, check_every=4)
    srv.keep_solutions = True
    try:
        rep = srv.run(trace)
        print(rep.summary())
        for rid, (xs, res) in sorted(srv.solutions.items()):
            for k, q in enumerate(res):
                if q[0] != 0:
                    print(f"not converged: request {rid} template {trace[rid][1]} system {k}: "
                          f"info={q[0]} iters={q[1]} refinements={q[2]} rel_resid={q[3]:.3e}")
        assert rep.requests == 24 and rep.not_converged == 0 and rep.max_rel_resid < 1e-8
        assert rep.peak_requests > 1 and rep.mean_members > 4     # work overlapped
        for rid, (t, ti, seed) in enumerate(trace):
            req = S.Request(rid, 0.0, templates[ti], seed)
            srv._admit(req, 0.0)
            for k, (A, b, x) in enumerate(req.systems()):
                # the pressure correction is solved by CG (FP64), the momentum
                # components by mixed-precision BiCGSTAB
                r = A.solve_dev(b.p, x.p, srv.opts_cg if k == 0 else srv.opts, setup=False)
                xs_mix, res_mix = srv.solutions[rid]
                assert r.iters == res_mix[k][1] and r.info == res_mix[k][0] == 0
                assert np.array_equal(x.download(), xs_mix[k]), (rid, k)
            req.release()
    finally:
        srv.close()


def test_serve_pooled_arrays_equal_fresh(ctx, templates):
    """Requests assembled into a finished request's pooled matrices and
    vectors give exactly the solutions and iteration counts of requests
    given freshly allocated ones; the pool is reused, and freed on close."""
    from kcfd_amd import serve as S
    trace = _trace(16, 150.0, 2, seed=4)       # two templates: later requests reuse arrays
    sols = {}
    for pool in (True, False):
        srv = S.Server(ctx, templates, precision="mixed", precond="ilu0", batched=True, check_every=4)
        srv.pool_arrays = pool
        srv.keep_solutions = True
        try:
            rep = srv.run(trace)
            print(pool, rep.summary(), "reused", srv.reused)
            assert rep.requests == 16 and rep.not_converged == 0
            assert (srv.reused > 0) == pool and (srv.pooled_bytes > 0) == pool
            sols[pool] = srv.solutions
        finally:
            srv.close()
        assert srv.pooled_bytes == 0 and not any(srv._pool.values())
    for rid in range(16):
        xa, ra = sols[True][rid]
        xb, rb = sols[False][rid]
        assert ra == rb, rid
        for k in range(4):
            assert np.array_equal(xa[k], xb[k]), (rid, k)


def test_dynamic_batch_reuses_detached_blocks(ctx, templates):
    """A detached system's vector block is reused by the next attach of the
    same size (the batch's device bytes stay the same), the solve from it
    equals the first one bit for bit, and kcfd_batch_trim frees the kept
    blocks."""
    from kcfd_amd import _lib as L, serve as S
    lib, check = L.lib(), L.check
    srv = S.Server(ctx, templates, precision="mixed", precond="ilu0", batched=True, check_every=4)
    req = S.Request(0, 0.0, templates[1], 7300)
    srv._admit(req, 0.0)
    A, b, x = req.systems()[1]                    # a momentum component, mixed BiCGSTAB
    fin = (L.ctypes.c_int * srv.capacity)()
    nfin = L.ctypes.c_int()
    res = L.SolveResult()
    by = L.ctypes.c_longlong()

    def solve_once():
        x.fill(0.0)
        slot = L.ctypes.c_int()
        check(lib.kcfd_batch_attach(srv.batch, A.h, b.p, x.p, L.ctypes.byref(slot)))
        check(lib.kcfd_batch_bytes(srv.batch, L.ctypes.byref(by)))
        attached = by.value
        for _ in range(10000):
            check(lib.kcfd_batch_step(srv.batch, fin, L.ctypes.byref(nfin)))
            if nfin.value:
                break
        assert nfin.value == 1 and fin[0] == slot.value
        check(lib.kcfd_batch_detach(srv.batch, slot.value, L.ctypes.byref(res)))
        return attached, (res.info, res.iters, res.refinements, res.rel_resid), x.download()

    try:
        check(lib.kcfd_batch_bytes(srv.batch, L.ctypes.byref(by)))
        empty = by.value
        b1, r1, x1 = solve_once()
        check(lib.kcfd_batch_bytes(srv.batch, L.ctypes.byref(by)))
        assert by.value == b1 > empty                 # detached: the block is kept
        b2, r2, x2 = solve_once()
        assert b2 == b1                               # reused, not allocated
        assert r1 == r2 and r1[0] == 0 and np.array_equal(x1, x2)
        check(lib.kcfd_batch_trim(srv.batch))
        check(lib.kcfd_batch_bytes(srv.batch, L.ctypes.byref(by)))
        assert by.value == empty
        b3, r3, x3 = solve_once()                     # a fresh block after the trim
        assert b3 == b1 and r3 == r1 and np.array_equal(x3, x1)
    finally:
        req.release()
        srv.close()


def test_serve_admission_by_hbm(ctx, templates):
    """A burst admitted under an HBM budget: the footprint in use never
    exceeds it, fewer requests overlap than without it, all complete."""
    from kcfd_amd import serve as S
    trace = _trace(10, 5000.0, len(templates), seed=2)
    reps = {}
    for budget in (None, int(2.5 * max(t.bytes for t in templates))):
        srv = S.Server(ctx, templates, precision="mixed", batched=True, hbm_budget=budget)
        try:
            reps[budget] = srv.run(trace)
        finally:
            srv.close()
        print(budget, reps[budget].summary())
        assert reps[budget].requests == 10 and reps[budget].not_converged == 0
    free, capped = reps[None], reps[budget]
    assert capped.peak_hbm_bytes <= budget
    assert capped.peak_requests < free.peak_requests
    assert capped.queue_ms.max() > free.queue_ms.max()


def test_serve_one_at_a_time_fp64(ctx, templates):
    from kcfd_amd import serve as S
    srv = S.Server(ctx, templates, precision="fp64", batched=False)
    # the templates were set up for mixed: the baseline sets up its own
    rep = srv.run(_trace(6, 1000.0, len(templates), seed=3))
    print(rep.summary())
    assert rep.requests == 6 and rep.not_converged == 0 and rep.peak_requests == 1


def test_serve_routes_stalled_kind_to_fp64(ctx, templates):
    """With every system solved by mixed BiCGSTAB (pressure_method='bicgstab'),
    a (template, system kind) marked as stalling in mixed precision is solved
    in FP64 from the start -- a pressure correction by the FP64 CG batch --;
    its results equal the FP64 CG solve of the same system alone, and the
    other systems stay mixed."""
    from kcfd_amd import device, serve as S
    trace = [(0.0, 0, 7100 + i) for i in range(3)] + [(0.0, 1, 7200)]
    srv = S.Server(ctx, templates, precision="mixed", precond="ilu0", batched=True, check_every=4,
                   pressure_method="bicgstab")
    srv.keep_solutions = True
    srv.force64.add((id(templates[0]), True))      # template 0's pressure solves
    try:
        rep = srv.run(trace)
        assert rep.requests == 4 and rep.not_converged == 0 and rep.max_rel_resid < 1e-8
        assert rep.fp64_direct == 3 and rep.fp64_fallbacks == 0
        o64 = device.default_opts(rtol=1e-8, precond="ilu0", precision="fp64", method="cg", check_every=4)
        for rid, (t, ti, seed) in enumerate(trace):
            req = S.Request(rid, 0.0, templates[ti], seed)
            srv._admit(req, 0.0)
            xs_mix, res_mix = srv.solutions[rid]
            for k, (A, b, x) in enumerate(req.systems()):
                fp64 = ti == 0 and k == 0
                r = A.solve_dev(b.p, x.p, o64 if fp64 else srv.opts, setup=False)
                assert r.iters == res_mix[k][1] and r.info == res_mix[k][0] == 0, (rid, k)
                assert np.array_equal(x.download(), xs_mix[k]), (rid, k)
            req.release()
    finally:
        srv.close()


def test_serve_fp64_momentum_on_segment_order(ctx, templates):
    """A template whose momentum kind is routed to FP64 (route_fp64) solves
    its momentum components in the FP64 batch on the segment ILU(0)
    ordering; each result equals the FP64 solve of the same system alone,
    bit for bit, and the other template stays mixed."""
    from kcfd_amd import device, serve as S
    T = templates[3]                                  # quad, 9,000 cells
    T.route_fp64(False, "segment")
    try:
        assert T.proto[1].ilu0_order()[1] == 2        # two segment colours
        trace = [(0.0, 3, 7400 + i) for i in range(3)] + [(0.0, 1, 7500)]
        srv = S.Server(ctx, templates, precision="mixed", precond="ilu0", batched=True, check_every=4)
        srv.keep_solutions = True
        try:
            rep = srv.run(trace)
            print(rep.summary())
            assert rep.requests == 4 and rep.not_converged == 0 and rep.max_rel_resid < 1e-8
            assert rep.fp64_direct == 9 and rep.fp64_fallbacks == 0
            o64 = device.default_opts(rtol=1e-8, precond="ilu0", precision="fp64", check_every=4)
            for rid, (t, ti, seed) in enumerate(trace):
                req = S.Request(rid, 0.0, templates[ti], seed)
                srv._admit(req, 0.0)
                xs, res = srv.solutions[rid]
                for k, (A, b, x) in enumerate(req.systems()):
                    o = srv.opts_cg if k == 0 else (o64 if ti == 3 else srv.opts)
                    r = A.solve_dev(b.p, x.p, o, setup=False)
                    assert r.iters == res[k][1] and r.info == res[k][0] == 0, (rid, k)
                    assert np.array_equal(x.download(), xs[k]), (rid, k)
                req.release()
        finally:
            srv.close()
    finally:
        T.fp64_kinds.discard(False)
        T.proto[1].setup_precond("ilu0", "mixed", order="color")
