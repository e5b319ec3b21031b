"""Config 5 on the GPU: the first half second of the serving trace
(Poisson arrivals at 200 requests/s, meshes of 1e3 to 1e6 cells, quad / hex
/ tet 50/30/20, each request a pressure correction plus three momentum
solves), served with continuous batching in mixed precision.

Every served solution is checked on the host: its FP64 relative residual
||b - A x|| / ||b|| below 1e-8 with the device-assembled matrix of its mesh
template (whose assembly equals the oracle's, tests/test_gpu_assembly.py).
Every pressure correction is a CG solve in FP64; on the templates of up to
20,000 cells each one is compared with the oracle's CG restated in the
device's dot-product order (oracle/solve.py: cg_restated, equal to scipy's
cg with numpy's dot): the same iterations and x bit for bit, whatever else
was on the device at the time.
"""

import numpy as np
import pytest

from helpers import restated_device

pytestmark = pytest.mark.gpu

SLICE_S = 0.5


@pytest.fixture(scope="module")
def served(ctx):
    from kcfd_amd import serve as S
    sizes = S.template_sizes(per_kind=8)
    trace = [x for x in S.make_trace(rate=200.0, duration=60.0) if x[0] < SLICE_S]
    kidx = {k: i for i, k in enumerate(S.KINDS)}
    tr = [(ta, kidx[k] * len(sizes) + S.nearest_template(sizes, n), sd) for ta, k, n, sd in trace]
    used = sorted({ti for _, ti, _ in tr})
    print(f"trace slice: {len(tr)} requests, {len(used)} templates, sizes {sizes}", flush=True)
    temps = {}
    for ti in used:
        kind, n = S.KINDS[ti // len(sizes)], sizes[ti % len(sizes)]
        pres, mom = S.build_template_cases(kind, n, 6000 + 100 * (ti // len(sizes)) + ti % len(sizes))
        temps[ti] = S.Template(ctx, kind, n, pres, mom, "ilu0", "mixed", "color")
    # the server indexes templates by position: give it a dense list
    order = list(temps)
    pos = {ti: j for j, ti in enumerate(order)}
    tlist = [temps[ti] for ti in order]
    for tp in tlist:
        tp.calibrate(pressure=False)
    trd = [(ta, pos[ti], sd) for ta, ti, sd in tr]
    srv = S.Server(ctx, tlist, precision="mixed", precond="ilu0", batched=True)
    srv.keep_solutions = True
    try:
        rep = srv.run(trd)
    finally:
        srv.close()
    print(rep.summary(), flush=True)
    yield tlist, trd, srv.solutions, rep
    for tp in tlist:
        tp.close()


def test_trace_slice_served_and_verified(served):
    from kcfd_amd import serve as S
    tlist, trd, sols, rep = served
    cells = [tlist[ti].n for _, ti, _ in trd]
    assert rep.requests == len(trd) and rep.not_converged == 0 and rep.max_rel_resid < 1e-8
    assert min(cells) < 2000 and max(cells) > 900_000          # the slice spans 1e3..1e6 cells
    mats = {}
    worst = 0.0
    for rid, (ta, ti, sd) in enumerate(trd):
        tp = tlist[ti]
        if ti not in mats:
            mats[ti] = [A.to_scipy() for A in tp.proto]
        P, M = mats[ti]
        bs = S.Request(rid, 0.0, tp, sd).rhs()
        xs, res = sols[rid]
        for k in range(4):
            A = P if k == 0 else M
            r = float(np.linalg.norm(bs[k] - A @ xs[k]) / np.linalg.norm(bs[k]))
            worst = max(worst, r)
            assert res[k][0] == 0 and r < 1e-8 * (1 + 1e-6), (rid, k, r, res[k])
    print(f"{len(trd)} requests, {4 * len(trd)} solutions: worst host FP64 relative residual {worst:.2e}")


def test_trace_slice_pressure_equals_restated_cg(served):
    from kcfd_amd import serve as S
    tlist, trd, sols, rep = served
    checked = 0
    for rid, (ta, ti, sd) in enumerate(trd):
        tp = tlist[ti]
        if tp.n > 20_000:
            continue
        A = tp.proto[0].to_scipy()
        perm = tp.proto[0].ilu0_order()[0]
        b = S.Request(rid, 0.0, tp, sd).rhs()[0]
        xs, res = sols[rid]
        xr, info, it = restated_device(A, b, solver="cg", perm=perm, rtol=1e-8, precond="ilu0_color")
        assert (res[0][0], res[0][1]) == (info, it), (rid, res[0], info, it)
        assert np.array_equal(xs[0], xr), rid
        checked += 1
    print(f"{checked} pressure corrections equal to the restated CG bit for bit")
    assert checked >= 10
