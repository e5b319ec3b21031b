"""The 512-system batch the bench times -- uneven quad / hex / tet systems of
2,000 to 250,000 cells (workloads.batch_specs(512), 24.5 million rows in
all) -- assembled on the device and solved as one batch with colour-ordered
ILU(0), in FP64 and in mixed precision.

Every system's solution is checked on the host: FP64 relative residual
||b - A x|| / ||b|| below 1e-8 with its device-assembled matrix.  For a
sample of each kind (the smallest system and the one nearest 40,000 cells)
the device assembly equals the oracle's, and the FP64 batch result equals
the oracle's BiCGSTAB restated in the device's dot product order
(helpers.restated_device): the same stop and iteration count and x bit for
bit, although 511 other systems share every launch and the tile tables are
compacted as systems finish.
"""

import numpy as np
import pytest

from helpers import oracle_system, restated_device

pytestmark = pytest.mark.gpu

NSYS = 512


@pytest.fixture(scope="module")
def batch512(ctx):
    from kcfd_amd import batch as B, workloads as W
    specs = W.batch_specs(NSYS)
    print(f"building {NSYS} host cases ({sum(s[1] for s in specs)} cells requested)", flush=True)
    cases = B.build_cases(specs, workers=16)
    sample = []
    for kind in ("quad", "hex", "tet"):
        ks = [k for k, s in enumerate(specs) if s[0] == kind]
        sample.append(min(ks, key=lambda k: specs[k][1]))
        sample.append(min(ks, key=lambda k: abs(specs[k][1] - 40_000)))
    kept = {k: cases[k] for k in sample}
    sb = B.SystemBatch.assemble(ctx, cases)
    del cases
    print(f"assembled {len(sb.mats)} systems, {sb.total_rows} rows", flush=True)
    yield sb, specs, kept
    sb.close()


def _host_residuals(sb):
    xs = sb.solutions()
    out = []
    for m, bb, x in zip(sb.mats, sb.bs, xs):
        A, b = m.to_scipy(), bb.download()
        out.append(float(np.linalg.norm(b - A @ x) / np.linalg.norm(b)))
    return np.array(out), xs


def test_batch512_fp64_verified_and_restated(batch512):
    sb, specs, kept = batch512
    r = sb.solve(precision="fp64", precond="ilu0", order="color")
    s = r.summary()
    print({k: s[k] for k in ("nsys", "solves_per_s", "converged", "iters_mean", "iters_max", "max_rel_resid")},
          flush=True)
    res, xs = _host_residuals(sb)
    bad = [(k, specs[k], q.info, q.iters, res[k]) for k, q in enumerate(r.results)
           if q.info != 0 or not res[k] < 1e-8]
    print(f"host FP64 residuals: worst {res.max():.2e}", flush=True)
    assert not bad, bad[:8]
    for k, case in kept.items():
        m = sb.mats[k]
        A = m.to_scipy()
        Ao, bo, _ = oracle_system(case)
        assert np.array_equal(A.indptr, Ao.indptr) and np.array_equal(A.indices, Ao.indices)
        assert abs(A - Ao).max() <= 1e-14 * abs(Ao).max()
        b = sb.bs[k].download()
        assert np.max(np.abs(b - bo)) <= 1e-14 * np.max(np.abs(bo))
        perm = m.ilu0_order()[0]
        xr, info, it, half = restated_device(A, b, perm=perm, rtol=1e-8, precond="ilu0_color")
        q = r.results[k]
        print(f"system {k} {specs[k]}: device {q.info}/{q.iters}, restated {info}/{it} (half step {half})",
              flush=True)
        assert (q.info, q.iters) == (info, it), (k, specs[k])
        assert np.array_equal(xs[k], xr), (k, specs[k])


def test_batch512_mixed_verified(batch512):
    sb, specs, _ = batch512
    r = sb.solve(precision="mixed", precond="ilu0", order="color")
    s = r.summary()
    print({k: s[k] for k in ("nsys", "solves_per_s", "converged", "iters_mean", "iters_max",
                             "refinements_total", "fp64_fallbacks", "max_rel_resid")}, flush=True)
    res, _ = _host_residuals(sb)
    bad = [(k, specs[k], q.info, q.iters, res[k]) for k, q in enumerate(r.results)
           if q.info != 0 or not res[k] < 1e-8]
    print(f"host FP64 residuals: worst {res.max():.2e}", flush=True)
    assert not bad, bad[:8]
