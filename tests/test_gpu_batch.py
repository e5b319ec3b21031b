"""Batched driver on the GPU: uneven quad/hex/tet systems assembled on the
device and solved in one batch, FP64 and mixed, plus the autotuner.  Each
FP64 batch result equals the oracle's BiCGSTAB restated in the device's dot
product order, bit for bit."""

import numpy as np
import pytest

from helpers import oracle_system, rel_l2, restated_device
from oracle import solve as osolve

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch(ctx):
    from kcfd_amd import batch as B, workloads as W
    specs = W.batch_specs(10, seed=4000, nmin=2000, nmax=20000)
    cases = B.build_cases(specs, workers=1)
    refs = [oracle_system(c)[:2] for c in cases]
    sb = B.SystemBatch.assemble(ctx, cases)
    yield sb, refs, specs
    sb.close()


@pytest.mark.parametrize("precond,order", [("jacobi", "color"), ("ilu0", "color"), ("ilu0", "segment")])
def test_batch_fp64_matches_oracle(small_batch, precond, order):
    """(the segment ordering: its quad and hex members take the segment
    apply kernel, the tet members the sync-free solves, in one batch)"""
    sb, refs, specs = small_batch
    r = sb.solve(precision="fp64", precond=precond, order=order)
    xs = sb.solutions()
    for k, (res, (A, b), x) in enumerate(zip(r.results, refs, xs)):
        # the device-assembled matrix is the oracle's
        Ad = sb.mats[k].to_scipy()
        assert abs(Ad - A).max() <= 1e-14 * abs(A).max()
        tr = osolve.true_rel_residual(A, x, b)
        print(f"{specs[k]} {precond} {order}: iters={res.iters} info={res.info} resid={res.rel_resid:.2e} "
              f"host={tr:.2e}")
        assert res.info == 0 and tr < 1e-8 and abs(tr - res.rel_resid) <= 1e-3 * tr + 1e-15
    # every system equals the restated solver in the device's dot order (each
    # system's tiles are summed as if it were alone in the batch)
    for k, res in enumerate(r.results):
        m = sb.mats[k]
        A, b = m.to_scipy(), sb.bs[k].download()
        if precond == "ilu0":
            out = restated_device(A, b, perm=m.ilu0_order()[0], rtol=1e-8, precond="ilu0_color")
        else:
            out = restated_device(A, b, rtol=1e-8, precond="jacobi")
        assert (res.info, res.iters) == out[1:3], (k, specs[k])
        assert np.array_equal(xs[k], out[0]), (k, specs[k])


def test_batch_mixed_error_vs_fp64(small_batch):
    sb, refs, specs = small_batch
    sb.solve(precision="fp64", precond="ilu0", order="color")
    x64 = sb.solutions()
    r = sb.solve(precision="mixed", precond="ilu0", order="color")
    xm = sb.solutions()
    for k, (res, (A, b)) in enumerate(zip(r.results, refs)):
        tr = osolve.true_rel_residual(A, xm[k], b)
        err = rel_l2(xm[k], x64[k])
        print(f"{specs[k]} mixed: iters={res.iters} refinements={res.refinements} resid={tr:.2e} "
              f"err vs fp64={err:.2e}")
        assert res.info == 0 and tr < 1e-8
    s = r.summary()
    assert s["converged"] == len(refs)


def test_autotune_small_batch(small_batch, tmp_path, monkeypatch):
    from kcfd_amd import tune
    sb, refs, _ = small_batch
    monkeypatch.setenv("KCFD_TUNE_CACHE", str(tmp_path / "tune.json"))
    choice = tune.autotune(sb, "mixed", "ilu0", "color", probe_iters=8)
    assert choice["grid_per_cu"] in tune.GRID_PER_CU and choice["trsv_per_cu"] in tune.TRSV_PER_CU
    assert choice["check_every"] in tune.CHECK_EVERY
    assert tune.autotune(sb, "mixed", "ilu0", "color") == choice      # from the cache
    r = sb.solve(precision="mixed", precond="ilu0", order="color", **choice)
    assert all(q.info == 0 for q in r.results)


def test_profile_leaves_batch_usable(small_batch):
    """kcfd_batch_profile runs real iterations (state restored afterwards):
    a solve after it reproduces the solve before it bit for bit."""
    from kcfd_amd import _lib as L, device
    sb, refs, _ = small_batch
    sb.setup("ilu0", "fp64", "color")
    b = sb.make_batch(device.default_opts(rtol=1e-8, precond="ilu0", precision="fp64"))
    try:
        r0 = sb.solve(precision="fp64", precond="ilu0", order="color", batch=b)
        x0 = sb.solutions()
        ms = L.ctypes.c_double()
        for phase in (1, 2, 3, 4, 5, 6):
            L.check(L.lib().kcfd_batch_profile(b.h, phase, 3, L.ctypes.byref(ms)))
            assert ms.value > 0
        r1 = sb.solve(precision="fp64", precond="ilu0", order="color", batch=b)
        x1 = sb.solutions()
    finally:
        b.close()
    assert [q.iters for q in r0.results] == [q.iters for q in r1.results]
    assert all(np.array_equal(a, c) for a, c in zip(x0, x1))
