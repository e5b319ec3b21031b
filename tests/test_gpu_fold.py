"""The folded BiCGSTAB product over a two-colour ILU(0)
(kcfd_solve_opts.fold_colour0).

With the colours ordered 0 then 1, the colour-0 rows of the ILU(0) factor U
are A's own rows (those rows have no L entries), so the colour-0 rows of
A M^-1 z are z itself and each product after a preconditioner apply runs
over the colour-1 rows only.  The identity is checked on the CPU against
the oracle's ILU(0) (test_fold_identity_oracle); on the GPU a folded solve
must converge to the same FP64-verified bar as the unfolded one, in mixed
precision (the default there) and in FP64 (where the default stays
unfolded, scipy's arithmetic).  Mixed results are FP64-verified on the
host; iteration counts of the two forms are printed.
"""

import numpy as np
import pytest
from scipy import sparse

from helpers import oracle_system, rel_l2
from oracle import solve as osolve


def _kc():
    from kcfd_amd import device, workloads
    return device, workloads


def test_fold_identity_oracle():
    """CPU: for the colour-ordered ILU(0) M = LU of a 7-point pressure
    matrix, (A M^-1 z)_i = z_i on the colour-0 rows, and the colour-1 rows
    are not the identity (M is not A)."""
    _, workloads = _kc()
    A, b, _ = oracle_system(workloads.hex_pressure(10, seed=1002))
    A = sparse.csr_matrix(A)
    perm, col, nc = osolve.color_order(A)
    assert nc == 2
    Ap = osolve.permuted(A, np.asarray(perm, np.int64))
    n0 = int(np.sum(col == 0))
    F = osolve.ILU0(Ap)
    z = np.random.default_rng(5).uniform(-1, 1, Ap.shape[0])
    v = Ap @ F.solve(z)
    assert np.max(np.abs(v[:n0] - z[:n0])) < 1e-12 * np.max(np.abs(z))
    assert np.max(np.abs(v[n0:] - z[n0:])) > 1e-3


def _case(which):
    _, workloads = _kc()
    if which == "hex24":
        return oracle_system(workloads.hex_pressure(24, seed=1002))[:2]
    if which == "quad96":
        return oracle_system(workloads.pressure_case(workloads.quad_mesh(96), seed=11))[:2]
    A, b, _ = oracle_system(workloads.hex_momentum(16, seed=1002))
    return A, b


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["hex24", "quad96", "hexmom16"])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_fold_converges(ctx, which, precision):
    device, _ = _kc()
    A, b = _case(which)
    out = {}
    for fold in (-1, 1):
        M = device.DeviceMatrix.from_csr(ctx, A)
        M.setup_precond("ilu0", precision, order="color")
        x, res = M.solve(b, rtol=1e-8, precond="ilu0", precision=precision, fold_colour0=fold)
        out[fold] = (x, res)
        M.close()
    (xu, ru), (xf, rf) = out[-1], out[1]
    print(f"{which} {precision}: unfolded {ru.iters} it / {ru.refinements} ref, "
          f"folded {rf.iters} it / {rf.refinements} ref; x apart {rel_l2(xf, xu):.2e}; "
          f"bytes {ru.bytes:.3e} / {rf.bytes:.3e}")
    for x, r in ((xu, ru), (xf, rf)):
        assert r.info == 0
        assert osolve.true_rel_residual(A, x, b) < 1e-8
    assert rf.iters <= 1.3 * ru.iters + 10
    # both are solutions to rtol 1e-8 of the same system
    assert rel_l2(xf, xu) < 1e-5


@pytest.mark.gpu
def test_fold_default_fp64_unfolded(ctx):
    """FP64 keeps scipy's arithmetic by default: fold_colour0 = 0 equals -1
    bit for bit; mixed folds by default: 0 equals 1 bit for bit."""
    device, _ = _kc()
    A, b = _case("hex24")
    for precision, same_as in (("fp64", -1), ("mixed", 1)):
        xs = {}
        for fold in (0, same_as):
            M = device.DeviceMatrix.from_csr(ctx, A)
            M.setup_precond("ilu0", precision, order="color")
            xs[fold] = M.solve(b, rtol=1e-8, precond="ilu0", precision=precision, fold_colour0=fold)
            M.close()
        (x0, r0), (x1, r1) = xs[0], xs[same_as]
        assert np.array_equal(x0, x1) and r0.iters == r1.iters and r0.refinements == r1.refinements


@pytest.mark.gpu
def test_fold_batch_equals_alone(ctx):
    """Folded members (two-colour hex and quad) beside an unfolded one (the
    three-colour hybrid mesh, solved by the sync-free triangular solves) in
    one mixed colour-ILU(0) batch: each result -- x, iterations, info --
    equals the same system solved alone, bit for bit.  Neither side
    continues a stalled mixed solve in FP64 (the small hybrid system stalls
    in FP32 and stops at its best iterate), so the mixed results themselves
    are compared; the converged ones are FP64-verified."""
    device, workloads = _kc()
    from kcfd_amd import batch as kbatch
    systems = [_case("hex24"), _case("quad96"),
               oracle_system(workloads.hybrid_pressure(12, 4, 4, seed=1007))[:2]]
    alone = []
    for A, b in systems:
        M = device.DeviceMatrix.from_csr(ctx, A)
        M.setup_precond("ilu0", "mixed", order="color")
        opts = device.default_opts(rtol=1e-8, precond="ilu0", precision="mixed")
        bb = device.DevBuf.from_array(ctx, b)
        xx = device.DevBuf.from_array(ctx, np.zeros(len(b)))
        r = M.solve_dev(bb.p, xx.p, opts, setup=False, fallback=False)
        alone.append((xx.download(), r))
        bb.free()
        xx.free()
        M.close()
    mats = [device.DeviceMatrix.from_csr(ctx, A) for A, _ in systems]
    bs = [device.DevBuf.from_array(ctx, b) for _, b in systems]
    sb = kbatch.SystemBatch(ctx, mats, bs)
    try:
        res = sb.solve(precision="mixed", precond="ilu0", order="color", rtol=1e-8, fallback=False)
        xs = sb.solutions()
    finally:
        sb.close()
    for k, ((x, r), (A, b)) in enumerate(zip(alone, systems)):
        rb = res.results[k]
        print(f"system {k}: alone {r.iters} it info {r.info}, batch {rb.iters} it info {rb.info}")
        assert np.array_equal(xs[k], x) and (rb.iters, rb.info) == (r.iters, r.info)
        if k < 2:
            assert r.info == 0 and osolve.true_rel_residual(A, xs[k], b) < 1e-8
