"""GPU parity of the BiCGSTAB path against the oracle.

The FP64 bar is bit equality.  The device's FP64 BiCGSTAB and CG follow
scipy's recurrences operation for operation; the SpMV reproduces scipy's
CSR row sums bit for bit and the ILU(0) apply the C oracle's triangular
solves.  The dot products are the one place where the device sums in
another order (per 256R-row tile, then over tiles).  So the oracle restates
scipy's solver with a pluggable dot product (oracle/solve.py:
bicgstab_restated, cg_restated; equal to scipy itself bit for bit with
numpy's dot, tests/test_oracle_restated.py), and each test here asserts
that the device's x, info, iteration count and half-step flag equal that
restatement evaluated in the device's dot-product order
(helpers.dot_tile_order) -- np.array_equal on x, not a tolerance.  scipy's
own result (numpy's dot order) is printed beside it: on long or
near-breakdown solves it differs by the amplified rounding of the other
summation order, which is not a parity defect.

Other bars: ILU(0) factors equal to the C oracle's to 1e-14 (same operation
order), mixed-precision results FP64-verified on the host.
"""

import numpy as np
import pytest

from helpers import oracle_system, rel_l2, restated_device
from oracle import golden, solve as osolve

pytestmark = pytest.mark.gpu


def _kc():
    from kcfd_amd import device, workloads
    return device, workloads


def _gpu_solve(ctx, A, b, **kw):
    device, _ = _kc()
    M = device.DeviceMatrix.from_csr(ctx, A)
    x, res = M.solve(b, **kw)
    return x, res


def _check_restated(x, res, A, b, label, solver="bicgstab", rows=2, perm=None, **kw):
    """Device FP64 result == the restated solver in the device's dot order,
    bit for bit (x, info, iterations, half-step flag)."""
    out = restated_device(A, b, solver=solver, rows_per_thread=rows, perm=perm, **kw)
    xr, info, it = out[:3]
    run = osolve.solve_cg if solver == "cg" else osolve.solve_bicgstab
    xs, info_s, it_s = run(A, b, **kw)
    print(f"{label}: gpu iters={res.iters} info={res.info} resid={res.rel_resid:.2e} | restated "
          f"(device order) {it} info {info} | scipy {it_s} info {info_s}, relL2 to scipy "
          f"{rel_l2(x, xs):.2e}")
    assert (res.info, res.iters) == (info, it)
    if solver == "bicgstab":
        assert res.half_step == out[3]
    assert np.array_equal(x, xr)
    return it


@pytest.mark.parametrize("pc", ["none", "jacobi"])
@pytest.mark.parametrize("tol", [(0.0, 1e-6), (1e-10, 0.0)])
def test_case16_golden(ctx, pc, tol):
    """The reference's figure (figures/case16_preconditioner.png): none 105,
    jacobi 95 at atol 1e-10, rtol 0 (verification/case16_preconditioner.py:
    170-173).  scipy reproduces them to one iteration (test_oracle_golden);
    the endgame is rounding-dominated (r~ = r0 lives on the inlet column and
    rho decays to ~1e-30; the Jacobi run ends in a rho breakdown), so the
    device is held to the restated solver in its own summation order."""
    atol, rtol = tol
    A, b = golden.case16_system()
    x, res = _gpu_solve(ctx, A, b, rtol=rtol, atol=atol, maxiter=2000, precond=pc)
    it = _check_restated(x, res, A, b, f"case16 {pc} atol {atol} rtol {rtol}", rtol=rtol, atol=atol,
                         maxiter=2000, precond=pc)
    if atol > 0:
        print(f"case16 {pc}: figure {golden.CASE16_ITERS[pc]}, device {it}")


@pytest.mark.parametrize("n", [32, 64, 128])
def test_quad_convdiff_fp64_jacobi(ctx, n):
    """Config 1 (quad 128^2, Jacobi, rtol 1e-8) and two smaller meshes."""
    device, workloads = _kc()
    case = workloads.quad_convdiff(n, seed=1001)
    A, b, _ = oracle_system(case)
    x, res = _gpu_solve(ctx, A, b, rtol=1e-8, precond="jacobi")
    assert res.info == 0 and res.rel_resid < 1e-8
    _check_restated(x, res, A, b, f"quad{n}", rtol=1e-8, precond="jacobi")


@pytest.mark.parametrize("n", [8, 16, 24])
def test_hex_pressure_ilu0(ctx, n):
    device, workloads = _kc()
    case = workloads.hex_pressure(n, seed=1002)
    A, b, _ = oracle_system(case)
    M = device.DeviceMatrix.from_csr(ctx, A)
    M.setup_precond("ilu0", "fp64")
    lu, nl, nu = M.ilu0_factors()
    ref = osolve.ILU0(A)
    err = np.max(np.abs(lu - ref.lu) / np.maximum(np.abs(ref.lu), 1e-300))
    print(f"hex{n} ilu0 factors: max rel diff {err:.2e}; levels L={nl} U={nu}")
    assert err <= 1e-14
    assert nl == 3 * n - 2 and nu == 3 * n - 2
    x, res = M.solve(b, rtol=1e-8, precond="ilu0")
    assert res.info == 0 and res.rel_resid < 1e-8
    _check_restated(x, res, A, b, f"hex{n} ilu0", rtol=1e-8, precond="ilu0_struct")


@pytest.mark.parametrize("which", ["hex12", "quad24", "tet6"])
def test_ilu0_color_order(ctx, which):
    """ILU(0) of P A P^T with the first-fit colour ordering: the same
    ordering as the oracle's, factors equal to the C oracle's on P A P^T,
    one solve level per colour, M^{-1} z equal to the oracle's, and the
    BiCGSTAB solve equal to the restated solver."""
    device, workloads = _kc()
    from kcfd_amd import _lib as L
    from kcfd_amd._lib import DevBuf
    case = {"hex12": lambda: workloads.hex_pressure(12, seed=1002),
            "quad24": lambda: workloads.quad_convdiff(24, seed=1001),
            "tet6": lambda: workloads.tet_twofluid(6, seed=1003)}[which]()
    A, b, _ = oracle_system(case)
    M = device.DeviceMatrix.from_csr(ctx, A)
    M.setup_precond("ilu0", "fp64", order="color")
    perm, nc = M.ilu0_order()
    operm, _, onc = osolve.color_order(A)
    assert nc == onc and np.array_equal(perm, operm)
    lu, nl, nu = M.ilu0_factors()
    ref = osolve.ILU0Ordered(A, operm)
    err = np.max(np.abs(lu - ref.f.lu) / np.maximum(np.abs(ref.f.lu), 1e-300))
    print(f"{which} colour ILU(0): {nc} colours, levels L={nl} U={nu}, factor rel diff {err:.2e}")
    assert err <= 1e-14
    assert nl == nc and 1 <= nu <= nc
    z = np.random.default_rng(7).uniform(-1, 1, A.shape[0])
    zd, yd = DevBuf.from_array(ctx, z), DevBuf(ctx, A.shape[0])
    L.check(L.lib().kcfd_precond_apply(M.h, zd.p, yd.p))
    y = yd.download()
    yo = ref.solve(z)
    assert rel_l2(y, yo) <= 1e-13
    x, res = M.solve(b, rtol=1e-8, precond="ilu0")
    assert res.info == 0 and res.rel_resid < 1e-8
    _check_restated(x, res, A, b, f"{which} ilu0 colour", perm=perm, rtol=1e-8, precond="ilu0_color")
    # back to the natural ordering on the same matrix
    M.setup_precond("ilu0", "fp64", order="natural")
    assert M.ilu0_order()[1] == 0


@pytest.mark.parametrize("which", ["hex10", "quad40", "tet5"])
def test_ilu0_segment_order(ctx, which):
    """ILU(0) with the segment ordering (KCFD_ORDER_SEGMENT): the same
    ordering as the oracle's restatement, factors equal to the C oracle's on
    P A P^T, the FP64 BiCGSTAB solve equal to the restated solver bit for
    bit, and the mixed-precision solve converged (FP64 residual on the
    host)."""
    device, workloads = _kc()
    case = {"hex10": lambda: workloads.hex_pressure(10, seed=1002),
            "quad40": lambda: workloads.quad_convdiff(40, seed=1001),
            "tet5": lambda: workloads.tet_twofluid(5, seed=1003)}[which]()
    A, b, _ = oracle_system(case)
    M = device.DeviceMatrix.from_csr(ctx, A)
    M.setup_precond("ilu0", "fp64", order="segment")
    perm, nc = M.ilu0_order()
    operm, _, onc = osolve.segment_order(A, seg_len=4)
    assert nc == onc and np.array_equal(perm, operm)
    lu, nl, nu = M.ilu0_factors()
    ref = osolve.ILU0Ordered(A, operm)
    err = np.max(np.abs(lu - ref.f.lu) / np.maximum(np.abs(ref.f.lu), 1e-300))
    print(f"{which} segment ILU(0): {nc} colours, levels L={nl} U={nu}, factor rel diff {err:.2e}")
    assert err <= 1e-14
    x, res = M.solve(b, rtol=1e-8, precond="ilu0")
    assert res.info == 0 and res.rel_resid < 1e-8
    _check_restated(x, res, A, b, f"{which} ilu0 segment", perm=perm, rtol=1e-8, precond="ilu0_segment")
    M.setup_precond("ilu0", "mixed", order="segment")
    xm, rm = M.solve(b, rtol=1e-8, precond="ilu0", precision="mixed")
    r = np.linalg.norm(b - A @ xm) / np.linalg.norm(b)
    print(f"{which} mixed segment: iters {rm.iters} info {rm.info} host resid {r:.2e}")
    assert rm.info == 0 and r < 1e-8


def test_mixed_color_ilu0(ctx):
    """Mixed precision with the colour-ordered ILU(0) (the fast ILU path)."""
    device, workloads = _kc()
    A, b, _ = oracle_system(workloads.hex_pressure(32, seed=1002))
    xo, _, _ = osolve.solve_bicgstab(A, b, rtol=1e-12, precond="ilu0_color")
    M = device.DeviceMatrix.from_csr(ctx, A)
    M.setup_precond("ilu0", "mixed", order="color")
    x, res = M.solve(b, rtol=1e-8, precond="ilu0", precision="mixed")
    true_res = osolve.true_rel_residual(A, x, b)
    print(f"mixed colour ilu0 hex32: iters={res.iters} refinements={res.refinements} "
          f"resid(gpu)={res.rel_resid:.2e} resid(cpu)={true_res:.2e} err vs fp64={rel_l2(x, xo):.2e}")
    assert res.info == 0 and true_res < 1e-8
    assert abs(true_res - res.rel_resid) <= 1e-3 * true_res + 1e-15


@pytest.mark.parametrize("n", [8, 16, 24])
def test_hex_pressure_jacobi(ctx, n):
    device, workloads = _kc()
    case = workloads.hex_pressure(n, seed=1002)
    A, b, _ = oracle_system(case)
    x, res = _gpu_solve(ctx, A, b, rtol=1e-8, precond="jacobi")
    assert res.info == 0 and res.rel_resid < 1e-8
    _check_restated(x, res, A, b, f"hex{n} jacobi", rtol=1e-8, precond="jacobi")


@pytest.mark.parametrize("pc", ["jacobi", "ilu0"])
@pytest.mark.parametrize("size", ["small", "large"])
def test_mixed_precision_refinement(ctx, pc, size):
    """FP32 SpMV/preconditioner/vectors with FP64 reliable updates: FP64-verified
    residual below 1e-8, the GPU's reported residual equal to the host's, and
    the error against the FP64 oracle solution reported."""
    device, workloads = _kc()
    if size == "small":
        case = workloads.hex_pressure(16, seed=1002) if pc == "ilu0" else workloads.quad_convdiff(64)
    else:
        case = workloads.hex_pressure(40, seed=1002)
    A, b, _ = oracle_system(case)
    xo, info_o, it_o = osolve.solve_bicgstab(A, b, rtol=1e-12,
                                             precond="ilu0_struct" if pc == "ilu0" else "jacobi")
    x, res = _gpu_solve(ctx, A, b, rtol=1e-8, precond=pc, precision="mixed")
    err = rel_l2(x, xo)
    true_res = osolve.true_rel_residual(A, x, b)
    print(f"mixed {pc}: iters={res.iters} refinements={res.refinements} info={res.info} "
          f"resid(gpu)={res.rel_resid:.2e} resid(cpu)={true_res:.2e} err vs fp64={err:.2e}")
    assert res.info == 0
    assert true_res < 1e-8 and res.rel_resid < 1e-8
    assert abs(true_res - res.rel_resid) <= 1e-3 * true_res + 1e-15


def test_batch_equals_alone(ctx):
    """Each system of a batch gives bit-identical results alone and in the mix.
    The last system's SELL slices span more than 16-bit column offsets reach,
    so in mixed precision the batch holds both FP32 layouts (narrow-index and
    packed) at once."""
    device, workloads = _kc()
    from kcfd_amd._lib import DevBuf
    cases = [workloads.quad_convdiff(20, seed=1), workloads.hex_pressure(10, seed=2),
             workloads.quad_convdiff(33, seed=3), workloads.hex_pressure(7, seed=4)]
    systems = [oracle_system(c)[:2] for c in cases] + [_wide_span_system()]
    for prec in ["fp64", "mixed"]:
        alone = []
        for A, b in systems:
            x, res = _gpu_solve(ctx, A, b, rtol=1e-8, precond="jacobi", precision=prec)
            alone.append((x, res.iters))
        mats = [device.DeviceMatrix.from_csr(ctx, A) for A, _ in systems]
        for m in mats:
            m.setup_precond("jacobi", prec)
        bs = [DevBuf.from_array(ctx, b) for _, b in systems]
        xs = [DevBuf.from_array(ctx, np.zeros(len(b))) for _, b in systems]
        bt = device.Batch(ctx, mats, bs, xs, device.default_opts(rtol=1e-8, precond="jacobi",
                                                                 precision=prec))
        res = bt.run()
        for k, (A, b) in enumerate(systems):
            xb = xs[k].download()
            print(f"batch {prec} sys{k}: iters alone={alone[k][1]} batch={res[k].iters}")
            assert res[k].iters == alone[k][1]
            assert np.array_equal(xb, alone[k][0])
        bt.close()


def test_zero_rhs_and_maxiter(ctx):
    device, workloads = _kc()
    case = workloads.quad_convdiff(16)
    A, b, _ = oracle_system(case)
    x, res = _gpu_solve(ctx, A, np.zeros(A.shape[0]), rtol=1e-8, precond="jacobi")
    assert res.info == 0 and res.iters == 0 and np.all(x == 0)
    x, res = _gpu_solve(ctx, A, b, rtol=1e-14, precond="jacobi", maxiter=3)
    xo, info_o, it_o = osolve.solve_bicgstab(A, b, rtol=1e-14, precond="jacobi", maxiter=3)
    assert res.info == info_o == 3 and res.iters == 3
    _check_restated(x, res, A, b, "quad16 maxiter 3", rtol=1e-14, precond="jacobi", maxiter=3)
    x0 = np.random.default_rng(5).uniform(-1, 1, len(b))
    x, res = _gpu_solve(ctx, A, b, x0=x0, rtol=1e-8, precond="jacobi")
    _check_restated(x, res, A, b, "quad16 x0", x0=x0, rtol=1e-8, precond="jacobi")


@pytest.mark.parametrize("pc,order", [("jacobi", None), ("ilu0", "natural"), ("ilu0", "color")])
def test_hybrid_mesh_solve(ctx, pc, order):
    """Mixed hex/tet duct (mesh/hybrid_mesh_generator.py): the device-assembled
    pressure system solved exactly as the restated solver in the device's
    order, natural and colour-ordered ILU(0) factors equal to the C
    oracle's.  The three-colour ILU(0) run is chaotic (over 40 one-ulp
    perturbations of b scipy itself takes 98 to 154 iterations), which bit
    equality with the restatement is not."""
    device, workloads = _kc()
    case = workloads.hybrid_pressure(12, 4, 4, seed=1007)
    A, b, _ = oracle_system(case)
    dm = device.DeviceMesh(case.mesh, ctx)
    Ad = workloads.assemble_device(dm, case)
    assert np.array_equal(Ad.to_scipy().indptr, A.indptr)
    M = device.DeviceMatrix.from_csr(ctx, A)
    M.setup_precond(pc, "fp64", order=order)
    opc = pc
    if pc == "ilu0":
        lu, nl, nu = M.ilu0_factors()
        if order == "color":
            perm, nc = M.ilu0_order()
            ref = osolve.ILU0Ordered(A, osolve.color_order(A)[0]).f
            assert np.array_equal(perm, osolve.color_order(A)[0])
            opc = "ilu0_color"
        else:
            ref = osolve.ILU0(A)
            opc = "ilu0_struct"
        assert np.max(np.abs(lu - ref.lu) / np.maximum(np.abs(ref.lu), 1e-300)) <= 1e-14
    x, res = M.solve(b, rtol=1e-8, precond=pc)
    assert res.info == 0 and res.rel_resid < 1e-8
    perm = M.ilu0_order()[0] if order == "color" else None
    _check_restated(x, res, A, b, f"hybrid {pc} {order}", perm=perm, rtol=1e-8, precond=opc)


def test_hybrid_full_size(ctx):
    """The mixed hex/tet duct at 64 x 32 x 32 (425,984 cells: 32,768 hexes,
    393,216 tets): device assembly equal to the oracle's to 1e-14, and the
    mixed-precision colour-ILU(0) solve FP64-verified below 1e-8."""
    device, workloads = _kc()
    case = workloads.hybrid_pressure(64, seed=1007)
    assert case.mesh.n_cells == 425_984
    A, b, _ = oracle_system(case)
    dm = device.DeviceMesh(case.mesh, ctx)
    Ad = workloads.assemble_device(dm, case)
    Ag = Ad.to_scipy()
    assert np.array_equal(Ag.indptr, A.indptr) and np.array_equal(Ag.indices, A.indices)
    ev = np.max(np.abs(Ag.data - A.data) / np.maximum(np.abs(A.data), 1e-300))
    M = device.DeviceMatrix.from_csr(ctx, A)
    M.setup_precond("ilu0", "mixed", order="color")
    x, res = M.solve(b, rtol=1e-8, precond="ilu0", precision="mixed")
    tr = osolve.true_rel_residual(A, x, b)
    print(f"hybrid 64x32x32: nnz={A.nnz} assembly max rel {ev:.2e}; mixed colour ILU(0) "
          f"{res.iters} it, {res.refinements} replacements, resid gpu {res.rel_resid:.2e} cpu {tr:.2e}, "
          f"{res.ms:.1f} ms")
    assert ev <= 1e-14
    assert res.info == 0 and tr < 1e-8


@pytest.mark.parametrize("rows", [1, 2, 4])
def test_rows_per_thread(ctx, rows):
    """Each tile-kernel shape (1, 2 or 4 rows per thread; tiles of 256, 512
    or 1024 rows): FP64 BiCGSTAB and CG equal to the restated solvers, the mixed
    solve FP64-verified, and a batch equal bit for bit to its systems solved
    alone with the same shape."""
    device, workloads = _kc()
    from kcfd_amd._lib import DevBuf
    A, b, _ = oracle_system(workloads.hex_pressure(14, seed=1002))
    M = device.DeviceMatrix.from_csr(ctx, A)
    M.setup_precond("ilu0", "fp64", order="color")
    perm = M.ilu0_order()[0]
    x, res = M.solve(b, rtol=1e-8, precond="ilu0", rows_per_thread=rows)
    assert res.info == 0
    _check_restated(x, res, A, b, f"rows {rows} bicgstab", rows=rows, perm=perm, rtol=1e-8,
                    precond="ilu0_color")
    x, res = M.solve(b, rtol=1e-8, precond="jacobi", rows_per_thread=rows)
    _check_restated(x, res, A, b, f"rows {rows} bicgstab jacobi", rows=rows, rtol=1e-8, precond="jacobi")
    M.setup_precond("ilu0", "fp64", order="color")
    x, res = M.solve(b, rtol=1e-8, precond="ilu0", method="cg", rows_per_thread=rows)
    assert res.info == 0
    _check_restated(x, res, A, b, f"rows {rows} cg", solver="cg", rows=rows, perm=perm, rtol=1e-8,
                    precond="ilu0_color")
    M.setup_precond("ilu0", "mixed", order="color")
    x, res = M.solve(b, rtol=1e-8, precond="ilu0", precision="mixed", rows_per_thread=rows)
    assert res.info == 0 and osolve.true_rel_residual(A, x, b) < 1e-8
    cases = [workloads.quad_convdiff(20, seed=1), workloads.hex_pressure(10, seed=2),
             workloads.quad_convdiff(33, seed=3)]
    systems = [oracle_system(c)[:2] for c in cases]
    alone = []
    for As, bs in systems:
        xs, rs = _gpu_solve(ctx, As, bs, rtol=1e-8, precond="jacobi", precision="mixed", rows_per_thread=rows)
        alone.append((xs, rs.iters))
    mats = [device.DeviceMatrix.from_csr(ctx, As) for As, _ in systems]
    for m in mats:
        m.setup_precond("jacobi", "mixed")
    bd = [DevBuf.from_array(ctx, bs) for _, bs in systems]
    xd = [DevBuf.from_array(ctx, np.zeros(len(bs))) for _, bs in systems]
    bt = device.Batch(ctx, mats, bd, xd, device.default_opts(rtol=1e-8, precond="jacobi", precision="mixed",
                                                             rows_per_thread=rows))
    out = bt.run()
    for k in range(len(systems)):
        assert out[k].iters == alone[k][1] and np.array_equal(xd[k].download(), alone[k][0])
    bt.close()


def test_degenerate_sizes(ctx):
    """n = 0: every method returns at once with info 0 (a batch rejects the
    empty system with a message); n = 1: every method solves a x = b with
    scipy's iteration count, alone and inside a batch next to larger
    systems."""
    device, workloads = _kc()
    from scipy import sparse
    from kcfd_amd._lib import DevBuf, KcfdError
    E = sparse.csr_matrix((0, 0))
    M0 = device.DeviceMatrix.from_csr(ctx, E)
    for pc in ("jacobi", "ilu0"):
        for method in ("bicgstab", "cg", "gmres"):
            x, res = M0.solve(np.zeros(0), rtol=1e-8, precond=pc, method=method)
            assert x.shape == (0,) and res.info == 0 and res.iters == 0
    one = sparse.csr_matrix(np.array([[4.0]]))
    M1 = device.DeviceMatrix.from_csr(ctx, one)
    for pc in ("none", "jacobi", "ilu0"):
        for method in ("bicgstab", "cg", "gmres"):
            x, res = M1.solve(np.array([2.0]), rtol=1e-12, precond=pc, method=method)
            print(f"n=1 {pc} {method}: x={x[0]!r} iters={res.iters} info={res.info}")
            assert res.info == 0 and abs(x[0] - 0.5) <= 1e-15
            # scipy's callback counts on 4 x = 2: bicgstab 0 (half-step exit), cg 1, gmres 1
            assert res.iters == {"bicgstab": 0, "cg": 1, "gmres": 1}[method]
        x, res = M1.solve(np.array([2.0]), rtol=1e-12, precond=pc, precision="mixed")
        assert res.info == 0 and abs(x[0] - 0.5) <= 1e-15
    A, b, _ = oracle_system(workloads.quad_convdiff(20, seed=1))
    mats = [device.DeviceMatrix.from_csr(ctx, one), device.DeviceMatrix.from_csr(ctx, A)]
    for m in mats:
        m.setup_precond("jacobi", "fp64")
    bs = [DevBuf.from_array(ctx, np.array([2.0])), DevBuf.from_array(ctx, b)]
    xs = [DevBuf.from_array(ctx, np.zeros(1)), DevBuf.from_array(ctx, np.zeros(len(b)))]
    bt = device.Batch(ctx, mats, bs, xs, device.default_opts(rtol=1e-8, precond="jacobi"))
    out = bt.run()
    bt.close()
    xa, ra = _gpu_solve(ctx, A, b, rtol=1e-8, precond="jacobi")
    assert out[0].info == 0 and abs(xs[0].download()[0] - 0.5) <= 1e-15
    assert out[1].iters == ra.iters and np.array_equal(xs[1].download(), xa)
    M0.setup_precond("jacobi", "fp64")
    with pytest.raises(KcfdError, match="empty"):
        device.Batch(ctx, [M0, mats[1]], [DevBuf(ctx, 1), bs[1]], [DevBuf(ctx, 1), xs[1]],
                     device.default_opts(rtol=1e-8, precond="jacobi"))


def test_ilu0_bad_pivots_raise(ctx):
    """A zero pivot (row 0 of [[0, 1], [1, 1]]) and a row without a diagonal
    entry make the device ILU(0) set-up raise a RuntimeError naming the row,
    as scipy's spilu raises RuntimeError for a singular factor; Jacobi keeps
    the reference's rule (1/diag where |diag| > 1e-30, else 0) and ends as
    scipy does (the same breakdown, iteration and x)."""
    device, _ = _kc()
    from scipy import sparse
    Z = sparse.csr_matrix(np.array([[0.0, 1.0], [1.0, 1.0]]))
    Z.eliminate_zeros()
    Zd = sparse.csr_matrix((np.array([0.0, 1.0, 1.0, 1.0]), np.array([0, 1, 0, 1]), np.array([0, 2, 4])),
                           shape=(2, 2))                     # explicit zero on the diagonal
    for A, msg in ((Zd, "pivot"), (Z, "no diagonal")):
        M = device.DeviceMatrix.from_csr(ctx, A)
        with pytest.raises(RuntimeError, match=msg):
            M.setup_precond("ilu0", "fp64")
        M.close()
    M = device.DeviceMatrix.from_csr(ctx, Zd)
    x, res = M.solve(np.array([1.0, 2.0]), rtol=1e-12, precond="jacobi", maxiter=50)
    xo, info_o, it_o = osolve.solve_bicgstab(Zd, np.array([1.0, 2.0]), rtol=1e-12, precond="jacobi",
                                             maxiter=50)
    print(f"jacobi with a zero diagonal: x={x}, info={res.info}, iters={res.iters}; scipy {xo}, {info_o}, {it_o}")
    # scipy: omega breakdown (-11) after 2 iterations at x = (0, 1.5)
    assert (res.info, res.iters) == (info_o, it_o) and np.array_equal(x, xo)


def _wide_span_system(n=70_000):
    """Tridiagonal diffusion plus a coupling of rows 0 and n-1: slice 0's
    off-diagonal columns span n - 1 > 65534."""
    from scipy import sparse
    main = np.full(n, 2.5)
    off = np.full(n - 1, -1.0)
    A = sparse.diags([off, main, off], [-1, 0, 1], format="lil")
    A[0, n - 1] = -0.25
    A[n - 1, 0] = -0.25
    A = A.tocsr()
    A.sort_indices()
    return A, np.random.default_rng(77).uniform(-1, 1, n)


@pytest.mark.parametrize("which", ["hex32_color", "tet_jacobi", "quad_natural", "hybrid_color", "wide"])
def test_mixed_narrow_index_sell(ctx, which, monkeypatch):
    """The mixed SELL copy with 16-bit column offsets from a per-slice base
    (6 bytes an entry) holds the same columns and values as the packed
    {column, value} form (KCFD_SELL16=0), so a mixed solve is bit-identical
    either way.  The narrow form is taken where every 64-row slice's
    off-diagonal columns span at most 65534, which the device matrix's
    footprint shows (2 bytes less per entry, 4 more per slice); a matrix with
    a wider slice keeps the packed form."""
    device, workloads = _kc()
    pc, order = "ilu0", "color"
    if which == "hex32_color":
        A, b, _ = oracle_system(workloads.hex_pressure(32, seed=1002))
    elif which == "tet_jacobi":
        A, b, _ = oracle_system(workloads.tet_twofluid(3000, seed=1003))
        pc, order = "jacobi", None
    elif which == "quad_natural":
        A, b, _ = oracle_system(workloads.quad_convdiff(64, seed=1001))
        order = "natural"
    elif which == "hybrid_color":
        A, b, _ = oracle_system(workloads.hybrid_pressure(12, 4, 4, seed=1007))
    else:
        A, b = _wide_span_system()
        pc, order = "jacobi", None
    out = {}
    for mode in ("packed", "narrow"):
        if mode == "packed":
            monkeypatch.setenv("KCFD_SELL16", "0")
        else:
            monkeypatch.delenv("KCFD_SELL16", raising=False)
        M = device.DeviceMatrix.from_csr(ctx, A)
        M.setup_precond(pc, "mixed", **({"order": order} if order else {}))
        x, res = M.solve(b, rtol=1e-8, precond=pc, precision="mixed")
        out[mode] = (x, res.iters, res.refinements, res.info, M.bytes())
        M.close()
    (xp, ip_, rp, fp, bp), (xn, in_, rn, fn, bn) = out["packed"], out["narrow"]
    print(f"{which}: n={A.shape[0]} iters {ip_}/{in_} refinements {rp}/{rn} info {fp}/{fn} "
          f"bytes packed {bp} narrow {bn}")
    assert np.array_equal(xp, xn) and (ip_, rp, fp) == (in_, rn, fn)
    assert fp == 0 and osolve.true_rel_residual(A, xn, b) < 1e-8
    if which == "wide":
        assert bn == bp
    else:
        assert bn < bp


def test_stale_values_refused(ctx):
    """The solvers read the SELL copy and the preconditioner that
    kcfd_precond_setup built.  A solve or a batch over a matrix whose values
    changed after that set-up (or that was never set up) is refused with
    KCFD_ESTATE and a message, and works again after a new set-up -- with
    the new values (checked against the restated solver)."""
    device, workloads = _kc()
    import ctypes
    from kcfd_amd import _lib as L
    from kcfd_amd._lib import DevBuf
    A, b, _ = oracle_system(workloads.quad_convdiff(16, seed=1001))
    M = device.DeviceMatrix.from_csr(ctx, A)
    bd, xd = DevBuf.from_array(ctx, b), DevBuf.from_array(ctx, np.zeros(A.shape[0]))
    opts = device.default_opts(rtol=1e-8, precond="jacobi")
    res = L.SolveResult()

    def solve():
        return L.lib().kcfd_solve(M.h, bd.p, xd.p, ctypes.byref(opts), ctypes.byref(res))

    assert solve() == L.KCFD_ESTATE                    # never set up
    assert "kcfd_precond_setup" in L.lib().kcfd_last_error().decode()
    L.check(L.lib().kcfd_precond_setup(M.h, L.PC_JACOBI, L.FP64))
    assert solve() == L.KCFD_OK
    A2 = A.copy()
    A2.data = A2.data * 2.0
    v = np.ascontiguousarray(A2.data)
    L.check(L.lib().kcfd_mat_set_values(M.h, v.ctypes.data))
    L.check(L.lib().kcfd_buf_upload(ctx.h, xd.p, np.zeros(A.shape[0]).ctypes.data, 8 * A.shape[0]))
    assert solve() == L.KCFD_ESTATE                    # values changed since the set-up
    with pytest.raises(L.KcfdError, match="values changed"):
        device.Batch(ctx, [M], [bd], [xd], opts)
    L.check(L.lib().kcfd_precond_setup(M.h, L.PC_JACOBI, L.FP64))
    assert solve() == L.KCFD_OK
    _check_restated(xd.download(), res, A2, b, "quad16 after new values", rtol=1e-8, precond="jacobi")
    # device assembly counts as a change too
    dm = device.DeviceMesh(workloads.quad_convdiff(16, seed=1001).mesh, ctx)
    Md = dm.new_system()
    L.check(L.lib().kcfd_precond_setup(Md.h, L.PC_JACOBI, L.FP64))
    L.check(L.lib().kcfd_asm_reset(Md.h))
    r2 = L.SolveResult()
    assert L.lib().kcfd_solve(Md.h, bd.p, xd.p, ctypes.byref(opts), ctypes.byref(r2)) == L.KCFD_ESTATE


def test_mixed_needs_diagonal(ctx):
    """The FP32 copy keeps each row's sum at its diagonal entry (difference
    form), so a mixed set-up of a matrix with a row that has no structural
    diagonal is refused with a message; FP64 on the same matrix works."""
    device, _ = _kc()
    from scipy import sparse
    Z = sparse.csr_matrix(np.array([[0.0, 1.0], [1.0, 1.0]]))
    Z.eliminate_zeros()
    M = device.DeviceMatrix.from_csr(ctx, Z)
    with pytest.raises(RuntimeError, match="diagonal entry in every row"):
        M.setup_precond("jacobi", "mixed")
    x, res = M.solve(np.array([1.0, 2.0]), rtol=1e-12, precond="none", maxiter=50)
    xr, info, it, half = restated_device(Z, np.array([1.0, 2.0]), rtol=1e-12, maxiter=50, precond="none")
    assert (res.info, res.iters) == (info, it) and np.array_equal(x, xr)


def test_trsv_wait_bound(ctx, monkeypatch):
    """The sync-free triangular solve (natural-order ILU(0)) bounds its wait
    for an operand.  With the bound set to zero polls from the host
    (KCFD_TRSV_POLL_LIMIT, read when the solver workspace is made) every
    wait gives up at once: the kernels still run to their end, and the solve
    returns KCFD_ESTATE with a message instead of a result.  A new workspace
    with the default bound solves the same system exactly as before."""
    device, workloads = _kc()
    A, b, _ = oracle_system(workloads.hex_pressure(16, seed=1002))
    monkeypatch.setenv("KCFD_TRSV_POLL_LIMIT", "0")
    M = device.DeviceMatrix.from_csr(ctx, A)
    M.setup_precond("ilu0", "fp64", order="natural")
    with pytest.raises(RuntimeError, match="exceeded 0 polls"):
        M.solve(b, rtol=1e-8, precond="ilu0")
    M.close()
    monkeypatch.delenv("KCFD_TRSV_POLL_LIMIT")
    M = device.DeviceMatrix.from_csr(ctx, A)
    M.setup_precond("ilu0", "fp64", order="natural")
    x, res = M.solve(b, rtol=1e-8, precond="ilu0")
    _check_restated(x, res, A, b, "hex16 ilu0 after the bounded-wait test", rtol=1e-8, precond="ilu0_struct")
