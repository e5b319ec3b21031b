"""Test helpers: replay a workload recipe through the CPU oracle."""

from __future__ import annotations

import numpy as np

from oracle import fvm as ofvm


def oracle_system(case):
    """Assemble ``case`` with the oracle; returns (A csr, rhs, diag tracker)."""
    return ofvm.system_of_case(case)


def pressure_correction(mesh, dP, rho, dirichlet, s):
    """The oracle's pressure-correction assembly (oracle/fvm.py)."""
    ofvm.pressure_correction_operator(mesh, dP, rho, dirichlet, s)


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb > 0 else 1.0))


def oracle_band(A, b, nrep=6, solver="bicgstab", **kw):
    """scipy BiCGSTAB (or CG, GMRES) on b and on nrep-1 copies of b perturbed
    by one ulp per entry: the reference's own rounding-noise floor.  Returns
    (iteration counts, infos, x of the unperturbed run, max relative L2
    spread of the perturbed solutions from it)."""
    from oracle import solve as osolve
    run = {"cg": osolve.solve_cg, "gmres": osolve.solve_gmres}.get(solver, osolve.solve_bicgstab)
    its, infos, xs = [], [], []
    for k in range(nrep):
        bb = b if k == 0 else b * (1 + np.random.default_rng(k).choice([-1, 1], len(b)) * 2.0 ** -52)
        x, info, it = run(A, bb, **kw)
        its.append(it)
        infos.append(info)
        xs.append(x)
    spread = max(rel_l2(x, xs[0]) for x in xs[1:]) if nrep > 1 else 0.0
    return its, infos, xs[0], spread


def _dot_fsum(a, b):
    import math
    return math.fsum((a * b).tolist())


def _dot_reversed(a, b):
    return float(np.dot(a[::-1], b[::-1]))


def _dot_chunked(a, b, chunk=256):
    p = a * b
    n = len(p)
    pad = (-n) % chunk
    if pad:
        p = np.concatenate([p, np.zeros(pad)])
    return float(np.sum(np.sum(p.reshape(-1, chunk), axis=1)))


DOT_ORDERS = {"fsum": _dot_fsum, "reversed": _dot_reversed, "chunked": _dot_chunked}


def _butterfly64(v):
    """wave64 xor-shuffle sum (v += shfl_xor(v, o), o = 32..1): lane 0's value.
    After the step with offset o, lanes l and l ^ o hold the same sum
    (IEEE addition commutes), so lane 0's value is the halving tree
    v[:32] + v[32:], then 16, ..., 1."""
    for o in (32, 16, 8, 4, 2, 1):
        v = v[..., :o] + v[..., o:2 * o]
    return v[..., 0]


def _block_sum256(vals):
    """256-lane workgroup sum: each wave's butterfly, then waves 0..3 in order."""
    ws = _butterfly64(vals.reshape(vals.shape[0], 4, 64))
    s = ws[:, 0]
    for q in range(1, 4):
        s = s + ws[:, q]
    return s


def dot_device_order(cus):
    """The device GMRES's dot product order: G = min(ceil(n/256), 2048,
    8 cus) workgroups of 256 lanes, each lane summing its grid-stride
    products in order, workgroup sums as _block_sum256, then the G partials
    summed the same way by one workgroup (lane j taking partials j, j+256,
    ...)."""
    def dot(a, b):
        n = len(a)
        G = max(1, min((n + 255) // 256, 2048, 8 * cus))
        p = a * b
        T = G * 256
        acc = np.zeros(T)
        for start in range(0, n, T):
            c = p[start:start + T]
            acc[:len(c)] = acc[:len(c)] + c
        part = _block_sum256(acc.reshape(G, 256))
        t = np.zeros(256)
        for start in range(0, G, 256):
            c = part[start:start + 256]
            t[:len(c)] = t[:len(c)] + c
        return float(_block_sum256(t.reshape(1, 256))[0])
    return dot


def _block_sum(vals, nt):
    """nt-lane workgroup sum (nt a multiple of 64): each wave's butterfly,
    then the wave sums added in wave order by thread 0."""
    ws = _butterfly64(vals.reshape(vals.shape[0], nt // 64, 64))
    s = ws[:, 0]
    for q in range(1, nt // 64):
        s = s + ws[:, q]
    return s


def dot_tile_order(rows_per_thread=2, perm=None, tiles_by=None):
    """The dot product order of the device BiCGSTAB and CG (one system, or
    one system of a batch -- a system's tiles start at its own row 0).
    Rows are cut into tiles of 256 R rows; thread t of a tile's 256-lane
    workgroup adds the products of rows t, t + 256, ..., t + 256 (R-1) in
    that order; the tile partial is the workgroup sum (_block_sum).  One
    1024-thread workgroup then finishes: thread j adds partials j, j + 1024,
    ... in the order of the system's tile list, and the threads are combined
    by the same fixed tree.

    With a permuted ILU(0) the solver's vectors are in P-order (``perm``:
    row of A at each position) and the tile list is sorted by the row of A
    each tile starts at (perm[256 R t]; the batch lists them so, to keep the
    tiles of both colours of a region in one XCD's range).  ``perm`` gathers
    vectors given in A's order into P-order and orders the tiles;
    ``tiles_by`` orders the tiles of vectors already in P-order."""
    R = int(rows_per_thread)
    T = 256 * R
    key_perm = perm if perm is not None else tiles_by

    def dot(a, b):
        p = np.asarray(a, np.float64) * np.asarray(b, np.float64)
        if perm is not None:
            p = p[perm]
        n = len(p)
        nt = max(1, (n + T - 1) // T)
        if n == nt * T:
            pp = p.reshape(nt, R, 256)
        else:
            pp = np.zeros(nt * T)
            pp[:n] = p
            pp = pp.reshape(nt, R, 256)
        acc = pp[:, 0, :]
        for r in range(1, R):
            acc = acc + pp[:, r, :]
        part = _block_sum(acc, 256)
        if key_perm is not None and n > 0:
            part = part[np.argsort(np.asarray(key_perm)[np.arange(nt) * T], kind="stable")]
        fin = np.zeros(1024)
        for q0 in range(0, nt, 1024):
            c = part[q0:q0 + 1024]
            fin[:len(c)] = fin[:len(c)] + c
        return float(_block_sum(fin.reshape(1, 1024), 1024)[0])
    return dot


def permuted_keep_order(A, perm):
    """P A P^T with row k = row perm[k] of A and each row's entries kept in
    A's CSR order (columns renumbered, not sorted): its CSR row sums are A's
    row sums term for term, as the device's P-order SELL copy computes them."""
    from scipy import sparse
    A = sparse.csr_matrix(A)
    perm = np.asarray(perm, np.int64)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    lens = np.diff(A.indptr)[perm]
    indptr = np.concatenate([[0], np.cumsum(lens)])
    nnz = int(indptr[-1])
    src = np.repeat(A.indptr[perm].astype(np.int64), lens) + (np.arange(nnz) - np.repeat(indptr[:-1], lens))
    Ap = sparse.csr_matrix((A.data[src], inv[A.indices[src]], indptr), shape=A.shape)
    Ap.has_sorted_indices = False
    return Ap


def restated_device(A, b, solver="bicgstab", rows_per_thread=2, perm=None, x0=None, precond="none",
                    M=None, **kw):
    """The oracle's restated BiCGSTAB (or CG) with the device's dot product
    order: the result the device's FP64 solve must equal bit for bit.
    With an ILU(0) row ordering ``perm`` ('ilu0_color' and a ``perm`` of the
    colour ordering), the solve runs in that order as the device's does:
    P A P^T with A's entry order in each row, b and x0 gathered, the ILU(0)
    factors of P A P^T applied directly, x scattered back at the end -- the
    same arithmetic as running in A's order with the products of every dot
    summed in P-order, without the gathers.  Returns (x, info,
    iterations[, half_step])."""
    from oracle import solve as osolve
    run = osolve.cg_restated if solver == "cg" else osolve.bicgstab_restated
    if perm is None and precond == "ilu0_color" and M is None:
        perm = osolve.color_order(A)[0]
    if perm is None:
        return run(A, b, dot=dot_tile_order(rows_per_thread), x0=x0, precond=precond, M=M, **kw)
    dot = dot_tile_order(rows_per_thread, tiles_by=perm)
    from scipy.sparse.linalg import LinearOperator
    perm = np.asarray(perm, np.int64)
    Ap = permuted_keep_order(A, perm)
    if M is None:
        if precond in ("ilu0_color", "ilu0_segment", "ilu0_perm"):
            f = osolve.ILU0(osolve.permuted(A, perm))
            M = LinearOperator(A.shape, matvec=f.solve)
        elif precond != "none":
            M = osolve.make_preconditioner(Ap, precond)
    else:
        raise ValueError("a permuted restatement builds its own preconditioner")
    bp = np.asarray(b, np.float64)[perm]
    xp0 = None if x0 is None else np.asarray(x0, np.float64)[perm]
    out = run(Ap, bp, dot=dot, x0=xp0, M=M, **kw)
    x = np.empty_like(out[0])
    x[perm] = out[0]
    return (x,) + tuple(out[1:])


def gmres_band(A, b, **kw):
    """scipy GMRES, and the restated GMRES (oracle.solve.gmres_restated)
    with the dot products summed in three other orders (exactly rounded,
    reversed, in chunks of 256) and the x update summed in k order: GMRES's
    own summation-order noise, which is the only arithmetic the device's
    solve does in another order.  Returns (inner iteration counts, cycle
    counts, infos, scipy's x, max relative L2 spread of the others from
    it)."""
    from oracle import solve as osolve
    x0, info0, it0, _, cyc0 = osolve.solve_gmres(A, b, counts=True, **kw)
    its, cycs, infos, spread = [it0], [cyc0], [info0], 0.0
    rk = {k: v for k, v in kw.items() if k in ("x0", "rtol", "atol", "maxiter", "precond", "restart")}
    if rk.get("restart") is None:
        rk["restart"] = 20
    for dot in DOT_ORDERS.values():
        x, info, it, cyc = osolve.gmres_restated(A, b, dot=dot, xupdate="sequential", **rk)
        its.append(it)
        cycs.append(cyc)
        infos.append(info)
        spread = max(spread, rel_l2(x, x0))
    return its, cycs, infos, x0, spread
