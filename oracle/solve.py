"""scipy harness with the reference's solve semantics, plus textbook ILU(0).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* ``solve_cg`` calls ``scipy.sparse.linalg.cg`` as ``solve_pressure_correction``
  does (core/linear_solver.py:88-93), with explicit ``rtol``/``atol`` and an
  iteration-counting callback.
* ``solve_gmres`` calls ``scipy.sparse.linalg.gmres`` as ``solve_linear_system``
  does for method='gmres' (core/linear_solver.py:64-68: restart 20, maxiter
  counting restart cycles), with explicit ``rtol``/``atol``; it counts inner
  iterations with a 'pr_norm' callback and restart cycles with an 'x'
  callback (two runs of the same deterministic solve).
* ``solve_bicgstab`` calls ``scipy.sparse.linalg.bicgstab`` exactly as
  ``solve_linear_system`` does (core/linear_solver.py:56-62) but with explicit
  ``rtol``/``atol`` and an iteration-counting callback (the convention of
  verification/case16_preconditioner.py:161-173).
* ``make_preconditioner`` restates ``create_preconditioner``
  (core/preconditioner.py:40-54): Jacobi ``where(|diag|>1e-30, 1/diag, 0)``
  and scipy ``spilu(A.tocsc(), fill_factor=1, drop_tol=0)``; ``'ilu0_struct'``
  is the textbook ILU(0) of ilu0.c that the GPU implements, ``'ilu0_color'``
  the same on P A P^T for the first-fit colour ordering, ``'ilu0_segment'``
  for the segment ordering (``segment_order``).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
from scipy import sparse
from scipy.sparse.linalg import LinearOperator, bicgstab, cg, gmres, spilu

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_build", "libkcfd_oracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        lib.kcfd_oracle_ilu0_factor.argtypes = [ctypes.c_int, P, P, P, P, P]
        lib.kcfd_oracle_ilu0_factor.restype = ctypes.c_int
        lib.kcfd_oracle_ilu0_solve.argtypes = [ctypes.c_int, P, P, P, P, P, P]
        lib.kcfd_oracle_ilu0_solve.restype = None
        lib.kcfd_oracle_color_order.argtypes = [ctypes.c_int, P, P, P, P]
        lib.kcfd_oracle_color_order.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class ILU0:
    """Structural ILU(0) factors on A's canonical CSR pattern (C oracle)."""

    def __init__(self, A):
        A = sparse.csr_matrix(A)
        A.sort_indices()
        self.n = A.shape[0]
        self.indptr = np.ascontiguousarray(A.indptr, np.int32)
        self.indices = np.ascontiguousarray(A.indices, np.int32)
        a = np.ascontiguousarray(A.data, np.float64)
        self.lu = np.empty_like(a)
        self.diagpos = np.full(self.n, -1, np.int32)
        st = _lib().kcfd_oracle_ilu0_factor(self.n, _p(self.indptr), _p(self.indices),
                                            _p(a), _p(self.lu), _p(self.diagpos))
        if st != 0:
            raise ValueError(f"ILU(0) breakdown at row {-st - 1}")

    def solve(self, b):
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty_like(b)
        _lib().kcfd_oracle_ilu0_solve(self.n, _p(self.indptr), _p(self.indices), _p(self.lu),
                                      _p(self.diagpos), _p(b), _p(x))
        return x

    def factors(self):
        """(L, U) as scipy CSR, L unit lower."""
        n = self.n
        rows = np.repeat(np.arange(n), np.diff(self.indptr))
        cols = self.indices
        low = cols < rows
        L = sparse.csr_matrix((self.lu[low], (rows[low], cols[low])), shape=(n, n)) + sparse.eye(n)
        up = cols >= rows
        U = sparse.csr_matrix((self.lu[up], (rows[up], cols[up])), shape=(n, n))
        return L.tocsr(), U.tocsr()


def ilu0_python(A):
    """Pure-Python ILU(0), IKJ (Saad Alg. 10.4), for tiny matrices; used to
    cross-check the C oracle bit for bit."""
    A = sparse.csr_matrix(A)
    A.sort_indices()
    n = A.shape[0]
    ip, ix = A.indptr, A.indices
    lu = A.data.astype(np.float64).copy()
    dpos = np.full(n, -1, np.int64)
    for i in range(n):
        pos = {int(ix[q]): q for q in range(ip[i], ip[i + 1])}
        for q in range(ip[i], ip[i + 1]):
            k = int(ix[q])
            if k >= i:
                break
            lik = lu[q] / lu[dpos[k]]
            lu[q] = lik
            for r in range(dpos[k] + 1, ip[k + 1]):
                j = int(ix[r])
                if j in pos:
                    lu[pos[j]] -= lik * lu[r]
        dpos[i] = pos[i]
    return lu, dpos


def color_order(A):
    """First-fit colouring of A + A^T (ilu0.c): (perm sorted by colour,
    colours per row, number of colours)."""
    A = sparse.csr_matrix(A)
    A.sort_indices()
    n = A.shape[0]
    ip = np.ascontiguousarray(A.indptr, np.int32)
    ix = np.ascontiguousarray(A.indices, np.int32)
    perm = np.empty(n, np.int32)
    col = np.empty(n, np.int32)
    nc = _lib().kcfd_oracle_color_order(n, _p(ip), _p(ix), _p(perm), _p(col))
    if nc < 0:
        raise MemoryError("colour ordering")
    return perm, col, nc


def segment_order(A, seg_len=4):
    """Segment ordering (the device's KCFD_ORDER_SEGMENT), restated in plain
    Python: rows in ascending order cut into segments of at most ``seg_len``
    rows, a new segment starting at a row not coupled (a_{i,i-1} or
    a_{i-1,i}) to the row before; segments coloured first-fit over A + A^T
    in ascending order.  Within each colour (colour 0 first) the segments,
    in the order of their first rows, are taken 64 at a time; a group's
    segments are sorted longest first (stable) and its rows numbered
    step-major (the first row of each segment, then the second rows, ...).
    Returns (perm, colour per row, number of colours).  A design choice of
    the GPU build with no reference counterpart: tests pin it by its
    invariants as well (coupled segments, no coupling inside a colour
    between segments, each segment's rows in order)."""
    A = sparse.csr_matrix(A)
    A.sort_indices()
    n = A.shape[0]
    S = (A + A.T).tocsr()
    S.sort_indices()
    seg = np.empty(n, np.int64)
    starts = []
    length = 0
    for i in range(n):
        if i == 0 or length == seg_len or not _coupled(A, i - 1, i):
            starts.append(i)
            length = 0
        seg[i] = len(starts) - 1
        length += 1
    starts.append(n)
    ns = len(starts) - 1
    scol = np.full(ns, -1, np.int64)
    for g in range(ns):
        taken = set()
        for i in range(starts[g], starts[g + 1]):
            for j in S.indices[S.indptr[i]:S.indptr[i + 1]]:
                h = seg[j]
                if h != g and scol[h] >= 0:
                    taken.add(int(scol[h]))
        c = 0
        while c in taken:
            c += 1
        scol[g] = c
    col = scol[seg]
    nc = int(scol.max()) + 1 if ns else 0
    perm = []
    for c in range(nc):
        segs = [g for g in range(ns) if scol[g] == c]
        for f in range(0, len(segs), 64):
            group = sorted(segs[f:f + 64], key=lambda g: -(starts[g + 1] - starts[g]))   # stable
            lens = [starts[g + 1] - starts[g] for g in group]
            for k in range(max(lens)):
                perm.extend(starts[g] + k for g, ln in zip(group, lens) if ln > k)
    return np.asarray(perm, np.int32), col, nc


def _coupled(A, i, j):
    """a_ij or a_ji in A's pattern (an explicit zero counts)."""
    return (j in A.indices[A.indptr[i]:A.indptr[i + 1]]) or (i in A.indices[A.indptr[j]:A.indptr[j + 1]])


def permuted(A, perm):
    """P A P^T as canonical CSR (row k = row perm[k] of A)."""
    A = sparse.csr_matrix(A)
    Ap = sparse.csr_matrix(A[perm][:, perm])
    Ap.sort_indices()
    return Ap


class ILU0Ordered:
    """M^{-1} z = P^T ILU0(P A P^T)^{-1} P z for a row ordering perm."""

    def __init__(self, A, perm):
        self.perm = np.asarray(perm, np.int64)
        self.inv = np.empty_like(self.perm)
        self.inv[self.perm] = np.arange(len(self.perm))
        self.f = ILU0(permuted(A, self.perm))

    def solve(self, z):
        return self.f.solve(np.asarray(z, np.float64)[self.perm])[self.inv]


def make_preconditioner(A, method: str):
    """core/preconditioner.py:40-54 (+ 'ilu0_struct', 'ilu0_color')."""
    n = A.shape[0]
    if method == "none":
        return None
    if method == "jacobi":
        d = A.diagonal()
        with np.errstate(divide="ignore"):
            dinv = np.where(np.abs(d) > 1e-30, 1.0 / d, 0.0)
        return LinearOperator((n, n), matvec=lambda x: dinv * x)
    if method == "ilu0":
        ilu = spilu(sparse.csc_matrix(A), fill_factor=1, drop_tol=0)
        return LinearOperator((n, n), matvec=ilu.solve)
    if method == "ilu0_struct":
        f = ILU0(A)
        return LinearOperator((n, n), matvec=f.solve)
    if method == "ilu0_color":
        f = ILU0Ordered(A, color_order(A)[0])
        return LinearOperator((n, n), matvec=f.solve)
    if method == "ilu0_segment":
        f = ILU0Ordered(A, segment_order(A)[0])
        return LinearOperator((n, n), matvec=f.solve)
    raise ValueError(f"Unknown preconditioner: {method}")


def solve_bicgstab(A, b, x0=None, rtol=1e-8, atol=0.0, maxiter=None, precond="jacobi", M=None,
                   on_iter=None):
    """scipy BiCGSTAB with the reference's call; returns (x, info, iterations)
    where iterations = number of callback calls = completed full iterations.
    ``on_iter(k)``, when given, is called after iteration k."""
    A = sparse.csr_matrix(A)
    n = A.shape[0]
    if x0 is None:
        x0 = np.zeros(n)
    if M is None:
        M = make_preconditioner(A, precond)
    count = [0]

    def cb(_xk):
        count[0] += 1
        if on_iter is not None:
            on_iter(count[0])

    x, info = bicgstab(A, b, x0=np.array(x0, np.float64), rtol=rtol, atol=atol,
                       maxiter=maxiter, M=M, callback=cb)
    return x, info, count[0]


def solve_cg(A, b, x0=None, rtol=1e-8, atol=0.0, maxiter=None, precond="jacobi", M=None):
    """scipy CG with the call of solve_pressure_correction
    (core/linear_solver.py:88-93); returns (x, info, iterations), iterations
    = callback calls = completed iterations."""
    A = sparse.csr_matrix(A)
    n = A.shape[0]
    if x0 is None:
        x0 = np.zeros(n)
    if M is None:
        M = make_preconditioner(A, precond)
    count = [0]

    def cb(_xk):
        count[0] += 1

    x, info = cg(A, b, x0=np.array(x0, np.float64), rtol=rtol, atol=atol, maxiter=maxiter, M=M,
                 callback=cb)
    return x, info, count[0]


def solve_gmres(A, b, x0=None, rtol=1e-8, atol=0.0, maxiter=None, precond="none", M=None,
                restart=None, counts=False):
    """scipy GMRES with the reference's call (core/linear_solver.py:64-68);
    returns (x, info, inner iterations) and, with counts=True, also the
    per-iteration preconditioned residual estimates and the restart cycles."""
    A = sparse.csr_matrix(A)
    n = A.shape[0]
    if x0 is None:
        x0 = np.zeros(n)
    if M is None:
        M = make_preconditioner(A, precond)
    pr = []
    x, info = gmres(A, b, x0=np.array(x0, np.float64), rtol=rtol, atol=atol, restart=restart,
                    maxiter=maxiter, M=M, callback=pr.append, callback_type="pr_norm")
    if not counts:
        return x, info, len(pr)
    cyc = [0]

    def cb(_xk):
        cyc[0] += 1

    x2, _ = gmres(A, b, x0=np.array(x0, np.float64), rtol=rtol, atol=atol, restart=restart,
                  maxiter=maxiter, M=M, callback=cb, callback_type="x")
    assert np.array_equal(x, x2)
    return x, info, len(pr), np.array(pr), cyc[0]


def lartg(f, g):
    """LAPACK xLARTG, the safe-scaling algorithm (Anderson, ACM TOMS 44(1)
    2017, Algorithm 978) that scipy's gmres calls: (c, s, r) with c >= 0 and
    r carrying the sign of f.  The device GMRES implements the same steps."""
    safmin = np.finfo(np.float64).tiny
    safmax = 1.0 / safmin
    rtmin, rtmax = np.sqrt(safmin), np.sqrt(safmax / 2)
    f, g = np.float64(f), np.float64(g)
    f1, g1 = abs(f), abs(g)
    if g == 0:
        return 1.0, 0.0, f
    if f == 0:
        return 0.0, np.copysign(1.0, g), g1
    if rtmin < f1 < rtmax and rtmin < g1 < rtmax:
        d = np.sqrt(f * f + g * g)
        r = np.copysign(d, f)
        return f1 / d, g / r, r
    u = min(safmax, max(safmin, f1, g1))
    fs, gs = f / u, g / u
    d = np.sqrt(fs * fs + gs * gs)
    r = np.copysign(d, f)
    return abs(fs) / d, gs / r, r * u


def gmres_restated(A, b, x0=None, rtol=1e-8, atol=0.0, maxiter=None, precond="none", restart=20,
                   dot=None, xupdate="blas", M=None):
    """GMRES(restart) step by step as the device runs it: left
    preconditioning, modified Gram-Schmidt, ``lartg`` rotations, inner stop
    at |S_{col+1}| <= ptol, the ptol adaptation between cycles, and a back
    substitution that skips zero components.  ``dot(a, b)`` replaces
    numpy's dot product (every inner product and norm goes through it), and
    ``xupdate='sequential'`` sums x += sum_k y_k v_k in k order instead of
    numpy's ``y @ V``, to measure what the summation order alone does.
    ``M`` (a LinearOperator) replaces ``precond``.  Returns (x, info, inner
    iterations, restart cycles)."""
    dot = np.dot if dot is None else dot

    def norm(v):
        return np.sqrt(dot(v, v))

    A = sparse.csr_matrix(A)
    n = A.shape[0]
    Mop = make_preconditioner(A, precond) if M is None else M
    psolve = (lambda v: v.copy()) if Mop is None else Mop.matvec
    b = np.asarray(b, np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    bnrm2 = norm(b)
    atol = max(float(atol), float(rtol) * bnrm2)
    if bnrm2 == 0:
        return b.copy(), 0, 0, 0
    eps = np.finfo(np.float64).eps
    maxiter = 10 * n if maxiter is None else maxiter
    m = min(restart, n)
    ptol_factor = 1.0
    ptol = norm(psolve(b)) * min(ptol_factor, atol / bnrm2)
    V = np.zeros((m + 1, n))
    H = np.zeros((m, m + 1))
    cs = np.zeros((m, 2))
    r = b - A @ x
    rnorm = norm(r)
    if rnorm < atol:
        return x, 0, 0, 0
    inner = cycles = 0
    for _ in range(maxiter):
        cycles += 1
        V[0] = psolve(r)
        beta = norm(V[0])
        V[0] *= 1.0 / beta
        S = np.zeros(m + 1)
        S[0] = beta
        broke = False
        for j in range(m):
            w = psolve(A @ V[j])
            h0 = norm(w)
            for k in range(j + 1):
                H[j, k] = dot(V[k], w)
                w -= H[j, k] * V[k]
            h1 = norm(w)
            broke = h1 <= eps * h0
            H[j, j + 1] = 0.0 if broke else h1
            V[j + 1] = w if broke else w * (1.0 / h1)
            for k in range(j):
                c, s = cs[k]
                a0, a1 = H[j, k], H[j, k + 1]
                H[j, k], H[j, k + 1] = c * a0 + s * a1, -s * a0 + c * a1
            c, s, H[j, j] = lartg(H[j, j], H[j, j + 1])
            H[j, j + 1] = 0.0
            cs[j] = c, s
            S[j], S[j + 1] = c * S[j], -s * S[j]
            presid = abs(S[j + 1])
            inner += 1
            if presid <= ptol or broke:
                break
        if H[j, j] == 0:
            S[j] = 0
        y = S[:j + 1].copy()
        for k in range(j, 0, -1):
            if y[k] != 0:
                y[k] /= H[k, k]
                y[:k] -= y[k] * H[k, :k]
        if y[0] != 0:
            y[0] /= H[0, 0]
        if xupdate == "sequential":
            acc = np.zeros(n)
            for k in range(j + 1):
                acc = acc + y[k] * V[k]
            x += acc
        else:
            x += y @ V[:j + 1]
        r = b - A @ x
        rnorm = norm(r)
        if rnorm <= atol or broke:
            break
        if presid <= ptol:
            ptol_factor = max(eps, 0.25 * ptol_factor)
        else:
            ptol_factor = min(1.0, 1.5 * ptol_factor)
        ptol = presid * min(ptol_factor, atol / rnorm)
    return x, (0 if rnorm <= atol else maxiter), inner, cycles


def _psolve_of(A, precond, M):
    Mop = make_preconditioner(A, precond) if M is None else M
    return (lambda v: v.copy()) if Mop is None else Mop.matvec


def bicgstab_restated(A, b, x0=None, rtol=1e-5, atol=0.0, maxiter=None, precond="none", M=None,
                      dot=None):
    """scipy 1.15.3 ``bicgstab`` (scipy/sparse/linalg/_isolve/iterative.py:
    157-302, the solver ``solve_linear_system`` calls at
    core/linear_solver.py:59) written out step by step, with ``dot(a, b)``
    in place of numpy's dot product.  Every inner product and every norm
    (``np.linalg.norm(v)`` is ``sqrt(v.dot(v))`` for a real vector) goes
    through ``dot``; everything else is scipy's arithmetic in scipy's
    order: r = b - A x (b itself when x0 = 0), the stop test on ||r|| at the
    top of each iteration, the rho and omega breakdown tests against eps^2,
    beta = (rho / rho_prev) (alpha / omega), p -= omega v; p *= beta;
    p += r, the rv == 0 test, the half-step exit on ||s||, omega = t.s / t.t,
    x += alpha M p; x += omega M s; r -= omega t.  With ``dot=np.dot`` it is
    scipy bit for bit.  Returns (x, info, iterations, half_step):
    iterations = completed full iterations (scipy's callback count),
    half_step = 1 when it stopped at the test on ||s||."""
    dot = np.dot if dot is None else dot
    A = sparse.csr_matrix(A)
    n = A.shape[0]
    psolve = _psolve_of(A, precond, M)
    b = np.asarray(b, np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    bnrm2 = np.sqrt(dot(b, b))
    atol = max(float(atol), float(rtol) * float(bnrm2))
    if bnrm2 == 0:
        return b.copy(), 0, 0, 0
    maxiter = 10 * n if maxiter is None else maxiter
    rhotol = np.finfo(np.float64).eps ** 2
    r = b - A @ x if x.any() else b.copy()
    rt = r.copy()
    rho_prev = omega = alpha = p = v = None
    done = 0
    for iteration in range(maxiter):
        if np.sqrt(dot(r, r)) < atol:
            return x, 0, done, 0
        rho = dot(rt, r)
        if np.abs(rho) < rhotol:
            return x, -10, done, 0
        if iteration > 0:
            if np.abs(omega) < rhotol:
                return x, -11, done, 0
            beta = (rho / rho_prev) * (alpha / omega)
            p -= omega * v
            p *= beta
            p += r
        else:
            p = r.copy()
        phat = psolve(p)
        v = A @ phat
        rv = dot(rt, v)
        if rv == 0:
            return x, -11, done, 0
        alpha = rho / rv
        r -= alpha * v
        s = r.copy()
        if np.sqrt(dot(s, s)) < atol:
            x += alpha * phat
            return x, 0, done, 1
        shat = psolve(s)
        t = A @ shat
        omega = dot(t, s) / dot(t, t)
        x += alpha * phat
        x += omega * shat
        r -= omega * t
        rho_prev = rho
        done += 1
    return x, maxiter, done, 0


def cg_restated(A, b, x0=None, rtol=1e-5, atol=0.0, maxiter=None, precond="none", M=None, dot=None):
    """scipy 1.15.3 ``cg`` (iterative.py, the solver of
    ``solve_pressure_correction``, core/linear_solver.py:88-93) step by step
    with a pluggable ``dot``, as ``bicgstab_restated``: the stop test on
    ||r|| at the top of each iteration, z = M r, rho = r.z, p = z
    (iteration 0) or p *= beta; p += z with beta = rho / rho_prev, q = A p,
    alpha = rho / p.q, x += alpha p, r -= alpha q.  A zero p.q (where scipy
    divides by zero) returns info -12.  Returns (x, info, iterations)."""
    dot = np.dot if dot is None else dot
    A = sparse.csr_matrix(A)
    n = A.shape[0]
    psolve = _psolve_of(A, precond, M)
    b = np.asarray(b, np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    bnrm2 = np.sqrt(dot(b, b))
    atol = max(float(atol), float(rtol) * float(bnrm2))
    if bnrm2 == 0:
        return b.copy(), 0, 0
    maxiter = 10 * n if maxiter is None else maxiter
    r = b - A @ x if x.any() else b.copy()
    rho_prev = p = None
    done = 0
    for iteration in range(maxiter):
        if np.sqrt(dot(r, r)) < atol:
            return x, 0, done
        z = psolve(r)
        rho = dot(r, z)
        if iteration > 0:
            p *= rho / rho_prev
            p += z
        else:
            p = z.copy()
        q = A @ p
        pq = dot(p, q)
        if pq == 0 or pq != pq:
            return x, -12, done
        alpha = rho / pq
        x += alpha * p
        r -= alpha * q
        rho_prev = rho
        done += 1
    return x, maxiter, done


def true_rel_residual(A, x, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(b - A @ x) / (nb if nb > 0 else 1.0))
