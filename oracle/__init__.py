"""CPU oracle for the FVM assembly + Green-Gauss + MUSCL + BiCGSTAB hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or as the timed CPU baseline).  The product package never imports it.

Modules
-------
``literal``  literal per-face/per-cell loop transcriptions of the reference
             functions (small meshes only; pins the vectorised restatement).
``fvm``      vectorised numpy restatement of the reference assembly,
             gradient, mass-flux and MUSCL functions.  It rebuilds the exact
             COO triplet order of the reference, so scipy's ``tocsr`` yields the
             same canonical CSR the reference would.
``solve``    scipy harness with the reference solve semantics (Jacobi
             LinearOperator, ``spilu(fill_factor=1, drop_tol=0)``) and a
             textbook structural ILU(0) (C, ``ilu0.c``; Python twin for tiny
             matrices).
``golden``   the case-16 system of the reference verification script, whose
             Jacobi BiCGSTAB iteration count (95) is printed in the reference's
             committed figure.

Pinning (what checks each function; see tests/test_oracle_*.py)
----------------------------------------------------------------
* mesh/assembly/gradient/MUSCL restatement: equal to the literal transcription
  on 2x2, 3x3, 4x4 quad, 2x2x2 / 3x3x3 hex and small tet meshes (brute force);
  closed forms on uniform grids (interior coeff = gamma, Dirichlet = 2 gamma,
  2D; gamma*h and 2*gamma*h, 3D); invariants (zero row sums, symmetry,
  M-matrix sign pattern, structure = diag U face pairs); Green-Gauss exact for
  linear fields at interior cells; MUSCL zero for linear fields (r = 0 quirk).
* Jacobi BiCGSTAB: case-16 golden iteration count 95 from the reference figure.
* ILU(0): (L U)_ij = A_ij on pattern(A); equals exact LU (scipy splu with
  natural ordering, no pivoting) when the pattern is closed under elimination
  (tridiagonal); the C and Python versions agree bit for bit.
"""
