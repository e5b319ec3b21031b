"""Write the restated-solver fixtures of the full-size pressure system.

TEST INFRASTRUCTURE.  The oracle's restated BiCGSTAB (oracle/solve.py:
bicgstab_restated, scipy's algorithm, pinned to scipy bit for bit in
tests/test_oracle_restated.py) evaluated in the device's dot-product order
(tests/helpers.py: restated_device) takes about two minutes on the 128^3
SIMPLE pressure-correction system, too long for a GPU test.  This script
runs it once and stores what the device's FP64 solve must reproduce bit for
bit: info, iterations, the half-step flag and the SHA-256 of x (float64,
little-endian), with SHA-256s of the oracle's matrix values and right-hand
side so that a test can check it starts from the same inputs.

    python tests/golden/make_restated_fixtures.py
"""

from __future__ import annotations

import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import __graft_entry__  # noqa: E402

__graft_entry__.load_package()

from helpers import oracle_system, restated_device  # noqa: E402
from oracle import solve as osolve  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, "<f8").tobytes()).hexdigest()


def main():
    from kcfd_amd import workloads
    out = {}
    n = 128
    case = workloads.hex_pressure(n, seed=1002)
    A, b, _ = oracle_system(case)
    perm = osolve.color_order(A)[0]
    for rows in (2,):
        t = time.time()
        x, info, it, half = restated_device(A, b, rows_per_thread=rows, perm=perm, rtol=1e-8,
                                            precond="ilu0_color")
        key = f"hex{n}_pressure_fp64_ilu0_color_rtol1e-8_rows{rows}"
        out[key] = dict(n=int(A.shape[0]), nnz=int(A.nnz), seed=1002, rtol=1e-8, rows_per_thread=rows,
                        precond="ilu0 colour order", info=int(info), iters=int(it), half_step=int(half),
                        x_sha256=sha(x), x_norm=float(np.linalg.norm(x)),
                        A_sha256=sha(A.data), b_sha256=sha(b),
                        rel_resid=osolve.true_rel_residual(A, x, b))
        print(key, out[key], f"{time.time() - t:.1f} s", flush=True)
    with open(os.path.join(HERE, "restated_hex128.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
