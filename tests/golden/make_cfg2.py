"""Writes the committed config-2 fixture from the CPU oracle (the C
restatement of LUDecomposition::new / solve, src/decomp/lu.rs:104-278, which
tests/test_oracle.py pins against the reference's own known-answer tests).

Config 2: A is 8192 x 8192 standard normal, B is 8192 x 16 standard normal,
seed 2024, streams 0 (A) and 2 (B), row-major draw order, Box-Muller over
index pairs (oracle.normal).  The fixture holds the full permutation vector,
the sign, the tie-freeness certificate (the smallest relative gap between the
two largest pivot candidates over all 8192 steps), and a sample of the
solution: every 32nd row of X (all 16 columns), the infinity norm of each
column and each column's normwise backward error.  The factorisation uses oracle.lu_new_colcopy (the same
operations in the same order as oracle.lu_new, with a contiguous copy of
the current column: bit-identical, minutes instead of hours at n = 8192)."""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle  # noqa: E402

SEED, N, NRHS, STRIDE = 2024, 8192, 16, 32


def main():
    a = oracle.normal(SEED, 0, (N, N))
    b = oracle.normal(SEED, 2, (N, NRHS))
    t0 = time.time()
    lu, piv, sign, gap = oracle.lu_new_colcopy(a, want_gap=True)
    t1 = time.time()
    x = oracle.lu_solve(lu, piv, b)
    t2 = time.time()
    rows = list(range(0, N, STRIDE))
    # normwise backward error of the oracle's solution, per column
    r = np.abs(b - a @ x).max(axis=0)
    eta = r / (np.abs(a).sum(axis=1).max() * np.abs(x).max(axis=0) + np.abs(b).max(axis=0))
    out = {
        "config": "cfg2: n=8192 standard normal A, 16 standard normal RHS, LU + solve",
        "seed": SEED, "streams": {"A": 0, "B": 2},
        "generator": "splitmix64 counter, Box-Muller over index pairs (oracle.normal)",
        "piv": [int(p) for p in piv], "pospivsign": bool(sign), "min_pivot_gap": gap,
        "x_rows": rows, "x_sample": x[rows].tolist(),
        "x_col_inf": np.abs(x).max(axis=0).tolist(),
        "x_backward_error": eta.tolist(),
        "oracle_seconds": {"lu": t1 - t0, "solve": t2 - t1},
    }
    with open(os.path.join(HERE, "cfg2_n8192.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
