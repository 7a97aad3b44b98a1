"""Writes the committed config-1 fixture from the CPU oracle (which is pinned
against the reference's own known-answer tests in tests/test_oracle.py).

Config 1: A (256x256), B (256x256) and b (256x1) uniform in [-1,1] from
seed 20240501, streams 0, 1, 2 in that draw order (row-major)."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle  # noqa: E402


def main():
    seed, n = 20240501, 256
    a = oracle.uniform(seed, 0, (n, n))
    bm = oracle.uniform(seed, 1, (n, n))
    b = oracle.uniform(seed, 2, (n, 1))
    lu, piv, sign, gap = oracle.lu_new(a, want_gap=True)
    x = oracle.lu_solve(lu, piv, b)
    c = oracle.mul(a, bm)
    idx = [[int(i), int(j)] for i, j in zip(np.arange(0, n, 17), np.arange(3, n, 23))]
    out = {
        "config": "cfg1: n=256 uniform[-1,1], LU+solve and C=A*B",
        "seed": seed, "streams": {"A": 0, "B": 1, "b": 2},
        "generator": "splitmix64 counter, uniform = 2*((bits>>11)*2^-53)-1",
        "piv": piv.tolist(), "pospivsign": bool(sign), "min_pivot_gap": gap,
        "x": x.ravel().tolist(),
        "c_fro": float(np.linalg.norm(c)), "c_index": idx,
        "c_sample": [float(c[i, j]) for i, j in idx],
        "det": oracle.lu_det(lu, sign),
    }
    with open(os.path.join(HERE, "cfg1_n256.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
