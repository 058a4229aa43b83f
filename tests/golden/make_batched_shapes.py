"""Writes batched_c4_shapes.json: the shape list of the batched uneven
workload (4096 GEMMs with m, n, k log-uniform in [16, 512]; 1024 lower
triangular TRSMs with n log-uniform in [32, 1024] and 64 right-hand sides),
drawn once from a fixed seed and committed so every run uses the same list."""
import json
import os

import numpy as np


def main():
    r = np.random.default_rng(2027)
    lu = lambda lo, hi, size: np.rint(np.exp2(r.uniform(np.log2(lo), np.log2(hi), size))).astype(int)
    gemm = lu(16, 512, (4096, 3)).tolist()
    trsm = lu(32, 1024, 1024).tolist()
    out = {"seed": 2027, "gemm_mnk": gemm, "trsm_n": trsm, "trsm_nrhs": 64}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "batched_c4_shapes.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(path)


if __name__ == "__main__":
    main()
