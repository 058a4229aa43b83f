"""The committed shape list of the batched uneven workload matches its
generator and the workload's definition (CPU)."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def test_batched_shape_list():
    with open(os.path.join(HERE, "golden", "batched_c4_shapes.json")) as f:
        d = json.load(f)
    g = np.array(d["gemm_mnk"])
    t = np.array(d["trsm_n"])
    assert g.shape == (4096, 3) and t.shape == (1024,) and d["trsm_nrhs"] == 64
    assert g.min() >= 16 and g.max() <= 512 and t.min() >= 32 and t.max() <= 1024
    # log-uniform: the median of log2 sits near the middle of the range
    assert abs(np.median(np.log2(g)) - 6.5) < 0.15 and abs(np.median(np.log2(t)) - 7.5) < 0.25
    r = np.random.default_rng(d["seed"])
    lu = lambda lo, hi, size: np.rint(np.exp2(r.uniform(np.log2(lo), np.log2(hi), size))).astype(int)
    assert np.array_equal(lu(16, 512, (4096, 3)), g)
    assert np.array_equal(lu(32, 1024, 1024), t)
