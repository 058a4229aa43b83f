"""Autotuner and learned tile-cost model.

The library exposes a timing service (``xkblas_tune_time``: one call of a
given shape and configuration on device-resident operands, HIP-event kernel
time) and the knobs the model chooses between: the GEMM macro tile, the block
order NB of TRSM/POTRF, the split count and the BF16 residue-GEMM variant.
This module

* collects timings of GEMM / SYRK / TRSM / POTRF calls with shapes drawn
  log-uniformly from [64, 32768] (``collect``);
* trains the 3-layer MLP of the library (in -> 64 -> 64 -> 1, FP32; forward,
  backward and Adam run as the library's own HIP kernels) to predict
  log(kernel time) from shape and configuration (``CostModel.fit``);
* picks the configuration with the lowest predicted time
  (``CostModel.pick``), falling back to exhaustive measurement when a shape
  lies outside the range the model was trained on (``choose``);
* judges the picks against exhaustive autotuning and against the method
  without the model, the reference's tile-size rule
  (``xkblas_auto_tilesize``, blas/xkblas.c:2437-2521) with the library's
  default kernels (``evaluate``).
"""
import ctypes
import math
import os
import time

import numpy as np

import xkblas_amd as xk

OPS = {"gemm": 0, "syrk": 1, "trsm": 2, "potrf": 3}
NFEAT = 14
LOG_LO, LOG_HI = 6.0, 15.0  # shapes in [64, 32768]
# xkblas_kernel_t of the reference rule used as the method without the
# model (POTRF has no reference rule; it gets the TRSM one)
KERN = {"gemm": 1, "syrk": 6, "trsm": 4, "potrf": 4}


SPLIT_MODES = (2, 3)  # XKBLAS_SPLIT_BF16_MATH, XKBLAS_SPLIT_FP8_MATH


def split_counts(mode, k):
    """The split count for inner dimension k: the library's default, the
    smallest count keeping >= 46 bits of every row's largest entry (the
    accuracy contract of the split modes).  More moduli only add planes, so
    the count is an accuracy setting, not a speed choice."""
    L = xk.lib()
    nm = ctypes.c_int()
    L.xkblas_split_info(mode, max(k, 2), 0, ctypes.byref(nm), None)
    return [nm.value]


def candidates(op, dtype, mode, m, n, k):
    """Configurations the model chooses between for one call."""
    if op in ("gemm", "syrk"):
        if mode in SPLIT_MODES and dtype == 0:
            kernels = (0, 1) if mode == 2 else (0,)
            return [{"tile": 0, "nb": 0, "split": c, "kernel": kv} for c in split_counts(mode, k) for kv in kernels]
        tiles = (32, 64, 128) if dtype == 0 else (64, 128)
        return [{"tile": t, "nb": 0, "split": 0, "kernel": 0} for t in tiles]
    dim = m
    nbs = [b for b in (64, 128, 256, 512, 1024, 2048, 4096) if b <= max(64, dim) and (b < 4096 or dim >= 16384)]
    return [{"tile": 0, "nb": b, "split": 0, "kernel": 0} for b in nbs]


def heuristic(op, dtype, mode, m, n, k):
    """The method without the model: the reference's tile-size rule for one
    GPU and the library's default kernels."""
    L = xk.lib()
    old = L.xkblas_get_param()
    L.xkblas_set_param(0, 0)
    nb = int(L.xkblas_auto_tilesize(None, KERN[op], m, n if op != "potrf" else m, k))
    L.xkblas_set_param(old, 0)
    if op in ("gemm", "syrk"):
        return {"tile": 0, "nb": 0, "split": 0, "kernel": 0}
    return {"tile": 0, "nb": min(nb, max(64, m)), "split": 0, "kernel": 0}


def features(op, dtype, mode, m, n, k, cfg):
    f = np.zeros(NFEAT, dtype=np.float32)
    lg = lambda v: (math.log2(max(v, 1)) - LOG_LO) / (LOG_HI - LOG_LO)
    f[0], f[1], f[2] = lg(m), lg(n), lg(k)
    f[3 + OPS[op]] = 1.0
    f[7] = float(dtype)
    f[8] = 1.0 if mode == 2 else 0.0
    f[9] = {0: 0.5, 32: -0.5, 64: 0.0, 128: 1.0}[cfg["tile"]]
    f[10] = (math.log2(cfg["nb"]) - 6.0) / 5.0 if cfg["nb"] else 0.0
    f[11] = cfg["split"] / 20.0
    f[12] = float(cfg["kernel"])
    f[13] = 1.0 if mode == 3 else 0.0
    return f


def measure(op, dtype, mode, m, n, k, cfg, iters=1):
    """Kernel time (ms) of one call through the library's timing service."""
    ms = xk.lib().xkblas_tune_time(OPS[op], dtype, mode, 111, 111, m, n, k, cfg["tile"], cfg["nb"], cfg["split"],
                                   cfg["kernel"], iters)
    if ms < 0:
        raise ValueError("bad tuning request %r" % ((op, dtype, mode, m, n, k, cfg),))
    return ms


def work(op, m, n, k):
    if op == "gemm":
        return 2.0 * m * n * k
    if op == "syrk":
        return float(n) * n * k
    if op == "trsm":
        return float(m) * m * n
    return m ** 3 / 3.0


def draw(r, max_flop=2e13, ops=("gemm", "syrk", "trsm", "potrf")):
    """One tuning problem: op, dtype, mode and a log-uniform shape."""
    while True:
        op = ops[int(r.integers(len(ops)))]
        dtype = int(r.random() < 0.25)
        mode = 0
        if dtype == 0 and r.random() < 0.4:
            mode = 2 if r.random() < 0.6 else 3
        m, n, k = (int(round(2 ** r.uniform(LOG_LO, LOG_HI))) for _ in range(3))
        if op == "potrf":
            n = k = m
        if op == "trsm":
            k = m
        if work(op, m, n, k) <= max_flop:
            return op, dtype, mode, m, n, k


def collect(n_samples, seed=0, time_budget=None, max_flop=2e13, progress=None):
    """Timings of random (problem, configuration) pairs: X (n x NFEAT), y = log ms."""
    r = np.random.default_rng(seed)
    X, y, meta = [], [], []
    t0 = time.time()
    while len(y) < n_samples:
        if time_budget is not None and time.time() - t0 > time_budget:
            break
        op, dtype, mode, m, n, k = draw(r, max_flop)
        cands = candidates(op, dtype, mode, m, n, k)
        cfg = cands[int(r.integers(len(cands)))]
        ms = measure(op, dtype, mode, m, n, k, cfg)
        X.append(features(op, dtype, mode, m, n, k, cfg))
        y.append(math.log(max(ms, 1e-4)))
        meta.append((OPS[op], dtype, mode, m, n, k, cfg["tile"], cfg["nb"], cfg["split"], cfg["kernel"]))
        if progress and len(y) % progress == 0:
            print("tuner: %d timings (%.0f s)" % (len(y), time.time() - t0), flush=True)
    return np.array(X, dtype=np.float32), np.array(y, dtype=np.float32), np.array(meta, dtype=np.int64)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class CostModel:
    """The library's MLP wrapped for the tuner."""

    def __init__(self, seed=1):
        self.L = xk.lib()
        self.h = self.L.xkblas_mlp_create(NFEAT, seed)
        if not self.h:
            raise RuntimeError("xkblas_mlp_create failed")
        self.mu = 0.0
        self.sd = 1.0

    def fit(self, X, y, epochs=400, batch=256, lr=2e-3, seed=0):
        r = np.random.default_rng(seed)
        self.mu, self.sd = float(y.mean()), float(y.std() or 1.0)
        t = ((y - self.mu) / self.sd).astype(np.float32)
        idx = r.permutation(len(y))
        Xs, ts = np.ascontiguousarray(X[idx]), np.ascontiguousarray(t[idx])
        loss = self.L.xkblas_mlp_train(self.h, _p(Xs), _p(ts), len(ts), epochs, batch, lr)
        return float(loss)

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros(len(X), dtype=np.float32)
        if len(X):
            self.L.xkblas_mlp_forward(self.h, _p(X), len(X), _p(out))
        return out * self.sd + self.mu

    def pick(self, op, dtype, mode, m, n, k):
        cands = candidates(op, dtype, mode, m, n, k)
        pred = self.predict(np.stack([features(op, dtype, mode, m, n, k, c) for c in cands]))
        return cands[int(np.argmin(pred))]

    def params(self):
        p = np.zeros(self.L.xkblas_mlp_num_params(self.h), dtype=np.float32)
        self.L.xkblas_mlp_get_params(self.h, _p(p))
        return p

    def load(self, params, mu, sd):
        params = np.ascontiguousarray(params, dtype=np.float32)
        self.L.xkblas_mlp_set_params(self.h, _p(params))
        self.mu, self.sd = float(mu), float(sd)

    def close(self):
        if self.h:
            self.L.xkblas_mlp_destroy(self.h)
            self.h = None


TRAINED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning")


def load_trained(path=TRAINED):
    """The model trained by `main` and shipped in xkblas_amd/tuning
    (cost_model_params.npy + cost_model.json), or None if absent."""
    import json
    pf, jf = os.path.join(path, "cost_model_params.npy"), os.path.join(path, "cost_model.json")
    if not (os.path.exists(pf) and os.path.exists(jf)):
        return None
    with open(jf) as f:
        meta = json.load(f)
    if meta.get("features") != NFEAT:
        return None
    model = CostModel()
    model.load(np.load(pf), meta["mu"], meta["sd"])
    return model


def in_range(m, n, k):
    return all(64 <= v <= 32768 for v in (m, n, k))


def choose(model, op, dtype, mode, m, n, k):
    """Model pick inside the trained range, exhaustive autotuning outside it."""
    if model is not None and in_range(m, n, k):
        return model.pick(op, dtype, mode, m, n, k), "model"
    cands = candidates(op, dtype, mode, m, n, k)
    times = [measure(op, dtype, mode, m, n, k, c, iters=2) for c in cands]
    return cands[int(np.argmin(times))], "exhaustive"


def evaluate(model, problems, iters=2):
    """Judge: time of the pick (the model inside its trained range, else the
    exhaustive fallback), the exhaustive best and the heuristic for each
    problem.  Returns per-problem rows and the geometric means of
    heuristic/pick (gain) and pick/exhaustive (regret)."""
    rows = []
    for (op, dtype, mode, m, n, k) in problems:
        cands = candidates(op, dtype, mode, m, n, k)
        times = {i: measure(op, dtype, mode, m, n, k, c, iters) for i, c in enumerate(cands)}
        best = min(times.values())
        pick, how = choose(model, op, dtype, mode, m, n, k)
        t_model = times[cands.index(pick)]
        h = heuristic(op, dtype, mode, m, n, k)
        t_heur = times[cands.index(h)] if h in cands else measure(op, dtype, mode, m, n, k, h, iters)
        rows.append({"op": op, "dtype": dtype, "mode": mode, "m": m, "n": n, "k": k, "model_ms": t_model,
                     "best_ms": best, "heuristic_ms": t_heur, "pick": pick, "picked_by": how, "heuristic": h})
    gain = float(np.exp(np.mean([math.log(r["heuristic_ms"] / r["model_ms"]) for r in rows])))
    regret = float(np.exp(np.mean([math.log(r["model_ms"] / r["best_ms"]) for r in rows])))
    return rows, gain, regret


def workload_problems(seed=4):
    """The problems the picks are judged on: the split-precision workload
    (DGEMM 16384^3 in split-BF16 and split-FP8 modes, DPOTRF 32768 in
    split-BF16 mode and 16384 in FP64) and a sample of the batched workload's
    shapes (GEMM m, n, k log-uniform in [16, 512]; TRSM n log-uniform in
    [32, 1024] with 64 right-hand sides), FP64 and FP32."""
    probs = [("gemm", 0, 2, 16384, 16384, 16384), ("gemm", 0, 3, 16384, 16384, 16384),
             ("potrf", 0, 2, 32768, 32768, 32768), ("potrf", 0, 0, 16384, 16384, 16384)]
    r = np.random.default_rng(seed)
    for i in range(12):
        dtype = i % 2
        m, n, k = (int(round(2 ** r.uniform(4, 9))) for _ in range(3))
        probs.append(("gemm", dtype, 0, m, n, k))
    for i in range(6):
        n = int(round(2 ** r.uniform(5, 10)))
        probs.append(("trsm", i % 2, 0, n, 64, n))
    return probs


def main(argv=None):
    """Offline job: collect timings, train, judge, save the model.

    python -m xkblas_amd.tuner --samples 20000 --out DIR [--budget SECONDS]
    """
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=20000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--budget", type=float, default=None, help="seconds of collection")
    ap.add_argument("--data", default=None, help="start from the timings in this .npz")
    ap.add_argument("--out", default="tuner_out", help="directory for the timings and the trained model")
    ap.add_argument("--collect-only", action="store_true")
    ap.add_argument("--eval", type=int, default=24, help="held-out problems to judge")
    args = ap.parse_args(argv)
    os.makedirs(args.out, exist_ok=True)
    L = xk.lib()
    L.xkblas_init()
    X = y = meta = None
    if args.data and os.path.exists(args.data):
        d = np.load(args.data)
        X, y, meta = d["X"], d["y"], d["meta"]
    t_start = time.time()
    save_to = os.path.join(args.out, "timings.npz")
    while y is None or len(y) < args.samples:
        left = None if args.budget is None else args.budget - (time.time() - t_start)
        if left is not None and left <= 0:
            break
        have = 0 if y is None else len(y)
        Xn, yn, mn = collect(min(1000, args.samples - have), seed=args.seed + have, time_budget=left)
        X = Xn if X is None else np.concatenate([X, Xn])
        y = yn if y is None else np.concatenate([y, yn])
        meta = mn if meta is None else np.concatenate([meta, mn])
        np.savez_compressed(save_to, X=X, y=y, meta=meta)  # checkpoint after every chunk
        print("tuner: %d timings (%.0f s)" % (len(y), time.time() - t_start), flush=True)
    if args.collect_only or len(y) < args.samples:
        return
    model = CostModel()
    t0 = time.time()
    loss = model.fit(X, y)
    t_fit = time.time() - t0
    r = np.random.default_rng(12345)
    probs = [draw(r, max_flop=5e12) for _ in range(args.eval)]
    rows, gain, regret = evaluate(model, probs)
    wrows, wgain, wregret = evaluate(model, workload_problems())
    res = {"timings": int(len(y)), "train_loss_std_units": loss, "fit_seconds": t_fit,
           "held_out": {"gain_vs_heuristic": gain, "regret_vs_exhaustive": regret, "rows": rows},
           "workloads": {"gain_vs_heuristic": wgain, "regret_vs_exhaustive": wregret, "rows": wrows},
           "mu": model.mu, "sd": model.sd}
    np.save(os.path.join(args.out, "cost_model_params.npy"), model.params())
    with open(os.path.join(args.out, "tuner_result.json"), "w") as f:
        json.dump(res, f, indent=1, default=str)
    print("tuner: loss %.4f  held-out gain %.3fx regret %.3fx  workloads gain %.3fx regret %.3fx"
          % (loss, gain, regret, wgain, wregret), flush=True)
    model.close()
    L.xkblas_tune_release()


if __name__ == "__main__":
    main()
