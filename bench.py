"""Benchmark of the tiled BLAS-3 hot path on one MI355X.

Headline metric: FP64-accurate DGEMM / DPOTRF TFLOP/s, split-BF16 vs plain
FP64 MFMA.  A step is one DGEMM C <- 1.5 A B - 0.5 C with M = N = K = 16384,
rows of A and columns of B scaled by 2^u, u ~ U[-8, 8] (base entries
U[-0.5, 0.5], seeded), called through xkblas_dgemm_async with A, B, C already
resident in the HBM tile cache.  `value` is the FP64-accurate split-BF16
rate; the plain FP64-MFMA rate and the split-FP8 rate are reported beside it,
each with the relative error on sampled entries against a double-double
oracle.  DPOTRF of A = G G^T + N I is timed at N = 16384 (FP64 MFMA) and at
N = 32768 in FP64 MFMA and split-BF16 modes (sampled residuals, and the
difference between the two factors).  The CPU baseline is the host BLAS
(dgemm_ via dlopen, the reference's native path) on a bounded 4096^3 sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 16384]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK = 78.6  # TFLOP/s, MI355X dense FP64 MFMA (AMD datasheet)
BF16_PEAK = 2500.0
FP8_PEAK = 5000.0
FP32_PEAK = 157.3  # TFLOP/s, dense FP32 MFMA
HBM_PEAK = 8000.0  # GB/s, HBM3E


def profiled_traffic(section="split-BF16"):
    """HBM bytes per launch of the residue GEMM at 16384^3, from the committed
    rocprofv3 summary (FETCH_SIZE x 2 + WRITE_SIZE x 1, factors from a 1 GiB
    device copy): the 'Residue GEMM per launch: **R GB read + W GB written'
    line of the given section; None when absent."""
    import re
    path = os.path.join(ROOT, "profiles", "split_dgemm16384_hbm_traffic.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    part = text.split("### " + section, 1)
    if len(part) < 2:
        return None
    m = re.search(r"Residue GEMM per launch: \*\*([0-9.]+) GB read \+ ([0-9.]+) GB written", part[1])
    return None if m is None else (float(m.group(1)) + float(m.group(2))) * 1e9


def make_inputs(n, seed):
    r = np.random.default_rng(seed)
    A = r.uniform(-0.5, 0.5, size=(n, n))
    A *= np.exp2(r.uniform(-8, 8, size=n))[:, None]        # rows of A
    B = r.uniform(-0.5, 0.5, size=(n, n))
    B *= np.exp2(r.uniform(-8, 8, size=n))[None, :]        # columns of B
    C = r.uniform(-0.5, 0.5, size=(n, n))
    return np.asfortranarray(A), np.asfortranarray(B), np.asfortranarray(C)


def time_dgemm(xk, A, B, C, steps, warmup, mode, nb, alpha=1.5, beta=-0.5):
    L = xk.lib()
    L.xkblas_set_modemath(mode)
    L.xkblas_set_param(nb, 0)
    n = A.shape[0]
    al = np.array([alpha])
    be = np.array([beta])
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)

    def step():
        rc = L.xkblas_dgemm_async(111, 111, n, n, n, p(al), p(A), n, p(B), n, p(be), p(C), n)
        assert rc == 0

    for _ in range(warmup):
        step()
    L.xkblas_sync()
    L.xkblas_profile_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    L.xkblas_sync()
    t1 = time.perf_counter()
    return (t1 - t0) / steps


def host_reference(A, B):
    """C = 1.5 A B - 0.5 C0 in FP64 by the host BLAS (the reference's CPU
    path) on the full matrices, for the full normwise comparison; None when
    no host BLAS is found."""
    import oracle
    if oracle.hostblas() is None:
        return None
    n = A.shape[0]
    r = np.random.default_rng(11)
    C = np.asfortranarray(r.uniform(-0.5, 0.5, (n, n)))
    oracle.host_dgemm(111, 111, 1.5, A, B, -0.5, C)
    return C


def accuracy(xk, A, B, mode, nb, samples=96, full_ref=None):
    """Relative error of one C <- 1.5 A B - 0.5 C0 against the double-double
    oracle on sampled entries: normwise over the samples, max elementwise,
    and max componentwise relative to (|alpha||A||B| + |beta||C0|); with
    full_ref (the host BLAS result) also the full normwise difference."""
    import oracle
    n = A.shape[0]
    r = np.random.default_rng(11)
    C0 = np.asfortranarray(r.uniform(-0.5, 0.5, (n, n)))
    C = C0.copy(order="F")
    L = xk.lib()
    L.xkblas_set_modemath(mode)
    L.xkblas_set_param(nb, 0)
    xk.gemm(111, 111, 1.5, A, B, -0.5, C)
    L.xkblas_memory_invalidate(C.ctypes.data_as(ctypes.c_void_p))
    rows = r.integers(0, n, samples)
    cols = r.integers(0, n, samples)
    hi, lo = oracle.dd_entries(111, 111, 1.5, A, B, -0.5, C0, rows, cols)
    got = C[rows, cols]
    err = np.abs((got - hi) - lo)
    absprod = np.array([1.5 * np.abs(A[i, :]) @ np.abs(B[:, j]) + 0.5 * abs(C0[i, j]) for i, j in zip(rows, cols)])
    out = {"normwise_sampled": float(np.linalg.norm(err) / np.linalg.norm(hi)),
           "max_elementwise": float(np.max(err / np.abs(hi))),
           "median_elementwise": float(np.median(err / np.abs(hi))),
           "max_componentwise": float(np.max(err / absprod)), "samples": int(samples),
           "oracle": "double-double dot products"}
    if full_ref is not None:
        out["normwise_full_vs_host_fp64"] = float(np.linalg.norm(C - full_ref) / np.linalg.norm(full_ref))
    return out


def host_to_host(xk, A, B, C, mode, nb):
    """One call from host buffers: invalidated caches, async + coherent +
    sync (the reference tests' timed sequence), so the host-to-device copies
    of A, B, C and the copy back of C are inside the time."""
    L = xk.lib()
    xk.invalidate()
    L.xkblas_set_modemath(mode)
    L.xkblas_set_param(nb, 0)
    n = A.shape[0]
    al, be = np.array([1.5]), np.array([-0.5])
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    L.xkblas_sync()
    t0 = time.perf_counter()
    L.xkblas_dgemm_async(111, 111, n, n, n, p(al), p(A), n, p(B), n, p(be), p(C), n)
    L.xkblas_memory_coherent_async(0, 0, n, n, p(C), n, 8)
    L.xkblas_sync()
    t = time.perf_counter() - t0
    L.xkblas_set_modemath(0)
    return t


def bench_c1(xk, steps=20):
    """The reference's own CPU-runnable case: DGEMM NN M = N = K = 2048, FP64,
    A, B, C ~ U[-0.5, 0.5] (seed 1), alpha 1.5, beta -0.5, tile 512.  Kernel
    rate with the operands resident, one call from host buffers, and the
    normwise difference to the host BLAS result."""
    import oracle
    L = xk.lib()
    n = 2048
    r = np.random.default_rng(1)
    A, B, C0 = (np.asfortranarray(r.uniform(-0.5, 0.5, (n, n))) for _ in range(3))
    C = C0.copy(order="F")
    t = time_dgemm(xk, A, B, C, steps, 2, 0, 512)
    cnt, ms, _ = kernel_stats(xk, "gemm_f64")
    flops = 2.0 * n ** 3
    C = C0.copy(order="F")
    xk.invalidate()
    h2h = host_to_host(xk, A, B, C, 0, 512)
    out = {"m": n, "n": n, "k": n, "tile": 512, "kernel_tflops": round(flops / (ms / cnt) / 1e9, 3),
           "step_ms": round(t * 1e3, 3), "host_to_host_ms": round(h2h * 1e3, 3)}
    if oracle.hostblas() is not None:
        ref = C0.copy(order="F")
        oracle.host_dgemm(111, 111, 1.5, A, B, -0.5, ref)
        out["normwise_vs_host_blas"] = float(np.linalg.norm(C - ref) / np.linalg.norm(ref))
    xk.invalidate()
    return out


def kernel_stats(xk, cls):
    cnt, ms, fl = xk.profile(cls)
    return cnt, ms, fl


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(budget_s=10.0):
    """The reference's CPU path: the host BLAS (MKL) through the oracle's
    restatement of the reference's dlopen path, all host cores.  `value` is
    dgemm_ NN 4096^3 (median of the runs that fit `budget_s`); beside it
    dpotrf_ (lower, order 8192, A = G G^T + n I), dsyrk_ (lower, n = k =
    4096) and dtrsm_ (left, lower, non-unit, 4096 x 4096), median of 3 each."""
    import oracle
    path = oracle.hostblas()
    if path is None:
        return None
    n = 4096
    r = np.random.default_rng(3)
    A = np.asfortranarray(r.uniform(-0.5, 0.5, (n, n)))
    B = np.asfortranarray(r.uniform(-0.5, 0.5, (n, n)))
    C = np.asfortranarray(r.uniform(-0.5, 0.5, (n, n)))
    oracle.host_dgemm(111, 111, 1.5, A, B, -0.5, C)  # warm-up (thread pool)
    times = []
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < budget_s and len(times) < 50:
        t0 = time.perf_counter()
        oracle.host_dgemm(111, 111, 1.5, A, B, -0.5, C)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))

    def med3(fn):
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    others = {}
    npo = 8192
    G = r.uniform(-1, 1, (npo, npo))
    S = np.asfortranarray(G @ G.T + npo * np.eye(npo))
    del G
    F = S.copy(order="F")

    def potrf():
        F[:] = S
        assert oracle.host_dpotrf(122, F) == 0
    t_copy = med3(lambda: F.__setitem__(slice(None), S))
    t = max(med3(potrf) - t_copy, 1e-9)
    others["dpotrf_8192"] = round((npo**3 / 3.0 + npo**2 / 2.0 + npo / 6.0) / t / 1e12, 4)
    del S, F
    t = med3(lambda: oracle.host_dsyrk(122, 111, 1.0, A, 0.0, C))
    others["dsyrk_4096"] = round(float(n) * n * (n + 1) / t / 1e12, 4)
    T = np.asfortranarray(np.tril(r.uniform(0, 1, (n, n))) / n + np.eye(n))
    X = B.copy(order="F")

    def trsm():
        X[:] = B
        oracle.host_dtrsm(141, 122, 111, 131, 1.0, T, X)
    t = max(med3(trsm) - med3(lambda: X.__setitem__(slice(None), B)), 1e-9)
    others["dtrsm_4096"] = round(float(n) ** 3 / t / 1e12, 4)
    cores = int(os.environ.get("MKL_NUM_THREADS", len(os.sched_getaffinity(0))))
    return {"value": round(2.0 * n**3 / med / 1e12, 4), "unit": "TFLOP/s", "cores": cores, "kind": "port",
            "sample": "host BLAS dgemm_ NN 4096^3, median of %d runs, through the oracle's restatement of the "
                      "reference's dlopen host-BLAS path" % len(times),
            "cpu": cpu_model(), "host_blas": oracle.hostblas_version(path), "library": os.path.basename(path),
            "other_routines_tflops": others}


def copy_probe():
    """A plain copy kernel of known size: 1 GiB read + 1 GiB written."""
    import torch
    a = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print("copy probe: %d bytes" % (a.numel() * 8), file=sys.stderr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--nb", type=int, default=4096)
    ap.add_argument("--potrf-n", type=int, default=16384)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-potrf", action="store_true")
    ap.add_argument("--no-fp8", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-full-check", action="store_true", help="skip the full host-BLAS normwise comparison")
    ap.add_argument("--potrf-split-n", type=int, default=32768,
                    help="DPOTRF size for the split-BF16 vs FP64 comparison (0: skip)")
    ap.add_argument("--only-split", type=int, default=0, metavar="MODE",
                    help="profile helper: two split DGEMMs in math mode MODE (2 BF16, 3 FP8), no JSON")
    ap.add_argument("--only-fp64", action="store_true",
                    help="profile helper: two plain FP64-MFMA DGEMMs of the headline workload, no JSON")
    ap.add_argument("--only-potrf", type=int, default=-1, metavar="MODE",
                    help="profile helper: DPOTRF of order --potrf-split-n (block order --potrf-nb) in math mode "
                         "MODE, twice, no JSON")
    ap.add_argument("--potrf-nb", type=int, default=2048)
    ap.add_argument("--only-posv-ir", action="store_true",
                    help="profile helper: xkblas_dposv_ir of order --potrf-split-n, 64 right-hand sides, block "
                         "order --potrf-nb, twice, no JSON")
    ap.add_argument("--only-batched", action="store_true", help="profile helper: the batched leg alone, no JSON")
    ap.add_argument("--batched-parts", action="store_true",
                    help="with --only-batched: also time the GEMM problems alone and the TRSM problems alone")
    ap.add_argument("--copy-probe", action="store_true",
                    help="profile helper: one 1 GiB device-to-device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if "MKL_NUM_THREADS" not in os.environ:
        os.environ["MKL_NUM_THREADS"] = str(len(os.sched_getaffinity(0)))

    import xkblas_amd as xk
    L = xk.lib()
    if L.xkblas_get_devicecount() < 1:
        raise RuntimeError("no GPU visible")
    L.xkblas_init()
    L.xkblas_profile_enable(1)
    if args.only_potrf >= 0:
        S, _ = make_spd(xk, args.potrf_split_n, 1024)
        print(json.dumps(bench_potrf(xk, S, args.potrf_nb, args.only_potrf)[0]), file=sys.stderr)
        L.xkblas_finalize()
        return
    if args.only_posv_ir:
        S, _ = make_spd(xk, args.potrf_split_n, 1024)
        B = np.asfortranarray(np.random.default_rng(3).uniform(-0.5, 0.5, (S.shape[0], 64)))
        L.xkblas_set_param(args.potrf_nb, 0)
        for _ in range(2):
            L.xkblas_profile_reset()
            X, info, it = xk.posv_ir(122, S, B)
            print(json.dumps({"info": info, "iterations": it, "ms": xk.profile("posv_ir_f64")[1]}), file=sys.stderr,
                  flush=True)
        L.xkblas_finalize()
        return
    if args.only_batched:
        print(json.dumps(bench_batched(xk, parts=args.batched_parts)), file=sys.stderr)
        L.xkblas_finalize()
        return
    n = args.n
    print("bench: generating inputs n=%d" % n, file=sys.stderr, flush=True)
    A, B, C = make_inputs(n, 2026)
    flops = 2.0 * n**3

    if args.only_split:
        time_dgemm(xk, A, B, C, 1, 1, args.only_split, args.nb)
        L.xkblas_finalize()
        return
    if args.only_fp64:
        time_dgemm(xk, A, B, C, 1, 1, 0, args.nb)
        L.xkblas_finalize()
        return

    # plain FP64 MFMA
    print("bench: fp64 dgemm", file=sys.stderr, flush=True)
    t_fp64 = time_dgemm(xk, A, B, C, args.steps, args.warmup, 0, args.nb)
    cnt, ms, _ = kernel_stats(xk, "gemm_f64")
    k_fp64 = ms / max(cnt, 1) / 1e3
    fp64_tflops = flops / t_fp64 / 1e12

    cfg = L.xkblas_get_configuration().decode()
    full_ref = None
    if not args.no_full_check:
        print("bench: host-BLAS reference product (full normwise check)", file=sys.stderr, flush=True)
        full_ref = host_reference(A, B)
    errs64 = accuracy(xk, A, B, 0, args.nb, full_ref=full_ref)

    # split modes (FP64-accurate): BF16 (the headline) and FP8 residue products
    def split_run(mode, cls, label, peak):
        print("bench: split-%s dgemm" % label, file=sys.stderr, flush=True)
        errs = accuracy(xk, A, B, mode, args.nb, full_ref=full_ref)
        t = time_dgemm(xk, A, B, C, args.steps, args.warmup, mode, args.nb)
        # kernel statistics of the timed resident steps only (the host-to-host
        # run below launches per row panel)
        c_mg, ms_mg, fl_mg = kernel_stats(xk, cls)
        c_pr, ms_pr, _ = kernel_stats(xk, "split_prep")
        c_cr, ms_cr, _ = kernel_stats(xk, "split_crt")
        h2h = host_to_host(xk, A, B, C, mode, args.nb)
        L.xkblas_set_modemath(0)
        nm, npl = ctypes.c_int(), ctypes.c_int()
        bits = L.xkblas_split_info(mode, n, L.xkblas_get_split_count(), ctypes.byref(nm), ctypes.byref(npl))
        return t, fl_mg / ms_mg / 1e9, ms_mg / max(c_mg, 1), {
            "tflops": round(flops / t / 1e12, 3), "ms_per_step": round(t * 1e3, 3),
            "speedup_vs_fp64_mfma": round(t_fp64 / t, 3),
            "moduli": nm.value, "residue_gemms": npl.value, "bits_per_side": bits,
            label + "_gemm_tflops": round(fl_mg / ms_mg / 1e9, 1),
            label + "_gemm_frac_of_peak": round(fl_mg / ms_mg / 1e9 / peak, 4),
            label + "_gemm_ms": round(ms_mg / max(c_mg, 1), 3),
            "prep_ms": round(ms_pr / max(c_pr, 1), 3), "crt_ms": round(ms_cr / max(c_cr, 1), 3),
            "host_to_host_ms": round(h2h * 1e3, 3), "host_to_host_tflops": round(flops / h2h / 1e12, 3),
            "rel_err": errs}

    t_split, bf16_achieved, bf16_ms, split = split_run(2, "split_mod_gemm", "bf16", BF16_PEAK)
    split["fp64_mfma_rel_err"] = errs64
    split_fp8 = None
    if not args.no_fp8:
        split_fp8 = split_run(3, "split_mod_gemm_fp8", "fp8", FP8_PEAK)[3]

    batched = None
    if not args.no_batched:
        print("bench: batched uneven gemm/trsm", file=sys.stderr, flush=True)
        batched = bench_batched(xk)

    potrf = None
    if not args.no_potrf:
        del A, B, C
        print("bench: dpotrf", file=sys.stderr, flush=True)
        S, syrk_rate = make_spd(xk, args.potrf_n, 1024)
        potrf = {"fp64_mfma": bench_potrf(xk, S, 1024, 0, syrk_rate)[0]}
        del S
        if args.potrf_split_n:
            # block order NB of both modes picked by the trained cost model
            nbig = args.potrf_split_n
            nb64 = nbsp = 1024
            picked = "default"
            from xkblas_amd import tuner
            model = tuner.load_trained()
            if model is not None:
                nb64 = tuner.choose(model, "potrf", 0, 0, nbig, nbig, nbig)[0]["nb"]
                nbsp = tuner.choose(model, "potrf", 0, 2, nbig, nbig, nbig)[0]["nb"]
                picked = "cost model (xkblas_amd/tuning)"
                model.close()
            print("bench: dpotrf n=%d fp64 (nb %d) and split-bf16 (nb %d)" % (nbig, nb64, nbsp), file=sys.stderr,
                  flush=True)
            S, syrk_rate = make_spd(xk, nbig, 1024)
            p64, F64 = bench_potrf(xk, S, nb64, 0, syrk_rate)
            psp, Fsp = bench_potrf(xk, S, nbsp, 2)
            p64["tile_picked_by"] = psp["tile_picked_by"] = picked
            psp["max_rel_diff_vs_fp64_factor"] = tril_rel_diff(Fsp, F64)
            potrf["large_fp64_mfma"] = p64
            potrf["large_split_bf16"] = psp
            potrf["large_split_speedup"] = round(p64["kernel_ms"] / psp["kernel_ms"], 3)
            del Fsp
            print("bench: dposv_ir n=%d nrhs 64" % nbig, file=sys.stderr, flush=True)
            potrf["posv_ir"] = bench_posv_ir(xk, S, F64, nbsp, p64["kernel_ms"], psp["kernel_ms"])
            del S, F64

    xk.invalidate()
    print("bench: dgemm 2048 tile 512 (the CPU-runnable configuration)", file=sys.stderr, flush=True)
    c1 = bench_c1(xk)
    print("bench: dgemmt/dsyr2k/dsymm/dtrmm/zgemm 8192", file=sys.stderr, flush=True)
    l3_more = bench_l3_more(xk)
    print("bench: zherk/ztrsm/zpotrf 4096", file=sys.stderr, flush=True)
    cplx = bench_complex(xk)
    cpu = None if args.no_cpu else cpu_baseline()

    value = split["tflops"]
    out = {
        "metric": "FP64-accurate DGEMM/DPOTRF TFLOP/s, split-BF16 vs plain FP64 MFMA, 1 GPU",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_split * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "dgemm_nn_split_bf16_fp64_accurate", "m": n, "n": n, "k": n, "alpha": 1.5,
                   "beta": -0.5,
                   "tile": args.nb, "inputs": "U[-0.5,0.5], rows of A / cols of B scaled 2^u, u~U[-8,8]",
                   "resident_in_hbm": True},
        "fp64_mfma_tflops": round(fp64_tflops, 3),
        "fp64_mfma_kernel": {"achieved": round(flops / k_fp64 / 1e12, 3), "peak": FP64_PEAK,
                             "frac": round(flops / k_fp64 / 1e12 / FP64_PEAK, 4)},
        "split_bf16": split,
        "split_fp8": split_fp8,
        "dgemm_2048_tile512": c1,
        "l3_more_fp64": l3_more,
        "complex_fp64": cplx,
        "dpotrf": potrf,
        "batched": batched,
        "roofline": {"bound": "mfma", "achieved": round(bf16_achieved, 1), "peak": BF16_PEAK, "unit": "TFLOP/s",
                     "frac": round(bf16_achieved / BF16_PEAK, 4),
                     "traffic": profiled_traffic() if n == 16384 else None,
                     "traffic_source": "profiles/split_dgemm16384_hbm_traffic.md (rocprofv3 FETCH_SIZE/WRITE_SIZE)",
                     # A and B planes read once (2 bytes each), int16 residues written
                     "traffic_algorithmic": split["residue_gemms"] * (2 * 2 * n * n + 2 * n * n),
                     "kernel": "mod_gemm_w8_kernel (BF16 residue GEMM, 256x256 tiles, 8 waves, all moduli in one "
                               "launch)",
                     "kernel_ms": round(bf16_ms, 3),
                     "fp64_equivalent_frac_of_fp64_peak": round(flops / t_split / 1e12 / FP64_PEAK, 4)},
        "cpu_baseline": cpu,
        "accuracy_target": {
            "bound": 1e-14,
            "dgemm": "split_bf16.rel_err.normwise_sampled: ||C - C_dd||_2 / ||C_dd||_2 over sampled entries, "
                     "C_dd from double-double dot products",
            "dpotrf_factor": "dpotrf.large_split_bf16.max_rel_diff_vs_fp64_factor: max|L - L_fp64| / max|L_fp64|",
            "dpotrf_solve": "dpotrf.posv_ir.rel_err_vs_refined_solution: max|x - x*| / max|x*| per sampled "
                            "right-hand side, x* the FP64 route's solution refined with an extended-precision "
                            "residual (the FP64 route's own error beside it; max_rel_diff_vs_fp64_route = "
                            "max|X - X_fp64| / max|X_fp64| mixes both routes' errors)"},
        "library": cfg,
    }
    L.xkblas_finalize()
    print(json.dumps(out))


def bench_batched(xk, seed=2027, reps=5, parts=False):
    """Batched uneven work: 4096 GEMMs (m, n, k log-uniform in [16, 512]) and
    1024 lower-triangular TRSMs (n log-uniform in [32, 1024], 64 right-hand
    sides, diagonally dominant L), all FP64 or all FP32, in one persistent
    launch replayed from a hipGraph.  The shapes are the committed list
    tests/golden/batched_c4_shapes.json; entries are drawn from `seed`.
    Operands live in one device pool.  The
    baseline runs the same problems as one launch each.  A few problems are
    checked against NumPy."""
    L = xk.lib()
    out = {"workload": "4096 gemm m,n,k~logU[16,512] + 1024 trsm lower n~logU[32,1024] nrhs 64"}
    with open(os.path.join(ROOT, "tests", "golden", "batched_c4_shapes.json")) as f:
        shapes = json.load(f)
    g = np.array(shapes["gemm_mnk"], dtype=int)
    t = np.array(shapes["trsm_n"], dtype=int)
    for dt, name, code in ((np.float64, "fp64", 0), (np.float32, "fp32", 1)):
        r = np.random.default_rng(seed)
        es = np.dtype(dt).itemsize
        sizes = [m * k + k * n + m * n for m, n, k in g] + [n * n + n * 64 for n in t]
        offs = np.concatenate([[0], np.cumsum([(x + 63) // 64 * 64 for x in sizes])])
        pool = L.xkblas_device_malloc(int(offs[-1]) * es)
        assert pool
        host = np.empty(int(offs[-1]), dtype=dt)
        probs, flops, checks = [], 0.0, []
        for i, (m, n, k) in enumerate(g):
            o = int(offs[i])
            A = r.uniform(-0.5, 0.5, m * k).astype(dt)
            B = r.uniform(-0.5, 0.5, k * n).astype(dt)
            C = r.uniform(-0.5, 0.5, m * n).astype(dt)
            host[o:o + m * k], host[o + m * k:o + m * k + k * n] = A, B
            host[o + m * k + k * n:o + m * k + k * n + m * n] = C
            m, n, k = int(m), int(n), int(k)
            base = int(pool) + o * es
            probs.append(xk.BatchProblem(xk.BATCH_GEMM, code, 111, 111, m, n, k, 1.5, -0.5,
                                         base, m, base + m * k * es, k, base + (m * k + k * n) * es, m))
            flops += 2.0 * m * n * k
            if i % 1024 == 0:
                ref = 1.5 * A.reshape(k, m).T.astype(np.float64) @ B.reshape(n, k).T.astype(np.float64) \
                    - 0.5 * C.reshape(n, m).T.astype(np.float64)
                checks.append((o + m * k + k * n, m, n, ref))
        for j, n in enumerate(t):
            n = int(n)
            o = int(offs[4096 + j])
            Lm = np.tril(r.uniform(-1, 1, (n, n)))
            np.fill_diagonal(Lm, 1.0 + np.abs(Lm).sum(axis=1))
            Bm = r.uniform(-1, 1, (n, 64))
            host[o:o + n * n] = Lm.T.reshape(-1).astype(dt)  # column-major
            host[o + n * n:o + n * n + n * 64] = Bm.T.reshape(-1).astype(dt)
            base = int(pool) + o * es
            probs.append(xk.BatchProblem(xk.BATCH_TRSM, code, 111, 111, n, 64, n, 1.0, 0.0,
                                         base, n, None, 0, base + n * n * es, n))
            flops += float(n) * n * 64
            if j % 256 == 0:
                checks.append((o + n * n, n, 64, np.linalg.solve(Lm, Bm)))
        arr = (xk.BatchProblem * len(probs))(*probs)
        p = ctypes.c_void_p

        def upload():
            assert L.xkblas_memcpy_htod(pool, host.ctypes.data_as(p), host.nbytes) == 0

        upload()
        # GEMM item shapes (xkblas_set_batch_tile): per problem, 32 x 32 only,
        # 64 x 64 only (the default plan below picks per precision: per problem
        # for FP64, 64 x 64 for FP32); each timed as a graph-replayed launch
        mode_ms = {}
        for mode in (1, 2, 3):
            L.xkblas_set_batch_tile(mode)
            mp = L.xkblas_batch_create(ctypes.cast(arr, p), len(arr))
            assert mp and L.xkblas_batch_use_graph(mp, 1) == 0
            L.xkblas_batch_run(mp)
            L.xkblas_sync()
            tm = []
            for _ in range(reps):
                upload()
                L.xkblas_sync()
                L.xkblas_profile_reset()
                L.xkblas_batch_run(mp)
                L.xkblas_sync()
                tm.append(xk.profile("batch")[1])
            mode_ms[{3: "tiles_64", 1: "tiles_per_problem", 2: "tiles_32"}[mode]] = round(min(tm), 3)
            L.xkblas_batch_use_graph(mp, 0)
            L.xkblas_batch_destroy(mp)
        L.xkblas_set_batch_tile(0)
        upload()
        plan = L.xkblas_batch_create(ctypes.cast(arr, p), len(arr))  # default mode 0
        assert plan
        nit, grid, pfl, pby = ctypes.c_int(), ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        L.xkblas_batch_info(plan, ctypes.byref(nit), ctypes.byref(grid), ctypes.byref(pfl), ctypes.byref(pby))
        assert L.xkblas_batch_use_graph(plan, 1) == 0
        L.xkblas_batch_run(plan)
        L.xkblas_sync()
        ms = []
        for _ in range(reps):
            upload()
            L.xkblas_sync()
            L.xkblas_profile_reset()
            L.xkblas_batch_run(plan)
            L.xkblas_sync()
            ms.append(xk.profile("batch")[1])
        res = np.empty(int(offs[-1]), dtype=dt)
        assert L.xkblas_memcpy_dtoh(res.ctypes.data_as(p), pool, res.nbytes) == 0
        part_ms = {}
        if parts:
            # the GEMM problems alone and the TRSM problems alone, same launch
            for part, lo, hi in (("gemm_only", 0, len(g)), ("trsm_only", len(g), len(arr))):
                sub = (xk.BatchProblem * (hi - lo))(*arr[lo:hi])
                sp = L.xkblas_batch_create(ctypes.cast(sub, p), hi - lo)
                assert sp and L.xkblas_batch_use_graph(sp, 1) == 0
                L.xkblas_batch_run(sp)
                L.xkblas_sync()
                part_t = []
                for _ in range(reps):
                    upload()
                    L.xkblas_sync()
                    L.xkblas_profile_reset()
                    L.xkblas_batch_run(sp)
                    L.xkblas_sync()
                    part_t.append(xk.profile("batch")[1])
                part_ms[part + "_ms"] = round(min(part_t), 3)
                L.xkblas_batch_use_graph(sp, 0)
                L.xkblas_batch_destroy(sp)
        err = 0.0
        for o, m, n, ref in checks:
            got = res[o:o + m * n].reshape(n, m).T.astype(np.float64)
            err = max(err, float(np.abs(got - ref).max() / max(1.0, np.abs(ref).max())))
        L.xkblas_batch_use_graph(plan, 0)
        L.xkblas_batch_destroy(plan)
        # baseline: one launch per problem
        upload()
        base_addr, psz = ctypes.addressof(arr), ctypes.sizeof(xk.BatchProblem)
        plans = [L.xkblas_batch_create(p(base_addr + i * psz), 1) for i in range(len(arr))]
        L.xkblas_sync()
        t0 = time.perf_counter()
        for pl in plans:
            L.xkblas_batch_run(pl)
        L.xkblas_sync()
        t_each = time.perf_counter() - t0
        for pl in plans:
            L.xkblas_batch_destroy(pl)
        L.xkblas_device_free(pool)
        best = min(ms)
        # roofline of the batch: its flops at the MFMA peak of the type or its
        # algorithmic bytes (operands read once, C read and written) at HBM
        # bandwidth, whichever is longer
        peak = FP64_PEAK if code == 0 else FP32_PEAK
        t_flop, t_byte = pfl.value / (peak * 1e9), pby.value / (HBM_PEAK * 1e6)  # ms
        # GEMM tiles of the default plan (FP64: 64 x 64, or 32 x 32 where that
        # covers <= 3/4 of the area; FP32: 64 x 64) with k in slices of 16:
        # the MFMA work the GEMM items execute
        def padded(m, n):
            a64 = -(-m // 64) * -(-n // 64) * 4096
            a32 = -(-m // 32) * -(-n // 32) * 1024
            return a32 if code == 0 and 4 * a32 <= 3 * a64 else a64
        gemm_exec = float(sum(2.0 * padded(m, n) * ((k + 15) // 16 * 16) for m, n, k in g))
        gemm_alg = float(sum(2.0 * m * n * k for m, n, k in g))
        out[name] = {"ms": round(best, 3), "gflops": round(flops / best / 1e6, 1), "problems": len(arr),
                     "work_items": nit.value, "workgroups": grid.value, "gemm_item_shapes_ms": mode_ms,
                     "roofline": {"bound": "mfma" if t_flop >= t_byte else "hbm",
                                  "flops": pfl.value, "bytes": pby.value,
                                  "roofline_ms": round(max(t_flop, t_byte), 4),
                                  "frac": round(max(t_flop, t_byte) / best, 4),
                                  "gemm_executed_over_algorithmic_flops": round(gemm_exec / gemm_alg, 3)},
                     "one_launch_per_problem_ms": round(t_each * 1e3, 3),
                     "speedup_vs_one_launch_per_problem": round(t_each * 1e3 / best, 2),
                     "max_rel_err_sampled": err, **part_ms}
    return out


def make_spd(xk, n, nb):
    """A = G G^T + n I (lower triangle), G ~ U[-1,1]; G G^T formed on the GPU
    by DSYRK (the DSYRK + DPOTRF workload).  Returns A and the DSYRK rate."""
    L = xk.lib()
    r = np.random.default_rng(7)
    G = np.asfortranarray(r.uniform(-1.0, 1.0, (n, n)))
    S = np.zeros((n, n), order="F")
    L.xkblas_set_modemath(0)
    L.xkblas_set_param(nb, 0)
    L.xkblas_profile_reset()
    xk.syrk(122, 111, 1.0, G, 0.0, S, sync=False)
    L.xkblas_sync()
    cs, ms_syrk, fl_syrk = xk.profile("syrk_f64")
    xk.coherent(S, 122)
    S[np.diag_indices(n)] += n
    xk.invalidate()
    return S, fl_syrk / ms_syrk / 1e9


def tril_rel_diff(X, Y):
    """max |X - Y| / max |Y| over the lower triangles, in column blocks."""
    n = X.shape[0]
    d = m = 0.0
    for j0 in range(0, n, 2048):
        Xb, Yb = np.tril(X[:, j0:j0 + 2048], -j0), np.tril(Y[:, j0:j0 + 2048], -j0)
        d = max(d, float(np.abs(Xb - Yb).max()))
        m = max(m, float(np.abs(Yb).max()))
    return d / m


def bench_potrf(xk, S, nb, mode, dsyrk_tflops=None, samples=96):
    """DPOTRF(lower) of S in math mode `mode`.  Each factorization starts from
    a fresh host copy (its host-to-device copy is outside the timed kernel
    window); the rate uses HIP-event kernel time.  The result's error is the
    sampled residual max |S_ij - (L L^T)_ij| / (||L_i|| ||L_j||) over random
    lower entries.  Returns (stats, factor)."""
    L = xk.lib()
    n = S.shape[0]
    L.xkblas_set_modemath(mode)
    L.xkblas_set_param(nb, 0)
    best, F = None, None
    for _ in range(2):
        F = S.copy(order="F")
        L.xkblas_profile_reset()
        rc = L.xkblas_dpotrf_async(122, n, F.ctypes.data_as(ctypes.c_void_p), n)
        assert rc == 0
        L.xkblas_sync()
        cnt, ms, fl = xk.profile("potrf_f64")
        xk.coherent(F, 122)
        info = L.xkblas_potrf_info(F.ctypes.data_as(ctypes.c_void_p))
        xk.invalidate()
        best = ms if best is None else min(best, ms)
    L.xkblas_set_modemath(0)
    r = np.random.default_rng(5)
    i = r.integers(0, n, samples)
    j = r.integers(0, n, samples)
    i, j = np.maximum(i, j), np.minimum(i, j)
    res = 0.0
    for a, b in zip(i, j):
        la, lb = F[a, :a + 1], F[b, :b + 1]
        dot = la[:b + 1] @ lb
        res = max(res, abs(S[a, b] - dot) / (np.linalg.norm(la) * np.linalg.norm(lb)))
    fl = n**3 / 3.0 + n**2 / 2.0 + n / 6.0
    out = {"n": n, "tile": nb, "mode": {0: "fp64_mfma", 2: "split_bf16", 3: "split_fp8"}[mode], "info": info,
           "tflops": round(fl / best / 1e9, 3), "kernel_ms": round(best, 3),
           "sampled_residual": float(res), "residual_samples": samples}
    if dsyrk_tflops is not None:
        out["dsyrk_tflops"] = round(dsyrk_tflops, 3)
    return out, F


def bench_l3_more(xk, n=8192):
    """DGEMMT (lower, NT), DSYR2K (lower, N), DSYMM (left, lower), DTRMM
    (left, lower, N, non-unit) and ZGEMM (N, C) of order n in FP64 MFMA,
    operands resident in HBM: the second of two calls, HIP-event kernel time.
    TFLOP/s from the textbook flop counts (triangle-only for GEMMT/SYR2K, m^2 n
    for TRMM, 8 n^3 real flop for ZGEMM)."""
    L = xk.lib()
    L.xkblas_set_modemath(0)
    L.xkblas_set_param(0, 0)
    r = np.random.default_rng(17)
    A, B, C = (np.asfortranarray(r.uniform(-0.5, 0.5, (n, n))) for _ in range(3))
    ZA, ZB, ZC = (np.asfortranarray(X + 0.5j * Y) for X, Y in ((A, B), (B, C), (C, A)))
    runs = {
        "dgemmt": ("gemmt_f64", float(n) * n * (n + 1), lambda: xk.gemmt(122, 111, 112, 1.0, A, B, 0.5, C, sync=False)),
        "dsyr2k": ("syr2k_f64", 2.0 * n * n * (n + 1), lambda: xk.syr2k(122, 111, 1.0, A, B, 0.5, C, sync=False)),
        "dsymm": ("symm_f64", 2.0 * n ** 3, lambda: xk.symm(141, 122, 1.0, A, B, 0.5, C, sync=False)),
        "dtrmm": ("trmm_f64", float(n) ** 3, lambda: xk.trmm(141, 122, 111, 131, 1.0, A, C, sync=False)),
        "zgemm": ("gemm_c128", 8.0 * n ** 3, lambda: xk.gemm(111, 113, 1.0 - 0.5j, ZA, ZB, 0.5j, ZC, sync=False)),
    }
    out = {"n": n}
    for name, (cls, fl, call) in runs.items():
        best = None
        for _ in range(2):
            L.xkblas_profile_reset()
            call()
            L.xkblas_sync()
            ms = xk.profile(cls)[1]
            best = ms if best is None else min(best, ms)
        out[name] = {"ms": round(best, 3), "tflops": round(fl / best / 1e9, 3),
                     "frac_of_fp64_peak": round(fl / best / 1e9 / FP64_PEAK, 4)}
    xk.invalidate()
    return out


def bench_complex(xk, n=4096):
    """ZHERK (lower, N), ZTRSM (left, lower, N, n RHS) and ZPOTRF (lower) of
    order n, operands resident (ZPOTRF: each factorization from a fresh host
    copy, its transfer outside the event window): HIP-event time, one untimed
    call, then the median of three (as the other legs: a warm-up, then timed
    steps; the call after the first still runs ~5 % slow).  TFLOP/s from real
    flop counts, 4x those of the real routines."""
    L = xk.lib()
    L.xkblas_set_modemath(0)
    L.xkblas_set_param(0, 0)
    r = np.random.default_rng(19)
    G = np.asfortranarray(r.uniform(-0.5, 0.5, (n, n)) + 1j * r.uniform(-0.5, 0.5, (n, n)))
    H = np.asfortranarray(G @ G.conj().T + n * np.eye(n))
    Tl = np.asfortranarray(np.tril(G) + n * np.eye(n))
    X = np.asfortranarray(G.copy())
    C = np.zeros((n, n), dtype=np.complex128, order="F")
    out = {"n": n}
    runs = {
        "zherk": ("herk_c128", 4.0 * n * n * (n + 1.0), lambda: xk.herk(122, 111, 1.0, G, 0.0, C, sync=False)),
        "ztrsm": ("trsm_c", 4.0 * n ** 3, lambda: xk.trsm(141, 122, 111, 131, 1.0, Tl, X, sync=False)),
    }
    for name, (cls, fl, call) in runs.items():
        times = []
        for _ in range(4):
            L.xkblas_profile_reset()
            call()
            L.xkblas_sync()
            times.append(xk.profile(cls)[1])
        ms = float(np.median(times[1:]))
        out[name] = {"ms": round(ms, 3), "tflops": round(fl / ms / 1e9, 3)}
    times = []
    for _ in range(4):
        F = H.copy(order="F")
        L.xkblas_profile_reset()
        xk.potrf(122, F, sync=False)
        L.xkblas_sync()
        times.append(xk.profile("potrf_c")[1])
        xk.invalidate()
    ms = float(np.median(times[1:]))
    fl = 4.0 * (n ** 3 / 3.0 + n * n / 2.0)
    out["zpotrf"] = {"ms": round(ms, 3), "tflops": round(fl / ms / 1e9, 3)}
    xk.invalidate()
    return out


def sampled_backward_error(S, X, B, samples=96, seed=11):
    """Componentwise (Oettli-Prager) backward error of X for A X = B over
    sampled rows: max |B - A X|_ij / (|A| |X|)_ij, A symmetric with its lower
    triangle in S."""
    n = S.shape[0]
    rows = np.random.default_rng(seed).integers(0, n, samples)
    worst = 0.0
    for i in rows:
        a = np.concatenate([S[i, :i + 1], S[i + 1:, i]])
        r = np.abs(B[i] - a @ X)
        d = np.abs(a) @ np.abs(X)
        worst = max(worst, float((r / d).max()))
    return worst


def refined_columns(S, F, B, Y, cols, block=1024):
    """Reference solutions of A x = b for the columns `cols` of B: Y's
    columns refined once with the residual b - A y evaluated in 80-bit long
    double (error ~ 2^-64 sqrt(n) |A| |y|) and the correction solved with the
    FP64 factor F (lower).  A is symmetric with its lower triangle in S."""
    from scipy.linalg import solve_triangular
    n = S.shape[0]
    y = Y[:, cols].astype(np.longdouble)
    r = B[:, cols].astype(np.longdouble)
    for i0 in range(0, n, block):
        i1 = min(n, i0 + block)
        ii = np.arange(i0, i1)[:, None]
        jj = np.arange(n)[None, :]
        rows = np.where(jj <= ii, S[i0:i1, :], S[:, i0:i1].T)  # A[i0:i1, :]
        r[i0:i1] -= rows.astype(np.longdouble) @ y
    d = solve_triangular(F, np.asarray(r, dtype=np.float64), lower=True, check_finite=False)
    d = solve_triangular(F, d, lower=True, trans=1, check_finite=False)
    return np.asarray(y + d.astype(np.longdouble), dtype=np.float64)


def bench_posv_ir(xk, S, F64, nb, p64_ms, psp_ms, nrhs=64, moduli=7):
    """xkblas_dposv_ir (split-BF16 factor with `moduli` CRT moduli, FP64
    residuals and corrections) on A = S (lower), B ~ U[-0.5,0.5] n x nrhs,
    against the FP64 route (DPOTRF kernel time p64_ms + two DTRSM with the FP64
    factor F64) and the accurate split route (psp_ms + the same DTRSMs).  The
    call is synchronous; its time is the HIP-event window around it with A
    already resident in HBM (second call)."""
    L = xk.lib()
    n = S.shape[0]
    B = np.asfortranarray(np.random.default_rng(3).uniform(-0.5, 0.5, (n, nrhs)))
    L.xkblas_set_modemath(0)
    L.xkblas_set_param(nb, 0)
    L.xkblas_set_posv_ir(moduli, 30)
    best = None
    for _ in range(2):
        L.xkblas_profile_reset()
        X, info, it = xk.posv_ir(122, S, B)
        ms = xk.profile("posv_ir_f64")[1]
        best = ms if best is None else min(best, ms)
    L.xkblas_set_posv_ir(0, -1)
    xk.invalidate()
    Y = B.copy(order="F")
    one = np.array([1.0])
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    L.xkblas_set_param(1024, 0)
    L.xkblas_profile_reset()
    for ta in (111, 112):
        L.xkblas_dtrsm_async(141, 122, ta, 131, n, nrhs, vp(one), vp(F64), n, vp(Y), n)
    L.xkblas_sync()
    trsm_ms = xk.profile("trsm_f64")[1]
    xk.coherent(Y)
    xk.invalidate()
    # forward errors of both routes on two sampled right-hand sides, against
    # the FP64 route's solution refined in extended precision
    cols = np.random.default_rng(17).choice(nrhs, 2, replace=False)
    Xs = refined_columns(S, F64, B, Y, cols)
    fwd = lambda Z: float(max(np.abs(Z[:, c] - Xs[:, q]).max() / np.abs(Xs[:, q]).max() for q, c in enumerate(cols)))
    fl = n**3 / 3.0 + 2.0 * n * n * nrhs
    return {"n": n, "nrhs": nrhs, "tile": nb, "factor": "split_bf16, %d moduli" % moduli, "info": info,
            "iterations": it, "ms": round(best, 3), "tflops_equivalent": round(fl / best / 1e9, 3),
            "fp64_route_ms": round(p64_ms + trsm_ms, 3), "split_accurate_route_ms": round(psp_ms + trsm_ms, 3),
            "two_dtrsm_ms": round(trsm_ms, 3), "speedup_vs_fp64_route": round((p64_ms + trsm_ms) / best, 3),
            "speedup_vs_split_accurate_route": round((psp_ms + trsm_ms) / best, 3),
            "sampled_backward_error": sampled_backward_error(S, X, B),
            "fp64_route_sampled_backward_error": sampled_backward_error(S, Y, B),
            "max_rel_diff_vs_fp64_route": float(np.abs(X - Y).max() / np.abs(Y).max()),
            "rel_err_vs_refined_solution": fwd(X), "fp64_route_rel_err_vs_refined_solution": fwd(Y),
            "refined_solution": "2 sampled right-hand sides: the FP64 route's x plus one correction from the "
                                "residual b - A x in 80-bit long double on the host"}


if __name__ == "__main__":
    main()
