"""Several host threads calling the library at once, each on its own
buffers: GEMM (+ coherent + sync), POTRF and SYRK per thread, every result
checked against the oracle, and each thread ending with
xkblas_memory_invalidate_caches.  The library has one context per process
(one compute stream, one tile cache, one lock around each API call): calls
from different threads are serialised in the order they take the lock, and
a thread's xkblas_sync waits for all enqueued work, its own included.  As in
the reference (blas/xkblas.c:1855-1908), invalidate_caches drops only the
calling thread's mirrors: a product the main thread enqueued before the
workers started is copied back after they have all invalidated.  ctypes
releases the GIL around foreign calls, so the threads do enter the library
concurrently."""
import threading

import numpy as np
import pytest

from helpers import rng, uniform, spd

pytestmark = pytest.mark.gpu


def worker(xk, orc, tid, errors):
    try:
        r = rng(900 + tid)
        for it in range(3):
            M, N, K = (int(x) for x in r.integers(100, 700, 3))
            A, B, C0 = uniform(r, M, K), uniform(r, K, N), uniform(r, M, N)
            C = C0.copy(order="F")
            xk.gemm(111, 111, 1.5, A, B, -0.5, C)
            Cref = orc.dgemm_tiled(111, 111, 1.5, A, B, -0.5, C0.copy(order="F"), 128)
            res = orc.gemm_residual(1.5, A, B, -0.5, C0, Cref, C)
            if not res <= 3.0:
                errors.append("thread %d gemm %d: residual %g" % (tid, it, res))
            n = int(r.integers(200, 600))
            S = spd(r, n)
            F = S.copy(order="F")
            info = xk.potrf(122, F)
            if info != 0 or not orc.potrf_residual(S, F, 122) <= 10.0:
                errors.append("thread %d potrf %d: info %d" % (tid, it, info))
            G = uniform(r, n, 150)
            W0 = orc.plgsy(0.0, n, 7 + tid)
            W = W0.copy(order="F")
            xk.syrk(121, 111, 0.5, G, 0.25, W)
            Wref = orc.dsyrk_tiled(121, 111, 0.5, G, 0.25, W0.copy(order="F"), 128)
            if not orc.syrk_residual(G, W0, Wref, W, 121) <= 10.0:
                errors.append("thread %d syrk %d" % (tid, it))
            xk.invalidate()
    except Exception as e:  # reported by the main thread
        errors.append("thread %d: %r" % (tid, e))


def test_threads_own_buffers(xk, orc):
    L = xk.lib()
    L.xkblas_set_param(128, 0)
    r = rng(899)
    A, B, C0 = uniform(r, 500, 300), uniform(r, 300, 400), uniform(r, 500, 400)
    C = C0.copy(order="F")
    xk.gemm(111, 111, 2.0, A, B, 0.5, C, sync=False)  # result still on the device
    errors = []
    threads = [threading.Thread(target=worker, args=(xk, orc, t, errors)) for t in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=240)
    assert not any(t.is_alive() for t in threads), "a thread did not finish"
    assert not errors, errors
    xk.coherent(C)
    Cref = orc.dgemm_tiled(111, 111, 2.0, A, B, 0.5, C0.copy(order="F"), 128)
    assert orc.gemm_residual(2.0, A, B, 0.5, C0, Cref, C) <= 3.0
    xk.invalidate()
