"""The f77 drop-in (libxkblas_amd_wrapper.so): threshold dispatch between the
host BLAS and the GPU, checked against the oracle.  Runs in a child process
so the thresholds (read when the library loads) are set per case."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import ctypes, json, sys
sys.path.insert(0, %(root)r)
sys.path.insert(0, %(root)r + "/tests")
import numpy as np, oracle
from helpers import rng, uniform, spd
W = ctypes.CDLL(%(root)r + "/xkblas_amd/libxkblas_amd_wrapper.so")
i = lambda v: ctypes.byref(ctypes.c_int(v))
d = lambda v: ctypes.byref(ctypes.c_double(v))
p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
r = rng(3)
M, N, K = 300, 200, 150
A, B, C0 = uniform(r, M, K), uniform(r, K, N), uniform(r, M, N)
C = C0.copy(order="F")
W.dgemm_(b"N", b"N", i(M), i(N), i(K), d(1.5), p(A), i(M), p(B), i(K), d(-0.5), p(C), i(M))
Cref = oracle.dgemm_tiled(111, 111, 1.5, A, B, -0.5, C0.copy(order="F"), 64)
res_gemm = oracle.gemm_residual(1.5, A, B, -0.5, C0, Cref, C)
n = 257
S = spd(r, n)
F = S.copy(order="F")
info = ctypes.c_int(-7)
W.dpotrf_(b"L", i(n), p(F), i(n), ctypes.byref(info))
res_potrf = oracle.potrf_residual(S, F, 122)
print(json.dumps({"gemm": res_gemm, "potrf": res_potrf, "info": info.value}))
'''


@pytest.mark.parametrize("threshold", ["0", "1000000000"])
def test_wrapper_dispatch(threshold):
    env = dict(os.environ, XKBLAS_THRESHOLD=threshold)
    out = subprocess.run([sys.executable, "-c", CHILD % {"root": ROOT}], capture_output=True, text=True, env=env,
                         timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    res = json.loads(out.stdout.strip().splitlines()[-1])
    assert res["info"] == 0
    assert res["gemm"] <= 3.0
    assert res["potrf"] <= 10.0
