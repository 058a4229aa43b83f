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
# dgemmt_, dsyr2k_, dsymm_, dtrmm_ against their textbook definitions
n, k = 260, 140
A, B, C0 = uniform(r, n, k), uniform(r, n, k), uniform(r, n, n)
nrm = lambda X: float(np.abs(X).max())
C = C0.copy(order="F")
W.dgemmt_(b"L", b"N", b"T", i(n), i(k), d(1.5), p(A), i(n), p(B), i(n), d(-0.5), p(C), i(n))
err = {"gemmt": nrm(np.tril(C - (1.5 * A @ B.T - 0.5 * C0))) / nrm(C0), "gemmt_other": nrm(np.triu(C - C0, 1))}
C = C0.copy(order="F")
W.dsyr2k_(b"U", b"N", i(n), i(k), d(0.5), p(A), i(n), p(B), i(n), d(2.0), p(C), i(n))
err["syr2k"] = nrm(np.triu(C - (0.5 * (A @ B.T + B @ A.T) + 2.0 * C0))) / nrm(C0)
S2 = np.asfortranarray(C0 + C0.T)
X = uniform(r, n, 70)
Y = np.zeros((n, 70), order="F")
W.dsymm_(b"L", b"L", i(n), i(70), d(1.0), p(np.asfortranarray(np.tril(S2))), i(n), p(X), i(n), d(0.0), p(Y), i(n))
err["symm"] = nrm(Y - S2 @ X) / nrm(S2 @ X)
T = np.asfortranarray(np.triu(C0))
Y = X.copy(order="F")
W.dtrmm_(b"L", b"U", b"T", b"N", i(n), i(70), d(2.0), p(T), i(n), p(Y), i(n))
err["trmm"] = nrm(Y - 2.0 * T.T @ X) / nrm(2.0 * T.T @ X)
# complex and single precision entries against their textbook definitions
zc = lambda m, n, dt=np.complex128: np.asfortranarray((r.uniform(-.5, .5, (m, n)) + 1j * r.uniform(-.5, .5, (m, n))).astype(dt))
za = lambda v, dt=np.complex128: np.array([v], dtype=dt).ctypes.data_as(ctypes.c_void_p)
ZA, ZB, ZC0 = zc(150, 90), zc(120, 90), zc(150, 120)
ZC = ZC0.copy(order="F")
W.zgemm_(b"N", b"C", i(150), i(120), i(90), za(1 - 1j), p(ZA), i(150), p(ZB), i(120), za(0.5j), p(ZC), i(150))
err["zgemm"] = nrm(ZC - ((1 - 1j) * ZA @ ZB.conj().T + 0.5j * ZC0)) / nrm(ZC0)
ZH0 = zc(150, 150)
ZH = ZH0.copy(order="F")
W.zherk_(b"L", b"N", i(150), i(90), d(1.0), p(ZA), i(150), d(0.5), p(ZH), i(150))
ref = ZA @ ZA.conj().T + 0.5 * ZH0
np.fill_diagonal(ref, ref.diagonal().real)
err["zherk"] = nrm(np.tril(ZH - ref)) / nrm(ref)
ZT = np.asfortranarray(np.triu(zc(150, 150)) + 3 * np.eye(150))
ZX = ZA.copy(order="F")
W.ztrsm_(b"L", b"U", b"C", b"N", i(150), i(90), za(2.0), p(ZT), i(150), p(ZX), i(150))
err["ztrsm"] = nrm(ZT.conj().T @ ZX - 2.0 * ZA) / nrm(ZA)
CA, CB, CC0 = zc(100, 80, np.complex64), zc(80, 60, np.complex64), zc(100, 60, np.complex64)
CC = CC0.copy(order="F")
W.cgemm_(b"N", b"N", i(100), i(60), i(80), za(1.0, np.complex64), p(CA), i(100), p(CB), i(80), za(1.0, np.complex64),
         p(CC), i(100))
err["cgemm_fp32"] = nrm(CC - (CA.astype(complex) @ CB.astype(complex) + CC0)) / nrm(CC0) * 1e-8
SA = np.asfortranarray(r.uniform(-.5, .5, (100, 50)).astype(np.float32))
SC0 = np.asfortranarray(r.uniform(-.5, .5, (100, 100)).astype(np.float32))
SC = SC0.copy(order="F")
fs = lambda v: ctypes.byref(ctypes.c_float(v))
W.ssyrk_(b"U", b"N", i(100), i(50), fs(1.0), p(SA), i(100), fs(1.0), p(SC), i(100))
err["ssyrk_fp32"] = nrm(np.triu(SC - (SA.astype(float) @ SA.T.astype(float) + SC0))) / nrm(SC0) * 1e-8
print(json.dumps({"gemm": res_gemm, "potrf": res_potrf, "info": info.value, "more": err}))
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
    assert res["more"]["gemmt_other"] == 0.0
    assert max(v for key, v in res["more"].items()) <= 1e-13, res["more"]
