"""Build libxkblas_amd.so (HIP, gfx950) in-tree.

Every translation unit under csrc/ is compiled by hipcc for gfx950 and linked
into xkblas_amd/libxkblas_amd.so.  `python -m xkblas_amd.build` or
`build(force=True)`; `force=False` recompiles only stale objects.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libxkblas_amd.so")
WRAPPER = os.path.join(HERE, "libxkblas_amd_wrapper.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: keep hipcc from emitting packed FP32 instructions.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + CSRC, "-I" + os.path.join(HERE, "..", "include")]

SOURCES = ["gemm.hip", "factor.hip", "split.hip", "batched.hip", "mlp.hip", "blas3.cpp", "runtime.cpp", "native.cpp", "tune.hip", "refine.hip", "complex.hip"]


def _compile(src, force):
    s = os.path.join(CSRC, src)
    o = os.path.join(OBJ, src + ".o")
    deps = [s] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    if not force and os.path.exists(o) and os.path.getmtime(o) >= max(os.path.getmtime(d) for d in deps):
        return o
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + FLAGS + lang + ["-c", s, "-o", o]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return o


def build(force=True, jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-ldl", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-6000:]))
    os.replace(tmp, LIB)
    # the f77 drop-in, a separate library linked against the main one
    wobj = _compile("wrapper.cpp", force)
    tmp = WRAPPER + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, wobj, "-L" + HERE, "-lxkblas_amd",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("wrapper link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-6000:]))
    os.replace(tmp, WRAPPER)
    return LIB


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
