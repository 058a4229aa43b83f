"""Trace output in the CSV schema of the reference's trace reader
(bin/ukilli.cpp:848-1113): exact headers, one kernel/task row per call,
copy rows whose sizes add up to the bytes the tile cache moved, and
consistent times (Start <= End, Duration = End - Start, rows in order)."""
import csv
import os

import numpy as np
import pytest

from helpers import rng, uniform, spd

pytestmark = pytest.mark.gpu

HEADERS = {
    "kern.csv": "Resource,Type,Stream,Push,Start,End,Duration,Name,TaskId",
    "task.csv": "Resource,Type,Push,Start,End,Duration,Name,TaskId",
    "cpy.csv": "Resource,Type,Stream,Push,Start,End,Duration,Size,Dir",
    "parallel.csv": "Resource,Type,Start,End,Duration,Name",
}


def test_trace_files(xk, tmp_path):
    L = xk.lib()
    xk.invalidate()
    assert L.xkblas_trace_start(str(tmp_path).encode()) == 0
    r = rng(53)
    A, B = uniform(r, 300, 200), uniform(r, 200, 250)
    C = np.zeros((300, 250), order="F")
    xk.gemm(111, 111, 1.0, A, B, 0.0, C)
    S = spd(r, 256)
    xk.potrf(122, S)
    xk.invalidate()
    assert L.xkblas_trace_stop() == 0
    rows = {}
    for name, header in HEADERS.items():
        with open(os.path.join(tmp_path, name)) as f:
            assert f.readline().strip() == header
            f.seek(0)
            rows[name] = list(csv.DictReader(f))
    kern = rows["kern.csv"]
    assert [k["Name"] for k in kern] == ["gemm_f64", "potrf_f64"]
    assert [k["Name"] for k in rows["task.csv"]] == ["gemm_f64", "potrf_f64"]
    for k in kern + rows["cpy.csv"]:
        t0, t1, d = float(k["Start"]), float(k["End"]), float(k["Duration"])
        assert 0.0 <= t0 <= t1 and abs(d - (t1 - t0)) < 1e-9
        assert float(k["Push"]) >= 0.0
    assert float(kern[0]["End"]) <= float(kern[1]["Start"]) + 1e-6
    cpy = rows["cpy.csv"]
    h2d = sum(int(c["Size"]) for c in cpy if c["Dir"] == "1")
    d2h = sum(int(c["Size"]) for c in cpy if c["Dir"] == "2")
    # A, B and the SPD matrix go up (C is write-only: beta 0); C and the factor come back
    assert h2d == (300 * 200 + 200 * 250 + 256 * 256) * 8
    assert d2h >= (300 * 250 + 256 * 257 // 2) * 8
    assert {c["Stream"] for c in cpy} <= {"2", "3"}
    par = rows["parallel.csv"]
    assert len(par) == 1 and float(par[0]["End"]) >= float(kern[-1]["End"]) - 1e-3
