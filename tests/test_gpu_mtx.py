"""Real-matrix ingest on the GPU: a Matrix Market file read by workloads/mtx.cpp
drives the SpMV path (checked against the oracle), and the plain C caller
apps/mtx_roofline.c (read, optimize, SpMV on device vectors beside the drop-in
call, SpMM on dense handles, each against the HBM roofline) runs on the
library."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from conftest import rel_err
from plsparse import workloads as W

pytestmark = pytest.mark.gpu

DRIVER = Path(__file__).resolve().parent.parent / "perf-libs-sparse_amd" / "lib" / "mtx_roofline"


def stencil_file(tmp_path, n=16, sym="general"):
    rp, ci, v = W.stencil27(n)
    m = n ** 3
    v = W.normal(len(v), 0x5EED0031)
    A = sp.csr_matrix((v, ci, rp), shape=(m, m))
    if sym == "symmetric":
        A = sp.tril(A + A.T)
    p = tmp_path / f"stencil_{sym}.mtx"
    scipy.io.mmwrite(str(p), A, symmetry=sym)
    return p, m


@pytest.mark.parametrize("sym", ["general", "symmetric"])
def test_spmv_on_read_matrix(pls, orc, tmp_path, sym):
    p, m = stencil_file(tmp_path, sym=sym)
    m2, n2, rp, ci, v = W.read_mtx(p)
    assert (m2, n2) == (m, m)
    x = W.uniform(m, 0x5EED0032)
    A = pls.SpMat.csr(m, m, rp, ci, v)
    A.hint(pls.HINT_SPMV_INVOCATIONS, pls.INVOCATIONS_MANY)
    A.optimize()
    y = np.zeros(m)
    A.spmv(x, y)
    assert rel_err(y, orc.spmv(m, rp, ci, v, x)) <= 1e-12


def _fields(line):
    head, *kv = line.split()
    return head, dict(f.split("=", 1) for f in kv)


def test_benchmark_driver(tmp_path):
    assert DRIVER.exists(), "build the driver with perf-libs-sparse_amd/build.py"
    p, m = stencil_file(tmp_path, n=12)
    print(f"running {DRIVER.name} on {p.name}", flush=True)
    r = subprocess.run([str(DRIVER), str(p), "1", "8", "33"], capture_output=True, text=True,
                       timeout=120)
    sys.stdout.write(r.stdout)
    sys.stderr.write(r.stderr)
    assert r.returncode == 0
    lines = [_fields(l) for l in r.stdout.splitlines()]
    nnz = len(W.stencil27(12)[1])
    assert lines[0] == ("matrix", {"m": str(m), "n": str(m), "nnz": str(nnz)})
    head, f = lines[1]
    assert head == "spmv" and f["variant"] in pls_variants()
    assert float(f["dropin_vs_device_rel"]) <= 1e-15  # same kernel on the same inputs
    head, f = lines[2]
    want = 12 * nnz + 4 * (m + 1) + 16 * m
    assert head == "spmv" and int(f["bytes"]) == want and float(f["ms"]) > 0
    assert abs(float(f["frac"]) - want / (float(f["ms"]) * 1e-3) / 8e12) <= 1e-3
    for (head, f), k in zip(lines[3:], (1, 8, 33)):
        assert head == "spmm" and int(f["k"]) == k and float(f["ms"]) > 0
        assert int(f["bytes"]) == 12 * nnz + 4 * (m + 1) + 16 * m * k
    assert len(lines) == 6


def pls_variants():
    import plsparse

    return plsparse.VARIANTS


def test_benchmark_driver_reports_reader_errors(tmp_path):
    p = tmp_path / "bad.mtx"
    p.write_text("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n")
    r = subprocess.run([str(DRIVER), str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "complex Matrix Market files not supported" in r.stderr
    assert r.stdout == ""
