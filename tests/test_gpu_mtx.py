"""Real-matrix ingest on the GPU: a Matrix Market file read by workloads/mtx.cpp
drives the SpMV path (checked against the oracle), and the C benchmark driver
apps/sparse_x_dense.c (the reference's bench/pl_bench_sparse_x_dense.c flow:
read, create A/B/C, spmm_optimize, 10 x spmm_exec) runs on the library."""
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

DRIVER = Path(__file__).resolve().parent.parent / "perf-libs-sparse_amd" / "lib" / "sparse_x_dense"


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


def test_benchmark_driver(tmp_path):
    assert DRIVER.exists(), "build the driver with perf-libs-sparse_amd/build.py"
    p, m = stencil_file(tmp_path, n=12)
    print(f"running {DRIVER.name} on {p.name}", flush=True)
    r = subprocess.run([str(DRIVER), str(p)], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    sys.stderr.write(r.stderr)
    assert r.returncode == 0
    lines = r.stdout.splitlines()
    nnz = len(W.stencil27(12)[1])
    assert lines[0] == f"A: {m} x {m} nnz={nnz}"
    assert lines[1] == f"B: {m} x 1024" and lines[2] == f"C: {m} x 1024"
    assert lines[3].startswith("iters=10 total_s=")
    assert r.stderr == "" or "error" not in r.stderr


def test_benchmark_driver_reports_reader_errors(tmp_path):
    p = tmp_path / "bad.mtx"
    p.write_text("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n")
    r = subprocess.run([str(DRIVER), str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "complex Matrix Market files not supported" in r.stderr
