"""Device products after in-place scaling and transposition of a matrix object.

Scaling and transposition change the host copy of an object and drop its
device plan; the next execution must run on the new values and shape, whatever
layout the plan had (CSR, SELL-C-sigma, column panels, BSR).  The reference's
own sequence (test/bsr_kernel_paths.c:87-132) is replayed from the fixture;
the rest is checked against numpy in FP64 to a normwise relative error of 1e-12.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def test_bsr_sequence_then_device_spmv(pls, golden):
    """Infinity norm 36, transpose, scale 0.5, export, then NOTRANS SpMV with x = 1."""
    g = golden["bsr_col_major_transposed_half"]
    A = pls.SpMat.bsr(g["layout"], 4, 4, 2, g["row_ptr"], g["col_indx"], np.array(g["vals"]))
    assert A.norm(pls.NORM_INF) == pytest.approx(g["inf_norm"], abs=g["tol"])
    A.transpose(pls.TRANS)
    A.scale(0.5)
    assert A.export_dense(pls.ROW_MAJOR).shape == (4, 4)
    y = np.zeros(4)
    A.spmv(np.array(g["x"]), y)
    assert np.max(np.abs(y - g["expected_trans"])) <= g["tol"]


def random_matrix(rng, m, n, mean):
    lens = rng.poisson(mean, m).clip(0, n)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(n, int(l), replace=False)) for l in lens]).astype(np.int32)
    return rp, ci, rng.standard_normal(int(rp[-1]))


@pytest.mark.parametrize("variant", [0, 1, 3, 8, 10, 12])
def test_plan_is_rebuilt_after_scale_and_transpose(pls, variant):
    rng = np.random.default_rng(1000 + variant)
    m, n = 3000, 2200
    rp, ci, v = random_matrix(rng, m, n, 9.0)
    D = np.zeros((m, n))
    for r in range(m):
        D[r, ci[rp[r]:rp[r + 1]]] = v[rp[r]:rp[r + 1]]
    A = pls.SpMat.csr(m, n, rp, ci, v)
    A.force_variant(variant)
    A.optimize()
    x = rng.standard_normal(n)
    y = np.zeros(m)
    A.spmv(x, y)
    assert rel_err(y, D @ x) <= TOL64
    A.scale(-1.5)
    y = np.zeros(m)
    A.spmv(x, y)
    assert rel_err(y, -1.5 * (D @ x)) <= TOL64
    A.transpose(pls.TRANS)
    assert A.query()[1:3] == (n, m)
    xt = rng.standard_normal(m)
    yt = np.zeros(n)
    A.spmv(xt, yt)
    assert rel_err(yt, -1.5 * (D.T @ xt)) <= TOL64
    A.optimize()  # a fresh selection on the transposed object
    yt = np.zeros(n)
    A.spmv(xt, yt)
    assert rel_err(yt, -1.5 * (D.T @ xt)) <= TOL64


def test_sell_object_norm_refused_and_transpose_converts(pls):
    rng = np.random.default_rng(77)
    m, n = 4096, 4096
    rp, ci, v = random_matrix(rng, m, n, 12.0)
    A = pls.SpMat.csr(m, n, rp, ci, v)
    A.force_sell(32, 64)
    A.optimize()  # the object is now SELL-C-sigma
    st, _ = A.norm(pls.NORM_FRB, raise_on_error=False)
    assert st == 1  # the reference computes no norm of a SELL-C-sigma object
    A.transpose(pls.TRANS)  # back to CSR, then transposed
    assert A.norm(pls.NORM_FRB) == pytest.approx(np.linalg.norm(v), rel=1e-12)
    D = np.zeros((m, n))
    for r in range(m):
        D[r, ci[rp[r]:rp[r + 1]]] = v[rp[r]:rp[r + 1]]
    x = rng.standard_normal(m)
    y = np.zeros(n)
    A.spmv(x, y)
    assert rel_err(y, D.T @ x) <= TOL64


def test_spmm_after_transpose_of_bsr(pls):
    rng = np.random.default_rng(5)
    b, mb, nb, k = 4, 64, 48, 16
    pattern = rng.random((mb, nb)) < 0.2
    pattern[:, 0] = True
    rp = np.concatenate([[0], np.cumsum(pattern.sum(1))]).astype(np.int32)
    ci = np.concatenate([np.flatnonzero(r) for r in pattern]).astype(np.int32)
    vals = rng.standard_normal(len(ci) * b * b)
    A = pls.SpMat.bsr(pls.ROW_MAJOR, mb * b, nb * b, b, rp, ci, vals)
    D = A.export_dense(pls.ROW_MAJOR)
    A.transpose(pls.TRANS)
    A.scale(2.0)
    B = rng.standard_normal((mb * b, k))
    Bm = pls.SpMat.dense(pls.ROW_MAJOR, mb * b, k, k, B)
    Cm = pls.SpMat.null(nb * b, k)
    pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, 1.0, A, Bm, 0.0, Cm)
    C = Cm.export_dense(pls.ROW_MAJOR)
    assert rel_err(C, 2.0 * (D.T @ B)) <= TOL64
