"""SpMM row blocks of clustered rows: rows grouped by shared columns and listed
per block, so the column-compressed kernel (k_spmm_lds) reads each block's
rows through that list.  Checked against the oracle (FP64 normwise 1e-12
against the compensated product; FP32 1e-6 against the exact product of the
stored FP32 values) on the 27-point stencil in its natural order and under a
random symmetric permutation (consecutive rows then share few columns, which
clustering recovers), with empty rows, a row too long for any list, lists of
256 columns, TRANS operands and value updates, beside the consecutive cut of
the same matrix."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
TOL64 = 1e-12


def permute(rp, ci, v, perm):
    """Symmetric permutation: row i of the result is row perm[i], column c is
    column inv[c]; columns sorted in each row."""
    m = len(rp) - 1
    inv = np.empty(m, np.int64)
    inv[perm] = np.arange(m)
    lens = np.diff(rp)[perm]
    nrp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    nci = np.empty(len(ci), np.int32)
    nv = np.empty(len(v), v.dtype)
    for i, r in enumerate(perm):
        cols = inv[ci[rp[r]:rp[r + 1]]]
        o = np.argsort(cols)
        nci[nrp[i]:nrp[i + 1]] = cols[o]
        nv[nrp[i]:nrp[i + 1]] = v[rp[r]:rp[r + 1]][o]
    return nrp, nci, nv


def transpose(m, n, rp, ci, v):
    """CSR of the n x m transpose, rows of the result in column order."""
    rows = np.repeat(np.arange(m, dtype=np.int32), np.diff(rp))
    o = np.lexsort((rows, ci))
    trp = np.concatenate([[0], np.cumsum(np.bincount(ci, minlength=n))]).astype(np.int32)
    return trp, rows[o], v[o]


def spmm_dev(A, Bv, C0, alpha, beta):
    import torch

    k = Bv.shape[1]
    Bd = torch.from_numpy(Bv).cuda()
    Cd = torch.from_numpy(C0.copy()).cuda()
    A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k, alpha, beta)
    torch.cuda.synchronize()
    return Cd.cpu().numpy()


def permuted_stencil(gen, n, seed):
    rp, ci, v = gen.stencil27(n)
    rng = np.random.default_rng(seed)
    v = v * rng.uniform(0.5, 1.5, len(v))
    return permute(rp, ci, v, rng.permutation(n ** 3))


@pytest.mark.parametrize("dtype,k", [(np.float64, 8), (np.float64, 32), (np.float64, 128),
                                     (np.float32, 8), (np.float32, 40)])
@pytest.mark.parametrize("beta", [0.0, -0.5])
def test_clustered_rows_on_permuted_stencil(pls, orc, gen, dtype, k, beta):
    rp, ci, v = permuted_stencil(gen, 14, 300 + k)
    m = len(rp) - 1
    rng = np.random.default_rng(k)
    Bv = rng.uniform(-1, 1, (m, k)).astype(dtype)
    C0 = rng.uniform(-1, 1, (m, k)).astype(dtype)
    vs = v.astype(dtype)
    A = pls.SpMat.csr(m, m, rp, ci, vs)
    got = spmm_dev(A, Bv, C0, 1.25, beta)
    info = A.spmm_info()
    assert info["blocks_state"] == 1 and A.spmm_rows_clustered()
    ref = orc.spmm(m, k, rp, ci, vs.astype(np.float64), Bv.astype(np.float64), 1.25, beta,
                   C0.astype(np.float64))
    tol = TOL64 if dtype == np.float64 else 1e-6
    assert rel_err(got, ref) <= tol
    # consecutive rows of the permuted matrix share few columns: the lists
    # are too long to be used, and the direct kernels give the same product
    A.set_spmm_rows(1)
    got1 = spmm_dev(A, Bv, C0, 1.25, beta)
    assert A.spmm_info()["blocks_state"] == -1 and not A.spmm_rows_clustered()
    assert rel_err(got1, ref) <= tol


@pytest.mark.parametrize("k", [8, 32])
def test_clustered_rows_shorten_natural_stencil_lists(pls, orc, gen, k):
    rp, ci, v = gen.stencil27(16)
    m = 16 ** 3
    rng = np.random.default_rng(70 + k)
    Bv = rng.uniform(-1, 1, (m, k))
    C0 = rng.uniform(-1, 1, (m, k))
    A = pls.SpMat.csr(m, m, rp, ci, v)
    A.set_spmm_rows(1)
    got1 = spmm_dev(A, Bv, C0, -0.75, 0.5)
    consecutive = A.spmm_info()["list_total"]
    assert not A.spmm_rows_clustered()
    A.set_spmm_rows(0)
    got0 = spmm_dev(A, Bv, C0, -0.75, 0.5)
    clustered = A.spmm_info()["list_total"]
    assert A.spmm_rows_clustered() and clustered < 0.7 * consecutive
    ref = orc.spmm(m, k, rp, ci, v, Bv, -0.75, 0.5, C0)
    assert rel_err(got0, ref) <= TOL64 and rel_err(got1, ref) <= TOL64


@pytest.mark.parametrize("k", [8, 32])
def test_clustered_rows_with_empty_and_long_rows(pls, orc, gen, k):
    """Forced clustering on a rectangular matrix with every 7th row empty (the
    empty rows form blocks of their own, C = beta * C there) and one row of 300
    distinct columns (a block without a list)."""
    prp, pci, pv = permuted_stencil(gen, 10, 41)
    m0 = len(prp) - 1
    n = m0 + 700
    rng = np.random.default_rng(5 + k)
    rows_c, rows_v = [], []
    for i in range(m0):
        if i % 7 == 3:
            rows_c.append(np.zeros(0, np.int32))
            rows_v.append(np.zeros(0))
        if i == m0 // 2:
            rows_c.append(np.sort(rng.choice(n, 300, replace=False)).astype(np.int32))
            rows_v.append(rng.standard_normal(300))
        rows_c.append(pci[prp[i]:prp[i + 1]])
        rows_v.append(pv[prp[i]:prp[i + 1]])
    m = len(rows_c)
    rp = np.concatenate([[0], np.cumsum([len(c) for c in rows_c])]).astype(np.int32)
    ci = np.concatenate(rows_c).astype(np.int32)
    v = np.concatenate(rows_v)
    Bv = rng.uniform(-1, 1, (n, k))
    C0 = rng.uniform(-1, 1, (m, k))
    A = pls.SpMat.csr(m, n, rp, ci, v)
    A.set_spmm_rows(2)
    got = spmm_dev(A, Bv, C0, 0.5, 0.5)
    assert A.spmm_info()["blocks_state"] == 1 and A.spmm_rows_clustered()
    assert rel_err(got, orc.spmm(m, k, rp, ci, v, Bv, 0.5, 0.5, C0)) <= TOL64
    got0 = spmm_dev(A, Bv, C0, 0.5, 0.0)
    assert rel_err(got0, orc.spmm(m, k, rp, ci, v, Bv, 0.5, 0.0)) <= TOL64


@pytest.mark.parametrize("dtype,k", [(np.float64, 8), (np.float64, 32), (np.float32, 40)])
def test_clustered_rows_lists_of_256(pls, orc, gen, dtype, k):
    """The 8 x 8-block operator (rows of 216 distinct columns: lists of 256)
    with its scalar rows permuted: clustering regroups the rows of a block row."""
    brp, bci, bv = gen.block_stencil(6, 5, 4, 8, 0x5EED0003)
    m = (len(brp) - 1) * 8
    rp, ci, v = gen.bsr_to_csr_rowmajor(brp, bci, bv, 8)
    rp, ci, v = permute(rp, ci, v, np.random.default_rng(8).permutation(m))
    rng = np.random.default_rng(80 + k)
    Bv = rng.uniform(-1, 1, (m, k)).astype(dtype)
    C0 = rng.uniform(-1, 1, (m, k)).astype(dtype)
    vs = v.astype(dtype)
    A = pls.SpMat.csr(m, m, rp, ci, vs)
    got = spmm_dev(A, Bv, C0, 1.0, 1.0)
    info = A.spmm_info()
    assert info["blocks_state"] == 1 and info["list_max"] == 256 and A.spmm_rows_clustered()
    ref = orc.spmm(m, k, rp, ci, vs.astype(np.float64), Bv.astype(np.float64), 1.0, 1.0,
                   C0.astype(np.float64))
    assert rel_err(got, ref) <= (TOL64 if dtype == np.float64 else 1e-6)


def test_clustered_rows_trans_operand_and_update(pls, orc, gen):
    """TRANS products through the drop-in call (the blocks are cut on the
    transposed operand), then a value update: the kept blocks multiply the new
    values."""
    rp, ci, v = permuted_stencil(gen, 12, 9)
    m = len(rp) - 1
    # an unsymmetric pattern: drop the entries of the lower triangle of every 5th row
    keep = np.ones(len(ci), bool)
    for r in range(0, m, 5):
        seg = slice(rp[r], rp[r + 1])
        keep[seg] = ci[seg] >= r
    lens = np.array([keep[rp[r]:rp[r + 1]].sum() for r in range(m)])
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci, v = ci[keep], v[keep]
    k = 24
    rng = np.random.default_rng(12)
    Bv = rng.uniform(-1, 1, (m, k))
    C0 = rng.uniform(-1, 1, (m, k))
    A = pls.SpMat.csr(m, m, rp, ci, v)
    trp, tci, tv = transpose(m, m, rp, ci, v)
    for it in range(2):
        B = pls.SpMat.dense(pls.ROW_MAJOR, m, k, k, Bv.reshape(-1))
        Cm = pls.SpMat.dense(pls.ROW_MAJOR, m, k, k, C0.reshape(-1))
        pls.spmm_exec(pls.TRANS, pls.NOTRANS, 0.75, A, B, -1.0, Cm)
        assert A.spmm_rows_clustered()
        ref = orc.spmm(m, k, trp, tci, tv, Bv, 0.75, -1.0, C0)
        assert rel_err(Cm.export_dense(pls.ROW_MAJOR), ref) <= TOL64
        if it == 0:
            pick = rng.choice(len(ci), 50, replace=False)
            rows = np.searchsorted(rp, pick, side="right") - 1
            newv = rng.standard_normal(50)
            A.update(rows, ci[pick], newv)
            v = v.copy()
            v[pick] = newv
            trp, tci, tv = transpose(m, m, rp, ci, v)


def test_scattered_columns_are_not_clustered(pls, gen):
    """Rows of uniformly random columns share none: the greedy is not run and
    the direct kernels serve the product."""
    import torch

    rp, ci, v = gen.powerlaw(20000, 3)
    m = len(rp) - 1
    k = 16
    A = pls.SpMat.csr(m, m, rp, ci, v)
    Bd = torch.zeros((m, k), dtype=torch.float64, device="cuda")
    Cd = torch.zeros((m, k), dtype=torch.float64, device="cuda")
    A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k)
    torch.cuda.synchronize()
    assert A.spmm_info()["blocks_state"] == -1 and not A.spmm_rows_clustered()
