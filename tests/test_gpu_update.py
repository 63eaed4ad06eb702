"""Value updates on optimized device plans (perflibs_spmat_update_{s,d}).

An update writes the changed values into the plan's device layout in place
(the plan, and so the chosen variant, survives the call).  The patched plan
must then compute exactly what a plan freshly built from the updated matrix
computes (bit-identical: same layout, same deterministic kernel), and FP64
results must match the oracle on the updated matrix.  Semantics follow the
reference's update (src/matvec.cpp:530-562, src/sell_c_sigma.cpp:966-1049):
updates apply in order, a later update of the same entry wins, the first
update that names no stored entry fails the call with the earlier ones kept,
and only the first stored copy of a duplicated (row, column) entry changes.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err
from test_gpu_spmv import powerlawish

pytestmark = pytest.mark.gpu

ALL_VARIANTS = list(range(13))
PANEL = 12


def host_update(rp, ci, v, rows, cols, vals, base=0):
    """Apply updates to CSR arrays the way the reference's CSR search does."""
    v = v.copy()
    for r, c, val in zip(rows, cols, vals):
        seg = ci[rp[r - base] - base:rp[r - base + 1] - base]
        hit = np.nonzero(seg == c)[0]
        assert hit.size, (r, c)
        v[rp[r - base] - base + hit[0]] = val
    return v


def matrix_with_duplicates(rng, m, n):
    rp, ci, v = powerlawish(rng, m, n, mean=10.0)
    # duplicate the first entry of a few rows (same column stored twice)
    rows = [r for r in range(0, m, 97) if rp[r + 1] > rp[r]][:12]
    lens = np.diff(rp).astype(np.int64)
    for r in rows:
        lens[r] += 1
    rp2 = np.zeros(m + 1, np.int64)
    rp2[1:] = np.cumsum(lens)
    ci2 = np.empty(int(rp2[-1]), np.int32)
    v2 = np.empty(int(rp2[-1]))
    for r in range(m):
        seg_c, seg_v = ci[rp[r]:rp[r + 1]], v[rp[r]:rp[r + 1]]
        if r in rows:
            seg_c = np.concatenate([seg_c[:1], seg_c])
            seg_v = np.concatenate([[0.25], seg_v])
        ci2[rp2[r]:rp2[r + 1]] = seg_c
        v2[rp2[r]:rp2[r + 1]] = seg_v
    return rp2.astype(np.int32), ci2, v2, rows


def pick_updates(rng, rp, ci, count, dup_rows):
    m = len(rp) - 1
    nz_rows = np.nonzero(np.diff(rp) > 0)[0]
    rows = rng.choice(nz_rows, count)
    cols = np.array([ci[rng.integers(rp[r], rp[r + 1])] for r in rows], np.int32)
    vals = rng.standard_normal(count)
    # the same entry twice (the later value wins) and duplicated entries
    rows = np.concatenate([rows, rows[:3], dup_rows[:4]]).astype(np.int32)
    cols = np.concatenate([cols, cols[:3], [ci[rp[r]] for r in dup_rows[:4]]]).astype(np.int32)
    vals = np.concatenate([vals, rng.standard_normal(3 + len(dup_rows[:4]))])
    assert rows.max() < m
    return rows, cols, vals


def make(pls, m, n, rp, ci, v, dtype, variant, precision=None):
    A = pls.SpMat.csr(m, n, rp, ci, v.astype(dtype))
    if precision is not None:
        A.set_precision(precision)
    A.force_variant(variant)
    if variant == PANEL:
        A.force_panel(7, 1)  # 24 panels over 3072 columns, several tiles per workgroup
    A.optimize()
    return A


def spmv(A, x, m, alpha, beta, y0, dtype):
    y = np.array(y0, dtype=dtype, copy=True)
    A.spmv(np.asarray(x, dtype=dtype), y, alpha, beta)
    return y


@pytest.mark.parametrize("variant", ALL_VARIANTS)
@pytest.mark.parametrize("prec", ["f64", "f32", "bf16"])
def test_update_patches_plan_in_place(pls, orc, variant, prec):
    rng = np.random.default_rng(7000 + variant)
    m, n = 4096, 3072
    rp, ci, v, dup_rows = matrix_with_duplicates(rng, m, n)
    dtype = np.float64 if prec == "f64" else np.float32
    precision = pls.VALUES_BF16 if prec == "bf16" else None
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, m)
    A = make(pls, m, n, rp, ci, v, dtype, variant, precision)
    spmv(A, x, m, 1.0, 0.0, y0, dtype)  # the plan exists on the device
    assert A.variant() == variant
    rows, cols, vals = pick_updates(rng, rp, ci, 200, dup_rows)
    vals = vals.astype(dtype)
    A.update(rows, cols, vals)
    assert A.variant() == variant, "the plan was dropped instead of patched"
    y = spmv(A, x, m, -1.5, 0.5, y0, dtype)
    v_new = host_update(rp, ci, v.astype(dtype), rows, cols, vals)
    F = make(pls, m, n, rp, ci, v_new, dtype, variant, precision)
    yf = spmv(F, x, m, -1.5, 0.5, y0, dtype)
    assert np.array_equal(y, yf)
    if prec == "f64":
        assert rel_err(y, orc.spmv(m, rp, ci, v_new, x, -1.5, 0.5, y0)) <= 1e-12


@pytest.mark.parametrize("variant", [0, 1, 8, 10, PANEL])
def test_failed_update_keeps_earlier_ones(pls, orc, variant):
    rng = np.random.default_rng(71)
    m, n = 4096, 3072
    rp, ci, v = powerlawish(rng, m, n, mean=10.0)
    x = rng.uniform(-1, 1, n)
    A = make(pls, m, n, rp, ci, v, np.float64, variant)
    spmv(A, x, m, 1.0, 0.0, np.zeros(m), np.float64)
    r = int(np.nonzero(np.diff(rp) > 1)[0][0])
    absent = int(np.setdiff1d(np.arange(n), ci[rp[r]:rp[r + 1]])[0])
    rows, cols, vals = [r, r], [int(ci[rp[r]]), absent], [3.5, 4.5]
    assert A.update(rows, cols, vals, raise_on_error=False) == 1
    assert A.variant() == variant
    v_new = host_update(rp, ci, v, rows[:1], cols[:1], vals[:1])
    y = spmv(A, x, m, 1.0, 0.0, np.zeros(m), np.float64)
    assert rel_err(y, orc.spmv(m, rp, ci, v_new, x)) <= 1e-12


@pytest.mark.parametrize("variant", [0, 1, 8, 10, PANEL])
def test_transposed_plan_is_patched(pls, variant):
    """A plan built for op(A) = A^T is patched in place too: each update is
    moved to the transposed operand (row c, at the position of original row r
    among its entries, duplicates in their original order)."""
    rng = np.random.default_rng(72 + variant)
    m, n = 2048, 3072
    rp, ci, v, dup_rows = matrix_with_duplicates(rng, m, n)
    x = rng.uniform(-1, 1, m)

    def build(vals):
        A = pls.SpMat.csr(m, n, rp, ci, vals)
        A.hint(pls.HINT_SPMV_OPERATION, pls.TRANS)
        A.force_variant(variant)
        if variant == PANEL:
            A.force_panel(7, 1)
        A.optimize()
        return A

    A = build(v)
    y = np.zeros(n)
    A.spmv(x, y, trans=pls.TRANS)
    rows, cols, vals = pick_updates(rng, rp, ci, 120, np.array(dup_rows))
    A.update(rows, cols, vals)
    assert A.variant() == variant, "the transposed plan was dropped instead of patched"
    y[:] = 0
    A.spmv(x, y, trans=pls.TRANS)
    v_new = host_update(rp, ci, v, rows, cols, vals)
    F = build(v_new)
    yf = np.zeros(n)
    F.spmv(x, yf, trans=pls.TRANS)
    assert np.array_equal(y, yf)
    D = np.zeros((m, n))
    for i in range(m):
        np.add.at(D[i], ci[rp[i]:rp[i + 1]], v_new[rp[i]:rp[i + 1]])
    assert rel_err(y, D.T @ x) <= 1e-12


@pytest.mark.parametrize("layout", ["row", "col"])
def test_bsr_handle_update(pls, orc, layout):
    rng = np.random.default_rng(73)
    b, mb, nb = 4, 300, 200
    brp = np.zeros(mb + 1, np.int32)
    cols = [np.sort(rng.choice(nb, rng.integers(1, 6), replace=False)) for _ in range(mb)]
    brp[1:] = np.cumsum([len(c) for c in cols])
    bci = np.concatenate(cols).astype(np.int32)
    bv = rng.standard_normal(len(bci) * b * b)
    lay = pls.ROW_MAJOR if layout == "row" else pls.COL_MAJOR
    A = pls.SpMat.bsr(lay, mb * b, nb * b, b, brp, bci, bv)
    A.force_variant(10)  # BSR b = 4 on the handle's own blocks
    A.optimize()
    x = rng.uniform(-1, 1, nb * b)
    y = np.zeros(mb * b)
    A.spmv(x, y)
    # dense image of the blocks
    D = np.zeros((mb * b, nb * b))
    for rb in range(mb):
        for k in range(brp[rb], brp[rb + 1]):
            blk = bv[k * b * b:(k + 1) * b * b].reshape(b, b)
            D[rb * b:(rb + 1) * b, bci[k] * b:(bci[k] + 1) * b] = blk if layout == "row" else blk.T
    ups = []
    for _ in range(40):
        rb = int(rng.integers(mb))
        k = int(rng.integers(brp[rb], brp[rb + 1]))
        ups.append((rb * b + int(rng.integers(b)), int(bci[k]) * b + int(rng.integers(b)),
                    float(rng.standard_normal())))
    rows, cols, vals = map(list, zip(*ups))
    A.update(rows, cols, vals)
    assert A.variant() == 10
    for r, c, val in ups:
        D[r, c] = val
    y[:] = 0
    A.spmv(x, y)
    assert rel_err(y, D @ x) <= 1e-12
    # a CSR-family plan of a BSR handle is rebuilt instead (its expansion drops zeros)
    A.force_variant(0)
    A.optimize()
    A.spmv(x, y)
    A.update(rows[:1], cols[:1], [0.0])
    assert A.variant() == -1
    D[rows[0], cols[0]] = 0.0
    y[:] = 0
    A.spmv(x, y)
    assert rel_err(y, D @ x) <= 1e-12


@pytest.mark.parametrize("variant", [1, 7, 11, PANEL])
def test_update_one_based_handle(pls, orc, variant):
    """Index base 1: updates name 1-based rows and columns; the device plan is
    patched at the same entries as for base 0."""
    rng = np.random.default_rng(74 + variant)
    m, n = 4096, 3072
    rp, ci, v, dup_rows = matrix_with_duplicates(rng, m, n)
    x = rng.uniform(-1, 1, n)
    A = make(pls, m, n, rp + 1, ci + 1, v, np.float64, variant)
    spmv(A, x, m, 1.0, 0.0, np.zeros(m), np.float64)
    rows, cols, vals = pick_updates(rng, rp, ci, 100, dup_rows)
    A.update(rows + 1, cols + 1, vals)
    assert A.variant() == variant
    v_new = host_update(rp, ci, v, rows, cols, vals)
    y = spmv(A, x, m, 1.0, 0.0, np.zeros(m), np.float64)
    assert rel_err(y, orc.spmv(m, rp, ci, v_new, x)) <= 1e-12


@pytest.mark.parametrize("k", [8, 32])
def test_update_then_spmm_column_compressed(pls, orc, gen, k):
    """SpMM plans with column-compressed row blocks share the CSR values, so an
    in-place update reaches the wide-B kernel too."""
    import torch

    rp, ci, v = gen.stencil27(16)
    m = 16 ** 3
    rng = np.random.default_rng(75)
    v = v * rng.uniform(0.5, 1.5, len(v))
    Bv = rng.uniform(-1, 1, (m, k))
    A = pls.SpMat.csr(m, m, rp, ci, v)
    Bd = torch.from_numpy(Bv).cuda()
    Cd = torch.zeros((m, k), dtype=torch.float64, device="cuda")
    A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k, 1.0, 0.0)
    rows, cols, vals = pick_updates(rng, rp, ci, 300, np.array([], np.int64))
    A.update(rows, cols, vals)
    A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k, 1.0, 0.0)
    torch.cuda.synchronize()
    v_new = host_update(rp, ci, v, rows, cols, vals)
    assert rel_err(Cd.cpu().numpy(), orc.spmm(m, k, rp, ci, v_new, Bv)) <= 1e-12


def test_transposed_forced_sell_plan_is_patched(pls, orc):
    """A SELL-C-sigma plan forced with force_sell for op(A) = A^T keeps the
    transposed operand's structure like every other TRANS plan, so an update
    patches its device layout in place (the plan survives) instead of dropping
    it and rebuilding the SELL layout on the host."""
    rng = np.random.default_rng(91)
    m, n = 2048, 3072
    rp, ci, v, dup_rows = matrix_with_duplicates(rng, m, n)
    x = rng.uniform(-1, 1, m)
    A = pls.SpMat.csr(m, n, rp, ci, v)
    A.hint(pls.HINT_SPMV_OPERATION, pls.TRANS)
    A.hint(pls.HINT_SPMV_INVOCATIONS, pls.INVOCATIONS_MANY)
    A.force_sell(32, 64)
    A.optimize()
    y = np.zeros(n)
    A.spmv(x, y, trans=pls.TRANS)
    before = A.variant()
    assert before >= 0
    rows, cols, vals = pick_updates(rng, rp, ci, 80, np.array(dup_rows))
    A.update(rows, cols, vals)
    assert A.variant() == before, "the transposed SELL plan was dropped instead of patched"
    y[:] = 0
    A.spmv(x, y, trans=pls.TRANS)
    v_new = host_update(rp, ci, v, rows, cols, vals)
    D = np.zeros((m, n))
    for i in range(m):
        np.add.at(D[i], ci[rp[i]:rp[i + 1]], v_new[rp[i]:rp[i + 1]])
    assert rel_err(y, D.T @ x) <= 1e-12
