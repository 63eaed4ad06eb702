"""GPU parity of CSR x dense SpMM through the C-ABI (FP64 normwise 1e-12
against the compensated oracle; reference SpMM goldens)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
TOL64 = 1e-12


def csr_of_dense(D):
    rp = np.concatenate([[0], np.cumsum((D != 0).sum(axis=1))]).astype(np.int32)
    ci = np.concatenate([np.nonzero(r)[0] for r in D]).astype(np.int32)
    v = np.concatenate([r[np.nonzero(r)[0]] for r in D])
    return rp, ci, v


def test_spmm_null_c_golden(pls, golden):
    g = golden["spmm_null_c"]
    A = pls.SpMat.csr(4, 4, g["row_ptr"], g["col_indx"], np.array(g["vals"]))
    B = pls.SpMat.dense(pls.ROW_MAJOR, 4, 4, 4, np.array(g["B_row_major"]))
    Cm = pls.SpMat.null(4, 4)
    pls.spmm_optimize(pls.NOTRANS, pls.NOTRANS, pls.SCALAR_ANY, A, B, pls.SCALAR_ZERO, Cm)
    pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, 1.0, A, B, 0.0, Cm)
    exp = np.array(g["dense_A"]).reshape(4, 4) @ np.array(g["B_row_major"]).reshape(4, 4)
    assert np.max(np.abs(Cm.export_dense() - exp)) <= g["tol"]


def test_spmm_variants_golden(pls, golden):
    g = golden["spmm_variants"]
    DA = np.array(g["dense_A"]).reshape(4, 4)
    DB = np.array(g["B_row_major"]).reshape(4, 4)
    A = pls.SpMat.csr(4, 4, g["row_ptr"], g["col_indx"], np.array(g["vals"]))
    B = pls.SpMat.dense(pls.ROW_MAJOR, 4, 4, 4, DB.reshape(-1))
    opt = pls.SpMat.null(4, 4)
    pls.spmm_optimize(pls.NOTRANS, pls.NOTRANS, pls.SCALAR_ANY, A, B, pls.SCALAR_ANY, opt)
    for ta in (pls.NOTRANS, pls.TRANS):
        for tb in (pls.NOTRANS, pls.TRANS):
            alpha = g["alpha_notrans"] if ta == pls.NOTRANS else g["alpha_trans"]
            Cm = pls.SpMat.null(4, 4)
            pls.spmm_exec(ta, tb, alpha, A, B, 0.0, Cm)
            opA = DA if ta == pls.NOTRANS else DA.T
            opB = DB if tb == pls.NOTRANS else DB.T
            assert np.max(np.abs(Cm.export_dense() - alpha * opA @ opB)) <= g["tol"]
    c0 = np.array(g["c_init_row_major"]).reshape(4, 4)
    Cd = pls.SpMat.dense(pls.ROW_MAJOR, 4, 4, 4, c0.reshape(-1))
    pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, g["alpha_beta"], A, B, g["beta"], Cd)
    exp = g["alpha_beta"] * DA @ DB + g["beta"] * c0
    assert np.max(np.abs(Cd.export_dense() - exp)) <= g["tol"]
    Bc = pls.SpMat.dense(pls.COL_MAJOR, 4, 4, 4, DB.T.reshape(-1))
    Cc = pls.SpMat.dense(pls.COL_MAJOR, 4, 4, 4, c0.T.reshape(-1))
    pls.spmm_optimize(pls.NOTRANS, pls.NOTRANS, pls.SCALAR_ANY, A, Bc, pls.SCALAR_ANY, Cc)
    pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, g["alpha_beta"], A, Bc, g["beta"], Cc)
    assert np.max(np.abs(Cc.export_dense() - exp)) <= g["tol"]


@pytest.mark.parametrize("k", [1, 3, 4, 5, 8, 17, 32, 100, 128, 300])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("beta", [0.0, 1.0, -0.5])
def test_spmm_parity(pls, orc, gen, k, dtype, beta):
    rng = np.random.default_rng(k)
    rp, ci, v = gen.stencil3d(16, 12, 10, 1, False)
    v = v * rng.uniform(0.5, 1.5, len(v))
    m = 16 * 12 * 10
    Bv = rng.uniform(-1, 1, (m, k))
    C0 = rng.uniform(-1, 1, (m, k))
    A = pls.SpMat.csr(m, m, rp, ci, v.astype(dtype))
    B = pls.SpMat.dense(pls.ROW_MAJOR, m, k, k, Bv.astype(dtype).reshape(-1))
    Cm = pls.SpMat.dense(pls.ROW_MAJOR, m, k, k, C0.astype(dtype).reshape(-1))
    pls.spmm_optimize(pls.NOTRANS, pls.NOTRANS, pls.SCALAR_ANY, A, B, pls.SCALAR_ANY, Cm)
    pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, 1.5, A, B, beta, Cm)
    if dtype == np.float64:
        ref = orc.spmm(m, k, rp, ci, v, Bv, 1.5, beta, C0)
        assert rel_err(Cm.export_dense(), ref) <= TOL64
    else:
        ref = orc.spmm(m, k, rp, ci, v.astype(np.float32).astype(np.float64),
                       Bv.astype(np.float32).astype(np.float64), 1.5, beta,
                       C0.astype(np.float32).astype(np.float64))
        assert rel_err(Cm.export_dense(), ref) <= 1e-6


def test_spmm_bsr_operand_and_transposes(pls, orc, gen):
    brp, bci, bv = gen.block_stencil(6, 5, 4, 4, 0x5EED0003)
    m = 6 * 5 * 4 * 4
    k = 16
    rng = np.random.default_rng(3)
    Bv = rng.uniform(-1, 1, (m, k))
    A = pls.SpMat.bsr(pls.ROW_MAJOR, m, m, 4, brp, bci, bv)
    B = pls.SpMat.dense(pls.ROW_MAJOR, m, k, k, Bv.reshape(-1))
    Cm = pls.SpMat.null(m, k)
    pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, 1.0, A, B, 0.0, Cm)
    rp, ci, v = orc.bsr2csr(91, m, m, 4, brp, bci, bv)
    assert rel_err(Cm.export_dense(), orc.spmm(m, k, rp, ci, v, Bv)) <= TOL64
    # op(A) = A^T, op(B) = B^T via a column-major B' of shape k x m
    D = np.zeros((m, m))
    for i in range(m):
        D[i, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    Bt = pls.SpMat.dense(pls.ROW_MAJOR, k, m, m, Bv.T.copy().reshape(-1))
    A2 = pls.SpMat.csr(m, m, rp, ci, v)
    C2 = pls.SpMat.null(m, k)
    pls.spmm_exec(pls.TRANS, pls.TRANS, 0.75, A2, Bt, 0.0, C2)
    assert rel_err(C2.export_dense(), 0.75 * D.T @ Bv) <= TOL64


def test_spmm_trivial_cases(pls):
    A = pls.SpMat.csr(3, 3, [0, 1, 2, 3], [0, 1, 2], np.ones(3))
    B = pls.SpMat.dense(pls.ROW_MAJOR, 3, 6, 6, np.ones(18))
    C0 = np.arange(18, dtype=np.float64)
    Cm = pls.SpMat.dense(pls.ROW_MAJOR, 3, 6, 6, C0)
    pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, 0.0, A, B, 2.0, Cm)
    assert np.array_equal(Cm.export_dense().reshape(-1), 2 * C0)
    pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, 0.0, A, B, 0.0, Cm)  # C becomes null
    assert np.array_equal(Cm.export_dense(), np.zeros((3, 6)))
    assert Cm.query()[3] == 0


def test_spmm_async_device(pls, orc, gen):
    import torch

    rp, ci, v = gen.stencil27(24)
    m = 24**3
    k = 32
    rng = np.random.default_rng(4)
    Bv = rng.uniform(-1, 1, (m, k))
    A = pls.SpMat.csr(m, m, rp, ci, v)
    Bd = torch.from_numpy(Bv).cuda()
    Cd = torch.zeros((m, k), dtype=torch.float64, device="cuda")
    A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k, 1.0, 0.0)
    torch.cuda.synchronize()
    assert rel_err(Cd.cpu().numpy(), orc.spmm(m, k, rp, ci, v, Bv)) <= TOL64


def test_stencil27_full_k8(pls, orc, gen):
    """The 128^3 operator at its real size with k = 8 right-hand sides."""
    rp, ci, v = gen.stencil27(128)
    m = 128**3
    assert rp[-1] == 55_742_968
    k = 8
    Bv = gen.uniform(m * k, 0x5EED0003).reshape(m, k)
    A = pls.SpMat.csr(m, m, rp, ci, v)
    B = pls.SpMat.dense(pls.ROW_MAJOR, m, k, k, Bv.reshape(-1))
    Cm = pls.SpMat.null(m, k)
    pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, 1.0, A, B, 0.0, Cm)
    got = Cm.export_dense()
    ref = orc.spmm(m, k, rp, ci, v, Bv)
    assert rel_err(got, ref) <= TOL64
    # closed form: with B = 1 interior rows vanish (26 - 26 neighbours)
    ones = pls.SpMat.dense(pls.ROW_MAJOR, m, k, k, np.ones(m * k))
    C1 = pls.SpMat.null(m, k)
    pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, 1.0, A, ones, 0.0, C1)
    g = C1.export_dense()[:, 0].reshape(128, 128, 128)
    assert np.all(g[1:-1, 1:-1, 1:-1] == 0.0) and np.all(g[0] > 0)


@pytest.mark.parametrize("k", [1, 2, 4, 8, 16])
def test_spmm_long_and_short_rows(pls, orc, k):
    """Row blocks too long to stage (rows of 3000 non-zeros), blocks of few long
    rows that several lane groups share, and empty rows, through the async path."""
    import torch

    rng = np.random.default_rng(50 + k)
    m, n = 3000, 7000
    lens = rng.integers(0, 40, m)
    lens[rng.random(m) < 0.1] = 0
    lens[::97] = 3000
    lens[1500:1540] = 400
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(n, int(l), replace=False)) for l in lens]).astype(np.int32)
    v = rng.standard_normal(int(rp[-1]))
    Bv = rng.uniform(-1, 1, (n, k))
    C0 = rng.uniform(-1, 1, (m, k))
    A = pls.SpMat.csr(m, n, rp, ci, v)
    Bd = torch.from_numpy(Bv).cuda()
    for beta in (0.0, 0.5):
        Cd = torch.from_numpy(C0.copy()).cuda()
        A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k, -1.25, beta)
        torch.cuda.synchronize()
        ref = orc.spmm(m, k, rp, ci, v, Bv, -1.25, beta, C0 if beta else None)
        assert rel_err(Cd.cpu().numpy(), ref) <= TOL64


def banded_with_long_rows(rng, m, n, bw=40, per_row=(0, 30)):
    """Banded rows (a B row shared by many rows of a block) plus rows with more
    than 128 distinct columns or 1024 non-zeros, runs of empty rows and
    duplicated columns."""
    rows = []
    for i in range(m):
        L = int(rng.integers(*per_row))
        lo, hi = max(0, i - bw), min(n, i + bw + 1)
        cols = np.sort(rng.choice(np.arange(lo, hi), min(L, hi - lo), replace=False))
        if i % 211 == 5:
            cols = np.sort(rng.choice(n, 300, replace=False))       # > 128 distinct columns
        if i % 503 == 7:
            cols = np.sort(rng.integers(0, 60, 1500))                 # > 1024 non-zeros, few columns
        if 900 <= i < 1000:
            cols = np.array([], np.int64)
        if i % 37 == 0 and len(cols):
            cols = np.sort(np.concatenate([cols, cols[:1]]))          # a duplicated column
        rows.append(cols)
    rp = np.concatenate([[0], np.cumsum([len(c) for c in rows])]).astype(np.int32)
    ci = np.concatenate(rows).astype(np.int32)
    return rp, ci, rng.standard_normal(int(rp[-1]))


@pytest.mark.parametrize("k", [32, 40, 64, 100, 129, 1024])
def test_spmm_wide_column_compressed_blocks(pls, orc, k):
    """Wide B through the column-compressed row-block kernel: list blocks,
    blocks without a list, empty rows, k not a multiple of the column chunk."""
    import torch

    rng = np.random.default_rng(700 + k)
    m, n = 3000, 3100
    rp, ci, v = banded_with_long_rows(rng, m, n)
    Bv = rng.uniform(-1, 1, (n, k))
    C0 = rng.uniform(-1, 1, (m, k))
    A = pls.SpMat.csr(m, n, rp, ci, v)
    Bd = torch.from_numpy(Bv).cuda()
    for beta in (0.0, 0.5):
        Cd = torch.from_numpy(C0.copy()).cuda()
        A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k, 0.75, beta)
        torch.cuda.synchronize()
        ref = orc.spmm(m, k, rp, ci, v, Bv, 0.75, beta, C0 if beta else None)
        assert rel_err(Cd.cpu().numpy(), ref) <= TOL64


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_spmm_wide_reduced_precision(pls, orc, prec):
    import torch

    rng = np.random.default_rng(91)
    m = n = 2500
    rp, ci, v = banded_with_long_rows(rng, m, n)
    k = 96
    v32 = v.astype(np.float32)
    A = pls.SpMat.csr(m, n, rp, ci, v32)
    if prec == "bf16":
        A.set_precision(pls.VALUES_BF16)
        u = v32.view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
        stored = (u.astype(np.uint32) << 16).view(np.float32)
    else:
        stored = v32
    Bv = rng.uniform(-1, 1, (n, k)).astype(np.float32)
    Bd = torch.from_numpy(Bv).cuda()
    Cd = torch.zeros((m, k), dtype=torch.float32, device="cuda")
    A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k, 1.0, 0.0)
    torch.cuda.synchronize()
    ref = orc.spmm(m, k, rp, ci, stored.astype(np.float64), Bv.astype(np.float64))
    assert rel_err(Cd.cpu().numpy(), ref) <= 1e-6


def test_spmm_k1024_random_columns(pls, orc):
    """k = 1024 (the reference benchmark's width) on random columns, where the
    block lists would not shrink B traffic: row-major column slices of 512."""
    import torch

    rng = np.random.default_rng(92)
    m, n, k = 1200, 50000, 1024
    lens = rng.integers(1, 20, m)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(n, int(l), replace=False)) for l in lens]).astype(np.int32)
    v = rng.standard_normal(int(rp[-1]))
    Bv = rng.uniform(-1, 1, (n, k))
    A = pls.SpMat.csr(m, n, rp, ci, v)
    Bd = torch.from_numpy(Bv).cuda()
    Cd = torch.zeros((m, k), dtype=torch.float64, device="cuda")
    A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k, 1.0, 0.0)
    torch.cuda.synchronize()
    assert rel_err(Cd.cpu().numpy(), orc.spmm(m, k, rp, ci, v, Bv)) <= TOL64


def test_spmm_drop_in_k1024_like_the_reference_benchmark(pls, orc, gen):
    """The reference benchmark's product through the drop-in objects: CSR A,
    row-major dense B with 1024 columns, null C, spmm_optimize then spmm_exec."""
    rp, ci, v = gen.stencil27(10)
    m = 10 ** 3
    k = 1024
    rng = np.random.default_rng(93)
    v = v * rng.uniform(0.5, 1.5, len(v))
    Bv = rng.uniform(-1, 1, (m, k))
    A = pls.SpMat.csr(m, m, rp, ci, v)
    B = pls.SpMat.dense(pls.ROW_MAJOR, m, k, k, Bv.reshape(-1))
    Cm = pls.SpMat.null(m, k)
    pls.spmm_optimize(pls.NOTRANS, pls.NOTRANS, pls.SCALAR_ONE, A, B, pls.SCALAR_ZERO, Cm)
    for _ in range(2):
        pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, 1.0, A, B, 0.0, Cm)
    assert rel_err(Cm.export_dense(), orc.spmm(m, k, rp, ci, v, Bv)) <= TOL64


def test_spmm_column_compressed_dispatch(pls, gen):
    """The wide-B blocks are used where B rows are shared (stencil: lists of
    128 columns; 8 x 8 blocks: lists of 256) and not on random columns; they are
    built on the first SpMM with k >= 8 only."""
    import torch

    def run(m, n, rp, ci, v, k):
        A = pls.SpMat.csr(m, n, rp, ci, v)
        Bd = torch.zeros((n, k), dtype=torch.float64, device="cuda")
        Cd = torch.zeros((m, k), dtype=torch.float64, device="cuda")
        A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k)
        torch.cuda.synchronize()
        return A.spmm_info()

    rp, ci, v = gen.stencil27(16)
    m = 16 ** 3
    assert run(m, m, rp, ci, v, 4)["blocks_state"] == 0
    info = run(m, m, rp, ci, v, 8)
    assert info["blocks_state"] == 1 and info["list_max"] == 128
    assert info["list_total"] < 0.6 * len(ci)
    brp, bci, bv = gen.block_stencil(6, 5, 4, 8, 0x5EED0003)
    mb = len(brp) - 1
    crp, cci, cv = gen.bsr_to_csr_rowmajor(brp, bci, bv, 8)
    info = run(mb * 8, mb * 8, crp, cci, cv, 32)
    assert info["blocks_state"] == 1 and info["list_max"] == 256
    rng = np.random.default_rng(94)
    m, n = 2000, 400000
    lens = rng.integers(1, 30, m)
    rrp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    rci = np.concatenate([np.sort(rng.choice(n, int(l), replace=False)) for l in lens]).astype(np.int32)
    assert run(m, n, rrp, rci, rng.standard_normal(int(rrp[-1])), 32)["blocks_state"] == -1


def block8_operand(gen, long_row=None):
    """The 8 x 8-block 27-point operator (dense N(0,1) blocks) as CSR; rows of
    216 distinct columns need the column-compressed lists of 256.  long_row:
    a row replaced by 300 distinct columns, which no list holds."""
    brp, bci, bv = gen.block_stencil(6, 5, 4, 8, 0x5EED0003)
    m = (len(brp) - 1) * 8
    rp, ci, v = gen.bsr_to_csr_rowmajor(brp, bci, bv, 8)
    if long_row is not None:
        rng = np.random.default_rng(7)
        cols = np.sort(rng.choice(m, 300, replace=False)).astype(np.int32)
        lens = np.diff(rp)
        lens[long_row] = 300
        parts_c = [ci[rp[i]:rp[i + 1]] if i != long_row else cols for i in range(m)]
        parts_v = [v[rp[i]:rp[i + 1]] if i != long_row else rng.standard_normal(300) for i in range(m)]
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        ci = np.concatenate(parts_c).astype(np.int32)
        v = np.concatenate(parts_v)
    return m, brp, bci, bv, rp, ci, v


@pytest.mark.parametrize("dtype,k", [(np.float64, 8), (np.float64, 32), (np.float64, 128),
                                     (np.float32, 8), (np.float32, 16), (np.float32, 40),
                                     (np.float32, 96)])
@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_spmm_lists_of_256_values(pls, orc, gen, dtype, k, beta):
    """Values of the column-compressed kernel with lists of 256 columns
    (k_spmm_lds<8|16|32, 256, ...>) on the 8 x 8-block operator, one row
    without a list: FP64 to 1e-12 against the compensated oracle, FP32 to
    1e-6 against the exact product of the stored FP32 values."""
    import torch

    m, _, _, _, rp, ci, v = block8_operand(gen, long_row=333)
    rng = np.random.default_rng(11 + k)
    Bv = rng.uniform(-1, 1, (m, k)).astype(dtype)
    C0 = rng.uniform(-1, 1, (m, k)).astype(dtype)
    vs = v.astype(dtype)
    A = pls.SpMat.csr(m, m, rp, ci, vs)
    Bd = torch.from_numpy(Bv).cuda()
    Cd = torch.from_numpy(C0.copy()).cuda()
    A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k, 1.25, beta)
    torch.cuda.synchronize()
    info = A.spmm_info()
    assert info["blocks_state"] == 1 and info["list_max"] == 256
    ref = orc.spmm(m, k, rp, ci, vs.astype(np.float64), Bv.astype(np.float64), 1.25, beta,
                   C0.astype(np.float64))
    tol = TOL64 if dtype == np.float64 else 1e-6
    assert rel_err(Cd.cpu().numpy(), ref) <= tol
    # a second call reuses the lists
    A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k, 1.25, 0.0)
    torch.cuda.synchronize()
    ref0 = orc.spmm(m, k, rp, ci, vs.astype(np.float64), Bv.astype(np.float64), 1.25, 0.0)
    assert rel_err(Cd.cpu().numpy(), ref0) <= tol


@pytest.mark.parametrize("k", [8, 32])
def test_spmm_bsr8_drop_in_native_blocks(pls, orc, gen, k):
    """The drop-in perflibs_spmm_exec_d on an 8 x 8 BSR handle converts the
    object to CSR like the reference's SpMM path (src/matmul.cpp:619-649) and
    runs the product on the kept blocks; it matches the oracle's bsr2csr +
    product."""
    m, brp, bci, bv, _, _, _ = block8_operand(gen)
    rng = np.random.default_rng(21 + k)
    Bv = rng.uniform(-1, 1, (m, k))
    C0 = rng.uniform(-1, 1, (m, k))
    A = pls.SpMat.bsr(pls.ROW_MAJOR, m, m, 8, brp, bci, bv)
    B = pls.SpMat.dense(pls.ROW_MAJOR, m, k, k, Bv.reshape(-1))
    Cm = pls.SpMat.dense(pls.ROW_MAJOR, m, k, k, C0.reshape(-1))
    pls.spmm_optimize(pls.NOTRANS, pls.NOTRANS, pls.SCALAR_ANY, A, B, pls.SCALAR_ANY, Cm)
    pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, -0.75, A, B, 0.5, Cm)
    rp, ci, v = orc.bsr2csr(91, m, m, 8, brp, bci, bv)
    assert rel_err(Cm.export_dense(), orc.spmm(m, k, rp, ci, v, Bv, -0.75, 0.5, C0)) <= TOL64
    info = A.spmm_info()
    assert info["blocks_state"] == 2 and info["list_max"] == 8
    assert A.query()[3] == len(ci)  # the object is CSR now, explicit zeros dropped


@pytest.mark.parametrize("k", [8, 40])
def test_spmm_leading_dimensions(pls, orc, gen, k):
    """Row-major B and C with leading dimensions wider than k (ldb = k + 3,
    ldc = k + 5) through the column-compressed kernel; the padding columns of C
    are left untouched."""
    import torch

    rp, ci, v = gen.stencil27(14)
    m = 14 ** 3
    rng = np.random.default_rng(95 + k)
    v = v * rng.uniform(0.5, 1.5, len(v))
    ldb, ldc = k + 3, k + 5
    Bfull = rng.uniform(-1, 1, (m, ldb))
    C0 = rng.uniform(-1, 1, (m, ldc))
    A = pls.SpMat.csr(m, m, rp, ci, v)
    Bd = torch.from_numpy(Bfull).cuda()
    Cd = torch.from_numpy(C0.copy()).cuda()
    A.spmm_async(k, Bd.data_ptr(), ldb, Cd.data_ptr(), ldc, 1.25, -0.5)
    torch.cuda.synchronize()
    got = Cd.cpu().numpy()
    ref = orc.spmm(m, k, rp, ci, v, np.ascontiguousarray(Bfull[:, :k]), 1.25, -0.5,
                   np.ascontiguousarray(C0[:, :k]))
    assert rel_err(got[:, :k], ref) <= TOL64
    assert np.array_equal(got[:, k:], C0[:, k:])
    assert A.spmm_info()["blocks_state"] == 1


@pytest.mark.parametrize("k", [1, 2, 3, 4])
@pytest.mark.parametrize("layout", ["row", "col"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spmm_few_columns(pls, orc, gen, k, layout, dtype):
    """k < 5 columns (the reference's repeated-SpMV range, src/matmul.cpp:42-89)
    in one launch that reads A once: B and C row- or column-major with leading
    dimensions wider than the matrix, beta 0 and 0.7, TRANS A, a null C
    (materialised column-major), and a SELL-C-sigma A (one SpMV per column)."""
    rng = np.random.default_rng(40 + k)
    m = 3000
    rp, ci, v = gen.powerlaw(m, 17)
    v = v.astype(dtype)
    vd = v.astype(np.float64)
    tol = TOL64 if dtype == np.float64 else 1e-6
    Bv = rng.uniform(-1, 1, (m, k)).astype(dtype).astype(np.float64)
    C0 = rng.uniform(-1, 1, (m, k)).astype(dtype).astype(np.float64)

    def dense(M):
        if layout == "row":
            ld = k + 3
            buf = np.zeros((m, ld), dtype)
            buf[:, :k] = M
            return pls.SpMat.dense(pls.ROW_MAJOR, m, k, ld, buf.reshape(-1))
        ld = m + 5
        buf = np.zeros((k, ld), dtype)
        buf[:, :m] = M.T
        return pls.SpMat.dense(pls.COL_MAJOR, m, k, ld, buf.reshape(-1))

    A = pls.SpMat.csr(m, m, rp, ci, v)
    B = dense(Bv)
    for beta in (0.0, 0.7):
        Cm = dense(C0)
        pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, 1.25, A, B, beta, Cm)
        ref = orc.spmm(m, k, rp, ci, vd, Bv, 1.25, beta, C0)
        assert rel_err(Cm.export_dense(), ref) <= tol, beta
    Cn = pls.SpMat.null(m, k, dtype)
    pls.spmm_optimize(pls.NOTRANS, pls.NOTRANS, pls.SCALAR_ANY, A, B, pls.SCALAR_ANY, Cn)
    pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, -2.0, A, B, 0.0, Cn)
    assert rel_err(Cn.export_dense(), orc.spmm(m, k, rp, ci, vd, Bv, -2.0)) <= tol
    # op(A) = A^T
    D = np.zeros((m, m))
    for i in range(m):
        D[i, ci[rp[i]:rp[i + 1]]] += vd[rp[i]:rp[i + 1]]
    At = pls.SpMat.csr(m, m, rp, ci, v)
    Ct = dense(C0)
    pls.spmm_exec(pls.TRANS, pls.NOTRANS, 0.5, At, B, 0.7, Ct)
    assert rel_err(Ct.export_dense(), 0.5 * D.T @ Bv + 0.7 * C0) <= tol
    # a SELL-C-sigma plan keeps one SpMV per column
    As = pls.SpMat.csr(m, m, rp, ci, v)
    As.force_sell(32, 1024)
    As.hint(pls.HINT_SPMV_INVOCATIONS, pls.INVOCATIONS_MANY)
    As.optimize()
    Cs = dense(C0)
    pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, 1.25, As, B, 0.7, Cs)
    assert rel_err(Cs.export_dense(), orc.spmm(m, k, rp, ci, vd, Bv, 1.25, 0.7, C0)) <= tol


def block_operand(gen, b, layout, seed):
    """Block-stencil operator with dense N(0,1) b x b blocks stored in `layout`;
    every stored block of block column bci[0] is all zero, so the b rows of B
    that column names meet stored zeros only."""
    brp, bci, bv = gen.block_stencil(6, 5, 4, b, 0x5EED0005 + seed)
    bv = bv.copy().reshape(len(bci), b * b)
    bv[bci == bci[0]] = 0.0
    bv = bv.reshape(-1)
    if layout == "col":
        nb = len(bci)
        bv = bv.reshape(nb, b, b).transpose(0, 2, 1).copy().reshape(-1)
    return brp, bci, bv


@pytest.mark.parametrize("b", [4, 8])
@pytest.mark.parametrize("layout", ["row", "col"])
@pytest.mark.parametrize("k", [5, 13, 32, 64, 100])
def test_spmm_native_bsr(pls, orc, gen, b, layout, k):
    """BSR x dense on the blocks (b = 4, 8; row- and column-major blocks),
    FP64 and FP32, beta 0 and 0.5, widths that take the vector and the
    per-element loads, against the oracle's bsr2csr + product; a stored zero
    block times Inf in B adds nothing (the CSR expansion drops it)."""
    lay = pls.ROW_MAJOR if layout == "row" else pls.COL_MAJOR
    brp, bci, bv = block_operand(gen, b, layout, k)
    m = (len(brp) - 1) * b
    rp, ci, v = orc.bsr2csr(lay, m, m, b, brp, bci, bv)
    rng = np.random.default_rng(100 + k)
    for dtype, tol in ((np.float64, TOL64), (np.float32, 1e-6)):
        Bv = rng.uniform(-1, 1, (m, k)).astype(dtype).astype(np.float64)
        # B rows that only stored zeros meet get an Inf
        zero_rows = np.arange(bci[0] * b, bci[0] * b + b)
        assert not np.isin(zero_rows, ci).any()
        Bv[zero_rows, 0] = np.inf
        C0 = rng.uniform(-1, 1, (m, k)).astype(dtype).astype(np.float64)
        for beta in (0.0, 0.5):
            A = pls.SpMat.bsr(lay, m, m, b, brp, bci, bv.astype(dtype))
            B = pls.SpMat.dense(pls.ROW_MAJOR, m, k, k, Bv.astype(dtype).reshape(-1))
            Cm = pls.SpMat.dense(pls.ROW_MAJOR, m, k, k, C0.astype(dtype).reshape(-1))
            pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, 1.5, A, B, beta, Cm)
            got = Cm.export_dense()
            assert np.isfinite(got).all()
            Bf = np.where(np.isinf(Bv), 0.0, Bv)
            ref = orc.spmm(m, k, rp, ci, v.astype(dtype).astype(np.float64), Bf, 1.5, beta, C0)
            assert rel_err(got, ref) <= tol, (dtype, beta)
            assert A.spmm_info()["blocks_state"] == 2


def test_spmm_native_bsr_async_update_and_precision(pls, orc, gen):
    """The asynchronous SpMM of a BSR handle runs on its blocks too; a value
    update or a precision change drops them (the CSR path then runs) and the
    results follow the updated values."""
    import torch

    brp, bci, bv = block_operand(gen, 8, "row", 3)
    m = (len(brp) - 1) * 8
    k = 24
    rng = np.random.default_rng(5)
    Bv = rng.uniform(-1, 1, (m, k))
    A = pls.SpMat.bsr(pls.ROW_MAJOR, m, m, 8, brp, bci, bv)
    Bd = torch.from_numpy(Bv).cuda()
    Cd = torch.zeros((m, k), dtype=torch.float64, device="cuda")
    A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k, 2.0, 0.0)
    torch.cuda.synchronize()
    rp, ci, v = orc.bsr2csr(pls.ROW_MAJOR, m, m, 8, brp, bci, bv)
    assert rel_err(Cd.cpu().numpy(), orc.spmm(m, k, rp, ci, v, Bv, 2.0)) <= TOL64
    assert A.spmm_info()["blocks_state"] == 2
    # update two stored entries of the (now CSR) object
    rows = np.array([10, 200], dtype=np.int32)
    cols = np.array([ci[rp[10]], ci[rp[200] + 3]], dtype=np.int32)
    A.update(rows, cols, np.array([7.0, -3.0]))
    v2 = v.copy()
    v2[rp[10]] = 7.0
    v2[rp[200] + 3] = -3.0
    A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k, 2.0, 0.0)
    torch.cuda.synchronize()
    assert A.spmm_info()["blocks_state"] != 2
    assert rel_err(Cd.cpu().numpy(), orc.spmm(m, k, rp, ci, v2, Bv, 2.0)) <= TOL64
