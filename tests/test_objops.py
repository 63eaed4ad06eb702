"""Norms, in-place scaling and transposition of matrix objects, CSC / COO export,
and the entry points outside the sparse x dense path.

The known-answer cases replay the reference's own tests (test/matrix_object_ops.c,
test/dense_update_paths.c, test/format_edge_ops.c, test/bsr_kernel_paths.c) from
the committed fixtures; the property tests pin the rest against numpy / scipy
(norms, transposes and CSC / COO forms of random matrices).  All of it is host
logic: no test here launches a kernel."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp


def dense_rm(A, pls, m, n):
    return A.export_dense(pls.ROW_MAJOR).reshape(m, n)


def random_csr(rng, m, n, density, dtype=np.float64, zeros=False):
    M = sp.random(m, n, density=density, random_state=rng, format="csr", dtype=np.float64)
    M.sort_indices()
    vals = rng.standard_normal(M.nnz).astype(dtype)
    if zeros and M.nnz:
        vals[:: max(1, M.nnz // 7)] = 0.0  # explicit zeros stay stored in a CSR object
    return M.indptr.astype(np.int32), M.indices.astype(np.int32), vals


def to_dense(m, n, rp, ci, v):
    D = np.zeros((m, n), dtype=np.float64)
    base = int(rp[0])
    for r in range(m):
        for j in range(rp[r] - base, rp[r + 1] - base):
            D[r, ci[j] - base] += v[j]
    return D


# ---------------------------------------------------------------------------
# Known answers of the reference's tests
# ---------------------------------------------------------------------------
def test_csr_object_sequence(pls, golden):
    """test/matrix_object_ops.c:74-274: update, scale 0.5, norms, transpose,
    CSC / COO export, NOCOPY refusal."""
    g = golden["objops_csr"]
    A = pls.SpMat.csr(g["m"], g["n"], g["row_ptr"], g["col_indx"], np.array(g["vals"]))
    _, _, ri, cp, v = A.export_csc(0)
    assert ri.tolist() == g["expected_csc_row_indx"] and cp.tolist() == g["expected_csc_col_ptr"]
    np.testing.assert_allclose(v, g["expected_csc_vals"], atol=g["tol"])
    m, n, r, c, v = A.export_coo()
    assert (m, n) == (4, 4)
    assert r.tolist() == g["expected_coo_row_indx"] and c.tolist() == g["expected_coo_col_indx"]
    np.testing.assert_allclose(v, g["vals"], atol=g["tol"])

    A.update(g["update_rows"], g["update_cols"], g["update_vals"])
    np.testing.assert_allclose(dense_rm(A, pls, 4, 4).ravel(), g["expected_updated_dense_row_major"],
                               atol=g["tol"])
    A.scale(g["scale"])
    assert A.norm(pls.NORM_INF) == pytest.approx(g["expected_inf_norm_after_scale"], abs=g["tol"])
    assert A.norm(pls.NORM_FRB) == pytest.approx(g["expected_frb_norm_after_scale"], abs=g["tol"])
    A.transpose(pls.TRANS)
    np.testing.assert_allclose(dense_rm(A, pls, 4, 4).ravel(),
                               g["expected_scaled_transposed_row_major"], atol=g["tol"])

    # a COO object exports the CSC arrays of the same matrix
    B = pls.SpMat.coo(4, 4, g["expected_coo_row_indx"], g["expected_coo_col_indx"], np.array(g["vals"]))
    _, _, ri, cp, v = B.export_csc(0)
    assert ri.tolist() == g["expected_csc_row_indx"] and cp.tolist() == g["expected_csc_col_ptr"]
    np.testing.assert_allclose(v, g["expected_csc_vals"], atol=g["tol"])

    keep = np.array(g["nocopy_dense_row_major"])
    N = pls.SpMat.dense(pls.ROW_MAJOR, 4, 4, 4, keep, flags=pls.CREATE_NOCOPY)
    assert N.scale(2.0, raise_on_error=False) == g["nocopy_scale_status"]
    assert "NO_COPY" in pls.lib().perflibs_amd_last_error().decode()
    assert N.transpose(pls.TRANS, raise_on_error=False) == g["nocopy_transpose_status"]
    assert keep.tolist() == g["nocopy_dense_row_major"]  # the user's values are untouched


def test_dense_object_sequence(pls, golden):
    """test/dense_update_paths.c:12-104: update, norms, scale, transpose, CSR round trip."""
    g = golden["dense_update_scale_transpose"]
    A = pls.SpMat.dense(g["layout"], g["m"], g["n"], g["lda"], np.array(g["vals"]))
    A.update(g["update_rows"], g["update_cols"], g["update_vals"])
    np.testing.assert_allclose(dense_rm(A, pls, 4, 4).ravel(), g["expected_after_update_row_major"],
                               atol=g["tol"])
    cm = A.export_dense(pls.COL_MAJOR)  # returned as the (m, n) matrix
    np.testing.assert_allclose(cm.T.ravel(), g["expected_after_update_col_major"], atol=g["tol"])
    assert A.norm(pls.NORM_INF) == pytest.approx(g["expected_inf_norm"], abs=g["tol"])
    assert A.norm(pls.NORM_FRB) == pytest.approx(g["expected_frb_norm"], abs=g["tol"])
    A.scale(g["scale"])
    np.testing.assert_allclose(dense_rm(A, pls, 4, 4).ravel(), g["expected_after_scale_row_major"],
                               atol=g["tol"])
    A.transpose(pls.TRANS)
    assert A.query()[1:3] == (4, 4)
    np.testing.assert_allclose(dense_rm(A, pls, 4, 4).ravel(),
                               g["expected_after_transpose_row_major"], atol=g["tol"])
    m, n, rp, ci, v = A.export_csr(0)
    R = pls.SpMat.csr(m, n, rp, ci, v)
    np.testing.assert_allclose(dense_rm(R, pls, 4, 4).ravel(),
                               g["expected_after_transpose_row_major"], atol=g["tol"])


def test_bsr_transpose_then_scale(pls, golden):
    """test/format_edge_ops.c:95-144: a 1-based column-major BSR object with
    explicit zeros in its blocks, transposed, then scaled."""
    g = golden["bsr_col_major_transpose_then_scale"]
    A = pls.SpMat.bsr(g["layout"], g["m"], g["n"], g["block_size"], g["row_ptr"], g["col_indx"],
                      np.array(g["vals"]))
    A.transpose(pls.TRANS)
    np.testing.assert_allclose(dense_rm(A, pls, 4, 4).ravel(), g["expected_transposed_row_major"],
                               atol=g["tol"])
    A.scale(g["scale"])
    np.testing.assert_allclose(dense_rm(A, pls, 4, 4).ravel(), g["expected_scaled_row_major"],
                               atol=g["tol"])


def test_bsr_norm_transpose_scale(pls, golden):
    """test/bsr_kernel_paths.c:87-132: infinity norm of a column-major BSR
    object, then transpose and scale by 0.5; the reference then multiplies by
    x = 1 (here by the exported matrix, on the host)."""
    g = golden["bsr_col_major_transposed_half"]
    A = pls.SpMat.bsr(g["layout"], g["m"], g["n"], g["block_size"], g["row_ptr"], g["col_indx"],
                      np.array(g["vals"]))
    assert A.norm(pls.NORM_INF) == pytest.approx(g["inf_norm"], abs=g["tol"])
    A.transpose(pls.TRANS)
    A.scale(0.5)
    assert A.query()[1:3] == (4, 4)
    D = dense_rm(A, pls, 4, 4)
    np.testing.assert_allclose(D @ np.array(g["x"]), g["expected_trans"], atol=g["tol"])


# ---------------------------------------------------------------------------
# Properties against numpy / scipy
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("base", [0, 1])
def test_csr_transpose_scale_norms_match_numpy(pls, dtype, base):
    rng = np.random.default_rng(0x0B1 + base)
    m, n = 37, 23
    rp, ci, v = random_csr(rng, m, n, 0.2, dtype, zeros=True)
    D = to_dense(m, n, rp, ci, v.astype(np.float64))
    A = pls.SpMat.csr(m, n, rp + base, ci + base, v)
    tol = 1e-12 if dtype == np.float64 else 1e-5
    assert A.norm(pls.NORM_INF) == pytest.approx(np.linalg.norm(D, np.inf), rel=tol)
    assert A.norm(pls.NORM_FRB) == pytest.approx(np.linalg.norm(D, "fro"), rel=tol)
    A.transpose(pls.TRANS)
    assert A.query()[1:] == (n, m, int(rp[-1]))
    np.testing.assert_allclose(dense_rm(A, pls, n, m), D.T, rtol=tol, atol=0)
    A.scale(-2.5)
    np.testing.assert_allclose(dense_rm(A, pls, n, m), -2.5 * D.T, rtol=tol, atol=0)
    assert A.norm(pls.NORM_INF) == pytest.approx(2.5 * np.linalg.norm(D.T, np.inf), rel=tol)
    A.transpose(pls.CONJTRANS)  # real values: the plain transpose
    np.testing.assert_allclose(dense_rm(A, pls, m, n), -2.5 * D, rtol=tol, atol=0)
    # the stored entries of a row of the transpose keep the order of the rows
    _, _, trp, tci, _ = A.export_csr(0)
    for r in range(m):
        seg = tci[trp[r]:trp[r + 1]]
        assert (np.diff(seg) > 0).all()


@pytest.mark.parametrize("layout_name", ["ROW_MAJOR", "COL_MAJOR"])
def test_bsr_scale_keeps_blocks_and_transpose_converts(pls, layout_name):
    rng = np.random.default_rng(7)
    layout = getattr(pls, layout_name)
    b, mb, nb = 3, 4, 5
    pattern = rng.random((mb, nb)) < 0.5
    pattern[:, 0] = True
    rp = np.concatenate([[0], np.cumsum(pattern.sum(1))]).astype(np.int32)
    ci = np.concatenate([np.flatnonzero(row) for row in pattern]).astype(np.int32)
    vals = rng.standard_normal(len(ci) * b * b)
    D = np.zeros((mb * b, nb * b))
    for r in range(mb):
        for p in range(rp[r], rp[r + 1]):
            blk = vals[p * b * b:(p + 1) * b * b].reshape(b, b)
            D[r * b:(r + 1) * b, ci[p] * b:(ci[p] + 1) * b] = blk if layout == pls.ROW_MAJOR else blk.T
    A = pls.SpMat.bsr(layout, mb * b, nb * b, b, rp, ci, vals)
    assert A.norm(pls.NORM_INF) == pytest.approx(np.linalg.norm(D, np.inf), rel=1e-12)
    assert A.norm(pls.NORM_FRB) == pytest.approx(np.linalg.norm(D, "fro"), rel=1e-12)
    A.scale(3.0)
    assert A.export_bsr(layout)[2] == b  # still a BSR object
    np.testing.assert_allclose(dense_rm(A, pls, mb * b, nb * b), 3.0 * D, rtol=1e-12)
    A.transpose(pls.TRANS)
    assert A.export_bsr(layout)[2] == 1  # now CSR: exported as 1 x 1 blocks
    np.testing.assert_allclose(dense_rm(A, pls, nb * b, mb * b), 3.0 * D.T, rtol=1e-12)


@pytest.mark.parametrize("layout_name", ["ROW_MAJOR", "COL_MAJOR"])
def test_dense_transpose_rectangular_with_padding(pls, layout_name):
    layout = getattr(pls, layout_name)
    m, n = 3, 5
    D = np.arange(1.0, 1.0 + m * n).reshape(m, n)
    lda = (n if layout == pls.ROW_MAJOR else m) + 2
    outer = m if layout == pls.ROW_MAJOR else n
    buf = np.full(outer * lda, -99.0)
    for i in range(m):
        for j in range(n):
            buf[i * lda + j if layout == pls.ROW_MAJOR else j * lda + i] = D[i, j]
    A = pls.SpMat.dense(layout, m, n, lda, buf)
    assert A.norm(pls.NORM_INF) == pytest.approx(np.linalg.norm(D, np.inf))
    A.transpose(pls.TRANS)
    assert A.query()[1:3] == (n, m)
    np.testing.assert_array_equal(A.export_dense(pls.ROW_MAJOR), D.T)
    np.testing.assert_array_equal(A.export_dense(pls.COL_MAJOR), D.T)
    A.scale(0.25)
    np.testing.assert_array_equal(A.export_dense(pls.ROW_MAJOR), 0.25 * D.T)


def test_frobenius_norm_does_not_overflow_or_underflow(pls):
    big = np.array([3e200, 4e200, 1e200])
    A = pls.SpMat.csr(1, 3, [0, 3], [0, 1, 2], big)
    assert A.norm(pls.NORM_FRB) == pytest.approx(np.sqrt(26.0) * 1e200, rel=1e-14)
    tiny = np.array([3e-200, 4e-200])
    T = pls.SpMat.csr(2, 2, [0, 1, 2], [0, 1], tiny)
    assert T.norm(pls.NORM_FRB) == pytest.approx(5e-200, rel=1e-14)
    f = np.array([3e30, 4e30], dtype=np.float32)  # squares overflow float32
    F = pls.SpMat.csr(1, 2, [0, 2], [0, 1], f)
    assert F.norm(pls.NORM_FRB) == pytest.approx(5e30, rel=1e-6)


def test_special_objects(pls):
    I = pls.SpMat.identity(9)
    assert I.norm(pls.NORM_INF) == 1.0
    assert I.norm(pls.NORM_FRB) == pytest.approx(3.0)
    I.scale(2.0)  # becomes an explicit CSR diagonal
    np.testing.assert_array_equal(I.export_dense(pls.ROW_MAJOR), 2.0 * np.eye(9))
    Z = pls.SpMat.null(3, 7)
    assert Z.norm(pls.NORM_FRB) == 0.0 and Z.norm(pls.NORM_INF) == 0.0
    Z.transpose(pls.TRANS)
    assert Z.query()[1:3] == (7, 3)
    E = pls.SpMat.csr(2, 2, [0, 0, 0], [], np.zeros(0))
    assert E.norm(pls.NORM_INF) == 0.0
    A = pls.SpMat.csr(2, 2, [0, 1, 2], [0, 1], np.array([1.0, 2.0]))
    st, _ = A.norm(999, raise_on_error=False)
    assert st == 1


def test_csc_and_coo_export_match_scipy(pls):
    rng = np.random.default_rng(11)
    m, n = 41, 29
    rp, ci, v = random_csr(rng, m, n, 0.15)
    S = sp.csr_matrix((v, ci, rp), shape=(m, n))
    for base in (0, 1):
        A = pls.SpMat.csr(m, n, rp + base, ci + base, v)
        for eb in (0, 1):
            mm, nn, ri, cp, cv = A.export_csc(eb)
            Sc = S.tocsc()
            Sc.sort_indices()
            assert (mm, nn) == (m, n)
            assert cp.tolist() == (Sc.indptr + eb).tolist()
            assert ri.tolist() == (Sc.indices + eb).tolist()
            np.testing.assert_array_equal(cv, Sc.data)
        mm, nn, r, c, cv = A.export_coo()
        Sq = S.tocoo()
        assert r.tolist() == (Sq.row + base).tolist() and c.tolist() == (Sq.col + base).tolist()
        np.testing.assert_array_equal(cv, Sq.data)
    # dense and BSR objects export their CSR form's triplets
    D = pls.SpMat.dense(pls.ROW_MAJOR, m, n, n, S.toarray())
    _, _, r, c, cv = D.export_coo()
    Sq = S.tocoo()
    assert r.tolist() == Sq.row.tolist() and c.tolist() == Sq.col.tolist()
    np.testing.assert_array_equal(cv, Sq.data)


def test_off_path_entry_points_fail_cleanly(pls):
    lib = pls.lib()
    h = C.c_void_p(12345)
    fn = lib.perflibs_spmat_create_csr_c
    fn.restype = C.c_int32
    assert fn(C.byref(h), 2, 2, None, None, None, 0) == 2
    assert h.value is None  # no object is created
    assert "perflibs_spmat_create_csr_c" in lib.perflibs_amd_last_error().decode()
    v = C.c_void_p(1)
    fn = lib.perflibs_spvec_create_d
    fn.restype = C.c_int32
    assert fn(C.byref(v), 0, 4, 0, None, None, 0) == 2 and v.value is None
    A = pls.SpMat.csr(2, 2, [0, 1, 2], [0, 1], np.array([1.0, 2.0]))
    B = pls.SpMat.csr(2, 2, [0, 1, 2], [0, 1], np.array([1.0, 2.0]))
    out = C.c_void_p()
    fn = lib.perflibs_spadd_exec_d
    fn.restype = C.c_int32
    fn.argtypes = [C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]
    assert fn(pls.NOTRANS, pls.NOTRANS, 1.0, A.h, 1.0, B.h, out) == 2
    assert "perflibs_spadd_exec_d" in lib.perflibs_amd_last_error().decode()
    # the handle still works for the sparse x dense path and its host logic
    assert A.norm(pls.NORM_FRB) == pytest.approx(np.sqrt(5.0))
