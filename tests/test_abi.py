"""The C-ABI library loads, exports every declared entry point, and its host
logic (parameter checks, hints, query, export, update) follows the reference.
None of these tests launches a kernel."""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(perflibs_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol(pls):
    lib = pls.lib()
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_csr_create_parameter_codes(pls):
    A, st = pls.SpMat.csr(-1, 3, [0, 0], [0], [1.0], raise_on_error=False)
    assert st == 1
    A, st = pls.SpMat.csr(2, -3, [0, 0, 0], [0], [1.0], raise_on_error=False)
    assert st == 1
    A, st = pls.SpMat.csr(2, 3, [2, 2, 2], [0], [1.0], raise_on_error=False)
    assert st == 1


def test_bsr_create_codes_match_oracle(pls, orc):
    cases = [(91, 4, 4, 2, [0, 1, 2]), (92, 4, 4, 2, [0, 0, 0]), (91, -4, 4, 2, [0, 0, 0]),
             (91, 4, -4, 2, [0, 0, 0]), (91, 4, 4, 3, [0, 0, 0]), (91, 4, 4, 2, [2, 2, 2])]
    for layout, m, n, b, rp in cases:
        _, st = pls.SpMat.bsr(layout, m, n, b, rp, [0, 1], np.ones(8), raise_on_error=False)
        code = orc.bsr_check(layout, m, n, b, rp)
        assert (st == 0) == (code == 0)
        if code:
            assert st == 1


def test_dense_create_codes(pls):
    ok = np.ones(6)
    assert pls.SpMat.dense(pls.ROW_MAJOR, 2, 3, 3, ok, raise_on_error=False)[1] == 0
    assert pls.SpMat.dense(95, 2, 3, 3, ok, raise_on_error=False)[1] == 1
    assert pls.SpMat.dense(pls.ROW_MAJOR, 2, 3, 2, ok, raise_on_error=False)[1] == 1  # lda < n
    assert pls.SpMat.dense(pls.COL_MAJOR, 2, 3, 1, ok, raise_on_error=False)[1] == 1  # lda < m


def test_query_and_index_base(pls):
    A = pls.SpMat.csr(3, 4, [1, 2, 4, 5], [1, 2, 4, 3], [1.0, 2.0, 3.0, 4.0])
    assert A.query() == (1, 3, 4, 4)
    B = pls.SpMat.bsr(pls.ROW_MAJOR, 4, 4, 2, [0, 1, 2], [0, 1], np.ones(8))
    assert B.query() == (0, 4, 4, 8)


def test_hint_validation(pls):
    A = pls.SpMat.csr(2, 2, [0, 1, 2], [0, 1], [1.0, 2.0])
    assert A.hint(pls.HINT_SPMV_OPERATION, pls.TRANS) == 0
    assert A.hint(pls.HINT_SPMV_OPERATION, pls.INVOCATIONS_MANY) == 1
    assert A.hint(pls.HINT_SPMV_INVOCATIONS, pls.INVOCATIONS_FEW) == 0
    assert A.hint(pls.HINT_STRUCTURE, pls.STRUCTURE_UNSTRUCTURED) == 0
    assert A.hint(pls.HINT_STRUCTURE, 180) == 0  # HPCG
    assert A.hint(pls.HINT_STRUCTURE, 999) == 1
    assert A.hint(pls.HINT_MEMORY, 200) == 0
    assert A.hint(12345, 300) == 1


def test_invalid_transpose_is_parameter_error(pls):
    A = pls.SpMat.csr(2, 2, [0, 1, 2], [0, 1], [1.0, 2.0])
    assert A.spmv(np.ones(2), np.zeros(2), trans=999, raise_on_error=False) == 1


@pytest.mark.parametrize("base", [0, 1])
def test_export_csr_roundtrip(pls, base):
    rp = np.array([0, 2, 2, 5]) + base
    ci = np.array([0, 3, 1, 2, 3]) + base
    v = np.array([1.0, 0.0, 3.0, 4.0, 5.0])
    A = pls.SpMat.csr(3, 4, rp, ci, v)
    for out_base in (0, 1):
        m, n, orp, oci, ov = A.export_csr(out_base)
        assert (m, n) == (3, 4)
        assert np.array_equal(orp, rp - base + out_base)
        assert np.array_equal(oci, ci - base + out_base)
        assert np.array_equal(ov, v)  # explicit zeros survive on a CSR object


def test_export_bsr_as_csr_matches_oracle(pls, orc):
    rng = np.random.default_rng(2)
    vals = rng.standard_normal(3 * 9)
    vals[[1, 5, 13]] = 0.0
    for layout in (pls.COL_MAJOR, pls.ROW_MAJOR):
        A = pls.SpMat.bsr(layout, 6, 6, 3, [0, 2, 2, 3], [0, 1, 2], vals)
        _, _, rp, ci, v = A.export_csr(0)
        erp, eci, ev = orc.bsr2csr(layout, 6, 6, 3, [0, 2, 2, 3], [0, 1, 2], vals)
        assert np.array_equal(rp, erp) and np.array_equal(ci, eci) and np.array_equal(v, ev)


def test_export_bsr_goldens(pls, golden):
    for g in golden["export_bsr"]:
        A = pls.SpMat.bsr(g["layout"], g["m"], g["n"], g["block_size"], g["row_ptr"],
                          g["col_indx"], np.array(g["vals"]))
        m, n, b, rp, ci, v = A.export_bsr(g["export_layout"], g["export_base"])
        assert (m, n, b) == (g["m"], g["n"], g["block_size"])
        assert rp.tolist() == g["expected_row_ptr"] and ci.tolist() == g["expected_col_indx"]
        assert v.tolist() == g["expected_vals"]
        if "expected_dense_row_major" in g:
            assert A.export_dense(pls.ROW_MAJOR).reshape(-1).tolist() == g["expected_dense_row_major"]


def test_export_bsr_layouts_and_other_formats(pls, orc):
    rng = np.random.default_rng(3)
    b, brp, bci = 3, [0, 2, 2, 3], [0, 1, 2]
    vals = rng.standard_normal(3 * 9)
    for layout in (pls.COL_MAJOR, pls.ROW_MAJOR):
        A = pls.SpMat.bsr(layout, 9, 9, b, brp, bci, vals)
        other = pls.ROW_MAJOR if layout == pls.COL_MAJOR else pls.COL_MAJOR
        _, _, bs, rp, ci, v = A.export_bsr(other, 0)
        assert bs == b and rp.tolist() == brp and ci.tolist() == bci
        # every block transposed into the other layout, same dense image
        assert np.array_equal(v.reshape(-1, b, b), vals.reshape(-1, b, b).transpose(0, 2, 1))
        B = pls.SpMat.bsr(other, 9, 9, b, rp, ci, v)
        assert np.array_equal(B.export_dense(), A.export_dense())
    # a CSR object exports as 1 x 1 blocks of its CSR arrays (explicit zeros kept)
    rp0, ci0, v0 = [1, 3, 3, 4], [1, 3, 2], [2.0, 0.0, -1.0]
    C = pls.SpMat.csr(3, 3, rp0, ci0, v0)
    m, n, bs, rp, ci, v = C.export_bsr(pls.COL_MAJOR, 0)
    assert (m, n, bs) == (3, 3, 1)
    assert rp.tolist() == [0, 2, 2, 3] and ci.tolist() == [0, 2, 1] and v.tolist() == v0
    # a dense object converts to CSR first (zeros dropped), then 1 x 1 blocks
    D = pls.SpMat.dense(pls.ROW_MAJOR, 2, 2, 2, np.array([0.0, 5.0, 6.0, 0.0]))
    m, n, bs, rp, ci, v = D.export_bsr(pls.ROW_MAJOR, 1)
    assert bs == 1 and rp.tolist() == [1, 2, 3] and ci.tolist() == [2, 1] and v.tolist() == [5.0, 6.0]
    lib = pls.lib()
    assert lib.perflibs_spmat_export_bsr_d(C.h, 77, 0, None, None, None, None, None, None) == 1


def test_export_dense_of_sparse_and_dense(pls):
    A = pls.SpMat.csr(2, 3, [0, 2, 3], [0, 2, 1], [1.0, 2.0, 3.0])
    D = np.array([[1.0, 0.0, 2.0], [0.0, 3.0, 0.0]])
    assert np.array_equal(A.export_dense(pls.ROW_MAJOR), D)
    assert np.array_equal(A.export_dense(pls.COL_MAJOR), D)
    vals = np.arange(12, dtype=np.float64)
    Bm = pls.SpMat.dense(pls.COL_MAJOR, 3, 2, 6, vals)  # lda 6 > m: padded input
    exp = np.stack([vals[0:3], vals[6:9]], axis=1)
    assert np.array_equal(Bm.export_dense(pls.ROW_MAJOR), exp)
    N = pls.SpMat.null(2, 3)
    assert np.array_equal(N.export_dense(), np.zeros((2, 3)))
    I = pls.SpMat.identity(3)
    assert np.array_equal(I.export_dense(), np.eye(3))


def test_update_semantics(pls):
    A = pls.SpMat.csr(3, 3, [1, 2, 4, 5], [1, 1, 3, 2], [1.0, 2.0, 3.0, 4.0])
    assert A.update([1], [1], [9.0], raise_on_error=False) == 0
    assert A.update([2], [3], [7.0], raise_on_error=False) == 0
    _, _, _, _, v = A.export_csr(1)
    assert v.tolist() == [9.0, 2.0, 7.0, 4.0]
    # not an existing non-zero -> parameter error
    assert A.update([1, 3], [1, 1], [5.0, 5.0], raise_on_error=False) == 1
    assert A.update([4], [1], [5.0], raise_on_error=False) == 1
    assert pls.lib().perflibs_spmat_update_d(A.h, -1, None, None, None) == 1
    assert pls.lib().perflibs_spmat_update_d(A.h, 0, None, None, None) == 0


def test_update_bsr(pls):
    A = pls.SpMat.bsr(pls.ROW_MAJOR, 4, 4, 2, [0, 1, 2], [0, 1], np.arange(8, dtype=np.float64))
    assert A.update([1, 2], [0, 3], [-1.0, -2.0], raise_on_error=False) == 0
    D = A.export_dense()
    assert D[1, 0] == -1.0 and D[2, 3] == -2.0
    assert A.update([0], [3], [1.0], raise_on_error=False) == 1  # block (0,1) absent


def test_spmm_parameter_checks(pls):
    A = pls.SpMat.csr(2, 3, [0, 1, 2], [0, 1], [1.0, 2.0])
    B = pls.SpMat.dense(pls.ROW_MAJOR, 3, 6, 6, np.ones(18))
    Cbad = pls.SpMat.null(3, 6)
    assert pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, 1.0, A, B, 0.0, Cbad, raise_on_error=False) == 1
    Cs = pls.SpMat.csr(2, 6, [0, 0, 0], [0], [0.0])
    # a beta hint of zero needs C null or dense (parameter 7, src/matmul.cpp:1290-1306)
    st = pls.spmm_optimize(pls.NOTRANS, pls.NOTRANS, pls.SCALAR_ANY, A, B, pls.SCALAR_ZERO, Cs,
                           raise_on_error=False)
    assert st == 1
    Bf = pls.SpMat.dense(pls.ROW_MAJOR, 3, 6, 6, np.ones(18, np.float32))
    C = pls.SpMat.null(2, 6)
    assert pls.spmm_exec(pls.NOTRANS, pls.NOTRANS, 1.0, A, Bf, 0.0, C, raise_on_error=False) == 1
    assert pls.spmm_exec(999, pls.NOTRANS, 1.0, A, B, 0.0, C, raise_on_error=False) == 1


def test_coo_ingest_sorted_like_coo2csr(pls):
    """COO entries become a 0-based CSR sorted by (row, column) like the
    reference's coo2csr (src/compressed_sparse_rows.cpp:590-640); duplicates stay."""
    rows = [2, 0, 2, 1, 0, 2]
    cols = [1, 3, 0, 2, 0, 1]
    vals = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    A = pls.SpMat.coo(3, 4, rows, cols, vals)
    assert A.query() == (0, 3, 4, 6)
    m, n, rp, ci, v = A.export_csr(0)
    assert (m, n) == (3, 4)
    assert rp.tolist() == [0, 2, 3, 6]
    assert ci.tolist() == [0, 3, 2, 0, 1, 1]
    assert v.tolist() == [5.0, 2.0, 4.0, 3.0, 1.0, 6.0]  # duplicate (2,1) keeps input order
    S = pls.SpMat.coo(3, 4, rows, cols, vals.astype(np.float32))
    assert S.export_csr(0)[4].dtype == np.float32
    assert pls.SpMat.coo(-1, 4, rows, cols, vals, raise_on_error=False)[1] == 1
    assert pls.SpMat.coo(3, -4, rows, cols, vals, raise_on_error=False)[1] == 1
    assert pls.SpMat.coo(3, 4, [3], [0], [1.0], raise_on_error=False)[1] == 1
    assert pls.SpMat.coo(3, 4, [0], [4], [1.0], raise_on_error=False)[1] == 1
    E = pls.SpMat.coo(2, 2, [], [], np.zeros(0))
    assert E.export_csr(0)[2].tolist() == [0, 0, 0]


@pytest.mark.parametrize("base", [0, 1])
def test_csc_ingest_matches_transpose(pls, base):
    """CSC arrays of A are the CSR arrays of A^T; the ingest rebuilds A's CSR
    (csc2csr, src/compressed_sparse_rows.cpp:749-811) and keeps the index base."""
    rng = np.random.default_rng(9 + base)
    m, n = 7, 5
    D = np.where(rng.random((m, n)) < 0.4, rng.standard_normal((m, n)), 0.0)
    col_ptr = np.concatenate([[0], np.cumsum((D != 0).sum(axis=0))]) + base
    row_indx = np.concatenate([np.nonzero(D[:, j])[0] for j in range(n)]) + base
    vals = np.concatenate([D[np.nonzero(D[:, j])[0], j] for j in range(n)])
    A = pls.SpMat.csc(m, n, row_indx, col_ptr, vals)
    assert A.query() == (base, m, n, int((D != 0).sum()))
    _, _, rp, ci, v = A.export_csr(base)
    E = np.zeros((m, n))
    for i in range(m):
        assert np.all(np.diff(ci[rp[i] - base:rp[i + 1] - base]) > 0)
        E[i, ci[rp[i] - base:rp[i + 1] - base] - base] = v[rp[i] - base:rp[i + 1] - base]
    assert np.array_equal(E, D)
    assert pls.SpMat.csc(m, n, row_indx, col_ptr + 2, vals, raise_on_error=False)[1] == 1
    assert pls.SpMat.csc(-1, n, row_indx, col_ptr, vals, raise_on_error=False)[1] == 1


def test_introspection_without_plans(pls):
    """Plan introspection on a handle that has no device plan yet is a
    parameter error (status 1), not a crash."""
    import ctypes as C

    A = pls.SpMat.csr(3, 3, [0, 1, 2, 3], [0, 1, 2], [1.0, 2.0, 3.0])
    v = [C.c_int32() for _ in range(5)]
    tot = C.c_int64()
    lib = pls.lib()
    assert lib.perflibs_amd_spmm_info(A.h, *[C.byref(x) for x in v[:3]], C.byref(tot)) == 1
    assert lib.perflibs_amd_panel_info(A.h, *[C.byref(x) for x in v]) == 1
    assert lib.perflibs_amd_spmm_info(None, None, None, None, None) == 1
    clustered = C.c_int(7)
    assert lib.perflibs_amd_spmm_rows_clustered(A.h, C.byref(clustered)) == 1
    assert clustered.value == 7  # left alone
    assert lib.perflibs_amd_spmm_rows_clustered(None, C.byref(clustered)) == 1
    # the row-block mode is checked and stored without a device plan
    for rows in (0, 1, 2):
        assert lib.perflibs_amd_set_spmm_rows(A.h, rows) == 0
    for rows in (-1, 3):
        assert lib.perflibs_amd_set_spmm_rows(A.h, rows) == 1
    assert lib.perflibs_amd_set_spmm_rows(None, 0) == 1


def test_noalloc_backward_refuses_sizes_beyond_its_scratch(pls):
    """The allocation-free MLP backward checks n against its static scratch
    (4096 samples) before touching the device; pointers are never read."""
    L = pls.lib()
    p = 0x1000
    for n in (4097, 1 << 20):
        assert L.perflibs_amd_mlp_backward_noalloc(p, p, n, p, p, p, p, p, p, p) == 1
        assert b"static scratch" in L.perflibs_amd_last_error()
    assert L.perflibs_amd_mlp_backward_noalloc(p, p, 0, p, p, p, p, p, p, p) == 1
