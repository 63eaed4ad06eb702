"""Matrix Market reader (workloads/mtx.cpp) against scipy.io.mmread and the
acceptance rules of the reference's benchmark reader
(bench/pl_bench_sparse_x_dense.c:124-240): header check, complex refused,
pattern entries = 1, symmetric / hermitian / skew-symmetric expansion,
(row, column)-sorted CSR with duplicates kept, the reference's error texts."""
from __future__ import annotations

import zlib

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from plsparse import workloads as W


def write(tmp_path, text, name="a.mtx"):
    p = tmp_path / name
    p.write_text(text)
    return p


def dense_of(m, n, rp, ci, v):
    D = np.zeros((m, n))
    for i in range(m):
        np.add.at(D[i], ci[rp[i]:rp[i + 1]], v[rp[i]:rp[i + 1]])
    return D


def check_sorted(m, rp, ci):
    assert rp[0] == 0 and np.all(np.diff(rp) >= 0)
    for i in range(m):
        seg = ci[rp[i]:rp[i + 1]]
        assert np.all(np.diff(seg) >= 0)


@pytest.mark.parametrize("sym", ["general", "symmetric", "skew-symmetric"])
@pytest.mark.parametrize("field", ["real", "integer", "pattern"])
def test_against_scipy(tmp_path, sym, field):
    if field == "pattern" and sym == "skew-symmetric":
        pytest.skip("the format has no skew-symmetric pattern matrices")
    rng = np.random.default_rng(zlib.crc32(f"{sym}/{field}".encode()))
    m = n = 57
    A = sp.random(m, n, density=0.08, random_state=rng, format="coo")
    if field == "integer":
        A.data = np.round(A.data * 20 - 10)
    if sym == "symmetric":
        A = sp.tril(A + A.T).tocoo()
    elif sym == "skew-symmetric":
        A = sp.tril(A - A.T, k=-1).tocoo()
    p = tmp_path / "m.mtx"
    scipy.io.mmwrite(str(p), A, field=field, symmetry=sym)
    m2, n2, rp, ci, v = W.read_mtx(p)
    assert (m2, n2) == (m, n)
    check_sorted(m, rp, ci)
    ref = scipy.io.mmread(str(p)).toarray()
    assert np.array_equal(dense_of(m, n, rp, ci, v), ref)
    # the expansion stores both triangles (no summing): nnz = off-diagonal count * 2 + diagonal
    nd = int(np.count_nonzero(A.row == A.col))
    assert len(ci) == (A.nnz if sym == "general" else 2 * A.nnz - nd)


def test_duplicates_kept_in_file_order_and_comments(tmp_path):
    p = write(tmp_path, "%%MatrixMarket matrix coordinate real general\n"
                        "% a comment\n%another\n"
                        "3 4 6\n"
                        "3 4 1.5\n1 2 2.0\n1 2 -7e-1\n"
                        "2 1 4\n  1   1   0.25\n3 1 9\n")
    m, n, rp, ci, v = W.read_mtx(p)
    assert (m, n) == (3, 4)
    assert rp.tolist() == [0, 3, 4, 6]
    assert ci.tolist() == [0, 1, 1, 0, 0, 3]
    assert v.tolist() == [0.25, 2.0, -0.7, 4.0, 9.0, 1.5]


def test_hermitian_real_mirrors_and_entries_across_lines(tmp_path):
    p = write(tmp_path, "%%MatrixMarket matrix coordinate real hermitian\n2 2 2\n1 1 3\n2\n1\n-5\n")
    m, n, rp, ci, v = W.read_mtx(p)
    assert dense_of(m, n, rp, ci, v).tolist() == [[3.0, -5.0], [-5.0, 0.0]]


def test_unknown_symmetry_keeps_entries(tmp_path):
    # like the reference, only the three known words expand the entries
    p = write(tmp_path, "%%MatrixMarket matrix coordinate real Symmetric\n2 2 1\n2 1 3\n")
    m, n, rp, ci, v = W.read_mtx(p)
    assert rp.tolist() == [0, 0, 1] and ci.tolist() == [0] and v.tolist() == [3.0]


@pytest.mark.parametrize("text,msg", [
    ("", "failed to read Matrix Market header"),
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", "expected MatrixMarket coordinate file"),
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", "complex Matrix Market files not supported"),
    ("%%MatrixMarket matrix coordinate real general\n% only comments\n", "missing size line"),
    ("%%MatrixMarket matrix coordinate real general\n\n2 2 1\n1 1 1\n", "bad size line"),
    ("%%MatrixMarket matrix coordinate real general\n2 0 1\n1 1 1\n", "invalid matrix dimensions"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n2 2\n", "bad numeric entry"),
    ("%%MatrixMarket matrix coordinate pattern general\n2 2 2\n1 1\n2\n", "bad pattern entry"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", "entry index out of range"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 0 1\n", "entry index out of range"),
    ("%%MatrixMarket matrix coordinate real symmetric\n3 2 1\n3 1 1\n", "entry index out of range"),
])
def test_errors(tmp_path, text, msg):
    p = write(tmp_path, text)
    with pytest.raises(ValueError, match=msg):
        W.read_mtx(p)


def test_missing_file(tmp_path):
    with pytest.raises(ValueError, match="failed to open .mtx file"):
        W.read_mtx(tmp_path / "absent.mtx")
