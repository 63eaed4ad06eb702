"""Row blocks of the wide-B SpMM plan, inspected on the host (no device):
consecutive and clustered cuts of CSR patterns checked against their limits
(every row in exactly one block, at most list_max distinct columns and 7 *
list_max non-zeros per list block of lists of 128, 8 * list_max for lists of
256, at most 128 rows), list totals recounted from the blocks, and the choice
between the cuts (clustering on stencils, in natural and scrambled row order;
nothing on scattered columns)."""
from __future__ import annotations

import numpy as np
import pytest


def blocks_ok(rp, ci, info):
    """Recount the blocks' lists; return the list total."""
    m = len(rp) - 1
    order, starts = info["row_order"], info["block_starts"]
    assert np.array_equal(np.sort(order), np.arange(m))
    assert starts[0] == 0 and starts[-1] == m and np.all(np.diff(starts) >= 1)
    umax = info["list_max"]
    nnz_max = (7 if umax == 128 else 8) * umax
    base = rp[0]
    total = 0
    for b in range(info["nblocks"]):
        rows = order[starts[b]:starts[b + 1]]
        assert len(rows) <= 128
        cols = [ci[rp[r] - base:rp[r + 1] - base] for r in rows]
        nz = sum(len(c) for c in cols)
        u = len(np.unique(np.concatenate(cols))) if nz else 0
        if u > umax or nz > nnz_max:  # a row too long for any list: a block of its own
            assert len(rows) == 1
            continue
        total += u
    assert total == info["list_total"]
    return total


def scramble(rp, ci, seed):
    """Symmetric random permutation of a square pattern."""
    m = len(rp) - 1
    perm = np.random.default_rng(seed).permutation(m)
    inv = np.empty(m, np.int64)
    inv[perm] = np.arange(m)
    lens = np.diff(rp)[perm]
    nrp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    nci = np.concatenate([np.sort(inv[ci[rp[r]:rp[r + 1]]]) for r in perm]).astype(np.int32)
    return nrp, nci


def test_stencil_clustered_and_consecutive(pls, gen):
    rp, ci, _ = gen.stencil27(12)
    m = len(rp) - 1
    c = pls.spmm_blocks_inspect(m, m, rp, ci, 0)
    s = pls.spmm_blocks_inspect(m, m, rp, ci, 1)
    assert c["used"] and c["clustered"] and c["list_max"] == 128
    assert s["used"] and not s["clustered"]
    assert np.array_equal(s["row_order"], np.arange(m))
    assert blocks_ok(rp, ci, c) < 0.7 * blocks_ok(rp, ci, s)
    # forced clustering gives the same cut as the automatic choice here
    f = pls.spmm_blocks_inspect(m, m, rp, ci, 2)
    assert f["list_total"] == c["list_total"] and np.array_equal(f["row_order"], c["row_order"])


def test_scrambled_stencil_needs_clustering(pls, gen):
    rp, ci, _ = gen.stencil27(12)
    m = len(rp) - 1
    rp, ci = scramble(rp, ci, 3)
    s = pls.spmm_blocks_inspect(m, m, rp, ci, 1)
    assert not s["used"]  # consecutive rows share too few columns
    c = pls.spmm_blocks_inspect(m, m, rp, ci, 0)
    assert c["used"] and c["clustered"]
    assert blocks_ok(rp, ci, c) <= 0.3 * len(ci)


def test_one_based_pattern_gives_the_same_blocks(pls, gen):
    rp, ci, _ = gen.stencil27(10)
    m = len(rp) - 1
    a = pls.spmm_blocks_inspect(m, m, rp, ci, 0)
    b = pls.spmm_blocks_inspect(m, m, rp + 1, ci + 1, 0)
    for key in ("used", "nblocks", "list_max", "list_total", "clustered"):
        assert a[key] == b[key]
    assert np.array_equal(a["row_order"], b["row_order"])
    assert np.array_equal(a["block_starts"], b["block_starts"])


def test_scattered_columns_have_no_blocks(pls, gen):
    rp, ci, _ = gen.random_rows(4000, 400000, 5, 40, 11)
    info = pls.spmm_blocks_inspect(4000, 400000, rp, ci, 0)
    assert not info["used"] and not info["clustered"]


def test_empty_and_long_rows_forced_clustering(pls, gen):
    rp0, ci0, _ = gen.stencil27(8)
    m0 = len(rp0) - 1
    rp0, ci0 = scramble(rp0, ci0, 5)
    n = m0 + 400
    rng = np.random.default_rng(2)
    rows = []
    for i in range(m0):
        if i % 5 == 1:
            rows.append(np.zeros(0, np.int32))
        if i == 100:
            long_row = len(rows)
            rows.append(np.sort(rng.choice(n, 300, replace=False)).astype(np.int32))
        rows.append(ci0[rp0[i]:rp0[i + 1]])
    m = len(rows)
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    ci = np.concatenate(rows).astype(np.int32)
    info = pls.spmm_blocks_inspect(m, n, rp, ci, 2)
    assert info["used"] and info["clustered"]
    blocks_ok(rp, ci, info)
    # the long row is a block of its own, the empty rows share blocks of empty rows
    order, starts = info["row_order"], info["block_starts"]
    lens = np.diff(rp)
    for b in range(info["nblocks"]):
        r = order[starts[b]:starts[b + 1]]
        if long_row in r:
            assert len(r) == 1
        if np.any(lens[r] == 0):
            assert np.all(lens[r] == 0)


def test_block_rows_of_eight_use_lists_of_256(pls, gen):
    brp, bci, bv = gen.block_stencil(5, 4, 3, 8, 0x5EED0003)
    m = (len(brp) - 1) * 8
    rp, ci, _ = gen.bsr_to_csr_rowmajor(brp, bci, bv, 8)
    info = pls.spmm_blocks_inspect(m, m, rp, ci, 0)
    assert info["used"] and info["list_max"] == 256
    blocks_ok(rp, ci, info)
    rp2, ci2 = scramble(rp, ci, 8)
    info2 = pls.spmm_blocks_inspect(m, m, rp2, ci2, 0)
    assert info2["used"] and info2["clustered"] and info2["list_max"] == 256
    assert blocks_ok(rp2, ci2, info2) <= 1.2 * info["list_total"]


@pytest.mark.parametrize("case", ["rows", "column", "row_ptr", "base"])
def test_invalid_patterns_are_parameter_errors(pls, case):
    import ctypes as C

    rp = np.array([0, 2, 3], np.int32)
    ci = np.array([0, 1, 1], np.int32)
    rows = 0
    if case == "rows":
        rows = 3
    elif case == "column":
        ci = np.array([0, 5, 1], np.int32)
    elif case == "row_ptr":
        rp = np.array([0, 3, 2], np.int32)
    else:
        rp = np.array([2, 4, 5], np.int32)
    used = C.c_int()
    st = pls.lib().perflibs_amd_spmm_blocks_inspect(2, 2, rp.ctypes.data, ci.ctypes.data, rows,
                                                    C.byref(used), None, None, None, None,
                                                    None, None)
    assert st == 1
