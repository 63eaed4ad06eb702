"""Pins the CPU oracle (oracle/) before anything is compared against it.

Each check is chosen so that a plausible mistake in the restatement (a wrong
tie-break, a dropped compensation term, a transposed block, an off-by-one in
the chunk bookkeeping) fails it: reference golden vectors, closed forms,
invariants of the SELL-C-sigma layout, brute force on tiny inputs and exact
rational arithmetic.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np
import pytest

from conftest import rel_err


def random_csr(rng, m, n, max_len, empty_frac=0.1, base=0):
    lens = rng.integers(0, max_len + 1, size=m)
    lens[rng.random(m) < empty_frac] = 0
    lens = np.minimum(lens, n)
    rp = np.zeros(m + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    ci = np.concatenate([np.sort(rng.choice(n, size=l, replace=False)) for l in lens]) if rp[-1] else np.zeros(0, np.int64)
    v = rng.standard_normal(int(rp[-1]))
    return (rp + base).astype(np.int32), (ci + base).astype(np.int32), v


def dense_of(m, n, rp, ci, v):
    base = int(rp[0])
    A = np.zeros((m, n))
    for i in range(m):
        for p in range(rp[i] - base, rp[i + 1] - base):
            A[i, ci[p] - base] += v[p]
    return A


# ---------------------------------------------------------------- goldens
def test_spmv_basic_golden(orc, golden):
    g = golden["spmv_basic_d"]
    y = orc.spmv(g["m"], g["row_ptr"], g["col_indx"], g["vals"], g["x"], g["alpha"], g["beta"])
    assert np.max(np.abs(y - g["expected"])) <= g["tol"]


def test_alpha_beta_golden(orc, golden):
    g = golden["alpha_beta_semantics"]
    y = orc.spmv(g["m"], g["row_ptr"], g["col_indx"], g["vals"], g["x_notrans"], g["alpha"], g["beta"], g["y_init"])
    assert np.max(np.abs(y - g["expected_notrans"])) <= g["tol"]
    At = dense_of(4, 4, np.array(g["row_ptr"]), np.array(g["col_indx"]), np.array(g["vals"])).T
    assert np.max(np.abs(At @ np.array(g["x_trans"]) - g["expected_trans"])) <= g["tol"]


def test_strategy_paths_closed_form(orc, golden):
    g = golden["strategy_paths"]
    m, w, s = g["m"], g["width"], g["col_stride"]
    rp = np.arange(m + 1, dtype=np.int32) * w
    ci = (np.arange(m)[:, None] * s + np.arange(w)[None, :]).reshape(-1).astype(np.int32)
    v = np.ones(m * w)
    y = orc.spmv(m, rp, ci, v, np.ones(s * m))
    assert np.all(np.abs(y - g["expected_y"]) <= g["tol_d"])
    # SELL round trip keeps the exported CSR of the reference test
    scs = orc.csr2scs(m, s * m, rp, ci, v, 16, 16 * 2000)
    crp, cci, cv = scs["csr"]
    e = g["export_after_optimize"]
    assert crp[m] == e["row_ptr_last"] and cci[0] == e["col_first"] and cci[-1] == e["col_last"]
    assert scs["beta"] == 1.0
    u = g["update"]
    p = rp[u["row"]] + list(ci[rp[u["row"]]:rp[u["row"] + 1]]).index(u["col"])
    v2 = v.copy()
    v2[p] = u["value"]
    y2 = orc.spmv(m, rp, ci, v2, np.ones(s * m))
    assert y2[u["row"]] == u["expected_row_value"]
    ab = g["alpha_beta"]
    y3 = orc.spmv(m, rp, ci, v2, np.ones(s * m), ab["alpha"], ab["beta"], np.full(m, ab["y0"]))
    assert y3[0] == ab["expected"] and y3[u["row"]] == ab["expected_updated_row"]


def omp_paths_case(golden):
    g = golden["csr_spmv_omp_paths_s"]
    m, n = g["m"], g["n"]
    lens = np.array([g["lengths"][i % 8] for i in range(m)])
    rp = np.zeros(m + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    ci = np.concatenate([(29 * i + 17) % (n - lens[i]) + np.arange(lens[i]) for i in range(m)])
    vals = (0.125 * ((np.arange(rp[-1]) % 13) + 1)).astype(np.float32)
    x = (((np.arange(n) % 11) - 5) * 0.5).astype(np.float32)
    y = ((np.arange(m) % 9) - 4).astype(np.float32)
    sums = np.array([np.sum(vals[rp[i]:rp[i + 1]].astype(np.float64) * x[ci[rp[i]:rp[i + 1]]].astype(np.float64))
                     for i in range(m)])
    expected = (np.float32(g["beta"]) * y + np.float32(g["alpha"]) * sums.astype(np.float32)).astype(np.float64)
    return g, rp.astype(np.int32), ci.astype(np.int32), vals, x, y, expected


def test_csr_omp_paths_fp32_golden(orc, golden):
    g, rp, ci, vals, x, y, expected = omp_paths_case(golden)
    out = orc.spmv_f32(g["m"], rp, ci, vals, x, g["alpha"], g["beta"], y)
    assert np.max(np.abs(out - expected)) <= g["tol"]


def test_bsr_goldens(orc, golden):
    g = golden["bsr_row_major"]
    y = orc.bsr_spmv(g["layout"], g["m"] // 2, 2, g["row_ptr"], g["col_indx"], g["vals"], g["x"])
    assert np.max(np.abs(y - g["expected_notrans"])) <= g["tol"]
    rp, ci, v = orc.bsr2csr(g["layout"], 4, 4, 2, g["row_ptr"], g["col_indx"], g["vals"])
    At = dense_of(4, 4, rp, ci, v).T
    assert np.max(np.abs(At @ np.array(g["x"]) - g["expected_trans"])) <= g["tol"]
    h = golden["bsr_col_major_transposed_half"]
    rp, ci, v = orc.bsr2csr(h["layout"], 4, 4, 2, h["row_ptr"], h["col_indx"], h["vals"])
    D = dense_of(4, 4, rp, ci, v)
    assert np.max(np.abs(h["alpha"] * D.T @ np.array(h["x"]) - h["expected_trans"])) <= h["tol"]
    assert np.max(np.abs(D).sum(axis=1)) == h["inf_norm"]  # the test's own norm check


# ---------------------------------------------------------------- Dot2
def test_dot2_cancellation_exact(orc):
    # naive evaluation gives 0; the exact value is 1
    rp = np.array([0, 3], np.int32)
    ci = np.array([0, 1, 2], np.int32)
    v = np.array([1e16, 1.0, -1e16])
    y = orc.spmv(1, rp, ci, v, np.ones(3))
    assert y[0] == 1.0


def test_dot2_matches_rational(orc):
    rng = np.random.default_rng(7)
    for trial in range(20):
        n = 40
        v = rng.standard_normal(n) * np.exp2(rng.integers(-30, 30, n))
        x = rng.standard_normal(n) * np.exp2(rng.integers(-30, 30, n))
        # make it ill-conditioned: append negated products
        v = np.concatenate([v, -v[:20] * (1 + 2.0**-40)])
        x = np.concatenate([x, x[:20]])
        rp = np.array([0, len(v)], np.int32)
        ci = np.arange(len(v), dtype=np.int32)
        alpha, beta, y0 = 1.25, -0.75, 3.5
        got = orc.spmv(1, rp, ci, v, x, alpha, beta, [y0])[0]
        exact = Fraction(alpha) * sum(Fraction(a) * Fraction(b) for a, b in zip(v, x)) + Fraction(beta) * Fraction(y0)
        assert abs(Fraction(got) - exact) <= abs(exact) * Fraction(2.0**-50) + Fraction(2.0**-1070)


def test_beta_zero_ignores_nan_and_alpha_zero(orc):
    rp = np.array([0, 1, 2], np.int32)
    ci = np.array([0, 1], np.int32)
    v = np.array([2.0, 3.0])
    y = orc.spmv(2, rp, ci, v, [1.0, 1.0], 1.0, 0.0, [np.nan, np.inf])
    assert np.array_equal(y, [2.0, 3.0])
    y = orc.spmv(2, rp, ci, v, [np.nan, 1.0], 0.0, 2.0, [1.0, 5.0])
    assert np.array_equal(y, [2.0, 10.0])


def test_spmm_dot2_matches_dense(orc):
    rng = np.random.default_rng(3)
    rp, ci, v = random_csr(rng, 30, 25, 6)
    B = rng.standard_normal((25, 7))
    C0 = rng.standard_normal((30, 7))
    got = orc.spmm(30, 7, rp, ci, v, B, 1.5, -0.5, C0)
    ref = 1.5 * dense_of(30, 25, rp, ci, v) @ B - 0.5 * C0
    assert rel_err(got, ref) < 1e-14


# ---------------------------------------------------------------- SELL-C-sigma
def bytes_rule(rng_max: int) -> int:
    """Closed form of the offset width: signed storage of the largest range."""
    if rng_max < 2**7:
        return 1
    if rng_max < 2**15:
        return 2
    return 4


def check_scs_invariants(orc, m, n, rp, ci, v, Cc, sigma):
    s = orc.csr2scs(m, n, rp, ci, v, Cc, sigma)
    base = int(rp[0])
    lens = np.diff(rp)
    nC = -(-m // Cc)
    assert s["nC"] == nC
    perm = s["row_permd2in"]
    inv = s["row_in2permd"]
    assert sorted(perm[:m].tolist()) == list(range(m))
    assert np.all(perm[m:] == -1)
    assert np.all(inv[perm[:m]] == np.arange(m))
    win = sigma if sigma else Cc
    for w0 in range(0, m, win):
        rows = perm[w0:min(w0 + win, m)]
        if sigma:
            keys = [(lens[r], r) for r in rows]
            assert keys == sorted(keys, reverse=True), "rows not sorted by (len, idx) descending"
        else:
            assert np.array_equal(rows, np.arange(w0, min(w0 + win, m)))
    for c in range(nC):
        rows = perm[c * Cc:min((c + 1) * Cc, m)]
        assert s["cl"][c] == lens[rows].max()
        cols = np.concatenate([ci[rp[r] - base:rp[r + 1] - base] for r in rows]) - base
        assert s["col_indx_min"][c] == (cols.min() if len(cols) else 2**31 - 1)
        assert s["chunk_min_nnz"][c] == lens[rows].sum()
    cs = np.concatenate([[0], np.cumsum(s["cl"] * Cc)[:-1]])
    assert np.array_equal(s["cs"], cs)
    assert s["padded"] == int((s["cl"] * Cc).sum())
    assert s["n_full_chunks"] == m // Cc and s["ntidyup"] == m % Cc
    rng_max = 0
    for c in range(nC):
        rows = perm[c * Cc:min((c + 1) * Cc, m)]
        cols = np.concatenate([ci[rp[r] - base:rp[r + 1] - base] for r in rows]) - base
        if len(cols):
            rng_max = max(rng_max, int(cols.max() - cols.min()))
    assert s["col_indx_bytes"] == bytes_rule(rng_max)
    # every slot: real entries reproduce A, padding is (0, 0)
    D = np.zeros((m, n))
    for c in range(nC):
        for r in range(Cc):
            row = perm[c * Cc + r]
            for j in range(s["cl"][c]):
                k = s["cs"][c] + j * Cc + r
                if row < 0 or j >= lens[row]:
                    assert s["vals"][k] == 0.0 and s["offsets"][k] == 0
                else:
                    D[row, s["offsets"][k] + s["col_indx_min"][c]] += s["vals"][k]
    assert np.array_equal(D, dense_of(m, n, rp, ci, v))
    # scs2csr drops zeros and otherwise returns the input (0-based)
    crp, cci, cv = s["csr"]
    keep = v != 0
    kept_rp = np.concatenate([[0], np.cumsum([keep[rp[i] - base:rp[i + 1] - base].sum() for i in range(m)])])
    assert np.array_equal(crp, kept_rp)
    assert np.array_equal(cci, (ci - base)[keep]) and np.array_equal(cv, v[keep])
    return s


@pytest.mark.parametrize("Cc", [1, 2, 3, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("mult", [0, 1, 2, 5])
def test_scs_invariants_random(orc, Cc, mult):
    rng = np.random.default_rng(Cc * 10 + mult)
    m, n = int(rng.integers(40, 300)), int(rng.integers(30, 400))
    rp, ci, v = random_csr(rng, m, n, 12, empty_frac=0.15)
    v[rng.random(len(v)) < 0.05] = 0.0  # explicit zeros
    check_scs_invariants(orc, m, n, rp, ci, v, Cc, Cc * mult)


def test_scs_one_based_and_wide_columns(orc):
    rng = np.random.default_rng(11)
    m, n = 150, 70000  # column ranges > 2^15 need 4-byte offsets
    rp, ci, v = random_csr(rng, m, n, 9, base=1)
    s = check_scs_invariants(orc, m, n, rp, ci, v, 8, 64)
    assert s["col_indx_bytes"] == 4


def test_scs_tie_break_descending_index(orc):
    # all rows equal length: the sort must put higher row indices first
    m = 8
    rp = np.arange(m + 1, dtype=np.int32) * 2
    ci = np.tile([0, 1], m).astype(np.int32)
    s = orc.csr2scs(m, 2, rp, ci, np.ones(2 * m), 4, 8)
    assert s["row_permd2in"].tolist() == [7, 6, 5, 4, 3, 2, 1, 0]
    s0 = orc.csr2scs(m, 2, rp, ci, np.ones(2 * m), 4, 0)
    assert s0["row_permd2in"].tolist() == list(range(m))


def test_scs_all_empty_chunk_min(orc):
    # a chunk of empty rows keeps col_indx_min = INT32_MAX (range ignored)
    rp = np.array([0, 0, 0, 0, 0, 3, 3, 3, 3], np.int32)
    ci = np.array([5, 9, 200], np.int32)
    s = orc.csr2scs(8, 300, rp, ci, np.ones(3), 4, 0)
    assert s["col_indx_min"].tolist() == [2**31 - 1, 5]
    assert s["cl"].tolist() == [0, 3] and s["col_indx_bytes"] == 2


def test_scs_brute_force_tiny(orc):
    """Every C and window on tiny matrices against a direct construction."""
    rng = np.random.default_rng(5)
    for trial in range(30):
        m, n = int(rng.integers(1, 24)), int(rng.integers(1, 20))
        rp, ci, v = random_csr(rng, m, n, 5, empty_frac=0.3)
        lens = np.diff(rp)
        for Cc in (1, 2, 3, 5, 8):
            for mult in (0, 1, 3):
                sigma = Cc * mult
                s = orc.csr2scs(m, n, rp, ci, v, Cc, sigma)
                win = sigma if sigma else Cc
                order = []
                for w0 in range(0, m, win):
                    rows = list(range(w0, min(w0 + win, m)))
                    if sigma:
                        rows = sorted(rows, key=lambda r: (lens[r], r), reverse=True)
                    order += rows
                assert s["row_permd2in"][:m].tolist() == order


def _csr_of_lengths(lens, n=64):
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.arange(l) for l in lens]).astype(np.int32) if sum(lens) else np.zeros(0, np.int32)
    return rp, ci


def test_scs_beta_hand_cases(orc):
    """get_beta / get_beta_threads on hand-computed padded and tidy-up cases
    (src/sell_c_sigma.cpp:369-452).  beta counts the padded slots of real rows
    only: cl[c] * C for full chunks, cl[last] * ntidyup for a partial last
    chunk; nnz / (cl * C summed over all chunks) would give 12/20 below."""
    # C = 2, sigma = 0, rows [4,1 | 1,1 | 2,0 | 3,3]: cl = [4,1,2,3], 15 nnz in 20 slots
    rp, ci = _csr_of_lengths([4, 1, 1, 1, 2, 0, 3, 3])
    r = orc.scs_beta_threads(8, 64, rp, ci, 2, 0, 2)
    assert r["cl"].tolist() == [4, 1, 2, 3] and r["ntidyup"] == 0
    assert r["beta"] == 15 / 20
    # split: thread 0 = chunks 0-1 (10 slots, 7 nnz), thread 1 = chunks 2-3 (10 slots, 8 nnz)
    assert r["chk_offset"].tolist() == [0, 2] and r["thread_min_nnz"].tolist() == [7, 8]
    assert (r["min_beta"], r["max_beta"]) == (7 / 10, 8 / 10)
    # drop the last row: last chunk holds 1 real row (ntidyup = 1)
    rp, ci = _csr_of_lengths([4, 1, 1, 1, 2, 0, 3])
    r = orc.scs_beta_threads(7, 64, rp, ci, 2, 0, 2)
    assert r["ntidyup"] == 1 and r["n_full_chunks"] == 3
    assert r["beta"] == 12 / 17  # slots 4*2 + 1*2 + 2*2 + 3*1
    # vals_per_thread = (14 - 1) // 2 = 6: chunk 0 alone, then chunks 1..3
    assert r["chk_offset"].tolist() == [0, 1] and r["thread_min_nnz"].tolist() == [5, 7]
    assert (r["min_beta"], r["max_beta"]) == (5 / 8, 7 / 9)
    # sigma = 4 sorts windows of 4 rows by (len, idx) descending:
    # window 0 lens [3,1,0,2] -> chunks {3,2}, {1,0}; window 1 = row 4 alone
    rp, ci = _csr_of_lengths([3, 1, 0, 2, 4])
    r = orc.scs_beta_threads(5, 64, rp, ci, 2, 4, 1)
    assert r["cl"].tolist() == [3, 1, 4] and r["ntidyup"] == 1
    assert r["beta"] == 10 / 12 and r["min_beta"] == r["max_beta"] == r["beta"]
    r0 = orc.scs_beta_threads(5, 64, rp, ci, 2, 0, 1)
    assert r0["cl"].tolist() == [3, 2, 4] and r0["beta"] == 10 / 14
    # C = 1, a 100-long row first, 3 threads: vals_per_thread = 34; the
    # thresholds (vals_next += nchunks * C + vals_per_thread) reach 69 and 104
    # while the slot count is already 101 and 104, so the reference would open
    # a fourth thread at chunk 3 (its assert, src/sell_c_sigma.cpp:527); the
    # chunk stays with the last thread
    rp, ci = _csr_of_lengths([100, 1, 3, 1], n=256)
    r = orc.scs_beta_threads(4, 256, rp, ci, 1, 0, 3)
    assert r["clamped"] and r["chk_offset"].tolist() == [0, 1, 2]
    assert r["nchunks_thread"].tolist() == [1, 1, 2] and r["thread_min_nnz"].tolist() == [100, 1, 4]
    assert r["beta"] == 1.0 and r["min_beta"] == r["max_beta"] == 1.0


@pytest.mark.parametrize("Cc,mult", [(1, 0), (3, 0), (4, 2), (8, 0), (8, 4), (32, 0), (32, 2), (64, 1)])
def test_scs_beta_brute_force(orc, Cc, mult):
    """beta = nnz / sum over chunks of cl[c] * (real rows of chunk c), counted
    directly from the permutation; the thread split partitions the chunks and
    its per-thread ratios bracket beta."""
    rng = np.random.default_rng(100 + Cc * 7 + mult)
    for trial in range(6):
        m = int(rng.integers(Cc, 40 * Cc + 3))
        lens = rng.integers(0, 30, m)
        lens[rng.random(m) < 0.2] = 0
        lens[rng.random(m) < 0.05] = 200  # long rows make padding
        rp, ci = _csr_of_lengths(lens.tolist(), n=256)
        s = orc.csr2scs(m, 256, rp, ci, np.ones(int(lens.sum())), Cc, Cc * mult, fill=False)
        perm = s["row_permd2in"]
        slots = sum(int(s["cl"][c]) * int((perm[c * Cc:(c + 1) * Cc] >= 0).sum()) for c in range(s["nC"]))
        for nth in (1, 2, 5, 16):
            r = orc.scs_beta_threads(m, 256, rp, ci, Cc, Cc * mult, nth)
            assert r["beta"] == (lens.sum() / slots if slots else r["beta"])
            chk, nct, tmn = r["chk_offset"], r["nchunks_thread"], r["thread_min_nnz"]
            assert chk[0] == 0 and np.all(np.diff(chk) >= 0) and np.all(chk <= s["nC"])
            assert int(tmn.sum()) == int(lens.sum())
            for t in range(nth):
                hi = chk[t + 1] if t + 1 < nth else s["nC"]
                assert tmn[t] == s["chunk_min_nnz"][chk[t]:hi].sum()
            if nth == 1:
                assert r["min_beta"] == r["max_beta"] == r["beta"]
            elif r["min_beta"] > 0:
                assert r["min_beta"] <= r["beta"] * (1 + 1e-15) and r["beta"] <= r["max_beta"] * (1 + 1e-15)


def test_scs_generic_kernel_matches_oracle(orc):
    rng = np.random.default_rng(9)
    m, n = 500, 400
    rp, ci, v = random_csr(rng, m, n, 30)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, m)
    ref = orc.spmv(m, rp, ci, v, x, 1.5, 0.5, y0)
    for Cc, sigma in ((16, 3200), (64, 1024), (32, 0)):
        y = y0.copy()
        orc.scs_spmv(m, n, rp, ci, v, Cc, sigma, x, 1.5, 0.5, y)
        assert rel_err(y, ref) < 1e-13


# ---------------------------------------------------------------- BSR
def test_bsr_check_codes(orc):
    assert orc.bsr_check(91, 4, 4, 2, [0, 1, 2]) == 0
    assert orc.bsr_check(92, 4, 4, 2, [0]) == 2
    assert orc.bsr_check(91, -4, 4, 2, [0]) == 3
    assert orc.bsr_check(91, 4, -4, 2, [0]) == 4
    assert orc.bsr_check(91, 4, 4, 3, [0]) == 5
    assert orc.bsr_check(91, 4, 4, 0, [0]) == 5
    assert orc.bsr_check(91, 4, 4, 2, [2]) == 6


@pytest.mark.parametrize("layout", [90, 91])
@pytest.mark.parametrize("b", [1, 2, 3, 4, 8])
def test_bsr2csr_dense_expansion(orc, layout, b):
    rng = np.random.default_rng(b * 100 + layout)
    mb, nb = 7, 6
    brp = np.zeros(mb + 1, np.int32)
    cols = []
    for i in range(mb):
        k = int(rng.integers(0, nb + 1))
        cols.append(np.sort(rng.choice(nb, size=k, replace=False)))
        brp[i + 1] = brp[i] + k
    bci = np.concatenate(cols).astype(np.int32)
    vals = rng.standard_normal(int(brp[-1]) * b * b)
    vals[rng.random(len(vals)) < 0.3] = 0.0
    D = np.zeros((mb * b, nb * b))
    for i in range(mb):
        for k in range(brp[i], brp[i + 1]):
            blk = vals[k * b * b:(k + 1) * b * b].reshape(b, b)
            D[i * b:(i + 1) * b, bci[k] * b:(bci[k] + 1) * b] = blk.T if layout == 90 else blk
    for base in (0, 1):
        rp, ci, v = orc.bsr2csr(layout, mb * b, nb * b, b, brp + base, bci + base, vals)
        assert rp[0] == base
        assert np.all(v != 0.0)
        assert np.array_equal(dense_of(mb * b, nb * b, rp, ci, v), D)
        # columns inside a row follow block order then j: ascending for sorted blocks
        for r in range(mb * b):
            seg = ci[rp[r] - base:rp[r + 1] - base]
            assert np.all(np.diff(seg) > 0)
    x = rng.standard_normal(nb * b)
    y = orc.bsr_spmv(layout, mb, b, brp, bci, vals, x, 2.0)
    assert rel_err(y, 2.0 * D @ x) < 1e-14


# ---------------------------------------------------------------- CSR decomposition
@pytest.mark.parametrize("strategy", [0, 1, 2])
@pytest.mark.parametrize("nthreads", [2, 3, 8, 16])
def test_csr_decomp_properties(orc, strategy, nthreads):
    rng = np.random.default_rng(strategy * 31 + nthreads)
    m = 2000
    lens = np.maximum(1, np.round(np.exp(rng.normal(2.5, 1.0, m)))).astype(np.int64)
    rp = np.zeros(m + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    rows, inds = orc.csr_decomp(m, rp, nthreads, strategy)
    nnz = int(rp[-1])
    assert rows[0] == 0 and inds[0] == 0 and rows[-1] == m and inds[-1] == nnz
    assert np.all(np.diff(inds) >= 0) and np.all(np.diff(rows) >= 0)
    for t in range(nthreads + 1):
        r = rows[t]
        if r < m:  # a thread starts inside (or at the start of) its first row
            assert rp[r] <= inds[t] < rp[r + 1] or inds[t] == nnz
    if strategy == 0:
        rpt = -(-m // nthreads)
        assert list(rows[:-1]) == [min(t * rpt, m) for t in range(nthreads)]


def test_csr_decomp_nnz_split_rows(orc):
    # one huge row: the nnz strategy must split it between threads
    rp = np.array([0, 1000, 1001, 1002, 1003], np.int32)
    rows, inds = orc.csr_decomp(4, rp, 4, 2)
    assert inds.tolist() == [0, 251, 502, 753, 1003]
    assert rows.tolist() == [0, 0, 0, 0, 4]


# ---------------------------------------------------------------- CPU baseline port
def test_cpu_baseline_kernels_agree(orc):
    import ctypes as C

    rng = np.random.default_rng(21)
    m, n = 3000, 2500
    lens = np.maximum(1, np.round(np.exp(rng.normal(2.0, 1.0, m)))).astype(np.int64)
    lens = np.minimum(lens, n)
    rp = np.zeros(m + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    ci = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in lens]).astype(np.int32)
    v = rng.standard_normal(int(rp[-1]))
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, m)
    ref = orc.spmv(m, rp, ci, v, x, 1.25, 0.5, y0)
    L = orc.lib()
    p = lambda a: a.ctypes.data
    for fn in (L.orc_cpu_csr_vanilla, L.orc_cpu_csr_tworow):
        y = y0.copy()
        fn(m, p(rp), p(ci), p(v), p(x), 1.25, 0.5, p(y))
        assert rel_err(y, ref) < 1e-13
    nt = 8
    rows, inds = orc.csr_decomp(m, rp, nt, 1)
    y = y0.copy()
    L.orc_cpu_csr_first_touch(m, p(rp), p(ci), p(v), p(x), 1.25, 0.5, p(y), nt, p(rows), p(inds))
    assert rel_err(y, ref) < 1e-13
    # SpMM row-blocked kernel
    k = 9
    B = rng.standard_normal((n, k))
    C0 = rng.standard_normal((m, k))
    Cm = C0.copy()
    L.orc_cpu_spmm_rowwise(m, k, p(rp), p(ci), p(v), p(B), k, 0.5, 2.0, p(Cm), k)
    assert rel_err(Cm, orc.spmm(m, k, rp, ci, v, B, 0.5, 2.0, C0)) < 1e-13


@pytest.mark.parametrize("b", [2, 4, 5, 8])
@pytest.mark.parametrize("layout", [90, 91])
def test_cpu_baseline_bsr(orc, b, layout):
    rng = np.random.default_rng(b + layout)
    mb = 50
    brp = np.zeros(mb + 1, np.int32)
    cols = [np.sort(rng.choice(mb, size=int(rng.integers(1, 6)), replace=False)) for _ in range(mb)]
    brp[1:] = np.cumsum([len(c) for c in cols])
    bci = np.concatenate(cols).astype(np.int32)
    vals = rng.standard_normal(int(brp[-1]) * b * b)
    x = rng.standard_normal(mb * b)
    y0 = rng.standard_normal(mb * b)
    ref = orc.bsr_spmv(layout, mb, b, brp, bci, vals, x, 1.5, -0.5, y0)
    y = y0.copy()
    p = lambda a: a.ctypes.data
    orc.lib().orc_cpu_bsr(layout, mb, b, p(brp), p(bci), p(vals), p(x), 1.5, -0.5, p(y))
    assert rel_err(y, ref) < 1e-13


def test_csc_coo_goldens_against_oracle(orc, golden):
    """The transcribed CSC / COO goldens agree with the oracle's SpMV of the
    same matrices (rebuilt as CSR and as the CSR of the transpose)."""
    for key, kind in (("csc_kernel_paths_d", "csc"), ("coo_spmv_paths_real_part", "coo")):
        g = golden[key]
        m, n = g["m"], g["n"]
        D = np.zeros((m, n))
        if kind == "csc":
            for j in range(n):
                for k in range(g["col_ptr"][j], g["col_ptr"][j + 1]):
                    D[g["row_indx"][k], j] += g["vals"][k]
        else:
            for r, c, v in zip(g["row_indx"], g["col_indx"], g["vals"]):
                D[r, c] += v
        for M, key_exp in ((D, "expected_notrans"), (D.T, "expected_trans")):
            rp = np.concatenate([[0], np.cumsum((M != 0).sum(axis=1))]).astype(np.int32)
            ci = np.concatenate([np.nonzero(M[i])[0] for i in range(M.shape[0])]).astype(np.int32)
            v = np.concatenate([M[i, np.nonzero(M[i])[0]] for i in range(M.shape[0])])
            y = orc.spmv(M.shape[0], rp, ci, v, np.array(g["x"], dtype=np.float64))
            assert np.max(np.abs(y - g[key_exp])) <= g["tol"]
