"""SELL-C-sigma padding efficiency on the GPU (get_beta / get_beta_threads,
src/sell_c_sigma.cpp:369-452) against the oracle, and the reference's
audition rule (beta >= 0.96, every thread above 0.85; src/matvec.cpp:303-325)
as the candidate filter of optimize()."""
from __future__ import annotations

import numpy as np
import pytest

from test_gpu_spmv import powerlawish

pytestmark = pytest.mark.gpu

REGISTRY = [(32, 0), (32, 64), (32, 1024), (32, 16384), (64, 0), (64, 64), (64, 1024), (64, 16384)]
SELL_VARIANTS = list(range(2, 10))


def lengths_csr(rng, lens, n):
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(n, int(l), replace=False)) for l in lens]).astype(np.int32)
    return rp, ci, rng.standard_normal(int(rp[-1]))


def matrices(gen):
    rng = np.random.default_rng(71)
    out = []
    m = 40_000 + 123  # partial last window and partial last chunk
    out.append(("powerlaw",) + (m, m) + powerlawish(rng, m, m, mean=12.0))
    lens = np.where(np.arange(30_001) % 7 == 0, 90, 3)  # long rows every 7th row: padding
    out.append(("spiky", 30_001, 5000) + lengths_csr(rng, lens, 5000))
    rp, ci, v = gen.laplace2d(150)
    out.append(("laplace", 150 * 150, 150 * 150, rp, ci, v))
    lens = rng.integers(0, 4, 5000)  # short rows, many empty
    out.append(("short", 5000, 300) + lengths_csr(rng, lens, 300))
    return out


@pytest.mark.parametrize("nthreads", [1, 7, 256])
def test_sell_beta_matches_oracle(pls, orc, gen, nthreads):
    for name, m, n, rp, ci, v in matrices(gen):
        A = pls.SpMat.csr(m, n, rp, ci, v)
        for Cc, sigma in REGISTRY + [(16, 48), (8, 0), (4, 12)]:  # the last three: host path
            got = A.sell_beta(Cc, sigma, nthreads)
            ref = orc.scs_beta_threads(m, n, rp, ci, Cc, sigma, nthreads)
            want = (ref["beta"], ref["min_beta"], ref["max_beta"])
            assert got == want, (name, Cc, sigma, nthreads, got, want)


def test_sell_beta_one_based_and_padding_cases(pls, orc):
    # hand case of the oracle tests, 1-based: beta 12 / 17, threads 5/8 and 7/9
    lens = [4, 1, 1, 1, 2, 0, 3]
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32) + 1
    ci = np.concatenate([np.arange(l) for l in lens]).astype(np.int32) + 1
    A = pls.SpMat.csr(7, 8, rp, ci, np.ones(12))
    assert A.sell_beta(2, 0, 2) == (12 / 17, 5 / 8, 7 / 9)
    assert A.sell_beta(2, 0, 1) == (12 / 17, 12 / 17, 12 / 17)


def test_candidate_rule_skips_padded_sell(pls, orc, gen):
    """Exhaustive selection times only SELL layouts that pass the rule; the
    same rule holds for the learned and heuristic modes (variant_candidate)."""
    rng = np.random.default_rng(5)
    # one long row per 64 in two whole sort windows of 16384 rows: sigma = 0 pads
    # every chunk, sigma = 16384 gathers the long rows into unpadded chunks
    lens = np.where(np.arange(2 * 16384) % 64 == 0, 200, 4)
    m = n = len(lens)
    rp, ci, v = lengths_csr(rng, lens, n)
    A = pls.SpMat.csr(m, n, rp, ci, v)
    ms = A.time_variants(1, 2)
    for var, (Cc, sigma) in zip(SELL_VARIANTS, REGISTRY):
        ref = orc.scs_beta_threads(m, n, rp, ci, Cc, sigma, 256)
        ok = ref["beta"] >= 0.96 and ref["min_beta"] > 0.85
        assert (ms[var] > 0) == ok, (Cc, sigma, ref["beta"], ref["min_beta"], ms[var])
    assert ms[2] < 0 and ms[6] < 0 and ms[5] > 0 and ms[9] > 0
    # a stencil pads only at the boundary rows: every SELL variant stays a candidate
    rp, ci, v = gen.laplace2d(200)
    L = pls.SpMat.csr(200 * 200, 200 * 200, rp, ci, v)
    msl = L.time_variants(1, 2)
    assert all(msl[var] > 0 for var in SELL_VARIANTS)
