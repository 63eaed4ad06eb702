"""GPU tests of the format selector: MLP kernels against the numpy oracle,
training, features, exhaustive timing and the selection modes."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu
K = 13  # PERFLIBS_AMD_NUM_VARIANTS


def dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def test_mlp_forward_backward_match_oracle(pls):
    import torch

    rng = np.random.default_rng(0)
    n = 300
    P = (rng.standard_normal(mlp.num_params(K)) * 0.2).astype(np.float32)
    X = rng.standard_normal((n, 12)).astype(np.float32)
    T = rng.standard_normal((n, K)).astype(np.float32)
    M = (rng.random((n, K)) < 0.8).astype(np.float32)
    assert pls.lib().perflibs_amd_mlp_num_params() == len(P)
    Pd, Xd, Td, Md = dev(P), dev(X), dev(T), dev(M)
    h1 = torch.empty((n, 64), device="cuda")
    h2 = torch.empty((n, 64), device="cuda")
    out = torch.empty((n, K), device="cuda")
    G = torch.empty(len(P), device="cuda")
    loss = torch.empty(2, device="cuda")
    L = pls.lib()
    pls.check("fwd", L.perflibs_amd_mlp_forward(Pd.data_ptr(), Xd.data_ptr(), n, h1.data_ptr(),
                                                h2.data_ptr(), out.data_ptr()))
    pls.check("bwd", L.perflibs_amd_mlp_backward(Pd.data_ptr(), Xd.data_ptr(), n, h1.data_ptr(),
                                                 h2.data_ptr(), out.data_ptr(), Td.data_ptr(),
                                                 Md.data_ptr(), G.data_ptr(), loss.data_ptr()))
    torch.cuda.synchronize()
    _, _, ref_out = mlp.forward(P.astype(np.float64), X, K)
    ref_loss, ref_g = mlp.backward(P.astype(np.float64), X, T.astype(np.float64), M, K)
    assert np.max(np.abs(out.cpu().numpy() - ref_out)) <= 1e-4 * max(1.0, np.abs(ref_out).max())
    assert abs(loss.cpu().numpy()[0] - ref_loss) <= 1e-4 * ref_loss
    g = G.cpu().numpy()
    assert np.linalg.norm(g - ref_g) <= 1e-4 * np.linalg.norm(ref_g)


def test_mlp_training_fits(pls):
    import torch

    rng = np.random.default_rng(1)
    n = 512
    X = rng.standard_normal((n, 12)).astype(np.float32)
    Wt = rng.standard_normal((12, K)).astype(np.float32)
    T = np.tanh(X @ Wt).astype(np.float32)
    M = np.ones((n, K), np.float32)
    P = (rng.standard_normal(mlp.num_params(K)) * 0.1).astype(np.float32)
    Pd, Xd, Td, Md = dev(P), dev(X), dev(T), dev(M)
    m1 = torch.zeros_like(Pd)
    v1 = torch.zeros_like(Pd)
    h1 = torch.empty((n, 64), device="cuda")
    h2 = torch.empty((n, 64), device="cuda")
    out = torch.empty((n, K), device="cuda")
    G = torch.empty_like(Pd)
    loss = torch.empty(2, device="cuda")
    L = pls.lib()
    losses = []
    for step in range(1, 801):
        L.perflibs_amd_mlp_forward(Pd.data_ptr(), Xd.data_ptr(), n, h1.data_ptr(), h2.data_ptr(), out.data_ptr())
        L.perflibs_amd_mlp_backward(Pd.data_ptr(), Xd.data_ptr(), n, h1.data_ptr(), h2.data_ptr(),
                                    out.data_ptr(), Td.data_ptr(), Md.data_ptr(), G.data_ptr(), loss.data_ptr())
        L.perflibs_amd_mlp_adam(Pd.data_ptr(), G.data_ptr(), m1.data_ptr(), v1.data_ptr(), step, 3e-3, 0.9, 0.999, 1e-8)
        if step in (1, 800):
            torch.cuda.synchronize()
            losses.append(float(loss.cpu()[0]))
    assert losses[1] < 0.1 * losses[0]


def test_features_of_known_matrices(pls, gen):
    rp, ci, v = gen.laplace2d(64)
    A = pls.SpMat.csr(4096, 4096, rp, ci, v)
    f = A.features()
    assert abs(f[0] - np.log(4096)) < 1e-5
    assert abs(np.expm1(f[3]) - rp[-1] / 4096) < 1e-4  # mean row length
    assert abs(f[5] - np.log1p(5)) < 1e-5  # longest row
    assert abs(f[7] - 64 / 4096) < 1e-6  # bandwidth / n
    assert f[11] == 0.0
    brp, bci, bv = gen.block_stencil(4, 4, 4, 8, 7)
    B = pls.SpMat.bsr(pls.ROW_MAJOR, 512, 512, 8, brp, bci, bv)
    fb = B.features()
    assert abs(fb[10] - 1.0) < 1e-6 and abs(fb[9] - 1.0) < 1e-6  # fully dense blocks
    E = pls.SpMat.csr(8, 8, [0, 0, 1, 1, 1, 2, 2, 2, 2], [3, 4], np.ones(2))
    assert abs(E.features()[11] - 6 / 8) < 1e-6


def cu_count():
    import torch

    return torch.cuda.get_device_properties(0).multi_processor_count


SELL_SHAPES = [(32, 0), (32, 64), (32, 1024), (32, 16384), (64, 0), (64, 64), (64, 1024), (64, 16384)]


def sell_rule(orc, m, n, rp, ci):
    """The reference's SELL audition rule per registry variant (beta >= 0.96,
    every thread above 0.85; src/matvec.cpp:303-325), threads = CUs."""
    out = []
    for Cc, sigma in SELL_SHAPES:
        r = orc.scs_beta_threads(m, n, rp, ci, Cc, sigma, cu_count())
        out.append(r["beta"] >= 0.96 and r["min_beta"] > 0.85)
    return np.array(out)


def test_time_variants_and_exhaustive_mode(pls, orc, gen):
    rp, ci, v = gen.powerlaw(1 << 16, 11)
    n = 1 << 16
    A = pls.SpMat.csr(n, n, rp, ci, v)
    ms = A.time_variants(2, 5)
    assert np.all(ms[:2] > 0) and ms[12] > 0
    # SELL layouts only where the padding rule passes
    assert np.array_equal(ms[2:10] > 0, sell_rule(orc, n, n, rp, ci))
    # random columns: 4x4 / 8x8 blocks would be ~1/16 full, so BSR is not a candidate
    assert np.all(ms[10:12] < 0)
    A.select_mode(pls.SELECT_EXHAUSTIVE)
    A.optimize()
    assert 0 <= A.variant() < K and ms[A.variant()] > 0
    assert A.select_info() == (pls.PATH_EXHAUSTIVE, int((ms > 0).sum()))
    ref = np.zeros(n)
    x = np.linspace(-1, 1, n)
    A.spmv(x, ref)
    B = pls.SpMat.csr(n, n, rp, ci, v)
    B.select_mode(pls.SELECT_FIXED_CSR)
    B.optimize()
    assert B.variant() == 0
    y = np.zeros(n)
    B.spmv(x, y)
    assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)


def test_learned_mode_uses_model(pls, gen):
    """A model whose output biases rank variant 4, then 7, then 2 first: the
    network mode takes 4; the learned mode times exactly those three and
    the two CSR kernels and keeps the fastest (the pick is one of them);
    features outside the stored training range run the exhaustive sweep."""
    rng = np.random.default_rng(3)
    P = (rng.standard_normal(mlp.num_params(K)) * 0.1).astype(np.float32)
    P[: -K] *= 1e-3  # hidden layers near zero: the output biases decide
    b3 = np.zeros(K, np.float32) + 5.0
    b3[4], b3[7], b3[2] = -5.0, -4.0, -3.0
    P[-K:] = b3
    L = pls.lib()
    rp, ci, v = gen.laplace2d(100)
    A = pls.SpMat.csr(10000, 10000, rp, ci, v)
    f = A.features()
    mean = f.astype(np.float32)  # these features standardise to 0
    std = np.ones(12, np.float32)
    pls.check("set_model", L.perflibs_amd_set_selector_model(P.ctypes.data, mean.ctypes.data, std.ctypes.data))
    try:
        A.select_mode(pls.SELECT_NETWORK)
        A.optimize()
        assert A.variant() == 4 and A.select_info() == (pls.PATH_NETWORK, 0)
        lo, hi = (f - 1).astype(np.float32), (f + 1).astype(np.float32)
        pls.check("range", L.perflibs_amd_set_selector_range(lo.ctypes.data, hi.ctypes.data))
        B = pls.SpMat.csr(10000, 10000, rp, ci, v)
        B.select_mode(pls.SELECT_LEARNED)
        B.optimize()
        assert B.select_info() == (pls.PATH_AUDITION, 5) and B.variant() in (4, 7, 2, 1, 0)
        # a matrix whose features lie outside the range: every candidate timed
        # (on power-law rows the padding rule leaves no SELL-C-sigma candidate)
        rp2, ci2, v2 = gen.powerlaw(1 << 15, 5)
        Cm = pls.SpMat.csr(1 << 15, 1 << 15, rp2, ci2, v2)
        Cm.select_mode(pls.SELECT_LEARNED)
        Cm.optimize()
        path, ntimed = Cm.select_info()
        assert path == pls.PATH_EXHAUSTIVE and ntimed >= 2
        ms = Cm.time_variants(1, 3)
        assert ms[Cm.variant()] > 0 and ntimed == int((ms > 0).sum())
    finally:
        L.perflibs_amd_clear_selector_model()


def test_shipped_model_picks_like_numpy(pls, orc, gen, tmp_path):
    from pathlib import Path

    from plsparse import selector as S

    path = Path(S.DEFAULT_MODEL)
    L = pls.lib()
    bad = tmp_path / "bad.txt"
    bad.write_text("plsparse-selector-mlp 1\nfeatures 12 hidden 32 outputs 13\n")
    assert L.perflibs_amd_load_selector_model(str(bad).encode()) == pls.STATUS_INPUT_PARAMETER_ERROR
    pls.check("load", L.perflibs_amd_load_selector_model(str(path).encode()))
    params, mean, std = S.load_model(path)
    for seed, (m, kind) in enumerate(((200_000, "powerlaw"), (90_000, "banded"), (40_000, "stencil"))):
        mm, n, rp, ci, v, _ = S.sample_matrix(kind, 900 + seed, rows=(m, m + 1))
        A = pls.SpMat.csr(mm, n, rp, ci, v)
        A.hint(pls.HINT_SPMV_INVOCATIONS, pls.INVOCATIONS_MANY)
        A.select_mode(pls.SELECT_NETWORK)
        f = A.features()
        A.optimize()
        _, _, pred = mlp.forward(params, ((f - mean) / std)[None, :], K)
        cand = np.ones(K, bool)
        cand[2:10] = sell_rule(orc, mm, n, rp, ci)
        cand[10] = f[9] >= 0.75 and mm % 4 == 0
        cand[11] = f[10] >= 0.75 and mm % 8 == 0
        want = int(np.argmin(np.where(cand, pred[0], np.inf)))
        assert A.variant() == want, (kind, pls.VARIANTS[A.variant()], pls.VARIANTS[want])
        # the default mode (a model is loaded) is the learned audition of the
        # best three and the two CSR kernels
        D = pls.SpMat.csr(mm, n, rp, ci, v)
        D.hint(pls.HINT_SPMV_INVOCATIONS, pls.INVOCATIONS_MANY)
        D.optimize()
        path, ntimed = D.select_info()
        timed = set(np.argsort(np.where(cand, pred[0], np.inf), kind="stable")[:min(3, int(cand.sum()))].tolist())
        timed |= {0, 1}
        assert (path == pls.PATH_AUDITION and ntimed == len(timed)
                and D.variant() in timed) or path == pls.PATH_EXHAUSTIVE


def test_block_fill_features_match_numpy(pls):
    """4x4 / 8x8 block fill against a direct count of distinct (block row, block
    column) pairs, with long rows that the fill kernel splits over a wave."""
    rng = np.random.default_rng(31)
    m, n = 4096, 3000
    lens = rng.integers(0, 30, m)
    lens[::37] = 2500
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(n, int(l), replace=False)) for l in lens]).astype(np.int32)
    A = pls.SpMat.csr(m, n, rp, ci, np.ones(len(ci)))
    f = A.features()
    rows = np.repeat(np.arange(m), lens)
    for b, idx in ((4, 9), (8, 10)):
        blocks = len(np.unique((rows // b).astype(np.int64) * (n // b + 1) + ci // b))
        assert abs(f[idx] - len(ci) / (blocks * b * b)) <= 1e-6 * f[idx]


def test_block_fill_features_unsorted_rows(pls):
    """Rows whose columns are not ascending (the reference accepts them) give
    the same block-fill features as the sorted matrix: the fill kernels run on
    a row-sorted copy when any row is unsorted."""
    rng = np.random.default_rng(32)
    m, n = 4096, 3000
    lens = rng.integers(0, 30, m)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    rows = [np.sort(rng.choice(n, int(l), replace=False)) for l in lens]
    ci_sorted = np.concatenate(rows).astype(np.int32)
    ci_shuffled = np.concatenate([rng.permutation(r) for r in rows]).astype(np.int32)
    fs = pls.SpMat.csr(m, n, rp, ci_sorted, np.ones(len(ci_sorted))).features()
    fu = pls.SpMat.csr(m, n, rp, ci_shuffled, np.ones(len(ci_sorted))).features()
    assert np.array_equal(fs, fu)


def test_block_fill_features_sampled_block_rows(pls, gen):
    """On operands with more than 65536 block rows the fill features merge every
    stride-th block row (stride = block rows // 65536) and divide the non-zeros
    of exactly those block rows by their distinct blocks."""
    m, n = 1_200_000, 50_000  # 300k / 150k block rows: strides 4 and 2
    rp, ci, v = gen.random_rows(m, n, 2, 9, 0x5EED0041)
    f = pls.SpMat.csr(m, n, rp, ci, v).features()
    rows = np.repeat(np.arange(m, dtype=np.int64), np.diff(rp))
    for b, idx in ((4, 9), (8, 10)):
        nbr = (m + b - 1) // b
        stride = max(1, nbr // 65536)
        keep = (rows // b) % stride == 0
        blocks = len(np.unique((rows[keep] // b) * (n // b + 1) + ci[keep] // b))
        want = keep.sum() / (blocks * b * b)
        assert stride > 1 and abs(f[idx] - want) <= 1e-6 * want, (b, f[idx], want)


def test_learned_selector_regret_on_regenerated_heldout(pls):
    """Regret on this GPU, not on stored timings: the 50 smallest matrices of the
    held-out set (every generator represented) are regenerated from their seeds,
    every candidate variant is timed here (the exhaustive sweep), and the default
    optimize() (the MLP ranks, its three best and the two CSR kernels are timed) is scored against it:
    regret = geomean(t(pick) / t(best)) - 1 <= 10%, below the hand heuristic and
    below every fixed variant that is a candidate on all of them."""
    import time

    from plsparse import selector as S

    d = np.load(S.DEFAULT_MODEL.parent / "heldout_mi355x.npz")
    order = np.argsort(d["sizes"][:, 1], kind="stable")[:50]
    assert len(set(d["kinds"][order])) == len(S.GENERATORS)
    t0 = time.time()
    times, learned, heur = [], [], []
    seeds = d["seeds"] if "seeds" in d else [S.matrix_seed(2, i) for i in range(len(d["times"]))]
    caps = d["max_nnz"] if "max_nnz" in d else [S.MAX_NNZ] * len(d["times"])
    for i in order:
        kind = str(d["kinds"][i])
        m, n, rp, ci, v, desc = S.sample_matrix(kind, int(seeds[i]), max_nnz=int(caps[i]))
        assert (m, int(rp[-1])) == tuple(d["sizes"][i]), (i, desc)
        A = pls.SpMat.csr(m, n, rp, ci, v, flags=pls.CREATE_NOCOPY)
        ms = A.time_variants(3, 10)
        A.destroy()
        picks = []
        for mode in (None, pls.SELECT_HEURISTIC):
            B = pls.SpMat.csr(m, n, rp, ci, v, flags=pls.CREATE_NOCOPY)
            B.hint(pls.HINT_SPMV_INVOCATIONS, pls.INVOCATIONS_MANY)
            if mode is not None:
                B.select_mode(mode)
            B.optimize()
            picks.append(B.variant())
            B.destroy()
        assert ms[picks[0]] > 0 and ms[picks[1]] > 0, (desc, picks, ms)
        times.append(ms)
        learned.append(picks[0])
        heur.append(picks[1])
        print(f"  {desc} m={m} nnz={int(rp[-1])}: best {pls.VARIANTS[int(np.argmin(np.where(ms > 0, ms, np.inf)))]}"
              f" learned {pls.VARIANTS[picks[0]]} heuristic {pls.VARIANTS[picks[1]]}", flush=True)
    T = np.array(times)
    best = np.where(T > 0, T, np.inf).min(axis=1)

    def regret(p):
        return float(np.exp(np.mean(np.log(T[np.arange(len(T)), p] / best))) - 1.0)

    r_learned, r_heur = regret(np.array(learned)), regret(np.array(heur))
    fixed = {pls.VARIANTS[v]: regret(np.full(len(T), v)) for v in range(K) if np.all(T[:, v] > 0)}
    print(f"  regret learned {r_learned:.4f} heuristic {r_heur:.4f} fixed {fixed} "
          f"({time.time() - t0:.0f} s)", flush=True)
    assert r_learned <= 0.10
    assert r_learned < r_heur
    assert fixed and r_learned < min(fixed.values())
