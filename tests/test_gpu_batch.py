"""GPU parity of the batched products (one persistent launch replayed from a
hipGraph) through the C-ABI: FP64 within 1e-12 of the compensated oracle per
problem, FP32 / BF16 storage with FP32 accumulation against the exact result of
the stored values and against the FP64 oracle."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
TOL64 = 1e-12


def to_bf16(a):
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def make(pls, gen, nprob, seed, rows=(64, 3000), ks=(1, 2, 4, 8, 16), dtype=np.float64,
         precision=0, empty=True):
    probs = gen.banded_batch(nprob, seed, rows=rows, ks=ks)
    if empty:  # a problem whose rows are all empty
        m = 77
        probs.append((m, np.zeros(m + 1, np.int32), np.zeros(1, np.int32), np.zeros(1), 4))
    mats = [pls.SpMat.csr(m, m, rp, ci, v.astype(dtype)) for (m, rp, ci, v, k) in probs]
    B = pls.Batch(mats, [p[4] for p in probs], precision)
    return probs, B


def operands(probs, seed, dtype):
    import torch

    rng = np.random.default_rng(seed)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    X = [torch.from_numpy(rng.uniform(-1, 1, (m, k)).astype(dtype)).cuda() for (m, _, _, _, k) in probs]
    Y0 = [rng.uniform(-1, 1, (m, k)).astype(dtype) for (m, _, _, _, k) in probs]
    Y = [torch.from_numpy(y).cuda() for y in Y0]
    return X, Y, Y0, tdt


def oracle_of(orc, probs, X, alpha, beta, Y0, vals=None):
    out = []
    for i, (m, rp, ci, v, k) in enumerate(probs):
        vv = v if vals is None else vals[i]
        x = X[i].cpu().numpy().astype(np.float64)
        out.append(orc.spmm(m, k, rp, ci, vv, x, alpha, beta, None if beta == 0 else Y0[i].astype(np.float64)))
    return out


@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-0.75, 1.5)])
def test_batch_fp64_parity(pls, orc, gen, alpha, beta):
    import torch

    probs, B = make(pls, gen, 40, 11)
    X, Y, Y0, _ = operands(probs, 12, np.float64)
    if beta == 0.0:
        for y in Y:
            y.fill_(float("nan"))
    B.set_operands([x.data_ptr() for x in X], [y.data_ptr() for y in Y])
    B.exec(alpha, beta)
    torch.cuda.synchronize()
    ref = oracle_of(orc, probs, X, alpha, beta, Y0)
    for i, (m, _, _, _, k) in enumerate(probs):
        got = Y[i].cpu().numpy()
        assert rel_err(got, ref[i]) <= TOL64, (i, m, k)
        if m == 77:
            assert np.array_equal(got, beta * Y0[i]) or beta == 0.0 and not got.any()


def test_batch_shapes_and_autotune(pls, orc, gen):
    import torch

    probs, B = make(pls, gen, 24, 21)
    X, Y, Y0, _ = operands(probs, 22, np.float64)
    B.set_operands([x.data_ptr() for x in X], [y.data_ptr() for y in Y])
    ref = oracle_of(orc, probs, X, 1.0, 0.0, Y0)
    shapes = [  # (tile_nnz, lanes per k class, columns per lane per k class)
        (64, [2, 2, 4, 8, 16], [1, 1, 1, 1, 1]),
        (1024, [64, 32, 16, 8, 64], [1, 1, 1, 1, 1]),
        (4096, [4, 1, 1, 2, 4], [1, 2, 4, 4, 4]),
        (100000, [16, 2, 2, 4, 8], [1, 1, 2, 2, 2]),
    ]
    for tile, lanes, cpl in shapes:
        B.set_shape(tile, lanes, cpl)
        for y in Y:
            y.fill_(float("nan"))
        B.exec(1.0, 0.0)
        torch.cuda.synchronize()
        info = B.info()
        assert info["tile_nnz"] == tile and list(info["lanes"].values()) == lanes
        for i in range(len(probs)):
            assert rel_err(Y[i].cpu().numpy(), ref[i]) <= TOL64, (tile, lanes, cpl, i)
    # invalid shapes are rejected
    for lanes, cpl in (([3, 2, 4, 8, 16], [1] * 5), ([2, 2, 4, 8, 8], [1] * 5), ([2, 2, 4, 8, 16], [1, 4, 1, 1, 1])):
        with pytest.raises(pls.PerflibsError):
            B.set_shape(2048, lanes, cpl)
    B.optimize(autotune=True)
    info = B.info()
    assert info["ntiles"] > 0 and info["grid"] > 0
    for y in Y:
        y.fill_(float("nan"))
    B.exec(1.0, 0.0)
    torch.cuda.synchronize()
    for i in range(len(probs)):
        assert rel_err(Y[i].cpu().numpy(), ref[i]) <= TOL64


def test_batch_graph_replays_and_new_operands(pls, orc, gen):
    import torch

    probs, B = make(pls, gen, 16, 31, ks=(1, 4, 16))
    X, Y, Y0, _ = operands(probs, 32, np.float64)
    B.set_operands([x.data_ptr() for x in X], [y.data_ptr() for y in Y])
    B.exec(2.0, 0.0)
    torch.cuda.synchronize()
    first = [y.cpu().numpy().copy() for y in Y]
    ref = oracle_of(orc, probs, X, 2.0, 0.0, Y0)
    for i in range(len(probs)):
        assert rel_err(first[i], ref[i]) <= TOL64
    for _ in range(3):  # replays of the same graph, eager work in between
        for y in Y:
            y.zero_()
        torch.cuda.synchronize()
        B.exec(2.0, 0.0)
        torch.cuda.synchronize()
        for i in range(len(probs)):
            assert np.array_equal(Y[i].cpu().numpy(), first[i])
    # new X buffers through the pointer table, same graph
    X2 = [torch.flip(x, [0]).contiguous() for x in X]
    B.set_operands([x.data_ptr() for x in X2], [y.data_ptr() for y in Y])
    B.exec(2.0, 0.0)
    torch.cuda.synchronize()
    ref2 = oracle_of(orc, probs, X2, 2.0, 0.0, Y0)
    for i in range(len(probs)):
        assert rel_err(Y[i].cpu().numpy(), ref2[i]) <= TOL64
    # alpha = 0 gives beta*Y and ignores X (NaN in X)
    for x in X2:
        x.fill_(float("nan"))
    for i, y in enumerate(Y):
        y.copy_(torch.from_numpy(Y0[i]))
    B.exec(0.0, 0.5)
    torch.cuda.synchronize()
    for i in range(len(probs)):
        assert np.array_equal(Y[i].cpu().numpy(), 0.5 * Y0[i])


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_batch_reduced_precision(pls, orc, gen, precision):
    import torch

    probs, B = make(pls, gen, 24, 41, dtype=np.float32, precision=2 if precision == "bf16" else 0,
                    empty=False)
    X, Y, Y0, _ = operands(probs, 42, np.float32)
    B.set_operands([x.data_ptr() for x in X], [y.data_ptr() for y in Y])
    B.exec(1.0, 0.0)
    torch.cuda.synchronize()
    stored = [to_bf16(v.astype(np.float32)) if precision == "bf16" else v.astype(np.float32)
              for (_, _, _, v, _) in probs]
    exact_stored = oracle_of(orc, probs, X, 1.0, 0.0, Y0, vals=[s.astype(np.float64) for s in stored])
    exact64 = oracle_of(orc, probs, X, 1.0, 0.0, Y0)
    for i in range(len(probs)):
        got = Y[i].cpu().numpy()
        assert rel_err(got, exact_stored[i]) <= 2e-6
        assert rel_err(got, exact64[i]) <= (1e-6 if precision == "fp32" else 6e-3)


def test_batch_rejects_bad_input(pls, gen):
    rp, ci, v = gen.banded(100, 32, 5, 10, 5)
    A = pls.SpMat.csr(100, 100, rp, ci, v)
    S = pls.SpMat.csr(100, 100, rp, ci, v.astype(np.float32))
    with pytest.raises(pls.PerflibsError):
        pls.Batch([A, S], [1, 1])
    with pytest.raises(pls.PerflibsError):
        pls.Batch([A], [3])
    with pytest.raises(pls.PerflibsError):
        pls.Batch([A], [1], precision=2)
    Bt = pls.Batch([A], [1])
    with pytest.raises(pls.PerflibsError):
        Bt.exec(1.0, 0.0)  # operands not set


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_multi_column_pieces_and_unaligned_operands(pls, orc, gen, dtype):
    """Row shapes with cpl = 2 and 4 columns per lane, on X_p that are 16-byte
    aligned and on X_p viewed one element into a larger buffer, switching
    between the two through the pointer table of the same batch."""
    import torch

    probs, B = make(pls, gen, 24, 41, ks=(2, 4, 8, 16), dtype=dtype, empty=False)
    vals = [v.astype(dtype).astype(np.float64) for (_, _, _, v, _) in probs]
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    rng = np.random.default_rng(42)
    Xh = [rng.uniform(-1, 1, (m, k)).astype(dtype) for (m, _, _, _, k) in probs]
    aligned = [torch.from_numpy(x).cuda() for x in Xh]
    # the same values one element into a larger buffer: 8 (FP64) or 4 (FP32) bytes off
    bufs = [torch.empty(x.size + 1, dtype=tdt, device="cuda") for x in Xh]
    for b, x in zip(bufs, Xh):
        b[1:] = torch.from_numpy(x.reshape(-1)).cuda()
    shifted = [b[1:] for b in bufs]
    assert all(s.data_ptr() % 16 != 0 for s in shifted)
    Y = [torch.zeros((m, k), dtype=tdt, device="cuda") for (m, _, _, _, k) in probs]
    ref = [orc.spmm(m, k, rp, ci, vals[i], Xh[i].astype(np.float64))
           for i, (m, rp, ci, _, k) in enumerate(probs)]
    tol = TOL64 if dtype == np.float64 else 1e-6
    for cpl in (2, 4):
        # lanes = k / cpl per row (at least one), cpl capped by k
        lanes = [max(1, (1 << c) // min(cpl, 1 << c)) for c in range(5)]
        cpls = [min(cpl, 1 << c) for c in range(5)]
        B.set_shape(1024, lanes, cpls)
        for X in (aligned, shifted, aligned):
            B.set_operands([x.data_ptr() for x in X], [y.data_ptr() for y in Y])
            for y in Y:
                y.zero_()
            B.exec(1.0, 0.0)
            torch.cuda.synchronize()
            for i in range(len(probs)):
                assert rel_err(Y[i].cpu().numpy(), ref[i]) <= tol, (cpl, i)


@pytest.mark.parametrize("Cc,sigma", [(32, 0), (32, 256), (64, 64), (64, 1024)])
@pytest.mark.parametrize("dtype,precision", [(np.float64, 0), (np.float32, 0), (np.float32, 2)])
def test_batch_sell_tiles(pls, orc, gen, Cc, sigma, dtype, precision):
    """k = 1 problems as SELL-C-sigma chunk tiles beside CSR row tiles of the
    k > 1 problems: every problem against the oracle (FP64 1e-12; FP32 / BF16
    storage within 2e-6 of the exact product of the stored values), for
    alpha / beta with and without reading Y, and a return to CSR tiles."""
    import torch

    probs, B = make(pls, gen, 30, 51 + Cc + sigma, rows=(40, 5000), ks=(1, 1, 4, 16), dtype=dtype,
                    precision=precision, empty=False)
    m0 = 9  # a k = 1 problem with empty rows and a partial last chunk
    rp = np.array([0, 0, 3, 3, 4, 9, 9, 9, 10, 10], np.int32)
    ci = np.array([0, 4, 8, 2, 1, 2, 3, 5, 7, 6], np.int32)
    probs.append((m0, rp, ci, np.arange(1.0, 11.0), 1))
    B = pls.Batch([pls.SpMat.csr(m, m, r, c, v.astype(dtype)) for (m, r, c, v, k) in probs],
                  [p[4] for p in probs], precision)
    X, Y, Y0, _ = operands(probs, 52, dtype)
    B.set_operands([x.data_ptr() for x in X], [y.data_ptr() for y in Y])
    B.set_sell(Cc, sigma)
    info = B.sell_info()
    assert (info["C"], info["sigma"]) == (Cc, sigma)
    if dtype == np.float64:
        stored = [v.astype(np.float64) for (_, _, _, v, _) in probs]
    elif precision == 2:
        stored = [to_bf16(v.astype(np.float32)).astype(np.float64) for (_, _, _, v, _) in probs]
    else:
        stored = [v.astype(np.float32).astype(np.float64) for (_, _, _, v, _) in probs]
    tol = TOL64 if dtype == np.float64 else 2e-6
    for alpha, beta in ((1.0, 0.0), (-0.5, 2.0)):
        for i, y in enumerate(Y):
            y.copy_(torch.from_numpy(Y0[i]))
        B.exec(alpha, beta)
        torch.cuda.synchronize()
        ref = oracle_of(orc, probs, X, alpha, beta, Y0, vals=stored)
        for i, (m, _, _, _, k) in enumerate(probs):
            assert rel_err(Y[i].cpu().numpy(), ref[i]) <= tol, (Cc, sigma, alpha, i, m, k)
    B.set_sell(0, 0)
    assert B.sell_info()["C"] == 0
    for y in Y:
        y.fill_(float("nan"))
    B.exec(1.0, 0.0)
    torch.cuda.synchronize()
    ref = oracle_of(orc, probs, X, 1.0, 0.0, Y0, vals=stored)
    for i in range(len(probs)):
        assert rel_err(Y[i].cpu().numpy(), ref[i]) <= tol


def test_batch_sell_padding_matches_layout_rule(pls, gen):
    """Padded entries of the batch's SELL-C-sigma layout = sum over its k = 1
    problems of sum_c width_c * C, widths the longest row of each chunk after
    sorting rows by length inside sigma windows (checked here from the row
    lengths alone), and sorting windows cut the padding of banded rows."""
    probs = gen.banded_batch(12, 61, rows=(100, 3000), ks=(1,))
    B = pls.Batch([pls.SpMat.csr(m, m, r, c, v) for (m, r, c, v, k) in probs], [1] * len(probs))

    def padded(Cc, sigma):
        tot = 0
        for (m, r, c, v, k) in probs:
            lens = np.diff(r).astype(np.int64)
            win = sigma if sigma else Cc
            nwin = -(-m // win)
            lp = np.zeros(nwin * win, np.int64)
            lp[:m] = lens
            if sigma:
                lp = -np.sort(-lp.reshape(nwin, win), axis=1).reshape(-1)
            nC = -(-m // Cc)
            tot += int(lp[: nC * Cc].reshape(nC, Cc).max(axis=1).sum()) * Cc
        return tot

    for Cc, sigma in ((32, 0), (32, 256), (64, 0), (64, 1024)):
        B.set_sell(Cc, sigma)
        assert B.sell_info()["padded"] == padded(Cc, sigma), (Cc, sigma)
    assert padded(64, 1024) < 0.7 * padded(64, 0)
    with pytest.raises(pls.PerflibsError):
        B.set_sell(48, 96)
    with pytest.raises(pls.PerflibsError):
        B.set_sell(32, 48)
