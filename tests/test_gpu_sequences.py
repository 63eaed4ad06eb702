"""Seeded random sequences of matrix-object calls on the GPU path, each step
checked against a dense FP64 model of the object.

Every sequence starts from a CSR or BSR object (rectangular shapes, empty rows)
and applies a random mix of hints, optimize() under each selection mode, a
forced variant, SELL-C-sigma shape or column-panel shape, BF16 value storage
switched on and off (FP32 objects), SpMV and SpMM (NOTRANS and TRANS, alpha / beta with and without
reading y), value updates, in-place scaling and transposition, norms and
export.  The model is numpy's dense FP64 matrix of the same stored values; results must
agree to a normwise relative error of 1e-12 (FP64 objects) or 2e-6 (FP32
objects, FP32 accumulation) after every call, whatever plans the earlier calls
built, patched or dropped.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def to_bf16(a):
    """Round to the nearest BF16 (ties to even), returned as FP64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def make_object(pls, rng, dt):
    if rng.random() < 0.6:
        m, n = int(rng.integers(1, 300)), int(rng.integers(1, 300))
        dens = float(rng.uniform(0.005, 0.2))
        mask = rng.random((m, n)) < dens
        mask[rng.random(m) < 0.1] = False  # empty rows
        D = np.where(mask, rng.standard_normal((m, n)).astype(dt).astype(np.float64), 0.0)
        rows, cols = np.nonzero(D)
        rp = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=m))]).astype(np.int32)
        base = int(rng.integers(0, 2))
        A = pls.SpMat.csr(m, n, rp + base, cols.astype(np.int32) + base, D[rows, cols].astype(dt))
        return A, D, "csr"
    b = int(rng.choice([2, 4, 8]))
    mb, nb = int(rng.integers(1, 30)), int(rng.integers(1, 30))
    pattern = rng.random((mb, nb)) < float(rng.uniform(0.05, 0.4))
    rp = np.concatenate([[0], np.cumsum(pattern.sum(1))]).astype(np.int32)
    ci = np.concatenate([np.flatnonzero(r) for r in pattern] + [np.zeros(0, int)]).astype(np.int32)
    vals = rng.standard_normal(len(ci) * b * b).astype(dt)
    layout = pls.ROW_MAJOR if rng.random() < 0.5 else pls.COL_MAJOR
    D = np.zeros((mb * b, nb * b))
    for r in range(mb):
        for p in range(rp[r], rp[r + 1]):
            blk = vals[p * b * b:(p + 1) * b * b].reshape(b, b).astype(np.float64)
            D[r * b:(r + 1) * b, ci[p] * b:(ci[p] + 1) * b] = blk if layout == pls.ROW_MAJOR else blk.T
    A = pls.SpMat.bsr(layout, mb * b, nb * b, b, rp, ci, vals)
    return A, D, "bsr"


def run_sequence(pls, seed, steps=14, dt=np.float64):
    """Returns the calls it checked, in order."""
    import torch

    tol = TOL if dt == np.float64 else 2e-6
    rng = np.random.default_rng(seed)
    A, D, kind = make_object(pls, rng, dt)
    log = [f"seed {seed} {kind} {D.shape}"]
    done = []
    bf16 = False  # BF16 value storage on the device (FP32 objects)
    ops = ["hint", "optimize", "spmv", "spmv", "spmm", "update", "scale", "transpose", "norm",
           "export", "layout"] + (["precision"] if dt == np.float32 else [])
    for _ in range(steps):
        op = rng.choice(ops)
        m, n = D.shape
        P = to_bf16(D) if bf16 else D  # the values the products use
        log.append(op)
        if op == "hint":
            A.hint(pls.HINT_SPMV_OPERATION, int(rng.choice([pls.NOTRANS, pls.TRANS])))
        elif op == "optimize":
            if rng.random() < 0.5:
                A.select_mode(int(rng.choice([pls.SELECT_LEARNED, pls.SELECT_EXHAUSTIVE,
                                              pls.SELECT_HEURISTIC, pls.SELECT_FIXED_CSR])))
                A.force_variant(-1)
            else:
                A.force_variant(int(rng.integers(0, pls.NUM_VARIANTS)))
            st = pls.lib().perflibs_spmv_optimize(A.h)
            assert st in (0, 1, 2), log
            if st != 0:  # a forced layout this object cannot take
                A.force_variant(-1)
        elif op == "spmv":
            trans = int(rng.choice([pls.NOTRANS, pls.TRANS]))
            rows, cols = (m, n) if trans == pls.NOTRANS else (n, m)
            x = rng.uniform(-1, 1, cols).astype(dt)
            alpha = float(rng.choice([1.0, -0.5, 0.0]))
            beta = float(rng.choice([0.0, 1.0, 2.5]))
            y0 = rng.uniform(-1, 1, rows).astype(dt)
            y = np.full(rows, np.nan, dtype=dt) if beta == 0.0 else y0.copy()
            device = rng.random() < 0.5
            if device:
                xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
                A.spmv(xd.data_ptr(), yd.data_ptr(), alpha, beta, trans)
                torch.cuda.synchronize()
                y = yd.cpu().numpy()
            else:
                A.spmv(x, y, alpha, beta, trans)
            opD = P if trans == pls.NOTRANS else P.T
            ref = alpha * (opD @ x.astype(np.float64)) + (beta * y0.astype(np.float64) if beta != 0.0 else 0.0)
            assert rel_err(y, ref) <= tol, log
        elif op == "spmm":
            k = int(rng.choice([1, 3, 8, 17]))
            transA = int(rng.choice([pls.NOTRANS, pls.TRANS]))
            rows, inner = (m, n) if transA == pls.NOTRANS else (n, m)
            Bv = rng.uniform(-1, 1, (inner, k)).astype(dt)
            C0 = rng.uniform(-1, 1, (rows, k)).astype(dt)
            alpha = float(rng.choice([1.0, 0.75]))
            beta = float(rng.choice([0.0, -1.0]))
            Bm = pls.SpMat.dense(pls.ROW_MAJOR, inner, k, k, Bv.reshape(-1))
            Cm = pls.SpMat.dense(pls.ROW_MAJOR, rows, k, k, C0.reshape(-1))
            pls.spmm_exec(transA, pls.NOTRANS, alpha, A, Bm, beta, Cm)
            opD = P if transA == pls.NOTRANS else P.T
            ref = alpha * opD @ Bv.astype(np.float64) + beta * C0.astype(np.float64)
            assert rel_err(Cm.export_dense(pls.ROW_MAJOR), ref) <= tol, log
        elif op == "update":
            if A.export_bsr()[2] != 1:  # updates are exercised on CSR-form objects here
                continue
            _, _, rp, ci, _ = A.export_csr(0)
            nnz = int(rp[-1])
            if nnz == 0:
                continue
            pick = rng.choice(nnz, size=min(nnz, int(rng.integers(1, 6))), replace=False)
            rows = np.searchsorted(rp, pick, side="right") - 1
            cols = ci[pick]
            vals = rng.standard_normal(len(pick)).astype(dt)
            base = A.query()[0]
            A.update(rows + base, cols + base, vals)
            D[rows, cols] = vals.astype(np.float64)
        elif op == "scale":
            alpha = float(rng.choice([2.0, -0.5]))
            A.scale(alpha)
            D = alpha * D
        elif op == "transpose":
            A.transpose(pls.TRANS)
            D = D.T.copy()
            assert A.query()[1:3] == D.shape, log
        elif op == "norm":
            st, v = A.norm(pls.NORM_FRB, raise_on_error=False)
            if st == 0:  # SELL-C-sigma objects have no norm (status 1, as in the reference)
                assert v == pytest.approx(np.linalg.norm(D), rel=tol, abs=1e-300), log
                assert A.norm(pls.NORM_INF) == pytest.approx(
                    np.abs(D).sum(axis=1).max() if D.size else 0.0, rel=tol, abs=1e-300), log
            else:
                assert st == 1, log
        elif op == "layout":  # a forced SELL-C-sigma shape or column-panel shape, then optimize
            if rng.random() < 0.5:
                Cc = int(rng.choice([1, 5, 16, 32, 64]))
                A.force_sell(Cc, int(rng.choice([0, Cc, 4 * Cc])))
            else:
                A.force_panel(int(rng.choice([0, 6, 9])), int(rng.choice([0, 2])))
                A.force_variant(12)
            st = pls.lib().perflibs_spmv_optimize(A.h)
            assert st in (0, 1, 2), log
            A.force_variant(-1)
        elif op == "precision":
            bf16 = not bf16
            A.set_precision(pls.VALUES_BF16 if bf16 else pls.VALUES_NATIVE)
        elif op == "export":
            mm, nn, rp, ci, v = A.export_csr(0)
            E = np.zeros((mm, nn))
            for r in range(mm):
                E[r, ci[rp[r]:rp[r + 1]]] += v[rp[r]:rp[r + 1]]
            assert np.array_equal(E, D), log
        done.append(str(op))
    return done


@pytest.mark.parametrize("block", range(12))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_random_call_sequences(pls, block, dtype):
    counts = {}
    for seed in range(block * 16, block * 16 + 16):
        for op in run_sequence(pls, 0x5E0000 + seed, steps=16, dt=dtype):
            counts[op] = counts.get(op, 0) + 1
    # every kind of call ran in this block of sequences
    expect = {"hint", "optimize", "spmv", "spmm", "update", "scale", "transpose", "norm",
              "export", "layout"} | ({"precision"} if dtype == np.float32 else set())
    assert set(counts) == expect, counts
