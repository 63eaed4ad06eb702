"""Native BSR SpMV (k_bsr_wave, variants 10 and 11) on the handle's own blocks:
b = 4 and 8, row- and column-major blocks, FP64 / FP32 / BF16 value storage,
x aligned for vector loads and x one element off that alignment (the kernel
then reads x element by element), empty block rows, beta 0 (y never read) and
beta != 0.  Checked against the dense image of the stored values in FP64:
normwise 1e-12 for FP64 handles, 2e-6 for FP32 accumulation."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def to_bf16(a):
    """Round to the nearest BF16 (ties to even), returned as FP64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("b", [4, 8])
@pytest.mark.parametrize("layout", ["row", "col"])
@pytest.mark.parametrize("prec", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("offset", [0, 1])
def test_native_bsr_spmv(pls, b, layout, prec, offset):
    import torch

    rng = np.random.default_rng(1000 * b + 10 * offset + len(prec) + (layout == "col"))
    mb, nb = 700, 500
    lens = rng.integers(0, 30, mb)
    lens[rng.random(mb) < 0.05] = 0  # empty block rows
    cols = [np.sort(rng.choice(nb, int(l), replace=False)) for l in lens]
    brp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    bci = np.concatenate(cols).astype(np.int32)
    dt = np.float64 if prec == "f64" else np.float32
    bv = rng.standard_normal(len(bci) * b * b).astype(dt)
    lay = pls.ROW_MAJOR if layout == "row" else pls.COL_MAJOR
    A = pls.SpMat.bsr(lay, mb * b, nb * b, b, brp, bci, bv)
    if prec == "bf16":
        A.set_precision(pls.VALUES_BF16)
    A.force_variant(10 if b == 4 else 11)
    A.optimize()
    assert A.variant() == (10 if b == 4 else 11)
    stored = to_bf16(bv) if prec == "bf16" else bv.astype(np.float64)
    D = np.zeros((mb * b, nb * b))
    for rb in range(mb):
        for k in range(brp[rb], brp[rb + 1]):
            blk = stored[k * b * b:(k + 1) * b * b].reshape(b, b)
            D[rb * b:(rb + 1) * b, bci[k] * b:(bci[k] + 1) * b] = blk if layout == "row" else blk.T
    tdt = torch.float64 if dt == np.float64 else torch.float32
    tol = 1e-12 if prec == "f64" else 2e-6
    x = rng.uniform(-1, 1, nb * b).astype(dt)
    y0 = rng.uniform(-1, 1, mb * b).astype(dt)
    xbuf = torch.zeros(nb * b + 1, dtype=tdt, device="cuda")
    xbuf[offset:offset + nb * b] = torch.from_numpy(x).cuda()
    es = xbuf.element_size()
    for alpha, beta in ((1.5, 0.0), (-0.75, 0.5)):
        ybuf = torch.full((mb * b + 1,), float("nan"), dtype=tdt, device="cuda")
        if beta != 0.0:
            ybuf[offset:offset + mb * b] = torch.from_numpy(y0).cuda()
        A.spmv(xbuf.data_ptr() + offset * es, ybuf.data_ptr() + offset * es, alpha, beta)
        torch.cuda.synchronize()
        got = ybuf[offset:offset + mb * b].cpu().numpy().astype(np.float64)
        ref = alpha * (D @ x.astype(np.float64)) + (beta * y0.astype(np.float64) if beta else 0.0)
        assert rel_err(got, ref) <= tol
        # the element beside the vector is left alone
        other = ybuf[mb * b if offset == 0 else 0].item()
        assert np.isnan(other)
