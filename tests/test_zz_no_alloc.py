"""perflibs_amd_mlp_backward_noalloc: the selector MLP's backward pass with its
scratch in the library's static device array, against the allocating
perflibs_amd_mlp_backward on the same inputs (the same kernels, so the
gradient and the loss must be bit-identical), on two calls in a row, the
second with fewer samples, and its size limit."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu
K = 13  # PERFLIBS_AMD_NUM_VARIANTS
MAX_SAMPLES = 4096  # PERFLIBS_AMD_MLP_NOALLOC_MAX_SAMPLES


def _backward_both(pls, P, n, seed):
    import torch

    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 12)).astype(np.float32)
    T = rng.standard_normal((n, K)).astype(np.float32)
    M = (rng.random((n, K)) < 0.7).astype(np.float32)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    Pd, Xd, Td, Md = dev(P), dev(X), dev(T), dev(M)
    h1 = torch.empty((n, 64), device="cuda")
    h2 = torch.empty((n, 64), device="cuda")
    out = torch.empty((n, K), device="cuda")
    L = pls.lib()
    pls.check("fwd", L.perflibs_amd_mlp_forward(Pd.data_ptr(), Xd.data_ptr(), n, h1.data_ptr(),
                                                h2.data_ptr(), out.data_ptr()))
    res = []
    for fn in (L.perflibs_amd_mlp_backward, L.perflibs_amd_mlp_backward_noalloc):
        G = torch.full((len(P),), float("nan"), device="cuda")
        loss = torch.full((2,), float("nan"), device="cuda")
        pls.check("bwd", fn(Pd.data_ptr(), Xd.data_ptr(), n, h1.data_ptr(), h2.data_ptr(),
                            out.data_ptr(), Td.data_ptr(), Md.data_ptr(), G.data_ptr(),
                            loss.data_ptr()))
        torch.cuda.synchronize()
        res.append((G.cpu().numpy(), loss.cpu().numpy()))
    return res, X, T, M


def test_noalloc_backward_matches_allocating_entry_point(pls):
    rng = np.random.default_rng(11)
    P = (rng.standard_normal(mlp.num_params(K)) * 0.2).astype(np.float32)
    for n, seed in ((3000, 1), (700, 2)):  # the second call smaller than the first
        ((g_ref, l_ref), (g_na, l_na)), X, T, M = _backward_both(pls, P, n, seed)
        assert np.isfinite(g_na).all() and np.isfinite(l_na).all()
        assert np.array_equal(g_ref.view(np.uint32), g_na.view(np.uint32)), n
        assert np.array_equal(l_ref.view(np.uint32), l_na.view(np.uint32)), n
        # and both are the gradient of the masked MSE (numpy oracle)
        ref_loss, ref_g = mlp.backward(P.astype(np.float64), X, T.astype(np.float64), M, K)
        assert abs(l_na[0] - ref_loss) <= 1e-4 * ref_loss
        assert np.linalg.norm(g_na - ref_g) <= 1e-4 * np.linalg.norm(ref_g)


def test_noalloc_backward_at_the_limit_and_beyond(pls):
    import torch

    rng = np.random.default_rng(12)
    P = (rng.standard_normal(mlp.num_params(K)) * 0.2).astype(np.float32)
    ((g_ref, l_ref), (g_na, l_na)), _, _, _ = _backward_both(pls, P, MAX_SAMPLES, 3)
    assert np.array_equal(g_ref.view(np.uint32), g_na.view(np.uint32))
    assert np.array_equal(l_ref.view(np.uint32), l_na.view(np.uint32))
    # one sample over the limit: refused before anything is launched
    L = pls.lib()
    buf = torch.zeros(64, device="cuda")
    p = buf.data_ptr()
    st = L.perflibs_amd_mlp_backward_noalloc(p, p, MAX_SAMPLES + 1, p, p, p, p, p, p, p)
    assert st == pls.STATUS_INPUT_PARAMETER_ERROR
    assert b"static scratch" in L.perflibs_amd_last_error()
