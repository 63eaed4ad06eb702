"""GPU parity of the SpMV path through the C-ABI.

FP64 results must agree with the compensated FP64 oracle to a normwise
relative error of 1e-12; integer layouts (SELL-C-sigma permutation, widths,
offsets, padding) must be bit-identical to the oracle's restatement of the
reference's csr2scs / gen_data_vectors.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
ALL_VARIANTS = list(range(13))
PANEL = 12


def powerlawish(rng, m, n, mean=8.0, empty=0.05, base=0):
    lens = np.minimum(n, np.maximum(0, np.round(np.exp(rng.normal(np.log(mean), 1.0, m))))).astype(np.int64)
    lens[rng.random(m) < empty] = 0
    rp = np.zeros(m + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    ci = np.concatenate([np.sort(rng.choice(n, int(l), replace=False)) for l in lens])
    return (rp + base).astype(np.int32), (ci + base).astype(np.int32), rng.standard_normal(int(rp[-1]))


def run_variant(pls, m, n, rp, ci, v, x, alpha, beta, y0, variant=None, sell=None, trans=300):
    A = pls.SpMat.csr(m, n, rp, ci, v)
    A.hint(pls.HINT_SPMV_OPERATION, trans)
    if variant is not None:
        A.force_variant(variant)
    if sell is not None:
        A.force_sell(*sell)
    A.optimize()
    y = np.array(y0, dtype=v.dtype, copy=True)
    A.spmv(np.asarray(x, dtype=v.dtype), y, alpha, beta, trans=trans)
    return y, A


# ---------------------------------------------------------------- goldens
@pytest.mark.parametrize("variant", ALL_VARIANTS)
def test_spmv_basic_golden_every_variant(pls, golden, variant):
    g = golden["spmv_basic_d"]
    A = pls.SpMat.csr(g["m"], g["n"], g["row_ptr"], g["col_indx"], np.array(g["vals"]))
    A.force_variant(variant)
    A.optimize()
    y = np.zeros(5)
    A.spmv(np.array(g["x"]), y, g["alpha"], g["beta"])
    assert np.max(np.abs(y - g["expected"])) <= g["tol"]


def test_alpha_beta_golden(pls, golden):
    g = golden["alpha_beta_semantics"]
    A = pls.SpMat.csr(4, 4, g["row_ptr"], g["col_indx"], np.array(g["vals"]))
    y = np.array(g["y_init"])
    A.spmv(np.array(g["x_notrans"]), y, g["alpha"], g["beta"])
    assert np.max(np.abs(y - g["expected_notrans"])) <= g["tol"]
    y = np.zeros(4)
    A.spmv(np.array(g["x_trans"]), y, 1.0, 0.0, trans=pls.TRANS)
    assert np.max(np.abs(y - g["expected_trans"])) <= g["tol"]


def test_bsr_goldens(pls, golden):
    g = golden["bsr_row_major"]
    A = pls.SpMat.bsr(g["layout"], 4, 4, 2, g["row_ptr"], g["col_indx"], np.array(g["vals"]))
    y = np.zeros(4)
    A.spmv(np.array(g["x"]), y)
    assert np.max(np.abs(y - g["expected_notrans"])) <= g["tol"]
    y = np.zeros(4)
    A.spmv(np.array(g["x"]), y, trans=pls.TRANS)  # converts to CSR like the reference
    assert np.max(np.abs(y - g["expected_trans"])) <= g["tol"]
    h = golden["bsr_col_major_transposed_half"]
    B = pls.SpMat.bsr(h["layout"], 4, 4, 2, h["row_ptr"], h["col_indx"], np.array(h["vals"]))
    y = np.zeros(4)
    B.spmv(np.array(h["x"]), y, h["alpha"], 0.0, trans=pls.TRANS)
    assert np.max(np.abs(y - h["expected_trans"])) <= h["tol"]


@pytest.mark.parametrize("dtype,tol_key", [(np.float64, "tol_d"), (np.float32, "tol_s")])
def test_strategy_paths_flow(pls, golden, dtype, tol_key):
    g = golden["strategy_paths"]
    m, w, s = g["m"], g["width"], g["col_stride"]
    tol = g[tol_key]
    rp = np.arange(m + 1, dtype=np.int32) * w
    ci = (np.arange(m)[:, None] * s + np.arange(w)[None, :]).reshape(-1).astype(np.int32)
    A = pls.SpMat.csr(m, s * m, rp, ci, np.ones(m * w, dtype))
    A.hint(pls.HINT_SPMV_OPERATION, pls.NOTRANS)
    A.hint(pls.HINT_SPMV_INVOCATIONS, pls.INVOCATIONS_MANY)
    A.optimize()
    _, _, erp, eci, _ = A.export_csr(0)
    e = g["export_after_optimize"]
    assert erp[m] == e["row_ptr_last"] and eci[0] == e["col_first"] and eci[-1] == e["col_last"]
    x = np.ones(s * m, dtype)
    y = np.zeros(m, dtype)
    A.spmv(x, y)
    assert np.all(np.abs(y - g["expected_y"]) <= tol)
    u = g["update"]
    A.update([u["row"]], [u["col"]], [u["value"]])
    y[:] = 0
    A.spmv(x, y)
    exp = np.full(m, g["expected_y"])
    exp[u["row"]] = u["expected_row_value"]
    assert np.all(np.abs(y - exp) <= tol)
    ab = g["alpha_beta"]
    y[:] = ab["y0"]
    A.spmv(x, y, ab["alpha"], ab["beta"])
    exp = np.full(m, ab["expected"])
    exp[u["row"]] = ab["expected_updated_row"]
    assert np.all(np.abs(y - exp) <= tol)


def test_csr_omp_paths_fp32_golden(pls, golden):
    from test_oracle import omp_paths_case

    g, rp, ci, vals, x, y, expected = omp_paths_case(golden)
    A = pls.SpMat.csr(g["m"], g["n"], rp, ci, vals)
    A.hint(pls.HINT_SPMV_INVOCATIONS, pls.INVOCATIONS_MANY)
    A.optimize()
    yy = y.copy()
    A.spmv(x, yy, g["alpha"], g["beta"])
    assert np.max(np.abs(yy - expected)) <= g["tol"]


# ---------------------------------------------------------------- parity
@pytest.mark.parametrize("variant", ALL_VARIANTS)
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.5, 0.5)])
def test_variant_parity_random(pls, orc, variant, alpha, beta):
    rng = np.random.default_rng(1000 + variant)
    m, n = 4096, 3072  # multiples of 8 so BSR variants apply
    rp, ci, v = powerlawish(rng, m, n, mean=10.0)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, m)
    ref = orc.spmv(m, rp, ci, v, x, alpha, beta, y0)
    y, A = run_variant(pls, m, n, rp, ci, v, x, alpha, beta, y0, variant=variant)
    assert A.variant() == variant
    assert rel_err(y, ref) <= TOL64


def test_one_based_and_long_rows(pls, orc):
    rng = np.random.default_rng(5)
    m, n = 3000, 50000
    rp, ci, v = powerlawish(rng, m, n, mean=30.0, base=1)
    # a few very long rows (one workgroup each in the adaptive kernel)
    x = rng.uniform(-1, 1, n)
    ref = orc.spmv(m, rp, ci, v, x, 2.0, 0.0)
    for variant in (0, 1, 5, 9, PANEL):
        y, _ = run_variant(pls, m, n, rp, ci, v, x, 2.0, 0.0, np.zeros(m), variant=variant)
        assert rel_err(y, ref) <= TOL64
    lens = np.full(8, 9000)
    rp2 = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci2 = np.concatenate([np.sort(rng.choice(n, 9000, replace=False)) for _ in lens]).astype(np.int32)
    v2 = rng.standard_normal(int(rp2[-1]))
    ref2 = orc.spmv(8, rp2, ci2, v2, x)
    for variant in (0, 1, 6, PANEL):
        y, _ = run_variant(pls, 8, n, rp2, ci2, v2, x, 1.0, 0.0, np.zeros(8), variant=variant)
        assert rel_err(y, ref2) <= TOL64


@pytest.mark.parametrize("Cc,sigma", [(1, 0), (2, 8), (3, 9), (7, 70), (16, 3200), (32, 64), (64, 0), (64, 128000)])
def test_sell_layout_bit_exact(pls, orc, Cc, sigma):
    rng = np.random.default_rng(Cc * 7 + sigma)
    m, n = 2077, 40000
    rp, ci, v = powerlawish(rng, m, n, mean=6.0, empty=0.1)
    v[rng.random(len(v)) < 0.03] = 0.0
    x = rng.uniform(-1, 1, n)
    y, A = run_variant(pls, m, n, rp, ci, v, x, 1.0, 0.0, np.zeros(m), sell=(Cc, sigma))
    got = A.sell_layout()
    exp = orc.csr2scs(m, n, rp, ci, v, Cc, sigma)
    for key in ("C", "sigma", "nC", "padded", "col_indx_bytes", "n_full_chunks", "ntidyup"):
        assert got[key] == exp[key], key
    for key in ("cs", "cl", "row_permd2in", "col_indx_min", "offsets", "vals"):
        assert np.array_equal(got[key], exp[key]), key
    assert rel_err(y, orc.spmv(m, rp, ci, v, x)) <= TOL64
    # export from the SELL object drops explicit zeros (scs2csr)
    _, _, erp, eci, ev = A.export_csr(0)
    crp, cci, cv = exp["csr"]
    assert np.array_equal(erp, crp) and np.array_equal(eci, cci) and np.array_equal(ev, cv)


def test_sell_generic_kernel_order_matches_reference_kernel(pls, orc):
    """C = 16 SELL on the GPU against the reference's generic SELL kernel order."""
    rng = np.random.default_rng(17)
    m, n = 1000, 900
    rp, ci, v = powerlawish(rng, m, n)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, m)
    ref = orc.scs_spmv(m, n, rp, ci, v, 16, 3200, x, 1.5, 0.25, y0.copy())
    y, _ = run_variant(pls, m, n, rp, ci, v, x, 1.5, 0.25, y0, sell=(16, 3200))
    assert rel_err(y, ref) <= TOL64


@pytest.mark.parametrize("pw,wpc", [(3, 0), (7, 2), (12, 1), (24, 1)])
def test_panel_shapes(pls, orc, pw, wpc):
    """Column-panel variant: many narrow panels, few resident workgroups and
    tiles capped by the packed-index width (several rounds per workgroup)."""
    rng = np.random.default_rng(pw * 31 + wpc)
    m, n = (600_000, 40_000) if pw == 24 else (5000, 3000)
    rp, ci, v = powerlawish(rng, m, n, mean=3.0 if pw == 24 else 9.0, empty=0.2)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, m)
    A = pls.SpMat.csr(m, n, rp, ci, v)
    A.force_panel(pw, wpc)
    A.optimize()
    assert A.variant() == PANEL
    info = A.panel_info()
    assert info["panel_log2"] == pw
    assert info["npanels"] == max(1, -(-n // (1 << pw)))
    assert info["ntiles"] % 4 == 0 and info["rows_per_tile_max"] << pw < 1 << 32
    if pw == 24:
        assert info["ntiles"] > 4 * info["grid"]  # more than one round of tiles
    for alpha, beta in ((1.0, 0.0), (-0.75, 1.5)):
        y = y0.copy()
        A.spmv(x, y, alpha, beta)
        assert rel_err(y, orc.spmv(m, rp, ci, v, x, alpha, beta, y0)) <= TOL64
    y = np.full(m, np.nan)
    A.spmv(x, y, 1.0, 0.0)
    assert rel_err(y, orc.spmv(m, rp, ci, v, x)) <= TOL64
    # fixed summation order: repeated launches give identical bits
    for _ in range(3):
        y2 = np.full(m, np.nan)
        A.spmv(x, y2, 1.0, 0.0)
        assert np.array_equal(y2, y)


# ---------------------------------------------------------------- semantics
def test_beta_zero_alpha_zero_and_empty(pls):
    A = pls.SpMat.csr(3, 3, [0, 1, 2, 3], [0, 1, 2], np.array([2.0, 3.0, 4.0]))
    y = np.array([np.nan, np.inf, 1.0])
    A.spmv(np.ones(3), y, 1.0, 0.0)
    assert np.array_equal(y, [2.0, 3.0, 4.0])
    y = np.array([1.0, 2.0, 3.0])
    A.spmv(np.array([np.nan] * 3), y, 0.0, 2.0)
    assert np.array_equal(y, [2.0, 4.0, 6.0])
    y = np.array([np.nan, 1.0, 1.0])
    A.spmv(np.ones(3), y, 0.0, 0.0)
    assert np.array_equal(y, [0.0, 0.0, 0.0])
    E = pls.SpMat.csr(3, 3, [0, 0, 0, 0], [0], np.array([0.0]))
    E.optimize()
    y = np.array([1.0, 2.0, 3.0])
    E.spmv(np.ones(3), y, 1.0, 0.5)
    assert np.array_equal(y, [0.5, 1.0, 1.5])
    Z = pls.SpMat.csr(0, 3, [0], [0], np.array([0.0]))
    assert Z.spmv(np.ones(3), np.zeros(1), raise_on_error=False) == 0
    W = pls.SpMat.csr(2, 0, [0, 0, 0], [0], np.array([0.0]))
    y = np.array([1.0, 2.0])
    W.spmv(np.zeros(1), y, 1.0, 3.0)
    assert np.array_equal(y, [3.0, 6.0])


def test_identity_and_null(pls):
    I = pls.SpMat.identity(4)
    y = np.array([1.0, 1.0, 1.0, 1.0])
    I.spmv(np.array([1.0, 2.0, 3.0, 4.0]), y, 2.0, 0.5)
    assert np.array_equal(y, [2.5, 4.5, 6.5, 8.5])
    N = pls.SpMat.null(4, 4)
    y = np.array([1.0, 2.0, 3.0, 4.0])
    N.spmv(np.ones(4), y, 1.0, 2.0)
    assert np.array_equal(y, [2.0, 4.0, 6.0, 8.0])


@pytest.mark.parametrize("trans", [310, 320])
def test_transpose_ops(pls, orc, trans):
    rng = np.random.default_rng(trans)
    m, n = 1500, 2100
    rp, ci, v = powerlawish(rng, m, n)
    x = rng.uniform(-1, 1, m)
    y0 = rng.uniform(-1, 1, n)
    D = np.zeros((m, n))
    for i in range(m):
        D[i, ci[rp[i]:rp[i + 1]]] += v[rp[i]:rp[i + 1]]
    # oracle of A^T: its CSR is the CSC of A
    Tcsr = np.asfortranarray(D)
    trp = np.concatenate([[0], np.cumsum((Tcsr != 0).sum(axis=0))]).astype(np.int32)
    tci = np.concatenate([np.nonzero(D[:, j])[0] for j in range(n)]).astype(np.int32)
    tv = np.concatenate([D[np.nonzero(D[:, j])[0], j] for j in range(n)])
    ref = orc.spmv(n, trp, tci, tv, x, 1.25, -0.5, y0)
    for variant in (0, 1, 8):
        y, A = run_variant(pls, m, n, rp, ci, v, x, 1.25, -0.5, y0, variant=variant, trans=trans)
        assert rel_err(y, ref) <= TOL64
    # an exec op that differs from the hint converts to CSR and still computes op(A)
    A = pls.SpMat.csr(m, n, rp, ci, v)
    A.force_variant(8)
    A.optimize()
    y = y0.copy()
    A.spmv(x, y, 1.25, -0.5, trans=trans)
    assert rel_err(y, ref) <= TOL64


def test_device_pointers_async(pls, orc):
    import torch

    rng = np.random.default_rng(8)
    m, n = 5000, 5000
    rp, ci, v = powerlawish(rng, m, n)
    x = rng.uniform(-1, 1, n)
    A = pls.SpMat.csr(m, n, rp, ci, v)
    A.force_variant(1)
    A.optimize()
    xd = torch.from_numpy(x).cuda()
    yd = torch.full((m,), 0.5, dtype=torch.float64, device="cuda")
    A.spmv_async(xd.data_ptr(), yd.data_ptr(), 1.0, 2.0)
    torch.cuda.synchronize()
    ref = orc.spmv(m, rp, ci, v, x, 1.0, 2.0, np.full(m, 0.5))
    assert rel_err(yd.cpu().numpy(), ref) <= TOL64
    # the synchronous drop-in call also takes device pointers
    yd.fill_(0.0)
    A.spmv(xd.data_ptr(), yd.data_ptr())
    assert rel_err(yd.cpu().numpy(), orc.spmv(m, rp, ci, v, x)) <= TOL64


# ---------------------------------------------------------------- precision
def to_bf16(a):
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("variant", [0, 1, 9, 10, PANEL])
def test_reduced_precision(pls, orc, precision, variant):
    rng = np.random.default_rng(77)
    m, n = 4096, 4096
    rp, ci, v = powerlawish(rng, m, n, mean=12.0)
    x = rng.uniform(-1, 1, n)
    v32, x32 = v.astype(np.float32), x.astype(np.float32)
    A = pls.SpMat.csr(m, n, rp, ci, v32)
    if precision == "bf16":
        A.set_precision(pls.VALUES_BF16)
    A.force_variant(variant)
    A.optimize()
    y = np.zeros(m, np.float32)
    A.spmv(x32, y)
    stored = to_bf16(v32) if precision == "bf16" else v32
    exact_rounded = orc.spmv_f32(m, rp, ci, stored, x32)
    # FP32 accumulation of the stored values: the arithmetic error alone
    assert rel_err(y, exact_rounded) <= 2e-6
    err64 = rel_err(y, orc.spmv(m, rp, ci, v, x))
    assert err64 <= (1e-6 if precision == "fp32" else 6e-3)


# ---------------------------------------------------------------- full-size configurations
def test_laplacian_316_full(pls, orc, gen):
    rp, ci, v = gen.laplace2d(316)
    n = 316 * 316
    assert rp[-1] == 498016
    x = gen.uniform(n, 0x5EED0001)
    y0 = gen.uniform(n, 0x5EED1001)
    ref = orc.spmv(n, rp, ci, v, x, 1.0, 0.5, y0)
    A = pls.SpMat.csr(n, n, rp, ci, v)
    A.hint(pls.HINT_SPMV_INVOCATIONS, pls.INVOCATIONS_MANY)
    A.optimize()
    y = y0.copy()
    A.spmv(x, y, 1.0, 0.5)
    assert rel_err(y, ref) <= TOL64
    # closed form: with x = 1, interior rows sum to 0 and boundary rows are positive
    y = np.zeros(n)
    A.spmv(np.ones(n), y)
    grid = y.reshape(316, 316)
    assert np.all(grid[1:-1, 1:-1] == 0.0) and np.all(grid[0] > 0) and np.all(grid[:, 0] > 0)


def test_powerlaw_4m_full(pls, orc, gen):
    n = 4_194_304
    rp, ci, v = gen.powerlaw(n, 0x5EED0002)
    x = gen.uniform(n, 0x5EED1002)
    ref = orc.spmv(n, rp, ci, v, x)
    A = pls.SpMat.csr(n, n, rp, ci, v)
    for variant in (1, PANEL, 9):
        A.force_variant(variant)
        A.optimize()
        y = np.zeros(n)
        A.spmv(x, y)
        assert rel_err(y, ref) <= TOL64
        if variant == 9:  # SELL object now: linearity A(2x) = 2 Ax
            y2 = np.zeros(n)
            A.spmv(2 * x, y2)
            assert np.array_equal(y2, 2 * y)


def test_coo_and_csc_objects(pls, orc):
    """SpMV on objects created from COO and CSC arrays (converted to CSR at creation)."""
    rng = np.random.default_rng(123)
    m, n = 3000, 2500
    rp, ci, v = powerlawish(rng, m, n, mean=9.0)
    rows = np.repeat(np.arange(m), np.diff(rp))
    perm = rng.permutation(len(v))
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, m)
    ref = orc.spmv(m, rp, ci, v, x, 0.5, 2.0, y0)
    A = pls.SpMat.coo(m, n, rows[perm], ci[perm], v[perm])
    A.optimize()
    y = y0.copy()
    A.spmv(x, y, 0.5, 2.0)
    assert rel_err(y, ref) <= TOL64
    # CSC of A = CSR of A^T
    order = np.lexsort((rows, ci))
    col_ptr = np.concatenate([[0], np.cumsum(np.bincount(ci, minlength=n))]).astype(np.int32) + 1
    Bc = pls.SpMat.csc(m, n, rows[order] + 1, col_ptr, v[order])
    Bc.optimize()
    y = y0.copy()
    Bc.spmv(x, y, 0.5, 2.0)
    assert rel_err(y, ref) <= TOL64


def test_csc_and_coo_goldens(pls, golden):
    g = golden["csc_kernel_paths_d"]
    A = pls.SpMat.csc(g["m"], g["n"], g["row_indx"], g["col_ptr"], np.array(g["vals"]))
    y = np.zeros(4)
    A.spmv(np.array(g["x"]), y)
    assert np.max(np.abs(y - g["expected_notrans"])) <= g["tol"]
    y = np.zeros(4)
    A.spmv(np.array(g["x"]), y, trans=pls.TRANS)
    assert np.max(np.abs(y - g["expected_trans"])) <= g["tol"]
    u = g["update"]
    A.update([u["row"]], [u["col"]], [u["value"]])
    _, _, rp, ci, v = A.export_csr(0)
    D = np.zeros((4, 4))
    for i in range(4):
        D[i, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    assert np.array_equal(D.reshape(-1), g["dense_after_update_row_major"])
    c = golden["coo_spmv_paths_real_part"]
    B = pls.SpMat.coo(c["m"], c["n"], c["row_indx"], c["col_indx"], np.array(c["vals"]))
    assert B.query() == (0, 4, 4, 5)
    y = np.zeros(4)
    B.spmv(np.array(c["x"]), y)
    assert np.max(np.abs(y - c["expected_notrans"])) <= c["tol"]
    y = np.zeros(4)
    B.spmv(np.array(c["x"]), y, trans=pls.TRANS)
    assert np.max(np.abs(y - c["expected_trans"])) <= c["tol"]
    u = c["update"]
    B.update([u["row"]], [u["col"]], [u["value"]])
    y = np.zeros(4)
    B.spmv(np.array(c["x"]), y)
    assert np.max(np.abs(y - c["expected_after_update"])) <= c["tol"]
    assert B.update([c["bad_update"]["row"]], [c["bad_update"]["col"]], [1.0], raise_on_error=False) == 1
