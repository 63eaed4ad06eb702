"""Seeded synthetic inputs of the benchmark configurations (numpy arrays), built
by libplsparse_workloads.so (workloads/gen.cpp).

Sizes and distributions:
  laplace2d(316)        5-point Laplacian, n = 99,856, nnz = 498,016 (diag 4, off -1)
  powerlaw(4_194_304)   lognormal(2.5, 1.0) row lengths clipped to [1, 20000],
                        uniform distinct sorted columns, N(0,1) values
  stencil27(128)        27-point operator on 128^3, nnz = 55,742,968 (diag 26, off -1)
  block_stencil(...)    27-point pattern on a block grid, dense N(0,1) b x b blocks
  banded(...)           banded random rows (the batched configuration)
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent.parent
LIB_PATH = PKG / "lib" / "libplsparse_workloads.so"
_lib = None

_P = C.c_void_p
_I = C.c_int32
_U = C.c_uint64
_D = C.c_double
_L = C.c_int64

_SIGS = {
    "pls_gen_uniform": (None, [_L, _U, _D, _D, _P]),
    "pls_gen_normal": (None, [_L, _U, _P]),
    "pls_gen_stencil2d": (_L, [_I, _I, _I, _I, _P, _P, _P]),
    "pls_gen_stencil3d": (_L, [_I, _I, _I, _I, _I, _P, _P, _P]),
    "pls_gen_powerlaw_rows": (_L, [_I, _I, _D, _D, _I, _I, _U, _P]),
    "pls_gen_powerlaw_fill": (None, [_I, _I, _P, _U, _P, _P]),
    "pls_gen_banded_rows": (_L, [_I, _I, _I, _I, _I, _U, _P]),
    "pls_gen_banded_fill": (None, [_I, _I, _I, _P, _U, _P, _P]),
    "pls_gen_random_rows": (_L, [_I, _I, _I, _I, _U, _P]),
    "pls_gen_random_fill": (None, [_I, _I, _P, _U, _P, _P]),
    "pls_gen_block_stencil": (_L, [_I, _I, _I, _I, _U, _P, _P, _P]),
    "pls_gen_block_random_rows": (_L, [_I, _I, _I, _I, _P]),
    "pls_gen_block_random_fill": (None, [_I, _I, _I, _I, _P, _U, _P, _P]),
    "pls_gen_threads": (C.c_int, []),
}


class _MtxCsr(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("nnz", C.c_int64),
                ("row_ptr", C.POINTER(C.c_int32)), ("col_indx", C.POINTER(C.c_int32)),
                ("vals", C.POINTER(C.c_double))]


_SIGS["pls_mtx_read_csr"] = (C.c_int, [C.c_char_p, C.POINTER(_MtxCsr), C.c_char_p, C.c_size_t])
_SIGS["pls_mtx_free"] = (None, [C.POINTER(_MtxCsr)])


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python {PKG / 'build.py'}`")
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def uniform(n: int, seed: int, lo=-1.0, hi=1.0) -> np.ndarray:
    out = np.empty(n, np.float64)
    lib().pls_gen_uniform(n, seed, lo, hi, _p(out))
    return out


def normal(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    lib().pls_gen_normal(n, seed, _p(out))
    return out


def _two_phase(m, rows_fn, fill_fn):
    rp = np.empty(m + 1, np.int32)
    nnz = rows_fn(rp)
    col = np.empty(max(nnz, 1), np.int32)
    val = np.empty(max(nnz, 1), np.float64)
    fill_fn(rp, col, val)
    return rp, col[:nnz], val[:nnz]


def stencil2d(nx: int, ny: int | None = None, r: int = 1, star: bool = True):
    ny = nx if ny is None else ny
    m = nx * ny
    return _two_phase(
        m,
        lambda rp: lib().pls_gen_stencil2d(nx, ny, r, int(star), _p(rp), None, None),
        lambda rp, c, v: lib().pls_gen_stencil2d(nx, ny, r, int(star), _p(rp), _p(c), _p(v)),
    )


def laplace2d(nx: int):
    """5-point Laplacian on nx x nx (diagonal 4, off-diagonals -1)."""
    return stencil2d(nx, nx, 1, True)


def stencil3d(nx: int, ny: int | None = None, nz: int | None = None, r: int = 1, star: bool = False):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    m = nx * ny * nz
    return _two_phase(
        m,
        lambda rp: lib().pls_gen_stencil3d(nx, ny, nz, r, int(star), _p(rp), None, None),
        lambda rp, c, v: lib().pls_gen_stencil3d(nx, ny, nz, r, int(star), _p(rp), _p(c), _p(v)),
    )


def stencil27(n: int):
    """27-point operator on n^3 (diagonal 26, neighbours -1)."""
    return stencil3d(n, n, n, 1, False)


def powerlaw(n: int, seed: int, mu=2.5, sigma=1.0, lo=1, hi=20000, m: int | None = None):
    m = n if m is None else m
    return _two_phase(
        m,
        lambda rp: lib().pls_gen_powerlaw_rows(m, n, mu, sigma, lo, hi, seed, _p(rp)),
        lambda rp, c, v: lib().pls_gen_powerlaw_fill(m, n, _p(rp), seed, _p(c), _p(v)),
    )


def banded(m: int, bw: int, lo: int, hi: int, seed: int, n: int | None = None):
    n = m if n is None else n
    return _two_phase(
        m,
        lambda rp: lib().pls_gen_banded_rows(m, n, bw, lo, hi, seed, _p(rp)),
        lambda rp, c, v: lib().pls_gen_banded_fill(m, n, bw, _p(rp), seed, _p(c), _p(v)),
    )


def random_rows(m: int, n: int, lo: int, hi: int, seed: int):
    return _two_phase(
        m,
        lambda rp: lib().pls_gen_random_rows(m, n, lo, hi, seed, _p(rp)),
        lambda rp, c, v: lib().pls_gen_random_fill(m, n, _p(rp), seed, _p(c), _p(v)),
    )


def block_stencil(bx: int, by: int, bz: int, b: int, seed: int):
    """BSR arrays (row_ptr, col, vals[nnzb*b*b], row-major blocks)."""
    mb = bx * by * bz
    brp = np.empty(mb + 1, np.int32)
    nnzb = lib().pls_gen_block_stencil(bx, by, bz, b, seed, _p(brp), None, None)
    bci = np.empty(max(nnzb, 1), np.int32)
    bval = np.empty(max(nnzb * b * b, 1), np.float64)
    lib().pls_gen_block_stencil(bx, by, bz, b, seed, _p(brp), _p(bci), _p(bval))
    return brp, bci[:nnzb], bval[: nnzb * b * b]


def block_random(mb: int, b: int, k: int, w: int, seed: int, nb: int | None = None):
    nb = mb if nb is None else nb
    brp = np.empty(mb + 1, np.int32)
    nnzb = lib().pls_gen_block_random_rows(mb, nb, k, w, _p(brp))
    bci = np.empty(max(nnzb, 1), np.int32)
    bval = np.empty(max(nnzb * b * b, 1), np.float64)
    lib().pls_gen_block_random_fill(mb, nb, b, w, _p(brp), seed, _p(bci), _p(bval))
    return brp, bci[:nnzb], bval[: nnzb * b * b]


def bsr_to_csr_rowmajor(brp, bci, bval, b: int):
    """Expand row-major BSR (0-based) to CSR keeping every stored value: row
    rb*b + i holds, block by block, the b values of row i of each block."""
    brp = np.asarray(brp, np.int64)
    bci = np.asarray(bci, np.int64)
    mb = len(brp) - 1
    nnzb = int(brp[-1])
    counts = np.diff(brp)
    rp = np.zeros(mb * b + 1, np.int64)
    rp[1:] = np.cumsum(np.repeat(counts * b, b))
    rb = np.repeat(np.arange(mb), counts)                # block row of every block
    k_in_row = np.arange(nnzb) - brp[rb]                 # position of the block in its block row
    i = np.arange(b)[None, :, None]
    j = np.arange(b)[None, None, :]
    base = (brp[rb] * b * b + k_in_row * b)[:, None, None]
    idx = base + i * (counts[rb] * b)[:, None, None] + j  # CSR slot of value (k, i, j)
    nnz = nnzb * b * b
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    col[idx.reshape(-1)] = np.broadcast_to(bci[:, None, None] * b + j, idx.shape).reshape(-1)
    val[idx.reshape(-1)] = np.asarray(bval, np.float64)[:nnz]
    return rp.astype(np.int32), col, val


def banded_batch(nprob: int, seed: int, rows=(256, 65536), bw=(32, 512), nnz=(5, 40),
                 ks=(1, 4, 16)):
    """Problems of the batched configuration: square banded random matrices
    with rows log-uniform in `rows`, bandwidth uniform in `bw`, non-zeros per
    row uniform in `nnz` (distinct columns inside the band, N(0,1) values) and
    a right-hand-side count drawn uniformly from `ks`.  Returns a list of
    (m, row_ptr, col_indx, vals, k)."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(rows[0]), np.log(rows[1])
    out = []
    for p in range(nprob):
        m = int(round(float(np.exp(rng.uniform(lo, hi)))))
        w = int(rng.integers(bw[0], bw[1] + 1))
        k = int(ks[int(rng.integers(0, len(ks)))])
        rp, ci, v = banded(m, w, nnz[0], nnz[1], (seed * 1_000_003 + p) & 0xFFFFFFFFFFFF)
        out.append((m, rp, ci, v, k))
    return out


def read_mtx(path) -> tuple[int, int, np.ndarray, np.ndarray, np.ndarray]:
    """Matrix Market coordinate file -> (m, n, row_ptr, col_indx, vals), 0-based
    CSR sorted by (row, column), through workloads/mtx.cpp (the acceptance
    rules of the reference's benchmark reader).  Raises ValueError with the
    reader's message on a malformed file."""
    a = _MtxCsr()
    err = C.create_string_buffer(256)
    if lib().pls_mtx_read_csr(str(path).encode(), C.byref(a), err, len(err)) != 0:
        raise ValueError(err.value.decode())
    try:
        m, nnz = a.m, a.nnz
        rp = np.ctypeslib.as_array(a.row_ptr, (m + 1,)).copy()
        ci = np.ctypeslib.as_array(a.col_indx, (max(nnz, 1),))[:nnz].copy()
        v = np.ctypeslib.as_array(a.vals, (max(nnz, 1),))[:nnz].copy()
        return m, a.n, rp, ci, v
    finally:
        lib().pls_mtx_free(C.byref(a))
