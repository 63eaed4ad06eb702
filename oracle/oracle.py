"""ctypes wrapper of liboracle.so — TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
always as the parity reference or the timed CPU baseline, never as the product path.
Each wrapped C function cites the reference file:line it restates (oracle/*.c).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
_lib = None

_P = C.c_void_p
_I = C.c_int32
_L = C.c_int64
_D = C.c_double


class _Scs(C.Structure):
    _fields_ = [
        ("m", _I), ("n", _I), ("C", _I), ("sigma", _I), ("nC", _L),
        ("cs", _P), ("cl", _P), ("row_in2permd", _P), ("row_permd2in", _P),
        ("col_indx_min", _P), ("chunk_min_nnz", _P),
        ("col_indx_bytes", _I), ("n_full_chunks", _I), ("ntidyup", _I),
        ("padded", _L), ("vals", _P), ("offsets", _P),
    ]


_SIGS = {
    "orc_csr2scs": (C.c_int, [_I, _I, _P, _P, _I, _I, C.POINTER(_Scs)]),
    "orc_scs_fill": (None, [C.POINTER(_Scs), _P, _P, _P]),
    "orc_scs_free": (None, [C.POINTER(_Scs)]),
    "orc_scs_beta": (_D, [C.POINTER(_Scs), _L]),
    "orc_scs_decomp_nnz": (C.c_int, [C.POINTER(_Scs), _I, _P, _P, _P]),
    "orc_scs_beta_threads": (C.c_int, [C.POINTER(_Scs), _I, _L, C.POINTER(_D), C.POINTER(_D)]),
    "orc_scs2csr": (_L, [C.POINTER(_Scs), C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    "orc_scs_spmv": (None, [C.POINTER(_Scs), _P, _P, _D, _D]),
    "orc_bsr_check": (C.c_int, [_I, _I, _I, _I, _P]),
    "orc_bsr2csr": (_L, [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P]),
    "orc_bsr_csr_nnz": (_L, [_I, _I, _I, _P, _P]),
    "orc_csr_decomp": (None, [_I, _P, _I, _I, _P, _P]),
    "orc_spmv_dot2": (None, [_I, _P, _P, _P, _P, _D, _D, _P]),
    "orc_bsr_spmv_dot2": (None, [_I, _I, _I, _P, _P, _P, _P, _D, _D, _P]),
    "orc_spmm_dot2": (None, [_I, _I, _P, _P, _P, _P, _L, _L, _D, _D, _P, _L, _L]),
    "orc_spmv_dot2_f32vals": (None, [_I, _P, _P, _P, _P, _D, _D, _P, _P]),
    "orc_omp_threads": (C.c_int, []),
    "orc_cpu_csr_vanilla": (None, [_I, _P, _P, _P, _P, _D, _D, _P]),
    "orc_cpu_csr_tworow": (None, [_I, _P, _P, _P, _P, _D, _D, _P]),
    "orc_cpu_csr_first_touch": (None, [_I, _P, _P, _P, _P, _D, _D, _P, _I, _P, _P]),
    "orc_cpu_bsr": (None, [_I, _I, _I, _P, _P, _P, _P, _D, _D, _P]),
    "orc_cpu_spmm_rowwise": (None, [_I, _I, _P, _P, _P, _P, _I, _D, _D, _P, _I]),
}


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    ct = {np.int64: C.c_int64, np.int32: C.c_int32, np.float64: C.c_double}[dtype]
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), (n,)).copy()


def csr2scs(m, n, row_ptr, col_indx, vals, Cc, sigma, fill=True) -> dict:
    """SELL-C-sigma structure (+ data) of src/sell_c_sigma.cpp:145-336, 584-705."""
    rp, ci = _i32(row_ptr), _i32(col_indx)
    s = _Scs()
    if lib().orc_csr2scs(m, n, _p(rp), _p(ci), Cc, sigma, C.byref(s)) != 0:
        raise ValueError("invalid SELL-C-sigma parameters")
    try:
        if fill:
            v = _f64(vals)
            lib().orc_scs_fill(C.byref(s), _p(rp), _p(ci), _p(v))
        nC = s.nC
        out = dict(
            C=s.C, sigma=s.sigma, nC=nC, padded=s.padded, col_indx_bytes=s.col_indx_bytes,
            n_full_chunks=s.n_full_chunks, ntidyup=s.ntidyup,
            cs=_arr(s.cs, nC, np.int64), cl=_arr(s.cl, nC, np.int64),
            row_in2permd=_arr(s.row_in2permd, nC * s.C, np.int64),
            row_permd2in=_arr(s.row_permd2in, nC * s.C, np.int64),
            col_indx_min=_arr(s.col_indx_min, nC, np.int64),
            chunk_min_nnz=_arr(s.chunk_min_nnz, nC, np.int32),
            beta=lib().orc_scs_beta(C.byref(s), int(rp[m] - rp[0])),
        )
        if fill:
            out["vals"] = _arr(s.vals, s.padded, np.float64)
            out["offsets"] = _arr(s.offsets, s.padded, np.int64)
            rpo, cio, vo = _P(), _P(), _P()
            nnz = lib().orc_scs2csr(C.byref(s), C.byref(rpo), C.byref(cio), C.byref(vo))
            out["csr"] = (_arr(rpo, m + 1, np.int32), _arr(cio, nnz, np.int32), _arr(vo, nnz, np.float64))
            libc = C.CDLL(None)
            for p in (rpo, cio, vo):
                libc.free(p)
        return out
    finally:
        lib().orc_scs_free(C.byref(s))


def scs_beta_threads(m, n, row_ptr, col_indx, Cc, sigma, nthreads) -> dict:
    """get_beta / get_beta_threads / gen_parallel_decomp (nnz division) of a
    fresh SELL-C-sigma object (src/sell_c_sigma.cpp:369-452, 493-552)."""
    rp, ci = _i32(row_ptr), _i32(col_indx)
    s = _Scs()
    if lib().orc_csr2scs(m, n, _p(rp), _p(ci), Cc, sigma, C.byref(s)) != 0:
        raise ValueError("invalid SELL-C-sigma parameters")
    try:
        nnz = int(rp[m] - rp[0])
        chk = np.zeros(nthreads, np.int64)
        nct = np.zeros(nthreads, np.int64)
        tmn = np.zeros(nthreads, np.int64)
        clamped = lib().orc_scs_decomp_nnz(C.byref(s), nthreads, _p(chk), _p(nct), _p(tmn))
        lo, hi = C.c_double(), C.c_double()
        lib().orc_scs_beta_threads(C.byref(s), nthreads, nnz, C.byref(lo), C.byref(hi))
        return dict(clamped=bool(clamped), beta=lib().orc_scs_beta(C.byref(s), nnz), min_beta=lo.value, max_beta=hi.value,
                    chk_offset=chk, nchunks_thread=nct, thread_min_nnz=tmn,
                    cl=_arr(s.cl, s.nC, np.int64), cs=_arr(s.cs, s.nC, np.int64),
                    chunk_min_nnz=_arr(s.chunk_min_nnz, s.nC, np.int32), ntidyup=s.ntidyup,
                    n_full_chunks=s.n_full_chunks)
    finally:
        lib().orc_scs_free(C.byref(s))


def scs_spmv(m, n, row_ptr, col_indx, vals, Cc, sigma, x, alpha, beta, y):
    """Generic SELL kernel (src/sell_c_sigma.cpp:719-797) on y in place."""
    rp, ci, v, xx = _i32(row_ptr), _i32(col_indx), _f64(vals), _f64(x)
    s = _Scs()
    lib().orc_csr2scs(m, n, _p(rp), _p(ci), Cc, sigma, C.byref(s))
    try:
        lib().orc_scs_fill(C.byref(s), _p(rp), _p(ci), _p(v))
        lib().orc_scs_spmv(C.byref(s), _p(xx), _p(y), alpha, beta)
    finally:
        lib().orc_scs_free(C.byref(s))
    return y


def bsr_check(layout, m, n, b, row_ptr) -> int:
    rp = _i32(row_ptr)
    return lib().orc_bsr_check(layout, m, n, b, _p(rp))


def bsr2csr(layout, m, n, b, row_ptr, col_indx, vals):
    """src/compressed_sparse_rows.cpp:966-1056 (zeros dropped, base kept)."""
    rp, ci, v = _i32(row_ptr), _i32(col_indx), _f64(vals)
    nnz = lib().orc_bsr_csr_nnz(layout, m, b, _p(rp), _p(v))
    crp = np.zeros(m + 1, np.int32)
    cci = np.zeros(max(nnz, 1), np.int32)
    cv = np.zeros(max(nnz, 1), np.float64)
    lib().orc_bsr2csr(layout, m, n, b, _p(rp), _p(ci), _p(v), _p(crp), _p(cci), _p(cv))
    return crp, cci[:nnz], cv[:nnz]


def csr_decomp(m, row_ptr, nthreads, strategy):
    """gen_parallel_decomp_mv (src/compressed_sparse_rows.cpp:388-506)."""
    rows = np.zeros(nthreads + 1, np.int32)
    inds = np.zeros(nthreads + 1, np.int32)
    rp = _i32(row_ptr)
    lib().orc_csr_decomp(m, _p(rp), nthreads, strategy, _p(rows), _p(inds))
    return rows, inds


def spmv(m, row_ptr, col_indx, vals, x, alpha=1.0, beta=0.0, y=None):
    """Compensated FP64 y = alpha*A*x + beta*y (Dot2)."""
    yy = np.zeros(m, np.float64) if y is None else np.array(y, dtype=np.float64, copy=True)
    rp, ci, v, xx = _i32(row_ptr), _i32(col_indx), _f64(vals), _f64(x)
    lib().orc_spmv_dot2(m, _p(rp), _p(ci), _p(v), _p(xx), alpha, beta, _p(yy))
    return yy


def spmv_f32(m, row_ptr, col_indx, vals32, x32, alpha=1.0, beta=0.0, y32=None):
    """Compensated result for float-stored inputs (exact up to the final rounding)."""
    out = np.zeros(m, np.float64)
    v = np.ascontiguousarray(vals32, dtype=np.float32)
    x = np.ascontiguousarray(x32, dtype=np.float32)
    yin = None if y32 is None else np.ascontiguousarray(y32, dtype=np.float32)
    rp, ci = _i32(row_ptr), _i32(col_indx)
    lib().orc_spmv_dot2_f32vals(m, _p(rp), _p(ci), _p(v), _p(x), alpha, beta, _p(yin), _p(out))
    return out


def bsr_spmv(layout, mb, b, row_ptr, col_indx, vals, x, alpha=1.0, beta=0.0, y=None):
    yy = np.zeros(mb * b, np.float64) if y is None else np.array(y, dtype=np.float64, copy=True)
    rp, ci, v, xx = _i32(row_ptr), _i32(col_indx), _f64(vals), _f64(x)
    lib().orc_bsr_spmv_dot2(layout, mb, b, _p(rp), _p(ci), _p(v), _p(xx), alpha, beta, _p(yy))
    return yy


def spmm(m, k, row_ptr, col_indx, vals, B, alpha=1.0, beta=0.0, Cm=None):
    """Compensated C = alpha*A*B + beta*C for row-major B (n x k) and C (m x k)."""
    Bc = _f64(B)
    out = np.zeros((m, k), np.float64) if Cm is None else np.array(Cm, dtype=np.float64, copy=True)
    rp, ci, v = _i32(row_ptr), _i32(col_indx), _f64(vals)
    lib().orc_spmm_dot2(m, k, _p(rp), _p(ci), _p(v), _p(Bc), k, 1, alpha, beta, _p(out), k, 1)
    return out


def omp_threads() -> int:
    return lib().orc_omp_threads()


class ScsHandle:
    """A SELL-C-sigma object of the oracle kept alive for repeated CPU SpMV
    (the timed baseline): structure + data built once, generic kernel run on
    demand."""

    def __init__(self, m, n, row_ptr, col_indx, vals, Cc, sigma):
        self._rp, self._ci, self._v = _i32(row_ptr), _i32(col_indx), _f64(vals)
        self.s = _Scs()
        if lib().orc_csr2scs(m, n, _p(self._rp), _p(self._ci), Cc, sigma, C.byref(self.s)) != 0:
            raise ValueError("invalid SELL-C-sigma parameters")
        lib().orc_scs_fill(C.byref(self.s), _p(self._rp), _p(self._ci), _p(self._v))

    def spmv(self, x, y, alpha=1.0, beta=0.0):
        lib().orc_scs_spmv(C.byref(self.s), _p(x), _p(y), alpha, beta)

    def close(self):
        if self.s is not None:
            lib().orc_scs_free(C.byref(self.s))
            self.s = None

    def __del__(self):
        self.close()


def scs_beta(m, n, row_ptr, col_indx, Cc, sigma) -> float:
    """Padding efficiency of a candidate layout (structure only)."""
    rp, ci = _i32(row_ptr), _i32(col_indx)
    s = _Scs()
    if lib().orc_csr2scs(m, n, _p(rp), _p(ci), Cc, sigma, C.byref(s)) != 0:
        raise ValueError("invalid SELL-C-sigma parameters")
    try:
        return lib().orc_scs_beta(C.byref(s), int(rp[m] - rp[0]))
    finally:
        lib().orc_scs_free(C.byref(s))
