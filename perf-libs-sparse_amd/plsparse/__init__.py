"""ctypes view of libplsparse_amd.so — the C-ABI declared in include/perflibs_sparse.h
and include/perflibs_sparse_amd.h.

This module only marshals arguments: every operation runs inside the library
(C++ host code and gfx950 HIP kernels).  Importing it fails loudly when the
library has not been built; there is no fallback.

Vectors may be numpy arrays (host memory: the library stages them through HBM)
or device pointers given as ints (e.g. ``torch_tensor.data_ptr()``), which the
library uses in place.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent.parent
LIB_PATH = PKG / "lib" / "libplsparse_amd.so"

# ---- enum values of perflibs_sparse.h ----
STATUS_SUCCESS, STATUS_INPUT_PARAMETER_ERROR, STATUS_EXECUTION_FAILURE = 0, 1, 2
HINT_STRUCTURE, HINT_SPMV_OPERATION, HINT_SPMM_OPERATION = 50, 60, 61
HINT_MEMORY, HINT_SPMV_INVOCATIONS, HINT_SPMM_INVOCATIONS = 70, 80, 81
COL_MAJOR, ROW_MAJOR = 90, 91
STRUCTURE_UNSTRUCTURED = 101
NOTRANS, TRANS, CONJTRANS = 300, 310, 320
INVOCATIONS_SINGLE, INVOCATIONS_FEW, INVOCATIONS_MANY = 400, 410, 420
SCALAR_ONE, SCALAR_ZERO, SCALAR_ANY = 500, 501, 502
CREATE_NOCOPY = 1
NORM_INF, NORM_FRB = 1001, 1002

# ---- perflibs_sparse_amd.h ----
VARIANTS = [
    "csr_vector", "csr_adaptive",
    "sell_c32_s0", "sell_c32_s64", "sell_c32_s1024", "sell_c32_s16384",
    "sell_c64_s0", "sell_c64_s64", "sell_c64_s1024", "sell_c64_s16384",
    "bsr_b4", "bsr_b8", "csr_panel",
]
NUM_VARIANTS = len(VARIANTS)
NUM_FEATURES = 12
MLP_HIDDEN = 64
SELECT_LEARNED, SELECT_EXHAUSTIVE, SELECT_HEURISTIC, SELECT_FIXED_CSR, SELECT_NETWORK = 0, 1, 2, 3, 4
PATH_NONE, PATH_FIXED, PATH_HEURISTIC, PATH_NETWORK, PATH_AUDITION, PATH_EXHAUSTIVE = -1, 0, 1, 2, 3, 4
VALUES_NATIVE, VALUES_BF16 = 0, 2


class PerflibsError(RuntimeError):
    def __init__(self, fn: str, status: int):
        super().__init__(f"{fn} returned status {status}: {lib().perflibs_amd_last_error().decode()}")
        self.status = status


_lib = None

_I = C.c_int32
_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_D = C.c_double
_F = C.c_float

_SIGS = {
    "perflibs_spmat_create_csr_s": (_I, [_PP, _I, _I, _P, _P, _P, _I]),
    "perflibs_spmat_create_csr_d": (_I, [_PP, _I, _I, _P, _P, _P, _I]),
    "perflibs_spmat_create_csc_s": (_I, [_PP, _I, _I, _P, _P, _P, _I]),
    "perflibs_spmat_create_csc_d": (_I, [_PP, _I, _I, _P, _P, _P, _I]),
    "perflibs_spmat_create_coo_s": (_I, [_PP, _I, _I, _I, _P, _P, _P, _I]),
    "perflibs_spmat_create_coo_d": (_I, [_PP, _I, _I, _I, _P, _P, _P, _I]),
    "perflibs_spmat_create_dense_s": (_I, [_PP, C.c_int, _I, _I, _I, _P, _I]),
    "perflibs_spmat_create_dense_d": (_I, [_PP, C.c_int, _I, _I, _I, _P, _I]),
    "perflibs_spmat_create_bsr_s": (_I, [_PP, C.c_int, _I, _I, _I, _P, _P, _P, _I]),
    "perflibs_spmat_create_bsr_d": (_I, [_PP, C.c_int, _I, _I, _I, _P, _P, _P, _I]),
    "perflibs_spmat_create_null": (_P, [_I, _I]),
    "perflibs_spmat_create_identity": (_P, [_I]),
    "perflibs_spmat_destroy": (_I, [_P]),
    "perflibs_spmat_query": (_I, [_P, _P, _P, _P, _P]),
    "perflibs_spmat_hint": (_I, [_P, C.c_int, C.c_int]),
    "perflibs_spmv_optimize": (_I, [_P]),
    "perflibs_spmm_optimize": (_I, [C.c_int, C.c_int, C.c_int, _P, _P, C.c_int, _P]),
    "perflibs_spmv_exec_s": (_I, [C.c_int, _F, _P, _P, _F, _P]),
    "perflibs_spmv_exec_d": (_I, [C.c_int, _D, _P, _P, _D, _P]),
    "perflibs_spmm_exec_s": (_I, [C.c_int, C.c_int, _F, _P, _P, _F, _P]),
    "perflibs_spmm_exec_d": (_I, [C.c_int, C.c_int, _D, _P, _P, _D, _P]),
    "perflibs_spmat_update_s": (_I, [_P, _I, _P, _P, _P]),
    "perflibs_spmat_update_d": (_I, [_P, _I, _P, _P, _P]),
    "perflibs_spmat_export_csr_s": (_I, [_P, _I, _P, _P, _PP, _PP, _PP]),
    "perflibs_spmat_export_csr_d": (_I, [_P, _I, _P, _P, _PP, _PP, _PP]),
    "perflibs_spmat_export_bsr_s": (_I, [_P, _I, _I, _P, _P, _P, _PP, _PP, _PP]),
    "perflibs_spmat_export_bsr_d": (_I, [_P, _I, _I, _P, _P, _P, _PP, _PP, _PP]),
    "perflibs_spmat_export_dense_s": (_I, [_P, C.c_int, _P, _P, _PP]),
    "perflibs_spmat_export_dense_d": (_I, [_P, C.c_int, _P, _P, _PP]),
    "perflibs_spmat_print_err": (None, [_P]),
    "perflibs_spmat_export_csc_s": (_I, [_P, _I, _P, _P, _PP, _PP, _PP]),
    "perflibs_spmat_export_csc_d": (_I, [_P, _I, _P, _P, _PP, _PP, _PP]),
    "perflibs_spmat_export_coo_s": (_I, [_P, _P, _P, _P, _PP, _PP, _PP]),
    "perflibs_spmat_export_coo_d": (_I, [_P, _P, _P, _P, _PP, _PP, _PP]),
    "perflibs_spscale_exec_s": (_I, [_F, _P]),
    "perflibs_spscale_exec_d": (_I, [_D, _P]),
    "perflibs_sptranspose_exec_s": (_I, [C.c_int, _P]),
    "perflibs_sptranspose_exec_d": (_I, [C.c_int, _P]),
    "perflibs_spnorm_exec_s": (_I, [_P, C.c_int, _P]),
    "perflibs_spnorm_exec_d": (_I, [_P, C.c_int, _P]),
    "perflibs_amd_set_stream": (_I, [_P]),
    "perflibs_amd_get_stream": (_P, []),
    "perflibs_amd_spmv_exec_async_d": (_I, [C.c_int, _D, _P, _P, _D, _P]),
    "perflibs_amd_spmv_exec_async_s": (_I, [C.c_int, _F, _P, _P, _F, _P]),
    "perflibs_amd_spmm_exec_async_d": (_I, [_D, _P, _I, _P, _I, _D, _P, _I]),
    "perflibs_amd_spmm_exec_async_s": (_I, [_F, _P, _I, _P, _I, _F, _P, _I]),
    "perflibs_amd_set_select_mode": (_I, [_P, C.c_int]),
    "perflibs_amd_force_variant": (_I, [_P, C.c_int]),
    "perflibs_amd_get_variant": (_I, [_P, _P]),
    "perflibs_amd_set_precision": (_I, [_P, C.c_int]),
    "perflibs_amd_force_sell": (_I, [_P, _I, _I]),
    "perflibs_amd_force_panel": (_I, [_P, _I, _I]),
    "perflibs_amd_panel_info": (_I, [_P, _P, _P, _P, _P, _P]),
    "perflibs_amd_spmm_info": (_I, [_P, _P, _P, _P, _P]),
    "perflibs_amd_set_spmm_rows": (_I, [_P, _I]),
    "perflibs_amd_spmm_rows_clustered": (_I, [_P, _P]),
    "perflibs_amd_spmm_blocks_inspect": (_I, [_I, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P]),
    "perflibs_amd_features": (_I, [_P, _P]),
    "perflibs_amd_time_variants": (_I, [_P, C.c_int, C.c_int, _P]),
    "perflibs_amd_sell_beta": (_I, [_P, _I, _I, _I, _P, _P, _P]),
    "perflibs_amd_mlp_num_params": (C.c_size_t, []),
    "perflibs_amd_set_selector_model": (_I, [_P, _P, _P]),
    "perflibs_amd_clear_selector_model": (_I, []),
    "perflibs_amd_load_selector_model": (_I, [C.c_char_p]),
    "perflibs_amd_set_selector_range": (_I, [_P, _P]),
    "perflibs_amd_select_info": (_I, [_P, _P, _P]),
    "perflibs_amd_mlp_forward": (_I, [_P, _P, C.c_int, _P, _P, _P]),
    "perflibs_amd_mlp_backward": (_I, [_P, _P, C.c_int, _P, _P, _P, _P, _P, _P, _P]),
    "perflibs_amd_mlp_backward_noalloc": (_I, [_P, _P, C.c_int, _P, _P, _P, _P, _P, _P, _P]),
    "perflibs_amd_mlp_adam": (_I, [_P, _P, _P, _P, C.c_int, _F, _F, _F, _F]),
    "perflibs_amd_sell_info": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "perflibs_amd_sell_export": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "perflibs_amd_batch_create": (_I, [_PP, C.c_int64, _P, _P, C.c_int]),
    "perflibs_amd_batch_destroy": (_I, [_P]),
    "perflibs_amd_batch_set_operands": (_I, [_P, _P, _P]),
    "perflibs_amd_batch_optimize": (_I, [_P, C.c_int]),
    "perflibs_amd_batch_set_shape": (_I, [_P, _I, _P, _P]),
    "perflibs_amd_batch_exec": (_I, [_P, _D, _D]),
    "perflibs_amd_batch_set_sell": (_I, [_P, _I, _I]),
    "perflibs_amd_batch_sell_info": (_I, [_P, _P, _P, _P]),
    "perflibs_amd_batch_info": (_I, [_P, _P, _P, _P, _P, _P]),
    "perflibs_amd_probe_panel_gathers": (_I, [_P, _P, _P]),
    "perflibs_amd_probe_random_gathers": (_I, [_P, C.c_int64, C.c_int64, _P]),
    "perflibs_amd_last_error": (C.c_char_p, []),
}


def lib() -> C.CDLL:
    """Load libplsparse_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python {PKG / 'build.py'}`")
        # One HIP runtime per process: PyTorch ships its own libamdhip64 (same
        # SONAME, different file).  Load it first so this library binds to it
        # instead of pulling /opt/rocm's copy in beside it.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = C.CDLL(str(LIB_PATH))
        _check_one_hip_runtime()
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def _check_one_hip_runtime() -> None:
    """Two different libamdhip64 files mapped in one process (PyTorch's and
    /opt/rocm's) each keep their own device state, and mixing their pointers
    and streams crashes (a segmentation fault in the first call that passes a
    torch buffer to this library).  Refuse to continue in that state."""
    try:
        maps = Path("/proc/self/maps").read_text()
    except OSError:
        return
    files = {line.split()[-1] for line in maps.splitlines()
             if "libamdhip64" in line and line.split()[-1].startswith("/")}
    real = {os.path.realpath(f) for f in files}
    if len(real) > 1:
        raise RuntimeError("two HIP runtimes are loaded in this process (" + ", ".join(sorted(real)) +
                           "): import torch before loading libplsparse_amd, or do not import torch")


def check(fn: str, status: int) -> None:
    if status != STATUS_SUCCESS:
        raise PerflibsError(fn, status)


def _ptr(a) -> int | None:
    """Host numpy array -> its address; int -> a device address as given."""
    if a is None:
        return None
    if isinstance(a, (int, np.integer)):
        return int(a)
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):  # torch tensor
        return int(a.data_ptr())
    raise TypeError(f"unsupported buffer {type(a)}")


class SpMat:
    """Owning wrapper of a perflibs_spmat_t handle."""

    def __init__(self, handle: int, dtype, keep=()):
        self.h = handle
        self.dtype = np.dtype(dtype)
        self._keep = keep  # NOCOPY arrays must outlive the handle

    # ---- creation ----
    @staticmethod
    def _sfx(dtype) -> str:
        dt = np.dtype(dtype)
        if dt == np.float64:
            return "d"
        if dt == np.float32:
            return "s"
        raise TypeError(f"unsupported dtype {dt}")

    @classmethod
    def csr(cls, m, n, row_ptr, col_indx, vals, flags=0, raise_on_error=True):
        vals = np.ascontiguousarray(vals)
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(col_indx, dtype=np.int32)
        h = C.c_void_p()
        fn = f"perflibs_spmat_create_csr_{cls._sfx(vals.dtype)}"
        st = getattr(lib(), fn)(C.byref(h), m, n, _ptr(rp), _ptr(ci), _ptr(vals), flags)
        obj = cls(h.value, vals.dtype, keep=(rp, ci, vals))
        if raise_on_error:
            check(fn, st)
            return obj
        return obj, st

    @classmethod
    def csc(cls, m, n, row_indx, col_ptr, vals, flags=0, raise_on_error=True):
        vals = np.ascontiguousarray(vals)
        ri = np.ascontiguousarray(row_indx, dtype=np.int32)
        cp = np.ascontiguousarray(col_ptr, dtype=np.int32)
        h = C.c_void_p()
        fn = f"perflibs_spmat_create_csc_{cls._sfx(vals.dtype)}"
        st = getattr(lib(), fn)(C.byref(h), m, n, _ptr(ri), _ptr(cp), _ptr(vals), flags)
        obj = cls(h.value, vals.dtype)
        if raise_on_error:
            check(fn, st)
            return obj
        return obj, st

    @classmethod
    def coo(cls, m, n, row_indx, col_indx, vals, flags=0, raise_on_error=True):
        vals = np.ascontiguousarray(vals)
        ri = np.ascontiguousarray(row_indx, dtype=np.int32)
        ci = np.ascontiguousarray(col_indx, dtype=np.int32)
        h = C.c_void_p()
        fn = f"perflibs_spmat_create_coo_{cls._sfx(vals.dtype)}"
        st = getattr(lib(), fn)(C.byref(h), m, n, len(ri), _ptr(ri), _ptr(ci), _ptr(vals), flags)
        obj = cls(h.value, vals.dtype)
        if raise_on_error:
            check(fn, st)
            return obj
        return obj, st

    @classmethod
    def bsr(cls, layout, m, n, b, row_ptr, col_indx, vals, flags=0, raise_on_error=True):
        vals = np.ascontiguousarray(vals)
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(col_indx, dtype=np.int32)
        h = C.c_void_p()
        fn = f"perflibs_spmat_create_bsr_{cls._sfx(vals.dtype)}"
        st = getattr(lib(), fn)(C.byref(h), layout, m, n, b, _ptr(rp), _ptr(ci), _ptr(vals), flags)
        obj = cls(h.value, vals.dtype, keep=(rp, ci, vals))
        if raise_on_error:
            check(fn, st)
            return obj
        return obj, st

    @classmethod
    def dense(cls, layout, m, n, lda, vals, flags=0, raise_on_error=True):
        vals = np.ascontiguousarray(vals)
        h = C.c_void_p()
        fn = f"perflibs_spmat_create_dense_{cls._sfx(vals.dtype)}"
        st = getattr(lib(), fn)(C.byref(h), layout, m, n, lda, _ptr(vals), flags)
        obj = cls(h.value, vals.dtype, keep=(vals,))
        if raise_on_error:
            check(fn, st)
            return obj
        return obj, st

    @classmethod
    def null(cls, m, n, dtype=np.float64):
        return cls(lib().perflibs_spmat_create_null(m, n), dtype)

    @classmethod
    def identity(cls, n, dtype=np.float64):
        return cls(lib().perflibs_spmat_create_identity(n), dtype)

    def destroy(self):
        if self.h:
            lib().perflibs_spmat_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # ---- state ----
    def query(self):
        b, m, n, nnz = (C.c_int32() for _ in range(4))
        check("query", lib().perflibs_spmat_query(self.h, C.byref(b), C.byref(m), C.byref(n), C.byref(nnz)))
        return b.value, m.value, n.value, nnz.value

    def hint(self, kind, value) -> int:
        return lib().perflibs_spmat_hint(self.h, kind, value)

    def optimize(self):
        check("perflibs_spmv_optimize", lib().perflibs_spmv_optimize(self.h))

    def force_variant(self, v: int):
        check("force_variant", lib().perflibs_amd_force_variant(self.h, v))

    def force_sell(self, Cc: int, sigma: int):
        check("force_sell", lib().perflibs_amd_force_sell(self.h, Cc, sigma))

    def force_panel(self, panel_log2: int = 0, wg_per_cu: int = 0):
        check("force_panel", lib().perflibs_amd_force_panel(self.h, panel_log2, wg_per_cu))

    def spmm_info(self) -> dict:
        v = [C.c_int32(), C.c_int32(), C.c_int32()]
        tot = C.c_int64()
        check("spmm_info", lib().perflibs_amd_spmm_info(self.h, *[C.byref(a) for a in v], C.byref(tot)))
        return {"blocks_state": v[0].value, "nblocks": v[1].value, "list_max": v[2].value,
                "list_total": tot.value}

    def set_spmm_rows(self, rows: int):
        """0: the better of consecutive and clustered row blocks; 1: consecutive
        only; 2: clustered whenever the greedy completes (next SpMM, k >= 8)."""
        check("set_spmm_rows", lib().perflibs_amd_set_spmm_rows(self.h, rows))

    def spmm_rows_clustered(self) -> bool:
        v = C.c_int()
        check("spmm_rows_clustered", lib().perflibs_amd_spmm_rows_clustered(self.h, C.byref(v)))
        return bool(v.value)

    def panel_info(self) -> dict:
        v = [C.c_int32() for _ in range(5)]
        check("panel_info", lib().perflibs_amd_panel_info(self.h, *[C.byref(a) for a in v]))
        return dict(zip(("panel_log2", "npanels", "ntiles", "rows_per_tile_max", "grid"),
                        (a.value for a in v)))

    def select_mode(self, mode: int):
        check("set_select_mode", lib().perflibs_amd_set_select_mode(self.h, mode))

    def set_precision(self, p: int):
        check("set_precision", lib().perflibs_amd_set_precision(self.h, p))

    def variant(self) -> int:
        v = C.c_int()
        check("get_variant", lib().perflibs_amd_get_variant(self.h, C.byref(v)))
        return v.value

    def select_info(self) -> tuple[int, int]:
        """(selection path of the last optimize(), number of variants it timed)."""
        p, t = C.c_int(), C.c_int()
        check("select_info", lib().perflibs_amd_select_info(self.h, C.byref(p), C.byref(t)))
        return p.value, t.value

    # ---- execution ----
    def spmv(self, x, y, alpha=1.0, beta=0.0, trans=NOTRANS, raise_on_error=True) -> int:
        fn = f"perflibs_spmv_exec_{self._sfx(self.dtype)}"
        st = getattr(lib(), fn)(trans, alpha, self.h, _ptr(x), beta, _ptr(y))
        if raise_on_error:
            check(fn, st)
        return st

    def spmv_async(self, x_dev: int, y_dev: int, alpha=1.0, beta=0.0, trans=NOTRANS):
        fn = f"perflibs_amd_spmv_exec_async_{self._sfx(self.dtype)}"
        check(fn, getattr(lib(), fn)(trans, alpha, self.h, _ptr(x_dev), beta, _ptr(y_dev)))

    def spmm_async(self, k, B_dev, ldb, C_dev, ldc, alpha=1.0, beta=0.0):
        fn = f"perflibs_amd_spmm_exec_async_{self._sfx(self.dtype)}"
        check(fn, getattr(lib(), fn)(alpha, self.h, k, _ptr(B_dev), ldb, beta, _ptr(C_dev), ldc))

    def update(self, rows, cols, vals, raise_on_error=True) -> int:
        r = np.ascontiguousarray(rows, dtype=np.int32)
        c = np.ascontiguousarray(cols, dtype=np.int32)
        v = np.ascontiguousarray(vals, dtype=self.dtype)
        fn = f"perflibs_spmat_update_{self._sfx(self.dtype)}"
        st = getattr(lib(), fn)(self.h, len(r), _ptr(r), _ptr(c), _ptr(v))
        if raise_on_error:
            check(fn, st)
        return st

    # ---- in-place scaling, transposition, norms (host copies of the object) ----
    def scale(self, alpha, raise_on_error=True) -> int:
        fn = f"perflibs_spscale_exec_{self._sfx(self.dtype)}"
        st = getattr(lib(), fn)(alpha, self.h)
        if raise_on_error:
            check(fn, st)
        return st

    def transpose(self, trans=TRANS, raise_on_error=True) -> int:
        fn = f"perflibs_sptranspose_exec_{self._sfx(self.dtype)}"
        st = getattr(lib(), fn)(trans, self.h)
        if raise_on_error:
            check(fn, st)
        return st

    def norm(self, kind=NORM_FRB, raise_on_error=True):
        """The norm as a Python float, or (status, None) with raise_on_error=False on failure."""
        fn = f"perflibs_spnorm_exec_{self._sfx(self.dtype)}"
        out = (C.c_double if self.dtype == np.float64 else C.c_float)()
        st = getattr(lib(), fn)(self.h, kind, C.byref(out))
        if raise_on_error:
            check(fn, st)
            return float(out.value)
        return (st, float(out.value) if st == 0 else None)

    # ---- export ----
    def export_csr(self, index_base=0):
        m, n = C.c_int32(), C.c_int32()
        rp, ci, v = C.c_void_p(), C.c_void_p(), C.c_void_p()
        fn = f"perflibs_spmat_export_csr_{self._sfx(self.dtype)}"
        check(fn, getattr(lib(), fn)(self.h, index_base, C.byref(m), C.byref(n), C.byref(rp), C.byref(ci), C.byref(v)))
        libc = C.CDLL(None)
        mm = m.value
        rpa = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_int32)), (mm + 1,)).copy()
        nnz = int(rpa[-1] - rpa[0])
        cia = np.ctypeslib.as_array(C.cast(ci, C.POINTER(C.c_int32)), (max(nnz, 1),))[:nnz].copy()
        ctype = C.c_double if self.dtype == np.float64 else C.c_float
        va = np.ctypeslib.as_array(C.cast(v, C.POINTER(ctype)), (max(nnz, 1),))[:nnz].copy()
        for p in (rp, ci, v):
            libc.free(p)
        return mm, n.value, rpa, cia, va

    def export_csc(self, index_base=0):
        """(m, n, row_indx, col_ptr, vals) of the object's CSC form."""
        m, n = C.c_int32(), C.c_int32()
        ri, cp, v = C.c_void_p(), C.c_void_p(), C.c_void_p()
        fn = f"perflibs_spmat_export_csc_{self._sfx(self.dtype)}"
        check(fn, getattr(lib(), fn)(self.h, index_base, C.byref(m), C.byref(n), C.byref(ri), C.byref(cp), C.byref(v)))
        cpa = np.ctypeslib.as_array(C.cast(cp, C.POINTER(C.c_int32)), (n.value + 1,)).copy()
        nnz = int(cpa[-1] - cpa[0])
        ria = np.ctypeslib.as_array(C.cast(ri, C.POINTER(C.c_int32)), (max(nnz, 1),))[:nnz].copy() if nnz else np.zeros(0, np.int32)
        ctype = C.c_double if self.dtype == np.float64 else C.c_float
        va = np.ctypeslib.as_array(C.cast(v, C.POINTER(ctype)), (max(nnz, 1),))[:nnz].copy() if nnz else np.zeros(0, self.dtype)
        libc = C.CDLL(None)
        for p in (ri, cp, v):
            libc.free(p)
        return m.value, n.value, ria, cpa, va

    def export_coo(self):
        """(m, n, row_indx, col_indx, vals) of the object's COO form."""
        m, n, nnz = C.c_int32(), C.c_int32(), C.c_int32()
        ri, ci, v = C.c_void_p(), C.c_void_p(), C.c_void_p()
        fn = f"perflibs_spmat_export_coo_{self._sfx(self.dtype)}"
        check(fn, getattr(lib(), fn)(self.h, C.byref(m), C.byref(n), C.byref(nnz), C.byref(ri), C.byref(ci), C.byref(v)))
        k = nnz.value
        ctype = C.c_double if self.dtype == np.float64 else C.c_float
        out = []
        for p, ct, dt in ((ri, C.c_int32, np.int32), (ci, C.c_int32, np.int32), (v, ctype, self.dtype)):
            out.append(np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), (k,)).copy() if k else np.zeros(0, dt))
        libc = C.CDLL(None)
        for p in (ri, ci, v):
            libc.free(p)
        return (m.value, n.value, *out)

    def export_bsr(self, block_layout=ROW_MAJOR, index_base=0):
        """(m, n, block_size, row_ptr, col_indx, vals[nnzb * b * b])."""
        m, n, b = C.c_int32(), C.c_int32(), C.c_int32()
        rp, ci, v = C.c_void_p(), C.c_void_p(), C.c_void_p()
        fn = f"perflibs_spmat_export_bsr_{self._sfx(self.dtype)}"
        check(fn, getattr(lib(), fn)(self.h, block_layout, index_base, C.byref(m), C.byref(n),
                                     C.byref(b), C.byref(rp), C.byref(ci), C.byref(v)))
        libc = C.CDLL(None)
        bs = b.value
        mb = m.value // bs
        rpa = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_int32)), (mb + 1,)).copy()
        nnzb = int(rpa[-1] - rpa[0])
        cia = np.ctypeslib.as_array(C.cast(ci, C.POINTER(C.c_int32)), (max(nnzb, 1),))[:nnzb].copy()
        ctype = C.c_double if self.dtype == np.float64 else C.c_float
        nv = nnzb * bs * bs
        va = np.ctypeslib.as_array(C.cast(v, C.POINTER(ctype)), (max(nv, 1),))[:nv].copy()
        for p in (rp, ci, v):
            libc.free(p)
        return m.value, n.value, bs, rpa, cia, va

    def export_dense(self, layout=ROW_MAJOR):
        m, n = C.c_int32(), C.c_int32()
        v = C.c_void_p()
        fn = f"perflibs_spmat_export_dense_{self._sfx(self.dtype)}"
        check(fn, getattr(lib(), fn)(self.h, layout, C.byref(m), C.byref(n), C.byref(v)))
        ctype = C.c_double if self.dtype == np.float64 else C.c_float
        size = m.value * n.value
        arr = np.ctypeslib.as_array(C.cast(v, C.POINTER(ctype)), (max(size, 1),))[:size].copy()
        C.CDLL(None).free(v)
        if layout == ROW_MAJOR:
            return arr.reshape(m.value, n.value)
        return arr.reshape(n.value, m.value).T

    # ---- introspection ----
    def features(self) -> np.ndarray:
        f = np.zeros(NUM_FEATURES, dtype=np.float32)
        check("features", lib().perflibs_amd_features(self.h, _ptr(f)))
        return f

    def time_variants(self, warmup=3, reps=10) -> np.ndarray:
        ms = np.zeros(NUM_VARIANTS, dtype=np.float64)
        check("time_variants", lib().perflibs_amd_time_variants(self.h, warmup, reps, _ptr(ms)))
        return ms

    def sell_beta(self, Cc: int, sigma: int, nthreads: int = 0) -> tuple[float, float, float]:
        """(beta, min per-thread beta, max per-thread beta) of SELL-C-sigma."""
        b, lo, hi = C.c_double(), C.c_double(), C.c_double()
        check("sell_beta", lib().perflibs_amd_sell_beta(self.h, Cc, sigma, nthreads, C.byref(b),
                                                        C.byref(lo), C.byref(hi)))
        return b.value, lo.value, hi.value

    def sell_layout(self) -> dict:
        Cc, sig, cb, nf, nt = (C.c_int32() for _ in range(5))
        nC, padded = C.c_int64(), C.c_int64()
        check("sell_info", lib().perflibs_amd_sell_info(
            self.h, C.byref(Cc), C.byref(sig), C.byref(nC), C.byref(padded), C.byref(cb), C.byref(nf), C.byref(nt)))
        out = dict(C=Cc.value, sigma=sig.value, nC=nC.value, padded=padded.value,
                   col_indx_bytes=cb.value, n_full_chunks=nf.value, ntidyup=nt.value)
        ncs = nC.value
        cs = np.zeros(ncs, np.int64)
        cl = np.zeros(ncs, np.int64)
        perm = np.zeros(ncs * Cc.value, np.int64)
        cmin = np.zeros(ncs, np.int64)
        offs = np.zeros(max(padded.value, 1), np.int64)
        vals = np.zeros(max(padded.value, 1), np.float64)
        check("sell_export", lib().perflibs_amd_sell_export(
            self.h, _ptr(cs), _ptr(cl), _ptr(perm), _ptr(cmin), _ptr(offs), _ptr(vals)))
        out.update(cs=cs, cl=cl, row_permd2in=perm, col_indx_min=cmin,
                   offsets=offs[: padded.value], vals=vals[: padded.value])
        return out


def set_stream(stream_ptr: int) -> None:
    check("set_stream", lib().perflibs_amd_set_stream(stream_ptr))


def spmm_exec(transA, transB, alpha, A: SpMat, B: SpMat, beta, Cm: SpMat, raise_on_error=True) -> int:
    fn = f"perflibs_spmm_exec_{SpMat._sfx(A.dtype)}"
    st = getattr(lib(), fn)(transA, transB, alpha, A.h, B.h, beta, Cm.h)
    if raise_on_error:
        check(fn, st)
    return st


def spmm_blocks_inspect(m: int, n: int, row_ptr, col_indx, rows: int = 0) -> dict:
    """The row blocks an SpMM plan would cut for this CSR pattern (host only, no
    device): used, nblocks, list_max, list_total, clustered, row_order (rows
    block by block) and block_starts (offsets into row_order)."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(col_indx, dtype=np.int32)
    order = np.empty(max(m, 1), np.int32)
    starts = np.empty(max(m, 1) + 1, np.int32)
    used, nb, lm, cl = C.c_int(), C.c_int32(), C.c_int32(), C.c_int()
    tot = C.c_int64()
    check("spmm_blocks_inspect", lib().perflibs_amd_spmm_blocks_inspect(
        m, n, _ptr(rp), _ptr(ci), rows, C.byref(used), C.byref(nb), C.byref(lm), C.byref(tot),
        C.byref(cl), _ptr(order), _ptr(starts)))
    return {"used": bool(used.value), "nblocks": nb.value, "list_max": lm.value,
            "list_total": tot.value, "clustered": bool(cl.value), "row_order": order[:m],
            "block_starts": starts[: nb.value + 1]}


def spmm_optimize(transA, transB, alpha_hint, A: SpMat, B: SpMat, beta_hint, Cm: SpMat,
                  raise_on_error=True) -> int:
    st = lib().perflibs_spmm_optimize(transA, transB, alpha_hint, A.h, B.h, beta_hint, Cm.h)
    if raise_on_error:
        check("perflibs_spmm_optimize", st)
    return st


class Batch:
    """A batch of CSR products Y_p = alpha*A_p*X_p + beta*Y_p run as one
    persistent launch replayed from a hipGraph (perflibs_amd_batch_*)."""

    K_CLASSES = (1, 2, 4, 8, 16)

    def __init__(self, mats, ks, precision=VALUES_NATIVE):
        self.mats = list(mats)  # keep the handles alive while the batch exists
        n = len(self.mats)
        hs = (C.c_void_p * max(n, 1))(*[m.h for m in self.mats])
        kk = np.ascontiguousarray(ks, dtype=np.int32)
        out = C.c_void_p()
        check("batch_create", lib().perflibs_amd_batch_create(C.byref(out), n, hs, kk.ctypes.data, precision))
        self.h = out.value
        self._ops = None

    def set_operands(self, X_ptrs, Y_ptrs):
        n = len(self.mats)
        xs = (C.c_void_p * max(n, 1))(*[int(p) for p in X_ptrs])
        ys = (C.c_void_p * max(n, 1))(*[int(p) for p in Y_ptrs])
        check("batch_set_operands", lib().perflibs_amd_batch_set_operands(self.h, xs, ys))
        self._ops = (xs, ys)

    def optimize(self, autotune=True):
        check("batch_optimize", lib().perflibs_amd_batch_optimize(self.h, int(bool(autotune))))

    def set_shape(self, tile_nnz, lanes=None, cpl=None):
        la = None if lanes is None else np.ascontiguousarray(lanes, dtype=np.int32)
        ca = None if cpl is None else np.ascontiguousarray(cpl, dtype=np.int32)
        check("batch_set_shape", lib().perflibs_amd_batch_set_shape(
            self.h, tile_nnz, None if la is None else la.ctypes.data, None if ca is None else ca.ctypes.data))

    def exec(self, alpha=1.0, beta=0.0):
        check("batch_exec", lib().perflibs_amd_batch_exec(self.h, alpha, beta))

    def set_sell(self, Cc: int, sigma: int):
        """Run the k = 1 problems as SELL-C-sigma tiles (Cc = 0: CSR row tiles)."""
        check("batch_set_sell", lib().perflibs_amd_batch_set_sell(self.h, Cc, sigma))

    def sell_info(self) -> dict:
        cc, sg, pad = C.c_int32(), C.c_int32(), C.c_int64()
        check("batch_sell_info", lib().perflibs_amd_batch_sell_info(self.h, C.byref(cc), C.byref(sg), C.byref(pad)))
        return {"C": cc.value, "sigma": sg.value, "padded": pad.value}

    def info(self) -> dict:
        nt = C.c_int64()
        tn = C.c_int32()
        gr = C.c_int32()
        la = np.zeros(5, np.int32)
        ca = np.zeros(5, np.int32)
        check("batch_info", lib().perflibs_amd_batch_info(self.h, C.byref(nt), C.byref(tn), C.byref(gr),
                                                          la.ctypes.data, ca.ctypes.data))
        return {"ntiles": nt.value, "tile_nnz": tn.value, "grid": gr.value,
                "lanes": dict(zip(self.K_CLASSES, la.tolist())), "cpl": dict(zip(self.K_CLASSES, ca.tolist()))}

    def destroy(self):
        if getattr(self, "h", None):
            lib().perflibs_amd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
