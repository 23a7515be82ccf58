"""ctypes binding of libfinchgpu.so (the C ABI declared in include/finchgpu.h).

The library is loaded from this package directory, where
``finch.jl_amd/build.py`` writes it.  There is no fallback: if the library is
missing or fails to load, every operation raises ``FinchGPUUnavailable``.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# FINCHGPU_LIB selects another build of the library (e.g. a tile-shape variant).
LIB_PATH = os.environ.get("FINCHGPU_LIB") or os.path.join(_HERE, "libfinchgpu.so")

# fgx_status
OK, EDIM, EPROTO, EUNSUP, EOOM, EADMIT, EHIP, EINVAL = range(8)
# fgx_dtype
F64, F32, BF16, F16, I32, I64 = range(6)
ROW_MAJOR, COL_MAJOR = 0, 1
OP_MUL, OP_ADD, OP_MAX = 0, 1, 2
PLUS_TIMES, MIN_PLUS = 0, 1
REQ_SPMV, REQ_SPMM, REQ_SDDMM = 0, 1, 2
BATCH_FORCE_ROWSPLIT, BATCH_FORCE_NNZSPLIT, BATCH_NO_GRAPH, BATCH_BF16_INPUTS = 1, 2, 4, 8
BATCH_MEASURE = 16
NOALLOC_SCRATCH_BYTES = 16 << 20     # FGX_NOALLOC_SCRATCH_BYTES


class FinchGPUError(RuntimeError):
    """A non-zero fgx_status."""

    def __init__(self, code, msg):
        super().__init__(f"[fgx status {code}] {msg}")
        self.code = code


class DimensionMismatch(FinchGPUError):
    """FGX_EDIM — Finch's DimensionMismatch."""


class FinchProtocolError(FinchGPUError):
    """FGX_EPROTO — Finch's FinchProtocolError (src/Finch.jl:70-72)."""


class UnsupportedError(FinchGPUError):
    """FGX_EUNSUP — format, dtype or program the backend does not handle."""


class AdmissionError(FinchGPUError):
    """FGX_EADMIT — serving admission rejected."""


class FinchGPUUnavailable(RuntimeError):
    """libfinchgpu.so is not built or cannot be loaded."""


_ERRORS = {EDIM: DimensionMismatch, EPROTO: FinchProtocolError, EUNSUP: UnsupportedError,
           EADMIT: AdmissionError}


class Compressed(ctypes.Structure):
    """struct fgx_compressed."""

    _fields_ = [
        ("n_outer", ctypes.c_int64),
        ("n_inner", ctypes.c_int64),
        ("nnz", ctypes.c_int64),
        ("ptr_type", ctypes.c_int32),
        ("idx_type", ctypes.c_int32),
        ("val_type", ctypes.c_int32),
        ("index_base", ctypes.c_int32),
        ("ptr", ctypes.c_void_p),
        ("idx", ctypes.c_void_p),
        ("val", ctypes.c_void_p),
    ]


class BatchItem(ctypes.Structure):
    """struct fgx_batch_item."""

    _fields_ = [("A", Compressed), ("n_rhs", ctypes.c_int64), ("B", ctypes.c_void_p),
                ("C", ctypes.c_void_p)]


class Request(ctypes.Structure):
    """struct fgx_request."""

    _fields_ = [("kind", ctypes.c_int32), ("operand", ctypes.c_int32), ("k", ctypes.c_int64),
                ("arrival_s", ctypes.c_double), ("in_", ctypes.c_void_p),
                ("in2", ctypes.c_void_p), ("out", ctypes.c_void_p)]


class ServeConfig(ctypes.Structure):
    """struct fgx_serve_config."""

    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("max_slots", ctypes.c_int32),
                ("max_batch", ctypes.c_int32), ("mode", ctypes.c_int32),
                ("input_type", ctypes.c_int32)]


class RequestResult(ctypes.Structure):
    """struct fgx_request_result."""

    _fields_ = [("admit_s", ctypes.c_double), ("done_s", ctypes.c_double),
                ("status", ctypes.c_int32), ("batch", ctypes.c_int32)]


_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_PC = ctypes.POINTER(Compressed)

# name -> (restype, argtypes); every symbol of include/finchgpu.h.
SIGNATURES = {
    "fgx_version": (_I, []),
    "fgx_ctx_create": (_I, [_I, ctypes.POINTER(_P)]),
    "fgx_ctx_destroy": (_I, [_P]),
    "fgx_ctx_stream": (_P, [_P]),
    "fgx_sync": (_I, [_P, _P]),
    "fgx_last_error": (ctypes.c_char_p, []),
    "fgx_plan_cache_clear": (_I, [_P]),
    "fgx_validate": (_I, [_P, _PC, _P]),
    "fgx_ctx_set_safe": (_I, [_P, _I]),
    "fgx_sequence": (_I, [_P, _I, _P, _I64, _I64, _P]),
    "fgx_spmv_csr": (_I, [_P, _PC, _P, _P, _P]),
    "fgx_spmv_csr_noalloc": (_I, [_P, _PC, _P, _P, _P]),
    "fgx_spmv_csc": (_I, [_P, _PC, _P, _P, _P]),
    "fgx_spmv_csc_ordered": (_I, [_P, _PC, _P, _P, _P]),
    "fgx_spmm_csr": (_I, [_P, _PC, _I64, _P, _I64, _I, _P, _I64, _I, _P]),
    "fgx_sddmm_csr": (_I, [_P, _PC, _I64, _P, _I64, _I, _P, _I64, _I, _P, _P]),
    "fgx_ewise_symbolic": (_I, [_P, _I, _PC, _PC, _P, ctypes.POINTER(_I64), _P]),
    "fgx_ewise_numeric": (_I, [_P, _I, _PC, _PC, _P, _P, _P, _P]),
    "fgx_ewise_async": (_I, [_P, _I, _PC, _PC, _P, _P, _P, _I64, _P]),
    "fgx_spgemm_symbolic": (_I, [_P, _I, _PC, _PC, _P, ctypes.POINTER(_I64), _P]),
    "fgx_spgemm_numeric": (_I, [_P, _I, _PC, _PC, _P, _P, _P, _P]),
    "fgx_convert": (_I, [_P, _P, _I, _P, _I, _I64, _P]),
    "fgx_transpose": (_I, [_P, _PC, _P, _P, _P, _P]),
    "fgx_coo_to_compressed": (_I, [_P, _I64, _I64, _I64, _I, _I, _I, _P, _P, _P, _P, _P, _P,
                                   ctypes.POINTER(_I64), _P]),
    "fgx_dropfills": (_I, [_P, _PC, ctypes.c_double, _P, _P, _P, ctypes.POINTER(_I64), _P]),
    "fgx_batch_plan_create": (_I, [_P, ctypes.POINTER(BatchItem), _I64, _I, ctypes.POINTER(_P)]),
    "fgx_batch_plan_run": (_I, [_P, _P, _P]),
    "fgx_batch_plan_stats": (_I, [_P, ctypes.POINTER(_I64)]),
    "fgx_batch_plan_destroy": (_I, [_P]),
    "fgx_batch_run_loop": (_I, [_P, ctypes.POINTER(BatchItem), _I64, _P]),
    "fgx_serve_run": (_I, [_P, _PC, _I64, ctypes.POINTER(Request), _I64,
                           ctypes.POINTER(ServeConfig), ctypes.POINTER(RequestResult)]),
}

_lib = None


def load():
    """The loaded library (raises FinchGPUUnavailable if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FinchGPUUnavailable(
            f"{LIB_PATH} is not built; run `python finch.jl_amd/build.py` first")
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        raise FinchGPUUnavailable(f"cannot load {LIB_PATH}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(code: int, what: str = ""):
    """Raises the exception matching a non-zero fgx_status."""
    if code == OK:
        return
    msg = load().fgx_last_error().decode(errors="replace")
    raise _ERRORS.get(code, FinchGPUError)(code, f"{what}: {msg}" if what else msg)
