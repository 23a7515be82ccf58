"""ctypes front end of the CPU oracle (finch_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of bench.py, as the checker / CPU baseline.
The product library never imports it.

Each wrapper takes numpy arrays in Finch's compressed layout (``ptr``/``idx``
with ``base`` 1 or 0, values) and returns numpy results.  The C functions
restate Finch.jl's generated loops; see the citations in finch_oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_INT = ctypes.c_int


def build() -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = ctypes.CDLL(_SO)
    return _lib


def _p(a):
    return a.ctypes.data_as(_P)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


_SFX = {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32", np.dtype(np.int32): "i32"}


def _sfx(dt):
    try:
        return _SFX[np.dtype(dt)]
    except KeyError as e:
        raise TypeError(f"oracle has no kernel for {dt}") from e


def max_threads() -> int:
    return int(lib().or_max_threads())


def spmv_csc(n_inner, ptr, idx, val, x, base=1):
    """y = A*x, A column-compressed (typical_spmv_csc.txt:26-83)."""
    val = np.ascontiguousarray(val)
    s = _sfx(val.dtype)
    x = np.ascontiguousarray(x, dtype=val.dtype)
    ptr, idx = _i64(ptr), _i64(idx)
    y = np.empty(n_inner, dtype=val.dtype)
    f = getattr(lib(), f"or_spmv_csc_{s}")
    f(_I64(n_inner), _I64(len(ptr) - 1), _p(ptr), _p(idx), _p(val), _INT(base), _p(x), _p(y))
    return y


def spmv_csc_atomic(n_inner, ptr, idx, val, x, base=1, nthreads=1):
    """y = A*x, A column-compressed, columns split over threads with atomic
    adds into y (benchmark/benchmarks.jl:402-413, AtomicElement)."""
    val = np.ascontiguousarray(val)
    s = _sfx(val.dtype)
    x = np.ascontiguousarray(x, dtype=val.dtype)
    ptr, idx = _i64(ptr), _i64(idx)
    y = np.empty(n_inner, dtype=val.dtype)
    f = getattr(lib(), f"or_spmv_csc_atomic_{s}")
    f(_I64(n_inner), _I64(len(ptr) - 1), _p(ptr), _p(idx), _p(val), _INT(base), _p(x), _p(y),
      _INT(nthreads))
    return y


def spmv_csc_mutex(n_inner, ptr, idx, val, x, base=1, nthreads=1):
    """y = A*x, A column-compressed, columns split over threads, each update
    of y[i] under y[i]'s lock (benchmark/benchmarks.jl:414-426, Mutex)."""
    val = np.ascontiguousarray(val)
    s = _sfx(val.dtype)
    x = np.ascontiguousarray(x, dtype=val.dtype)
    ptr, idx = _i64(ptr), _i64(idx)
    y = np.empty(n_inner, dtype=val.dtype)
    f = getattr(lib(), f"or_spmv_csc_mutex_{s}")
    f(_I64(n_inner), _I64(len(ptr) - 1), _p(ptr), _p(idx), _p(val), _INT(base), _p(x), _p(y),
      _INT(nthreads))
    return y


def spmv_csr_greedy(ptr, idx, val, x, chunk, base=1, nthreads=1):
    """CSR SpMV with Finch's greedy schedule (src/architecture.jl:721-767)."""
    val = np.ascontiguousarray(val)
    s = _sfx(val.dtype)
    x = np.ascontiguousarray(x, dtype=val.dtype)
    ptr, idx = _i64(ptr), _i64(idx)
    m = len(ptr) - 1
    y = np.empty(m, dtype=val.dtype)
    f = getattr(lib(), f"or_spmv_csr_greedy_{s}")
    f(_I64(m), _p(ptr), _p(idx), _p(val), _INT(base), _p(x), _p(y), _I64(chunk), _INT(nthreads))
    return y


def spmm_csr_greedy(ptr, idx, val, B, chunk, base=1, nthreads=1):
    """CSR SpMM with Finch's greedy schedule (src/architecture.jl:721-767)."""
    val = np.ascontiguousarray(val)
    s = _sfx(val.dtype)
    B = np.ascontiguousarray(B, dtype=val.dtype)
    ptr, idx = _i64(ptr), _i64(idx)
    m = len(ptr) - 1
    k = B.shape[1]
    C = np.empty((m, k), dtype=val.dtype)
    f = getattr(lib(), f"or_spmm_csr_greedy_{s}")
    f(_I64(m), _I64(k), _p(ptr), _p(idx), _p(val), _INT(base), _p(B), _p(C), _I64(chunk),
      _INT(nthreads))
    return C


def spmv_csr(ptr, idx, val, x, base=1, nthreads=1):
    """y = A*x, A row-compressed (parallel_spmv.txt:36-76)."""
    val = np.ascontiguousarray(val)
    s = _sfx(val.dtype)
    x = np.ascontiguousarray(x, dtype=val.dtype)
    ptr, idx = _i64(ptr), _i64(idx)
    m = len(ptr) - 1
    y = np.empty(m, dtype=val.dtype)
    f = getattr(lib(), f"or_spmv_csr_{s}")
    f(_I64(m), _I64(len(x)), _p(ptr), _p(idx), _p(val), _INT(base), _p(x), _p(y), _INT(nthreads))
    return y


def spmm_csr(ptr, idx, val, B, base=1, nthreads=1):
    """C = A*B, A row-compressed, B row-major (n_inner x k)."""
    val = np.ascontiguousarray(val)
    s = _sfx(val.dtype)
    B = np.ascontiguousarray(B, dtype=val.dtype)
    ptr, idx = _i64(ptr), _i64(idx)
    m = len(ptr) - 1
    k = B.shape[1]
    C = np.empty((m, k), dtype=val.dtype)
    f = getattr(lib(), f"or_spmm_csr_{s}")
    f(_I64(m), _I64(B.shape[0]), _I64(k), _p(ptr), _p(idx), _p(val), _INT(base), _p(B), _p(C),
      _INT(nthreads))
    return C


def sddmm_csr(ptr, idx, val, X, Y, base=1):
    """out[q] = A[q] * X[i,:].Y[j,:] on pattern(A)."""
    val = np.ascontiguousarray(val)
    s = _sfx(val.dtype)
    X = np.ascontiguousarray(X, dtype=val.dtype)
    Y = np.ascontiguousarray(Y, dtype=val.dtype)
    ptr, idx = _i64(ptr), _i64(idx)
    m = len(ptr) - 1
    out = np.empty(len(val), dtype=val.dtype)
    f = getattr(lib(), f"or_sddmm_csr_{s}")
    f(_I64(m), _I64(Y.shape[0]), _I64(X.shape[1]), _p(ptr), _p(idx), _p(val), _INT(base),
      _p(X), _p(Y), _p(out))
    return out


EWISE_OPS = {"mul": 0, "add": 1, "max": 2}


def ewise(op, aptr, aidx, aval, bptr, bidx, bval, base=1):
    """Fiber-wise two-finger merge (typical_merge_twofinger.txt:45-186)."""
    opc = EWISE_OPS[op] if isinstance(op, str) else int(op)
    aval = np.ascontiguousarray(aval)
    s = _sfx(aval.dtype)
    bval = np.ascontiguousarray(bval, dtype=aval.dtype)
    aptr, aidx, bptr, bidx = _i64(aptr), _i64(aidx), _i64(bptr), _i64(bidx)
    n_outer = len(aptr) - 1
    cap = max(1, len(aidx) + len(bidx))
    cptr = np.empty(n_outer + 1, dtype=np.int64)
    cidx = np.empty(cap, dtype=np.int64)
    cval = np.empty(cap, dtype=aval.dtype)
    f = getattr(lib(), f"or_ewise_{s}")
    f.restype = _I64
    nnz = f(_INT(opc), _I64(n_outer), _p(aptr), _p(aidx), _p(aval), _p(bptr), _p(bidx),
            _p(bval), _INT(base), _p(cptr), _p(cidx), _p(cval))
    return cptr, cidx[:nnz].copy(), cval[:nnz].copy()


SEMIRINGS = {"plus_times": 0, "min_plus": 1}


def spgemm(semiring, m, n, aptr, aidx, aval, bptr, bidx, bval, base=1):
    """Gustavson C = A*B row by row (kernels/gustavsons.jl:52-182)."""
    sr = SEMIRINGS[semiring] if isinstance(semiring, str) else int(semiring)
    aval = np.ascontiguousarray(aval)
    bval = np.ascontiguousarray(bval, dtype=aval.dtype)
    aptr, aidx, bptr, bidx = _i64(aptr), _i64(aidx), _i64(bptr), _i64(bidx)
    L = lib()
    cptr = np.empty(m + 1, dtype=np.int64)
    L.or_spgemm_symbolic.restype = _I64
    nnz = L.or_spgemm_symbolic(_I64(m), _I64(n), _p(aptr), _p(aidx), _p(bptr), _p(bidx),
                               _INT(base), _p(cptr))
    cidx = np.empty(max(1, nnz), dtype=np.int64)
    cval = np.empty(max(1, nnz), dtype=aval.dtype)
    s = _sfx(aval.dtype)
    if s == "i32":
        if sr != 0:
            raise TypeError("min-plus over Int32 is not a configured semiring")
        L.or_spgemm_i32(_I64(m), _I64(n), _p(aptr), _p(aidx), _p(aval), _p(bptr), _p(bidx),
                        _p(bval), _INT(base), _p(cptr), _p(cidx), _p(cval))
    else:
        getattr(L, f"or_spgemm_{s}")(_INT(sr), _I64(m), _I64(n), _p(aptr), _p(aidx), _p(aval),
                                     _p(bptr), _p(bidx), _p(bval), _INT(base), _p(cptr),
                                     _p(cidx), _p(cval))
    return cptr, cidx[:nnz].copy(), cval[:nnz].copy()


def transpose(n_outer, n_inner, ptr, idx, val, base=1):
    """Counting transpose: permutedims of a 2-D Dense(SparseList) (copy.jl:64-71)."""
    val = np.ascontiguousarray(val)
    ptr, idx = _i64(ptr), _i64(idx)
    nnz = int(ptr[n_outer] - ptr[0])
    tptr = np.empty(n_inner + 1, dtype=np.int64)
    tidx = np.empty(max(1, nnz), dtype=np.int64)
    tval = np.empty(max(1, nnz), dtype=val.dtype)
    lib().or_transpose(_I64(n_outer), _I64(n_inner), _p(ptr), _p(idx), _p(val),
                       _INT(val.dtype.itemsize), _INT(base), _p(tptr), _p(tidx), _p(tval))
    return tptr, tidx[:nnz].copy(), tval[:nnz].copy()


def coo_to_compressed(n_outer, n_inner, outer, inner, val, base=1):
    """fsparse then Dense(SparseList): (outer, inner)-sorted, duplicates summed
    left to right in input order (fsparse.jl:64-97)."""
    val = np.ascontiguousarray(val)
    s = _sfx(val.dtype)
    outer, inner = _i64(outer), _i64(inner)
    nnz = len(val)
    cptr = np.empty(n_outer + 1, dtype=np.int64)
    cidx = np.empty(max(1, nnz), dtype=np.int64)
    cval = np.empty(max(1, nnz), dtype=val.dtype)
    f = getattr(lib(), f"or_coo_{s}")
    f.restype = _I64
    u = f(_I64(n_outer), _I64(n_inner), _I64(nnz), _p(outer), _p(inner), _p(val), _INT(base),
          _p(cptr), _p(cidx), _p(cval))
    return cptr, cidx[:u].copy(), cval[:u].copy()


def dropfills(n_outer, ptr, idx, val, fill=0.0, base=1):
    """Keep the entries that are not isequal to `fill` (copy.jl:73-117)."""
    val = np.ascontiguousarray(val)
    s = _sfx(val.dtype)
    ptr, idx = _i64(ptr), _i64(idx)
    nnz = int(ptr[n_outer] - ptr[0])
    cptr = np.empty(n_outer + 1, dtype=np.int64)
    cidx = np.empty(max(1, nnz), dtype=np.int64)
    cval = np.empty(max(1, nnz), dtype=val.dtype)
    f = getattr(lib(), f"or_dropfills_{s}")
    f.restype = _I64
    ctype = {"f64": ctypes.c_double, "f32": ctypes.c_float, "i32": ctypes.c_int32}[s]
    u = f(_I64(n_outer), _p(ptr), _p(idx), _p(val), _INT(base), ctype(fill), _p(cptr),
          _p(cidx), _p(cval))
    return cptr, cidx[:u].copy(), cval[:u].copy()
