"""Device operations over the C ABI: the hot loops of Finch's sparse kernels.

Operands are ``Compressed`` device operands (``ptr``/``idx``/``val`` torch
tensors on a HIP device, Finch's 1-based layout by default) and dense torch
tensors.  Every call goes through libfinchgpu.so on the caller's current torch
stream; nothing here computes on the host.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _lib as L

_TORCH_DT = {
    torch.float64: L.F64, torch.float32: L.F32, torch.bfloat16: L.BF16, torch.float16: L.F16,
    torch.int32: L.I32, torch.int64: L.I64,
}
_FROM_DT = {v: k for k, v in _TORCH_DT.items()}


def fgx_dtype(dt: torch.dtype) -> int:
    try:
        return _TORCH_DT[dt]
    except KeyError as e:
        raise L.UnsupportedError(L.EUNSUP, f"dtype {dt} has no fgx_dtype") from e


def acc_dtype(val_dtype: torch.dtype) -> torch.dtype:
    """Dense/accumulation dtype paired with a value dtype (FP64 stays FP64)."""
    return torch.float64 if val_dtype == torch.float64 else torch.float32


class Context:
    """An fgx_ctx on one HIP device (one per process and device)."""

    def __init__(self, device: int = 0):
        lib = L.load()
        h = ctypes.c_void_p()
        L.check(lib.fgx_ctx_create(int(device), ctypes.byref(h)), "fgx_ctx_create")
        self.handle = h
        self.device = int(device)

    def close(self):
        if self.handle:
            L.load().fgx_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts: dict[int, Context] = {}


def context(device: int | torch.device | None = None) -> Context:
    if device is None:
        device = torch.cuda.current_device()
    if isinstance(device, torch.device):
        device = device.index if device.index is not None else torch.cuda.current_device()
    ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        _contexts[device] = ctx
    return ctx


def set_safe_mode(safe: bool = True, device: int | torch.device | None = None):
    """Finch's mode=:safe for this device's context: every entry point checks
    each compressed operand (fgx_validate) before a kernel reads it."""
    L.check(L.load().fgx_ctx_set_safe(context(device).handle, int(bool(safe))),
            "fgx_ctx_set_safe")


def _stream(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t: torch.Tensor | None) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr() if t is not None and t.numel() > 0 else 0)


@dataclass
class Compressed:
    """Device buffers of Dense(SparseList(Element)): n_outer fibers over 1:n_inner.

    Read row-wise (fibers = rows) it is CSR; read column-wise it is CSC.
    """

    ptr: torch.Tensor
    idx: torch.Tensor
    val: torch.Tensor
    n_outer: int
    n_inner: int
    base: int = 1

    @property
    def nnz(self) -> int:
        return int(self.idx.numel())

    @property
    def device(self):
        return self.ptr.device

    def desc(self) -> L.Compressed:
        for name in ("ptr", "idx", "val"):
            t = getattr(self, name)
            if not t.is_cuda or not t.is_contiguous():
                raise L.UnsupportedError(L.EUNSUP, f"{name} must be a contiguous device tensor")
        if self.ptr.numel() != self.n_outer + 1:
            raise L.DimensionMismatch(L.EDIM, f"ptr has {self.ptr.numel()} entries, "
                                              f"expected {self.n_outer + 1}")
        if self.val.numel() != self.idx.numel():
            raise L.DimensionMismatch(L.EDIM, "idx and val lengths differ")
        return L.Compressed(self.n_outer, self.n_inner, self.nnz, fgx_dtype(self.ptr.dtype),
                            fgx_dtype(self.idx.dtype), fgx_dtype(self.val.dtype), self.base,
                            _ptr(self.ptr), _ptr(self.idx), _ptr(self.val))

    def with_values(self, val: torch.Tensor) -> "Compressed":
        return Compressed(self.ptr, self.idx, val, self.n_outer, self.n_inner, self.base)

    def to_values(self, dtype: torch.dtype) -> "Compressed":
        """Same pattern with values converted on the device (e.g. BF16 storage)."""
        return self.with_values(convert(self.val, dtype))


def sequence(n: int, dtype: torch.dtype, device, start: int = 1) -> torch.Tensor:
    """start, start+1, ..., start+n-1 on the device (fgx_sequence): the
    positions of a frozen SparseDict filled in coordinate order."""
    out = torch.empty(n, dtype=dtype, device=device)
    L.check(L.load().fgx_sequence(context(out.device).handle, fgx_dtype(dtype), _ptr(out), n,
                                  start, _stream(out.device)), "fgx_sequence")
    return out


def validate(A: Compressed) -> None:
    """Checks A against the SparseList invariants on the device (fgx_validate):
    raises FinchProtocolError for a malformed ptr or unsorted fibre and
    DimensionMismatch for coordinates outside the shape or nnz != ptr[end]-base."""
    L.check(L.load().fgx_validate(context(A.device).handle, ctypes.byref(A.desc()),
                                  _stream(A.device)), "fgx_validate")


def _check_dense(t: torch.Tensor, dt: torch.dtype, name: str):
    if not t.is_cuda:
        raise L.UnsupportedError(L.EUNSUP, f"{name} must be a device tensor")
    if t.dtype != dt:
        raise L.UnsupportedError(L.EUNSUP, f"{name} has dtype {t.dtype}, expected {dt}")


def convert(src: torch.Tensor, dtype: torch.dtype) -> torch.Tensor:
    out = torch.empty(src.shape, dtype=dtype, device=src.device)
    L.check(L.load().fgx_convert(context(src.device).handle, _ptr(src), fgx_dtype(src.dtype),
                                 _ptr(out), fgx_dtype(dtype), src.numel(), _stream(src.device)),
            "fgx_convert")
    return out


def spmv_csr(A: Compressed, x: torch.Tensor, y: torch.Tensor | None = None) -> torch.Tensor:
    """y = A*x with A read row-wise (fgx_spmv_csr)."""
    dt = acc_dtype(A.val.dtype)
    _check_dense(x, dt, "x")
    if x.numel() != A.n_inner:
        raise L.DimensionMismatch(L.EDIM, f"x has {x.numel()} entries, A has {A.n_inner} columns")
    if y is None:
        y = torch.empty(A.n_outer, dtype=dt, device=A.device)
    L.check(L.load().fgx_spmv_csr(context(A.device).handle, ctypes.byref(A.desc()), _ptr(x),
                                  _ptr(y), _stream(A.device)), "fgx_spmv_csr")
    return y


def spmv_csr_noalloc(A: Compressed, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """y = A*x like spmv_csr, through the entry point that allocates nothing
    (fgx_spmv_csr_noalloc): the caller passes y, the tile table and carries
    live in the library's fixed per-GPU scratch array."""
    dt = acc_dtype(A.val.dtype)
    _check_dense(x, dt, "x")
    _check_dense(y, dt, "y")
    if x.numel() != A.n_inner or y.numel() != A.n_outer:
        raise L.DimensionMismatch(L.EDIM, f"x has {x.numel()} / y has {y.numel()} entries for "
                                          f"A of {A.n_outer} x {A.n_inner}")
    L.check(L.load().fgx_spmv_csr_noalloc(context(A.device).handle, ctypes.byref(A.desc()),
                                          _ptr(x), _ptr(y), _stream(A.device)),
            "fgx_spmv_csr_noalloc")
    return y


def spmv_csc(A: Compressed, x: torch.Tensor, y: torch.Tensor | None = None) -> torch.Tensor:
    """y = A*x with A read column-wise (fgx_spmv_csc)."""
    dt = acc_dtype(A.val.dtype)
    _check_dense(x, dt, "x")
    if x.numel() != A.n_outer:
        raise L.DimensionMismatch(L.EDIM, f"x has {x.numel()} entries, A has {A.n_outer} columns")
    if y is None:
        y = torch.empty(A.n_inner, dtype=dt, device=A.device)
    L.check(L.load().fgx_spmv_csc(context(A.device).handle, ctypes.byref(A.desc()), _ptr(x),
                                  _ptr(y), _stream(A.device)), "fgx_spmv_csc")
    return y


def spmv_csc_ordered(A: Compressed, x: torch.Tensor, y: torch.Tensor | None = None) -> torch.Tensor:
    """y = A*x with A read column-wise, run-to-run deterministic
    (fgx_spmv_csc_ordered: device transpose, then row sums)."""
    dt = acc_dtype(A.val.dtype)
    _check_dense(x, dt, "x")
    if x.numel() != A.n_outer:
        raise L.DimensionMismatch(L.EDIM, f"x has {x.numel()} entries, A has {A.n_outer} columns")
    if y is None:
        y = torch.empty(A.n_inner, dtype=dt, device=A.device)
    L.check(L.load().fgx_spmv_csc_ordered(context(A.device).handle, ctypes.byref(A.desc()),
                                          _ptr(x), _ptr(y), _stream(A.device)),
            "fgx_spmv_csc_ordered")
    return y


def _layout(t: torch.Tensor):
    """(layout, ld) of a 2-D tensor: row-major or column-major strides."""
    r, c = t.shape
    sr, sc = t.stride()
    if sc == 1 and (sr >= c or r <= 1):
        return L.ROW_MAJOR, max(sr, c)
    if sr == 1 and (sc >= r or c <= 1):
        return L.COL_MAJOR, max(sc, r)
    raise L.UnsupportedError(L.EUNSUP, f"dense operand strides {t.stride()} unsupported")


def spmm_csr(A: Compressed, B: torch.Tensor, C: torch.Tensor | None = None) -> torch.Tensor:
    """C = A*B, A CSR, B dense (n_inner x k) row- or column-major (fgx_spmm_csr)."""
    dt = acc_dtype(A.val.dtype)
    _check_dense(B, dt, "B")
    if B.dim() != 2 or B.shape[0] != A.n_inner:
        raise L.DimensionMismatch(L.EDIM, f"B has shape {tuple(B.shape)}, A has {A.n_inner} columns")
    k = B.shape[1]
    if C is None:
        C = torch.empty((A.n_outer, k), dtype=dt, device=A.device)
    bl, ldb = _layout(B)
    cl, ldc = _layout(C)
    L.check(L.load().fgx_spmm_csr(context(A.device).handle, ctypes.byref(A.desc()), k, _ptr(B),
                                  ldb, bl, _ptr(C), ldc, cl, _stream(A.device)), "fgx_spmm_csr")
    return C


def sddmm_csr(A: Compressed, X: torch.Tensor, Y: torch.Tensor,
              out: torch.Tensor | None = None) -> torch.Tensor:
    """out[q] = A[q] * X[i,:].Y[j,:] on pattern(A) (fgx_sddmm_csr)."""
    dt = acc_dtype(A.val.dtype)
    _check_dense(X, dt, "X")
    _check_dense(Y, dt, "Y")
    if X.shape[0] != A.n_outer or Y.shape[0] != A.n_inner or X.shape[1] != Y.shape[1]:
        raise L.DimensionMismatch(L.EDIM, f"X {tuple(X.shape)} / Y {tuple(Y.shape)} do not match "
                                          f"A ({A.n_outer} x {A.n_inner})")
    if out is None:
        out = torch.empty(A.nnz, dtype=dt, device=A.device)
    xl, ldx = _layout(X)
    yl, ldy = _layout(Y)
    L.check(L.load().fgx_sddmm_csr(context(A.device).handle, ctypes.byref(A.desc()), X.shape[1],
                                   _ptr(X), ldx, xl, _ptr(Y), ldy, yl, _ptr(out),
                                   _stream(A.device)), "fgx_sddmm_csr")
    return out


EWISE_OPS = {"mul": L.OP_MUL, "add": L.OP_ADD, "max": L.OP_MAX}
SEMIRINGS = {"plus_times": L.PLUS_TIMES, "min_plus": L.MIN_PLUS}


def ewise(op: str, A: Compressed, B: Compressed) -> Compressed:
    """C = op(A, B) fiber by fiber: symbolic -> allocate -> numeric."""
    if (A.n_outer, A.n_inner) != (B.n_outer, B.n_inner):
        raise L.DimensionMismatch(L.EDIM, f"shapes ({A.n_outer}, {A.n_inner}) and "
                                          f"({B.n_outer}, {B.n_inner}) differ")
    lib = L.load()
    ctx = context(A.device).handle
    s = _stream(A.device)
    opc = EWISE_OPS[op]
    da, db = A.desc(), B.desc()
    cptr = torch.empty(A.n_outer + 1, dtype=A.ptr.dtype, device=A.device)
    nnz = ctypes.c_int64()
    L.check(lib.fgx_ewise_symbolic(ctx, opc, ctypes.byref(da), ctypes.byref(db), _ptr(cptr),
                                   ctypes.byref(nnz), s), "fgx_ewise_symbolic")
    cidx = torch.empty(nnz.value, dtype=A.idx.dtype, device=A.device)
    cval = torch.empty(nnz.value, dtype=A.val.dtype, device=A.device)
    L.check(lib.fgx_ewise_numeric(ctx, opc, ctypes.byref(da), ctypes.byref(db), _ptr(cptr),
                                  _ptr(cidx), _ptr(cval), s), "fgx_ewise_numeric")
    return Compressed(cptr, cidx, cval, A.n_outer, A.n_inner, A.base)


def ewise_async(op: str, A: Compressed, B: Compressed, out=None):
    """C = op(A, B) in one stream-ordered call with no host synchronisation
    (fgx_ewise_async).  Returns (cptr, cidx, cval) with cidx/cval of the
    bound's length (intersection: min(nnz(A), nnz(B)); union: the sum), or
    fills the buffers `out` = (cptr, cidx, cval) given; nnz(C) is
    cptr[-1] - base once the stream reaches it.  Capturable in a graph
    after one call outside the capture."""
    if (A.n_outer, A.n_inner) != (B.n_outer, B.n_inner):
        raise L.DimensionMismatch(L.EDIM, f"shapes ({A.n_outer}, {A.n_inner}) and "
                                          f"({B.n_outer}, {B.n_inner}) differ")
    opc = EWISE_OPS[op]
    if out is None:
        cap = min(A.nnz, B.nnz) if op == "mul" else A.nnz + B.nnz
        out = (torch.empty(A.n_outer + 1, dtype=A.ptr.dtype, device=A.device),
               torch.empty(cap, dtype=A.idx.dtype, device=A.device),
               torch.empty(cap, dtype=A.val.dtype, device=A.device))
    cptr, cidx, cval = out
    if cptr.numel() != A.n_outer + 1 or cidx.numel() != cval.numel():
        raise L.DimensionMismatch(L.EDIM, "ewise_async: output buffer lengths")
    da, db = A.desc(), B.desc()
    L.check(L.load().fgx_ewise_async(context(A.device).handle, opc, ctypes.byref(da),
                                     ctypes.byref(db), _ptr(cptr), _ptr(cidx), _ptr(cval),
                                     cidx.numel(), _stream(A.device)), "fgx_ewise_async")
    return cptr, cidx, cval


def spgemm(semiring: str, A: Compressed, B: Compressed) -> Compressed:
    """Gustavson C = A*B over CSR operands: symbolic -> allocate -> numeric."""
    if A.n_inner != B.n_outer:
        raise L.DimensionMismatch(L.EDIM, f"A has {A.n_inner} columns, B has {B.n_outer} rows")
    lib = L.load()
    ctx = context(A.device).handle
    s = _stream(A.device)
    sr = SEMIRINGS[semiring]
    da, db = A.desc(), B.desc()
    cptr = torch.empty(A.n_outer + 1, dtype=A.ptr.dtype, device=A.device)
    nnz = ctypes.c_int64()
    L.check(lib.fgx_spgemm_symbolic(ctx, sr, ctypes.byref(da), ctypes.byref(db), _ptr(cptr),
                                    ctypes.byref(nnz), s), "fgx_spgemm_symbolic")
    cidx = torch.empty(nnz.value, dtype=A.idx.dtype, device=A.device)
    cval = torch.empty(nnz.value, dtype=A.val.dtype, device=A.device)
    L.check(lib.fgx_spgemm_numeric(ctx, sr, ctypes.byref(da), ctypes.byref(db), _ptr(cptr),
                                   _ptr(cidx), _ptr(cval), s), "fgx_spgemm_numeric")
    return Compressed(cptr, cidx, cval, A.n_outer, B.n_inner, A.base)


def transpose(A: Compressed) -> Compressed:
    """The same matrix stored along its other dimension (CSC <-> CSR): fibre j
    of the result lists, in ascending order, the fibres of A that store j."""
    lib = L.load()
    da = A.desc()
    tptr = torch.empty(A.n_inner + 1, dtype=A.ptr.dtype, device=A.device)
    tidx = torch.empty(A.nnz, dtype=A.idx.dtype, device=A.device)
    tval = torch.empty(A.nnz, dtype=A.val.dtype, device=A.device)
    L.check(lib.fgx_transpose(context(A.device).handle, ctypes.byref(da), _ptr(tptr), _ptr(tidx),
                              _ptr(tval), _stream(A.device)), "fgx_transpose")
    return Compressed(tptr, tidx, tval, A.n_inner, A.n_outer, A.base)


def coo_to_compressed(outer: torch.Tensor, inner: torch.Tensor, val: torch.Tensor, n_outer: int,
                      n_inner: int, base: int = 1) -> Compressed:
    """Coordinate lists -> compressed fibres (fsparse, then Dense(SparseList)):
    sorted by (outer, inner), duplicate coordinates summed in input order."""
    if not (outer.numel() == inner.numel() == val.numel()):
        raise L.DimensionMismatch(L.EDIM, "coordinate and value lists differ in length")
    if outer.dtype != inner.dtype:
        raise L.UnsupportedError(L.EUNSUP, "coordinate lists must share one integer type")
    for name, t in (("outer", outer), ("inner", inner), ("val", val)):
        if not t.is_cuda or not t.is_contiguous():
            raise L.UnsupportedError(L.EUNSUP, f"{name} must be a contiguous device tensor")
    lib = L.load()
    dev = val.device
    n = val.numel()
    cptr = torch.empty(n_outer + 1, dtype=outer.dtype, device=dev)
    cidx = torch.empty(n, dtype=outer.dtype, device=dev)
    cval = torch.empty(n, dtype=val.dtype, device=dev)
    nnz = ctypes.c_int64()
    L.check(lib.fgx_coo_to_compressed(context(dev).handle, int(n_outer), int(n_inner), n,
                                      fgx_dtype(outer.dtype), fgx_dtype(val.dtype), int(base),
                                      _ptr(outer), _ptr(inner), _ptr(val), _ptr(cptr),
                                      _ptr(cidx), _ptr(cval), ctypes.byref(nnz), _stream(dev)),
            "fgx_coo_to_compressed")
    u = nnz.value
    return Compressed(cptr, cidx[:u].clone(), cval[:u].clone(), int(n_outer), int(n_inner),
                      int(base))


def dropfills(A: Compressed, fill: float = 0.0) -> Compressed:
    """The stored entries of A that are not isequal to `fill`."""
    lib = L.load()
    da = A.desc()
    cptr = torch.empty(A.n_outer + 1, dtype=A.ptr.dtype, device=A.device)
    cidx = torch.empty(A.nnz, dtype=A.idx.dtype, device=A.device)
    cval = torch.empty(A.nnz, dtype=A.val.dtype, device=A.device)
    nnz = ctypes.c_int64()
    L.check(lib.fgx_dropfills(context(A.device).handle, ctypes.byref(da), float(fill), _ptr(cptr),
                              _ptr(cidx), _ptr(cval), ctypes.byref(nnz), _stream(A.device)),
            "fgx_dropfills")
    u = nnz.value
    return Compressed(cptr, cidx[:u].clone(), cval[:u].clone(), A.n_outer, A.n_inner, A.base)


def to_device(ptr, idx, val, n_outer, n_inner, base=1, device="cuda") -> Compressed:
    """Uploads host (numpy) compressed buffers."""
    t = lambda a: torch.as_tensor(a).to(device).contiguous()
    return Compressed(t(ptr), t(idx), t(val), int(n_outer), int(n_inner), int(base))
