"""Finch programs routed to the backend's kernels.

``einsum(program, **tensors)`` takes the index notation of Finch's ``@einsum``
(src/interface/einsum.jl:64-82,190-201) and ``@finch`` increments, recognises
the programs the backend implements, and runs them on the device through the C
ABI.  Which kernel a program needs follows from the operands' storage, the way
Finch's lowering follows the level formats:

==========================================  ==========================================
program                                     kernel
==========================================  ==========================================
``y[i] += A[i, j] * x[j]``, A CSC           CSC SpMV (typical_spmv_csc.txt:26-83)
``y[i] += A[i, j] * x[j]``, A swizzled CSR   CSR SpMV (parallel_spmv.txt)
``C[i, k] += A[i, j] * B[j, k]``, A sparse   SpMM (A swizzled CSR), B/C dense
``C[i, k] += A[i, j] * B[j, k]``, both sparse  Gustavson SpGEMM, plus-times
``C[i, k] <<min>>= A[i, j] + B[j, k]``       Gustavson SpGEMM, min-plus (fill Inf)
``C[i, j] = A[i, j] * B[i, j]`` (``+``, ``max``)  element-wise coiteration
==========================================  ==========================================

When the storage order of an operand disagrees with the program (a CSC
operand of SpMM, or sparse operands of different orders), the operand is
transposed on the device first (``fgx_transpose``), as Finch's concordize
pass inserts a transpose (src/scheduler/optimize.jl:396-452); the transposed
copy is cached on the tensor.  ``permutedims``, ``fsparse`` and ``dropfills``
are the device forms of src/interface/copy.jl:64-117 and
src/interface/fsparse.jl:64-97.

``compute(lazy(A) * (lazy(X) @ lazy(Y)))`` is the fused SDDMM plan of
``compute(lazy(M) .* (lazy(A)*lazy(B)))`` (benchmark/benchmarks.jl:141-168).
Programs outside this set raise ``UnsupportedError``; nothing falls back to
the host.
"""
from __future__ import annotations

import re

import numpy as np
import torch

from . import _lib as L
from . import ops
from .tensors import Dense, Element, SparseList, SwizzleArray, Tensor

_TERM = r"(\w+)\[\s*([\w\s,]*)\]"
_PROG = re.compile(r"^\s*" + _TERM + r"\s*(\+=|=|<<\s*(\w+)\s*>>=)\s*(.+?)\s*$")


def _term(s):
    m = re.fullmatch(r"\s*" + _TERM + r"\s*", s)
    if not m:
        return None
    return m.group(1), tuple(x.strip() for x in m.group(2).split(",") if x.strip())


def parse(program: str):
    """(out, out_indices, op, reduce_op, [(name, indices)...], combine) of a program."""
    m = _PROG.match(program)
    if not m:
        raise L.UnsupportedError(L.EUNSUP, f"cannot parse program {program!r}")
    out, out_idx = m.group(1), tuple(x.strip() for x in m.group(2).split(",") if x.strip())
    assign = m.group(3).replace(" ", "")
    reduce_op = m.group(4) or ("+" if assign == "+=" else None)
    rhs = m.group(5)
    fm = re.fullmatch(r"\s*max\s*\((\s*\w+\[[^\]]*\]\s*),(\s*\w+\[[^\]]*\]\s*)\)\s*", rhs)
    if fm:
        combine, parts = "max", [fm.group(1), fm.group(2)]
    else:
        for c in ("*", "+"):
            parts = rhs.split(c)
            if len(parts) == 2:
                combine = c
                break
        else:
            raise L.UnsupportedError(L.EUNSUP, f"unsupported right-hand side {rhs!r}")
    terms = [_term(p) for p in parts]
    if any(t is None for t in terms):
        raise L.UnsupportedError(L.EUNSUP, f"unsupported right-hand side {rhs!r}")
    return out, out_idx, reduce_op, terms, combine


# ---- operand views -----------------------------------------------------------------

def _compressed(t: Tensor) -> ops.Compressed:
    sl = t.lvl.lvl
    return ops.Compressed(sl.ptr, sl.idx, t.element.val, t.lvl.shape, sl.shape, 1)


def _other_order(t: Tensor) -> ops.Compressed:
    """t's matrix stored along the other dimension (device transpose), cached
    on the tensor while its buffers stay the same objects."""
    key = (id(t.lvl.lvl.ptr), id(t.lvl.lvl.idx), id(t.element.val))
    cached = getattr(t, "_other_order", None)
    if cached is None or cached[0] != key:
        cached = (key, ops.transpose(_compressed(t)))
        t._other_order = cached
    return cached[1]


def _kind(x):
    """('csc', Tensor) | ('csr', Tensor) | ('dense', torch) for a device operand."""
    if isinstance(x, SwizzleArray) and x.dims == (2, 1) and x.body.format == "csc":
        return "csr", x.body
    if isinstance(x, Tensor):
        f = x.format
        if f == "csc":
            return "csc", x
        if f in ("dense", "dense2"):
            return "dense", _dense_view(x)
    if isinstance(x, torch.Tensor):
        return "dense", x
    raise L.UnsupportedError(L.EUNSUP, f"operand {x!r} is not a backend format")


def _dense_view(t: Tensor) -> torch.Tensor:
    v = t.element.val
    if not isinstance(v, torch.Tensor) or not v.is_cuda:
        raise L.UnsupportedError(L.EUNSUP, "transfer(GPU(), tensor) the operands first")
    if t.format == "dense":
        return v
    m, n = t.shape
    return v.view(n, m).t()          # column-major (Finch's Dense(Dense)) as a strided view


def _dense_tensor(v: torch.Tensor) -> Tensor:
    """Dense(Element) / Dense(Dense(Element)) tensor over a device result."""
    if v.dim() == 1:
        t = Tensor(Dense(Element(0.0), v.numel()))
        t.element.val = v
        return t
    colmajor = v.t().contiguous()    # Finch stores matrices column-major
    t = Tensor(Dense(Dense(Element(0.0), v.shape[0]), v.shape[1]))
    t.element.val = colmajor.view(-1)
    return t


def _sparse_tensor(C: ops.Compressed, fill=0.0) -> Tensor:
    t = Tensor(Dense(SparseList(Element(fill), C.n_inner), C.n_outer))
    t.lvl.lvl.ptr, t.lvl.lvl.idx = C.ptr, C.idx
    t.element.val = C.val
    return t


# ---- einsum --------------------------------------------------------------------------

def einsum(program: str, **tensors):
    out, oidx, red, terms, comb = parse(program)
    (a_name, a_idx), (b_name, b_idx) = terms
    A, B = tensors[a_name], tensors[b_name]
    ka, va = _kind(A)
    kb, vb = _kind(B)

    # SpMV: y[i] += A[i, j] * x[j]
    if red == "+" and comb == "*" and len(oidx) == 1 and len(a_idx) == 2 and len(b_idx) == 1:
        i, j = a_idx
        if oidx != (i,) or b_idx != (j,):
            raise L.UnsupportedError(L.EUNSUP, f"index pattern of {program!r}")
        if ka == "csc":
            return _dense_tensor(ops.spmv_csc(_compressed(va), vb))
        if ka == "csr":
            return _dense_tensor(ops.spmv_csr(_compressed(va), vb))
    # Two-operand matrix programs.
    if len(a_idx) == 2 and len(b_idx) == 2 and len(oidx) == 2:
        i, j = a_idx
        j2, k = b_idx
        if red is not None and j == j2 and oidx == (i, k):
            if comb == "*" and red == "+" and ka in ("csr", "csc") and kb == "dense":
                Ar = _compressed(va) if ka == "csr" else _other_order(va)
                return _dense_tensor(ops.spmm_csr(Ar, vb))
            if kb in ("csc", "csr") and ka in ("csc", "csr") and ka != kb:
                # Bring B to A's storage order.
                vb = _sparse_tensor(_other_order(vb))
                kb = ka
            if kb in ("csc", "csr") and ka == kb:
                semiring = {("*", "+"): "plus_times", ("+", "min"): "min_plus"}.get((comb, red))
                if semiring is None:
                    raise L.UnsupportedError(L.EUNSUP, f"semiring ({comb}, {red}) unsupported")
                if ka == "csr":  # rows of A times rows of B
                    C = ops.spgemm(semiring, _compressed(va), _compressed(vb))
                    return SwizzleArray(_sparse_tensor(C, _fill(semiring)), (2, 1))
                # Column-compressed: C = A*B column by column is (B^T A^T)^T.
                C = ops.spgemm(semiring, _compressed(vb), _compressed(va))
                return _sparse_tensor(C, _fill(semiring))
        if red is None and a_idx == b_idx == oidx and ka in ("csc", "csr") and kb in ("csc", "csr"):
            if ka != kb:
                vb = _sparse_tensor(_other_order(vb))
            op = {"*": "mul", "+": "add", "max": "max"}[comb]
            C = ops.ewise(op, _compressed(va), _compressed(vb))
            t = _sparse_tensor(C)
            return SwizzleArray(t, (2, 1)) if ka == "csr" else t
    raise L.UnsupportedError(L.EUNSUP, f"program {program!r} over ({ka}, {kb}) operands is not "
                                       "routed to the backend")


def _fill(semiring):
    return float("inf") if semiring == "min_plus" else 0.0


# ---- lazy / compute ------------------------------------------------------------------------

class LazyTensor:
    """A deferred expression: leaves, `@` (tensordot over the shared index) and
    element-wise `*`, `+`, `maximum`."""

    def __init__(self, op, args):
        self.op = op
        self.args = args

    def __matmul__(self, other):
        return LazyTensor("matmul", (self, _lazy(other)))

    def __mul__(self, other):
        return LazyTensor("mul", (self, _lazy(other)))

    def __add__(self, other):
        return LazyTensor("add", (self, _lazy(other)))


def _lazy(x):
    return x if isinstance(x, LazyTensor) else LazyTensor("leaf", (x,))


def lazy(x) -> LazyTensor:
    return _lazy(x)


def maximum(a, b) -> LazyTensor:
    return LazyTensor("max", (_lazy(a), _lazy(b)))


def compute(expr: LazyTensor):
    """Runs a recognised plan (src/interface/lazy.jl:707-730)."""
    if expr.op == "leaf":
        return expr.args[0]
    a, b = expr.args
    # Fused SDDMM: A .* (X @ Y) with A sparse and X, Y dense.
    if expr.op == "mul" and a.op == "leaf" and b.op == "matmul":
        x, y = b.args
        if x.op == "leaf" and y.op == "leaf":
            ka, va = _kind(a.args[0])
            kx, vx = _kind(x.args[0])
            ky, vy = _kind(y.args[0])
            if ka == "csr" and kx == ky == "dense":
                # X is m x k; Y holds the k x n factor; the kernel takes Y^T (n x k).
                out = ops.sddmm_csr(_compressed(va), vx, vy.t())
                C = _compressed(va).with_values(out)
                return SwizzleArray(_sparse_tensor(C), (2, 1))
            if ka == "csc" and kx == ky == "dense":
                # Column fibres j over rows i: the same row dots with the
                # factors' roles exchanged, on A's own pattern.
                out = ops.sddmm_csr(_compressed(va), vy.t(), vx)
                return _sparse_tensor(_compressed(va).with_values(out))
    names = {"matmul": None, "mul": "*", "add": "+", "max": "max"}
    if expr.op in names and a.op == "leaf" and b.op == "leaf":
        A, B = a.args[0], b.args[0]
        if expr.op == "matmul":
            kb, _ = _kind(B)
            if kb == "dense" and (B.dim() if isinstance(B, torch.Tensor) else len(B.shape)) == 1:
                return einsum("y[i] += A[i, j] * x[j]", A=A, x=B)
            return einsum("C[i, k] += A[i, j] * B[j, k]", A=A, B=B)
        c = names[expr.op]
        rhs = "max(A[i, j], B[i, j])" if c == "max" else f"A[i, j] {c} B[i, j]"
        return einsum(f"C[i, j] = {rhs}", A=A, B=B)
    raise L.UnsupportedError(L.EUNSUP, f"plan {expr.op} is not routed to the backend")


# ---- format conversion ---------------------------------------------------------------------

def permutedims(t, perm=(2, 1)):
    """permutedims of a 2-D device matrix (src/interface/copy.jl:64-71).

    A Dense(SparseList) tensor gives the transposed matrix in the same format,
    built on the device; a swizzled CSR view gives its underlying tensor."""
    perm = tuple(perm)
    if perm == (1, 2):
        return t
    if perm != (2, 1):
        raise L.UnsupportedError(L.EUNSUP, f"permutation {perm} of a matrix")
    k, v = _kind(t)
    if k == "csc":
        return _sparse_tensor(ops.transpose(_compressed(v)), v.fill_value)
    if k == "csr":
        return v
    raise L.UnsupportedError(L.EUNSUP, "permutedims of a dense operand is not routed")


def fsparse(I: torch.Tensor, J: torch.Tensor, V: torch.Tensor, shape=None) -> Tensor:
    """fsparse(I, J, V, (m, n)) (src/interface/fsparse.jl:64-97) copied into
    Tensor(Dense(SparseList(Element(0)))), assembled on the device.

    I and J are 1-based device integer tensors.  Entries are ordered column
    by column, rows ascending; duplicates are summed left to right in input
    order."""
    if shape is None:
        shape = (int(I.max()), int(J.max())) if I.numel() else (0, 0)
    m, n = (int(x) for x in shape)
    C = ops.coo_to_compressed(J.contiguous(), I.contiguous(), V.contiguous(), n, m, 1)
    return _sparse_tensor(C)


def dropfills(t, fill=None):
    """dropfills(t) (src/interface/copy.jl:73-117): the same format without
    the stored entries that are isequal to the fill value."""
    k, v = _kind(t)
    if k not in ("csc", "csr"):
        raise L.UnsupportedError(L.EUNSUP, "dropfills of a dense operand is not routed")
    f = v.fill_value if fill is None else fill
    out = _sparse_tensor(ops.dropfills(_compressed(v), float(f)), f)
    return SwizzleArray(out, (2, 1)) if k == "csr" else out


# ---- SpGEMM formulations (docs/examples/spgemm.jl) ------------------------------------------

def _gemm_csc(A, B):
    ka, va = _kind(A)
    kb, vb = _kind(B)
    if ka != "csc" or kb != "csc":
        raise L.UnsupportedError(L.EUNSUP, "the SpGEMM examples take Dense(SparseList) operands")
    return einsum("C[i, k] += A[i, j] * B[j, k]", A=A, B=B)


def spgemm_inner(A, B):
    """docs/examples/spgemm.jl:1-25.  Each C[i, j] sums A[i, k]*B[k, j] over
    the k of both fibres in ascending order, and C stores the (i, j) with at
    least one product; that is the column-order Gustavson result exactly,
    which is how it runs here."""
    return _gemm_csc(A, B)


def spgemm_outer(A, B):
    """docs/examples/spgemm.jl:27-57.  The outer-product loop adds the k-th
    outer product for k ascending, so each C[i, j] again sums over k in
    ascending order: the same values and pattern as spgemm_gustavson."""
    return _gemm_csc(A, B)


def spgemm_gustavson(A, B):
    """docs/examples/spgemm.jl:59-76, on the device (fgx_spgemm_*)."""
    return _gemm_csc(A, B)
