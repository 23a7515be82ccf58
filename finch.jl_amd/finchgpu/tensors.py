"""Finch's level formats on the host and on the device.

Mirrors the parts of Finch's tensor interface that the backend's kernels read
and write:

* ``Element(fill)`` — leaf values (src/tensors/levels/element_levels.jl:20-22);
* ``SparseList(lvl, shape)`` — 1-based ``ptr``/``idx`` compressed level
  (src/tensors/levels/sparse_list_levels.jl:37-42);
* ``SparseDict(lvl, shape)`` — the hash-table level the ``compute`` scheduler
  gives sparse results (src/interface/traits.jl:509-517), held in its frozen
  form (src/tensors/levels/sparse_dict_levels.jl:356-400): ``ptr`` prefix
  sums, ``idx`` sorted within each fibre, ``val`` the 1-based positions of
  the child level's entries;
* ``Dense(lvl, shape)`` — dense level (src/tensors/levels/dense_levels.jl:26-29);
* ``AtomicElement(fill)``, ``Mutex(lvl)``, ``Separate(lvl)`` — the levels
  Finch's threaded loops write through (atomic_element_levels.jl:15-37,
  mutex_levels.jl:19-36, separate_levels.jl:22-31): as outputs of the routed
  programs, the device kernels' own update protocol (device atomics for the
  scatter-style loops, one writer per row for the row-parallel ones) stands
  in for their per-element atomics, locks and per-position memory;
* ``Tensor(lvl, arr)`` — a tensor over a level tree, built from a dense array
  the way ``dropfills!`` does (src/interface/copy.jl:113-116): column-major
  traversal, fill values dropped, every other value stored;
* ``swizzle(A, 2, 1)`` — the zero-copy index permutation Finch uses to read
  column-compressed buffers row-wise (src/tensors/combinators/swizzle.jl:1-64);
* ``GPU(id)`` and ``transfer(device, tensor)`` — the device and the move of
  level buffers onto it (src/architecture.jl:163,269-294; per level
  sparse_list_levels.jl:64-69, dense_levels.jl:47-49, element_levels.jl:46-48).

Buffers are numpy arrays on the host and torch tensors on a HIP device.  The
formats the kernels take are Dense(Element) (vectors and column-major
matrices) and Dense(SparseList(Element)) (compressed matrices).
"""
from __future__ import annotations

import numpy as np

try:  # torch is only needed for device buffers
    import torch
except ImportError:  # pragma: no cover
    torch = None


def _is_device(a) -> bool:
    return torch is not None and isinstance(a, torch.Tensor) and a.is_cuda


class Element:
    def __init__(self, fill=0.0, dtype=None, val=None):
        self.fill = fill
        self.dtype = np.dtype(dtype) if dtype is not None else np.asarray(fill).dtype
        self.val = val if val is not None else np.zeros(0, self.dtype)

    def __repr__(self):
        return f"Element{{{self.fill}, {self.dtype}}}({_show(self.val)})"


class AtomicElement(Element):
    """Leaf values updated with atomic read-modify-writes
    (src/tensors/levels/atomic_element_levels.jl:15-37).  The values are
    one buffer, as in Element; the device kernels that write an
    AtomicElement output update it with device atomics."""

    def __repr__(self):
        return f"AtomicElement{{{self.fill}, {self.dtype}}}({_show(self.val)})"


class Mutex:
    """A lock per position over the child level
    (src/tensors/levels/mutex_levels.jl:19-36).  The child holds the values;
    the locks guard host threads only and are not materialised here: on the
    device every update of a Mutex output is a device atomic, which is the
    protocol the locks provide (one read-modify-write at a time per position)."""

    def __init__(self, lvl):
        self.lvl = lvl

    def __repr__(self):
        return f"Mutex({self.lvl!r})"


class Separate:
    """Each position owns a separate copy of the child level, in its own
    memory (src/tensors/levels/separate_levels.jl:1-31): ``lvl`` is the
    prototype and ``val`` the per-position levels.  Only Separate(Element)
    is mirrored: ``val`` holds one single-value Element per position."""

    def __init__(self, lvl, val=None):
        self.lvl = lvl
        self.val = list(val) if val is not None else []

    def __repr__(self):
        return f"Separate({self.lvl!r}, {len(self.val)} positions)"


class SparseList:
    def __init__(self, lvl, shape=None, ptr=None, idx=None, Ti=np.int64):
        self.lvl = lvl
        self.shape = shape
        self.Ti = np.dtype(Ti)
        self.ptr = ptr if ptr is not None else np.ones(1, self.Ti)
        self.idx = idx if idx is not None else np.zeros(0, self.Ti)

    def __repr__(self):
        return f"SparseList{{{self.Ti}}}({self.lvl!r}, {self.shape}, {_show(self.ptr)}, {_show(self.idx)})"


class SparseDict:
    """Finch's SparseDict level after freeze.  Freezing counts each fibre's
    entries into ``ptr``, then places every (fibre, coordinate) -> position
    entry of the table in coordinate order, so ``idx`` is sorted within each
    fibre and ``val[q]`` is the position in the child level of the entry
    stored at q.  Entries inserted fibre by fibre in coordinate order (a
    copy from another format, and every result the device kernels write)
    give ``val == 1:nnz``; ``in_order`` records that."""

    def __init__(self, lvl, shape=None, ptr=None, idx=None, val=None, Ti=np.int64,
                 in_order=None):
        self.lvl = lvl
        self.shape = shape
        self.Ti = np.dtype(Ti)
        self.ptr = ptr if ptr is not None else np.ones(1, self.Ti)
        self.idx = idx if idx is not None else np.zeros(0, self.Ti)
        n = len(self.idx) if not _is_device(self.idx) else int(self.idx.numel())
        if val is None:
            val = np.arange(1, n + 1, dtype=self.Ti)
            in_order = True
        self.val = val
        if in_order is None:
            if _is_device(val):
                ar = torch.arange(1, n + 1, device=val.device, dtype=val.dtype)
                in_order = bool(torch.equal(val, ar))
            else:
                in_order = bool(np.array_equal(np.asarray(val), np.arange(1, n + 1)))
        self.in_order = in_order

    def __repr__(self):
        return (f"SparseDict{{{self.Ti}}}({self.lvl!r}, {self.shape}, {_show(self.ptr)}, "
                f"{_show(self.idx)}, {_show(self.val)})")


class Dense:
    def __init__(self, lvl, shape=None):
        self.lvl = lvl
        self.shape = shape

    def __repr__(self):
        return f"Dense({self.lvl!r}, {self.shape})"


def _show(a):
    if _is_device(a):
        return f"<{a.device} {tuple(a.shape)} {a.dtype}>"
    return np.asarray(a).tolist()


class GPU:
    """A HIP device: Finch's `AbstractDevice` for the backend (one per process)."""

    def __init__(self, index: int = 0):
        self.index = int(index)

    @property
    def torch_device(self):
        return torch.device("cuda", self.index)

    def __repr__(self):
        return f"GPU({self.index})"


class Tensor:
    """A tensor over a level tree: Dense(Element), SparseList(Element) or
    Dense(SparseList(Element))."""

    def __init__(self, lvl, arr=None):
        self.lvl = lvl
        if arr is not None:
            self._fill_from(arr)

    # ---- formats --------------------------------------------------------------
    @property
    def format(self) -> str:
        l = self.lvl
        if isinstance(l, Dense) and _leaf(l.lvl):
            return "dense"
        if isinstance(l, Dense) and isinstance(l.lvl, Dense) and _leaf(l.lvl.lvl):
            return "dense2"
        if (isinstance(l, Dense) and isinstance(l.lvl, Separate)
                and type(l.lvl.lvl) is Element):
            return "separate"
        if isinstance(l, SparseList) and isinstance(l.lvl, Element):
            return "sparsevec"
        if isinstance(l, Dense) and isinstance(l.lvl, SparseList) and isinstance(l.lvl.lvl, Element):
            return "csc"
        if isinstance(l, Dense) and isinstance(l.lvl, SparseDict) and isinstance(l.lvl.lvl, Element):
            return "cscdict"
        raise NotImplementedError(f"level tree {l!r} is not a backend format")

    @property
    def shape(self):
        f = self.format
        if f in ("dense", "sparsevec", "separate"):
            return (self.lvl.shape,)
        if f == "dense2":
            return (self.lvl.lvl.shape, self.lvl.shape)
        return (self.lvl.lvl.shape, self.lvl.shape)

    @property
    def element(self) -> Element:
        l = self.lvl
        while not isinstance(l, Element):
            l = l.lvl
        return l

    @property
    def fill_value(self):
        return self.element.fill

    @property
    def on_device(self) -> bool:
        return _is_device(self.element.val)

    def separate_values(self):
        """The values of a Dense(Separate(Element)) tensor, one per position
        (a host array, or a device tensor gathered from the positions)."""
        vals = [e.val for e in self.lvl.lvl.val]
        if vals and _is_device(vals[0]):
            return torch.cat(vals)
        return np.concatenate(vals) if vals else np.zeros(0, self.element.dtype)

    def countstored(self) -> int:
        f = self.format
        if f in ("csc", "cscdict"):
            return int(self.lvl.lvl.idx.numel() if self.on_device else len(self.lvl.lvl.idx))
        if f == "sparsevec":
            return int(len(self.lvl.idx))
        return int(np.prod(self.shape))

    # ---- construction ----------------------------------------------------------
    def _fill_from(self, arr):
        f = self.format
        el = self.element
        fill = el.fill
        if hasattr(arr, "tocsc"):  # scipy.sparse (zero-copy in Finch; a copy here)
            if f not in ("csc", "cscdict"):
                raise NotImplementedError("scipy input needs Dense(SparseList(Element))")
            S = arr.tocsc()
            S.sort_indices()
            sl = self.lvl.lvl
            sl.shape, self.lvl.shape = S.shape
            sl.ptr = S.indptr.astype(sl.Ti) + 1
            sl.idx = S.indices.astype(sl.Ti) + 1
            el.val = S.data.astype(el.dtype if el.val.size == 0 else el.val.dtype)
            if f == "cscdict":
                sl.val, sl.in_order = np.arange(1, S.nnz + 1, dtype=sl.Ti), True
            return
        a = np.asarray(arr)
        dt = np.result_type(a.dtype, np.asarray(fill).dtype) if el.val.size == 0 else el.val.dtype
        if f == "separate":
            self.lvl.shape = a.shape[0]
            self.lvl.lvl.val = [Element(fill, dt, np.array([v], dt)) for v in a.astype(dt)]
            el.dtype = np.dtype(dt)
            return
        if f == "dense":
            self.lvl.shape = a.shape[0]
            el.val = a.astype(dt).copy()
        elif f == "dense2":
            self.lvl.lvl.shape, self.lvl.shape = a.shape
            el.val = np.asfortranarray(a).astype(dt).ravel(order="F")
        elif f == "sparsevec":
            keep = a != fill
            self.lvl.shape = a.shape[0]
            self.lvl.ptr = np.array([1, keep.sum() + 1], self.lvl.Ti)
            self.lvl.idx = (np.nonzero(keep)[0] + 1).astype(self.lvl.Ti)
            el.val = a[keep].astype(dt)
        else:  # csc: column-major traversal of a 2-D array
            m, n = a.shape
            sl = self.lvl.lvl
            sl.shape, self.lvl.shape = m, n
            keep = a != fill
            counts = keep.sum(axis=0)
            sl.ptr = np.concatenate([[1], 1 + np.cumsum(counts)]).astype(sl.Ti)
            rows, cols = np.nonzero(keep.T)           # column-major order
            sl.idx = (cols + 1).astype(sl.Ti)
            el.val = a.T[rows, cols].astype(dt)
            if f == "cscdict":   # inserted column by column, rows ascending
                sl.val, sl.in_order = np.arange(1, len(sl.idx) + 1, dtype=sl.Ti), True
        el.dtype = np.dtype(dt) if not _is_device(el.val) else el.dtype

    @classmethod
    def from_buffers(cls, ptr, idx, val, m, n, fill=0.0, base=1):
        """Dense(SparseList(Element)) over existing buffers (1-based, or 0-based
        with base=0 as a PlusOneVector view)."""
        t = cls(Dense(SparseList(Element(fill, np.asarray(val).dtype if not _is_device(val) else None),
                                 m), n))
        sl = t.lvl.lvl
        if base == 0:
            ptr = ptr + 1
            idx = idx + 1
        sl.ptr, sl.idx = ptr, idx
        t.element.val = val
        return t

    # ---- views -------------------------------------------------------------------
    def todense(self) -> np.ndarray:
        """Host dense copy (for inspection and tests)."""
        f = self.format
        el = self.element
        host = lambda a: a.cpu().numpy() if _is_device(a) else np.asarray(a)
        if f == "separate":
            return host(self.separate_values()).copy()
        val = host(el.val)
        if f == "dense":
            return val.copy()
        if f == "dense2":
            return val.reshape(self.shape, order="F").copy()
        if f == "sparsevec":
            out = np.full(self.shape, el.fill, dtype=val.dtype)
            out[host(self.lvl.idx) - 1] = val
            return out
        m, n = self.shape
        sl = self.lvl.lvl
        ptr, idx = host(sl.ptr).astype(np.int64), host(sl.idx).astype(np.int64)
        out = np.full((m, n), el.fill, dtype=val.dtype)
        cols = np.repeat(np.arange(n), np.diff(ptr))
        if f == "cscdict":
            val = val[host(sl.val).astype(np.int64) - 1]
        out[idx - 1, cols] = val
        return out

    def __repr__(self):
        return f"Tensor({self.lvl!r})"


class SwizzleArray:
    """swizzle(A, dims...): A read with permuted indices (no data moves)."""

    def __init__(self, body: Tensor, dims):
        self.body = body
        self.dims = tuple(dims)

    @property
    def shape(self):
        s = self.body.shape
        return tuple(s[d - 1] for d in self.dims)

    def todense(self):
        return np.transpose(self.body.todense(), [d - 1 for d in self.dims])

    def __repr__(self):
        return f"swizzle({self.body!r}, {', '.join(map(str, self.dims))})"


def swizzle(A, *dims) -> SwizzleArray:
    if isinstance(A, SwizzleArray):  # compose permutations
        dims = tuple(A.dims[d - 1] for d in dims)
        A = A.body
    if sorted(dims) != list(range(1, len(dims) + 1)):
        raise ValueError(f"{dims} is not a permutation")
    return SwizzleArray(A, dims)


def transfer(device: GPU, t):
    """Moves every level buffer of `t` to `device` (torch tensors)."""
    if isinstance(t, SwizzleArray):
        return SwizzleArray(transfer(device, t.body), t.dims)
    dev = device.torch_device
    mv = lambda a: a.to(dev) if _is_device(a) else torch.as_tensor(np.ascontiguousarray(a)).to(dev)
    out = Tensor(_copy_levels(t.lvl))
    l = out.lvl
    while True:
        if isinstance(l, SparseList):
            l.ptr, l.idx = mv(l.ptr), mv(l.idx)
        if isinstance(l, SparseDict):
            l.ptr, l.idx, l.val = mv(l.ptr), mv(l.idx), mv(l.val)
        if isinstance(l, Separate):
            for e in l.val:
                e.val = mv(e.val)
        if isinstance(l, Element):
            l.val = mv(l.val)
            break
        l = l.lvl
    return out


def _leaf(l) -> bool:
    """A value leaf: Element or AtomicElement, possibly under Mutex locks."""
    while isinstance(l, Mutex):
        l = l.lvl
    return isinstance(l, Element)


def _copy_levels(l):
    if isinstance(l, Element):
        return type(l)(l.fill, l.dtype, l.val)
    if isinstance(l, Mutex):
        return Mutex(_copy_levels(l.lvl))
    if isinstance(l, Separate):
        return Separate(_copy_levels(l.lvl), [Element(e.fill, e.dtype, e.val) for e in l.val])
    if isinstance(l, SparseList):
        return SparseList(_copy_levels(l.lvl), l.shape, l.ptr, l.idx, l.Ti)
    if isinstance(l, SparseDict):
        return SparseDict(_copy_levels(l.lvl), l.shape, l.ptr, l.idx, l.val, l.Ti, l.in_order)
    return Dense(_copy_levels(l.lvl), l.shape)


def fsprand(m, n, p, seed=0, dtype=np.float64, Ti=np.int64):
    """A Dense(SparseList(Element(0))) m x n tensor with density p (the
    distribution of Finch's fsprand; see finchgpu.synth)."""
    from .synth import fsprand as _fsprand
    ptr, idx, val = _fsprand(m, n, p, seed=seed, dtype=dtype)
    t = Tensor(Dense(SparseList(Element(dtype(0)), m, Ti=Ti), n))
    t.lvl.lvl.ptr, t.lvl.lvl.idx = ptr.astype(Ti), idx.astype(Ti)
    t.element.val = val
    return t
