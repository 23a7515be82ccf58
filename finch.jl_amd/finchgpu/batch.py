"""Batched form: many independent SpMV/SpMM products in one persistent launch.

``BatchPlan`` wraps ``fgx_batch_plan_*``: the library picks row-split or
nnz-split per item from its shape and row-length statistics (the autotuner),
cuts all items into a cost-balanced work queue for one persistent kernel and
captures that launch in a hipGraph, so ``run()`` is a single graph replay.
``run_loop`` is the plain path: one library call per item.

``synth_batch`` draws the batched workload on the device: square CSR matrices
with log-uniform row counts, Poisson(mu) nonzeros per row (mu log-uniform),
uniform distinct columns, N(0,1) values, each with a dense FP32 right-hand
block of n_rhs columns.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib as L
from .ops import Compressed, _ptr, _stream, context


class BatchItem:
    """C = A * B for one operand (B: n_inner x n_rhs, C: n_outer x n_rhs, row-major; C FP32,
    B FP32, or BF16 when A's values are BF16)."""

    def __init__(self, A: Compressed, B: torch.Tensor, C: torch.Tensor | None = None):
        if B.dim() == 1:
            B = B.view(-1, 1)
        if C is None:
            C = torch.empty((A.n_outer, B.shape[1]), dtype=torch.float32, device=A.device)
        if C.dim() == 1:
            C = C.view(-1, 1)
        if B.shape[0] != A.n_inner or C.shape[0] != A.n_outer or B.shape[1] != C.shape[1]:
            raise L.DimensionMismatch(L.EDIM, "batch item shapes do not match")
        if not (B.is_contiguous() and C.is_contiguous()):
            raise L.UnsupportedError(L.EUNSUP, "batch dense operands must be contiguous row-major")
        if C.dtype != torch.float32 or B.dtype not in (torch.float32, torch.bfloat16):
            raise L.UnsupportedError(L.EUNSUP, "batch: C must be FP32 and B FP32 or BF16")
        self.A, self.B, self.C = A, B, C

    @property
    def n_rhs(self) -> int:
        return int(self.B.shape[1])

    def struct(self) -> L.BatchItem:
        return L.BatchItem(self.A.desc(), self.n_rhs, _ptr(self.B), _ptr(self.C))


def _items_array(items):
    arr = (L.BatchItem * len(items))()
    for i, it in enumerate(items):
        arr[i] = it.struct()
    return arr


class BatchPlan:
    """One autotuned persistent launch (replayed from a hipGraph) over all items.

    BF16 right-hand sides (every item's B, with BF16 values) select the
    BF16-input kernels (``BATCH_BF16_INPUTS``); mixing FP32 and BF16 B raises."""

    def __init__(self, items, flags: int = 0):
        self.items = list(items)
        bf16 = {it.B.dtype == torch.bfloat16 for it in self.items}
        if len(bf16) > 1:
            raise L.UnsupportedError(L.EUNSUP, "batch: items mix FP32 and BF16 right-hand sides")
        if bf16 == {True}:
            flags |= L.BATCH_BF16_INPUTS
        self.device = self.items[0].A.device if self.items else torch.device("cuda")
        self._ctx = context(self.device)
        self._arr = _items_array(self.items)
        h = ctypes.c_void_p()
        L.check(L.load().fgx_batch_plan_create(self._ctx.handle, self._arr, len(self.items),
                                               int(flags), ctypes.byref(h)),
                "fgx_batch_plan_create")
        self.handle = h

    def run(self):
        L.check(L.load().fgx_batch_plan_run(self._ctx.handle, self.handle, _stream(self.device)),
                "fgx_batch_plan_run")

    def stats(self) -> dict:
        st = (ctypes.c_int64 * 6)()
        L.check(L.load().fgx_batch_plan_stats(self.handle, st), "fgx_batch_plan_stats")
        keys = ("work_units", "items_rowsplit", "items_nnzsplit", "grid", "nnz", "alg_bytes")
        return dict(zip(keys, [int(v) for v in st]))

    def close(self):
        if self.handle:
            L.load().fgx_batch_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_loop(items):
    """Plain path: one fgx_spmv_csr / fgx_spmm_csr call per item."""
    if not items:
        return
    if any(it.B.dtype != torch.float32 for it in items):
        raise L.UnsupportedError(L.EUNSUP, "batch loop: right-hand sides must be FP32")
    arr = _items_array(items)
    dev = items[0].A.device
    L.check(L.load().fgx_batch_run_loop(context(dev).handle, arr, len(items), _stream(dev)),
            "fgx_batch_run_loop")


def synth_matrix(m, mu, seed, device, value_dtype=torch.float32, itype=torch.int32):
    """Square m x m CSR with Poisson(mu) nonzeros per row, drawn on the device."""
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    lengths = torch.poisson(torch.full((m,), float(mu), device=device), generator=g).to(torch.int64)
    lengths.clamp_(max=m)
    rows = torch.repeat_interleave(torch.arange(m, device=device), lengths)
    cols = torch.randint(0, m, (rows.numel(),), device=device, generator=g)
    keys = torch.unique(rows * m + cols)           # sorted, distinct
    rows = keys // m
    cols = keys - rows * m
    ptr = torch.zeros(m + 1, dtype=torch.int64, device=device)
    ptr[1:] = torch.cumsum(torch.bincount(rows, minlength=m), 0)
    val = torch.randn(keys.numel(), device=device, generator=g).to(value_dtype)
    return Compressed((ptr + 1).to(itype), (cols + 1).to(itype), val.contiguous(), m, m, 1)


def synth_batch(count, seed=0, device="cuda", rows=(256, 65536), mu=(4.0, 64.0),
                rhs=(1, 8, 32, 128), value_dtype=torch.float32):
    """The batched workload: `count` items with log-uniform rows and mean nnz/row."""
    rng = np.random.default_rng(seed)
    items = []
    for i in range(count):
        m = int(round(math.exp(rng.uniform(math.log(rows[0]), math.log(rows[1])))))
        mu_i = math.exp(rng.uniform(math.log(mu[0]), math.log(mu[1])))
        k = int(rng.choice(rhs))
        A = synth_matrix(m, mu_i, seed * 100003 + i, device, value_dtype)
        g = torch.Generator(device=device)
        g.manual_seed(seed * 7919 + i)
        B = torch.randn((m, k), device=device, generator=g)
        items.append(BatchItem(A, B))
    return items
