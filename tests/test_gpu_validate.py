"""GPU: structural validation of compressed operands (fgx_validate) and the
context's safe mode.

A SparseList level stores fibre p at positions ptr[p]-base .. ptr[p+1]-base-1,
coordinates strictly increasing inside a fibre, ptr[1] == 1
(src/tensors/levels/sparse_list_levels.jl:37-42).  Broken ptr arrays and
unsorted fibres raise FinchProtocolError; coordinates outside the shape and an
nnz that disagrees with ptr raise DimensionMismatch.  The check kernels read
only ptr[0..n_outer] and idx[0..nnz), so every malformed operand here is safe
to hand to fgx_validate itself.  Safe-mode tests through a compute entry point
use only malformations whose reads would stay in bounds.
"""
import numpy as np
import pytest
import torch

from finchgpu import _lib as L
from finchgpu import ops, synth

pytestmark = pytest.mark.gpu


def operand(ptr, idx, val, n_outer, n_inner, dev, itype=torch.int32, base=1):
    t = lambda a, dt: torch.as_tensor(np.asarray(a)).to(dt).to(dev)
    return ops.Compressed(t(ptr, itype), t(idx, itype), t(val, torch.float32), n_outer, n_inner,
                          base)


GOOD = ([1, 3, 3, 5], [1, 4, 2, 3], [1.0, 2.0, 3.0, 4.0])   # 3 fibres over 1:4


@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
@pytest.mark.parametrize("base", [0, 1])
def test_valid_operands_pass(dev, itype, base):
    p, i, v = GOOD
    sh = 1 - base
    ops.validate(operand(np.array(p) - sh, np.array(i) - sh, v, 3, 4, dev, itype, base))
    # Empty fibres everywhere, and no stored entries at all.
    ops.validate(operand(np.full(6, base), [], [], 5, 7, dev, itype, base))
    # A larger power-law operand with long and empty fibres.
    ptr, idx, val = synth.powerlaw_csr(20_000, 30_000, 300_000, cap=20_000, seed=3)
    ops.validate(operand(ptr - sh, idx - sh, val, 20_000, 30_000, dev, itype, base))


BAD = [
    # (label, ptr, idx, exception)
    ("ptr[1] != 1", [2, 3, 3, 5], [1, 4, 2, 3], L.FinchProtocolError),
    ("ptr decreases", [1, 4, 3, 5], [1, 4, 2, 3], L.FinchProtocolError),
    ("ptr beyond nnz", [1, 3, 9, 5], [1, 4, 2, 3], L.FinchProtocolError),
    ("unsorted fibre", [1, 3, 3, 5], [4, 1, 2, 3], L.FinchProtocolError),
    ("duplicate coordinate", [1, 3, 3, 5], [1, 4, 3, 3], L.FinchProtocolError),
    ("nnz != ptr[end]-1", [1, 3, 3, 4], [1, 4, 2, 3], L.DimensionMismatch),
    ("coordinate 0", [1, 3, 3, 5], [0, 4, 2, 3], L.DimensionMismatch),
    ("coordinate n+1", [1, 3, 3, 5], [1, 5, 2, 3], L.DimensionMismatch),
]


@pytest.mark.parametrize("label,ptr,idx,exc", BAD, ids=[b[0] for b in BAD])
@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
def test_malformed_operands_raise(dev, label, ptr, idx, exc, itype):
    A = operand(ptr, idx, [1.0] * len(idx), 3, 4, dev, itype)
    with pytest.raises(exc):
        ops.validate(A)


def test_unsorted_across_fibre_boundary_is_valid(dev):
    # A fibre may start below where the previous one ended.
    ops.validate(operand([1, 3, 5], [3, 4, 1, 2], [1.0] * 4, 2, 4, dev))


def test_safe_mode_rejects_before_compute(dev):
    # Unsorted and duplicate coordinates stay in bounds, so even a kernel run
    # on them could not fault; safe mode must refuse them before any kernel.
    x = torch.ones(4, device=dev)
    y = torch.full((3,), 7.0, device=dev)
    bad = operand([1, 3, 3, 5], [4, 1, 2, 2], [1.0] * 4, 3, 4, dev)
    ops.spmv_csr(bad, x, y)                       # fast mode: trusted, runs
    torch.cuda.synchronize()
    ops.set_safe_mode(True, dev)
    try:
        y.fill_(7.0)
        with pytest.raises(L.FinchProtocolError):
            ops.spmv_csr(bad, x, y)
        torch.cuda.synchronize()
        assert y.tolist() == [7.0, 7.0, 7.0]      # nothing was written
        good = operand(*GOOD, 3, 4, dev)
        assert ops.spmv_csr(good, x).tolist() == [3.0, 0.0, 7.0]
        with pytest.raises(L.FinchProtocolError):
            ops.ewise("add", bad, good)
    finally:
        ops.set_safe_mode(False, dev)
    ops.spmv_csr(bad, x, y)
    torch.cuda.synchronize()
