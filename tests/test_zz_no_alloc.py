"""GPU: fgx_spmv_csr_noalloc, the CSR SpMV entry point that allocates nothing.

It runs the same kernels as fgx_spmv_csr with its tile table and carries in
the library's fixed per-GPU scratch array, so its result must equal
fgx_spmv_csr's bit for bit.  Two calls in a row (the second on a smaller
operand) check that the shared array's leftover contents from the first call
do not leak into the second.  Device free memory must not change across the
calls, and an operand needing more than the array holds is refused on the host.
"""
import ctypes

import numpy as np
import pytest
import torch

from finchgpu import _lib as L
from finchgpu import ops, synth

pytestmark = pytest.mark.gpu


def _free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _operand(m, n, nnz, seed, dev, itype, vtype):
    ptr, idx, val = synth.powerlaw_csr(m, n, nnz, cap=max(2, n // 4), seed=seed)
    A = ops.to_device(ptr.astype(np.int64), idx.astype(np.int64), val, m, n, device=dev)
    A = ops.Compressed(A.ptr.to(itype[0]), A.idx.to(itype[1]), A.val.to(vtype), m, n, 1)
    x = torch.as_tensor(synth.dense(n, 1, seed=seed + 1)[:, 0].copy())
    return A, x.to(ops.acc_dtype(vtype)).to(dev)


@pytest.mark.parametrize("itype", [(torch.int32, torch.int32), (torch.int64, torch.int64),
                                   (torch.int64, torch.int32)])
@pytest.mark.parametrize("vtype", [torch.float32, torch.float64, torch.bfloat16])
def test_noalloc_matches_spmv_csr_two_calls(dev, itype, vtype):
    big = _operand(200_000, 150_000, 2_500_000, 5, dev, itype, vtype)
    small = _operand(3_000, 9_000, 40_000, 7, dev, itype, vtype)
    refs = [ops.spmv_csr(A, x) for A, x in (big, small)]
    outs = [torch.full_like(r, float("nan")) for r in refs]
    torch.cuda.synchronize()
    before = _free_bytes()
    for (A, x), y in zip((big, small), outs):
        ops.spmv_csr_noalloc(A, x, y)
        torch.cuda.synchronize()
    assert _free_bytes() == before
    for r, y in zip(refs, outs):
        assert torch.equal(r, y)


def test_noalloc_full_workload_and_repeat(dev):
    # The 1M x 1M, 16M-entry Zipf(1.5) operand fits the array (62,512 tiles).
    ptr, idx, val = synth.powerlaw_csr(1_000_000, 1_000_000, 16_000_000, cap=100_000, seed=11)
    A = ops.to_device(ptr.astype(np.int32), idx.astype(np.int32), val, 1_000_000, 1_000_000,
                      device=dev)
    x = torch.as_tensor(synth.dense(1_000_000, 1, seed=12)[:, 0].copy()).to(dev)
    ref = ops.spmv_csr(A, x)
    y = torch.empty_like(ref)
    for _ in range(2):
        y.fill_(float("nan"))
        ops.spmv_csr_noalloc(A, x, y)
        torch.cuda.synchronize()
        assert torch.equal(ref, y)


def test_noalloc_refuses_operands_beyond_the_array(dev):
    # A descriptor claiming 300M FP32 entries needs ~23 MiB of tile table and
    # carries: refused on the host before any kernel runs (the buffers behind
    # it are never read).
    p = torch.zeros(2, dtype=torch.int64, device=dev)
    A = ops.Compressed(p, p, torch.zeros(2, device=dev), 1, 10, 1)
    d = A.desc()
    d.nnz = 300_000_000
    x = torch.zeros(10, device=dev)
    y = torch.zeros(1, device=dev)
    rc = L.load().fgx_spmv_csr_noalloc(ops.context(dev).handle, ctypes.byref(d), ops._ptr(x),
                                      ops._ptr(y), ops._stream(dev))
    assert rc == L.EINVAL
    assert b"limit" in L.load().fgx_last_error()
