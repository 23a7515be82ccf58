"""GPU: coiteration in one stream-ordered call (fgx_ewise_async).

The single call runs the count, the device scan and the fill back to back,
with no host synchronisation, into outputs sized to the bound.  Its `ptr`,
and its first nnz coordinates and values, must equal the two-phase call's
bit for bit (which the coiteration tests pin to the oracle); a capacity
below the bound is refused; and the call replays correctly from a captured
HIP graph, with the program's own eager work between replays.
"""
import numpy as np
import pytest
import torch

from finchgpu import _lib as L
from finchgpu import ops, synth

pytestmark = pytest.mark.gpu


def _operand(ptr, idx, val, m, n, dev, itype, base):
    sh = 1 - base
    return ops.Compressed(torch.from_numpy(np.asarray(ptr) - sh).to(itype).to(dev),
                          torch.from_numpy(np.asarray(idx) - sh).to(itype).to(dev),
                          torch.from_numpy(val).to(dev), m, n, base)


def _pair(dev, m, n, la, lb, dtype, itype, base, seed=5):
    rng = np.random.default_rng(seed)
    ap, ai, av = synth.poisson_csr(m, n, la, rng=rng, dtype=np.float64)
    bp, bi, bv = synth.poisson_csr(m, n, lb, rng=rng, dtype=np.float64)
    if dtype == np.int32:
        av, bv = np.rint(av * 10), np.rint(bv * 10)
    A = _operand(ap, ai, av.astype(dtype), m, n, dev, itype, base)
    B = _operand(bp, bi, bv.astype(dtype), m, n, dev, itype, base)
    return A, B


def _same(two_phase: ops.Compressed, one_call):
    cptr, cidx, cval = one_call
    torch.cuda.synchronize()
    nnz = int(cptr[-1].item()) - two_phase.base
    assert torch.equal(cptr, two_phase.ptr)
    assert nnz == two_phase.idx.numel()
    assert torch.equal(cidx[:nnz], two_phase.idx)
    assert torch.equal(cval[:nnz].view(torch.uint8) if cval.dtype.is_floating_point else cval[:nnz],
                       two_phase.val.view(torch.uint8) if cval.dtype.is_floating_point
                       else two_phase.val)


@pytest.mark.parametrize("op", ["mul", "add", "max"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32])
@pytest.mark.parametrize("itype,base", [(torch.int32, 1), (torch.int64, 0)])
@pytest.mark.parametrize("shape", [(2000, 800, 3.0, 9.0), (200, 30000, 70.0, 140.0)])
def test_one_call_equals_two_phase(dev, op, dtype, itype, base, shape):
    A, B = _pair(dev, *shape, dtype, itype, base)
    _same(ops.ewise(op, A, B), ops.ewise_async(op, A, B))


def test_empty_and_capacity(dev):
    A, B = _pair(dev, 500, 300, 2.0, 4.0, np.float32, torch.int32, 1)
    Z = ops.Compressed(torch.ones(501, dtype=torch.int32, device=dev),
                       torch.empty(0, dtype=torch.int32, device=dev),
                       torch.empty(0, dtype=torch.float32, device=dev), 500, 300, 1)
    _same(ops.ewise("add", A, Z), ops.ewise_async("add", A, Z))
    _same(ops.ewise("mul", A, Z), ops.ewise_async("mul", A, Z))
    small = (torch.empty(501, dtype=torch.int32, device=dev),
             torch.empty(A.nnz + B.nnz - 1, dtype=torch.int32, device=dev),
             torch.empty(A.nnz + B.nnz - 1, dtype=torch.float32, device=dev))
    with pytest.raises(L.FinchGPUError):
        ops.ewise_async("add", A, B, out=small)


def test_graph_replays_match_eager(dev):
    A, B = _pair(dev, 20_000, 20_000, 8.0, 24.0, np.float32, torch.int32, 1, seed=9)
    cap = A.nnz + B.nnz
    out = (torch.empty(A.n_outer + 1, dtype=torch.int32, device=dev),
           torch.empty(cap, dtype=torch.int32, device=dev),
           torch.empty(cap, dtype=torch.float32, device=dev))
    s = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(s):
        ops.ewise_async("max", A, B, out=out)     # sizes this stream's scratch
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        ops.ewise_async("max", A, B, out=out)
    for rep in range(3):
        A.val.mul_(-1.5)                          # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _same(ops.ewise("max", A, B), out)
