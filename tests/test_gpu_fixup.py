"""GPU parity of the carry fix-up on constructed row layouts.

Rows that cross tile boundaries leave carries that a fix-up launch adds in
fixed order.  The matrix below is laid out in position space so that every
case of that launch occurs: more than eight rows spanning several tiles
inside one 64-tile chunk (several rounds of 8-lane subgroups), rows crossing
the boundary between two chunks, rows ending exactly on a tile boundary,
empty rows between long ones, and a long row whose run ends at the last
tile.  SpMV (FP32 and FP64), SpMM (16-lane groups and 2-lane groups), the
batched plan (nnz-split forced) and the serving loop must all match the
oracle, and repeated calls must give the same bits.
"""
import numpy as np
import pytest
import torch

import oracle
from finchgpu import _lib as L
from finchgpu import batch, ops, serve

pytestmark = pytest.mark.gpu

FP64_REL = 1e-12
FP32_NORMWISE = 1e-5


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300)


def carry_heavy_lengths():
    """Row lengths: 20 rows of 600 (each spans 2-3 tiles of 256 positions),
    empty rows between some of them, rows of exactly 256 and 512 on tile
    boundaries, a 9,000-position row across the 64-tile chunk boundary,
    mixed short rows, and a final row of 5,000."""
    rng = np.random.default_rng(5)
    lens = []
    for i in range(20):
        lens.append(600)
        if i % 3 == 0:
            lens += [0, 0]
    # Align to a tile boundary, then rows that end exactly on boundaries.
    lens.append((-sum(lens)) % 256)
    lens += [256, 512, 256, 1, 255, 257]
    lens.append(9000)
    lens += list(rng.integers(0, 300, size=400))
    lens += [0] * 50
    lens.append(5000)
    return np.array(lens, np.int64)


def build(n=20_000, seed=7):
    lens = carry_heavy_lengths()
    rng = np.random.default_rng(seed)
    ptr = np.concatenate([[1], 1 + np.cumsum(lens)]).astype(np.int64)
    idx = np.concatenate([np.sort(rng.choice(n, size=int(l), replace=False)) + 1
                          for l in lens]).astype(np.int64)
    val = rng.standard_normal(idx.size)
    return ptr, idx, val, lens.size, n


@pytest.fixture(scope="module")
def heavy():
    return build()


def device_operand(ptr, idx, val, m, n, dev, vdtype):
    return ops.Compressed(torch.from_numpy(ptr.astype(np.int32)).to(dev),
                          torch.from_numpy(idx.astype(np.int32)).to(dev),
                          torch.from_numpy(val.astype(vdtype)).to(dev), m, n, 1)


def test_layout_covers_the_cases(heavy):
    ptr, idx, val, m, n = heavy
    p0 = ptr[:-1] - 1
    p1 = ptr[1:] - 1
    W = 256
    crossed = np.where(p1 > p0, (p1 - 1) // W - p0 // W, 0)
    first_tile = p0 // W
    long_rows = crossed >= 2
    assert np.bincount(first_tile[long_rows] // 64).max() > 8     # several subgroup rounds
    assert ((p0 // W) // 64 != ((p1 - 1) // W) // 64)[p1 > p0].any()  # crosses a chunk
    assert (p1 % W == 0).sum() >= 2                               # ends on a tile boundary
    assert long_rows[-1]                                          # run ends at the last tile


@pytest.mark.parametrize("vdtype,tol", [(np.float32, FP32_NORMWISE), (np.float64, FP64_REL)])
def test_spmv(dev, heavy, vdtype, tol):
    ptr, idx, val, m, n = heavy
    A = device_operand(ptr, idx, val, m, n, dev, vdtype)
    x = np.random.default_rng(1).standard_normal(n).astype(vdtype)
    xd = torch.from_numpy(x).to(dev)
    y = ops.spmv_csr(A, xd)
    ref = oracle.spmv_csr(ptr, idx, val.astype(vdtype).astype(np.float64), x.astype(np.float64))
    assert rel_err(y.cpu().numpy(), ref) <= tol
    for _ in range(2):
        assert torch.equal(ops.spmv_csr(A, xd), y)


@pytest.mark.parametrize("k", [64, 8])
def test_spmm(dev, heavy, k):
    ptr, idx, val, m, n = heavy
    A = device_operand(ptr, idx, val, m, n, dev, np.float32)
    B = np.random.default_rng(2).standard_normal((n, k)).astype(np.float32)
    Bd = torch.from_numpy(B).to(dev)
    C = ops.spmm_csr(A, Bd)
    ref = oracle.spmm_csr(ptr, idx, val.astype(np.float32).astype(np.float64),
                          B.astype(np.float64))
    assert rel_err(C.cpu().numpy(), ref) <= FP32_NORMWISE
    assert torch.equal(ops.spmm_csr(A, Bd), C)


def test_batched_nnz_split(dev, heavy):
    ptr, idx, val, m, n = heavy
    A = device_operand(ptr, idx, val, m, n, dev, np.float32)
    v64 = val.astype(np.float32).astype(np.float64)
    items = [batch.BatchItem(A, torch.randn((n, k), device=dev)) for k in (1, 64, 8)]
    plan = batch.BatchPlan(items, L.BATCH_FORCE_NNZSPLIT)
    assert plan.stats()["items_nnzsplit"] == len(items)
    plan.run()
    torch.cuda.synchronize()
    for it in items:
        Bh = it.B.cpu().numpy().astype(np.float64)
        ref = (oracle.spmv_csr(ptr, idx, v64, Bh[:, 0].copy())[:, None] if Bh.shape[1] == 1
               else oracle.spmm_csr(ptr, idx, v64, Bh))
        assert rel_err(it.C.cpu().numpy(), ref) <= FP32_NORMWISE
    plan.close()


def test_serving_segmented(dev, heavy):
    ptr, idx, val, m, n = heavy
    A = device_operand(ptr, idx, val, m, n, dev, np.float32)
    v64 = val.astype(np.float32).astype(np.float64)
    g = torch.Generator(device=dev)
    g.manual_seed(4)
    reqs = []
    for i, (kind, k) in enumerate([("spmv", 1), ("spmm", 16), ("spmm", 64), ("spmv", 1)]):
        shape = (n,) if kind == "spmv" else (n, k)
        inp = torch.randn(shape, device=dev, generator=g)
        out = torch.full((m,) if kind == "spmv" else (m, k), float("nan"), device=dev)
        reqs.append(serve.Request(kind, 0, k, 1e-5 * i, inp, None, out))
    res = serve.serve([A], reqs, "batched")
    assert (res.status == 0).all()
    for r in reqs:
        x = r.inp.cpu().numpy().astype(np.float64)
        ref = (oracle.spmv_csr(ptr, idx, v64, x) if r.kind == "spmv"
               else oracle.spmm_csr(ptr, idx, v64, x))
        assert rel_err(r.out.cpu().numpy(), ref) <= FP32_NORMWISE
