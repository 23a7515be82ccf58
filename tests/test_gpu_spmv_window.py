"""SpMV tiles at the edges of the LDS fibre window.

A wave tile caches the ends of its fibres r_lo .. r_hi in LDS when there are
at most 512 of them, packs them to the fibres that hold positions (zeroing the
empty ones it owns on the way) and walks only those; a wider window is read
from `ptr` instead.  These layouts put a run of K empty rows in front of a row
that fills the first tile, so the first tile's window holds K + 2 fibres:
both sides of the 512 cap, empty rows before the first position, and rows
ending exactly at a tile boundary.  `y` starts as NaN, so a row the kernel
fails to write shows.  FP32 (256 positions per tile) and FP64 (128) runs;
checked against the oracle at the FP64 1e-12 / FP32 1e-5 bars.
"""
import numpy as np
import pytest
import torch

import oracle
from finchgpu import ops

pytestmark = pytest.mark.gpu


def layout(k_empty, tile, rng):
    """K empty rows, a row holding the first `tile` positions, a row
    continuing past the tile, then a stretch of short and empty rows."""
    lens = [0] * k_empty + [tile, 10]
    tail = rng.poisson(0.6, 3000)
    tail[rng.random(tail.size) < 0.5] = 0
    lens = np.concatenate([lens, tail, [tile, 0, 0]]).astype(np.int64)
    return lens


@pytest.mark.parametrize("vdtype", [np.float32, np.float64])
@pytest.mark.parametrize("k_empty", [0, 1, 509, 510, 511, 600])
@pytest.mark.parametrize("base", [0, 1])
def test_window_cap_and_empty_rows(dev, vdtype, k_empty, base):
    rng = np.random.default_rng(k_empty * 4 + base)
    tile = 256 if vdtype == np.float32 else 128
    lens = layout(k_empty, tile, rng)
    m, n = lens.size, 5000
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    idx = np.concatenate([np.sort(rng.choice(n, int(l), replace=False)) for l in lens if l])
    val = rng.standard_normal(idx.size).astype(vdtype)
    x = rng.standard_normal(n).astype(vdtype)
    A = ops.Compressed(torch.from_numpy(ptr + base).to(torch.int32).to(dev),
                       torch.from_numpy(idx + base).to(torch.int32).to(dev),
                       torch.from_numpy(val).to(dev), m, n, base)
    y = torch.full((m,), float("nan"), dtype=torch.from_numpy(x).dtype, device=dev)
    got = ops.spmv_csr(A, torch.from_numpy(x).to(dev), y).cpu().numpy().astype(np.float64)
    ref = oracle.spmv_csr(ptr + base, idx + base, val.astype(np.float64), x.astype(np.float64),
                          base=base)
    assert not np.isnan(got).any()
    assert np.all(got[lens == 0] == 0)
    tol = 1e-12 if vdtype == np.float64 else 1e-5
    assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(ref)


@pytest.mark.parametrize("m", [300, 700])
def test_no_positions(dev, m):
    # nnz = 0: one tile, every row empty; 300 rows fit the LDS window, 700 do not.
    A = ops.Compressed(torch.zeros(m + 1, dtype=torch.int32, device=dev),
                       torch.zeros(0, dtype=torch.int32, device=dev),
                       torch.zeros(0, dtype=torch.float32, device=dev), m, 50, 0)
    y = torch.full((m,), float("nan"), dtype=torch.float32, device=dev)
    got = ops.spmv_csr(A, torch.ones(50, dtype=torch.float32, device=dev), y).cpu().numpy()
    assert np.all(got == 0)
