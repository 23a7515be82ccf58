"""GPU parity: CSC SpMV through row buckets (count, scan, scatter of products
into per-bucket runs, LDS accumulation), the form fgx_spmv_csc takes for
large operands, against the oracle's restatement of
test/reference64/typical/typical_spmv_csc.txt:26-83.

FGX_CSC_BUCKETS=1 forces the bucket path on small operands and =0 forces the
atomic kernel, so the same cases run through both.  The order of the adds
into each y[i] is not fixed on either path (Finch's AtomicElement form), so
FP results are held to the FP64 / FP32 tolerances; y[i] with at most two
terms is exact whatever the order.
"""
import numpy as np
import pytest
import torch

import oracle
from conftest import dense_to_csc
from finchgpu import ops, synth

pytestmark = pytest.mark.gpu

FP64_REL = 1e-12
FP32_NORMWISE = 1e-5


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.linalg.norm(ref)
    return np.linalg.norm(got - ref) / (den if den > 0 else 1.0)


def up(a, dev, dtype=None):
    t = torch.as_tensor(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(dev)


def csc(ptr, idx, val, n_cols, n_rows, dev, itype=torch.int32, ptype=None, base=1):
    ptype = ptype or itype
    return ops.Compressed(up(np.asarray(ptr) - (1 - base), dev, ptype),
                          up(np.asarray(idx) - (1 - base), dev, itype), up(val, dev),
                          n_cols, n_rows, base)


@pytest.fixture(params=["1", "0"], ids=["buckets", "atomic"])
def path(request, monkeypatch):
    monkeypatch.setenv("FGX_CSC_BUCKETS", request.param)
    return request.param


def test_golden(dev, goldens, path):
    g = goldens["spmv_csc"]
    ptr, idx, val = dense_to_csc(np.array(g["A_dense"]))
    y = ops.spmv_csc(csc(ptr, idx, val, 3, 4, dev, torch.int64), up(np.array(g["x"]), dev))
    assert y.cpu().tolist() == g["y"]      # typical_spmv_csc.txt:92, at most two terms per row


@pytest.mark.parametrize("itype,ptype", [(torch.int32, torch.int32), (torch.int64, torch.int64),
                                         (torch.int32, torch.int64)])
@pytest.mark.parametrize("base", [1, 0])
def test_fsprand_10k_fp64(dev, path, itype, ptype, base):
    n = 10_000
    ptr, idx, val = synth.fsprand(n, n, 1e-3, seed=0)
    x = 2.0 * np.random.default_rng(0).random(n) - 1.0
    y = ops.spmv_csc(csc(ptr, idx, val, n, n, dev, itype, ptype, base), up(x, dev)).cpu().numpy()
    ref = oracle.spmv_csc(n, ptr, idx, val, x)
    assert rel_err(y, ref) <= FP64_REL


def _shapes():
    rng = np.random.default_rng(71)
    out = {}
    # Rows not a multiple of the bucket, one row per bucket edge, empty rows.
    m, n = 3 * 8192 + 17, 900
    lens = rng.integers(0, 60, n)
    ptr = np.concatenate([[0], np.cumsum(lens)]) + 1
    idx = np.concatenate([np.sort(rng.choice(m, l, replace=False)) for l in lens]) + 1
    out["ragged_buckets"] = (ptr, idx, rng.standard_normal(idx.size), n, m)
    # Fewer rows than one bucket; a 5000-long column between empty ones.
    L, m = 5000, 5000
    ptr = np.array([1, 1, 1 + L, 1 + L], np.int64)
    out["long_column"] = (ptr, np.arange(1, L + 1), rng.standard_normal(L), 3, m)
    # Only the last row of the last bucket is touched.
    m = 2 * 8192 + 1
    ptr = np.array([1, 3, 3, 4], np.int64)
    out["last_row"] = (ptr, np.array([1, m, m]), np.array([1.5, -2.0, 4.0]), 3, m)
    # Every column empty: y is all zeros.
    out["no_entries"] = (np.ones(11, np.int64), np.zeros(0, np.int64), np.zeros(0), 10, 40_000)
    return out


SHAPES = _shapes()


@pytest.mark.parametrize("name", sorted(SHAPES))
@pytest.mark.parametrize("dtype,tol", [(np.float64, FP64_REL), (np.float32, FP32_NORMWISE)])
def test_shapes(dev, path, name, dtype, tol):
    ptr, idx, val, n, m = SHAPES[name]
    x = np.random.default_rng(5).standard_normal(n).astype(dtype)
    A = csc(ptr, idx, val.astype(dtype), n, m, dev)
    y = ops.spmv_csc(A, up(x, dev)).cpu().numpy()
    ref = oracle.spmv_csc(m, ptr, idx, val.astype(np.float64), x.astype(np.float64))
    assert y.shape == (m,)
    assert rel_err(y, ref) <= tol
    touched = np.zeros(m, bool)
    touched[np.asarray(idx, np.int64) - 1] = True
    assert np.all(y[~touched] == 0)


def test_powerlaw_1m_default_path(dev, monkeypatch):
    # The bench's 1M x 1M Zipf(1.5) operand read column-wise (16M entries):
    # large enough that fgx_spmv_csc takes the bucket path by itself.
    monkeypatch.delenv("FGX_CSC_BUCKETS", raising=False)
    n = 1_000_000
    ptr, idx, val = synth.powerlaw_csr(n, n, 16_000_000, seed=0)
    x = synth.dense(n, 1, seed=1)[:, 0].copy()
    A = csc(ptr, idx, val, n, n, dev)
    y = ops.spmv_csc(A, up(x, dev)).cpu().numpy()
    ref = oracle.spmv_csc(n, ptr, idx, val.astype(np.float64), x.astype(np.float64))
    assert rel_err(y, ref) <= FP32_NORMWISE
    # The ordered form over the same buffers agrees within tolerance.
    yo = ops.spmv_csc_ordered(A, up(x, dev)).cpu().numpy()
    assert rel_err(yo, ref) <= FP32_NORMWISE


def test_too_many_buckets_falls_back(dev, monkeypatch):
    # 20M rows are more buckets than the scatter's LDS slots: the atomic
    # kernel runs even when the bucket path is asked for.
    monkeypatch.setenv("FGX_CSC_BUCKETS", "1")
    m, n = 20_000_000, 4
    rows = np.array([1, 19_999_999, 7, 20_000_000, 8_192 * 1000 + 1])
    ptr = np.array([1, 3, 3, 5, 6], np.int64)
    val = np.array([1.0, 2.0, 3.0, 4.0, 5.0], np.float32)
    x = np.array([1.0, -1.0, 2.0, 0.5], np.float32)
    y = ops.spmv_csc(csc(ptr, rows, val, n, m, dev), up(x, dev)).cpu().numpy()
    ref = oracle.spmv_csc(m, ptr, rows, val.astype(np.float64), x.astype(np.float64))
    assert np.array_equal(y, ref.astype(np.float32))
