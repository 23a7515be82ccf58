"""GPU parity: SpMV (CSR/CSC), SpMM and SDDMM through the C ABI against the oracle.

Tolerances (stated by the backend's contract): FP64 results relative error
<= 1e-12 against the FP64 oracle; FP32 results normwise error <= 1e-5 against
an FP64 oracle run on the same (FP32-rounded) inputs; integer-exact goldens
bit for bit.
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


def operand(ptr, idx, val, n_outer, n_inner, dev, itype=torch.int64, ptype=None, base=1):
    ptype = ptype or itype
    return ops.Compressed(up(np.asarray(ptr) - (1 - base), dev, ptype),
                          up(np.asarray(idx) - (1 - base), dev, itype), up(val, dev),
                          n_outer, n_inner, base)


IDX = [(torch.int32, torch.int32), (torch.int64, torch.int64), (torch.int32, torch.int64)]


# ---- goldens -------------------------------------------------------------------

def test_spmv_csc_golden(dev, goldens):
    g = goldens["spmv_csc"]
    ptr, idx, val = dense_to_csc(np.array(g["A_dense"]))
    A = operand(ptr, idx, val, 3, 4, dev)
    y = ops.spmv_csc(A, up(np.array(g["x"]), dev)).cpu().numpy()
    # typical_spmv_csc.txt:92; each y[i] here is summed by one lane in column order.
    assert y.tolist() == g["y"]


def test_spmv_csr_golden_transpose(dev, goldens):
    g = goldens["spmv_transpose"]
    ptr, idx, val = dense_to_csc(np.array(g["A_dense"]))
    A = operand(ptr, idx, val, 2, 2, dev)
    assert ops.spmv_csr(A, up(np.array(g["x"]), dev)).cpu().tolist() == g["y"]


def test_zero_based_buffers(dev, goldens):
    g = goldens["zero_based"]
    A = ops.Compressed(up(np.array(g["ptr"]), dev), up(np.array(g["idx"]), dev),
                       up(np.array(g["val"]), dev), g["n"], g["m"], base=0)
    dense = np.array(g["A_dense"])
    for j in range(g["n"]):
        e = up(np.eye(g["n"])[j], dev)
        assert ops.spmv_csc(A, e).cpu().tolist() == dense[:, j].tolist()


# ---- 10k x 10k FP64 at density 1e-3 (fsprand), CSC and swizzled CSR -------------

@pytest.mark.parametrize("itype,ptype", IDX)
@pytest.mark.parametrize("base", [1, 0])
def test_fsprand_10k_fp64(dev, itype, ptype, base):
    n = 10_000
    ptr, idx, val = synth.fsprand(n, n, 1e-3, seed=0)
    rng = np.random.default_rng(0)
    x = 2.0 * rng.random(n) - 1.0
    A = operand(ptr, idx, val, n, n, dev, itype, ptype, base)
    # CSC: y = A x (typical_spmv_csc.txt order).
    y = ops.spmv_csc(A, up(x, dev)).cpu().numpy()
    ref = oracle.spmv_csc(n, ptr, idx, val, x)
    assert rel_err(y, ref) <= FP64_REL
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # CSR = swizzle(A, 2, 1) over the same buffers: y = A^T x.
    yt = ops.spmv_csr(A, up(x, dev)).cpu().numpy()
    reft = oracle.spmv_csr(ptr, idx, val, x)
    assert rel_err(yt, reft) <= FP64_REL
    np.testing.assert_allclose(yt, reft, rtol=1e-12, atol=1e-12 * np.abs(reft).max())


# ---- edge shapes ---------------------------------------------------------------------

def _edge_cases():
    rng = np.random.default_rng(11)
    cases = {}
    cases["empty"] = (np.ones(6, np.int64), np.zeros(0, np.int64), np.zeros(0), 5, 7)
    # One 100k-long row between empty rows: spans many tiles (carry fix-up).
    L = 100_000
    ptr = np.array([1, 1, 1 + L, 1 + L, 1 + L], np.int64)
    idx = np.sort(rng.choice(200_000, L, replace=False)) + 1
    cases["long_row"] = (ptr, idx, rng.standard_normal(L), 4, 200_000)
    # 30k empty rows, then a few dense ones, then 20k trailing empty rows:
    # windows larger than the LDS cache take the global path.
    m = 50_003
    lens = np.zeros(m, np.int64)
    lens[30_000:30_003] = [5, 700, 1]
    ptr = np.concatenate([[0], np.cumsum(lens)]) + 1
    idx = np.concatenate([np.sort(rng.choice(1000, l, replace=False)) for l in lens if l]) + 1
    cases["sparse_rows"] = (ptr, idx, rng.standard_normal(idx.size), m, 1000)
    # Every row length from 0 to 300 (runs and tiles end at every offset).
    lens = np.arange(301)
    rng.shuffle(lens)
    ptr = np.concatenate([[0], np.cumsum(lens)]) + 1
    idx = np.concatenate([np.sort(rng.choice(400, l, replace=False)) for l in lens if l]) + 1
    cases["ragged"] = (ptr, idx, rng.standard_normal(idx.size), lens.size, 400)
    cases["single"] = (np.array([1, 2]), np.array([3]), np.array([2.5]), 1, 3)
    return cases


EDGE = _edge_cases()


@pytest.mark.parametrize("name", sorted(EDGE))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spmv_edges(dev, name, dtype):
    ptr, idx, val, m, n = EDGE[name]
    val = val.astype(dtype)
    x = np.random.default_rng(1).standard_normal(n).astype(dtype)
    A = operand(ptr, idx, val, m, n, dev, torch.int32)
    y = ops.spmv_csr(A, up(x, dev)).cpu().numpy()
    ref64 = oracle.spmv_csr(ptr, idx, val.astype(np.float64), x.astype(np.float64))
    tol = FP64_REL if dtype == np.float64 else FP32_NORMWISE
    assert y.shape == (m,)
    assert rel_err(y, ref64) <= tol
    assert np.all(y[np.diff(ptr) == 0] == 0)
    # CSC reading of the same buffers.
    xc = np.random.default_rng(2).standard_normal(m).astype(dtype)
    yc = ops.spmv_csc(A, up(xc, dev)).cpu().numpy()
    refc = oracle.spmv_csc(n, ptr, idx, val.astype(np.float64), xc.astype(np.float64))
    assert rel_err(yc, refc) <= tol


def test_spmv_deterministic(dev):
    ptr, idx, val, m, n = EDGE["long_row"]
    A = operand(ptr, idx, val.astype(np.float32), m, n, dev, torch.int32)
    x = up(np.random.default_rng(3).standard_normal(n).astype(np.float32), dev)
    y0 = ops.spmv_csr(A, x)
    for _ in range(3):
        assert torch.equal(ops.spmv_csr(A, x), y0)


# ---- SpMM -------------------------------------------------------------------------------

@pytest.mark.parametrize("k", [1, 3, 8, 16, 32, 64, 100, 128, 300])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_spmm_widths(dev, k, dtype):
    ptr, idx, val, m, n = EDGE["ragged"]
    val = val.astype(dtype)
    B = synth.dense(n, k, seed=k, dtype=dtype)
    A = operand(ptr, idx, val, m, n, dev, torch.int32)
    C = ops.spmm_csr(A, up(B, dev)).cpu().numpy()
    ref = oracle.spmm_csr(ptr, idx, val.astype(np.float64), B.astype(np.float64))
    tol = FP64_REL if dtype == np.float64 else FP32_NORMWISE
    assert rel_err(C, ref) <= tol


@pytest.mark.parametrize("layout", ["row", "col"])
def test_spmm_layouts_and_long_rows(dev, layout):
    ptr, idx, val, m, n = EDGE["long_row"]
    val = val.astype(np.float32)
    B = synth.dense(n, 64, seed=5)
    A = operand(ptr, idx, val, m, n, dev, torch.int64)
    Bd = up(B, dev) if layout == "row" else up(np.asfortranarray(B).T, dev).t()
    C = ops.spmm_csr(A, Bd).cpu().numpy()
    ref = oracle.spmm_csr(ptr, idx, val.astype(np.float64), B.astype(np.float64))
    assert rel_err(C, ref) <= FP32_NORMWISE


@pytest.mark.parametrize("k", [3, 64, 100])
def test_spmm_column_major_in_and_out(dev, k):
    # Finch's Dense(Dense) operands are column-major on both sides: widths of
    # 4 and more are transposed to k-fast rows, narrower ones run strided.
    ptr, idx, val, m, n = EDGE["ragged"]
    val = val.astype(np.float32)
    B = synth.dense(n, k, seed=6)
    A = operand(ptr, idx, val, m, n, dev, torch.int32)
    Bd = up(np.ascontiguousarray(B.T), dev).t()
    Cd = torch.full((k, m), float("nan"), device=dev).t()
    ops.spmm_csr(A, Bd, Cd)
    ref = oracle.spmm_csr(ptr, idx, val.astype(np.float64), B.astype(np.float64))
    assert rel_err(Cd.cpu().numpy(), ref) <= FP32_NORMWISE
    if k >= 4:  # same bits as the row-major call
        assert torch.equal(Cd.contiguous(), ops.spmm_csr(A, up(B, dev)))


# ---- SDDMM ------------------------------------------------------------------------------

@pytest.mark.parametrize("k", [6, 32, 64])
def test_sddmm_column_major(dev, k):
    ptr, idx, val, m, n = EDGE["long_row"]
    val = val.astype(np.float32)
    X = synth.dense(m, k, seed=1)
    Y = synth.dense(n, k, seed=2)
    A = operand(ptr, idx, val, m, n, dev, torch.int32)
    Xc = up(np.ascontiguousarray(X.T), dev).t()
    Yc = up(np.ascontiguousarray(Y.T), dev).t()
    out = ops.sddmm_csr(A, Xc, Yc).cpu().numpy()
    ref = oracle.sddmm_csr(ptr, idx, val.astype(np.float64), X.astype(np.float64),
                           Y.astype(np.float64))
    assert rel_err(out, ref) <= FP32_NORMWISE

@pytest.mark.parametrize("k", [4, 32, 36, 64, 100, 7])
@pytest.mark.parametrize("name", ["ragged", "long_row", "sparse_rows"])
def test_sddmm(dev, k, name):
    ptr, idx, val, m, n = EDGE[name]
    val = val.astype(np.float32)
    X = synth.dense(m, k, seed=1)
    Y = synth.dense(n, k, seed=2)
    A = operand(ptr, idx, val, m, n, dev, torch.int32)
    out = ops.sddmm_csr(A, up(X, dev), up(Y, dev)).cpu().numpy()
    ref = oracle.sddmm_csr(ptr, idx, val.astype(np.float64), X.astype(np.float64),
                           Y.astype(np.float64))
    assert rel_err(out, ref) <= FP32_NORMWISE


def test_sddmm_fp64(dev):
    ptr, idx, val, m, n = EDGE["ragged"]
    X = synth.dense(m, 16, seed=1, dtype=np.float64)
    Y = synth.dense(n, 16, seed=2, dtype=np.float64)
    A = operand(ptr, idx, val, m, n, dev, torch.int64)
    out = ops.sddmm_csr(A, up(X, dev), up(Y, dev)).cpu().numpy()
    ref = oracle.sddmm_csr(ptr, idx, val, X, Y)
    assert rel_err(out, ref) <= FP64_REL


@pytest.mark.parametrize("k", [256, 300, 1000])
@pytest.mark.parametrize("layout", ["row", "col"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sddmm_few_positions_wide_rows(dev, k, layout, dtype):
    # Finch's documented example shape (getting_started.md:126-134): 100
    # stored positions of a 1000 x 1000 pattern, dense factors of width k, in
    # Finch's column-major Dense(Dense) layout or k-fast rows.  Few positions
    # and wide rows take a wave per position.
    rng = np.random.default_rng(k)
    m = n = 1000
    flat = np.sort(rng.choice(m * n, 100, replace=False))
    rows, cols = flat // n, flat % n
    ptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=m))]) + 1
    idx = cols + 1
    val = rng.random(100).astype(dtype)
    X = rng.random((m, k)).astype(dtype)
    Y = rng.random((n, k)).astype(dtype)
    A = operand(ptr, idx, val, m, n, dev, torch.int32)
    if layout == "row":
        Xd, Yd = up(X, dev), up(Y, dev)
    else:
        Xd = up(np.ascontiguousarray(X.T), dev).t()
        Yd = up(np.ascontiguousarray(Y.T), dev).t()
    out = ops.sddmm_csr(A, Xd, Yd).cpu().numpy()
    ref = oracle.sddmm_csr(ptr, idx, val.astype(np.float64), X.astype(np.float64),
                           Y.astype(np.float64))
    assert rel_err(out, ref) <= (FP64_REL if dtype == np.float64 else FP32_NORMWISE)


@pytest.mark.parametrize("vtype", [torch.bfloat16, torch.float16])
def test_sddmm_few_positions_wide_rows_half_values(dev, vtype):
    # The wave-per-position path with 16-bit stored values: FP32 dense
    # factors and accumulation, checked against FP64 on the rounded values.
    rng = np.random.default_rng(3)
    m = n = 500
    k = 384
    flat = np.sort(rng.choice(m * n, 60, replace=False))
    rows, cols = flat // n, flat % n
    ptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=m))]) + 1
    idx = cols + 1
    v16 = torch.from_numpy(rng.standard_normal(60).astype(np.float32)).to(vtype)
    X = rng.standard_normal((m, k)).astype(np.float32)
    Y = rng.standard_normal((n, k)).astype(np.float32)
    A = ops.Compressed(up(ptr.astype(np.int32), dev), up(idx.astype(np.int32), dev),
                       v16.to(dev), m, n, 1)
    out = ops.sddmm_csr(A, up(X, dev), up(Y, dev)).cpu().numpy()
    ref = oracle.sddmm_csr(ptr, idx, v16.float().numpy().astype(np.float64), X.astype(np.float64),
                           Y.astype(np.float64))
    assert rel_err(out, ref) <= FP32_NORMWISE


# ---- 1M x 1M, ~16M nnz, Zipf(1.5) rows, at full size --------------------------------------

@pytest.fixture(scope="module")
def powerlaw_1m():
    ptr, idx, val = synth.powerlaw_csr(1_000_000, 1_000_000, 16_000_000, seed=0)
    return ptr, idx, val


def test_powerlaw_1m_spmv_spmm_sddmm(dev, powerlaw_1m):
    ptr, idx, val = powerlaw_1m
    m = n = 1_000_000
    A = operand(ptr, idx, val, m, n, dev, torch.int32)
    nth = max(1, min(16, oracle.max_threads()))
    x = synth.dense(n, 1, seed=1)[:, 0].copy()
    yd = ops.spmv_csr(A, up(x, dev))
    # The entry point that allocates nothing (table and carries in the
    # library's static array) gives the same bits at full size.
    assert torch.equal(yd, ops.spmv_csr_noalloc(A, up(x, dev), torch.empty_like(yd)))
    y = yd.cpu().numpy()
    ref = oracle.spmv_csr(ptr, idx, val.astype(np.float64), x.astype(np.float64), nthreads=nth)
    assert rel_err(y, ref) <= FP32_NORMWISE
    B = synth.dense(n, 64, seed=2)
    C = ops.spmm_csr(A, up(B, dev)).cpu().numpy()
    refC = oracle.spmm_csr(ptr, idx, val.astype(np.float64), B.astype(np.float64), nthreads=nth)
    assert rel_err(C, refC) <= FP32_NORMWISE
    del refC, C
    X = synth.dense(m, 64, seed=3)
    out = ops.sddmm_csr(A, up(X, dev), up(B, dev)).cpu().numpy()
    refS = oracle.sddmm_csr(ptr, idx, val.astype(np.float64), X.astype(np.float64),
                            B.astype(np.float64))
    assert rel_err(out, refS) <= FP32_NORMWISE


def test_bf16_values_fp32_accumulate(dev, powerlaw_1m):
    ptr, idx, val = powerlaw_1m
    n = 1_000_000
    A = operand(ptr, idx, val, n, n, dev, torch.int32)
    Ab = A.to_values(torch.bfloat16)
    B = up(synth.dense(n, 64, seed=2), dev)
    C32 = ops.spmm_csr(A, B).cpu().numpy()
    Cb = ops.spmm_csr(Ab, B).cpu().numpy()
    # Storage rounding of values only: normwise error vs FP32 values <= 1e-2.
    assert rel_err(Cb, C32) <= 1e-2
    # Against the oracle on the bf16-rounded values it is an FP32 sum again.
    vb = Ab.val.float().cpu().numpy().astype(np.float64)
    refb = oracle.spmv_csr(ptr, idx, vb, B[:, 0].cpu().numpy().astype(np.float64), nthreads=8)
    assert rel_err(Cb[:, 0], refb) <= FP32_NORMWISE


# ---- FP16-stored values, FP32 accumulation ---------------------------------------

@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
@pytest.mark.parametrize("k", [1, 8, 64])
def test_fp16_values_fp32_accumulate(dev, itype, k):
    # Values stored as IEEE half, widened on load, summed in FP32.  Against the
    # FP64 oracle on the half-rounded values the result is an FP32 sum again
    # (1e-5); against the unrounded values only storage rounding remains (1e-2).
    m, n = 120_000, 90_000
    ptr, idx, val = synth.powerlaw_csr(m, n, 1_500_000, cap=20_000, seed=41)
    A = operand(ptr, idx, val, m, n, dev, itype)
    Ah = A.to_values(torch.float16)
    assert Ah.val.dtype == torch.float16
    vh = Ah.val.float().cpu().numpy().astype(np.float64)
    # The conversion kernel rounds to nearest even like numpy's float16 cast.
    assert np.array_equal(vh, val.astype(np.float16).astype(np.float64))
    Bn = synth.dense(n, k, seed=42)
    if k == 1:
        got = ops.spmv_csr(Ah, up(Bn[:, 0].copy(), dev)).cpu().numpy()
        ref_h = oracle.spmv_csr(ptr, idx, vh, Bn[:, 0].astype(np.float64), nthreads=8)
        ref_v = oracle.spmv_csr(ptr, idx, val.astype(np.float64), Bn[:, 0].astype(np.float64),
                                nthreads=8)
    else:
        got = ops.spmm_csr(Ah, up(Bn, dev)).cpu().numpy()
        ref_h = oracle.spmm_csr(ptr, idx, vh, Bn.astype(np.float64), nthreads=8)
        ref_v = oracle.spmm_csr(ptr, idx, val.astype(np.float64), Bn.astype(np.float64),
                                nthreads=8)
    assert rel_err(got, ref_h) <= FP32_NORMWISE
    assert rel_err(got, ref_v) <= 1e-2


def test_fp16_values_sddmm_and_csc(dev):
    m, n, k = 40_000, 30_000, 32
    ptr, idx, val = synth.powerlaw_csr(m, n, 400_000, cap=5_000, seed=43)
    A = operand(ptr, idx, val, m, n, dev, torch.int32).to_values(torch.float16)
    vh = A.val.float().cpu().numpy().astype(np.float64)
    X, Y = synth.dense(m, k, seed=44), synth.dense(n, k, seed=45)
    out = ops.sddmm_csr(A, up(X, dev), up(Y, dev)).cpu().numpy()
    ref = oracle.sddmm_csr(ptr, idx, vh, X.astype(np.float64), Y.astype(np.float64))
    assert rel_err(out, ref) <= FP32_NORMWISE
    # The same buffers read column-wise (CSC SpMV: y = A^T x here).
    x = synth.dense(m, 1, seed=46)[:, 0].copy()
    y = ops.spmv_csc(A, up(x, dev)).cpu().numpy()
    ref_c = oracle.spmv_csc(n, ptr, idx, vh, x.astype(np.float64))
    assert rel_err(y, ref_c) <= FP32_NORMWISE


# ---- deterministic CSC SpMV --------------------------------------------------------

@pytest.mark.parametrize("dtype,tol", [(np.float64, FP64_REL), (np.float32, FP32_NORMWISE)])
@pytest.mark.parametrize("itype", [(torch.int32, torch.int32), (torch.int64, torch.int64),
                                   (torch.int64, torch.int32)])
def test_spmv_csc_ordered_matches_oracle_and_repeats(dev, dtype, tol, itype):
    # 10k x 10k at density 1e-3 (the FP64 parity workload) and a power-law
    # operand with long columns: same bits on every call, unlike the atomic
    # scatter, and within tolerance of the serial oracle.
    cases = [synth.fsprand(10_000, 10_000, 1e-3, seed=0) + (10_000, 10_000),
             synth.powerlaw_csr(30_000, 20_000, 500_000, cap=15_000, seed=61) + (30_000, 20_000)]
    for ptr, idx, val, n_outer, n_inner in cases:
        A = ops.Compressed(up(ptr, dev, itype[0]), up(idx, dev, itype[1]),
                           up(val.astype(dtype), dev), n_outer, n_inner, 1)
        x = synth.dense(n_outer, 1, seed=62)[:, 0].astype(dtype)
        ref = oracle.spmv_csc(n_inner, ptr, idx, val.astype(np.float64), x.astype(np.float64))
        y1 = ops.spmv_csc_ordered(A, up(x, dev)).cpu().numpy()
        y2 = ops.spmv_csc_ordered(A, up(x, dev)).cpu().numpy()
        assert np.array_equal(y1, y2)
        assert rel_err(y1, ref) <= tol
        ya = ops.spmv_csc(A, up(x, dev)).cpu().numpy()
        assert rel_err(ya, ref) <= tol


def test_spmv_csc_ordered_empty_and_golden(dev, goldens):
    g = goldens["spmv_csc"]
    ptr, idx, val = dense_to_csc(np.array(g["A_dense"]))
    A = operand(ptr, idx, val, 3, 4, dev)
    y = ops.spmv_csc_ordered(A, up(np.array(g["x"]), dev)).cpu().numpy()
    assert y.tolist() == g["y"]           # at most two terms per row: exact
    E = operand(np.ones(6, np.int64), np.zeros(0, np.int64), np.zeros(0), 5, 7, dev)
    assert ops.spmv_csc_ordered(E, up(np.ones(5), dev)).cpu().tolist() == [0.0] * 7
