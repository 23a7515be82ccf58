"""Seeded randomized parity sweep of every entry point against the oracle.

Each case draws a shape (1 to 20,000 on a side, log-uniform; every fourth
dense-operand case 5,000 to 50,000), a row-length
mix (Poisson rows, empty rows and a few long rows that span many tiles), an
index width (Int32 / Int64), an index base (0 / 1) and a value type, and runs
SpMV (CSR and CSC readings), SpMM at an odd and a wide width, SDDMM,
coiteration, SpGEMM, transpose, COO assembly and dropfills on the same buffers.  The bars
are the ones the backend states: FP64 relative 1e-12, FP32 normwise 1e-5,
and bit-exact structure, integer, min-plus, coiteration and conversion
results.
"""
import numpy as np
import pytest
import torch

import oracle
from finchgpu import ops

pytestmark = pytest.mark.gpu

N_CASES = 64


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300)


def random_csr(rng, m, n, base):
    mu = rng.uniform(0.0, min(40.0, n))
    lens = np.minimum(rng.poisson(mu, m), n)
    lens[rng.random(m) < 0.2] = 0                       # empty rows
    if m > 4 and rng.random() < 0.5:                    # a few long rows
        lens[rng.choice(m, 2, replace=False)] = rng.integers(n // 2, n + 1, 2)
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    idx = np.concatenate([np.sort(rng.choice(n, int(l), replace=False)) for l in lens if l] or
                         [np.zeros(0, np.int64)]).astype(np.int64)
    return ptr + base, idx + base


def device(ptr, idx, val, m, n, base, itype, dev):
    up = lambda a, dt=None: (torch.as_tensor(np.ascontiguousarray(a)) if dt is None else
                             torch.as_tensor(np.ascontiguousarray(a)).to(dt)).to(dev)
    return ops.Compressed(up(ptr, itype), up(idx, itype), up(val), m, n, base)


def draw_case(seed, large=False):
    rng = np.random.default_rng(1000 + seed)
    lo = np.log(5_000) if large else 0.0
    hi = np.log(50_000) if large else np.log(20_000)
    m = int(np.exp(rng.uniform(lo, hi)))
    n = int(np.exp(rng.uniform(lo, hi)))
    base = int(rng.integers(0, 2))
    itype = [torch.int32, torch.int64][int(rng.integers(0, 2))]
    vdtype = [np.float32, np.float64][int(rng.integers(0, 2))]
    return rng, m, n, base, itype, vdtype


@pytest.mark.parametrize("seed", range(N_CASES))
def test_dense_operand_kernels(dev, seed):
    # Every fourth case is large (5k-50k a side): many tiles, long carry runs.
    rng, m, n, base, itype, vdtype = draw_case(seed, large=seed % 4 == 0)
    ptr, idx = random_csr(rng, m, n, base)
    val = rng.standard_normal(idx.size).astype(vdtype)
    v64 = val.astype(np.float64)
    A = device(ptr, idx, val, m, n, base, itype, dev)
    tol = 1e-12 if vdtype == np.float64 else 1e-5
    case = f"m={m} n={n} nnz={idx.size} base={base} {itype} {vdtype.__name__}"
    # SpMV, CSR reading (y = A x) and CSC reading (y = A^T x).
    x = rng.standard_normal(n).astype(vdtype)
    y = ops.spmv_csr(A, torch.from_numpy(x).to(dev)).cpu().numpy()
    assert rel_err(y, oracle.spmv_csr(ptr, idx, v64, x.astype(np.float64), base=base)) <= tol, case
    xc = rng.standard_normal(m).astype(vdtype)
    yc = ops.spmv_csc(A, torch.from_numpy(xc).to(dev)).cpu().numpy()
    refc = oracle.spmv_csc(n, ptr, idx, v64, xc.astype(np.float64), base=base)
    assert rel_err(yc, refc) <= tol, case
    # SpMM at an odd width and a wide one.
    for k in (int(rng.choice([1, 3, 5])), int(rng.choice([64, 100, 130]))):
        B = rng.standard_normal((n, k)).astype(vdtype)
        C = ops.spmm_csr(A, torch.from_numpy(B).to(dev)).cpu().numpy()
        ref = oracle.spmm_csr(ptr, idx, v64, B.astype(np.float64), base=base)
        assert rel_err(C, ref) <= tol, f"{case} k={k}"
    # SDDMM.
    k = int(rng.choice([4, 16, 32, 64]))
    X = rng.standard_normal((m, k)).astype(vdtype)
    Y = rng.standard_normal((n, k)).astype(vdtype)
    out = ops.sddmm_csr(A, torch.from_numpy(X).to(dev), torch.from_numpy(Y).to(dev)).cpu().numpy()
    ref = oracle.sddmm_csr(ptr, idx, v64, X.astype(np.float64), Y.astype(np.float64), base=base)
    assert rel_err(out, ref) <= tol, f"{case} sddmm k={k}"


@pytest.mark.parametrize("seed", range(N_CASES))
def test_sparse_output_kernels(dev, seed):
    rng, m, n, base, itype, vdtype = draw_case(seed)
    ptr, idx = random_csr(rng, m, n, base)
    bptr, bidx = random_csr(rng, m, n, base)
    case = f"m={m} n={n} nnz={idx.size}/{bidx.size} base={base} {itype} {vdtype.__name__}"
    # Coiteration: structure and values bit-exact (same operations, same operands).
    val = rng.standard_normal(idx.size).astype(vdtype)
    bval = rng.standard_normal(bidx.size).astype(vdtype)
    A = device(ptr, idx, val, m, n, base, itype, dev)
    Bm = device(bptr, bidx, bval, m, n, base, itype, dev)
    for op in ("mul", "add", "max"):
        C = ops.ewise(op, A, Bm)
        rp, ri, rv = oracle.ewise(op, ptr, idx, val, bptr, bidx, bval, base=base)
        assert np.array_equal(C.ptr.cpu().numpy(), rp), f"{case} {op}"
        assert np.array_equal(C.idx.cpu().numpy(), ri), f"{case} {op}"
        assert np.array_equal(C.val.cpu().numpy(), rv), f"{case} {op}"
    # Transpose: exact.
    T = ops.transpose(A)
    tp, ti, tv = oracle.transpose(m, n, ptr, idx, val, base=base)
    assert np.array_equal(T.ptr.cpu().numpy(), tp) and np.array_equal(T.idx.cpu().numpy(), ti), case
    assert np.array_equal(T.val.cpu().numpy(), tv), case
    # dropfills: a third of the stored values set to the fill.
    dv = val.copy()
    dv[rng.random(dv.size) < 0.33] = 0
    D = ops.dropfills(device(ptr, idx, dv, m, n, base, itype, dev), 0.0)
    dp, di, dvv = oracle.dropfills(m, ptr, idx, dv, 0.0, base=base)
    assert np.array_equal(D.ptr.cpu().numpy(), dp) and np.array_equal(D.idx.cpu().numpy(), di), case
    assert np.array_equal(D.val.cpu().numpy(), dvv), case
    # COO assembly of the same entries plus ~10 % duplicates, shuffled: sorted
    # fibres, duplicates folded left to right in input order, bit-exact.
    rows = np.repeat(np.arange(m, dtype=np.int64), np.diff(ptr)) + base
    dup = rng.random(idx.size) < 0.1
    co = np.concatenate([rows, rows[dup]])
    ci = np.concatenate([idx, idx[dup]])
    cv = np.concatenate([val, rng.standard_normal(int(dup.sum())).astype(vdtype)])
    perm = rng.permutation(co.size)
    co, ci, cv = co[perm], ci[perm], cv[perm]
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a)).to(dt).to(dev)
    K = ops.coo_to_compressed(t(co, itype), t(ci, itype), t(cv, torch.from_numpy(cv).dtype), m, n,
                              base)
    kp, ki, kv = oracle.coo_to_compressed(m, n, co, ci, cv, base=base)
    assert np.array_equal(K.ptr.cpu().numpy(), kp) and np.array_equal(K.idx.cpu().numpy(), ki), case
    assert np.array_equal(K.val.cpu().numpy(), kv), case
    # SpGEMM A (m x n) * B' (n x n2): Int32 plus-times and min-plus, bit-exact.
    n2 = int(rng.integers(1, 3000))
    gp, gi = random_csr(rng, n, n2, base)
    ai32 = rng.integers(1, 10, idx.size).astype(np.int32)
    gi32 = rng.integers(1, 10, gi.size).astype(np.int32)
    G = ops.spgemm("plus_times", device(ptr, idx, ai32, m, n, base, itype, dev),
                   device(gp, gi, gi32, n, n2, base, itype, dev))
    rp, ri, rv = oracle.spgemm("plus_times", m, n2, ptr, idx, ai32, gp, gi, gi32, base=base)
    assert np.array_equal(G.ptr.cpu().numpy(), rp) and np.array_equal(G.idx.cpu().numpy(), ri), case
    assert np.array_equal(G.val.cpu().numpy(), rv), case
    gv = rng.standard_normal(gi.size).astype(np.float32)
    av = val.astype(np.float32)
    G = ops.spgemm("min_plus", device(ptr, idx, av, m, n, base, itype, dev),
                   device(gp, gi, gv, n, n2, base, itype, dev))
    rp, ri, rv = oracle.spgemm("min_plus", m, n2, ptr, idx, av, gp, gi, gv, base=base)
    assert np.array_equal(G.ptr.cpu().numpy(), rp) and np.array_equal(G.idx.cpu().numpy(), ri), case
    assert np.array_equal(G.val.cpu().numpy(), rv), case
