"""GPU parity: format conversion on the device (transpose CSC <-> CSR, COO ->
compressed assembly, dropfills) against the oracle.

Everything here is index and byte work, so the bar is bit-exact: ptr, idx and
the value bits must equal the oracle's, including the left-to-right sum of
duplicate coordinates.  The 1M-row power-law matrix is compared against the
oracle in full, and the transposed-twice round trip must give back the input
buffers exactly.
"""
import numpy as np
import pytest
import torch

import oracle
from finchgpu import _lib, ops, synth

pytestmark = pytest.mark.gpu


def up(a, dev, dtype=None):
    t = torch.as_tensor(np.ascontiguousarray(a))
    return (t.to(dtype) if dtype is not None else t).to(dev)


def host(C):
    return C.ptr.cpu().numpy(), C.idx.cpu().numpy(), C.val.cpu()


def bits(t: torch.Tensor) -> np.ndarray:
    """Value bits as unsigned integers (compares NaN payloads and -0.0 too)."""
    view = {2: torch.int16, 4: torch.int32, 8: torch.int64}[t.element_size()]
    return t.contiguous().view(view).numpy()


def random_csr(m, n, density, seed, dtype=np.float64):
    rng = np.random.default_rng(seed)
    import scipy.sparse as sp
    A = sp.random(m, n, density=density, random_state=rng, format="csr")
    A.sort_indices()
    vals = rng.standard_normal(A.nnz).astype(dtype) if dtype != np.int32 else \
        rng.integers(-9, 10, A.nnz).astype(np.int32)
    return A.indptr.astype(np.int64), A.indices.astype(np.int64), vals


def check_transpose(dev, m, n, ptr0, idx0, val, base, itype, ptype, vdtype=None):
    """ptr0/idx0 are 0-based host arrays; the device operand uses `base`."""
    tv = torch.as_tensor(val) if vdtype is None else torch.as_tensor(val).to(vdtype)
    A = ops.Compressed(up(ptr0 + base, dev, ptype), up(idx0 + base, dev, itype), tv.to(dev), m, n,
                       base)
    print(f"transpose m={m} n={n} nnz={A.nnz} base={base} idx={itype} val={tv.dtype}", flush=True)
    T = ops.transpose(A)
    tp, ti, tval = host(T)
    ep, ei, ev = oracle.transpose(m, n, ptr0 + base, idx0 + base, bits(tv), base)
    assert T.n_outer == n and T.n_inner == m
    assert tp.dtype == np.dtype(str(ptype).replace("torch.", ""))
    assert np.array_equal(tp, ep) and np.array_equal(ti, ei)
    assert np.array_equal(bits(tval), ev)
    return A, T


@pytest.mark.parametrize("base", [1, 0])
@pytest.mark.parametrize("itype,ptype", [(torch.int32, torch.int32), (torch.int64, torch.int64),
                                         (torch.int32, torch.int64)])
def test_transpose_small_matches_oracle(dev, base, itype, ptype):
    ptr, idx, val = random_csr(300, 517, 0.03, 1)
    check_transpose(dev, 300, 517, ptr, idx, val, base, itype, ptype)


@pytest.mark.parametrize("vdtype", [torch.float32, torch.bfloat16, torch.int32])
def test_transpose_value_types(dev, vdtype):
    ptr, idx, val = random_csr(700, 260, 0.05, 2, np.float32 if vdtype != torch.int32 else np.int32)
    check_transpose(dev, 700, 260, ptr, idx, val, 1, torch.int32, torch.int32, vdtype)


def test_transpose_edge_shapes(dev):
    # Empty matrix; a single inner coordinate (no key bits); empty fibres;
    # one dense fibre; stored explicit zeros.
    check_transpose(dev, 5, 7, np.zeros(6, np.int64), np.zeros(0, np.int64), np.zeros(0), 1,
                    torch.int32, torch.int32)
    ptr = np.array([0, 1, 1, 3, 3]); idx = np.array([0, 0, 0]); val = np.array([1.0, 0.0, -2.0])
    check_transpose(dev, 4, 1, ptr, idx, val, 1, torch.int32, torch.int32)
    n = 5000
    ptr = np.array([0, 0, n, n]); idx = np.arange(n); val = np.arange(n, dtype=np.float64) - 7
    check_transpose(dev, 3, n, ptr, idx, val, 1, torch.int64, torch.int64)


def test_transpose_three_digit_keys_and_round_trip(dev):
    # 200k inner coordinates: keys of 18 bits, three radix passes.
    m, n = 120_000, 200_000
    ptr, idx, val = synth.powerlaw_csr(m, n, 1_500_000, cap=20_000, seed=4)
    A, T = check_transpose(dev, m, n, ptr.astype(np.int64) - 1, idx.astype(np.int64) - 1, val, 1,
                           torch.int32, torch.int32)
    B = ops.transpose(T)
    assert torch.equal(B.ptr, A.ptr) and torch.equal(B.idx, A.idx)
    assert np.array_equal(bits(B.val.cpu()), bits(A.val.cpu()))


def test_transpose_powerlaw_1m_matches_oracle(dev):
    # The 1M x 1M Zipf(1.5) operand (16M stored entries) in full.
    m = 1_000_000
    ptr, idx, val = synth.powerlaw_csr(m, m, 16_000_000, seed=0)
    check_transpose(dev, m, m, ptr.astype(np.int64) - 1, idx.astype(np.int64) - 1, val, 1,
                    torch.int32, torch.int32)


# ---- COO -> compressed --------------------------------------------------------------

def check_coo(dev, n_outer, n_inner, outer, inner, val, base=1, itype=torch.int32):
    print(f"coo n_outer={n_outer} n_inner={n_inner} nnz={len(val)} base={base} idx={itype} "
          f"val={np.asarray(val).dtype}", flush=True)
    C = ops.coo_to_compressed(up(outer, dev, itype), up(inner, dev, itype), up(val, dev),
                              n_outer, n_inner, base)
    cp, ci, cv = host(C)
    sh = 1 - base
    ep, ei, ev = oracle.coo_to_compressed(n_outer, n_inner, np.asarray(outer) + sh,
                                          np.asarray(inner) + sh, np.asarray(val), 1)
    assert np.array_equal(cp + sh, ep) and np.array_equal(ci + sh, ei)
    assert np.array_equal(bits(cv), bits(torch.as_tensor(ev)))
    return C


def test_coo_representation_goldens(dev, goldens):
    # Dense{SparseList}_representation.txt: Finch's buffers for each array,
    # rebuilt from its nonzeros listed in a shuffled order.
    rng = np.random.default_rng(0)
    for c in goldens["representation_dense_sparselist"]["cases"]:
        A = np.array(c["dense"], dtype=np.float64)
        m, n = A.shape
        I, J = np.nonzero(A)
        p = rng.permutation(len(I))
        C = check_coo(dev, n, m, J[p] + 1, I[p] + 1, A[I[p], J[p]])
        cp, ci, cv = host(C)
        assert (cp.tolist(), ci.tolist(), cv.tolist()) == (c["ptr"], c["idx"], c["val"]), c["name"]


def test_coo_left_fold_of_duplicates_bit_exact(dev):
    # 1e16 + 1 + -1e16 folded left is 0.0; any other order gives 1 or 2.
    C = check_coo(dev, 2, 3, np.array([2, 1, 2, 2, 1]), np.array([3, 1, 3, 3, 2]),
                  np.array([1e16, 5.0, 1.0, -1e16, 4.0]))
    assert host(C)[2].tolist() == [5.0, 4.0, 0.0]


@pytest.mark.parametrize("base", [1, 0])
@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
@pytest.mark.parametrize("vdtype", [np.float32, np.float64, np.int32])
def test_coo_random_with_duplicates(dev, base, itype, vdtype):
    rng = np.random.default_rng(7)
    m, n, k = 3000, 70_000, 200_000  # 17-bit inner keys, many duplicates per row
    outer = rng.integers(0, m, k) + base
    inner = rng.integers(0, n // 50, k) * 50 + base
    val = (rng.standard_normal(k).astype(vdtype) if vdtype != np.int32
           else rng.integers(-2**30, 2**30, k).astype(np.int32))
    check_coo(dev, m, n, outer, inner, val, base, itype)


def test_coo_empty_and_out_of_shape(dev):
    C = check_coo(dev, 4, 4, np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0))
    assert C.nnz == 0 and C.ptr.tolist() == [1] * 5
    with pytest.raises(_lib.DimensionMismatch):
        ops.coo_to_compressed(up([1, 5], dev, torch.int32), up([1, 1], dev, torch.int32),
                              up([1.0, 2.0], dev), 4, 4, 1)
    with pytest.raises(_lib.DimensionMismatch):
        ops.coo_to_compressed(up([1, 2], dev, torch.int32), up([0, 1], dev, torch.int32),
                              up([1.0, 2.0], dev), 4, 4, 1)


# ---- dropfills ------------------------------------------------------------------------

def test_dropfills_representation_goldens(dev, goldens):
    for c in goldens["representation_dense_sparselist"]["cases"]:
        A = np.array(c["dense"], dtype=np.float64)
        m, n = A.shape
        ptr = np.arange(0, m * n + 1, m) + 1
        idx = np.tile(np.arange(m), n) + 1
        F = ops.Compressed(up(ptr, dev, torch.int64), up(idx, dev, torch.int64),
                           up(A.T.reshape(-1).copy(), dev), n, m, 1)
        cp, ci, cv = host(ops.dropfills(F, 0.0))
        assert (cp.tolist(), ci.tolist(), cv.tolist()) == (c["ptr"], c["idx"], c["val"]), c["name"]


@pytest.mark.parametrize("fill", [0.0, float("nan"), float("inf")])
def test_dropfills_isequal_and_random(dev, fill):
    rng = np.random.default_rng(11)
    ptr, idx, val = random_csr(2000, 1500, 0.02, 3)
    pick = rng.random(val.size)
    val[pick < 0.2] = 0.0
    val[(pick >= 0.2) & (pick < 0.3)] = -0.0
    val[(pick >= 0.3) & (pick < 0.35)] = np.nan
    val[(pick >= 0.35) & (pick < 0.4)] = np.inf
    A = ops.Compressed(up(ptr + 1, dev, torch.int32), up(idx + 1, dev, torch.int32), up(val, dev),
                       2000, 1500, 1)
    for v in (val, val.astype(np.float32)):
        A = A.with_values(up(v, dev))
        cp, ci, cv = host(ops.dropfills(A, fill))
        ep, ei, ev = oracle.dropfills(2000, ptr + 1, idx + 1, v, fill)
        assert np.array_equal(cp, ep) and np.array_equal(ci, ei)
        assert np.array_equal(bits(cv), bits(torch.as_tensor(ev)))
