"""GPU parity: element-wise coiteration (.*, +, max.) and Gustavson SpGEMM
(plus-times Int32/FP32/FP64, min-plus FP32) against the oracle.

Output structure (ptr, idx) and integer / min-plus values must match the
oracle bit for bit; FP plus-times values within FP64 1e-12 / FP32 1e-5.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from conftest import read_mtx
from finchgpu import ops, synth

pytestmark = pytest.mark.gpu


def up(a, dev, dtype=None):
    t = torch.as_tensor(np.ascontiguousarray(a))
    return (t.to(dtype) if dtype is not None else t).to(dev)


def operand(ptr, idx, val, n_outer, n_inner, dev, itype=torch.int32, ptype=None, base=1):
    ptype = ptype or itype
    sh = 1 - base
    return ops.Compressed(up(np.asarray(ptr) - sh, dev, ptype), up(np.asarray(idx) - sh, dev, itype),
                          up(val, dev), n_outer, n_inner, base)


def host(C: ops.Compressed):
    return C.ptr.cpu().numpy(), C.idx.cpu().numpy(), C.val.cpu().numpy()


# ---- coiteration ---------------------------------------------------------------------

def test_merge_golden(dev, goldens):
    g = goldens["merge_add"]
    x, y = g["x"], g["y"]
    X = operand(x["ptr"], x["idx"], np.array(x["val"]), 1, g["shape"], dev, torch.int64)
    Y = operand(y["ptr"], y["idx"], np.array(y["val"]), 1, g["shape"], dev, torch.int64)
    cptr, cidx, cval = host(ops.ewise("add", X, Y))
    z = g["z"]["twofinger"]
    assert cptr.tolist() == z["ptr"] and cidx.tolist() == z["idx"] and cval.tolist() == z["val"]


def _pair(seed, m, n, la, lb, dtype):
    rng = np.random.default_rng(seed)
    a = synth.poisson_csr(m, n, la, rng=rng, dtype=dtype)
    b = synth.poisson_csr(m, n, lb, rng=rng, dtype=dtype)
    return a, b


@pytest.mark.parametrize("op", ["mul", "add", "max"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32])
@pytest.mark.parametrize("shape", [(1000, 500, 3.0, 9.0), (300, 20000, 60.0, 150.0)])
@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
def test_ewise_bit_exact(dev, op, dtype, shape, itype):
    m, n, la, lb = shape
    (ap, ai, av), (bp, bi, bv) = _pair(7, m, n, la, lb, np.float64)
    if dtype == np.int32:
        av = np.rint(av * 10).astype(np.int32)
        bv = np.rint(bv * 10).astype(np.int32)
    av, bv = av.astype(dtype), bv.astype(dtype)
    A = operand(ap, ai, av, m, n, dev, itype)
    B = operand(bp, bi, bv, m, n, dev, itype)
    cptr, cidx, cval = host(ops.ewise(op, A, B))
    rptr, ridx, rval = oracle.ewise(op, ap, ai, av, bp, bi, bv)
    assert np.array_equal(cptr, rptr)
    assert np.array_equal(cidx, ridx)
    assert np.array_equal(cval, rval)      # same operations, same operands: bit-exact


def test_ewise_edges(dev):
    # Empty fibres on either side, one empty operand, identical patterns.
    m, n = 50, 40
    (ap, ai, av), _ = _pair(3, m, n, 2.0, 2.0, np.float32)
    ep = np.ones(m + 1, np.int64)
    E = operand(ep, np.zeros(0, np.int64), np.zeros(0, np.float32), m, n, dev)
    A = operand(ap, ai, av, m, n, dev)
    for op in ("mul", "add", "max"):
        for (X, xs), (Y, ys) in [((A, (ap, ai, av)), (E, (ep, [], np.zeros(0, np.float32)))),
                                 ((E, (ep, [], np.zeros(0, np.float32))), (A, (ap, ai, av))),
                                 ((A, (ap, ai, av)), (A, (ap, ai, av)))]:
            cptr, cidx, cval = host(ops.ewise(op, X, Y))
            rptr, ridx, rval = oracle.ewise(op, *xs, *ys)
            assert np.array_equal(cptr, rptr) and np.array_equal(cidx, ridx)
            assert np.array_equal(cval, rval)


_SPECIALS = [1.0, -1.0, 0.0, -0.0, float("nan"), float("inf"), -float("inf"), 2.5, -2.5]


def _bits(v):
    return v.view(np.uint32 if v.dtype.itemsize == 4 else np.uint64)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
def test_ewise_max_special_values_bit_exact(dev, dtype, itype):
    # Julia's max on NaN, +-0 and +-Inf at both-sided and one-sided positions,
    # including signalling and negative NaN payloads: bits must match the oracle.
    vals = np.array(_SPECIALS, dtype)
    if dtype == np.float32:
        sn = np.array([0x7FA00001, 0xFFC00002], np.uint32).view(np.float32)
        vals = np.concatenate([vals, sn])
    k = vals.size
    # Row 0: every pair (x, y) both stored.  Row 1: A-only coordinates (odd)
    # and B-only coordinates (even), each value once on each side.
    xs, ys = np.meshgrid(vals, vals, indexing="ij")
    xs, ys = xs.ravel(), ys.ravel()
    n = max(xs.size, 2 * k)
    ai = np.concatenate([np.arange(1, xs.size + 1), np.arange(1, 2 * k, 2)])
    bi = np.concatenate([np.arange(1, ys.size + 1), np.arange(2, 2 * k + 1, 2)])
    ap = np.array([1, xs.size + 1, xs.size + k + 1])
    bp = np.array([1, ys.size + 1, ys.size + k + 1])
    av = np.concatenate([xs, vals]).astype(dtype)
    bv = np.concatenate([ys, vals]).astype(dtype)
    A = operand(ap, ai, av, 2, n, dev, itype)
    B = operand(bp, bi, bv, 2, n, dev, itype)
    cptr, cidx, cval = host(ops.ewise("max", A, B))
    rptr, ridx, rval = oracle.ewise("max", ap, ai, av, bp, bi, bv)
    assert np.array_equal(cptr, rptr) and np.array_equal(cidx, ridx)
    assert np.array_equal(_bits(cval), _bits(rval))
    # Spot checks of Julia's definition on the device result itself.
    both = cval[:xs.size]
    j = np.flatnonzero((xs == 0) & (ys == 0) & (np.signbit(xs) != np.signbit(ys)))
    assert j.size == 2 and not np.signbit(both[j]).any()          # max(-0.0, 0.0) == 0.0
    assert np.isnan(both[np.isnan(xs) | np.isnan(ys)]).all()       # NaN from either side


@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
def test_ewise_int32_wraps_bit_exact(dev, itype):
    big = np.array([2**31 - 1, 65536, 46341, -(2**31), 7], np.int32)
    two = np.array([1, 65536, 46341, -1, -9], np.int32)
    idx = np.arange(1, 6)
    A = operand([1, 6], idx, big, 1, 5, dev, itype)
    B = operand([1, 6], idx, two, 1, 5, dev, itype)
    for op in ("add", "mul", "max"):
        _, _, got = host(ops.ewise(op, A, B))
        _, _, want = oracle.ewise(op, [1, 6], idx, big, [1, 6], idx, two)
        assert got.tolist() == want.tolist(), op


def _rows_csr(lengths, n, rng, dtype):
    """1-based CSR with the given row lengths, distinct sorted columns."""
    ptr = np.concatenate([[1], 1 + np.cumsum(lengths)]).astype(np.int64)
    idx = np.concatenate([np.sort(rng.choice(n, size=l, replace=False)) + 1 if l else
                          np.zeros(0, np.int64) for l in lengths]).astype(np.int64)
    val = rng.standard_normal(idx.size).astype(dtype)
    return ptr, idx, val


@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
def test_ewise_long_fibres_beside_short(dev, itype):
    # Short rows pick 8-lane groups whose LDS slots hold 64 coordinates per
    # side; a few rows of hundreds of entries exceed them and are searched in
    # global memory.  Both kinds must give the oracle's output bit for bit.
    rng = np.random.default_rng(11)
    m, n = 400, 6000
    la = rng.poisson(3, m)
    lb = rng.poisson(5, m)
    la[[7, 100, 399]] = [500, 64, 65]
    lb[[7, 100, 250]] = [900, 65, 64]
    (ap, ai, av) = _rows_csr(la, n, rng, np.float32)
    (bp, bi, bv) = _rows_csr(lb, n, rng, np.float32)
    # Row 7 of B shares 100 coordinates with row 7 of A, so the long rows'
    # intersection is not empty.
    a7 = ai[ap[7] - 1: ap[8] - 1]
    others = np.setdiff1d(np.arange(1, n + 1), a7)
    bi[bp[7] - 1: bp[8] - 1] = np.sort(np.concatenate(
        [a7[:100], rng.choice(others, size=lb[7] - 100, replace=False)]))
    A = operand(ap, ai, av, m, n, dev, itype)
    B = operand(bp, bi, bv, m, n, dev, itype)
    for op in ("mul", "add", "max"):
        cptr, cidx, cval = host(ops.ewise(op, A, B))
        rptr, ridx, rval = oracle.ewise(op, ap, ai, av, bp, bi, bv)
        assert np.array_equal(cptr, rptr) and np.array_equal(cidx, ridx)
        assert np.array_equal(cval, rval)
    assert oracle.ewise("mul", ap, ai, av, bp, bi, bv)[1].size >= 100


def test_ewise_200k_poisson(dev):
    # 200k x 200k, Poisson(8) and Poisson(24) rows, FP32 values.
    m = n = 200_000
    (ap, ai, av), (bp, bi, bv) = _pair(0, m, n, 8.0, 24.0, np.float32)
    A = operand(ap, ai, av, m, n, dev)
    B = operand(bp, bi, bv, m, n, dev)
    for op in ("mul", "add", "max"):
        cptr, cidx, cval = host(ops.ewise(op, A, B))
        rptr, ridx, rval = oracle.ewise(op, ap, ai, av, bp, bi, bv)
        assert np.array_equal(cptr, rptr) and np.array_equal(cidx, ridx)
        assert np.array_equal(cval, rval)


# ---- SpGEMM ------------------------------------------------------------------------------

def _gemm_check(dev, semiring, ap, ai, av, m, l, bp, bi, bv, n, itype=torch.int32, exact=True,
                tol=0.0):
    A = operand(ap, ai, av, m, l, dev, itype)
    B = operand(bp, bi, bv, l, n, dev, itype)
    cptr, cidx, cval = host(ops.spgemm(semiring, A, B))
    rptr, ridx, rval = oracle.spgemm(semiring, m, n, ap, ai, av, bp, bi, bv)
    assert np.array_equal(cptr, rptr)
    assert np.array_equal(cidx, ridx)
    if exact:
        assert np.array_equal(cval, rval)
    else:
        err = np.linalg.norm(cval.astype(np.float64) - rval) / max(np.linalg.norm(rval), 1e-300)
        assert err <= tol


@pytest.mark.parametrize("name", ["west0132.mtx", "lpi_itest6.mtx", "lp_blend.mtx", "Trec4.mtx"])
def test_spgemm_finch_fixtures(dev, name):
    # kernel_tests.jl:89-130: Gustavson B == A*A on Finch's Matrix Market data.
    S = read_mtx(name)
    T = S if S.shape[0] == S.shape[1] else S.T.tocsr()
    if S.dtype.kind == "i":
        S, T = S.astype(np.int32), T.astype(np.int32)
    S.sort_indices()
    T.sort_indices()
    # Same k order as the reference's update sequence: FP64 sums bit-exact.
    _gemm_check(dev, "plus_times", S.indptr + 1, S.indices + 1, S.data, S.shape[0], S.shape[1],
                T.indptr + 1, T.indices + 1, T.data, T.shape[1], itype=torch.int64)


@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
def test_spgemm_200k_poisson_int32_and_min_plus(dev, itype):
    # 200k x 200k, Poisson(8) x Poisson(24) rows; Int32 U{1..9}
    # plus-times and FP32 min-plus, both bit-exact.
    m = 200_000
    rng = np.random.default_rng(0)
    ap, ai, av = synth.poisson_csr(m, m, 8.0, rng=rng, dtype=np.int32, values="int19")
    bp, bi, bv = synth.poisson_csr(m, m, 24.0, rng=rng, dtype=np.int32, values="int19")
    _gemm_check(dev, "plus_times", ap, ai, av, m, m, bp, bi, bv, m, itype)
    af = rng.standard_normal(ai.size).astype(np.float32)
    bf = rng.standard_normal(bi.size).astype(np.float32)
    _gemm_check(dev, "min_plus", ap, ai, af, m, m, bp, bi, bf, m, itype)


def _skewed(seed, dtype, values):
    """Rows whose product counts land in every bin: wave (<=256), block
    (<=1536) and the dense-window path (longer)."""
    rng = np.random.default_rng(seed)
    l, n = 3000, 20_000
    lens_a = np.array([0, 1, 5, 20, 40, 150, 700, 2500, 3, 0])
    m = lens_a.size
    ap = np.concatenate([[0], np.cumsum(lens_a)]) + 1
    ai = np.concatenate([np.sort(rng.choice(l, k, replace=False)) for k in lens_a if k]) + 1
    bp, bi, bv = synth.poisson_csr(l, n, 6.0, rng=rng, dtype=dtype, values=values)
    av = (rng.integers(1, 10, ai.size) if values == "int19" else rng.standard_normal(ai.size))
    return ap, ai, av.astype(dtype), m, l, bp, bi, bv, n


def test_spgemm_all_bins_exact(dev):
    ap, ai, av, m, l, bp, bi, bv, n = _skewed(1, np.int32, "int19")
    _gemm_check(dev, "plus_times", ap, ai, av, m, l, bp, bi, bv, n)
    ap, ai, av, m, l, bp, bi, bv, n = _skewed(2, np.float32, "randn")
    _gemm_check(dev, "min_plus", ap, ai, av, m, l, bp, bi, bv, n)
    ap, ai, av, m, l, bp, bi, bv, n = _skewed(3, np.float64, "randn")
    _gemm_check(dev, "min_plus", ap, ai, av, m, l, bp, bi, bv, n)


def test_spgemm_fp_plus_times_all_bins(dev):
    # Every bin folds each column's products in the reference's order (A
    # positions ascending, then B positions), the dense window included, with
    # each product rounded before its add: FP sums are bit-exact.
    ap, ai, av, m, l, bp, bi, bv, n = _skewed(4, np.float32, "randn")
    _gemm_check(dev, "plus_times", ap, ai, av, m, l, bp, bi, bv, n)
    ap, ai, av, m, l, bp, bi, bv, n = _skewed(5, np.float64, "randn")
    _gemm_check(dev, "plus_times", ap, ai, av, m, l, bp, bi, bv, n)


@pytest.mark.parametrize("n", [(1 << 22) + 3, (1 << 23) + 7])
def test_spgemm_wide_columns_merge_paths(dev, n):
    # Column counts past the register sorts' keys (31 bits hold a column and
    # an 8- or 9-bit product index): above 2^22 the 257..512-product rows go
    # to the block merge, above 2^23 the short rows go to the wave merge.  FP
    # plus-times folds stay in the reference's order there too.
    rng = np.random.default_rng(40 + n % 11)
    l = 2000
    lens_a = np.concatenate([rng.integers(1, 30, 200), rng.integers(60, 80, 30),
                             rng.integers(150, 200, 4), [0, 0]])
    m = lens_a.size
    ap = np.concatenate([[0], np.cumsum(lens_a)]) + 1
    ai = np.concatenate([np.sort(rng.choice(l, k, replace=False)) for k in lens_a if k]) + 1
    prods = None
    for dtype, semiring in ((np.float32, "plus_times"), (np.float32, "min_plus")):
        bp, bi, bv = synth.poisson_csr(l, n, 6.0, rng=rng, dtype=dtype, values="randn")
        if prods is None:
            blen = np.diff(bp)
            prods = np.array([blen[ai[ap[r] - 1: ap[r + 1] - 1] - 1].sum() for r in range(m)])
            assert (prods <= 256).any() and ((prods > 256) & (prods <= 512)).any()
            assert ((prods > 512) & (prods <= 1536)).any() and prods.max() <= 1536
        av = rng.standard_normal(ai.size).astype(dtype)
        _gemm_check(dev, semiring, ap, ai, av, m, l, bp, bi, bv, n)


def test_spgemm_sorted_block_layout_edges(dev):
    # Rows of at most 256 products whose lists all fit 32-product blocks (at
    # most 8 lists) are sorted from one block per list; a list of 33, or a
    # ninth list, takes the full network.  A row's key count is the larger of
    # its products and its lists.  FP plus-times is order-sensitive, so a
    # wrong fold order or a lost product changes the bits.
    rng = np.random.default_rng(77)
    l, n = 200, 5000
    lens_b = np.array([32] * 8 + [33, 1, 0, 31, 16, 32] + [5] * 26 + [1, 0] * 80)
    bp = np.concatenate([[0], np.cumsum(lens_b)]) + 1
    bi = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for k in lens_b if k]) + 1
    rows_a = [list(range(8)),                 # 8 full blocks: 256 products
              [8, 9, 11, 12],                 # a 33-product list among short ones
              [9, 10, 11, 12, 14, 15, 16, 17, 18],  # nine lists, 73 products
              [10, 11, 12],                   # an empty list among short ones
              [13, 14, 15, 16, 17, 18, 19, 20],
              list(range(14, 40)),            # 26 lists of 5: past 8 lists
              list(range(8)) + [9],           # 257 products: the 512-key sort
              [0, 1, 2, 3],                   # 4 full blocks of the 8: 128 products
              [0, 1, 2, 8],                   # 129 products, a 33-product list
              [9, 10, 11, 12, 14],            # five lists, 53 products
              [8, 9],                         # a 33-product list, 34 products
              [9, 11, 12, 13],                # four blocks, 80 products
              [10],                           # one empty list: no products
              list(range(40, 168)),           # 128 lists, 64 products
              list(range(40, 200))]           # 160 lists, 80 products
    ap = np.concatenate([[0], np.cumsum([len(r) for r in rows_a])]) + 1
    ai = np.concatenate([np.array(r) for r in rows_a]) + 1
    for dtype, semiring in ((np.float32, "plus_times"), (np.float64, "plus_times"),
                            (np.float32, "min_plus")):
        av = rng.standard_normal(ai.size).astype(dtype)
        bv = rng.standard_normal(bi.size).astype(dtype)
        _gemm_check(dev, semiring, ap, ai, av, len(rows_a), l, bp, bi, bv, n)


def test_spgemm_empty(dev):
    ep = np.ones(6, np.int64)
    e = np.zeros(0, np.int64)
    _gemm_check(dev, "plus_times", ep, e, np.zeros(0, np.float32), 5, 4, np.ones(5, np.int64), e,
                np.zeros(0, np.float32), 7)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_spgemm_min_plus_julia_min(dev, dtype):
    # Julia's `min`: -0.0 is below 0.0 and NaN wins, in every product order
    # (wave rows here; the all-bins test covers the block and window bins).
    B_ptr, B_idx = np.array([1, 2, 3]), np.array([1, 1])
    cases = [((-0.0, 0.0), (-0.0, 0.0)), ((0.0, -0.0), (0.0, -0.0)),
             ((1.0, np.nan), (0.0, 0.0)), ((np.nan, 1.0), (0.0, 0.0))]
    for a, b in cases:
        A = operand([1, 3], [1, 2], np.array(a, dtype), 1, 2, dev)
        B = operand(B_ptr, B_idx, np.array(b, dtype), 2, 1, dev)
        C = ops.spgemm("min_plus", A, B)
        p, i, v = oracle.spgemm("min_plus", 1, 1, np.array([1, 3]), np.array([1, 2]),
                                np.array(a, dtype), B_ptr, B_idx, np.array(b, dtype))
        got = C.val.cpu().numpy()
        assert C.ptr.cpu().tolist() == p.tolist() and C.idx.cpu().tolist() == i.tolist()
        if np.isnan(v[0]):
            assert np.isnan(got[0])
        else:
            assert got[0] == v[0] and np.signbit(got[0]) == np.signbit(v[0])


@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
def test_symbolic_ptr_into_unaligned_buffer(dev, itype):
    # The symbolic passes write ptr with 16-byte vector stores where the
    # caller's buffer allows; a ptr that starts one element into an allocation
    # (a view, as Julia's or torch's slicing can hand over) takes the scalar
    # path.  Both must give the same ptr and nnz.
    from finchgpu import _lib as L
    from finchgpu.ops import _ptr, _stream, context
    import ctypes
    m = n = 3000
    (ap, ai, av), (bp, bi, bv) = _pair(9, m, n, 5.0, 9.0, np.float32)
    A = operand(ap, ai, av, m, n, dev, itype)
    B = operand(bp, bi, bv, m, n, dev, itype)
    lib = L.load()
    ctx = context(dev).handle
    da, db = A.desc(), B.desc()
    for call, args in (("fgx_ewise_symbolic", (L.OP_ADD,)), ("fgx_spgemm_symbolic", (L.PLUS_TIMES,))):
        got = []
        for offset in (0, 1):
            big = torch.full((m + 1 + offset,), -7, dtype=itype, device=dev)
            cptr = big[offset:]
            nnz = ctypes.c_int64()
            L.check(getattr(lib, call)(ctx, *args, ctypes.byref(da), ctypes.byref(db), _ptr(cptr),
                                       ctypes.byref(nnz), _stream(dev)), call)
            torch.cuda.synchronize()
            if offset:
                assert int(big[0]) == -7                  # nothing written before the view
            got.append((cptr.cpu().numpy().copy(), nnz.value))
        assert np.array_equal(got[0][0], got[1][0]) and got[0][1] == got[1][1]
        assert got[0][0][0] == 1 and got[0][0][-1] == 1 + got[0][1]


def _bin_edges(seed, dtype, values):
    """A rows built on B rows of known lengths (1, 2, 0 and 4 entries) so that
    their product counts sit on every bin edge: 256 / 257 (wave sort of 4
    keys per lane / of 8), 512 / 513 (of 8 / block merge), 1536 / 1537 (block
    merge / dense window), and rows whose lists point at empty B rows, where
    the list count (300, 512, 513) picks the bin with no products at all."""
    rng = np.random.default_rng(seed)
    n = 20_000
    blen = np.concatenate([np.full(1000, 1), np.full(1000, 2), np.zeros(1000, int), np.full(1000, 4)])
    l = blen.size
    bp = np.concatenate([[0], np.cumsum(blen)]) + 1
    bi = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for k in blen if k]) + 1
    one, two, empty, four = (np.arange(1000) + 1000 * g for g in range(4))
    pick = lambda pool, k: list(rng.choice(pool, k, replace=False))
    rows = [pick(one, 256), pick(one, 257), pick(two, 256), pick(two, 256) + pick(one, 1),
            pick(empty, 300), pick(empty, 512), pick(empty, 513), pick(four, 384),
            pick(four, 384) + pick(one, 1), pick(empty, 200) + pick(one, 256), [], pick(two, 3)]
    ap = np.concatenate([[0], np.cumsum([len(r) for r in rows])]) + 1
    ai = np.concatenate([np.sort(np.array(r, dtype=np.int64)) for r in rows if r]) + 1
    draw = (lambda k: rng.integers(1, 10, k)) if values == "int19" else rng.standard_normal
    return (ap, ai, draw(ai.size).astype(dtype), len(rows), l, bp, bi,
            draw(bi.size).astype(dtype), n)


@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
def test_spgemm_bin_edges_exact(dev, itype):
    ap, ai, av, m, l, bp, bi, bv, n = _bin_edges(5, np.int32, "int19")
    prods = [int(np.diff(bp)[ai[ap[r] - 1: ap[r + 1] - 1] - 1].sum()) for r in range(m)]
    assert prods[:4] == [256, 257, 512, 513] and prods[7:9] == [1536, 1537]
    _gemm_check(dev, "plus_times", ap, ai, av, m, l, bp, bi, bv, n, itype=itype)
    ap, ai, av, m, l, bp, bi, bv, n = _bin_edges(6, np.float32, "randn")
    _gemm_check(dev, "min_plus", ap, ai, av, m, l, bp, bi, bv, n, itype=itype)
