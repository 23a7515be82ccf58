"""Pins the CPU oracle (oracle/finch_oracle.c) before it is trusted as the checker.

Each oracle kernel is checked against Finch's own golden vectors where the
reference holds one, and otherwise against independent computations: scipy
sparse products, brute-force dense loops and closed forms.  The checks are
chosen so that plausible mistakes (a transposed operand, a dropped term, a
wrong fill, a reordered sum, a lost index base) fail at least one of them.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from conftest import dense_to_csc, read_mtx


# ---- SpMV -------------------------------------------------------------------

def test_spmv_csc_golden_bit_exact(goldens):
    g = goldens["spmv_csc"]
    ptr, idx, val = dense_to_csc(np.array(g["A_dense"]))
    y = oracle.spmv_csc(len(g["A_dense"]), ptr, idx, val, np.array(g["x"]))
    # typical_spmv_csc.txt:92 — bit-exact, including 15.400000000000002
    # (x*a rounded before the add; an FMA would give a different last bit).
    assert y.tolist() == g["y"]


def test_spmv_row_dot_on_csc_buffers_is_transpose(goldens):
    # parallel_spmv_atomics.txt:117 — reading CSC buffers fibre-wise with a
    # dot product gives A^T x = [4, 6]; the scatter form gives A x = [3, 7].
    g = goldens["spmv_transpose"]
    A = np.array(g["A_dense"])
    ptr, idx, val = dense_to_csc(A)
    assert oracle.spmv_csr(ptr, idx, val, np.array(g["x"])).tolist() == g["y"]
    assert oracle.spmv_csc(2, ptr, idx, val, np.array(g["x"])).tolist() == (A @ g["x"]).tolist()
    assert oracle.spmv_csc(2, ptr, idx, val, np.array(g["x"])).tolist() != g["y"]


def test_zero_based_buffers(goldens):
    # issue_tests.jl:388-404 — PlusOneVector views of 0-based buffers.
    g = goldens["zero_based"]
    A = np.array(g["A_dense"])
    ptr, idx, val = np.array(g["ptr"]), np.array(g["idx"]), np.array(g["val"])
    for j in range(g["n"]):
        e = np.eye(g["n"])[j]
        np.testing.assert_array_equal(oracle.spmv_csc(g["m"], ptr, idx, val, e, base=0), A[:, j])
    ptr1, idx1, val1 = dense_to_csc(A)
    np.testing.assert_array_equal(ptr + 1, ptr1)
    np.testing.assert_array_equal(idx + 1, idx1)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 1e-5)])
@pytest.mark.parametrize("nthreads", [1, 3])
def test_spmv_csr_vs_scipy(dtype, tol, nthreads):
    rng = np.random.default_rng(1)
    S = sp.random(300, 200, density=0.05, format="csr", random_state=rng, dtype=np.float64)
    S.data = S.data.astype(dtype)
    x = rng.standard_normal(200).astype(dtype)
    y = oracle.spmv_csr(S.indptr + 1, S.indices + 1, S.data, x, nthreads=nthreads)
    ref = S.astype(np.float64) @ x.astype(np.float64)
    assert np.linalg.norm(y - ref) <= tol * np.linalg.norm(ref)
    # Threaded static blocks give the same bits as serial (same per-row order).
    assert np.array_equal(y, oracle.spmv_csr(S.indptr + 1, S.indices + 1, S.data, x))


def test_spmv_csc_vs_scipy_and_empty_columns():
    rng = np.random.default_rng(2)
    S = sp.random(150, 90, density=0.03, format="csc", random_state=rng)
    S[:, 5] = 0
    S.eliminate_zeros()
    x = rng.standard_normal(90)
    y = oracle.spmv_csc(150, S.indptr + 1, S.indices + 1, S.data, x)
    np.testing.assert_allclose(y, S @ x, rtol=1e-13, atol=1e-13)


# ---- SpMM / SDDMM -------------------------------------------------------------

@pytest.mark.parametrize("nthreads", [1, 4])
def test_spmm_vs_scipy(nthreads):
    rng = np.random.default_rng(3)
    S = sp.random(120, 80, density=0.1, format="csr", random_state=rng)
    B = rng.standard_normal((80, 7))
    C = oracle.spmm_csr(S.indptr + 1, S.indices + 1, S.data, B, nthreads=nthreads)
    np.testing.assert_allclose(C, S @ B, rtol=1e-12, atol=1e-12)
    # Column c of SpMM equals SpMV with B[:, c] (linearity check on the oracle itself).
    for c in range(7):
        np.testing.assert_array_equal(C[:, c], oracle.spmv_csr(S.indptr + 1, S.indices + 1,
                                                               S.data, B[:, c].copy()))


def test_sddmm_closed_form():
    rng = np.random.default_rng(4)
    S = sp.random(60, 50, density=0.1, format="csr", random_state=rng)
    X = rng.standard_normal((60, 9))
    Y = rng.standard_normal((50, 9))
    out = oracle.sddmm_csr(S.indptr + 1, S.indices + 1, S.data, X, Y)
    dense = (X @ Y.T) * S.toarray()
    np.testing.assert_allclose(out, dense[S.nonzero()], rtol=1e-12, atol=1e-12)
    # Dropping the A scale, or swapping the roles of X and Y, must fail.
    assert not np.allclose(out, (X @ Y.T)[S.nonzero()])
    assert not np.allclose(out, (S.toarray() * (X @ X[:50].T))[S.nonzero()])


# ---- coiteration ---------------------------------------------------------------

def test_merge_goldens(goldens):
    g = goldens["merge_add"]
    x, y = g["x"], g["y"]
    cptr, cidx, cval = oracle.ewise("add", x["ptr"], x["idx"], np.array(x["val"]),
                                    y["ptr"], y["idx"], np.array(y["val"]))
    for kind, z in g["z"].items():
        assert cptr.tolist() == z["ptr"], kind
        assert cidx.tolist() == z["idx"], kind
        assert cval.tolist() == z["val"], kind


def _set_algebra(op, S, T):
    """Expected coordinate set and values from dense algebra on stored patterns."""
    A, B = S.toarray(), T.toarray()
    sa, sb = S.toarray() != 0, T.toarray() != 0
    mask = (sa & sb) if op == "mul" else (sa | sb)
    val = {"mul": A * B, "add": A + B, "max": np.maximum(A, B)}[op]
    return mask, val


@pytest.mark.parametrize("op", ["mul", "add", "max"])
def test_ewise_set_algebra(op):
    rng = np.random.default_rng(5)
    S = sp.random(40, 30, density=0.2, format="csr", random_state=rng)
    T = sp.random(40, 30, density=0.3, format="csr", random_state=rng)
    S.data -= 0.5
    T.data -= 0.5
    cptr, cidx, cval = oracle.ewise(op, S.indptr + 1, S.indices + 1, S.data,
                                    T.indptr + 1, T.indices + 1, T.data)
    mask, val = _set_algebra(op, S, T)
    rows, cols = np.nonzero(mask)
    assert cptr[-1] - 1 == mask.sum()
    np.testing.assert_array_equal(np.diff(cptr), mask.sum(axis=1))
    np.testing.assert_array_equal(cidx - 1, cols)
    np.testing.assert_array_equal(cval, val[rows, cols])


def test_ewise_max_stores_zero_from_negative_one_sided():
    # max.(A, B) over the union: a one-sided -2.0 becomes max(-2, 0) = 0, stored.
    cptr, cidx, cval = oracle.ewise("max", [1, 2], [3], np.array([-2.0]), [1, 1], [],
                                    np.array([]))
    assert cptr.tolist() == [1, 2] and cidx.tolist() == [3] and cval.tolist() == [0.0]


# Julia's Base.max on floats (base/math.jl): a NaN on either side gives NaN,
# -0.0 orders below 0.0, otherwise the larger value.  Expected values are
# written from that definition, not computed by the oracle's formula.
_NAN = float("nan")
_INF = float("inf")
JULIA_MAX_TABLE = [
    # (x, y, max(x, y))
    (1.0, 2.0, 2.0), (2.0, 1.0, 2.0), (-3.0, -5.0, -3.0),
    (1.0, _NAN, _NAN), (_NAN, 1.0, _NAN), (_NAN, _NAN, _NAN),
    (_INF, _NAN, _NAN), (_NAN, -_INF, _NAN),
    (-0.0, 0.0, 0.0), (0.0, -0.0, 0.0), (-0.0, -0.0, -0.0), (0.0, 0.0, 0.0),
    (-_INF, -0.0, -0.0), (-_INF, _INF, _INF), (_INF, -_INF, _INF), (-_INF, -_INF, -_INF),
    (-1.0, -0.0, -0.0), (-0.0, -1.0, -0.0),
]


def _same(a, b):
    """Equal including the sign of zero; NaN equals NaN."""
    if np.isnan(a) or np.isnan(b):
        return bool(np.isnan(a) and np.isnan(b))
    return a == b and np.signbit(a) == np.signbit(b)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ewise_max_is_julia_max_both_sided(dtype):
    # Both sides stored at every coordinate: C[i] = max(A[i], B[i]).
    k = len(JULIA_MAX_TABLE)
    x = np.array([t[0] for t in JULIA_MAX_TABLE], dtype)
    y = np.array([t[1] for t in JULIA_MAX_TABLE], dtype)
    idx = np.arange(1, k + 1)
    cptr, cidx, cval = oracle.ewise("max", [1, k + 1], idx, x, [1, k + 1], idx, y)
    assert cptr.tolist() == [1, k + 1] and cidx.tolist() == idx.tolist()
    for (a, b, want), got in zip(JULIA_MAX_TABLE, cval):
        assert _same(got, dtype(want)), (a, b, got, want)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ewise_max_is_julia_max_one_sided(dtype):
    # A missing side reads as the fill 0: an A-only entry a gives max(a, 0)
    # and a B-only entry b gives max(0, b).
    vals = [-0.0, 0.0, _NAN, -_INF, _INF, -2.0, 3.0]
    want = [0.0, 0.0, _NAN, 0.0, _INF, 0.0, 3.0]
    k = len(vals)
    a_idx = np.arange(1, 2 * k, 2)          # odd coordinates: A only
    b_idx = np.arange(2, 2 * k + 1, 2)      # even coordinates: B only
    v = np.array(vals, dtype)
    cptr, cidx, cval = oracle.ewise("max", [1, k + 1], a_idx, v, [1, k + 1], b_idx, v)
    assert cidx.tolist() == list(range(1, 2 * k + 1))
    for j, w in enumerate(want):
        assert _same(cval[2 * j], dtype(w)), ("A-only", vals[j], cval[2 * j])
        assert _same(cval[2 * j + 1], dtype(w)), ("B-only", vals[j], cval[2 * j + 1])


def test_ewise_max_nan_bits_follow_x_minus_y():
    # Julia returns `x - y` when a NaN is present; on x86-64 that is the first
    # NaN operand with its quiet bit set.  Signalling payloads come back quiet.
    sn = np.array([0x7FA00001, 0xFFC00002], np.uint32).view(np.float32)   # sNaN, -qNaN
    x = np.array([sn[0], 1.0, sn[1], sn[0]], np.float32)
    y = np.array([1.0, sn[1], sn[0], np.float32(2.0)], np.float32)
    idx = np.arange(1, 5)
    _, _, c = oracle.ewise("max", [1, 5], idx, x, [1, 5], idx, y)
    got = c.view(np.uint32).tolist()
    assert got == [0x7FE00001, 0xFFC00002, 0xFFC00002, 0x7FE00001]


def test_ewise_int32_wraps():
    # Julia's Int32 `+` and `*` wrap modulo 2^32.
    big = np.array([2**31 - 1, 65536, 46341, -(2**31)], np.int32)
    two = np.array([1, 65536, 46341, -1], np.int32)
    idx = np.arange(1, 5)
    _, _, s = oracle.ewise("add", [1, 5], idx, big, [1, 5], idx, two)
    _, _, p = oracle.ewise("mul", [1, 5], idx, big, [1, 5], idx, two)
    _, _, m = oracle.ewise("max", [1, 5], idx, big, [1, 5], idx, two)
    assert s.tolist() == [-(2**31), 131072, 92682, 2**31 - 1]
    assert p.tolist() == [2**31 - 1, 0, 46341 * 46341 - 2**32, -(2**31)]
    assert m.tolist() == [2**31 - 1, 65536, 46341, -1]


# ---- SpGEMM --------------------------------------------------------------------

@pytest.mark.parametrize("name", ["west0132.mtx", "lpi_itest6.mtx", "lp_blend.mtx", "Trec4.mtx"])
def test_gustavson_matches_scipy_on_finch_fixtures(name):
    # kernel_tests.jl:89-130 checks B == A*A on these matrices (square ones);
    # here A*A^T covers the rectangular ones too.
    A = read_mtx(name)
    if A.dtype.kind == "i":
        A = A.astype(np.int32)  # integer fixture: exact Int32 plus-times
    Bm = A if A.shape[0] == A.shape[1] else A.T.tocsr()
    A.sort_indices()
    Bm.sort_indices()
    cptr, cidx, cval = oracle.spgemm("plus_times", A.shape[0], Bm.shape[1], A.indptr + 1,
                                     A.indices + 1, A.data, Bm.indptr + 1, Bm.indices + 1,
                                     Bm.data)
    ref = (A @ Bm).tocsr()
    ref.sort_indices()
    # Structure: every touched coordinate is stored, even where the sum cancels.
    pa, pb = A.copy(), Bm.copy()
    pa.data[:] = 1
    pb.data[:] = 1
    touched = (pa.astype(np.int64) @ pb.astype(np.int64)).tocsr()
    touched.sort_indices()
    np.testing.assert_array_equal(cptr - 1, touched.indptr)
    np.testing.assert_array_equal(cidx - 1, touched.indices)
    dense = np.zeros(ref.shape)
    rows = np.repeat(np.arange(ref.shape[0]), np.diff(cptr))
    dense[rows, cidx - 1] = cval
    np.testing.assert_allclose(dense, ref.toarray(), rtol=1e-12, atol=1e-12)


def test_gustavson_int32_exact_and_wrapping():
    rng = np.random.default_rng(6)
    S = sp.random(50, 40, density=0.1, format="csr", random_state=rng)
    T = sp.random(40, 60, density=0.1, format="csr", random_state=rng)
    S.data = rng.integers(1, 10, S.nnz).astype(np.int32)
    T.data = rng.integers(1, 10, T.nnz).astype(np.int32)
    cptr, cidx, cval = oracle.spgemm("plus_times", 50, 60, S.indptr + 1, S.indices + 1, S.data,
                                     T.indptr + 1, T.indices + 1, T.data)
    ref = (S.astype(np.int64) @ T.astype(np.int64)).toarray()
    rows = np.repeat(np.arange(50), np.diff(cptr))
    assert np.array_equal(cval.astype(np.int64), ref[rows, cidx - 1])
    # Int32 overflow wraps as in Julia: 4*2^30 + 4*2^30 = 2^33 == 0 (mod 2^32).
    _, _, v = oracle.spgemm("plus_times", 1, 1, [1, 3], [1, 2], np.array([4, 4], np.int32),
                            [1, 2, 3], [1, 1], np.array([2 ** 30, 2 ** 30], np.int32))
    assert v.tolist() == [0]


def test_min_plus_brute_force():
    # interface_tests.jl:70-84 checks the max-plus product by brute force; the
    # min-plus product with fill Inf is checked the same way.
    rng = np.random.default_rng(7)
    m, l, n = 12, 9, 11
    S = sp.random(m, l, density=0.3, format="csr", random_state=rng, dtype=np.float32)
    T = sp.random(l, n, density=0.3, format="csr", random_state=rng, dtype=np.float32)
    S.data = rng.standard_normal(S.nnz).astype(np.float32)
    T.data = rng.standard_normal(T.nnz).astype(np.float32)
    cptr, cidx, cval = oracle.spgemm("min_plus", m, n, S.indptr + 1, S.indices + 1, S.data,
                                     T.indptr + 1, T.indices + 1, T.data)
    A = np.full((m, l), np.inf, np.float32)
    A[S.nonzero()] = S.data
    B = np.full((l, n), np.inf, np.float32)
    B[T.nonzero()] = T.data
    R = np.full((m, n), np.inf, np.float32)
    for i in range(m):
        for j in range(n):
            for k in range(l):
                R[i, j] = min(R[i, j], np.float32(A[i, k] + B[k, j]))
    rows = np.repeat(np.arange(m), np.diff(cptr))
    assert np.array_equal(cval, R[rows, cidx - 1])           # bit-exact
    assert np.isfinite(R).sum() == cptr[-1] - 1             # pattern = finite entries


# ---- format conversion --------------------------------------------------------------

def _full_storage(A):
    """Every entry of a dense matrix stored, column-major (CSC, 1-based)."""
    m, n = A.shape
    ptr = np.arange(0, m * n + 1, m, dtype=np.int64) + 1
    idx = np.tile(np.arange(m, dtype=np.int64), n) + 1
    return ptr, idx, A.T.reshape(-1).copy()


def test_dropfills_and_coo_reproduce_representation_goldens(goldens):
    # Dense{SparseList}_representation.txt: the exact ptr/idx/val Finch builds
    # from each dense array.  Reached two independent ways: dropping the fills
    # from full storage, and assembling the nonzeros from shuffled COO lists.
    rng = np.random.default_rng(0)
    for c in goldens["representation_dense_sparselist"]["cases"]:
        A = np.array(c["dense"], dtype=np.float64)
        m, n = A.shape
        ptr, idx, val = _full_storage(A)
        dp, di, dv = oracle.dropfills(n, ptr, idx, val, 0.0)
        assert (dp.tolist(), di.tolist(), dv.tolist()) == (c["ptr"], c["idx"], c["val"]), c["name"]
        I, J = np.nonzero(A)
        p = rng.permutation(len(I))
        cp, ci, cv = oracle.coo_to_compressed(n, m, J[p] + 1, I[p] + 1, A[I[p], J[p]])
        assert (cp.tolist(), ci.tolist(), cv.tolist()) == (c["ptr"], c["idx"], c["val"]), c["name"]


def test_dropfills_isequal_semantics():
    # copy.jl:98-101 drops only values isequal to the fill: -0.0 and NaN stay
    # when the fill is 0.0; with a NaN fill the NaNs go.
    ptr = np.array([1, 4, 6])
    idx = np.array([1, 2, 3, 1, 3])
    val = np.array([0.0, -0.0, np.nan, 2.0, 0.0])
    p, i, v = oracle.dropfills(2, ptr, idx, val, 0.0)
    assert p.tolist() == [1, 3, 4] and i.tolist() == [2, 3, 1]
    assert np.signbit(v[0]) and np.isnan(v[1]) and v[2] == 2.0
    p, i, v = oracle.dropfills(2, ptr, idx, val, np.nan)
    assert p.tolist() == [1, 3, 5] and i.tolist() == [1, 2, 1, 3]


@pytest.mark.parametrize("base", [1, 0])
def test_transpose_is_dense_transpose(base):
    # A permutation of the stored entries: the result, densified, is A^T
    # exactly, fibres are sorted, and transposing twice is the identity.
    rng = np.random.default_rng(3)
    A = sp.random(57, 41, density=0.12, random_state=rng, format="csr")
    A.data = rng.standard_normal(A.nnz)
    A.sort_indices()
    ptr, idx = A.indptr.astype(np.int64) + base, A.indices.astype(np.int64) + base
    tp, ti, tv = oracle.transpose(57, 41, ptr, idx, A.data, base)
    T = sp.csr_matrix((tv, ti - base, tp - base), shape=(41, 57))
    assert np.array_equal(T.toarray(), A.toarray().T)
    for o in range(41):
        assert np.all(np.diff(ti[tp[o] - base:tp[o + 1] - base]) > 0)
    bp, bi, bv = oracle.transpose(41, 57, tp, ti, tv, base)
    assert np.array_equal(bp, ptr) and np.array_equal(bi, idx) and np.array_equal(bv, A.data)


def test_transpose_keeps_explicit_zeros_and_empty_fibres():
    # Written coordinates are stored even when equal to the fill
    # (scalars.jl:209-218), so a stored 0.0 survives the transpose.
    ptr = np.array([1, 1, 3, 3, 4])
    idx = np.array([2, 5, 2])
    val = np.array([0.0, 7.0, 1.5])
    tp, ti, tv = oracle.transpose(4, 6, ptr, idx, val)
    assert tp.tolist() == [1, 1, 3, 3, 3, 4, 4]
    assert ti.tolist() == [2, 4, 2] and tv.tolist() == [0.0, 1.5, 7.0]


def test_coo_duplicates_fold_left_in_input_order():
    # fsparse.jl:85-90: foldl(+, V[start:stop-1]) over each run of equal
    # coordinates, in input order.  1e16 + 1 + -1e16 (left fold) is 0.0,
    # while any other order gives 1.0 or 2.0: the order is pinned.
    outer = np.array([2, 1, 2, 2, 1])
    inner = np.array([3, 1, 3, 3, 2])
    val = np.array([1e16, 5.0, 1.0, -1e16, 4.0])
    cp, ci, cv = oracle.coo_to_compressed(2, 3, outer, inner, val)
    assert cp.tolist() == [1, 3, 4] and ci.tolist() == [1, 2, 3]
    assert cv.tolist() == [5.0, 4.0, 0.0]
    # Int32 sums wrap like Julia's Int32.
    cp, ci, cv = oracle.coo_to_compressed(1, 1, np.array([1, 1]), np.array([1, 1]),
                                          np.array([2**31 - 1, 2], dtype=np.int32))
    assert cv.tolist() == [-(2**31) + 1]


def test_coo_matches_scipy_sum_duplicates():
    rng = np.random.default_rng(5)
    m, n, k = 300, 200, 5000
    I = rng.integers(0, m, k)
    J = rng.integers(0, n, k)
    V = rng.integers(-9, 10, k).astype(np.float64)  # small integers: every order is exact
    cp, ci, cv = oracle.coo_to_compressed(n, m, J + 1, I + 1, V)
    S = sp.coo_matrix((V, (I, J)), shape=(m, n)).tocsc()
    S.sum_duplicates()
    assert np.array_equal(cp - 1, S.indptr) and np.array_equal(ci - 1, S.indices)
    assert np.array_equal(cv, S.data)


def test_min_plus_uses_julia_min():
    # `<<min>>=` folds with Julia's `min` (base/math.jl): NaN wins and -0.0 is
    # below 0.0, so the fold gives the same result in either product order.
    B_ptr, B_idx = np.array([1, 2, 3]), np.array([1, 1])
    for order in ((-0.0, 0.0), (0.0, -0.0)):
        # Products A[0,k] + B[k,0]: -0.0 + -0.0 and 0.0 + 0.0, in both orders.
        p, i, v = oracle.spgemm("min_plus", 1, 1, np.array([1, 3]), np.array([1, 2]),
                                np.array(order), B_ptr, B_idx, np.array(order))
        assert p.tolist() == [1, 2] and i.tolist() == [1]
        assert v[0] == 0.0 and np.signbit(v[0])
    for order in ((1.0, np.nan), (np.nan, 1.0)):
        p, i, v = oracle.spgemm("min_plus", 1, 1, np.array([1, 3]), np.array([1, 2]),
                                np.array(order), B_ptr, B_idx, np.array([0.0, 0.0]))
        assert np.isnan(v[0])


# ---- threaded baselines: AtomicElement CSC and the greedy schedule --------------------

def test_spmv_csc_atomic_golden_and_scipy(goldens):
    # benchmark/benchmarks.jl:402-413 (AtomicElement y, parallel columns).  On
    # the golden of typical_spmv_csc.txt:92 every y[i] has at most two terms,
    # and two FP adds onto 0 give the same bits in either order.
    g = goldens["spmv_csc"]
    ptr, idx, val = dense_to_csc(np.array(g["A_dense"]))
    for nth in (1, 2, 3):
        y = oracle.spmv_csc_atomic(4, ptr, idx, val, np.array(g["x"], np.float64), nthreads=nth)
        assert y.tolist() == g["y"]
    rng = np.random.default_rng(51)
    S = sp.random(3000, 2000, density=0.01, format="csc", random_state=rng)
    x = rng.standard_normal(2000)
    ref = S @ x
    serial = oracle.spmv_csc(3000, S.indptr + 1, S.indices + 1, S.data, x)
    assert np.array_equal(oracle.spmv_csc_atomic(3000, S.indptr + 1, S.indices + 1, S.data, x,
                                                 nthreads=1), serial)
    y8 = oracle.spmv_csc_atomic(3000, S.indptr + 1, S.indices + 1, S.data, x, nthreads=8)
    assert np.linalg.norm(y8 - ref) <= 1e-12 * np.linalg.norm(ref)


def test_spmv_csc_mutex_golden_and_scipy(goldens):
    # benchmark/benchmarks.jl:414-426 (Mutex(Element) y, parallel columns):
    # the golden bit for bit at any thread count, the serial loop's bits at
    # one thread, and scipy at eight.
    g = goldens["spmv_csc"]
    ptr, idx, val = dense_to_csc(np.array(g["A_dense"]))
    for nth in (1, 2, 3):
        y = oracle.spmv_csc_mutex(4, ptr, idx, val, np.array(g["x"], np.float64), nthreads=nth)
        assert y.tolist() == g["y"]
    rng = np.random.default_rng(53)
    S = sp.random(3000, 2000, density=0.01, format="csc", random_state=rng)
    x = rng.standard_normal(2000)
    ref = S @ x
    serial = oracle.spmv_csc(3000, S.indptr + 1, S.indices + 1, S.data, x)
    assert np.array_equal(oracle.spmv_csc_mutex(3000, S.indptr + 1, S.indices + 1, S.data, x,
                                                nthreads=1), serial)
    y8 = oracle.spmv_csc_mutex(3000, S.indptr + 1, S.indices + 1, S.data, x, nthreads=8)
    assert np.linalg.norm(y8 - ref) <= 1e-12 * np.linalg.norm(ref)
    y32 = oracle.spmv_csc_mutex(3000, S.indptr + 1, S.indices + 1, S.data.astype(np.float32),
                                x.astype(np.float32), nthreads=4)
    assert np.linalg.norm(y32 - ref) <= 1e-5 * np.linalg.norm(ref)


@pytest.mark.parametrize("chunk", [1, 7, 1000])
@pytest.mark.parametrize("nth", [1, 3, 8])
def test_greedy_schedule_equals_serial_bit_for_bit(chunk, nth):
    rng = np.random.default_rng(52)
    S = sp.random(5000, 4000, density=0.003, format="csr", random_state=rng)
    S.sort_indices()
    x = rng.standard_normal(4000)
    B = rng.standard_normal((4000, 5))
    p, i = S.indptr + 1, S.indices + 1
    y0 = oracle.spmv_csr(p, i, S.data, x)
    C0 = oracle.spmm_csr(p, i, S.data, B)
    assert np.array_equal(oracle.spmv_csr_greedy(p, i, S.data, x, chunk, nthreads=nth), y0)
    assert np.array_equal(oracle.spmm_csr_greedy(p, i, S.data, B, chunk, nthreads=nth), C0)
    assert np.linalg.norm(y0 - S @ x) <= 1e-13 * np.linalg.norm(S @ x)
