"""GPU tests of the Finch interface mirror: programs written as in Finch's own
tests, run through einsum / lazy / compute, checked against the oracle and
Finch's goldens."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from conftest import read_mtx
from finchgpu import interface as fi
from finchgpu.tensors import (GPU, Dense, Element, SparseDict, SparseList, SwizzleArray, Tensor,
                              swizzle, transfer)

pytestmark = pytest.mark.gpu


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300)


def test_spmv_csc_golden(dev, goldens):
    # test/suites/typical_tests.jl:47-73 with typical_spmv_csc.txt:92.
    g = goldens["spmv_csc"]
    A = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), np.array(g["A_dense"])))
    x = transfer(GPU(0), Tensor(Dense(Element(0.0)), np.array(g["x"])))
    y = fi.einsum("y[i] += A[i, j] * x[j]", A=A, x=x)
    assert y.todense().tolist() == g["y"]


def test_spmv_swizzled_csr(dev, goldens):
    g = goldens["spmv_transpose"]
    A = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), np.array(g["A_dense"])))
    x = transfer(GPU(0), Tensor(Dense(Element(0.0)), np.array(g["x"])))
    # swizzle(A, 2, 1)[i, j] = A[j, i]: row-wise read of the column buffers.
    y = fi.einsum("y[i] += S[i, j] * x[j]", S=swizzle(A, 2, 1), x=x)
    assert y.todense().tolist() == g["y"]


def _rand(m, n, density, seed, dtype=np.float64):
    rng = np.random.default_rng(seed)
    S = sp.random(m, n, density=density, format="csc", random_state=rng, dtype=np.float64)
    S.data = rng.standard_normal(S.nnz)
    return S.astype(dtype)


def test_spmm_with_finch_column_major_dense(dev):
    # @einsum C[i,k] += A[i,j]*B[j,k] (interface_tests.jl:22-24) with A read
    # row-wise (swizzle of the transposed storage) and B in Finch's
    # column-major Dense(Dense(Element)).
    S = _rand(300, 200, 0.05, 1)
    At = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), S.T.tocsc()))  # stores A^T
    Bn = np.random.default_rng(2).standard_normal((200, 7))
    B = transfer(GPU(0), Tensor(Dense(Dense(Element(0.0))), Bn))
    C = fi.einsum("C[i, k] += A[i, j] * B[j, k]", A=swizzle(At, 2, 1), B=B)
    assert C.shape == (300, 7)
    assert rel_err(C.todense(), S @ Bn) <= 1e-12


@pytest.mark.parametrize("name", ["west0132.mtx", "lp_blend.mtx"])
def test_gustavson_column_compressed(dev, name):
    # kernel_tests.jl:89-130: B == A*A (or A*A^T) on Finch's fixtures, with the
    # operands in Finch's default column-compressed format.
    S = read_mtx(name).tocsc()
    T = S if S.shape[0] == S.shape[1] else S.T.tocsc()
    A = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), S))
    B = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), T))
    C = fi.einsum("C[i, k] += A[i, j] * B[j, k]", A=A, B=B)
    ref = (S @ T).toarray()
    assert rel_err(C.todense(), ref) <= 1e-12
    # Same update order as the reference's column-wise Gustavson: bit-exact
    # against the oracle run column by column (rows of T^T x rows of S^T).
    Tt, St = T.T.tocsr(), S.T.tocsr()
    Tt.sort_indices()
    St.sort_indices()
    cptr, cidx, cval = oracle.spgemm("plus_times", Tt.shape[0], St.shape[1], Tt.indptr + 1,
                                     Tt.indices + 1, Tt.data, St.indptr + 1, St.indices + 1,
                                     St.data)
    sl = C.lvl.lvl
    assert np.array_equal(sl.ptr.cpu().numpy(), cptr)
    assert np.array_equal(sl.idx.cpu().numpy(), cidx)
    assert np.array_equal(C.element.val.cpu().numpy(), cval)


def test_min_plus_einsum(dev):
    # interface_tests.jl:70-84 (max-plus with fill -Inf) in its min-plus form.
    rng = np.random.default_rng(3)
    An = np.where(rng.random((9, 8)) < 0.4, rng.standard_normal((9, 8)), np.inf)
    Bn = np.where(rng.random((8, 7)) < 0.4, rng.standard_normal((8, 7)), np.inf)
    A = transfer(GPU(0), Tensor(Dense(SparseList(Element(np.inf))), An))
    B = transfer(GPU(0), Tensor(Dense(SparseList(Element(np.inf))), Bn))
    C = fi.einsum("C[i, k] <<min>>= A[i, j] + B[j, k]", A=A, B=B)
    ref = np.full((9, 7), np.inf)
    for i in range(9):
        for k in range(7):
            for j in range(8):
                ref[i, k] = min(ref[i, k], An[i, j] + Bn[j, k])
    assert np.array_equal(C.todense(), ref)


@pytest.mark.parametrize("prog,fn", [("C[i, j] = A[i, j] * B[i, j]", lambda a, b: a * b),
                                     ("C[i, j] = A[i, j] + B[i, j]", lambda a, b: a + b),
                                     ("C[i, j] = max(A[i, j], B[i, j])", np.maximum)])
def test_elementwise(dev, prog, fn):
    S, T = _rand(60, 50, 0.1, 4), _rand(60, 50, 0.2, 5)
    A = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), S))
    B = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), T))
    C = fi.einsum(prog, A=A, B=B)
    assert np.array_equal(C.todense(), fn(S.toarray(), T.toarray()))


def test_fused_sddmm_compute(dev):
    # compute(lazy(M) .* (lazy(A) * lazy(B))) (benchmark/benchmarks.jl:141-168).
    S = _rand(80, 70, 0.1, 6, np.float32)
    Mt = transfer(GPU(0), Tensor(Dense(SparseList(Element(np.float32(0)))), S.T.tocsc()))
    rng = np.random.default_rng(7)
    Xn = rng.standard_normal((80, 16)).astype(np.float32)
    Yn = rng.standard_normal((16, 70)).astype(np.float32)
    X = transfer(GPU(0), Tensor(Dense(Dense(Element(np.float32(0)))), Xn))
    Y = transfer(GPU(0), Tensor(Dense(Dense(Element(np.float32(0)))), Yn))
    C = fi.compute(fi.lazy(swizzle(Mt, 2, 1)) * (fi.lazy(X) @ fi.lazy(Y)))
    ref = S.toarray().astype(np.float64) * (Xn.astype(np.float64) @ Yn.astype(np.float64))
    assert rel_err(C.todense(), ref) <= 1e-5


# ---- storage order and format conversion ---------------------------------------------

def test_spmm_with_column_compressed_operand(dev):
    # @einsum C[i,k] += A[i,j]*B[j,k] with A in Finch's default CSC: the
    # operand is transposed on the device to row order, then SpMM runs.
    S = _rand(300, 200, 0.05, 11)
    A = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), S))
    Bn = np.random.default_rng(12).standard_normal((200, 9))
    B = transfer(GPU(0), Tensor(Dense(Dense(Element(0.0))), Bn))
    C = fi.einsum("C[i, k] += A[i, j] * B[j, k]", A=A, B=B)
    assert C.shape == (300, 9)
    assert rel_err(C.todense(), S @ Bn) <= 1e-12
    # The row-ordered copy is cached on the tensor.
    assert getattr(A, "_other_order", None) is not None


@pytest.mark.parametrize("prog,fn", [("C[i, j] = A[i, j] * B[i, j]", lambda a, b: a * b),
                                     ("C[i, j] = A[i, j] + B[i, j]", lambda a, b: a + b)])
def test_elementwise_mixed_storage_orders(dev, prog, fn):
    S, T = _rand(60, 50, 0.1, 13), _rand(60, 50, 0.2, 14)
    A = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), S))
    Bt = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), T.T.tocsc()))
    C = fi.einsum(prog, A=A, B=swizzle(Bt, 2, 1))
    assert np.array_equal(C.todense(), fn(S.toarray(), T.toarray()))


def test_spgemm_mixed_storage_orders(dev):
    S, T = _rand(70, 40, 0.1, 15), _rand(40, 55, 0.1, 16)
    At = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), S.T.tocsc()))
    B = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), T))
    C = fi.einsum("C[i, k] += A[i, j] * B[j, k]", A=swizzle(At, 2, 1), B=B)
    assert rel_err(C.todense(), (S @ T).toarray()) <= 1e-12


def test_fused_sddmm_on_column_compressed_pattern(dev):
    S = _rand(80, 70, 0.1, 17, np.float32)
    M = transfer(GPU(0), Tensor(Dense(SparseList(Element(np.float32(0)))), S))
    rng = np.random.default_rng(18)
    Xn = rng.standard_normal((80, 16)).astype(np.float32)
    Yn = rng.standard_normal((16, 70)).astype(np.float32)
    X = transfer(GPU(0), Tensor(Dense(Dense(Element(np.float32(0)))), Xn))
    Y = transfer(GPU(0), Tensor(Dense(Dense(Element(np.float32(0)))), Yn))
    C = fi.compute(fi.lazy(M) * (fi.lazy(X) @ fi.lazy(Y)))
    ref = S.toarray().astype(np.float64) * (Xn.astype(np.float64) @ Yn.astype(np.float64))
    assert C.format == "cscdict"          # compute() results are Dense(SparseDict)
    assert rel_err(C.todense(), ref) <= 1e-5


def test_permutedims_fsparse_dropfills(dev, goldens):
    # permutedims (copy.jl:64-71) of a CSC tensor is the transpose in the
    # same format; fsparse (fsparse.jl:64-97) of a representation golden's
    # nonzeros rebuilds Finch's buffers; dropfills removes stored zeros.
    S = _rand(45, 30, 0.2, 19)
    A = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), S))
    P = fi.permutedims(A)
    assert P.shape == (30, 45) and P.format == "csc"
    assert np.array_equal(P.todense(), S.toarray().T)
    assert np.array_equal(fi.permutedims(P).todense(), S.toarray())
    for c in goldens["representation_dense_sparselist"]["cases"]:
        D = np.array(c["dense"], dtype=np.float64)
        I, J = np.nonzero(D)
        p = np.random.default_rng(20).permutation(len(I))
        dev_i = lambda a: torch.as_tensor(a + 1, dtype=torch.int64).to(dev)
        T = fi.fsparse(dev_i(I[p]), dev_i(J[p]), torch.as_tensor(D[I[p], J[p]]).to(dev), D.shape)
        sl = T.lvl.lvl
        assert sl.ptr.tolist() == c["ptr"] and sl.idx.tolist() == c["idx"], c["name"]
        assert T.element.val.tolist() == c["val"], c["name"]
    E = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), S))
    E.element.val = E.element.val * (torch.arange(E.element.val.numel(), device=dev) % 3 != 0)
    F = fi.dropfills(E)
    kept = E.element.val.cpu().numpy()
    assert F.countstored() == int(np.count_nonzero(kept != 0.0) + np.count_nonzero(
        (kept == 0.0) & np.signbit(kept)))
    assert np.array_equal(F.todense(), E.todense())


def test_spgemm_formulations_match_the_examples_test(dev):
    # test/suites/examples_tests.jl:103-119: 32x32 Int matrices at p = 0.1;
    # inner, Gustavson and outer products all equal A*B.
    rng = np.random.default_rng(21)
    S = sp.random(32, 32, density=0.1, random_state=rng, format="csc")
    T = sp.random(32, 32, density=0.1, random_state=rng, format="csc")
    S.data = rng.integers(-1000, 1000, S.nnz).astype(np.float64)
    T.data = rng.integers(-1000, 1000, T.nnz).astype(np.float64)
    Si, Ti = S.astype(np.int32), T.astype(np.int32)
    A = transfer(GPU(0), Tensor(Dense(SparseList(Element(np.int32(0)))), Si))
    B = transfer(GPU(0), Tensor(Dense(SparseList(Element(np.int32(0)))), Ti))
    ref = (S @ T).toarray()
    outs = [fn(A, B) for fn in (fi.spgemm_inner, fi.spgemm_gustavson, fi.spgemm_outer)]
    for C in outs:
        assert np.array_equal(C.todense().astype(np.float64), ref)
    sl0 = outs[0].lvl.lvl
    for C in outs[1:]:
        assert torch.equal(C.lvl.lvl.ptr, sl0.ptr) and torch.equal(C.lvl.lvl.idx, sl0.idx)


# ---- @finch with declared outputs, and SparseDict results --------------------------------

def test_finch_declared_dense_output_written_in_place(dev, goldens):
    # @finch (y .= 0; for j=_, i=_; y[i] += A[i,j] * x[j]) with y declared by
    # the caller (typical_tests.jl:47-73): y's own buffer receives the result.
    g = goldens["spmv_csc"]
    A = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), np.array(g["A_dense"])))
    x = transfer(GPU(0), Tensor(Dense(Element(0.0)), np.array(g["x"])))
    y = transfer(GPU(0), Tensor(Dense(Element(0.0)), np.full(4, 123.0)))
    buf = y.element.val
    out = fi.finch("y[i] += A[i, j] * x[j]", y=y, A=A, x=x)
    assert out["y"] is y and y.element.val is buf
    assert y.todense().tolist() == g["y"]
    # CSR read of the same buffers (benchmarks.jl:365-376: y[i] += A[j, i] * x[j]).
    z = transfer(GPU(0), Tensor(Dense(Element(0.0)), np.zeros(3)))
    fi.finch("z[i] += A[j, i] * w[j]", z=z, A=A,
             w=transfer(GPU(0), Tensor(Dense(Element(0.0)), np.ones(4))))
    assert np.allclose(z.todense(), np.array(g["A_dense"]).sum(axis=0))


def test_finch_declared_sparse_outputs_keep_their_format(dev):
    S, T = _rand(60, 50, 0.1, 31), _rand(60, 50, 0.2, 32)
    A = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), S))
    B = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), T))
    C = Tensor(Dense(SparseList(Element(0.0))))
    fi.finch("C[i, j] = A[i, j] + B[i, j]", C=C, A=A, B=B)
    assert C.format == "csc" and C.shape == (60, 50)
    assert np.array_equal(C.todense(), (S + T).toarray())
    # A swizzled declared output stores the result's rows as its fibres.
    R = SwizzleArray(Tensor(Dense(SparseList(Element(0.0)))), (2, 1))
    fi.finch("R[i, j] = A[i, j] * B[i, j]", R=R, A=A, B=B)
    assert np.array_equal(R.todense(), S.multiply(T).toarray())
    assert R.body.shape == (50, 60)


def test_compute_results_are_sparse_dict_in_frozen_layout(dev):
    # Finch's scheduler returns Dense(SparseDict) (traits.jl:509-517); after
    # freeze (sparse_dict_levels.jl:356-400) ptr/idx are the sorted
    # coordinates and val the positions of the values, 1:nnz when the
    # entries were inserted in coordinate order.
    S, T = _rand(70, 40, 0.1, 33), _rand(70, 40, 0.15, 34)
    A = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), S))
    B = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), T))
    D = fi.compute(fi.maximum(fi.lazy(A), fi.lazy(B)))
    assert D.format == "cscdict" and isinstance(D.lvl.lvl, SparseDict)
    C = Tensor(Dense(SparseList(Element(0.0))))
    fi.finch("C[i, j] = max(A[i, j], B[i, j])", C=C, A=A, B=B)
    sd = D.lvl.lvl
    assert torch.equal(sd.ptr, C.lvl.lvl.ptr) and torch.equal(sd.idx, C.lvl.lvl.idx)
    assert sd.val.tolist() == list(range(1, C.countstored() + 1)) and sd.in_order
    assert np.array_equal(D.todense(), C.todense())
    # A host-built SparseDict of the same matrix freezes to the same buffers
    # (a sum of two random patterns stores no zeros).
    P = fi.compute(fi.lazy(A) + fi.lazy(B))
    H = Tensor(Dense(SparseDict(Element(0.0))), P.todense())
    assert H.lvl.lvl.ptr.tolist() == P.lvl.lvl.ptr.tolist()
    assert H.lvl.lvl.idx.tolist() == P.lvl.lvl.idx.tolist()
    assert H.lvl.lvl.val.tolist() == P.lvl.lvl.val.tolist()
    # SparseDict results feed later programs (values in coordinate order).
    E = fi.einsum("C[i, j] = A[i, j] * B[i, j]", A=D, B=A)
    assert np.array_equal(E.todense(), np.maximum(S.toarray(), T.toarray()) * S.toarray())


def test_unrouted_device_programs_raise(dev):
    S = _rand(20, 20, 0.2, 35)
    A = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), S))
    B = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), S))
    from finchgpu._lib import UnsupportedError
    with pytest.raises(UnsupportedError):
        fi.einsum("C[i, k] <<max>>= A[i, j] + B[j, k]", A=A, B=B)


# ---- outputs through Finch's threaded-write levels ---------------------------------------

def test_threaded_write_levels_as_outputs(dev):
    # test/suites/parallel_tests.jl:627-690 (and benchmark/benchmarks.jl:402-426):
    # y = Dense(AtomicElement(0.0)) or Dense(Mutex(Element(0.0))) under
    # `for j = parallel(_), i = _; y[i] += A[i, j] * x[j]`, and
    # y = Dense(Separate(Element(0.0))) under
    # `for i = parallel(_), j = _; y[i] += A[j, i] * x[j]`; the reference
    # checks norm(y - A*x) / norm(A*x) < 1e-10 (and permutedims(A) * x).
    from finchgpu.tensors import AtomicElement, Mutex, Separate
    S = _rand(42, 42, 0.1, 41)
    xh = np.random.default_rng(42).random(42)
    A = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), S))
    x = transfer(GPU(0), Tensor(Dense(Element(0.0)), xh))
    for lvl in (AtomicElement(0.0), Mutex(Element(0.0))):
        y = transfer(GPU(0), Tensor(Dense(lvl), np.zeros(42)))
        fi.finch("y[i] += A[i, j] * x[j]", y=y, A=A, x=x)
        assert rel_err(y.todense(), S @ xh) < 1e-10, lvl
    y = Tensor(Dense(Separate(Element(0.0))))
    fi.finch("y[i] += A[j, i] * x[j]", y=y, A=A, x=x)
    assert y.shape == (42,) and len(y.lvl.lvl.val) == 42
    assert rel_err(y.todense(), S.T @ xh) < 1e-10
    # A declared Separate output with another number of positions is a
    # DimensionMismatch, as Finch's dimension check raises.
    from finchgpu._lib import DimensionMismatch
    z = transfer(GPU(0), Tensor(Dense(Separate(Element(0.0))), np.zeros(41)))
    with pytest.raises(DimensionMismatch):
        fi.finch("z[i] += A[j, i] * x[j]", z=z, A=A, x=x)
    # SpMM into Dense(Dense(Mutex(Element))) (parallel_tests.jl:331).
    Bh = np.random.default_rng(43).random((42, 42))
    B = transfer(GPU(0), Tensor(Dense(Dense(Element(0.0))), Bh))
    Ct = transfer(GPU(0), Tensor(Dense(Dense(Mutex(Element(0.0)))), np.zeros((42, 42))))
    fi.finch("Ct[i, j] += A[i, k] * B[k, j]", Ct=Ct, A=A, B=B)
    assert rel_err(Ct.todense(), S @ Bh) < 1e-10
