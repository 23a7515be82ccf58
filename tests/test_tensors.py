"""Host side of the Finch interface mirror: level formats built from dense
arrays must match Finch's representation goldens, and programs must parse to
the kernels they route to (no device needed)."""
import numpy as np
import pytest

from finchgpu import interface
from finchgpu._lib import UnsupportedError
from finchgpu.tensors import Dense, Element, SparseList, Tensor, swizzle


def test_dense_sparselist_representation_goldens(goldens):
    # test/reference64/representation/Dense{SparseList}_representation.txt
    cases = goldens["representation_dense_sparselist"]["cases"]
    assert len(cases) >= 5
    for c in cases:
        A = np.array(c["dense"])
        t = Tensor(Dense(SparseList(Element(0.0))), A)
        assert t.lvl.lvl.ptr.tolist() == c["ptr"], c["name"]
        assert t.lvl.lvl.idx.tolist() == c["idx"], c["name"]
        assert t.element.val.tolist() == c["val"], c["name"]
        assert np.array_equal(t.todense(), A)
        assert t.countstored() == len(c["idx"])


def test_zero_based_buffers_view(goldens):
    # issue_tests.jl:388-404: PlusOneVector views of 0-based buffers.
    g = goldens["zero_based"]
    t = Tensor.from_buffers(np.array(g["ptr"]), np.array(g["idx"]), np.array(g["val"]),
                            g["m"], g["n"], base=0)
    assert np.array_equal(t.todense(), np.array(g["A_dense"]))


def test_spmv_golden_operand_layout(goldens):
    g = goldens["spmv_csc"]
    t = Tensor(Dense(SparseList(Element(0.0))), np.array(g["A_dense"]))
    assert t.shape == (4, 3)
    assert t.lvl.lvl.ptr.tolist() == [1, 3, 3, 6]
    assert t.lvl.lvl.idx.tolist() == [2, 3, 1, 3, 4]


def test_swizzle_is_transpose_view():
    A = np.array([[0, 2.0, 0], [1.0, 0, 3.0]])
    t = Tensor(Dense(SparseList(Element(0.0))), A)
    s = swizzle(t, 2, 1)
    assert s.shape == (3, 2)
    assert np.array_equal(s.todense(), A.T)
    assert swizzle(s, 2, 1).dims == (1, 2)


@pytest.mark.parametrize("prog,expect", [
    ("y[i] += A[i, j] * x[j]", ("y", ("i",), "+", "*")),
    ("C[i,k] += A[i,j] * B[j,k]", ("C", ("i", "k"), "+", "*")),
    ("C[i, k] <<min>>= A[i, j] + B[j, k]", ("C", ("i", "k"), "min", "+")),
    ("C[i, j] = max(A[i, j], B[i, j])", ("C", ("i", "j"), None, "max")),
    ("C[i, j] = A[i, j] * B[i, j]", ("C", ("i", "j"), None, "*")),
])
def test_parse(prog, expect):
    out, oidx, red, terms, comb = interface.parse(prog)
    assert (out, oidx, red, comb) == expect
    assert len(terms) == 2


def test_unrecognised_programs_raise():
    with pytest.raises(UnsupportedError):
        interface.parse("y[i] = sin(A[i, j])")


def test_threaded_write_levels_are_dense_outputs():
    # The output levels of Finch's threaded SpMV/SpMM tests
    # (test/suites/parallel_tests.jl:603-690; benchmark/benchmarks.jl:402-426):
    # AtomicElement and Mutex(Element) hold one value buffer like Element;
    # Separate(Element) gives each position its own Element.
    from finchgpu.tensors import AtomicElement, Mutex, Separate
    x = np.array([1.5, 0.0, -2.0])
    for lvl in (AtomicElement(0.0), Mutex(Element(0.0)), Mutex(AtomicElement(0.0))):
        t = Tensor(Dense(lvl), x)
        assert t.format == "dense" and t.shape == (3,)
        assert np.array_equal(t.todense(), x)
    M = np.arange(6.0).reshape(2, 3)
    t = Tensor(Dense(Dense(Mutex(Element(0.0)))), M)
    assert t.format == "dense2" and t.shape == (2, 3)
    assert np.array_equal(t.todense(), M)
    s = Tensor(Dense(Separate(Element(0.0))), x)
    assert s.format == "separate" and s.shape == (3,)
    assert [e.val.tolist() for e in s.lvl.lvl.val] == [[1.5], [0.0], [-2.0]]
    assert np.array_equal(s.todense(), x)
    # A Separate over anything but Element is not a backend format.
    with pytest.raises(NotImplementedError):
        Tensor(Dense(Separate(Dense(Element(0.0))))).format
