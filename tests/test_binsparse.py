"""Binsparse ``.bspnpy`` files (src/interface/fileio/binsparse.jl, ext/NPZExt.jl).

The reference tests binsparse by round trips (test/suites/fileio_tests.jl:71-125:
write, read, same structure).  Here the same round trips, plus the layout the
format fixes: 0-based ``pointers_to_1`` / ``indices_1``, which for Finch's
0-based fixture (test/suites/issue_tests.jl:388-404) are exactly its
``ptr_c`` / ``idx_c`` buffers.
"""
import json
import os

import numpy as np
import pytest
import torch

import oracle
from finchgpu import binsparse, ops
from finchgpu.tensors import GPU, Dense, Element, SparseList, Tensor, swizzle


def test_zero_based_fixture_is_the_csc_layout(tmp_path, goldens):
    g = goldens["zero_based"]
    A = np.array(g["A_dense"])
    t = Tensor(Dense(SparseList(Element(0.0))), A)
    p = binsparse.bspwrite(str(tmp_path / "A.bspnpy"), t)
    desc = json.load(open(os.path.join(p, "binsparse.json")))["binsparse"]
    assert desc["format"] == "CSC" and desc["shape"] == [4, 3] and desc["version"] == "0.1"
    assert desc["number_of_stored_values"] == 5
    assert desc["data_types"] == {"values": "float64", "fill_value": "float64",
                                  "pointers_to_1": "int64", "indices_1": "int64"}
    assert np.load(os.path.join(p, "pointers_to_1.npy")).tolist() == g["ptr"]
    assert np.load(os.path.join(p, "indices_1.npy")).tolist() == g["idx"]
    assert np.load(os.path.join(p, "values.npy")).tolist() == g["val"]
    r = binsparse.bspread(p)
    assert r.format == "csc" and np.array_equal(r.todense(), A)


def test_round_trips_of_representation_goldens(tmp_path, goldens):
    for i, c in enumerate(goldens["representation_dense_sparselist"]["cases"]):
        A = np.array(c["dense"], dtype=np.float64)
        t = Tensor(Dense(SparseList(Element(0.0))), A)
        for name, obj in (("csc", t), ("csr", swizzle(t, 2, 1))):
            p = binsparse.bspwrite(str(tmp_path / f"{i}_{name}.bspnpy"), obj)
            r = binsparse.bspread(p)
            body = r.body if name == "csr" else r
            assert body.lvl.lvl.ptr.tolist() == c["ptr"], c["name"]
            assert body.lvl.lvl.idx.tolist() == c["idx"], c["name"]
            assert body.element.val.tolist() == c["val"], c["name"]
            assert np.array_equal(r.todense(), A if name == "csc" else A.T)


def test_dense_vector_and_fill_value(tmp_path):
    v = Tensor(Dense(Element(0.0)), np.array([1.5, 0.0, -2.0]))
    r = binsparse.bspread(binsparse.bspwrite(str(tmp_path / "v.bspnpy"), v))
    assert r.todense().tolist() == [1.5, 0.0, -2.0]
    M = np.where(np.eye(3) > 0, 1.0, np.inf)
    t = Tensor(Dense(SparseList(Element(np.inf))), M)
    r = binsparse.bspread(binsparse.bspwrite(str(tmp_path / "inf.bspnpy"), t))
    assert r.fill_value == np.inf and np.array_equal(r.todense(), M)


def test_rejects_other_forms(tmp_path):
    from finchgpu import _lib
    t = Tensor(Dense(SparseList(Element(0.0))), np.eye(2))
    with pytest.raises(_lib.UnsupportedError):
        binsparse.bspwrite(str(tmp_path / "a.bsp.h5"), t)


@pytest.mark.gpu
def test_device_read_zero_based_and_coo(tmp_path, dev, goldens):
    # The file's 0-based arrays go to the kernels unchanged (index_base 0).
    g = goldens["zero_based"]
    A = np.array(g["A_dense"])
    p = binsparse.bspwrite(str(tmp_path / "A.bspnpy"),
                           Tensor(Dense(SparseList(Element(0.0))), A))
    kind, C = binsparse.read_compressed(p, dev)
    assert kind == "csc" and C.base == 0
    x = np.array([1.0, 2.0, 3.0])
    y = ops.spmv_csc(C, torch.as_tensor(x).to(dev)).cpu().numpy()
    assert y.tolist() == oracle.spmv_csc(4, np.array(g["ptr"]), np.array(g["idx"]),
                                         np.array(g["val"]), x, base=0).tolist()
    # A COO file (binsparse spec: indices_0 rows, indices_1 columns), entries
    # in any order, assembled on the device.
    q = tmp_path / "B.bspnpy"
    os.makedirs(q)
    rows, cols = np.nonzero(A)
    order = np.random.default_rng(0).permutation(rows.size)
    arrays = {"indices_0": rows[order].astype(np.uint32), "indices_1": cols[order].astype(np.uint32),
              "values": A[rows, cols][order], "fill_value": np.array([0.0])}
    for k, v in arrays.items():
        np.save(q / f"{k}.npy", v)
    desc = {"format": "COO", "shape": [4, 3], "version": "0.1", "number_of_stored_values": 5,
            "data_types": {"indices_0": "uint32", "indices_1": "uint32", "values": "float64",
                           "fill_value": "float64"}}
    json.dump({"binsparse": desc}, open(q / "binsparse.json", "w"))
    r = binsparse.bspread(str(q), GPU(0))
    assert np.array_equal(r.todense(), A)
