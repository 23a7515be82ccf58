"""Shared test setup: import paths, the `gpu` marker and common helpers."""
import json
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "finch.jl_amd"), os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN_DIR = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libfinchgpu.so")


@pytest.fixture(scope="session")
def goldens():
    with open(os.path.join(GOLDEN_DIR, "finch_goldens.json")) as f:
        return json.load(f)


def dense_to_csc(A, base=1):
    """Dense -> Dense(SparseList(Element)) buffers (column fibers), storing nonzeros."""
    A = np.asarray(A)
    m, n = A.shape
    ptr = [base]
    idx, val = [], []
    for j in range(n):
        for i in range(m):
            if A[i, j] != 0:
                idx.append(i + base)
                val.append(A[i, j])
        ptr.append(len(idx) + base)
    return np.array(ptr, np.int64), np.array(idx, np.int64), np.array(val, A.dtype)


def read_mtx(name):
    """A Matrix Market fixture from Finch's test/data as a scipy CSR matrix."""
    import scipy.io
    return scipy.io.mmread(os.path.join(GOLDEN_DIR, "mtx", name)).tocsr()


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from finchgpu import _lib
    _lib.load()
    return torch.device("cuda:0")
