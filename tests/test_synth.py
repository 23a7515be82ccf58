"""Host-side checks of the synthetic operand generators (no GPU).

`fsprand` follows Finch's fsprand semantics (src/interface/fsparse.jl:173-290):
nnz ~ Binomial(m*n, p), distinct uniform coordinates, column-major order,
values from rand.  The power-law and Poisson generators define the benchmark
workloads.  Each property is checked on the buffers themselves.
"""
import numpy as np
import pytest

from finchgpu import synth


def check_compressed(ptr, idx, n_outer, n_inner, base=1):
    """Finch's SparseList invariants: ptr[0] = base, non-decreasing, and each
    fibre's coordinates strictly increasing inside [base, base + n_inner)."""
    assert ptr.size == n_outer + 1 and ptr[0] == base and ptr[-1] - base == idx.size
    assert np.all(np.diff(ptr) >= 0)
    if idx.size:
        assert idx.min() >= base and idx.max() < base + n_inner
        fib = np.repeat(np.arange(n_outer), np.diff(ptr))
        same = fib[1:] == fib[:-1]
        assert np.all(np.diff(idx)[same] > 0)


@pytest.mark.parametrize("m,n,p", [(10_000, 10_000, 1e-3), (300, 2000, 0.01), (50, 40, 0.3)])
def test_fsprand_density_and_layout(m, n, p):
    ptr, idx, val = synth.fsprand(m, n, p, seed=0)
    check_compressed(ptr, idx, n, m)          # CSC: n columns of row indices in 1..m
    mean, sd = m * n * p, np.sqrt(m * n * p * (1 - p))
    assert abs(idx.size - mean) <= 5 * sd    # nnz ~ Binomial(m*n, p)
    assert val.dtype == np.float64 and np.all((val >= 0) & (val < 1))   # rand


def test_fsprand_uniform_coordinates():
    # Uniform over the m*n cells: column counts and row counts are flat.
    m, n = 2000, 500
    ptr, idx, _ = synth.fsprand(m, n, 0.05, seed=3)
    cols = np.diff(ptr)
    rows = np.bincount(idx - 1, minlength=m)
    assert abs(cols.mean() - m * 0.05) < 3 * np.sqrt(m * 0.05 / n)
    assert cols.std() < 2 * np.sqrt(m * 0.05)
    assert rows.std() < 2 * np.sqrt(n * 0.05)


def test_fsprand_is_seeded_and_zero_based_option():
    a = synth.fsprand(500, 400, 0.02, seed=11)
    b = synth.fsprand(500, 400, 0.02, seed=11)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    p0, i0, v0 = synth.fsprand(500, 400, 0.02, seed=11, base=0)
    assert np.array_equal(p0 + 1, a[0]) and np.array_equal(i0 + 1, a[1])
    assert np.array_equal(v0, a[2])
    check_compressed(p0, i0, 400, 500, base=0)


def test_powerlaw_rows_scaled_and_capped():
    m, target, cap = 200_000, 3_000_000, 20_000
    ptr, idx, val = synth.powerlaw_csr(m, m, target, cap=cap, seed=0)
    check_compressed(ptr, idx, m, m)
    lens = np.diff(ptr)
    assert abs(idx.size - target) < 0.02 * target      # scaled to the target
    assert lens.max() <= cap
    # Heavy tail: the longest rows hold a large share of the entries, and most
    # rows are short or empty (Zipf(1.5)).
    top = np.sort(lens)[::-1][: m // 1000].sum()
    assert top > 0.1 * idx.size
    assert np.median(lens) < lens.mean()
    assert val.dtype == np.float32 and abs(float(val.mean())) < 0.01


def test_poisson_rows():
    m, n, mu = 50_000, 200_000, 24.0
    ptr, idx, val = synth.poisson_csr(m, n, mu, seed=4)
    check_compressed(ptr, idx, m, n)
    lens = np.diff(ptr)
    assert abs(lens.mean() - mu) < 0.1
    assert abs(lens.var() - mu) < 0.5                   # Poisson: variance = mean
