"""Seeded synthetic operands for the benchmark and parity workloads.

Finch draws its test matrices with ``fsprand`` (src/interface/fsparse.jl:173-290):
nnz ~ Binomial(m*n, p), distinct uniform coordinates, sorted column-major,
values ``rand``/``randn``.  Julia's Xoshiro/Binomial streams cannot be
reproduced outside Julia, so these generators follow the same distributions
with numpy's PCG64 (``numpy.random.default_rng(seed)``); the seeds name the
operands, not bit-identical copies of Finch's.

All generators return host (numpy) buffers in Finch's 1-based compressed layout
unless ``base=0`` is passed: ``ptr`` (n_outer+1), ``idx`` (nnz), ``val`` (nnz).
"""
from __future__ import annotations

import numpy as np


def _distinct_sorted(rng, fiber_lengths, n_inner):
    """Distinct uniform coordinates per fiber, sorted within each fiber.

    Draws each fiber's coordinates with replacement, sorts (fiber, coord)
    keys and drops repeats; a fiber then redraws its shortfall until it has
    min(length, n_inner) distinct coordinates.
    """
    lengths = np.minimum(np.asarray(fiber_lengths, dtype=np.int64), n_inner)
    n_outer = lengths.size
    fibers = np.repeat(np.arange(n_outer, dtype=np.int64), lengths)
    cols = rng.integers(0, n_inner, size=fibers.size, dtype=np.int64)
    keys = np.unique(fibers * n_inner + cols)
    for _ in range(64):
        have = np.bincount(keys // n_inner, minlength=n_outer)
        short = lengths - have
        if not short.any():
            break
        f2 = np.repeat(np.arange(n_outer, dtype=np.int64), short)
        c2 = rng.integers(0, n_inner, size=f2.size, dtype=np.int64)
        keys = np.unique(np.concatenate([keys, f2 * n_inner + c2]))
        # A fiber may now hold more than asked: trim the surplus at random.
        have = np.bincount(keys // n_inner, minlength=n_outer)
        extra = have - lengths
        if extra.any():
            keep = np.ones(keys.size, dtype=bool)
            starts = np.concatenate([[0], np.cumsum(have)[:-1]])
            for f in np.nonzero(extra > 0)[0]:
                drop = rng.choice(have[f], size=extra[f], replace=False)
                keep[starts[f] + drop] = False
            keys = keys[keep]
    fib = keys // n_inner
    idx = keys - fib * n_inner
    ptr = np.zeros(n_outer + 1, dtype=np.int64)
    np.cumsum(np.bincount(fib, minlength=n_outer), out=ptr[1:])
    return ptr, idx


def fsprand(m, n, p, seed=0, dtype=np.float64, values="rand", base=1, rng=None):
    """Column-compressed (CSC) m x n matrix with density p, like Finch's fsprand.

    nnz ~ Binomial(m*n, p); coordinates distinct and uniform; values U[0,1)
    (``rand``) or N(0,1) (``randn``).  Returns (ptr, idx, val) over n columns.
    """
    rng = rng or np.random.default_rng(seed)
    nnz = int(rng.binomial(m * n, p))
    # Uniform distinct coordinates == a multinomial split of nnz over columns
    # followed by distinct rows within each column.
    counts = rng.multinomial(nnz, np.full(n, 1.0 / n)) if n > 0 else np.zeros(0, np.int64)
    ptr, idx = _distinct_sorted(rng, counts, m)
    nz = idx.size
    val = (rng.random(nz) if values == "rand" else rng.standard_normal(nz)).astype(dtype)
    return ptr + base, idx + base, val


def zipf_lengths(m, target_nnz, s=1.5, cap=100_000, seed=0, rng=None):
    """Power-law row lengths: Zipf(s) draws truncated at `cap`, scaled to ~target_nnz.

    Each row draws Z ~ Zipf(s) conditioned on Z <= cap; lengths are
    round(c * Z) with c found by bisection so that the total is within 0.1 %
    of `target_nnz` (a raw truncated Zipf(1.5) averages about 242 per row).
    """
    rng = rng or np.random.default_rng(seed)
    z = rng.zipf(s, size=m)
    bad = z > cap
    while bad.any():
        z[bad] = rng.zipf(s, size=int(bad.sum()))
        bad = z > cap
    z = z.astype(np.float64)
    lo, hi = 0.0, 4.0 * target_nnz / max(z.sum(), 1.0) + 1.0
    for _ in range(80):
        c = 0.5 * (lo + hi)
        tot = np.rint(z * c).sum()
        if tot < target_nnz:
            lo = c
        else:
            hi = c
        if abs(tot - target_nnz) <= 1e-3 * target_nnz:
            break
    return np.rint(z * c).astype(np.int64)


def powerlaw_csr(m, n, target_nnz, s=1.5, cap=100_000, seed=0, dtype=np.float32, base=1):
    """CSR m x n with Zipf(s) row lengths (see zipf_lengths), uniform distinct
    columns and N(0,1) values."""
    rng = np.random.default_rng(seed)
    lengths = zipf_lengths(m, target_nnz, s, cap, rng=rng)
    ptr, idx = _distinct_sorted(rng, lengths, n)
    val = rng.standard_normal(idx.size, dtype=np.float32).astype(dtype)
    return ptr + base, idx + base, val


def poisson_csr(m, n, mean, seed=0, dtype=np.float32, values="randn", base=1, rng=None):
    """CSR m x n with Poisson(mean) nonzeros per row and uniform distinct columns.

    values: "randn" (N(0,1)), "int19" (uniform integers 1..9) or "rand" (U[0,1)).
    """
    rng = rng or np.random.default_rng(seed)
    lengths = rng.poisson(mean, size=m)
    ptr, idx = _distinct_sorted(rng, lengths, n)
    nz = idx.size
    if values == "int19":
        val = rng.integers(1, 10, size=nz).astype(dtype)
    elif values == "rand":
        val = rng.random(nz).astype(dtype)
    else:
        val = rng.standard_normal(nz).astype(dtype)
    return ptr + base, idx + base, val


def dense(rows, cols, seed=0, dtype=np.float32, kind="randn"):
    rng = np.random.default_rng(seed)
    if kind == "randn":
        return rng.standard_normal((rows, cols)).astype(dtype)
    if kind == "sym":  # U(-1, 1)
        return (2.0 * rng.random((rows, cols)) - 1.0).astype(dtype)
    return rng.random((rows, cols)).astype(dtype)


def transpose_compressed(ptr, idx, val, n_outer, n_inner, base=1):
    """The same matrix compressed along the other dimension (CSC <-> CSR),
    coordinates sorted within each fiber (a stable counting sort)."""
    ptr0 = np.asarray(ptr, dtype=np.int64) - base
    idx0 = np.asarray(idx, dtype=np.int64) - base
    outer = np.repeat(np.arange(n_outer, dtype=np.int64), np.diff(ptr0))
    order = np.lexsort((outer, idx0))
    tptr = np.zeros(n_inner + 1, dtype=np.int64)
    np.cumsum(np.bincount(idx0, minlength=n_inner), out=tptr[1:])
    return tptr + base, outer[order] + base, np.asarray(val)[order]
