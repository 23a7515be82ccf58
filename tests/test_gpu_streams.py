"""GPU: calls from several host threads on several streams of one context.

Each stream gets its own device scratch in the context, and a tile table
built on one stream is waited for before another stream reads it, so results
on concurrent streams must equal the serial ones bit for bit (the SpMV, SpMM
and coiteration kernels are deterministic).
"""
import threading

import numpy as np
import pytest
import torch

from finchgpu import ops, synth

pytestmark = pytest.mark.gpu


def _operands(dev):
    out = []
    for s, (m, nnz) in enumerate([(40_000, 600_000), (3_000, 20_000), (120_000, 1_500_000),
                                  (9_000, 90_000)]):
        ptr, idx, val = synth.powerlaw_csr(m, m, nnz, cap=m // 2, seed=10 + s)
        A = ops.to_device(ptr.astype(np.int32), idx.astype(np.int32), val, m, m, device=dev)
        x = torch.as_tensor(synth.dense(m, 1, seed=20 + s)[:, 0].copy()).to(dev)
        B = torch.as_tensor(synth.dense(m, 32, seed=30 + s)).to(dev)
        out.append((A, x, B))
    return out


def _work(A, x, B):
    y = ops.spmv_csr(A, x)
    C = ops.spmm_csr(A, B)
    E = ops.ewise("add", A, A)
    return y, C, E.ptr, E.idx, E.val


def test_four_threads_four_streams_match_serial(dev):
    opnds = _operands(dev)
    # Threads 0 and 2 share operand 0, so one tile table is built on one
    # stream and read from another.
    assign = [0, 1, 0, 2, 3, 2]
    serial = [[t.cpu() for t in _work(*opnds[a])] for a in assign]
    torch.cuda.synchronize()
    ctx = ops.context(dev)
    from finchgpu import _lib as L
    L.check(L.load().fgx_plan_cache_clear(ctx.handle))   # tables rebuilt under concurrency
    streams = [torch.cuda.Stream(device=dev) for _ in assign]
    results = [None] * len(assign)
    errors = []

    def run(t):
        try:
            with torch.cuda.stream(streams[t]):
                outs = None
                for _ in range(12):
                    outs = _work(*opnds[assign[t]])
                streams[t].synchronize()
                results[t] = [o.cpu() for o in outs]
        except Exception as e:  # reported below
            errors.append((t, repr(e)))

    threads = [threading.Thread(target=run, args=(t,)) for t in range(len(assign))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    torch.cuda.synchronize()
    assert not errors, errors
    for t, (want, got) in enumerate(zip(serial, results)):
        for w, g in zip(want, got):
            assert torch.equal(w, g), f"thread {t} differs from the serial result"
