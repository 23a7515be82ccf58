"""GPU parity at full scale for the batched and serving forms.

* The whole 4,096-item batched workload (rows log-uniform in [256, 65,536],
  Poisson rows with mean log-uniform in [4, 64], n_rhs in {1, 8, 32, 128};
  about 10^9 stored entries) runs as one graph-replayed plan; every 64th item
  and the 8 largest are checked against the FP64 oracle (FP32 normwise 1e-5),
  and the BF16-value plan against FP64 within 1e-2.
* The serving loop runs SpMV, SpMM (k 16, 64) and SDDMM (k 32) requests on a
  1M x 1M Zipf(1.5) operand (16M stored entries) mixed with small requests:
  each large result matches the oracle, the same bits alone and in the mix.
"""
import numpy as np
import pytest
import torch

import oracle
from finchgpu import batch, serve, synth
from finchgpu.ops import Compressed

pytestmark = pytest.mark.gpu

NTH = 16


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300)


def _host(A: Compressed, values=None):
    v = (A.val if values is None else values).float().cpu().numpy().astype(np.float64)
    return A.ptr.cpu().numpy(), A.idx.cpu().numpy(), v


def test_full_batched_plan_sample_vs_oracle(dev):
    items = batch.synth_batch(4096, seed=0, device=dev)
    nnz = sum(it.A.nnz for it in items)
    assert nnz > 9e8
    plan = batch.BatchPlan(items)
    plan.run()
    torch.cuda.synchronize()
    st = plan.stats()
    assert st["items_rowsplit"] + st["items_nnzsplit"] == 4096
    by_size = sorted(range(len(items)), key=lambda i: -items[i].A.nnz)
    sample = sorted(set(range(0, len(items), 64)) | set(by_size[:8]))
    refs = {}
    for i in sample:
        it = items[i]
        p, ix, v = _host(it.A)
        refs[i] = oracle.spmm_csr(p, ix, v, it.B.cpu().numpy().astype(np.float64), nthreads=NTH)
        assert rel_err(it.C.cpu().numpy(), refs[i]) <= 1e-5, i
    plan.close()
    # BF16-stored values, FP32 accumulation: within 1e-2 of FP64 on the FP32
    # values, and an FP32 sum of the BF16-rounded values.
    items16 = [batch.BatchItem(it.A.to_values(torch.bfloat16), it.B) for it in items]
    plan16 = batch.BatchPlan(items16)
    plan16.run()
    torch.cuda.synchronize()
    for i in sample:
        it = items16[i]
        got = it.C.cpu().numpy()
        assert rel_err(got, refs[i]) <= 1e-2, i
        p, ix, v16 = _host(it.A)
        ref16 = oracle.spmm_csr(p, ix, v16, it.B.cpu().numpy().astype(np.float64), nthreads=NTH)
        assert rel_err(got, ref16) <= 1e-5, i
    plan16.close()


@pytest.fixture(scope="module")
def serving_ops(dev):
    small = serve.serving_workload(dev, n_small=24, n_large=0, seed=3)
    ptr, idx, val = synth.powerlaw_csr(1_000_000, 1_000_000, 16_000_000, seed=0)
    big = Compressed(torch.from_numpy(ptr.astype(np.int32)).to(dev),
                     torch.from_numpy(idx.astype(np.int32)).to(dev),
                     torch.from_numpy(val).to(dev), 1_000_000, 1_000_000, 1)
    return small + [big]


def _large_requests(ops, inputs, t0):
    b = len(ops) - 1
    A = ops[b]
    out = []
    for j, (kind, k) in enumerate((("spmv", 1), ("spmm", 64), ("spmm", 16), ("sddmm", 32))):
        ins = inputs.get(b, kind, k)
        shape = (A.n_outer,) if kind == "spmv" else ((A.nnz,) if kind == "sddmm"
                                                      else (A.n_outer, k))
        out.append(serve.Request(kind, b, k, t0 + 3e-4 * j, ins[0],
                                 ins[1] if len(ins) > 1 else None,
                                 torch.full(shape, float("nan"), device=A.device)))
    return out


def _reference(A, r, values=None):
    p, ix, v = _host(A, values)
    x = r.inp.float().cpu().numpy().astype(np.float64)
    if r.kind == "spmv":
        return oracle.spmv_csr(p, ix, v, x, nthreads=NTH)
    if r.kind == "spmm":
        return oracle.spmm_csr(p, ix, v, x, nthreads=NTH)
    return oracle.sddmm_csr(p, ix, v, x, r.inp2.float().cpu().numpy().astype(np.float64))


@pytest.mark.parametrize("form", ["fp32", "bf16_values_and_inputs"])
def test_serving_with_a_1m_row_operand(dev, serving_ops, form):
    ops = serving_ops
    dt = torch.float32 if form == "fp32" else torch.bfloat16
    O = ops if form == "fp32" else [A.to_values(torch.bfloat16) for A in ops]
    inputs = serve.Inputs(ops, seed=5, dtype=dt)
    small = serve.make_trace(ops[:-1], inputs, qps=40_000, n=200, seed=6, outputs=True)
    large = _large_requests(ops, inputs, 1e-3)
    res = serve.serve(O, small + large, "batched", input_dtype=dt)
    assert (res.status == 0).all()
    big = ops[-1]
    for r in large:
        got = r.out.cpu().numpy()
        if form == "fp32":
            assert rel_err(got, _reference(big, r)) <= 1e-5, r.kind
        else:
            # within 1e-2 of FP64 on the FP32 values and FP32 draws of the inputs
            ins32 = serve.Inputs(ops, seed=5).get(r.operand, r.kind, r.k)
            r32 = serve.Request(r.kind, r.operand, r.k, 0.0, ins32[0],
                                ins32[1] if len(ins32) > 1 else None)
            assert rel_err(got, _reference(big, r32)) <= 1e-2, r.kind
            # and an FP32 sum of the BF16-rounded operands
            assert rel_err(got, _reference(big, r, O[-1].val)) <= 1e-5, r.kind
    mixed = [r.out.clone() for r in large]
    for r, m in zip(large, mixed):
        alone = serve.Request(r.kind, r.operand, r.k, 0.0, r.inp, r.inp2,
                              torch.full_like(r.out, float("nan")))
        one = serve.serve(O, [alone], "batched", input_dtype=dt)
        assert (one.status == 0).all()
        assert torch.equal(alone.out, m), r.kind
