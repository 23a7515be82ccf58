"""GPU tests of the serving loop: continuous batching gives every request the
oracle's result (FP32 tolerance), the same bits alone as in the mix, the plain
per-call path agrees within tolerance, BF16-stored values stay within 1e-2 of
FP32, and admission rejects requests larger than the HBM budget."""
import numpy as np
import pytest
import torch

import oracle
from finchgpu import _lib as L
from finchgpu import batch, serve, synth
from finchgpu.ops import Compressed

pytestmark = pytest.mark.gpu

FP32_NORMWISE = 1e-5


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300)


@pytest.fixture(scope="module")
def operands(dev):
    ops = [batch.synth_matrix(int(m), float(mu), 100 + i, dev)
           for i, (m, mu) in enumerate(zip(np.geomspace(256, 4096, 20).astype(int),
                                           np.geomspace(4, 64, 20)))]
    for s in range(2):
        ptr, idx, val = synth.powerlaw_csr(20_000, 20_000, 400_000, cap=20_000, seed=50 + s)
        ops.append(Compressed(torch.from_numpy(ptr.astype(np.int32)).to(dev),
                              torch.from_numpy(idx.astype(np.int32)).to(dev),
                              torch.from_numpy(val).to(dev), 20_000, 20_000, 1))
    return ops


def reference(op: Compressed, r: serve.Request):
    ptr, idx = op.ptr.cpu().numpy(), op.idx.cpu().numpy()
    val = op.val.float().cpu().numpy().astype(np.float64)
    x = r.inp.cpu().numpy().astype(np.float64)
    if r.kind == "spmv":
        return oracle.spmv_csr(ptr, idx, val, x)
    if r.kind == "spmm":
        return oracle.spmm_csr(ptr, idx, val, x)
    return oracle.sddmm_csr(ptr, idx, val, x, r.inp2.cpu().numpy().astype(np.float64))


def test_batched_results_and_alone_vs_mix(dev, operands):
    inputs = serve.Inputs(operands, seed=1)
    reqs = serve.make_trace(operands, inputs, qps=50_000, n=300, seed=2, outputs=True)
    res = serve.serve(operands, reqs, "batched", max_batch=64)
    assert (res.status == 0).all()
    assert res.summary()["launches"] < len(reqs)          # requests were fused
    for r in reqs:
        assert rel_err(r.out.cpu().numpy(), reference(operands[r.operand], r)) <= FP32_NORMWISE
    # Each request alone gives the same bits as in the mix.
    mixed = [r.out.clone() for r in reqs]
    for i in range(0, len(reqs), 10):
        r = reqs[i]
        alone = serve.Request(r.kind, r.operand, r.k, 0.0, r.inp, r.inp2,
                              torch.full_like(r.out, float("nan")))
        one = serve.serve(operands, [alone], "batched")
        assert (one.status == 0).all()
        assert torch.equal(alone.out, mixed[i])


def test_plain_path_agrees(dev, operands):
    inputs = serve.Inputs(operands, seed=1)
    reqs = serve.make_trace(operands, inputs, qps=20_000, n=120, seed=3, outputs=True)
    serve.serve(operands, reqs, "batched")
    got = [r.out.clone() for r in reqs]
    res = serve.serve(operands, reqs, "plain")
    assert (res.status == 0).all()
    assert (np.diff(res.batch[np.argsort(res.arrival_s, kind="stable")]) > 0).all()
    for r, g in zip(reqs, got):
        assert rel_err(r.out.cpu().numpy(), g.cpu().numpy()) <= FP32_NORMWISE


def test_bf16_operands_within_1e2(dev, operands):
    inputs = serve.Inputs(operands, seed=1)
    reqs = serve.make_trace(operands, inputs, qps=50_000, n=150, seed=4, outputs=True)
    serve.serve(operands, reqs, "batched")
    ref = [r.out.clone() for r in reqs]
    ops16 = [A.to_values(torch.bfloat16) for A in operands]
    res = serve.serve(ops16, reqs, "batched")
    assert (res.status == 0).all()
    for r, g in zip(reqs, ref):
        assert rel_err(r.out.cpu().numpy(), g.cpu().numpy()) <= 1e-2


def test_bf16_values_and_inputs_within_1e2_of_fp64(dev, operands):
    # The all-BF16 form: BF16 operand values and BF16 dense inputs, FP32
    # accumulation and FP32 outputs.  Each result within 1e-2 (normwise) of
    # the FP64 oracle on the unrounded FP32 inputs, within FP32 tolerance of
    # the oracle on the BF16-rounded ones, and the same bits alone and in the mix.
    in32 = serve.Inputs(operands, seed=1)
    in16 = serve.Inputs(operands, seed=1, dtype=torch.bfloat16)
    r32 = serve.make_trace(operands, in32, qps=50_000, n=200, seed=6, outputs=True)
    r16 = serve.make_trace(operands, in16, qps=50_000, n=200, seed=6, outputs=True)
    ops16 = [A.to_values(torch.bfloat16) for A in operands]
    res = serve.serve(ops16, r16, "batched", max_batch=64, input_dtype=torch.bfloat16)
    assert (res.status == 0).all()
    assert res.summary()["launches"] < len(r16)
    for a, b in zip(r32, r16):
        assert (a.kind, a.operand, a.k) == (b.kind, b.operand, b.k)
        assert rel_err(b.out.cpu().numpy(), reference(operands[a.operand], a)) <= 1e-2
        # The same sums in FP64 on the BF16-rounded values and inputs.
        r64 = serve.Request(b.kind, b.operand, b.k, 0.0, b.inp.float(),
                            b.inp2.float() if b.inp2 is not None else None)
        assert rel_err(b.out.cpu().numpy(), reference(ops16[b.operand], r64)) <= FP32_NORMWISE
    mixed = [r.out.clone() for r in r16]
    for i in range(0, len(r16), 15):
        r = r16[i]
        alone = serve.Request(r.kind, r.operand, r.k, 0.0, r.inp, r.inp2,
                              torch.full_like(r.out, float("nan")))
        one = serve.serve(ops16, [alone], "batched", input_dtype=torch.bfloat16)
        assert (one.status == 0).all()
        assert torch.equal(alone.out, mixed[i])


def test_bf16_inputs_refused_where_unsupported(dev, operands):
    in16 = serve.Inputs(operands, seed=1, dtype=torch.bfloat16)
    r16 = serve.make_trace(operands, in16, qps=50_000, n=5, seed=7, outputs=True)
    with pytest.raises(L.UnsupportedError):          # FP32 values with BF16 inputs
        serve.serve(operands, r16, "batched", input_dtype=torch.bfloat16)
    ops16 = [A.to_values(torch.bfloat16) for A in operands]
    with pytest.raises(L.UnsupportedError):          # the plain path is FP32 only
        serve.serve(ops16, r16, "plain", input_dtype=torch.bfloat16)



@pytest.mark.parametrize("bf16", [False, True])
def test_widths_of_four_and_eight_columns_per_lane(dev, operands, bf16):
    # SpMM widths 12 (4 columns per lane), 24 and 40 (8 per lane) and SDDMM
    # 24, in FP32 and in the all-BF16 form, each equal to the oracle.
    ops_ = [A.to_values(torch.bfloat16) for A in operands] if bf16 else operands
    dt = torch.bfloat16 if bf16 else torch.float32
    inp = serve.Inputs(operands, seed=3, dtype=dt)
    reqs = serve.make_trace(operands, inp, qps=50_000, n=120, seed=8, outputs=True,
                            spmm_k=(8, 12, 24, 40), sddmm_k=24)
    res = serve.serve(ops_, reqs, "batched", max_batch=64, input_dtype=dt)
    assert (res.status == 0).all()
    for r in reqs:
        r64 = serve.Request(r.kind, r.operand, r.k, 0.0, r.inp.float(),
                            r.inp2.float() if r.inp2 is not None else None)
        assert rel_err(r.out.cpu().numpy(), reference(ops_[r.operand], r64)) <= FP32_NORMWISE


def test_admission_budget_and_slots(dev, operands):
    inputs = serve.Inputs(operands, seed=1)
    reqs = serve.make_trace(operands, inputs, qps=50_000, n=80, seed=5, outputs=True)
    sizes = []
    for r in reqs:
        A = operands[r.operand]
        fp = r.inp.numel() * 4 + (r.inp2.numel() * 4 if r.inp2 is not None else 0) + r.out.numel() * 4
        sizes.append(fp)
    budget = int(np.median(sizes))
    res = serve.serve(operands, reqs, "batched", hbm_budget_bytes=budget, max_slots=2)
    big = np.array(sizes) > budget
    assert (res.status[big] == L.EADMIT).all()
    assert (res.status[~big] == 0).all()
    for r, b in zip(reqs, big):
        if not b:
            assert rel_err(r.out.cpu().numpy(), reference(operands[r.operand], r)) <= FP32_NORMWISE
    # With two slots, at most two requests are ever in flight.
    ok = res.status == 0
    for t in res.admit_s[ok]:
        inflight = ((res.admit_s[ok] <= t) & (res.done_s[ok] > t)).sum()
        assert inflight <= 2
