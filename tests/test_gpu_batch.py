"""GPU parity of the batched persistent form against the oracle, per item.

The batched launch (row-split or nnz-split per item, replayed from a
hipGraph) must give every item the same product as the oracle within the
FP32 tolerance, the same bits on every replay and with or without the graph,
and BF16-stored values must stay within 1e-2 normwise of the FP32 result.
"""
import numpy as np
import pytest
import torch

import oracle
from finchgpu import _lib as L
from finchgpu import batch, synth
from finchgpu.ops import Compressed

pytestmark = pytest.mark.gpu

FP32_NORMWISE = 1e-5


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300)


def host_ref(item):
    A = item.A
    ptr, idx = A.ptr.cpu().numpy(), A.idx.cpu().numpy()
    val = A.val.float().cpu().numpy().astype(np.float64)
    B = item.B.float().cpu().numpy().astype(np.float64)
    if B.shape[1] == 1:
        return oracle.spmv_csr(ptr, idx, val, B[:, 0].copy())[:, None]
    return oracle.spmm_csr(ptr, idx, val, B)


@pytest.fixture(scope="module")
def small_batch(dev):
    # The batched distribution at reduced row counts so the oracle stays fast.
    return batch.synth_batch(40, seed=3, device=dev, rows=(256, 4096))


def skewed_items(dev):
    """Power-law rows (the autotuner must pick nnz-split for these)."""
    items = []
    for s, k in enumerate((1, 8, 32)):
        ptr, idx, val = synth.powerlaw_csr(3000, 3000, 60_000, cap=2500, seed=s)
        A = Compressed(torch.from_numpy(ptr.astype(np.int32)).to(dev),
                       torch.from_numpy(idx.astype(np.int32)).to(dev),
                       torch.from_numpy(val).to(dev), 3000, 3000)
        B = torch.randn((3000, k), device=dev)
        items.append(batch.BatchItem(A, B))
    return items


def wide_even_items(dev):
    """Wide dense operands over even, long rows: the autotuner picks row-split."""
    return [batch.BatchItem(batch.synth_matrix(1500, 40.0, 70 + i, dev),
                            torch.randn((1500, 128), device=dev)) for i in range(3)]


def test_autotuner_picks_rowsplit_for_wide_even_rows(dev, small_batch):
    items = wide_even_items(dev) + skewed_items(dev)
    plan = batch.BatchPlan(items)
    st = plan.stats()
    assert st["items_rowsplit"] == 3 and st["items_nnzsplit"] == 3
    plan.run()
    torch.cuda.synchronize()
    for it in items:
        assert rel_err(it.C.cpu().numpy(), host_ref(it)) <= FP32_NORMWISE
    plan.close()


@pytest.mark.parametrize("flags", [0, L.BATCH_NO_GRAPH, L.BATCH_FORCE_NNZSPLIT,
                                   L.BATCH_FORCE_ROWSPLIT])
def test_batch_matches_oracle(dev, small_batch, flags):
    items = small_batch + skewed_items(dev)
    plan = batch.BatchPlan(items, flags)
    st = plan.stats()
    assert st["items_rowsplit"] + st["items_nnzsplit"] == len(items)
    if flags == 0:
        assert st["items_nnzsplit"] >= 3          # the power-law items
    plan.run()
    torch.cuda.synchronize()
    for it in items:
        assert rel_err(it.C.cpu().numpy(), host_ref(it)) <= FP32_NORMWISE
    plan.close()


def test_graph_replays_are_identical(dev, small_batch):
    items = small_batch + skewed_items(dev)
    plan = batch.BatchPlan(items)
    plan.run()
    first = [it.C.clone() for it in items]
    for _ in range(3):
        for it in items:
            it.C.zero_()                 # eager work between replays
        torch.cuda.synchronize()
        plan.run()
        torch.cuda.synchronize()
        for it, f in zip(items, first):
            assert torch.equal(it.C, f)
    eager = batch.BatchPlan(items, L.BATCH_NO_GRAPH)
    eager.run()
    torch.cuda.synchronize()
    for it, f in zip(items, first):
        assert torch.equal(it.C, f)
    plan.close()
    eager.close()


def test_loop_path_matches(dev, small_batch):
    batch.run_loop(small_batch)
    torch.cuda.synchronize()
    for it in small_batch:
        assert rel_err(it.C.cpu().numpy(), host_ref(it)) <= FP32_NORMWISE


def test_bf16_values(dev, small_batch):
    items32 = small_batch
    plan32 = batch.BatchPlan(items32)
    plan32.run()
    items16 = [batch.BatchItem(it.A.to_values(torch.bfloat16), it.B) for it in items32]
    plan16 = batch.BatchPlan(items16)
    plan16.run()
    torch.cuda.synchronize()
    for a, b in zip(items32, items16):
        assert rel_err(b.C.cpu().numpy(), a.C.cpu().numpy()) <= 1e-2
        # FP32 accumulation of the BF16-rounded values: oracle on those values.
        assert rel_err(b.C.cpu().numpy(), host_ref(b)) <= FP32_NORMWISE
    plan32.close()
    plan16.close()


@pytest.mark.parametrize("flags", [0, L.BATCH_FORCE_NNZSPLIT, L.BATCH_FORCE_ROWSPLIT])
def test_bf16_values_and_inputs(dev, small_batch, flags):
    """BF16 values and BF16 right-hand sides, FP32 accumulation: every item equals
    the FP32 oracle on the BF16-rounded operands, and stays within 1e-2 of the
    FP32 plan."""
    items = small_batch + skewed_items(dev)
    plan32 = batch.BatchPlan(items)
    plan32.run()
    items16 = [batch.BatchItem(it.A.to_values(torch.bfloat16), it.B.to(torch.bfloat16))
               for it in items]
    plan16 = batch.BatchPlan(items16, flags)
    plan16.run()
    torch.cuda.synchronize()
    for a, b in zip(items, items16):
        assert rel_err(b.C.cpu().numpy(), host_ref(b)) <= FP32_NORMWISE
        assert rel_err(b.C.cpu().numpy(), a.C.cpu().numpy()) <= 1e-2
    first = [it.C.clone() for it in items16]
    plan16.run()
    torch.cuda.synchronize()
    for it, f in zip(items16, first):
        assert torch.equal(it.C, f)
    plan32.close()
    plan16.close()


def test_bf16_inputs_refused_without_bf16_values(dev, small_batch):
    it = small_batch[0]
    with pytest.raises(L.UnsupportedError):
        batch.BatchPlan([batch.BatchItem(it.A, it.B.to(torch.bfloat16))])
    a16 = it.A.to_values(torch.bfloat16)
    mixed = [batch.BatchItem(a16, it.B.to(torch.bfloat16)), batch.BatchItem(a16, it.B)]
    with pytest.raises(L.UnsupportedError):
        batch.BatchPlan(mixed)
    with pytest.raises(L.UnsupportedError):
        batch.run_loop([batch.BatchItem(a16, it.B.to(torch.bfloat16))])



@pytest.mark.parametrize("bf16", [False, True])
def test_widths_of_four_and_eight_columns_per_lane(dev, bf16):
    """Widths of 16 and more that are multiples of 8 read 8 columns per lane
    (k = 20: 4 per lane; BF16 rows of width 8: one 8-column lane per
    position), the others 4: every item equals the oracle, on both the
    nnz-split and the row-split bodies."""
    items = []
    for i, k in enumerate((4, 8, 12, 16, 20, 24, 40, 136)):
        A = batch.synth_matrix(700 + 37 * i, 30.0, 200 + i, dev)
        if bf16:
            A = A.to_values(torch.bfloat16)
        B = torch.randn((A.n_inner, k), device=dev)
        items.append(batch.BatchItem(A, B.to(torch.bfloat16) if bf16 else B))
    for flags in (L.BATCH_FORCE_NNZSPLIT, L.BATCH_FORCE_ROWSPLIT):
        plan = batch.BatchPlan(items, flags)
        plan.run()
        torch.cuda.synchronize()
        for it in items:
            assert rel_err(it.C.cpu().numpy(), host_ref(it)) <= FP32_NORMWISE, (flags, it.n_rhs)
        plan.close()


def test_measured_tuner_matches_oracle(dev, small_batch):
    """FGX_BATCH_MEASURE times both forms on each bucket of row-split
    candidates before choosing; whatever it picks, every item equals the
    oracle, and the plan replays to the same bits."""
    items = small_batch + wide_even_items(dev) + skewed_items(dev)
    plan = batch.BatchPlan(items, L.BATCH_MEASURE)
    st = plan.stats()
    assert st["items_nnzsplit"] >= 3          # skewed rows are never measured
    plan.run()
    torch.cuda.synchronize()
    first = [it.C.clone() for it in items]
    for it in items:
        assert rel_err(it.C.cpu().numpy(), host_ref(it)) <= FP32_NORMWISE
    plan.run()
    torch.cuda.synchronize()
    for it, f in zip(items, first):
        assert torch.equal(it.C, f)
    plan.close()
