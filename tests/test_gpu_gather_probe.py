"""GPU: the benchmark's gather-floor probe (libfgx_probe.so) reads the rows it
is asked for.

The probe stands beside SpMM's timing as the floor its B-row gathers impose,
so it must gather exactly SpMM's rows: the sum of its per-lane outputs equals
the sum of B[idx[q] - base, :] over all positions q (an FP64 reduction of the
same rows, compared at FP32 tolerance).  Shapes the probe does not build are
rejected on the host before any launch.
"""
import ctypes
import os

import numpy as np
import pytest
import torch

from conftest import REPO

pytestmark = pytest.mark.gpu

PROBE = os.path.join(REPO, "finch.jl_amd", "finchgpu", "libfgx_probe.so")


@pytest.fixture(scope="module")
def probe(dev):
    if not os.path.exists(PROBE):
        pytest.fail(f"{PROBE} is not built")
    lib = ctypes.CDLL(PROBE)
    lib.fgx_probe_gather_out_len.restype = ctypes.c_int64
    lib.fgx_probe_gather_out_len.argtypes = [ctypes.c_int]
    lib.fgx_probe_gather_rows.restype = ctypes.c_int
    lib.fgx_probe_gather_rows.argtypes = [
        ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
        ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
        ctypes.c_int, ctypes.c_void_p]
    return lib


def run(lib, B, idx, base, lanes, u, bpc):
    n = lib.fgx_probe_gather_out_len(bpc)
    out = torch.zeros(n, dtype=torch.float32, device=B.device)
    rc = lib.fgx_probe_gather_rows(B.data_ptr(), B.stride(0) // 4, lanes, idx.data_ptr(),
                                   idx.element_size(), base, idx.numel(), out.data_ptr(), n, u,
                                   bpc, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return rc, out


@pytest.mark.parametrize("lanes,u,bpc,itype,base", [
    (16, 8, 8, torch.int32, 1), (16, 16, 16, torch.int32, 0), (16, 32, 8, torch.int64, 1),
    (8, 16, 8, torch.int32, 1), (32, 8, 16, torch.int64, 0)])
def test_probe_sums_the_requested_rows(probe, dev, lanes, u, bpc, itype, base):
    rng = np.random.default_rng(lanes * 100 + u)
    rows, npos = 3000, 50_001           # a ragged last chunk
    width = lanes * 4
    Bh = rng.standard_normal((rows, width)).astype(np.float32)
    ih = rng.integers(0, rows, npos)
    B = torch.from_numpy(Bh).to(dev)
    idx = torch.from_numpy(ih + base).to(itype).to(dev)
    rc, out = run(probe, B, idx, base, lanes, u, bpc)
    assert rc == 0
    want = Bh.astype(np.float64)[ih].sum()
    got = out.double().sum().item()
    scale = np.abs(Bh.astype(np.float64)[ih]).sum()
    assert abs(got - want) <= 1e-5 * scale, (got, want)


def test_probe_rejects_unbuilt_shapes(probe, dev):
    B = torch.zeros((16, 64), dtype=torch.float32, device=dev)
    idx = torch.ones(4, dtype=torch.int32, device=dev)
    assert run(probe, B, idx, 1, 16, 7, 8)[0] != 0      # u not built
    assert run(probe, B, idx, 1, 12, 8, 8)[0] != 0      # lanes not built
    n = probe.fgx_probe_gather_out_len(8)
    out = torch.zeros(n - 1, dtype=torch.float32, device=dev)
    rc = probe.fgx_probe_gather_rows(B.data_ptr(), 16, 16, idx.data_ptr(), 4, 1, 4,
                                     out.data_ptr(), n - 1, 8, 8,
                                     torch.cuda.current_stream().cuda_stream)
    assert rc != 0                                      # out too short
