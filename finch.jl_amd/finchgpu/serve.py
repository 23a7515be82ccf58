"""Serving form: SpMV / SpMM / SDDMM requests against device-resident operands.

``serve()`` replays an open-loop request trace through ``fgx_serve_run``:
continuous batching (mode "batched": requests admitted by free slots and HBM
footprint are fused into persistent launches on several streams) or the plain
per-call path (mode "plain": one library call and synchronisation per request).
Latencies are measured by the library's scheduler on the host clock, from a
request's arrival time to the moment its launch is seen complete.

``serving_workload`` builds the operand set of the serving benchmark (batched
size distribution plus 1M-row power-law matrices); ``make_trace`` draws Poisson
arrivals with the 60 % SpMV / 30 % SpMM (k 16 or 64) / 10 % SDDMM (k 32) mix.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L
from .ops import Compressed, _ptr, context

KINDS = {"spmv": L.REQ_SPMV, "spmm": L.REQ_SPMM, "sddmm": L.REQ_SDDMM}


@dataclass
class Request:
    kind: str                 # "spmv", "spmm" or "sddmm"
    operand: int
    k: int
    arrival_s: float
    inp: torch.Tensor         # x / B / X
    inp2: torch.Tensor | None = None   # Y (SDDMM)
    out: torch.Tensor | None = None    # None: the server's own output arena


@dataclass
class Results:
    arrival_s: np.ndarray
    admit_s: np.ndarray
    done_s: np.ndarray
    status: np.ndarray
    batch: np.ndarray
    extra: dict = field(default_factory=dict)

    @property
    def latency_s(self) -> np.ndarray:
        return self.done_s - self.arrival_s

    def summary(self) -> dict:
        ok = self.status == 0
        lat = self.latency_s[ok]
        span = float(self.done_s[ok].max() - self.arrival_s.min()) if ok.any() else float("nan")
        return {
            "requests": int(ok.sum()),
            "rejected": int((~ok).sum()),
            "req_per_s": float(ok.sum() / span) if span > 0 else float("nan"),
            "p50_ms": float(np.percentile(lat, 50) * 1e3) if lat.size else float("nan"),
            "p99_ms": float(np.percentile(lat, 99) * 1e3) if lat.size else float("nan"),
            "launches": int(np.unique(self.batch[ok]).size),
        }


def serve(operands: list[Compressed], requests: list[Request], mode: str = "batched",
          hbm_budget_bytes: int = 200 * 10**9, max_slots: int = 4096,
          max_batch: int = 512, input_dtype: torch.dtype = torch.float32) -> Results:
    """input_dtype torch.bfloat16: the requests' dense inputs are BF16 (with
    BF16-stored operand values, batched mode), read at half the bytes and
    accumulated in FP32; outputs stay FP32."""
    lib = L.load()
    in_type = {torch.float32: 0, torch.bfloat16: L.BF16}[input_dtype]
    for r in requests:
        for t in (r.inp, r.inp2):
            if t is not None and t.dtype != input_dtype:
                raise L.UnsupportedError(L.EUNSUP, f"request input is {t.dtype}, "
                                                   f"expected {input_dtype}")
    dev = operands[0].device
    ops_arr = (L.Compressed * len(operands))(*[A.desc() for A in operands])
    req_arr = (L.Request * len(requests))()
    for i, r in enumerate(requests):
        req_arr[i] = L.Request(KINDS[r.kind], r.operand, r.k if r.kind != "spmv" else 1,
                               float(r.arrival_s), _ptr(r.inp), _ptr(r.inp2), _ptr(r.out))
    cfg = L.ServeConfig(int(hbm_budget_bytes), int(max_slots), int(max_batch),
                        0 if mode == "batched" else 1, in_type)
    res = (L.RequestResult * len(requests))()
    torch.cuda.synchronize(dev)
    L.check(lib.fgx_serve_run(context(dev).handle, ops_arr, len(operands), req_arr, len(requests),
                              ctypes.byref(cfg), res), "fgx_serve_run")
    return Results(np.array([r.arrival_s for r in requests]),
                   np.array([x.admit_s for x in res]), np.array([x.done_s for x in res]),
                   np.array([x.status for x in res]), np.array([x.batch for x in res]))


class Inputs:
    """Shared, device-resident dense inputs per (operand, kind, width); with
    dtype torch.bfloat16 they are the FP32 draws rounded to BF16 on the
    device (fgx_convert)."""

    def __init__(self, operands, seed=0, dtype=torch.float32):
        self.operands = operands
        self.seed = seed
        self.dtype = dtype
        self._cache = {}

    def get(self, op, kind, k):
        key = (op, kind, k)
        t = self._cache.get(key)
        if t is None:
            A = self.operands[op]
            g = torch.Generator(device=A.device)
            g.manual_seed(self.seed * 1_000_003 + op * 97 + k * 7 + len(kind))
            if kind == "spmv":
                t = (torch.randn(A.n_inner, device=A.device, generator=g),)
            elif kind == "spmm":
                t = (torch.randn((A.n_inner, k), device=A.device, generator=g),)
            else:
                t = (torch.randn((A.n_outer, k), device=A.device, generator=g),
                     torch.randn((A.n_inner, k), device=A.device, generator=g))
            if self.dtype != torch.float32:
                from .ops import convert
                t = tuple(convert(u.contiguous(), self.dtype) for u in t)
            self._cache[key] = t
        return t


def make_trace(operands, inputs: Inputs, qps: float, n: int, seed: int = 0,
               mix=(0.6, 0.3, 0.1), spmm_k=(16, 64), sddmm_k=32, outputs=False):
    """n requests with Poisson(qps) arrivals, operands uniform over the set."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / qps, size=n))
    t -= t[0]
    reqs = []
    for i in range(n):
        op = int(rng.integers(len(operands)))
        u = rng.random()
        if u < mix[0]:
            kind, k = "spmv", 1
        elif u < mix[0] + mix[1]:
            kind, k = "spmm", int(rng.choice(spmm_k))
        else:
            kind, k = "sddmm", sddmm_k
        ins = inputs.get(op, kind, k)
        A = operands[op]
        out = None
        if outputs:
            shape = (A.n_outer,) if kind == "spmv" else ((A.nnz,) if kind == "sddmm"
                                                          else (A.n_outer, k))
            out = torch.empty(shape, dtype=torch.float32, device=A.device)
        reqs.append(Request(kind, op, k, float(t[i]), ins[0], ins[1] if len(ins) > 1 else None,
                            out))
    return reqs


def serving_workload(device, n_small=256, n_large=4, seed=0, value_dtype=torch.float32,
                     large_rows=1_000_000, large_nnz=16_000_000):
    """Operands: n_small matrices of the batched size distribution plus n_large
    power-law (Zipf 1.5) matrices of large_rows rows."""
    from . import batch, synth
    rng = np.random.default_rng(seed)
    ops = []
    for i in range(n_small):
        m = int(round(np.exp(rng.uniform(np.log(256), np.log(65536)))))
        mu = float(np.exp(rng.uniform(np.log(4.0), np.log(64.0))))
        ops.append(batch.synth_matrix(m, mu, seed * 7777 + i, device, value_dtype))
    for j in range(n_large):
        ptr, idx, val = synth.powerlaw_csr(large_rows, large_rows, large_nnz, seed=seed + 11 + j)
        ops.append(Compressed(torch.from_numpy(ptr.astype(np.int32)).to(device),
                              torch.from_numpy(idx.astype(np.int32)).to(device),
                              torch.from_numpy(val).to(device).to(value_dtype),
                              large_rows, large_rows, 1))
    return ops
