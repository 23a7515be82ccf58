#!/usr/bin/env python3
"""Benchmark: SpMM / SpMV / SDDMM HBM throughput on one MI355X.

Workload (the configuration the headline metric is quoted on): A is a
1,000,000 x 1,000,000 CSR matrix with ~16M stored FP32 values, Zipf(1.5) row
lengths (truncated at 100k, scaled to 16M total), uniform column indices, N(0,1)
values, Int32 ptr/idx; B (= Y) and X are dense 1M x 64 FP32 N(0,1), row-major.
A "step" is one SpMM C = A*B (the headline kernel); SpMV y = A*x and SDDMM
out = A .* (X*Y^T) are timed the same way and reported beside it.

`value` is algorithmic bytes / time for SpMM (GB/s) with every operand resident
in HBM before the timed region; `roofline` compares it with the 8 TB/s HBM
peak.  `cpu_baseline` times the oracle's threaded CSR SpMM (the C restatement
of Finch's row-parallel loop) on this host.

The line also carries the batched form (4,096 uneven matrices: one persistent
hipGraph launch vs a per-matrix loop, BF16 vs FP32 values) and the serving
loop (req/s within a p99 target, batched FP32 / BF16 vs the plain per-call path).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batch]
                       [--no-serve]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "finch.jl_amd"))

HBM_PEAK_GBS = 8000.0
M = N = 1_000_000
NNZ_TARGET = 16_000_000
K = 64


def algorithmic_bytes(m, n, nnz, k, kind, sv=4, si=4, sp=4, sd=4):
    """Bytes a kernel must move at least once (each operand read/written once)."""
    if kind == "spmv":
        return nnz * (sv + si) + (m + 1) * sp + n * sd + m * sd
    if kind == "spmm":
        return nnz * (sv + si) + (m + 1) * sp + n * k * sd + m * k * sd
    if kind == "sddmm":
        return nnz * (2 * sv + si) + (m + 1) * sp + (m + n) * k * sd
    raise ValueError(kind)


def measured_traffic():
    """HBM bytes per SpMM launch, from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE scaled by a copy of known size); None if absent."""
    path = os.path.join(HERE, "profiles", "spmm_csr_powerlaw_1m_k64.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except (OSError, ValueError):
        return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="approximate CPU time budget of the baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-form section")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving section")
    ap.add_argument("--no-sparse", action="store_true",
                    help="skip the coiteration / SpGEMM and 10k FP64 sections")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--serve-slo-ms", type=float, default=5.0,
                    help="p99 latency target at which served req/s are compared")
    ap.add_argument("--kernels-only", action="store_true",
                    help="time only SpMM, SpMV, SDDMM and the vendor comparison (for profiler "
                         "passes): no oracle errors, variants, CPU leg or other sections")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched-form section (FP32, BF16 values, BF16 values and "
                         "inputs) and print its JSON (for per-class profiler passes)")
    ap.add_argument("--no-loop", action="store_true",
                    help="skip the batched section's per-matrix loop (for profiler passes)")
    ap.add_argument("--atomic-probe", action="store_true",
                    help="only time the scatter-add probes (device atomics, LDS atomics)")
    ap.add_argument("--serve-kinds", action="store_true",
                    help="only measure the serving mix's device time per request by operand "
                         "class and kind (saturated batched runs) and print its JSON")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first copy B (256 MB) once: a kernel of known traffic that calibrates "
                         "rocprofv3 byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only: the sparse kernels are "
                                   "measured at their full single-GPU shape; a multi-GPU run "
                                   "would shard rows by nnz and is not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    from finchgpu import ops, synth

    dev = torch.device("cuda:0")
    if args.atomic_probe:
        print(json.dumps({"atomic_probe": bench_atomic_probe(dev)}), flush=True)
        return
    if args.batch_only:
        print(json.dumps({"batched": bench_batched(dev, args.batch_count, not args.no_loop)}),
              flush=True)
        return
    t0 = time.time()
    ptr, idx, val = synth.powerlaw_csr(M, N, NNZ_TARGET, seed=0)
    nnz = int(idx.size)
    print(f"# generated A: nnz={nnz} in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    A = ops.Compressed(torch.from_numpy(ptr.astype(np.int32)).to(dev),
                       torch.from_numpy(idx.astype(np.int32)).to(dev),
                       torch.from_numpy(val).to(dev), M, N, base=1)
    if args.serve_kinds:
        print(json.dumps({"serving_by_kind": bench_serving_by_kind(dev, [A])}), flush=True)
        return
    B = torch.from_numpy(synth.dense(N, K, seed=2)).to(dev)
    X = torch.from_numpy(synth.dense(M, K, seed=3)).to(dev)
    x = B[:, 0].contiguous()
    C = torch.empty((M, K), dtype=torch.float32, device=dev)
    y = torch.empty(M, dtype=torch.float32, device=dev)
    out = torch.empty(nnz, dtype=torch.float32, device=dev)

    if args.copy_probe:
        probe = torch.empty_like(B)
        probe.copy_(B)
        torch.cuda.synchronize()
        del probe

    # Between timed calls a 512 MB read-modify-write pass (outside the timed
    # events) evicts the 256 MB Infinity Cache, so each call reads its
    # operands from HBM.
    flushbuf = torch.zeros(128 << 20, dtype=torch.float32, device=dev)
    flush = lambda: flushbuf.add_(1.0)

    spmm = timed(lambda: ops.spmm_csr(A, B, C), args.steps, args.warmup, flush)
    spmm_warm = timed(lambda: ops.spmm_csr(A, B, C), args.steps, 1)
    print(f"# spmm done: {spmm['event_ms']:.4f} ms/call (flushed)", file=sys.stderr, flush=True)
    spmv = timed(lambda: ops.spmv_csr(A, x, y), args.steps, args.warmup, flush)
    spmv_warm = timed(lambda: ops.spmv_csr(A, x, y), args.steps, 1)
    sddmm = timed(lambda: ops.sddmm_csr(A, X, B, out), args.steps, args.warmup, flush)
    sddmm_warm = timed(lambda: ops.sddmm_csr(A, X, B, out), args.steps, 1)
    print("# spmv/sddmm done", file=sys.stderr, flush=True)
    gather_floor = bench_gather_floor(A, B, flush, min(args.steps, 10))
    print(f"# gather floor: {gather_floor and gather_floor['ms']} ms", file=sys.stderr, flush=True)

    by_spmm = algorithmic_bytes(M, N, nnz, K, "spmm")
    by_spmv = algorithmic_bytes(M, N, nnz, K, "spmv")
    by_sddmm = algorithmic_bytes(M, N, nnz, K, "sddmm")
    gbs = lambda b, ms: b / (ms * 1e-3) / 1e9
    value = gbs(by_spmm, spmm["event_ms"])

    if args.kernels_only:
        args.no_cpu = args.no_sparse = args.no_batch = args.no_serve = True
    # Normwise error of each timed result against the FP64 oracle on the same
    # inputs (threaded C restatement; checker only, outside every timed region).
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle
    nth = max(1, min(16, oracle.max_threads()))
    v64, Bh = val.astype(np.float64), B.cpu().numpy()
    errors = None if args.kernels_only else {
        "spmm": _normwise(C, oracle.spmm_csr(ptr, idx, v64, Bh.astype(np.float64), nthreads=nth)),
        "spmv": _normwise(y, oracle.spmv_csr(ptr, idx, v64, Bh[:, 0].astype(np.float64),
                                             nthreads=nth)),
        "sddmm": _normwise(out, oracle.sddmm_csr(ptr, idx, v64, X.cpu().numpy().astype(np.float64),
                                                 Bh.astype(np.float64))),
    }
    print(f"# errors vs FP64: {errors}", file=sys.stderr, flush=True)

    variants = None
    if not args.kernels_only:
        variants = bench_variants(dev, A, ptr, idx, val, B, x, flush, args.steps, oracle, nth)
        print("# variants done", file=sys.stderr, flush=True)
    vendor = bench_vendor(dev, A, B, x, flush)
    print("# vendor done", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = bench_cpu(oracle, ptr, idx, val, Bh, nnz, by_spmm, by_spmv, args.cpu_seconds)
        print("# cpu done", file=sys.stderr, flush=True)

    sparse = None
    if not args.no_sparse:
        sparse = bench_sparse_sparse(dev, with_cpu=not args.no_cpu)
        sparse["fp64_10k"] = bench_fp64_10k(dev)
        sparse["docs_sddmm_example"] = bench_docs_sddmm(dev)
        sparse["format_conversion"] = bench_format(dev, A, ptr, idx, val, with_cpu=not args.no_cpu)
        print("# sparse-sparse done", file=sys.stderr, flush=True)
    batched = None
    if not args.no_batch:
        del C, X, out
        torch.cuda.empty_cache()
        batched = bench_batched(dev, args.batch_count)
        print("# batched done", file=sys.stderr, flush=True)
    served = None
    if not args.no_serve:
        torch.cuda.empty_cache()
        served = bench_serving(dev, args.serve_slo_ms, [A])
        print("# serving done", file=sys.stderr, flush=True)

    traffic = measured_traffic()
    # The rate at which the call moves its measured HBM traffic (the B-row
    # gathers that the algorithmic count credits once): how close the kernel
    # runs to the bandwidth limit, beside the algorithmic fraction.
    traffic_gbs = traffic / (spmm["event_ms"] * 1e-3) / 1e9 if traffic else None
    line = {
        "metric": "SpMM/SpMV HBM GB/s (% of peak); served req/s at p99, batched vs plain",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(spmm["event_ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded numpy PCG64; Zipf(1.5) rows, uniform columns, N(0,1))",
        "config": {"workload": "SpMM C=A*B: A 1Mx1M CSR FP32 Int32 idx, "
                               f"nnz={nnz}, Zipf(1.5) rows; B 1Mx64 FP32 row-major",
                   "m": M, "n": N, "nnz": nnz, "k": K, "index": "int32"},
        "timing": "HIP events per call on the calling stream; a 512 MB pass between calls "
                  "evicts the 256 MB Infinity Cache (value, ms_per_step); warm_* are "
                  "back-to-back calls",
        "pct_of_hbm_peak": round(100.0 * value / HBM_PEAK_GBS, 2),
        "roofline": {"bound": "hbm", "achieved": round(value, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(value / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_gbs": round(traffic_gbs, 1) if traffic_gbs else None,
                     "traffic_frac": round(traffic_gbs / HBM_PEAK_GBS, 4) if traffic_gbs else None,
                     "traffic_source": "profiles/spmm_csr_powerlaw_1m_k64.md (rocprofv3 "
                                       "FETCH_SIZE x2.00 + WRITE_SIZE x1.00, copy-calibrated)",
                     "kernel": "spmm_csr_kernel + fixup_chunk_kernel (one fgx_spmm_csr call)",
                     "algorithmic_bytes": by_spmm, "launch_ms": round(spmm["event_ms"], 5),
                     "warm_launch_ms": round(spmm_warm["event_ms"], 5),
                     "gather_floor": gather_floor,
                     "launch_over_gather_floor": (round(spmm["event_ms"] / gather_floor["ms"], 3)
                                                  if gather_floor else None)},
        "normwise_error_vs_fp64": errors,
        "spmv": {"gbs": round(gbs(by_spmv, spmv["event_ms"]), 2),
                 "event_ms": round(spmv["event_ms"], 5),
                 "warm_event_ms": round(spmv_warm["event_ms"], 5),
                 "warm_gbs": round(gbs(by_spmv, spmv_warm["event_ms"]), 2),
                 "algorithmic_bytes": by_spmv,
                 "pct_of_hbm_peak": round(100 * gbs(by_spmv, spmv["event_ms"]) / HBM_PEAK_GBS, 2)},
        "sddmm": {"gbs": round(gbs(by_sddmm, sddmm["event_ms"]), 2),
                  "event_ms": round(sddmm["event_ms"], 5),
                  "warm_event_ms": round(sddmm_warm["event_ms"], 5),
                  "algorithmic_bytes": by_sddmm,
                  "gflops": round(2 * nnz * K / (sddmm["event_ms"] * 1e-3) / 1e9, 1),
                  "pct_of_hbm_peak": round(100 * gbs(by_sddmm, sddmm["event_ms"]) / HBM_PEAK_GBS,
                                           2)},
        "spmm_warm": {"event_ms": round(spmm_warm["event_ms"], 5),
                      "gbs": round(gbs(by_spmm, spmm_warm["event_ms"]), 2)},
        "variants": variants,
        "vendor": vendor,
        "cpu_baseline": cpu,
        "sparse_sparse": sparse,
        "batched": batched,
        "serving": served,
    }
    print(json.dumps(line), flush=True)


def timed(fn, steps, warmup, flush=None):
    """Per-call HIP event times of fn() on the current stream (ms).  With
    `flush`, it runs before each call, outside the events."""
    import numpy as np
    import torch

    for _ in range(warmup):
        if flush is not None:
            flush()
        fn()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t = time.perf_counter()
    for s, e in ev:
        if flush is not None:
            flush()
        s.record(stream)
        fn()
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    per = [s.elapsed_time(e) for s, e in ev]
    return {"event_ms": float(np.mean(per)), "event_median_ms": float(np.median(per)),
            "wall_ms": wall / steps * 1e3}


def bench_gather_floor(A, B, flush, steps):
    """The gather floor of SpMM's B-row gathers on this device.

    The probe library gathers exactly the rows an SpMM call gathers (row
    idx[q] of B for every stored position q, 256 bytes each) and nothing
    else, flushed between calls like the SpMM timing.  The best time over a
    few launch shapes is the floor the gathers alone impose; the SpMM call
    is reported against it.  None if the probe library is not built.
    """
    import ctypes

    import torch

    path = os.path.join(HERE, "finch.jl_amd", "finchgpu", "libfgx_probe.so")
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    lib.fgx_probe_gather_out_len.restype = ctypes.c_int64
    lib.fgx_probe_gather_out_len.argtypes = [ctypes.c_int]
    lib.fgx_probe_gather_rows.restype = ctypes.c_int
    lib.fgx_probe_gather_rows.argtypes = [
        ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
        ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
        ctypes.c_int, ctypes.c_void_p]
    assert B.dtype == torch.float32 and B.shape[1] == 64 and B.stride(1) == 1
    assert B.shape[0] == A.n_inner and A.n_inner < 2 ** 31
    nnz = A.nnz

    def probe(idx, Bv, u, bpc):
        n = lib.fgx_probe_gather_out_len(bpc)
        out = torch.empty(n, dtype=torch.float32, device=B.device)

        def run():
            rc = lib.fgx_probe_gather_rows(
                Bv.data_ptr(), Bv.stride(0) // 4, 16, idx.data_ptr(), idx.element_size(),
                A.base, nnz, out.data_ptr(), n, u, bpc, torch.cuda.current_stream().cuda_stream)
            if rc != 0:
                raise RuntimeError(f"gather probe failed ({rc})")
        return timed(run, steps, 2, flush)["event_ms"]

    shapes = {f"u{u}_b{bpc}": probe(A.idx, B, u, bpc) for u in (8, 16, 32) for bpc in (8, 16)}
    best = min(shapes.values())
    # The same gathers folded onto the first `rows` rows of B (idx mod rows):
    # the rate once the gathered rows fit the Infinity Cache (262,144 rows,
    # 64 MB) or each XCD's 4 MB L2 (16,384 rows).
    folded = {}
    for rows in (262_144, 16_384):
        fidx = torch.remainder(A.idx - A.base, rows) + A.base
        t = min(probe(fidx, B[:rows], u, 16) for u in (8, 16, 32))
        folded[str(rows)] = {"ms": round(t, 5), "gathered_tbs": round(nnz * 256 / (t * 1e-3) / 1e12, 3)}
        del fidx
    return {"ms": round(best, 5), "rows_gathered": nnz, "bytes_gathered": nnz * 256,
            "gathered_tbs": round(nnz * 256 / (best * 1e-3) / 1e12, 3),
            "by_shape_ms": {k: round(v, 5) for k, v in shapes.items()},
            "rows_of_B_folded_to": folded}


def bench_atomic_probe(dev, n=16_000_000, m=1_000_000):
    """The two accumulation forms a CSC SpMV can use, apart from any SpMV:
    n random float adds into a 1M-entry y through device atomics (device
    scope; workgroup or device scope into one copy of y per XCD), and n
    (row, value) entries added into an 8,192-entry LDS array (float atomics,
    integer atomics, plain writes).  Every float add is 1.0, so y sums to n."""
    import ctypes

    import numpy as np
    import torch

    lib = ctypes.CDLL(os.path.join(HERE, "finch.jl_amd", "finchgpu", "libfgx_probe.so"))
    lib.fgx_probe_atomic_scatter.restype = ctypes.c_int
    lib.fgx_probe_atomic_scatter.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                             ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_void_p]
    lib.fgx_probe_lds_accumulate.restype = ctypes.c_int
    lib.fgx_probe_lds_accumulate.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                             ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                             ctypes.c_void_p]
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    rows = torch.randint(0, m, (n,), dtype=torch.int32, device=dev, generator=g)
    vals = torch.ones(n, dtype=torch.float32, device=dev)
    y = torch.zeros(8 * m, dtype=torch.float32, device=dev)
    flushbuf = torch.zeros(128 << 20, dtype=torch.float32, device=dev)
    stream = lambda: torch.cuda.current_stream().cuda_stream
    out = {}
    for mode, name in ((0, "device_scope"), (1, "per_xcd_copy_workgroup_scope"),
                       (2, "per_xcd_copy_device_scope")):
        def run():
            if lib.fgx_probe_atomic_scatter(rows.data_ptr(), vals.data_ptr(), n, y.data_ptr(), m,
                                            mode, 16, stream()) != 0:
                raise RuntimeError("atomic probe failed")
        y.zero_()
        run()
        total = float(y.double().sum())
        t = timed(run, 10, 1, lambda: flushbuf.add_(1.0))["event_ms"]
        out[f"global_atomic_{name}"] = {"ms": round(t, 4), "g_adds_per_s": round(n / t / 1e6, 1),
                                        "sum_ok": total == float(n)}
    entries = torch.stack([rows & 8191, vals.view(torch.int32)], dim=1).contiguous()
    blocks = 512
    res = torch.empty(blocks * 1024, dtype=torch.float32, device=dev)
    for mode, name in ((0, "float_atomic"), (1, "int_atomic"), (2, "plain_write")):
        def run():
            if lib.fgx_probe_lds_accumulate(entries.data_ptr(), n, mode, res.data_ptr(),
                                            res.numel(), blocks, stream()) != 0:
                raise RuntimeError("LDS probe failed")
        t = timed(run, 10, 1)["event_ms"]
        out[f"lds_{name}"] = {"ms": round(t, 4), "g_ops_per_s": round(n / t / 1e6, 1)}
    del flushbuf
    return {"entries": n, "rows": m, **out}


def _cpu_median(fn, reps=10):
    """Median wall time (s) of `reps` warm calls of a host function."""
    import numpy as np
    fn()
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        t.append(time.perf_counter() - t0)
    return float(np.median(t))


def _normwise(got, ref):
    import numpy as np
    g = got.float().cpu().numpy().astype(np.float64).reshape(ref.shape) if hasattr(got, "cpu") \
        else np.asarray(got, np.float64)
    den = float(np.linalg.norm(ref))
    return float(f"{float(np.linalg.norm(g - ref)) / (den if den > 0 else 1.0):.3e}")


def bench_variants(dev, A, ptr, idx, val, B, x, flush, steps, oracle, nth):
    """The same 1M x 1M SpMV and SpMM with Finch's default Int64 ptr/idx, and
    with values stored as BF16 / FP16 (FP32 accumulation), each with its
    algorithmic bytes at its own element sizes and its error against FP64 on
    the original FP32 values; plus the CSC SpMV forms (atomic scatter and the
    deterministic transpose-then-rows form) on the same buffers read
    column-wise."""
    import numpy as np
    import torch

    from finchgpu import ops

    gbs = lambda b, ms: round(b / (ms * 1e-3) / 1e9, 1)
    m, n, nnz = A.n_outer, A.n_inner, A.nnz
    v64 = val.astype(np.float64)
    ref_v = oracle.spmv_csr(ptr, idx, v64, x.cpu().numpy().astype(np.float64), nthreads=nth)
    ref_m = oracle.spmm_csr(ptr, idx, v64, B.cpu().numpy().astype(np.float64), nthreads=nth)
    out = {}
    y = torch.empty(m, dtype=torch.float32, device=dev)
    C = torch.empty((m, B.shape[1]), dtype=torch.float32, device=dev)
    A64 = ops.Compressed(A.ptr.to(torch.int64), A.idx.to(torch.int64), A.val, m, n, 1)
    forms = [("int64_idx", A64, 4, 8, 8), ("bf16_values", A.to_values(torch.bfloat16), 2, 4, 4),
             ("fp16_values", A.to_values(torch.float16), 2, 4, 4)]
    for name, Av, sv, si, sp in forms:
        tv = timed(lambda: ops.spmv_csr(Av, x, y), steps, 2, flush)
        tm = timed(lambda: ops.spmm_csr(Av, B, C), max(5, steps // 4), 2, flush)
        bv = algorithmic_bytes(m, n, nnz, B.shape[1], "spmv", sv=sv, si=si, sp=sp)
        bm = algorithmic_bytes(m, n, nnz, B.shape[1], "spmm", sv=sv, si=si, sp=sp)
        out[name] = {"spmv_ms": round(tv["event_ms"], 5), "spmv_gbs": gbs(bv, tv["event_ms"]),
                     "spmm_ms": round(tm["event_ms"], 5), "spmm_gbs": gbs(bm, tm["event_ms"]),
                     "spmv_err_vs_fp64": _normwise(y, ref_v),
                     "spmm_err_vs_fp64": _normwise(C, ref_m),
                     "spmv_algorithmic_bytes": bv, "spmm_algorithmic_bytes": bm}
    del A64, C
    # CSC: the same buffers as 1M column fibres; y = A^T x.
    yc = torch.empty(n, dtype=torch.float32, device=dev)
    # fgx_spmv_csc takes its row-bucket path at this size; FGX_CSC_BUCKETS=0
    # times the atomic scatter kernel on the same call.
    tb = timed(lambda: ops.spmv_csc(A, x, yc), steps, 2, flush)
    yb = yc.clone()
    os.environ["FGX_CSC_BUCKETS"] = "0"
    try:
        ta = timed(lambda: ops.spmv_csc(A, x, yc), steps, 2, flush)
    finally:
        del os.environ["FGX_CSC_BUCKETS"]
    ya = yc.clone()
    to = timed(lambda: ops.spmv_csc_ordered(A, x, yc), max(5, steps // 4), 2, flush)
    yo1 = yc.clone()
    ops.spmv_csc_ordered(A, x, yc)
    refc = oracle.spmv_csc_atomic(n, ptr, idx, v64, x.cpu().numpy().astype(np.float64),
                                  nthreads=nth)
    bc = algorithmic_bytes(n, m, nnz, 1, "spmv")
    out["csc_spmv"] = {"buckets_ms": round(tb["event_ms"], 5),
                       "buckets_gbs": gbs(bc, tb["event_ms"]),
                       "buckets_err_vs_fp64": _normwise(yb, refc),
                       "atomic_ms": round(ta["event_ms"], 5), "atomic_gbs": gbs(bc, ta["event_ms"]),
                       "ordered_ms": round(to["event_ms"], 5),
                       "ordered_gbs": gbs(bc, to["event_ms"]),
                       "atomic_err_vs_fp64": _normwise(ya, refc),
                       "ordered_err_vs_fp64": _normwise(yo1, refc),
                       "ordered_repeat_bit_identical": bool(torch.equal(yo1, yc))}
    return out


def bench_vendor(dev, A, B, x, flush):
    """Vendor kernels on the same operands, for comparison only: rocSPARSE
    SpMV/SpMM through torch's sparse CSR tensors, and torch's embedding_bag
    (the same row gathers with per-entry weights, summed per row).  Each is
    timed like the library calls (flushed, HIP events)."""
    import torch

    gbs = lambda b, ms: round(b / (ms * 1e-3) / 1e9, 1)
    m, n, nnz = A.n_outer, A.n_inner, A.nnz
    out = {}
    crow = (A.ptr - 1).contiguous()
    col = (A.idx - 1).contiguous()
    try:
        S = torch.sparse_csr_tensor(crow, col, A.val, (m, n))
        t = timed(lambda: torch.mv(S, x), 10, 2, flush)
        out["rocsparse_spmv"] = {"ms": round(t["event_ms"], 5),
                                 "gbs": gbs(algorithmic_bytes(m, n, nnz, 1, "spmv"), t["event_ms"])}
        t = timed(lambda: torch.mm(S, B), 5, 1, flush)
        out["rocsparse_spmm"] = {"ms": round(t["event_ms"], 5),
                                 "gbs": gbs(algorithmic_bytes(m, n, nnz, B.shape[1], "spmm"),
                                            t["event_ms"])}
        del S
    except Exception as e:  # reported, not fatal: a comparison point only
        out["rocsparse_error"] = repr(e)[:200]
    try:
        offs = crow[:-1].to(torch.int64)
        colw = col.to(torch.int64)
        f = lambda: torch.nn.functional.embedding_bag(colw, B, offs, mode="sum",
                                                      per_sample_weights=A.val)
        t = timed(f, 5, 1, flush)
        out["embedding_bag_spmm"] = {"ms": round(t["event_ms"], 5),
                                     "gbs": gbs(algorithmic_bytes(m, n, nnz, B.shape[1], "spmm"),
                                                t["event_ms"])}
    except Exception as e:
        out["embedding_bag_error"] = repr(e)[:200]
    torch.cuda.empty_cache()
    return out


def bench_cpu(oracle, ptr, idx, val, Bh, nnz, by_spmm, by_spmv, seconds):
    """The oracle's C restatement of Finch's threaded loops on this host.

    SpMM (k=64) runs on rows 0..m/8 of A against the whole B and is scaled to
    the full matrix by stored entries (time x nnz / nnz_sample); SpMV runs on
    the whole matrix.  Each figure is the median of at least 10 warm
    repetitions, at 1 thread and at every available thread, with Finch's
    static fld blocks (the `parallel` default) and with its greedy schedule
    at its default chunk of 1 row; CSC SpMV with AtomicElement adds and with a
    Mutex per y element (the Mutex leg, 16M lock round trips, is the median
    of 3)."""
    import numpy as np

    nproc = os.cpu_count()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or oracle.max_threads()
    rows = M // 8
    q1 = int(ptr[rows] - 1)
    sp, si, sv = ptr[: rows + 1], idx[:q1], val[:q1]
    xh = Bh[:, 0].copy()
    scale = nnz / q1

    def med(fn, reps=10):
        fn()
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            t.append(time.perf_counter() - t0)
        return float(np.median(t))

    gbs = lambda b, s: round(b / s / 1e9, 3)
    res = {}
    budget = time.perf_counter() + seconds
    for name, nt in (("all_threads", threads), ("one_thread", 1)):
        mm = med(lambda: oracle.spmm_csr(sp, si, sv, Bh, nthreads=nt)) * scale
        mv = med(lambda: oracle.spmv_csr(ptr, idx, val, xh, nthreads=nt))
        res[name] = {"threads": nt, "spmm_full_equiv_ms": round(mm * 1e3, 2),
                     "spmm_gbs": gbs(by_spmm, mm), "spmv_ms": round(mv * 1e3, 2),
                     "spmv_gbs": gbs(by_spmv, mv)}
        if time.perf_counter() > budget:
            break
    g = med(lambda: oracle.spmv_csr_greedy(ptr, idx, val, xh, 1, nthreads=threads))
    res["all_threads"]["spmv_greedy_chunk1_ms"] = round(g * 1e3, 2)
    c = med(lambda: oracle.spmv_csc_atomic(M, ptr, idx, val, xh, nthreads=threads))
    res["all_threads"]["csc_atomic_spmv_ms"] = round(c * 1e3, 2)
    c = med(lambda: oracle.spmv_csc_mutex(M, ptr, idx, val, xh, nthreads=threads), reps=3)
    res["all_threads"]["csc_mutex_spmv_ms"] = round(c * 1e3, 2)
    allt = res["all_threads"]
    return {"value": allt["spmm_gbs"], "unit": "GB/s", "cores": threads, "kind": "port",
            "nproc": nproc,
            "sample": f"oracle CSR SpMM (FP32, Finch's static row blocks) on rows 0..{rows} of A "
                      f"({q1} of {nnz} stored entries) x the full 1M x 64 B, median of 10 reps, "
                      f"scaled to the full matrix by stored entries; SpMV on the full matrix",
            "detail": res}


def _events_ms(fn, reps, warmup=1):
    """Mean event time of fn() on the current stream (ms)."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def _csr_bytes(A, vbytes=4, ibytes=4, pbytes=4):
    return (A.n_outer + 1) * pbytes + A.nnz * (ibytes + vbytes)


def _wall_ms(fn, reps):
    """Mean wall time of fn() with a device synchronisation after each call
    (the two-phase calls read nnz back to the host between their phases)."""
    import torch

    fn()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        fn()
        torch.cuda.synchronize()
    return (time.perf_counter() - t) / reps * 1e3


def bench_sparse_sparse(dev, with_cpu=True):
    """Coiteration and Gustavson SpGEMM: A, B 200,000 x 200,000 CSR with
    Poisson(8) / Poisson(24) nnz per row and uniform columns.  A.*B, A+B and
    max.(A,B) on FP32 values; C = A*B with Int32 values U{1..9} (plus-times)
    and with FP32 values (min-plus).  Each call is the whole two-phase
    operation (symbolic, nnz read back, allocation, numeric).  The CPU leg is
    the oracle's serial loops (Finch runs these serially: SparseList outputs
    are not concurrent)."""
    import numpy as np
    import torch

    from finchgpu import ops, synth

    n = 200_000
    ap, ai, av = synth.poisson_csr(n, n, 8, seed=21)
    bp, bi, bv = synth.poisson_csr(n, n, 24, seed=22)
    up = lambda p_, i_, v_: ops.Compressed(torch.from_numpy(p_.astype(np.int32)).to(dev),
                                           torch.from_numpy(i_.astype(np.int32)).to(dev),
                                           torch.from_numpy(v_).to(dev), n, n, 1)
    A, B = up(ap, ai, av), up(bp, bi, bv)
    gbs = lambda b, ms: round(b / (ms * 1e-3) / 1e9, 1)
    out = {"workload": f"200k x 200k CSR, A Poisson(8) nnz={A.nnz}, B Poisson(24) nnz={B.nnz}, "
                       "Int32 ptr/idx"}
    if with_cpu:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle
    for op in ("mul", "add", "max"):
        C = ops.ewise(op, A, B)
        ms = _wall_ms(lambda: ops.ewise(op, A, B), 10)
        by = _csr_bytes(A) + _csr_bytes(B) + _csr_bytes(C)
        row = {"ms": round(ms, 3), "nnz_out": C.nnz, "algorithmic_bytes": by, "gbs": gbs(by, ms)}
        # The one-call form (fgx_ewise_async: no host synchronisation, outputs
        # sized to the bound), timed by HIP events over back-to-back calls.
        bufs = ops.ewise_async(op, A, B)
        ams = _events_ms(lambda: ops.ewise_async(op, A, B, out=bufs), 20, warmup=2)
        row["one_call_ms"] = round(ams, 4)
        row["one_call_gbs"] = gbs(by, ams)
        if with_cpu:
            row["cpu_ms_1thread"] = round(
                _cpu_median(lambda: oracle.ewise(op, ap, ai, av, bp, bi, bv)) * 1e3, 2)
        out[f"ewise_{op}"] = row
    rng = np.random.default_rng(23)
    Ai = up(ap, ai, rng.integers(1, 10, size=ai.size).astype(np.int32))
    Bi = up(bp, bi, rng.integers(1, 10, size=bi.size).astype(np.int32))
    blen = np.diff(bp)
    products = int(blen[ai - 1].sum())
    for name, sr, X, Y in (("spgemm_plus_times_i32", "plus_times", Ai, Bi),
                           ("spgemm_min_plus_f32", "min_plus", A, B)):
        C = ops.spgemm(sr, X, Y)
        ms = _wall_ms(lambda: ops.spgemm(sr, X, Y), 5)
        by = _csr_bytes(X) + _csr_bytes(Y) + _csr_bytes(C)
        row = {"ms": round(ms, 3), "nnz_out": C.nnz, "products": products,
               "gproducts_per_s": round(products / (ms * 1e-3) / 1e9, 3),
               "algorithmic_bytes": by, "gbs": gbs(by, ms)}
        if with_cpu:
            rows = n // 10   # bounded sample: the first tenth of the rows
            xv = X.val.cpu().numpy()
            yv = Y.val.cpu().numpy()
            q1 = int(ap[rows] - 1)
            row["cpu_ms_1thread_first_tenth_rows"] = round(_cpu_median(
                lambda: oracle.spgemm(sr, rows, n, ap[: rows + 1], ai[:q1], xv[:q1], bp, bi, yv))
                * 1e3, 2)
        out[name] = row
    return out


def bench_format(dev, A, ptr, idx, val, with_cpu=True):
    """Format conversion on the 1M x 1M power-law operand (16M stored
    entries): the device transpose (CSR -> CSC, stable radix sort), COO
    assembly of the same entries listed in a shuffled order, and dropfills.
    Algorithmic bytes: the input read once plus the output written once."""
    import numpy as np
    import torch

    from finchgpu import ops

    gbs = lambda b, ms: round(b / (ms * 1e-3) / 1e9, 1)
    nnz = A.nnz
    out = {"workload": f"1M x 1M CSR, nnz={nnz}, Zipf(1.5) rows, Int32 ptr/idx, FP32 values"}
    ms = _events_ms(lambda: ops.transpose(A), 10)
    by = 2 * _csr_bytes(A)
    out["transpose"] = {"ms": round(ms, 3), "algorithmic_bytes": by, "gbs": gbs(by, ms)}
    rng = np.random.default_rng(31)
    perm = torch.from_numpy(rng.permutation(nnz)).to(dev)
    rows = torch.repeat_interleave(torch.arange(A.n_outer, device=dev, dtype=torch.int32),
                                   (A.ptr[1:] - A.ptr[:-1]).long())
    ro = (rows + 1)[perm].contiguous()
    co = A.idx[perm].contiguous()
    vo = A.val[perm].contiguous()
    del rows
    ms = _wall_ms(lambda: ops.coo_to_compressed(ro, co, vo, A.n_outer, A.n_inner, 1), 5)
    by = nnz * 12 + _csr_bytes(A)
    out["coo_to_compressed"] = {"ms": round(ms, 3), "algorithmic_bytes": by, "gbs": gbs(by, ms),
                                "input": "the same entries as (row, col, val) lists in a "
                                         "random order"}
    ms = _wall_ms(lambda: ops.dropfills(A, 0.0), 10)
    by = 2 * _csr_bytes(A)
    out["dropfills"] = {"ms": round(ms, 3), "algorithmic_bytes": by, "gbs": gbs(by, ms)}
    del perm, ro, co, vo
    if with_cpu:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle
        out["transpose"]["cpu_ms_1thread"] = round(
            _cpu_median(lambda: oracle.transpose(A.n_outer, A.n_inner, ptr, idx, val, 1)) * 1e3, 2)
    return out


def bench_docs_sddmm(dev):
    """The one timing the reference publishes on this path
    (docs/src/getting_started.md:126-134): compute(lazy(A) .* (lazy(B) *
    lazy(C))) with A = fsprand(1000, 1000, 100) (100 stored entries, Finch's
    default column-compressed format, U[0,1) values) and B, C dense 1000x1000
    Float64 U[0,1), quoted at 694.666 us on unstated CPU hardware.  Here the
    same program runs through the Python router (interface.compute, the fused
    SDDMM route, FP64) end to end with a device synchronisation per call,
    and the device kernel alone is timed with events; the result is checked
    against the dense FP64 product on the host."""
    import numpy as np
    import scipy.sparse as sp
    import torch

    from finchgpu import interface as fi
    from finchgpu import ops
    from finchgpu.tensors import Dense, Element, GPU, SparseList, Tensor, transfer

    rng = np.random.default_rng(5)
    m = n = k = 1000
    flat = np.sort(rng.choice(m * n, 100, replace=False))
    S = sp.csc_matrix((rng.random(100), (flat % m, flat // m)), shape=(m, n))
    Bn, Cn = rng.random((m, k)), rng.random((k, n))
    A = transfer(GPU(0), Tensor(Dense(SparseList(Element(0.0))), S))
    B = transfer(GPU(0), Tensor(Dense(Dense(Element(0.0))), Bn))
    C = transfer(GPU(0), Tensor(Dense(Dense(Element(0.0))), Cn))
    run = lambda: fi.compute(fi.lazy(A) * (fi.lazy(B) @ fi.lazy(C)))
    R = run()
    torch.cuda.synchronize()
    ref = S.toarray() * (Bn @ Cn)
    got = R.todense()
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    times = []
    for _ in range(50):
        t = time.perf_counter()
        run()
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t) * 1e6)
    # The kernel alone: the same product on the row-ordered copy of A.
    St = S.tocsr()
    Ar = ops.Compressed(torch.from_numpy(St.indptr.astype(np.int32) + 1).to(dev),
                        torch.from_numpy(St.indices.astype(np.int32) + 1).to(dev),
                        torch.from_numpy(St.data).to(dev), m, n, 1)
    X = torch.from_numpy(Bn).to(dev)
    Y = torch.from_numpy(np.ascontiguousarray(Cn.T)).to(dev)
    out = torch.empty(100, dtype=torch.float64, device=dev)
    kern_us = _events_ms(lambda: ops.sddmm_csr(Ar, X, Y, out), 200, warmup=5) * 1e3
    return {"program": "compute(lazy(A) .* (lazy(B) * lazy(C))), A = fsprand(1000, 1000, 100), "
                       "B, C dense 1000x1000 Float64",
            "published_us": 694.666, "published_hardware": "unstated (Finch CPU)",
            "router_us_median": round(float(np.median(times)), 1),
            "kernel_us": round(kern_us, 2), "rel_err_vs_fp64": err}


def bench_fp64_10k(dev):
    """The parity configuration: 10,000 x 10,000 FP64 at density 1e-3 (fsprand
    semantics), CSC SpMV and the swizzled CSR ordering; single calls are
    latency-bound (1.6 MB), so this reports latency."""
    import numpy as np
    import torch

    from finchgpu import ops, synth

    ptr, idx, val = synth.fsprand(10_000, 10_000, 1e-3, seed=0)
    A = ops.Compressed(torch.from_numpy(ptr).to(dev), torch.from_numpy(idx).to(dev),
                       torch.from_numpy(val).to(dev), 10_000, 10_000, 1)
    x = torch.from_numpy(synth.dense(10_000, 1, seed=1, dtype=np.float64, kind="sym")[:, 0]).to(dev)
    y = torch.empty(10_000, dtype=torch.float64, device=dev)
    csc = _events_ms(lambda: ops.spmv_csc(A, x, y), 50)
    csco = _events_ms(lambda: ops.spmv_csc_ordered(A, x, y), 50)
    csr = _events_ms(lambda: ops.spmv_csr(A, x, y), 50)
    return {"nnz": A.nnz, "csc_spmv_us": round(csc * 1e3, 2),
            "csc_ordered_spmv_us": round(csco * 1e3, 2), "csr_spmv_us": round(csr * 1e3, 2)}


def bench_batched(dev, count, with_loop=True):
    """Batched form: `count` independent square CSR matrices (rows log-uniform in
    [256, 65536], Poisson nnz/row with mean log-uniform in [4, 64], n_rhs in
    {1, 8, 32, 128}): one autotuned persistent hipGraph launch vs a per-matrix
    loop, and BF16-stored values (FP32 accumulate) vs FP32."""
    import torch

    from finchgpu import _lib as L
    from finchgpu import batch

    t0 = time.time()
    items = batch.synth_batch(count, seed=0, device=dev)
    nnz = sum(it.A.nnz for it in items)
    alg = sum(algorithmic_bytes(it.A.n_outer, it.A.n_inner, it.A.nnz, it.n_rhs, "spmm")
              for it in items)
    print(f"# batch generated: {count} matrices, nnz={nnz} in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan = batch.BatchPlan(items)
    torch.cuda.synchronize()
    create_ms = (time.perf_counter() - t0) * 1e3
    plan_ms = _events_ms(plan.run, 5)
    ref = [it.C.clone() for it in items]
    loop_ms = _events_ms(lambda: batch.run_loop(items), 3) if with_loop else None
    stats = plan.stats()
    plan.close()
    # The measured autotuner: row-split vs nnz-split timed per bucket of
    # similar items at plan creation.
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mplan = batch.BatchPlan(items, L.BATCH_MEASURE)
    torch.cuda.synchronize()
    mcreate_ms = (time.perf_counter() - t0) * 1e3
    mplan_ms = _events_ms(mplan.run, 5)
    mstats = mplan.stats()
    mplan.close()
    den = sum(float(torch.linalg.vector_norm(r.double()) ** 2) for r in ref)

    def err(its):
        num = sum(float(torch.linalg.vector_norm((b.C - r).double()) ** 2)
                  for b, r in zip(its, ref))
        return float(f"{(num / max(den, 1e-300)) ** 0.5:.3e}")

    # BF16 values (FP32 right-hand sides), then BF16 values and BF16 right-hand
    # sides; both accumulate in FP32 and write FP32 C.
    items16 = [batch.BatchItem(it.A.to_values(torch.bfloat16), it.B) for it in items]
    alg16 = alg - nnz * 2
    plan16 = batch.BatchPlan(items16)
    plan16_ms = _events_ms(plan16.run, 5)
    plan16.close()
    err16 = err(items16)
    items16i = [batch.BatchItem(a.A, a.B.to(torch.bfloat16), a.C) for a in items16]
    alg16i = alg16 - sum(it.A.n_inner * it.n_rhs for it in items) * 2
    plan16i = batch.BatchPlan(items16i)
    plan16i_ms = _events_ms(plan16i.run, 5)
    plan16i.close()
    err16i = err(items16i)
    # Per class (one kernel per right-hand-side width): items, nnz and the
    # algorithmic bytes of each variant, for per-class rooflines from a kernel trace.
    classes = {}
    for it in items:
        c = classes.setdefault(str(it.n_rhs), {"items": 0, "nnz": 0, "alg_fp32": 0,
                                               "alg_bf16_values": 0, "alg_bf16_values_inputs": 0})
        a32 = algorithmic_bytes(it.A.n_outer, it.A.n_inner, it.A.nnz, it.n_rhs, "spmm")
        c["items"] += 1
        c["nnz"] += it.A.nnz
        c["alg_fp32"] += a32
        c["alg_bf16_values"] += a32 - 2 * it.A.nnz
        c["alg_bf16_values_inputs"] += a32 - 2 * it.A.nnz - 2 * it.A.n_inner * it.n_rhs
    gbs = lambda b, ms: round(b / (ms * 1e-3) / 1e9, 1)
    out = {"workload": f"{count} CSR matrices, nnz={nnz}, FP32 values, Int32 idx",
           "plan_graph_ms": round(plan_ms, 3), "plan_graph_gbs": gbs(alg, plan_ms),
           "loop_ms": round(loop_ms, 3) if loop_ms else None,
           "loop_gbs": gbs(alg, loop_ms) if loop_ms else None,
           "speedup_vs_loop": round(loop_ms / plan_ms, 2) if loop_ms else None,
           "bf16_plan_ms": round(plan16_ms, 3), "bf16_plan_gbs": gbs(alg16, plan16_ms),
           "bf16_normwise_err_vs_fp32": err16,
           "bf16_inputs_plan_ms": round(plan16i_ms, 3),
           "bf16_inputs_plan_gbs": gbs(alg16i, plan16i_ms),
           "bf16_inputs_normwise_err_vs_fp32": err16i,
           "plan_create_ms": round(create_ms, 1),
           "measured_tuner": {"plan_graph_ms": round(mplan_ms, 3), "plan_create_ms":
                              round(mcreate_ms, 1), "items_rowsplit": mstats["items_rowsplit"]},
           "algorithmic_bytes": alg, "plan_stats": stats, "classes": classes}
    del items, items16, items16i, ref
    torch.cuda.empty_cache()
    return out


def _serving_operands(dev, large):
    """The serving mix's operands: 256 matrices of the batched size
    distribution plus four 1M-row Zipf(1.5) matrices (the timed matrix and
    three row permutations of it), with FP32 and BF16 values, and the FP32 and
    BF16 dense inputs of every (operand, kind, width)."""
    import torch

    from finchgpu import ops, serve

    def permuted(A, seed):
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        ptr = A.ptr.to(torch.int64) - A.base
        lens = ptr[1:] - ptr[:-1]
        perm = torch.randperm(A.n_outer, device=dev, generator=g)
        nl = lens[perm]
        nptr = torch.zeros(A.n_outer + 1, dtype=torch.int64, device=dev)
        nptr[1:] = torch.cumsum(nl, 0)
        row_of = torch.repeat_interleave(torch.arange(A.n_outer, device=dev), nl)
        src = ptr[perm][row_of] + (torch.arange(A.nnz, device=dev) - nptr[row_of])
        return ops.Compressed((nptr + A.base).to(A.ptr.dtype), A.idx[src].contiguous(),
                              A.val[src].contiguous(), A.n_outer, A.n_inner, A.base)

    small = serve.serving_workload(dev, n_small=256, n_large=0, seed=0)
    big = [large[0]] + [permuted(large[0], 100 + j) for j in range(3)]
    operands32 = small + big
    operands16 = [A.to_values(torch.bfloat16) for A in operands32]
    inputs = serve.Inputs(operands32, seed=1)
    inputs16 = serve.Inputs(operands32, seed=1, dtype=torch.bfloat16)
    return operands32, operands16, inputs, inputs16


def bench_serving_by_kind(dev, large):
    """Device time per request of the serving mix, by request kind and operand
    class: for each kind (SpMV, SpMM k=16, SpMM k=64, SDDMM k=32) and each
    operand class (the 256 small operands, the four 1M-row operands), a trace
    of requests that all arrive at once is served in batched mode; requests /
    span is the saturated rate.  The mix-weighted sum is the device time of an
    average request, and each term's share of it shows where the device goes."""
    import torch

    from finchgpu import serve

    operands32, operands16, inputs, inputs16 = _serving_operands(dev, large)
    n_small = 256
    p_large = (len(operands32) - n_small) / len(operands32)
    kinds = (("spmv", (1.0, 0.0, 0.0), (16,), 0.6), ("spmm16", (0.0, 1.0, 0.0), (16,), 0.15),
             ("spmm64", (0.0, 1.0, 0.0), (64,), 0.15), ("sddmm", (0.0, 0.0, 1.0), (16,), 0.1))

    def alg_bytes(r, vb, db):
        A = operands32[r.operand]
        if r.kind == "sddmm":
            io = (A.n_outer + A.n_inner) * r.k * db + A.nnz * 4
        elif r.kind == "spmv":
            io = A.n_inner * db + A.n_outer * 4
        else:
            io = A.n_inner * r.k * db + A.n_outer * r.k * 4
        return A.nnz * (vb + 4) + 4 * (A.n_outer + 1) + io

    variants = (("batched_fp32", operands32, inputs, torch.float32, 4, 4),
                ("batched_bf16", operands16, inputs, torch.float32, 2, 4),
                ("batched_bf16_values_and_inputs", operands16, inputs16, torch.bfloat16, 2, 2))
    out = {}
    for name, ops_, src, dt, vb, db in variants:
        serve.serve(ops_, serve.make_trace(operands32, src, qps=20_000, n=2_000, seed=1), "batched",
                    input_dtype=dt)  # warm-up (tile tables, plans)
        rows, mean_us = {}, 0.0
        for cls, n, p_cls in (("small", 8_000, 1.0 - p_large), ("large", 40_000, p_large)):
            for kname, mix, ks, p_kind in kinds:
                reqs = [r for r in serve.make_trace(operands32, src, qps=5e6, n=n, seed=3, mix=mix,
                                                    spmm_k=ks)
                        if (r.operand < n_small) == (cls == "small")]
                res = serve.serve(ops_, reqs, "batched", input_dtype=dt)
                span = float(res.done_s.max() - res.arrival_s.min())
                us = span / len(reqs) * 1e6
                mean_us += p_cls * p_kind * us
                rows[f"{cls}_{kname}"] = {
                    "requests": len(reqs), "us_per_request": round(us, 3),
                    "alg_gbs": round(sum(alg_bytes(r, vb, db) for r in reqs) / span / 1e9, 1),
                    "weight": round(p_cls * p_kind, 5)}
        for r in rows.values():
            r["share_of_device_time"] = round(r["weight"] * r["us_per_request"] / mean_us, 4)
        out[name] = {"by_class_and_kind": rows, "mix_us_per_request": round(mean_us, 3),
                     "saturated_req_per_s_estimate": round(1e6 / mean_us)}
        print(f"# serving by kind {name}: {mean_us:.3f} us per request", file=sys.stderr,
              flush=True)
    out["method"] = ("per (operand class, kind): all requests arrive at t=0 and are served in "
                     "batched mode; us_per_request = span / requests; weights: 60 % SpMV, "
                     "15 % SpMM k=16, 15 % SpMM k=64, 10 % SDDMM k=32, operands uniform over "
                     "256 small + 4 large")
    return out


def bench_serving(dev, slo_ms, large):
    """Serving form: open-loop Poisson arrivals, 60 % SpMV / 30 % SpMM (k 16 or
    64) / 10 % SDDMM (k 32) over 256 operands of the batched size distribution
    plus four 1M-row Zipf(1.5) matrices (the timed matrix and three row
    permutations of it), HBM admission budget 200 GB.  For each path, the
    offered rate is raised until p99 latency exceeds the target; the largest
    achieved req/s within it is reported, and for the batched paths also the
    largest within the plain path's own p99 (the comparison at equal p99)."""
    import numpy as np
    import torch

    from finchgpu import serve

    operands32, operands16, inputs, inputs16 = _serving_operands(dev, large)
    in_of = {"fp32": (inputs, torch.float32), "bf16": (inputs16, torch.bfloat16)}
    qps_grid = [1_000, 2_000, 5_000, 10_000, 20_000, 30_000, 50_000, 70_000, 100_000, 140_000,
                200_000, 300_000]

    def run(ops_, mode, q, seed, inp, target):
        n = int(min(20_000, max(2_000, q * 0.25)))
        src, dt = in_of[inp]
        reqs = serve.make_trace(operands32, src, qps=q, n=n, seed=seed)
        s = serve.serve(ops_, reqs, mode, input_dtype=dt).summary()
        return {"offered_qps": q, "req_per_s": round(s["req_per_s"], 1),
                "p50_ms": round(s["p50_ms"], 3), "p99_ms": round(s["p99_ms"], 3),
                "launches": s["launches"], "ok": s["p99_ms"] <= target and s["rejected"] == 0}

    def passes(ops_, mode, q, points, inp, target):
        # A rate passes when two runs (different arrival draws) both keep p99
        # within the target; the slower of the two is recorded.
        a = run(ops_, mode, q, 7, inp, target)
        points.append(a)
        if not a["ok"]:
            return None
        b = run(ops_, mode, q, 8, inp, target)
        points.append(b)
        if not b["ok"]:
            return None
        return min((a, b), key=lambda r: r["req_per_s"]) | {"p99_ms": max(a["p99_ms"], b["p99_ms"])}

    def serving_error(ops_, mode, inp):
        """Worst and mean normwise error vs the FP64 oracle (on the FP32
        operands and inputs) over 60 requests of the mix, at least one per
        large operand; the checker runs on the host after serving."""
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle
        src, dt = in_of[inp]
        rng = np.random.default_rng(9)
        reqs = serve.make_trace(operands32, src, qps=20_000, n=56, seed=9, outputs=True)
        for j in range(4):
            kind = ["spmv", "spmm", "sddmm", "spmm"][j]
            k = {"spmv": 1, "spmm": 64, "sddmm": 32}[kind]
            ins = src.get(256 + j, kind, k)
            A = operands32[256 + j]
            shape = (A.n_outer,) if kind == "spmv" else ((A.nnz,) if kind == "sddmm"
                                                          else (A.n_outer, k))
            reqs.append(serve.Request(kind, 256 + j, k, float(0.0028 + 1e-4 * j), ins[0],
                                      ins[1] if len(ins) > 1 else None,
                                      torch.empty(shape, dtype=torch.float32, device=dev)))
        serve.serve(ops_, reqs, mode, input_dtype=dt)
        errs = []
        for r in reqs:
            A = operands32[r.operand]
            p, i = A.ptr.cpu().numpy(), A.idx.cpu().numpy()
            v = A.val.cpu().numpy().astype(np.float64)
            x = inputs.get(r.operand, r.kind, r.k)
            x0 = x[0].cpu().numpy().astype(np.float64)
            if r.kind == "spmv":
                ref = oracle.spmv_csr(p, i, v, x0, nthreads=16)
            elif r.kind == "spmm":
                ref = oracle.spmm_csr(p, i, v, x0, nthreads=16)
            else:
                ref = oracle.sddmm_csr(p, i, v, x0, x[1].cpu().numpy().astype(np.float64))
            errs.append(_normwise(r.out, ref))
        del rng
        return {"max": max(errs), "mean": float(f"{float(np.mean(errs)):.3e}"), "requests": len(errs)}

    def sweep(ops_, mode, inp, target, grid):
        run(ops_, mode, 20_000, 1, inp, target)     # warm-up (plans, tables, scratch)
        best, points, lo, hi = None, [], None, None
        for q in grid:
            r = passes(ops_, mode, q, points, inp, target)
            if r is None:
                hi = q
                break
            best, lo = r, q
        # Bisect between the last passing and the first failing rate.
        if lo is not None and hi is not None:
            for _ in range(4):
                mid = (lo + hi) // 2
                r = passes(ops_, mode, mid, points, inp, target)
                if r is None:
                    hi = mid
                else:
                    best, lo = r, mid
        return best, points

    variants = (("plain_fp32", operands32, "plain", "fp32"),
                ("batched_fp32", operands32, "batched", "fp32"),
                ("batched_bf16", operands16, "batched", "fp32"),
                ("batched_bf16_values_and_inputs", operands16, "batched", "bf16"))
    res = {}
    for name, ops_, mode, inp in variants:
        best, points = sweep(ops_, mode, inp, slo_ms, qps_grid)
        res[name] = {"best": best, "sweep": points,
                     "normwise_err_vs_fp64": serving_error(ops_, mode, inp)}
        print(f"# serving {name}: {best}", file=sys.stderr, flush=True)
    base = (res["plain_fp32"]["best"] or {}).get("req_per_s")
    # The comparison at equal p99: each batched path swept again with the
    # plain path's measured p99 (at its best rate) as the target.
    plain_p99 = (res["plain_fp32"]["best"] or {}).get("p99_ms")
    for name, ops_, mode, inp in variants[1:]:
        b = (res[name]["best"] or {}).get("req_per_s")
        res[name]["x_plain"] = round(b / base, 2) if (b and base) else None
        if plain_p99:
            eq, eq_points = sweep(ops_, mode, inp, plain_p99, [q for q in qps_grid if q >= 50_000])
            e = (eq or {}).get("req_per_s")
            res[name]["at_plain_p99"] = {"p99_target_ms": plain_p99, "best": eq,
                                         "x_plain": round(e / base, 2) if (e and base) else None,
                                         "sweep": eq_points}
            print(f"# serving {name} at the plain p99 ({plain_p99} ms): {eq}", file=sys.stderr,
                  flush=True)
    res["p99_target_ms"] = slo_ms
    res["operands"] = ("256 batched-distribution CSR + 4 x 1M-row Zipf(1.5); FP32 values, BF16 "
                       "values, and BF16 values with BF16 dense inputs (FP32 accumulation and "
                       "outputs throughout)")
    return res


if __name__ == "__main__":
    main()
