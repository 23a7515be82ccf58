// fgx_scan.hpp — device-wide exclusive scan of per-fibre counts into a
// compressed `ptr` array: ptr[0] = base, ptr[i+1] = base + sum(cnt[0..i]).
// This is the `freeze` step of the SparseList level (the prefix sum of
// src/tensors/levels/sparse_list_levels.jl:310-327), done in three launches:
// per-block sums, a one-block scan of those sums, and the final per-block
// scan that writes ptr.
#pragma once
#include "fgx_device.hpp"
#include "fgx_internal.hpp"

namespace fgx {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

// The kScanItems consecutive counts cnt[q0 .. q0 + kScanItems) (zero past n),
// with vector loads when the run is whole and aligned.
template <typename TC>
__device__ __forceinline__ void load_counts(const TC* __restrict__ cnt, int64_t q0, int64_t n,
                                            int64_t (&v)[kScanItems]) {
  static_assert(kScanItems == 8, "vector loads below take 8 counts");
  const TC* p = cnt + q0;
  constexpr uintptr_t kAlign = sizeof(TC) == 1 ? 7 : 15;  // 8-byte loads for byte counts
  const bool whole = q0 + kScanItems <= n && (reinterpret_cast<uintptr_t>(p) & kAlign) == 0;
  if constexpr (sizeof(TC) == 1) {
    if (whole) {
      const uint64_t u = *reinterpret_cast<const uint64_t*>(p);
#pragma unroll
      for (int i = 0; i < 8; i++) v[i] = static_cast<int64_t>(static_cast<TC>(u >> (8 * i)));
      return;
    }
  } else if constexpr (sizeof(TC) == 4) {
    typedef int v4i __attribute__((ext_vector_type(4)));
    if (whole) {
      const v4i x = reinterpret_cast<const v4i*>(p)[0], y = reinterpret_cast<const v4i*>(p)[1];
      v[0] = static_cast<TC>(x.x); v[1] = static_cast<TC>(x.y);
      v[2] = static_cast<TC>(x.z); v[3] = static_cast<TC>(x.w);
      v[4] = static_cast<TC>(y.x); v[5] = static_cast<TC>(y.y);
      v[6] = static_cast<TC>(y.z); v[7] = static_cast<TC>(y.w);
      return;
    }
  } else if constexpr (sizeof(TC) == 8) {
    typedef long long v2l __attribute__((ext_vector_type(2)));
    if (whole) {
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const v2l x = reinterpret_cast<const v2l*>(p)[c];
        v[2 * c] = static_cast<int64_t>(x.x);
        v[2 * c + 1] = static_cast<int64_t>(x.y);
      }
      return;
    }
  }
#pragma unroll
  for (int i = 0; i < kScanItems; i++) v[i] = (q0 + i < n) ? static_cast<int64_t>(cnt[q0 + i]) : 0;
}

// Stores the kScanItems values o[] at dst, as vectors when the run is whole
// and 16-byte aligned.
template <typename TP>
__device__ __forceinline__ void store_run(TP* __restrict__ dst, int valid,
                                          const TP (&o)[kScanItems]) {
  const bool whole = valid == kScanItems && (reinterpret_cast<uintptr_t>(dst) & 15) == 0;
  if constexpr (sizeof(TP) == 4) {
    typedef int v4i __attribute__((ext_vector_type(4)));
    if (whole) {
      reinterpret_cast<v4i*>(dst)[0] = v4i{static_cast<int>(o[0]), static_cast<int>(o[1]),
                                           static_cast<int>(o[2]), static_cast<int>(o[3])};
      reinterpret_cast<v4i*>(dst)[1] = v4i{static_cast<int>(o[4]), static_cast<int>(o[5]),
                                           static_cast<int>(o[6]), static_cast<int>(o[7])};
      return;
    }
  } else if constexpr (sizeof(TP) == 8) {
    typedef long long v2l __attribute__((ext_vector_type(2)));
    if (whole) {
#pragma unroll
      for (int c = 0; c < 4; c++)
        reinterpret_cast<v2l*>(dst)[c] =
            v2l{static_cast<long long>(o[2 * c]), static_cast<long long>(o[2 * c + 1])};
      return;
    }
  }
#pragma unroll
  for (int i = 0; i < kScanItems; i++)
    if (i < valid) dst[i] = o[i];
}

namespace {  // kernels local to each translation unit

__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t* lds, int64_t* total) {
  const int t = threadIdx.x;
  const int lane = t & 63, w = t >> 6;
  // Inclusive wave scan.
  int64_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int64_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) lds[w] = x;
  __syncthreads();
  if (t == 0) {
    int64_t s = 0;
    for (int i = 0; i < static_cast<int>(blockDim.x) / 64; i++) {
      int64_t c = lds[i];
      lds[i] = s;
      s += c;
    }
    lds[16] = s;
  }
  __syncthreads();
  const int64_t r = lds[w] + x - v;
  *total = lds[16];
  __syncthreads();
  return r;
}

template <typename TC>
__global__ __launch_bounds__(kScanThreads) void scan_block_sums(const TC* __restrict__ cnt,
                                                                int64_t n,
                                                                int64_t* __restrict__ sums) {
  __shared__ int64_t lds[17];
  const int64_t b0 = static_cast<int64_t>(blockIdx.x) * kScanTile;
  // Each thread sums kScanItems consecutive counts (the same runs as
  // scan_write_ptr; a block sum does not depend on the order).
  int64_t v[kScanItems];
  load_counts(cnt, b0 + static_cast<int64_t>(threadIdx.x) * kScanItems, n, v);
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) s += v[i];
  int64_t tot;
  block_exclusive_scan(s, lds, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void scan_sums(int64_t* __restrict__ sums, int64_t nb,
                                                          int64_t* __restrict__ grand) {
  __shared__ int64_t lds[17];
  int64_t carry = 0;
  for (int64_t c0 = 0; c0 < nb; c0 += kScanThreads) {
    const int64_t q = c0 + threadIdx.x;
    const int64_t v = q < nb ? sums[q] : 0;
    int64_t tot;
    const int64_t ex = block_exclusive_scan(v, lds, &tot);
    if (q < nb) sums[q] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *grand = carry;
}

}  // namespace

template <typename TP, typename TC>
__global__ __launch_bounds__(kScanThreads) void scan_write_ptr(const TC* __restrict__ cnt,
                                                               int64_t n,
                                                               const int64_t* __restrict__ sums,
                                                               int64_t base, TP* __restrict__ ptr) {
  __shared__ int64_t lds[17];
  const int64_t b0 = static_cast<int64_t>(blockIdx.x) * kScanTile;
  // Each thread owns kScanItems consecutive entries.
  const int64_t q0 = b0 + static_cast<int64_t>(threadIdx.x) * kScanItems;
  int64_t v[kScanItems];
  load_counts(cnt, q0, n, v);
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) s += v[i];
  int64_t tot;
  int64_t run = block_exclusive_scan(s, lds, &tot) + sums[blockIdx.x];
  // ptr[q] = base + (counts before q) for q in [q0, q0 + kScanItems): the
  // thread's slots start at q0, so whole aligned runs go out as vectors; the
  // thread holding the last count also writes ptr[n].
  TP o[kScanItems];
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    o[i] = static_cast<TP>(base + run);
    run += v[i];
  }
  if (q0 < n) store_run(ptr + q0, static_cast<int>(n - q0 < kScanItems ? n - q0 : kScanItems), o);
  if (q0 < n && q0 + kScanItems >= n) ptr[n] = static_cast<TP>(base + run);
  if (n == 0 && blockIdx.x == 0 && threadIdx.x == 0) ptr[0] = static_cast<TP>(base);
}

// Writes ptr (n+1 entries) from cnt (n entries) and returns the total in
// *total_host (synchronises `s`).  `work` needs scan_work_bytes(n) bytes.
inline size_t scan_work_bytes(int64_t n) {
  const int64_t nb = (n + kScanTile - 1) / kScanTile + 1;
  return static_cast<size_t>(nb + 1) * sizeof(int64_t);
}

// Counts may be any integer type (e.g. one-byte flags); sums are int64.
template <typename TP, typename TC = int64_t>
int scan_counts_to_ptr(const TC* cnt, int64_t n, int64_t base, TP* ptr, void* work,
                       int64_t* pinned, int64_t* total_host, hipStream_t s) {
  const int64_t nb = n > 0 ? (n + kScanTile - 1) / kScanTile : 1;
  int64_t* sums = static_cast<int64_t*>(work);
  int64_t* grand = sums + nb;
  hipLaunchKernelGGL((scan_block_sums<TC>), dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0,
                     s, cnt, n, sums);
  FGX_TRY(launch_check("scan_block_sums"));
  hipLaunchKernelGGL(scan_sums, dim3(1), dim3(kScanThreads), 0, s, sums, nb, grand);
  FGX_TRY(launch_check("scan_sums"));
  hipLaunchKernelGGL((scan_write_ptr<TP, TC>), dim3(static_cast<unsigned>(nb)), dim3(kScanThreads),
                     0, s, cnt, n, sums, base, ptr);
  FGX_TRY(launch_check("scan_write_ptr"));
  if (total_host) {
    FGX_HIP(hipMemcpyAsync(pinned, grand, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    FGX_HIP(hipStreamSynchronize(s));
    *total_host = *pinned;
  }
  return FGX_OK;
}

}  // namespace fgx
