// fgx_device.hpp — device-side building blocks shared by the gfx950 kernels:
// value conversions, streaming loads, wave64 primitives, and the fiber
// search/walk that maps a contiguous range of stored positions back to the
// fibers (rows of CSR, columns of CSC) that own them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fgx {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Value types.  BF16 and F16 are storage-only: they are widened to FP32 on
// load and every accumulation runs in FP32.
struct bf16_t { uint16_t bits; };
struct f16_t { uint16_t bits; };

__device__ __forceinline__ float to_f32(float v) { return v; }
__device__ __forceinline__ double to_f64(double v) { return v; }
__device__ __forceinline__ float to_f32(bf16_t v) {
  return __uint_as_float(static_cast<uint32_t>(v.bits) << 16);
}
__device__ __forceinline__ float to_f32(f16_t v) {
  _Float16 h;
  __builtin_memcpy(&h, &v.bits, 2);
  return static_cast<float>(h);
}

// Accumulator type of a value type: F64 stays F64, everything else is FP32.
template <typename TV> struct acc_of { using type = float; };
template <> struct acc_of<double> { using type = double; };

template <typename TA, typename TV>
__device__ __forceinline__ TA widen(TV v) { return static_cast<TA>(to_f32(v)); }
template <>
__device__ __forceinline__ double widen<double, double>(double v) { return v; }

// ---------------------------------------------------------------------------
// Streaming loads for the operand's ptr/idx/val arrays: each element is read
// once, so they bypass the caches' retention policy (nontemporal) and leave
// L2/MALL capacity to the gathered dense operand.
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}
template <>
__device__ __forceinline__ bf16_t ld_stream<bf16_t>(const bf16_t* p) {
  bf16_t r;
  r.bits = __builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(p));
  return r;
}
template <>
__device__ __forceinline__ f16_t ld_stream<f16_t>(const f16_t* p) {
  f16_t r;
  r.bits = __builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(p));
  return r;
}

// V consecutive elements starting at p.  When `vec` holds (host-checked
// 16-byte alignment of the base and of every lane's start), the run is read
// with 16-byte vector loads.
template <typename T, int V>
__device__ __forceinline__ void ld_run(const T* __restrict__ p, bool vec, T (&out)[V]) {
  constexpr int bytes = static_cast<int>(sizeof(T)) * V;
  if constexpr (bytes % 16 == 0) {
    if (vec) {
      typedef int v4i __attribute__((ext_vector_type(4)));
      const v4i* p4 = reinterpret_cast<const v4i*>(p);
      v4i tmp[bytes / 16];
#pragma unroll
      for (int c = 0; c < bytes / 16; c++) tmp[c] = __builtin_nontemporal_load(p4 + c);
      __builtin_memcpy(&out[0], &tmp[0], bytes);
      return;
    }
  }
#pragma unroll
  for (int v = 0; v < V; v++) out[v] = ld_stream(p + v);
}

// ---------------------------------------------------------------------------
// Wave64 primitives.
__device__ __forceinline__ int lane_id() { return static_cast<int>(__lane_id()); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int popc(uint64_t m) { return __popcll(m); }

template <typename T>
__device__ __forceinline__ T shfl(T v, int src) { return __shfl(v, src, kWave); }
template <typename T>
__device__ __forceinline__ T shfl_up(T v, int d) { return __shfl_up(v, d, kWave); }

// ---------------------------------------------------------------------------
// Fiber positions.  pos(r) = ptr[r] - base is the first stored position of
// fiber r; fiber r owns [pos(r), pos(r+1)).
template <typename TP>
__device__ __forceinline__ int64_t pos_at(const TP* ptr, int64_t base, int64_t r) {
  return static_cast<int64_t>(ptr[r]) - base;
}

// Index of the calling wave within its block, as a wave-uniform (scalar)
// value: everything derived from it (tile numbers, work units, descriptors)
// then lives in scalar registers.
__device__ __forceinline__ int wave_in_block() {
  return __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
}

// Smallest r in [lo, hi) with pos(r+1) > target, or hi if none — the fiber
// that owns stored position `target` (skipping empty fibers).  Cooperative
// over a group of `G` lanes (G = 32 or 64, aligned) that all pass the same
// arguments: each round samples G fiber ends and narrows the range G-fold,
// so a million fibers take four dependent loads instead of twenty.
template <int G, typename TP>
__device__ int64_t group_find_fiber(const TP* __restrict__ ptr, int64_t base, int64_t lo,
                                    int64_t hi, int64_t target) {
  const int lane = lane_id();
  const int gl = lane & (G - 1);
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane & ~(G - 1)));
  while (hi > lo) {
    const int64_t span = hi - lo;
    if (span <= G) {
      bool p = gl < span && pos_at(ptr, base, lo + gl + 1) <= target;
      return lo + popc(ballot(p) & gmask);
    }
    const int64_t step = (span + G - 1) / G;
    int64_t off = (gl + 1) * step;
    if (off > span) off = span;
    bool p = pos_at(ptr, base, lo + off - 1 + 1) <= target;
    const int c = popc(ballot(p) & gmask);
    int64_t nlo = lo + ((int64_t)c * step < span ? (int64_t)c * step : span);
    int64_t nhi = hi;
    if (c < G) {
      int64_t e = ((int64_t)c + 1) * step;
      nhi = lo + (e < span ? e : span) - 1;
    }
    lo = nlo;
    hi = nhi;
  }
  return lo;
}

// Per-wave window of fiber ends cached in LDS: win[t] = pos(r0 + t + 1) for
// t < n.  Lanes find and walk their fibers in LDS instead of global memory.
template <typename TP>
struct FiberWindow {
  const TP* ptr;     // global fallback
  int64_t base;
  int64_t r0;        // first fiber of the window
  int64_t n;         // fibers in the window
  const int64_t* win;  // LDS copy (valid when cached)
  bool cached;

  __device__ __forceinline__ int64_t end_of(int64_t r) const {  // pos(r+1)
    return cached ? win[r - r0] : pos_at(ptr, base, r + 1);
  }
  // Smallest r in [r0, r0+n) with pos(r+1) > target (r0+n if none).
  __device__ __forceinline__ int64_t find(int64_t target) const {
    int64_t lo = r0, hi = r0 + n;
    while (lo < hi) {
      int64_t mid = lo + ((hi - lo) >> 1);
      if (end_of(mid) <= target) lo = mid + 1; else hi = mid;
    }
    return lo;
  }
};

// Loads the window for fibers [r0, r0+n) into `lds` (one wave, coalesced).
template <typename TP>
__device__ __forceinline__ FiberWindow<TP> load_window(const TP* __restrict__ ptr, int64_t base,
                                                       int64_t r0, int64_t n, int64_t* lds,
                                                       int cap) {
  FiberWindow<TP> w{ptr, base, r0, n, lds, n <= cap};
  if (w.cached) {
    for (int64_t t = lane_id(); t < n; t += kWave) lds[t] = pos_at(ptr, base, r0 + t + 1);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  return w;
}

}  // namespace fgx
