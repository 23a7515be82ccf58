// fgx_gather_floor.hip — measurement probes, built as their own library
// (libfgx_probe.so) and loaded only by the benchmark; they are not part of
// the C ABI in include/finchgpu.h.
//
// The probe gathers exactly the dense rows that a CSR SpMM C = A*B gathers
// (one row of B per stored position, in position order) and does nothing
// else: no fibre walk, no output rows, no carries.  Its time is the floor
// that the gathers alone impose on this device, so an SpMM kernel's time
// can be read against it rather than only against the HBM peak.
//
// Each row is `lanes` x 16 bytes (256 bytes for 64 FP32 columns).  A wave
// takes chunks of 64/lanes * U positions: the chunk's coordinates are read
// with coalesced loads, broadcast by shuffles, and the U row gathers of
// each lane group are all issued before the first is used.  A grid-stride
// loop over chunks keeps `blocks_per_cu` blocks of 4 waves per CU.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

template <int G, int U, typename TI>
__global__ __launch_bounds__(256) void gather_floor_kernel(const float4* __restrict__ B,
                                                           int64_t ld4,
                                                           const TI* __restrict__ idx,
                                                           int64_t base, int64_t npos,
                                                           float* __restrict__ out) {
  constexpr int NG = 64 / G;       // rows per wave step
  constexpr int P = NG * U;        // positions per chunk
  constexpr int NS = (P + 63) / 64;
  const int lane = threadIdx.x & 63;
  const int g = lane / G;
  const int gl = lane % G;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * 4;
  float acc = 0.f;
  for (int64_t c0 = wave * P; c0 < npos; c0 += nwaves * P) {
    int ids[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const int64_t q = c0 + s * 64 + lane;
      ids[s] = (s * 64 + lane < P && q < npos) ? static_cast<int>(static_cast<int64_t>(idx[q]) - base)
                                               : -1;
    }
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      // Position u*NG + g of the chunk sits in ids[(u*NG)/64] of lane
      // (u*NG)%64 + g (NG divides 64, so g never crosses a 64-block).
      const int r = __shfl(ids[(u * NG) / 64], (u * NG) % 64 + g);
      v[u] = r >= 0 ? B[static_cast<int64_t>(r) * ld4 + gl] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < U; u++) acc += (v[u].x + v[u].y) + (v[u].z + v[u].w);
  }
  out[wave * 64 + lane] = acc;
}

template <int G, typename TI>
int launch(int u, const float4* B, int64_t ld4, const TI* idx, int64_t base, int64_t npos,
           float* out, unsigned blocks, hipStream_t s) {
  switch (u) {
    case 8:
      hipLaunchKernelGGL((gather_floor_kernel<G, 8, TI>), dim3(blocks), dim3(256), 0, s, B, ld4,
                         idx, base, npos, out);
      break;
    case 16:
      hipLaunchKernelGGL((gather_floor_kernel<G, 16, TI>), dim3(blocks), dim3(256), 0, s, B, ld4,
                         idx, base, npos, out);
      break;
    case 32:
      hipLaunchKernelGGL((gather_floor_kernel<G, 32, TI>), dim3(blocks), dim3(256), 0, s, B, ld4,
                         idx, base, npos, out);
      break;
    default:
      return 1;
  }
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

}  // namespace

extern "C" {

// Launch-shape of the probe for `blocks_per_cu` on the current device: the
// number of floats `out` must hold (one per lane of the grid).
int64_t fgx_probe_gather_out_len(int blocks_per_cu) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return -1;
  return static_cast<int64_t>(cus) * blocks_per_cu * 256;
}

// Gathers row (idx[q] - base) of B for q in [0, npos): rows of `lanes` * 16
// bytes, `ld4` 16-byte units apart; idx_bytes 4 or 8; u in {8, 16, 32} rows
// per lane group in flight; lanes in {8, 16, 32}.  `out` holds
// fgx_probe_gather_out_len(blocks_per_cu) floats (per-lane sums of the
// gathered values).  Every idx[q] - base must lie in [0, rows of B) and below
// 2^31.  Returns 0, or non-zero for an unsupported shape or a launch error.
int fgx_probe_gather_rows(const void* B, int64_t ld4, int lanes, const void* idx, int idx_bytes,
                          int64_t base, int64_t npos, float* out, int64_t out_len, int u,
                          int blocks_per_cu, hipStream_t s) {
  if (blocks_per_cu < 1 || npos < 0 || ld4 < lanes) return 1;
  const int64_t need = fgx_probe_gather_out_len(blocks_per_cu);
  if (need < 0 || out_len < need) return 1;
  const unsigned blocks = static_cast<unsigned>(need / 256);
  const float4* Bv = static_cast<const float4*>(B);
  if (idx_bytes == 4) {
    const int32_t* ix = static_cast<const int32_t*>(idx);
    if (lanes == 8) return launch<8>(u, Bv, ld4, ix, base, npos, out, blocks, s);
    if (lanes == 16) return launch<16>(u, Bv, ld4, ix, base, npos, out, blocks, s);
    if (lanes == 32) return launch<32>(u, Bv, ld4, ix, base, npos, out, blocks, s);
  } else if (idx_bytes == 8) {
    const int64_t* ix = static_cast<const int64_t*>(idx);
    if (lanes == 8) return launch<8>(u, Bv, ld4, ix, base, npos, out, blocks, s);
    if (lanes == 16) return launch<16>(u, Bv, ld4, ix, base, npos, out, blocks, s);
    if (lanes == 32) return launch<32>(u, Bv, ld4, ix, base, npos, out, blocks, s);
  }
  return 1;
}

}  // extern "C"

// ---- scatter-add probes -------------------------------------------------------------
// The rate of the two accumulation forms a CSC SpMV can use, measured apart
// from any SpMV: n float adds of vals[q] into y[rows[q]] through device
// atomics, and n adds of (row, value) entries into an LDS array.

namespace {

// The XCD (accelerator complex die) this wave runs on: HW_REG_XCC_ID, bits [3:0].
__device__ __forceinline__ int xcc_id() {
  return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 7;
}

// mode 0: device-scope atomic add into y; 1: workgroup-scope atomic add
// into the copy of y of this wave's XCD (y holds 8 copies of m); 2: device
// scope into that copy.
template <int MODE>
__global__ __launch_bounds__(256) void atomic_scatter_kernel(const int* __restrict__ rows,
                                                             const float* __restrict__ vals,
                                                             int64_t n, float* y, int64_t m) {
  float* yc = MODE == 0 ? y : y + static_cast<int64_t>(xcc_id()) * m;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * 256;
  for (int64_t q = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; q < n; q += stride) {
    if (MODE == 1)
      __hip_atomic_fetch_add(yc + rows[q], vals[q], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
    else
      unsafeAtomicAdd(yc + rows[q], vals[q]);
  }
}

// Entries (row & 8191, value) added into an 8,192-entry LDS array by
// 1024-thread blocks, each taking an equal share of the n entries.
// mode 0: LDS float atomic add; 1: LDS integer atomic add; 2: plain LDS write.
template <int MODE>
__global__ __launch_bounds__(1024) void lds_accumulate_kernel(const uint2* __restrict__ in,
                                                              int64_t n, float* __restrict__ out) {
  __shared__ float acc[8192];
  __shared__ int iacc[8192];
  for (int i = threadIdx.x; i < 8192; i += 1024) {
    acc[i] = 0.f;
    iacc[i] = 0;
  }
  __syncthreads();
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = static_cast<int64_t>(blockIdx.x) * per;
  const int64_t hi = lo + per < n ? lo + per : n;
  constexpr int U = 8;
  int64_t q = lo + threadIdx.x;
  for (; q + (U - 1) * 1024 < hi; q += U * 1024) {
    uint2 t[U];
#pragma unroll
    for (int u = 0; u < U; u++) t[u] = in[q + u * 1024];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const unsigned r = t[u].x & 8191u;
      const float v = __uint_as_float(t[u].y);
      if (MODE == 0) atomicAdd(&acc[r], v);
      else if (MODE == 1) atomicAdd(&iacc[r], static_cast<int>(v));
      else acc[r] = v;
    }
  }
  for (; q < hi; q += 1024) {
    const uint2 t = in[q];
    atomicAdd(&acc[t.x & 8191u], __uint_as_float(t.y));
  }
  __syncthreads();
  float s = 0.f;
  for (int i = threadIdx.x; i < 8192; i += 1024) s += acc[i] + static_cast<float>(iacc[i]);
  out[static_cast<int64_t>(blockIdx.x) * 1024 + threadIdx.x] = s;
}

}  // namespace

extern "C" {

// n adds of vals[q] into y[rows[q]], rows in [0, m); y holds m floats (mode
// 0) or 8 * m (modes 1 and 2, one copy per XCD).  Returns 0 or non-zero.
int fgx_probe_atomic_scatter(const int* rows, const float* vals, int64_t n, float* y, int64_t m,
                             int mode, int blocks_per_cu, hipStream_t s) {
  const int64_t words = fgx_probe_gather_out_len(blocks_per_cu);
  if (words < 0 || n < 0 || m <= 0) return 1;
  const unsigned blocks = static_cast<unsigned>(words / 256);
  if (mode == 0)
    hipLaunchKernelGGL(atomic_scatter_kernel<0>, dim3(blocks), dim3(256), 0, s, rows, vals, n, y, m);
  else if (mode == 1)
    hipLaunchKernelGGL(atomic_scatter_kernel<1>, dim3(blocks), dim3(256), 0, s, rows, vals, n, y, m);
  else if (mode == 2)
    hipLaunchKernelGGL(atomic_scatter_kernel<2>, dim3(blocks), dim3(256), 0, s, rows, vals, n, y, m);
  else
    return 1;
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

// n entries (row, float bits) added into LDS by `blocks` blocks of 1024
// threads; `out` holds blocks * 1024 floats.  Returns 0 or non-zero.
int fgx_probe_lds_accumulate(const void* entries, int64_t n, int mode, float* out, int64_t out_len,
                             int blocks, hipStream_t s) {
  if (blocks < 1 || n < 0 || out_len < static_cast<int64_t>(blocks) * 1024) return 1;
  const uint2* in = static_cast<const uint2*>(entries);
  if (mode == 0)
    hipLaunchKernelGGL(lds_accumulate_kernel<0>, dim3(blocks), dim3(1024), 0, s, in, n, out);
  else if (mode == 1)
    hipLaunchKernelGGL(lds_accumulate_kernel<1>, dim3(blocks), dim3(1024), 0, s, in, n, out);
  else if (mode == 2)
    hipLaunchKernelGGL(lds_accumulate_kernel<2>, dim3(blocks), dim3(1024), 0, s, in, n, out);
  else
    return 1;
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

}  // extern "C"
