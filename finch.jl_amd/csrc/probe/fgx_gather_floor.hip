// fgx_gather_floor.hip — measurement probe, built as its own library
// (libfgx_probe.so) and loaded only by the benchmark; it is not part of the
// C ABI in include/finchgpu.h.
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
