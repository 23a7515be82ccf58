// fgx_launch.hpp — launch helpers shared by the CSR tile kernels: the
// carry fix-up kernel and the value-type dispatch.
#pragma once
#include "fgx_csr_tiles.hpp"
#include "fgx_internal.hpp"

namespace fgx {

// One chunk of fixup_chunk_tiles(width) tiles per wave.
template <typename TA>
__global__ __launch_bounds__(256) void fixup_chunk_kernel(const int64_t* __restrict__ carry_row,
                                                          const TA* __restrict__ carry_val,
                                                          int64_t ntiles, int width,
                                                          TA* __restrict__ out, int64_t ld,
                                                          int64_t kofs, int64_t k) {
  const int64_t chunk = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wave_in_block();
  const int64_t t0 = chunk * fixup_chunk_tiles(width);
  if (t0 >= ntiles) return;  // wave-uniform
  fixup_chunk<TA>(carry_row, carry_val, ntiles, t0, t0 + fixup_chunk_tiles(width), width, out,
                  ld, kofs, k);
}

template <typename TA>
static inline int launch_fixup(hipStream_t s, const int64_t* crow, const TA* cval, int64_t ntiles,
                               int width, TA* out, int64_t ld, int64_t kofs, int64_t k) {
  const int64_t chunks = (ntiles + fixup_chunk_tiles(width) - 1) / fixup_chunk_tiles(width);
  const int64_t blocks = (chunks + kWavesPerBlock - 1) / kWavesPerBlock;
  hipLaunchKernelGGL((fixup_chunk_kernel<TA>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                     s, crow, cval, ntiles, width, out, ld, kofs, k);
  return launch_check("fixup_chunk_kernel");
}

// dst[r * ldd + c] = src[r + c * lds] for r < rows, c < cols: a column-major
// matrix to row-major (and, viewed transposed, row-major to column-major).
// 64 x 64 tiles through LDS, both sides coalesced; a 1-D grid of tiles.
template <typename T>
__global__ __launch_bounds__(256) void col_to_row_kernel(const T* __restrict__ src, int64_t rows,
                                                         int64_t cols, int64_t lds,
                                                         T* __restrict__ dst, int64_t ldd) {
  __shared__ T tile[64][65];
  const int64_t ntr = (rows + 63) / 64;
  const int64_t id = blockIdx.x;
  const int64_t r0 = (id % ntr) * 64, c0 = (id / ntr) * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int i = ty; i < 64; i += 4) {
    const int64_t r = r0 + tx, c = c0 + i;
    if (r < rows && c < cols) tile[i][tx] = src[r + c * lds];
  }
  __syncthreads();
  for (int i = ty; i < 64; i += 4) {
    const int64_t r = r0 + i, c = c0 + tx;
    if (r < rows && c < cols) dst[r * ldd + c] = tile[tx][i];
  }
}

template <typename T>
static int col_to_row(const T* src, int64_t rows, int64_t cols, int64_t lds, T* dst, int64_t ldd,
                      hipStream_t s) {
  if (rows == 0 || cols == 0) return FGX_OK;
  const int64_t tiles = ((rows + 63) / 64) * ((cols + 63) / 64);
  hipLaunchKernelGGL((col_to_row_kernel<T>), dim3(static_cast<unsigned>(tiles)), dim3(256), 0, s,
                     src, rows, cols, lds, dst, ldd);
  return launch_check("col_to_row_kernel");
}

}  // namespace fgx

#define FGX_DISPATCH_VAL(vt, ...)                                              \
  switch (vt) {                                                                \
    case FGX_F64: { using TV = double; __VA_ARGS__; break; }                   \
    case FGX_F32: { using TV = float; __VA_ARGS__; break; }                    \
    case FGX_BF16: { using TV = bf16_t; __VA_ARGS__; break; }                  \
    case FGX_F16: { using TV = f16_t; __VA_ARGS__; break; }                    \
    default: return set_error(FGX_EUNSUP, "unsupported value type %d", vt);    \
  }

