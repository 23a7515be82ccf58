// fgx_launch.hpp — launch helpers shared by the CSR tile kernels: the
// carry fix-up kernel and the value-type dispatch.
#pragma once
#include "fgx_csr_tiles.hpp"
#include "fgx_internal.hpp"

namespace fgx {

// Width 1 (SpMV): one lane per tile.  A run of one tile, the common case (a
// row split over two tiles), is completed by its lane with one add.  Longer
// runs (a row spanning several tiles; about 8 per 64 tiles on a Zipf(1.5)
// matrix) are taken eight at a time by 8-lane subgroups: each finds its run's
// length 8 tiles per step and sums the carries lane-strided, then by a fixed
// xor tree, so the order is fixed.  Every lane reaches each ballot.
template <typename TA>
__global__ __launch_bounds__(256) void fixup_lane_kernel(const int64_t* __restrict__ carry_row,
                                                         const TA* __restrict__ carry_val,
                                                         int64_t ntiles, TA* __restrict__ out,
                                                         int64_t ld, int64_t kofs, int64_t k) {
  const int lane = lane_id();
  const int64_t wave0 = static_cast<int64_t>(blockIdx.x) * 256 + (threadIdx.x & ~63);
  const int64_t t = wave0 + lane;
  bool longrun = false;
  if (t < ntiles) {
    const int64_t r = carry_row[t];
    if (r >= 0) {
      if (t + 1 < ntiles && carry_row[t + 1] == -(r + 2)) longrun = true;
      else if (kofs < k) out[r * ld + kofs] += carry_val[t];
    }
  }
  uint64_t runs = ballot(longrun);
  const int sub = lane >> 3, sl = lane & 7;
  while (runs) {
    // Subgroup `sub` takes the sub-th remaining run, if there is one.
    uint64_t m = runs;
    for (int i = 0; i < sub && m; i++) m &= m - 1;
    const bool have = m != 0;
    const int l = have ? __ffsll(static_cast<unsigned long long>(m)) - 1 : 0;
    for (int i = 0; i < 8 && runs; i++) runs &= runs - 1;
    if (have) {
      const int64_t t0 = wave0 + l;
      const int64_t r = carry_row[t0];
      const int64_t cont = -(r + 2);
      int64_t len = 1;
      for (int64_t c = t0 + 1;; c += 8) {
        const int64_t t2 = c + sl;
        const bool in = t2 < ntiles && carry_row[t2] == cont;
        const uint32_t msk = static_cast<uint32_t>(ballot(in) >> (sub * 8)) & 0xffu;
        const int run = msk == 0xffu ? 8 : __builtin_ctz(~msk);
        len += run;
        if (run < 8) break;
      }
      TA s = 0;
      for (int64_t i = sl; i < len; i += 8) s += carry_val[t0 + i];
      s += __shfl_xor(s, 4, kWave);
      s += __shfl_xor(s, 2, kWave);
      s += __shfl_xor(s, 1, kWave);
      if (sl == 0 && kofs < k) out[r * ld + kofs] += s;
    }
  }
}

// Width > 1 (SpMM): 16 lanes per tile, 4 tiles per wave.  A one-tile run is
// added by its 16 lanes (columns strided by 16); longer runs are taken by the
// whole wave through fixup_tile.
constexpr int kFixupGroup = 16;
template <typename TA>
__global__ __launch_bounds__(256) void fixup_group_kernel(const int64_t* __restrict__ carry_row,
                                                          const TA* __restrict__ carry_val,
                                                          int64_t ntiles, int width,
                                                          TA* __restrict__ out, int64_t ld,
                                                          int64_t kofs, int64_t k) {
  constexpr int GL = kFixupGroup;
  const int lane = lane_id();
  const int gl = lane % GL;
  const int64_t wave_t0 =
      (static_cast<int64_t>(blockIdx.x) * 256 + (threadIdx.x & ~63)) / GL;
  const int64_t t = wave_t0 + lane / GL;
  bool longrun = false;
  if (t < ntiles) {
    const int64_t r = carry_row[t];
    if (r >= 0) {
      if (t + 1 < ntiles && carry_row[t + 1] == -(r + 2)) {
        longrun = true;
      } else {
        for (int col = gl; col < width; col += GL)
          if (kofs + col < k) out[r * ld + kofs + col] += carry_val[t * width + col];
      }
    }
  }
  uint64_t runs = ballot(longrun && gl == 0);
  while (runs) {
    const int l = __ffsll(static_cast<unsigned long long>(runs)) - 1;
    runs &= runs - 1;
    fixup_tile<TA>(carry_row, carry_val, ntiles, wave_t0 + l / GL, width, out, ld, kofs, k);
  }
}

template <typename TA>
static inline int launch_fixup(hipStream_t s, const int64_t* crow, const TA* cval, int64_t ntiles,
                               int width, TA* out, int64_t ld, int64_t kofs, int64_t k) {
  if (width == 1) {
    const int64_t blocks = (ntiles + 255) / 256;
    hipLaunchKernelGGL((fixup_lane_kernel<TA>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                       s, crow, cval, ntiles, out, ld, kofs, k);
    return launch_check("fixup_lane_kernel");
  }
  const int64_t blocks = (ntiles * kFixupGroup + 255) / 256;
  hipLaunchKernelGGL((fixup_group_kernel<TA>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                     s, crow, cval, ntiles, width, out, ld, kofs, k);
  return launch_check("fixup_group_kernel");
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

