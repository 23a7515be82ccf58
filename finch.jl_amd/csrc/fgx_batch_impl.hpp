// fgx_batch_impl.hpp — device side of the batched persistent kernel: the
// work-unit bodies (row-split SpMV/SpMM, nnz-split tiles, SDDMM sub-tiles)
// and the kernel that walks the work queue.  Instantiated per value type in
// fgx_batch_k_*.hip.
#pragma once
#include "fgx_batch.hpp"
#include "fgx_sddmm_core.hpp"

namespace fgx {

// ---- ROWS units ----------------------------------------------------------------------
// SpMV: kSpmvRowGroup lanes per row, strided over the row, then a shuffle tree.
template <typename TP, typename TI, typename TV>
__device__ __forceinline__ void rows_spmv(const BatchDesc& d, int64_t r0, int64_t r1) {
  constexpr int G = kSpmvRowGroup;
  const int lane = lane_id();
  const int g = lane / G, gl = lane % G;
  const TP* ptr = static_cast<const TP*>(d.ptr);
  const TI* idx = static_cast<const TI*>(d.idx);
  const TV* val = static_cast<const TV*>(d.val);
  const int64_t base = d.base;
  for (int64_t r = r0 + g; r < r1; r += kWave / G) {
    const int64_t q0 = pos_at(ptr, base, r), q1 = pos_at(ptr, base, r + 1);
    float acc = 0.f;
    for (int64_t q = q0 + gl; q < q1; q += G)
      acc += widen<float>(ld_stream(val + q)) * d.B[static_cast<int64_t>(ld_stream(idx + q)) - base];
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
    if (gl == 0) d.C[r] = acc;
  }
}

// SpMM: G lanes per row, each lane a 16-byte slice of the dense width; four
// stored positions in flight per lane.
template <int G, typename TP, typename TI, typename TV>
__device__ __forceinline__ void rows_spmm(const BatchDesc& d, int64_t r0, int64_t r1) {
  const int lane = lane_id();
  const int g = lane / G, gl = lane % G;
  const TP* ptr = static_cast<const TP*>(d.ptr);
  const TI* idx = static_cast<const TI*>(d.idx);
  const TV* val = static_cast<const TV*>(d.val);
  const int64_t base = d.base;
  const int64_t col0 = static_cast<int64_t>(gl) * 4;
  const int vc = static_cast<int>(col0 < d.k ? (d.k - col0 < 4 ? d.k - col0 : 4) : 0);
  for (int64_t r = r0 + g; r < r1; r += kWave / G) {
    const int64_t q0 = pos_at(ptr, base, r), q1 = pos_at(ptr, base, r + 1);
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    int64_t q = q0;
    for (; q + 4 <= q1; q += 4) {
      float b[4][4];
      float v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t j = static_cast<int64_t>(ld_stream(idx + q + u)) - base;
        v[u] = widen<float>(ld_stream(val + q + u));
        if (vc > 0) ld_dense<4, float>(d.B + j * d.ldb + col0, vc, b[u]);
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[c] += v[u] * b[u][c];
    }
    for (; q < q1; q++) {
      const int64_t j = static_cast<int64_t>(ld_stream(idx + q)) - base;
      const float v = widen<float>(ld_stream(val + q));
      float b[4] = {0.f, 0.f, 0.f, 0.f};
      if (vc > 0) ld_dense<4, float>(d.B + j * d.ldb + col0, vc, b);
#pragma unroll
      for (int c = 0; c < 4; c++) acc[c] += v * b[c];
    }
    if (vc > 0) st_dense<4, float>(d.C + r * d.ldc + col0, vc, acc);
  }
}

// ---- SDDMM units: 64 positions per sub-tile, one per lane ----------------------------------
// `lds` holds the wave's fibre window (kSdWin entries) followed by the X and
// Y staging tiles of the matrix-core path.
template <typename TP, typename TI, typename TV>
__device__ __forceinline__ void sddmm_unit(const BatchDesc& d, int64_t st0, int64_t st1,
                                           unsigned char* lds) {
  const TP* ptr = static_cast<const TP*>(d.ptr);
  const TI* idx = static_cast<const TI*>(d.idx);
  const TV* val = static_cast<const TV*>(d.val);
  const int64_t base = d.base;
  const int lane = lane_id();
  int64_t* win_lds = reinterpret_cast<int64_t*>(lds);
  float(*xs)[kSdPitch] = reinterpret_cast<float(*)[kSdPitch]>(lds + kSdWin * sizeof(int64_t));
  float(*ys)[kSdPitch] = xs + kWave;
  for (int64_t st = st0; st < st1; st++) {
    const int64_t q0 = st * kWave;
    const int64_t qend = (q0 + kWave < d.nnz) ? q0 + kWave : d.nnz;
    int64_t r_lo, r_hi, pos_lo;
    FiberWindow<TP> win = tile_setup(ptr, base, d.m, q0, qend, d.table, st, win_lds, &r_lo, &r_hi,
                                     &pos_lo, kSdWin);
    const int64_t q = q0 + lane;
    const bool in = q < qend;
    int64_t i = 0, j = 0;
    float a = 0.f;
    if (in) {
      j = static_cast<int64_t>(ld_stream(idx + q)) - base;
      a = widen<float>(ld_stream(val + q));
      i = win.find(q);
    }
    float dot = 0.f;
    if (d.mfma) {
      f32x4 c = {0.f, 0.f, 0.f, 0.f};
      sddmm_sub_mfma(d.B, d.ldb, d.B2, d.ldc, d.k, in, i, j, xs, ys, c);
      dot = mfma_diag(c);
    } else if (in) {
      const float* xr = d.B + i * d.ldb;
      const float* yr = d.B2 + j * d.ldc;
      for (int64_t kk = 0; kk < d.k; kk++) dot += xr[kk] * yr[kk];
    }
    if (in) d.C[q] = a * dot;
  }
}

// ---- TILES units ------------------------------------------------------------------------------
template <int G, typename TP, typename TI, typename TV>
__device__ __forceinline__ void tiles_spmm(const BatchDesc& d, int64_t t0, int64_t t1,
                                           int64_t* lds, int64_t* carry_row, float* carry_val) {
  constexpr int U = G < 8 ? G : 8;
  for (int64_t t = t0; t < t1; t++)
    spmm_tile<TP, TI, TV, float, G, 4, tile_e_for_g(G), U, false>(
        static_cast<const TP*>(d.ptr), static_cast<const TI*>(d.idx),
        static_cast<const TV*>(d.val), d.base, d.m, d.nnz, d.k, 0, d.B, d.ldb, d.C, d.ldc, t,
        d.table, lds,
        carry_row + d.carry_tile0 + t, carry_val + d.carry_val0 + t * (G * 4));
}

template <typename TP, typename TI, typename TV>
__device__ __forceinline__ void tiles_spmv(const BatchDesc& d, int64_t t0, int64_t t1,
                                           int64_t* lds, int64_t* carry_row, float* carry_val) {
  for (int64_t t = t0; t < t1; t++)
    spmv_tile<TP, TI, TV, float, 8>(static_cast<const TP*>(d.ptr), static_cast<const TI*>(d.idx),
                                    static_cast<const TV*>(d.val), d.base, d.m, d.nnz, d.B, d.C,
                                    t, d.vec != 0, d.table, lds, carry_row + d.carry_tile0 + t,
                                    carry_val + d.carry_val0 + t);
}

// Blocks are dispatched round-robin over the XCDs; renumber them so that
// consecutive logical blocks (and so consecutive work units of one item)
// share an XCD and its L2.
__device__ __forceinline__ int64_t xcd_block() {
  constexpr int kXcd = 8;
  const int64_t nb = gridDim.x;
  const int64_t b = blockIdx.x;
  if (nb % kXcd != 0) return b;
  return (b % kXcd) * (nb / kXcd) + b / kXcd;
}

// Per-wave LDS of the batch kernel: the fibre window of ROWS/TILES units, or
// (with SDDMM units) the SDDMM window plus X/Y staging tiles.
constexpr int kLdsPlain = kWinCap * static_cast<int>(sizeof(int64_t));
constexpr int kLdsSddmm = kSdWin * static_cast<int>(sizeof(int64_t)) + 2 * kWave * kSdPitch * 4;

template <typename TP, typename TI, typename TV, bool SDDMM>
__global__ __launch_bounds__(kBatchThreads) void batch_kernel(
    const BatchDesc* __restrict__ desc, const BatchWork* __restrict__ work, int64_t n_work,
    int64_t* __restrict__ carry_row, float* __restrict__ carry_val) {
  constexpr int kWaveLds = SDDMM ? (kLdsSddmm > kLdsPlain ? kLdsSddmm : kLdsPlain) : kLdsPlain;
  __shared__ __align__(16) unsigned char smem[kWavesPerBlock][kWaveLds];
  int64_t* lds[kWavesPerBlock];
#pragma unroll
  for (int i = 0; i < kWavesPerBlock; i++) lds[i] = reinterpret_cast<int64_t*>(smem[i]);
  const int w = threadIdx.x / kWave;
  const int64_t nw = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  for (int64_t u = xcd_block() * kWavesPerBlock + w; u < n_work; u += nw) {
    const BatchWork wk = work[u];
    const BatchDesc d = desc[wk.item];
    if (wk.kind == BK_ROWS) {
      if (d.k == 1) {
        rows_spmv<TP, TI, TV>(d, wk.begin, wk.end);
      } else {
        switch (group_for_k(d.k)) {
          case 2: rows_spmm<2, TP, TI, TV>(d, wk.begin, wk.end); break;
          case 4: rows_spmm<4, TP, TI, TV>(d, wk.begin, wk.end); break;
          case 8: rows_spmm<8, TP, TI, TV>(d, wk.begin, wk.end); break;
          case 16: rows_spmm<16, TP, TI, TV>(d, wk.begin, wk.end); break;
          case 32: rows_spmm<32, TP, TI, TV>(d, wk.begin, wk.end); break;
          default: rows_spmm<64, TP, TI, TV>(d, wk.begin, wk.end); break;
        }
      }
    } else if (wk.kind == BK_TILES) {
      if (d.k == 1) {
        tiles_spmv<TP, TI, TV>(d, wk.begin, wk.end, lds[w], carry_row, carry_val);
      } else {
        switch (group_for_k(d.k)) {
          case 2: tiles_spmm<2, TP, TI, TV>(d, wk.begin, wk.end, lds[w], carry_row, carry_val); break;
          case 4: tiles_spmm<4, TP, TI, TV>(d, wk.begin, wk.end, lds[w], carry_row, carry_val); break;
          case 8: tiles_spmm<8, TP, TI, TV>(d, wk.begin, wk.end, lds[w], carry_row, carry_val); break;
          case 16: tiles_spmm<16, TP, TI, TV>(d, wk.begin, wk.end, lds[w], carry_row, carry_val); break;
          case 32: tiles_spmm<32, TP, TI, TV>(d, wk.begin, wk.end, lds[w], carry_row, carry_val); break;
          default: tiles_spmm<64, TP, TI, TV>(d, wk.begin, wk.end, lds[w], carry_row, carry_val); break;
        }
      }
    } else if constexpr (SDDMM) {
      sddmm_unit<TP, TI, TV>(d, wk.begin, wk.end, smem[w]);
    }
  }
}

// Launches batch_kernel<., ., TV, SDDMM> for the plan's index types.
#define FGX_BATCH_DEFINE(NAME, TYPE, SDDMM_)                                                 \
  int NAME(const fgx_batch_plan* p, hipStream_t s) {                                         \
    using TV = TYPE;                                                                         \
    FGX_DISPATCH_IDX(static_cast<IdxCombo>(p->combo),                                        \
                     hipLaunchKernelGGL((batch_kernel<TP, TI, TV, SDDMM_>),                  \
                                        dim3(static_cast<unsigned>(p->grid)),                \
                                        dim3(kBatchThreads), 0, s, p->d_desc, p->d_work,     \
                                        p->n_work, p->d_carry_row, p->d_carry_val));         \
    return FGX_OK;                                                                           \
  }

int batch_kernel_f32(const fgx_batch_plan* p, hipStream_t s);
int batch_kernel_f32_sddmm(const fgx_batch_plan* p, hipStream_t s);
int batch_kernel_bf16(const fgx_batch_plan* p, hipStream_t s);
int batch_kernel_bf16_sddmm(const fgx_batch_plan* p, hipStream_t s);

}  // namespace fgx
