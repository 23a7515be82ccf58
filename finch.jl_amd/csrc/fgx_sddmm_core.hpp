// fgx_sddmm_core.hpp — the matrix-core dot products of SDDMM, shared by the
// plain SDDMM kernel and the batched/serving kernel.
//
// A wave takes 64 stored positions at a time, one per lane, and computes the
// 64 dot products X[i,:].Y[j,:] with V_MFMA_F32_4X4X1_16B_F32: its 16 blocks
// each form the 4x4 outer product of four lanes' X and Y values; summed over
// the width, the diagonal of block b holds the dot products of lanes
// 4b..4b+3.  Only the diagonal is used (a sampled product has no reuse across
// positions), a quarter of the MFMA rate, which still runs ahead of the row
// gathers that bound the kernel.
#pragma once
#include "fgx_csr_tiles.hpp"

namespace fgx {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kSddmmSub = 4;     // 64-position sub-tiles per wave tile
constexpr int kSdChunk = 32;     // dense columns staged per pass: one 128-byte line per row
constexpr int kSdPitch = 36;     // LDS row pitch (floats): conflict-free 16-byte reads
constexpr int kSdWin = 256;      // fibre ends cached per wave

// One 64-position sub-tile: lane t owns position q0+t.  The Y rows (of the
// positions' columns: random rows) are staged through LDS 32 columns (one
// 128-byte line) at a time with coalesced 16-byte loads (8 lanes per row),
// so each line is requested from L2 once; then each lane reads
// back its own row's values.  The X rows (of the positions' rows) are read
// by each lane directly: positions of one row share its address, so a wave
// touches only the few rows its positions fall in.  Each lane then feeds its
// X and Y values to the matrix core.
// TD is the dense factors' storage type: float, or BF16 (read 8 bytes per
// 4-column slice and widened before staging).
// `whole`: k is a multiple of 32 and the sub-tile's wave holds at least one
// position.  The gathers then take no branch: a lane without a position reads
// row 0 of X and Y (its i and j are 0), and only its own dot product, which is
// never stored, sees those values (the diagonal of a 4x4 block takes each
// lane's X and Y alone).
template <typename TD>
__device__ __forceinline__ void sddmm_sub_mfma(const TD* __restrict__ X, int64_t ldx,
                                               const TD* __restrict__ Y, int64_t ldy,
                                               int64_t k, bool in, int64_t i, int64_t j,
                                               float (*ys)[kSdPitch], f32x4& c,
                                               bool whole = false) {
  constexpr int kLanesPerRow = kSdChunk / 4;          // 16-byte slices per staged row
  constexpr int kRowsPerLoad = kWave / kLanesPerRow;   // rows one wave load covers
  const int lane = lane_id();
  const int part = lane % kLanesPerRow;
#ifdef FGX_SDDMM_LDS_DIRECT
  // Measurement build only (staging comparison): FP32 Y rows go straight from
  // memory to LDS with global_load_lds_dwordx4 (no VGPR round trip, no LDS
  // write instructions).  Each load writes lane l's 16 bytes at byte 16*l of
  // its 1 KB block, so rows are packed 128 bytes apart; lane l of row t loads
  // slice (l % 8) ^ (t % 8), and the read-back of slice q of row t finds it at
  // slot q ^ (t % 8), which spreads eight consecutive rows' reads over all
  // banks.  Only when every 32-column chunk is full (no zero padding needed).
  if constexpr (std::is_same<TD, float>::value) {
    if (k % kSdChunk == 0) {
      float* flat = &ys[0][0];
      for (int64_t c0 = 0; c0 < k; c0 += kSdChunk) {
        float4 xv[kSdChunk / 4];
#pragma unroll
        for (int q = 0; q < kSdChunk / 4; q++)
          xv[q] = in ? *reinterpret_cast<const float4*>(X + i * ldx + c0 + 4 * q)
                     : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int s = 0; s < kLanesPerRow; s++) {
          const int t = lane / kLanesPerRow + kRowsPerLoad * s;
          const int64_t jt = shfl(j, t);
          const int tin = shfl(static_cast<int>(in), t);
          const int slice = part ^ (t % kLanesPerRow);
          if (tin)
            __builtin_amdgcn_global_load_lds(
                static_cast<const void*>(Y + jt * ldy + c0 + 4 * slice),
                (__attribute__((address_space(3))) void*)(flat + s * kWave * 4),
                16, 0, 0);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int q = 0; q < kSdChunk / 4; q++) {
          const float4 a4 = xv[q];
          const float4 b4 = *reinterpret_cast<const float4*>(
              flat + lane * kSdChunk + 4 * (q ^ (lane % kLanesPerRow)));
          c = __builtin_amdgcn_mfma_f32_4x4x1f32(a4.x, b4.x, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_4x4x1f32(a4.y, b4.y, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_4x4x1f32(a4.z, b4.z, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_4x4x1f32(a4.w, b4.w, c, 0, 0, 0);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      return;
    }
  }
#endif
  for (int64_t c0 = 0; c0 < k; c0 += kSdChunk) {
    // Loads first (storage form), conversions after every load is issued.
    DenseRaw<4, TD> yr[kLanesPerRow];
    DenseRaw<4, TD> xr[kSdChunk / 4];
    if (whole) {
#pragma unroll
      for (int s = 0; s < kLanesPerRow; s++) {
        const int t = lane / kLanesPerRow + kRowsPerLoad * s;
        ld_raw<4, TD>(Y + shfl(j, t) * ldy + c0 + 4 * part, 4, yr[s]);
      }
#pragma unroll
      for (int q = 0; q < kSdChunk / 4; q++) ld_raw<4, TD>(X + i * ldx + c0 + 4 * q, 4, xr[q]);
    } else {
#pragma unroll
      for (int s = 0; s < kLanesPerRow; s++) {
        const int t = lane / kLanesPerRow + kRowsPerLoad * s;
        const int64_t jt = shfl(j, t);
        const int tin = shfl(static_cast<int>(in), t);
        const int64_t col = c0 + 4 * part;
        if (tin && col < k) ld_raw<4, TD>(Y + jt * ldy + col, 4, yr[s]);
        else zero_raw<4, TD>(yr[s]);
      }
#pragma unroll
      for (int q = 0; q < kSdChunk / 4; q++) {
        const int64_t col = c0 + 4 * q;
        if (in && col < k) ld_raw<4, TD>(X + i * ldx + col, 4, xr[q]);
        else zero_raw<4, TD>(xr[q]);
      }
    }
    float4 xv[kSdChunk / 4];
#pragma unroll
    for (int q = 0; q < kSdChunk / 4; q++) {
      float o[4];
      widen_raw<4, TD, float>(xr[q], o);
      xv[q] = make_float4(o[0], o[1], o[2], o[3]);
    }
#pragma unroll
    for (int s = 0; s < kLanesPerRow; s++) {
      const int t = lane / kLanesPerRow + kRowsPerLoad * s;
      float o[4];
      widen_raw<4, TD, float>(yr[s], o);
      *reinterpret_cast<float4*>(&ys[t][4 * part]) = make_float4(o[0], o[1], o[2], o[3]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int q = 0; q < kSdChunk / 4; q++) {
      const float4 a4 = xv[q];
      const float4 b4 = *reinterpret_cast<const float4*>(&ys[lane][4 * q]);
#ifdef FGX_SDDMM_VALU_DOT
      // Measurement build only (MFMA-vs-VALU comparison): the same staging,
      // with the dot accumulated by VALU FMAs into this lane's diagonal slot.
      const int d = lane & 3;
      float acc = d == 0 ? c[0] : (d == 1 ? c[1] : (d == 2 ? c[2] : c[3]));
      acc = __builtin_fmaf(a4.x, b4.x, acc);
      acc = __builtin_fmaf(a4.y, b4.y, acc);
      acc = __builtin_fmaf(a4.z, b4.z, acc);
      acc = __builtin_fmaf(a4.w, b4.w, acc);
      c[0] = acc; c[1] = acc; c[2] = acc; c[3] = acc;
#else
      c = __builtin_amdgcn_mfma_f32_4x4x1f32(a4.x, b4.x, c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_4x4x1f32(a4.y, b4.y, c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_4x4x1f32(a4.z, b4.z, c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_4x4x1f32(a4.w, b4.w, c, 0, 0, 0);
#endif
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Lane t's dot product: the diagonal entry of its 4x4 block, held in
// accumulator register t % 4.
__device__ __forceinline__ float mfma_diag(const f32x4& c) {
  const int d = lane_id() & 3;
  return d == 0 ? c[0] : (d == 1 ? c[1] : (d == 2 ? c[2] : c[3]));
}

// One wave tile of kWave * kSddmmSub stored positions: the fibre window is
// set up once, then each 64-position sub-tile computes its dot products (on
// the matrix cores when `mfma`, else on the VALU in column order) and
// stores out[q] = A[q] * dot.  Shared by the SDDMM kernel and the batch /
// serving kernels, so a position's result is the same in either.
template <typename TP, typename TI, typename TV, typename TD = float>
__device__ __forceinline__ void sddmm_tile(const TP* __restrict__ ptr, const TI* __restrict__ idx,
                                           const TV* __restrict__ val, int64_t base, int64_t m,
                                           int64_t nnz, int64_t k, const TD* __restrict__ X,
                                           int64_t ldx, const TD* __restrict__ Y, int64_t ldy,
                                           float* __restrict__ out, int64_t tile,
                                           const int64_t* __restrict__ table, int64_t* lds_win,
                                           float (*ys)[kSdPitch], bool mfma) {
  constexpr int64_t W = kWave * kSddmmSub;
  const int lane = lane_id();
  const int64_t q0 = tile * W;
  const int64_t qend = (q0 + W < nnz) ? q0 + W : nnz;
  int64_t r_lo, r_hi, pos_lo;
  FiberWindow<TP> win =
      tile_setup(ptr, base, m, q0, qend, table, tile, lds_win, &r_lo, &r_hi, &pos_lo, kSdWin);
  const bool whole = k % kSdChunk == 0 && qend > q0;  // wave-uniform
#pragma unroll 1
  for (int sub = 0; sub < kSddmmSub; sub++) {
    const int64_t q = q0 + sub * kWave + lane;
    const bool in = q < qend;
    int64_t i = 0, j = 0;
    float a = 0.f;
    if (in) {
      j = static_cast<int64_t>(ld_stream(idx + q)) - base;
      a = widen<float>(ld_stream(val + q));
      i = win.find(q);
    }
    float dot = 0.f;
    if (mfma) {
      f32x4 c = {0.f, 0.f, 0.f, 0.f};
      sddmm_sub_mfma<TD>(X, ldx, Y, ldy, k, in, i, j, ys, c, whole);
      dot = mfma_diag(c);
    } else if (in) {
      const TD* xr = X + i * ldx;
      const TD* yr = Y + j * ldy;
      for (int64_t kk = 0; kk < k; kk++) dot += widen<float>(xr[kk]) * widen<float>(yr[kk]);
    }
    if (in) __builtin_nontemporal_store(a * dot, out + q);
    if (q0 + (sub + 1) * kWave >= qend) break;
  }
}

}  // namespace fgx
