// fgx_batch_impl.hpp — device side of the batched persistent kernel: the
// work-unit bodies (row-split SpMV/SpMM, nnz-split tiles, SDDMM sub-tiles)
// and the kernel that walks the work queue.  Instantiated per value type in
// fgx_batch_k_*.hip.
#pragma once
#include <algorithm>
#include <atomic>
#include "fgx_batch.hpp"
#include "fgx_sddmm_core.hpp"

namespace fgx {

// ---- ROWS units ----------------------------------------------------------------------
// Row-split: a lane group owns a row at a time, with no carries, tile table
// or fix-up.  Each row's coordinates/values are loaded a chunk at a time in
// one round trip and the dependent gathers are issued together, so a short
// row costs two or three round trips.

// SpMV: kSpmvRowGroup lanes per row; each lane takes every G-th position of
// a chunk of kRowsSpmvRun * G positions, then a shuffle tree per row.
constexpr int kRowsSpmvRun = 4;

template <typename TP, typename TI, typename TV, typename TD>
__device__ __forceinline__ void rows_spmv(const BatchDesc& d, int64_t r0, int64_t r1) {
  constexpr int G = kSpmvRowGroup;
  constexpr int S = kRowsSpmvRun;
  const int lane = lane_id();
  const int g = lane / G, gl = lane % G;
  const TP* ptr = static_cast<const TP*>(d.ptr);
  const TI* idx = static_cast<const TI*>(d.idx);
  const TV* val = static_cast<const TV*>(d.val);
  const TD* x = static_cast<const TD*>(d.B);
  float* y = static_cast<float*>(d.C);
  const int64_t base = d.base;
  for (int64_t r = r0 + g; r < r1; r += kWave / G) {
    const int64_t q0 = pos_at(ptr, base, r), q1 = pos_at(ptr, base, r + 1);
    float acc = 0.f;
    for (int64_t c = q0; c < q1; c += S * G) {
      TI jj[S];
      TV vv[S];
#pragma unroll
      for (int u = 0; u < S; u++) {
        const int64_t q = c + gl + u * G;
        jj[u] = q < q1 ? ld_stream(idx + q) : static_cast<TI>(base);
        vv[u] = q < q1 ? ld_stream(val + q) : TV{};
      }
      // Unconditional gathers (past-the-row slots read x[0] and are not
      // accumulated), so they issue back to back.
      float xg[S];
#pragma unroll
      for (int u = 0; u < S; u++) xg[u] = widen<float>(x[static_cast<int64_t>(jj[u]) - base]);
#pragma unroll
      for (int u = 0; u < S; u++)
        if (c + gl + u * G < q1) acc += widen<float>(vv[u]) * xg[u];
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
    if (gl == 0) y[r] = acc;
  }
}

// Gathers in flight per lane: 4 slices of 32 bytes (8 FP32 columns), or of
// 16 bytes for groups of 8 lanes and more 8, a one-lane group 4, and narrower
// groups one per lane of the group (4 for two lanes measured no faster).
template <int G, int KV, typename TD>
constexpr int spmm_u() {
  return KV * static_cast<int>(sizeof(TD)) > 16
             ? 4
             : (G >= 8 ? 8 : (G == 1 ? 4 : (G < 4 ? G : 4)));
}

// SpMM: G lanes per row, each lane KV columns of the dense width.  A row is
// walked in chunks of G positions (lane gl loads position c + gl), broadcast
// within the group, with U rows of B gathered at a time; the sum runs over
// the row's positions in order.
template <int G, int KV, typename TP, typename TI, typename TV, typename TD>
__device__ __forceinline__ void rows_spmm(const BatchDesc& d, int64_t r0, int64_t r1) {
  constexpr int U = spmm_u<G, KV, TD>();
  const int lane = lane_id();
  const int g = lane / G, gl = lane % G;
  const int gbase = g * G;
  const TP* ptr = static_cast<const TP*>(d.ptr);
  const TI* idx = static_cast<const TI*>(d.idx);
  const TV* val = static_cast<const TV*>(d.val);
  const int64_t base = d.base;
  const TD* Bd = static_cast<const TD*>(d.B);
  float* Cd = static_cast<float*>(d.C);
  const int64_t col0 = static_cast<int64_t>(gl) * KV;
  const int vc = static_cast<int>(col0 < d.k ? (d.k - col0 < KV ? d.k - col0 : KV) : 0);
  for (int64_t r = r0 + g; r < r1; r += kWave / G) {
    const int64_t q0 = pos_at(ptr, base, r), q1 = pos_at(ptr, base, r + 1);
    float acc[KV];
#pragma unroll
    for (int cc = 0; cc < KV; cc++) acc[cc] = 0.f;
    for (int64_t c = q0; c < q1; c += G) {
      const bool inc = c + gl < q1;
      const int64_t jc = inc ? static_cast<int64_t>(ld_stream(idx + c + gl)) - base : 0;
      const float vcv = inc ? widen<float>(ld_stream(val + c + gl)) : 0.f;
      const int cnt = static_cast<int>(q1 - c < G ? q1 - c : G);
      for (int e0 = 0; e0 < cnt; e0 += U) {
        DenseRaw<KV, TD> bv[U];
        float ve[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int src = gbase + ((e0 + u) % G);
          const int64_t je = shfl(jc, src);
          ve[u] = shfl(vcv, src);
          if (e0 + u < cnt && vc > 0) ld_raw<KV, TD>(Bd + je * d.ldb + col0, vc, bv[u]);
          else zero_raw<KV, TD>(bv[u]);
        }
#pragma unroll
        for (int u = 0; u < U; u++)
          if (e0 + u < cnt) {
            float bw[KV];
            widen_raw<KV, TD, float>(bv[u], bw);
#pragma unroll
            for (int cc = 0; cc < KV; cc++) acc[cc] += ve[u] * bw[cc];
          }
      }
    }
    if (vc > 0) st_dense<KV, float>(Cd + r * d.ldc + col0, vc, acc);
  }
}

// ---- SDDMM units: wave tiles of kWave * kSddmmSub positions ---------------------------------
// `lds` holds the wave's fibre window (kSdWin entries) followed by the Y
// staging tile of the matrix-core path.
template <typename TP, typename TI, typename TV, typename TD>
__device__ __forceinline__ void sddmm_unit(const BatchDesc& d, int64_t t0, int64_t t1,
                                           unsigned char* lds) {
  int64_t* win_lds = reinterpret_cast<int64_t*>(lds);
  float(*ys)[kSdPitch] = reinterpret_cast<float(*)[kSdPitch]>(lds + kSdWin * sizeof(int64_t));
  for (int64_t t = t0; t < t1; t++)
    sddmm_tile<TP, TI, TV, TD>(static_cast<const TP*>(d.ptr), static_cast<const TI*>(d.idx),
                               static_cast<const TV*>(d.val), d.base, d.m, d.nnz, d.k,
                               static_cast<const TD*>(d.B), d.ldb, static_cast<const TD*>(d.B2),
                               d.ldc, static_cast<float*>(d.C), t, d.table, win_lds, ys,
                               d.mfma != 0);
}

// Measurement switch: FP32 right-hand sides through the branch-free gathers too.
#ifndef FGX_BATCH_BRANCHFREE_F32
#define FGX_BATCH_BRANCHFREE_F32 0
#endif

// ---- TILES units ------------------------------------------------------------------------------
// Positions per TILES tile for G lanes per position.
template <int G>
constexpr int tile_w_for_g() { return (kWave / G) * tile_e_for_g(G); }

// The tile's coordinates and values are staged in LDS after the fibre window
// (`lds` + kWinCap entries), as in the single-operand SpMM kernel.
template <int G, int KV, typename TP, typename TI, typename TV, typename TD>
__device__ __forceinline__ void tiles_spmm(const BatchDesc& d, int64_t t0, int64_t t1,
                                           int64_t* lds, int64_t* carry_row, float* carry_val) {
  constexpr int U = spmm_u<G, KV, TD>();
  constexpr int W = tile_w_for_g<G>();
  static_assert(sizeof(TI) + sizeof(float) <= 12, "TILES staging holds 12 bytes per position");
  TI* lj = reinterpret_cast<TI*>(lds + kWinCap);
  float* lv = reinterpret_cast<float*>(lj + W);
  for (int64_t t = t0; t < t1; t++)
    spmm_tile<TP, TI, TV, float, G, KV, tile_e_for_g(G), U, TD>(
        static_cast<const TP*>(d.ptr), static_cast<const TI*>(d.idx),
        static_cast<const TV*>(d.val), d.base, d.m, d.nnz, d.k, 0, static_cast<const TD*>(d.B),
        d.ldb, static_cast<float*>(d.C), d.ldc, t,
        d.table, lds, carry_row + d.carry_tile0 + t, carry_val + d.carry_val0 + t * (G * KV),
        lj, lv,
        // Branch-free gathers for BF16 right-hand sides only: with FP32 ones
        // the batch bodies (held to 4 waves per SIMD) measured 4 % slower.
        (std::is_same<TD, bf16_t>::value || FGX_BATCH_BRANCHFREE_F32) &&
            d.n * d.ldb * static_cast<int64_t>(sizeof(TD)) < (int64_t(1) << 32) - 4096);
}

template <typename TP, typename TI, typename TV, typename TD>
__device__ __forceinline__ void tiles_spmv(const BatchDesc& d, int64_t t0, int64_t t1,
                                           int64_t* lds, int64_t* carry_row, float* carry_val) {
  for (int64_t t = t0; t < t1; t++)
    spmv_tile<TP, TI, TV, float, kBatchSpmvRun, TD>(static_cast<const TP*>(d.ptr), static_cast<const TI*>(d.idx),
                                    static_cast<const TV*>(d.val), d.base, d.m, d.nnz,
                                    static_cast<const TD*>(d.B), static_cast<float*>(d.C),
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

// Per-wave LDS of a class's kernel: the fibre window (ROWS/TILES), plus the
// staged coordinates/values of a TILES tile (12 bytes per position covers
// 8-byte indices), or the SDDMM window plus X/Y staging tiles.
template <int CLS>
constexpr int lds_bytes_for() {
  return CLS == kClsSddmm
             ? kSdWin * static_cast<int>(sizeof(int64_t)) + kWave * kSdPitch * 4
             : (CLS == kClsSpmv ? kWinCap * static_cast<int>(sizeof(int64_t))
                                : kWinCap * static_cast<int>(sizeof(int64_t)) +
                                      tile_w_for_g<cls_g(CLS)>() * 12);
}

// Largest s in [0, n) with seg_off[s] <= t (seg_off nondecreasing,
// seg_off[0] = 0 <= t): the segment holding tile t.  A 64-ary search, one
// coalesced probe per step; the result is wave-uniform.
__device__ __forceinline__ int64_t find_segment(const int64_t* __restrict__ seg_off, int64_t n,
                                                int64_t t) {
  int64_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const int64_t step = (hi - lo + kWave - 1) / kWave;
    const int64_t p = lo + static_cast<int64_t>(lane_id()) * step;
    const bool le = p < hi && seg_off[p] <= t;
    const int c = __popcll(ballot(le)) - 1;
    lo += static_cast<int64_t>(c) * step;
    hi = lo + step < hi ? lo + step : hi;
  }
  return lo;
}

template <typename TP, typename TI, typename TV, typename TD, int CLS>
__device__ __forceinline__ void run_unit(const BatchDesc& d, int32_t kind, int64_t b, int64_t e,
                                         unsigned char* smem, int64_t* carry_row,
                                         float* carry_val) {
  int64_t* lds = reinterpret_cast<int64_t*>(smem);
  if constexpr (CLS == kClsSddmm) {
    sddmm_unit<TP, TI, TV, TD>(d, b, e, smem);
  } else if constexpr (CLS == kClsSpmv) {
    if (kind == BK_ROWS) rows_spmv<TP, TI, TV, TD>(d, b, e);
    else tiles_spmv<TP, TI, TV, TD>(d, b, e, lds, carry_row, carry_val);
  } else {
    if (kind == BK_ROWS) rows_spmm<cls_g(CLS), cls_kv(CLS), TP, TI, TV, TD>(d, b, e);
    else tiles_spmm<cls_g(CLS), cls_kv(CLS), TP, TI, TV, TD>(d, b, e, lds, carry_row, carry_val);
  }
}

// One kernel per unit class (SDDMM, SpMV, SpMM of group width G), over the
// work entries [w0, w1) of that class: each kernel holds the registers and
// LDS of one unit body only.  Two launch forms:
//   * persistent (seg_off == nullptr): a resident grid walks the entries as
//     work units in a fixed stride;
//   * segmented (seg_off != nullptr): entry u is a whole request's tiles
//     [0, n_u) with its descriptor at desc[u], seg_off[i] the tiles before
//     entry w0 + i, and one wave runs each of the `total` tiles, so the
//     hardware can interleave the blocks of launches on different streams.
// The two forms are separate instantiations (SEG), so the single-tile
// segmented body is not held to the registers of the persistent unit loop.
// Every instantiation is held to 128 VGPRs, 4 waves per SIMD: the FP32
// persistent G=8 and G=16 bodies otherwise take 131-133 and run 3 (the few
// values that no longer fit are loop invariants, spilled once and reloaded
// per work unit).
#ifndef FGX_BATCH_MIN_WAVES
#define FGX_BATCH_MIN_WAVES 4
#endif
template <typename TP, typename TI, typename TV, typename TD, int CLS, bool SEG>
__global__ __launch_bounds__(kBatchThreads)
__attribute__((amdgpu_waves_per_eu(FGX_BATCH_MIN_WAVES))) void batch_kernel(
    const BatchDesc* __restrict__ desc, const BatchWork* __restrict__ work, int64_t w0,
    int64_t w1, const int64_t* __restrict__ seg_off, int64_t total,
    int64_t* __restrict__ carry_row, float* __restrict__ carry_val, SegInline inl) {
  __shared__ __align__(16) unsigned char smem[kWavesPerBlock][lds_bytes_for<CLS>()];
  const int w = wave_in_block();
  if constexpr (SEG) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + w;
    if (t >= total) return;
    if (inl.n > 0) {
      // Few entries (e.g. a large request launched alone): offsets and
      // descriptors are kernel arguments, read from the kernarg segment.
      int64_t u = 0;
      for (int64_t e = 1; e < inl.n; e++) u += inl.seg[e] <= t ? 1 : 0;
      const int64_t b = t - inl.seg[u];
      run_unit<TP, TI, TV, TD, CLS>(inl.desc[u], inl.desc[u].kind, b, b + 1, smem[w], carry_row,
                                carry_val);
      return;
    }
    // Segmented entries start at 0 and desc[u] is entry u's descriptor.
    const int64_t u = w0 + find_segment(seg_off, w1 - w0, t);
    const BatchDesc d = desc[u];
    const int64_t b = t - seg_off[u - w0];
    run_unit<TP, TI, TV, TD, CLS>(d, d.kind, b, b + 1, smem[w], carry_row, carry_val);
    return;
  }
  const int64_t nw = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  for (int64_t u = w0 + xcd_block() * kWavesPerBlock + w; u < w1; u += nw) {
    const BatchWork wk = work[u];
    run_unit<TP, TI, TV, TD, CLS>(desc[wk.item], wk.kind, wk.begin, wk.end, smem[w], carry_row,
                              carry_val);
  }
}

// Persistent grid: as many blocks as can be resident at once (so no block
// starts after others have drained the queue), a multiple of the 8 XCDs,
// and no more than the work needs.
inline int64_t persistent_grid(int64_t n_work, int num_cus, int blocks_per_cu) {
  const int64_t want = (n_work + kWavesPerBlock - 1) / kWavesPerBlock;
  int64_t grid = std::min<int64_t>(want, static_cast<int64_t>(num_cus) *
                                             std::max(1, blocks_per_cu));
  if (grid >= 8) grid -= grid % 8;
  return std::max<int64_t>(grid, 1);
}

template <typename TP, typename TI, typename TV, typename TD, int CLS, bool SEG>
int launch_form(fgx_batch_plan* p, int64_t w0, int64_t w1, const int64_t* seg_off, int64_t total,
                const SegInline* inl, hipStream_t s) {
  int64_t grid;
  if constexpr (SEG) {
    grid = std::max<int64_t>(1, (total + kWavesPerBlock - 1) / kWavesPerBlock);
  } else {
    static std::atomic<int> occ{0};
    int o = occ.load();
    if (o == 0) {
      int n = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_kernel<TP, TI, TV, TD, CLS, false>,
                                                       kBatchThreads, 0) != hipSuccess ||
          n < 1) {
        (void)hipGetLastError();
        n = 1;
      }
      occ.store(n);
      o = n;
    }
    grid = persistent_grid(w1 - w0, p->num_cus, o);
  }
  p->grid = std::max<int64_t>(p->grid, grid);
  hipLaunchKernelGGL((batch_kernel<TP, TI, TV, TD, CLS, SEG>), dim3(static_cast<unsigned>(grid)),
                     dim3(kBatchThreads), 0, s, p->d_desc, p->d_work, w0, w1, seg_off, total,
                     p->d_carry_row, p->d_carry_val, inl ? *inl : SegInline{});
  return launch_check("batch_kernel");
}

template <typename TP, typename TI, typename TV, typename TD, int CLS>
int launch_class(fgx_batch_plan* p, int64_t w0, int64_t w1, const int64_t* seg_off,
                 int64_t total, const SegInline* inl, hipStream_t s) {
  return seg_off ? launch_form<TP, TI, TV, TD, CLS, true>(p, w0, w1, seg_off, total, inl, s)
                 : launch_form<TP, TI, TV, TD, CLS, false>(p, w0, w1, nullptr, 0, nullptr, s);
}

// Launchers of one value type: NAME(p, cls, w0, w1, seg_off, total, inl, s)
// runs class `cls` over work entries [w0, w1) for the plan's index types
// (`inl`: the entries of a segmented class passed by value, or null).
#define FGX_BATCH_DEFINE_SPMM(NAME, TYPE, DTYPE)                                            \
  int NAME(fgx_batch_plan* p, int cls, int64_t w0, int64_t w1, const int64_t* seg_off,     \
           int64_t total, const SegInline* inl, hipStream_t s) {                            \
    using TV = TYPE;                                                                        \
    using TD = DTYPE;                                                                       \
    int rc = set_error(FGX_EUNSUP, "batch: unit class %d", cls);                           \
    FGX_DISPATCH_IDX(static_cast<IdxCombo>(p->combo), {                                     \
      switch (cls) {                                                                        \
        case kClsSpmv: rc = launch_class<TP, TI, TV, TD, kClsSpmv>(p, w0, w1, seg_off, total, inl, s); break;        \
        case 2: rc = launch_class<TP, TI, TV, TD, 2>(p, w0, w1, seg_off, total, inl, s); break;                      \
        case 4: rc = launch_class<TP, TI, TV, TD, 4>(p, w0, w1, seg_off, total, inl, s); break;                      \
        case 8: rc = launch_class<TP, TI, TV, TD, 8>(p, w0, w1, seg_off, total, inl, s); break;                      \
        case 16: rc = launch_class<TP, TI, TV, TD, 16>(p, w0, w1, seg_off, total, inl, s); break;                    \
        case 32: rc = launch_class<TP, TI, TV, TD, 32>(p, w0, w1, seg_off, total, inl, s); break;                    \
        case 64: rc = launch_class<TP, TI, TV, TD, 64>(p, w0, w1, seg_off, total, inl, s); break;                    \
        case kClsWide | 1: rc = launch_class<TP, TI, TV, TD, kClsWide | 1>(p, w0, w1, seg_off, total, inl, s); break;   \
        case kClsWide | 2: rc = launch_class<TP, TI, TV, TD, kClsWide | 2>(p, w0, w1, seg_off, total, inl, s); break;   \
        case kClsWide | 4: rc = launch_class<TP, TI, TV, TD, kClsWide | 4>(p, w0, w1, seg_off, total, inl, s); break;   \
        case kClsWide | 8: rc = launch_class<TP, TI, TV, TD, kClsWide | 8>(p, w0, w1, seg_off, total, inl, s); break;   \
        case kClsWide | 16: rc = launch_class<TP, TI, TV, TD, kClsWide | 16>(p, w0, w1, seg_off, total, inl, s); break; \
        case kClsWide | 32: rc = launch_class<TP, TI, TV, TD, kClsWide | 32>(p, w0, w1, seg_off, total, inl, s); break; \
        default: break;                                                                     \
      }                                                                                     \
    });                                                                                     \
    if (rc == FGX_OK) clear_error();                                                        \
    return rc;                                                                              \
  }

#define FGX_BATCH_DEFINE_SDDMM(NAME, TYPE, DTYPE)                                           \
  int NAME(fgx_batch_plan* p, int cls, int64_t w0, int64_t w1, const int64_t* seg_off,     \
           int64_t total, const SegInline* inl, hipStream_t s) {                            \
    using TV = TYPE;                                                                        \
    using TD = DTYPE;                                                                       \
    if (cls != kClsSddmm) return set_error(FGX_EUNSUP, "batch: unit class %d", cls);       \
    int rc = FGX_OK;                                                                        \
    FGX_DISPATCH_IDX(static_cast<IdxCombo>(p->combo),                                       \
                     rc = (launch_class<TP, TI, TV, TD, kClsSddmm>(p, w0, w1, seg_off, total, inl, s)));             \
    return rc;                                                                              \
  }

#define FGX_BATCH_LAUNCH_ARGS \
  fgx_batch_plan *p, int cls, int64_t w0, int64_t w1, const int64_t *seg_off, int64_t total, \
      const SegInline *inl, hipStream_t s
int batch_spmm_f32(FGX_BATCH_LAUNCH_ARGS);
int batch_sddmm_f32(FGX_BATCH_LAUNCH_ARGS);
int batch_spmm_bf16(FGX_BATCH_LAUNCH_ARGS);
int batch_sddmm_bf16(FGX_BATCH_LAUNCH_ARGS);
// BF16 values with BF16 dense inputs (the serving loop's all-BF16 form).
int batch_spmm_bf16_bf16in(FGX_BATCH_LAUNCH_ARGS);
int batch_sddmm_bf16_bf16in(FGX_BATCH_LAUNCH_ARGS);

}  // namespace fgx
