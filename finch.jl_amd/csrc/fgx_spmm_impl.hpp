// fgx_spmm_impl.hpp — SpMM kernels: C = A*B with A CSR and B, C dense.
//
// Row-major B/C take the nnz-balanced lane-group tiles of fgx_csr_tiles.hpp
// (16-byte column slices per lane, rows of B gathered once per stored
// position, C rows written once); other layouts take a strided kernel.
#pragma once
#include "fgx_launch.hpp"

#ifndef FGX_SPMM_E
#define FGX_SPMM_E 64
#endif
#ifndef FGX_SPMM_U
#define FGX_SPMM_U 8
#endif
// B rows per gather batch for wide groups when B is larger than
// FGX_SPMM_SMALL_B bytes.  Measured on the 1M-row Zipf(1.5) operand with its
// columns folded onto fewer rows of B (k = 64): with B at 4 MB (L2-resident)
// 8 rows per batch take 0.297 ms and 16 take 0.357 ms; with B at 64 MB
// 0.589 vs 0.573 ms, and at 256 MB 0.617 vs 0.604 ms.  Larger batches keep
// more gathers in flight per wave where they come from HBM; smaller ones
// leave room for more waves where L2 serves them.
#ifndef FGX_SPMM_U_LARGE_B
#define FGX_SPMM_U_LARGE_B 16
#endif
#ifndef FGX_SPMM_SMALL_B
#define FGX_SPMM_SMALL_B (int64_t(16) << 20)
#endif

namespace fgx {

// FGX_SPMM_WAVES (measurement builds): resident waves per SIMD to hold the
// tile kernel's register budget to.
#ifdef FGX_SPMM_WAVES
#define FGX_SPMM_OCCUPANCY __attribute__((amdgpu_waves_per_eu(FGX_SPMM_WAVES)))
#else
#define FGX_SPMM_OCCUPANCY
#endif

template <typename TP, typename TI, typename TV, typename TA, int G, int KV, int E, int U>
__global__ __launch_bounds__(256) FGX_SPMM_OCCUPANCY void spmm_csr_kernel(
    const TP* __restrict__ ptr, const TI* __restrict__ idx, const TV* __restrict__ val,
    int64_t base, int64_t m, int64_t nnz, int64_t k, int64_t kofs, const TA* __restrict__ B,
    int64_t ldb, TA* __restrict__ C, int64_t ldc, int64_t ntiles, const int64_t* __restrict__ table,
    int64_t* __restrict__ carry_row, TA* __restrict__ carry_val, int b_off32) {
  __shared__ int64_t lds[kWavesPerBlock][kWinCap];
  const int w = wave_in_block();
  const int64_t tile = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + w;
  if (tile >= ntiles) return;
  constexpr int W = (kWave / G) * E;
  __shared__ TI lds_j[kWavesPerBlock][W];
  __shared__ TA lds_v[kWavesPerBlock][W];
  spmm_tile<TP, TI, TV, TA, G, KV, E, U>(ptr, idx, val, base, m, nnz, k, kofs, B, ldb, C, ldc,
                                         tile, table, lds[w], carry_row + tile,
                                         carry_val + tile * (G * KV), lds_j[w], lds_v[w],
                                         b_off32 != 0);
}

// General-layout SpMM (any strides): one wave per row, lanes over columns.
// Serves column-major B/C below width 4; wider ones are transposed to k-fast
// rows first.
template <typename TP, typename TI, typename TV, typename TA>
__global__ __launch_bounds__(256) void spmm_csr_strided_kernel(
    const TP* __restrict__ ptr, const TI* __restrict__ idx, const TV* __restrict__ val,
    int64_t base, int64_t m, int64_t k, const TA* __restrict__ B, int64_t b_rs, int64_t b_cs,
    TA* __restrict__ C, int64_t c_rs, int64_t c_cs) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wave_in_block();
  if (row >= m) return;
  const int lane = lane_id();
  const int64_t q0 = pos_at(ptr, base, row), q1 = pos_at(ptr, base, row + 1);
  for (int64_t c = lane; c < k; c += kWave) {
    TA acc = 0;
    for (int64_t q = q0; q < q1; q++) {
      const int64_t j = static_cast<int64_t>(idx[q]) - base;
      acc += widen<TA>(val[q]) * B[j * b_rs + c * b_cs];
    }
    C[row * c_rs + c * c_cs] = acc;
  }
}

// SpMM dispatch over the dense width: lane groups of G lanes x KV columns,
// U rows of B per gather batch.
template <typename TP, typename TI, typename TV, typename TA, int G, int KV, int U>
static int spmm_launch_u(fgx_ctx* ctx, const fgx_compressed* A, int64_t k, int64_t kofs,
                         const TA* B, int64_t ldb, TA* C, int64_t ldc, hipStream_t s) {
  constexpr int NG = kWave / G;
  // Positions per lane group (E); wide groups take FGX_SPMM_E.
  constexpr int E = (G >= 32) ? tile_e_for_g(G) : ((G >= 16) ? FGX_SPMM_E : (G >= 4 ? 32 : 16));
  constexpr int64_t W = static_cast<int64_t>(NG) * E;
  const int64_t m = A->n_outer;
  const int64_t ntiles = A->nnz > 0 ? (A->nnz + W - 1) / W : 1;
  const int width = G * KV;
  void* scr;
  FGX_TRY(scratch(ctx, s, static_cast<size_t>(ntiles) * (sizeof(int64_t) + sizeof(TA) * width), &scr));
  int64_t* crow = static_cast<int64_t*>(scr);
  TA* cval = reinterpret_cast<TA*>(crow + ntiles);
  const int64_t* table = nullptr;
  FGX_TRY(tile_table(ctx, A, W, s, &table));
  const int64_t blocks = (ntiles + kWavesPerBlock - 1) / kWavesPerBlock;
  // Whether every row of B starts within 4 GB of the slice's start.
  const int b_off32 = (A->n_inner > 0 ? (A->n_inner - 1) * ldb * static_cast<int64_t>(sizeof(TA))
                                       : 0) < (int64_t(1) << 32) - 4096;
  hipLaunchKernelGGL((spmm_csr_kernel<TP, TI, TV, TA, G, KV, E, U>),
                     dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                     static_cast<const TP*>(A->ptr), static_cast<const TI*>(A->idx),
                     static_cast<const TV*>(A->val), static_cast<int64_t>(A->index_base), m,
                     A->nnz, k, kofs, B, ldb, C, ldc, ntiles, table, crow, cval, b_off32);
  FGX_TRY(launch_check("spmm_csr_kernel"));
  if (ntiles > 1) FGX_TRY(launch_fixup<TA>(s, crow, cval, ntiles, width, C, ldc, kofs, k));
  return FGX_OK;
}

template <typename TP, typename TI, typename TV, typename TA, int G, int KV>
static int spmm_launch(fgx_ctx* ctx, const fgx_compressed* A, int64_t k, int64_t kofs,
                       const TA* B, int64_t ldb, TA* C, int64_t ldc, hipStream_t s) {
  constexpr int U = G < 8 ? G : (FGX_SPMM_U < G ? FGX_SPMM_U : G);
  constexpr int UL = G < 16 ? U : (FGX_SPMM_U_LARGE_B < G ? FGX_SPMM_U_LARGE_B : G);
  const int64_t b_bytes = A->n_inner * ldb * static_cast<int64_t>(sizeof(TA));
  if (UL != U && b_bytes > FGX_SPMM_SMALL_B)
    return spmm_launch_u<TP, TI, TV, TA, G, KV, UL>(ctx, A, k, kofs, B, ldb, C, ldc, s);
  return spmm_launch_u<TP, TI, TV, TA, G, KV, U>(ctx, A, k, kofs, B, ldb, C, ldc, s);
}

template <typename TP, typename TI, typename TV>
static int spmm_csr_typed(fgx_ctx* ctx, const fgx_compressed* A, int64_t k, const void* Bv,
                          int64_t ldb, int b_layout, void* Cv, int64_t ldc, int c_layout,
                          hipStream_t s) {
  using TA = typename acc_of<TV>::type;
  const TA* B = static_cast<const TA*>(Bv);
  TA* C = static_cast<TA*>(Cv);
  const int64_t m = A->n_outer;
  if (m == 0 || k == 0) return FGX_OK;
  if ((b_layout != FGX_ROW_MAJOR || c_layout != FGX_ROW_MAJOR) && k >= 4) {
    // Column-major dense operands (Finch's Dense(Dense) layout): relay them
    // out as k-fast rows in the context's dense arena, run the tiled kernels,
    // and transpose the result back.  Costs one extra pass over B and C.
    const int64_t n = A->n_inner;
    const size_t bt = b_layout == FGX_ROW_MAJOR ? 0 : ((static_cast<size_t>(n * k) * sizeof(TA) + 255) & ~size_t(255));
    const size_t ct = c_layout == FGX_ROW_MAJOR ? 0 : static_cast<size_t>(m * k) * sizeof(TA);
    void* arena = nullptr;
    FGX_TRY(dense_scratch(ctx, s, bt + ct, &arena));
    const TA* Bf = B;
    int64_t ldbf = ldb;
    if (bt) {
      TA* Bt = static_cast<TA*>(arena);
      FGX_TRY(col_to_row<TA>(B, n, k, ldb, Bt, k, s));
      Bf = Bt;
      ldbf = k;
    }
    TA* Cf = c_layout == FGX_ROW_MAJOR ? C : reinterpret_cast<TA*>(static_cast<char*>(arena) + bt);
    const int64_t ldcf = c_layout == FGX_ROW_MAJOR ? ldc : k;
    FGX_TRY((spmm_csr_typed<TP, TI, TV>(ctx, A, k, Bf, ldbf, FGX_ROW_MAJOR, Cf, ldcf,
                                        FGX_ROW_MAJOR, s)));
    // Row-major m x k (ld k) is column-major k x m; its row-major form with
    // ld ldc is the column-major m x k output.
    if (ct) FGX_TRY(col_to_row<TA>(Cf, k, m, k, C, ldc, s));
    return FGX_OK;
  }
  if (b_layout != FGX_ROW_MAJOR || c_layout != FGX_ROW_MAJOR) {
    const int64_t b_rs = b_layout == FGX_ROW_MAJOR ? ldb : 1;
    const int64_t b_cs = b_layout == FGX_ROW_MAJOR ? 1 : ldb;
    const int64_t c_rs = c_layout == FGX_ROW_MAJOR ? ldc : 1;
    const int64_t c_cs = c_layout == FGX_ROW_MAJOR ? 1 : ldc;
    const int64_t blocks = (m + kWavesPerBlock - 1) / kWavesPerBlock;
    hipLaunchKernelGGL((spmm_csr_strided_kernel<TP, TI, TV, TA>),
                       dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                       static_cast<const TP*>(A->ptr), static_cast<const TI*>(A->idx),
                       static_cast<const TV*>(A->val), static_cast<int64_t>(A->index_base), m, k,
                       B, b_rs, b_cs, C, c_rs, c_cs);
    return launch_check("spmm_csr_strided_kernel");
  }
  // Vector width: 16-byte lanes when every row start is 16-byte aligned.
  constexpr int KVV = sizeof(TA) == 4 ? 4 : 2;
  const bool vec = aligned16(B) && aligned16(C) && (ldb % KVV == 0) && (ldc % KVV == 0);
  const int64_t slice = vec ? 64 * KVV : 64;
  for (int64_t kofs = 0; kofs < k; kofs += slice) {
    const int64_t kw = (k - kofs < slice) ? k - kofs : slice;
    if (vec) {
      const int64_t lanes = (kw + KVV - 1) / KVV;
      if (lanes <= 2) FGX_TRY((spmm_launch<TP, TI, TV, TA, 2, KVV>(ctx, A, k, kofs, B, ldb, C, ldc, s)));
      else if (lanes <= 4) FGX_TRY((spmm_launch<TP, TI, TV, TA, 4, KVV>(ctx, A, k, kofs, B, ldb, C, ldc, s)));
      else if (lanes <= 8) FGX_TRY((spmm_launch<TP, TI, TV, TA, 8, KVV>(ctx, A, k, kofs, B, ldb, C, ldc, s)));
      else if (lanes <= 16) FGX_TRY((spmm_launch<TP, TI, TV, TA, 16, KVV>(ctx, A, k, kofs, B, ldb, C, ldc, s)));
      else if (lanes <= 32) FGX_TRY((spmm_launch<TP, TI, TV, TA, 32, KVV>(ctx, A, k, kofs, B, ldb, C, ldc, s)));
      else FGX_TRY((spmm_launch<TP, TI, TV, TA, 64, KVV>(ctx, A, k, kofs, B, ldb, C, ldc, s)));
    } else {
      if (kw <= 4) FGX_TRY((spmm_launch<TP, TI, TV, TA, 4, 1>(ctx, A, k, kofs, B, ldb, C, ldc, s)));
      else if (kw <= 16) FGX_TRY((spmm_launch<TP, TI, TV, TA, 16, 1>(ctx, A, k, kofs, B, ldb, C, ldc, s)));
      else FGX_TRY((spmm_launch<TP, TI, TV, TA, 64, 1>(ctx, A, k, kofs, B, ldb, C, ldc, s)));
    }
  }
  return FGX_OK;
}

// Per-value-type instantiations live in their own translation units so they
// compile in parallel.
#define FGX_SPMM_ARGS fgx_ctx* ctx, const fgx_compressed* A, IdxCombo ic, int64_t k, \
    const void* B, int64_t ldb, int b_layout, void* C, int64_t ldc, int c_layout, hipStream_t s
int spmm_csr_f64(FGX_SPMM_ARGS);
int spmm_csr_f32(FGX_SPMM_ARGS);
int spmm_csr_bf16(FGX_SPMM_ARGS);
int spmm_csr_f16(FGX_SPMM_ARGS);

#define FGX_SPMM_DEFINE(NAME, TYPE)                                                        \
  int NAME(FGX_SPMM_ARGS) {                                                                \
    using TV = TYPE;                                                                       \
    int rc = FGX_OK;                                                                       \
    FGX_DISPATCH_IDX(ic, rc = (spmm_csr_typed<TP, TI, TV>(ctx, A, k, B, ldb, b_layout, C,  \
                                                          ldc, c_layout, s)));             \
    return rc;                                                                             \
  }

}  // namespace fgx
