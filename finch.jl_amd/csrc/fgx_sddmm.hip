// fgx_sddmm.hip — SDDMM: out[q] = A[q] * X[i,:].Y[j,:] on the pattern of A.
//
// Fast path (FP32 dense, width a multiple of 4): nnz-balanced wave tiles of
// 4 x 64 positions, dot products on the matrix cores (fgx_sddmm_core.hpp),
// Y rows staged through LDS.  Column-major X/Y (Finch's Dense(Dense)) are
// first transposed to k-fast rows in the context's dense arena.
//
// General path (FP64, other widths): VALU dot products, one lane per position.
//
// Few positions, wide rows (at most kSplitMaxNnz positions, k >= kSplitMinK):
// a wave per position, its lanes striding over k in either layout, then a
// butterfly sum, so a handful of long dot products still spreads over the
// device (Finch's documented example, 100 positions of width 1000, takes one
// wave each instead of two waves in all).
#include "fgx_launch.hpp"
#include "fgx_sddmm_core.hpp"
#include "fgx_internal.hpp"

namespace fgx {

template <typename TP, typename TI, typename TV>
__global__ __launch_bounds__(256) void sddmm_mfma_kernel(
    const TP* __restrict__ ptr, const TI* __restrict__ idx, const TV* __restrict__ val,
    int64_t base, int64_t m, int64_t nnz, int64_t k, const float* __restrict__ X, int64_t ldx,
    const float* __restrict__ Y, int64_t ldy, float* __restrict__ out, int64_t ntiles,
    const int64_t* __restrict__ table) {
  __shared__ int64_t lds[kWavesPerBlock][kSdWin];
  __shared__ __align__(16) float ys[kWavesPerBlock][kWave][kSdPitch];
  const int w = wave_in_block();
  const int64_t tile = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + w;
  if (tile >= ntiles) return;
  sddmm_tile<TP, TI, TV>(ptr, idx, val, base, m, nnz, k, X, ldx, Y, ldy, out, tile, table, lds[w],
                         ys[w], true);
}

template <typename TP, typename TI, typename TV, typename TA>
__global__ __launch_bounds__(256) void sddmm_valu_kernel(
    const TP* __restrict__ ptr, const TI* __restrict__ idx, const TV* __restrict__ val,
    int64_t base, int64_t m, int64_t nnz, int64_t k, const TA* __restrict__ X, int64_t x_rs,
    int64_t x_cs, const TA* __restrict__ Y, int64_t y_rs, int64_t y_cs, TA* __restrict__ out,
    int64_t ntiles) {
  __shared__ int64_t lds[kWavesPerBlock][kWinCap];
  const int w = wave_in_block();
  const int64_t tile = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + w;
  if (tile >= ntiles) return;
  const int lane = lane_id();
  const int64_t q0 = tile * kWave;
  const int64_t qend = (q0 + kWave < nnz) ? q0 + kWave : nnz;
  int64_t r_lo, r_hi;
  tile_fibers(ptr, base, m, q0, qend, &r_lo, &r_hi);
  const int64_t r_last = r_hi < m ? r_hi : m - 1;
  FiberWindow<TP> win = load_window(ptr, base, r_lo, r_last - r_lo + 1, lds[w], kWinCap);
  const int64_t q = q0 + lane;
  if (q >= qend) return;
  const int64_t j = static_cast<int64_t>(idx[q]) - base;
  const int64_t i = win.find(q);
  TA acc = 0;
  for (int64_t kk = 0; kk < k; kk++) acc += X[i * x_rs + kk * x_cs] * Y[j * y_rs + kk * y_cs];
  out[q] = widen<TA>(val[q]) * acc;
}

constexpr int64_t kSplitMaxNnz = 4096;
constexpr int64_t kSplitMinK = 256;

template <typename TP, typename TI, typename TV, typename TA>
__global__ __launch_bounds__(256) void sddmm_ksplit_kernel(
    const TP* __restrict__ ptr, const TI* __restrict__ idx, const TV* __restrict__ val,
    int64_t base, int64_t m, int64_t nnz, int64_t k, const TA* __restrict__ X, int64_t x_rs,
    int64_t x_cs, const TA* __restrict__ Y, int64_t y_rs, int64_t y_cs, TA* __restrict__ out) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wave_in_block();
  if (q >= nnz) return;  // wave-uniform
  const int lane = lane_id();
  const int64_t i = group_find_fiber<kWave>(ptr, base, 0, m, q);
  const int64_t j = static_cast<int64_t>(idx[q]) - base;
  const TA* x = X + i * x_rs;
  const TA* y = Y + j * y_rs;
  TA acc = 0;
  for (int64_t c = lane; c < k; c += kWave) acc += x[c * x_cs] * y[c * y_cs];
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, kWave);
  if (lane == 0) out[q] = widen<TA>(val[q]) * acc;
}

template <typename TP, typename TI, typename TV>
static int sddmm_typed(fgx_ctx* ctx, const fgx_compressed* A, int64_t k, const void* Xv,
                       int64_t ldx, int x_layout, const void* Yv, int64_t ldy, int y_layout,
                       void* outv, hipStream_t s) {
  using TA = typename acc_of<TV>::type;
  const int64_t m = A->n_outer, nnz = A->nnz;
  if (nnz == 0 || m == 0) return FGX_OK;
  const TA* X = static_cast<const TA*>(Xv);
  const TA* Y = static_cast<const TA*>(Yv);
  TA* out = static_cast<TA*>(outv);
  if (nnz <= kSplitMaxNnz && k >= kSplitMinK) {
    const int64_t x_rs = x_layout == FGX_ROW_MAJOR ? ldx : 1;
    const int64_t x_cs = x_layout == FGX_ROW_MAJOR ? 1 : ldx;
    const int64_t y_rs = y_layout == FGX_ROW_MAJOR ? ldy : 1;
    const int64_t y_cs = y_layout == FGX_ROW_MAJOR ? 1 : ldy;
    const int64_t blocks = (nnz + kWavesPerBlock - 1) / kWavesPerBlock;
    hipLaunchKernelGGL((sddmm_ksplit_kernel<TP, TI, TV, TA>), dim3(static_cast<unsigned>(blocks)),
                       dim3(256), 0, s, static_cast<const TP*>(A->ptr),
                       static_cast<const TI*>(A->idx), static_cast<const TV*>(A->val),
                       static_cast<int64_t>(A->index_base), m, nnz, k, X, x_rs, x_cs, Y, y_rs,
                       y_cs, out);
    return launch_check("sddmm_ksplit_kernel");
  }
  bool mfma = false;
  if constexpr (sizeof(TA) == 4) {
    if ((x_layout != FGX_ROW_MAJOR || y_layout != FGX_ROW_MAJOR) && k % 4 == 0) {
      const int64_t n = A->n_inner;
      const size_t xb = x_layout == FGX_ROW_MAJOR
                            ? 0 : ((static_cast<size_t>(m * k) * sizeof(TA) + 255) & ~size_t(255));
      const size_t yb = y_layout == FGX_ROW_MAJOR ? 0 : static_cast<size_t>(n * k) * sizeof(TA);
      void* arena = nullptr;
      FGX_TRY(dense_scratch(ctx, s, xb + yb, &arena));
      const TA* Xf = X;
      const TA* Yf = Y;
      int64_t ldxf = ldx, ldyf = ldy;
      if (xb) {
        TA* Xt = static_cast<TA*>(arena);
        FGX_TRY(col_to_row<TA>(X, m, k, ldx, Xt, k, s));
        Xf = Xt;
        ldxf = k;
      }
      if (yb) {
        TA* Yt = reinterpret_cast<TA*>(static_cast<char*>(arena) + xb);
        FGX_TRY(col_to_row<TA>(Y, n, k, ldy, Yt, k, s));
        Yf = Yt;
        ldyf = k;
      }
      return sddmm_typed<TP, TI, TV>(ctx, A, k, Xf, ldxf, FGX_ROW_MAJOR, Yf, ldyf, FGX_ROW_MAJOR,
                                     outv, s);
    }
    mfma = x_layout == FGX_ROW_MAJOR && y_layout == FGX_ROW_MAJOR && k % 4 == 0 && ldx % 4 == 0 &&
           ldy % 4 == 0 && aligned16(X) && aligned16(Y);
  }
  if (mfma) {
    if constexpr (sizeof(TA) == 4) {
      const int64_t W = kWave * kSddmmSub;
      const int64_t ntiles = (nnz + W - 1) / W;
      const int64_t blocks = (ntiles + kWavesPerBlock - 1) / kWavesPerBlock;
      const int64_t* table = nullptr;
      FGX_TRY(tile_table(ctx, A, W, s, &table));
      hipLaunchKernelGGL((sddmm_mfma_kernel<TP, TI, TV>), dim3(static_cast<unsigned>(blocks)),
                         dim3(256), 0, s, static_cast<const TP*>(A->ptr),
                         static_cast<const TI*>(A->idx), static_cast<const TV*>(A->val),
                         static_cast<int64_t>(A->index_base), m, nnz, k, X, ldx, Y, ldy, out,
                         ntiles, table);
      return launch_check("sddmm_mfma_kernel");
    }
  }
  const int64_t ntiles = (nnz + kWave - 1) / kWave;
  const int64_t blocks = (ntiles + kWavesPerBlock - 1) / kWavesPerBlock;
  const int64_t x_rs = x_layout == FGX_ROW_MAJOR ? ldx : 1, x_cs = x_layout == FGX_ROW_MAJOR ? 1 : ldx;
  const int64_t y_rs = y_layout == FGX_ROW_MAJOR ? ldy : 1, y_cs = y_layout == FGX_ROW_MAJOR ? 1 : ldy;
  hipLaunchKernelGGL((sddmm_valu_kernel<TP, TI, TV, TA>), dim3(static_cast<unsigned>(blocks)),
                     dim3(256), 0, s, static_cast<const TP*>(A->ptr),
                     static_cast<const TI*>(A->idx), static_cast<const TV*>(A->val),
                     static_cast<int64_t>(A->index_base), m, nnz, k, X, x_rs, x_cs, Y, y_rs, y_cs,
                     out, ntiles);
  return launch_check("sddmm_valu_kernel");
}

}  // namespace fgx

using namespace fgx;

extern "C" int fgx_sddmm_csr(fgx_ctx* ctx, const fgx_compressed* A, int64_t k, const void* X,
                             int64_t ldx, int x_layout, const void* Y, int64_t ldy, int y_layout,
                             void* out_val, void* stream) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_sddmm_csr: NULL ctx");
  FGX_TRY(check_compressed(A, "fgx_sddmm_csr"));
  if (k < 0) return set_error(FGX_EINVAL, "fgx_sddmm_csr: negative k");
  if (A->nnz > 0 && (!out_val || (k > 0 && (!X || !Y))))
    return set_error(FGX_EINVAL, "fgx_sddmm_csr: NULL buffer");
  if ((x_layout != FGX_ROW_MAJOR && x_layout != FGX_COL_MAJOR) ||
      (y_layout != FGX_ROW_MAJOR && y_layout != FGX_COL_MAJOR))
    return set_error(FGX_EINVAL, "fgx_sddmm_csr: bad layout");
  if (ldx < (x_layout == FGX_ROW_MAJOR ? k : A->n_outer) ||
      ldy < (y_layout == FGX_ROW_MAJOR ? k : A->n_inner))
    return set_error(FGX_EDIM, "fgx_sddmm_csr: leading dimension too small");
  IdxCombo ic;
  FGX_TRY(idx_combo(A, &ic));
  std::lock_guard<std::mutex> lk(ctx->mu);
  FGX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = pick_stream(ctx, stream);
  FGX_TRY(admit_operand(ctx, A, s, "fgx_sddmm_csr"));
  int rc = FGX_OK;
  switch (A->val_type) {
    case FGX_F64: { using TV = double; FGX_DISPATCH_IDX(ic, rc = (sddmm_typed<TP, TI, TV>(ctx, A, k, X, ldx, x_layout, Y, ldy, y_layout, out_val, s))); break; }
    case FGX_F32: { using TV = float; FGX_DISPATCH_IDX(ic, rc = (sddmm_typed<TP, TI, TV>(ctx, A, k, X, ldx, x_layout, Y, ldy, y_layout, out_val, s))); break; }
    case FGX_BF16: { using TV = bf16_t; FGX_DISPATCH_IDX(ic, rc = (sddmm_typed<TP, TI, TV>(ctx, A, k, X, ldx, x_layout, Y, ldy, y_layout, out_val, s))); break; }
    case FGX_F16: { using TV = f16_t; FGX_DISPATCH_IDX(ic, rc = (sddmm_typed<TP, TI, TV>(ctx, A, k, X, ldx, x_layout, Y, ldy, y_layout, out_val, s))); break; }
    default: return set_error(FGX_EUNSUP, "fgx_sddmm_csr: unsupported value type %d", A->val_type);
  }
  return rc;
}
