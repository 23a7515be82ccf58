// fgx_spmm.hip — SpMM entry point (fgx_spmm_csr).
#include "fgx_spmm_impl.hpp"


using namespace fgx;

extern "C" {

int fgx_spmm_csr(fgx_ctx* ctx, const fgx_compressed* A, int64_t k, const void* B, int64_t ldb,
                 int b_layout, void* C, int64_t ldc, int c_layout, void* stream) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_spmm_csr: NULL ctx");
  FGX_TRY(check_compressed(A, "fgx_spmm_csr"));
  if (k < 0) return set_error(FGX_EINVAL, "fgx_spmm_csr: negative k");
  if (k > 0 && A->n_outer > 0 && !C) return set_error(FGX_EINVAL, "fgx_spmm_csr: NULL C");
  if (k > 0 && A->nnz > 0 && !B) return set_error(FGX_EINVAL, "fgx_spmm_csr: NULL B");
  if ((b_layout != FGX_ROW_MAJOR && b_layout != FGX_COL_MAJOR) ||
      (c_layout != FGX_ROW_MAJOR && c_layout != FGX_COL_MAJOR))
    return set_error(FGX_EINVAL, "fgx_spmm_csr: bad layout");
  const int64_t b_need = b_layout == FGX_ROW_MAJOR ? k : A->n_inner;
  const int64_t c_need = c_layout == FGX_ROW_MAJOR ? k : A->n_outer;
  if (ldb < b_need || ldc < c_need)
    return set_error(FGX_EDIM, "fgx_spmm_csr: leading dimension too small (ldb %lld, ldc %lld)",
                     (long long)ldb, (long long)ldc);
  IdxCombo ic;
  FGX_TRY(idx_combo(A, &ic));
  std::lock_guard<std::mutex> lk(ctx->mu);
  FGX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = pick_stream(ctx, stream);
  FGX_TRY(admit_operand(ctx, A, s, "fgx_spmm_csr"));
  switch (A->val_type) {
    case FGX_F64: return spmm_csr_f64(ctx, A, ic, k, B, ldb, b_layout, C, ldc, c_layout, s);
    case FGX_F32: return spmm_csr_f32(ctx, A, ic, k, B, ldb, b_layout, C, ldc, c_layout, s);
    case FGX_BF16: return spmm_csr_bf16(ctx, A, ic, k, B, ldb, b_layout, C, ldc, c_layout, s);
    case FGX_F16: return spmm_csr_f16(ctx, A, ic, k, B, ldb, b_layout, C, ldc, c_layout, s);
    default: return set_error(FGX_EUNSUP, "fgx_spmm_csr: unsupported value type %d", A->val_type);
  }
}

}  // extern "C"
