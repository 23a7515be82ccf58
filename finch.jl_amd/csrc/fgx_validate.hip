// fgx_validate.hip — structural check of a compressed operand and the
// context's safe mode.
//
// A SparseList level stores fibre p at positions ptr[p]-base .. ptr[p+1]-base-1
// with coordinates strictly increasing inside the fibre and ptr[1] == 1
// (src/tensors/levels/sparse_list_levels.jl:37-42).  Finch's generated code
// trusts these invariants in :fast mode and guards its own writes in :safe
// mode (src/execute.jl:1-11; sparse_list_levels.jl:538-547).  The kernels here
// trust them too, so an operand that breaks them could send a tile walk or a
// gather outside its buffers.  fgx_validate checks every invariant on the
// device (two bandwidth-bound passes and one synchronisation); with safe mode
// on, every entry point runs it on each compressed operand before any kernel
// reads the operand.
#include "fgx_internal.hpp"
#include "fgx_device.hpp"

#include <cstring>

namespace fgx {

// Violation bits accumulated by the check kernels.
enum : uint32_t {
  kBadFirstPtr = 1u << 0,   // ptr[0] != base
  kPtrDecreases = 1u << 1,  // ptr[p] > ptr[p+1]
  kPtrOutside = 1u << 2,    // ptr[p] outside [base, base + nnz]
  kNnzMismatch = 1u << 3,   // ptr[n_outer] - base != nnz
  kIdxOutside = 1u << 4,    // idx[q] outside [base, base + n_inner)
  kIdxOrder = 1u << 5,      // idx not strictly increasing inside a fibre
};

template <typename TP>
__global__ __launch_bounds__(256) void validate_ptr_kernel(const TP* __restrict__ ptr, int64_t base,
                                                           int64_t n_outer, int64_t nnz,
                                                           uint32_t* __restrict__ flags) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  uint32_t bad = 0;
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p <= n_outer;
       p += stride) {
    const int64_t v = static_cast<int64_t>(ptr[p]);
    if (v < base || v > base + nnz) bad |= kPtrOutside;
    if (p < n_outer && v > static_cast<int64_t>(ptr[p + 1])) bad |= kPtrDecreases;
    if (p == 0 && v != base) bad |= kBadFirstPtr;
    if (p == n_outer && v - base != nnz) bad |= kNnzMismatch;
  }
  if (bad) atomicOr(flags, bad);
}

// True when position t (0-based) starts some fibre, i.e. ptr[p] - base == t
// for some p.  Reads stay inside ptr[0..n_outer] whatever ptr holds; the
// answer is meaningful once ptr is monotone, which the ptr pass reports.
template <typename TP>
__device__ __forceinline__ bool starts_fibre(const TP* __restrict__ ptr, int64_t base,
                                             int64_t n_outer, int64_t t) {
  int64_t lo = 0, hi = n_outer + 1;
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (static_cast<int64_t>(ptr[mid]) - base < t) lo = mid + 1; else hi = mid;
  }
  return lo <= n_outer && static_cast<int64_t>(ptr[lo]) - base == t;
}

template <typename TP, typename TI>
__global__ __launch_bounds__(256) void validate_idx_kernel(const TP* __restrict__ ptr,
                                                           const TI* __restrict__ idx, int64_t base,
                                                           int64_t n_outer, int64_t n_inner,
                                                           int64_t nnz,
                                                           uint32_t* __restrict__ flags) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  uint32_t bad = 0;
  for (int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < nnz;
       q += stride) {
    const int64_t v = static_cast<int64_t>(idx[q]);
    if (v < base || v >= base + n_inner) bad |= kIdxOutside;
    if (q + 1 < nnz && static_cast<int64_t>(idx[q + 1]) <= v &&
        !starts_fibre(ptr, base, n_outer, q + 1))
      bad |= kIdxOrder;
  }
  if (bad) atomicOr(flags, bad);
}

static unsigned grid_of(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  return static_cast<unsigned>(b);
}

__global__ void clear_flags_kernel(uint32_t* flags) { *flags = 0u; }

int validate_operand(fgx_ctx* ctx, const fgx_compressed* A, hipStream_t s, const char* name,
                     uint32_t* flags) {
  FGX_TRY(check_compressed(A, name));
  IdxCombo ic;
  FGX_TRY(idx_combo(A, &ic));
  if (!flags) {
    void* scr;
    FGX_TRY(scratch(ctx, s, 256, &scr));
    flags = static_cast<uint32_t*>(scr);
  }
  hipLaunchKernelGGL(clear_flags_kernel, dim3(1), dim3(1), 0, s, flags);
  FGX_TRY(launch_check("clear_flags_kernel"));
  const int64_t base = A->index_base;
  FGX_DISPATCH_IDX(ic, {
    hipLaunchKernelGGL((validate_ptr_kernel<TP>), dim3(grid_of(A->n_outer + 1)), dim3(256), 0, s,
                       static_cast<const TP*>(A->ptr), base, A->n_outer, A->nnz, flags);
    FGX_TRY(launch_check("validate_ptr_kernel"));
    if (A->nnz > 0) {
      hipLaunchKernelGGL((validate_idx_kernel<TP, TI>), dim3(grid_of(A->nnz)), dim3(256), 0, s,
                         static_cast<const TP*>(A->ptr), static_cast<const TI*>(A->idx), base,
                         A->n_outer, A->n_inner, A->nnz, flags);
      FGX_TRY(launch_check("validate_idx_kernel"));
    }
  });
  FGX_HIP(hipMemcpyAsync(ctx->host_i64, flags, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  FGX_HIP(hipStreamSynchronize(s));
  uint32_t bad;
  std::memcpy(&bad, ctx->host_i64, sizeof(bad));
  if (bad == 0) return FGX_OK;
  // Protocol violations (the level's buffers are not a SparseList) come
  // first; a coordinate or count that disagrees with the shape is a
  // dimension mismatch.
  const uint32_t proto = bad & (kBadFirstPtr | kPtrDecreases | kPtrOutside | kIdxOrder);
  return set_error(proto ? FGX_EPROTO : FGX_EDIM,
                   "%s: malformed operand (flags 0x%x:%s%s%s%s%s%s)", name, bad,
                   (bad & kBadFirstPtr) ? " ptr[0] != base;" : "",
                   (bad & kPtrDecreases) ? " ptr decreases;" : "",
                   (bad & kPtrOutside) ? " ptr outside [base, base+nnz];" : "",
                   (bad & kNnzMismatch) ? " ptr[n]-base != nnz;" : "",
                   (bad & kIdxOutside) ? " idx outside 1:n_inner;" : "",
                   (bad & kIdxOrder) ? " idx not increasing within a fibre;" : "");
}

int admit_operand(fgx_ctx* ctx, const fgx_compressed* A, hipStream_t s, const char* name) {
  if (!ctx->safe) return FGX_OK;
  return validate_operand(ctx, A, s, name);
}

}  // namespace fgx

using namespace fgx;

extern "C" {

int fgx_validate(fgx_ctx* ctx, const fgx_compressed* A, void* stream) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_validate: NULL ctx");
  std::lock_guard<std::mutex> lk(ctx->mu);
  FGX_HIP(hipSetDevice(ctx->device));
  return validate_operand(ctx, A, pick_stream(ctx, stream), "fgx_validate");
}

int fgx_ctx_set_safe(fgx_ctx* ctx, int safe) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_ctx_set_safe: NULL ctx");
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->safe = safe != 0;
  return FGX_OK;
}

}  // extern "C"
