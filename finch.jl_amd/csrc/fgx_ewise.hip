// fgx_ewise.hip — element-wise coiteration of two compressed operands,
// C = op(A, B) fibre by fibre (Finch's two-finger merge,
// test/reference64/typical/typical_merge_twofinger.txt:45-186), as a
// symbolic/numeric pair.
//
// A group of G lanes owns one fibre.  Instead of walking two fingers, every
// stored coordinate finds its partner by a binary search in the other fibre,
// so a group handles G coordinates per step and the output position of each
// is known in closed form:
//   union (ADD, MAX):   pos(a in A) = #A<a + #B<a - #(A∩B)<a
//                       pos(b in B\A) = #B<b + #A<b - #(A∩B)<b
//   intersection (MUL): pos(a in A∩B) = #(A∩B)<a
// where the #(A∩B)< terms are running ballot counts of found partners.
// Values follow the reference: both stored -> b + a, max(a, b) or a * b;
// one side stored -> the value itself (ADD) or max against the fill 0 (MAX).
#include "fgx_scan.hpp"

namespace fgx {

template <typename TV>
__device__ __forceinline__ TV ew_max(TV a, TV b) { return a < b ? b : a; }

template <typename TV>
__device__ __forceinline__ TV ew_both(int op, TV a, TV b) {
  return op == FGX_OP_MUL ? a * b : (op == FGX_OP_ADD ? b + a : ew_max(a, b));
}
template <typename TV>
__device__ __forceinline__ TV ew_only_a(int op, TV a) { return op == FGX_OP_ADD ? a : ew_max(a, TV(0)); }
template <typename TV>
__device__ __forceinline__ TV ew_only_b(int op, TV b) { return op == FGX_OP_ADD ? b : ew_max(TV(0), b); }

// First position p in [lo, hi) with idx[p] >= key (hi if none).
template <typename TI>
__device__ __forceinline__ int64_t lower_bound(const TI* __restrict__ idx, int64_t lo, int64_t hi,
                                               int64_t key) {
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (static_cast<int64_t>(idx[mid]) < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

template <int G>
__device__ __forceinline__ uint64_t group_mask() {
  const int lane = lane_id();
  return (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane & ~(G - 1)));
}

template <int G, typename TP, typename TI>
__global__ __launch_bounds__(256) void ewise_count_kernel(
    int op, int64_t m, int64_t base, const TP* __restrict__ aptr, const TI* __restrict__ aidx,
    const TP* __restrict__ bptr, const TI* __restrict__ bidx, int64_t* __restrict__ cnt) {
  const int64_t row = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
  if (row >= m) return;
  const int gl = lane_id() & (G - 1);
  const uint64_t gm = group_mask<G>();
  const int64_t a0 = pos_at(aptr, base, row), a1 = pos_at(aptr, base, row + 1);
  const int64_t b0 = pos_at(bptr, base, row), b1 = pos_at(bptr, base, row + 1);
  int64_t found = 0;
  if (b1 > b0) {
    for (int64_t c = a0; c < a1; c += G) {
      const int64_t p = c + gl;
      bool f = false;
      if (p < a1) {
        const int64_t key = aidx[p];
        const int64_t lb = lower_bound(bidx, b0, b1, key);
        f = lb < b1 && static_cast<int64_t>(bidx[lb]) == key;
      }
      found += popc(ballot(f) & gm);
    }
  }
  if (gl == 0) cnt[row] = (op == FGX_OP_MUL) ? found : (a1 - a0) + (b1 - b0) - found;
}

template <int G, typename TP, typename TI, typename TV>
__global__ __launch_bounds__(256) void ewise_fill_kernel(
    int op, int64_t m, int64_t base, const TP* __restrict__ aptr, const TI* __restrict__ aidx,
    const TV* __restrict__ aval, const TP* __restrict__ bptr, const TI* __restrict__ bidx,
    const TV* __restrict__ bval, const TP* __restrict__ cptr, TI* __restrict__ cidx,
    TV* __restrict__ cval) {
  const int64_t row = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
  if (row >= m) return;
  const int lane = lane_id();
  const int gl = lane & (G - 1);
  const uint64_t gm = group_mask<G>();
  const uint64_t lt = (1ull << lane) - 1;
  const int64_t a0 = pos_at(aptr, base, row), a1 = pos_at(aptr, base, row + 1);
  const int64_t b0 = pos_at(bptr, base, row), b1 = pos_at(bptr, base, row + 1);
  const int64_t c0 = pos_at(cptr, base, row);
  const bool uni = op != FGX_OP_MUL;
  // A side: every stored a (union) or every a with a partner (intersection).
  int64_t dup = 0;
  for (int64_t c = a0; c < a1; c += G) {
    const int64_t p = c + gl;
    bool f = false;
    int64_t lb = b0;
    int64_t key = 0;
    if (p < a1) {
      key = aidx[p];
      lb = lower_bound(bidx, b0, b1, key);
      f = lb < b1 && static_cast<int64_t>(bidx[lb]) == key;
    }
    const uint64_t fb = ballot(f) & gm;
    const int64_t before = dup + popc(fb & lt);
    if (p < a1) {
      if (uni) {
        const int64_t pos = c0 + (p - a0) + (lb - b0) - before;
        cidx[pos] = static_cast<TI>(key);
        cval[pos] = f ? ew_both<TV>(op, aval[p], bval[lb]) : ew_only_a<TV>(op, aval[p]);
      } else if (f) {
        const int64_t pos = c0 + before;
        cidx[pos] = static_cast<TI>(key);
        cval[pos] = ew_both<TV>(op, aval[p], bval[lb]);
      }
    }
    dup += popc(fb);
  }
  if (!uni) return;
  // B side: the b without a partner in A.
  dup = 0;
  for (int64_t c = b0; c < b1; c += G) {
    const int64_t q = c + gl;
    bool f = false;
    int64_t la = a0;
    int64_t key = 0;
    if (q < b1) {
      key = bidx[q];
      la = lower_bound(aidx, a0, a1, key);
      f = la < a1 && static_cast<int64_t>(aidx[la]) == key;
    }
    const uint64_t fb = ballot(f) & gm;
    const int64_t before = dup + popc(fb & lt);
    if (q < b1 && !f) {
      const int64_t pos = c0 + (q - b0) + (la - a0) - before;
      cidx[pos] = static_cast<TI>(key);
      cval[pos] = ew_only_b<TV>(op, bval[q]);
    }
    dup += popc(fb);
  }
}

static int pick_group(int64_t m, int64_t nnz_a, int64_t nnz_b) {
  const double avg = m > 0 ? static_cast<double>(nnz_a + nnz_b) / static_cast<double>(m) : 0.0;
  if (avg <= 24.0) return 8;
  if (avg <= 96.0) return 16;
  return 64;
}

static int check_pair(const fgx_compressed* A, const fgx_compressed* B, const char* what) {
  FGX_TRY(check_compressed(A, what));
  FGX_TRY(check_compressed(B, what));
  if (A->n_outer != B->n_outer || A->n_inner != B->n_inner)
    return set_error(FGX_EDIM, "%s: shapes (%lld, %lld) and (%lld, %lld) differ", what,
                     (long long)A->n_outer, (long long)A->n_inner, (long long)B->n_outer,
                     (long long)B->n_inner);
  if (A->ptr_type != B->ptr_type || A->idx_type != B->idx_type || A->val_type != B->val_type ||
      A->index_base != B->index_base)
    return set_error(FGX_EUNSUP, "%s: operands must share ptr/idx/value types and index base", what);
  if (A->val_type != FGX_F64 && A->val_type != FGX_F32 && A->val_type != FGX_I32)
    return set_error(FGX_EUNSUP, "%s: value type %d unsupported", what, A->val_type);
  return FGX_OK;
}

template <int G, typename TP, typename TI>
static void launch_count(int op, const fgx_compressed* A, const fgx_compressed* B, int64_t* cnt,
                         hipStream_t s) {
  const int64_t m = A->n_outer;
  const int64_t blocks = (m * G + 255) / 256;
  hipLaunchKernelGGL((ewise_count_kernel<G, TP, TI>), dim3(static_cast<unsigned>(blocks)),
                     dim3(256), 0, s, op, m, static_cast<int64_t>(A->index_base),
                     static_cast<const TP*>(A->ptr), static_cast<const TI*>(A->idx),
                     static_cast<const TP*>(B->ptr), static_cast<const TI*>(B->idx), cnt);
}

template <int G, typename TP, typename TI, typename TV>
static void launch_fill(int op, const fgx_compressed* A, const fgx_compressed* B, const void* cptr,
                        void* cidx, void* cval, hipStream_t s) {
  const int64_t m = A->n_outer;
  const int64_t blocks = (m * G + 255) / 256;
  hipLaunchKernelGGL((ewise_fill_kernel<G, TP, TI, TV>), dim3(static_cast<unsigned>(blocks)),
                     dim3(256), 0, s, op, m, static_cast<int64_t>(A->index_base),
                     static_cast<const TP*>(A->ptr), static_cast<const TI*>(A->idx),
                     static_cast<const TV*>(A->val), static_cast<const TP*>(B->ptr),
                     static_cast<const TI*>(B->idx), static_cast<const TV*>(B->val),
                     static_cast<const TP*>(cptr), static_cast<TI*>(cidx),
                     static_cast<TV*>(cval));
}

}  // namespace fgx

using namespace fgx;

extern "C" {

int fgx_ewise_symbolic(fgx_ctx* ctx, int op, const fgx_compressed* A, const fgx_compressed* B,
                       void* c_ptr, int64_t* c_nnz, void* stream) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_ewise_symbolic: NULL ctx");
  if (op < FGX_OP_MUL || op > FGX_OP_MAX) return set_error(FGX_EINVAL, "fgx_ewise_symbolic: bad op");
  FGX_TRY(check_pair(A, B, "fgx_ewise_symbolic"));
  if (!c_ptr || !c_nnz) return set_error(FGX_EINVAL, "fgx_ewise_symbolic: NULL output");
  IdxCombo ic;
  FGX_TRY(idx_combo(A, &ic));
  std::lock_guard<std::mutex> lk(ctx->mu);
  FGX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = pick_stream(ctx, stream);
  const int64_t m = A->n_outer;
  void* scr;
  FGX_TRY(scratch(ctx, static_cast<size_t>(m + 1) * sizeof(int64_t) + scan_work_bytes(m), &scr));
  int64_t* cnt = static_cast<int64_t*>(scr);
  void* work = cnt + m + 1;
  const int G = pick_group(m, A->nnz, B->nnz);
  if (m > 0) {
    FGX_DISPATCH_IDX(ic, {
      if (G == 8) launch_count<8, TP, TI>(op, A, B, cnt, s);
      else if (G == 16) launch_count<16, TP, TI>(op, A, B, cnt, s);
      else launch_count<64, TP, TI>(op, A, B, cnt, s);
    });
    FGX_TRY(launch_check("ewise_count_kernel"));
  }
  int64_t total = 0;
  const int64_t base = A->index_base;
  FGX_DISPATCH_IDX(ic, FGX_TRY(scan_counts_to_ptr<TP>(cnt, m, base, static_cast<TP*>(c_ptr), work,
                                                      ctx->host_i64, &total, s)));
  *c_nnz = total;
  return FGX_OK;
}

int fgx_ewise_numeric(fgx_ctx* ctx, int op, const fgx_compressed* A, const fgx_compressed* B,
                      const void* c_ptr, void* c_idx, void* c_val, void* stream) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_ewise_numeric: NULL ctx");
  if (op < FGX_OP_MUL || op > FGX_OP_MAX) return set_error(FGX_EINVAL, "fgx_ewise_numeric: bad op");
  FGX_TRY(check_pair(A, B, "fgx_ewise_numeric"));
  if (!c_ptr) return set_error(FGX_EINVAL, "fgx_ewise_numeric: NULL c_ptr");
  IdxCombo ic;
  FGX_TRY(idx_combo(A, &ic));
  std::lock_guard<std::mutex> lk(ctx->mu);
  FGX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = pick_stream(ctx, stream);
  const int64_t m = A->n_outer;
  if (m == 0) return FGX_OK;
  const int G = pick_group(m, A->nnz, B->nnz);
#define FGX_EW_FILL(TV_)                                                              \
  FGX_DISPATCH_IDX(ic, {                                                              \
    if (G == 8) launch_fill<8, TP, TI, TV_>(op, A, B, c_ptr, c_idx, c_val, s);        \
    else if (G == 16) launch_fill<16, TP, TI, TV_>(op, A, B, c_ptr, c_idx, c_val, s); \
    else launch_fill<64, TP, TI, TV_>(op, A, B, c_ptr, c_idx, c_val, s);              \
  })
  switch (A->val_type) {
    case FGX_F64: FGX_EW_FILL(double); break;
    case FGX_F32: FGX_EW_FILL(float); break;
    case FGX_I32: FGX_EW_FILL(int32_t); break;
    default: return set_error(FGX_EUNSUP, "fgx_ewise_numeric: value type");
  }
#undef FGX_EW_FILL
  return launch_check("ewise_fill_kernel");
}

}  // extern "C"
