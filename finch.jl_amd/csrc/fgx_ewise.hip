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

#include <algorithm>

namespace fgx {

// Julia's `max` (base/math.jl): for floats, a NaN on either side gives NaN
// and -0.0 orders below 0.0.  Julia returns `x - y` when a NaN is present,
// which on x86-64 is the first NaN operand with its quiet bit set; the same
// bits are returned here.
__device__ __forceinline__ float ew_max(float x, float y) {
  if (x != x) return __uint_as_float(__float_as_uint(x) | 0x00400000u);
  if (y != y) return __uint_as_float(__float_as_uint(y) | 0x00400000u);
  return (y > x || (y == x && !__builtin_signbit(y) && __builtin_signbit(x))) ? y : x;
}
__device__ __forceinline__ double ew_max(double x, double y) {
  if (x != x)
    return __longlong_as_double(__double_as_longlong(x) | 0x0008000000000000ll);
  if (y != y)
    return __longlong_as_double(__double_as_longlong(y) | 0x0008000000000000ll);
  return (y > x || (y == x && !__builtin_signbit(y) && __builtin_signbit(x))) ? y : x;
}
__device__ __forceinline__ int32_t ew_max(int32_t x, int32_t y) { return y > x ? y : x; }

// `+` and `*` on Int32 wrap modulo 2^32, as Julia's Int32 arithmetic does.
__device__ __forceinline__ float ew_add(float a, float b) { return a + b; }
__device__ __forceinline__ double ew_add(double a, double b) { return a + b; }
__device__ __forceinline__ int32_t ew_add(int32_t a, int32_t b) {
  return static_cast<int32_t>(static_cast<uint32_t>(a) + static_cast<uint32_t>(b));
}
__device__ __forceinline__ float ew_mul(float a, float b) { return a * b; }
__device__ __forceinline__ double ew_mul(double a, double b) { return a * b; }
__device__ __forceinline__ int32_t ew_mul(int32_t a, int32_t b) {
  return static_cast<int32_t>(static_cast<uint32_t>(a) * static_cast<uint32_t>(b));
}

template <typename TV>
__device__ __forceinline__ TV ew_both(int op, TV a, TV b) {
  return op == FGX_OP_MUL ? ew_mul(a, b) : (op == FGX_OP_ADD ? ew_add(b, a) : ew_max(a, b));
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

// Coordinates staged in LDS per side per group.  A fibre pair that fits is
// loaded once, coalesced, and searched in LDS; a longer pair is searched in
// global memory (same results, more round trips).
template <int G>
__host__ __device__ constexpr int ew_cap() { return 8 * G; }

// Either side of a fibre pair: its coordinates staged in LDS (positions in
// the fibre are ints: a staged fibre holds at most ew_cap<G>() of them) or
// read from global memory (64-bit positions).  The kernels instantiate their
// walk once per kind, so no access decides between the two.
template <typename TI>
struct EwLds {
  const TI* s;
  int n;
  __device__ __forceinline__ TI at(int i) const { return s[i]; }
  // First i in [0, n) with coordinate >= key (n if none).
  __device__ __forceinline__ int lower(TI key) const {
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (s[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
  }
};
template <typename TI>
struct EwGlobal {
  const TI* g;  // idx + fibre start
  int64_t n;
  __device__ __forceinline__ TI at(int64_t i) const { return g[i]; }
  __device__ __forceinline__ int64_t lower(TI key) const { return lower_bound(g, 0, n, key); }
};

// Loads the group's two fibres into its LDS slots when both fit; the group's
// lanes share a wave, so a wave-scope fence orders the stores and the reads.
template <int G, typename TI>
__device__ __forceinline__ bool ew_stage(const TI* __restrict__ aidx, int64_t a0, int64_t la,
                                         const TI* __restrict__ bidx, int64_t b0, int64_t lb,
                                         TI* sa, TI* sb, int gl) {
  constexpr int CAP = ew_cap<G>();
  const bool staged = la <= CAP && lb <= CAP;
  if (staged) {
    for (int64_t i = gl; i < la; i += G) sa[i] = ld_stream(aidx + a0 + i);
    for (int64_t i = gl; i < lb; i += G) sb[i] = ld_stream(bidx + b0 + i);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return staged;
}

template <int G, typename TP, typename TI>
__global__ __launch_bounds__(256) void ewise_count_kernel(
    int op, int64_t m, int64_t base, const TP* __restrict__ aptr, const TI* __restrict__ aidx,
    const TP* __restrict__ bptr, const TI* __restrict__ bidx, int64_t* __restrict__ cnt) {
  constexpr int CAP = ew_cap<G>();
  __shared__ TI s_a[256 / G][CAP];
  __shared__ TI s_b[256 / G][CAP];
  const int64_t row = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
  if (row >= m) return;
  const int grp = threadIdx.x / G;
  const int gl = lane_id() & (G - 1);
  const uint64_t gm = group_mask<G>();
  const int64_t a0 = pos_at(aptr, base, row), a1 = pos_at(aptr, base, row + 1);
  const int64_t b0 = pos_at(bptr, base, row), b1 = pos_at(bptr, base, row + 1);
  const int64_t la = a1 - a0, lb = b1 - b0;
  const bool staged = ew_stage<G>(aidx, a0, la, bidx, b0, lb, s_a[grp], s_b[grp], gl);
  // The common coordinates: each coordinate of the shorter fibre looks for
  // its partner in the longer one.
  int64_t found = 0;
  auto walk = [&](const auto& A, const auto& B) {
    const auto& S = la <= lb ? A : B;
    const auto& L = la <= lb ? B : A;
    using Ix = decltype(S.n);
    if (L.n == 0) return;
    for (Ix c = 0; c < S.n; c += G) {
      const Ix i = c + gl;
      bool f = false;
      if (i < S.n) {
        const TI key = S.at(i);
        const Ix p = L.lower(key);
        f = p < L.n && L.at(p) == key;
      }
      found += popc(ballot(f) & gm);
    }
  };
  if (staged)
    walk(EwLds<TI>{s_a[grp], static_cast<int>(la)}, EwLds<TI>{s_b[grp], static_cast<int>(lb)});
  else
    walk(EwGlobal<TI>{aidx + a0, la}, EwGlobal<TI>{bidx + b0, lb});
  if (gl == 0) cnt[row] = (op == FGX_OP_MUL) ? found : la + lb - found;
}

template <int G, typename TP, typename TI, typename TV>
__global__ __launch_bounds__(256) void ewise_fill_kernel(
    int op, int64_t m, int64_t base, const TP* __restrict__ aptr, const TI* __restrict__ aidx,
    const TV* __restrict__ aval, const TP* __restrict__ bptr, const TI* __restrict__ bidx,
    const TV* __restrict__ bval, const TP* __restrict__ cptr, TI* __restrict__ cidx,
    TV* __restrict__ cval) {
  constexpr int CAP = ew_cap<G>();
  __shared__ TI s_a[256 / G][CAP];
  __shared__ TI s_b[256 / G][CAP];
  const int64_t row = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
  if (row >= m) return;
  const int grp = threadIdx.x / G;
  const int lane = lane_id();
  const int gl = lane & (G - 1);
  const uint64_t gm = group_mask<G>();
  const uint64_t lt = (1ull << lane) - 1;
  const int64_t a0 = pos_at(aptr, base, row), a1 = pos_at(aptr, base, row + 1);
  const int64_t b0 = pos_at(bptr, base, row), b1 = pos_at(bptr, base, row + 1);
  const int64_t c0 = pos_at(cptr, base, row);
  const int64_t la = a1 - a0, lb = b1 - b0;
  const bool uni = op != FGX_OP_MUL;
  // Intersection: the symbolic pass's ptr already says which rows have no
  // common coordinate (almost all, for sparse random patterns); they write
  // nothing and skip the staging and the searches.
  if (!uni && (la == 0 || lb == 0 || pos_at(cptr, base, row + 1) == c0)) return;
  // Union: the values of the group's first NPRE steps on each side are
  // loaded together with the coordinates (one round trip); each step then
  // loads the value NPRE steps ahead, so the stores never wait on a fresh
  // load.  Intersection reads only the values of matched coordinates.
  constexpr int NPRE = 4;
  TV pa[NPRE], pb[NPRE];
#pragma unroll
  for (int j = 0; j < NPRE; j++) {
    const int64_t ia = static_cast<int64_t>(j) * G + gl;
    pa[j] = (uni && ia < la) ? aval[a0 + ia] : TV(0);
    const int64_t ib = static_cast<int64_t>(j) * G + gl;
    pb[j] = (uni && ib < lb) ? bval[b0 + ib] : TV(0);
  }
  const bool staged = ew_stage<G>(aidx, a0, la, bidx, b0, lb, s_a[grp], s_b[grp], gl);
  const TV* av_f = aval + a0;
  const TV* bv_f = bval + b0;
  TI* ci = cidx + c0;
  TV* cv = cval + c0;
  auto walk = [&](const auto& A, const auto& B) {
    using Ix = decltype(A.n);
    const Ix na = A.n, nb = B.n;
    // A side: every stored a (union) or every a with a partner (intersection).
    Ix dup = 0;
    for (Ix c = 0; c < na; c += G) {
      const Ix i = c + gl;
      const TV av = pa[0];
#pragma unroll
      for (int j = 0; j + 1 < NPRE; j++) pa[j] = pa[j + 1];
      {
        const Ix ia = i + static_cast<Ix>(NPRE) * G;
        pa[NPRE - 1] = (uni && ia < na) ? av_f[ia] : TV(0);
      }
      bool f = false;
      Ix rb = 0;
      TI key = 0;
      if (i < na) {
        key = A.at(i);
        rb = B.lower(key);
        f = rb < nb && B.at(rb) == key;
      }
      const uint64_t fb = ballot(f) & gm;
      const Ix before = dup + popc(fb & lt);
      if (i < na) {
        if (uni) {
          const Ix pos = i + rb - before;
          ci[pos] = key;
          cv[pos] = f ? ew_both<TV>(op, av, bv_f[rb]) : ew_only_a<TV>(op, av);
        } else if (f) {
          ci[before] = key;
          cv[before] = ew_both<TV>(op, av_f[i], bv_f[rb]);
        }
      }
      dup += popc(fb);
    }
    if (!uni) return;
    // B side: the b without a partner in A.
    dup = 0;
    for (Ix c = 0; c < nb; c += G) {
      const Ix i = c + gl;
      const TV bv = pb[0];
#pragma unroll
      for (int j = 0; j + 1 < NPRE; j++) pb[j] = pb[j + 1];
      {
        const Ix ib = i + static_cast<Ix>(NPRE) * G;
        pb[NPRE - 1] = ib < nb ? bv_f[ib] : TV(0);
      }
      bool f = false;
      Ix ra = 0;
      TI key = 0;
      if (i < nb) {
        key = B.at(i);
        ra = A.lower(key);
        f = ra < na && A.at(ra) == key;
      }
      const uint64_t fb = ballot(f) & gm;
      const Ix before = dup + popc(fb & lt);
      if (i < nb && !f) {
        const Ix pos = i + ra - before;
        ci[pos] = key;
        cv[pos] = ew_only_b<TV>(op, bv);
      }
      dup += popc(fb);
    }
  };
  if (staged)
    walk(EwLds<TI>{s_a[grp], static_cast<int>(la)}, EwLds<TI>{s_b[grp], static_cast<int>(lb)});
  else
    walk(EwGlobal<TI>{aidx + a0, la}, EwGlobal<TI>{bidx + b0, lb});
}

static int pick_group(int64_t m, int64_t nnz_a, int64_t nnz_b) {
  const double avg = m > 0 ? static_cast<double>(nnz_a + nnz_b) / static_cast<double>(m) : 0.0;
  if (avg <= 48.0) return 8;
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
  FGX_TRY(admit_operand(ctx, A, s, "fgx_ewise_symbolic"));
  FGX_TRY(admit_operand(ctx, B, s, "fgx_ewise_symbolic"));
  const int64_t m = A->n_outer;
  void* scr;
  FGX_TRY(scratch(ctx, s, static_cast<size_t>(m + 1) * sizeof(int64_t) + scan_work_bytes(m), &scr));
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

int fgx_ewise_async(fgx_ctx* ctx, int op, const fgx_compressed* A, const fgx_compressed* B,
                    void* c_ptr, void* c_idx, void* c_val, int64_t c_cap, void* stream) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_ewise_async: NULL ctx");
  if (op < FGX_OP_MUL || op > FGX_OP_MAX) return set_error(FGX_EINVAL, "fgx_ewise_async: bad op");
  FGX_TRY(check_pair(A, B, "fgx_ewise_async"));
  if (!c_ptr) return set_error(FGX_EINVAL, "fgx_ewise_async: NULL c_ptr");
  // The intersection keeps at most the shorter list of each fibre pair, the
  // union at most both.
  const int64_t bound = op == FGX_OP_MUL ? std::min(A->nnz, B->nnz) : A->nnz + B->nnz;
  if (c_cap < bound || (bound > 0 && (!c_idx || !c_val)))
    return set_error(FGX_EINVAL, "fgx_ewise_async: output capacity %lld below the bound %lld",
                     (long long)c_cap, (long long)bound);
  IdxCombo ic;
  FGX_TRY(idx_combo(A, &ic));
  std::lock_guard<std::mutex> lk(ctx->mu);
  FGX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = pick_stream(ctx, stream);
  FGX_TRY(admit_operand(ctx, A, s, "fgx_ewise_async"));
  FGX_TRY(admit_operand(ctx, B, s, "fgx_ewise_async"));
  const int64_t m = A->n_outer;
  void* scr;
  FGX_TRY(scratch(ctx, s, static_cast<size_t>(m + 1) * sizeof(int64_t) + scan_work_bytes(m), &scr));
  int64_t* cnt = static_cast<int64_t*>(scr);
  void* work = cnt + m + 1;
  const int G = pick_group(m, A->nnz, B->nnz);
  const int64_t base = A->index_base;
  if (m > 0) {
    FGX_DISPATCH_IDX(ic, {
      if (G == 8) launch_count<8, TP, TI>(op, A, B, cnt, s);
      else if (G == 16) launch_count<16, TP, TI>(op, A, B, cnt, s);
      else launch_count<64, TP, TI>(op, A, B, cnt, s);
    });
    FGX_TRY(launch_check("ewise_count_kernel"));
  }
  // ptr from the counts on the device, with no read-back of the total.
  FGX_DISPATCH_IDX(ic, FGX_TRY(scan_counts_to_ptr<TP>(cnt, m, base, static_cast<TP*>(c_ptr), work,
                                                      ctx->host_i64, nullptr, s)));
  if (m == 0) return FGX_OK;
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
    default: return set_error(FGX_EUNSUP, "fgx_ewise_async: value type");
  }
#undef FGX_EW_FILL
  return launch_check("ewise_fill_kernel");
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
  FGX_TRY(admit_operand(ctx, A, s, "fgx_ewise_numeric"));
  FGX_TRY(admit_operand(ctx, B, s, "fgx_ewise_numeric"));
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
