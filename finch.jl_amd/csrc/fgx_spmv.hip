// fgx_spmv.hip — SpMV entry points (CSR and CSC).
//
// CSR SpMV / SpMM use the nnz-balanced wave tiles of fgx_csr_tiles.hpp plus a
// fix-up launch for rows that span tiles.  CSC SpMV scatters into y with
// hardware float atomics over the same nnz-balanced tiles (the form Finch
// parallelises with AtomicElement, benchmark/benchmarks.jl:402-413).
#include "fgx_launch.hpp"

namespace fgx {

// Positions per lane run for SpMV (FGX_SPMV_RUN for 4-byte and smaller
// values; half of it for FP64).  On the 1M-row Zipf(1.5) matrix (16M
// positions, FP32), per call: 2 -> 118.1 us, 4 -> 107.4 us, 8 -> 112.8 us.
#ifndef FGX_SPMV_RUN
#define FGX_SPMV_RUN 4
#endif
template <typename TV> struct spmv_v { static constexpr int value = FGX_SPMV_RUN; };
template <> struct spmv_v<double> { static constexpr int value = FGX_SPMV_RUN / 2; };

template <typename TP, typename TI, typename TV, typename TA, int V>
__global__ __launch_bounds__(256) void spmv_csr_kernel(
    const TP* __restrict__ ptr, const TI* __restrict__ idx, const TV* __restrict__ val,
    int64_t base, int64_t m, int64_t nnz, const TA* __restrict__ x, TA* __restrict__ y,
    int64_t ntiles, int vec, const int64_t* __restrict__ table, int64_t* __restrict__ carry_row,
    TA* __restrict__ carry_val) {
  __shared__ int64_t lds[kWavesPerBlock][kWinCap];
  const int w = wave_in_block();
  const int64_t tile = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + w;
  if (tile >= ntiles) return;
  spmv_tile<TP, TI, TV, TA, V>(ptr, idx, val, base, m, nnz, x, y, tile, vec != 0, table, lds[w],
                               carry_row + tile, carry_val + tile);
}

// ---- CSC SpMV ---------------------------------------------------------------
template <typename TA>
__global__ void fill_kernel(TA* __restrict__ p, int64_t n, TA v) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

template <typename TA>
__device__ __forceinline__ void atomic_add(TA* p, TA v) {
  unsafeAtomicAdd(p, v);
}

template <typename TP, typename TI, typename TV, typename TA, int V>
__global__ __launch_bounds__(256) void spmv_csc_kernel(
    const TP* __restrict__ ptr, const TI* __restrict__ idx, const TV* __restrict__ val,
    int64_t base, int64_t n_outer, int64_t nnz, const TA* __restrict__ x, TA* __restrict__ y,
    int64_t ntiles, int vec) {
  __shared__ int64_t lds[kWavesPerBlock][kWinCap];
  const int w = wave_in_block();
  const int64_t tile = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + w;
  if (tile >= ntiles) return;
  constexpr int W = kWave * V;
  const int lane = lane_id();
  const int64_t q0 = tile * W;
  const int64_t qend = (q0 + W < nnz) ? q0 + W : nnz;
  int64_t c_lo, c_hi;
  tile_fibers(ptr, base, n_outer, q0, qend, &c_lo, &c_hi);
  const int64_t c_last = c_hi < n_outer ? c_hi : n_outer - 1;
  FiberWindow<TP> win = load_window(ptr, base, c_lo, c_last - c_lo + 1, lds[w], kWinCap);
  const int64_t a = q0 + static_cast<int64_t>(lane) * V;
  const int64_t b = (a + V < qend) ? a + V : qend;
  if (a >= qend) return;
  TV vv[V];
  TI jj[V];
  if (b - a == V) {
    ld_run<TV, V>(val + a, vec != 0, vv);
    ld_run<TI, V>(idx + a, vec != 0, jj);
  } else {
#pragma unroll
    for (int s = 0; s < V; s++) {
      const bool in = a + s < b;
      vv[s] = in ? ld_stream(val + a + s) : TV{};
      jj[s] = in ? ld_stream(idx + a + s) : TI{};
    }
  }
  int64_t c = win.find(a);
  TA xc = x[c];
#pragma unroll
  for (int s = 0; s < V; s++) {
    const int64_t q = a + s;
    if (q < b) {
      while (c < c_last && win.end_of(c) <= q) { c++; xc = x[c]; }
      atomic_add<TA>(y + (static_cast<int64_t>(jj[s]) - base), widen<TA>(vv[s]) * xc);
    }
  }
}

// ---------------------------------------------------------------------------
// Tiles of one CSR SpMV call: W positions per wave tile, ntiles tiles.
template <typename TV>
static void spmv_csr_shape(const fgx_compressed* A, int64_t* W, int64_t* ntiles) {
  *W = kWave * spmv_v<TV>::value;
  *ntiles = A->nnz > 0 ? (A->nnz + *W - 1) / *W : 1;
}

// The kernel and its fix-up over caller-provided buffers: the tile table
// (ntiles + 1 entries, already filled) and one carry slot per tile.
template <typename TP, typename TI, typename TV>
static int spmv_csr_launch(const fgx_compressed* A, const void* x, void* y, hipStream_t s,
                           const int64_t* table, int64_t* crow,
                           typename acc_of<TV>::type* cval) {
  using TA = typename acc_of<TV>::type;
  constexpr int V = spmv_v<TV>::value;
  int64_t W, ntiles;
  spmv_csr_shape<TV>(A, &W, &ntiles);
  const int vec = aligned16(A->val) && aligned16(A->idx);
  const int64_t blocks = (ntiles + kWavesPerBlock - 1) / kWavesPerBlock;
  hipLaunchKernelGGL((spmv_csr_kernel<TP, TI, TV, TA, V>), dim3(static_cast<unsigned>(blocks)),
                     dim3(256), 0, s, static_cast<const TP*>(A->ptr),
                     static_cast<const TI*>(A->idx), static_cast<const TV*>(A->val),
                     static_cast<int64_t>(A->index_base), A->n_outer, A->nnz,
                     static_cast<const TA*>(x), static_cast<TA*>(y), ntiles, vec, table, crow,
                     cval);
  FGX_TRY(launch_check("spmv_csr_kernel"));
  if (ntiles > 1) FGX_TRY(launch_fixup<TA>(s, crow, cval, ntiles, 1, static_cast<TA*>(y), 1, 0, 1));
  return FGX_OK;
}

template <typename TP, typename TI, typename TV>
static int spmv_csr_typed(fgx_ctx* ctx, const fgx_compressed* A, const void* x, void* y,
                          hipStream_t s) {
  using TA = typename acc_of<TV>::type;
  if (A->n_outer == 0) return FGX_OK;
  int64_t W, ntiles;
  spmv_csr_shape<TV>(A, &W, &ntiles);
  void* scr;
  FGX_TRY(scratch(ctx, s, static_cast<size_t>(ntiles) * (sizeof(int64_t) + sizeof(TA)), &scr));
  int64_t* crow = static_cast<int64_t*>(scr);
  TA* cval = reinterpret_cast<TA*>(crow + ntiles);
  const int64_t* table = nullptr;
  FGX_TRY(tile_table(ctx, A, W, s, &table));
  return spmv_csr_launch<TP, TI, TV>(A, x, y, s, table, crow, cval);
}

// ---- CSR SpMV that allocates nothing ----------------------------------------------
// Device memory for the tile table and the carries of fgx_spmv_csr_noalloc:
// one array per GPU, held until the process ends.  Host code never uses this
// name as an address; it uses the device address that noalloc_setup obtains
// with hipGetSymbolAddress (kept per context, i.e. per device).
static __device__ __align__(256) unsigned char noalloc_scratch[FGX_NOALLOC_SCRATCH_BYTES];

static size_t round256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// Bytes of noalloc_scratch one call needs: table, carry rows, carry values.
template <typename TV>
static size_t noalloc_bytes(const fgx_compressed* A) {
  using TA = typename acc_of<TV>::type;
  int64_t W, ntiles;
  spmv_csr_shape<TV>(A, &W, &ntiles);
  const size_t nt = static_cast<size_t>(ntiles);
  return round256((nt + 1) * sizeof(int64_t)) + round256(nt * sizeof(int64_t)) +
         round256(nt * sizeof(TA));
}

int noalloc_setup(fgx_ctx* ctx) {
  void* p = nullptr;
  FGX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)));
  ctx->noalloc_scratch = p;
  return FGX_OK;
}

template <typename TP, typename TI, typename TV>
static int spmv_csr_noalloc_typed(fgx_ctx* ctx, const fgx_compressed* A, const void* x, void* y,
                                  hipStream_t s) {
  using TA = typename acc_of<TV>::type;
  if (A->n_outer == 0) return FGX_OK;
  const size_t need = noalloc_bytes<TV>(A);
  if (need > FGX_NOALLOC_SCRATCH_BYTES - 256)
    return set_error(FGX_EINVAL,
                     "fgx_spmv_csr_noalloc: %lld stored entries need %zu bytes of scratch; the "
                     "limit is %zu (use fgx_spmv_csr)",
                     static_cast<long long>(A->nnz), need,
                     static_cast<size_t>(FGX_NOALLOC_SCRATCH_BYTES - 256));
  int64_t W, ntiles;
  spmv_csr_shape<TV>(A, &W, &ntiles);
  char* base = static_cast<char*>(ctx->noalloc_scratch);
  int64_t* table = reinterpret_cast<int64_t*>(base);
  base += round256((static_cast<size_t>(ntiles) + 1) * sizeof(int64_t));
  int64_t* crow = reinterpret_cast<int64_t*>(base);
  base += round256(static_cast<size_t>(ntiles) * sizeof(int64_t));
  TA* cval = reinterpret_cast<TA*>(base);
  FGX_TRY(fill_tile_table(A, W, s, table));
  return spmv_csr_launch<TP, TI, TV>(A, x, y, s, table, crow, cval);
}

template <typename TP, typename TI, typename TV>
static int spmv_csc_typed(fgx_ctx* ctx, const fgx_compressed* A, const void* x, void* y,
                          hipStream_t s) {
  (void)ctx;
  using TA = typename acc_of<TV>::type;
  constexpr int V = spmv_v<TV>::value;
  const int64_t W = kWave * V;
  const int64_t ni = A->n_inner;
  if (ni > 0) {
    const int64_t blocks = ((ni + 255) / 256) < 4096 ? (ni + 255) / 256 : 4096;
    hipLaunchKernelGGL((fill_kernel<TA>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                       static_cast<TA*>(y), ni, TA(0));
    FGX_TRY(launch_check("fill_kernel"));
  }
  if (A->nnz == 0 || A->n_outer == 0) return FGX_OK;
  const int64_t ntiles = (A->nnz + W - 1) / W;
  const int vec = aligned16(A->val) && aligned16(A->idx);
  const int64_t blocks = (ntiles + kWavesPerBlock - 1) / kWavesPerBlock;
  hipLaunchKernelGGL((spmv_csc_kernel<TP, TI, TV, TA, V>), dim3(static_cast<unsigned>(blocks)),
                     dim3(256), 0, s, static_cast<const TP*>(A->ptr),
                     static_cast<const TI*>(A->idx), static_cast<const TV*>(A->val),
                     static_cast<int64_t>(A->index_base), A->n_outer, A->nnz,
                     static_cast<const TA*>(x), static_cast<TA*>(y), ntiles, vec);
  return launch_check("spmv_csc_kernel");
}

int spmv_csr_into(fgx_ctx* ctx, const fgx_compressed* A, const void* x, void* y, hipStream_t s) {
  IdxCombo ic;
  FGX_TRY(idx_combo(A, &ic));
  int rc = FGX_OK;
  FGX_DISPATCH_VAL(A->val_type, FGX_DISPATCH_IDX(ic, rc = (spmv_csr_typed<TP, TI, TV>(ctx, A, x, y, s))));
  return rc;
}

static size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

}  // namespace fgx

using namespace fgx;

extern "C" {

int fgx_spmv_csr(fgx_ctx* ctx, const fgx_compressed* A, const void* x, void* y, void* stream) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_spmv_csr: NULL ctx");
  FGX_TRY(check_compressed(A, "fgx_spmv_csr"));
  if ((A->n_inner > 0 && !x && A->nnz > 0) || (A->n_outer > 0 && !y))
    return set_error(FGX_EINVAL, "fgx_spmv_csr: NULL x/y");
  IdxCombo ic;
  FGX_TRY(idx_combo(A, &ic));
  std::lock_guard<std::mutex> lk(ctx->mu);
  FGX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = pick_stream(ctx, stream);
  FGX_TRY(admit_operand(ctx, A, s, "fgx_spmv_csr"));
  int rc = FGX_OK;
  FGX_DISPATCH_VAL(A->val_type, FGX_DISPATCH_IDX(ic, rc = (spmv_csr_typed<TP, TI, TV>(ctx, A, x, y, s))));
  return rc;
}

int fgx_spmv_csr_noalloc(fgx_ctx* ctx, const fgx_compressed* A, const void* x, void* y,
                         void* stream) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_spmv_csr_noalloc: NULL ctx");
  FGX_TRY(check_compressed(A, "fgx_spmv_csr_noalloc"));
  if ((A->n_inner > 0 && !x && A->nnz > 0) || (A->n_outer > 0 && !y))
    return set_error(FGX_EINVAL, "fgx_spmv_csr_noalloc: NULL x/y");
  if (!ctx->noalloc_scratch)
    return set_error(FGX_EHIP, "fgx_spmv_csr_noalloc: scratch address not set up");
  IdxCombo ic;
  FGX_TRY(idx_combo(A, &ic));
  std::lock_guard<std::mutex> lk(ctx->mu);
  FGX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = pick_stream(ctx, stream);
  if (ctx->safe) {
    // The flag word sits past the largest region a call may use.
    uint32_t* flags = reinterpret_cast<uint32_t*>(static_cast<char*>(ctx->noalloc_scratch) +
                                                  FGX_NOALLOC_SCRATCH_BYTES - 256);
    FGX_TRY(validate_operand(ctx, A, s, "fgx_spmv_csr_noalloc", flags));
  }
  int rc = FGX_OK;
  FGX_DISPATCH_VAL(A->val_type,
                   FGX_DISPATCH_IDX(ic, rc = (spmv_csr_noalloc_typed<TP, TI, TV>(ctx, A, x, y, s))));
  return rc;
}

int fgx_spmv_csc_ordered(fgx_ctx* ctx, const fgx_compressed* A, const void* x, void* y,
                         void* stream) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_spmv_csc_ordered: NULL ctx");
  FGX_TRY(check_compressed(A, "fgx_spmv_csc_ordered"));
  if ((A->n_outer > 0 && !x && A->nnz > 0) || (A->n_inner > 0 && !y))
    return set_error(FGX_EINVAL, "fgx_spmv_csc_ordered: NULL x/y");
  IdxCombo ic;
  FGX_TRY(idx_combo(A, &ic));
  std::lock_guard<std::mutex> lk(ctx->mu);
  FGX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = pick_stream(ctx, stream);
  FGX_TRY(admit_operand(ctx, A, s, "fgx_spmv_csc_ordered"));
  if (A->n_inner == 0) return FGX_OK;
  // A's rows as fibres, in this stream's dense arena: ptr, idx, val.
  const size_t pb = align256(static_cast<size_t>(A->n_inner + 1) * elem_size(A->ptr_type));
  const size_t ib = align256(static_cast<size_t>(A->nnz) * elem_size(A->idx_type));
  const size_t vb = align256(static_cast<size_t>(A->nnz) * elem_size(A->val_type));
  void* arena;
  FGX_TRY(dense_scratch(ctx, s, pb + ib + vb, &arena));
  char* base = static_cast<char*>(arena);
  fgx_compressed T = *A;
  T.n_outer = A->n_inner;
  T.n_inner = A->n_outer;
  T.ptr = base;
  T.idx = base + pb;
  T.val = base + pb + ib;
  FGX_TRY(transpose_into(ctx, A, const_cast<void*>(T.ptr), const_cast<void*>(T.idx),
                         const_cast<void*>(T.val), s));
  return spmv_csr_into(ctx, &T, x, y, s);
}

int fgx_spmv_csc(fgx_ctx* ctx, const fgx_compressed* A, const void* x, void* y, void* stream) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_spmv_csc: NULL ctx");
  FGX_TRY(check_compressed(A, "fgx_spmv_csc"));
  if ((A->n_outer > 0 && !x && A->nnz > 0) || (A->n_inner > 0 && !y))
    return set_error(FGX_EINVAL, "fgx_spmv_csc: NULL x/y");
  IdxCombo ic;
  FGX_TRY(idx_combo(A, &ic));
  std::lock_guard<std::mutex> lk(ctx->mu);
  FGX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = pick_stream(ctx, stream);
  FGX_TRY(admit_operand(ctx, A, s, "fgx_spmv_csc"));
  int rc = FGX_OK;
  FGX_DISPATCH_VAL(A->val_type, FGX_DISPATCH_IDX(ic, rc = (spmv_csc_typed<TP, TI, TV>(ctx, A, x, y, s))));
  return rc;
}

}  // extern "C"
