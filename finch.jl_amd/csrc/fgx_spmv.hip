// fgx_spmv.hip — SpMV entry points (CSR and CSC).
//
// CSR SpMV / SpMM use the nnz-balanced wave tiles of fgx_csr_tiles.hpp plus a
// fix-up launch for rows that span tiles.  CSC SpMV (the form Finch
// parallelises with AtomicElement, benchmark/benchmarks.jl:402-413) scatters
// into y through row buckets for large operands and with hardware float
// atomics over the same nnz-balanced tiles for small ones.
#include "fgx_launch.hpp"
#include "fgx_scan.hpp"

#include <cstdlib>

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

// ---- CSC SpMV by row buckets --------------------------------------------------
// Large CSC operands scatter through row buckets instead of global atomics:
// a device float atomic add on MI355X retires at ~21 G/s whatever its scope
// or target copy, so 16M scattered adds cost ~0.77 ms, while the same
// products written once to per-bucket runs and summed in LDS cost a few
// passes over 8-16 bytes per stored entry.
//   1. count: per block of 256 * csc_run positions, how many rows fall in
//      each bucket of 2^RB rows (LDS histogram);
//   2. scan: bucket-major exclusive scan of the counts, so each bucket's
//      entries are one contiguous run and each block owns a slice of it;
//   3. scatter: the same blocks walk their columns (the wave-tile walk of
//      the atomic kernel), form each product A[i,j]*x[j], sort it into the
//      block's bucket order in LDS and write the block's slices out as
//      contiguous runs;
//   4. accumulate: S blocks per bucket add a share of its run into an LDS
//      copy of the bucket's rows (LDS float atomics); with S = 1 they write
//      those rows of y, otherwise S partial rows that a last pass sums in
//      order.
// Like the atomic kernel (Finch's AtomicElement form), the order of the adds
// into each y[i] is not fixed.
constexpr int64_t kCscMaxBuckets = 1024;   // scatter LDS: a cursor and a base per bucket
constexpr int kCscWinCap = 256;            // scatter LDS: fibre window entries per wave
constexpr int64_t kCscMinNnz = int64_t(1) << 18;  // below: atomic kernel

template <typename TA>
struct CscEntry {
  uint32_t r;  // row within the bucket (in the scatter's LDS stage: the row)
  TA v;        // A[i,j] * x[j]
};
// Rows per bucket (the accumulate block's LDS copy is 32 KB) and positions
// per lane (the scatter's LDS stage is 32 KB).
template <typename TA>
constexpr int csc_row_bits() { return sizeof(TA) == 8 ? 12 : 13; }
template <typename TA>
constexpr int csc_run() { return sizeof(TA) == 8 ? 8 : 16; }

// Rows outside [0, m) are skipped by both the count and the scatter, so the
// slices stay consistent and no LDS slot outside the buckets is touched.
template <typename TI, int RB, int RUN>
__global__ __launch_bounds__(256) void csc_bucket_count_kernel(const TI* __restrict__ idx,
                                                               int64_t base, int64_t nnz,
                                                               int64_t m, int64_t nb, int64_t nblk,
                                                               int32_t* __restrict__ counts) {
  __shared__ int32_t h[kCscMaxBuckets];
  for (int64_t b = threadIdx.x; b < nb; b += blockDim.x) h[b] = 0;
  __syncthreads();
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * 256 * RUN;
#pragma unroll 4
  for (int s = 0; s < RUN; s++) {
    const int64_t q = q0 + s * 256 + threadIdx.x;
    if (q < nnz) {
      const int64_t row = static_cast<int64_t>(ld_stream(idx + q)) - base;
      if (static_cast<uint64_t>(row) < static_cast<uint64_t>(m)) atomicAdd(&h[row >> RB], 1);
    }
  }
  __syncthreads();
  for (int64_t b = threadIdx.x; b < nb; b += blockDim.x) counts[b * nblk + blockIdx.x] = h[b];
}

template <typename TP, typename TI, typename TV, typename TA, int RB, int RUN>
__global__ __launch_bounds__(256) void csc_bucket_scatter_kernel(
    const TP* __restrict__ ptr, const TI* __restrict__ idx, const TV* __restrict__ val,
    int64_t base, int64_t n_outer, int64_t nnz, const TA* __restrict__ x, int64_t m, int64_t nb,
    int64_t nblk, const int32_t* __restrict__ counts, const int64_t* __restrict__ offsets,
    const int64_t* __restrict__ table, int vec, CscEntry<TA>* __restrict__ out) {
  constexpr int BP = 256 * RUN;
  constexpr int64_t W = int64_t(kWave) * RUN;
  constexpr int PER = static_cast<int>(kCscMaxBuckets / 256);
  __shared__ CscEntry<TA> stage[BP];
  __shared__ int32_t cur[kCscMaxBuckets];      // next local slot of each bucket
  __shared__ int64_t gbase[kCscMaxBuckets];    // global slot of local slot 0, per bucket
  __shared__ int64_t lds[kWavesPerBlock][kCscWinCap];
  __shared__ int32_t wsum[kWavesPerBlock + 1];
  const int tid = threadIdx.x;
  const int w = wave_in_block();
  const int lane = lane_id();
  // Loads that depend only on the block index go out first: this lane's run
  // of coordinates and values, then the block's bucket counts and slice
  // starts; the fibre window is located while they are in flight.
  const int64_t tile = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + w;
  const int64_t q0 = tile * W;
  const int64_t qend = (q0 + W < nnz) ? q0 + W : nnz;
  const int64_t a = q0 + static_cast<int64_t>(lane) * RUN;
  const int64_t b = (a + RUN < qend) ? a + RUN : qend;
  TV vv[RUN];
  TI jj[RUN];
  if (b - a == RUN) {
    ld_run<TV, RUN>(val + a, vec != 0, vv);
    ld_run<TI, RUN>(idx + a, vec != 0, jj);
  } else {
#pragma unroll
    for (int s = 0; s < RUN; s++) {
      const bool in = a + s < b;
      vv[s] = in ? ld_stream(val + a + s) : TV{};
      jj[s] = in ? ld_stream(idx + a + s) : static_cast<TI>(base);
    }
  }
  int cnt[PER];
  int64_t off[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int64_t bk = static_cast<int64_t>(tid) * PER + k;
    cnt[k] = bk < nb ? counts[bk * nblk + blockIdx.x] : 0;
    off[k] = bk < nb ? offsets[bk * nblk + blockIdx.x] : 0;
  }
  FiberWindow<TP> win{};
  int64_t c_lo = 0, c_hi = 0, pos_lo = 0;
  if (q0 < nnz) win = tile_setup(ptr, base, n_outer, q0, qend, table, tile, lds[w], &c_lo, &c_hi,
                                 &pos_lo, kCscWinCap);
  // The block's bucket counts, scanned into local slots (bucket order).
  int sum = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) sum += cnt[k];
  int incl = sum;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int t = shfl_up(incl, d);
    if (lane >= d) incl += t;
  }
  if (lane == kWave - 1) wsum[w] = incl;
  __syncthreads();
  int ex = incl - sum;
  for (int v = 0; v < w; v++) ex += wsum[v];
  if (tid == 255) wsum[kWavesPerBlock] = ex + sum;  // the block's entry count
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int64_t bk = static_cast<int64_t>(tid) * PER + k;
    if (bk < nb) {
      cur[bk] = ex;
      gbase[bk] = off[k] - ex;
    }
    ex += cnt[k];
  }
  __syncthreads();
  // Products, staged in bucket order.
  if (a < qend) {
    // Each position's column from the window (LDS only), then every x
    // gather issued before the first is used.
    const int64_t c_last = c_hi < n_outer ? c_hi : n_outer - 1;
    int64_t c = win.find(a);
    int64_t col[RUN];
#pragma unroll
    for (int s = 0; s < RUN; s++) {
      while (c < c_last && win.end_of(c) <= a + s) c++;
      col[s] = c;
    }
    TA xs[RUN];
#pragma unroll
    for (int s = 0; s < RUN; s++) xs[s] = x[col[s]];
#pragma unroll
    for (int s = 0; s < RUN; s++) {
      const int64_t row = static_cast<int64_t>(jj[s]) - base;
      if (a + s < b && static_cast<uint64_t>(row) < static_cast<uint64_t>(m)) {
        const int at = atomicAdd(&cur[row >> RB], 1);
        stage[at].r = static_cast<uint32_t>(row);
        stage[at].v = widen<TA>(vv[s]) * xs[s];
      }
    }
  }
  __syncthreads();
  // The block's slices, each a contiguous run of its bucket.
  const int total = wsum[kWavesPerBlock];
  constexpr uint32_t kMask = (1u << RB) - 1;
  for (int i = tid; i < total; i += 256) {
    const CscEntry<TA> e = stage[i];
    CscEntry<TA> o;
    o.r = e.r & kMask;
    o.v = e.v;
    out[gbase[e.r >> RB] + i] = o;
  }
}

template <typename TA, int RB>
__global__ __launch_bounds__(1024) void csc_bucket_accum_kernel(
    const CscEntry<TA>* __restrict__ in, const int64_t* __restrict__ offsets, int64_t nblk,
    int64_t m, int split, TA* __restrict__ y, TA* __restrict__ partial) {
  __shared__ TA acc[1 << RB];
  for (int i = threadIdx.x; i < (1 << RB); i += blockDim.x) acc[i] = TA(0);
  __syncthreads();
  const int64_t b = blockIdx.x / split;
  const int64_t sp = blockIdx.x % split;
  const int64_t e0 = offsets[b * nblk], e1 = offsets[(b + 1) * nblk];
  const int64_t lo = e0 + (e1 - e0) * sp / split, hi = e0 + (e1 - e0) * (sp + 1) / split;
  int64_t q = lo + threadIdx.x;
  constexpr int U = 8;
  for (; q + (U - 1) * 1024 < hi; q += U * 1024) {
    CscEntry<TA> t[U];
#pragma unroll
    for (int u = 0; u < U; u++) t[u] = in[q + u * 1024];
#pragma unroll
    for (int u = 0; u < U; u++) atomicAdd(&acc[t[u].r], t[u].v);
  }
  for (; q < hi; q += 1024) {
    const CscEntry<TA> t = in[q];
    atomicAdd(&acc[t.r], t.v);
  }
  __syncthreads();
  TA* dst = split == 1 ? y : partial + sp * m;
  const int64_t r0 = b << RB;
  for (int i = threadIdx.x; i < (1 << RB) && r0 + i < m; i += blockDim.x) dst[r0 + i] = acc[i];
}

// y[i] = sum of the split partial rows, in split order.
template <typename TA>
__global__ __launch_bounds__(256) void csc_partial_sum_kernel(const TA* __restrict__ partial,
                                                              int split, int64_t m,
                                                              TA* __restrict__ y) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += stride) {
    TA acc = partial[i];
    for (int s = 1; s < split; s++) acc += partial[s * m + i];
    y[i] = acc;
  }
}

// Whether a CSC SpMV takes the bucket path: large enough to amortise its
// launches, and few enough buckets for the scatter's LDS slots.  The
// environment variable FGX_CSC_BUCKETS forces it off (0) or on (1, when the
// bucket count allows) for measurement and tests.
template <typename TA>
static bool csc_use_buckets(const fgx_compressed* A) {
  const int64_t nb = (A->n_inner + (int64_t(1) << csc_row_bits<TA>()) - 1) >> csc_row_bits<TA>();
  if (nb > kCscMaxBuckets || A->nnz == 0 || A->n_outer == 0 || A->n_inner == 0) return false;
  const char* env = getenv("FGX_CSC_BUCKETS");
  if (env && env[0] == '0') return false;
  if (env && env[0] == '1') return true;
  return A->nnz >= kCscMinNnz;
}

template <typename TP, typename TI, typename TV>
static int spmv_csc_buckets(fgx_ctx* ctx, const fgx_compressed* A, const void* x, void* y,
                            hipStream_t s) {
  using TA = typename acc_of<TV>::type;
  constexpr int RB = csc_row_bits<TA>();
  constexpr int RUN = csc_run<TA>();
  constexpr int64_t BP = 256 * RUN;
  const int64_t m = A->n_inner;
  const int64_t nb = (m + (int64_t(1) << RB) - 1) >> RB;
  const int64_t nblk = (A->nnz + BP - 1) / BP;
  const int64_t ncnt = nb * nblk;
  // Enough accumulate blocks for two per CU.
  const int64_t cus = ctx->num_cus > 0 ? ctx->num_cus : 256;
  const int64_t want = (2 * cus + nb - 1) / nb;
  const int split = static_cast<int>(want < 1 ? 1 : (want > 16 ? 16 : want));
  auto up = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
  const size_t cb = up(static_cast<size_t>(ncnt) * sizeof(int32_t));
  const size_t ob = up(static_cast<size_t>(ncnt + 1) * sizeof(int64_t));
  const size_t wb = up(scan_work_bytes(ncnt));
  const size_t pb = split > 1 ? up(static_cast<size_t>(split) * m * sizeof(TA)) : 0;
  const size_t eb = static_cast<size_t>(A->nnz) * sizeof(CscEntry<TA>);
  void* arena;
  FGX_TRY(dense_scratch(ctx, s, cb + ob + wb + pb + eb, &arena));
  char* p = static_cast<char*>(arena);
  int32_t* counts = reinterpret_cast<int32_t*>(p);
  int64_t* offsets = reinterpret_cast<int64_t*>(p + cb);
  void* work = p + cb + ob;
  TA* partial = reinterpret_cast<TA*>(p + cb + ob + wb);
  CscEntry<TA>* entries = reinterpret_cast<CscEntry<TA>*>(p + cb + ob + wb + pb);
  const int64_t ib = static_cast<int64_t>(A->index_base);
  hipLaunchKernelGGL((csc_bucket_count_kernel<TI, RB, RUN>), dim3(static_cast<unsigned>(nblk)),
                     dim3(256), 0, s, static_cast<const TI*>(A->idx), ib, A->nnz, m, nb, nblk,
                     counts);
  FGX_TRY(launch_check("csc_bucket_count_kernel"));
  FGX_TRY((scan_counts_to_ptr<int64_t, int32_t>(counts, ncnt, 0, offsets, work, nullptr, nullptr,
                                                s)));
  const int vec = aligned16(A->val) && aligned16(A->idx);
  const int64_t* table = nullptr;
  FGX_TRY(tile_table(ctx, A, int64_t(kWave) * RUN, s, &table));
  hipLaunchKernelGGL((csc_bucket_scatter_kernel<TP, TI, TV, TA, RB, RUN>),
                     dim3(static_cast<unsigned>(nblk)), dim3(256), 0, s,
                     static_cast<const TP*>(A->ptr), static_cast<const TI*>(A->idx),
                     static_cast<const TV*>(A->val), ib, A->n_outer, A->nnz,
                     static_cast<const TA*>(x), m, nb, nblk, counts, offsets, table, vec,
                     entries);
  FGX_TRY(launch_check("csc_bucket_scatter_kernel"));
  hipLaunchKernelGGL((csc_bucket_accum_kernel<TA, RB>), dim3(static_cast<unsigned>(nb * split)),
                     dim3(1024), 0, s, entries, offsets, nblk, m, split, static_cast<TA*>(y),
                     partial);
  FGX_TRY(launch_check("csc_bucket_accum_kernel"));
  if (split == 1) return FGX_OK;
  const int64_t blocks = ((m + 255) / 256) < 2048 ? (m + 255) / 256 : 2048;
  hipLaunchKernelGGL((csc_partial_sum_kernel<TA>), dim3(static_cast<unsigned>(blocks)), dim3(256),
                     0, s, partial, split, m, static_cast<TA*>(y));
  return launch_check("csc_partial_sum_kernel");
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
  using TA = typename acc_of<TV>::type;
  constexpr int V = spmv_v<TV>::value;
  const int64_t W = kWave * V;
  const int64_t ni = A->n_inner;
  // Every row of y is written by the bucket path's accumulate step.
  if (csc_use_buckets<TA>(A)) return spmv_csc_buckets<TP, TI, TV>(ctx, A, x, y, s);
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
