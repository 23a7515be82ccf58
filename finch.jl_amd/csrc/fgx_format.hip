// fgx_format.hip — format conversion on the device.
//
//   * fgx_transpose: a compressed matrix stored along its other dimension
//     (CSC <-> CSR), every stored entry kept, each output fibre sorted;
//   * fgx_coo_to_compressed: coordinate lists -> Dense(SparseList), ordered
//     by (outer, inner), duplicates combined with + in input order;
//   * fgx_dropfills: the stored entries whose value differs from the fill.
//
// All three are deterministic.  The ordering step is a stable LSD radix sort
// written for wave64: 8-bit digits; per pass, per-tile digit counts, one
// exclusive scan of the (digit, tile) table, then a stable scatter whose
// ranks inside a tile come from one ballot per digit bit.  No atomics decide
// where an element goes, so a run reproduces bit for bit.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "fgx_device.hpp"
#include "fgx_internal.hpp"
#include "fgx_scan.hpp"

namespace fgx {
namespace {

constexpr int kSortThreads = 512;
constexpr int kSortItems = 8;
constexpr int64_t kSortTile = static_cast<int64_t>(kSortThreads) * kSortItems;
constexpr int kDigitBits = 8;
constexpr int kBins = 1 << kDigitBits;
constexpr int kSortWaves = kSortThreads / kWave;
static_assert(kSortThreads >= kBins, "the digit tables give one digit to each of the first threads");

// Raw value words: values are moved, never interpreted, by the sort.
template <int B> struct word_of;
template <> struct word_of<2> { using type = uint16_t; };
template <> struct word_of<4> { using type = uint32_t; };
template <> struct word_of<8> { using type = uint64_t; };

template <typename T>
__device__ __forceinline__ uint32_t rebased(const T* __restrict__ a, int64_t base, int64_t e) {
  return static_cast<uint32_t>(static_cast<int64_t>(a[e]) - base);
}

// Each wave of a sort block owns a contiguous sub-tile of kWaveSpan
// elements: element e of the tile sits in wave e / kWaveSpan, round
// (e % kWaveSpan) / 64, lane e % 64.
constexpr int kWaveSpan = kSortItems * kWave;

// Lanes of the wave whose digit (dbits wide) equals this lane's (among lanes
// with `in`).
__device__ __forceinline__ uint64_t same_digit(uint32_t d, bool in, int dbits) {
  uint64_t same = ballot(in);
  for (int b = 0; b < dbits; b++) {
    const bool bit = ((d >> b) & 1u) != 0;
    const uint64_t m = ballot(bit);
    same &= bit ? m : ~m;
  }
  return same;
}

// Digit counts of one tile of keys: hist[d * ntiles + tile], for digits of
// dbits (<= kDigitBits) bits starting at bit `shift`.
template <typename TK>
__global__ __launch_bounds__(kSortThreads) void sort_hist_kernel(const TK* __restrict__ keys,
                                                                 int64_t kbase, int64_t n,
                                                                 int shift, int dbits,
                                                                 int64_t ntiles,
                                                                 int64_t* __restrict__ hist) {
  __shared__ uint32_t h[kBins];
  const int tid = threadIdx.x;
  const int nbins = 1 << dbits;
  const uint32_t dmask = static_cast<uint32_t>(nbins - 1);
  if (tid < kBins) h[tid] = 0;
  const int64_t t0 = static_cast<int64_t>(blockIdx.x) * kSortTile;
  uint32_t k[kSortItems];
#pragma unroll
  for (int i = 0; i < kSortItems; i++) {
    const int64_t e = t0 + static_cast<int64_t>(i) * kSortThreads + tid;
    k[i] = e < n ? (rebased(keys, kbase, e) >> shift) & dmask : kBins;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kSortItems; i++)
    if (k[i] < kBins) atomicAdd(&h[k[i]], 1u);
  __syncthreads();
  if (tid < nbins) hist[static_cast<int64_t>(tid) * ntiles + blockIdx.x] = h[tid];
}

// Tile processed by this block.  Blocks go round-robin over the 8 XCDs;
// numbering the tiles so that each XCD takes a contiguous range keeps the
// boundary cache lines shared by consecutive tiles' output runs in one L2.
// A bijection for any grid size.
__device__ __forceinline__ int64_t xcd_tile() {
  constexpr int kXcd = 8;
  const int64_t nb = gridDim.x, b = blockIdx.x;
  const int64_t q = nb / kXcd, r = nb % kXcd, x = b % kXcd;
  return x * q + (x < r ? x : r) + b / kXcd;
}

// Exclusive scan of one value per thread over the block; returns the prefix.
__device__ __forceinline__ uint32_t block_prefix(uint32_t v, uint32_t* wsum) {
  const int lane = lane_id();
  const int w = wave_in_block();
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t y = shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane == kWave - 1) wsum[w] = x;
  __syncthreads();
  uint32_t before = 0;
  for (int v2 = 0; v2 < w; v2++) before += wsum[v2];
  return before + x - v;
}

// Stable scatter of one tile, in three steps.
//   1. Wave ranks: each wave walks its sub-tile in element order, 64
//      elements a round; the lanes sharing an element's digit are found with
//      dbits ballots, and the wave keeps its own running digit counts in
//      LDS (read by every lane, then advanced by each group's lowest lane: one
//      instruction stream, so no barrier is needed).
//   2. Tile ranks: the waves' counts are combined per digit (one barrier) into
//      each wave's first slot for each digit, after the tile's earlier digits;
//      every element is parked in LDS at its rank in the tile sorted stably by
//      digit.
//   3. Writes: thread t takes the t-th element of the sorted tile to
//      offs[d * ntiles + tile] + (t - start[d]), so consecutive threads write
//      consecutive addresses of one digit's run (coalesced stores).
// The sub-tile is loaded before the first round, so its loads are in flight
// together.
template <typename TK, typename TQ, typename TW, typename TQO>
__global__ __launch_bounds__(kSortThreads) void sort_scatter_kernel(
    const TK* __restrict__ keys, int64_t kbase, const TQ* __restrict__ pay, int64_t pbase,
    const TW* __restrict__ val, int64_t n, int shift, int dbits, int64_t ntiles,
    const int64_t* __restrict__ offs, uint32_t* __restrict__ keys_out, TQO* __restrict__ pay_out,
    int64_t pout_base, TW* __restrict__ val_out) {
  // Small tables (int32 / uint16 so that three blocks fit a CU's LDS):
  // positions are below 2^31 and tile counts below 2^16.
  __shared__ int32_t gbase[kBins];
  __shared__ uint16_t wcnt[kSortWaves][kBins];
  __shared__ uint32_t wsum[kSortWaves];
  __shared__ uint32_t sk[kSortTile];
  __shared__ uint32_t sq[kSortTile];
  __shared__ TW sv[kSortTile];
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int w = wave_in_block();
  const int nbins = 1 << dbits;
  const uint32_t dmask = static_cast<uint32_t>(nbins - 1);
  const int64_t tile = xcd_tile();
  const int64_t t0 = tile * kSortTile;
  const int64_t e0 = t0 + w * kWaveSpan + lane;
  uint32_t k[kSortItems], p[kSortItems];
  TW x[kSortItems];
#pragma unroll
  for (int i = 0; i < kSortItems; i++) {
    const int64_t e = e0 + i * kWave;
    if (e < n) {
      k[i] = rebased(keys, kbase, e);
      p[i] = rebased(pay, pbase, e);
      x[i] = val[e];
    } else {
      k[i] = 0;
      p[i] = 0;
      x[i] = TW(0);
    }
  }
  if (tid < kBins) {
#pragma unroll
    for (int v = 0; v < kSortWaves; v++) wcnt[v][tid] = 0;
  }
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
  __syncthreads();
  // 1. Ranks within the wave's sub-tile.
  uint16_t* cnt = wcnt[w];
  uint32_t r[kSortItems];
#pragma unroll
  for (int i = 0; i < kSortItems; i++) {
    const bool in = e0 + i * kWave < n;
    const uint32_t d = (k[i] >> shift) & dmask;
    const uint64_t same = same_digit(d, in, dbits);
    const uint32_t rank = static_cast<uint32_t>(popc(same & below));
    const uint32_t c = cnt[d];
    r[i] = c + rank;
    __builtin_amdgcn_wave_barrier();
    if (in && rank == 0) cnt[d] = static_cast<uint16_t>(c + static_cast<uint32_t>(popc(same)));
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // 2. Tile ranks: digit tid's start in the tile, then each wave's first slot.
  {
    const bool owner = tid < nbins;
    const int64_t cell = static_cast<int64_t>(owner ? tid : 0) * ntiles + tile;
    const int64_t g0 = owner ? offs[cell] : 0;
    const uint32_t tcnt = owner ? static_cast<uint32_t>(offs[cell + 1] - g0) : 0;
    const uint32_t start = block_prefix(tcnt, wsum);
    if (owner) {
      gbase[tid] = static_cast<int32_t>(g0 - start);
      uint32_t s0 = start;
#pragma unroll
      for (int v = 0; v < kSortWaves; v++) {
        const uint32_t c = wcnt[v][tid];
        wcnt[v][tid] = static_cast<uint16_t>(s0);
        s0 += c;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kSortItems; i++) {
    if (e0 + i * kWave < n) {
      const uint32_t pos = cnt[(k[i] >> shift) & dmask] + r[i];
      sk[pos] = k[i];
      sq[pos] = p[i];
      sv[pos] = x[i];
    }
  }
  __syncthreads();
  // 3. Coalesced writes of the sorted tile.
  const int64_t tile_n = n - t0 < kSortTile ? n - t0 : kSortTile;
#pragma unroll
  for (int i = 0; i < kSortItems; i++) {
    const int t = i * kSortThreads + tid;
    if (t < tile_n) {
      const uint32_t kk = sk[t];
      const int64_t pos = static_cast<int64_t>(gbase[(kk >> shift) & dmask]) + t;
      keys_out[pos] = kk;
      pay_out[pos] = static_cast<TQO>(static_cast<int64_t>(sq[t]) + pout_base);
      val_out[pos] = sv[t];
    }
  }
}

// fib[q] = the fibre that owns stored position q.  A wave takes 512
// consecutive positions (8 per lane): the fibres owning its first and last
// position are found by a cooperative 32-ary search (four dependent loads
// for a million fibres), their ends are loaded into LDS, and each lane finds
// its first position's fibre there and walks forward.
constexpr int kExpandRun = 8;
constexpr int kExpandTile = kWave * kExpandRun;
constexpr int kExpandWin = 512;
template <typename TP>
__global__ __launch_bounds__(256) void expand_fibres_kernel(const TP* __restrict__ ptr,
                                                            int64_t base, int64_t n_outer,
                                                            int64_t nnz, uint32_t* __restrict__ fib) {
  __shared__ int64_t lds[4][kExpandWin];
  const int w = wave_in_block();
  const int64_t q0 = (static_cast<int64_t>(blockIdx.x) * 4 + w) * kExpandTile;
  if (q0 >= nnz) return;  // wave-uniform
  const int64_t qlast = (q0 + kExpandTile < nnz ? q0 + kExpandTile : nnz) - 1;
  const int lane = lane_id();
  const int64_t r = group_find_fiber<32>(ptr, base, 0, n_outer, lane < 32 ? q0 : qlast);
  const int64_t r_lo = shfl(r, 0), r_hi = shfl(r, 32);
  const FiberWindow<TP> win = load_window(ptr, base, r_lo, r_hi - r_lo + 1, lds[w], kExpandWin);
  const int64_t a = q0 + static_cast<int64_t>(lane) * kExpandRun;
  if (a > qlast) return;
  int64_t o = win.find(a);
  typedef unsigned v4u __attribute__((ext_vector_type(4)));
  const bool whole = a + kExpandRun <= nnz;
#pragma unroll
  for (int c = 0; c < kExpandRun / 4; c++) {
    uint32_t f[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int64_t q = a + c * 4 + i;
      while (o < r_hi && win.end_of(o) <= q) o++;
      f[i] = static_cast<uint32_t>(o);
    }
    if (whole) {
      *reinterpret_cast<v4u*>(fib + a + c * 4) = v4u{f[0], f[1], f[2], f[3]};
    } else {
      for (int i = 0; i < 4; i++)
        if (a + c * 4 + i < nnz) fib[a + c * 4 + i] = f[i];
    }
  }
}

// ptr[j] = base + (first position p with keys[p] >= j) for j in [0, nk],
// from keys sorted ascending: position p writes the fibres that start at p
// (those after keys[p-1] up to keys[p]).  Keys are clamped to nk so that no
// write leaves ptr.  Each thread takes 4 consecutive positions with one
// 16-byte load (the key before them is one more load, mostly an L1 hit).
template <typename TP>
__device__ __forceinline__ void ptr_at_position(const uint32_t* __restrict__ keys, int64_t n,
                                                int64_t nk, int64_t base, int64_t p,
                                                int64_t prev, int64_t cur, TP* __restrict__ ptr) {
  const int64_t lo = p == 0 ? 0 : std::min<int64_t>(prev, nk) + 1;
  const int64_t hi = p == n ? nk : std::min<int64_t>(cur, nk);
  for (int64_t j = lo; j <= hi; j++) ptr[j] = static_cast<TP>(base + p);
}

template <typename TP>
__global__ __launch_bounds__(256) void ptr_from_keys_kernel(const uint32_t* __restrict__ keys,
                                                            int64_t n, int64_t nk, int64_t base,
                                                            TP* __restrict__ ptr) {
  typedef unsigned v4u __attribute__((ext_vector_type(4)));
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x * 4;
  for (int64_t p0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 4; p0 <= n;
       p0 += stride) {
    int64_t k[5];
    k[0] = p0 == 0 ? 0 : keys[p0 - 1];
    if (p0 + 4 <= n && (reinterpret_cast<uintptr_t>(keys) & 15) == 0) {
      const v4u q = *reinterpret_cast<const v4u*>(keys + p0);
      k[1] = q.x; k[2] = q.y; k[3] = q.z; k[4] = q.w;
    } else {
      for (int i = 0; i < 4; i++) k[i + 1] = p0 + i < n ? keys[p0 + i] : 0;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int64_t p = p0 + i;
      if (p <= n) ptr_at_position<TP>(keys, n, nk, base, p, k[i], k[i + 1], ptr);
    }
  }
}

// flag[0] |= 1 if any coordinate lies outside [base, base + dim).
template <typename TI>
__global__ __launch_bounds__(256) void coord_check_kernel(const TI* __restrict__ a, int64_t n,
                                                          int64_t base, int64_t dim,
                                                          int* __restrict__ flag) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  bool bad = false;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += stride) {
    const int64_t c = static_cast<int64_t>(a[e]) - base;
    bad = bad || c < 0 || c >= dim;
  }
  if (ballot(bad) != 0 && lane_id() == 0) atomicOr(flag, 1);
}

// head[p] = 1 where sorted entry p starts a run of equal (outer, inner).
// head[p] = 1 where a run of equal (outer, inner) coordinates starts: one
// byte per entry, four entries per thread stored as one 32-bit word.
__global__ __launch_bounds__(256) void run_heads_kernel(const uint32_t* __restrict__ ko,
                                                        const uint32_t* __restrict__ ki, int64_t n,
                                                        uint8_t* __restrict__ head) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x * 4;
  const bool aligned = (reinterpret_cast<uintptr_t>(head) & 3) == 0;
  for (int64_t p0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 4; p0 < n;
       p0 += stride) {
    uint32_t w = 0;
    for (int i = 0; i < 4 && p0 + i < n; i++) {
      const int64_t p = p0 + i;
      const bool h = p == 0 || ko[p] != ko[p - 1] || ki[p] != ki[p - 1];
      w |= (h ? 1u : 0u) << (8 * i);
    }
    if (p0 + 4 <= n && aligned) {
      *reinterpret_cast<uint32_t*>(head + p0) = w;
    } else {
      for (int i = 0; i < 4 && p0 + i < n; i++) head[p0 + i] = static_cast<uint8_t>((w >> (8 * i)) & 1u);
    }
  }
}

template <typename T>
__device__ __forceinline__ T combine_add(T a, T b) { return a + b; }
template <>
__device__ __forceinline__ int32_t combine_add<int32_t>(int32_t a, int32_t b) {
  return static_cast<int32_t>(static_cast<uint32_t>(a) + static_cast<uint32_t>(b));
}

// One thread per run head: the run's values summed left to right (foldl),
// written at the run's rank u = uidx[p].
template <typename TI, typename TV>
__global__ __launch_bounds__(256) void combine_runs_kernel(
    const uint32_t* __restrict__ ko, const uint32_t* __restrict__ ki, const TV* __restrict__ v,
    const int32_t* __restrict__ uidx, int64_t n, int64_t base, uint32_t* __restrict__ ou,
    TI* __restrict__ c_idx, TV* __restrict__ c_val) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < n; p += stride) {
    const int32_t u = uidx[p];
    if (uidx[p + 1] == u) continue;  // not a head
    TV acc = v[p];
    for (int64_t s = p + 1; s < n && uidx[s + 1] == uidx[s]; s++) acc = combine_add(acc, v[s]);
    ou[u] = ko[p];
    c_idx[u] = static_cast<TI>(static_cast<int64_t>(ki[p]) + base);
    c_val[u] = acc;
  }
}

// isequal(v, fill): NaN equals NaN; -0.0 differs from 0.0 (bits compared).
__device__ __forceinline__ bool is_fill(double v, double f) {
  if (v != v || f != f) return (v != v) && (f != f);
  return __double_as_longlong(v) == __double_as_longlong(f);
}
__device__ __forceinline__ bool is_fill(float v, float f) {
  if (v != v || f != f) return (v != v) && (f != f);
  return __float_as_uint(v) == __float_as_uint(f);
}
__device__ __forceinline__ bool is_fill(int32_t v, int32_t f) { return v == f; }

// One byte per entry: 1 if it is kept (not the fill).
// Four entries per thread, stored as one 32-bit word.
template <typename TV>
__global__ __launch_bounds__(256) void keep_flags_kernel(const TV* __restrict__ val, int64_t n,
                                                         TV fill, uint8_t* __restrict__ keep) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x * 4;
  const bool aligned = (reinterpret_cast<uintptr_t>(keep) & 3) == 0;
  for (int64_t q0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 4; q0 < n;
       q0 += stride) {
    if (q0 + 4 <= n && aligned) {
      uint32_t w = 0;
#pragma unroll
      for (int i = 0; i < 4; i++) w |= (is_fill(val[q0 + i], fill) ? 0u : 1u) << (8 * i);
      *reinterpret_cast<uint32_t*>(keep + q0) = w;
    } else {
      for (int64_t q = q0; q < n && q < q0 + 4; q++) keep[q] = is_fill(val[q], fill) ? 0 : 1;
    }
  }
}

// c_ptr[o] = base + (kept entries before fibre o).  kept[q] = kept entries
// before position q (TK: int32 when the entry count allows, else int64).
template <typename TP, typename TK>
__global__ __launch_bounds__(256) void kept_ptr_kernel(const TP* __restrict__ ptr, int64_t base,
                                                       int64_t n_outer,
                                                       const TK* __restrict__ kept,
                                                       TP* __restrict__ c_ptr) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t o = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; o <= n_outer;
       o += stride)
    c_ptr[o] = static_cast<TP>(base + static_cast<int64_t>(kept[pos_at(ptr, base, o)]));
}

template <typename TI, typename TV, typename TK>
__global__ __launch_bounds__(256) void compact_kernel(const TI* __restrict__ idx,
                                                      const TV* __restrict__ val, int64_t n,
                                                      const TK* __restrict__ kept,
                                                      TI* __restrict__ c_idx, TV* __restrict__ c_val) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < n; q += stride) {
    const TK u = kept[q];
    if (kept[q + 1] != u) {
      c_idx[u] = idx[q];
      c_val[u] = val[q];
    }
  }
}

inline unsigned grid_for(int64_t n, int64_t per_block, int64_t cap = 65536) {
  const int64_t b = (n + per_block - 1) / per_block;
  return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(b, cap)));
}

// Bits needed for keys in [0, nk).
inline int key_bits(int64_t nk) {
  int b = 0;
  while (b < 63 && (int64_t(1) << b) < nk) b++;
  return b;
}

// Carves 256-byte aligned pieces out of one scratch block.
struct Carver {
  char* p;
  size_t used = 0;
  template <typename T>
  T* take(int64_t count) {
    T* r = reinterpret_cast<T*>(p + used);
    used += (static_cast<size_t>(std::max<int64_t>(count, 1)) * sizeof(T) + 255) & ~size_t(255);
    return r;
  }
};

// Scratch of a sort over n elements with TW value words (two ping-pong sets
// of keys / payloads / values, the digit table and its scan).
template <typename TW>
size_t sort_scratch_bytes(int64_t n) {
  const int64_t ntiles = (n + kSortTile - 1) / kSortTile;
  const int64_t cells = static_cast<int64_t>(kBins) * ntiles;
  const size_t per = 2 * (static_cast<size_t>(n) * (4 + 4 + sizeof(TW)) + 3 * 256);
  return per + 2 * (static_cast<size_t>(cells + 1) * 8 + 256) + scan_work_bytes(cells) + 256;
}

template <typename TW>
struct SortBufs {
  uint32_t* k[2];
  uint32_t* q[2];
  TW* v[2];
  int64_t* hist;
  int64_t* offs;
  void* work;
  void carve(Carver& c, int64_t n) {
    const int64_t ntiles = (n + kSortTile - 1) / kSortTile;
    const int64_t cells = static_cast<int64_t>(kBins) * ntiles;
    for (int i = 0; i < 2; i++) {
      k[i] = c.take<uint32_t>(n);
      q[i] = c.take<uint32_t>(n);
      v[i] = c.take<TW>(n);
    }
    hist = c.take<int64_t>(cells + 1);
    offs = c.take<int64_t>(cells + 1);
    work = c.take<char>(static_cast<int64_t>(scan_work_bytes(cells)));
  }
};

// One radix pass on the dbits-wide digit at bit `shift`: counts, scan,
// stable scatter.
template <typename TK, typename TQ, typename TW, typename TQO>
int sort_pass(const TK* keys, int64_t kbase, const TQ* pay, int64_t pbase, const TW* val, int64_t n,
              int shift, int dbits, SortBufs<TW>& S, uint32_t* kout, TQO* pout, int64_t pout_base,
              TW* vout, hipStream_t s) {
  const int64_t ntiles = (n + kSortTile - 1) / kSortTile;
  hipLaunchKernelGGL((sort_hist_kernel<TK>), dim3(static_cast<unsigned>(ntiles)),
                     dim3(kSortThreads), 0, s, keys, kbase, n, shift, dbits, ntiles, S.hist);
  FGX_TRY(launch_check("sort_hist_kernel"));
  FGX_TRY(scan_counts_to_ptr<int64_t>(S.hist, (int64_t(1) << dbits) * ntiles, 0, S.offs, S.work,
                                      nullptr, nullptr, s));
  hipLaunchKernelGGL((sort_scatter_kernel<TK, TQ, TW, TQO>), dim3(static_cast<unsigned>(ntiles)),
                     dim3(kSortThreads), 0, s, keys, kbase, pay, pbase, val, n, shift, dbits,
                     ntiles, S.offs, kout, pout, pout_base, vout);
  return launch_check("sort_scatter_kernel");
}

// Stable sort of n (key, payload, value) triples by key in [0, 2^bits).
// Keys and payloads are read as (a[e] - base); the sorted keys go to kout,
// the payloads to pout (+ pout_base, as TQO) and the values to vout.
template <typename TK, typename TQ, typename TW, typename TQO>
int radix_sort(const TK* keys, int64_t kbase, const TQ* pay, int64_t pbase, const TW* val,
               int64_t n, int bits, SortBufs<TW>& S, uint32_t* kout, TQO* pout, int64_t pout_base,
               TW* vout, hipStream_t s) {
  // As few passes as 8-bit digits need, with the bits spread evenly over
  // them: a 20-bit key takes 7+7+6 rather than 8+8+4, so each digit's run in
  // a 4096-element tile averages 32 elements (a full 128-byte line of keys)
  // instead of 16.
  const int passes = std::max(1, (bits + kDigitBits - 1) / kDigitBits);
  const int dbits = std::max(1, (bits + passes - 1) / passes);
  if (passes == 1)
    return sort_pass<TK, TQ, TW, TQO>(keys, kbase, pay, pbase, val, n, 0, dbits, S, kout, pout,
                                      pout_base, vout, s);
  FGX_TRY((sort_pass<TK, TQ, TW, uint32_t>(keys, kbase, pay, pbase, val, n, 0, dbits, S, S.k[0],
                                           S.q[0], 0, S.v[0], s)));
  int cur = 0;
  for (int pass = 1; pass < passes - 1; pass++) {
    FGX_TRY((sort_pass<uint32_t, uint32_t, TW, uint32_t>(S.k[cur], 0, S.q[cur], 0, S.v[cur], n,
                                                          pass * dbits, dbits, S, S.k[cur ^ 1],
                                                          S.q[cur ^ 1], 0, S.v[cur ^ 1], s)));
    cur ^= 1;
  }
  return sort_pass<uint32_t, uint32_t, TW, TQO>(S.k[cur], 0, S.q[cur], 0, S.v[cur], n,
                                                (passes - 1) * dbits, dbits, S, kout, pout,
                                                pout_base, vout, s);
}

template <typename TP>
int write_ptr_from_keys(const uint32_t* keys, int64_t n, int64_t nk, int64_t base, TP* ptr,
                        hipStream_t s) {
  hipLaunchKernelGGL((ptr_from_keys_kernel<TP>), dim3(grid_for(n + 1, 256 * 4, 16384)), dim3(256),
                     0, s, keys, n, nk, base, ptr);
  return launch_check("ptr_from_keys_kernel");
}

// ---- transpose ------------------------------------------------------------------------------
template <typename TP, typename TI, int VB>
int transpose_typed(fgx_ctx* ctx, const fgx_compressed* A, void* t_ptr, void* t_idx, void* t_val,
                    hipStream_t s) {
  using TW = typename word_of<VB>::type;
  const int64_t n = A->nnz, base = A->index_base;
  if (n == 0) return write_ptr_from_keys<TP>(nullptr, 0, A->n_inner, base, static_cast<TP*>(t_ptr), s);
  const size_t bytes = sort_scratch_bytes<TW>(n) + 2 * (static_cast<size_t>(n) * 4 + 256);
  void* scr;
  FGX_TRY(scratch(ctx, s, bytes, &scr));
  Carver c{static_cast<char*>(scr)};
  uint32_t* fib = c.take<uint32_t>(n);
  uint32_t* kfin = c.take<uint32_t>(n);
  SortBufs<TW> S;
  S.carve(c, n);
  hipLaunchKernelGGL((expand_fibres_kernel<TP>), dim3(grid_for(n, 4 * kExpandTile, INT32_MAX)),
                     dim3(256), 0, s, static_cast<const TP*>(A->ptr), base, A->n_outer, n, fib);
  FGX_TRY(launch_check("expand_fibres_kernel"));
  FGX_TRY((radix_sort<TI, uint32_t, TW, TI>(static_cast<const TI*>(A->idx), base, fib, 0,
                                            static_cast<const TW*>(A->val), n,
                                            key_bits(A->n_inner), S, kfin,
                                            static_cast<TI*>(t_idx), base,
                                            static_cast<TW*>(t_val), s)));
  return write_ptr_from_keys<TP>(kfin, n, A->n_inner, base, static_cast<TP*>(t_ptr), s);
}

// ---- COO -> compressed ---------------------------------------------------------------------
template <typename TI, typename TV>
int coo_typed(fgx_ctx* ctx, int64_t n_outer, int64_t n_inner, int64_t n, int64_t base,
              const TI* outer, const TI* inner, const TV* val, TI* c_ptr, TI* c_idx, TV* c_val,
              int64_t* c_nnz, hipStream_t s) {
  if (n == 0) {
    *c_nnz = 0;
    return write_ptr_from_keys<TI>(nullptr, 0, n_outer, base, c_ptr, s);
  }
  const size_t n4 = static_cast<size_t>(n) * 4 + 256, nv = static_cast<size_t>(n) * sizeof(TV) + 256;
  const size_t bytes = sort_scratch_bytes<TV>(n) + 5 * n4 + 2 * nv +
                       (2 * static_cast<size_t>(n) + 1) * 8 + 512 + scan_work_bytes(n) + 1024;
  void* scr;
  FGX_TRY(scratch(ctx, s, bytes, &scr));
  Carver c{static_cast<char*>(scr)};
  uint32_t* k1 = c.take<uint32_t>(n);  // inner, after the sort by inner
  uint32_t* q1 = c.take<uint32_t>(n);  // outer, after the sort by inner
  TV* v1 = c.take<TV>(n);
  uint32_t* ko = c.take<uint32_t>(n);  // outer, in (outer, inner) order
  uint32_t* ki = c.take<uint32_t>(n);  // inner, in (outer, inner) order
  TV* v2 = c.take<TV>(n);
  uint32_t* ou = c.take<uint32_t>(n);  // outer of each distinct coordinate
  // One-byte run heads and an int32 run index (nnz is below 2^31 here).
  uint8_t* head = c.take<uint8_t>(n);
  int32_t* uidx = c.take<int32_t>(n + 1);
  void* work = c.take<char>(static_cast<int64_t>(scan_work_bytes(n)));
  int* flag = c.take<int>(1);
  SortBufs<TV> S;
  S.carve(c, n);
  // Coordinates outside the shape are an argument error (fsparse requires
  // 1 <= I[n][q] <= M[n]); checked before anything is placed.
  FGX_HIP(hipMemsetAsync(flag, 0, sizeof(int), s));
  hipLaunchKernelGGL((coord_check_kernel<TI>), dim3(grid_for(n, 256, 4096)), dim3(256), 0, s,
                     outer, n, base, n_outer, flag);
  FGX_TRY(launch_check("coord_check_kernel"));
  hipLaunchKernelGGL((coord_check_kernel<TI>), dim3(grid_for(n, 256, 4096)), dim3(256), 0, s,
                     inner, n, base, n_inner, flag);
  FGX_TRY(launch_check("coord_check_kernel"));
  int* h_flag = reinterpret_cast<int*>(ctx->host_i64);
  FGX_HIP(hipMemcpyAsync(h_flag, flag, sizeof(int), hipMemcpyDeviceToHost, s));
  FGX_HIP(hipStreamSynchronize(s));
  if (*h_flag) return set_error(FGX_EDIM, "fgx_coo_to_compressed: coordinate outside the shape");
  // Stable by inner, then stable by outer: (outer, inner) order with ties
  // (duplicate coordinates) in input order.
  FGX_TRY((radix_sort<TI, TI, TV, uint32_t>(inner, base, outer, base, val, n, key_bits(n_inner), S,
                                            k1, q1, 0, v1, s)));
  FGX_TRY((radix_sort<uint32_t, uint32_t, TV, uint32_t>(q1, 0, k1, 0, v1, n, key_bits(n_outer), S,
                                                         ko, ki, 0, v2, s)));
  hipLaunchKernelGGL(run_heads_kernel, dim3(grid_for(n, 256 * 4)), dim3(256), 0, s, ko, ki, n,
                     head);
  FGX_TRY(launch_check("run_heads_kernel"));
  int64_t u = 0;
  FGX_TRY((scan_counts_to_ptr<int32_t, uint8_t>(head, n, 0, uidx, work, ctx->host_i64, &u, s)));
  hipLaunchKernelGGL((combine_runs_kernel<TI, TV>), dim3(grid_for(n, 256)), dim3(256), 0, s, ko, ki,
                     v2, uidx, n, base, ou, c_idx, c_val);
  FGX_TRY(launch_check("combine_runs_kernel"));
  FGX_TRY(write_ptr_from_keys<TI>(ou, u, n_outer, base, c_ptr, s));
  *c_nnz = u;
  return FGX_OK;
}

// ---- dropfills ------------------------------------------------------------------------------
template <typename TP, typename TI, typename TV>
int dropfills_typed(fgx_ctx* ctx, const fgx_compressed* A, TV fill, void* c_ptr, void* c_idx,
                    void* c_val, int64_t* c_nnz, hipStream_t s) {
  const int64_t n = A->nnz, base = A->index_base;
  if (n == 0) {
    *c_nnz = 0;
    return write_ptr_from_keys<TP>(nullptr, 0, A->n_outer, base, static_cast<TP*>(c_ptr), s);
  }
  // Keep flags are one byte each; the prefix of kept entries is int32 when
  // it fits (fewer bytes through the scan and the compaction), else int64.
  auto run = [&](auto kept_tag) -> int {
    using TK = decltype(kept_tag);
    void* scr;
    FGX_TRY(scratch(ctx, s, static_cast<size_t>(n) + 256 + (static_cast<size_t>(n) + 1) * sizeof(TK) +
                             256 + scan_work_bytes(n) + 256,
                    &scr));
    Carver c{static_cast<char*>(scr)};
    uint8_t* keep = c.take<uint8_t>(n);
    TK* kept = c.take<TK>(n + 1);
    void* work = c.take<char>(static_cast<int64_t>(scan_work_bytes(n)));
    const TV* val = static_cast<const TV*>(A->val);
    hipLaunchKernelGGL((keep_flags_kernel<TV>), dim3(grid_for(n, 256 * 4)), dim3(256), 0, s, val,
                       n, fill, keep);
    FGX_TRY(launch_check("keep_flags_kernel"));
    int64_t u = 0;
    FGX_TRY((scan_counts_to_ptr<TK, uint8_t>(keep, n, 0, kept, work, ctx->host_i64, &u, s)));
    hipLaunchKernelGGL((kept_ptr_kernel<TP, TK>), dim3(grid_for(A->n_outer + 1, 256)), dim3(256),
                       0, s, static_cast<const TP*>(A->ptr), base, A->n_outer, kept,
                       static_cast<TP*>(c_ptr));
    FGX_TRY(launch_check("kept_ptr_kernel"));
    hipLaunchKernelGGL((compact_kernel<TI, TV, TK>), dim3(grid_for(n, 256)), dim3(256), 0, s,
                       static_cast<const TI*>(A->idx), val, n, kept, static_cast<TI*>(c_idx),
                       static_cast<TV*>(c_val));
    FGX_TRY(launch_check("compact_kernel"));
    *c_nnz = u;
    return FGX_OK;
  };
  return n < INT32_MAX ? run(int32_t{}) : run(int64_t{});
}

}  // namespace
}  // namespace fgx

using namespace fgx;

extern "C" {

int fgx_transpose(fgx_ctx* ctx, const fgx_compressed* A, void* t_ptr, void* t_idx, void* t_val,
                  void* stream) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_transpose: NULL ctx");
  FGX_TRY(check_compressed(A, "fgx_transpose"));
  if (!t_ptr || (A->nnz > 0 && (!t_idx || !t_val)))
    return set_error(FGX_EINVAL, "fgx_transpose: NULL output");
  if (A->n_outer > UINT32_MAX || A->n_inner > UINT32_MAX || A->nnz > INT32_MAX)
    return set_error(FGX_EUNSUP, "fgx_transpose: dimensions beyond 2^32 or nnz beyond 2^31");
  if (A->n_outer == 0 && A->nnz > 0)
    return set_error(FGX_EINVAL, "fgx_transpose: stored entries without fibres");
  IdxCombo ic;
  FGX_TRY(idx_combo(A, &ic));
  const int vb = elem_size(A->val_type);
  if (vb == 0) return set_error(FGX_EUNSUP, "fgx_transpose: value type");
  std::lock_guard<std::mutex> lk(ctx->mu);
  FGX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = pick_stream(ctx, stream);
  FGX_TRY(admit_operand(ctx, A, s, "fgx_transpose"));
  int rc = FGX_OK;
  FGX_DISPATCH_IDX(ic, {
    if (vb == 2) rc = transpose_typed<TP, TI, 2>(ctx, A, t_ptr, t_idx, t_val, s);
    else if (vb == 4) rc = transpose_typed<TP, TI, 4>(ctx, A, t_ptr, t_idx, t_val, s);
    else rc = transpose_typed<TP, TI, 8>(ctx, A, t_ptr, t_idx, t_val, s);
  });
  return rc;
}

}  // extern "C"

namespace fgx {

// fgx_transpose's work for callers that already hold ctx->mu and checked A.
int transpose_into(fgx_ctx* ctx, const fgx_compressed* A, void* t_ptr, void* t_idx, void* t_val,
                   hipStream_t s) {
  if (A->n_outer > UINT32_MAX || A->n_inner > UINT32_MAX || A->nnz > INT32_MAX)
    return set_error(FGX_EUNSUP, "transpose: dimensions beyond 2^32 or nnz beyond 2^31");
  IdxCombo ic;
  FGX_TRY(idx_combo(A, &ic));
  const int vb = elem_size(A->val_type);
  if (vb == 0) return set_error(FGX_EUNSUP, "transpose: value type");
  int rc = FGX_OK;
  FGX_DISPATCH_IDX(ic, {
    if (vb == 2) rc = transpose_typed<TP, TI, 2>(ctx, A, t_ptr, t_idx, t_val, s);
    else if (vb == 4) rc = transpose_typed<TP, TI, 4>(ctx, A, t_ptr, t_idx, t_val, s);
    else rc = transpose_typed<TP, TI, 8>(ctx, A, t_ptr, t_idx, t_val, s);
  });
  return rc;
}

}  // namespace fgx

extern "C" {

int fgx_coo_to_compressed(fgx_ctx* ctx, int64_t n_outer, int64_t n_inner, int64_t nnz,
                          int idx_type, int val_type, int index_base, const void* outer,
                          const void* inner, const void* val, void* c_ptr, void* c_idx,
                          void* c_val, int64_t* c_nnz, void* stream) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_coo_to_compressed: NULL ctx");
  if (n_outer < 0 || n_inner < 0 || nnz < 0)
    return set_error(FGX_EINVAL, "fgx_coo_to_compressed: negative size");
  if (index_base != 0 && index_base != 1)
    return set_error(FGX_EINVAL, "fgx_coo_to_compressed: index_base must be 0 or 1");
  if (!c_ptr || !c_nnz || (nnz > 0 && (!outer || !inner || !val || !c_idx || !c_val)))
    return set_error(FGX_EINVAL, "fgx_coo_to_compressed: NULL argument");
  if (n_outer > UINT32_MAX || n_inner > UINT32_MAX || nnz > INT32_MAX)
    return set_error(FGX_EUNSUP, "fgx_coo_to_compressed: dimensions beyond 2^32 or nnz beyond 2^31");
  if (idx_type != FGX_I32 && idx_type != FGX_I64)
    return set_error(FGX_EUNSUP, "fgx_coo_to_compressed: coordinates must be I32 or I64");
  std::lock_guard<std::mutex> lk(ctx->mu);
  FGX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = pick_stream(ctx, stream);
#define FGX_COO(TI_, TV_)                                                                       \
  coo_typed<TI_, TV_>(ctx, n_outer, n_inner, nnz, index_base, static_cast<const TI_*>(outer),   \
                      static_cast<const TI_*>(inner), static_cast<const TV_*>(val),             \
                      static_cast<TI_*>(c_ptr), static_cast<TI_*>(c_idx), static_cast<TV_*>(c_val), \
                      c_nnz, s)
#define FGX_COO_VAL(TI_)                                                                  \
  switch (val_type) {                                                                     \
    case FGX_F64: return FGX_COO(TI_, double);                                            \
    case FGX_F32: return FGX_COO(TI_, float);                                             \
    case FGX_I32: return FGX_COO(TI_, int32_t);                                           \
    default: return set_error(FGX_EUNSUP, "fgx_coo_to_compressed: value type");           \
  }
  if (idx_type == FGX_I32) {
    FGX_COO_VAL(int32_t);
  } else {
    FGX_COO_VAL(int64_t);
  }
#undef FGX_COO_VAL
#undef FGX_COO
}

int fgx_dropfills(fgx_ctx* ctx, const fgx_compressed* A, double fill, void* c_ptr, void* c_idx,
                  void* c_val, int64_t* c_nnz, void* stream) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_dropfills: NULL ctx");
  FGX_TRY(check_compressed(A, "fgx_dropfills"));
  if (!c_ptr || !c_nnz || (A->nnz > 0 && (!c_idx || !c_val)))
    return set_error(FGX_EINVAL, "fgx_dropfills: NULL output");
  IdxCombo ic;
  FGX_TRY(idx_combo(A, &ic));
  std::lock_guard<std::mutex> lk(ctx->mu);
  FGX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = pick_stream(ctx, stream);
  FGX_TRY(admit_operand(ctx, A, s, "fgx_dropfills"));
  int rc = FGX_OK;
  switch (A->val_type) {
    case FGX_F64:
      FGX_DISPATCH_IDX(ic, rc = (dropfills_typed<TP, TI, double>(ctx, A, fill, c_ptr, c_idx, c_val,
                                                                 c_nnz, s)));
      break;
    case FGX_F32:
      FGX_DISPATCH_IDX(ic, rc = (dropfills_typed<TP, TI, float>(ctx, A, static_cast<float>(fill),
                                                                c_ptr, c_idx, c_val, c_nnz, s)));
      break;
    case FGX_I32:
      if (!(fill >= INT32_MIN && fill <= INT32_MAX) ||
          fill != static_cast<double>(static_cast<int32_t>(fill)))
        return set_error(FGX_EINVAL, "fgx_dropfills: fill is not an Int32 value");
      FGX_DISPATCH_IDX(ic, rc = (dropfills_typed<TP, TI, int32_t>(ctx, A, static_cast<int32_t>(fill),
                                                                  c_ptr, c_idx, c_val, c_nnz, s)));
      break;
    default:
      return set_error(FGX_EUNSUP, "fgx_dropfills: value type (F64, F32 or I32)");
  }
  return rc;
}

}  // extern "C"
