// fgx_spgemm.hip — Gustavson SpGEMM C = A*B over CSR operands, row by row,
// as a symbolic/numeric pair (docs/examples/spgemm.jl:59-76; generated code
// test/reference64/kernels/gustavsons.jl).
//
// Each output row C_i = sum over stored k of A_i of A[i,k] * B_k.  The
// reference accumulates into a SparseByteMap workspace in k order and sorts
// the touched coordinates at the end.  Here a team of lanes (a wave for rows
// with at most 256 products, a 256-thread block up to 1536) instead
//   1. gathers the row's products into LDS, list by list (list t is B_k of the
//      t-th stored k, already sorted by column);
//   2. merges the lists pairwise (log2 of their count levels); every product
//      finds its place in the merged list by a binary search in the partner
//      list, and ties keep the left list first, so equal columns stay in k
//      order;
//   3. folds each run of equal columns in that order, from the semiring's
//      identity: w = a*b + w (plus-times) or w = min(w, a+b) (min-plus) —
//      the reference's update sequence, so even FP32/FP64 plus-times sums
//      round exactly as the oracle's.
// The symbolic pass needs only each row's count of distinct columns, which no
// order affects: there a team inserts the row's columns into an
// open-addressing hash table in LDS (linear probing by LDS compare-and-swap)
// and counts the inserts that fill a slot.  (Hash accumulation was also
// measured for the numeric pass of the order-free semirings, Int32
// plus-times and min-plus: with the sort of each row's distinct columns it
// ran at 1.1-1.7 ms per kernel against 0.6-0.9 ms for the ordered merge.)
// Rows with more products or lists take a windowed dense accumulator in LDS
// (exact for Int32 and min-plus; FP plus-times sums there use LDS atomics).
#include "fgx_scan.hpp"

#include <type_traits>

namespace fgx {

constexpr int kWaveCap = 256;    // products (and lists) per wave-team row
constexpr int kWide = 512;       // products per wide wave row (numeric register sort)
constexpr int kBlockCap = 1536;  // products (and lists) per block-team row
constexpr int kWindow = 4096;    // columns per pass of the dense fallback

template <int NT>
__device__ __forceinline__ void team_sync() {
  if constexpr (NT == 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Exclusive scan over a team; `tmp` holds NT/64 + 1 ints of LDS.
// Inclusive sum over the wave's 64 lanes by DPP register moves (no LDS
// round trips): shifts by 1, 2, 4 and 8 inside each 16-lane row (lanes
// without a source add 0), then lane 15 of rows 0 and 2 into rows 1 and 3,
// then lane 31 into rows 2 and 3.
__device__ __forceinline__ int wave_inclusive_sum(int x) {
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, true);   // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, true);   // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, true);   // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, true);   // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);  // row_bcast:15
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false);  // row_bcast:31
  return x;
}

template <int NT>
__device__ __forceinline__ int team_scan(int v, int tt, int* tmp, int* total) {
  const int lane = tt & 63;
  const int x = wave_inclusive_sum(v);
  if constexpr (NT == 64) {
    *total = __builtin_amdgcn_readlane(x, 63);
    return x - v;
  } else {
    const int w = tt >> 6;
    if (lane == 63) tmp[w] = x;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; i++) {
      const int c = tmp[i];
      off += i < w ? c : 0;
      tot += c;
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
  }
}

template <typename TV, int SR>
__device__ __forceinline__ TV sr_mul(TV a, TV b) {
  if constexpr (SR == FGX_MIN_PLUS) return a + b;
  else if constexpr (sizeof(TV) == 4 && TV(0.5) == TV(0))  // Int32: wrap like Julia
    return static_cast<TV>(static_cast<uint32_t>(a) * static_cast<uint32_t>(b));
  else return a * b;
}
// Julia's `min` on floats: NaN wins (as the quiet NaN), -0.0 is below 0.0.
template <typename TV>
__device__ __forceinline__ TV jl_min(TV x, TV y) {
  if (x != x || y != y) return static_cast<TV>(__builtin_nan(""));
  if (y < x || (y == x && __builtin_signbit(y) && !__builtin_signbit(x))) return y;
  return x;
}

template <typename TV, int SR>
__device__ __forceinline__ TV sr_add(TV acc, TV p) {
  if constexpr (SR == FGX_MIN_PLUS) return jl_min(acc, p);
  else if constexpr (sizeof(TV) == 4 && TV(0.5) == TV(0))
    return static_cast<TV>(static_cast<uint32_t>(p) + static_cast<uint32_t>(acc));
  else return p + acc;
}
template <typename TV, int SR>
__device__ __forceinline__ TV sr_zero() {
  if constexpr (SR == FGX_MIN_PLUS) return static_cast<TV>(__builtin_huge_val());
  else return TV(0);
}

// Last position p in [0, n) with off[p] <= x (the list holding product x).
__device__ __forceinline__ int list_of(const int* off, int n, int x) {
  int lo = 0, hi = n;  // off[0] == 0 <= x
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

// Count of keys[lo, hi) < key (strict) or <= key.
template <bool LE>
__device__ __forceinline__ int count_below(const int* keys, int lo, int hi, int key) {
  int a = lo, b = hi;
  while (a < b) {
    const int mid = (a + b) >> 1;
    const bool go = LE ? keys[mid] <= key : keys[mid] < key;
    if (go) a = mid + 1; else b = mid;
  }
  return a - lo;
}

// Per-team LDS workspace for rows of up to CAP products and CAP lists.
template <int CAP, typename TV>
struct TeamLds {
  int keys[2][CAP];
  TV vals[2][CAP];
  int off[2][CAP + 1];
  int lk[CAP];      // B row of list t
  TV la[CAP];       // A value of list t
  int tmp[8];
};

// Processes one row with a team of NT lanes.  NUMERIC = false: writes the
// row's distinct-column count to cnt[row].  NUMERIC = true: writes the row's
// sorted columns and folded values at cptr[row].
template <int NT, int CAP, bool NUMERIC, int SR, typename TP, typename TI, typename TV>
__device__ void team_row(int64_t row, int tt, int64_t base, const TP* __restrict__ aptr,
                         const TI* __restrict__ aidx, const TV* __restrict__ aval,
                         const TP* __restrict__ bptr, const TI* __restrict__ bidx,
                         const TV* __restrict__ bval, int64_t* __restrict__ cnt,
                         const TP* __restrict__ cptr, TI* __restrict__ cidx,
                         TV* __restrict__ cval, TeamLds<CAP, TV>& L) {
  const int64_t a0 = pos_at(aptr, base, row), a1 = pos_at(aptr, base, row + 1);
  const int nl = static_cast<int>(a1 - a0);
  // 1. List offsets (prefix sum of the B row lengths).
  int run = 0;
  for (int c = 0; c < nl; c += NT) {
    const int t = c + tt;
    int len = 0, k = 0;
    if (t < nl) {
      k = static_cast<int>(static_cast<int64_t>(aidx[a0 + t]) - base);
      len = static_cast<int>(pos_at(bptr, base, k + 1) - pos_at(bptr, base, k));
      L.lk[t] = k;
      if constexpr (NUMERIC) L.la[t] = aval[a0 + t];
    }
    int tot;
    const int ex = team_scan<NT>(len, tt, L.tmp, &tot);
    if (t < nl) L.off[0][t] = run + ex;
    run += tot;
  }
  const int np = run;
  if (tt == 0) L.off[0][nl] = np;
  team_sync<NT>();
  // 2. Gather the products.
  for (int p = tt; p < np; p += NT) {
    const int t = list_of(L.off[0], nl, p);
    const int k = L.lk[t];
    const int64_t q = pos_at(bptr, base, k) + (p - L.off[0][t]);
    L.keys[0][p] = static_cast<int>(static_cast<int64_t>(bidx[q]) - base);
    if constexpr (NUMERIC) L.vals[0][p] = sr_mul<TV, SR>(L.la[t], bval[q]);
  }
  team_sync<NT>();
  // 3. Pairwise merges, stable (left list first on ties).
  int cur = 0, nlist = nl;
  while (nlist > 1) {
    const int* off = L.off[cur];
    for (int p = tt; p < np; p += NT) {
      const int t = list_of(off, nlist, p);
      const int key = L.keys[cur][p];
      int np_ = p;
      if ((t ^ 1) < nlist) {
        if ((t & 1) == 0)
          np_ = p + count_below<false>(L.keys[cur], off[t + 1], off[t + 2], key);
        else
          np_ = off[t - 1] + (p - off[t]) +
                count_below<true>(L.keys[cur], off[t - 1], off[t], key);
      }
      L.keys[cur ^ 1][np_] = key;
      if constexpr (NUMERIC) L.vals[cur ^ 1][np_] = L.vals[cur][p];
    }
    const int nn = (nlist + 1) >> 1;
    for (int s = tt; s <= nn; s += NT) L.off[cur ^ 1][s] = s < nn ? off[2 * s] : np;
    team_sync<NT>();
    cur ^= 1;
    nlist = nn;
  }
  // 4. Runs of equal columns.
  int seen = 0;
  for (int c = 0; c < np; c += NT) {
    const int p = c + tt;
    const bool head = p < np && (p == 0 || L.keys[cur][p] != L.keys[cur][p - 1]);
    int tot;
    const int ex = team_scan<NT>(head ? 1 : 0, tt, L.tmp, &tot);
    if constexpr (NUMERIC) {
      if (head) {
        const int key = L.keys[cur][p];
        TV w = sr_zero<TV, SR>();
        for (int r = p; r < np && L.keys[cur][r] == key; r++) w = sr_add<TV, SR>(w, L.vals[cur][r]);
        const int64_t o = pos_at(cptr, base, row) + seen + ex;
        cidx[o] = static_cast<TI>(key + base);
        cval[o] = w;
      }
    }
    seen += tot;
  }
  if constexpr (!NUMERIC) {
    if (tt == 0) cnt[row] = seen;
  }
  team_sync<NT>();
}

// Wave teams: 4 rows per 256-thread block.
template <bool NUMERIC, int SR, typename TP, typename TI, typename TV>
__global__ __launch_bounds__(256) void spgemm_wave_kernel(
    const int64_t* __restrict__ rows, const int64_t* __restrict__ nrows_p, int64_t base,
    const TP* __restrict__ aptr, const TI* __restrict__ aidx, const TV* __restrict__ aval,
    const TP* __restrict__ bptr, const TI* __restrict__ bidx, const TV* __restrict__ bval,
    int64_t* __restrict__ cnt, const TP* __restrict__ cptr, TI* __restrict__ cidx,
    TV* __restrict__ cval) {
  __shared__ TeamLds<kWaveCap, TV> lds[4];
  const int w = wave_in_block();
  const int64_t nrows = *nrows_p;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + w; r < nrows;
       r += static_cast<int64_t>(gridDim.x) * 4) {
    team_row<64, kWaveCap, NUMERIC, SR>(rows[r], threadIdx.x & 63, base, aptr, aidx, aval, bptr,
                                        bidx, bval, cnt, cptr, cidx, cval, lds[w]);
    team_sync<64>();  // the wave's LDS is reused by its next row
  }
}

template <bool NUMERIC, int SR, typename TP, typename TI, typename TV>
__global__ __launch_bounds__(256) void spgemm_block_kernel(
    const int64_t* __restrict__ rows, const int64_t* __restrict__ nrows_p, int64_t base,
    const TP* __restrict__ aptr, const TI* __restrict__ aidx, const TV* __restrict__ aval,
    const TP* __restrict__ bptr, const TI* __restrict__ bidx, const TV* __restrict__ bval,
    int64_t* __restrict__ cnt, const TP* __restrict__ cptr, TI* __restrict__ cidx,
    TV* __restrict__ cval) {
  extern __shared__ __align__(16) unsigned char dyn[];
  TeamLds<kBlockCap, TV>& L = *reinterpret_cast<TeamLds<kBlockCap, TV>*>(dyn);
  const int64_t nrows = *nrows_p;
  for (int64_t r = blockIdx.x; r < nrows; r += gridDim.x) {
    team_row<256, kBlockCap, NUMERIC, SR>(rows[r], threadIdx.x, base, aptr, aidx, aval, bptr,
                                          bidx, bval, cnt, cptr, cidx, cval, L);
    __syncthreads();  // the block's LDS is reused by its next row
  }
}

// ---- wave rows, numeric: register sort ------------------------------------------------
// A row of at most 64*PER products (PER = 4 or 8) as (key, value) pairs,
// PER per lane (element e = PER * lane + r): key = column << PB | product
// index, where the product index p (PB bits) counts the row's products in
// the reference's order (A position, then B position).  A bitonic sort of
// the keys in registers (shuffles across lanes, swaps within a lane) leaves
// each column's products in that order, so folding each run left to right
// is the reference's update sequence, bit for bit.  The key is 31 bits, so
// columns must be below 2^(31 - PB).
template <int PER>
struct SortShape {
  static constexpr int kCap = 64 * PER;
  static constexpr int kPBits = PER == 4 ? 8 : 9;
  static constexpr int64_t kMaxCols = int64_t(1) << (31 - kPBits);
};

// The product lists (steps 1-2) and the sorted keys (step 4) are never live
// together, so they share the wave's LDS: half the footprint, twice the
// resident waves for the 512-product rows.
template <int PER, typename TV, typename TP>
struct WaveSortLds {
  union {
    struct {
      int off[SortShape<PER>::kCap + 1];
      TP bs[SortShape<PER>::kCap];  // 0-based start of each list's B row
    } in;
    uint32_t sk[SortShape<PER>::kCap];  // sorted keys
  } u;
  TV pv[SortShape<PER>::kCap];  // product values by product index
};

// Value of lane (lane ^ d) for the bitonic network.  Distances 1 and 2 are
// quad permutations and 8 is a rotation by 8 within each 16-lane row (the
// same as xor 8 there), all DPP register moves.  Distance 4 takes two row
// rotations (row_ror:n hands lane i the value of lane (i - n) mod 16): lanes
// with bit 2 clear want lane i + 4 (rotation 12), the others lane i - 4
// (rotation 4).  The other distances (16, 31, 32, 63) go through the LDS
// crossbar.  d is a constant once the network is unrolled.
__device__ __forceinline__ uint32_t xor_lanes(uint32_t v, int d) {
  switch (d) {
    // Mirrors (lane ^ (2^m - 1)) of the sort's flip steps: within a quad,
    // within 8 lanes (row_half_mirror) and within a 16-lane row (row_mirror).
    case 3: return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x1B, 0xF, 0xF, false));
    case 7: return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x141, 0xF, 0xF, false));
    case 15: return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x140, 0xF, 0xF, false));
    case 4: {
      const int up = __builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x12C, 0xF, 0xF, false);
      const int dn = __builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x124, 0xF, 0xF, false);
      return static_cast<uint32_t>((lane_id() & 4) ? dn : up);
    }
    case 1: return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0xB1, 0xF, 0xF, false));
    case 2: return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x4E, 0xF, 0xF, false));
    case 8: return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x128, 0xF, 0xF, false));
    default: return __shfl_xor(v, d, 64);
  }
}

// Stages K0, 2*K0, .., K1 of the bitonic sort of the 64*PER keys (element
// e = PER*lane + r in register r of lane `lane`), in the form whose
// compare-exchanges all put the minimum at the lower position: each merge
// stage of size k first pairs e with its mirror e ^ (k - 1) inside the
// k-block, then e with e ^ j for j = k/4 .. 1.  Pairs inside a lane are then
// a min and a max with no direction select; across lanes the lower lane of a
// pair keeps the minimum.  Stage k leaves every k-block sorted when every
// k/2-block was.
template <int PER, int K0, int K1>
__device__ __forceinline__ void bitonic_stages(uint32_t (&key)[PER], int lane) {
#pragma unroll
  for (int k = K0; k <= K1; k <<= 1) {
    if (k <= PER) {
#pragma unroll
      for (int r = 0; r < PER; r++) {
        const int r2 = r ^ (k - 1);
        if (r2 <= r) continue;
        const uint32_t lo = key[r] < key[r2] ? key[r] : key[r2];
        const uint32_t hi = key[r] < key[r2] ? key[r2] : key[r];
        key[r] = lo;
        key[r2] = hi;
      }
    } else {
      // The mirror of element PER*lane + r is register PER-1-r of lane
      // lane ^ (k/PER - 1).
      const bool lower = (lane & (k / (2 * PER))) == 0;
      uint32_t pk[PER];
#pragma unroll
      for (int r = 0; r < PER; r++) pk[r] = xor_lanes(key[PER - 1 - r], k / PER - 1);
#pragma unroll
      for (int r = 0; r < PER; r++) {
        const uint32_t lo = pk[r] < key[r] ? pk[r] : key[r];
        const uint32_t hi = pk[r] < key[r] ? key[r] : pk[r];
        key[r] = lower ? lo : hi;
      }
    }
#pragma unroll
    for (int j = k >> 2; j > 0; j >>= 1) {
      if (j >= PER) {
        const int d = j / PER;
        const bool lower = (lane & d) == 0;
#pragma unroll
        for (int r = 0; r < PER; r++) {
          const uint32_t pk = xor_lanes(key[r], d);
          const uint32_t lo = pk < key[r] ? pk : key[r];
          const uint32_t hi = pk < key[r] ? key[r] : pk;
          key[r] = lower ? lo : hi;
        }
      } else {
#pragma unroll
        for (int r = 0; r < PER; r++) {
          if (r & j) continue;
          const int r2 = r | j;
          const uint32_t lo = key[r] < key[r2] ? key[r] : key[r2];
          const uint32_t hi = key[r] < key[r2] ? key[r2] : key[r];
          key[r] = lo;
          key[r2] = hi;
        }
      }
    }
  }
}

template <int PER, int SR, typename TP, typename TI, typename TV>
__device__ void sort_row(int64_t row, int lane, int64_t base, const TP* __restrict__ aptr,
                         const TI* __restrict__ aidx, const TV* __restrict__ aval,
                         const TP* __restrict__ bptr, const TI* __restrict__ bidx,
                         const TV* __restrict__ bval, const TP* __restrict__ cptr,
                         TI* __restrict__ cidx, TV* __restrict__ cval, WaveSortLds<PER, TV, TP>& L) {
  constexpr int N = SortShape<PER>::kCap;
  constexpr int PB = SortShape<PER>::kPBits;
  const int64_t a0 = pos_at(aptr, base, row), a1 = pos_at(aptr, base, row + 1);
  const int nl = static_cast<int>(a1 - a0);
  // 1. List offsets.
  constexpr int kRun = 32;
  int run = 0;
  bool long_list = false;
  for (int c = 0; c < nl; c += kWave) {
    const int t = c + lane;
    int len = 0;
    if (t < nl) {
      const int k = static_cast<int>(static_cast<int64_t>(aidx[a0 + t]) - base);
      const int64_t s0 = pos_at(bptr, base, k);
      len = static_cast<int>(pos_at(bptr, base, k + 1) - s0);
      L.u.in.bs[t] = static_cast<TP>(s0);
    }
    long_list = long_list || ballot(len > kRun) != 0;
    int tot;
    const int ex = team_scan<64>(len, lane, nullptr, &tot);
    if (t < nl) L.u.in.off[t] = run + ex;
    run += tot;
  }
  const int np = run;
  if (lane == 0) L.u.in.off[nl] = np;
  team_sync<64>();
  // Each B row arrives sorted by column (and its products by index), so when
  // every list fits a 32-key block and the blocks fit the network, list t is
  // laid out from element 32*t, padded with keys that sort last.
  // (Only the 256-key sort: the 512-key one measured slower with the
  // second layout, 88 -> 96 us on rows of 257..512 products.)
  const bool blocked = PER == 4 && nl <= N / kRun && !long_list;
  // 2. This lane's products (padding keys sort last).  Values stay in LDS
  // under their product index, which the key carries: the sort moves keys
  // only.
  // The lane's products are consecutive: one search finds the first one's
  // list, and each next product steps forward past the lists it has left
  // (off[nl] = np > p stops the walk).  Positions are found first, then all
  // of the lane's loads are issued, then used: the loads of one product do
  // not wait for those of the one before.
  int tl[PER], pl[PER];
  int64_t ql[PER];
  if (PER == 4 && blocked) {
#pragma unroll
    for (int r = 0; r < PER; r++) {
      const int e = PER * lane + r;
      const int t = e / kRun, j = e % kRun;
      const int o = t < nl ? L.u.in.off[t] : 0;
      const bool in = t < nl && o + j < L.u.in.off[t + 1];
      tl[r] = t;
      pl[r] = in ? o + j : np;
      ql[r] = in ? static_cast<int64_t>(L.u.in.bs[t]) + j : 0;
    }
  } else {
    int t = PER * lane < np ? list_of(L.u.in.off, nl, PER * lane) : 0;
#pragma unroll
    for (int r = 0; r < PER; r++) {
      const int p = PER * lane + r;
      if (p < np && r > 0)
        while (L.u.in.off[t + 1] <= p) t++;
      tl[r] = t;
      pl[r] = p;
      ql[r] = static_cast<int64_t>(L.u.in.bs[t]) + (p - L.u.in.off[t]);
    }
  }
  TI bi[PER];
  TV bv[PER], av[PER];
#pragma unroll
  for (int r = 0; r < PER; r++) {
    if (pl[r] < np) {
      bi[r] = bidx[ql[r]];
      bv[r] = bval[ql[r]];
      av[r] = aval[a0 + tl[r]];
    }
  }
  uint32_t key[PER];
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int p = pl[r];
    key[r] = 0xffffffffu;
    if (p < np) {
      const uint32_t col = static_cast<uint32_t>(static_cast<int64_t>(bi[r]) - base);
      key[r] = (col << PB) | static_cast<uint32_t>(p);
      L.pv[p] = sr_mul<TV, SR>(av[r], bv[r]);
    }
  }
  // 3. Bitonic sort of the N keys, ascending.  Laid out in sorted 32-key
  // blocks, the keys skip the stages that sort those blocks.
  if (PER != 4 || !blocked) bitonic_stages<PER, 2, kRun>(key, lane);
  bitonic_stages<PER, 2 * kRun, N>(key, lane);
  // 4. Fold each column's run in order and write it at its rank.  The
  // sorted keys overwrite the product lists: every lane's reads of those
  // are complete (their values fed the sort) before the wave's stores.
  team_sync<64>();
#pragma unroll
  for (int r = 0; r < PER; r++) L.u.sk[PER * lane + r] = key[r];
  team_sync<64>();
  constexpr uint32_t kPMask = (1u << PB) - 1;
  const int64_t o = pos_at(cptr, base, row);
  int seen = 0;
#pragma unroll
  for (int r = 0; r < PER; r++) {
    // Rounds of 64 consecutive sorted positions.
    const int e = r * kWave + lane;
    const uint32_t col = L.u.sk[e] >> PB;
    const bool head = e < np && (e == 0 || (L.u.sk[e - 1] >> PB) != col);
    // Heads before this lane in the round: a ballot and a masked popcount.
    const uint64_t heads = ballot(head);
    const int ex = popc(heads & ((uint64_t(1) << lane) - 1));
    const int tot = popc(heads);
    if (head) {
      TV w = sr_zero<TV, SR>();
      for (int x = e; x < np && (L.u.sk[x] >> PB) == col; x++)
        w = sr_add<TV, SR>(w, L.pv[L.u.sk[x] & kPMask]);
      cidx[o + seen + ex] = static_cast<TI>(static_cast<int64_t>(col) + base);
      cval[o + seen + ex] = w;
    }
    seen += tot;
  }
}

// Four wave rows per 256-thread block.
template <int PER, int SR, typename TP, typename TI, typename TV>
__global__ __launch_bounds__(256) void spgemm_wave_sort_kernel(
    const int64_t* __restrict__ rows, const int64_t* __restrict__ nrows_p, int64_t base,
    const TP* __restrict__ aptr, const TI* __restrict__ aidx, const TV* __restrict__ aval,
    const TP* __restrict__ bptr, const TI* __restrict__ bidx, const TV* __restrict__ bval,
    const TP* __restrict__ cptr, TI* __restrict__ cidx, TV* __restrict__ cval) {
  __shared__ WaveSortLds<PER, TV, TP> lds[4];
  const int w = wave_in_block();
  const int64_t nrows = *nrows_p;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + w; r < nrows;
       r += static_cast<int64_t>(gridDim.x) * 4) {
    sort_row<PER, SR>(rows[r], threadIdx.x & 63, base, aptr, aidx, aval, bptr, bidx, bval, cptr,
                      cidx, cval, lds[w]);
    team_sync<64>();  // the wave's LDS is reused by its next row
  }
}

// ---- hash teams (symbolic pass) ---------------------------------------------------------
constexpr int kWaveSlots = 512;    // 2 x kWaveCap: at most half full
constexpr int kBlockSlots = 4096;  // >= 2 x kBlockCap
constexpr int kEmptyKey = -1;

template <int SLOTS>
__device__ __forceinline__ int hash_slot(int key) {
  constexpr int kBits = __builtin_ctz(SLOTS);
  return static_cast<int>((static_cast<uint32_t>(key) * 2654435761u) >> (32 - kBits));
}

// Per-team LDS of a counted row: the table and the row's lists.
template <int CAP, int SLOTS, typename TP>
struct HashLds {
  int key[SLOTS];
  int off[CAP + 1];
  TP bs[CAP];  // 0-based start of each list's B row
  int tmp[8];
};

// cnt[row] = the number of distinct columns of row `row` of A*B, counted by
// a team of NT lanes: each product's column is inserted into the LDS table
// (linear probing by compare-and-swap) and the lane whose insert fills an
// empty slot counts it.
template <int NT, int CAP, int SLOTS, typename TP, typename TI>
__device__ void hash_count_row(int64_t row, int tt, int64_t base, const TP* __restrict__ aptr,
                               const TI* __restrict__ aidx, const TP* __restrict__ bptr,
                               const TI* __restrict__ bidx, int64_t* __restrict__ cnt,
                               HashLds<CAP, SLOTS, TP>& L) {
  for (int i = tt; i < SLOTS; i += NT) L.key[i] = kEmptyKey;
  const int64_t a0 = pos_at(aptr, base, row), a1 = pos_at(aptr, base, row + 1);
  const int nl = static_cast<int>(a1 - a0);
  int run = 0;
  for (int c = 0; c < nl; c += NT) {
    const int t = c + tt;
    int len = 0;
    if (t < nl) {
      const int k = static_cast<int>(static_cast<int64_t>(aidx[a0 + t]) - base);
      const int64_t s0 = pos_at(bptr, base, k);
      len = static_cast<int>(pos_at(bptr, base, k + 1) - s0);
      L.bs[t] = static_cast<TP>(s0);
    }
    int tot;
    const int ex = team_scan<NT>(len, tt, L.tmp, &tot);
    if (t < nl) L.off[t] = run + ex;
    run += tot;
  }
  const int np = run;
  if (tt == 0) L.off[nl] = np;
  team_sync<NT>();
  int fresh = 0;
  for (int p = tt; p < np; p += NT) {
    const int t = list_of(L.off, nl, p);
    const int64_t q = static_cast<int64_t>(L.bs[t]) + (p - L.off[t]);
    const int key = static_cast<int>(static_cast<int64_t>(bidx[q]) - base);
    int slot = hash_slot<SLOTS>(key);
    for (;;) {
      const int prev = atomicCAS(&L.key[slot], kEmptyKey, key);
      if (prev == kEmptyKey || prev == key) {
        fresh += prev == kEmptyKey ? 1 : 0;
        break;
      }
      slot = (slot + 1) & (SLOTS - 1);
    }
  }
  int distinct;
  team_scan<NT>(fresh, tt, L.tmp, &distinct);
  if (tt == 0) cnt[row] = distinct;
}

// Wave hash teams: 4 rows per 256-thread block; CAP = 256 products with 512
// slots, or 512 products with 1024 slots (one wave per row instead of a
// 256-thread block for the rows of 257..512 products).
template <int CAP, int SLOTS, typename TP, typename TI>
__global__ __launch_bounds__(256) void spgemm_wave_count_kernel(
    const int64_t* __restrict__ rows, const int64_t* __restrict__ nrows_p, int64_t base,
    const TP* __restrict__ aptr, const TI* __restrict__ aidx, const TP* __restrict__ bptr,
    const TI* __restrict__ bidx, int64_t* __restrict__ cnt) {
  static_assert(SLOTS >= 2 * CAP, "hash table at most half full");
  __shared__ HashLds<CAP, SLOTS, TP> lds[4];
  const int w = wave_in_block();
  const int64_t nrows = *nrows_p;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + w; r < nrows;
       r += static_cast<int64_t>(gridDim.x) * 4) {
    hash_count_row<64, CAP, SLOTS>(rows[r], threadIdx.x & 63, base, aptr, aidx, bptr, bidx, cnt,
                                   lds[w]);
    team_sync<64>();  // the wave's LDS is reused by its next row
  }
}

template <typename TP, typename TI>
__global__ __launch_bounds__(256) void spgemm_block_count_kernel(
    const int64_t* __restrict__ rows, const int64_t* __restrict__ nrows_p, int64_t base,
    const TP* __restrict__ aptr, const TI* __restrict__ aidx, const TP* __restrict__ bptr,
    const TI* __restrict__ bidx, int64_t* __restrict__ cnt) {
  __shared__ HashLds<kBlockCap, kBlockSlots, TP> L;
  const int64_t nrows = *nrows_p;
  for (int64_t r = blockIdx.x; r < nrows; r += gridDim.x) {
    hash_count_row<256, kBlockCap, kBlockSlots>(rows[r], threadIdx.x, base, aptr, aidx, bptr,
                                                bidx, cnt, L);
    __syncthreads();  // the block's LDS is reused by its next row
  }
}

// ---- dense-window fallback ---------------------------------------------------------
template <typename TV, int SR>
__device__ __forceinline__ void lds_accumulate(TV* p, TV v) {
  if constexpr (SR == FGX_MIN_PLUS) {
    // Compare-and-swap until the stored value is min(stored, v).
    if constexpr (sizeof(TV) == 8) {
      unsigned long long* a = reinterpret_cast<unsigned long long*>(p);
      unsigned long long old = *a;
      for (;;) {
        const unsigned long long want = static_cast<unsigned long long>(
            __double_as_longlong(jl_min(__longlong_as_double(static_cast<long long>(old)), v)));
        if (want == old) break;
        const unsigned long long prev = atomicCAS(a, old, want);
        if (prev == old) break;
        old = prev;
      }
    } else {
      unsigned int* a = reinterpret_cast<unsigned int*>(p);
      unsigned int old = *a;
      for (;;) {
        const unsigned int want = __float_as_uint(jl_min(__uint_as_float(old), v));
        if (want == old) break;
        const unsigned int prev = atomicCAS(a, old, want);
        if (prev == old) break;
        old = prev;
      }
    }
  } else {
    atomicAdd(p, v);
  }
}

// acc = a*b + acc with the product rounded before the add, as the
// reference's `w[i] = A[i,k]*B[k,j] + w[i]` without a fused multiply-add.
template <typename TV>
__device__ __forceinline__ void add_rounded_product(TV* acc, TV a, TV b) {
#pragma clang fp contract(off)
  const TV prod = a * b;
  *acc = prod + *acc;
}

// One row of the dense-window fallback, by the whole block (the LDS window
// is the block's; every path through a window ends at a barrier, so the
// next row may clear it).
template <bool NUMERIC, int SR, typename TP, typename TI, typename TV>
__device__ void window_row(int64_t row, int64_t base, int64_t n_inner, const TP* __restrict__ aptr,
                           const TI* __restrict__ aidx, const TV* __restrict__ aval,
                           const TP* __restrict__ bptr, const TI* __restrict__ bidx,
                           const TV* __restrict__ bval, int64_t* __restrict__ cnt,
                           const TP* __restrict__ cptr, TI* __restrict__ cidx,
                           TV* __restrict__ cval, TV* acc, int* flag, int* tmp) {
  const int tt = threadIdx.x;
  const int64_t a0 = pos_at(aptr, base, row), a1 = pos_at(aptr, base, row + 1);
  int64_t seen = 0;
  for (int64_t w0 = 0; w0 < n_inner; w0 += kWindow) {
    const int64_t w1 = w0 + kWindow < n_inner ? w0 + kWindow : n_inner;
    for (int c = tt; c < kWindow; c += 256) {
      flag[c] = 0;
      acc[c] = sr_zero<TV, SR>();
    }
    __syncthreads();
    if constexpr (NUMERIC && SR == FGX_PLUS_TIMES && !std::is_integral<TV>::value) {
      // Floating-point sums in the reference's order (docs/examples/spgemm.jl
      // :59-76): A positions ascending, then each B row's positions.  One A
      // position per step, its B row's columns (distinct) spread over the
      // block, so every column's terms are added in that order.
      for (int64_t q = a0; q < a1; q++) {
        const int64_t k = static_cast<int64_t>(aidx[q]) - base;
        const TV a = aval[q];
        const int64_t b0 = pos_at(bptr, base, k), b1 = pos_at(bptr, base, k + 1);
        int64_t lo = b0, hi = b1;
        while (lo < hi) {
          const int64_t mid = lo + ((hi - lo) >> 1);
          if (static_cast<int64_t>(bidx[mid]) - base < w0) lo = mid + 1; else hi = mid;
        }
        for (int64_t p = lo + tt; p < b1; p += 256) {
          const int64_t j = static_cast<int64_t>(bidx[p]) - base;
          if (j >= w1) break;
          flag[j - w0] = 1;
          add_rounded_product(&acc[j - w0], a, bval[p]);
        }
        __syncthreads();
      }
    } else {
      // Counting, Int32 sums (wrapping adds commute) and min-plus (Julia's
      // `min` is order-free): every thread takes A positions in parallel.
      for (int64_t q = a0 + tt; q < a1; q += 256) {
        const int64_t k = static_cast<int64_t>(aidx[q]) - base;
        const TV a = NUMERIC ? aval[q] : TV(0);
        const int64_t b0 = pos_at(bptr, base, k), b1 = pos_at(bptr, base, k + 1);
        int64_t lo = b0, hi = b1;
        while (lo < hi) {
          const int64_t mid = lo + ((hi - lo) >> 1);
          if (static_cast<int64_t>(bidx[mid]) - base < w0) lo = mid + 1; else hi = mid;
        }
        for (int64_t p = lo; p < b1; p++) {
          const int64_t j = static_cast<int64_t>(bidx[p]) - base;
          if (j >= w1) break;
          flag[j - w0] = 1;
          if constexpr (NUMERIC) lds_accumulate<TV, SR>(&acc[j - w0], sr_mul<TV, SR>(a, bval[p]));
        }
      }
      __syncthreads();
    }
    // Ordered emit of the window's touched columns.
    for (int c = 0; c < kWindow; c += 256) {
      const int j = c + tt;
      const int f = (w0 + j < w1) ? flag[j] : 0;
      int tot;
      const int ex = team_scan<256>(f, tt, tmp, &tot);
      if constexpr (NUMERIC) {
        if (f) {
          const int64_t o = pos_at(cptr, base, row) + seen + ex;
          cidx[o] = static_cast<TI>(w0 + j + base);
          cval[o] = acc[j];
        }
      }
      seen += tot;
    }
    __syncthreads();
  }
  if constexpr (!NUMERIC) {
    if (tt == 0) cnt[row] = seen;
  }
}

template <bool NUMERIC, int SR, typename TP, typename TI, typename TV>
__global__ __launch_bounds__(256) void spgemm_window_kernel(
    const int64_t* __restrict__ rows, const int64_t* __restrict__ nrows_p, int64_t base,
    int64_t n_inner,
    const TP* __restrict__ aptr, const TI* __restrict__ aidx, const TV* __restrict__ aval,
    const TP* __restrict__ bptr, const TI* __restrict__ bidx, const TV* __restrict__ bval,
    int64_t* __restrict__ cnt, const TP* __restrict__ cptr, TI* __restrict__ cidx,
    TV* __restrict__ cval) {
  __shared__ TV acc[kWindow];
  __shared__ int flag[kWindow];
  __shared__ int tmp[8];
  const int64_t nrows = *nrows_p;
  for (int64_t rr = blockIdx.x; rr < nrows; rr += gridDim.x)
    window_row<NUMERIC, SR>(rows[rr], base, n_inner, aptr, aidx, aval, bptr, bidx, bval, cnt, cptr,
                            cidx, cval, acc, flag, tmp);
}

// Row bins by products: 0 up to 256 (wave), 1 up to 512 (wide wave), 2 up
// to 1536 (block), 3 beyond (dense window).  Slots are taken with one
// atomic per bin and 1024-row block: the waves' counts (ballots) are
// combined in LDS first.  One atomic per row, or per wave, on the four
// shared counters serialises the pass (one cache line takes every update).
// The order of rows within a bin does not affect any output.
constexpr int kBinThreads = 1024;
template <typename TP, typename TI>
__global__ __launch_bounds__(kBinThreads) void spgemm_bin_kernel(int64_t m, int64_t base,
                                                                 const TP* __restrict__ aptr,
                                                                 const TI* __restrict__ aidx,
                                                                 const TP* __restrict__ bptr,
                                                                 int64_t* __restrict__ bin_count,
                                                                 int64_t* __restrict__ bin_rows) {
  constexpr int kWaves = kBinThreads / kWave;
  __shared__ int wcnt[kWaves][4];
  __shared__ int64_t wbase[kWaves][4];
  const int64_t row = static_cast<int64_t>(blockIdx.x) * kBinThreads + threadIdx.x;
  const bool valid = row < m;
  int bin = 4;
  if (valid) {
    const int64_t a0 = pos_at(aptr, base, row), a1 = pos_at(aptr, base, row + 1);
    int64_t prod = 0;
    for (int64_t q = a0; q < a1 && prod <= kBlockCap; q++) {
      const int64_t k = static_cast<int64_t>(aidx[q]) - base;
      prod += pos_at(bptr, base, k + 1) - pos_at(bptr, base, k);
    }
    const int64_t key = prod > a1 - a0 ? prod : a1 - a0;
    bin = key <= kWaveCap ? 0 : (key <= kWide ? 1 : (key <= kBlockCap ? 2 : 3));
  }
  const int lane = lane_id();
  const int w = wave_in_block();
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint64_t mask[4];
#pragma unroll
  for (int b = 0; b < 4; b++) {
    mask[b] = ballot(bin == b);
    if (lane == 0) wcnt[w][b] = __popcll(mask[b]);
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int b = threadIdx.x;
    int tot = 0;
    for (int v = 0; v < kWaves; v++) tot += wcnt[v][b];
    int64_t first = 0;
    if (tot > 0)
      first = static_cast<int64_t>(atomicAdd(reinterpret_cast<unsigned long long*>(bin_count + b),
                                             static_cast<unsigned long long>(tot)));
    for (int v = 0; v < kWaves; v++) {
      wbase[v][b] = first;
      first += wcnt[v][b];
    }
  }
  __syncthreads();
  if (bin < 4)
    bin_rows[bin * m + wbase[w][bin] + __popcll(mask[bin] & below)] = row;
}

// The bins stay on the device: count[b] is read by the row kernels, which
// loop over their bin's rows, so no pass waits on the host for the sizes.
struct Bins {
  int64_t* rows;          // 4 * m
  const int64_t* count;   // 4, on the device
};

// Grids of the row kernels: enough blocks to fill the device several times
// over, at most one row (block kernels) or four rows (wave kernels) each.
constexpr int64_t kWaveRowBlocks = 8192;
constexpr int64_t kBlockRowBlocks = 2048;
inline unsigned wave_row_grid(int64_t m) {
  return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((m + 3) / 4, kWaveRowBlocks)));
}
inline unsigned block_row_grid(int64_t m) {
  return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(m, kBlockRowBlocks)));
}

template <typename TP, typename TI>
static int make_bins(const fgx_compressed* A, const fgx_compressed* B, int64_t* dev, Bins* bins,
                     hipStream_t s) {
  const int64_t m = A->n_outer;
  int64_t* bin_count = dev;
  bins->count = bin_count;
  bins->rows = dev + 4;
  FGX_HIP(hipMemsetAsync(bin_count, 0, 4 * sizeof(int64_t), s));
  if (m > 0) {
    hipLaunchKernelGGL((spgemm_bin_kernel<TP, TI>),
                       dim3(static_cast<unsigned>((m + kBinThreads - 1) / kBinThreads)),
                       dim3(kBinThreads), 0, s, m, static_cast<int64_t>(A->index_base),
                       static_cast<const TP*>(A->ptr), static_cast<const TI*>(A->idx),
                       static_cast<const TP*>(B->ptr), bin_count, bins->rows);
    FGX_TRY(launch_check("spgemm_bin_kernel"));
  }
  return FGX_OK;
}

template <bool NUMERIC, int SR, typename TP, typename TI, typename TV>
static int run_bins(const fgx_compressed* A, const fgx_compressed* B, const Bins& bins,
                    int64_t* cnt, const void* cptr, void* cidx, void* cval, hipStream_t s) {
  const int64_t m = A->n_outer;
  if (m == 0) return FGX_OK;
  const int64_t base = A->index_base;
  auto aptr = static_cast<const TP*>(A->ptr);
  auto aidx = static_cast<const TI*>(A->idx);
  auto aval = static_cast<const TV*>(A->val);
  auto bptr = static_cast<const TP*>(B->ptr);
  auto bidx = static_cast<const TI*>(B->idx);
  auto bval = static_cast<const TV*>(B->val);
  auto cp = static_cast<const TP*>(cptr);
  auto ci = static_cast<TI*>(cidx);
  auto cv = static_cast<TV*>(cval);
  const unsigned wg = wave_row_grid(m), bg = block_row_grid(m);
  if constexpr (!NUMERIC) {
    // Symbolic: distinct-column counts from the LDS hash teams.
    hipLaunchKernelGGL((spgemm_wave_count_kernel<kWaveCap, kWaveSlots, TP, TI>), dim3(wg),
                       dim3(256), 0, s, bins.rows, bins.count, base, aptr, aidx, bptr, bidx, cnt);
    FGX_TRY(launch_check("spgemm_wave_count_kernel"));
    hipLaunchKernelGGL((spgemm_wave_count_kernel<kWide, 2 * kWide, TP, TI>), dim3(wg), dim3(256), 0,
                       s, bins.rows + m, bins.count + 1, base, aptr, aidx, bptr, bidx, cnt);
    FGX_TRY(launch_check("spgemm_wave_count_kernel"));
    hipLaunchKernelGGL((spgemm_block_count_kernel<TP, TI>), dim3(bg), dim3(256), 0, s,
                       bins.rows + 2 * m, bins.count + 2, base, aptr, aidx, bptr, bidx, cnt);
    FGX_TRY(launch_check("spgemm_block_count_kernel"));
  } else {
    // Numeric: rows of up to 256 / 512 products by the register sort (when
    // the columns fit its key), the rest by the ordered merge.
    const bool sort4 = B->n_inner <= SortShape<4>::kMaxCols;
    const bool sort8 = B->n_inner <= SortShape<8>::kMaxCols;
    if (sort4) {
      hipLaunchKernelGGL((spgemm_wave_sort_kernel<4, SR, TP, TI, TV>), dim3(wg), dim3(256), 0, s,
                         bins.rows, bins.count, base, aptr, aidx, aval, bptr, bidx, bval, cp, ci,
                         cv);
      FGX_TRY(launch_check("spgemm_wave_sort_kernel"));
    } else {
      hipLaunchKernelGGL((spgemm_wave_kernel<NUMERIC, SR, TP, TI, TV>), dim3(wg), dim3(256), 0, s,
                         bins.rows, bins.count, base, aptr, aidx, aval, bptr, bidx, bval, cnt, cp,
                         ci, cv);
      FGX_TRY(launch_check("spgemm_wave_kernel"));
    }
    if (sort8) {
      hipLaunchKernelGGL((spgemm_wave_sort_kernel<8, SR, TP, TI, TV>), dim3(wg), dim3(256), 0, s,
                         bins.rows + m, bins.count + 1, base, aptr, aidx, aval, bptr, bidx, bval,
                         cp, ci, cv);
      FGX_TRY(launch_check("spgemm_wave_sort_kernel"));
    }
    const size_t lds = sizeof(TeamLds<kBlockCap, TV>);
    for (int b = sort8 ? 2 : 1; b <= 2; b++) {
      hipLaunchKernelGGL((spgemm_block_kernel<NUMERIC, SR, TP, TI, TV>), dim3(bg), dim3(256), lds,
                         s, bins.rows + b * m, bins.count + b, base, aptr, aidx, aval, bptr, bidx,
                         bval, cnt, cp, ci, cv);
      FGX_TRY(launch_check("spgemm_block_kernel"));
    }
  }
  // The dense-window rows are rare: a smaller grid keeps the launch cheap
  // when there are none.
  const unsigned window_grid = std::min<unsigned>(bg, 1024);
  hipLaunchKernelGGL((spgemm_window_kernel<NUMERIC, SR, TP, TI, TV>), dim3(window_grid), dim3(256),
                     0, s,
                     bins.rows + 3 * m, bins.count + 3, base, B->n_inner, aptr, aidx, aval, bptr,
                     bidx, bval, cnt, cp, ci, cv);
  FGX_TRY(launch_check("spgemm_window_kernel"));
  return FGX_OK;
}

static int check_gemm(int semiring, const fgx_compressed* A, const fgx_compressed* B,
                      const char* what) {
  FGX_TRY(check_compressed(A, what));
  FGX_TRY(check_compressed(B, what));
  if (A->n_inner != B->n_outer)
    return set_error(FGX_EDIM, "%s: A has %lld columns, B has %lld rows", what,
                     (long long)A->n_inner, (long long)B->n_outer);
  if (A->ptr_type != B->ptr_type || A->idx_type != B->idx_type || A->val_type != B->val_type ||
      A->index_base != B->index_base)
    return set_error(FGX_EUNSUP, "%s: operands must share ptr/idx/value types and index base", what);
  if (B->n_inner > INT32_MAX || A->n_inner > INT32_MAX)
    return set_error(FGX_EUNSUP, "%s: dimensions beyond Int32 range", what);
  if (semiring == FGX_PLUS_TIMES) {
    if (A->val_type != FGX_I32 && A->val_type != FGX_F32 && A->val_type != FGX_F64)
      return set_error(FGX_EUNSUP, "%s: plus-times value type %d", what, A->val_type);
  } else if (semiring == FGX_MIN_PLUS) {
    if (A->val_type != FGX_F32 && A->val_type != FGX_F64)
      return set_error(FGX_EUNSUP, "%s: min-plus needs F32 or F64 values", what);
  } else {
    return set_error(FGX_EINVAL, "%s: bad semiring %d", what, semiring);
  }
  return FGX_OK;
}

}  // namespace fgx

using namespace fgx;

extern "C" {

int fgx_spgemm_symbolic(fgx_ctx* ctx, int semiring, const fgx_compressed* A,
                        const fgx_compressed* B, void* c_ptr, int64_t* c_nnz, void* stream) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_spgemm_symbolic: NULL ctx");
  FGX_TRY(check_gemm(semiring, A, B, "fgx_spgemm_symbolic"));
  if (!c_ptr || !c_nnz) return set_error(FGX_EINVAL, "fgx_spgemm_symbolic: NULL output");
  IdxCombo ic;
  FGX_TRY(idx_combo(A, &ic));
  std::lock_guard<std::mutex> lk(ctx->mu);
  FGX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = pick_stream(ctx, stream);
  FGX_TRY(admit_operand(ctx, A, s, "fgx_spgemm_symbolic"));
  FGX_TRY(admit_operand(ctx, B, s, "fgx_spgemm_symbolic"));
  const int64_t m = A->n_outer;
  void* scr;
  const size_t need = static_cast<size_t>(4 + 4 * m + m + 1) * sizeof(int64_t) + scan_work_bytes(m);
  FGX_TRY(scratch(ctx, s, need, &scr));
  int64_t* binmem = static_cast<int64_t*>(scr);
  int64_t* cnt = binmem + 4 + 4 * m;
  void* work = cnt + m + 1;
  int64_t total = 0;
  FGX_DISPATCH_IDX(ic, {
    Bins bins;
    FGX_TRY((make_bins<TP, TI>(A, B, binmem, &bins, s)));
    // Structure does not depend on values: count with the FP32 instantiation.
    FGX_TRY((run_bins<false, FGX_PLUS_TIMES, TP, TI, float>(A, B, bins, cnt, nullptr, nullptr,
                                                            nullptr, s)));
    FGX_TRY(scan_counts_to_ptr<TP>(cnt, m, A->index_base, static_cast<TP*>(c_ptr), work,
                                   ctx->host_i64, &total, s));
  });
  *c_nnz = total;
  return FGX_OK;
}

int fgx_spgemm_numeric(fgx_ctx* ctx, int semiring, const fgx_compressed* A,
                       const fgx_compressed* B, const void* c_ptr, void* c_idx, void* c_val,
                       void* stream) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_spgemm_numeric: NULL ctx");
  FGX_TRY(check_gemm(semiring, A, B, "fgx_spgemm_numeric"));
  if (!c_ptr) return set_error(FGX_EINVAL, "fgx_spgemm_numeric: NULL c_ptr");
  IdxCombo ic;
  FGX_TRY(idx_combo(A, &ic));
  std::lock_guard<std::mutex> lk(ctx->mu);
  FGX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = pick_stream(ctx, stream);
  FGX_TRY(admit_operand(ctx, A, s, "fgx_spgemm_numeric"));
  FGX_TRY(admit_operand(ctx, B, s, "fgx_spgemm_numeric"));
  const int64_t m = A->n_outer;
  void* scr;
  FGX_TRY(scratch(ctx, s, static_cast<size_t>(4 + 4 * m) * sizeof(int64_t), &scr));
  int64_t* binmem = static_cast<int64_t*>(scr);
#define FGX_GEMM_NUM(SR_, TV_)                                                            \
  FGX_DISPATCH_IDX(ic, {                                                                  \
    Bins bins;                                                                            \
    FGX_TRY((make_bins<TP, TI>(A, B, binmem, &bins, s)));                                 \
    FGX_TRY((run_bins<true, SR_, TP, TI, TV_>(A, B, bins, nullptr, c_ptr, c_idx, c_val, s))); \
  })
  if (semiring == FGX_PLUS_TIMES) {
    switch (A->val_type) {
      case FGX_I32: FGX_GEMM_NUM(FGX_PLUS_TIMES, int32_t); break;
      case FGX_F32: FGX_GEMM_NUM(FGX_PLUS_TIMES, float); break;
      case FGX_F64: FGX_GEMM_NUM(FGX_PLUS_TIMES, double); break;
    }
  } else {
    switch (A->val_type) {
      case FGX_F32: FGX_GEMM_NUM(FGX_MIN_PLUS, float); break;
      case FGX_F64: FGX_GEMM_NUM(FGX_MIN_PLUS, double); break;
    }
  }
#undef FGX_GEMM_NUM
  return FGX_OK;
}

}  // extern "C"
