// fgx_csr_tiles.hpp — nnz-balanced ("merge-path") tiles over a CSR operand.
//
// The stored positions [0, nnz) are cut into equal wave tiles, so a wave's
// work does not depend on how the positions are spread over rows: a 100k-nnz
// row is split over many waves and a run of empty rows costs one store each.
// Within a tile, each lane (SpMV) or lane group (SpMM) walks a contiguous run
// of positions and the rows they belong to:
//
//   * a lane finalises every row whose end pos(r+1) falls in its run (a, b];
//     rows that end exactly at a run boundary, and empty rows, are finalised
//     by the run that contains that boundary, so every row is written once;
//   * the first row a run finalises may have started in earlier runs: its
//     value is held back and completed by a segmented scan of the runs'
//     carries (in position order), then stored;
//   * a row that started in an earlier *wave* is completed by a second, tiny
//     fix-up launch that adds the earlier waves' carries in wave order.
//
// Sums are therefore deterministic (no atomics) and independent of what
// else runs on the device, which the serving path relies on.
#pragma once
#include <type_traits>

#include "fgx_device.hpp"

namespace fgx {

constexpr int kWavesPerBlock = 4;
constexpr int kWinCap = 512;  // fibre ends cached in LDS per wave

// Positions per lane group of an SpMM tile, for G lanes per position (the
// batch and serving kernels use the same tiles).
#ifndef FGX_TILE_E_WIDE
#define FGX_TILE_E_WIDE 128  // positions per lane group for G >= 32
#endif
// One-lane groups take 8 positions each, so a tile (512 positions) stages no
// more in LDS than a two-lane group's.
__host__ __device__ constexpr int tile_e_for_g(int g) {
  return g >= 32 ? FGX_TILE_E_WIDE : (g >= 16 ? 64 : (g >= 4 ? 32 : (g >= 2 ? 16 : 8)));
}

// Fibres that a wave tile [q0, qend) touches: r_lo owns q0 (0 for the first
// tile), r_hi owns qend (the row a carry continues into; m if none).
template <typename TP>
__device__ __forceinline__ void tile_fibers(const TP* __restrict__ ptr, int64_t base, int64_t m,
                                            int64_t q0, int64_t qend, int64_t* r_lo,
                                            int64_t* r_hi) {
  const int lane = lane_id();
  const int64_t target = lane < 32 ? q0 : qend;
  int64_t r = group_find_fiber<32>(ptr, base, 0, m, target);
  int64_t lo = shfl(r, 0);
  int64_t hi = shfl(r, 32);
  *r_lo = q0 == 0 ? 0 : lo;
  *r_hi = hi;
}

// Fibres of a tile from a precomputed table (table[t] = the fibre owning
// position t*W, table[ntiles] = m), verified against ptr before use: the
// table is only a hint, so a stale one (ptr rewritten in place) costs a
// search, never a wrong answer.  Loads the LDS window and returns pos(r_lo).
template <typename TP>
__device__ __forceinline__ FiberWindow<TP> tile_setup(const TP* __restrict__ ptr, int64_t base,
                                                      int64_t m, int64_t q0, int64_t qend,
                                                      const int64_t* __restrict__ table,
                                                      int64_t tile, int64_t* lds,
                                                      int64_t* r_lo_out, int64_t* r_hi_out,
                                                      int64_t* pos_lo_out, int cap = kWinCap) {
  int64_t r_lo = 0, r_hi = m;
  bool good = false;
  FiberWindow<TP> win;
  if (table) {
    r_lo = tile == 0 ? 0 : table[tile];
    r_hi = table[tile + 1];
    good = r_lo >= 0 && r_lo <= r_hi && r_hi <= m && r_lo < m;
    if (good) {
      const int64_t r_last = r_hi < m ? r_hi : m - 1;
      const int64_t p_lo = pos_at(ptr, base, r_lo);
      win = load_window(ptr, base, r_lo, r_last - r_lo + 1, lds, cap);
      if (tile > 0) good = good && p_lo <= q0 && win.end_of(r_lo) > q0;
      if (r_hi < m) good = good && win.end_of(r_hi) > qend;
      else good = good && pos_at(ptr, base, m) <= qend;
      if (r_hi > r_lo) good = good && win.end_of(r_hi - 1) <= qend;
      *pos_lo_out = p_lo;
    }
  }
  if (!good) {
    tile_fibers(ptr, base, m, q0, qend, &r_lo, &r_hi);
    const int64_t r_last = r_hi < m ? r_hi : m - 1;
    *pos_lo_out = pos_at(ptr, base, r_lo);
    win = load_window(ptr, base, r_lo, r_last - r_lo + 1, lds, cap);
  }
  *r_lo_out = r_lo;
  *r_hi_out = r_hi;
  return win;
}

// Carry slot of a tile: -1 (no carry); r (the tile starts a run of tiles
// carrying into row r); -(r + 2) (the tile continues the run of the tile
// before it, i.e. the whole tile lies inside row r, which began earlier).
__device__ __forceinline__ int64_t encode_carry(bool carries, bool continues, int64_t r) {
  return carries ? (continues ? -(r + 2) : r) : -1;
}

// The end of an SpMV tile walk: a segmented inclusive scan of the runs'
// carries (a run that finalised a row starts a new segment) completes each
// run's held-back first row; the last lane leaves the tile's carry.
// `start_hi` = pos(r_hi) on the last lane, -1 when r_hi == r_lo.
template <typename TA>
__device__ __forceinline__ void spmv_runs_scan(TA acc, bool has_emit, int64_t first_row,
                                               TA first_val, TA* __restrict__ y, int64_t m,
                                               int64_t q0, int64_t qend, int64_t tile,
                                               int64_t r_lo, int64_t r_hi, int64_t pos_lo,
                                               int64_t start_hi, int64_t* carry_row_out,
                                               TA* carry_val_out) {
  const int lane = lane_id();
  TA v = acc;
  bool f = has_emit;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    TA v_up = shfl_up(v, d);
    int f_up = shfl_up(static_cast<int>(f), d);
    if (lane >= d) {
      if (!f) v = v_up + v;
      f = f || f_up;
    }
  }
  TA carry_in = shfl_up(v, 1);
  if (lane == 0) carry_in = 0;
  if (has_emit) y[first_row] = first_val + carry_in;

  if (lane == kWave - 1) {
    // Does the continuing row r_hi hold positions of this tile?
    const bool carries = r_hi < m && qend > q0 && start_hi < qend;
    const bool continues = tile > 0 && r_lo == r_hi && pos_lo < q0;
    *carry_row_out = encode_carry(carries, continues, r_hi);
    *carry_val_out = carries ? v : TA(0);
  }
}

// The walk of an SpMV tile over a fibre window too wide for LDS (read from
// ptr): finalise the rows that end in each lane's run, zero the tile's empty rows,
// complete held-back rows by a segmented scan, and leave the carry.
template <typename TA, int V, typename Win, typename TV>
__device__ __forceinline__ void spmv_tile_walk(const Win win, int64_t m, int64_t q0, int64_t qend,
                                               int64_t a, int64_t b, bool active,
                                               const TV (&vv)[V], const TA (&xg)[V],
                                               TA* __restrict__ y, int64_t tile, int64_t r_lo,
                                               int64_t r_hi, int64_t pos_lo,
                                               int64_t* carry_row_out, TA* carry_val_out) {
  const int lane = lane_id();
  const int64_t r_win_last = r_hi < m ? r_hi : m - 1;

  TA acc = 0;
  bool has_emit = false;
  int64_t first_row = -1;
  TA first_val = 0;
  int64_t r = active ? (a == 0 ? 0 : win.find(a)) : r_hi;
  // Rows that end inside the lane's run are stored as they end, except empty
  // rows (same end as the row before): those are zeroed below by the whole
  // wave, with coalesced stores, instead of one scattered store per lane.
  int64_t last_end = -1;
  // Finalise row r, which ends at e: the first row a run finalises is held
  // back for the scan, later ones are stored unless empty.  (r <= r_win_last
  // implies r < m.)
  auto finish = [&](int64_t e) {
    const bool store = has_emit && e != last_end;
    first_row = has_emit ? first_row : r;
    first_val = has_emit ? first_val : acc;
    has_emit = true;
    if (store) y[r] = acc;
    last_end = e;
    acc = 0;
    r++;
  };
  if (active) {
#pragma unroll
    for (int s = 0; s < V; s++) {
      const int64_t q = a + s;
      if (q < b) {
        int64_t e;
        while (r <= r_win_last && (e = win.end_of(r)) <= q) finish(e);
        acc += widen<TA>(vv[s]) * xg[s];
      }
    }
    int64_t e;
    while (r <= r_win_last && (e = win.end_of(r)) <= b) finish(e);
  }
  // Empty rows the tile owns: rows ending in (q0, qend] (from position 0 on
  // for the first tile), i.e. [r_lo, r_hi); row r is empty when it ends where
  // row r - 1 does (row r_lo of a later tile holds position q0).
  for (int64_t rr = r_lo + (tile > 0 ? 1 : 0) + lane; rr < r_hi && rr < m; rr += kWave) {
    const int64_t prev = rr == 0 ? 0 : win.end_of(rr - 1);
    if (win.end_of(rr) == prev) y[rr] = TA(0);
  }
  // pos(r_hi): where the continuing row starts (-1 when the tile lies in it).
  const int64_t start_hi =
      (lane == kWave - 1 && r_hi < m && qend > q0 && r_hi > r_lo) ? win.end_of(r_hi - 1) : -1;
  spmv_runs_scan(acc, has_emit, first_row, first_val, y, m, q0, qend, tile, r_lo, r_hi, pos_lo,
                 start_hi, carry_row_out, carry_val_out);
}

// The non-empty fibres of an LDS-cached window, packed in place over it:
// pk[c] = (end - q0) << 32 | (r - r0) for the c-th non-empty fibre r, ends
// clamped to W + 1 (they are only compared with positions of the tile).
// Empty fibres hold no position, so a run's walk never visits them.  The
// entries keep the window's type, so every access to the buffer (including
// the next tile's window, in the batched loop) is ordered as may-alias.
struct PackedRows {
  const int64_t* pk;
  int nc;
  int64_t r0;
  __device__ __forceinline__ int rel_end(int c) const { return static_cast<int>(pk[c] >> 32); }
  __device__ __forceinline__ int64_t row(int c) const { return r0 + (pk[c] & 0xFFFFFFFF); }
  // First c with rel_end(c) > t (nc if none).
  __device__ __forceinline__ int find(int t) const {
    int lo = 0, hi = nc;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (rel_end(mid) <= t) lo = mid + 1; else hi = mid;
    }
    return lo;
  }
};

// Packs the window `lds` (n <= kWinCap ends of fibres r0 .. r0+n-1) of the
// tile starting at q0 into PackedRows, and zeroes y over the empty fibres r0+t
// with t in [t_zero0, t_zero1) (the empty rows the tile owns).  pos_lo =
// pos(r0).  Returns in *start_hi the end of window entry t_hi (-1 if t_hi < 0),
// read before the packing overwrites it.
template <int W, typename TA>
__device__ __forceinline__ PackedRows pack_window(int64_t* lds, int64_t n, int64_t r0, int64_t q0,
                                                  int64_t pos_lo, int64_t t_zero0,
                                                  int64_t t_zero1, TA* __restrict__ y,
                                                  int64_t t_hi, int64_t* start_hi) {
  constexpr int R = kWinCap / kWave;
  const int lane = lane_id();
  const uint64_t below = (uint64_t{1} << lane) - 1;
  int64_t pkv[R];
  int slot[R];
  int nc = 0;
  // Read every entry (and its predecessor) before any packed write.
#pragma unroll
  for (int i = 0; i < R; i++) {
    slot[i] = -1;
    pkv[i] = 0;
    if (i * kWave < n) {
      const int t = i * kWave + lane;
      const bool in = t < n;
      const int64_t e = in ? lds[t] : 0;
      const int64_t p = in ? (t == 0 ? pos_lo : lds[t - 1]) : 0;
      const bool keep = in && e > p;
      if (in && !keep && t >= t_zero0 && t < t_zero1) y[r0 + t] = TA(0);
      const uint64_t bal = ballot(keep);
      const int64_t rel = e - q0 < W + 1 ? e - q0 : W + 1;
      pkv[i] = static_cast<int64_t>((static_cast<uint64_t>(rel) << 32) | static_cast<uint32_t>(t));
      slot[i] = keep ? nc + popc(bal & below) : -1;
      nc += popc(bal);
    }
  }
  *start_hi = t_hi >= 0 ? lds[t_hi] : -1;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int i = 0; i < R; i++)
    if (slot[i] >= 0) lds[slot[i]] = pkv[i];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return PackedRows{lds, nc, r0};
}

// The walk of an SpMV tile over its packed non-empty fibres: each lane
// finalises the rows that end in its run (a, b] (in tile-relative positions),
// and every row it visits holds positions, so each visit stores a result.
template <typename TA, int V, typename TV>
__device__ __forceinline__ void spmv_tile_walk_packed(
    const PackedRows rows, int64_t m, int64_t q0, int64_t qend, int64_t a, int64_t b, bool active,
    const TV (&vv)[V], const TA (&xg)[V], TA* __restrict__ y, int64_t tile, int64_t r_lo,
    int64_t r_hi, int64_t pos_lo, int64_t start_hi, int64_t* carry_row_out, TA* carry_val_out) {
  TA acc = 0;
  bool has_emit = false;
  int64_t first_row = -1;
  TA first_val = 0;
  const int ra = static_cast<int>(a - q0), rb = static_cast<int>(b - q0);
  int c = active ? rows.find(ra) : rows.nc;
  // Finalise row c: the first row a run finalises is held back for the scan.
  auto finish = [&]() {
    const int64_t r = rows.row(c);
    if (has_emit) y[r] = acc;
    first_row = has_emit ? first_row : r;
    first_val = has_emit ? first_val : acc;
    has_emit = true;
    acc = 0;
    c++;
  };
  if (active) {
#pragma unroll
    for (int s = 0; s < V; s++) {
      if (ra + s < rb) {
        while (c < rows.nc && rows.rel_end(c) <= ra + s) finish();
        acc += widen<TA>(vv[s]) * xg[s];
      }
    }
    while (c < rows.nc && rows.rel_end(c) <= rb) finish();
  }
  spmv_runs_scan(acc, has_emit, first_row, first_val, y, m, q0, qend, tile, r_lo, r_hi, pos_lo,
                 start_hi, carry_row_out, carry_val_out);
}

// ---------------------------------------------------------------------------
// SpMV tile: lane-level runs of V positions.  Writes y for every row the tile
// finalises and (carry_row, carry_val) of the row continuing into the next
// tile (carry_row = -1 when no positions of this tile belong to it).
template <typename TP, typename TI, typename TV, typename TA, int V, typename TX = TA>
__device__ __forceinline__ void spmv_tile(const TP* __restrict__ ptr, const TI* __restrict__ idx,
                                          const TV* __restrict__ val, int64_t base, int64_t m,
                                          int64_t nnz, const TX* __restrict__ x,
                                          TA* __restrict__ y, int64_t tile, bool vec,
                                          const int64_t* __restrict__ table, int64_t* lds_win,
                                          int64_t* carry_row_out, TA* carry_val_out) {
  constexpr int W = kWave * V;
  const int lane = lane_id();
  const int64_t q0 = tile * W;
  const int64_t qend = (q0 + W < nnz) ? q0 + W : nnz;
  const int64_t a = q0 + static_cast<int64_t>(lane) * V;
  const int64_t b = (a + V < qend) ? a + V : qend;
  const bool active = a < qend || (a == 0 && m > 0);

  // Stream the run's values and coordinates first: these loads depend only
  // on the tile index, so they are in flight while the tile's fibres are
  // located (the window set-up ends in a fence that would otherwise hold
  // them back).
  TV vv[V];
  TI jj[V];
  if (active && b - a == V) {
    ld_run<TV, V>(val + a, vec, vv);
    ld_run<TI, V>(idx + a, vec, jj);
  } else {
#pragma unroll
    for (int s = 0; s < V; s++) {
      const bool in = a + s < b;
      vv[s] = in ? ld_stream(val + a + s) : TV{};
      jj[s] = in ? ld_stream(idx + a + s) : TI{};
    }
  }
  // Gather x (independent loads, issued together), then locate the tile's
  // fibres while the gathers are in flight.  The gathers are unconditional:
  // positions past the run read x[0] (valid whenever the tile holds a
  // position) and are never accumulated.  A load under a per-position branch
  // makes the compiler wait for every earlier load before issuing the next.
  TA xg[V];
  if (qend > q0) {
#pragma unroll
    for (int s = 0; s < V; s++) {
      const bool in = a + s < b;
      const int64_t j = in ? static_cast<int64_t>(jj[s]) - base : 0;
      xg[s] = widen<TA>(x[j]);
    }
  } else {
#pragma unroll
    for (int s = 0; s < V; s++) xg[s] = TA(0);
  }
  int64_t r_lo, r_hi, pos_lo;
  FiberWindow<TP> win =
      tile_setup(ptr, base, m, q0, qend, table, tile, lds_win, &r_lo, &r_hi, &pos_lo);
  // The window is almost always cached in LDS.  That case packs it to the
  // non-empty rows (zeroing the empty ones the tile owns on the way), so a
  // run's walk steps only over rows that hold positions: on power-law
  // matrices most rows are empty, and a lane crossing a stretch of them would
  // hold its wave back one LDS read per row.
  if (win.cached) {
    const int64_t t_hi = (r_hi < m && qend > q0 && r_hi > r_lo) ? r_hi - 1 - r_lo : -1;
    const int64_t t_zero1 = (r_hi < m ? r_hi : m) - r_lo;
    int64_t start_hi;
    const PackedRows rows = pack_window<W, TA>(lds_win, win.n, r_lo, q0, pos_lo,
                                               tile > 0 ? 1 : 0, t_zero1, y, t_hi, &start_hi);
    spmv_tile_walk_packed<TA, V>(rows, m, q0, qend, a, b, active, vv, xg, y, tile, r_lo, r_hi,
                                 pos_lo, start_hi, carry_row_out, carry_val_out);
  } else
    spmv_tile_walk<TA, V>(win, m, q0, qend, a, b, active, vv, xg, y, tile, r_lo, r_hi, pos_lo,
                          carry_row_out, carry_val_out);
}
// ---------------------------------------------------------------------------
// SpMM tile: groups of G lanes, each lane owning KV consecutive dense columns
// of the [kofs, kofs + G*KV) slice; each group walks E consecutive positions.
// Rows of B and C are `ldb`/`ldc` apart (row-major).
template <int KV, typename TA>
struct accv {
  TA v[KV];
};

template <int KV, typename TA>
__device__ __forceinline__ void ld_dense(const TA* __restrict__ p, int valid_cols, TA (&o)[KV]) {
  if constexpr (KV == 4 && sizeof(TA) == 4) {
    if (valid_cols >= 4) {
      float4 t = *reinterpret_cast<const float4*>(p);
      o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
      return;
    }
  }
  if constexpr (KV == 8 && sizeof(TA) == 4) {
    if (valid_cols >= 8) {
      float4 t0 = *reinterpret_cast<const float4*>(p);
      float4 t1 = *reinterpret_cast<const float4*>(p + 4);
      o[0] = t0.x; o[1] = t0.y; o[2] = t0.z; o[3] = t0.w;
      o[4] = t1.x; o[5] = t1.y; o[6] = t1.z; o[7] = t1.w;
      return;
    }
  }
  if constexpr (KV == 2 && sizeof(TA) == 8) {
    if (valid_cols >= 2) {
      double2 t = *reinterpret_cast<const double2*>(p);
      o[0] = t.x; o[1] = t.y;
      return;
    }
  }
#pragma unroll
  for (int c = 0; c < KV; c++) o[c] = c < valid_cols ? p[c] : TA(0);
}

// A gather of KV dense input values kept in their storage form until use:
// converting right after each load would put the conversion (and a wait for
// that load) inside the load's branch, so consecutive gathers would no longer
// be in flight together.  ld_raw loads, widen_raw converts.
template <int KV, typename TD>
struct DenseRaw {
  TD v[KV];
};
template <int KV>
struct DenseRaw<KV, bf16_t> {
  uint32_t w[(KV + 1) / 2];  // two BF16 values per word, low half first
};

template <int KV, typename TD>
__device__ __forceinline__ void ld_raw(const TD* __restrict__ p, int valid_cols,
                                       DenseRaw<KV, TD>& o) {
#ifdef FGX_GATHER_NT
  // Measurement build only (cache-policy comparison): full 4-column slices
  // gathered with the non-temporal (streaming) policy instead of the default.
  if constexpr (KV == 4 && std::is_same<TD, float>::value) {
    if (valid_cols >= 4) {
      typedef float f4n __attribute__((ext_vector_type(4)));
      const f4n t = __builtin_nontemporal_load(reinterpret_cast<const f4n*>(p));
      o.v[0] = t[0]; o.v[1] = t[1]; o.v[2] = t[2]; o.v[3] = t[3];
      return;
    }
  }
  if constexpr (KV == 4 && std::is_same<TD, bf16_t>::value) {
    if (valid_cols >= 4) {
      typedef unsigned u2n __attribute__((ext_vector_type(2)));
      const u2n u = __builtin_nontemporal_load(reinterpret_cast<const u2n*>(p));
      o.w[0] = u[0];
      o.w[1] = u[1];
      return;
    }
  }
#endif
  if constexpr (std::is_same<TD, bf16_t>::value) {
    if constexpr (KV == 4) {
      if (valid_cols >= 4) {
        const uint2 u = *reinterpret_cast<const uint2*>(p);
        o.w[0] = u.x;
        o.w[1] = u.y;
        return;
      }
    }
    if constexpr (KV == 8) {
      if (valid_cols >= 8) {
        const uint4 u = *reinterpret_cast<const uint4*>(p);
        o.w[0] = u.x;
        o.w[1] = u.y;
        o.w[2] = u.z;
        o.w[3] = u.w;
        return;
      }
    }
#pragma unroll
    for (int c = 0; c < (KV + 1) / 2; c++) {
      const uint32_t lo = 2 * c < valid_cols ? p[2 * c].bits : 0u;
      const uint32_t hi = 2 * c + 1 < valid_cols ? p[2 * c + 1].bits : 0u;
      o.w[c] = lo | (hi << 16);
    }
  } else {
    ld_dense<KV, TD>(p, valid_cols, o.v);
  }
}

template <int KV, typename TD>
__device__ __forceinline__ void zero_raw(DenseRaw<KV, TD>& o) {
  if constexpr (std::is_same<TD, bf16_t>::value) {
#pragma unroll
    for (int c = 0; c < (KV + 1) / 2; c++) o.w[c] = 0u;
  } else {
#pragma unroll
    for (int c = 0; c < KV; c++) o.v[c] = TD(0);
  }
}

template <int KV, typename TD, typename TA>
__device__ __forceinline__ void widen_raw(const DenseRaw<KV, TD>& r, TA (&o)[KV]) {
  if constexpr (std::is_same<TD, bf16_t>::value) {
#pragma unroll
    for (int c = 0; c < KV; c++)
      o[c] = static_cast<TA>(__uint_as_float((c & 1) ? (r.w[c / 2] & 0xFFFF0000u)
                                                     : (r.w[c / 2] << 16)));
  } else {
#pragma unroll
    for (int c = 0; c < KV; c++) o[c] = static_cast<TA>(r.v[c]);
  }
}

template <int KV, typename TA>
__device__ __forceinline__ void st_dense(TA* __restrict__ p, int valid_cols, const TA (&o)[KV]) {
  if constexpr (KV == 4 && sizeof(TA) == 4) {
    if (valid_cols >= 4) {
      *reinterpret_cast<float4*>(p) = make_float4(o[0], o[1], o[2], o[3]);
      return;
    }
  }
  if constexpr (KV == 8 && sizeof(TA) == 4) {
    if (valid_cols >= 8) {
      *reinterpret_cast<float4*>(p) = make_float4(o[0], o[1], o[2], o[3]);
      *reinterpret_cast<float4*>(p + 4) = make_float4(o[4], o[5], o[6], o[7]);
      return;
    }
  }
  if constexpr (KV == 2 && sizeof(TA) == 8) {
    if (valid_cols >= 2) {
      *reinterpret_cast<double2*>(p) = make_double2(o[0], o[1]);
      return;
    }
  }
#pragma unroll
  for (int c = 0; c < KV; c++)
    if (c < valid_cols) p[c] = o[c];
}

// Output rows are written once and not read again by the kernel: stream
// them past the caches so they do not evict the gathered dense rows.
template <int KV, typename TA>
__device__ __forceinline__ void st_dense_nt(TA* __restrict__ p, int valid_cols, const TA (&o)[KV]) {
  if constexpr (KV == 4 && sizeof(TA) == 4) {
    if (valid_cols >= 4) {
      typedef float f4v __attribute__((ext_vector_type(4)));
      f4v t = {o[0], o[1], o[2], o[3]};
      __builtin_nontemporal_store(t, reinterpret_cast<f4v*>(p));
      return;
    }
  }
  if constexpr (KV == 8 && sizeof(TA) == 4) {
    if (valid_cols >= 8) {
      typedef float f4v __attribute__((ext_vector_type(4)));
      f4v t0 = {o[0], o[1], o[2], o[3]};
      f4v t1 = {o[4], o[5], o[6], o[7]};
      __builtin_nontemporal_store(t0, reinterpret_cast<f4v*>(p));
      __builtin_nontemporal_store(t1, reinterpret_cast<f4v*>(p + 4));
      return;
    }
  }
  if constexpr (KV == 2 && sizeof(TA) == 8) {
    if (valid_cols >= 2) {
      typedef double d2v __attribute__((ext_vector_type(2)));
      d2v t = {o[0], o[1]};
      __builtin_nontemporal_store(t, reinterpret_cast<d2v*>(p));
      return;
    }
  }
  st_dense<KV, TA>(p, valid_cols, o);
}

template <typename TP, typename TI, typename TV, typename TA, int G, int KV, int E, int U,
          typename TB = TA>
__device__ __forceinline__ void spmm_tile(const TP* __restrict__ ptr, const TI* __restrict__ idx,
                                          const TV* __restrict__ val, int64_t base, int64_t m,
                                          int64_t nnz, int64_t k, int64_t kofs,
                                          const TB* __restrict__ B, int64_t ldb,
                                          TA* __restrict__ C, int64_t ldc, int64_t tile,
                                          const int64_t* __restrict__ table, int64_t* lds_win,
                                          int64_t* carry_row_out,
                                          TA* carry_val_out /* G*KV entries */,
                                          TI* lds_j, TA* lds_v, bool b_off32) {
  constexpr int NG = kWave / G;
  constexpr int W = NG * E;
  static_assert(E % G == 0, "E must be a multiple of G");
  static_assert(G % U == 0 || U % G == 0, "U and G must nest");
  const int lane = lane_id();
  const int g = lane / G;
  const int gl = lane % G;
  const int64_t q0 = tile * W;
  const int64_t qend = (q0 + W < nnz) ? q0 + W : nnz;
  // The tile's coordinates and values, loaded coalesced before the fibre
  // set-up (so the loads overlap it) and parked in LDS afterwards; the gather
  // loop then reads them as LDS broadcasts, with no global round trip.
  constexpr int NS = W / kWave;
  TI sj[NS];
  TA sv[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int64_t q = q0 + s * kWave + lane;
    sj[s] = q < qend ? ld_stream(idx + q) : static_cast<TI>(base);
    sv[s] = q < qend ? widen<TA>(ld_stream(val + q)) : TA(0);
  }
  int64_t r_lo, r_hi, pos_lo;
  FiberWindow<TP> win =
      tile_setup(ptr, base, m, q0, qend, table, tile, lds_win, &r_lo, &r_hi, &pos_lo);
  const int64_t r_win_last = r_hi < m ? r_hi : m - 1;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    lds_j[s * kWave + lane] = sj[s];
    lds_v[s * kWave + lane] = sv[s];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  const int64_t col0 = kofs + static_cast<int64_t>(gl) * KV;
  const int valid_cols = static_cast<int>(col0 < k ? (k - col0 < KV ? k - col0 : KV) : 0);

  const int64_t a = q0 + static_cast<int64_t>(g) * E;
  const int64_t b = (a + E < qend) ? a + E : qend;
  const bool active = a < qend || (a == 0 && m > 0);

  TA acc[KV];
#pragma unroll
  for (int c = 0; c < KV; c++) acc[c] = 0;
  bool has_emit = false;
  int64_t first_row = -1;
  TA first_val[KV];
#pragma unroll
  for (int c = 0; c < KV; c++) first_val[c] = 0;
  int64_t r = active ? (a == 0 ? 0 : win.find(a)) : r_hi;

  auto emit = [&](int64_t row) {
    if (!has_emit) {
      has_emit = true;
      first_row = row;
#pragma unroll
      for (int c = 0; c < KV; c++) first_val[c] = acc[c];
    } else if (valid_cols > 0) {
      st_dense_nt<KV, TA>(C + row * ldc + col0, valid_cols, acc);
    }
#pragma unroll
    for (int c = 0; c < KV; c++) acc[c] = 0;
  };

  // The end of the row the walk is in, kept in a register so that a
  // position that ends no row costs one compare (no LDS read).  Groups
  // without positions never advance.
  constexpr int64_t kNoEnd = INT64_MAX;
  auto end_at = [&](int64_t row) {
    return (row < m && row <= r_win_last) ? win.end_of(row) : kNoEnd;
  };
  int64_t r_end = active ? end_at(r) : kNoEnd;
  auto advance = [&](int64_t q) {
    while (r_end <= q) {
      emit(r);
      r++;
      r_end = end_at(r);
    }
  };

  // Batches of U positions; the group's lanes read each position's
  // coordinate and value from LDS (broadcast) and gather their B slices.
  // When every lane's slice is whole and the tile holds positions, the
  // gathers take no branch (b_off32: every row of B starts within 4 GB of
  // B, so byte offsets fit 32 bits): padding past the tile end is staged as
  // coordinate `base` with value 0, so it reads B's first row, whose values
  // are then masked to 0 (an Inf or NaN there must not reach acc).  Only
  // padding positions lie past `b`, and only in the last tile, where every
  // row ending there is the tile's to emit.
  const TI* gj = lds_j + g * E;
  const TA* gv = lds_v + g * E;
  const bool whole = kofs + static_cast<int64_t>(G) * KV <= k && qend > q0 && b_off32;
#pragma unroll 1
  for (int e0 = 0; e0 < E; e0 += U) {
    DenseRaw<KV, TB> bv[U];
    TA ve[U];
    if (whole) {
      // Byte offsets from the slice's start fit 32 bits here, so each
      // gather takes one register for its address.
      const char* Bs = reinterpret_cast<const char*>(B + col0);
      const uint32_t row_bytes = static_cast<uint32_t>(ldb) * static_cast<uint32_t>(sizeof(TB));
#pragma unroll
      for (int u = 0; u < U; u++) {
        const uint32_t je = static_cast<uint32_t>(static_cast<int64_t>(gj[e0 + u]) - base);
        ve[u] = gv[e0 + u];
        ld_raw<KV, TB>(reinterpret_cast<const TB*>(Bs + je * row_bytes), KV, bv[u]);
      }
    } else {
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int64_t je = static_cast<int64_t>(gj[e0 + u]) - base;
        ve[u] = gv[e0 + u];
        const bool in = (a + e0 + u < b) && valid_cols > 0;
        if (in) ld_raw<KV, TB>(B + je * ldb + col0, valid_cols, bv[u]);
        else zero_raw<KV, TB>(bv[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t q = a + e0 + u;
      advance(q);
      TA bw[KV];
      widen_raw<KV, TB, TA>(bv[u], bw);
      const bool in = q < b;
#pragma unroll
      for (int c = 0; c < KV; c++) acc[c] += ve[u] * (in ? bw[c] : TA(0));
    }
  }
  advance(b);

  // Segmented scan of the groups' carries (same rule as spmv_tile).
  TA v[KV];
#pragma unroll
  for (int c = 0; c < KV; c++) v[c] = acc[c];
  bool f = has_emit;
#pragma unroll
  for (int d = G; d < kWave; d <<= 1) {
    int f_up = shfl_up(static_cast<int>(f), d);
    TA v_up[KV];
#pragma unroll
    for (int c = 0; c < KV; c++) v_up[c] = shfl_up(v[c], d);
    if (lane >= d) {
      if (!f) {
#pragma unroll
        for (int c = 0; c < KV; c++) v[c] = v_up[c] + v[c];
      }
      f = f || f_up;
    }
  }
  TA cin[KV];
#pragma unroll
  for (int c = 0; c < KV; c++) {
    cin[c] = shfl_up(v[c], G);
    if (lane < G) cin[c] = 0;
  }
  if (has_emit && valid_cols > 0) {
    TA o[KV];
#pragma unroll
    for (int c = 0; c < KV; c++) o[c] = first_val[c] + cin[c];
    st_dense_nt<KV, TA>(C + first_row * ldc + col0, valid_cols, o);
  }
  bool carries = false;
  if (r_hi < m && qend > q0) {
    const int64_t start_hi = (r_hi > r_lo) ? win.end_of(r_hi - 1) : -1;
    carries = start_hi < qend;
  }
  const bool continues = tile > 0 && r_lo == r_hi && pos_lo < q0;
  if (lane == kWave - 1) *carry_row_out = encode_carry(carries, continues, r_hi);
  if (g == NG - 1 && carries) {
#pragma unroll
    for (int c = 0; c < KV; c++) carry_val_out[gl * KV + c] = v[c];
  }
}

// Completes the row that a run of tiles carries into, if tile t starts such
// a run (carry slot >= 0).  Called by a whole wave.  The run's length is
// found 64 slots at a time with a ballot; then SpMV (width 1) sums the run's
// carries with lanes over tiles and a fixed shuffle tree, and SpMM sums them
// in tile order with lanes over columns.  Deterministic.
template <typename TA>
__device__ __forceinline__ void fixup_tile(const int64_t* __restrict__ carry_row,
                                           const TA* __restrict__ carry_val, int64_t ntiles,
                                           int64_t t, int width, TA* __restrict__ out, int64_t ld,
                                           int64_t kofs, int64_t k) {
  const int lane = lane_id();
  const int64_t r = carry_row[t];
  if (r < 0) return;
  const int64_t cont = -(r + 2);
  int64_t len = 1;
  for (int64_t c = t + 1;; c += kWave) {
    const int64_t t2 = c + lane;
    const bool in = t2 < ntiles && carry_row[t2] == cont;
    const uint64_t msk = ballot(in);
    const int run = (~msk == 0) ? kWave : __ffsll(static_cast<unsigned long long>(~msk)) - 1;
    len += run;
    if (run < kWave) break;
  }
  if (width == 1) {
    TA s = 0;
    for (int64_t i = lane; i < len; i += kWave) s += carry_val[t + i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
    if (lane == 0 && kofs < k) out[r * ld + kofs] += s;
  } else {
    for (int col = lane; col < width; col += kWave) {
      if (kofs + col >= k) continue;
      TA s = 0;
      // Unrolled so that eight loads are in flight; the adds stay in tile
      // order.
#pragma unroll 8
      for (int64_t i = 0; i < len; i++) s += carry_val[(t + i) * width + col];
      out[r * ld + kofs + col] += s;
    }
  }
}

// Carry fix-up over a chunk of tiles [t0, t1) of one operand, by the whole
// wave (every lane reaches each ballot).
//   Width 1 (SpMV): one lane per tile, up to 64 tiles.  A run of one tile,
//   the common case (a row split over two tiles), is completed by its lane
//   with one add.  Longer runs (a row spanning several tiles; about 8 per 64
//   tiles on a Zipf(1.5) matrix) are taken eight at a time by 8-lane
//   subgroups: each finds its run's length 8 tiles per step and sums the
//   carries lane-strided, then by a fixed xor tree.
//   Width > 1 (SpMM): 16 lanes per tile, up to 4 tiles; a one-tile run is
//   added by its 16 lanes (columns strided by 16), longer runs are taken by
//   the whole wave through fixup_tile.
// Each run is summed in a fixed order, so results do not depend on timing.
constexpr int kFixupGroup = 16;
__host__ __device__ inline int fixup_chunk_tiles(int width) {
  return width == 1 ? kWave : kWave / kFixupGroup;
}

template <typename TA>
__device__ __forceinline__ void fixup_chunk(const int64_t* __restrict__ carry_row,
                                            const TA* __restrict__ carry_val, int64_t ntiles,
                                            int64_t t0, int64_t t1, int width,
                                            TA* __restrict__ out, int64_t ld, int64_t kofs,
                                            int64_t k) {
  const int lane = lane_id();
  if (t1 > ntiles) t1 = ntiles;
  if (width == 1) {
    const int64_t t = t0 + lane;
    bool longrun = false;
    if (t < t1) {
      const int64_t r = carry_row[t];
      if (r >= 0) {
        if (t + 1 < ntiles && carry_row[t + 1] == -(r + 2)) longrun = true;
        else if (kofs < k) out[r * ld + kofs] += carry_val[t];
      }
    }
    uint64_t runs = ballot(longrun);
    const int sub = lane >> 3, sl = lane & 7;
    while (runs) {
      // Subgroup `sub` takes the sub-th remaining run, if there is one.
      uint64_t m = runs;
      for (int i = 0; i < sub && m; i++) m &= m - 1;
      const bool have = m != 0;
      const int l = have ? __ffsll(static_cast<unsigned long long>(m)) - 1 : 0;
      for (int i = 0; i < 8 && runs; i++) runs &= runs - 1;
      if (have) {
        const int64_t ts = t0 + l;
        const int64_t r = carry_row[ts];
        const int64_t cont = -(r + 2);
        int64_t len = 1;
        for (int64_t c = ts + 1;; c += 8) {
          const int64_t t2 = c + sl;
          const bool in = t2 < ntiles && carry_row[t2] == cont;
          const uint32_t msk = static_cast<uint32_t>(ballot(in) >> (sub * 8)) & 0xffu;
          const int run = msk == 0xffu ? 8 : __builtin_ctz(~msk);
          len += run;
          if (run < 8) break;
        }
        TA s = 0;
        for (int64_t i = sl; i < len; i += 8) s += carry_val[ts + i];
        s += __shfl_xor(s, 4, kWave);
        s += __shfl_xor(s, 2, kWave);
        s += __shfl_xor(s, 1, kWave);
        if (sl == 0 && kofs < k) out[r * ld + kofs] += s;
      }
    }
  } else {
    constexpr int GL = kFixupGroup;
    const int gl = lane % GL;
    const int64_t t = t0 + lane / GL;
    bool longrun = false;
    if (t < t1) {
      const int64_t r = carry_row[t];
      if (r >= 0) {
        if (t + 1 < ntiles && carry_row[t + 1] == -(r + 2)) {
          longrun = true;
        } else {
          for (int col = gl; col < width; col += GL)
            if (kofs + col < k) out[r * ld + kofs + col] += carry_val[t * width + col];
        }
      }
    }
    uint64_t runs = ballot(longrun && gl == 0);
    while (runs) {
      const int l = __ffsll(static_cast<unsigned long long>(runs)) - 1;
      runs &= runs - 1;
      fixup_tile<TA>(carry_row, carry_val, ntiles, t0 + l / GL, width, out, ld, kofs, k);
    }
  }
}

// Builds the tile table: table[t] = smallest r with pos(r+1) > t*W (m if
// none), for t in [0, ntiles]; table[ntiles] uses position nnz.
template <typename TP>
__global__ void tile_table_kernel(const TP* __restrict__ ptr, int64_t base, int64_t m,
                                  int64_t nnz, int64_t W, int64_t ntiles,
                                  int64_t* __restrict__ table) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  const int64_t q = t < ntiles ? t * W : nnz;
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (pos_at(ptr, base, mid + 1) <= q) lo = mid + 1; else hi = mid;
  }
  table[t] = lo;
}

}  // namespace fgx
