// fgx_batch.hpp — the batched persistent form: one launch runs the SpMV,
// SpMM and SDDMM work of many independent operands from a work queue.
//
// The host cuts every item into work units of similar cost (bytes moved),
// in one of three kinds:
//   ROWS  — rows [begin, end) of a SpMM/SpMV item, one lane group per row
//           (row-split: no carries, no tile table);
//   TILES — nnz tiles [begin, end) of a SpMM/SpMV item, merge-path tiles of
//           fgx_csr_tiles.hpp with carries finished by the fix-up launch;
//   SDDMM — 64-position sub-tiles [begin, end) of an SDDMM item.
// Units are grouped by class (SDDMM, SpMV, SpMM per lane-group width) and
// each class runs as one persistent kernel whose resident grid walks the
// class's units in a fixed stride, so the launches (and their hipGraph) are
// the same on every replay and results do not depend on timing.
#pragma once
#include <vector>
#include "fgx_csr_tiles.hpp"
#include "fgx_internal.hpp"

namespace fgx {

enum BatchKind : int32_t { BK_ROWS = 0, BK_TILES = 1, BK_SDDMM = 2 };
enum BatchOp : int32_t { BO_SPMM = 0, BO_SDDMM = 2 };

struct BatchDesc {
  const void* ptr;
  const void* idx;
  const void* val;
  const void* B;    // SpMM/SpMV: dense input (n x k); SDDMM: X (m x k); FP32 or BF16
  const void* B2;   // SDDMM: Y (n x k), same type as B
  float* C;         // SpMM/SpMV: output (m x k); SDDMM: output values (nnz)
  int64_t m, n, nnz, k, ldb, ldc;
  const int64_t* table; // tile table hint of TILES items (may be null)
  int64_t carry_tile0;  // first carry slot of this item (TILES items)
  int64_t carry_val0;   // first carry value of this item
  int32_t base;
  int32_t op;
  int32_t vec;     // 16-byte aligned ptr/idx/val runs (SpMV tiles)
  int32_t mfma;    // SDDMM on matrix cores (k % 4 == 0, aligned rows)
  int32_t kind;    // segmented launches: the BatchKind of this entry
  int32_t cpl;     // dense columns per lane of SpMM: 4 (FP32 rows), 8 (BF16 rows)
};

// Entries of a segmented launch passed by value as a kernel argument (when a
// class has at most kSegInline of them), so a wave finds its entry and
// descriptor without a memory round trip.  n == 0: read them from memory.
constexpr int kSegInline = 12;
struct SegInline {
  int64_t n;
  int64_t seg[kSegInline + 1];  // tiles before entry e
  BatchDesc desc[kSegInline];
};

struct BatchWork {
  int32_t item;
  int32_t kind;
  int64_t begin;
  int64_t end;
};

constexpr int kBatchThreads = 256;
constexpr int kSpmvRowGroup = 8;   // lanes per row, row-split SpMV

// Lanes per row (row-split) or per position (TILES) for a dense width k:
// 16-byte column slices of cpl columns (4 FP32 or 8 BF16 values), rounded up
// to a power of two.
__host__ __device__ inline int group_for_k(int64_t k, int cpl = 4) {
  const int64_t lanes = (k + cpl - 1) / cpl;
  int g = (cpl == 8 && lanes <= 1) ? 1 : 2;  // one 8-column lane covers k = 8
  while (g < lanes && g < 64) g <<= 1;
  return g;
}

// Dense columns per lane of SpMM for a width k and dense type: 8 (a lane
// reads 32 bytes of an FP32 row or 16 of a BF16 row) when k >= 16 is a
// multiple of 8, and for BF16 rows of width 8 (one lane per position); else
// 4.  Measured on the 4,096-item plan: the FP32 n_rhs=32 class 5.37 -> 3.43
// ms and the n_rhs=128 class 19.7 -> 17.0 ms with 8; at n_rhs=8 one 8-column
// lane per position is 9 % faster than two 4-column lanes for BF16 rows and
// 10 % slower for FP32 rows.
__host__ __device__ inline int cols_per_lane_for(int64_t k, int dense_type = FGX_F32) {
  if (k % 8 != 0) return 4;
  return (k >= 16 || (k == 8 && dense_type == FGX_BF16)) ? 8 : 4;
}



// Work-unit classes: the batch runs one persistent kernel per class, so each
// kernel carries the registers and LDS of one unit body.  SpMM classes are
// the lane-group width G (2..64), plus kClsWide for 8 columns per lane.
constexpr int kClsSddmm = 0;
constexpr int kClsSpmv = 1;
constexpr int kClsWide = 128;
constexpr int kMaxClasses = 16;
__host__ __device__ constexpr int cls_g(int cls) { return cls & (kClsWide - 1); }
__host__ __device__ constexpr int cls_kv(int cls) { return (cls & kClsWide) ? 8 : 4; }

__host__ __device__ inline int unit_class(const BatchDesc& d) {
  return d.op == BO_SDDMM
             ? kClsSddmm
             : (d.k == 1 ? kClsSpmv : group_for_k(d.k, d.cpl) | (d.cpl == 8 ? kClsWide : 0));
}

// Positions per lane run of a batched SpMV (width 1) tile.
#ifndef FGX_BATCH_SPMV_RUN
#define FGX_BATCH_SPMV_RUN 8
#endif
constexpr int kBatchSpmvRun = FGX_BATCH_SPMV_RUN;

// Positions per TILES tile for a width k.
__host__ __device__ inline int64_t tile_width_for_k(int64_t k, int cpl = 4) {
  if (k == 1) return kWave * kBatchSpmvRun;
  const int g = group_for_k(k, cpl);
  return static_cast<int64_t>(kWave / g) * tile_e_for_g(g);
}

// Carry values per tile of a width-k item: one per column slot of a group.
__host__ __device__ inline int fix_width_for_k(int64_t k, int cpl = 4) {
  return k == 1 ? 1 : group_for_k(k, cpl) * cpl;
}

}  // namespace fgx

// Host side of the plan (fgx_batch.hip).
struct fgx_batch_plan {
  int device = 0;
  int val_type = FGX_F32;
  int dense_type = FGX_F32;   // dense inputs: FGX_F32, or FGX_BF16 (with BF16 values)
  int combo = 0;
  int64_t count = 0;
  int64_t grid = 0;
  int num_cus = 0;
  // Segmented launches (the serving loop): each work entry is a request's
  // whole range and one wave runs each tile, so the hardware interleaves
  // the blocks of launches on different streams; batch plans use resident
  // (persistent) grids instead.  The staged arrays live in one device
  // buffer filled by one copy per launch.
  bool segmented = false;
  int64_t* d_segoff = nullptr;                    // view into d_stage
  int64_t cls_seg0[fgx::kMaxClasses] = {};        // class c's prefix array in d_segoff
  int64_t cls_total[fgx::kMaxClasses] = {};       // tiles of class c
  int64_t fix_seg0 = 0, fix_total = 0;            // the fix-up entries' prefix array
  fgx::SegInline cls_inl[fgx::kMaxClasses] = {};  // class c's entries, if few enough
  fgx::SegInline fix_inl = {};                     // the fix-up entries, if few enough
  void* d_stage = nullptr;
  void* h_stage = nullptr;                        // pinned
  int64_t cap_stage = 0, cap_hstage = 0;
  int64_t n_work = 0, n_rows = 0, n_tiles = 0, n_sddmm = 0;
  // Work units sorted by class: class cls_id[c] owns units [cls_w0[c], cls_w1[c]).
  int n_cls = 0;
  int cls_id[fgx::kMaxClasses] = {};
  int64_t cls_w0[fgx::kMaxClasses] = {}, cls_w1[fgx::kMaxClasses] = {};
  int64_t* d_tables = nullptr;  // tile tables of the plan's TILES items
  int64_t total_nnz = 0;
  int64_t alg_bytes = 0;
  fgx::BatchDesc* d_desc = nullptr;
  fgx::BatchWork* d_work = nullptr;
  fgx::BatchWork* d_fix = nullptr;  // TILES work units, for the fix-up launch
  int64_t n_fix = 0;
  int64_t* d_carry_row = nullptr;
  float* d_carry_val = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool use_graph = false;
  bool has_sddmm = false;
  int64_t cap_desc = 0, cap_work = 0, cap_fix = 0, cap_crow = 0, cap_cval = 0;
  hipStream_t capture_stream = nullptr;
  // Buffers replaced by a larger one while the plan is in use.  They are
  // freed with the plan, not at the replacement: hipFree waits for the whole
  // device, which would stall every other stream of a serving loop.
  std::vector<void*> retired_dev, retired_host;
};

namespace fgx {
// The autotuner's choice for one item: row-split (ROWS) or nnz-split
// (TILES), and rows or tiles per work unit.
struct TunePick {
  int kind;
  int64_t unit;
};
TunePick autotune_pick(int64_t m, int64_t nnz, int64_t k, int64_t max_row, int vbytes, int flags,
                       int cpl = 4);
// Uploads host descriptor/work arrays into p's (grown) device buffers on
// stream s and launches; the previous launch on p must have finished.
int batch_launch_staged(fgx_batch_plan* p, const BatchDesc* h_desc, int64_t nd,
                        BatchWork* h_work, int64_t nw, const BatchWork* h_fix, int64_t nf,
                        int64_t carry_tiles, int64_t carry_vals, bool has_sddmm, int num_cus,
                        hipStream_t s);
void batch_free(fgx_batch_plan* p);
// Sizes a plan's carry and staging buffers up front (for a launch of up to
// `entries` work entries, `carry_tiles` carries and `carry_vals` carry
// values), so that later launches allocate nothing: a reallocation frees
// device memory, and hipFree waits for the device.
int batch_reserve(fgx_batch_plan* p, int64_t entries, int64_t carry_tiles, int64_t carry_vals);
// Maximum row length of each compressed operand (one block per operand).
int max_row_lengths(const fgx_compressed* ops, int64_t n, int combo, int64_t* out, hipStream_t s);
// Sorts work units by class (stable, so each class keeps the host's order)
// and records the class ranges in p.
int order_by_class(fgx_batch_plan* p, const BatchDesc* desc, BatchWork* work, int64_t n);
// Builds (but does not launch) a plan's device tables from descriptors and
// work units already on the host; used by batch plans and the server.
int batch_launch(fgx_batch_plan* p, hipStream_t s);
int batch_upload(fgx_batch_plan* p, const BatchDesc* desc, int64_t n_desc, const BatchWork* work,
                 int64_t n_work, int64_t carry_tiles, int64_t carry_vals, int num_cus,
                 hipStream_t s);
}  // namespace fgx
