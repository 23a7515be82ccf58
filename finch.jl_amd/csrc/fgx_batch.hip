// fgx_batch.hip — batched persistent launches (fgx_batch_*) and the device
// side shared with the serving loop.
#include "fgx_batch_impl.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

namespace fgx {

// Carries of TILES units, finished by fixup_chunk (a chunk of 64 tiles per
// wave for SpMV items, 4 for SpMM items).  Persistent form (seg_off ==
// nullptr): one wave per unit, over the unit's chunks.  Segmented form:
// entries are whole requests' tile ranges, one wave per chunk.
__device__ __forceinline__ int item_fix_width(const BatchDesc& d) {
  return fix_width_for_k(d.k, d.cpl);
}

__device__ __forceinline__ void fixup_item_chunk(const BatchDesc& d, int64_t t0, int64_t t1,
                                                 const int64_t* __restrict__ carry_row,
                                                 const float* __restrict__ carry_val) {
  const int width = item_fix_width(d);
  const int64_t W = tile_width_for_k(d.k, d.cpl);
  const int64_t ntiles = d.nnz > 0 ? (d.nnz + W - 1) / W : 1;
  const int64_t ld = d.k == 1 ? 1 : d.ldc;
  fixup_chunk<float>(carry_row + d.carry_tile0, carry_val + d.carry_val0, ntiles, t0, t1, width,
                     d.C, ld, 0, d.k);
}

// Chunk `c` of a segmented fix-up entry covering tiles [begin, end).
__device__ __forceinline__ void fixup_entry_chunk(const BatchDesc& d, int64_t begin, int64_t end,
                                                  int64_t c, const int64_t* __restrict__ carry_row,
                                                  const float* __restrict__ carry_val) {
  const int64_t ct = fixup_chunk_tiles(item_fix_width(d));
  const int64_t t0 = begin + c * ct;
  const int64_t t1 = t0 + ct < end ? t0 + ct : end;
  fixup_item_chunk(d, t0, t1, carry_row, carry_val);
}

__global__ __launch_bounds__(kBatchThreads) void batch_fixup_kernel(
    const BatchDesc* __restrict__ desc, const BatchWork* __restrict__ fix, int64_t n_fix,
    const int64_t* __restrict__ seg_off, int64_t total, const int64_t* __restrict__ carry_row,
    const float* __restrict__ carry_val, SegInline inl) {
  const int w = wave_in_block();
  if (seg_off) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + w;
    if (c >= total) return;  // wave-uniform
    if (inl.n > 0) {  // few entries: offsets and descriptors are arguments
      int64_t u = 0;
      for (int64_t e = 1; e < inl.n; e++) u += inl.seg[e] <= c ? 1 : 0;
      // Inline fix-up entries cover whole requests: tiles [0, ntiles).
      const BatchDesc& d = inl.desc[u];
      const int64_t W = tile_width_for_k(d.k, d.cpl);
      const int64_t nt = d.nnz > 0 ? (d.nnz + W - 1) / W : 1;
      fixup_entry_chunk(d, 0, nt, c - inl.seg[u], carry_row, carry_val);
      return;
    }
    const int64_t u = find_segment(seg_off, n_fix, c);
    const BatchWork wk = fix[u];
    fixup_entry_chunk(desc[wk.item], wk.begin, wk.end, c - seg_off[u], carry_row, carry_val);
    return;
  }
  const int64_t nw = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  for (int64_t u = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + w; u < n_fix; u += nw) {
    const BatchWork wk = fix[u];
    const BatchDesc d = desc[wk.item];
    const int64_t ct = fixup_chunk_tiles(item_fix_width(d));
    for (int64_t t = wk.begin; t < wk.end; t += ct)
      fixup_item_chunk(d, t, t + ct < wk.end ? t + ct : wk.end, carry_row, carry_val);
  }
}

// The tile tables of a plan's TILES items in one launch: thread e builds
// entry e of the concatenated tables, where item slot j owns entries
// [off[j], off[j+1]) (its ntiles + 1 entries, laid out as the plan's tables).
// Entry t of an item is the fibre owning position t*W (nnz for the last),
// as tile_table_kernel computes it for one operand.
template <typename TP>
__global__ __launch_bounds__(256) void plan_tables_kernel(const BatchDesc* __restrict__ desc,
                                                          const int64_t* __restrict__ slots,
                                                          int64_t n, int64_t total,
                                                          int64_t* __restrict__ tables) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int64_t* off = slots + n;  // n + 1 offsets after the n item ids
  int64_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (off[mid] <= e) lo = mid; else hi = mid;
  }
  const BatchDesc& d = desc[slots[lo]];
  const int64_t W = tile_width_for_k(d.k, d.cpl);
  const int64_t ntiles = off[lo + 1] - off[lo] - 1;
  const int64_t t = e - off[lo];
  const int64_t q = t < ntiles ? t * W : d.nnz;
  const TP* ptr = static_cast<const TP*>(d.ptr);
  int64_t r0 = 0, r1 = d.m;
  while (r0 < r1) {
    const int64_t mid = r0 + ((r1 - r0) >> 1);
    if (pos_at(ptr, static_cast<int64_t>(d.base), mid + 1) <= q) r0 = mid + 1; else r1 = mid;
  }
  tables[e] = r0;
}

template <typename T>
static int dev_alloc_copy(T** out, const T* src, int64_t n, hipStream_t s);

static int build_plan_tables(fgx_batch_plan* p, const std::vector<int64_t>& ids,
                             const std::vector<int64_t>& off, int64_t total, hipStream_t s) {
  const int64_t n = static_cast<int64_t>(ids.size());
  if (n == 0 || total == 0) return FGX_OK;
  std::vector<int64_t> host(ids);
  host.insert(host.end(), off.begin(), off.end());
  int64_t* d = nullptr;
  FGX_TRY(dev_alloc_copy(&d, host.data(), static_cast<int64_t>(host.size()), s));
  const unsigned blocks = static_cast<unsigned>((total + 255) / 256);
  if (p->combo == IDX_64_64 || p->combo == IDX_64_32)
    hipLaunchKernelGGL((plan_tables_kernel<int64_t>), dim3(blocks), dim3(256), 0, s, p->d_desc, d,
                       n, total, p->d_tables);
  else
    hipLaunchKernelGGL((plan_tables_kernel<int32_t>), dim3(blocks), dim3(256), 0, s, p->d_desc, d,
                       n, total, p->d_tables);
  const int rc = launch_check("plan_tables_kernel");
  const hipError_t e = hipStreamSynchronize(s);
  (void)hipFree(d);
  FGX_TRY(rc);
  return e == hipSuccess ? FGX_OK : set_error(FGX_EHIP, "plan tables: %s", hipGetErrorString(e));
}

// Per-item row-length statistics for the autotuner: [0] max row length.
template <typename TP>
__global__ __launch_bounds__(256) void row_stats_kernel(const BatchDesc* __restrict__ desc,
                                                        int64_t* __restrict__ out) {
  __shared__ int64_t red[4];
  const BatchDesc d = desc[blockIdx.x];
  const TP* ptr = static_cast<const TP*>(d.ptr);
  int64_t mx = 0;
  for (int64_t r = threadIdx.x; r < d.m; r += blockDim.x) {
    const int64_t len = pos_at(ptr, d.base, r + 1) - pos_at(ptr, d.base, r);
    mx = len > mx ? len : mx;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int64_t y = __shfl_xor(mx, o, 64);
    mx = y > mx ? y : mx;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t m2 = red[0];
    for (int i = 1; i < 4; i++) m2 = red[i] > m2 ? red[i] : m2;
    out[blockIdx.x] = m2;
  }
}

// ---- host ----------------------------------------------------------------------------------
static int launch_raw(fgx_batch_plan* p, hipStream_t s) {
  for (int c = 0; c < p->n_cls; c++) {
    const int cls = p->cls_id[c];
    const int64_t w0 = p->cls_w0[c], w1 = p->cls_w1[c];
    if (w1 <= w0) continue;
    const int64_t* seg = p->segmented ? p->d_segoff + p->cls_seg0[c] : nullptr;
    const int64_t total = p->segmented ? p->cls_total[c] : 0;
    if (p->segmented && total == 0) continue;
    const SegInline* one = (p->segmented && p->cls_inl[c].n > 0) ? &p->cls_inl[c] : nullptr;
    int rc;
    if (p->dense_type == FGX_BF16) {
      if (p->val_type != FGX_BF16)
        return set_error(FGX_EUNSUP, "batch: BF16 dense inputs need BF16 values");
      rc = cls == kClsSddmm ? batch_sddmm_bf16_bf16in(p, cls, w0, w1, seg, total, one, s)
                            : batch_spmm_bf16_bf16in(p, cls, w0, w1, seg, total, one, s);
    } else if (p->val_type == FGX_F32)
      rc = cls == kClsSddmm ? batch_sddmm_f32(p, cls, w0, w1, seg, total, one, s)
                            : batch_spmm_f32(p, cls, w0, w1, seg, total, one, s);
    else if (p->val_type == FGX_BF16)
      rc = cls == kClsSddmm ? batch_sddmm_bf16(p, cls, w0, w1, seg, total, one, s)
                            : batch_spmm_bf16(p, cls, w0, w1, seg, total, one, s);
    else
      return set_error(FGX_EUNSUP, "batch: value type %d (F32 or BF16)", p->val_type);
    FGX_TRY(rc);
  }
  if (p->n_fix > 0 && (!p->segmented || p->fix_total > 0)) {
    const int64_t g = p->segmented
                          ? (p->fix_total + kWavesPerBlock - 1) / kWavesPerBlock
                          : std::min<int64_t>((p->n_fix + kWavesPerBlock - 1) / kWavesPerBlock, 4096);
    const bool inl = p->segmented && p->fix_inl.n > 0;
    hipLaunchKernelGGL(batch_fixup_kernel, dim3(static_cast<unsigned>(g)), dim3(kBatchThreads), 0,
                       s, p->d_desc, p->d_fix, p->n_fix,
                       p->segmented ? p->d_segoff + p->fix_seg0 : nullptr, p->fix_total,
                       p->d_carry_row, p->d_carry_val, inl ? p->fix_inl : SegInline{});
    FGX_TRY(launch_check("batch_fixup_kernel"));
  }
  return FGX_OK;
}

int batch_launch(fgx_batch_plan* p, hipStream_t s) {
  if (p->use_graph && p->exec) {
    FGX_HIP(hipGraphLaunch(p->exec, s));
    return FGX_OK;
  }
  return launch_raw(p, s);
}

static void free_plan_buffers(fgx_batch_plan* p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->capture_stream) (void)hipStreamDestroy(p->capture_stream);
  if (p->d_stage) {
    // Segmented launches: desc/work/fix/seg_off are views into d_stage.
    (void)hipFree(p->d_stage);
    p->d_desc = nullptr;
    p->d_work = nullptr;
    p->d_fix = nullptr;
  }
  if (p->h_stage) (void)hipHostFree(p->h_stage);
  p->d_stage = nullptr;
  p->h_stage = nullptr;
  p->d_segoff = nullptr;
  p->cap_stage = p->cap_hstage = 0;
  if (p->d_desc) (void)hipFree(p->d_desc);
  if (p->d_work) (void)hipFree(p->d_work);
  if (p->d_fix) (void)hipFree(p->d_fix);
  if (p->d_carry_row) (void)hipFree(p->d_carry_row);
  if (p->d_carry_val) (void)hipFree(p->d_carry_val);
  if (p->d_tables) (void)hipFree(p->d_tables);
  p->d_tables = nullptr;
  for (void* b : p->retired_dev) (void)hipFree(b);
  for (void* b : p->retired_host) (void)hipHostFree(b);
  p->retired_dev.clear();
  p->retired_host.clear();
  p->exec = nullptr;
  p->graph = nullptr;
  p->capture_stream = nullptr;
  p->d_desc = nullptr;
  p->d_work = nullptr;
  p->d_fix = nullptr;
  p->d_carry_row = nullptr;
  p->d_carry_val = nullptr;
}

template <typename T>
static int dev_alloc_copy(T** out, const T* src, int64_t n, hipStream_t s) {
  const size_t bytes = static_cast<size_t>(std::max<int64_t>(n, 1)) * sizeof(T);
  if (hipMalloc(reinterpret_cast<void**>(out), bytes) != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return set_error(FGX_EOOM, "batch: device allocation of %zu bytes failed", bytes);
  }
  if (n > 0 && src) FGX_HIP(hipMemcpyAsync(*out, src, static_cast<size_t>(n) * sizeof(T),
                                           hipMemcpyHostToDevice, s));
  return FGX_OK;
}

int order_by_class(fgx_batch_plan* p, const BatchDesc* desc, BatchWork* work, int64_t n) {
  std::stable_sort(work, work + n, [&](const BatchWork& a, const BatchWork& b) {
    return unit_class(desc[a.item]) < unit_class(desc[b.item]);
  });
  p->n_cls = 0;
  for (int64_t u = 0; u < n;) {
    const int cls = unit_class(desc[work[u].item]);
    int64_t v = u;
    while (v < n && unit_class(desc[work[v].item]) == cls) v++;
    if (p->n_cls >= kMaxClasses)
      return set_error(FGX_EUNSUP, "batch: more than %d unit classes", kMaxClasses);
    p->cls_id[p->n_cls] = cls;
    p->cls_w0[p->n_cls] = u;
    p->cls_w1[p->n_cls] = v;
    p->n_cls++;
    u = v;
  }
  return FGX_OK;
}

int batch_upload(fgx_batch_plan* p, const BatchDesc* desc, int64_t n_desc, const BatchWork* work_in,
                 int64_t n_work, int64_t carry_tiles, int64_t carry_vals, int num_cus,
                 hipStream_t s) {
  std::vector<BatchWork> work(work_in, work_in + n_work);
  FGX_TRY(order_by_class(p, desc, work.data(), n_work));
  std::vector<BatchWork> fix;
  for (int64_t u = 0; u < n_work; u++)
    if (work[static_cast<size_t>(u)].kind == BK_TILES) fix.push_back(work[static_cast<size_t>(u)]);
  FGX_TRY(dev_alloc_copy(&p->d_desc, desc, n_desc, s));
  FGX_TRY(dev_alloc_copy(&p->d_work, work.data(), n_work, s));
  FGX_TRY(dev_alloc_copy(&p->d_fix, fix.data(), static_cast<int64_t>(fix.size()), s));
  FGX_TRY(dev_alloc_copy<int64_t>(&p->d_carry_row, nullptr, carry_tiles, s));
  FGX_TRY(dev_alloc_copy<float>(&p->d_carry_val, nullptr, carry_vals, s));
  FGX_HIP(hipStreamSynchronize(s));
  p->n_work = n_work;
  p->n_fix = static_cast<int64_t>(fix.size());
  p->num_cus = num_cus;  // the launch sizes its persistent grid from this
  return FGX_OK;
}

template <typename T>
static int grow_device(fgx_batch_plan* p, T** buf, int64_t* cap, int64_t need) {
  if (need <= *cap && *buf) return FGX_OK;
  int64_t want = std::max<int64_t>(need, std::max<int64_t>(2 * *cap, 64));
  if (*buf) p->retired_dev.push_back(*buf);
  *buf = nullptr;
  *cap = 0;
  if (hipMalloc(reinterpret_cast<void**>(buf), static_cast<size_t>(want) * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(FGX_EOOM, "batch: device allocation failed");
  }
  *cap = want;
  return FGX_OK;
}

static size_t align16(size_t n) { return (n + 15) & ~static_cast<size_t>(15); }

int batch_launch_staged(fgx_batch_plan* p, const BatchDesc* h_desc, int64_t nd,
                        BatchWork* h_work, int64_t nw, const BatchWork* h_fix, int64_t nf,
                        int64_t carry_tiles, int64_t carry_vals, bool has_sddmm, int num_cus,
                        hipStream_t s) {
  FGX_TRY(grow_device(p, &p->d_carry_row, &p->cap_crow, carry_tiles));
  FGX_TRY(grow_device(p, &p->d_carry_val, &p->cap_cval, carry_vals));
  FGX_TRY(order_by_class(p, h_desc, h_work, nw));
  p->segmented = true;
  // Descriptors in entry order (entry u's at u, with its kind), so a wave
  // reads its descriptor with one load; fix-up entries follow the renumbering.
  std::vector<BatchDesc> sdesc(static_cast<size_t>(nw));
  std::vector<int32_t> renum(static_cast<size_t>(nd), -1);
  for (int64_t u = 0; u < nw; u++) {
    sdesc[static_cast<size_t>(u)] = h_desc[h_work[u].item];
    sdesc[static_cast<size_t>(u)].kind = h_work[u].kind;
    if (h_work[u].begin != 0)
      return set_error(FGX_EINVAL, "batch: segmented entries must start at 0");
    renum[static_cast<size_t>(h_work[u].item)] = static_cast<int32_t>(u);
    h_work[u].item = static_cast<int32_t>(u);
  }
  std::vector<BatchWork> sfix(h_fix, h_fix + nf);
  for (auto& f : sfix) {
    f.item = renum[static_cast<size_t>(f.item)];
    if (f.item < 0) return set_error(FGX_EINVAL, "batch: fix-up entry without a work entry");
  }
  h_desc = sdesc.data();
  nd = nw;
  h_fix = sfix.data();
  // Per-class prefix counts of the entries' tiles, then the fix-up's.
  std::vector<int64_t> seg;
  seg.reserve(static_cast<size_t>(nw + nf + p->n_cls + 1));
  for (int c = 0; c < p->n_cls; c++) {
    p->cls_seg0[c] = static_cast<int64_t>(seg.size());
    int64_t acc = 0;
    for (int64_t u = p->cls_w0[c]; u < p->cls_w1[c]; u++) {
      seg.push_back(acc);
      acc += h_work[u].end - h_work[u].begin;
    }
    seg.push_back(acc);
    p->cls_total[c] = acc;
  }
  // Classes (and fix-up lists) with few entries pass them as kernel arguments.
  for (int c = 0; c < p->n_cls; c++) {
    SegInline& in = p->cls_inl[c];
    const int64_t n = p->cls_w1[c] - p->cls_w0[c];
    in.n = n <= kSegInline ? n : 0;
    for (int64_t e = 0; e < in.n; e++) {
      in.seg[e] = seg[static_cast<size_t>(p->cls_seg0[c] + e)];
      in.desc[e] = h_desc[p->cls_w0[c] + e];
    }
  }
  p->fix_inl.n = nf <= kSegInline ? nf : 0;
  p->fix_seg0 = static_cast<int64_t>(seg.size());
  int64_t facc = 0;
  for (int64_t u = 0; u < nf; u++) {
    seg.push_back(facc);
    // One wave per chunk of the entry's tiles.
    const BatchDesc& fd = h_desc[h_fix[u].item];
    const int64_t ct = fixup_chunk_tiles(fix_width_for_k(fd.k, fd.cpl));
    facc += (h_fix[u].end - h_fix[u].begin + ct - 1) / ct;
  }
  seg.push_back(facc);
  p->fix_total = facc;
  for (int64_t e = 0; e < p->fix_inl.n; e++) {
    p->fix_inl.seg[e] = seg[static_cast<size_t>(p->fix_seg0 + e)];
    p->fix_inl.desc[e] = h_desc[h_fix[e].item];
  }
  // One pinned staging buffer, one copy: [desc | work | fix | seg_off].
  const size_t o_desc = 0;
  const size_t o_work = align16(o_desc + static_cast<size_t>(nd) * sizeof(BatchDesc));
  const size_t o_fix = align16(o_work + static_cast<size_t>(nw) * sizeof(BatchWork));
  const size_t o_seg = align16(o_fix + static_cast<size_t>(nf) * sizeof(BatchWork));
  const size_t bytes = o_seg + seg.size() * sizeof(int64_t);
  if (static_cast<int64_t>(bytes) > p->cap_hstage) {
    const size_t want = std::max(bytes, static_cast<size_t>(2 * p->cap_hstage) + 4096);
    if (p->h_stage) p->retired_host.push_back(p->h_stage);
    p->h_stage = nullptr;
    p->cap_hstage = 0;
    if (hipHostMalloc(&p->h_stage, want) != hipSuccess) {
      (void)hipGetLastError();
      p->h_stage = nullptr;
      return set_error(FGX_EOOM, "batch: pinned staging allocation failed");
    }
    p->cap_hstage = static_cast<int64_t>(want);
  }
  if (static_cast<int64_t>(bytes) > p->cap_stage) {
    const size_t want = std::max(bytes, static_cast<size_t>(2 * p->cap_stage) + 4096);
    if (p->d_stage) p->retired_dev.push_back(p->d_stage);
    p->d_stage = nullptr;
    p->cap_stage = 0;
    if (hipMalloc(&p->d_stage, want) != hipSuccess) {
      (void)hipGetLastError();
      p->d_stage = nullptr;
      return set_error(FGX_EOOM, "batch: staging allocation failed");
    }
    p->cap_stage = static_cast<int64_t>(want);
  }
  char* h = static_cast<char*>(p->h_stage);
  if (nd > 0) std::memcpy(h + o_desc, h_desc, static_cast<size_t>(nd) * sizeof(BatchDesc));
  if (nw > 0) std::memcpy(h + o_work, h_work, static_cast<size_t>(nw) * sizeof(BatchWork));
  if (nf > 0) std::memcpy(h + o_fix, h_fix, static_cast<size_t>(nf) * sizeof(BatchWork));
  std::memcpy(h + o_seg, seg.data(), seg.size() * sizeof(int64_t));
  // When every class (and the fix-up) passes its entries as arguments, the
  // kernels read nothing staged: skip the copy.
  bool all_inline = nf == 0 || p->fix_inl.n > 0;
  for (int c = 0; c < p->n_cls && all_inline; c++) all_inline = p->cls_inl[c].n > 0;
  if (!all_inline)
    FGX_HIP(hipMemcpyAsync(p->d_stage, p->h_stage, bytes, hipMemcpyHostToDevice, s));
  char* d = static_cast<char*>(p->d_stage);
  p->d_desc = reinterpret_cast<BatchDesc*>(d + o_desc);
  p->d_work = reinterpret_cast<BatchWork*>(d + o_work);
  p->d_fix = reinterpret_cast<BatchWork*>(d + o_fix);
  p->d_segoff = reinterpret_cast<int64_t*>(d + o_seg);
  p->n_work = nw;
  p->n_fix = nf;
  p->has_sddmm = has_sddmm;
  p->num_cus = num_cus;
  return launch_raw(p, s);
}

void batch_free(fgx_batch_plan* p) { free_plan_buffers(p); }

int batch_reserve(fgx_batch_plan* p, int64_t entries, int64_t carry_tiles, int64_t carry_vals) {
  FGX_TRY(grow_device(p, &p->d_carry_row, &p->cap_crow, carry_tiles));
  FGX_TRY(grow_device(p, &p->d_carry_val, &p->cap_cval, carry_vals));
  // Staging: descriptors, work and fix-up entries and the segment offsets
  // of `entries` entries (the layout of batch_launch_staged, with slack for
  // its 16-byte alignment).
  const size_t bytes = static_cast<size_t>(entries) *
                           (sizeof(BatchDesc) + 2 * sizeof(BatchWork) + 2 * sizeof(int64_t)) +
                       4096;
  if (static_cast<int64_t>(bytes) > p->cap_hstage) {
    if (p->h_stage) p->retired_host.push_back(p->h_stage);
    p->h_stage = nullptr;
    p->cap_hstage = 0;
    if (hipHostMalloc(&p->h_stage, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return set_error(FGX_EOOM, "batch: pinned staging allocation failed");
    }
    p->cap_hstage = static_cast<int64_t>(bytes);
  }
  if (static_cast<int64_t>(bytes) > p->cap_stage) {
    if (p->d_stage) p->retired_dev.push_back(p->d_stage);
    p->d_stage = nullptr;
    p->cap_stage = 0;
    if (hipMalloc(&p->d_stage, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return set_error(FGX_EOOM, "batch: staging allocation failed");
    }
    p->cap_stage = static_cast<int64_t>(bytes);
  }
  return FGX_OK;
}

int max_row_lengths(const fgx_compressed* ops, int64_t n, int combo, int64_t* out,
                    hipStream_t s) {
  if (n <= 0) return FGX_OK;
  std::vector<BatchDesc> desc(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; i++) {
    BatchDesc& d = desc[static_cast<size_t>(i)];
    d = BatchDesc{};
    d.ptr = ops[i].ptr;
    d.m = ops[i].n_outer;
    d.base = ops[i].index_base;
  }
  BatchDesc* dd = nullptr;
  int64_t* dstat = nullptr;
  int rc = dev_alloc_copy(&dd, desc.data(), n, s);
  if (rc == FGX_OK) rc = dev_alloc_copy<int64_t>(&dstat, nullptr, n, s);
  if (rc == FGX_OK) {
    for (int64_t c0 = 0; c0 < n; c0 += 65535) {
      const int64_t nb = std::min<int64_t>(65535, n - c0);
      if (combo == IDX_32_32)
        hipLaunchKernelGGL((row_stats_kernel<int32_t>), dim3(static_cast<unsigned>(nb)), dim3(256),
                           0, s, dd + c0, dstat + c0);
      else
        hipLaunchKernelGGL((row_stats_kernel<int64_t>), dim3(static_cast<unsigned>(nb)), dim3(256),
                           0, s, dd + c0, dstat + c0);
    }
    rc = launch_check("row_stats_kernel");
  }
  if (rc == FGX_OK && (hipMemcpyAsync(out, dstat, n * sizeof(int64_t), hipMemcpyDeviceToHost, s) !=
                           hipSuccess ||
                       hipStreamSynchronize(s) != hipSuccess))
    rc = set_error(FGX_EHIP, "row statistics copy failed");
  if (dd) (void)hipFree(dd);
  if (dstat) (void)hipFree(dstat);
  return rc;
}

// Picks row-split vs nnz-split and the work-unit size from the item's shape
// and row-length statistics.  Units aim at ~256 KB of traffic so that
// thousands of them spread over the device.
TunePick autotune_pick(int64_t m, int64_t nnz, int64_t k, int64_t max_row, int vbytes,
                              int flags, int cpl) {
  const double mean = m > 0 ? static_cast<double>(nnz) / static_cast<double>(m) : 0.0;
  // Measured on MI355X (graph replays of same-class items, row-split time /
  // tile time): row-split wins only for wide dense operands (k >= 128) over
  // even rows of mean length 12 and more: 0.92-0.96 on the batched workload's
  // Poisson items (mean 12-64), 0.77-0.95 on calibration sets with mean
  // 32-128; at mean 4-12 tiles win (1.07), at k = 32 by 1.26-1.58, and on
  // power-law rows by up to 40x, where one long row serialises a lane group.
  const bool even = static_cast<double>(max_row) <= 2.0 * mean + 16.0;
  bool tiles = !(k >= 128 && mean >= 12.0 && even);
  if (flags & FGX_BATCH_FORCE_ROWSPLIT) tiles = false;
  if (flags & FGX_BATCH_FORCE_NNZSPLIT) tiles = true;
  const double target = 256.0 * 1024.0;
  TunePick pk;
  if (tiles) {
    const double tile_bytes =
        static_cast<double>(tile_width_for_k(k, cpl)) * (4.0 + vbytes + 4.0 * static_cast<double>(k));
    pk.kind = BK_TILES;
    pk.unit = std::max<int64_t>(1, static_cast<int64_t>(target / tile_bytes));
  } else {
    const double row_bytes = mean * (4.0 + vbytes + 4.0 * static_cast<double>(k)) + 4.0 * k + 4.0;
    const int64_t per_iter = k == 1 ? kWave / kSpmvRowGroup : kWave / group_for_k(k, cpl);
    int64_t rows = static_cast<int64_t>(target / std::max(row_bytes, 1.0));
    rows = std::max<int64_t>(per_iter, (rows / per_iter) * per_iter);
    pk.kind = BK_ROWS;
    pk.unit = rows;
  }
  return pk;
}

}  // namespace fgx

using namespace fgx;

extern "C" {

// The plan of fgx_batch_plan_create, with the context's lock already held.
// force[i] (optional): BK_ROWS or BK_TILES overrides the autotuner's choice
// for item i (-1: the rule's); maxrow_in (optional): the items' longest rows.
static int plan_create(fgx_ctx* ctx, const fgx_batch_item* items, int64_t count, int flags,
                       const signed char* force, const int64_t* maxrow_in,
                       fgx_batch_plan** out) {
  *out = nullptr;
  hipStream_t s = ctx->stream;
  auto* p = new fgx_batch_plan();
  p->device = ctx->device;
  p->count = count;
  std::vector<BatchDesc> desc(static_cast<size_t>(count));
  IdxCombo combo = IDX_32_32;
  const bool bf16_in = (flags & FGX_BATCH_BF16_INPUTS) != 0;
  p->dense_type = bf16_in ? FGX_BF16 : FGX_F32;
  for (int64_t i = 0; i < count; i++) {
    const fgx_batch_item& it = items[i];
    int rc = check_compressed(&it.A, "fgx_batch_plan_create");
    IdxCombo ic;
    if (rc == FGX_OK) rc = idx_combo(&it.A, &ic);
    if (rc == FGX_OK && i == 0) {
      combo = ic;
      p->val_type = it.A.val_type;
    }
    if (rc == FGX_OK && (ic != combo || it.A.val_type != p->val_type))
      rc = set_error(FGX_EUNSUP, "fgx_batch_plan_create: items must share index and value types");
    if (rc == FGX_OK && it.A.val_type != FGX_F32 && it.A.val_type != FGX_BF16)
      rc = set_error(FGX_EUNSUP, "fgx_batch_plan_create: values must be F32 or BF16");
    if (rc == FGX_OK && (it.n_rhs < 1 || it.n_rhs > 256 || (it.n_rhs > 1 && it.n_rhs % 4 != 0)))
      rc = set_error(FGX_EUNSUP, "fgx_batch_plan_create: n_rhs must be 1 or a multiple of 4 up to 256");
    if (rc == FGX_OK && (!it.C || (it.A.nnz > 0 && !it.B)))
      rc = set_error(FGX_EINVAL, "fgx_batch_plan_create: NULL dense operand");
    if (rc == FGX_OK && bf16_in && it.A.val_type != FGX_BF16)
      rc = set_error(FGX_EUNSUP, "fgx_batch_plan_create: BF16 inputs need BF16 values");
    // Rows are read up to 32 bytes per lane (cols_per_lane_for).
    if (rc == FGX_OK && it.n_rhs > 1 && (!aligned16(it.C) || !aligned16(it.B)))
      rc = set_error(FGX_EUNSUP, "fgx_batch_plan_create: dense operands must be 16-byte aligned");
    if (rc != FGX_OK) {
      delete p;
      return rc;
    }
    BatchDesc& d = desc[static_cast<size_t>(i)];
    d = BatchDesc{};
    d.ptr = it.A.ptr;
    d.idx = it.A.idx;
    d.val = it.A.val;
    d.B = static_cast<const float*>(it.B);
    d.B2 = nullptr;
    d.C = static_cast<float*>(it.C);
    d.m = it.A.n_outer;
    d.n = it.A.n_inner;
    d.nnz = it.A.nnz;
    d.k = it.n_rhs;
    d.ldb = it.n_rhs;
    d.ldc = it.n_rhs;
    d.base = it.A.index_base;
    d.op = BO_SPMM;
    d.vec = aligned16(it.A.val) && aligned16(it.A.idx);
    d.mfma = 0;
    d.cpl = cols_per_lane_for(it.n_rhs, p->dense_type);
    d.table = nullptr;
  }
  p->combo = combo;
  // Row-length statistics (one block per item).
  std::vector<int64_t> maxrow(static_cast<size_t>(count), 0);
  if (maxrow_in) {
    std::copy(maxrow_in, maxrow_in + count, maxrow.begin());
  } else if (count > 0) {
    std::vector<fgx_compressed> ops(static_cast<size_t>(count));
    for (int64_t i = 0; i < count; i++) ops[static_cast<size_t>(i)] = items[i].A;
    int rc = max_row_lengths(ops.data(), count, combo, maxrow.data(), s);
    if (rc != FGX_OK) {
      delete p;
      return rc;
    }
  }
  // Work units.
  const int vbytes = elem_size(p->val_type);
  const int ibytes = elem_size(items && count ? items[0].A.idx_type : FGX_I32);
  std::vector<BatchWork> work;
  std::vector<char> is_tiles(static_cast<size_t>(count), 0);
  int64_t carry_tiles = 0, carry_vals = 0;
  // Units are no coarser than needed to give every resident wave of the
  // device work many times over (small plans get one-tile units).
  int64_t total_tiles = 0;
  for (int64_t i = 0; i < count; i++) {
    const BatchDesc& d = desc[static_cast<size_t>(i)];
    const int64_t W = tile_width_for_k(d.k, d.cpl);
    total_tiles += d.nnz > 0 ? (d.nnz + W - 1) / W : 1;
  }
  const int64_t unit_cap = std::max<int64_t>(
      1, total_tiles / std::max<int64_t>(1, static_cast<int64_t>(ctx->num_cus) * 256));
  for (int64_t i = 0; i < count; i++) {
    BatchDesc& d = desc[static_cast<size_t>(i)];
    p->total_nnz += d.nnz;
    p->alg_bytes += d.nnz * (vbytes + ibytes) + (d.m + 1) * elem_size(items[i].A.ptr_type) +
                    d.n * d.k * (bf16_in ? 2 : 4) + d.m * d.k * 4;
    if (d.m == 0) continue;
    int fl = flags;
    if (force && force[i] == BK_ROWS) fl = (fl & ~FGX_BATCH_FORCE_NNZSPLIT) | FGX_BATCH_FORCE_ROWSPLIT;
    if (force && force[i] == BK_TILES) fl = (fl & ~FGX_BATCH_FORCE_ROWSPLIT) | FGX_BATCH_FORCE_NNZSPLIT;
    TunePick pk =
        autotune_pick(d.m, d.nnz, d.k, maxrow[static_cast<size_t>(i)], vbytes, fl, d.cpl);
    if (pk.kind == BK_TILES) pk.unit = std::min(pk.unit, unit_cap);
    if (pk.kind == BK_TILES) {
      const int64_t W = tile_width_for_k(d.k, d.cpl);
      const int64_t nt = d.nnz > 0 ? (d.nnz + W - 1) / W : 1;
      const int width = fix_width_for_k(d.k, d.cpl);
      d.carry_tile0 = carry_tiles;
      d.carry_val0 = carry_vals;
      carry_tiles += nt;
      carry_vals += nt * width;
      for (int64_t t = 0; t < nt; t += pk.unit)
        work.push_back(BatchWork{static_cast<int32_t>(i), BK_TILES, t, std::min(nt, t + pk.unit)});
      is_tiles[static_cast<size_t>(i)] = 1;
      p->n_tiles++;
    } else {
      for (int64_t r = 0; r < d.m; r += pk.unit)
        work.push_back(BatchWork{static_cast<int32_t>(i), BK_ROWS, r, std::min(d.m, r + pk.unit)});
      p->n_rows++;
    }
  }
  // Tile tables of the TILES items, in one allocation.
  int64_t table_entries = 0;
  std::vector<int64_t> table_off(static_cast<size_t>(count), -1);
  for (int64_t i = 0; i < count; i++) {
    const BatchDesc& d = desc[static_cast<size_t>(i)];
    if (!is_tiles[static_cast<size_t>(i)]) continue;
    const int64_t W = tile_width_for_k(d.k, d.cpl);
    table_off[static_cast<size_t>(i)] = table_entries;
    table_entries += (d.nnz > 0 ? (d.nnz + W - 1) / W : 1) + 1;
  }
  int rc = FGX_OK;
  if (table_entries > 0) {
    if (hipMalloc(reinterpret_cast<void**>(&p->d_tables),
                  static_cast<size_t>(table_entries) * sizeof(int64_t)) != hipSuccess) {
      (void)hipGetLastError();
      p->d_tables = nullptr;
      rc = set_error(FGX_EOOM, "fgx_batch_plan_create: tile tables");
    }
    for (int64_t i = 0; i < count && rc == FGX_OK; i++) {
      const int64_t off = table_off[static_cast<size_t>(i)];
      if (off < 0) continue;
      desc[static_cast<size_t>(i)].table = p->d_tables + off;
    }
  }
  if (rc == FGX_OK)
    rc = batch_upload(p, desc.data(), count, work.data(), static_cast<int64_t>(work.size()),
                      carry_tiles, carry_vals, ctx->num_cus, s);
  // All tile tables in one launch (from the uploaded descriptors).
  if (rc == FGX_OK && table_entries > 0) {
    std::vector<int64_t> ids, offs;
    for (int64_t i = 0; i < count; i++) {
      if (table_off[static_cast<size_t>(i)] < 0) continue;
      ids.push_back(i);
      offs.push_back(table_off[static_cast<size_t>(i)]);
    }
    offs.push_back(table_entries);
    rc = build_plan_tables(p, ids, offs, table_entries, s);
  }
  if (rc != FGX_OK) {
    free_plan_buffers(p);
    delete p;
    return rc;
  }
  if (!(flags & FGX_BATCH_NO_GRAPH)) {
    // Capture the persistent launch (and its fix-up) once; replays reuse it.
    rc = FGX_OK;
    if (hipStreamCreateWithFlags(&p->capture_stream, hipStreamNonBlocking) != hipSuccess)
      rc = set_error(FGX_EHIP, "batch: capture stream");
    if (rc == FGX_OK && hipStreamBeginCapture(p->capture_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      rc = set_error(FGX_EHIP, "batch: begin capture");
    if (rc == FGX_OK) {
      int lrc = launch_raw(p, p->capture_stream);
      hipGraph_t g = nullptr;
      if (hipStreamEndCapture(p->capture_stream, &g) != hipSuccess || lrc != FGX_OK)
        rc = set_error(FGX_EHIP, "batch: end capture");
      p->graph = g;
    }
    if (rc == FGX_OK && hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess)
      rc = set_error(FGX_EHIP, "batch: graph instantiate");
    if (rc != FGX_OK) {
      free_plan_buffers(p);
      delete p;
      return rc;
    }
    p->use_graph = true;
  }
  *out = p;
  return FGX_OK;
}

// FGX_BATCH_MEASURE: the autotuner's choice, measured.  Items are bucketed
// by lane-group class and row statistics (mean row length by powers of two,
// even or skewed rows); for every bucket of even rows at k >= 32 (where
// row-split can win) both forms are timed on up to 512 of its items, each as
// a plan of its own (one warm-up and three timed runs, HIP events on the
// context's stream), and the faster form is used for the whole bucket.
// Skewed rows and narrow widths keep nnz-split tiles, which the rule's
// calibration shows winning there by 1.3-40x.  The timed runs write those
// items' C.
static int measure_choices(fgx_ctx* ctx, const fgx_batch_item* items, int64_t count, int flags,
                           std::vector<signed char>* force, std::vector<int64_t>* maxrow) {
  force->assign(static_cast<size_t>(count), -1);
  maxrow->assign(static_cast<size_t>(count), 0);
  if (count <= 0) return FGX_OK;
  IdxCombo combo;
  FGX_TRY(idx_combo(&items[0].A, &combo));
  std::vector<fgx_compressed> ops(static_cast<size_t>(count));
  for (int64_t i = 0; i < count; i++) ops[static_cast<size_t>(i)] = items[i].A;
  hipStream_t s = ctx->stream;
  FGX_TRY(max_row_lengths(ops.data(), count, combo, maxrow->data(), s));
  const int dense_type = (flags & FGX_BATCH_BF16_INPUTS) ? FGX_BF16 : FGX_F32;
  std::map<int64_t, std::vector<int64_t>> buckets;
  for (int64_t i = 0; i < count; i++) {
    const fgx_batch_item& it = items[i];
    const int64_t m = it.A.n_outer;
    if (it.n_rhs < 32 || m == 0 || it.A.nnz == 0) continue;
    const double mean = static_cast<double>(it.A.nnz) / static_cast<double>(m);
    if (static_cast<double>((*maxrow)[static_cast<size_t>(i)]) > 2.0 * mean + 16.0) continue;
    const int cpl = cols_per_lane_for(it.n_rhs, dense_type);
    const int64_t cls = group_for_k(it.n_rhs, cpl) | (cpl == 8 ? kClsWide : 0);
    const int64_t mb = static_cast<int64_t>(std::floor(std::log2(std::max(mean, 1.0))));
    buckets[cls * 64 + mb].push_back(i);
  }
  hipEvent_t e0, e1;
  FGX_HIP(hipEventCreate(&e0));
  FGX_HIP(hipEventCreate(&e1));
  int rc = FGX_OK;
  const int base = flags & ~(FGX_BATCH_FORCE_ROWSPLIT | FGX_BATCH_FORCE_NNZSPLIT | FGX_BATCH_MEASURE);
  for (auto it = buckets.begin(); it != buckets.end() && rc == FGX_OK; ++it) {
    const std::vector<int64_t>& members = it->second;
    const size_t ns = std::min<size_t>(members.size(), 512);
    std::vector<fgx_batch_item> sample(ns);
    std::vector<int64_t> smax(ns);
    for (size_t j = 0; j < ns; j++) {
      sample[j] = items[members[j]];
      smax[j] = (*maxrow)[static_cast<size_t>(members[j])];
    }
    float ms[2] = {0.f, 0.f};
    const int forms[2] = {FGX_BATCH_FORCE_NNZSPLIT, FGX_BATCH_FORCE_ROWSPLIT};
    for (int f = 0; f < 2 && rc == FGX_OK; f++) {
      fgx_batch_plan* sp = nullptr;
      rc = plan_create(ctx, sample.data(), static_cast<int64_t>(ns), base | forms[f] | FGX_BATCH_NO_GRAPH,
                       nullptr, smax.data(), &sp);
      if (rc == FGX_OK) rc = batch_launch(sp, s);
      if (rc == FGX_OK && hipEventRecord(e0, s) != hipSuccess) rc = set_error(FGX_EHIP, "measure: event");
      for (int r = 0; r < 3 && rc == FGX_OK; r++) rc = batch_launch(sp, s);
      if (rc == FGX_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                           hipEventElapsedTime(&ms[f], e0, e1) != hipSuccess))
        rc = set_error(FGX_EHIP, "measure: timing");
      if (sp) fgx_batch_plan_destroy(sp);
    }
    if (rc != FGX_OK) break;
    const signed char pick = ms[1] < ms[0] ? BK_ROWS : BK_TILES;
    for (int64_t i : members) (*force)[static_cast<size_t>(i)] = pick;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int fgx_batch_plan_create(fgx_ctx* ctx, const fgx_batch_item* items, int64_t count, int flags,
                          fgx_batch_plan** out) {
  clear_error();
  if (!ctx || !out) return set_error(FGX_EINVAL, "fgx_batch_plan_create: NULL argument");
  if (count < 0 || (count > 0 && !items)) return set_error(FGX_EINVAL, "fgx_batch_plan_create: bad items");
  *out = nullptr;
  std::lock_guard<std::mutex> lk(ctx->mu);
  FGX_HIP(hipSetDevice(ctx->device));
  if (!(flags & FGX_BATCH_MEASURE) ||
      (flags & (FGX_BATCH_FORCE_ROWSPLIT | FGX_BATCH_FORCE_NNZSPLIT)))
    return plan_create(ctx, items, count, flags, nullptr, nullptr, out);
  std::vector<signed char> force;
  std::vector<int64_t> maxrow;
  FGX_TRY(measure_choices(ctx, items, count, flags, &force, &maxrow));
  return plan_create(ctx, items, count, flags & ~FGX_BATCH_MEASURE, force.data(), maxrow.data(),
                     out);
}

int fgx_batch_plan_run(fgx_ctx* ctx, fgx_batch_plan* plan, void* stream) {
  clear_error();
  if (!ctx || !plan) return set_error(FGX_EINVAL, "fgx_batch_plan_run: NULL argument");
  FGX_HIP(hipSetDevice(ctx->device));
  return batch_launch(plan, pick_stream(ctx, stream));
}

int fgx_batch_plan_stats(const fgx_batch_plan* plan, int64_t* st) {
  if (!plan || !st) return set_error(FGX_EINVAL, "fgx_batch_plan_stats: NULL argument");
  st[0] = plan->n_work;
  st[1] = plan->n_rows;
  st[2] = plan->n_tiles;
  st[3] = plan->grid;
  st[4] = plan->total_nnz;
  st[5] = plan->alg_bytes;
  return FGX_OK;
}

int fgx_batch_plan_destroy(fgx_batch_plan* plan) {
  if (!plan) return FGX_OK;
  (void)hipSetDevice(plan->device);
  (void)hipDeviceSynchronize();
  free_plan_buffers(plan);
  delete plan;
  return FGX_OK;
}

int fgx_batch_run_loop(fgx_ctx* ctx, const fgx_batch_item* items, int64_t count, void* stream) {
  clear_error();
  if (!ctx || (count > 0 && !items)) return set_error(FGX_EINVAL, "fgx_batch_run_loop: NULL argument");
  for (int64_t i = 0; i < count; i++) {
    const fgx_batch_item& it = items[i];
    int rc = it.n_rhs == 1
                 ? fgx_spmv_csr(ctx, &it.A, it.B, it.C, stream)
                 : fgx_spmm_csr(ctx, &it.A, it.n_rhs, it.B, it.n_rhs, FGX_ROW_MAJOR, it.C, it.n_rhs,
                                FGX_ROW_MAJOR, stream);
    if (rc != FGX_OK) return rc;
  }
  return FGX_OK;
}

}  // extern "C"
