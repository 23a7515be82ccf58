// fgx_serve.hip — the serving loop: SpMV / SpMM / SDDMM requests against
// device-resident operands, replayed from an open-loop arrival trace.
//
// Continuous batching (mode 0).  One scheduler thread (the caller) runs:
//   * release: requests whose arrival time has passed join a FIFO queue;
//   * completion: finished launches (hipEventQuery) retire their requests,
//     giving back their slots, HBM footprint and output space;
//   * admission and launch: while a stream is idle, requests are admitted in
//     FIFO order as long as the in-flight slots and the HBM budget allow
//     (footprint = the bytes of a request's inputs and output), and fused
//     into one launch of the batch kernels (fgx_batch.hip) on that stream,
//     one wave per tile of every admitted request.  Several streams keep launches overlapping on the
//     device: requests that arrive while others run go into the next launch.
// Each request's work units depend only on its operand and width, never on
// the rest of the batch, so a result is bit-identical alone and in any mix.
//
// Plain path (mode 1): each request, in arrival order, is one library call
// followed by a stream synchronisation — the per-call GPU path.
#include "fgx_batch.hpp"
#include "fgx_sddmm_core.hpp"

#include <algorithm>
#include <chrono>
#include <list>
#include <thread>
#include <vector>

namespace fgx {

namespace {

using Clock = std::chrono::steady_clock;

#ifndef FGX_SERVE_STREAMS
#define FGX_SERVE_STREAMS 4
#endif
constexpr int kServeStreams = FGX_SERVE_STREAMS;
// A request whose algorithmic bytes exceed kLargeRequestBytes is large: it
// only ever opens a launch (small requests may follow it in that launch); a
// large request met while a launch is filling keeps its place for the next
// free stream, and the small requests behind it still join the launch.  A
// launch stops taking requests once it holds kLaunchBytes of work.
// Measured on the serving mix (p99 target 5 ms):
//   * launch cap: 32 MB saturated near 112k req/s, 128 MB near 130k, 512 MB
//     near 134k (fewer, larger launches under a deep queue); 2 GB no better;
//   * capping the launches that hold large requests at 1 or 2 of the 4
//     streams starved the large requests, whose latency sets the p99 (they
//     are 1.5 % of requests): ~110k / ~120k req/s;
//   * opening each launch with the oldest waiting large request: no change.
constexpr double kLargeRequestBytes = 32.0e6;
#ifndef FGX_SERVE_LAUNCH_MB
#define FGX_SERVE_LAUNCH_MB 512
#endif
constexpr double kLaunchBytes = FGX_SERVE_LAUNCH_MB * 1.0e6;

// First-fit allocator over a device arena, for request outputs the server owns.
class Arena {
 public:
  int init(size_t bytes) {
    size_ = bytes;
    if (bytes == 0) return FGX_OK;
    if (hipMalloc(&base_, bytes) != hipSuccess) {
      (void)hipGetLastError();
      base_ = nullptr;
      return set_error(FGX_EOOM, "serve: output arena of %zu bytes", bytes);
    }
    free_.push_back({0, bytes});
    return FGX_OK;
  }
  ~Arena() {
    if (base_) (void)hipFree(base_);
  }
  void* alloc(size_t bytes) {
    bytes = (bytes + 255) & ~static_cast<size_t>(255);
    for (size_t i = 0; i < free_.size(); i++) {
      if (free_[i].second >= bytes) {
        const size_t off = free_[i].first;
        free_[i].first += bytes;
        free_[i].second -= bytes;
        if (free_[i].second == 0) free_.erase(free_.begin() + static_cast<long>(i));
        return static_cast<char*>(base_) + off;
      }
    }
    return nullptr;
  }
  void release(void* p, size_t bytes) {
    bytes = (bytes + 255) & ~static_cast<size_t>(255);
    const size_t off = static_cast<size_t>(static_cast<char*>(p) - static_cast<char*>(base_));
    auto it = std::lower_bound(free_.begin(), free_.end(), std::make_pair(off, size_t(0)));
    it = free_.insert(it, {off, bytes});
    // Coalesce with the neighbours.
    size_t i = static_cast<size_t>(it - free_.begin());
    if (i + 1 < free_.size() && free_[i].first + free_[i].second == free_[i + 1].first) {
      free_[i].second += free_[i + 1].second;
      free_.erase(free_.begin() + static_cast<long>(i) + 1);
    }
    if (i > 0 && free_[i - 1].first + free_[i - 1].second == free_[i].first) {
      free_[i - 1].second += free_[i].second;
      free_.erase(free_.begin() + static_cast<long>(i));
    }
  }
  size_t size() const { return size_; }

 private:
  void* base_ = nullptr;
  size_t size_ = 0;
  std::vector<std::pair<size_t, size_t>> free_;
};

struct OperandInfo {
  int64_t max_row = 0;
  // Tile tables of the operand, one per tile width it is requested at.
  std::vector<std::pair<int64_t, int64_t*>> tables;  // (W, table)
  const int64_t* table_for(int64_t W) const {
    for (auto& t : tables)
      if (t.first == W) return t.second;
    return nullptr;
  }
};

struct Launch {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  int32_t seq = 0;
  std::vector<int64_t> reqs;
  fgx_batch_plan plan;
};


struct Sizes {
  int64_t in = 0, in2 = 0, out = 0;
  int64_t footprint() const { return in + in2 + out; }
};

// Bytes of a request's dense inputs (`dbytes` per element: 4 for FP32, 2
// for BF16 inputs) and of its FP32 output.
Sizes request_sizes(const fgx_request& r, const fgx_compressed& A, int dbytes) {
  Sizes z;
  const int64_t k = r.kind == FGX_REQ_SPMV ? 1 : r.k;
  if (r.kind == FGX_REQ_SDDMM) {
    z.in = A.n_outer * k * dbytes;
    z.in2 = A.n_inner * k * dbytes;
    z.out = A.nnz * 4;
  } else {
    z.in = A.n_inner * k * dbytes;
    z.out = A.n_outer * k * 4;
  }
  return z;
}

// Algorithmic bytes of one request (its operand read once, dense in/out once).
double request_cost(const fgx_request& r, const fgx_compressed& A, int vbytes, int dbytes) {
  const Sizes z = request_sizes(r, A, dbytes);
  return static_cast<double>(A.nnz) * (vbytes + 4) + 4.0 * (A.n_outer + 1) +
         static_cast<double>(z.footprint());
}

}  // namespace

}  // namespace fgx

using namespace fgx;

extern "C" int fgx_serve_run(fgx_ctx* ctx, const fgx_compressed* operands, int64_t n_operands,
                             const fgx_request* reqs, int64_t n_reqs, const fgx_serve_config* cfg,
                             fgx_request_result* results) {
  clear_error();
  if (!ctx || !cfg || (n_reqs > 0 && (!reqs || !results)) || (n_operands > 0 && !operands))
    return set_error(FGX_EINVAL, "fgx_serve_run: NULL argument");
  if (n_operands <= 0 && n_reqs > 0) return set_error(FGX_EINVAL, "fgx_serve_run: no operands");
  if (cfg->mode != 0 && cfg->mode != 1) return set_error(FGX_EINVAL, "fgx_serve_run: bad mode");
  // Dense inputs: FP32 (0 or FGX_F32), or BF16 with BF16 values in batched mode.
  const int dense_type = cfg->input_type == 0 ? FGX_F32 : cfg->input_type;
  if (dense_type != FGX_F32 && dense_type != FGX_BF16)
    return set_error(FGX_EUNSUP, "fgx_serve_run: input type %d (F32 or BF16)", cfg->input_type);
  if (dense_type == FGX_BF16 && cfg->mode != 0)
    return set_error(FGX_EUNSUP, "fgx_serve_run: BF16 inputs in batched mode only");
  const int dbytes = dense_type == FGX_BF16 ? 2 : 4;
  const int max_slots = std::max(1, cfg->max_slots);
  const int max_batch = std::max(1, cfg->max_batch);
  const int64_t budget = cfg->hbm_budget_bytes > 0 ? cfg->hbm_budget_bytes : INT64_MAX;
  // Operand checks: one value type and index combination for the whole set.
  IdxCombo combo = IDX_32_32;
  for (int64_t i = 0; i < n_operands; i++) {
    FGX_TRY(check_compressed(&operands[i], "fgx_serve_run"));
    IdxCombo ic;
    FGX_TRY(idx_combo(&operands[i], &ic));
    if (i == 0) combo = ic;
    if (ic != combo || operands[i].val_type != operands[0].val_type)
      return set_error(FGX_EUNSUP, "fgx_serve_run: operands must share index and value types");
    if (operands[i].val_type != FGX_F32 && operands[i].val_type != FGX_BF16)
      return set_error(FGX_EUNSUP, "fgx_serve_run: values must be F32 or BF16");
    if (dense_type == FGX_BF16 && operands[i].val_type != FGX_BF16)
      return set_error(FGX_EUNSUP, "fgx_serve_run: BF16 inputs need BF16 values");
  }
  for (int64_t r = 0; r < n_reqs; r++) {
    const fgx_request& q = reqs[r];
    if (q.operand < 0 || q.operand >= n_operands || q.kind < FGX_REQ_SPMV || q.kind > FGX_REQ_SDDMM)
      return set_error(FGX_EINVAL, "fgx_serve_run: request %lld: bad operand or kind", (long long)r);
    const int64_t k = q.kind == FGX_REQ_SPMV ? 1 : q.k;
    if (k < 1 || k > 256 || (k > 1 && k % 4 != 0))
      return set_error(FGX_EUNSUP, "fgx_serve_run: request %lld: width must be 1 or a multiple "
                                   "of 4 up to 256", (long long)r);
    if (!q.in || (q.kind == FGX_REQ_SDDMM && !q.in2))
      return set_error(FGX_EINVAL, "fgx_serve_run: request %lld: NULL input", (long long)r);
    // SpMM rows are read and written 16 bytes per lane (4 FP32 or 8 BF16
    // columns); BF16 SDDMM rows are read 8 bytes per 4 columns.
    if (q.kind == FGX_REQ_SPMM && (!aligned16(q.in) || !aligned16(q.out)))
      return set_error(FGX_EINVAL, "fgx_serve_run: request %lld: SpMM rows must be 16-byte "
                                   "aligned", (long long)r);
    if (dense_type == FGX_BF16 && q.kind == FGX_REQ_SDDMM &&
        ((reinterpret_cast<uintptr_t>(q.in) & 7u) || (reinterpret_cast<uintptr_t>(q.in2) & 7u)))
      return set_error(FGX_EINVAL, "fgx_serve_run: request %lld: BF16 inputs must be 8-byte "
                                   "aligned", (long long)r);
  }
  FGX_HIP(hipSetDevice(ctx->device));
  const int vbytes = n_operands > 0 ? elem_size(operands[0].val_type) : 4;

  // ---- set-up (untimed): streams, operand statistics, tile tables, arena ----
  std::vector<Launch> L(kServeStreams);
  int rc = FGX_OK;
  for (auto& l : L) {
    if (hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&l.done, hipEventDisableTiming) != hipSuccess)
      rc = set_error(FGX_EHIP, "fgx_serve_run: stream/event creation failed");
  }
  std::vector<OperandInfo> info(static_cast<size_t>(n_operands));
  std::vector<int64_t> maxrow(static_cast<size_t>(n_operands), 0);
  if (rc == FGX_OK) rc = max_row_lengths(operands, n_operands, combo, maxrow.data(), L[0].stream);
  // Tile tables for the widths each operand is requested at, where the
  // autotuner picks nnz-split tiles, and for SDDMM.
  if (rc == FGX_OK) {
    for (int64_t r = 0; r < n_reqs && rc == FGX_OK; r++) {
      const fgx_request& q = reqs[r];
      const fgx_compressed& A = operands[q.operand];
      OperandInfo& oi = info[static_cast<size_t>(q.operand)];
      oi.max_row = maxrow[static_cast<size_t>(q.operand)];
      const int64_t k = q.kind == FGX_REQ_SPMV ? 1 : q.k;
      int64_t W = 0;
      if (q.kind == FGX_REQ_SDDMM) {
        W = kWave * kSddmmSub;
      } else if (autotune_pick(A.n_outer, A.nnz, k, oi.max_row, vbytes, 0,
                               cols_per_lane_for(k, dense_type)).kind == BK_TILES) {
        W = tile_width_for_k(k, cols_per_lane_for(k, dense_type));
      }
      if (W > 0 && !oi.table_for(W)) {
        int64_t* t = nullptr;
        rc = build_tile_table(&A, W, L[0].stream, &t);
        if (rc == FGX_OK) oi.tables.push_back({W, t});
      }
    }
  }
  for (int64_t i = 0; i < n_operands; i++) info[static_cast<size_t>(i)].max_row = maxrow[static_cast<size_t>(i)];
  // Launch buffers sized before the clock starts: the largest single
  // request's carries, or the 64 largest of the small requests (those that
  // share launches) together, whichever is more.  A launch that needs more
  // grows its plan without freeing (fgx_batch_plan::retired_dev).
  if (rc == FGX_OK && cfg->mode == 0) {
    std::vector<std::pair<int64_t, int64_t>> need;  // small requests: (carry tiles, values)
    int64_t top_t = 0, top_v = 0;
    for (int64_t r = 0; r < n_reqs; r++) {
      const fgx_request& q = reqs[r];
      if (q.kind == FGX_REQ_SDDMM) continue;
      const fgx_compressed& A = operands[q.operand];
      const int64_t k = q.kind == FGX_REQ_SPMV ? 1 : q.k;
      const int cpl = cols_per_lane_for(k, dense_type);
      if (autotune_pick(A.n_outer, A.nnz, k, maxrow[static_cast<size_t>(q.operand)], vbytes, 0,
                        cpl)
              .kind != BK_TILES)
        continue;
      const int64_t W = tile_width_for_k(k, cpl);
      const int64_t nt = A.nnz > 0 ? (A.nnz + W - 1) / W : 1;
      const int64_t nv = nt * fix_width_for_k(k, cpl);
      top_t = std::max(top_t, nt);
      top_v = std::max(top_v, nv);
      if (request_cost(q, A, vbytes, dbytes) <= kLargeRequestBytes) need.push_back({nt, nv});
    }
    std::sort(need.begin(), need.end(),
              [](const std::pair<int64_t, int64_t>& a, const std::pair<int64_t, int64_t>& b) {
                return a.second > b.second;
              });
    int64_t ct = 0, cv = 0;
    for (size_t i = 0; i < need.size() && i < 64; i++) {
      ct += need[i].first;
      cv += need[i].second;
    }
    for (auto& l : L) {
      if (rc != FGX_OK) break;
      rc = batch_reserve(&l.plan, max_batch, std::max(ct, top_t), std::max(cv, top_v));
    }
  }
  // Output arena for requests without an `out` buffer.
  Arena arena;
  int64_t max_out = 0;
  bool needs_arena = false;
  for (int64_t r = 0; r < n_reqs; r++) {
    if (!reqs[r].out) {
      needs_arena = true;
      max_out = std::max(max_out, request_sizes(reqs[r], operands[reqs[r].operand], dbytes).out);
    }
  }
  if (rc == FGX_OK && needs_arena) {
    size_t want = static_cast<size_t>(std::min<int64_t>(budget, int64_t(8) << 30));
    want = std::max(want, static_cast<size_t>(max_out) + 256);
    rc = arena.init(want);
  }
  if (rc == FGX_OK && hipStreamSynchronize(L[0].stream) != hipSuccess)
    rc = set_error(FGX_EHIP, "fgx_serve_run: set-up failed");

  std::vector<int64_t> order(static_cast<size_t>(n_reqs));
  for (int64_t i = 0; i < n_reqs; i++) order[static_cast<size_t>(i)] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
  std::vector<void*> outp(static_cast<size_t>(n_reqs), nullptr);
  for (int64_t r = 0; r < n_reqs; r++) {
    results[r].admit_s = results[r].done_s = -1.0;
    results[r].status = FGX_OK;
    results[r].batch = -1;
  }

  const auto t0 = Clock::now();
  auto now_s = [&]() { return std::chrono::duration<double>(Clock::now() - t0).count(); };

  if (rc == FGX_OK && cfg->mode == 1) {
    // ---- plain path: one call per request, synchronised ----
    hipStream_t s = L[0].stream;
    int32_t seq = 0;
    for (size_t oi = 0; oi < order.size() && rc == FGX_OK; oi++) {
      const int64_t r = order[oi];
      const fgx_request& q = reqs[r];
      const fgx_compressed& A = operands[q.operand];
      while (now_s() < q.arrival_s) std::this_thread::yield();
      const Sizes z = request_sizes(q, A, dbytes);
      if (z.footprint() > budget) {
        results[r].status = FGX_EADMIT;
        results[r].admit_s = results[r].done_s = now_s();
        continue;
      }
      void* out = q.out;
      if (!out) out = arena.alloc(static_cast<size_t>(z.out));
      results[r].admit_s = now_s();
      if (q.kind == FGX_REQ_SPMV) rc = fgx_spmv_csr(ctx, &A, q.in, out, s);
      else if (q.kind == FGX_REQ_SPMM)
        rc = fgx_spmm_csr(ctx, &A, q.k, q.in, q.k, FGX_ROW_MAJOR, out, q.k, FGX_ROW_MAJOR, s);
      else
        rc = fgx_sddmm_csr(ctx, &A, q.k, q.in, q.k, FGX_ROW_MAJOR, q.in2, q.k, FGX_ROW_MAJOR, out, s);
      if (rc == FGX_OK && hipStreamSynchronize(s) != hipSuccess)
        rc = set_error(FGX_EHIP, "fgx_serve_run: plain call failed");
      results[r].done_s = now_s();
      results[r].batch = seq++;
      if (!q.out) arena.release(out, static_cast<size_t>(z.out));
    }
  } else if (rc == FGX_OK) {
    // ---- continuous batching ----
    std::list<int64_t> pending;
    size_t next = 0;
    int64_t finished = 0, inflight_reqs = 0, inflight_bytes = 0;
    int32_t seq = 0;
    std::vector<BatchDesc> hdesc;
    std::vector<BatchWork> hwork, hfix;
    while (finished < n_reqs && rc == FGX_OK) {
      bool progress = false;
      const double tn = now_s();
      while (next < order.size() && reqs[order[next]].arrival_s <= tn) {
        pending.push_back(order[next++]);
        progress = true;
      }
      // Completions.
      for (auto& l : L) {
        if (!l.busy) continue;
        const hipError_t e = hipEventQuery(l.done);
        if (e == hipErrorNotReady) continue;
        if (e != hipSuccess) {
          rc = set_error(FGX_EHIP, "fgx_serve_run: launch failed: %s", hipGetErrorString(e));
          break;
        }
        const double td = now_s();
        for (int64_t r : l.reqs) {
          results[r].done_s = td;
          const Sizes z = request_sizes(reqs[r], operands[reqs[r].operand], dbytes);
          inflight_bytes -= z.footprint();
          inflight_reqs--;
          if (!reqs[r].out) arena.release(outp[static_cast<size_t>(r)], static_cast<size_t>(z.out));
          finished++;
        }
        l.reqs.clear();
        l.busy = false;
        progress = true;
      }
      if (rc != FGX_OK) break;
      // Admission and launch.
      for (auto& l : L) {
        if (l.busy || pending.empty()) continue;
        hdesc.clear();
        hwork.clear();
        hfix.clear();
        int64_t carry_tiles = 0, carry_vals = 0;
        bool has_sddmm = false;
        double batch_cost = 0.0;
        auto it = pending.begin();

        while (it != pending.end() && static_cast<int>(l.reqs.size()) < max_batch &&
               inflight_reqs < max_slots) {
          const int64_t r = *it;
          const fgx_request& q = reqs[r];
          const fgx_compressed& A = operands[q.operand];
          const Sizes z = request_sizes(q, A, dbytes);
          if (z.footprint() > budget || (!q.out && static_cast<size_t>(z.out) + 256 > arena.size())) {
            it = pending.erase(it);
            results[r].status = FGX_EADMIT;
            results[r].admit_s = results[r].done_s = now_s();
            finished++;
            continue;
          }
          if (inflight_bytes + z.footprint() > budget) break;
          const double cost = request_cost(q, A, vbytes, dbytes);
          if (!l.reqs.empty() && batch_cost > kLaunchBytes) break;
          // A large request opens a launch; met later, it waits for the
          // next free stream and the requests behind it go ahead.
          if (cost > kLargeRequestBytes && !l.reqs.empty()) {
            ++it;
            continue;
          }
          void* out = q.out;
          if (!out) {
            out = arena.alloc(static_cast<size_t>(z.out));
            if (!out) break;
          }
          it = pending.erase(it);
          outp[static_cast<size_t>(r)] = out;
          results[r].admit_s = now_s();
          results[r].batch = seq;
          inflight_reqs++;
          inflight_bytes += z.footprint();
          batch_cost += cost;
          l.reqs.push_back(r);
          // Descriptor and work units of this request.
          const OperandInfo& oi = info[static_cast<size_t>(q.operand)];
          const int64_t k = q.kind == FGX_REQ_SPMV ? 1 : q.k;
          BatchDesc d{};
          d.ptr = A.ptr;
          d.idx = A.idx;
          d.val = A.val;
          d.B = q.in;
          d.B2 = q.in2;
          d.C = static_cast<float*>(out);
          d.m = A.n_outer;
          d.n = A.n_inner;
          d.nnz = A.nnz;
          d.k = k;
          d.ldb = k;
          d.ldc = k;
          d.base = A.index_base;
          d.vec = aligned16(A.val) && aligned16(A.idx);
          const int cpl = cols_per_lane_for(k, dense_type);
          d.cpl = cpl;
          const int32_t item = static_cast<int32_t>(hdesc.size());
          if (q.kind == FGX_REQ_SDDMM) {
            d.op = BO_SDDMM;
            d.mfma = (k % 4 == 0) && aligned16(q.in) && aligned16(q.in2);
            d.table = oi.table_for(kWave * kSddmmSub);
            has_sddmm = true;
            // One entry per request: the launch runs a wave per SDDMM tile.
            const int64_t W = kWave * kSddmmSub;
            hwork.push_back(BatchWork{item, BK_SDDMM, 0, (A.nnz + W - 1) / W});
          } else {
            d.op = BO_SPMM;
            const TunePick pk = autotune_pick(A.n_outer, A.nnz, k, oi.max_row, vbytes, 0, cpl);
            if (pk.kind == BK_TILES) {
              const int64_t W = tile_width_for_k(k, cpl);
              const int64_t nt = A.nnz > 0 ? (A.nnz + W - 1) / W : 1;
              const int width = fix_width_for_k(k, cpl);
              d.table = oi.table_for(W);
              d.carry_tile0 = carry_tiles;
              d.carry_val0 = carry_vals;
              carry_tiles += nt;
              carry_vals += nt * width;
              const BatchWork w{item, BK_TILES, 0, nt};  // a wave per tile
              hwork.push_back(w);
              hfix.push_back(w);
            } else if (A.n_outer > 0) {
              hwork.push_back(BatchWork{item, BK_ROWS, 0, A.n_outer});  // a wave per row
            }
          }
          hdesc.push_back(d);
        }
        if (l.reqs.empty()) continue;
        l.plan.val_type = operands[0].val_type;
        l.plan.dense_type = dense_type;
        l.plan.combo = combo;
        rc = batch_launch_staged(&l.plan, hdesc.data(), static_cast<int64_t>(hdesc.size()),
                                 hwork.data(), static_cast<int64_t>(hwork.size()), hfix.data(),
                                 static_cast<int64_t>(hfix.size()), carry_tiles, carry_vals,
                                 has_sddmm, ctx->num_cus, l.stream);
        if (rc == FGX_OK && hipEventRecord(l.done, l.stream) != hipSuccess)
          rc = set_error(FGX_EHIP, "fgx_serve_run: event record failed");
        if (rc != FGX_OK) break;
        l.busy = true;
        l.seq = seq++;
        progress = true;
      }
      if (!progress) std::this_thread::yield();
    }
  }

  // ---- tear-down ----
  for (auto& l : L) {
    if (l.stream) (void)hipStreamSynchronize(l.stream);
    batch_free(&l.plan);
    if (l.done) (void)hipEventDestroy(l.done);
    if (l.stream) (void)hipStreamDestroy(l.stream);
  }
  for (auto& oi : info)
    for (auto& t : oi.tables) (void)hipFree(t.second);
  return rc;
}
