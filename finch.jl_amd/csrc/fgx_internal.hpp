// fgx_internal.hpp — host-side internals of libfinchgpu: the context,
// error reporting, scratch arena and operand validation.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/finchgpu.h"

// Cached tile table of one operand (see tile_setup): keyed by the ptr
// buffer, its shape and the tile width.  Tables are hints verified by every
// tile, so a ptr rewritten in place is still handled correctly.
// A table built on one stream is used from another only after that stream
// waits for `ready` (recorded after the build); `ready_seen` is set once the
// event has completed, after which no wait is needed.
struct fgx_tile_table {
  const void* ptr;
  int64_t n_outer, nnz, W;
  int32_t ptr_type, base;
  int64_t* d_table;
  uint64_t stamp;
  hipStream_t built_on;
  hipEvent_t ready;
  bool ready_seen;
};

// Device scratch of one stream: calls on different streams of a context run
// concurrently on the device, so each stream gets its own arenas.  Work on
// one stream is ordered, which makes reuse across its calls safe.
struct fgx_arena {
  hipStream_t stream = nullptr;
  // Carries, counters, bins, scan workspace.
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // Dense operands relaid out by a call (column-major dense operands
  // transposed to k-fast rows).
  void* dscratch = nullptr;
  size_t dscratch_bytes = 0;
  uint64_t stamp = 0;
};

struct fgx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  // Growable device scratch, one set of arenas per stream that has called
  // into this context (least recently used evicted beyond kMaxArenas).  Host
  // state is serialised by `mu`.
  std::vector<fgx_arena> arenas;
  // Safe mode (fgx_ctx_set_safe): validate every compressed operand before
  // a kernel reads it.
  bool safe = false;
  // Device address of the no-allocation scratch array on this context's
  // device (fgx_spmv_csr_noalloc), from hipGetSymbolAddress at creation.
  void* noalloc_scratch = nullptr;
  // Small pinned host buffer for data-dependent sizes (symbolic nnz).
  int64_t* host_i64 = nullptr;
  std::vector<fgx_tile_table> tables;
  uint64_t clock = 0;
  std::mutex mu;
};

namespace fgx {

// Records the calling thread's last error; returns `code`.
int set_error(int code, const char* fmt, ...);
void clear_error();

#define FGX_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return ::fgx::set_error(FGX_EHIP, "%s failed: %s (%s:%d)", #expr,                \
                              hipGetErrorString(e_), __FILE__, __LINE__);              \
  } while (0)

#define FGX_TRY(expr)                 \
  do {                                \
    int s_ = (expr);                  \
    if (s_ != FGX_OK) return s_;      \
  } while (0)

// Scratch of at least `bytes` for work on stream `s` (grows, never
// shrinks).  Caller holds ctx->mu.
int scratch(fgx_ctx* ctx, hipStream_t s, size_t bytes, void** out);
int dense_scratch(fgx_ctx* ctx, hipStream_t s, size_t bytes, void** out);
void free_arenas(fgx_ctx* ctx);

// The stream a call runs on: the caller's hipStream_t; NULL is the HIP null
// (default) stream, as in rocBLAS, so calls order with the caller's other
// default-stream work.
inline hipStream_t pick_stream(fgx_ctx* ctx, void* s) {
  (void)ctx;
  return static_cast<hipStream_t>(s);
}

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Tile table (ntiles+1 entries) of A for tiles of W positions, built on
// first use and cached in the context.  Caller holds ctx->mu.
int tile_table(fgx_ctx* ctx, const fgx_compressed* A, int64_t W, hipStream_t s,
               const int64_t** out);
void free_tile_tables(fgx_ctx* ctx);
// Builds an uncached tile table; *out is hipMalloc'd and owned by the caller.
int build_tile_table(const fgx_compressed* A, int64_t W, hipStream_t s, int64_t** out);
// Writes the tile table of A for width W (ntiles + 1 entries) into d.
int fill_tile_table(const fgx_compressed* A, int64_t W, hipStream_t s, int64_t* d);

// Checks the descriptor (non-null buffers, supported type combination).
int check_compressed(const fgx_compressed* A, const char* name);

// Device check of A's SparseList invariants (synchronises s): FGX_EPROTO or
// FGX_EDIM on a violation.  admit_operand runs it only in safe mode.
// `flags` is a device word to use instead of the stream's scratch arena
// (callers that must not allocate).
int validate_operand(fgx_ctx* ctx, const fgx_compressed* A, hipStream_t s, const char* name,
                     uint32_t* flags = nullptr);
int admit_operand(fgx_ctx* ctx, const fgx_compressed* A, hipStream_t s, const char* name);

// Transpose (fgx_transpose) and CSR SpMV (fgx_spmv_csr) for entry points
// that already hold ctx->mu and checked their operands.
int transpose_into(fgx_ctx* ctx, const fgx_compressed* A, void* t_ptr, void* t_idx, void* t_val,
                   hipStream_t s);
int spmv_csr_into(fgx_ctx* ctx, const fgx_compressed* A, const void* x, void* y, hipStream_t s);

// Looks up the no-allocation scratch address for ctx's (current) device.
int noalloc_setup(fgx_ctx* ctx);

// Index-type combinations the kernels are instantiated for.
enum IdxCombo { IDX_32_32 = 0, IDX_64_64 = 1, IDX_64_32 = 2 };
int idx_combo(const fgx_compressed* A, IdxCombo* out);

inline int elem_size(int dtype) {
  switch (dtype) {
    case FGX_F64: case FGX_I64: return 8;
    case FGX_F32: case FGX_I32: return 4;
    case FGX_BF16: case FGX_F16: return 2;
    default: return 0;
  }
}

// Dense element type paired with a value type: F64 for F64, F32 otherwise.
inline int dense_type_of(int val_type) { return val_type == FGX_F64 ? FGX_F64 : FGX_F32; }

int launch_check(const char* what);

}  // namespace fgx

// Dispatch helpers: call F.template operator()<TP, TI>() for the operand's
// index combination.
#define FGX_DISPATCH_IDX(combo, ...)                                   \
  switch (combo) {                                                     \
    case ::fgx::IDX_32_32: { using TP = int32_t; using TI = int32_t; __VA_ARGS__; break; } \
    case ::fgx::IDX_64_64: { using TP = int64_t; using TI = int64_t; __VA_ARGS__; break; } \
    case ::fgx::IDX_64_32: { using TP = int64_t; using TI = int32_t; __VA_ARGS__; break; } \
  }
