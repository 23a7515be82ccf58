// fgx_ctx.hip — context lifecycle, error reporting, scratch arena and the
// value-precision conversion kernel.
#include "fgx_internal.hpp"
#include "fgx_csr_tiles.hpp"

#include <cstring>

namespace fgx {

static thread_local char g_last_error[512] = "";

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

void clear_error() { g_last_error[0] = '\0'; }

// Grows one of a stream's device arenas to at least `bytes`.  The old
// arena may still be in use by queued work, so the device is drained first.
static int grow_arena(void** buf, size_t* cur, size_t bytes, void** out) {
  if (bytes == 0) bytes = 256;
  if (bytes > *cur) {
    if (*buf) {
      FGX_HIP(hipDeviceSynchronize());
      FGX_HIP(hipFree(*buf));
      *buf = nullptr;
      *cur = 0;
    }
    size_t want = bytes + bytes / 4;
    want = (want + 4095) & ~static_cast<size_t>(4095);
    if (hipMalloc(buf, want) != hipSuccess) {
      (void)hipGetLastError();
      *buf = nullptr;
      return set_error(FGX_EOOM, "scratch allocation of %zu bytes failed", want);
    }
    *cur = want;
  }
  *out = *buf;
  return FGX_OK;
}

// Streams with arenas of their own; beyond this the least recently used
// stream's arenas are freed (after draining the device) and reused.
constexpr size_t kMaxArenas = 16;

static void release_arena(fgx_arena* a) {
  if (a->scratch) (void)hipFree(a->scratch);
  if (a->dscratch) (void)hipFree(a->dscratch);
  *a = fgx_arena{};
}

static int arena_of(fgx_ctx* ctx, hipStream_t s, fgx_arena** out) {
  for (auto& a : ctx->arenas) {
    if (a.stream == s) {
      a.stamp = ++ctx->clock;
      *out = &a;
      return FGX_OK;
    }
  }
  if (ctx->arenas.size() >= kMaxArenas) {
    size_t oldest = 0;
    for (size_t i = 1; i < ctx->arenas.size(); i++)
      if (ctx->arenas[i].stamp < ctx->arenas[oldest].stamp) oldest = i;
    // Queued work on the evicted stream may still read its arenas.
    FGX_HIP(hipDeviceSynchronize());
    release_arena(&ctx->arenas[oldest]);
    ctx->arenas.erase(ctx->arenas.begin() + static_cast<long>(oldest));
  }
  fgx_arena a;
  a.stream = s;
  a.stamp = ++ctx->clock;
  ctx->arenas.push_back(a);
  *out = &ctx->arenas.back();
  return FGX_OK;
}

int scratch(fgx_ctx* ctx, hipStream_t s, size_t bytes, void** out) {
  fgx_arena* a;
  FGX_TRY(arena_of(ctx, s, &a));
  return grow_arena(&a->scratch, &a->scratch_bytes, bytes, out);
}

int dense_scratch(fgx_ctx* ctx, hipStream_t s, size_t bytes, void** out) {
  fgx_arena* a;
  FGX_TRY(arena_of(ctx, s, &a));
  return grow_arena(&a->dscratch, &a->dscratch_bytes, bytes, out);
}

void free_arenas(fgx_ctx* ctx) {
  if (ctx->arenas.empty()) return;
  (void)hipDeviceSynchronize();
  for (auto& a : ctx->arenas) release_arena(&a);
  ctx->arenas.clear();
}

// Enough for a serving-sized operand set (hundreds of operands, a few tile
// widths each): a table evicted and rebuilt on every call would also free
// device memory, which waits for the device.
constexpr size_t kMaxTables = 1024;

void free_tile_tables(fgx_ctx* ctx) {
  if (ctx->tables.empty()) return;
  (void)hipDeviceSynchronize();
  for (auto& t : ctx->tables) {
    (void)hipFree(t.d_table);
    if (t.ready) (void)hipEventDestroy(t.ready);
  }
  ctx->tables.clear();
}

int fill_tile_table(const fgx_compressed* A, int64_t W, hipStream_t s, int64_t* d) {
  const int64_t ntiles = A->nnz > 0 ? (A->nnz + W - 1) / W : 1;
  const unsigned blocks = static_cast<unsigned>((ntiles + 1 + 255) / 256);
  if (A->ptr_type == FGX_I32)
    hipLaunchKernelGGL((tile_table_kernel<int32_t>), dim3(blocks), dim3(256), 0, s,
                       static_cast<const int32_t*>(A->ptr), static_cast<int64_t>(A->index_base),
                       A->n_outer, A->nnz, W, ntiles, d);
  else
    hipLaunchKernelGGL((tile_table_kernel<int64_t>), dim3(blocks), dim3(256), 0, s,
                       static_cast<const int64_t*>(A->ptr), static_cast<int64_t>(A->index_base),
                       A->n_outer, A->nnz, W, ntiles, d);
  return launch_check("tile_table_kernel");
}

int build_tile_table(const fgx_compressed* A, int64_t W, hipStream_t s, int64_t** out) {
  const int64_t ntiles = A->nnz > 0 ? (A->nnz + W - 1) / W : 1;
  int64_t* d = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&d), (ntiles + 1) * sizeof(int64_t)) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(FGX_EOOM, "tile table allocation failed");
  }
  const int rc = fill_tile_table(A, W, s, d);
  if (rc != FGX_OK) {
    (void)hipFree(d);
    return rc;
  }
  *out = d;
  return FGX_OK;
}

int tile_table(fgx_ctx* ctx, const fgx_compressed* A, int64_t W, hipStream_t s,
               const int64_t** out) {
  for (auto& t : ctx->tables) {
    if (t.ptr == A->ptr && t.n_outer == A->n_outer && t.nnz == A->nnz && t.W == W &&
        t.ptr_type == A->ptr_type && t.base == A->index_base) {
      t.stamp = ++ctx->clock;
      if (!t.ready_seen && t.built_on != s) {
        // Built on another stream: order this stream after the build.
        const hipError_t q = hipEventQuery(t.ready);
        if (q == hipSuccess) {
          t.ready_seen = true;
        } else {
          if (q != hipErrorNotReady) (void)hipGetLastError();
          FGX_HIP(hipStreamWaitEvent(s, t.ready, 0));
        }
      }
      *out = t.d_table;
      return FGX_OK;
    }
  }
  if (ctx->tables.size() >= kMaxTables) {
    size_t oldest = 0;
    for (size_t i = 1; i < ctx->tables.size(); i++)
      if (ctx->tables[i].stamp < ctx->tables[oldest].stamp) oldest = i;
    FGX_HIP(hipDeviceSynchronize());
    (void)hipFree(ctx->tables[oldest].d_table);
    if (ctx->tables[oldest].ready) (void)hipEventDestroy(ctx->tables[oldest].ready);
    ctx->tables.erase(ctx->tables.begin() + static_cast<long>(oldest));
  }
  fgx_tile_table t{A->ptr, A->n_outer, A->nnz, W, A->ptr_type, A->index_base, nullptr,
                   ++ctx->clock, s, nullptr, false};
  FGX_HIP(hipEventCreateWithFlags(&t.ready, hipEventDisableTiming));
  const int rc = build_tile_table(A, W, s, &t.d_table);
  if (rc == FGX_OK && hipEventRecord(t.ready, s) == hipSuccess) {
    ctx->tables.push_back(t);
  } else {
    (void)hipGetLastError();
    (void)hipEventDestroy(t.ready);
    if (t.d_table) {
      (void)hipStreamSynchronize(s);
      (void)hipFree(t.d_table);
    }
    return rc != FGX_OK ? rc : set_error(FGX_EHIP, "tile table event record failed");
  }
  *out = t.d_table;
  return FGX_OK;
}

int check_compressed(const fgx_compressed* A, const char* name) {
  if (!A) return set_error(FGX_EINVAL, "%s: NULL descriptor", name);
  if (A->n_outer < 0 || A->n_inner < 0 || A->nnz < 0)
    return set_error(FGX_EINVAL, "%s: negative size", name);
  if (A->index_base != 0 && A->index_base != 1)
    return set_error(FGX_EINVAL, "%s: index_base must be 0 or 1", name);
  if (!A->ptr) return set_error(FGX_EINVAL, "%s: NULL ptr", name);
  if (A->nnz > 0 && (!A->idx || !A->val)) return set_error(FGX_EINVAL, "%s: NULL idx/val", name);
  if (A->ptr_type == FGX_I32 && A->nnz > INT32_MAX)
    return set_error(FGX_EUNSUP, "%s: nnz exceeds Int32 ptr", name);
  return FGX_OK;
}

int idx_combo(const fgx_compressed* A, IdxCombo* out) {
  if (A->ptr_type == FGX_I32 && A->idx_type == FGX_I32) *out = IDX_32_32;
  else if (A->ptr_type == FGX_I64 && A->idx_type == FGX_I64) *out = IDX_64_64;
  else if (A->ptr_type == FGX_I64 && A->idx_type == FGX_I32) *out = IDX_64_32;
  else
    return set_error(FGX_EUNSUP, "unsupported ptr/idx types (%d, %d)", A->ptr_type, A->idx_type);
  return FGX_OK;
}

int launch_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(FGX_EHIP, "%s launch failed: %s", what, hipGetErrorString(e));
  return FGX_OK;
}

// ---- conversion kernel ----------------------------------------------------
__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return static_cast<uint16_t>((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

template <typename S, typename D>
__device__ __forceinline__ D conv(S s);

template <int ST, int DT>
__global__ void convert_kernel(const void* __restrict__ src, void* __restrict__ dst, int64_t n) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) {
    float f;
    double d = 0.0;
    if constexpr (ST == FGX_F32) f = static_cast<const float*>(src)[i];
    else if constexpr (ST == FGX_F64) { d = static_cast<const double*>(src)[i]; f = static_cast<float>(d); }
    else if constexpr (ST == FGX_BF16) f = to_f32(static_cast<const bf16_t*>(src)[i]);
    else f = to_f32(static_cast<const f16_t*>(src)[i]);
    if constexpr (DT == FGX_F32) static_cast<float*>(dst)[i] = f;
    else if constexpr (DT == FGX_F64) static_cast<double*>(dst)[i] = (ST == FGX_F64) ? d : static_cast<double>(f);
    else if constexpr (DT == FGX_BF16) static_cast<uint16_t*>(dst)[i] = f32_to_bf16_rne(f);
    else {
      _Float16 h = static_cast<_Float16>(f);
      uint16_t b;
      __builtin_memcpy(&b, &h, 2);
      static_cast<uint16_t*>(dst)[i] = b;
    }
  }
}

template <typename T>
__global__ void sequence_kernel(T* __restrict__ out, int64_t n, int64_t start) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) out[i] = static_cast<T>(start + i);
}

}  // namespace fgx

using namespace fgx;

extern "C" {

int fgx_version(void) { return FGX_ABI_VERSION; }

const char* fgx_last_error(void) { return g_last_error; }

int fgx_ctx_create(int device, fgx_ctx** out) {
  clear_error();
  if (!out) return set_error(FGX_EINVAL, "fgx_ctx_create: NULL out");
  int n = 0;
  FGX_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return set_error(FGX_EINVAL, "fgx_ctx_create: device %d of %d", device, n);
  FGX_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  FGX_HIP(hipGetDeviceProperties(&prop, device));
  fgx_ctx* c = new fgx_ctx();
  c->device = device;
  c->num_cus = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return set_error(FGX_EHIP, "fgx_ctx_create: stream creation failed");
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&c->host_i64), 64 * sizeof(int64_t)) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return set_error(FGX_EHIP, "fgx_ctx_create: pinned host buffer allocation failed");
  }
  // With `device` current: the no-allocation entry point's scratch address
  // on this device.
  const int rc = noalloc_setup(c);
  if (rc != FGX_OK) {
    (void)hipHostFree(c->host_i64);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  *out = c;
  return FGX_OK;
}

int fgx_ctx_destroy(fgx_ctx* ctx) {
  if (!ctx) return FGX_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  free_tile_tables(ctx);
  free_arenas(ctx);
  if (ctx->host_i64) (void)hipHostFree(ctx->host_i64);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return FGX_OK;
}

void* fgx_ctx_stream(fgx_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int fgx_sync(fgx_ctx* ctx, void* stream) {
  if (!ctx) return set_error(FGX_EINVAL, "fgx_sync: NULL ctx");
  FGX_HIP(hipSetDevice(ctx->device));
  FGX_HIP(hipStreamSynchronize(pick_stream(ctx, stream)));
  return FGX_OK;
}

int fgx_plan_cache_clear(fgx_ctx* ctx) {
  if (!ctx) return set_error(FGX_EINVAL, "fgx_plan_cache_clear: NULL ctx");
  std::lock_guard<std::mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  free_tile_tables(ctx);
  return FGX_OK;
}

int fgx_sequence(fgx_ctx* ctx, int dtype, void* out, int64_t n, int64_t start, void* stream) {
  clear_error();
  if (!ctx) return set_error(FGX_EINVAL, "fgx_sequence: NULL ctx");
  if (n < 0) return set_error(FGX_EINVAL, "fgx_sequence: negative n");
  if (n == 0) return FGX_OK;
  if (!out) return set_error(FGX_EINVAL, "fgx_sequence: NULL out");
  if (dtype != FGX_I32 && dtype != FGX_I64)
    return set_error(FGX_EUNSUP, "fgx_sequence: dtype %d is not an index type", dtype);
  if (dtype == FGX_I32 && (start < INT32_MIN || start + n - 1 > INT32_MAX))
    return set_error(FGX_EINVAL, "fgx_sequence: values exceed Int32");
  FGX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = pick_stream(ctx, stream);
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (dtype == FGX_I32)
    hipLaunchKernelGGL((sequence_kernel<int32_t>), dim3(static_cast<unsigned>(blocks)), dim3(256),
                       0, s, static_cast<int32_t*>(out), n, start);
  else
    hipLaunchKernelGGL((sequence_kernel<int64_t>), dim3(static_cast<unsigned>(blocks)), dim3(256),
                       0, s, static_cast<int64_t*>(out), n, start);
  return launch_check("fgx_sequence");
}

int fgx_convert(fgx_ctx* ctx, const void* src, int src_type, void* dst, int dst_type, int64_t n,
                void* stream) {
  if (!ctx) return set_error(FGX_EINVAL, "fgx_convert: NULL ctx");
  if (n < 0) return set_error(FGX_EINVAL, "fgx_convert: negative n");
  if (n == 0) return FGX_OK;
  if (!src || !dst) return set_error(FGX_EINVAL, "fgx_convert: NULL buffer");
  FGX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = pick_stream(ctx, stream);
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  dim3 g(static_cast<unsigned>(blocks)), b(256);
#define FGX_CONV(ST, DT)                                                       \
  if (src_type == ST && dst_type == DT) {                                      \
    hipLaunchKernelGGL((convert_kernel<ST, DT>), g, b, 0, s, src, dst, n);     \
    return launch_check("fgx_convert");                                        \
  }
  FGX_CONV(FGX_F32, FGX_BF16)
  FGX_CONV(FGX_F32, FGX_F16)
  FGX_CONV(FGX_BF16, FGX_F32)
  FGX_CONV(FGX_F16, FGX_F32)
  FGX_CONV(FGX_F32, FGX_F64)
  FGX_CONV(FGX_F64, FGX_F32)
#undef FGX_CONV
  return set_error(FGX_EUNSUP, "fgx_convert: %d -> %d unsupported", src_type, dst_type);
}

}  // extern "C"
