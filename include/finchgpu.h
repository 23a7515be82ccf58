/*
 * finchgpu.h — C ABI of the MI355X (gfx950) execution backend for Finch.jl's
 * compiled sparse kernels.  One shared library, libfinchgpu.so, exports every
 * symbol declared here.  Signatures use plain pointers, sizes and small POD
 * structs only, so Julia `ccall`, Python `ctypes` and C callers bind them
 * directly.
 *
 * Conventions
 * -----------
 * - Every buffer argument is a device pointer (hipMalloc'd or equivalent) owned
 *   by the caller; the library never frees caller memory.  Finch's `Tensor`
 *   wraps its level vectors by reference (src/tensors/tensors.jl:22) and
 *   outputs are resized in place by `freeze` — the two-phase
 *   symbolic/numeric entry points below mirror that protocol
 *   (src/tensors/levels/sparse_list_levels.jl:301-327,509-594).
 * - A compressed operand (`fgx_compressed`) is Finch's
 *   `Dense(SparseList(Element(z)))`: `n_outer` fibers (the Dense level),
 *   each a sorted list of coordinates in `1:n_inner` (the SparseList level,
 *   src/tensors/levels/sparse_list_levels.jl:37-42) with values in the
 *   Element level (src/tensors/levels/element_levels.jl:20-22).  Read
 *   row-wise it is CSR; Finch spells that `swizzle(A, 2, 1)` over the same
 *   buffers (src/tensors/combinators/swizzle.jl:1-64).
 * - `index_base` is 1 for Finch's native 1-based `ptr`/`idx` (ptr[1] == 1) and
 *   0 for `PlusOneVector` views of 0-based buffers (src/util/vectors.jl:3-37).
 * - Dense operands are row-major ("k-fast") unless a layout argument says
 *   otherwise; `ld` is the element stride between consecutive rows (row-major)
 *   or columns (column-major).  Finch's own `Dense(Dense(Element))` is
 *   column-major; `swizzle(Bt, 2, 1)` of the transpose gives row-major.
 * - `stream` is a hipStream_t; NULL means the HIP null (default) stream, as in
 *   rocBLAS.  Calls are stream-ordered; entry points that must return a
 *   data-dependent size (`*_symbolic`) synchronise that stream before
 *   returning.
 * - Return value: 0 (FGX_OK) or an fgx_status code; fgx_last_error() gives the
 *   message of the calling thread's last failure.
 */
#ifndef FINCHGPU_H
#define FINCHGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FGX_ABI_VERSION 1

/* Status codes.  The Julia glue rethrows them as the matching Finch
 * exceptions. */
typedef enum fgx_status {
  FGX_OK = 0,
  FGX_EDIM = 1,    /* DimensionMismatch (e.g. test/reference64/typical/typical_spmv_csc.txt:42) */
  FGX_EPROTO = 2,  /* FinchProtocolError (src/Finch.jl:70-72, sparse_list_levels.jl:541-547) */
  FGX_EUNSUP = 3,  /* unsupported format/dtype/program (src/lower.jl:289-291) */
  FGX_EOOM = 4,    /* device allocation failed */
  FGX_EADMIT = 5,  /* serving admission rejected (HBM budget or slots) */
  FGX_EHIP = 6,    /* HIP runtime error */
  FGX_EINVAL = 7   /* invalid argument (NULL pointer, negative size, ...) */
} fgx_status;

/* Element types.  BF16/F16 are backend-private storage forms for values
 * (accumulation stays FP32); Finch has no BFloat16 element type. */
typedef enum fgx_dtype {
  FGX_F64 = 0,
  FGX_F32 = 1,
  FGX_BF16 = 2,
  FGX_F16 = 3,
  FGX_I32 = 4,
  FGX_I64 = 5
} fgx_dtype;

typedef enum fgx_layout { FGX_ROW_MAJOR = 0, FGX_COL_MAJOR = 1 } fgx_layout;

/* Element-wise coiteration operators (src/interface/eager.jl:28-30,91-111).
 * MUL keeps the intersection of stored coordinates (0 annihilates `*`,
 * src/symbolic/symbolic.jl:147); ADD and MAX keep the union. */
typedef enum fgx_ewise_op { FGX_OP_MUL = 0, FGX_OP_ADD = 1, FGX_OP_MAX = 2 } fgx_ewise_op;

/* SpGEMM semirings.  PLUS_TIMES: C[i,k] += A[i,j]*B[j,k] with fill 0
 * (docs/examples/spgemm.jl:59-76).  MIN_PLUS: C[i,k] <<min>>= A[i,j]+B[j,k]
 * with fill +Inf (`+` annihilated by Inf, src/symbolic/symbolic.jl:146). */
typedef enum fgx_semiring { FGX_PLUS_TIMES = 0, FGX_MIN_PLUS = 1 } fgx_semiring;

/* A read-only compressed operand: Dense(SparseList(Element)). */
typedef struct fgx_compressed {
  int64_t n_outer;   /* fibers (rows for CSR, columns for CSC); ptr has n_outer+1 entries */
  int64_t n_inner;   /* SparseList shape: coordinates lie in [base, base+n_inner) */
  int64_t nnz;       /* ptr[n_outer] - ptr[0] */
  int32_t ptr_type;  /* FGX_I32 or FGX_I64 */
  int32_t idx_type;  /* FGX_I32 or FGX_I64 */
  int32_t val_type;  /* FGX_F64, FGX_F32, FGX_BF16, FGX_F16 or FGX_I32 */
  int32_t index_base;/* 1 (Finch) or 0 (PlusOneVector data) */
  const void* ptr;
  const void* idx;
  const void* val;
} fgx_compressed;

typedef struct fgx_ctx fgx_ctx;

/* ---- context ------------------------------------------------------------ */

/* ABI version (FGX_ABI_VERSION).  No reference counterpart. */
int fgx_version(void);

/* Create a context on HIP device `device` (one process per GPU).  The context
 * owns a stream, device scratch arenas (one set per calling stream, so calls
 * on different streams may run concurrently) and the per-matrix plan cache.  Stands in for
 * a `GPU <: AbstractDevice` instance (src/architecture.jl:6-64,170-173). */
int fgx_ctx_create(int device, fgx_ctx** out);
int fgx_ctx_destroy(fgx_ctx* ctx);
/* A non-blocking stream owned by the context (a hipStream_t) for callers that
 * want one; it is not used unless passed explicitly. */
void* fgx_ctx_stream(fgx_ctx* ctx);
/* Wait for all work queued on `stream` (NULL: the null stream). */
int fgx_sync(fgx_ctx* ctx, void* stream);
/* Message of the calling thread's last failing call ("" if none). */
const char* fgx_last_error(void);
/* Release cached per-matrix plans (nnz tile tables, autotune picks). */
int fgx_plan_cache_clear(fgx_ctx* ctx);

/* Structural check of a compressed operand against the SparseList level's
 * invariants (src/tensors/levels/sparse_list_levels.jl:37-42): ptr[0] == base,
 * ptr non-decreasing and inside [base, base+nnz], ptr[n_outer]-base == nnz,
 * every idx in [base, base+n_inner), and idx strictly increasing inside each
 * fibre.  Returns FGX_EPROTO (FinchProtocolError) when ptr or the order is
 * broken and FGX_EDIM (DimensionMismatch) when a coordinate lies outside the
 * shape or nnz disagrees with ptr.  Runs two device passes and synchronises
 * `stream`.  Finch's counterpart is the checking done in mode=:safe
 * (src/execute.jl:1-11; sparse_list_levels.jl:538-547). */
int fgx_validate(fgx_ctx* ctx, const fgx_compressed* A, void* stream);
/* safe != 0: every entry point validates each compressed operand it is given
 * (fgx_validate) before a kernel reads it, as Finch's mode=:safe checks its
 * programs; safe == 0 (the default) trusts the operands, as mode=:fast does.
 * Applies to calls made through this context from then on. */
int fgx_ctx_set_safe(fgx_ctx* ctx, int safe);

/* out[q] = start + q for q in [0, n), dtype FGX_I32 or FGX_I64.  The `val`
 * (positions) array of a frozen SparseDict level whose entries were inserted
 * fibre by fibre in coordinate order is 1:nnz (freeze_level! of
 * src/tensors/levels/sparse_dict_levels.jl:356-400); the routers build the
 * SparseDict results of `compute` with it. */
int fgx_sequence(fgx_ctx* ctx, int dtype, void* out, int64_t n, int64_t start, void* stream);

/* ---- SpMV / SpMM / SDDMM ------------------------------------------------ */

/* y = A*x with A read row-wise (CSR): y .= 0; for i, j: y[i] += A[j,i]*x[j]
 * over swizzle(A,2,1).  Replaces the loop nest of
 * test/reference64/parallel/parallel_spmv.txt (program benchmark/benchmarks.jl:365-400).
 * x has n_inner entries, y has n_outer; both of the accumulation type
 * (F64 for F64 values, F32 otherwise). */
int fgx_spmv_csr(fgx_ctx* ctx, const fgx_compressed* A, const void* x, void* y, void* stream);

/* y = A*x with A read column-wise, deterministically: A is transposed on the
 * device into the stream's scratch (its rows as fibres), then summed row by
 * row as fgx_spmv_csr does, so every y[i] is the same bits from run to run
 * and in any mix of other work.  fgx_spmv_csc instead scatters with atomic
 * adds, as Finch's AtomicElement form does (benchmark/benchmarks.jl:402-413),
 * whose order of adds varies.  Replaces the serial
 * test/reference64/typical/typical_spmv_csc.txt:26-83 where run-to-run
 * identical results are needed.  Limits of fgx_transpose apply (nnz < 2^31). */
int fgx_spmv_csc_ordered(fgx_ctx* ctx, const fgx_compressed* A, const void* x, void* y,
                         void* stream);

/* fgx_spmv_csr for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments, kernels and results
 * (bit for bit) as fgx_spmv_csr; the tile table and carries live in a fixed
 * device array of FGX_NOALLOC_SCRATCH_BYTES per GPU instead of the context's
 * arenas and plan cache.  The call allocates nothing; it returns FGX_EINVAL
 * when A needs more scratch than that (about 200M stored FP32 entries, 85M
 * FP64).  Every call on one GPU shares the array: calls that overlap on one
 * GPU (other streams, other threads, other contexts) race and must be
 * ordered by the caller.  Safe mode synchronises `stream` and cannot be used
 * under graph capture. */
#define FGX_NOALLOC_SCRATCH_BYTES (16u << 20)
int fgx_spmv_csr_noalloc(fgx_ctx* ctx, const fgx_compressed* A, const void* x, void* y,
                         void* stream);
/* y = A*x with A read column-wise (CSC): y .= 0; for j, i: y[i] += A[i,j]*x[j].
 * Replaces test/reference64/typical/typical_spmv_csc.txt:26-83
 * (program test/suites/typical_tests.jl:47-73).  x has n_outer entries, y has
 * n_inner.  The adds into each y[i] come in no fixed order, as with Finch's
 * AtomicElement form (benchmark/benchmarks.jl:402-413): operands of at least
 * 2^18 stored entries (and at most 1024 buckets of 8192 FP32 / 4096 FP64
 * rows) go through row buckets summed in LDS, smaller ones through device
 * atomic adds.  The bucket path uses nnz * 8 (FP32) or 16 (FP64) bytes of the
 * stream's scratch. */
int fgx_spmv_csc(fgx_ctx* ctx, const fgx_compressed* A, const void* x, void* y, void* stream);

/* C = A*B, A CSR (n_outer x n_inner), B n_inner x k, C n_outer x k:
 * @einsum C[i,k] += A[i,j]*B[j,k] (src/interface/einsum.jl:190; tested
 * test/suites/interface_tests.jl:22-24,38-48; parallel goldens
 * test/reference64/parallel/parallel_spmms_no_atomics_*.txt).
 * Dense element type: F64 for F64 values, F32 otherwise. */
int fgx_spmm_csr(fgx_ctx* ctx, const fgx_compressed* A, int64_t k,
                 const void* B, int64_t ldb, int b_layout,
                 void* C, int64_t ldc, int c_layout, void* stream);

/* out_val[q] = A.val[q] * sum_k X[i,k]*Y[j,k] for every stored (i,j) of CSR A:
 * compute(lazy(A) .* (lazy(X)*lazy(Yt))) (benchmark/benchmarks.jl:141-168,
 * docs/src/getting_started.md:126-134).  The output shares A's ptr/idx
 * (pattern(A): 0 annihilates `*`).  X is n_outer x k, Y is n_inner x k. */
int fgx_sddmm_csr(fgx_ctx* ctx, const fgx_compressed* A, int64_t k,
                  const void* X, int64_t ldx, int x_layout,
                  const void* Y, int64_t ldy, int y_layout,
                  void* out_val, void* stream);

/* ---- sparse-sparse: two-phase (symbolic -> caller resizes -> numeric) ---- */

/* Element-wise C = op(A, B) over two compressed operands of equal shape,
 * fiber by fiber: the sorted-list two-finger merge of
 * test/reference64/typical/typical_merge_twofinger.txt:1-194.
 * Symbolic: writes c_ptr (n_outer+1 entries, A's ptr_type and index_base) and
 * *c_nnz (synchronises).  Numeric: fills c_idx (A's idx_type) and c_val
 * (A's val_type) of length *c_nnz. */
int fgx_ewise_symbolic(fgx_ctx* ctx, int op, const fgx_compressed* A, const fgx_compressed* B,
                       void* c_ptr, int64_t* c_nnz, void* stream);
int fgx_ewise_numeric(fgx_ctx* ctx, int op, const fgx_compressed* A, const fgx_compressed* B,
                      const void* c_ptr, void* c_idx, void* c_val, void* stream);
/* Both phases in one stream-ordered call, with no host synchronisation: the
 * caller sizes c_idx / c_val to c_cap >= the bound (FGX_OP_MUL: min(nnz(A),
 * nnz(B)); FGX_OP_ADD / FGX_OP_MAX: nnz(A) + nnz(B); FGX_EINVAL below it),
 * and nnz(C) = c_ptr[n_outer] - index_base once the stream reaches it.  The
 * same results, bit for bit, as symbolic + numeric.  For stream-ordered
 * pipelines and graph capture (call once outside the capture first, so the
 * context's scratch is sized).  Finch's counterpart is the same assemble
 * loop into an output whose capacity the caller reserved (`resize_if_smaller!`,
 * src/util/shims.jl:8-12), frozen without a shrink. */
int fgx_ewise_async(fgx_ctx* ctx, int op, const fgx_compressed* A, const fgx_compressed* B,
                    void* c_ptr, void* c_idx, void* c_val, int64_t c_cap, void* stream);

/* Gustavson SpGEMM C = A*B over CSR operands (A: m x l, B: l x n), row by row:
 * spgemm_gustavson (docs/examples/spgemm.jl:59-76; generated code
 * test/reference64/kernels/gustavsons.jl:1-257; tested
 * test/suites/kernel_tests.jl:89-130).  Output coordinates of each fiber are
 * sorted (golden gustavsons.jl:182 `sort!`) and every touched coordinate is
 * stored.  Value types: FGX_PLUS_TIMES with I32 (exact, wrapping) or F32/F64;
 * FGX_MIN_PLUS with F32/F64 (exact).  Same two-phase protocol as ewise: the
 * symbolic call waits for nnz(C); the numeric call is stream-ordered only
 * (no host synchronisation). */
int fgx_spgemm_symbolic(fgx_ctx* ctx, int semiring, const fgx_compressed* A,
                        const fgx_compressed* B, void* c_ptr, int64_t* c_nnz, void* stream);
int fgx_spgemm_numeric(fgx_ctx* ctx, int semiring, const fgx_compressed* A,
                       const fgx_compressed* B, const void* c_ptr, void* c_idx,
                       void* c_val, void* stream);

/* ---- format conversion on the device ------------------------------------- */

/* Transpose: the same matrix stored along its other dimension (CSC <-> CSR).
 * permutedims(A, (2, 1)) of a 2-D Dense(SparseList) tensor and the swizzled
 * copy it runs (src/interface/copy.jl:32-40,64-71); Finch's concordize pass
 * inserts it when loop order and storage order disagree
 * (src/scheduler/optimize.jl:396-452).  Outputs, all in A's types and
 * index_base: t_ptr (A.n_inner+1 entries), t_idx (A.nnz: A's fibre numbers)
 * and t_val (A.nnz).  Every stored entry is kept, explicit zeros included;
 * each output fibre is sorted.  Deterministic (stable radix sort). */
int fgx_transpose(fgx_ctx* ctx, const fgx_compressed* A, void* t_ptr, void* t_idx, void* t_val,
                  void* stream);

/* COO -> compressed: fsparse(I, J, V, (m, n)) (src/interface/fsparse.jl:64-97)
 * copied into Dense(SparseList).  Entry q is (outer[q], inner[q], val[q])
 * with coordinates in [base, base+n_outer) x [base, base+n_inner) (idx_type
 * I32 or I64; val_type F64, F32 or I32).  Output: entries ordered by
 * (outer, inner); each run of equal coordinates combined with `+` left to
 * right in input order (fsparse.jl:85-90; I32 wraps); c_ptr (n_outer+1,
 * idx_type) and c_idx/c_val of capacity nnz; *c_nnz = distinct coordinates
 * (synchronises).  A coordinate outside the shape returns FGX_EDIM. */
int fgx_coo_to_compressed(fgx_ctx* ctx, int64_t n_outer, int64_t n_inner, int64_t nnz,
                          int idx_type, int val_type, int index_base, const void* outer,
                          const void* inner, const void* val, void* c_ptr, void* c_idx,
                          void* c_val, int64_t* c_nnz, void* stream);

/* dropfills(A) into the same format (src/interface/copy.jl:73-117): the
 * stored entries whose value is not isequal to `fill` (NaN equals NaN, -0.0
 * differs from 0.0).  Values F64, F32 or I32.  c_ptr (A.n_outer+1, A's
 * ptr_type), c_idx/c_val of capacity A.nnz; *c_nnz = kept entries
 * (synchronises). */
int fgx_dropfills(fgx_ctx* ctx, const fgx_compressed* A, double fill, void* c_ptr, void* c_idx,
                  void* c_val, int64_t* c_nnz, void* stream);

/* ---- value-precision conversion ----------------------------------------- */

/* dst[i] = (dst_type) src[i] for n elements (F32 <-> BF16/F16/F64), round to
 * nearest even.  Builds the BF16/F16 value-storage form of an operand.  No
 * reference counterpart (Finch has no BFloat16 element type). */
int fgx_convert(fgx_ctx* ctx, const void* src, int src_type, void* dst, int dst_type,
                int64_t n, void* stream);

/* ---- batched form ------------------------------------------------------- */

/* One item of a batch: C = A*B with B n_inner x n_rhs and C n_outer x n_rhs,
 * both row-major and 16-byte aligned (n_rhs > 1); C is F32, B is F32 (or BF16
 * under FGX_BATCH_BF16_INPUTS); A's values F32 or BF16. */
typedef struct fgx_batch_item {
  fgx_compressed A;
  int64_t n_rhs;
  const void* B;
  void* C;
} fgx_batch_item;

typedef struct fgx_batch_plan fgx_batch_plan;

/* Batch plan flags. */
#define FGX_BATCH_FORCE_ROWSPLIT 1  /* autotuner may only pick row-split */
#define FGX_BATCH_FORCE_NNZSPLIT 2  /* autotuner may only pick nnz-split */
#define FGX_BATCH_NO_GRAPH 4        /* launch directly instead of replaying a hipGraph */
#define FGX_BATCH_BF16_INPUTS 8     /* every B is BF16; needs BF16 values; products
                                       accumulate in F32 */
#define FGX_BATCH_MEASURE 16        /* time row-split against nnz-split on a sample of each
                                       bucket of similar items where row-split can win, and
                                       use the faster (the timed runs write those items' C) */

/* Build the work queue of one persistent launch over all items: per-item
 * algorithm (row-split vs nnz-split) and tile width picked by the autotuner
 * from shape and row-length statistics; device descriptor tables; and, unless
 * FGX_BATCH_NO_GRAPH, the hipGraph that replays the launch.  Items (and the
 * buffers they point to) must stay valid while the plan is used.  The Finch
 * analogue is a loop of `@finch` SpMM calls; there is no batched form in the
 * reference. */
int fgx_batch_plan_create(fgx_ctx* ctx, const fgx_batch_item* items, int64_t count,
                          int flags, fgx_batch_plan** out);
int fgx_batch_plan_run(fgx_ctx* ctx, fgx_batch_plan* plan, void* stream);
/* Statistics of the plan: [0] work items, [1] items on row-split,
 * [2] items on nnz-split, [3] persistent grid size, [4] total nnz,
 * [5] algorithmic bytes per run. */
int fgx_batch_plan_stats(const fgx_batch_plan* plan, int64_t* stats6);
int fgx_batch_plan_destroy(fgx_batch_plan* plan);
/* Plain path: one fgx_spmm_csr (or fgx_spmv_csr for n_rhs == 1) launch per item. */
int fgx_batch_run_loop(fgx_ctx* ctx, const fgx_batch_item* items, int64_t count, void* stream);

/* ---- serving loop ------------------------------------------------------- */

typedef enum fgx_req_kind { FGX_REQ_SPMV = 0, FGX_REQ_SPMM = 1, FGX_REQ_SDDMM = 2 } fgx_req_kind;

/* One request against a device-resident operand.  SpMV/SpMM: out = A*in
 * (in: n_inner x k, out: n_outer x k, row-major F32).  SDDMM: out[q] =
 * A[q] * X[i,:].Y[j,:] with in = X (n_outer x k), in2 = Y (n_inner x k),
 * out of length nnz.  arrival_s is the arrival time relative to the start of
 * fgx_serve_run. */
typedef struct fgx_request {
  int32_t kind;
  int32_t operand;   /* index into the operand table */
  int64_t k;
  double arrival_s;
  const void* in;
  const void* in2;
  void* out;
} fgx_request;

typedef struct fgx_serve_config {
  int64_t hbm_budget_bytes;  /* admission budget over in-flight footprints */
  int32_t max_slots;         /* max requests in flight */
  int32_t max_batch;         /* max requests fused into one launch */
  int32_t mode;              /* 0: continuous batching; 1: plain per-call path */
  int32_t input_type;        /* dense inputs: 0 or FGX_F32; FGX_BF16 (batched mode, BF16
                                values; 8-byte aligned inputs) */
} fgx_serve_config;

typedef struct fgx_request_result {
  double admit_s;     /* admitted to the device (seconds since start) */
  double done_s;      /* result complete on the device */
  int32_t status;     /* FGX_OK or FGX_EADMIT when it can never fit */
  int32_t batch;      /* launch sequence number it ran in */
} fgx_request_result;

/* Replay an open-loop request trace against `operands`: requests are released
 * at their arrival times, admitted while slots and the HBM budget allow,
 * fused into persistent launches while others run, and timed to completion.
 * No reference counterpart (Finch has no serving path). */
int fgx_serve_run(fgx_ctx* ctx, const fgx_compressed* operands, int64_t n_operands,
                  const fgx_request* reqs, int64_t n_reqs, const fgx_serve_config* cfg,
                  fgx_request_result* results);

#ifdef __cplusplus
}
#endif
#endif /* FINCHGPU_H */
