/*
 * finch_oracle.c — CPU oracle: a C restatement of the loop nests Finch.jl
 * generates for the sparse kernels this backend replaces.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for parity tests and the
 * timed CPU baseline ("port") in bench.py; the product path (libfinchgpu.so)
 * never links, loads or calls it.
 *
 * Every function follows the generated Julia code it cites (paths relative to
 * the Finch.jl v1.2.12 tree), in the same accumulation order:
 *   - 1-based (or 0-based, `base`) ptr/idx; fiber p occupies
 *     q in [ptr[p]-base, ptr[p+1]-base);
 *   - `y[i] = x*a + y[i]` products are rounded before the add: built with
 *     -ffp-contract=off, which reproduces the FP64 golden
 *     test/reference64/typical/typical_spmv_csc.txt:92 bit for bit;
 *   - every coiterated coordinate is stored even when its value equals the
 *     fill (src/tensors/scalars.jl:209-218).
 * Indices are taken as int64 (Finch's default `Int`); callers widen int32
 * buffers before calling.
 *
 * Pinning (tests/test_oracle.py): each kernel is checked against the
 * reference's golden vectors where one exists, against scipy/numpy or a
 * brute-force dense loop on small inputs, and against closed forms.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;

/* Julia's fld for non-negative operands. */
static inline i64 fld(i64 a, i64 b) { return a / b; }

/* ---------------------------------------------------------------------------
 * CSC SpMV:  y .= 0; for j = _, i = _; y[i] += A[i, j] * x[j]
 * test/reference64/typical/typical_spmv_csc.txt:26-83 — outer loop over the
 * Dense level (columns j ascending), inner walk of the SparseList fiber,
 * `y[i] = x[j] * A[i,j] + y[i]` (line 62).
 * ------------------------------------------------------------------------- */
#define DEF_SPMV_CSC(T, SFX)                                                       \
  void or_spmv_csc_##SFX(i64 n_inner, i64 n_outer, const i64* ptr, const i64* idx, \
                         const T* val, int base, const T* x, T* y) {              \
    for (i64 i = 0; i < n_inner; i++) y[i] = (T)0;                                 \
    for (i64 j = 0; j < n_outer; j++) {                                            \
      T xj = x[j];                                                                 \
      for (i64 q = ptr[j] - base; q < ptr[j + 1] - base; q++) {                    \
        i64 i = idx[q] - base;                                                     \
        y[i] = xj * val[q] + y[i];                                                 \
      }                                                                            \
    }                                                                              \
  }
DEF_SPMV_CSC(double, f64)
DEF_SPMV_CSC(float, f32)

/* ---------------------------------------------------------------------------
 * CSR SpMV over swizzle(A, 2, 1):  y .= 0; for i = parallel(_), j = _;
 * y[i] += A[j, i] * x[j]
 * test/reference64/parallel/parallel_spmv.txt:36-76 — thread tid (1-based)
 * of n takes rows fld(m*(tid-1), n)+1 .. fld(m*tid, n) (static schedule,
 * src/architecture.jl:689-719); per row, `y[i] = A[q]*x[idx[q]] + y[i]` with
 * q ascending (line 58).  nthreads <= 1 is the serial program
 * (benchmark/benchmarks.jl:365-376).
 * ------------------------------------------------------------------------- */
#define DEF_SPMV_CSR(T, SFX)                                                          \
  static void spmv_csr_rows_##SFX(i64 r0, i64 r1, const i64* ptr, const i64* idx,     \
                                  const T* val, int base, const T* x, T* y) {         \
    for (i64 i = r0; i < r1; i++) {                                                   \
      T acc = (T)0;                                                                   \
      for (i64 q = ptr[i] - base; q < ptr[i + 1] - base; q++)                         \
        acc = val[q] * x[idx[q] - base] + acc;                                        \
      y[i] = acc;                                                                     \
    }                                                                                 \
  }                                                                                   \
  void or_spmv_csr_##SFX(i64 n_outer, i64 n_inner, const i64* ptr, const i64* idx,    \
                         const T* val, int base, const T* x, T* y, int nthreads) {    \
    (void)n_inner;                                                                    \
    if (nthreads <= 1) {                                                              \
      spmv_csr_rows_##SFX(0, n_outer, ptr, idx, val, base, x, y);                     \
      return;                                                                         \
    }                                                                                 \
    _Pragma("omp parallel for schedule(static, 1) num_threads(nthreads)")            \
    for (int tid = 1; tid <= nthreads; tid++) {                                       \
      i64 r0 = fld(n_outer * (tid - 1), nthreads);                                    \
      i64 r1 = fld(n_outer * tid, nthreads);                                          \
      spmv_csr_rows_##SFX(r0, r1, ptr, idx, val, base, x, y);                         \
    }                                                                                 \
  }
DEF_SPMV_CSR(double, f64)
DEF_SPMV_CSR(float, f32)

/* ---------------------------------------------------------------------------
 * CSR SpMM:  @einsum C[i,k] += A[i,j]*B[j,k]  (src/interface/einsum.jl:190),
 * with A read row-wise.  Rows are split across threads like parallel_spmv.txt
 * (test/reference64/parallel/parallel_spmms_no_atomics_1.txt: the parallel
 * loop is over the output rows, no atomics).  For each (i, k) the sum runs
 * over the stored j of row i in ascending order: C[i,k] = A[q]*B[j,k] + C[i,k].
 * B is n_inner x k and C is n_outer x k, both row-major.
 * ------------------------------------------------------------------------- */
#define DEF_SPMM_CSR(T, SFX)                                                           \
  static void spmm_csr_rows_##SFX(i64 r0, i64 r1, i64 k, const i64* ptr, const i64* idx,\
                                  const T* val, int base, const T* B, T* C) {          \
    for (i64 i = r0; i < r1; i++) {                                                    \
      T* c = C + i * k;                                                                \
      for (i64 kk = 0; kk < k; kk++) c[kk] = (T)0;                                     \
      for (i64 q = ptr[i] - base; q < ptr[i + 1] - base; q++) {                        \
        const T a = val[q];                                                            \
        const T* b = B + (idx[q] - base) * k;                                          \
        for (i64 kk = 0; kk < k; kk++) c[kk] = a * b[kk] + c[kk];                      \
      }                                                                                \
    }                                                                                  \
  }                                                                                    \
  void or_spmm_csr_##SFX(i64 n_outer, i64 n_inner, i64 k, const i64* ptr, const i64* idx,\
                         const T* val, int base, const T* B, T* C, int nthreads) {     \
    (void)n_inner;                                                                     \
    if (nthreads <= 1) {                                                               \
      spmm_csr_rows_##SFX(0, n_outer, k, ptr, idx, val, base, B, C);                   \
      return;                                                                          \
    }                                                                                  \
    _Pragma("omp parallel for schedule(static, 1) num_threads(nthreads)")             \
    for (int tid = 1; tid <= nthreads; tid++) {                                        \
      i64 r0 = fld(n_outer * (tid - 1), nthreads);                                     \
      i64 r1 = fld(n_outer * tid, nthreads);                                           \
      spmm_csr_rows_##SFX(r0, r1, k, ptr, idx, val, base, B, C);                       \
    }                                                                                  \
  }
DEF_SPMM_CSR(double, f64)
DEF_SPMM_CSR(float, f32)

/* ---------------------------------------------------------------------------
 * Threaded CSC SpMV with atomic adds:
 *   y = Tensor(Dense(AtomicElement(0.0))); y .= 0;
 *   for j = parallel(_), i = _; y[i] += A[i, j] * x[j]
 * benchmark/benchmarks.jl:402-413; AtomicElement's update is an atomic
 * read-modify-write of y[i] (src/tensors/levels/atomic_element_levels.jl:
 * 203-224, atomic_modify! src/architecture.jl:487-547).  Columns are split in
 * Finch's static fld blocks; each product x[j]*A[i,j] is rounded, then added
 * atomically.  The order of the adds into one y[i] depends on the threads'
 * interleaving, so the result is deterministic only at nthreads <= 1.
 * ------------------------------------------------------------------------- */
#define DEF_SPMV_CSC_ATOMIC(T, SFX)                                                      \
  void or_spmv_csc_atomic_##SFX(i64 n_inner, i64 n_outer, const i64* ptr, const i64* idx,\
                                const T* val, int base, const T* x, T* y, int nthreads) { \
    for (i64 i = 0; i < n_inner; i++) y[i] = (T)0;                                       \
    if (nthreads < 1) nthreads = 1;                                                      \
    _Pragma("omp parallel for schedule(static, 1) num_threads(nthreads)")               \
    for (int tid = 1; tid <= nthreads; tid++) {                                          \
      i64 j0 = fld(n_outer * (tid - 1), nthreads);                                       \
      i64 j1 = fld(n_outer * tid, nthreads);                                             \
      for (i64 j = j0; j < j1; j++) {                                                    \
        const T xj = x[j];                                                               \
        for (i64 q = ptr[j] - base; q < ptr[j + 1] - base; q++) {                        \
          const T p = xj * val[q];                                                       \
          T* dst = &y[idx[q] - base];                                                    \
          _Pragma("omp atomic")                                                          \
          *dst += p;                                                                     \
        }                                                                                \
      }                                                                                  \
    }                                                                                    \
  }
DEF_SPMV_CSC_ATOMIC(double, f64)
DEF_SPMV_CSC_ATOMIC(float, f32)

/* ---------------------------------------------------------------------------
 * CSC SpMV with a Mutex output level (benchmark/benchmarks.jl:414-426:
 * y = Tensor(Dense(Mutex(Element(0.0))))): the same static column blocks,
 * and each update of y[i] is done under y[i]'s own lock, taken before the
 * read and released after the write (src/tensors/levels/mutex_levels.jl:
 * 304-352, aquire_lock! / release_lock! around the assignment).
 * ------------------------------------------------------------------------- */
#define DEF_SPMV_CSC_MUTEX(T, SFX)                                                       \
  void or_spmv_csc_mutex_##SFX(i64 n_inner, i64 n_outer, const i64* ptr, const i64* idx, \
                               const T* val, int base, const T* x, T* y, int nthreads) {  \
    omp_lock_t* locks = (omp_lock_t*)malloc(sizeof(omp_lock_t) * (size_t)(n_inner > 0 ? n_inner : 1)); \
    for (i64 i = 0; i < n_inner; i++) {                                                  \
      y[i] = (T)0;                                                                       \
      omp_init_lock(&locks[i]);                                                          \
    }                                                                                    \
    if (nthreads < 1) nthreads = 1;                                                      \
    _Pragma("omp parallel for schedule(static, 1) num_threads(nthreads)")               \
    for (int tid = 1; tid <= nthreads; tid++) {                                          \
      i64 j0 = fld(n_outer * (tid - 1), nthreads);                                       \
      i64 j1 = fld(n_outer * tid, nthreads);                                             \
      for (i64 j = j0; j < j1; j++) {                                                    \
        const T xj = x[j];                                                               \
        for (i64 q = ptr[j] - base; q < ptr[j + 1] - base; q++) {                        \
          const i64 i = idx[q] - base;                                                   \
          omp_set_lock(&locks[i]);                                                       \
          y[i] = xj * val[q] + y[i];                                                     \
          omp_unset_lock(&locks[i]);                                                     \
        }                                                                                \
      }                                                                                  \
    }                                                                                    \
    for (i64 i = 0; i < n_inner; i++) omp_destroy_lock(&locks[i]);                       \
    free(locks);                                                                         \
  }
DEF_SPMV_CSC_MUTEX(double, f64)
DEF_SPMV_CSC_MUTEX(float, f32)

/* ---------------------------------------------------------------------------
 * Greedy schedule of the threaded CSR loops (src/architecture.jl:721-767):
 * threads repeatedly take the next chunk of `chk` rows from a shared atomic
 * counter until the rows run out; each row is summed as in the static
 * programs above, so the result equals the serial one bit for bit.
 * ------------------------------------------------------------------------- */
#define DEF_GREEDY(T, SFX)                                                               \
  void or_spmv_csr_greedy_##SFX(i64 n_outer, const i64* ptr, const i64* idx,            \
                                const T* val, int base, const T* x, T* y, i64 chk,       \
                                int nthreads) {                                          \
    if (chk < 1) chk = 1;                                                                \
    i64 next = 0;                                                                        \
    _Pragma("omp parallel num_threads(nthreads)")                                        \
    for (;;) {                                                                           \
      i64 c;                                                                             \
      _Pragma("omp atomic capture")                                                      \
      c = next++;                                                                        \
      const i64 r0 = c * chk;                                                            \
      if (r0 >= n_outer) break;                                                          \
      const i64 r1 = r0 + chk < n_outer ? r0 + chk : n_outer;                            \
      spmv_csr_rows_##SFX(r0, r1, ptr, idx, val, base, x, y);                            \
    }                                                                                    \
  }                                                                                      \
  void or_spmm_csr_greedy_##SFX(i64 n_outer, i64 k, const i64* ptr, const i64* idx,     \
                                const T* val, int base, const T* B, T* C, i64 chk,       \
                                int nthreads) {                                          \
    if (chk < 1) chk = 1;                                                                \
    i64 next = 0;                                                                        \
    _Pragma("omp parallel num_threads(nthreads)")                                        \
    for (;;) {                                                                           \
      i64 c;                                                                             \
      _Pragma("omp atomic capture")                                                      \
      c = next++;                                                                        \
      const i64 r0 = c * chk;                                                            \
      if (r0 >= n_outer) break;                                                          \
      const i64 r1 = r0 + chk < n_outer ? r0 + chk : n_outer;                            \
      spmm_csr_rows_##SFX(r0, r1, k, ptr, idx, val, base, B, C);                         \
    }                                                                                    \
  }
DEF_GREEDY(double, f64)
DEF_GREEDY(float, f32)

/* ---------------------------------------------------------------------------
 * SDDMM:  C[i,j] = A[i,j] * sum_k X[i,k]*Y[j,k] on pattern(A)
 * The fused plan of compute(lazy(A) .* (lazy(X)*lazy(Yt)))
 * (benchmark/benchmarks.jl:141-168) is the @finch nest
 *   for i, j (stored in A), k: C[i,j] += A[i,j] * X[i,k] * Y[j,k].
 * There is no golden for it; this restatement sums the dot over k ascending
 * and scales once (Finch's fused plan fixes no order a test can observe), and
 * is pinned by the closed form A o (X Y^T) in FP64.  X is n_outer x k and Y
 * n_inner x k, row-major.  Serial: Finch's SparseList output is not
 * concurrent (src/tensors/levels/sparse_list_levels.jl:185-188).
 * ------------------------------------------------------------------------- */
#define DEF_SDDMM(T, SFX)                                                               \
  void or_sddmm_csr_##SFX(i64 n_outer, i64 n_inner, i64 k, const i64* ptr, const i64* idx,\
                          const T* val, int base, const T* X, const T* Y, T* out) {     \
    (void)n_inner;                                                                      \
    for (i64 i = 0; i < n_outer; i++) {                                                 \
      const T* xr = X + i * k;                                                          \
      for (i64 q = ptr[i] - base; q < ptr[i + 1] - base; q++) {                         \
        const T* yr = Y + (idx[q] - base) * k;                                          \
        T acc = (T)0;                                                                   \
        for (i64 kk = 0; kk < k; kk++) acc = xr[kk] * yr[kk] + acc;                     \
        out[q] = val[q] * acc;                                                          \
      }                                                                                 \
    }                                                                                   \
  }
DEF_SDDMM(double, f64)
DEF_SDDMM(float, f32)

/* ---------------------------------------------------------------------------
 * Element-wise coiteration, fiber by fiber:
 *   C .= 0; for p = _, i = _; C[i, p] = A[i, p] op B[i, p]
 * test/reference64/typical/typical_merge_twofinger.txt:45-181 (two-finger
 * merge of two SparseList fibers, appending to the output SparseList with
 * `z_lvl_qos += 1`, then the ptr prefix sum of lines 182-186).
 *   op 0 (MUL): 0 annihilates `*` (src/symbolic/symbolic.jl:147), so only
 *               coordinates stored in both survive: value a*b;
 *   op 1 (ADD): union; both -> b + a (golden line 63 writes y + x), one-sided
 *               -> the stored value (a + 0 == a);
 *   op 2 (MAX): union; both -> max(a, b), one-sided -> max(a, 0) for an
 *               A-only entry and max(0, b) for a B-only one (the missing side
 *               reads as the fill 0, src/interface/eager.jl:28-30).
 * `max` is Julia's Base.max (base/math.jl): for floats a NaN on either side
 * gives NaN, and -0.0 orders below 0.0.  Julia returns `x - y` when a NaN is
 * present; on x86-64 that is the first NaN operand with its quiet bit set,
 * which JL_MAX returns bit for bit.  Int32 `+` and `*` wrap modulo 2^32 as
 * Julia's Int32 does (computed in uint32_t, no signed overflow).
 * Returns nnz(C); cptr gets n_outer+1 entries in `base`; cidx/cval need room
 * for nnz(A)+nnz(B).
 * ------------------------------------------------------------------------- */
static inline float JL_MAX_f32(float x, float y) {
  uint32_t b;
  if (isnan(x)) { memcpy(&b, &x, 4); b |= 0x00400000u; memcpy(&x, &b, 4); return x; }
  if (isnan(y)) { memcpy(&b, &y, 4); b |= 0x00400000u; memcpy(&y, &b, 4); return y; }
  if (y > x || (y == x && !signbit(y) && signbit(x))) return y;
  return x;
}
static inline double JL_MAX_f64(double x, double y) {
  uint64_t b;
  if (isnan(x)) { memcpy(&b, &x, 8); b |= 0x0008000000000000ull; memcpy(&x, &b, 8); return x; }
  if (isnan(y)) { memcpy(&b, &y, 8); b |= 0x0008000000000000ull; memcpy(&y, &b, 8); return y; }
  if (y > x || (y == x && !signbit(y) && signbit(x))) return y;
  return x;
}
static inline int32_t JL_MAX_i32(int32_t x, int32_t y) { return y > x ? y : x; }
static inline float EW_ADD_f32(float a, float b) { return a + b; }
static inline double EW_ADD_f64(double a, double b) { return a + b; }
static inline int32_t EW_ADD_i32(int32_t a, int32_t b) {
  return (int32_t)((uint32_t)a + (uint32_t)b);
}
static inline float EW_MUL_f32(float a, float b) { return a * b; }
static inline double EW_MUL_f64(double a, double b) { return a * b; }
static inline int32_t EW_MUL_i32(int32_t a, int32_t b) {
  return (int32_t)((uint32_t)a * (uint32_t)b);
}
#define DEF_EWISE(T, SFX)                                                                 \
  i64 or_ewise_##SFX(int op, i64 n_outer, const i64* aptr, const i64* aidx, const T* aval,\
                     const i64* bptr, const i64* bidx, const T* bval, int base, i64* cptr,\
                     i64* cidx, T* cval) {                                                \
    i64 qos = 0;                                                                          \
    cptr[0] = base;                                                                       \
    for (i64 p = 0; p < n_outer; p++) {                                                   \
      i64 qa = aptr[p] - base, qa_stop = aptr[p + 1] - base;                              \
      i64 qb = bptr[p] - base, qb_stop = bptr[p + 1] - base;                              \
      while (qa < qa_stop && qb < qb_stop) {                                              \
        i64 ia = aidx[qa], ib = bidx[qb];                                                 \
        i64 i = ia < ib ? ia : ib;                                                        \
        if (ia == i && ib == i) {                                                         \
          T a = aval[qa], b = bval[qb];                                                   \
          cidx[qos] = i;                                                                  \
          cval[qos] = op == 0 ? EW_MUL_##SFX(a, b)                                           \
                      : (op == 1 ? EW_ADD_##SFX(b, a) : JL_MAX_##SFX(a, b));             \
          qos++; qa++; qb++;                                                              \
        } else if (ia == i) {                                                             \
          if (op != 0) {                                                                  \
            cidx[qos] = i;                                                                \
            cval[qos] = op == 1 ? aval[qa] : JL_MAX_##SFX(aval[qa], (T)0);                \
            qos++;                                                                        \
          }                                                                               \
          qa++;                                                                           \
        } else {                                                                          \
          if (op != 0) {                                                                  \
            cidx[qos] = i;                                                                \
            cval[qos] = op == 1 ? bval[qb] : JL_MAX_##SFX((T)0, bval[qb]);                \
            qos++;                                                                        \
          }                                                                               \
          qb++;                                                                           \
        }                                                                                 \
      }                                                                                   \
      if (op != 0) {                                                                      \
        for (; qa < qa_stop; qa++, qos++) {                                               \
          cidx[qos] = aidx[qa];                                                           \
          cval[qos] = op == 1 ? aval[qa] : JL_MAX_##SFX(aval[qa], (T)0);                  \
        }                                                                                 \
        for (; qb < qb_stop; qb++, qos++) {                                               \
          cidx[qos] = bidx[qb];                                                           \
          cval[qos] = op == 1 ? bval[qb] : JL_MAX_##SFX((T)0, bval[qb]);                  \
        }                                                                                 \
      }                                                                                   \
      cptr[p + 1] = qos + base;                                                           \
    }                                                                                     \
    return qos;                                                                           \
  }
DEF_EWISE(double, f64)
DEF_EWISE(float, f32)
DEF_EWISE(int32_t, i32)

/* ---------------------------------------------------------------------------
 * Gustavson SpGEMM, one output fiber at a time:
 *   C .= 0; for j = _; w .= 0; for k = _, i = _; w[i] += A[i,k]*B[k,j];
 *                      for i = _; C[i,j] = w[i]
 * docs/examples/spgemm.jl:59-76; generated code
 * test/reference64/kernels/gustavsons.jl:52-182: a SparseByteMap workspace
 * (value vector w, touched flags tbl, list srt of touched coordinates), the
 * update `w[i] = A[i,k]*A[k,j] + w[i]` (line 141) for the stored k of the
 * fiber in ascending order and the stored i of fiber k, `sort!(srt)` (line
 * 182), then a copy of the touched entries into the output SparseList.  Here
 * fibers are CSR rows (the same nest over swizzled operands): for row i of A,
 * for stored k (ascending), for stored j of row k of B.
 * Semiring 0: plus-times (w starts at 0; `w = a*b + w`).
 * Semiring 1: min-plus (w starts at +Inf; `w = min(w, a + b)` with Julia's
 *             `min`, JL_MIN below), the
 *             tropical product with fill Inf (src/symbolic/symbolic.jl:146).
 * Int32 arithmetic wraps, as Julia's Int32 does.
 * ------------------------------------------------------------------------- */
static int cmp_i64(const void* a, const void* b) {
  i64 x = *(const i64*)a, y = *(const i64*)b;
  return (x > y) - (x < y);
}

/* Symbolic pass: structural nnz of each output row; cptr gets the prefix sum. */
i64 or_spgemm_symbolic(i64 m, i64 n, const i64* aptr, const i64* aidx, const i64* bptr,
                       const i64* bidx, int base, i64* cptr) {
  unsigned char* tbl = (unsigned char*)calloc((size_t)(n > 0 ? n : 1), 1);
  i64* srt = (i64*)malloc(sizeof(i64) * (size_t)(n > 0 ? n : 1));
  i64 total = 0;
  cptr[0] = base;
  for (i64 i = 0; i < m; i++) {
    i64 fill = 0;
    for (i64 qa = aptr[i] - base; qa < aptr[i + 1] - base; qa++) {
      i64 k = aidx[qa] - base;
      for (i64 qb = bptr[k] - base; qb < bptr[k + 1] - base; qb++) {
        i64 j = bidx[qb] - base;
        if (!tbl[j]) { tbl[j] = 1; srt[fill++] = j; }
      }
    }
    for (i64 t = 0; t < fill; t++) tbl[srt[t]] = 0;
    total += fill;
    cptr[i + 1] = total + base;
  }
  free(tbl);
  free(srt);
  return total;
}

/* Julia's `min` on floats (base/math.jl): NaN wins (returned as the quiet
 * NaN), -0.0 is below 0.0, otherwise the smaller value.  With these rules
 * `min` is associative and commutative, so a min-plus fold does not depend
 * on the order of its products. */
static inline float JL_MIN_f32(float x, float y) {
  if (isnan(x) || isnan(y)) return NAN;
  if (y < x || (y == x && signbit(y) && !signbit(x))) return y;
  return x;
}
static inline double JL_MIN_f64(double x, double y) {
  if (isnan(x) || isnan(y)) return (double)NAN;
  if (y < x || (y == x && signbit(y) && !signbit(x))) return y;
  return x;
}
#define DEF_SPGEMM(T, SFX, INF)                                                          \
  void or_spgemm_##SFX(int semiring, i64 m, i64 n, const i64* aptr, const i64* aidx,     \
                       const T* aval, const i64* bptr, const i64* bidx, const T* bval,   \
                       int base, const i64* cptr, i64* cidx, T* cval) {                  \
    T* w = (T*)malloc(sizeof(T) * (size_t)(n > 0 ? n : 1));                              \
    unsigned char* tbl = (unsigned char*)calloc((size_t)(n > 0 ? n : 1), 1);             \
    i64* srt = (i64*)malloc(sizeof(i64) * (size_t)(n > 0 ? n : 1));                      \
    const T init = semiring == 0 ? (T)0 : (INF);                                         \
    for (i64 j = 0; j < n; j++) w[j] = init;                                             \
    for (i64 i = 0; i < m; i++) {                                                        \
      i64 fill = 0;                                                                      \
      for (i64 qa = aptr[i] - base; qa < aptr[i + 1] - base; qa++) {                     \
        i64 k = aidx[qa] - base;                                                         \
        T a = aval[qa];                                                                  \
        for (i64 qb = bptr[k] - base; qb < bptr[k + 1] - base; qb++) {                   \
          i64 j = bidx[qb] - base;                                                       \
          if (semiring == 0) {                                                           \
            w[j] = SPG_MUL_##SFX(a, bval[qb]) + w[j];                                    \
          } else {                                                                       \
            w[j] = JL_MIN_##SFX(w[j], a + bval[qb]);                                     \
          }                                                                              \
          if (!tbl[j]) { tbl[j] = 1; srt[fill++] = j; }                                  \
        }                                                                                \
      }                                                                                  \
      qsort(srt, (size_t)fill, sizeof(i64), cmp_i64);                                    \
      i64 q = cptr[i] - base;                                                            \
      for (i64 t = 0; t < fill; t++, q++) {                                              \
        i64 j = srt[t];                                                                  \
        cidx[q] = j + base;                                                              \
        cval[q] = w[j];                                                                  \
        w[j] = init;                                                                     \
        tbl[j] = 0;                                                                      \
      }                                                                                  \
    }                                                                                    \
    free(w);                                                                             \
    free(tbl);                                                                           \
    free(srt);                                                                           \
  }
/* Int32 products and sums wrap modulo 2^32 like Julia's Int32. */
#define SPG_MUL_i32(a, b) ((int32_t)((uint32_t)(a) * (uint32_t)(b)))
#define SPG_MUL_f32(a, b) ((a) * (b))
#define SPG_MUL_f64(a, b) ((a) * (b))
static inline int32_t wrap_add_i32(int32_t a, int32_t b) {
  return (int32_t)((uint32_t)a + (uint32_t)b);
}
DEF_SPGEMM(float, f32, INFINITY)
DEF_SPGEMM(double, f64, INFINITY)

/* Int32 plus-times with explicit wrapping adds (signed overflow is undefined
 * in C, defined modulo 2^32 in Julia). */
void or_spgemm_i32(i64 m, i64 n, const i64* aptr, const i64* aidx, const int32_t* aval,
                   const i64* bptr, const i64* bidx, const int32_t* bval, int base,
                   const i64* cptr, i64* cidx, int32_t* cval) {
  int32_t* w = (int32_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t));
  unsigned char* tbl = (unsigned char*)calloc((size_t)(n > 0 ? n : 1), 1);
  i64* srt = (i64*)malloc(sizeof(i64) * (size_t)(n > 0 ? n : 1));
  for (i64 i = 0; i < m; i++) {
    i64 fill = 0;
    for (i64 qa = aptr[i] - base; qa < aptr[i + 1] - base; qa++) {
      i64 k = aidx[qa] - base;
      int32_t a = aval[qa];
      for (i64 qb = bptr[k] - base; qb < bptr[k + 1] - base; qb++) {
        i64 j = bidx[qb] - base;
        w[j] = wrap_add_i32(SPG_MUL_i32(a, bval[qb]), w[j]);
        if (!tbl[j]) { tbl[j] = 1; srt[fill++] = j; }
      }
    }
    qsort(srt, (size_t)fill, sizeof(i64), cmp_i64);
    i64 q = cptr[i] - base;
    for (i64 t = 0; t < fill; t++, q++) {
      i64 j = srt[t];
      cidx[q] = j + base;
      cval[q] = w[j];
      w[j] = 0;
      tbl[j] = 0;
    }
  }
  free(w);
  free(tbl);
  free(srt);
}

/* Number of host threads OpenMP would use (for reporting). */
int or_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---------------------------------------------------------------------------
 * Format conversion.
 *
 * Transpose: permutedims(A, (2, 1)) of a 2-D Dense(SparseList) tensor
 * (src/interface/copy.jl:64-71), i.e. the swizzled copy of copy.jl:32-40
 * into the other storage order.  Every stored entry is kept (a written
 * coordinate is stored even when it equals the fill, src/tensors/scalars.jl:
 * 209-218); output fiber j lists the input fibers that store j in ascending
 * order.  The textbook counting transpose: count, prefix sum, then place
 * the entries in input order.  Values are copied as `vbytes`-byte words.
 * ------------------------------------------------------------------------- */
void or_transpose(i64 n_outer, i64 n_inner, const i64* ptr, const i64* idx, const void* val,
                  int vbytes, int base, i64* tptr, i64* tidx, void* tval) {
  i64* next = (i64*)calloc((size_t)(n_inner + 1), sizeof(i64));
  const i64 nnz = ptr[n_outer] - ptr[0];
  for (i64 q = 0; q < nnz; q++) next[idx[q] - base + 1]++;
  for (i64 j = 0; j < n_inner; j++) next[j + 1] += next[j];
  for (i64 j = 0; j <= n_inner; j++) tptr[j] = next[j] + base;
  for (i64 o = 0; o < n_outer; o++) {
    for (i64 q = ptr[o] - base; q < ptr[o + 1] - base; q++) {
      const i64 j = idx[q] - base;
      const i64 p = next[j]++;
      tidx[p] = o + base;
      memcpy((char*)tval + p * vbytes, (const char*)val + q * vbytes, (size_t)vbytes);
    }
  }
  free(next);
}

/* Stable counting sort of perm[0..n) by key[perm[.]] in [0, nk). */
static void stable_count_sort(i64 n, i64 nk, const i64* key, i64* perm, i64* tmp) {
  i64* c = (i64*)calloc((size_t)(nk + 1), sizeof(i64));
  for (i64 t = 0; t < n; t++) c[key[perm[t]] + 1]++;
  for (i64 j = 0; j < nk; j++) c[j + 1] += c[j];
  for (i64 t = 0; t < n; t++) tmp[c[key[perm[t]]]++] = perm[t];
  memcpy(perm, tmp, sizeof(i64) * (size_t)n);
  free(c);
}

/* ---------------------------------------------------------------------------
 * COO -> compressed: fsparse(I, J, V, (m, n)) (src/interface/fsparse.jl:
 * 64-97) copied into Dense(SparseList).  Entries are ordered by (outer,
 * inner) with ties kept in input order (two stable counting sorts, inner
 * first), and each run of equal coordinates is combined with `+` as
 * foldl(combine, V[start:stop-1]) (fsparse.jl:85-90), left to right.  Int32
 * sums wrap.  Returns the number of distinct coordinates; cptr gets n_outer+1
 * entries.  (The reference's fsparse_lt, fsparse.jl:1-7, seeds its
 * comparison with `I[1][j] < I[1][j]`, so it orders by the trailing
 * coordinate only; the two agree whenever the leading coordinates arrive
 * sorted within each fibre.  The full (outer, inner) order here is that of
 * SparseArrays.sparse, which the fsparse docstring refers to.)
 * ------------------------------------------------------------------------- */
#define DEF_COO(T, SFX, ADD)                                                               \
  i64 or_coo_##SFX(i64 n_outer, i64 n_inner, i64 nnz, const i64* outer, const i64* inner, \
                   const T* val, int base, i64* cptr, i64* cidx, T* cval) {                \
    i64* perm = (i64*)malloc(sizeof(i64) * (size_t)(nnz > 0 ? nnz : 1));                 \
    i64* tmp = (i64*)malloc(sizeof(i64) * (size_t)(nnz > 0 ? nnz : 1));                  \
    i64* ko = (i64*)calloc((size_t)(nnz > 0 ? nnz : 1), sizeof(i64));                    \
    i64* ki = (i64*)calloc((size_t)(nnz > 0 ? nnz : 1), sizeof(i64));                    \
    for (i64 q = 0; q < nnz; q++) {                                                        \
      perm[q] = q;                                                                         \
      ko[q] = outer[q] - base;                                                             \
      ki[q] = inner[q] - base;                                                             \
    }                                                                                      \
    stable_count_sort(nnz, n_inner, ki, perm, tmp);                                        \
    stable_count_sort(nnz, n_outer, ko, perm, tmp);                                        \
    i64 u = 0;                                                                             \
    for (i64 o = 0; o <= n_outer; o++) cptr[o] = 0;                                        \
    for (i64 t = 0; t < nnz;) {                                                            \
      const i64 q = perm[t];                                                               \
      T acc = val[q];                                                                      \
      i64 s = t + 1;                                                                       \
      while (s < nnz && ko[perm[s]] == ko[q] && ki[perm[s]] == ki[q]) {                    \
        acc = ADD(acc, val[perm[s]]);                                                      \
        s++;                                                                               \
      }                                                                                    \
      cidx[u] = ki[q] + base;                                                              \
      cval[u] = acc;                                                                       \
      cptr[ko[q] + 1]++;                                                                   \
      u++;                                                                                 \
      t = s;                                                                               \
    }                                                                                      \
    cptr[0] = base;                                                                        \
    for (i64 o = 0; o < n_outer; o++) cptr[o + 1] += cptr[o];                              \
    free(perm);                                                                            \
    free(tmp);                                                                             \
    free(ko);                                                                              \
    free(ki);                                                                              \
    return u;                                                                              \
  }
#define COO_ADD(a, b) ((a) + (b))
DEF_COO(double, f64, COO_ADD)
DEF_COO(float, f32, COO_ADD)
DEF_COO(int32_t, i32, wrap_add_i32)

/* ---------------------------------------------------------------------------
 * dropfills(A) into the same format (src/interface/copy.jl:73-117): keep each
 * stored entry unless isequal(value, fill) — `isequal` treats NaN as equal
 * to NaN and -0.0 as different from 0.0 (copy.jl:98-101).  Returns the kept
 * count; cptr gets n_outer+1 entries.
 * ------------------------------------------------------------------------- */
#define DEF_DROPFILLS(T, SFX, ISEQ)                                                       \
  i64 or_dropfills_##SFX(i64 n_outer, const i64* ptr, const i64* idx, const T* val,      \
                         int base, T fill, i64* cptr, i64* cidx, T* cval) {               \
    i64 u = 0;                                                                            \
    cptr[0] = base;                                                                       \
    for (i64 o = 0; o < n_outer; o++) {                                                   \
      for (i64 q = ptr[o] - base; q < ptr[o + 1] - base; q++) {                           \
        if (ISEQ(val[q], fill)) continue;                                                 \
        cidx[u] = idx[q];                                                                 \
        cval[u] = val[q];                                                                 \
        u++;                                                                              \
      }                                                                                   \
      cptr[o + 1] = u + base;                                                             \
    }                                                                                     \
    return u;                                                                             \
  }
static inline int iseq_f64(double a, double b) {
  if (isnan(a) || isnan(b)) return isnan(a) && isnan(b);
  return a == b && signbit(a) == signbit(b);
}
static inline int iseq_f32(float a, float b) {
  if (isnan(a) || isnan(b)) return isnan(a) && isnan(b);
  return a == b && signbit(a) == signbit(b);
}
#define ISEQ_I32(a, b) ((a) == (b))
DEF_DROPFILLS(double, f64, iseq_f64)
DEF_DROPFILLS(float, f32, iseq_f32)
DEF_DROPFILLS(int32_t, i32, ISEQ_I32)
