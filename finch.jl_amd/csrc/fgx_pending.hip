// fgx_pending.hip — entry points whose kernels are not built yet; they
// report FGX_EUNSUP.
#include "fgx_internal.hpp"

using namespace fgx;

extern "C" {
int fgx_ewise_symbolic(fgx_ctx*, int, const fgx_compressed*, const fgx_compressed*, void*, int64_t*, void*) { return set_error(FGX_EUNSUP, "fgx_ewise_symbolic: not built"); }
int fgx_ewise_numeric(fgx_ctx*, int, const fgx_compressed*, const fgx_compressed*, const void*, void*, void*, void*) { return set_error(FGX_EUNSUP, "fgx_ewise_numeric: not built"); }
int fgx_spgemm_symbolic(fgx_ctx*, int, const fgx_compressed*, const fgx_compressed*, void*, int64_t*, void*) { return set_error(FGX_EUNSUP, "not built"); }
int fgx_spgemm_numeric(fgx_ctx*, int, const fgx_compressed*, const fgx_compressed*, const void*, void*, void*, void*) { return set_error(FGX_EUNSUP, "not built"); }
int fgx_batch_plan_create(fgx_ctx*, const fgx_batch_item*, int64_t, int, fgx_batch_plan**) { return set_error(FGX_EUNSUP, "not built"); }
int fgx_batch_plan_run(fgx_ctx*, fgx_batch_plan*, void*) { return set_error(FGX_EUNSUP, "not built"); }
int fgx_batch_plan_stats(const fgx_batch_plan*, int64_t*) { return set_error(FGX_EUNSUP, "not built"); }
int fgx_batch_plan_destroy(fgx_batch_plan*) { return FGX_OK; }
int fgx_batch_run_loop(fgx_ctx*, const fgx_batch_item*, int64_t, void*) { return set_error(FGX_EUNSUP, "not built"); }
int fgx_serve_run(fgx_ctx*, const fgx_compressed*, int64_t, const fgx_request*, int64_t, const fgx_serve_config*, fgx_request_result*) { return set_error(FGX_EUNSUP, "not built"); }
}
