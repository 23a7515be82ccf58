// fgx_pending.hip — entry points whose kernels are not built yet; they
// report FGX_EUNSUP.
#include "fgx_internal.hpp"

using namespace fgx;

extern "C" {
int fgx_serve_run(fgx_ctx*, const fgx_compressed*, int64_t, const fgx_request*, int64_t, const fgx_serve_config*, fgx_request_result*) { return set_error(FGX_EUNSUP, "not built"); }
}
