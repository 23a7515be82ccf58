// fgx_spmm_f16.hip — SpMM kernels instantiated for f16_t values.
#include "fgx_spmm_impl.hpp"

namespace fgx {
FGX_SPMM_DEFINE(spmm_csr_f16, f16_t)
}  // namespace fgx
