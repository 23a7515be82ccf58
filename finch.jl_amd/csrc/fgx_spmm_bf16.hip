// fgx_spmm_bf16.hip — SpMM kernels instantiated for bf16_t values.
#include "fgx_spmm_impl.hpp"

namespace fgx {
FGX_SPMM_DEFINE(spmm_csr_bf16, bf16_t)
}  // namespace fgx
