// fgx_spmm_f32.hip — SpMM kernels instantiated for float values.
#include "fgx_spmm_impl.hpp"

namespace fgx {
FGX_SPMM_DEFINE(spmm_csr_f32, float)
}  // namespace fgx
