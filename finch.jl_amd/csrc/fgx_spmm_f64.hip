// fgx_spmm_f64.hip — SpMM kernels instantiated for double values.
#include "fgx_spmm_impl.hpp"

namespace fgx {
FGX_SPMM_DEFINE(spmm_csr_f64, double)
}  // namespace fgx
