// fgx_batch_k_bf16.hip — batch kernel instantiated for bf16_t values.
#include "fgx_batch_impl.hpp"

namespace fgx {
FGX_BATCH_DEFINE_SPMM(batch_spmm_bf16, bf16_t, float)
}  // namespace fgx
