// fgx_batch_k_bf16_in.hip — batch kernel instantiated for bf16_t values and bf16_t dense inputs.
#include "fgx_batch_impl.hpp"

namespace fgx {
FGX_BATCH_DEFINE_SPMM(batch_spmm_bf16_bf16in, bf16_t, bf16_t)
}  // namespace fgx
