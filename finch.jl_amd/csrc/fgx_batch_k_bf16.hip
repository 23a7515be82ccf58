// fgx_batch_k_bf16.hip — batch kernel instantiated for bf16_t values.
#include "fgx_batch_impl.hpp"

namespace fgx {
FGX_BATCH_DEFINE(batch_kernel_bf16, bf16_t, false)
}  // namespace fgx
