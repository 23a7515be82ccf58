// fgx_batch_k_bf16_in_sddmm.hip — batch kernel instantiated for bf16_t values and bf16_t dense
// factors with SDDMM units.
#include "fgx_batch_impl.hpp"

namespace fgx {
FGX_BATCH_DEFINE_SDDMM(batch_sddmm_bf16_bf16in, bf16_t, bf16_t)
}  // namespace fgx
