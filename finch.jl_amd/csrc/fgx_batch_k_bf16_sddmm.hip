// fgx_batch_k_bf16_sddmm.hip — batch kernel instantiated for bf16_t values with SDDMM units.
#include "fgx_batch_impl.hpp"

namespace fgx {
FGX_BATCH_DEFINE_SDDMM(batch_sddmm_bf16, bf16_t, float)
}  // namespace fgx
