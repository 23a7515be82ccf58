// fgx_batch_k_f32_sddmm.hip — batch kernel instantiated for float values with SDDMM units.
#include "fgx_batch_impl.hpp"

namespace fgx {
FGX_BATCH_DEFINE_SDDMM(batch_sddmm_f32, float, float)
}  // namespace fgx
