// fgx_batch_k_f32.hip — batch kernel instantiated for float values.
#include "fgx_batch_impl.hpp"

namespace fgx {
FGX_BATCH_DEFINE(batch_kernel_f32, float, false)
}  // namespace fgx
