// fgx_batch_k_f32.hip — batch kernel instantiated for float values.
#include "fgx_batch_impl.hpp"

namespace fgx {
FGX_BATCH_DEFINE_SPMM(batch_spmm_f32, float, float)
}  // namespace fgx
