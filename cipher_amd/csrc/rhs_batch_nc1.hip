// rhs_batch_nc1.hip - batched right-hand-side kernels for 1 component(s).
#include "rhs_batch.h"

namespace cipher {
template cipher_status launch_batch_NC<1>(const BatchLaunch &L, int precision, hipStream_t st);
}  // namespace cipher
