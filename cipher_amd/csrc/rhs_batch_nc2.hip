// rhs_batch_nc2.hip - batched right-hand-side kernels for 2 component(s).
#include "rhs_batch.h"

namespace cipher {
template cipher_status launch_batch_NC<2>(const BatchLaunch &L, int precision, hipStream_t st);
}  // namespace cipher
