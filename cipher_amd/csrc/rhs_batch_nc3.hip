// rhs_batch_nc3.hip - batched right-hand-side kernels for 3 component(s).
#include "rhs_batch.h"

namespace cipher {
template cipher_status launch_batch_NC<3>(const BatchLaunch &L, int precision, hipStream_t st);
}  // namespace cipher
