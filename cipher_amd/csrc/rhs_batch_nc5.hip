// rhs_batch_nc5.hip - batched right-hand-side kernels for 5 component(s).
#include "rhs_batch.h"

namespace cipher {
template cipher_status launch_batch_NC<5>(const BatchLaunch &L, int precision, hipStream_t st);
}  // namespace cipher
