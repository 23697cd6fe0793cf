// rhs_batch_nc4.hip - batched right-hand-side kernels for 4 component(s).
#include "rhs_batch.h"

namespace cipher {
template cipher_status launch_batch_NC<4>(const BatchLaunch &L, int precision, hipStream_t st);
}  // namespace cipher
