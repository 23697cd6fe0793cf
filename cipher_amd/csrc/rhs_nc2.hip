// rhs_nc2.hip - right-hand-side kernels instantiated for 2 component(s).
#include "rhs_impl.h"

namespace cipher {
template cipher_status launch_NC<2>(const RhsLaunch &L, int precision, hipStream_t st);
}  // namespace cipher
