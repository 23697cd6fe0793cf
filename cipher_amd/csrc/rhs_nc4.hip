// rhs_nc4.hip - right-hand-side kernels instantiated for 4 component(s).
#include "rhs_impl.h"

namespace cipher {
template cipher_status launch_NC<4>(const RhsLaunch &L, int precision, hipStream_t st);
}  // namespace cipher
