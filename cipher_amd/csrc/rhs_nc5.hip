// rhs_nc5.hip - right-hand-side kernels instantiated for 5 component(s).
#include "rhs_impl.h"

namespace cipher {
template cipher_status launch_NC<5>(const RhsLaunch &L, int precision, hipStream_t st);
}  // namespace cipher
