// rhs_nc2.hip - right-hand-side kernels instantiated for 2 component(s).
#include "rhs_impl.h"

namespace cipher {
template cipher_status launch_NC<2>(const RhsLaunch &L, int precision, hipStream_t st);
template cipher_status launch_output_NC<2>(const DevCfg &c, const double *X, double *cm, long long N,
                                             const OutFields &f, double *out, hipStream_t st);
}  // namespace cipher
