// rhs_inst.hip - explicit instantiation of the single-mesh right-hand-side
// kernels for one component count CIPHER_NC, compiled once per lane capacity
// CIPHER_G (4, 8 or 16: that capacity's kernels) and once with CIPHER_G = 0
// (the per-capacity dispatcher and the output-field kernels), so the build
// spreads the kernel bodies over many translation units.
#include "rhs_impl.h"

namespace cipher {
#if CIPHER_G == 0
extern template cipher_status launch_G<4, CIPHER_NC>(const RhsLaunch &L, int precision, hipStream_t st);
extern template cipher_status launch_G<8, CIPHER_NC>(const RhsLaunch &L, int precision, hipStream_t st);
extern template cipher_status launch_G<16, CIPHER_NC>(const RhsLaunch &L, int precision, hipStream_t st);
template cipher_status launch_NC<CIPHER_NC>(const RhsLaunch &L, int precision, hipStream_t st);
template cipher_status launch_output_NC<CIPHER_NC>(const DevCfg &c, const double *X, double *cm, long long N,
                                                   const OutFields &f, double *out, hipStream_t st);
#else
template cipher_status launch_G<CIPHER_G, CIPHER_NC>(const RhsLaunch &L, int precision, hipStream_t st);
#endif
}  // namespace cipher
