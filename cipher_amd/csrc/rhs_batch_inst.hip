// rhs_batch_inst.hip - explicit instantiation of the batched right-hand-side
// kernels for one component count CIPHER_NC, compiled once per lane capacity
// CIPHER_G (4, 8 or 16) and once with CIPHER_G = 0 (the dispatcher).
#include "rhs_batch.h"

namespace cipher {
#if CIPHER_G == 0
extern template cipher_status batch_G<4, CIPHER_NC>(const BatchLaunch &L, int precision, hipStream_t st);
extern template cipher_status batch_G<8, CIPHER_NC>(const BatchLaunch &L, int precision, hipStream_t st);
extern template cipher_status batch_G<16, CIPHER_NC>(const BatchLaunch &L, int precision, hipStream_t st);
template cipher_status launch_batch_NC<CIPHER_NC>(const BatchLaunch &L, int precision, hipStream_t st);
#else
template cipher_status batch_G<CIPHER_G, CIPHER_NC>(const BatchLaunch &L, int precision, hipStream_t st);
#endif
}  // namespace cipher
