// rhs.hip - dispatch of the right-hand side on the number of components
// (the kernels live in rhs_impl.h, instantiated once per component count in
// rhs_nc1.hip .. rhs_nc5.hip).
#include "kernels.h"

namespace cipher {

template <int NC>
cipher_status launch_NC(const RhsLaunch &L, int precision, hipStream_t st);

cipher_status launch_rhs(const RhsLaunch &L, int precision, hipStream_t st) {
    switch (L.cfg->ncp) {
    case 1: return launch_NC<1>(L, precision, st);
    case 2: return launch_NC<2>(L, precision, st);
    case 3: return launch_NC<3>(L, precision, st);
    case 4: return launch_NC<4>(L, precision, st);
    case 5: return launch_NC<5>(L, precision, st);
    default: return CIPHER_ERR_UNSUPPORTED;
    }
}

template <int NC>
cipher_status launch_batch_NC(const BatchLaunch &L, int precision, hipStream_t st);

cipher_status launch_rhs_batch(const BatchLaunch &L, int precision, hipStream_t st) {
    switch (L.ncp) {
    case 1: return launch_batch_NC<1>(L, precision, st);
    case 2: return launch_batch_NC<2>(L, precision, st);
    case 3: return launch_batch_NC<3>(L, precision, st);
    case 4: return launch_batch_NC<4>(L, precision, st);
    case 5: return launch_batch_NC<5>(L, precision, st);
    default: return CIPHER_ERR_UNSUPPORTED;
    }
}

template <int NC>
cipher_status launch_output_NC(const DevCfg &c, const double *X, double *cm, long long N, const OutFields &f,
                               double *out, hipStream_t st);

cipher_status launch_output(const DevCfg &c, const double *X, double *cm, long long ncells, const OutFields &f,
                            double *out, hipStream_t st) {
    switch (c.ncp) {
    case 1: return launch_output_NC<1>(c, X, cm, ncells, f, out, st);
    case 2: return launch_output_NC<2>(c, X, cm, ncells, f, out, st);
    case 3: return launch_output_NC<3>(c, X, cm, ncells, f, out, st);
    case 4: return launch_output_NC<4>(c, X, cm, ncells, f, out, st);
    case 5: return launch_output_NC<5>(c, X, cm, ncells, f, out, st);
    default: return CIPHER_ERR_UNSUPPORTED;
    }
}

// fold every T-series at the temperature of the first cell (deterministic:
// the same code as the direct evaluation)
__global__ void k_tsfold(const double *X, long long ncells, DevCfg c, double *tsv) {
    const double T = ncells > 0 ? X[c.tmo] : 0.0;
    for (int i = threadIdx.x; i < c.ntseries; i += blockDim.x) tsv[i] = dtseries<double>(&c.ts[i], T);
    if (threadIdx.x == 0) tsv[c.ntseries] = T;
}

cipher_status launch_tsfold(const double *X, long long ncells, const DevCfg &c, double *tsv, hipStream_t st) {
    hipLaunchKernelGGL(k_tsfold, dim3(1), dim3(256), 0, st, X, ncells, c, tsv);
    return hipGetLastError() == hipSuccess ? CIPHER_OK : CIPHER_ERR_HIP;
}

// the same per member (one block each) at the member's first cell
__global__ void k_tsfold_batch(const BatchMember *__restrict__ mem, const int32_t *__restrict__ live,
                               const int32_t *__restrict__ seg, const int32_t *__restrict__ segmem, int nseg,
                               const double *X) {
    const int mi = segmem[seg_find(seg, nseg, blockIdx.x)];
    if (live && !live[mi]) return;
    const BatchMember &m = mem[mi];
    if (!m.fold) return;
    const DevCfg &c = m.c;
    const double T = m.a.ncells > 0 ? X[m.xoff + c.tmo] : 0.0;
    double *tsv = const_cast<double *>(c.tsv);
    for (int i = threadIdx.x; i < c.ntseries; i += blockDim.x) tsv[i] = dtseries<double>(&c.ts[i], T);
    if (threadIdx.x == 0) tsv[c.ntseries] = T;
}

cipher_status launch_fold_batch(const BatchLaunch &L, hipStream_t st) {
    if (L.fold.nblocks == 0) return CIPHER_OK;
    hipLaunchKernelGGL(k_tsfold_batch, dim3((unsigned)L.fold.nblocks), dim3(256), 0, st, L.mem, L.live, L.fold.seg,
                       L.fold.segmem, L.fold.nseg, L.X);
    return hipGetLastError() == hipSuccess ? CIPHER_OK : CIPHER_ERR_HIP;
}

size_t prep_record_bytes(int ncp) { return (size_t)(2 * ncp + 1) * sizeof(double); }

}  // namespace cipher
