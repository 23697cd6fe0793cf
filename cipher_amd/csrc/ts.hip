// ts.hip - FP64 time-integration arithmetic on gfx950.
//
// k_wrms_partial/k_wrms_final: weighted RMS error norm of PETSc's
//   TSErrorWeightedNorm (2-norm form, scalar atol/rtol):
//   sqrt( (1/N) sum_i ( |u_i - y_i| / (atol + rtol*max(|u_i|,|y_i|)) )^2 ),
//   as a deterministic two-pass reduction (fixed block partition, wavefront
//   shuffles, fixed-order final sum) so repeated calls give identical bits.
// k_axpy_n: Y = X + (((a0 K0 + a1 K1) + a2 K2) + ...), the stage/update
//   combination of an explicit RK step, streamed once through HBM.
#include <cmath>
#include <mutex>
#include <string>

#include "kernels.h"

int cipher_fail(int code, const std::string &msg);

namespace cipher {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// block blk of nblk: grid-stride sum over n doubles into partial[blk]
__device__ __forceinline__ void wrms_part(const double *__restrict__ U, const double *__restrict__ Y, long long n,
                                          double atol, double rtol, double *partial, long long blk, long long nblk) {
    __shared__ double red[4];
    double s = 0.0;
    const long long stride = nblk * blockDim.x;
    for (long long i = blk * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double u = U[i], y = Y[i];
        const double au = fabs(u), ay = fabs(y);
        const double tol = atol + rtol * (au > ay ? au : ay);
        const double e = fabs(u - y) / tol;
        s += e * e;
    }
    s = wave_sum(s);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blk] = ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ __launch_bounds__(256) void k_wrms_partial(const double *__restrict__ U, const double *__restrict__ Y,
                                                      long long n, double atol, double rtol, double *partial) {
    wrms_part(U, Y, n, atol, rtol, partial, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void wrms_fin(const double *partial, int npart, long long n, double *out) {
    __shared__ double red[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < npart; i += blockDim.x) s += partial[i];
    s = wave_sum(s);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double tot = ((red[0] + red[1]) + red[2]) + red[3];
        out[0] = n > 0 ? sqrt(tot / (double)n) : 0.0;
    }
}

__global__ __launch_bounds__(256) void k_wrms_final(const double *partial, int npart, long long n, double *out) {
    wrms_fin(partial, npart, n, out);
}

int wrms_parts(long long n) {
    long long p = (n + 4095) / 4096;
    return (int)(p < 1 ? 1 : (p > 1024 ? 1024 : p));
}

// per-member norms of a batch: the member's partial sums over its own
// segment of blocks (the same partition as a single-mesh call), then one
// final block per member
__global__ __launch_bounds__(256) void k_wrms_partial_batch(const BatchMember *__restrict__ mem,
                                                            const int32_t *__restrict__ live,
                                                            const int32_t *__restrict__ seg,
                                                            const int32_t *__restrict__ segmem, int nseg,
                                                            const double *U, const double *Y, double atol, double rtol) {
    const int k = seg_find(seg, nseg, blockIdx.x);
    const int mi = segmem[k];
    if (live && !live[mi]) return;
    const BatchMember &m = mem[mi];
    const long long n = m.a.ncells * (long long)m.c.D;
    wrms_part(U + m.xoff, Y + m.xoff, n, atol, rtol, m.partial, (long long)blockIdx.x - seg[k],
              (long long)(seg[k + 1] - seg[k]));
}

__global__ __launch_bounds__(256) void k_wrms_final_batch(const BatchMember *__restrict__ mem,
                                                          const int32_t *__restrict__ live,
                                                          const int32_t *__restrict__ segmem, double *norms) {
    const int mi = segmem[blockIdx.x];
    if (live && !live[mi]) return;
    const BatchMember &m = mem[mi];
    wrms_fin(m.partial, m.npart, m.a.ncells * (long long)m.c.D, norms + mi);
}

cipher_status launch_wrms_batch(const BatchLaunch &L, const double *U, const double *Y, double atol, double rtol,
                                double *norms, hipStream_t st) {
    const BatchSegs &s = L.dofs;
    if (s.nblocks == 0) return CIPHER_OK;
    hipLaunchKernelGGL(k_wrms_partial_batch, dim3((unsigned)s.nblocks), dim3(256), 0, st, L.mem, L.live, s.seg,
                       s.segmem, s.nseg, U, Y, atol, rtol);
    hipLaunchKernelGGL(k_wrms_final_batch, dim3((unsigned)s.nseg), dim3(256), 0, st, L.mem, L.live, s.segmem, norms);
    return hipGetLastError() == hipSuccess ? CIPHER_OK : CIPHER_ERR_HIP;
}

cipher_status launch_wrms(const double *U, const double *Y, long long n, double atol, double rtol,
                          double *partial, int npart, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_wrms_partial, dim3(npart), dim3(256), 0, st, U, Y, n, atol, rtol, partial);
    hipLaunchKernelGGL(k_wrms_final, dim3(1), dim3(256), 0, st, partial, npart, n, out);
    return hipGetLastError() == hipSuccess ? CIPHER_OK : CIPHER_ERR_HIP;
}

struct AxpyArgs {
    const double *K[8];
    double a[8];
};

template <int NK>
__global__ __launch_bounds__(256) void k_axpy(long long n, double *__restrict__ Y, const double *__restrict__ X,
                                              AxpyArgs p) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double t = 0.0;
        if (NK > 0) {
            t = p.a[0] * p.K[0][i];
#pragma unroll
            for (int j = 1; j < NK; j++) t += p.a[j] * p.K[j][i];
            Y[i] = X[i] + t;
        } else {
            Y[i] = X[i];
        }
    }
}

cipher_status launch_axpy_n(long long n, double *Y, const double *X, int nk, const double *a,
                            const double *const *K, hipStream_t st) {
    if (nk < 0 || nk > 8) return CIPHER_ERR_ARG;
    AxpyArgs p;
    for (int j = 0; j < 8; j++) {
        p.K[j] = j < nk ? K[j] : nullptr;
        p.a[j] = j < nk ? a[j] : 0.0;
    }
    long long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    switch (nk) {
    case 0: hipLaunchKernelGGL((k_axpy<0>), dim3((unsigned)blocks), dim3(256), 0, st, n, Y, X, p); break;
    case 1: hipLaunchKernelGGL((k_axpy<1>), dim3((unsigned)blocks), dim3(256), 0, st, n, Y, X, p); break;
    case 2: hipLaunchKernelGGL((k_axpy<2>), dim3((unsigned)blocks), dim3(256), 0, st, n, Y, X, p); break;
    case 3: hipLaunchKernelGGL((k_axpy<3>), dim3((unsigned)blocks), dim3(256), 0, st, n, Y, X, p); break;
    case 4: hipLaunchKernelGGL((k_axpy<4>), dim3((unsigned)blocks), dim3(256), 0, st, n, Y, X, p); break;
    case 5: hipLaunchKernelGGL((k_axpy<5>), dim3((unsigned)blocks), dim3(256), 0, st, n, Y, X, p); break;
    case 6: hipLaunchKernelGGL((k_axpy<6>), dim3((unsigned)blocks), dim3(256), 0, st, n, Y, X, p); break;
    case 7: hipLaunchKernelGGL((k_axpy<7>), dim3((unsigned)blocks), dim3(256), 0, st, n, Y, X, p); break;
    default: hipLaunchKernelGGL((k_axpy<8>), dim3((unsigned)blocks), dim3(256), 0, st, n, Y, X, p); break;
    }
    return hipGetLastError() == hipSuccess ? CIPHER_OK : CIPHER_ERR_HIP;
}

// Y = X + (((dt*A0) K0 + (dt*A1) K1) + ...) with dt read from device memory,
// so a captured stage sequence stays valid when the step size changes
template <int NK>
__global__ __launch_bounds__(256) void k_axpy_dt(long long n, double *__restrict__ Y, const double *__restrict__ X,
                                                 const double *__restrict__ dtp, AxpyArgs p) {
    const double dt = dtp[0];
    // a zero step is an attempt that does not run (the device-driven loop
    // has ended); the integrator never takes one otherwise
    if (dt == 0.0) return;
    double a[NK];
#pragma unroll
    for (int j = 0; j < NK; j++) a[j] = dt * p.a[j];
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double t = a[0] * p.K[0][i];
#pragma unroll
        for (int j = 1; j < NK; j++) t += a[j] * p.K[j][i];
        Y[i] = X[i] + t;
    }
}

// per-member step sizes: member mi's slice uses dtp[mi]
template <int NK>
__global__ __launch_bounds__(256) void k_axpy_dt_batch(const BatchMember *__restrict__ mem,
                                                       const int32_t *__restrict__ live,
                                                       const int32_t *__restrict__ seg,
                                                       const int32_t *__restrict__ segmem, int nseg, double *Y,
                                                       const double *X, const double *__restrict__ dtp, AxpyArgs p) {
    const int k = seg_find(seg, nseg, blockIdx.x);
    const int mi = segmem[k];
    if (live && !live[mi]) return;
    const BatchMember &m = mem[mi];
    const long long n = m.a.ncells * (long long)m.c.D;
    const long long o = m.xoff;
    const double dt = dtp[mi];
    double a[NK];
#pragma unroll
    for (int j = 0; j < NK; j++) a[j] = dt * p.a[j];
    const long long stride = (long long)(seg[k + 1] - seg[k]) * blockDim.x;
    for (long long i = ((long long)blockIdx.x - seg[k]) * blockDim.x + threadIdx.x; i < n; i += stride) {
        double t = a[0] * p.K[0][o + i];
#pragma unroll
        for (int j = 1; j < NK; j++) t += a[j] * p.K[j][o + i];
        Y[o + i] = X[o + i] + t;
    }
}

cipher_status launch_axpy_dt_batch(const BatchLaunch &L, double *Y, const double *X, const double *dt_dev, int nk,
                                   const double *A, const double *const *K, hipStream_t st) {
    if (nk < 1 || nk > 4) return CIPHER_ERR_ARG;
    const BatchSegs &s = L.dofs;
    if (s.nblocks == 0) return CIPHER_OK;
    AxpyArgs p;
    for (int j = 0; j < 8; j++) {
        p.K[j] = j < nk ? K[j] : nullptr;
        p.a[j] = j < nk ? A[j] : 0.0;
    }
#define CIPHER_AXB(NK)                                                                                       \
    hipLaunchKernelGGL((k_axpy_dt_batch<NK>), dim3((unsigned)s.nblocks), dim3(256), 0, st, L.mem, L.live, s.seg, \
                       s.segmem, s.nseg, Y, X, dt_dev, p)
    switch (nk) {
    case 1: CIPHER_AXB(1); break;
    case 2: CIPHER_AXB(2); break;
    case 3: CIPHER_AXB(3); break;
    default: CIPHER_AXB(4); break;
    }
#undef CIPHER_AXB
    return hipGetLastError() == hipSuccess ? CIPHER_OK : CIPHER_ERR_HIP;
}

// Y = X on the slices of the members whose mask entry is set
__global__ __launch_bounds__(256) void k_copy_batch(const BatchMember *__restrict__ mem,
                                                    const int32_t *__restrict__ seg,
                                                    const int32_t *__restrict__ segmem, int nseg, double *Y,
                                                    const double *X, const int32_t *__restrict__ mask) {
    const int k = seg_find(seg, nseg, blockIdx.x);
    const int mi = segmem[k];
    if (!mask[mi]) return;
    const BatchMember &m = mem[mi];
    const long long n = m.a.ncells * (long long)m.c.D;
    const long long o = m.xoff;
    const long long stride = (long long)(seg[k + 1] - seg[k]) * blockDim.x;
    for (long long i = ((long long)blockIdx.x - seg[k]) * blockDim.x + threadIdx.x; i < n; i += stride)
        Y[o + i] = X[o + i];
}

cipher_status launch_copy_batch(const BatchLaunch &L, double *Y, const double *X, const int32_t *mask,
                                hipStream_t st) {
    const BatchSegs &s = L.dofs;
    if (s.nblocks == 0) return CIPHER_OK;
    hipLaunchKernelGGL(k_copy_batch, dim3((unsigned)s.nblocks), dim3(256), 0, st, L.mem, s.seg, s.segmem, s.nseg, Y,
                       X, mask);
    return hipGetLastError() == hipSuccess ? CIPHER_OK : CIPHER_ERR_HIP;
}

cipher_status launch_axpy_dt(long long n, double *Y, const double *X, const double *dt_dev, int nk, const double *A,
                             const double *const *K, hipStream_t st) {
    if (nk < 1 || nk > 4) return CIPHER_ERR_ARG;
    AxpyArgs p;
    for (int j = 0; j < 8; j++) {
        p.K[j] = j < nk ? K[j] : nullptr;
        p.a[j] = j < nk ? A[j] : 0.0;
    }
    long long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    switch (nk) {
    case 1: hipLaunchKernelGGL((k_axpy_dt<1>), dim3((unsigned)blocks), dim3(256), 0, st, n, Y, X, dt_dev, p); break;
    case 2: hipLaunchKernelGGL((k_axpy_dt<2>), dim3((unsigned)blocks), dim3(256), 0, st, n, Y, X, dt_dev, p); break;
    case 3: hipLaunchKernelGGL((k_axpy_dt<3>), dim3((unsigned)blocks), dim3(256), 0, st, n, Y, X, dt_dev, p); break;
    default: hipLaunchKernelGGL((k_axpy_dt<4>), dim3((unsigned)blocks), dim3(256), 0, st, n, Y, X, dt_dev, p); break;
    }
    return hipGetLastError() == hipSuccess ? CIPHER_OK : CIPHER_ERR_HIP;
}

// ---------------------------------------------------------------------------
// cipher_wrms_norm_noalloc: the weighted RMS norm with its partial sums and
// result in a file-scope device array instead of a context's buffers, for
// callers that cannot allocate.  wrms_parts() never exceeds WRMS_MAX_PARTS,
// so the array covers every n; every call overwrites the partials it reads
// (nothing needs clearing).  Calls that overlap on one GPU share the array
// and race.
constexpr int WRMS_MAX_PARTS = 1024;
constexpr int WRMS_SCRATCH = WRMS_MAX_PARTS + 32;  // partials, then the norm
static __device__ __align__(256) double scratch[WRMS_SCRATCH];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch[MAX_DEVICES];

// the array's device address on the current device (looked up once per
// device; each GPU holds its own copy)
static int scratch_address(double **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return cipher_fail(CIPHER_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES) return cipher_fail(CIPHER_ERR_UNSUPPORTED, "device index beyond 64");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return cipher_fail(CIPHER_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return CIPHER_OK;
}

}  // namespace cipher

extern "C" int cipher_wrms_norm_noalloc(const double *U, const double *Y, int64_t n, double atol, double rtol,
                                        void *stream, double *norm_out) {
    using namespace cipher;
    if (!U || !Y || !norm_out || n < 0) return cipher_fail(CIPHER_ERR_ARG, "bad argument");
    const int npart = wrms_parts(n);
    if (npart > WRMS_MAX_PARTS)
        return cipher_fail(CIPHER_ERR_ARG, "error norm needs more partial sums than the static scratch holds");
    double *ptr = nullptr;
    int rc = scratch_address(&ptr);
    if (rc != CIPHER_OK) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    cipher_status s = launch_wrms(U, Y, n, atol, rtol, ptr, npart, ptr + WRMS_MAX_PARTS, st);
    if (s != CIPHER_OK) return cipher_fail(s, "wrms launch");
    hipError_t e = hipMemcpyAsync(norm_out, ptr + WRMS_MAX_PARTS, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return cipher_fail(CIPHER_ERR_HIP, std::string("norm copy: ") + hipGetErrorString(e));
    if (!std::isfinite(*norm_out)) return cipher_fail(CIPHER_ERR_NONFINITE, "error norm is not finite");
    return CIPHER_OK;
}
