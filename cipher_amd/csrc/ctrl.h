// ctrl.h - step-size controller of the adaptive integrator, shared by the
// host (cipher_ts_adapt) and the device (the step kernels that decide
// acceptance and the next step size without a host round trip).
//
// Both sides must take the same decisions from the same error norms, so the
// controller uses IEEE binary64 additions, multiplications, divisions,
// comparisons and bit manipulation only: no libm function, whose host and
// device implementations may differ in the last bit.  log and exp are
// evaluated from their series after an exact argument reduction, arctan from
// its series after the tan(pi/12) reduction (see the functions).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cipher_amd.h"

namespace cipher {

__host__ __device__ inline double ctl_from_bits(uint64_t u) { return __builtin_bit_cast(double, u); }
__host__ __device__ inline uint64_t ctl_to_bits(double d) { return __builtin_bit_cast(uint64_t, d); }

// natural logarithm of a positive normal x: x = m 2^e with m in
// [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m - 1)/(m + 1), |s| < 0.172,
// the atanh series 2 (s + s^3/3 + ... + s^23/23) (next term < 1e-19)
__host__ __device__ inline double ctl_log(double x) {
    const uint64_t u = ctl_to_bits(x);
    int e = (int)((u >> 52) & 0x7ff) - 1023;
    double m = ctl_from_bits((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 1.4142135623730951) {
        m *= 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    double p = 1.0 / 23.0;
    for (int k = 21; k >= 1; k -= 2) p = p * s2 + 1.0 / (double)k;
    return (double)e * 0.6931471805599453 + 2.0 * s * p;
}

// exponential, |y| clamped to 700: y = k ln2 + r with k the nearest integer
// of y/ln2, exp r from its Taylor series to r^20 (|r| < 0.35, next term
// < 1e-28), times 2^k built from its exponent bits
__host__ __device__ inline double ctl_exp(double y) {
    if (y > 700.0) y = 700.0;
    if (y < -700.0) y = -700.0;
    const double k = floor(y * 1.4426950408889634 + 0.5);
    const double r = y - k * 0.6931471805599453;
    double p = 1.0;
    for (int i = 20; i >= 1; i--) p = 1.0 + p * r / (double)i;
    return p * ctl_from_bits((uint64_t)((int)k + 1023) << 52);
}

// x^a for x > 0 (x clamped to [1e-300, 1e300])
__host__ __device__ inline double ctl_pow(double x, double a) {
    if (x < 1e-300) x = 1e-300;
    if (x > 1e300) x = 1e300;
    return ctl_exp(a * ctl_log(x));
}

// arctangent: odd; atan x = pi/2 - atan(1/x) for x > 1; atan x = pi/6 +
// atan((sqrt3 x - 1)/(sqrt3 + x)) for x > tan(pi/12); then the series
// x (1 - x^2/3 + x^4/5 - ... - x^28/29) for |x| <= tan(pi/12) (next term
// < 1e-18)
__host__ __device__ inline double ctl_atan(double x) {
    const bool neg = x < 0.0;
    if (neg) x = -x;
    const bool inv = x > 1.0;
    if (inv) x = 1.0 / x;
    const bool red = x > 0.2679491924311227;
    if (red) x = (x * 1.7320508075688772 - 1.0) / (1.7320508075688772 + x);
    const double u = -(x * x);
    double p = 1.0 / 29.0;
    for (int k = 27; k >= 1; k -= 2) p = p * u + 1.0 / (double)k;
    double r = x * p;
    if (red) r += 0.5235987755982988;
    if (inv) r = 1.5707963267948966 - r;
    return neg ? -r : r;
}

// Acceptance and next step size for an attempt of size dt with error norm
// err (PETSc's TSAdapt: accept if err <= 1 or dt is at the minimum; a PI
// controller, Soederlind's PI.4.2 with k*beta1 = 3/5, k*beta2 = -1/5, k the
// order of the method (3 for Bogacki-Shampine 3(2), 2 for Heun 2(1)), the
// elementary controller (exponent 1/k) without history, and the arctan
// limiter rho = 1 + atan(r - 1);
// safety factor, extra factor after consecutive rejections, ratio clip and
// step limits).  Returns CIPHER_ERR_NONFINITE for a non-finite err at the
// minimum step (the step is then shrunk as far as allowed).
__host__ __device__ inline int ctl_adapt(const cipher_ts_params &p, cipher_ts_ctrl &st, double dt, double err,
                                         int *accept, double *dt_next) {
    const bool at_min = dt <= (1.0 + 1.4901161193847656e-08) * p.dt_min;
    double safety = st.safety > 0.0 ? st.safety : p.safety;
    if (!(err - err == 0.0)) {
        // a stage blew up: shrink as far as allowed, fail at the minimum
        *accept = 0;
        st.rejected_last = 1;
        const double v = dt * p.clip_lo;
        *dt_next = v > p.dt_min ? v : p.dt_min;
        return at_min ? CIPHER_ERR_NONFINITE : CIPHER_OK;
    }
    const bool acc = err <= 1.0 || at_min;
    if (!acc && st.rejected_last) safety *= p.reject_safety;
    const double k = p.method == CIPHER_RK_2A ? 2.0 : 3.0;
    double rho;
    if (err > 0.0) {
        double r = ctl_pow(1.0 / err, (st.err_prev > 0.0 ? 0.6 : 1.0) / k);
        if (st.err_prev > 0.0) r *= ctl_pow(1.0 / st.err_prev, -0.2 / k);
        rho = 1.0 + ctl_atan(r - 1.0);
    } else {
        rho = p.clip_hi;
    }
    double ratio = safety * rho;
    if (ratio < p.clip_lo) ratio = p.clip_lo;
    if (ratio > p.clip_hi) ratio = p.clip_hi;
    double dn = dt * ratio;
    if (dn < p.dt_min) dn = p.dt_min;
    if (dn > p.dt_max) dn = p.dt_max;
    *dt_next = dn;
    *accept = acc ? 1 : 0;
    if (acc) {
        st.err_prev = err;
        st.safety = p.safety;
        st.rejected_last = 0;
    } else {
        st.safety = safety;
        st.rejected_last = 1;
    }
    return CIPHER_OK;
}

// Device-resident state of a solve: the time loop of cipher_ts_solve
// (exact final time, step and rejection counts, controller memory) advanced
// by the step kernels, read by the host once per group of attempts.
struct TsDev {
    double t, t_end, dt, dt_try, err;
    cipher_ts_ctrl ctrl;
    int64_t steps, rejected, rhs_evals, max_steps;
    int32_t done;       // the loop has ended (time reached, step limit, error)
    int32_t active;     // the current attempt runs
    int32_t status;     // CIPHER_OK, CIPHER_ERR_CAPACITY or CIPHER_ERR_NONFINITE
    int32_t empty;      // no cells: nothing to integrate
};

// start of an attempt: the step size, with the exact final time
__device__ inline void ts_begin(TsDev &s, int32_t post_err, double *dt_dev, int32_t *live) {
    if (!s.done && post_err) {
        s.done = 1;
        s.status = CIPHER_ERR_CAPACITY;
    }
    double dt_try = 0.0;
    int act = 0;
    if (!s.done) {
        const bool more = s.t < s.t_end && (s.max_steps <= 0 || s.steps < s.max_steps) && !s.empty;
        if (more) {
            dt_try = (s.t + s.dt > s.t_end) ? s.t_end - s.t : s.dt;
            act = dt_try > 0.0 ? 1 : 0;
        }
        if (!act) {
            s.done = 1;
            dt_try = 0.0;
        }
    }
    s.dt_try = dt_try;
    s.active = act;
    *dt_dev = dt_try;
    *live = act;
}

// end of an attempt: acceptance, time, counts and the next step size
__device__ inline void ts_decide(TsDev &s, const cipher_ts_params &p, double err, int32_t *acc) {
    *acc = 0;
    if (!s.active) return;
    s.err = err;
    s.rhs_evals += p.method == CIPHER_RK_2A ? 2 : 4;
    int a = 0;
    double dtn = s.dt_try;
    const int rc = ctl_adapt(p, s.ctrl, s.dt_try, err, &a, &dtn);
    if (rc) {
        s.status = rc;
        s.done = 1;
        return;
    }
    if (a) {
        s.t = (s.dt_try == s.t_end - s.t) ? s.t_end : s.t + s.dt_try;
        s.steps++;
    } else {
        s.rejected++;
    }
    s.dt = dtn;
    *acc = a;
}

}  // namespace cipher
