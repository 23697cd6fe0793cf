// integrator.hip - adaptive explicit Runge-Kutta on device vectors.
//
// The integrator CIPHER runs in PETSc (TSRK 3bs, TSADAPTDSP, scalar
// tolerances and step limits, src/cipher.c:1097-1115; the load-case loop
// with an exact final time, src/cipher.c:1589-1606; PostStep after each
// accepted step, src/cipher.c:1108):
//
//   Bogacki-Shampine 3(2):  K1 = f(X)
//                           K2 = f(X + dt/2 K1)
//                           K3 = f(X + 3dt/4 K2)
//                           Xn = X + dt (2/9 K1 + 1/3 K2 + 4/9 K3),  K4 = f(Xn)
//                           Xh = X + dt (7/24 K1 + 1/4 K2 + 1/3 K3 + 1/8 K4)
//   error  = WRMS(Xn, Xh) with tol_i = atol + rtol max(|Xn_i|, |Xh_i|)
//   accept = error <= 1 (or dt at dt_min), then X = Xn and PostStep.
//
// The first stage is always evaluated: PostStep changes the solution after
// the step, so the last stage of the previous step is not f(X).
//
// An attempt is a fixed sequence of launches whose only varying input is dt,
// kept in device memory, so it can be captured once in a HIP graph and
// replayed; the only host synchronisation per attempt is the error norm.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "ctx.h"

using namespace cipher;

#define TSCHK(expr)                                                                                \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return cipher_fail(CIPHER_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_));   \
    } while (0)

// Butcher tableau of Bogacki and Shampine (1989)
static const double A2[1] = {1.0 / 2.0};
static const double A3[1] = {3.0 / 4.0};
static const double B[3] = {2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0};
static const double BH[4] = {7.0 / 24.0, 1.0 / 4.0, 1.0 / 3.0, 1.0 / 8.0};

void cipher_ts_invalidate(cipher_ctx *c);

void cipher_ts_free(cipher_ctx *c) {
    TsState &s = c->ts;
    cipher_ts_invalidate(c);
    double *v[] = {s.K[0], s.K[1], s.K[2], s.K[3], s.Y, s.Xn, s.Xh, s.dt_dev};
    for (double *p : v)
        if (p) (void)hipFree(p);
    for (int i = 0; i < 4; i++) s.K[i] = nullptr;
    s.Y = s.Xn = s.Xh = s.dt_dev = nullptr;
    if (s.h_dt) (void)hipHostFree(s.h_dt);
    s.h_dt = nullptr;
    if (s.own) (void)hipStreamDestroy(s.own);
    s.own = nullptr;
    s.n = 0;
    s.ready = false;
}

void cipher_ts_invalidate(cipher_ctx *c) {
    TsState &s = c->ts;
    if (s.exec) (void)hipGraphExecDestroy(s.exec);
    if (s.graph) (void)hipGraphDestroy(s.graph);
    s.exec = nullptr;
    s.graph = nullptr;
    s.graph_X = nullptr;
}

extern "C" int cipher_ts_adapt(const cipher_ts_params *p, cipher_ts_ctrl *st, double dt, double err, int *accept,
                               double *dt_next) {
    if (!p || !st || !accept || !dt_next || !(dt > 0.0)) return cipher_fail(CIPHER_ERR_ARG, "bad argument");
    const bool at_min = dt <= (1.0 + 1.4901161193847656e-08) * p->dt_min;
    double safety = st->safety > 0.0 ? st->safety : p->safety;
    if (!std::isfinite(err)) {
        // a stage blew up: shrink as far as allowed, fail at the minimum
        *accept = 0;
        st->rejected_last = 1;
        *dt_next = std::fmax(dt * p->clip_lo, p->dt_min);
        if (at_min) return cipher_fail(CIPHER_ERR_NONFINITE, "error norm is not finite at the minimum step");
        return CIPHER_OK;
    }
    const bool acc = err <= 1.0 || at_min;
    if (!acc && st->rejected_last) safety *= p->reject_safety;
    // PI controller (Soederlind's PI.4.2: k*beta1 = 3/5, k*beta2 = -1/5,
    // k = 3 for the second-order error estimate), elementary controller
    // (exponent 1/k) without history, arctan limiter rho = 1 + atan(r - 1)
    const double k = 3.0;
    double rho;
    if (err > 0.0) {
        double r = std::pow(1.0 / err, (st->err_prev > 0.0 ? 0.6 : 1.0) / k);
        if (st->err_prev > 0.0) r *= std::pow(1.0 / st->err_prev, -0.2 / k);
        rho = 1.0 + std::atan(r - 1.0);
    } else {
        rho = p->clip_hi;
    }
    double ratio = safety * rho;
    ratio = std::fmin(std::fmax(ratio, p->clip_lo), p->clip_hi);
    *dt_next = std::fmin(std::fmax(dt * ratio, p->dt_min), p->dt_max);
    *accept = acc ? 1 : 0;
    if (acc) {
        st->err_prev = err;
        st->safety = p->safety;
        st->rejected_last = 0;
    } else {
        st->safety = safety;
        st->rejected_last = 1;
    }
    return CIPHER_OK;
}

extern "C" int cipher_ts_setup(cipher_ctx *c, const cipher_ts_params *p) {
    if (!c || !p) return cipher_fail(CIPHER_ERR_ARG, "null argument");
    if (!c->cell_off) return cipher_fail(CIPHER_ERR_STATE, "no mesh uploaded");
    if (!(p->atol >= 0.0) || !(p->rtol >= 0.0) || !(p->atol > 0.0 || p->rtol > 0.0) || !(p->dt_min > 0.0) ||
        !(p->dt_max >= p->dt_min) || !(p->safety > 0.0) || !(p->reject_safety > 0.0) || !(p->clip_lo > 0.0) ||
        !(p->clip_hi >= p->clip_lo))
        return cipher_fail(CIPHER_ERR_ARG, "integrator parameters out of range");
    cipher_ts_free(c);
    TsState &s = c->ts;
    s.p = *p;
    s.ctrl = cipher_ts_ctrl{};
    s.n = c->ncells * (long long)c->L.ndof;
    const size_t bytes = sizeof(double) * (size_t)(s.n > 0 ? s.n : 1);
    for (int i = 0; i < 4; i++) TSCHK(hipMalloc(&s.K[i], bytes));
    TSCHK(hipMalloc(&s.Y, bytes));
    TSCHK(hipMalloc(&s.Xn, bytes));
    TSCHK(hipMalloc(&s.Xh, bytes));
    TSCHK(hipMalloc(&s.dt_dev, sizeof(double)));
    TSCHK(hipHostMalloc(&s.h_dt, sizeof(double), hipHostMallocDefault));
    s.ready = true;
    return CIPHER_OK;
}

// the launches of one attempt, in order, on the context's stream
static int enqueue_attempt(cipher_ctx *c, const double *X) {
    TsState &s = c->ts;
    hipStream_t st = c->stream;
    const long long n = s.n;
    int rc;
    const double *k1[1] = {s.K[0]};
    const double *k2[1] = {s.K[1]};
    const double *k123[3] = {s.K[0], s.K[1], s.K[2]};
    const double *k1234[4] = {s.K[0], s.K[1], s.K[2], s.K[3]};
    // the stage values keep X's active sets (the rate of an AS entry is 0),
    // so the cells are binned once per attempt and the lists reused
    if ((rc = cipher_rhs_impl(c, X, s.K[0], nullptr, 1))) return rc;
    if (launch_axpy_dt(n, s.Y, X, s.dt_dev, 1, A2, k1, st)) return cipher_fail(CIPHER_ERR_HIP, "stage 2");
    if ((rc = cipher_rhs_impl(c, s.Y, s.K[1], nullptr, 2))) return rc;
    if (launch_axpy_dt(n, s.Y, X, s.dt_dev, 1, A3, k2, st)) return cipher_fail(CIPHER_ERR_HIP, "stage 3");
    if ((rc = cipher_rhs_impl(c, s.Y, s.K[2], nullptr, 2))) return rc;
    if (launch_axpy_dt(n, s.Xn, X, s.dt_dev, 3, B, k123, st)) return cipher_fail(CIPHER_ERR_HIP, "stage 4");
    if ((rc = cipher_rhs_impl(c, s.Xn, s.K[3], nullptr, 3))) return rc;
    if (launch_axpy_dt(n, s.Xh, X, s.dt_dev, 4, BH, k1234, st)) return cipher_fail(CIPHER_ERR_HIP, "embedded");
    if (launch_wrms(s.Xn, s.Xh, n, s.p.atol, s.p.rtol, c->partial, wrms_parts(n), c->norm, st))
        return cipher_fail(CIPHER_ERR_HIP, "norm");
    return CIPHER_OK;
}

static int capture(cipher_ctx *c, const double *X) {
    TsState &s = c->ts;
    cipher_ts_invalidate(c);
    TSCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    int rc = enqueue_attempt(c, X);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return cipher_fail(CIPHER_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    s.graph = g;
    TSCHK(hipGraphInstantiate(&s.exec, g, nullptr, nullptr, 0));
    s.graph_X = X;
    return CIPHER_OK;
}

extern "C" int cipher_ts_attempt(cipher_ctx *c, const double *X, double dt, double *err_out) {
    if (!c || !X || !err_out || !(dt > 0.0)) return cipher_fail(CIPHER_ERR_ARG, "bad argument");
    TsState &s = c->ts;
    if (!s.ready) return cipher_fail(CIPHER_ERR_STATE, "cipher_ts_setup not called for this mesh");
    bool restore = false;
    if (s.p.use_graph && !c->stream) {
        // the legacy default stream cannot be captured: use a stream of our
        // own that synchronises with it
        if (!s.own) TSCHK(hipStreamCreate(&s.own));
        c->stream = s.own;
        restore = true;
    }
    int rc = CIPHER_OK;
    *s.h_dt = dt;
    hipError_t e = hipMemcpyAsync(s.dt_dev, s.h_dt, sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) rc = cipher_fail(CIPHER_ERR_HIP, "dt upload");
    if (!rc) {
        if (s.p.use_graph) {
            if (!s.exec || s.graph_X != X) rc = capture(c, X);
            if (!rc && hipGraphLaunch(s.exec, c->stream) != hipSuccess) rc = cipher_fail(CIPHER_ERR_HIP, "graph launch");
        } else {
            rc = enqueue_attempt(c, X);
        }
    }
    if (!rc) {
        e = hipMemcpyAsync(c->h_norm, c->norm, sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = cipher_fail(CIPHER_ERR_HIP, std::string("attempt: ") + hipGetErrorString(e));
        else *err_out = *c->h_norm;
    }
    if (restore) c->stream = nullptr;
    return rc;
}

extern "C" int cipher_ts_accept(cipher_ctx *c, double *X) {
    if (!c || !X) return cipher_fail(CIPHER_ERR_ARG, "null argument");
    TsState &s = c->ts;
    if (!s.ready) return cipher_fail(CIPHER_ERR_STATE, "cipher_ts_setup not called for this mesh");
    TSCHK(hipMemcpyAsync(X, s.Xn, sizeof(double) * (size_t)s.n, hipMemcpyDeviceToDevice, c->stream));
    return cipher_post_common(c, X, nullptr, false);
}

extern "C" int cipher_ts_solve(cipher_ctx *c, double *X, double t0, double t_end, double *dt_inout, int64_t max_steps,
                               cipher_ts_stats *stats) {
    if (!c || !X || !dt_inout || !stats || !(t_end >= t0)) return cipher_fail(CIPHER_ERR_ARG, "bad argument");
    TsState &s = c->ts;
    if (!s.ready) return cipher_fail(CIPHER_ERR_STATE, "cipher_ts_setup not called for this mesh");
    cipher_ts_stats out{};
    double t = t0;
    double dt = *dt_inout > 0.0 ? *dt_inout : s.p.dt_min;
    int rc = CIPHER_OK;
    while (t < t_end && (max_steps <= 0 || out.steps < max_steps)) {
        // exact final time: the last step ends at t_end
        const double dt_try = (t + dt > t_end) ? t_end - t : dt;
        if (!(dt_try > 0.0)) break;
        double err = 0.0;
        if ((rc = cipher_ts_attempt(c, X, dt_try, &err))) break;
        out.rhs_evals += 4;
        out.err = err;
        int acc = 0;
        double dt_next = dt_try;
        if ((rc = cipher_ts_adapt(&s.p, &s.ctrl, dt_try, err, &acc, &dt_next))) break;
        if (acc) {
            if ((rc = cipher_ts_accept(c, X))) break;
            t = (dt_try == t_end - t) ? t_end : t + dt_try;
            out.steps++;
        } else {
            out.rejected++;
        }
        dt = dt_next;
    }
    out.t = t;
    out.dt = dt;
    *stats = out;
    *dt_inout = dt;
    return rc;
}

// Load cases, src/cipher.c:1589-1606: the case's end time is the running
// sum loadcasetime + time[k], each case restarts at timestepmin.  The
// reference splits a case into TSSolve calls of amrinterval steps and re-reads
// the step size between them; without remeshing that is one solve.
extern "C" int cipher_ts_loadcases(cipher_ctx *c, double *X, int32_t ncases, const double *times,
                                   const double *rates, double t0, cipher_ts_stats *stats) {
    if (!c || !X || !times || !stats || ncases < 0) return cipher_fail(CIPHER_ERR_ARG, "bad argument");
    if (!c->ts.ready) return cipher_fail(CIPHER_ERR_STATE, "cipher_ts_setup not called for this mesh");
    for (int32_t k = 0; k < ncases; k++)
        if (!(times[k] >= 0.0)) return cipher_fail(CIPHER_ERR_ARG, "load-case time must be >= 0");
    double start = t0, t = t0;
    for (int32_t k = 0; k < ncases; start += times[k++]) {
        int rc = cipher_set_temperature_rate(c, rates != nullptr, rates ? rates[k] : 0.0);
        if (rc) return rc;
        const double end = start + times[k];
        double dt = c->ts.p.dt_min;
        cipher_ts_stats st{};
        st.t = t;
        st.dt = dt;
        if (t < end && (rc = cipher_ts_solve(c, X, t, end, &dt, 0, &st))) return rc;
        t = st.t;
        stats[k] = st;
    }
    return CIPHER_OK;
}
