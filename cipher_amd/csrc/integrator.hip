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
//   Heun 2(1) ("2a"):       K1 = f(X),  Xh = X + dt K1
//                           K2 = f(Xh), Xn = X + dt (1/2 K1 + 1/2 K2)
//   error  = WRMS(Xn, Xh) with tol_i = atol + rtol max(|Xn_i|, |Xh_i|)
//   accept = error <= 1 (or dt at dt_min), then X = Xn and PostStep.
//
// The first stage is always evaluated: PostStep changes the solution after
// the step, so the last stage of the previous step is not f(X).
//
// An attempt is a fixed sequence of launches whose only varying input is dt,
// kept in device memory, so it can be captured once in a HIP graph and
// replayed.  cipher_ts_solve runs whole steps on the device: a one-thread
// kernel starts each attempt (step size with the exact final time, or no
// work once the loop has ended), the attempt's launches are predicated on
// it, a second one-thread kernel takes the controller's decision from the
// error norm (ctrl.h: the same IEEE arithmetic as cipher_ts_adapt), and the
// copy and PostStep of an accepted step are predicated on that decision; the
// host synchronises once per group of attempts.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include <algorithm>

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
// Heun's method with the explicit Euler step as its embedded solution
static const double H2A[1] = {1.0};
static const double H2B[2] = {1.0 / 2.0, 1.0 / 2.0};

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
    if (s.dev) (void)hipFree(s.dev);
    if (s.live) (void)hipFree(s.live);
    if (s.acc) (void)hipFree(s.acc);
    if (s.h_dev) (void)hipHostFree(s.h_dev);
    s.dev = nullptr;
    s.h_dev = nullptr;
    s.live = s.acc = nullptr;
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
    if (s.sexec) (void)hipGraphExecDestroy(s.sexec);
    if (s.sgraph) (void)hipGraphDestroy(s.sgraph);
    s.sexec = nullptr;
    s.sgraph = nullptr;
    s.sgraph_X = nullptr;
}

extern "C" int cipher_ts_adapt(const cipher_ts_params *p, cipher_ts_ctrl *st, double dt, double err, int *accept,
                               double *dt_next) {
    if (!p || !st || !accept || !dt_next || !(dt > 0.0)) return cipher_fail(CIPHER_ERR_ARG, "bad argument");
    if (ctl_adapt(*p, *st, dt, err, accept, dt_next) != CIPHER_OK)
        return cipher_fail(CIPHER_ERR_NONFINITE, "error norm is not finite at the minimum step");
    return CIPHER_OK;
}

extern "C" int cipher_ts_setup(cipher_ctx *c, const cipher_ts_params *p) {
    if (!c || !p) return cipher_fail(CIPHER_ERR_ARG, "null argument");
    if (!c->cell_off) return cipher_fail(CIPHER_ERR_STATE, "no mesh uploaded");
    if (!(p->atol >= 0.0) || !(p->rtol >= 0.0) || !(p->atol > 0.0 || p->rtol > 0.0) || !(p->dt_min > 0.0) ||
        !(p->dt_max >= p->dt_min) || !(p->safety > 0.0) || !(p->reject_safety > 0.0) || !(p->clip_lo > 0.0) ||
        !(p->clip_hi >= p->clip_lo) || (p->method != CIPHER_RK_3BS && p->method != CIPHER_RK_2A))
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
    TSCHK(hipMalloc(&s.dev, sizeof(TsDev)));
    TSCHK(hipHostMalloc(&s.h_dev, sizeof(TsDev), hipHostMallocDefault));
    TSCHK(hipMalloc(&s.live, sizeof(int32_t)));
    TSCHK(hipMalloc(&s.acc, sizeof(int32_t)));
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
    if (s.p.method == CIPHER_RK_2A) {
        const double *k12[2] = {s.K[0], s.K[1]};
        if (launch_axpy_dt(n, s.Xh, X, s.dt_dev, 1, H2A, k1, st)) return cipher_fail(CIPHER_ERR_HIP, "Euler step");
        if ((rc = cipher_rhs_impl(c, s.Xh, s.K[1], nullptr, 2))) return rc;
        if (launch_axpy_dt(n, s.Xn, X, s.dt_dev, 2, H2B, k12, st)) return cipher_fail(CIPHER_ERR_HIP, "Heun step");
        if (launch_wrms(s.Xn, s.Xh, n, s.p.atol, s.p.rtol, c->partial, wrms_parts(n), c->norm, st))
            return cipher_fail(CIPHER_ERR_HIP, "norm");
        return CIPHER_OK;
    }
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

// ---- device-driven steps ---------------------------------------------------
// start of an attempt (one thread): the step size and the attempt's run flag
__global__ void k_ts_begin(TsDev *s, double *dt_dev, const int32_t *post_err, int32_t *live) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ts_begin(*s, *post_err, dt_dev, live);
}

// end of an attempt (one thread): the controller's decision from the norm
__global__ void k_ts_decide(TsDev *s, const double *norm, cipher_ts_params p, int32_t *acc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ts_decide(*s, p, *norm, acc);
}

// X = Xn where the attempt was accepted
__global__ __launch_bounds__(256) void k_copy_if(long long n, double *X, const double *Xn, const int32_t *acc) {
    if (*acc == 0) return;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        X[i] = Xn[i];
}

// one step on the device: begin, the attempt (predicated on its run flag),
// the decision, and the copy and PostStep of an accepted step (predicated on
// the decision); a capacity overflow of PostStep ends the loop at the next
// begin
static int enqueue_step(cipher_ctx *c, double *X) {
    TsState &s = c->ts;
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(k_ts_begin, dim3(1), dim3(64), 0, st, s.dev, s.dt_dev, (const int32_t *)c->err, s.live);
    c->pred = s.live;
    int rc = enqueue_attempt(c, X);
    c->pred = nullptr;
    if (rc) return rc;
    hipLaunchKernelGGL(k_ts_decide, dim3(1), dim3(64), 0, st, s.dev, (const double *)c->norm, s.p, s.acc);
    long long blocks = (s.n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_copy_if, dim3((unsigned)blocks), dim3(256), 0, st, s.n, X, (const double *)s.Xn,
                       (const int32_t *)s.acc);
    PostLaunch L;
    L.cfg = &c->dc;
    L.X = X;
    L.cell_off = c->cell_off;
    L.adj = c->adj;
    L.glist = c->glist;
    L.err = c->err;
    L.phase = nullptr;
    L.ncells = c->ncells;
    L.pred = s.acc;
    if (launch_poststep(L, st) != CIPHER_OK) return cipher_fail(CIPHER_ERR_HIP, "poststep launch");
    return hipGetLastError() == hipSuccess ? CIPHER_OK : cipher_fail(CIPHER_ERR_HIP, "step launch");
}

static int capture_step(cipher_ctx *c, double *X) {
    TsState &s = c->ts;
    if (s.sexec) (void)hipGraphExecDestroy(s.sexec);
    if (s.sgraph) (void)hipGraphDestroy(s.sgraph);
    s.sexec = nullptr;
    s.sgraph = nullptr;
    TSCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    int rc = enqueue_step(c, X);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return cipher_fail(CIPHER_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    s.sgraph = g;
    TSCHK(hipGraphInstantiate(&s.sexec, g, nullptr, nullptr, 0));
    s.sgraph_X = X;
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
    bool restore = false;
    if (s.p.use_graph && !c->stream) {
        // the legacy default stream cannot be captured: use a stream of our
        // own that synchronises with it
        if (!s.own) TSCHK(hipStreamCreate(&s.own));
        c->stream = s.own;
        restore = true;
    }
    TsDev &h = *s.h_dev;
    h = TsDev{};
    h.t = t0;
    h.t_end = t_end;
    h.dt = *dt_inout > 0.0 ? *dt_inout : s.p.dt_min;
    h.ctrl = s.ctrl;
    h.max_steps = max_steps;
    h.empty = s.n > 0 ? 0 : 1;
    int rc = CIPHER_OK;
    int64_t syncs = 0;
    hipError_t e = hipMemcpyAsync(s.dev, s.h_dev, sizeof(TsDev), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->err, 0, sizeof(int32_t), c->stream);
    if (e != hipSuccess) rc = cipher_fail(CIPHER_ERR_HIP, "solve set-up");
    int group = 1;
    while (!rc) {
        for (int k = 0; k < group && !rc; k++) {
            if (s.p.use_graph) {
                if (!s.sexec || s.sgraph_X != X) rc = capture_step(c, X);
                if (!rc && hipGraphLaunch(s.sexec, c->stream) != hipSuccess)
                    rc = cipher_fail(CIPHER_ERR_HIP, "graph launch");
            } else {
                rc = enqueue_step(c, X);
            }
        }
        if (rc) break;
        e = hipMemcpyAsync(s.h_dev, s.dev, sizeof(TsDev), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        syncs++;
        if (e != hipSuccess) {
            rc = cipher_fail(CIPHER_ERR_HIP, std::string("step: ") + hipGetErrorString(e));
            break;
        }
        if (h.done) break;
        // the next group: the attempts that reach t_end at the current step
        // size (one more for a rejection), at most twice the last group and 16
        const double rem = (h.t_end - h.t) / h.dt;
        int want = rem < 15.0 ? (int)rem + 2 : 16;
        if (h.max_steps > 0 && h.max_steps - h.steps + 1 < want) want = (int)(h.max_steps - h.steps + 1);
        group = std::min(std::min(2 * group, 16), std::max(want, 1));
    }
    if (!rc && h.status == CIPHER_ERR_CAPACITY)
        rc = cipher_fail(CIPHER_ERR_CAPACITY, "active-set union exceeds the capacity PF_SIZE");
    if (!rc && h.status == CIPHER_ERR_NONFINITE)
        rc = cipher_fail(CIPHER_ERR_NONFINITE, "error norm is not finite at the minimum step");
    s.ctrl = h.ctrl;
    cipher_ts_stats out{};
    out.steps = h.steps;
    out.rejected = h.rejected;
    out.rhs_evals = h.rhs_evals;
    out.t = h.t;
    out.dt = h.dt;
    out.err = h.err;
    out.host_syncs = syncs;
    *stats = out;
    *dt_inout = h.dt;
    if (restore) c->stream = nullptr;
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
