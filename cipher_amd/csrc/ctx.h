// ctx.h - the context behind the opaque cipher_ctx handle (host side only),
// shared by the C ABI (api.hip) and the integrator (integrator.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "kernels.h"

// adaptive explicit RK state: stage vectors, device step size, HIP graph of
// one attempt (captured for one solution pointer)
struct TsState {
    bool ready = false;
    cipher_ts_params p{};
    cipher_ts_ctrl ctrl{};
    double *K[4] = {nullptr, nullptr, nullptr, nullptr};
    double *Y = nullptr, *Xn = nullptr, *Xh = nullptr;
    double *dt_dev = nullptr;
    double *h_dt = nullptr;           // pinned
    long long n = 0;                  // doubles per vector
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const double *graph_X = nullptr;  // solution pointer the graph was captured for
    hipStream_t own = nullptr;        // capture stream when the context has none
};

struct cipher_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int precision = CIPHER_FP64;
    cipher_layout L{};
    DevCfg dc{};
    std::vector<void *> allocs;   // config tables
    // mesh
    long long ncells = 0;
    int32_t *cell_off = nullptr, *adj = nullptr, *bcell_off = nullptr, *bface_id = nullptr;
    double *cellgeom = nullptr;
    void *cm = nullptr;
    int32_t *glist = nullptr;
    int32_t *list = nullptr;        // FP32-mixed recomputation lists (5 * ncells)
    int32_t *bins = nullptr;        // cells by active-set size (5 * ncells)
    int32_t *counts = nullptr;      // list and recomputation-list sizes (RHS_COUNTERS)
    hipStream_t aux[4] = {nullptr, nullptr, nullptr, nullptr};   // default priority
    hipEvent_t ev_fork = nullptr, ev_join[4] = {nullptr, nullptr, nullptr, nullptr};
    int32_t *err = nullptr;
    double *partial = nullptr;
    double *norm = nullptr;
    double *tsv = nullptr;          // folded T-series (ntseries + 1)
    bool fold = false;
    double *h_norm = nullptr;
    int32_t *h_err = nullptr;
    int npart = 1024;
    TsState ts;                     // adaptive integrator (integrator.hip)
};


// shared between the translation units of the C ABI
int cipher_fail(int code, const std::string &msg);
int cipher_rhs_impl(cipher_ctx *c, const double *X, double *F, uint32_t *active);
int cipher_post_common(cipher_ctx *c, double *X, const int32_t *phase, bool setup);
void cipher_ts_free(cipher_ctx *c);
// a captured attempt bakes in the precision, stream and temperature rate
void cipher_ts_invalidate(cipher_ctx *c);
