// ctx.h - the context behind the opaque cipher_ctx handle (host side only),
// shared by the C ABI (api.hip) and the integrator (integrator.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "ctrl.h"
#include "kernels.h"

// One coordinate step of the launch-shape search: set one field of a
// ListShape (0 lanes, 1 block order, 2 grid cap).
struct ShapeStep {
    int field, value;
};

// Steps tried for list q of a mesh with lane capacity G: lanes 2^q..4*2^q
// (single-phase cells without chemistry keep their one-lane path), both
// block orders, and for the single-phase list of a mesh with chemistry both
// with and without a launch of its own (SHAPE_OWN), caps of 16..65536 blocks
// (a cap below what the list needs makes its blocks stride: slower, never
// wrong).
static inline std::vector<ShapeStep> shape_steps(int q, int G, int ncp) {
    std::vector<ShapeStep> out;
    const int base = 1 << q;
    if (!(q == 0 && ncp == 1))
        for (int l = base; l <= std::min(G, base * 4); l *= 2) out.push_back({0, l});
    for (int m : {0, 1}) out.push_back({1, m});
    if (q == 0 && ncp > 1)
        for (int m : {cipher::SHAPE_OWN, cipher::SHAPE_OWN | 1}) out.push_back({1, m});
    for (int cap : {16, 128, 1024, 4096, 16384, 65536}) out.push_back({2, cap});
    return out;
}

static inline cipher::ListShape apply_step(cipher::ListShape s, const ShapeStep &st) {
    if (st.field == 0) s.lanes = st.value;
    else if (st.field == 1) s.map = st.value;
    else s.cap = st.value;
    return s;
}

// a shape for list q of a mesh with lane capacity G and ncp components
static inline bool shape_valid(const cipher_list_shape &s, int q, int G, int ncp) {
    const int base = 1 << q;
    const bool lanes_ok = s.lanes == 0 || ((s.lanes & (s.lanes - 1)) == 0 && s.lanes >= base && s.lanes <= G &&
                                           s.lanes <= 16);
    const bool map_ok = s.map == 0 || s.map == 1 ||
                        (q == 0 && ncp > 1 && (s.map == cipher::SHAPE_OWN || s.map == (cipher::SHAPE_OWN | 1)));
    return lanes_ok && s.cap >= 1 && map_ok;
}

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
    // device-driven solve: loop state, the attempt's run flag and the
    // accept flag, and the captured step (attempt, decision, accept)
    cipher::TsDev *dev = nullptr;
    cipher::TsDev *h_dev = nullptr;   // pinned
    int32_t *live = nullptr, *acc = nullptr;
    hipGraph_t sgraph = nullptr;
    hipGraphExec_t sexec = nullptr;
    const double *sgraph_X = nullptr;
};

struct cipher_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int precision = CIPHER_FP64;
    cipher_layout L{};
    DevCfg dc{};
    // a device copy of dc, which the list and prep kernels read through a
    // pointer (a by-value kernel argument of ~50 fields is hoisted into
    // scalar registers and spilled); refreshed whenever dc changes
    DevCfg *d_cfg = nullptr;
    std::vector<void *> allocs;   // config tables
    // mesh
    long long ncells = 0;
    int32_t *cell_off = nullptr, *adj = nullptr, *bcell_off = nullptr, *bface_id = nullptr;
    double *cellgeom = nullptr;
    double *fgeo = nullptr;         // flux factor per face-list entry
    void *cm = nullptr;
    int32_t *glist = nullptr;
    int32_t *list = nullptr;        // FP32-mixed recomputation lists (5 * ncells)
    int32_t *bins = nullptr;        // cells by active-set size (5 * ncells)
    // list and recomputation-list sizes: RHS_RING + 1 sets of RHS_COUNTERS.
    // Eager cipher_rhs calls take the ring's sets in turn, all of them
    // cleared by one fill every RHS_RING calls; stream capture and the
    // integrator's stages use the last set, cleared by each binning call.
    int32_t *counts = nullptr;
    int ring_pos = cipher::RHS_RING;
    int32_t *err = nullptr;
    double *partial = nullptr;
    double *norm = nullptr;
    double *tsv = nullptr;          // folded T-series (ntseries + 1)
    bool fold = false;
    double *h_norm = nullptr;
    int32_t *h_err = nullptr;
    int npart = 1024;
    TsState ts;                     // adaptive integrator (integrator.hip)
    cipher::ListShape shape[5];     // launch shapes of the per-size lists
    const int32_t *pred = nullptr;  // set while enqueuing a device-driven step: RHS launches run while *pred != 0
};


// shared between the translation units of the C ABI
int cipher_fail(int code, const std::string &msg);
// copy c->dc to its device copy (after the context's stream has drained, so no
// launch in flight reads a half-written configuration)
int cipher_sync_cfg(cipher_ctx *c);
// bins_mode: see RhsLaunch (0 for a right-hand side on its own)
int cipher_rhs_impl(cipher_ctx *c, const double *X, double *F, uint32_t *active, int bins_mode = 0);
int cipher_post_common(cipher_ctx *c, double *X, const int32_t *phase, bool setup);
void cipher_ts_free(cipher_ctx *c);
// a captured attempt bakes in the precision, stream and temperature rate
void cipher_ts_invalidate(cipher_ctx *c);
