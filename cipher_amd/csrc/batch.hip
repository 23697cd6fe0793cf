// batch.hip - batched ensembles: many independent contexts per launch.
//
// The reference integrates one microstructure per process.  A batch takes
// contexts that each hold their own configuration and mesh and runs them
// together: their states are concatenated (member i's records start at
// xoff[i] doubles), members with the same component count and lane capacity
// form a class, and every kernel of the right-hand side, PostStep, the error
// norm and the stage combinations is one launch per class (per cell list)
// whose grid is cut into one segment of blocks per member.  Each member's
// results are those of its context run alone: the kernel bodies are the
// single-mesh ones, and the error norm keeps the single-mesh partition.
//
// The integrator gives every member its own step size and controller.  One
// attempt advances all live members together (a member is live until it
// reaches its end time); the attempt is a fixed launch sequence whose inputs
// (step sizes, live flags) sit in device memory, so it is captured once in
// a HIP graph and replayed.
//
// Launch shapes (lanes per cell, grid cap and block order of each class's
// cell lists) come from cipher_batch_autotune, which times candidates on a
// given state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <string>
#include <vector>

#include "ctx.h"

using namespace cipher;

#define BCHK(expr)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return cipher_fail(CIPHER_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_));   \
    } while (0)

static const double A2[1] = {1.0 / 2.0};
static const double A3[1] = {3.0 / 4.0};
static const double B[3] = {2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0};
static const double BH[4] = {7.0 / 24.0, 1.0 / 4.0, 1.0 / 3.0, 1.0 / 8.0};
// Heun's method with the explicit Euler step as its embedded solution
static const double H2A[1] = {1.0};
static const double H2B[2] = {1.0 / 2.0, 1.0 / 2.0};

namespace {

struct BatchClass {
    int ncp = 1, G = 8;
    std::vector<int> members;
    ListShape shape[5];
    BatchLaunch L{};
    double cost = 0.0;        // estimated time of the class's launches alone (ms)
    int lane = 0;             // stream slot of the class (0: the caller's stream)
};

// streams a batched right-hand side spreads its classes over: the caller's
// and BATCH_STREAMS - 1 forked ones (a process has 4 hardware queues by
// default; a fifth stream would share one and serialise behind it)
constexpr int BATCH_STREAMS = 4;

struct BatchTs {
    bool ready = false;
    cipher_ts_params p{};
    double *K[4] = {nullptr, nullptr, nullptr, nullptr};
    double *Y = nullptr, *Xn = nullptr, *Xh = nullptr;
    double *dt_dev = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const double *graph_X = nullptr;
    hipStream_t own = nullptr;
    // per-member step-controller memory, carried from call to call as a
    // context's is (cleared by cipher_batch_ts_setup)
    std::vector<cipher_ts_ctrl> ctrl;
    // device-driven solve: per-member loop state and the captured step
    TsDev *dev = nullptr;
    TsDev *h_dev = nullptr;           // pinned, nmembers
    hipGraph_t sgraph = nullptr;
    hipGraphExec_t sexec = nullptr;
    const double *sgraph_X = nullptr;
};

}  // namespace

struct cipher_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int precision = CIPHER_FP64;
    std::vector<cipher_ctx *> m;
    std::vector<long long> xoff, celloff;   // nmembers + 1
    BatchMember *d_mem = nullptr;
    int32_t *d_counts = nullptr;            // RHS_COUNTERS per member
    int32_t *d_err = nullptr;               // PostStep flag per member
    double *d_partial = nullptr;            // error-norm partial sums
    double *d_norms = nullptr;              // per-member error norms
    int32_t *d_live = nullptr, *d_acc = nullptr;
    double *h_norms = nullptr, *h_dt = nullptr;  // pinned, nmembers
    int32_t *h_err = nullptr, *h_live = nullptr, *h_acc = nullptr;
    std::vector<BatchClass> cls;
    std::vector<int> sched;                 // class enqueue order (longest first)
    BatchLaunch all{};                      // fold and per-dof launches over every member
    std::vector<void *> segs;               // device segment tables
    hipStream_t aux[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t ev_fork = nullptr, ev_join[4] = {nullptr, nullptr, nullptr, nullptr};
    BatchTs ts;
};

static int lanes_for(int P) { return P <= 4 ? 4 : (P <= 8 ? 8 : (P <= 16 ? 16 : 0)); }

// Longest-processing-time schedule of the classes over the streams: classes
// in decreasing cost, each to the stream with the least cost so far, so the
// streams finish together instead of one stream's queue of small classes
// running alone at the end.
static void schedule_classes(cipher_batch *b) {
    const int nk = (int)b->cls.size();
    b->sched.resize(nk);
    for (int k = 0; k < nk; k++) b->sched[k] = k;
    std::stable_sort(b->sched.begin(), b->sched.end(),
                     [&](int x, int y) { return b->cls[x].cost > b->cls[y].cost; });
    double load[BATCH_STREAMS] = {0.0};
    for (int k : b->sched) {
        int s = 0;
        for (int j = 1; j < BATCH_STREAMS; j++)
            if (load[j] < load[s]) s = j;
        b->cls[k].lane = s;
        load[s] += b->cls[k].cost;
    }
}

// cost model used until the classes are timed: cells times a per-cell cost
// that grows with the component count and the lane capacity
static double model_cost(const cipher_batch *b, const BatchClass &k) {
    long long cells = 0;
    for (int i : k.members) cells += b->m[i]->ncells;
    const double w = (k.ncp == 1 ? 0.5 : 0.8 * k.ncp) * (k.G >= 16 ? 1.5 : 1.0);
    return 1e-6 * w * (double)cells;
}
static bool needs_prep(const cipher_ctx *c) { return c->L.dp_size > 0 || c->dc.has_tcond; }

static void free_segs(cipher_batch *b) {
    for (void *p : b->segs)
        if (p) (void)hipFree(p);
    b->segs.clear();
}

// segment table over `members`: member i gets nb(i) blocks (0: left out)
template <class F>
static int make_segs(cipher_batch *b, BatchSegs &s, const std::vector<int> &members, F nb) {
    std::vector<int32_t> seg{0}, segmem;
    long long tot = 0;
    for (int i : members) {
        const long long n = nb(i);
        if (n <= 0) continue;
        tot += n;
        if (tot > 0x7fffffffLL) return cipher_fail(CIPHER_ERR_UNSUPPORTED, "batch launch exceeds 2^31 blocks");
        seg.push_back((int32_t)tot);
        segmem.push_back(i);
    }
    s.nseg = (int)segmem.size();
    s.nblocks = (int)tot;
    s.seg = nullptr;
    s.segmem = nullptr;
    if (s.nseg == 0) return CIPHER_OK;
    int32_t *d = nullptr;
    BCHK(hipMalloc(&d, sizeof(int32_t) * (seg.size() + segmem.size())));
    b->segs.push_back(d);
    BCHK(hipMemcpy(d, seg.data(), sizeof(int32_t) * seg.size(), hipMemcpyHostToDevice));
    BCHK(hipMemcpy(d + seg.size(), segmem.data(), sizeof(int32_t) * segmem.size(), hipMemcpyHostToDevice));
    s.seg = d;
    s.segmem = d + seg.size();
    return CIPHER_OK;
}

// segment table of a merged list launch: for each list q in `order`, every
// member i gets nb(i, q) blocks
template <class F>
static int make_list_segs(cipher_batch *b, BatchSegs &s, const std::vector<int> &members, const std::vector<int> &order,
                          F nb) {
    std::vector<int32_t> seg{0}, segmem, segq;
    long long tot = 0;
    for (int q : order)
        for (int i : members) {
            const long long n = nb(i, q);
            if (n <= 0) continue;
            tot += n;
            if (tot > 0x7fffffffLL) return cipher_fail(CIPHER_ERR_UNSUPPORTED, "batch launch exceeds 2^31 blocks");
            seg.push_back((int32_t)tot);
            segmem.push_back(i);
            segq.push_back(q);
        }
    s.nseg = (int)segmem.size();
    s.nblocks = (int)tot;
    s.seg = s.segmem = s.segq = nullptr;
    if (s.nseg == 0) return CIPHER_OK;
    int32_t *d = nullptr;
    BCHK(hipMalloc(&d, sizeof(int32_t) * (seg.size() + 2 * segmem.size())));
    b->segs.push_back(d);
    BCHK(hipMemcpy(d, seg.data(), sizeof(int32_t) * seg.size(), hipMemcpyHostToDevice));
    BCHK(hipMemcpy(d + seg.size(), segmem.data(), sizeof(int32_t) * segmem.size(), hipMemcpyHostToDevice));
    BCHK(hipMemcpy(d + seg.size() + segmem.size(), segq.data(), sizeof(int32_t) * segq.size(), hipMemcpyHostToDevice));
    s.seg = d;
    s.segmem = d + seg.size();
    s.segq = d + seg.size() + segmem.size();
    return CIPHER_OK;
}

static long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

// blocks of member i in a grid-striding list launch: enough for every cell
// up to the launch's cap, shared out by cell count
static long long list_blocks(long long N, long long Nclass, int cpb, int cap) {
    if (N <= 0) return 0;
    long long full = cdiv(N, cpb);
    long long share = (long long)std::llround((double)cap * (double)N / (double)std::max(Nclass, 1LL));
    return std::max(1LL, std::min(full, std::max(1LL, share)));
}

static int build_tables(cipher_batch *b) {
    free_segs(b);
    const int n = (int)b->m.size();
    std::vector<int> everyone(n);
    for (int i = 0; i < n; i++) everyone[i] = i;
    int rc;
    BatchLaunch &A = b->all;
    A.mem = b->d_mem;
    if ((rc = make_segs(b, A.fold, everyone, [&](int i) { return b->m[i]->fold && b->m[i]->ncells > 0 ? 1LL : 0LL; })))
        return rc;
    if ((rc = make_segs(b, A.dofs, everyone, [&](int i) {
             const long long nd = b->m[i]->ncells * (long long)b->m[i]->L.ndof;
             return nd > 0 ? (long long)wrms_parts(nd) : 0LL;
         })))
        return rc;
    for (BatchClass &k : b->cls) {
        BatchLaunch &L = k.L;
        L.mem = b->d_mem;
        L.ncp = k.ncp;
        L.G = k.G;
        L.nbins = k.G >= 16 ? 5 : (k.G >= 8 ? 4 : 3);
        long long Nclass = 0;
        for (int i : k.members) Nclass += b->m[i]->ncells;
        const int cpbG = 256 / k.G;
        // (prep records are written list by list; this table only marks the
        // members that need them)
        if ((rc = make_segs(b, L.prep, k.members,
                            [&](int i) { return needs_prep(b->m[i]) ? cdiv(b->m[i]->ncells, cpbG) : 0LL; })))
            return rc;
        if ((rc = make_segs(b, L.bin, k.members, [&](int i) { return cdiv(b->m[i]->ncells, BIN_TILE); }))) return rc;
        if ((rc = make_segs(b, L.post, k.members, [&](int i) { return cdiv(b->m[i]->ncells, cpbG); }))) return rc;
        // all lists of all members in one launch, largest active sets first
        std::vector<int> order;
        L.lists.lds = 0;
        for (int q = 0; q < 5; q++) {
            const int lanes = k.shape[q].lanes > 0 ? k.shape[q].lanes : (1 << q);
            L.shape.lanes[q] = lanes;
            L.shape.map[q] = k.shape[q].map & 1;
        }
        for (int q = L.nbins - 1; q >= 0; q--) {
            order.push_back(q);
            for (int i : k.members)
                L.lists.lds = std::max(L.lists.lds, rhs_lds_bytes(b->m[i]->dc, L.shape.lanes[q], 256 / L.shape.lanes[q]));
        }
        if ((rc = make_list_segs(b, L.lists, k.members, order, [&](int i, int q) {
                 return list_blocks(b->m[i]->ncells, Nclass, 256 / L.shape.lanes[q], k.shape[q].cap);
             })))
            return rc;
        // the single-phase list in a launch of its own: segments of lists
        // 1.. and of list 0 (the merged table above still drives the prep
        // records)
        L.listsN = BatchSegs{};
        L.lists0 = BatchSegs{};
        if (own_launch(k.shape[0], k.ncp)) {
            const std::vector<int> multi(order.begin(), order.end() - 1), single{0};
            auto nb = [&](int i, int q) {
                return list_blocks(b->m[i]->ncells, Nclass, 256 / L.shape.lanes[q], k.shape[q].cap);
            };
            if ((rc = make_list_segs(b, L.listsN, k.members, multi, nb))) return rc;
            if ((rc = make_list_segs(b, L.lists0, k.members, single, nb))) return rc;
            L.listsN.lds = L.lists.lds;
            L.lists0.lds = L.lists.lds;
        }
        L.relists.lds = L.lists.lds;
        if ((rc = make_list_segs(b, L.relists, k.members, order, [&](int i, int q) {
                 return list_blocks(b->m[i]->ncells, Nclass, 256 / L.shape.lanes[q],
                                    std::min(k.shape[q].cap, RHS_REDO_CAP));
             })))
            return rc;
    }
    return CIPHER_OK;
}

static void ts_invalidate(cipher_batch *b) {
    BatchTs &s = b->ts;
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

static void ts_free(cipher_batch *b) {
    BatchTs &s = b->ts;
    ts_invalidate(b);
    double *v[] = {s.K[0], s.K[1], s.K[2], s.K[3], s.Y, s.Xn, s.Xh, s.dt_dev};
    for (double *p : v)
        if (p) (void)hipFree(p);
    for (int i = 0; i < 4; i++) s.K[i] = nullptr;
    s.Y = s.Xn = s.Xh = s.dt_dev = nullptr;
    if (s.dev) (void)hipFree(s.dev);
    if (s.h_dev) (void)hipHostFree(s.h_dev);
    s.dev = nullptr;
    s.h_dev = nullptr;
    if (s.own) (void)hipStreamDestroy(s.own);
    s.own = nullptr;
    s.ready = false;
}

static void batch_free(cipher_batch *b) {
    ts_free(b);
    free_segs(b);
    void *d[] = {b->d_mem, b->d_counts, b->d_err, b->d_partial, b->d_norms, b->d_live, b->d_acc};
    for (void *p : d)
        if (p) (void)hipFree(p);
    void *h[] = {b->h_norms, b->h_dt, b->h_err, b->h_live, b->h_acc};
    for (void *p : h)
        if (p) (void)hipHostFree(p);
    for (int i = 0; i < 4; i++) {
        if (b->aux[i]) (void)hipStreamDestroy(b->aux[i]);
        if (b->ev_join[i]) (void)hipEventDestroy(b->ev_join[i]);
    }
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
}

__global__ void k_fill_batch(int32_t *p, long long n, int32_t v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

static int fill(cipher_batch *b, int32_t *p, long long n, int32_t v, hipStream_t st) {
    if (n <= 0) return CIPHER_OK;
    const long long nb = std::min<long long>(cdiv(n, 256), 1024);
    hipLaunchKernelGGL(k_fill_batch, dim3((unsigned)nb), dim3(256), 0, st, p, n, v);
    BCHK(hipGetLastError());
    return CIPHER_OK;
}

extern "C" int cipher_batch_create(cipher_ctx *const *members, int32_t nmembers, cipher_batch **out) {
    if (!members || !out || nmembers <= 0) return cipher_fail(CIPHER_ERR_ARG, "bad argument");
    *out = nullptr;
    for (int i = 0; i < nmembers; i++) {
        const cipher_ctx *c = members[i];
        if (!c) return cipher_fail(CIPHER_ERR_ARG, "null member");
        if (!c->cell_off) return cipher_fail(CIPHER_ERR_STATE, "member without an uploaded mesh");
        if (c->device != members[0]->device) return cipher_fail(CIPHER_ERR_ARG, "members on different devices");
        if (!lanes_for(c->L.pf_size)) return cipher_fail(CIPHER_ERR_UNSUPPORTED, "capacity above 16");
        for (int j = 0; j < i; j++)
            if (members[j] == c) return cipher_fail(CIPHER_ERR_ARG, "a context appears twice");
    }
    cipher_batch *b = new cipher_batch();
    b->device = members[0]->device;
    b->m.assign(members, members + nmembers);
    int rc = CIPHER_OK;
    auto bail = [&](int code, const std::string &msg) {
        batch_free(b);
        delete b;
        return cipher_fail(code, msg);
    };
    if (hipSetDevice(b->device) != hipSuccess) return bail(CIPHER_ERR_HIP, "hipSetDevice");
    const int n = nmembers;
    b->xoff.assign(n + 1, 0);
    b->celloff.assign(n + 1, 0);
    long long npart = 0;
    for (int i = 0; i < n; i++) {
        const cipher_ctx *c = b->m[i];
        b->xoff[i + 1] = b->xoff[i] + c->ncells * (long long)c->L.ndof;
        b->celloff[i + 1] = b->celloff[i] + c->ncells;
        npart += wrms_parts(std::max(1LL, c->ncells * (long long)c->L.ndof));
    }
    bool ok = hipMalloc(&b->d_mem, sizeof(BatchMember) * n) == hipSuccess &&
              hipMalloc(&b->d_counts, sizeof(int32_t) * RHS_COUNTERS * n) == hipSuccess &&
              hipMalloc(&b->d_err, sizeof(int32_t) * n) == hipSuccess &&
              hipMalloc(&b->d_partial, sizeof(double) * npart) == hipSuccess &&
              hipMalloc(&b->d_norms, sizeof(double) * n) == hipSuccess &&
              hipMalloc(&b->d_live, sizeof(int32_t) * n) == hipSuccess &&
              hipMalloc(&b->d_acc, sizeof(int32_t) * n) == hipSuccess &&
              hipHostMalloc(&b->h_norms, sizeof(double) * n, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc(&b->h_dt, sizeof(double) * n, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc(&b->h_err, sizeof(int32_t) * n, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc(&b->h_live, sizeof(int32_t) * n, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc(&b->h_acc, sizeof(int32_t) * n, hipHostMallocDefault) == hipSuccess;
    for (int i = 0; ok && i < 4; i++)
        ok = hipStreamCreateWithFlags(&b->aux[i], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&b->ev_join[i], hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) == hipSuccess;
    if (!ok) return bail(CIPHER_ERR_HIP, "batch allocation");
    std::vector<BatchMember> hm(n);
    long long poff = 0;
    for (int i = 0; i < n; i++) {
        cipher_ctx *c = b->m[i];
        BatchMember &m = hm[i];
        m = BatchMember{};
        RhsArgs &a = m.a;
        a.X = nullptr;
        a.F = nullptr;
        a.active = nullptr;
        a.cellgeom = c->cellgeom;
        a.cell_off = c->cell_off;
        a.adj = c->adj;
        a.fgeo = c->fgeo;
        a.bcell_off = c->bcell_off;
        a.bface_id = c->bface_id;
        a.cm = (const double *)c->cm;
        a.cells = nullptr;
        a.cells_count = nullptr;
        a.redo = c->list;
        a.redo_count = nullptr;
        a.ncells = c->ncells;
        m.c = c->dc;
        m.xoff = b->xoff[i];
        m.celloff = b->celloff[i];
        m.bins = c->bins;
        m.counts = b->d_counts + RHS_COUNTERS * i;
        m.cm = (double *)c->cm;
        m.glist = c->glist;
        m.err = b->d_err + i;
        m.npart = wrms_parts(std::max(1LL, c->ncells * (long long)c->L.ndof));
        m.partial = b->d_partial + poff;
        poff += m.npart;
        m.fold = c->fold ? 1 : 0;
        // class of the member
        const int G = lanes_for(c->L.pf_size);
        const int ncp = c->dc.ncp;
        auto it = std::find_if(b->cls.begin(), b->cls.end(),
                               [&](const BatchClass &k) { return k.ncp == ncp && k.G == G; });
        if (it == b->cls.end()) {
            BatchClass k;
            k.ncp = ncp;
            k.G = G;
            b->cls.push_back(k);
            it = b->cls.end() - 1;
        }
        it->members.push_back(i);
    }
    if (hipMemcpy(b->d_mem, hm.data(), sizeof(BatchMember) * n, hipMemcpyHostToDevice) != hipSuccess)
        return bail(CIPHER_ERR_HIP, "member table upload");
    if ((rc = build_tables(b))) {
        std::string msg = cipher_last_error();
        return bail(rc, msg);
    }
    for (BatchClass &k : b->cls) k.cost = model_cost(b, k);
    schedule_classes(b);
    *out = b;
    return CIPHER_OK;
}

extern "C" int cipher_batch_destroy(cipher_batch *b) {
    if (!b) return CIPHER_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    else (void)hipDeviceSynchronize();
    batch_free(b);
    delete b;
    return CIPHER_OK;
}

extern "C" int cipher_batch_offsets(const cipher_batch *b, int64_t *offsets, int64_t *cell_offsets) {
    if (!b || !offsets) return cipher_fail(CIPHER_ERR_ARG, "null argument");
    for (size_t i = 0; i < b->xoff.size(); i++) {
        offsets[i] = b->xoff[i];
        if (cell_offsets) cell_offsets[i] = b->celloff[i];
    }
    return CIPHER_OK;
}

extern "C" int cipher_batch_set_precision(cipher_batch *b, int precision) {
    if (!b) return cipher_fail(CIPHER_ERR_ARG, "null batch");
    if (precision != CIPHER_FP64 && precision != CIPHER_FP32_MIXED)
        return cipher_fail(CIPHER_ERR_ARG, "unknown precision");
    b->precision = precision;
    ts_invalidate(b);
    return CIPHER_OK;
}

extern "C" int cipher_batch_set_stream(cipher_batch *b, void *stream) {
    if (!b) return cipher_fail(CIPHER_ERR_ARG, "null batch");
    b->stream = (hipStream_t)stream;
    ts_invalidate(b);
    return CIPHER_OK;
}

extern "C" int cipher_batch_set_temperature_rate(cipher_batch *b, int32_t member, int has_rate, double rate) {
    if (!b) return cipher_fail(CIPHER_ERR_ARG, "null batch");
    if (member < 0 || member >= (int32_t)b->m.size()) return cipher_fail(CIPHER_ERR_ARG, "member out of range");
    cipher_ctx *c = b->m[member];
    int rc = cipher_set_temperature_rate(c, has_rate, rate);
    if (rc) return rc;
    // the two fields, from the context's updated configuration into the
    // member's copy on the device, ordered after the batch's earlier work
    const size_t lo = offsetof(DevCfg, trate), hi = offsetof(DevCfg, has_trate) + sizeof(int);
    char *dst = (char *)(b->d_mem + member) + offsetof(BatchMember, c) + lo;
    const hipStream_t st = b->stream;
    if (hipMemcpyAsync(dst, (const char *)&c->dc + lo, hi - lo, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return cipher_fail(CIPHER_ERR_HIP, "member table update");
    ts_invalidate(b);
    return CIPHER_OK;
}

// zero the recomputation counters (8..15) of every member, keeping the lists
__global__ void k_fill_redo(int32_t *counts, int nmem) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nmem * 8) counts[(i / 8) * RHS_COUNTERS + 8 + (i % 8)] = 0;
}

// the launches of one batched right-hand side, in order, on b->stream; with
// reuse set, the cells keep the lists of the previous call (stages of one
// attempt share the active sets)
static int batch_rhs_impl(cipher_batch *b, const double *X, double *F, uint32_t *active, const int32_t *live,
                          bool reuse = false) {
    hipStream_t st = b->stream;
    int rc;
    const int nmem = (int)b->m.size();
    if (!reuse) {
        if ((rc = fill(b, b->d_counts, (long long)RHS_COUNTERS * (long long)nmem, 0, st))) return rc;
    } else if (b->precision == CIPHER_FP32_MIXED) {
        hipLaunchKernelGGL(k_fill_redo, dim3((unsigned)((nmem * 8 + 255) / 256)), dim3(256), 0, st, b->d_counts, nmem);
        BCHK(hipGetLastError());
    }
    b->all.X = X;
    b->all.live = live;
    if (launch_fold_batch(b->all, st)) return cipher_fail(CIPHER_ERR_HIP, "batch fold launch");
    // classes are independent: they run on the caller's stream and four
    // forked streams, each class's launches in order on its stream
    // (in the schedule's order: longest first, each on its stream slot)
    const bool fork = b->cls.size() > 1;
    const int nfork = std::min<int>(BATCH_STREAMS - 1, (int)b->cls.size() - 1);
    if (fork) {
        BCHK(hipEventRecord(b->ev_fork, st));
        for (int i = 0; i < nfork; i++) BCHK(hipStreamWaitEvent(b->aux[i], b->ev_fork, 0));
    }
    for (int ki : b->sched) {
        BatchClass &k = b->cls[ki];
        k.L.X = X;
        k.L.F = F;
        k.L.active = active;
        k.L.live = live;
        k.L.do_bin = reuse ? 0 : 1;
        hipStream_t s = (!fork || k.lane == 0) ? st : b->aux[std::min(k.lane, nfork) - 1];
        if (launch_rhs_batch(k.L, b->precision, s)) return cipher_fail(CIPHER_ERR_HIP, "batch list launch");
    }
    if (fork) {
        for (int i = 0; i < nfork; i++) {
            BCHK(hipEventRecord(b->ev_join[i], b->aux[i]));
            BCHK(hipStreamWaitEvent(st, b->ev_join[i], 0));
        }
    }
    return CIPHER_OK;
}

extern "C" int cipher_batch_rhs(cipher_batch *b, double t, const double *X, double *F) {
    (void)t;
    if (!b || !X || !F) return cipher_fail(CIPHER_ERR_ARG, "null argument");
    return batch_rhs_impl(b, X, F, nullptr, nullptr);
}

extern "C" int cipher_batch_rhs_active(cipher_batch *b, double t, const double *X, double *F, uint32_t *active) {
    (void)t;
    if (!b || !X || !F || !active) return cipher_fail(CIPHER_ERR_ARG, "null argument");
    return batch_rhs_impl(b, X, F, active, nullptr);
}

// PostStep of the members whose live flag is set (all when live is null)
static int batch_post_impl(cipher_batch *b, double *X, const int32_t *live, int32_t *failed) {
    hipStream_t st = b->stream;
    const int n = (int)b->m.size();
    int rc;
    if ((rc = fill(b, b->d_err, n, 0, st))) return rc;
    for (BatchClass &k : b->cls) {
        k.L.live = live;
        if (launch_post_batch(k.L, X, st)) return cipher_fail(CIPHER_ERR_HIP, "batch poststep launch");
    }
    BCHK(hipMemcpyAsync(b->h_err, b->d_err, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    BCHK(hipStreamSynchronize(st));
    for (int i = 0; i < n; i++)
        if (b->h_err[i]) {
            if (failed) *failed = i;
            return cipher_fail(CIPHER_ERR_CAPACITY,
                               "active-set union exceeds the capacity PF_SIZE in member " + std::to_string(i));
        }
    if (failed) *failed = -1;
    return CIPHER_OK;
}

extern "C" int cipher_batch_poststep(cipher_batch *b, double *X, int32_t *failed_member) {
    if (!b || !X) return cipher_fail(CIPHER_ERR_ARG, "null argument");
    return batch_post_impl(b, X, nullptr, failed_member);
}

extern "C" int cipher_batch_wrms_norm(cipher_batch *b, const double *U, const double *Y, double atol, double rtol,
                                      double *norms) {
    if (!b || !U || !Y || !norms) return cipher_fail(CIPHER_ERR_ARG, "null argument");
    hipStream_t st = b->stream;
    const int n = (int)b->m.size();
    b->all.live = nullptr;
    if (launch_wrms_batch(b->all, U, Y, atol, rtol, b->d_norms, st)) return cipher_fail(CIPHER_ERR_HIP, "batch norm");
    BCHK(hipMemcpyAsync(b->h_norms, b->d_norms, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    BCHK(hipStreamSynchronize(st));
    for (int i = 0; i < n; i++) norms[i] = b->m[i]->ncells > 0 ? b->h_norms[i] : 0.0;
    return CIPHER_OK;
}

// ---- autotuner -------------------------------------------------------------
// Time each class's launches alone on the caller's stream (median of reps,
// counters cleared and T-series folded before each, untimed) and schedule the
// classes over the streams by those times.
static int measure_class_costs(cipher_batch *b, const double *X, double *F, int reps, hipEvent_t e0, hipEvent_t e1) {
    hipStream_t st = b->stream;
    const int nmem = (int)b->m.size();
    int rc;
    for (BatchClass &k : b->cls) {
        std::vector<float> v;
        for (int r = 0; r < reps; r++) {
            if ((rc = fill(b, b->d_counts, (long long)RHS_COUNTERS * (long long)nmem, 0, st))) return rc;
            b->all.X = X;
            b->all.live = nullptr;
            if (launch_fold_batch(b->all, st)) return cipher_fail(CIPHER_ERR_HIP, "batch fold launch");
            k.L.X = X;
            k.L.F = F;
            k.L.active = nullptr;
            k.L.live = nullptr;
            k.L.do_bin = 1;
            BCHK(hipEventRecord(e0, st));
            if (launch_rhs_batch(k.L, b->precision, st)) return cipher_fail(CIPHER_ERR_HIP, "batch list launch");
            BCHK(hipEventRecord(e1, st));
            BCHK(hipEventSynchronize(e1));
            float t = 0.f;
            BCHK(hipEventElapsedTime(&t, e0, e1));
            v.push_back(t);
        }
        std::sort(v.begin(), v.end());
        k.cost = v[v.size() / 2];
    }
    schedule_classes(b);
    return CIPHER_OK;
}

// Coordinate descent over the launch shape of each class's lists: for every
// (class, list) in turn, lanes per cell, block order and grid cap are varied
// one at a time, each candidate timed as the median of `reps`
// whole batched right-hand sides on X, and the fastest kept.  The choice
// changes only launch shapes, never results.
extern "C" int cipher_batch_autotune(cipher_batch *b, const double *X, double *F, int32_t reps, double *ms_out) {
    if (!b || !X || !F) return cipher_fail(CIPHER_ERR_ARG, "null argument");
    if (reps < 1) reps = 1;
    hipStream_t st = b->stream;
    hipEvent_t e0, e1;
    BCHK(hipEventCreate(&e0));
    BCHK(hipEventCreate(&e1));
    int rc = CIPHER_OK;
    auto time_once = [&](float &ms) -> int {
        std::vector<float> v;
        for (int r = 0; r < reps; r++) {
            BCHK(hipEventRecord(e0, st));
            int q = batch_rhs_impl(b, X, F, nullptr, nullptr);
            if (q) return q;
            BCHK(hipEventRecord(e1, st));
            BCHK(hipEventSynchronize(e1));
            float t = 0.f;
            BCHK(hipEventElapsedTime(&t, e0, e1));
            v.push_back(t);
        }
        std::sort(v.begin(), v.end());
        ms = v[v.size() / 2];
        return CIPHER_OK;
    };
    float best = 0.f;
    if ((rc = batch_rhs_impl(b, X, F, nullptr, nullptr)) == CIPHER_OK) rc = measure_class_costs(b, X, F, reps, e0, e1);
    if (rc == CIPHER_OK) rc = time_once(best);
    for (size_t ki = 0; ki < b->cls.size() && rc == CIPHER_OK; ki++) {
        BatchClass &k = b->cls[ki];
        for (int q = 0; q < k.L.nbins && rc == CIPHER_OK; q++) {
            for (const ShapeStep &step : shape_steps(q, k.G, k.ncp)) {
                const ListShape keep = k.shape[q];
                k.shape[q] = apply_step(keep, step);
                const int kl = keep.lanes > 0 ? keep.lanes : (1 << q);
                const int cl = k.shape[q].lanes > 0 ? k.shape[q].lanes : (1 << q);
                if (kl == cl && k.shape[q].map == keep.map && k.shape[q].cap == keep.cap) {
                    k.shape[q] = keep;
                    continue;
                }
                if ((rc = build_tables(b))) break;
                float ms = 0.f;
                if ((rc = time_once(ms))) break;
                if (ms < best) best = ms;
                else k.shape[q] = keep;
            }
            if (rc == CIPHER_OK) rc = build_tables(b);
        }
    }
    // the stream schedule of the tuned shapes (kept only if it is faster)
    if (rc == CIPHER_OK) {
        std::vector<double> keep_cost;
        for (const BatchClass &k : b->cls) keep_cost.push_back(k.cost);
        const std::vector<int> keep_sched = b->sched;
        std::vector<int> keep_lane;
        for (const BatchClass &k : b->cls) keep_lane.push_back(k.lane);
        float ms = 0.f;
        if ((rc = measure_class_costs(b, X, F, reps, e0, e1)) == CIPHER_OK && (rc = time_once(ms)) == CIPHER_OK) {
            if (ms < best) {
                best = ms;
            } else {
                b->sched = keep_sched;
                for (size_t k = 0; k < b->cls.size(); k++) {
                    b->cls[k].cost = keep_cost[k];
                    b->cls[k].lane = keep_lane[k];
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ts_invalidate(b);
    if (rc == CIPHER_OK && ms_out) *ms_out = best;
    return rc;
}

extern "C" int cipher_batch_shape(const cipher_batch *b, int32_t k, int32_t q, int32_t *ncp, int32_t *capacity_lanes,
                                  cipher_list_shape *shape) {
    if (!b || k < 0 || k >= (int)b->cls.size() || q < 0 || q >= 5)
        return cipher_fail(CIPHER_ERR_ARG, "no such class or list");
    const BatchClass &c = b->cls[k];
    if (ncp) *ncp = c.ncp;
    if (capacity_lanes) *capacity_lanes = c.G;
    if (shape) {
        shape->lanes = c.shape[q].lanes > 0 ? c.shape[q].lanes : (1 << q);
        shape->cap = c.shape[q].cap;
        shape->map = c.shape[q].map;
        shape->pad_ = 0;
    }
    return CIPHER_OK;
}

// ---- per-member adaptive integration ----------------------------------------
extern "C" int cipher_batch_ts_setup(cipher_batch *b, const cipher_ts_params *p) {
    if (!b || !p) return cipher_fail(CIPHER_ERR_ARG, "null argument");
    if (!(p->atol >= 0.0) || !(p->rtol >= 0.0) || !(p->atol > 0.0 || p->rtol > 0.0) || !(p->dt_min > 0.0) ||
        !(p->dt_max >= p->dt_min) || !(p->safety > 0.0) || !(p->reject_safety > 0.0) || !(p->clip_lo > 0.0) ||
        !(p->clip_hi >= p->clip_lo) || (p->method != CIPHER_RK_3BS && p->method != CIPHER_RK_2A))
        return cipher_fail(CIPHER_ERR_ARG, "integrator parameters out of range");
    ts_free(b);
    BatchTs &s = b->ts;
    s.p = *p;
    const size_t bytes = sizeof(double) * (size_t)std::max(1LL, b->xoff.back());
    for (int i = 0; i < 4; i++) BCHK(hipMalloc(&s.K[i], bytes));
    BCHK(hipMalloc(&s.Y, bytes));
    BCHK(hipMalloc(&s.Xn, bytes));
    BCHK(hipMalloc(&s.Xh, bytes));
    BCHK(hipMalloc(&s.dt_dev, sizeof(double) * b->m.size()));
    BCHK(hipMalloc(&s.dev, sizeof(TsDev) * b->m.size()));
    BCHK(hipHostMalloc(&s.h_dev, sizeof(TsDev) * b->m.size(), hipHostMallocDefault));
    s.ctrl.assign(b->m.size(), cipher_ts_ctrl{});
    s.ready = true;
    return CIPHER_OK;
}

static int enqueue_attempt(cipher_batch *b, const double *X) {
    BatchTs &s = b->ts;
    hipStream_t st = b->stream;
    int rc;
    const double *k1[1] = {s.K[0]};
    const double *k2[1] = {s.K[1]};
    const double *k123[3] = {s.K[0], s.K[1], s.K[2]};
    const double *k1234[4] = {s.K[0], s.K[1], s.K[2], s.K[3]};
    const int32_t *live = b->d_live;
    BatchLaunch &A = b->all;
    A.live = live;
    // stages 2-4 keep the active sets of X: bin once per attempt
    if ((rc = batch_rhs_impl(b, X, s.K[0], nullptr, live))) return rc;
    A.live = live;
    if (s.p.method == CIPHER_RK_2A) {
        const double *k12[2] = {s.K[0], s.K[1]};
        if (launch_axpy_dt_batch(A, s.Xh, X, s.dt_dev, 1, H2A, k1, st)) return cipher_fail(CIPHER_ERR_HIP, "Euler step");
        if ((rc = batch_rhs_impl(b, s.Xh, s.K[1], nullptr, live, true))) return rc;
        if (launch_axpy_dt_batch(A, s.Xn, X, s.dt_dev, 2, H2B, k12, st)) return cipher_fail(CIPHER_ERR_HIP, "Heun step");
        A.live = live;
        if (launch_wrms_batch(A, s.Xn, s.Xh, s.p.atol, s.p.rtol, b->d_norms, st))
            return cipher_fail(CIPHER_ERR_HIP, "norm");
        return CIPHER_OK;
    }
    if (launch_axpy_dt_batch(A, s.Y, X, s.dt_dev, 1, A2, k1, st)) return cipher_fail(CIPHER_ERR_HIP, "stage 2");
    if ((rc = batch_rhs_impl(b, s.Y, s.K[1], nullptr, live, true))) return rc;
    if (launch_axpy_dt_batch(A, s.Y, X, s.dt_dev, 1, A3, k2, st)) return cipher_fail(CIPHER_ERR_HIP, "stage 3");
    if ((rc = batch_rhs_impl(b, s.Y, s.K[2], nullptr, live, true))) return rc;
    if (launch_axpy_dt_batch(A, s.Xn, X, s.dt_dev, 3, B, k123, st)) return cipher_fail(CIPHER_ERR_HIP, "stage 4");
    if ((rc = batch_rhs_impl(b, s.Xn, s.K[3], nullptr, live, true))) return rc;
    if (launch_axpy_dt_batch(A, s.Xh, X, s.dt_dev, 4, BH, k1234, st)) return cipher_fail(CIPHER_ERR_HIP, "embedded");
    A.live = live;
    if (launch_wrms_batch(A, s.Xn, s.Xh, s.p.atol, s.p.rtol, b->d_norms, st)) return cipher_fail(CIPHER_ERR_HIP, "norm");
    return CIPHER_OK;
}

static int capture(cipher_batch *b, const double *X) {
    BatchTs &s = b->ts;
    ts_invalidate(b);
    BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    int rc = enqueue_attempt(b, X);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return cipher_fail(CIPHER_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    s.graph = g;
    BCHK(hipGraphInstantiate(&s.exec, g, nullptr, nullptr, 0));
    s.graph_X = X;
    return CIPHER_OK;
}

// ---- device-driven steps: one thread per member ----------------------------
__global__ void k_bts_begin(TsDev *s, int n, double *dt_dev, const int32_t *err, int32_t *live) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) ts_begin(s[i], err[i], dt_dev + i, live + i);
}

__global__ void k_bts_decide(TsDev *s, int n, const double *norms, cipher_ts_params p, int32_t *acc) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) ts_decide(s[i], p, norms[i], acc + i);
}

// one step of every member on the device: per-member begin, the attempt
// (members that do not run skip every launch), per-member decision, copy
// and PostStep of the accepted members
static int enqueue_step(cipher_batch *b, double *X) {
    BatchTs &s = b->ts;
    hipStream_t st = b->stream;
    const int n = (int)b->m.size();
    const unsigned nb = (unsigned)((n + 63) / 64);
    hipLaunchKernelGGL(k_bts_begin, dim3(nb), dim3(64), 0, st, s.dev, n, s.dt_dev, (const int32_t *)b->d_err,
                       b->d_live);
    int rc = enqueue_attempt(b, X);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bts_decide, dim3(nb), dim3(64), 0, st, s.dev, n, (const double *)b->d_norms, s.p, b->d_acc);
    if (launch_copy_batch(b->all, X, s.Xn, b->d_acc, st)) return cipher_fail(CIPHER_ERR_HIP, "accept copy");
    for (BatchClass &k : b->cls) {
        k.L.live = b->d_acc;
        if (launch_post_batch(k.L, X, st)) return cipher_fail(CIPHER_ERR_HIP, "batch poststep launch");
    }
    BCHK(hipGetLastError());
    return CIPHER_OK;
}

static int capture_step(cipher_batch *b, double *X) {
    BatchTs &s = b->ts;
    if (s.sexec) (void)hipGraphExecDestroy(s.sexec);
    if (s.sgraph) (void)hipGraphDestroy(s.sgraph);
    s.sexec = nullptr;
    s.sgraph = nullptr;
    BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    int rc = enqueue_step(b, X);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return cipher_fail(CIPHER_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    s.sgraph = g;
    BCHK(hipGraphInstantiate(&s.sexec, g, nullptr, nullptr, 0));
    s.sgraph_X = X;
    return CIPHER_OK;
}

extern "C" int cipher_batch_ts_solve(cipher_batch *b, double *X, const double *t0, const double *t_end,
                                     double *dt_inout, int64_t max_steps, cipher_ts_stats *stats) {
    if (!b || !X || !t0 || !t_end || !dt_inout || !stats) return cipher_fail(CIPHER_ERR_ARG, "null argument");
    BatchTs &s = b->ts;
    if (!s.ready) return cipher_fail(CIPHER_ERR_STATE, "cipher_batch_ts_setup not called");
    const int n = (int)b->m.size();
    for (int i = 0; i < n; i++)
        if (!(t_end[i] >= t0[i])) return cipher_fail(CIPHER_ERR_ARG, "t_end before t0");
    bool restore = false;
    if (s.p.use_graph && !b->stream) {
        if (!s.own) BCHK(hipStreamCreate(&s.own));
        b->stream = s.own;
        restore = true;
    }
    hipStream_t st = b->stream;
    for (int i = 0; i < n; i++) {
        TsDev &h = s.h_dev[i];
        h = TsDev{};
        h.t = t0[i];
        h.t_end = t_end[i];
        h.dt = dt_inout[i] > 0.0 ? dt_inout[i] : s.p.dt_min;
        h.ctrl = s.ctrl[i];
        h.max_steps = max_steps;
        h.empty = b->m[i]->ncells > 0 ? 0 : 1;
    }
    int rc = CIPHER_OK;
    int64_t syncs = 0;
    hipError_t e = hipMemcpyAsync(s.dev, s.h_dev, sizeof(TsDev) * n, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) rc = cipher_fail(CIPHER_ERR_HIP, "solve set-up");
    if (!rc) rc = fill(b, b->d_err, n, 0, st);
    int group = 1;
    while (!rc) {
        for (int k = 0; k < group && !rc; k++) {
            if (s.p.use_graph) {
                if (!s.sexec || s.sgraph_X != X) rc = capture_step(b, X);
                if (!rc && hipGraphLaunch(s.sexec, st) != hipSuccess) rc = cipher_fail(CIPHER_ERR_HIP, "graph launch");
            } else {
                rc = enqueue_step(b, X);
            }
        }
        if (rc) break;
        e = hipMemcpyAsync(s.h_dev, s.dev, sizeof(TsDev) * n, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        syncs++;
        if (e != hipSuccess) {
            rc = cipher_fail(CIPHER_ERR_HIP, std::string("step: ") + hipGetErrorString(e));
            break;
        }
        // the next group: the attempts the slowest running member needs to
        // reach its end time at its step size (one more for a rejection),
        // at most twice the last group and 16
        bool all_done = true;
        int want = 1;
        for (int i = 0; i < n; i++) {
            const TsDev &h = s.h_dev[i];
            if (h.done) continue;
            all_done = false;
            const double rem = (h.t_end - h.t) / h.dt;
            int w = rem < 15.0 ? (int)rem + 2 : 16;
            if (h.max_steps > 0 && h.max_steps - h.steps + 1 < w) w = (int)(h.max_steps - h.steps + 1);
            want = std::max(want, w);
        }
        if (all_done) break;
        group = std::min(std::min(2 * group, 16), std::max(want, 1));
    }
    for (int i = 0; i < n && !rc; i++) {
        const TsDev &h = s.h_dev[i];
        if (h.status == CIPHER_ERR_CAPACITY)
            rc = cipher_fail(CIPHER_ERR_CAPACITY,
                             "active-set union exceeds the capacity PF_SIZE in member " + std::to_string(i));
        else if (h.status == CIPHER_ERR_NONFINITE)
            rc = cipher_fail(CIPHER_ERR_NONFINITE,
                             "error norm is not finite at the minimum step (member " + std::to_string(i) + ")");
    }
    for (int i = 0; i < n; i++) {
        const TsDev &h = s.h_dev[i];
        s.ctrl[i] = h.ctrl;
        cipher_ts_stats o{};
        o.steps = h.steps;
        o.rejected = h.rejected;
        o.rhs_evals = h.rhs_evals;
        o.t = h.t;
        o.dt = h.dt;
        o.err = h.err;
        o.host_syncs = syncs;
        stats[i] = o;
        dt_inout[i] = h.dt;
    }
    if (restore) b->stream = nullptr;
    return rc;
}
