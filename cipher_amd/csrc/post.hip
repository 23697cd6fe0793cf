// post.hip - active-set maintenance on gfx950.
//
// k_project    PostStep's projection and pruning (src/cipher.c:875-909):
//              Euclidean projection of the slot fractions onto the Gibbs
//              simplex (SimplexProjection, src/utility.c:587-603: sort
//              descending, threshold, clip) and the pruned list {phi > TOL}.
// k_union      superset = own pruned list U every face neighbour's pruned list
//              (SetUnion, src/utility.c:195-230), new slots initialised from
//              the phase's c0 (src/cipher.c:953-957), surviving slots moved
//              through SetIntersection (src/cipher.c:960-966).  Overflow of
//              PF_SIZE raises the capacity flag (src/cipher.c:968 asserts).
// k_setup      SetUpProblem (src/init.c:1721-1830).
//
// G lanes per cell, lane g owns slot g; sorted-set merges are done with
// group shuffles on ranks, so lists stay ascending and bit-exact.
#include "kernels.h"

namespace cipher {

template <int G, class T>
__device__ __forceinline__ T psh(T v, int src) {
    return __shfl(v, src, G);
}

// Merge two ascending duplicate-free lists held one element per lane (A: na,
// B: nb) into A.  Ranks: a_s goes to s + #{non-duplicate b < a_s}; a
// non-duplicate b_t goes to #{non-duplicate b before t} + #{a < b_t}.
// Returns the union count (may exceed G: the caller flags overflow; lanes
// at or beyond the count hold BIG).
template <int G>
__device__ __forceinline__ int group_union(int &a, int na, int b, int nb) {
    const int slot = threadIdx.x & (G - 1);
    const int BIG = 0x7fffffff;
    const int av = slot < na ? a : BIG;
    const int bv = slot < nb ? b : BIG;
    int lessA = 0, dupB = 0;
#pragma unroll
    for (int t = 0; t < G; t++) {
        const int at = psh<G>(av, t);
        if (t < na && at < bv) lessA++;
        if (t < na && at == bv) dupB = 1;
    }
    int lessB = 0, before = 0, ndup = 0;
#pragma unroll
    for (int t = 0; t < G; t++) {
        const int bt = psh<G>(bv, t);
        const int dt = psh<G>(dupB, t);
        if (t < nb) {
            ndup += dt;
            if (!dt) {
                if (bt < av) lessB++;
                if (t < slot) before++;
            }
        }
    }
    const int rA = slot < na ? slot + lessB : -1;
    const int rB = (slot < nb && !dupB) ? before + lessA : -1;
    const int cnt = na + nb - ndup;
    int out = BIG;
#pragma unroll
    for (int t = 0; t < G; t++) {
        const int ra = psh<G>(rA, t), rb = psh<G>(rB, t);
        const int va = psh<G>(av, t), vb = psh<G>(bv, t);
        if (ra == slot) out = va;
        if (rb == slot) out = vb;
    }
    a = out;
    return cnt;
}

template <int G>
__device__ __forceinline__ void project_cells(double *X, int32_t *glist, long long ncells, const DevCfg &c,
                                              long long blk) {
    const long long cell = (blk * blockDim.x + threadIdx.x) / G;
    const int slot = threadIdx.x & (G - 1);
    if (cell >= ncells) return;
    double *x = X + cell * c.D;
    const int n = clamp_count(x[c.aso], c.P);
    const bool on = slot < n;
    const int myid = on ? clamp_id(x[c.aso + 1 + slot], c.npf) : 0;
    int32_t *gl = glist + cell * (c.P + 1);
    if (n > 1) {
        const double v = on ? x[c.pfo + slot] : -INFINITY;
        // stable descending rank
        int rank = 0;
#pragma unroll
        for (int t = 0; t < G; t++) {
            const double vt = psh<G>(v, t);
            if (t < n && on && (vt > v || (vt == v && t < slot))) rank++;
        }
        int inv = 0;
#pragma unroll
        for (int t = 0; t < G; t++) {
            const int rt = psh<G>(on ? rank : -1, t);
            if (rt == slot) inv = t;
        }
        double tsum = 0.0, tmax = 0.0;
        bool found = false;
        for (int i = 0; i < n - 1; i++) {
            const double si = psh<G>(v, psh<G>(inv, i));
            const double sn = psh<G>(v, psh<G>(inv, i + 1));
            tsum += si;
            tmax = (tsum - 1.0) / ((double)(i + 1));
            if (tmax >= sn) { found = true; break; }
        }
        if (!found) {
            const double sl = psh<G>(v, psh<G>(inv, n - 1));
            tmax = (tsum + sl - 1.0) / ((double)n);
        }
        const double out = v - tmax > 0.0 ? v - tmax : 0.0;
        if (on) x[c.pfo + slot] = out;
        const bool keep = on && out > CIPHER_TOL;
        const unsigned long long bal = __ballot(keep);
        const int lane = threadIdx.x & 63;
        const int gbase = lane & ~(G - 1);
        const unsigned long long gm = (G == 64) ? ~0ull : (((1ull << G) - 1) << gbase);
        const int pos = __popcll(bal & gm & ((1ull << lane) - 1));
        const int cnt = __popcll(bal & gm);
        if (keep) gl[1 + pos] = myid;
        if (slot == 0) gl[0] = cnt;
    } else {
        if (slot == 0) {
            x[c.pfo] = 1.0;
            gl[0] = 1;
            gl[1] = clamp_id(x[c.aso + 1], c.npf);
        }
    }
}

template <int G>
__global__ __launch_bounds__(256) void k_project(double *X, int32_t *glist, long long ncells, DevCfg c,
                                                  const int32_t *pred) {
    if (pred && *pred == 0) return;
    project_cells<G>(X, glist, ncells, c, blockIdx.x);
}

template <int G, int NC>
__device__ __forceinline__ void union_cells(double *X, const int32_t *glist, const int32_t *cell_off,
                                            const int32_t *adj, int32_t *err, long long ncells, const DevCfg &c,
                                            long long blk) {
    constexpr int ND = NC - 1;
    constexpr int SPV = 2 * Cnt<ND>::v;
    const long long cell = (blk * blockDim.x + threadIdx.x) / G;
    const int slot = threadIdx.x & (G - 1);
    if (cell >= ncells) return;
    double *x = X + cell * c.D;
    const int n = clamp_count(x[c.aso], c.P);
    const bool on = slot < n;
    const int oldid = on ? clamp_id(x[c.aso + 1 + slot], c.npf) : -1;
    const double oldphi = on ? x[c.pfo + slot] : 0.0;
    double oldex[SPV];
#pragma unroll
    for (int k = 0; k < SPV; k++) oldex[k] = (on && k < c.SP) ? x[c.exo + slot * c.SP + k] : 0.0;
    const double T = x[c.tmo];
    const int32_t *gl = glist + cell * (c.P + 1);
    int cnt = min(max(gl[0], 0), c.P);
    int ss = slot < cnt ? gl[1 + slot] : 0;
    bool over = false;
    for (int k = cell_off[cell]; k < cell_off[cell + 1]; k++) {
        const long long nb = adj[k] >> 1;
        const int32_t *gn = glist + nb * (c.P + 1);
        const int nn = min(max(gn[0], 0), c.P);
        const int bv = slot < nn ? gn[1 + slot] : 0;
        cnt = group_union<G>(ss, cnt, bv, nn);
        if (cnt > c.P) { over = true; break; }
    }
    if (over) {
        if (slot == 0) atomicOr(err, 1);
        return;
    }
    // reorder surviving slots (SetIntersection(old, superset)), init new ones
    int pos = -1;
    for (int t = 0; t < n; t++) {
        const int v = psh<G>(oldid, t);
        if (slot < cnt && v == ss) pos = t;
    }
    const double nphi = psh<G>(oldphi, pos < 0 ? 0 : pos);
    double nex[SPV];
#pragma unroll
    for (int k = 0; k < SPV; k++) nex[k] = psh<G>(oldex[k], pos < 0 ? 0 : pos);
    if (slot < cnt) {
        double phiv = 0.0;
        if (pos >= 0) {
            phiv = nphi;
        } else if constexpr (NC > 1) {
            const cipher_material &m = c.mat[c.phase_material[min(max(ss, 0), c.npf - 1)]];
            double y[2 * NC];
#pragma unroll
            for (int k = 0; k < 2 * NC; k++) y[k] = m.c0[k];
            double sp[SPV];
#pragma unroll
            for (int k = 0; k < SPV; k++) sp[k] = 0.0;
            d_sitepot_explicit<NC, double>(c, m, sp, y, T);
#pragma unroll
            for (int k = 0; k < SPV; k++) nex[k] = sp[k];
        }
        x[c.pfo + slot] = phiv;
#pragma unroll
        for (int k = 0; k < SPV; k++)
            if (k < c.SP) x[c.exo + slot * c.SP + k] = nex[k];
        x[c.aso + 1 + slot] = (double)ss;
    }
    if (slot == 0) x[c.aso] = (double)cnt;
}

template <int G, int NC>
__global__ __launch_bounds__(256) void k_union(double *X, const int32_t *glist, const int32_t *cell_off,
                                               const int32_t *adj, int32_t *err, long long ncells, DevCfg c,
                                               const int32_t *pred) {
    if (pred && *pred == 0) return;
    union_cells<G, NC>(X, glist, cell_off, adj, err, ncells, c, blockIdx.x);
}

// batched PostStep: one segment of blocks per member (members of one class)
template <int G>
__global__ __launch_bounds__(256) void k_project_batch(const BatchMember *__restrict__ mem,
                                                       const int32_t *__restrict__ live,
                                                       const int32_t *__restrict__ seg,
                                                       const int32_t *__restrict__ segmem, int nseg, double *X) {
    const int k = seg_find(seg, nseg, blockIdx.x);
    const int mi = segmem[k];
    if (live && !live[mi]) return;
    const BatchMember &m = mem[mi];
    const DevCfg &c = m.c;
    project_cells<G>(X + m.xoff, m.glist, m.a.ncells, c, (long long)blockIdx.x - seg[k]);
}

template <int G, int NC>
__global__ __launch_bounds__(256) void k_union_batch(const BatchMember *__restrict__ mem,
                                                     const int32_t *__restrict__ live,
                                                     const int32_t *__restrict__ seg,
                                                     const int32_t *__restrict__ segmem, int nseg, double *X) {
    const int k = seg_find(seg, nseg, blockIdx.x);
    const int mi = segmem[k];
    if (live && !live[mi]) return;
    const BatchMember &m = mem[mi];
    const DevCfg &c = m.c;
    union_cells<G, NC>(X + m.xoff, m.glist, m.a.cell_off, m.a.adj, m.err, m.a.ncells, c,
                       (long long)blockIdx.x - seg[k]);
}

template <int G, int NC>
__global__ __launch_bounds__(256) void k_setup(double *X, const int32_t *phase, const int32_t *cell_off,
                                               const int32_t *adj, int32_t *err, long long ncells, DevCfg c) {
    constexpr int ND = NC - 1;
    constexpr int SPV = 2 * Cnt<ND>::v;
    const long long cell = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int slot = threadIdx.x & (G - 1);
    if (cell >= ncells) return;
    double *x = X + cell * c.D;
    const int ph = min(max(phase[cell], 0), c.npf - 1);
    int cnt = 1;
    int ss = ph;
    bool over = false;
    for (int k = cell_off[cell]; k < cell_off[cell + 1]; k++) {
        const int pn = min(max(phase[adj[k] >> 1], 0), c.npf - 1);
        cnt = group_union<G>(ss, cnt, pn, 1);
        if (cnt > c.P) { over = true; break; }
    }
    if (over) {
        if (slot == 0) atomicOr(err, 1);
        return;
    }
    const cipher_material &mown = c.mat[c.phase_material[ph]];
    const double T = mown.temperature0;
    if (slot < cnt) {
        const cipher_material &m = c.mat[c.phase_material[min(max(ss, 0), c.npf - 1)]];
        x[c.aso + 1 + slot] = (double)ss;
        x[c.pfo + slot] = (ss == ph) ? 1.0 : 0.0;
        if constexpr (NC > 1) {
            double y[2 * NC];
#pragma unroll
            for (int k = 0; k < 2 * NC; k++) y[k] = m.c0[k];
            double sp[SPV];
#pragma unroll
            for (int k = 0; k < SPV; k++) sp[k] = 0.0;
            d_sitepot_explicit<NC, double>(c, m, sp, y, T);
#pragma unroll
            for (int k = 0; k < SPV; k++)
                if (k < c.SP) x[c.exo + slot * c.SP + k] = sp[k];
            if (ss == ph) {
                double spi[SPV], spe[SPV];
#pragma unroll
                for (int k = 0; k < SPV; k++) { spi[k] = 0.0; spe[k] = 0.0; }
                d_sitepot_implicit<NC, double>(c, m, spi, y, T);
                d_sitepot_explicit<NC, double>(c, m, spe, y, T);
#pragma unroll
                for (int q = 0; q < ND; q++) {
                    double d = 0.0;
#pragma unroll
                    for (int s = 0; s < 2; s++)
                        if (s < m.nsites) d += spi[s * ND + q] + spe[s * ND + q];
                    x[c.dpo + q] = d;
                }
            }
        }
    }
    if (slot == 0) {
        x[c.aso] = (double)cnt;
        x[c.tmo] = T;
    }
}

template <int G>
static cipher_status post_G(const PostLaunch &L, hipStream_t st, bool setup) {
    constexpr int CPB = 256 / G;
    const long long nb = (L.ncells + CPB - 1) / CPB;
    if (nb == 0) return CIPHER_OK;
    const DevCfg &c = *L.cfg;
#define CIPHER_POST_NC(NC)                                                                                   \
    case NC:                                                                                                 \
        if (setup)                                                                                           \
            hipLaunchKernelGGL((k_setup<G, NC>), dim3((unsigned)nb), dim3(256), 0, st, L.X, L.phase,         \
                               L.cell_off, L.adj, L.err, L.ncells, c);                                       \
        else                                                                                                 \
            hipLaunchKernelGGL((k_union<G, NC>), dim3((unsigned)nb), dim3(256), 0, st, L.X, L.glist,         \
                               L.cell_off, L.adj, L.err, L.ncells, c, L.pred);                               \
        break;
    if (!setup)
        hipLaunchKernelGGL((k_project<G>), dim3((unsigned)nb), dim3(256), 0, st, L.X, L.glist, L.ncells, c, L.pred);
    switch (c.ncp) {
        CIPHER_POST_NC(1)
        CIPHER_POST_NC(2)
        CIPHER_POST_NC(3)
        CIPHER_POST_NC(4)
        CIPHER_POST_NC(5)
    default: return CIPHER_ERR_UNSUPPORTED;
    }
#undef CIPHER_POST_NC
    return hipGetLastError() == hipSuccess ? CIPHER_OK : CIPHER_ERR_HIP;
}

static cipher_status post_dispatch(const PostLaunch &L, hipStream_t st, bool setup) {
    const int P = L.cfg->P;
    if (P <= 4) return post_G<4>(L, st, setup);
    if (P <= 8) return post_G<8>(L, st, setup);
    if (P <= 16) return post_G<16>(L, st, setup);
    return CIPHER_ERR_UNSUPPORTED;
}

template <int G>
static cipher_status post_batch_G(const BatchLaunch &L, double *X, hipStream_t st) {
    const BatchSegs &s = L.post;
    if (s.nblocks == 0) return CIPHER_OK;
    hipLaunchKernelGGL((k_project_batch<G>), dim3((unsigned)s.nblocks), dim3(256), 0, st, L.mem, L.live, s.seg,
                       s.segmem, s.nseg, X);
#define CIPHER_POSTB_NC(NC)                                                                                 \
    case NC:                                                                                                \
        hipLaunchKernelGGL((k_union_batch<G, NC>), dim3((unsigned)s.nblocks), dim3(256), 0, st, L.mem, L.live, \
                           s.seg, s.segmem, s.nseg, X);                                                     \
        break;
    switch (L.ncp) {
        CIPHER_POSTB_NC(1)
        CIPHER_POSTB_NC(2)
        CIPHER_POSTB_NC(3)
        CIPHER_POSTB_NC(4)
        CIPHER_POSTB_NC(5)
    default: return CIPHER_ERR_UNSUPPORTED;
    }
#undef CIPHER_POSTB_NC
    return hipGetLastError() == hipSuccess ? CIPHER_OK : CIPHER_ERR_HIP;
}

cipher_status launch_post_batch(const BatchLaunch &L, double *X, hipStream_t st) {
    if (L.G == 4) return post_batch_G<4>(L, X, st);
    if (L.G == 8) return post_batch_G<8>(L, X, st);
    if (L.G == 16) return post_batch_G<16>(L, X, st);
    return CIPHER_ERR_UNSUPPORTED;
}

cipher_status launch_poststep(const PostLaunch &L, hipStream_t st) { return post_dispatch(L, st, false); }
cipher_status launch_setup(const PostLaunch &L, hipStream_t st) { return post_dispatch(L, st, true); }

}  // namespace cipher
