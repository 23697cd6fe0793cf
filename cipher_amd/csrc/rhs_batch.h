// rhs_batch.h - the right-hand side over a batch of independent meshes.
//
// The kernels are the single-mesh bodies of rhs_impl.h behind a member
// lookup: the grid of every launch is cut into one segment per member of a
// class (members with the same component count and lane capacity), each
// block finds its member by a binary search over the segment offsets, loads
// that member's mesh pointers and configuration, and processes its share of
// the member's cells.  Per-size lists are per member, so the ensemble keeps
// the single-mesh binning; one launch per class and list covers all members.
#pragma once
#include "rhs_impl.h"

namespace cipher {

template <int U>
__global__ __launch_bounds__(256) void k_bin_batch(const BatchMember *__restrict__ mem, const int32_t *__restrict__ live,
                                                   const int32_t *__restrict__ seg, const int32_t *__restrict__ segmem,
                                                   int nseg, int nbins, const double *__restrict__ X) {
    const int k = seg_find(seg, nseg, blockIdx.x);
    const int mi = segmem[k];
    if (live && !live[mi]) return;
    const BatchMember &m = mem[mi];
    const DevCfg &c = m.c;
    bin_tile(X + m.xoff, m.a.ncells, c, nbins, m.bins, m.a.ncells, m.counts, (long long)blockIdx.x - seg[k]);
}

// All lists of all members of a class: segment k runs list segq[k] of
// member segmem[k] (FP32-mixed appends the cells it leaves to that list's
// recomputation list); with redo set, the recomputation lists instead.
template <int GMAX, int NC, class R>
__global__ __launch_bounds__(256, NC == 1 ? CIPHER_RHS_MINB1 : CIPHER_RHS_MINBN) void k_rhs_batch(const BatchMember *__restrict__ mem,
                                                                     const int32_t *__restrict__ live,
                                                                     const int32_t *__restrict__ seg,
                                                                     const int32_t *__restrict__ segmem,
                                                                     const int32_t *__restrict__ segq, int nseg,
                                                                     BatchLists bl, int redo, const double *X,
                                                                     double *F, uint32_t *active) {
    const int k = seg_find(seg, nseg, blockIdx.x);
    const int mi = segmem[k];
    if (live && !live[mi]) return;
    const int q = segq[k];
    const BatchMember &m = mem[mi];
    RhsArgs a = m.a;
    a.X = X + m.xoff;
    a.F = F + m.xoff;
    a.active = active ? active + m.celloff : nullptr;
    a.map = bl.map[q];
    a.redo = m.a.redo + (long long)q * a.ncells;
    a.redo_count = m.counts + 8 + q;
    a.cells = redo ? a.redo : m.bins + (long long)q * a.ncells;
    a.cells_count = redo ? a.redo_count : m.counts + q;
    const DevCfg &c = m.c;
    const long long blk = (long long)blockIdx.x - seg[k], nblk = seg[k + 1] - seg[k];
    switch (bl.lanes[q]) {
    case 1: rhs_cells<1, GMAX, NC, R>(a, c, blk, nblk); break;
    case 2: rhs_cells<2, GMAX, NC, R>(a, c, blk, nblk); break;
    case 4: rhs_cells<4, GMAX, NC, R>(a, c, blk, nblk); break;
    case 8: if constexpr (GMAX >= 8) rhs_cells<8, GMAX, NC, R>(a, c, blk, nblk); break;
    default: if constexpr (GMAX >= 16) rhs_cells<16, GMAX, NC, R>(a, c, blk, nblk); break;
    }
}

// The single-phase lists (list 0) of all members of a chemistry class in a
// launch of their own, compiled for more resident waves (k_rhs_one).
template <int GMAX, int NC, class R>
__global__ __launch_bounds__(256, CIPHER_RHS_MINB_ONE(NC)) void k_rhs_batch_one(const BatchMember *__restrict__ mem,
                                                                            const int32_t *__restrict__ live,
                                                                            const int32_t *__restrict__ seg,
                                                                            const int32_t *__restrict__ segmem,
                                                                            int nseg, int map, const double *X,
                                                                            double *F, uint32_t *active) {
    const int k = seg_find(seg, nseg, blockIdx.x);
    const int mi = segmem[k];
    if (live && !live[mi]) return;
    const BatchMember &m = mem[mi];
    RhsArgs a = m.a;
    a.X = X + m.xoff;
    a.F = F + m.xoff;
    a.active = active ? active + m.celloff : nullptr;
    a.map = map;
    a.redo = m.a.redo;
    a.redo_count = m.counts + 8;
    a.cells = m.bins;
    a.cells_count = m.counts;
    const long long blk = (long long)blockIdx.x - seg[k], nblk = seg[k + 1] - seg[k];
    rhs_cells<1, GMAX, NC, R>(a, m.c, blk, nblk);
}

// the prep records of all lists of all members (the list launch's
// segments), each cell with its list's lane count
template <int GMAX, int NC>
__global__ __launch_bounds__(256) void k_prep_batch_lists(const BatchMember *__restrict__ mem,
                                                          const int32_t *__restrict__ live,
                                                          const int32_t *__restrict__ seg,
                                                          const int32_t *__restrict__ segmem,
                                                          const int32_t *__restrict__ segq, int nseg, BatchLists bl,
                                                          const double *X) {
    const int k = seg_find(seg, nseg, blockIdx.x);
    const int mi = segmem[k];
    if (live && !live[mi]) return;
    const BatchMember &m = mem[mi];
    const DevCfg &c = m.c;
    if (!(NC > 1 || c.has_tcond)) return;
    const int q = segq[k];
    const long long N = m.a.ncells;
    const int32_t *cells = m.bins + (long long)q * N;
    const int32_t *cnt = m.counts + q;
    const double *Xm = X + m.xoff;
    const long long blk = (long long)blockIdx.x - seg[k], nblk = seg[k + 1] - seg[k];
    switch (bl.lanes[q]) {
    case 1: prep_list<1, NC>(Xm, m.cm, N, c, cells, cnt, blk, nblk); break;
    case 2: prep_list<2, NC>(Xm, m.cm, N, c, cells, cnt, blk, nblk); break;
    case 4: prep_list<4, NC>(Xm, m.cm, N, c, cells, cnt, blk, nblk); break;
    case 8: if constexpr (GMAX >= 8) prep_list<8, NC>(Xm, m.cm, N, c, cells, cnt, blk, nblk); break;
    default: if constexpr (GMAX >= 16) prep_list<16, NC>(Xm, m.cm, N, c, cells, cnt, blk, nblk); break;
    }
}

template <int G, int NC>
cipher_status batch_G(const BatchLaunch &L, int precision, hipStream_t st) {
    if (L.do_bin && L.bin.nblocks)
        hipLaunchKernelGGL(k_bin_batch<0>, dim3((unsigned)L.bin.nblocks), dim3(256), 0, st, L.mem, L.live, L.bin.seg,
                           L.bin.segmem, L.bin.nseg, L.nbins, L.X);
    const BatchSegs &s = L.lists;
    // prep records after binning, list by list (members that need them)
    if (L.prep.nblocks && s.nblocks)
        hipLaunchKernelGGL((k_prep_batch_lists<G, NC>), dim3((unsigned)s.nblocks), dim3(256), 0, st, L.mem, L.live,
                           s.seg, s.segmem, s.segq, s.nseg, L.shape, L.X);
    // with the single-phase lists in a launch of their own, the merged
    // launch covers lists 1.. (then list 0 follows on the same stream)
    const bool own0 = NC > 1 && L.lists0.nblocks > 0;
    const BatchSegs &sm = own0 ? L.listsN : s;
    const BatchSegs &s0 = L.lists0;
    if (s.nblocks) {
        if (precision == CIPHER_FP32_MIXED) {
            if (sm.nblocks)
                hipLaunchKernelGGL((k_rhs_batch<G, NC, float>), dim3((unsigned)sm.nblocks), dim3(256), sm.lds, st,
                                   L.mem, L.live, sm.seg, sm.segmem, sm.segq, sm.nseg, L.shape, 0, L.X, L.F, L.active);
            if constexpr (NC > 1) {
                if (own0)
                    hipLaunchKernelGGL((k_rhs_batch_one<G, NC, float>), dim3((unsigned)s0.nblocks), dim3(256), s0.lds,
                                       st, L.mem, L.live, s0.seg, s0.segmem, s0.nseg, L.shape.map[0], L.X, L.F,
                                       L.active);
            }
            const BatchSegs &r = L.relists;
            hipLaunchKernelGGL((k_rhs_batch<G, NC, double>), dim3((unsigned)r.nblocks), dim3(256), r.lds, st, L.mem,
                               L.live, r.seg, r.segmem, r.segq, r.nseg, L.shape, 1, L.X, L.F, L.active);
        } else {
            if (sm.nblocks)
                hipLaunchKernelGGL((k_rhs_batch<G, NC, double>), dim3((unsigned)sm.nblocks), dim3(256), sm.lds, st,
                                   L.mem, L.live, sm.seg, sm.segmem, sm.segq, sm.nseg, L.shape, 0, L.X, L.F, L.active);
            if constexpr (NC > 1) {
                if (own0)
                    hipLaunchKernelGGL((k_rhs_batch_one<G, NC, double>), dim3((unsigned)s0.nblocks), dim3(256), s0.lds,
                                       st, L.mem, L.live, s0.seg, s0.segmem, s0.nseg, L.shape.map[0], L.X, L.F,
                                       L.active);
            }
        }
    }
    return hipGetLastError() == hipSuccess ? CIPHER_OK : CIPHER_ERR_HIP;
}

template <int NC>
cipher_status launch_batch_NC(const BatchLaunch &L, int precision, hipStream_t st) {
    if (L.G == 4) return batch_G<4, NC>(L, precision, st);
    if (L.G == 8) return batch_G<8, NC>(L, precision, st);
    if (L.G == 16) return batch_G<16, NC>(L, precision, st);
    return CIPHER_ERR_UNSUPPORTED;
}

}  // namespace cipher
