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

template <int G, int NC>
__global__ __launch_bounds__(256) void k_prep_batch(const BatchMember *__restrict__ mem, const int32_t *__restrict__ live,
                                                    const int32_t *__restrict__ seg, const int32_t *__restrict__ segmem,
                                                    int nseg, const double *__restrict__ X) {
    const int k = seg_find(seg, nseg, blockIdx.x);
    const int mi = segmem[k];
    if (live && !live[mi]) return;
    const BatchMember &m = mem[mi];
    const DevCfg c = m.c;
    prep_cells<G, NC>(X + m.xoff, m.cm, m.a.ncells, c, (long long)blockIdx.x - seg[k]);
}

template <int U>
__global__ __launch_bounds__(256) void k_bin_batch(const BatchMember *__restrict__ mem, const int32_t *__restrict__ live,
                                                   const int32_t *__restrict__ seg, const int32_t *__restrict__ segmem,
                                                   int nseg, int nbins, const double *__restrict__ X) {
    const int k = seg_find(seg, nseg, blockIdx.x);
    const int mi = segmem[k];
    if (live && !live[mi]) return;
    const BatchMember &m = mem[mi];
    const DevCfg c = m.c;
    bin_tile(X + m.xoff, m.a.ncells, c, nbins, m.bins, m.a.ncells, m.counts, (long long)blockIdx.x - seg[k]);
}

// slot 0..4: the member's list of cells that need 2^slot lanes (FP32-mixed
// appends the cells it leaves to recomputation list slot); slot 5..9: the
// recomputation list of list slot-5
template <int G, int GMAX, int NC, class R>
__global__ __launch_bounds__(256, NC == 1 ? 4 : 1) void k_rhs_batch(const BatchMember *__restrict__ mem,
                                                                     const int32_t *__restrict__ live,
                                                                     const int32_t *__restrict__ seg,
                                                                     const int32_t *__restrict__ segmem, int nseg,
                                                                     int slot, const double *X, double *F,
                                                                     uint32_t *active) {
    const int k = seg_find(seg, nseg, blockIdx.x);
    const int mi = segmem[k];
    if (live && !live[mi]) return;
    const BatchMember &m = mem[mi];
    RhsArgs a = m.a;
    a.X = X + m.xoff;
    a.F = F + m.xoff;
    a.active = active ? active + m.celloff : nullptr;
    const int q = slot < 5 ? slot : slot - 5;
    a.redo = m.a.redo + (long long)q * a.ncells;
    a.redo_count = m.counts + 8 + q;
    a.cells = slot < 5 ? m.bins + (long long)q * a.ncells : a.redo;
    a.cells_count = slot < 5 ? m.counts + q : a.redo_count;
    const DevCfg c = m.c;
    rhs_cells<G, GMAX, NC, R>(a, c, (long long)blockIdx.x - seg[k], (long long)(seg[k + 1] - seg[k]));
}

template <int G, int GMAX, int NC, class R>
static void batch_list(const BatchLaunch &L, int slot, const BatchSegs &s, hipStream_t st) {
    if (s.nblocks == 0) return;
    hipLaunchKernelGGL((k_rhs_batch<G, GMAX, NC, R>), dim3((unsigned)s.nblocks), dim3(s.threads), s.lds, st, L.mem, L.live,
                       s.seg, s.segmem, s.nseg, slot, L.X, L.F, L.active);
}

// list q, and in FP32-mixed its FP64 recomputation list, on stream st
template <int G, int GMAX, int NC, class R>
static void batch_bin(const BatchLaunch &L, int q, hipStream_t st) {
    batch_list<G, GMAX, NC, R>(L, q, L.list[q], st);
    if constexpr (sizeof(R) == 4) batch_list<G, GMAX, NC, double>(L, 5 + q, L.list[q], st);
}

template <int GMAX, int NC, class R>
static void batch_lists(const BatchLaunch &L, const hipStream_t *st) {
    if constexpr (GMAX >= 16) batch_bin<16, GMAX, NC, R>(L, 4, st[4]);
    if constexpr (GMAX >= 8) batch_bin<8, GMAX, NC, R>(L, 3, st[3]);
    batch_bin<4, GMAX, NC, R>(L, 2, st[2]);
    batch_bin<2, GMAX, NC, R>(L, 1, st[1]);
    batch_bin<1, GMAX, NC, R>(L, 0, st[0]);
}

template <int G, int NC>
static cipher_status batch_G(const BatchLaunch &L, int phase, int precision, const hipStream_t *st) {
    if (phase == 0) {
        if (L.prep.nblocks)
            hipLaunchKernelGGL((k_prep_batch<G, NC>), dim3((unsigned)L.prep.nblocks), dim3(256), L.prep.lds, st[0],
                               L.mem, L.live, L.prep.seg, L.prep.segmem, L.prep.nseg, L.X);
        if (L.bin.nblocks)
            hipLaunchKernelGGL(k_bin_batch<0>, dim3((unsigned)L.bin.nblocks), dim3(256), 0, st[0], L.mem, L.live,
                               L.bin.seg, L.bin.segmem, L.bin.nseg, L.nbins, L.X);
    } else if (phase == 1) {
        if (precision == CIPHER_FP32_MIXED) batch_lists<G, NC, float>(L, st);
        else batch_lists<G, NC, double>(L, st);
    }
    return hipGetLastError() == hipSuccess ? CIPHER_OK : CIPHER_ERR_HIP;
}

template <int NC>
cipher_status launch_batch_NC(const BatchLaunch &L, int phase, int precision, const hipStream_t *st) {
    if (L.G == 4) return batch_G<4, NC>(L, phase, precision, st);
    if (L.G == 8) return batch_G<8, NC>(L, phase, precision, st);
    if (L.G == 16) return batch_G<16, NC>(L, phase, precision, st);
    return CIPHER_ERR_UNSUPPORTED;
}

}  // namespace cipher
