// rhs_impl.h - the phase-field right-hand side on gfx950 (kernels and launch).
//
// Restates RHSFunction of micmog/CIPHER (src/cipher.c:27-681).  One RHS is
//
//   k_bin       the cells sorted into five lists by active-set size n;
//   k_prep_all  per-cell composition, volume-ref mobility and thermal
//               conductivity (src/cipher.c:149-233) into a compact FP64
//               record, only when the face fluxes need them (chemistry or
//               conduction), list by list;
//   k_rhs_all   per-cell gather over the cell's faces in face order (the
//               reference scatters faces in the same order, so each cell
//               sums its Laplacian contributions in the same sequence,
//               src/cipher.c:236-389) fused with the cell update
//               (src/cipher.c:392-664); all lists in one launch, one block
//               range per list.
//
// Mapping: list q is processed with G = 2^q lanes per cell (or more, a
// launch-shape choice), lane g owning active slot g, so a two-phase cell
// takes two lanes whatever PF_SIZE is.  Per-cell sums over slots are replayed
// by every lane in the reference's loop order through group shuffles, so
// FP64 results follow the reference's rounding sequence.  Faces are handled G
// at a time; the loads of the next chunk's faces run ahead of their use.
//
// FP32-mixed: the phase Laplacian (Kahan-compensated), interpolants,
// capillary terms and pair rates run in FP32; chemical fluxes, site
// potentials, site fractions, chemical energy, driving force, explicit
// potentials and the mu rate stay FP64.  Every cell with a slot whose active
// flag is decided by the sign of an FP32 value inside its rounding band goes
// to its list's recomputation list, which an FP64 launch recomputes, so the
// active sets equal the FP64 path's.
#pragma once
#include <algorithm>

#include "kernels.h"

// minimum resident 256-thread blocks per CU of the right-hand-side kernels
// without (MINB1) and with chemistry (MINBN); this bounds their registers
// per lane: 3 blocks = 3 waves per SIMD = at most 168 VGPRs
#ifndef CIPHER_RHS_MINB1
#define CIPHER_RHS_MINB1 3
#endif
#ifndef CIPHER_RHS_MINBN
#define CIPHER_RHS_MINBN 2
#endif
// the same for the launch of the single-phase cells of a mesh with
// chemistry (one lane per cell, no pair loops): without spills, 2 components
// fit 128 VGPRs (4 waves per SIMD) and 3 components 168 (3 waves); 4 and 5
// components take 200-240 (2 waves, where the merged launch spills 70-150)
#define CIPHER_RHS_MINB_ONE(NC) ((NC) <= 2 ? 4 : ((NC) == 3 ? 3 : 2))

namespace cipher {

template <int G, class T>
__device__ __forceinline__ T gsh(T v, int src) {
    if constexpr (G == 1) return v;
    else return __shfl(v, src, G);
}

// Loops over a group's slots g < n (or a face chunk's faces t < nfc) that sum
// shuffled values.  A group of up
// to four lanes runs all G iterations, unrolled, with the arithmetic guarded
// by g < n: the shuffles of every iteration (which read lanes of the group
// only) issue together instead of one shuffle latency per iteration, and the
// guarded sums keep the loop's order.  Wider groups loop over n.  A guarded
// iteration g >= n must not index a table with a shuffled phase id (the id
// of a slot beyond n is not a phase).  The list kernels unroll (S = G); the
// prep kernels keep the loops over n (S = 64), which measured faster there.
template <int S>
__device__ __forceinline__ constexpr int slot_end(int n) { return S <= 4 ? S : n; }
#define CIPHER_PRAGMA(x) _Pragma(#x)
#define CIPHER_SLOT_UNROLL_OF(S) CIPHER_PRAGMA(unroll S <= 4 ? S : 1)

// interface energy of an interface type at T (src/cipher.c:425, 462-465)
template <class R>
__device__ __forceinline__ R itf_energy(const DevCfg &c, int it, R T) {
    const cipher_interface &f = c.itf[it];
    R e = R(1);
    e *= (R)f.e0;
    if (c.has_iact && c.ts[f.e_act].n) e *= dexp<R>(dts<R>(c, f.e_act, T) / (R)CRG / T);
    return e;
}
// interface mobility of an interface type at T (src/cipher.c:426, 466-469)
template <class R>
__device__ __forceinline__ R itf_mobility(const DevCfg &c, int it, R T) {
    const cipher_interface &f = c.itf[it];
    R m = R(1);
    m *= (R)f.m0;
    if (c.has_iact && c.ts[f.m_act].n) m *= dexp<R>(dts<R>(c, f.m_act, T) / (R)CRG / T);
    return m;
}

// every active lane of the wave holds a cell at the folding temperature of
// the T-series table (then the thermodynamics read folded values without a
// per-lane test); the result is uniform over the wave
__device__ __forceinline__ bool wave_iso(const DevCfg &c, double T) {
    return c.tsv != nullptr && __all(T == c.tsv[c.ntseries]);
}
__device__ __forceinline__ double dts_w(const DevCfg &c, int i, double T, bool iso) {
    return iso ? c.tsv[i] : dts<double>(c, i, T);
}

// Kahan-compensated accumulator in FP32; plain sum (reference order) in FP64
template <class R>
struct Acc {
    R s = R(0), comp = R(0);
    __device__ __forceinline__ void add(R v) {
        if constexpr (sizeof(R) == 4) {
            R y = v - comp;
            R t = s + y;
            comp = (t - s) - y;
            s = t;
        } else {
            s += v;
        }
    }
};

// per-cell interpolants of the active slots (src/utility.c:470-535); S the
// slot-loop bound of slot_end
template <int G, class R, int S = G>
__device__ __forceinline__ void interpolant(const DevCfg &c, int n, R phi, R &hh, R &h) {
    hh = (c.interp == CIPHER_INTERP_QUADRATIC) ? phi * phi : phi * phi * (R(3) - R(2) * phi);
    R hsum = R(0);
    CIPHER_SLOT_UNROLL_OF(S)
    for (int g = 0; g < slot_end<S>(n); g++) {
        const R v = gsh<G>(hh, g);
        if (g < n) hsum += v;
    }
    if (c.interp == CIPHER_INTERP_LINEAR) h = phi;
    else if (n == 1) h = R(1);
    else h = hh / hsum;
}

// in <- dh/dphi^T in for the slot-distributed vector (src/utility.c:478-557)
template <int G, class R, int S = G>
__device__ __forceinline__ R interp_deriv(const DevCfg &c, int n, R phi, R hh, R in) {
    if (c.interp == CIPHER_INTERP_LINEAR) return in;
    if (n == 1) return R(0);
    R avg = R(0), sum = R(0);
    CIPHER_SLOT_UNROLL_OF(S)
    for (int g = 0; g < slot_end<S>(n); g++) {
        const R hg = gsh<G>(hh, g);
        const R ig = gsh<G>(in, g);
        if (g < n) {
            sum += hg;
            avg += hg * ig;
        }
    }
    if (c.interp == CIPHER_INTERP_QUADRATIC) return R(2) * phi * (in - avg / sum) / sum;
    return R(6) * phi * (R(1) - phi) * (in - avg / sum) / sum;
}

// segregation-potential interface sum mu_int[c] = sum_{k<j} h_k h_j V_kj[c]
template <int G, int NC, class R>
__device__ __forceinline__ void interface_mu(const DevCfg &c, int n, int myid, R h, R *mui) {
    constexpr int ND = NC - 1;
#pragma unroll
    for (int k = 0; k < Cnt<ND>::v; k++) mui[k] = R(0);
    if (!c.has_potential) return;
    #pragma unroll 1
    for (int gk = 0; gk < n; gk++) {
        const int idk = gsh<G>(myid, gk);
        const R hk = gsh<G>(h, gk);
        #pragma unroll 1
        for (int gj = gk + 1; gj < n; gj++) {
            const int idj = gsh<G>(myid, gj);
            const R hj = gsh<G>(h, gj);
            const cipher_interface &f = c.itf[c.ilist[idk * c.npf + idj]];
            if (f.has_potential) {
#pragma unroll
                for (int k = 0; k < ND; k++) mui[k] += hk * hj * (R)f.potential[k];
            }
        }
    }
}

// implicit site potentials and site fractions of the slot's phase, in FP64
// (src/cipher.c:170-189 and 584-609): spim = nu_s (mu - mu_int) - EX
template <int G, int NC, bool ISO>
__device__ __forceinline__ void slot_sitefrac(const DevCfg &c, const cipher_material &m, bool on, int ns, int n,
                                              int myid, double h, const double *x, int slot, double T, double *spim,
                                              double *y) {
    constexpr int ND = NC - 1;
    double mui[Cnt<ND>::v];
    interface_mu<G, NC, double>(c, n, myid, h, mui);
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int k = 0; k < ND; k++)
            spim[s * ND + k] = (s < ns) ? m.stoichiometry[s] * (x[c.dpo + k] - mui[k]) - x[c.exo + slot * c.SP + s * ND + k]
                                        : 0.0;
#pragma unroll
    for (int k = 0; k < 2 * NC; k++) y[k] = 0.0;
    if (on) d_sitefrac<NC, double, ISO>(c, m, y, spim, T);
}

// ---------------------------------------------------------------------------
// block staging of the AoS records of consecutive cells
__device__ __forceinline__ void stage_in(double *lds, const double *X, long long c0, int ncell_blk, int D) {
    const long long base = c0 * D;
    const int tot = ncell_blk * D;
    for (int i = threadIdx.x; i < tot; i += blockDim.x) lds[i] = X[base + i];
}

// ---------------------------------------------------------------------------
// k_prep: composition, mobility, conductivity per cell (src/cipher.c:149-233)
// the prep record of one cell, one group of G lanes (slot = the lane's
// active slot); x is the cell's record (staged in LDS or in HBM)
template <int G, int NC>
__device__ __forceinline__ void prep_one(const double *x, long long cell, double *__restrict__ cm, const DevCfg &c,
                                         int slot) {
    constexpr int ND = NC - 1;
    const int n = clamp_count(x[c.aso], c.P);
    const bool on = slot < n;
    const int myid = on ? clamp_id(x[c.aso + 1 + slot], c.npf) : 0;
    const double phi = on ? x[c.pfo + slot] : 0.0;
    const double T = x[c.tmo];
    const bool iso = wave_iso(c, T);
    double hh, h;
    interpolant<G, double, 64>(c, n, phi, hh, h);
    if (!on) h = 0.0;
    const int mat = c.phase_material[myid];
    const cipher_material &m = c.mat[mat];
    double comp[NC], mob[NC];
    double tcond = 0.0;
    if constexpr (NC > 1) {
        double spim[2 * Cnt<ND>::v], y[2 * NC], me[NC];
        const int ns = on ? m.nsites : 0;
        if (iso) slot_sitefrac<G, NC, true>(c, m, on, ns, n, myid, h, x, slot, T, spim, y);
        else slot_sitefrac<G, NC, false>(c, m, on, ns, n, myid, h, x, slot, T, spim, y);
        if (on) {
            if (iso) d_mobility<NC, double, true>(c, m, me, y, T);
            else d_mobility<NC, double, false>(c, m, me, y, T);
        } else {
#pragma unroll
            for (int k = 0; k < NC; k++) me[k] = 0.0;
        }
        // per-slot terms h*me and h*nu_s*y, summed in slot order
#pragma unroll
        for (int k = 0; k < NC; k++) { comp[k] = 0.0; mob[k] = 0.0; }
        CIPHER_SLOT_UNROLL_OF(64)
        for (int g = 0; g < slot_end<64>(n); g++) {
            const int nsg = gsh<G>(ns, g);
#pragma unroll
            for (int k = 0; k < NC; k++) {
                const double v = gsh<G>(h * me[k], g);
                if (g < n) mob[k] += v;
            }
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const double st = on && s < ns ? m.stoichiometry[s] : 0.0;
#pragma unroll
                for (int k = 0; k < NC; k++) {
                    const double v = gsh<G>(h * st * y[s * NC + k], g);
                    if (g < n && s < nsg) comp[k] += v;
                }
            }
        }
        if (c.has_imobc) {
            #pragma unroll 1
            for (int gk = 0; gk < n; gk++) {
                const int idk = gsh<G>(myid, gk);
                #pragma unroll 1
                for (int gj = gk + 1; gj < n; gj++) {
                    const int idj = gsh<G>(myid, gj);
                    const cipher_interface &f = c.itf[c.ilist[idk * c.npf + idj]];
                    if (f.has_mobilityc) {
#pragma unroll
                        for (int k = 0; k < NC; k++) mob[k] *= exp(f.mobilityc[k] / CRG / T);
                    }
                }
            }
        }
    }
    if (c.has_tcond) {
        const double tv = (on && c.ts[m.tconductivity].n) ? h * dts_w(c, m.tconductivity, T, iso) : 0.0;
        #pragma unroll 1
        for (int g = 0; g < n; g++) {
            const int idg = gsh<G>(myid, g);
            const double v = gsh<G>(tv, g);
            if (c.ts[c.mat[c.phase_material[idg]].tconductivity].n) tcond += v;
        }
    }
    if (slot == 0) {
        double *o = cm + cell * (2 * NC + 1);
        if constexpr (NC > 1) {
#pragma unroll
            for (int k = 0; k < NC; k++) { o[k] = comp[k]; o[NC + k] = mob[k]; }
        }
        o[2 * NC] = tcond;
    }
}

// (one block of CPB cells; blk is the block's index within the mesh)
template <int G, int NC>
__device__ __forceinline__ void prep_cells(const double *__restrict__ X, double *__restrict__ cm, long long ncells,
                                           const DevCfg &c, long long blk) {
    constexpr int CPB = 256 / G;
    extern __shared__ double lds[];
    const long long c0 = blk * CPB;
    const int nblk = (int)min((long long)CPB, ncells - c0);
    stage_in(lds, X, c0, nblk, c.D);
    __syncthreads();
    const int lc = threadIdx.x / G;
    const int slot = threadIdx.x & (G - 1);
    if (lc >= nblk) return;
    prep_one<G, NC>(lds + lc * c.D, c0 + lc, cm, c, slot);
}

// the prep records of one per-size list (G lanes per cell, cells read from
// HBM), blocks blk, blk + nblk, ... of it
template <int G, int NC>
__device__ __forceinline__ void prep_list(const double *__restrict__ X, double *__restrict__ cm, long long ncells,
                                          const DevCfg &c, const int32_t *cells, const int32_t *count, long long blk,
                                          long long nblk) {
    const int CPB = blockDim.x / G;
    const int lc = threadIdx.x / G;
    const int slot = threadIdx.x & (G - 1);
    const long long limit = min((long long)*count, ncells);
    for (long long j = blk * CPB + lc; j < limit; j += nblk * CPB) {
        const long long cell = cells[j];
        prep_one<G, NC>(X + cell * c.D, cell, cm, c, slot);
    }
}

template <int G, int NC>
__global__ __launch_bounds__(256) void k_prep(const double *__restrict__ X, double *__restrict__ cm,
                                              long long ncells, DevCfg c) {
    prep_cells<G, NC>(X, cm, ncells, c, blockIdx.x);
}

// ---------------------------------------------------------------------------
// k_output: the fields CIPHER writes at an output step (src/cipher.c:1005-1080)
// for one cell per group of G lanes: the largest interpolant in slot order
// (strict '>' from -1e6, src/typedef.h:25), the last active pair's interface,
// and the composition of the prep record (the reference's output composition
// is the RHS precompute's arithmetic, src/cipher.c:170-200 and 1038-1066)
template <int G, int NC>
__global__ __launch_bounds__(256) void k_output(const double *__restrict__ X, const double *__restrict__ cm,
                                                long long ncells, DevCfg c, OutFields f, double *__restrict__ out) {
    const long long cell = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int slot = threadIdx.x & (G - 1);
    if (cell >= ncells) return;
    const double *x = X + cell * c.D;
    const int n = clamp_count(x[c.aso], c.P);
    const bool on = slot < n;
    const int myid = on ? clamp_id(x[c.aso + 1 + slot], c.npf) : 0;
    const double phi = on ? x[c.pfo + slot] : 0.0;
    double hh, h;
    interpolant<G, double>(c, n, phi, hh, h);
    double best = -1.0e+6;
    int bid = 0;
    #pragma unroll 1
    for (int g = 0; g < n; g++) {
        const double hg = gsh<G>(h, g);
        const int idg = gsh<G>(myid, g);
        if (hg > best) {
            best = hg;
            bid = idg;
        }
    }
    const int ia = gsh<G>(myid, n >= 2 ? n - 2 : 0), ib = gsh<G>(myid, n >= 2 ? n - 1 : 0);
    // composition of a single-component cell: every site holds the one
    // component (y = 1), so it is sum_g h_g sum_s nu_s
    double comp1 = 0.0;
    if constexpr (NC == 1) {
        const cipher_material &m = c.mat[c.phase_material[myid]];
        double sv = 0.0;
        for (int s = 0; s < m.nsites; s++) sv += h * m.stoichiometry[s];
        #pragma unroll 1
        for (int g = 0; g < n; g++) comp1 += gsh<G>(on ? sv : 0.0, g);
    }
    for (int k = 0; k < f.n; k++) {
        const int idx = f.index[k];
        double v = 0.0;
        switch (f.kind[k]) {
        case CIPHER_OUT_PHASEID: v = (double)bid; break;
        case CIPHER_OUT_MATID: v = (double)c.phase_material[bid]; break;
        case CIPHER_OUT_INTERFACEID: v = n >= 2 ? (double)c.ilist[ia * c.npf + ib] : -1.0; break;
        case CIPHER_OUT_CHEMPOT: v = (idx >= 0 && idx < c.ncp - 1) ? x[c.dpo + idx] : 0.0; break;
        case CIPHER_OUT_COMP:
            if (idx >= 0 && idx < c.ncp) v = NC > 1 ? cm[cell * (2 * NC + 1) + idx] : comp1;
            break;
        case CIPHER_OUT_PHI:
            #pragma unroll 1
            for (int g = 0; g < n; g++) {
                const double pg = gsh<G>(phi, g);
                if (gsh<G>(myid, g) == idx) v = pg;
            }
            break;
        case CIPHER_OUT_TEMPERATURE: v = x[c.tmo]; break;
        default: break;
        }
        if (slot == 0) out[cell * f.n + k] = v;
    }
}

template <int G, int NC>
static cipher_status output_G(const DevCfg &c, const double *X, double *cm, long long N, const OutFields &f, double *out,
                              hipStream_t st) {
    constexpr int CPB = 256 / G;
    const long long nb = (N + CPB - 1) / CPB;
    if (nb == 0) return CIPHER_OK;
    if ((NC > 1 || c.has_tcond) && f.need_comp) {
        const size_t sh = (size_t)CPB * c.D * sizeof(double);
        hipLaunchKernelGGL((k_prep<G, NC>), dim3((unsigned)nb), dim3(256), sh, st, X, cm, N, c);
    }
    hipLaunchKernelGGL((k_output<G, NC>), dim3((unsigned)nb), dim3(256), 0, st, X, (const double *)cm, N, c, f, out);
    return hipGetLastError() == hipSuccess ? CIPHER_OK : CIPHER_ERR_HIP;
}

template <int NC>
cipher_status launch_output_NC(const DevCfg &c, const double *X, double *cm, long long N, const OutFields &f,
                               double *out, hipStream_t st) {
    if (c.P <= 4) return output_G<4, NC>(c, X, cm, N, f, out, st);
    if (c.P <= 8) return output_G<8, NC>(c, X, cm, N, f, out, st);
    if (c.P <= 16) return output_G<16, NC>(c, X, cm, N, f, out, st);
    return CIPHER_ERR_UNSUPPORTED;
}

// ---------------------------------------------------------------------------
// interface energy / mobility of a pair from the compact [e0, width, m0]
// table; an activation factor exp(Q(T)/RT) is evaluated in FP64 (Q/RT is
// often 10-20, so an FP32 exponent would lose that many ulps)
template <class R>
__device__ __forceinline__ R pair_e(const DevCfg &c, const double *itab, int it, double T) {
    if (c.has_iact) return (R)itf_energy<double>(c, it, T);
    R e = R(1);
    e *= (R)itab[3 * it];
    return e;
}
template <class R>
__device__ __forceinline__ R pair_m(const DevCfg &c, const double *itab, int it, double T) {
    if (c.has_iact) return (R)itf_mobility<double>(c, it, T);
    R m = R(1);
    m *= (R)itab[3 * it + 2];
    return m;
}

// pairs gk < gj of G slots, and the position of pair (gk, gj) in a
// row-major upper triangle
__host__ __device__ constexpr int pair_count(int G) { return G > 1 ? G * (G - 1) / 2 : 1; }
template <int G>
__device__ __forceinline__ int pair_index(int gk, int gj) {
    return gk * (2 * G - gk - 1) / 2 + (gj - gk - 1);
}

// make LDS writes of this wave visible to its other lanes
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// k_rhs: fused face gather + cell right-hand side over a list of cells whose
// active sets fit G slots (cells are binned by active-set size first, so a
// two-phase cell occupies two lanes, not PF_SIZE).  Each group of G lanes
// stages its cell's record in LDS, computes, and writes the cell's F record
// itself.  In FP32-mixed, cells whose active flags fall inside the FP32
// rounding band are appended to a.redo for an FP64 pass over the same list.
// Blocks are dealt to the 8 XCDs round-robin, each XCD with its own L2.
// Block blk of nblk is given position x*q + min(x, r) + blk/8 of the list
// order (x = blk mod 8, q = nblk/8, r = nblk mod 8), so each XCD takes one
// contiguous range of the (Morton-ordered) list and the neighbour records a
// block reads were mostly fetched by blocks of the same XCD.
__device__ __forceinline__ long long xcd_block(long long blk, long long nblk) {
    const long long q = nblk >> 3, r = nblk & 7, x = blk & 7;
    return x * q + (x < r ? x : r) + (blk >> 3);
}

// (blocks blk, blk + nblk, ... of the list; a block is blockDim/G groups)
template <int G, int GMAX, int NC, class R>
__device__ __forceinline__ void rhs_cells(const RhsArgs &a, const DevCfg &c, long long blk, long long nblk) {
    constexpr int RM = (GMAX + G - 1) / G;                 // neighbour entries per lane
    constexpr int ND = NC - 1;
    constexpr int NDV = Cnt<ND>::v;
    constexpr bool MIXED = sizeof(R) == 4;
    constexpr int SG = G;                                  // slot loops (slot_end)
    extern __shared__ int ptype[];                         // CPB*G*G ints
    const int CPB = blockDim.x / G;
    const int D = c.D;
    const double *itab = c.itab;
    const int lc = threadIdx.x / G;
    const int slot = threadIdx.x & (G - 1);
    int *pt = ptype + lc * G * G;
    const long long limit = min((long long)*a.cells_count, a.ncells);
    const long long stride = nblk * CPB;
    if (a.map) blk = xcd_block(blk, nblk);
    if constexpr (G == 1 && NC == 1) {
        if (!c.has_tcond && !c.has_theat) {
            // Single-phase cells without chemistry, conduction or heat terms:
            // zero phase rate and Laplacian, so F is 0 except the imposed
            // temperature rate (src/cipher.c:558-568).  A wave writes its 64
            // cells' records together: one contiguous range when the cells
            // are consecutive, else as many whole records per step as the
            // wave's lanes cover.  Records of an even length D start on 16
            // bytes (F 16-byte aligned), so they are written as double2.
            const int lane = (int)(threadIdx.x & 63);
            const double tr = c.has_trate ? 0.0 + c.trate : 0.0;
            const int D2 = D >> 1;
            const bool wide = (D & 1) == 0 && D2 <= 64 && ((uintptr_t)a.F & 15) == 0;
            for (long long j0 = blk * CPB + (lc & ~63); j0 < limit; j0 += stride) {
                const long long j = j0 + lane;
                const int cell = j < limit ? a.cells[j] : -1;
                if (a.active && cell >= 0) a.active[cell] = 0u;
                const int c0 = __shfl(cell, 0, 64);
                const int cnt = (int)min((long long)64, limit - j0);
                const bool contig = __all(cell < 0 || cell == c0 + lane);
                if (wide) {
                    if (contig) {
                        double2 *F0 = reinterpret_cast<double2 *>(a.F + (long long)c0 * D);
                        const int tot = cnt * D2;
                        const int step = 64 % D2;
                        int rem = lane % D2;           // double2 index within the record
                        for (int k = lane; k < tot; k += 64) {
                            double2 v = {0.0, 0.0};
                            if (c.has_trate && 2 * rem == c.tmo) v.x = tr;
                            if (c.has_trate && 2 * rem + 1 == c.tmo) v.y = tr;
                            F0[k] = v;
                            rem += step;
                            if (rem >= D2) rem -= D2;
                        }
                    } else {
                        const int ppr = 64 / D2;           // records per step
                        const int r = lane / D2, f = lane - r * D2;
                        double2 v = {0.0, 0.0};
                        if (c.has_trate && 2 * f == c.tmo) v.x = tr;
                        if (c.has_trate && 2 * f + 1 == c.tmo) v.y = tr;
                        for (int i0 = 0; i0 < cnt; i0 += ppr) {
                            const int i = i0 + r;
                            const int ci = __shfl(cell, min(i, 63), 64);
                            if (r < ppr && i < cnt) reinterpret_cast<double2 *>(a.F + (long long)ci * D)[f] = v;
                        }
                    }
                } else if (contig) {
                    double *F0 = a.F + (long long)c0 * D;
                    const int tot = cnt * D;
                    const int step = 64 % D;
                    int rem = lane % D;
                    for (int k = lane; k < tot; k += 64) {
                        F0[k] = (c.has_trate && rem == c.tmo) ? tr : 0.0;
                        rem += step;
                        if (rem >= D) rem -= D;
                    }
                } else {
                    for (int i = 0; i < cnt; i++) {
                        const int ci = __shfl(cell, i, 64);
                        if (lane < D) a.F[(long long)ci * D + lane] = (c.has_trate && lane == c.tmo) ? tr : 0.0;
                    }
                }
            }
            return;
        }
    }
    // the G lanes of a group share the loop trip count; groups of a wave may
    // differ, so nothing below synchronises more than the group's wave
    for (long long j = blk * CPB + lc; j < limit; j += stride) {
        const long long cell = (long long)a.cells[j];
        // the group's lanes read the cell's record together (same or
        // consecutive addresses), so it is served by one or two cache lines
        const double *x = a.X + cell * D;
        wave_sync();
        const int n = clamp_count(x[c.aso], c.P);
        const bool on = slot < n;
        const int myid = on ? clamp_id(x[c.aso + 1 + slot], c.npf) : 0xFFFF;
        const int id0 = clamp_id(x[c.aso + 1], c.npf);
        const double phid = on ? x[c.pfo + slot] : 0.0;
        const R phi = (R)phid;
        const double Td = x[c.tmo];
        const R T = (R)Td;
        const bool iso = wave_iso(c, Td);
        const double hS = a.cellgeom[cell];
        const int dim = c.dim;

        // ---- faces (src/cipher.c:236-317), in face order --------------------
        // Each chunk of up to G faces is first handled face-parallel (lane t
        // owns face f0+t: neighbour scalars, geometry factor, chemical and
        // thermal flux terms), then accumulated into the cell's sums in face
        // order, so every sum keeps the reference's sequence.
        Acc<R> lap;
        Acc<double> dlap[NDV];   // chemical fluxes stay FP64 in both modes
        Acc<R> tlap;
        R lapabs = R(0);
        double cmS[2 * NC + 1];
        if (NC > 1 || c.has_tcond) {
#pragma unroll
            for (int q = 0; q < 2 * NC + 1; q++) cmS[q] = a.cm[cell * (2 * NC + 1) + q];
        }
        // a single-phase cell uses its faces only for chemistry and conduction
        // (its phase-fraction rate is zero), so grain-growth cells skip them
        const bool need_lap = G > 1 && n > 1;   // n <= G: a one-lane list is single-phase
        const int kb = a.cell_off[cell];
        const int ke = (need_lap || NC > 1 || c.has_tcond) ? a.cell_off[cell + 1] : kb;
        // least-squares gradient of the slot's phase fraction
        // (src/cipher.c:103-131): minus the weighted own value over every
        // neighbour first, then the common neighbours' values in face order
        const bool grad_on = c.has_grad && need_lap;
        double gr[3] = {0.0, 0.0, 0.0};
        if (grad_on) {
            for (int k = kb; k < ke; k++)
                for (int d = 0; d < c.dim; d++) gr[d] -= c.gw[(long long)k * c.dim + d] * phid;
        }
        // the face-parallel inputs run ahead of their use: the adjacency
        // entries two chunks ahead, the neighbour's size, count and first id
        // one chunk ahead, so a chunk's dependent loads overlap the previous
        // chunk's work
        int e_n1 = (kb + slot < ke) ? a.adj[kb + slot] : 0;
        int e_n2 = (kb + G + slot < ke) ? a.adj[kb + G + slot] : 0;
        // the cell's volume (the L cell's when it is L, the R cell's when R)
        const double vS = dfastpow(hS, (unsigned)dim);
        const R vol = (R)vS;
        const R ivol = MIXED ? (R)(1.0 / vS) : R(0);    // FP32-mixed: one division per cell
        double ff_n = 0.0, nbnd_n = 0.0, id0d_n = 0.0;
        if (kb + slot < ke) {
            const double *xq = a.X + (long long)(e_n1 >> 1) * D;
            ff_n = a.fgeo[kb + slot];
            nbnd_n = xq[c.aso];
            id0d_n = xq[c.aso + 1];
        }
        for (int f0 = kb; f0 < ke; f0 += G) {
            const int nfc = min(G, ke - f0);
            // -- face-parallel part
            const bool fon = slot < nfc;
            const int e = fon ? e_n1 : 0;
            const double ffd = fon ? ff_n : 0.0;
            const double nbnd = fon ? nbnd_n : 0.0;
            const double id0d = fon ? id0d_n : 0.0;
            // next chunk: its scalars from the entries loaded one chunk ago,
            // and the entries of the chunk after it
            e_n1 = e_n2;
            if (f0 + G + slot < ke) {
                const double *xq = a.X + (long long)(e_n1 >> 1) * D;
                ff_n = a.fgeo[f0 + G + slot];
                nbnd_n = xq[c.aso];
                id0d_n = xq[c.aso + 1];
            }
            e_n2 = (f0 + 2 * G + slot < ke) ? a.adj[f0 + 2 * G + slot] : 0;
            const int nbf = e >> 1;
            const bool selfL = (e & 1) == 0;
            const double *xf = a.X + (long long)nbf * D;
            const R ff = (R)ffd;
            const int nbn = clamp_count(nbnd, c.P);
            const int Ln = selfL ? n : nbn;
            const int Lid0 = selfL ? id0 : clamp_id(id0d, c.npf);
            const int lguard = (fon && need_lap && (Ln > 1 || Lid0 > 1)) ? 1 : 0;
            const int guard = (lguard || (fon && grad_on)) ? 1 : 0;
            // chemical flux (src/cipher.c:290-308) in FP64 in both modes: the
            // divergence of nearly balanced fluxes loses digits that an FP32
            // flux would not have
            double cterm[NDV];
#pragma unroll
            for (int q = 0; q < NDV; q++) cterm[q] = 0.0;
            if constexpr (NC > 1) {
                if (fon) {
                    double cmN[2 * NC + 1];
#pragma unroll
                    for (int q = 0; q < 2 * NC + 1; q++) cmN[q] = a.cm[(long long)nbf * (2 * NC + 1) + q];
                    double me[NC], ca[NC], MB[ND * ND], dmu[ND], fd[ND];
#pragma unroll
                    for (int q = 0; q < NC; q++) {
                        const double mL = selfL ? cmS[NC + q] : cmN[NC + q];
                        const double mR = selfL ? cmN[NC + q] : cmS[NC + q];
                        if (fabs(mL) > 1e-32 && fabs(mR) > 1e-32) me[q] = 2.0 * mL * mR / (mL + mR);
                        else me[q] = 0.0;
                        const double cL = selfL ? cmS[q] : cmN[q], cR = selfL ? cmN[q] : cmS[q];
                        ca[q] = 2.0 * cL * cR / (cL + cR);
                    }
#pragma unroll
                    for (int i = 0; i < ND; i++)
#pragma unroll
                        for (int jj = 0; jj < ND; jj++) {
                            double v = 0.0;
#pragma unroll
                            for (int q = 0; q < NC; q++)
                                v += ((jj == q ? 1.0 : 0.0) - ca[jj]) * ((q == i ? 1.0 : 0.0) - ca[i]) * me[q];
                            MB[i * ND + jj] = v;
                        }
#pragma unroll
                    for (int q = 0; q < ND; q++) {
                        const double mus = x[c.dpo + q], mun = xf[c.dpo + q];
                        dmu[q] = selfL ? (mun - mus) : (mus - mun);
                    }
                    dmatvec<ND, double>(fd, MB, dmu);
                    const double vold = vS;
#pragma unroll
                    for (int q = 0; q < ND; q++) cterm[q] = selfL ? ffd * fd[q] / vold : -(ffd * fd[q] / vold);
                }
            }
            R tterm = R(0);
            int tvalid = 0;
            if (c.has_tcond && fon) {
                const double tSd = cmS[2 * NC];
                const double tNd = a.cm[(long long)nbf * (2 * NC + 1) + 2 * NC];
                const double tLd = selfL ? tSd : tNd, tRd = selfL ? tNd : tSd;
                if (fabs(tLd) > 1e-32 && fabs(tRd) > 1e-32) {
                    const R tL = (R)tLd, tR = (R)tRd;
                    const double TN = xf[c.tmo];
                    const double TL = selfL ? Td : TN, TR = selfL ? TN : Td;
                    const R ft = R(2) * tL * tR / (tL + tR) * (R)(TR - TL);
                    tterm = selfL ? ff * ft / vol : -(ff * ft / vol);
                    tvalid = 1;
                }
            }
            // -- in face order: phase Laplacian over the common active phases
            // the neighbour's active list may be longer than G: lane s holds
            // its entries s, s+G, s+2G, ...
            int gnext = gsh<G>(guard, 0);
            double idn[RM], phn[RM];
            {
                const double *xt = a.X + (long long)gsh<G>(nbf, 0) * D;
                const int cnt = gsh<G>(nbn, 0);
#pragma unroll
                for (int r = 0; r < RM; r++) {
                    const int k = r * G + slot;
                    const bool ld = gnext && k < cnt;
                    idn[r] = ld ? xt[c.aso + 1 + k] : 0.0;
                    phn[r] = ld ? xt[c.pfo + k] : 0.0;
                }
            }
            // (faces t >= nfc of an unrolled chunk do nothing)
            CIPHER_SLOT_UNROLL_OF(SG)
            for (int t = 0; t < slot_end<SG>(nfc); t++) {
                const bool tv = t < nfc;
                const int gt = tv ? gnext : 0;
                double idc[RM], phc[RM];
#pragma unroll
                for (int r = 0; r < RM; r++) { idc[r] = idn[r]; phc[r] = phn[r]; }
                if (t + 1 < nfc) {
                    gnext = gsh<G>(guard, t + 1);
                    const double *xt = a.X + (long long)gsh<G>(nbf, t + 1) * D;
                    const int cnt = gsh<G>(nbn, t + 1);
#pragma unroll
                    for (int r = 0; r < RM; r++) {
                        const int k = r * G + slot;
                        const bool ld = gnext && k < cnt;
                        idn[r] = ld ? xt[c.aso + 1 + k] : 0.0;
                        phn[r] = ld ? xt[c.pfo + k] : 0.0;
                    }
                }
                if (gt) {
                    const int nbnt = gsh<G>(nbn, t);
                    int pos = -1;
#pragma unroll
                    for (int r = 0; r < RM; r++) {
                        const int k0 = r * G;
                        if (k0 < nbnt) {
                            const int nbid = (k0 + slot < nbnt) ? clamp_id(idc[r], c.npf) : 0xFFFE;
                            const int ue = min(G, nbnt - k0);
#pragma unroll
                            for (int u = 0; u < G; u++)
                                if (u < ue && gsh<G>(nbid, u) == myid) pos = k0 + u;
                        }
                    }
                    R pv = R(0);
                    double pvd = 0.0;
#pragma unroll
                    for (int r = 0; r < RM; r++) {
                        if (r * G < nbnt) {
                            const double cand = gsh<G>(phc[r], pos < 0 ? 0 : (pos & (G - 1)));
                            if (pos >= 0 && pos / G == r) pvd = cand;
                        }
                    }
                    pv = (R)pvd;
                    if (grad_on && on && pos >= 0)
                        for (int d = 0; d < c.dim; d++) gr[d] += c.gw[(long long)(f0 + t) * c.dim + d] * pvd;
                    const R fft = gsh<G>(ff, t);
                    const int sL = gsh<G>((int)selfL, t);
                    const int lgt = gsh<G>(lguard, t);     // every lane of the group shuffles
                    if (on && pos >= 0 && lgt) {
                        R term;
                        if constexpr (MIXED) {
                            // the difference is taken in FP64 and rounded once,
                            // so each term carries relative errors only (at most
                            // 5 roundings: difference, two factors, two products)
                            const R fl = sL ? (R)(pvd - phid) : (R)(phid - pvd);
                            term = sL ? fft * fl * ivol : -(fft * fl * ivol);
                        } else if (sL) {
                            const R fl = pv - phi;
                            term = fft * fl / vol;
                        } else {
                            const R fl = phi - pv;
                            term = -(fft * fl / vol);
                        }
                        lap.add(term);
                        if constexpr (MIXED) lapabs += fabsf(term);
                    }
                }
                if constexpr (NC > 1) {
#pragma unroll
                    for (int q = 0; q < ND; q++) {
                        const double v = gsh<G>(cterm[q], t);
                        if (tv) dlap[q].add(v);
                    }
                }
                if (c.has_tcond) {
                    const R tt = gsh<G>(tterm, t);
                    if (tv && gsh<G>(tvalid, t)) tlap.add(tt);
                }
            }
        }
        // ---- boundary faces (src/cipher.c:320-389), per condition ----------
        if (c.nbcs > 0) {
            const int bb = a.bcell_off[cell], be = a.bcell_off[cell + 1];
            for (int b = 0; b < c.nbcs; b++) {
                const cipher_bc &bc = c.bc[b];
                for (int k = bb; k < be; k++) {
                    if (a.bface_id[k] != bc.boundary_id) continue;
                    const R dL = (R)hS;
                    double fd[NDV];
                    R ft = R(0);
                    if (bc.chem_type != CIPHER_BC_NONE) {
#pragma unroll
                        for (int q = 0; q < ND; q++) {
                            fd[q] = 0.0;
                            if (bc.chem_mask & (1 << q)) {
                                if (bc.chem_type == CIPHER_BC_NEUMANN) fd[q] = bc.chem_val[q] * hS;
                                else fd[q] = cmS[NC + q] * (bc.chem_val[q] - x[c.dpo + q]);
                            }
                        }
                    }
                    const R tc = c.has_tcond ? (R)cmS[2 * NC] : R(0);
                    if (bc.thermal_type == CIPHER_BC_NEUMANN) ft = tc * (R)bc.thermal_val * dL;
                    else if (bc.thermal_type == CIPHER_BC_DIRICHLET) ft = tc * ((R)bc.thermal_val - T);
                    const R cf = R(1) / dfastpow(dL, 2u);
                    if (bc.chem_type != CIPHER_BC_NONE) {
                        const double cfd = 1.0 / dfastpow(hS, 2u);
#pragma unroll
                        for (int q = 0; q < ND; q++) dlap[q].add(cfd * fd[q]);
                    }
                    if (bc.thermal_type != CIPHER_BC_NONE) tlap.add(cf * ft);
                }
            }
        }
        const R plap = lap.s;

        // ---- cell (src/cipher.c:392-664) ------------------------------------
        // interpolants in FP64 in both modes: the chemical driving force and
        // the composition rate need them (chemical energies of ~1e9 J/m^3
        // nearly cancel between phases, beyond FP32's reach); FP32-mixed
        // rounds them for its FP32 terms
        double hhd, hd;
        interpolant<G, double, SG>(c, n, phid, hhd, hd);
        if (!on) { hd = 0.0; hhd = 0.0; }
        const R h = (R)hd;
        const int mat = on ? c.phase_material[myid] : 0;
        const cipher_material &m = c.mat[mat];
        const int ns = on ? m.nsites : 0;

        // site fractions of the slot's phase, FP64 (same arithmetic as k_prep)
        double yd[2 * NC], spimd[2 * NDV];
        if (iso) slot_sitefrac<G, NC, true>(c, m, on, ns, n, myid, hd, x, slot, Td, spimd, yd);
        else slot_sitefrac<G, NC, false>(c, m, on, ns, n, myid, hd, x, slot, Td, spimd, yd);
        // interface type of every active pair, row gk of the group's table
        {
            int row[G];
#pragma unroll
            for (int gj = 0; gj < G; gj++) {
                const int idj = gsh<G>(myid, gj);
                row[gj] = (on && gj > slot && gj < n) ? c.ilist[myid * c.npf + idj] : 0;
            }
#pragma unroll
            for (int gj = 0; gj < G; gj++)
                if (on && gj > slot && gj < n) pt[slot * G + gj] = row[gj];
        }
        // energy, width and mobility of every active pair gk < gj, row gk
        // written by lane gk (the pair loops below read them from LDS
        // instead of chasing interface type -> table entry through memory)
        // (FP32-mixed also keeps 1/width and width/pi^2, rounded once from
        // FP64, in place of the FP32 divisions of the pair loops)
        constexpr int PT = MIXED ? 5 : 3;                  // entries per pair
        R *ptab = reinterpret_cast<R *>(ptype + CPB * G * G) + lc * PT * pair_count(G);
        #pragma unroll 1
        for (int gj = slot + 1; gj < n; gj++) {
            const int it = pt[slot * G + gj];
            R *pp = ptab + PT * pair_index<G>(slot, gj);
            const double wd = itab[3 * it + 1];
            pp[1] = (R)wd;
            if (!grad_on) {
                pp[0] = pair_e<R>(c, itab, it, Td);
                pp[2] = pair_m<R>(c, itab, it, Td);
            }
            if constexpr (MIXED) {
                pp[3] = (R)(1.0 / wd);
                pp[4] = (R)(wd / (CPI * CPI));
            }
        }
        // anisotropic interface energy and mobility of every active pair
        // (src/cipher.c:429-469): 1 + sum_dir (n.dir)^4 val along the unit
        // difference of the two gradients, then the constant and the
        // activation factor; row gk of the group's tables
        if (grad_on) {
            #pragma unroll 1
            for (int gj = 0; gj < n; gj++) {
                double gjv[3];
                for (int d = 0; d < c.dim; d++) gjv[d] = gsh<G>(gr[d], gj);
                if (on && gj > slot) {
                    const cipher_interface &f = c.itf[pt[slot * G + gj]];
                    double nv[3], rn = 0.0;
                    for (int d = 0; d < c.dim; d++) {
                        nv[d] = gr[d] - gjv[d];
                        rn += nv[d] * nv[d];
                    }
                    rn = sqrt(rn);
                    if (rn > 1e-16)
                        for (int d = 0; d < c.dim; d++) nv[d] /= rn;
                    double ev = 1.0, mv = 1.0;
                    for (int r = 0; r < f.e_naniso; r++) {
                        double dot = 0.0;
                        for (int d = 0; d < c.dim; d++) dot += nv[d] * c.aniso[f.e_aniso + c.dim * r + d];
                        ev += dfastpow(dot, 4u) * c.aniso[f.e_aniso + c.dim * f.e_naniso + r];
                    }
                    for (int r = 0; r < f.m_naniso; r++) {
                        double dot = 0.0;
                        for (int d = 0; d < c.dim; d++) dot += nv[d] * c.aniso[f.m_aniso + c.dim * r + d];
                        mv += dfastpow(dot, 4u) * c.aniso[f.m_aniso + c.dim * f.m_naniso + r];
                    }
                    ev *= f.e0;
                    if (c.has_iact && c.ts[f.e_act].n) ev *= dexp<double>(dts_w(c, f.e_act, Td, iso) / CRG / Td);
                    mv *= f.m0;
                    if (c.has_iact && c.ts[f.m_act].n) mv *= dexp<double>(dts_w(c, f.m_act, Td, iso) / CRG / Td);
                    ptab[PT * pair_index<G>(slot, gj)] = (R)ev;
                    ptab[PT * pair_index<G>(slot, gj) + 2] = (R)mv;
                }
            }
        }
        wave_sync();

        double pfrhs = 0.0;
        bool act = false;
        bool unsure = false;
        if (G > 1 && n > 1) {
            // capillary driving force (src/cipher.c:474-497), replayed per slot
            R cap = R(0), capabs = R(0);
            // (pairs and junctions of slots beyond n read pair-table entries
            // of the group that hold no pair: in bounds, and discarded)
            CIPHER_SLOT_UNROLL_OF(SG)
            for (int gk = 0; gk < slot_end<SG>(n); gk++) {
                const R phk = gsh<G>(phi, gk), lapk = gsh<G>(plap, gk);
                CIPHER_SLOT_UNROLL_OF(SG)
                for (int gj = gk + 1; gj < slot_end<SG>(n); gj++) {
                    const R phj = gsh<G>(phi, gj), lapj = gsh<G>(plap, gj);
                    const R *pkj = ptab + PT * pair_index<G>(gk, gj);
                    const R e = pkj[0];
                    const R w = pkj[1];
                    const bool pj = gj < n;
                    if constexpr (MIXED) {
                        const R iw = pkj[3], wp = pkj[4];
                        const R lapaj = gsh<G>(lapabs, gj), lapak = gsh<G>(lapabs, gk);
                        if (pj && slot == gk) {
                            cap -= e * (phj * iw + lapj * wp);
                            capabs += fabsf(e) * (fabsf(phj * iw) + lapaj * fabsf(wp));
                        }
                        if (pj && slot == gj) {
                            cap -= e * (phk * iw + lapk * wp);
                            capabs += fabsf(e) * (fabsf(phk * iw) + lapak * fabsf(wp));
                        }
                    } else {
                        if (pj && slot == gk) cap -= e * (phj / w + lapj * w / (R)CPI / (R)CPI);
                        if (pj && slot == gj) cap -= e * (phk / w + lapk * w / (R)CPI / (R)CPI);
                    }
                    CIPHER_SLOT_UNROLL_OF(SG)
                    for (int gi = gj + 1; gi < slot_end<SG>(n); gi++) {
                        const R phi_i = gsh<G>(phi, gi);
                        const R *pki = ptab + PT * pair_index<G>(gk, gi);
                        const R *pji = ptab + PT * pair_index<G>(gj, gi);
                        R tj = R(0);
                        const R r1 = MIXED ? e * pkj[3] : e / w;
                        tj = tj > r1 ? tj : r1;
                        const R r2 = MIXED ? pki[0] * pki[3] : pki[0] / pki[1];
                        tj = tj > r2 ? tj : r2;
                        const R r3 = MIXED ? pji[0] * pji[3] : pji[0] / pji[1];
                        tj = tj > r3 ? tj : r3;
                        tj *= (R)c.junctionpenalty;
                        R t3 = R(0);
                        if (slot == gk) t3 = tj * phj * phi_i;
                        if (slot == gj) t3 = tj * phk * phi_i;
                        if (slot == gi) t3 = tj * phk * phj;
                        if (gi < n && (slot == gk || slot == gj || slot == gi)) {
                            cap -= t3;
                            if constexpr (MIXED) capabs += fabsf(t3);
                        }
                    }
                }
                if (gk < n && slot == gk) {
                    cap *= R(8);
                    cap += lapk;
                    if constexpr (MIXED) capabs = R(8) * capabs + lapabs;
                }
            }
            // chemical driving force (src/cipher.c:499-522; Chemenergy, src/material.c:480-490), FP64
            double chs = 0.0;
            if (on) {
                double ch = iso ? d_chemenergy<NC, double, true>(c, m, yd, Td) : d_chemenergy<NC, double, false>(c, m, yd, Td);
                ch /= m.molarvolume;
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    if (s >= ns) break;
#pragma unroll
                    for (int k = 0; k < ND; k++) ch -= x[c.dpo + k] * m.stoichiometry[s] * yd[s * NC + k];
                }
                chs = ch;
            }
            if constexpr (NC > 1) {
                if (c.has_potential) {
                    double wdp[NDV];
#pragma unroll
                    for (int k = 0; k < ND; k++) wdp[k] = 0.0;
                    #pragma unroll 1
                    for (int g = 0; g < n; g++) {
                        const int nsg = gsh<G>(ns, g);
#pragma unroll
                        for (int s = 0; s < 2; s++) {
                            const double st = (s < ns) ? m.stoichiometry[s] : 0.0;
#pragma unroll
                            for (int k = 0; k < ND; k++) {
                                const double v = gsh<G>(hd * st * yd[s * NC + k], g);
                                if (s < nsg) wdp[k] += v;
                            }
                        }
                    }
                    #pragma unroll 1
                    for (int gk = 0; gk < n; gk++) {
                        const int idk = gsh<G>(myid, gk);
                        const double hk = gsh<G>(hd, gk);
                        #pragma unroll 1
                        for (int gj = gk + 1; gj < n; gj++) {
                            const int idj = gsh<G>(myid, gj);
                            const double hj = gsh<G>(hd, gj);
                            const cipher_interface &f = c.itf[c.ilist[idk * c.npf + idj]];
                            if (f.has_potential) {
                                double rv = 0.0;
#pragma unroll
                                for (int k = 0; k < ND; k++) rv += f.potential[k] * wdp[k];
                                if (slot == gk) chs += rv * hj;
                                if (slot == gj) chs += rv * hk;
                            }
                        }
                    }
                }
            }
            chs = interp_deriv<G, double, SG>(c, n, phid, hhd, chs);
            // unconstrained rate and active flags (src/cipher.c:524-539); the
            // pair rates are summed in FP64 in both modes
            double ru = 0.0;
            R ruabs = R(0);
            R nact = R(0);
            CIPHER_SLOT_UNROLL_OF(SG)
            for (int gk = 0; gk < slot_end<SG>(n); gk++) {
                const R capk = gsh<G>(cap, gk);
                const double chk = gsh<G>(chs, gk), hk = gsh<G>(hd, gk);
                R capak = R(0);
                if constexpr (MIXED) capak = gsh<G>(capabs, gk);
                CIPHER_SLOT_UNROLL_OF(SG)
                for (int gj = gk + 1; gj < slot_end<SG>(n); gj++) {
                    const R capj = gsh<G>(cap, gj);
                    const double chj = gsh<G>(chs, gj), hj = gsh<G>(hd, gj);
                    const R mkj = ptab[PT * pair_index<G>(gk, gj) + 2];
                    const double rin = (double)(capk - capj) - (chk - chj) * 8.0 * sqrt(hk * hj) / CPI;
                    const double rv = (double)mkj * rin;
                    const bool pj = gj < n;
                    if (pj && slot == gk) ru += rv;
                    if (pj && slot == gj) ru -= rv;
                    if constexpr (MIXED) {
                        // the chemical term is FP64 in both modes (identical
                        // bits); the FP32 errors are those of the two capillary
                        // terms (bounded by their magnitudes), of their
                        // difference and of the mobility's rounding
                        const R capaj = gsh<G>(capabs, gj);
                        if (pj && (slot == gk || slot == gj)) ruabs += fabsf(mkj) * (capak + capaj + (float)fabs(rin));
                    }
                }
                if (gk < n && slot == gk) {
                    act = (phid > CIPHER_TOL && phid < 1.0 - CIPHER_TOL) || (phid < CIPHER_TOL && ru > 0.0)
                          || (phid > 1.0 - CIPHER_TOL && ru < 0.0);
                    // the decision reads the sign of ru: is it inside the FP32
                    // band?  |ru32 - ru64| <= (25 + m) u ruabs, u = 2^-24, m the
                    // accumulation steps of a capillary term (pairs and
                    // junctions of the cell): 7.1u from the compensated
                    // Laplacian, 16.3u from a pair term, m u from summing them,
                    // u from the difference and the mobility; the band takes
                    // (32 + 2m) u
                    if constexpr (MIXED) {
                        const int msum = (n - 1) + (n - 1) * (n - 2) / 2;
                        const float band = (float)(32 + 2 * msum) * 5.9604644775390625e-8f;
                        unsure = !(phid > CIPHER_TOL && phid < 1.0 - CIPHER_TOL) && (float)fabs(ru) <= band * ruabs;
                    }
                }
            }
            CIPHER_SLOT_UNROLL_OF(SG)
            for (int g = 0; g < slot_end<SG>(n); g++) {
                const int ag = gsh<G>((int)act, g);
                if (g < n) nact += ag ? R(1) : R(0);
            }
            // constrained rate (src/cipher.c:541-555); a slot beyond n is
            // never active
            CIPHER_SLOT_UNROLL_OF(SG)
            for (int gk = 0; gk < slot_end<SG>(n); gk++) {
                const int ak = gsh<G>((int)act, gk);
                const R capk = gsh<G>(cap, gk);
                const double chk = gsh<G>(chs, gk), hk = gsh<G>(hd, gk);
                CIPHER_SLOT_UNROLL_OF(SG)
                for (int gj = gk + 1; gj < slot_end<SG>(n); gj++) {
                    const int aj = gsh<G>((int)act, gj);
                    const R capj = gsh<G>(cap, gj);
                    const double chj = gsh<G>(chs, gj), hj = gsh<G>(hd, gj);
                    const R mkj = ptab[PT * pair_index<G>(gk, gj) + 2];
                    const double rv = (double)mkj * ((double)(capk - capj) - (chk - chj) * 8.0 * sqrt(hk * hj) / CPI);
                    const bool use = ak && aj && gj < n;
                    if (use && slot == gk) pfrhs += rv;
                    if (use && slot == gj) pfrhs -= rv;
                }
                if (ak && gk < n && slot == gk) pfrhs /= (double)nact;
            }
        }
        const int gbase = (threadIdx.x & 63) & ~(G - 1);
        const unsigned long long gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << gbase);
        if constexpr (MIXED) {
            if (__ballot(unsure) & gmask) {
                // leave this cell to the FP64 pass
                if (slot == 0) {
                    const int k = atomicAdd(a.redo_count, 1);
                    a.redo[k] = (int32_t)cell;
                }
            }
        }
        if (a.active) {
            const unsigned long long bal = __ballot(act);
            if (slot == 0) a.active[cell] = (uint32_t)((bal & gmask) >> gbase);
        }

        // temperature rate (src/cipher.c:558-568)
        R tmrhs = tlap.s;
        if (c.has_theat || c.has_trate) {
            R cpv = R(0);
            const R cpt = (on && c.ts[m.specific_heat].n) ? h * (R)dts_w(c, m.specific_heat, Td, iso) : R(0);
            const R lht = (on && c.ts[m.latent_heat].n) ? (R)pfrhs * (R)dts_w(c, m.latent_heat, Td, iso) : R(0);
            #pragma unroll 1
            for (int g = 0; g < n; g++) {
                const int idg = gsh<G>(myid, g);
                const cipher_material &mg = c.mat[c.phase_material[idg]];
                const R v1 = gsh<G>(cpt, g), v2 = gsh<G>(lht, g);
                if (c.ts[mg.specific_heat].n) cpv += v1;
                if (c.ts[mg.latent_heat].n) tmrhs += v2;
            }
            if (cpv != R(0)) tmrhs /= cpv;
            if (c.has_trate) tmrhs += (R)c.trate;
        }

        double dprhs[NDV];
#pragma unroll
        for (int k = 0; k < NDV; k++) dprhs[k] = 0.0;
        double cprd[2 * NDV];
#pragma unroll
        for (int k = 0; k < 2 * NDV; k++) cprd[k] = 0.0;
        if constexpr (NC > 1) {
            // explicit potential rate (src/cipher.c:572-581), FP64
            if (on) {
                double spw[2 * NDV];
                if (iso) d_sitepot_explicit<NC, double, true>(c, m, spw, yd, Td);
                else d_sitepot_explicit<NC, double, false>(c, m, spw, yd, Td);
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    if (s >= ns) break;
#pragma unroll
                    for (int k = 0; k < ND; k++)
                        cprd[s * ND + k] = m.chempot_ex_kineticcoeff * (spw[s * ND + k] - x[c.exo + slot * c.SP + s * ND + k]);
                }
            }
            {
                // the EX entries of the cell's F record (unused slots and sites 0)
                double *Fx = a.F + cell * D + c.exo;
                const int msites = c.SP / ND;
                for (int p = slot; p < c.P; p += G) {
                    const bool own = (p == slot) && on;
#pragma unroll
                    for (int s = 0; s < 2; s++) {
                        if (s >= msites) break;
#pragma unroll
                        for (int q = 0; q < ND; q++) Fx[p * c.SP + s * ND + q] = (own && s < ns) ? cprd[s * ND + q] : 0.0;
                    }
                }
            }
            // composition rate (src/cipher.c:584-662), FP64 in both modes
            const double hq = hd, hhq = hhd, pfq = (double)pfrhs, tmq = (double)tmrhs;
            double wcp[NDV];
#pragma unroll
            for (int k = 0; k < ND; k++) wcp[k] = 0.0;
            if (c.has_potential) {
                #pragma unroll 1
                for (int gk = 0; gk < n; gk++) {
                    const double hk = gsh<G>(hq, gk);
                    const int idk = gsh<G>(myid, gk);
                    #pragma unroll 1
                    for (int gj = gk + 1; gj < n; gj++) {
                        const double hj = gsh<G>(hq, gj);
                        const int idj = gsh<G>(myid, gj);
                        const cipher_interface &f = c.itf[c.ilist[idk * c.npf + idj]];
                        if (f.has_potential) {
#pragma unroll
                            for (int k = 0; k < ND; k++) {
                                if (slot == gk) wcp[k] += f.potential[k] * hj;
                                if (slot == gj) wcp[k] += f.potential[k] * hk;
                            }
                        }
                    }
                }
            }
            double tpot[2 * ND * ND], tT[2 * ND];
            double exs = 0.0;
#pragma unroll
            for (int k = 0; k < 2 * ND * ND; k++) tpot[k] = 0.0;
#pragma unroll
            for (int k = 0; k < 2 * ND; k++) tT[k] = 0.0;
            if (on) {
                if (iso) d_sitefrac_tangent<NC, double, true>(c, m, tpot, tT, exs, yd, spimd, Td);
                else d_sitefrac_tangent<NC, double, false>(c, m, tpot, tT, exs, yd, spimd, Td);
            }
            // the lane's explicit-potential tangent terms T_ex . EXdot per
            // sublattice (evaluated once; the tangent arrays die after the
            // composition tangent below)
            double tv[2 * ND];
#pragma unroll
            for (int s = 0; s < 2; s++) {
                double tex[ND * ND];
#pragma unroll
                for (int q = 0; q < ND * ND; q++) tex[q] = exs == 0.0 ? 0.0 : -tpot[s * ND * ND + q];
                double cps[ND];
#pragma unroll
                for (int q = 0; q < ND; q++) cps[q] = cprd[s * ND + q];
                dmatvec<ND, double>(tv + s * ND, tex, cps);
            }
            double dcdm[ND * ND];
#pragma unroll
            for (int q = 0; q < ND * ND; q++) dcdm[q] = 0.0;
            CIPHER_SLOT_UNROLL_OF(SG)
            for (int g = 0; g < slot_end<SG>(n); g++) {
                const int nsg = gsh<G>(ns, g);
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    const double st = (s < ns) ? m.stoichiometry[s] : 0.0;
#pragma unroll
                    for (int q = 0; q < ND * ND; q++) {
                        const double v = gsh<G>(hq * st * st * tpot[s * ND * ND + q], g);
                        if (g < n && s < nsg) dcdm[q] += v;
                    }
                }
            }
            {
                double tmp[ND];
#pragma unroll
                for (int k = 0; k < ND; k++) tmp[k] = wcp[k];
                dmatvec<ND, double>(wcp, dcdm, tmp);
            }
            double dmdc[ND * ND];
            dinvert<ND, double>(dmdc, dcdm);
            double wdp[ND];
#pragma unroll
            for (int k = 0; k < ND; k++) wdp[k] = 0.0;
            if (G > 1 && n > 1) {
#pragma unroll
                for (int q = 0; q < ND; q++) {
                    double wpf = 0.0;
#pragma unroll
                    for (int s = 0; s < 2; s++)
                        if (s < ns) wpf += m.stoichiometry[s] * yd[s * NC + q];
                    wpf -= wcp[q];
                    if (!on) wpf = 0.0;
                    wpf = interp_deriv<G, double, SG>(c, n, phid, hhq, wpf);
                    CIPHER_SLOT_UNROLL_OF(SG)
                    for (int g = 0; g < slot_end<SG>(n); g++) {
                        const double v = gsh<G>(pfq * wpf, g);
                        if (g < n) wdp[q] -= v;
                    }
                }
            }
            // Chemsource (src/material.c:321-351), called with the interpolants
            double csrc[ND];
#pragma unroll
            for (int k = 0; k < ND; k++) csrc[k] = 0.0;
            if (c.has_sources) {
                #pragma unroll 1
                for (int gk = 0; gk < n; gk++) {
                    const int idk = gsh<G>(myid, gk);
                    const double hk = gsh<G>(hq, gk);
                    const cipher_material &mk = c.mat[c.phase_material[idk]];
                    for (int s = 0; s < mk.nsources; s++) {
                        const cipher_source &src = c.src[mk.source + s];
                        if (src.model != CIPHER_SOURCE_SINK) continue;
#pragma unroll
                        for (int q = 0; q < ND; q++)
                            csrc[q] += hk * hk * (src.rate[q] * (src.ceq[q] - cmS[q]));
                    }
                    #pragma unroll 1
                    for (int gj = gk + 1; gj < n; gj++) {
                        const int idj = gsh<G>(myid, gj);
                        const double hj = gsh<G>(hq, gj);
                        const cipher_interface &f = c.itf[c.ilist[idk * c.npf + idj]];
                        for (int s = 0; s < f.nsources; s++) {
                            const cipher_source &src = c.src[f.source + s];
                            if (src.model != CIPHER_SOURCE_SINK) continue;
#pragma unroll
                            for (int q = 0; q < ND; q++)
                                csrc[q] += 2.0 * hk * hj * (src.rate[q] * (src.ceq[q] - cmS[q]));
                        }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < ND; q++) wdp[q] += ((double)dlap[q].s + csrc[q]);
            // sum_g sum_s h nu (T_ex . EXdot + T_T * Tdot)
            CIPHER_SLOT_UNROLL_OF(SG)
            for (int g = 0; g < slot_end<SG>(n); g++) {
                const int nsg = gsh<G>(ns, g);
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    const double st = (s < ns) ? m.stoichiometry[s] : 0.0;
#pragma unroll
                    for (int q = 0; q < ND; q++) {
                        const double v = gsh<G>(hq * st * (tv[s * ND + q] + tT[s * ND + q] * tmq), g);
                        if (g < n && s < nsg) wdp[q] -= v;
                    }
                }
            }
            dmatvec<ND, double>(dprhs, dmdc, wdp);
        }

        // ---- write the cell's F record (AS entries and unused slots are 0) ----
        double *Fc = a.F + cell * D;
        for (int k = slot; k < c.pfo; k += G) Fc[k] = 0.0;
        for (int p = slot; p < c.P; p += G) Fc[c.pfo + p] = (p == slot && on) ? (double)pfrhs : 0.0;
        if (slot == 0) {
            Fc[c.tmo] = (double)tmrhs;
#pragma unroll
            for (int q = 0; q < ND; q++) Fc[c.dpo + q] = dprhs[q];
        }
    }  // cell loop
}

// All per-size lists of one right-hand side in one launch: the grid is cut
// into block ranges, one per list (largest active sets first, so the
// longest-latency cells start earliest), each range striding over its list
// with that list's lane count.  One launch replaces a fork of concurrent
// per-list launches and the cross-stream waits it needs.
template <int GMAX, int NC, class R>
__global__ __launch_bounds__(256, NC == 1 ? CIPHER_RHS_MINB1 : CIPHER_RHS_MINBN) void k_rhs_all(RhsArgs a, const DevCfg *__restrict__ cfg,
                                                                   const int32_t *bins, int32_t *redo, int32_t *counts,
                                                                   ListRanges rg) {
    if (a.pred && *a.pred == 0) return;
    const DevCfg &c = *cfg;
    int r = 0;
    while (r + 1 < rg.n && (int)blockIdx.x >= rg.start[r + 1]) r++;
    const int q = rg.list[r];
    const long long N = a.ncells;
    a.redo = redo + q * N;
    a.redo_count = counts + 8 + q;
    a.cells = rg.redo ? a.redo : bins + q * N;
    a.cells_count = rg.redo ? a.redo_count : counts + q;
    a.map = rg.map[r];
    const long long blk = (long long)blockIdx.x - rg.start[r], nblk = rg.start[r + 1] - rg.start[r];
    switch (rg.lanes[r]) {
    case 1: rhs_cells<1, GMAX, NC, R>(a, c, blk, nblk); break;
    case 2: rhs_cells<2, GMAX, NC, R>(a, c, blk, nblk); break;
    case 4: rhs_cells<4, GMAX, NC, R>(a, c, blk, nblk); break;
    case 8: if constexpr (GMAX >= 8) rhs_cells<8, GMAX, NC, R>(a, c, blk, nblk); break;
    default: if constexpr (GMAX >= 16) rhs_cells<16, GMAX, NC, R>(a, c, blk, nblk); break;
    }
}

// The single-phase list (list 0, one lane per cell) of a mesh with
// chemistry in a launch of its own: its body has no pair loops and needs
// fewer registers than the merged launch, whose register count is set by the
// multi-phase paths, so more of its waves are resident.
template <int GMAX, int NC, class R>
__global__ __launch_bounds__(256, CIPHER_RHS_MINB_ONE(NC)) void k_rhs_one(RhsArgs a, const DevCfg *__restrict__ cfg,
                                                                      const int32_t *bins, int32_t *redo,
                                                                      int32_t *counts, int map) {
    if (a.pred && *a.pred == 0) return;
    a.redo = redo;
    a.redo_count = counts + 8;
    a.cells = bins;
    a.cells_count = counts;
    a.map = map;
    rhs_cells<1, GMAX, NC, R>(a, *cfg, blockIdx.x, gridDim.x);
}

// The prep records of all per-size lists in one launch (block ranges as in
// k_rhs_all, 2^q lanes per cell of list q), so a cell occupies as many lanes
// as its active set needs.
template <int GMAX, int NC>
__global__ __launch_bounds__(256) void k_prep_all(const double *__restrict__ X, double *__restrict__ cm,
                                                  long long ncells, const DevCfg *__restrict__ cfg,
                                                  const int32_t *bins, const int32_t *counts, ListRanges rg,
                                                  const int32_t *pred) {
    if (pred && *pred == 0) return;
    const DevCfg &c = *cfg;
    int r = 0;
    while (r + 1 < rg.n && (int)blockIdx.x >= rg.start[r + 1]) r++;
    const int q = rg.list[r];
    const long long blk = (long long)blockIdx.x - rg.start[r], nblk = rg.start[r + 1] - rg.start[r];
    const int32_t *cells = bins + q * ncells;
    const int32_t *cnt = counts + q;
    switch (rg.lanes[r]) {
    case 1: prep_list<1, NC>(X, cm, ncells, c, cells, cnt, blk, nblk); break;
    case 2: prep_list<2, NC>(X, cm, ncells, c, cells, cnt, blk, nblk); break;
    case 4: prep_list<4, NC>(X, cm, ncells, c, cells, cnt, blk, nblk); break;
    case 8: if constexpr (GMAX >= 8) prep_list<8, NC>(X, cm, ncells, c, cells, cnt, blk, nblk); break;
    default: if constexpr (GMAX >= 16) prep_list<16, NC>(X, cm, ncells, c, cells, cnt, blk, nblk); break;
    }
}

// bin b holds the cells whose active-set size needs 2^b lanes
__device__ __forceinline__ int lane_bin(int n) {
    return n <= 1 ? 0 : (n == 2 ? 1 : (n <= 4 ? 2 : (n <= 8 ? 3 : 4)));
}

// Sort the cells into per-size lists.  A block takes a tile of BIN_TILE
// consecutive cells, each thread BIN_CPT consecutive cells of it; a scan of
// the per-thread counts (shuffles within each wave, then the four wave totals)
// gives every cell its place, and one atomic per bin per tile reserves the
// tile's range, so each list keeps cell order within a tile (and the Morton
// locality of the tiles).  Needs blockDim.x == 256.
__device__ __forceinline__ void bin_tile(const double *__restrict__ X, long long ncells, const DevCfg &c, int nbins,
                                         int32_t *__restrict__ bins, long long binstride,
                                         int32_t *__restrict__ counts, long long blk) {
    __shared__ int wtot[4][5];
    __shared__ int base[4][5];
    const long long c0 = blk * BIN_TILE + (long long)threadIdx.x * BIN_CPT;
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    int bq[BIN_CPT];
    int cnt[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < BIN_CPT; k++) {
        const long long cell = c0 + k;
        bq[k] = cell < ncells ? lane_bin(clamp_count(X[cell * c.D + c.aso], c.P)) : -1;
    }
#pragma unroll
    for (int k = 0; k < BIN_CPT; k++)
#pragma unroll
        for (int b = 0; b < 5; b++) cnt[b] += (bq[k] == b) ? 1 : 0;
    // inclusive scan over the wave's lanes, per bin
    int inc[5];
#pragma unroll
    for (int b = 0; b < 5; b++) inc[b] = cnt[b];
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
        for (int b = 0; b < 5; b++) {
            const int v = __shfl_up(inc[b], (unsigned)off, 64);
            if (lane >= off) inc[b] += v;
        }
    }
    if (lane == 63) {
#pragma unroll
        for (int b = 0; b < 5; b++) wtot[wave][b] = inc[b];
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)nbins) {
        const int b = (int)threadIdx.x;
        const int t0 = wtot[0][b], t1 = wtot[1][b], t2 = wtot[2][b], t3 = wtot[3][b];
        const int tot = t0 + t1 + t2 + t3;
        const int g = tot ? atomicAdd(&counts[b], tot) : 0;
        base[0][b] = g;
        base[1][b] = g + t0;
        base[2][b] = g + t0 + t1;
        base[3][b] = g + t0 + t1 + t2;
    }
    __syncthreads();
    int pos[5];
#pragma unroll
    for (int b = 0; b < 5; b++) pos[b] = (b < nbins ? base[wave][b] : 0) + inc[b] - cnt[b];
#pragma unroll
    for (int k = 0; k < BIN_CPT; k++) {
        const int b = bq[k];
#pragma unroll
        for (int q = 0; q < 5; q++)
            if (b == q && q < nbins) bins[q * binstride + pos[q]++] = (int32_t)(c0 + k);
    }
}

template <int U>
__global__ __launch_bounds__(256) void k_bin(const double *__restrict__ X, long long ncells, DevCfg c, int nbins,
                                              int32_t *__restrict__ bins, long long binstride,
                                              int32_t *__restrict__ counts, const int32_t *pred) {
    if (pred && *pred == 0) return;
    bin_tile(X, ncells, c, nbins, bins, binstride, counts, blockIdx.x);
}

template <int U>
__global__ void k_fill_i32(int32_t *p, int n, int32_t v) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = v;
}


// blocks of list q's range: enough for every cell, at most its cap
static inline long long list_range_blocks(const ListShape &sh, int q, long long N) {
    const int lanes = sh.lanes > 0 ? sh.lanes : (1 << q);
    const int cpb = 256 / lanes;
    long long nb = (N + cpb - 1) / cpb;
    if (nb > sh.cap) nb = sh.cap;
    if (nb < 1) nb = 1;
    return nb;
}

// Ranges of the merged launch for lists qlo..nbins-1 (launch order: largest
// active sets first); *lds: the largest LDS need over the ranges.
static inline ListRanges make_ranges(const ListShape *shape, long long N, const DevCfg &c, int nbins, size_t *lds,
                                     long long *nblocks, int qlo = 0) {
    ListRanges rg{};
    rg.n = nbins - qlo;
    long long tot = 0;
    size_t l = 0;
    for (int r = 0; r < nbins - qlo; r++) {
        const int q = nbins - 1 - r;
        const ListShape &sh = shape[q];
        const int lanes = sh.lanes > 0 ? sh.lanes : (1 << q);
        const int cpb = 256 / lanes;
        const long long nb = list_range_blocks(sh, q, N);
        rg.start[r] = (int32_t)tot;
        rg.list[r] = q;
        rg.lanes[r] = lanes;
        rg.map[r] = sh.map & 1;
        tot += nb;
        l = std::max(l, rhs_lds_bytes(c, lanes, cpb));
    }
    rg.start[nbins - qlo] = (int32_t)tot;
    *lds = l;
    *nblocks = tot;
    return rg;
}

// ---------------------------------------------------------------------------
template <int G, int NC>
cipher_status launch_G(const RhsLaunch &L, int precision, hipStream_t st) {
    constexpr int CPB = 256 / G;
    const long long N = L.ncells;
    const long long nb = (N + CPB - 1) / CPB;
    if (nb == 0) return CIPHER_OK;
    const DevCfg &c = *L.cfg;
    const bool need_prep = (NC > 1) || c.has_tcond;
    RhsArgs a;
    a.X = L.X; a.F = L.F; a.active = L.active; a.cellgeom = L.cellgeom;
    a.cell_off = L.cell_off; a.adj = L.adj; a.fgeo = L.fgeo; a.bcell_off = L.bcell_off; a.bface_id = L.bface_id;
    a.cm = (const double *)L.cm; a.ncells = N;
    a.cells = nullptr; a.cells_count = nullptr;
    a.redo = nullptr; a.redo_count = nullptr; a.map = 0; a.pad_ = 0;
    a.pred = L.pred;
    // counts[0..4]: list sizes, counts[8..12]: their FP32-mixed recomputation
    // lists.  A call that bins finds them cleared (L.fill covers them: this
    // call's own set, or a whole ring of sets cleared once for many calls);
    // a call that reuses the lists of the previous one (same active sets)
    // clears only the recomputation counters, and only in FP32-mixed.
    const int nbins = G >= 16 ? 5 : (G >= 8 ? 4 : 3);
    if (L.bins_mode < 2) {
        if (L.nfill > 0)
            hipLaunchKernelGGL(k_fill_i32<0>, dim3((unsigned)((L.nfill + 255) / 256)), dim3(256), 0, st, L.fill,
                               L.nfill, 0);
        hipLaunchKernelGGL(k_bin<0>, dim3((unsigned)((N + BIN_TILE - 1) / BIN_TILE)), dim3(256), 0, st, L.X, N, c,
                           nbins, L.bins, N, L.counts, L.pred);
    } else if (precision == CIPHER_FP32_MIXED) {
        hipLaunchKernelGGL(k_fill_i32<0>, dim3(1), dim3(64), 0, st, L.counts + 8, RHS_COUNTERS - 8, 0);
    }
    if (need_prep) {
        // composition/mobility/conductivity records, list by list with 2^q
        // lanes per cell (after the lists are known)
        ListShape ps[5];
        for (int q = 0; q < 5; q++) {
            ps[q] = L.shape[q];
            ps[q].lanes = 0;
        }
        size_t l0 = 0;
        long long tp = 0;
        const ListRanges pr = make_ranges(ps, N, c, nbins, &l0, &tp);
        hipLaunchKernelGGL((k_prep_all<G, NC>), dim3((unsigned)tp), dim3(256), 0, st, L.X, (double *)L.cm, N, L.dcfg,
                           (const int32_t *)L.bins, (const int32_t *)L.counts, pr, L.pred);
    }
    size_t lds = 0;
    long long tot = 0;
    // the single-phase list in a launch of its own (after the merged launch
    // of the multi-phase lists, on the same stream)
    const bool own0 = own_launch(L.shape[0], NC);
    ListRanges rg = make_ranges(L.shape, N, c, nbins, &lds, &tot, own0 ? 1 : 0);
    const long long nb0 = own0 ? list_range_blocks(L.shape[0], 0, N) : 0;
    const size_t lds0 = rhs_lds_bytes(c, 1, 256);
    if (precision == CIPHER_FP32_MIXED) {
        hipLaunchKernelGGL((k_rhs_all<G, NC, float>), dim3((unsigned)tot), dim3(256), lds, st, a, L.dcfg, L.bins, L.redo,
                           L.counts, rg);
        if constexpr (NC > 1) {
            if (own0)
                hipLaunchKernelGGL((k_rhs_one<G, NC, float>), dim3((unsigned)nb0), dim3(256), lds0, st, a, L.dcfg,
                                   L.bins, L.redo, L.counts, L.shape[0].map & 1);
        }
        // FP64 recomputation of the cells whose flags sat in the FP32 band;
        // the recomputation lists are short, so each range gets at most
        // RHS_REDO_CAP blocks (they stride over longer lists)
        ListShape rs[5];
        for (int q = 0; q < 5; q++) {
            rs[q] = L.shape[q];
            rs[q].cap = std::min(rs[q].cap, RHS_REDO_CAP);
        }
        size_t lds2 = 0;
        long long tot2 = 0;
        ListRanges rr = make_ranges(rs, N, c, nbins, &lds2, &tot2);
        rr.redo = 1;
        hipLaunchKernelGGL((k_rhs_all<G, NC, double>), dim3((unsigned)tot2), dim3(256), lds2, st, a, L.dcfg, L.bins, L.redo,
                           L.counts, rr);
    } else {
        hipLaunchKernelGGL((k_rhs_all<G, NC, double>), dim3((unsigned)tot), dim3(256), lds, st, a, L.dcfg, L.bins, L.redo,
                           L.counts, rg);
        if constexpr (NC > 1) {
            if (own0)
                hipLaunchKernelGGL((k_rhs_one<G, NC, double>), dim3((unsigned)nb0), dim3(256), lds0, st, a, L.dcfg,
                                   L.bins, L.redo, L.counts, L.shape[0].map & 1);
        }
    }
    return hipGetLastError() == hipSuccess ? CIPHER_OK : CIPHER_ERR_HIP;
}

template <int NC>
cipher_status launch_NC(const RhsLaunch &L, int precision, hipStream_t st) {
    const int P = L.cfg->P;
    if (P <= 4) return launch_G<4, NC>(L, precision, st);
    if (P <= 8) return launch_G<8, NC>(L, precision, st);
    if (P <= 16) return launch_G<16, NC>(L, precision, st);
    return CIPHER_ERR_UNSUPPORTED;
}

}  // namespace cipher
