// rhs.hip - the phase-field right-hand side on gfx950.
//
// Restates RHSFunction of micmog/CIPHER (src/cipher.c:27-681) as two kernels:
//
//   k_prep  per-cell composition, volume-ref mobility and thermal
//           conductivity (src/cipher.c:149-233) into a compact scratch record,
//           only when the face fluxes need them (chemistry or conduction);
//   k_rhs   per-cell gather over the cell's faces in face order (the
//           reference scatters faces in the same order, so each cell sums its
//           Laplacian contributions in the same sequence, src/cipher.c:236-389)
//           fused with the cell update (src/cipher.c:392-664).
//
// Mapping: G lanes per cell (G = next power of two >= PF_SIZE), lane g owns
// active slot g.  Per-cell sums over slots are replayed by every lane in the
// reference's loop order through group shuffles, so FP64 results follow the
// reference's rounding sequence.  A block's contiguous X records are staged in
// LDS with coalesced loads and its F records are assembled in LDS and stored
// coalesced, so the AoS state layout costs no extra HBM traffic.
#include "kernels.h"

namespace cipher {

template <int G, class T>
__device__ __forceinline__ T gsh(T v, int src) {
    return __shfl(v, src, G);
}

// interface energy of an interface type at T (src/cipher.c:425, 462-465)
template <class R>
__device__ __forceinline__ R itf_energy(const DevCfg &c, int it, R T) {
    const cipher_interface &f = c.itf[it];
    R e = R(1);
    e *= (R)f.e0;
    if (c.has_iact && c.ts[f.e_act].n) e *= dexp<R>(dtseries<R>(&c.ts[f.e_act], T) / (R)CRG / T);
    return e;
}
// interface mobility of an interface type at T (src/cipher.c:426, 466-469)
template <class R>
__device__ __forceinline__ R itf_mobility(const DevCfg &c, int it, R T) {
    const cipher_interface &f = c.itf[it];
    R m = R(1);
    m *= (R)f.m0;
    if (c.has_iact && c.ts[f.m_act].n) m *= dexp<R>(dtseries<R>(&c.ts[f.m_act], T) / (R)CRG / T);
    return m;
}

// Kahan-compensated accumulator (FP32 mode); plain sum in FP64 mode
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

// ---------------------------------------------------------------------------
// block staging of contiguous AoS records
__device__ __forceinline__ void stage_in(double *lds, const double *X, long long c0, int ncell_blk, int D) {
    const long long base = c0 * D;
    const int tot = ncell_blk * D;
    for (int i = threadIdx.x; i < tot; i += blockDim.x) lds[i] = X[base + i];
}
__device__ __forceinline__ void stage_out(double *F, const double *lds, long long c0, int ncell_blk, int D) {
    const long long base = c0 * D;
    const int tot = ncell_blk * D;
    for (int i = threadIdx.x; i < tot; i += blockDim.x) F[base + i] = lds[i];
}

// ---------------------------------------------------------------------------
// per-cell interpolants of the active slots (src/utility.c:470-535)
template <int G, class R>
__device__ __forceinline__ void interpolant(const DevCfg &c, int n, R phi, R &hh, R &hsum, R &h) {
    hh = (c.interp == CIPHER_INTERP_QUADRATIC) ? phi * phi : phi * phi * (R(3) - R(2) * phi);
    hsum = R(0);
    for (int g = 0; g < n; g++) hsum += gsh<G>(hh, g);
    if (c.interp == CIPHER_INTERP_LINEAR) h = phi;
    else if (n == 1) h = R(1);
    else h = hh / hsum;
}

// in <- dh/dphi^T in for the slot-distributed vector (src/utility.c:478-557)
template <int G, class R>
__device__ __forceinline__ R interp_deriv(const DevCfg &c, int n, R phi, R hh, R in) {
    if (c.interp == CIPHER_INTERP_LINEAR) return in;
    if (n == 1) return R(0);
    R avg = R(0), sum = R(0);
    for (int g = 0; g < n; g++) {
        const R hg = gsh<G>(hh, g);
        const R ig = gsh<G>(in, g);
        sum += hg;
        avg += hg * ig;
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
    for (int gk = 0; gk < n; gk++) {
        const int idk = gsh<G>(myid, gk);
        const R hk = gsh<G>(h, gk);
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

// ---------------------------------------------------------------------------
// k_prep: composition, mobility, conductivity per cell (src/cipher.c:149-233)
template <int G, int NC, class R, class S>
__global__ __launch_bounds__(256) void k_prep(const double *__restrict__ X, S *__restrict__ cm,
                                              long long ncells, DevCfg c) {
    constexpr int ND = NC - 1;
    constexpr int CPB = 256 / G;
    extern __shared__ double lds[];
    const long long c0 = (long long)blockIdx.x * CPB;
    const int nblk = (int)min((long long)CPB, ncells - c0);
    stage_in(lds, X, c0, nblk, c.D);
    __syncthreads();
    const int lc = threadIdx.x / G;
    const int slot = threadIdx.x & (G - 1);
    if (lc >= nblk) return;
    const long long cell = c0 + lc;
    const double *x = lds + lc * c.D;
    const int n = clamp_count(x[c.aso], c.P);
    const bool on = slot < n;
    const int myid = on ? clamp_id(x[c.aso + 1 + slot], c.npf) : 0;
    const R phi = on ? (R)x[c.pfo + slot] : R(0);
    const R T = (R)x[c.tmo];
    R hh, hsum, h;
    interpolant<G, R>(c, n, phi, hh, hsum, h);
    if (!on) h = R(0);
    const int mat = c.phase_material[myid];
    const cipher_material &m = c.mat[mat];
    R comp[NC], mob[NC];
    R tcond = R(0);
    if constexpr (NC > 1) {
        R mui[Cnt<ND>::v];
        interface_mu<G, NC, R>(c, n, myid, h, mui);
        R spim[2 * Cnt<ND>::v], y[2 * NC], me[NC];
        const int ns = on ? m.nsites : 0;
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int k = 0; k < ND; k++) {
                const R mu = (R)x[c.dpo + k];
                spim[s * ND + k] = (s < ns) ? (R)m.stoichiometry[s] * (mu - mui[k]) - (R)x[c.exo + slot * c.SP + s * ND + k]
                                            : R(0);
            }
#pragma unroll
        for (int k = 0; k < 2 * NC; k++) y[k] = R(0);
        if (on) {
            d_sitefrac<NC, R>(c, m, y, spim, T);
            d_mobility<NC, R>(c, m, me, y, T);
        } else {
#pragma unroll
            for (int k = 0; k < NC; k++) me[k] = R(0);
        }
        // per-slot terms h*me and h*nu_s*y, summed in slot order
#pragma unroll
        for (int k = 0; k < NC; k++) { comp[k] = R(0); mob[k] = R(0); }
        for (int g = 0; g < n; g++) {
            const int nsg = gsh<G>(ns, g);
#pragma unroll
            for (int k = 0; k < NC; k++) mob[k] += gsh<G>(h * me[k], g);
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const R st = on && s < ns ? (R)m.stoichiometry[s] : R(0);
#pragma unroll
                for (int k = 0; k < NC; k++) {
                    const R v = gsh<G>(h * st * y[s * NC + k], g);
                    if (s < nsg) comp[k] += v;
                }
            }
        }
        if (c.has_imobc) {
            for (int gk = 0; gk < n; gk++) {
                const int idk = gsh<G>(myid, gk);
                for (int gj = gk + 1; gj < n; gj++) {
                    const int idj = gsh<G>(myid, gj);
                    const cipher_interface &f = c.itf[c.ilist[idk * c.npf + idj]];
                    if (f.has_mobilityc) {
#pragma unroll
                        for (int k = 0; k < NC; k++) mob[k] *= dexp<R>((R)f.mobilityc[k] / (R)CRG / T);
                    }
                }
            }
        }
    }
    if (c.has_tcond) {
        const cipher_material &mm = c.mat[mat];
        R tv = (on && c.ts[mm.tconductivity].n) ? h * dtseries<R>(&c.ts[mm.tconductivity], T) : R(0);
        for (int g = 0; g < n; g++) {
            const int idg = gsh<G>(myid, g);
            const R v = gsh<G>(tv, g);
            if (c.ts[c.mat[c.phase_material[idg]].tconductivity].n) tcond += v;
        }
    }
    if (slot == 0) {
        S *o = cm + cell * (2 * NC + 1);
        if constexpr (NC > 1) {
#pragma unroll
            for (int k = 0; k < NC; k++) { o[k] = (S)comp[k]; o[NC + k] = (S)mob[k]; }
        }
        o[2 * NC] = (S)tcond;
    }
}

// ---------------------------------------------------------------------------
// k_rhs: fused face gather + cell right-hand side
template <int G, int NC, class R, class S>
__global__ __launch_bounds__(256) void k_rhs(RhsArgs<S> a, DevCfg c) {
    constexpr int ND = NC - 1;
    constexpr int NDV = Cnt<ND>::v;
    constexpr int CPB = 256 / G;
    extern __shared__ double lds[];
    double *ldsX = lds;
    double *ldsF = lds + CPB * c.D;
    const long long c0 = (long long)blockIdx.x * CPB;
    const int nblk = (int)min((long long)CPB, a.ncells - c0);
    stage_in(ldsX, a.X, c0, nblk, c.D);
    for (int i = threadIdx.x; i < nblk * c.D; i += blockDim.x) ldsF[i] = 0.0;
    __syncthreads();
    const int lc = threadIdx.x / G;
    const int slot = threadIdx.x & (G - 1);
    if (lc < nblk) {
        const long long cell = c0 + lc;
        const double *x = ldsX + lc * c.D;
        double *fr = ldsF + lc * c.D;
        const int D = c.D;
        const int n = clamp_count(x[c.aso], c.P);
        const bool on = slot < n;
        const int myid = on ? clamp_id(x[c.aso + 1 + slot], c.npf) : 0xFFFF;
        const int id0 = clamp_id(x[c.aso + 1], c.npf);
        const double phid = on ? x[c.pfo + slot] : 0.0;
        const R phi = (R)phid;
        const R T = (R)x[c.tmo];
        const double hS = a.cellgeom[cell];
        const int dim = c.dim;

        // ---- faces (src/cipher.c:236-317), in face order --------------------
        Acc<R> lap;
        Acc<R> dlap[NDV];
        Acc<R> tlap;
        const int kb = a.cell_off[cell], ke = a.cell_off[cell + 1];
        for (int k = kb; k < ke; k++) {
            const int e = a.adj[k];
            const long long nb = e >> 1;
            const bool selfL = (e & 1) == 0;
            const double hN = a.cellgeom[nb];
            const double dL = selfL ? hS : hN, dR = selfL ? hN : hS;
            const double vL = dfastpow(dL, (unsigned)dim), vR = dfastpow(dR, (unsigned)dim);
            const double ffd = 2.0 * (dL < dR ? vL / dL : vR / dR) / (dL + dR);
            const R ff = (R)ffd, volL = (R)vL, volR = (R)vR;
            const double *xn = a.X + nb * D;
            const int nbn = clamp_count(xn[c.aso], c.P);
            const int Ln = selfL ? n : nbn;
            const int Lid0 = selfL ? id0 : clamp_id(xn[c.aso + 1], c.npf);
            if (Ln > 1 || Lid0 > 1) {
                const int nbid = slot < nbn ? clamp_id(xn[c.aso + 1 + slot], c.npf) : 0xFFFE;
                const R nbphi = slot < nbn ? (R)xn[c.pfo + slot] : R(0);
                int pos = -1;
                for (int t = 0; t < nbn; t++) {
                    const int v = gsh<G>(nbid, t);
                    if (v == myid) pos = t;
                }
                const R pv = gsh<G>(nbphi, pos < 0 ? 0 : pos);
                if (on && pos >= 0) {
                    if (selfL) {
                        const R fl = pv - phi;
                        lap.add(ff * fl / volL);
                    } else {
                        const R fl = phi - pv;
                        lap.add(-(ff * fl / volR));
                    }
                }
            }
            if constexpr (NC > 1) {
                const S *cS = a.cm + cell * (2 * NC + 1), *cN = a.cm + nb * (2 * NC + 1);
                const S *cmL = selfL ? cS : cN, *cmR = selfL ? cN : cS;
                R me[NC], ca[NC], MB[ND * ND], dmu[ND], fd[ND];
#pragma unroll
                for (int q = 0; q < NC; q++) {
                    const R mL = (R)cmL[NC + q], mR = (R)cmR[NC + q];
                    if (fabs((double)mL) > 1e-32 && fabs((double)mR) > 1e-32) me[q] = R(2) * mL * mR / (mL + mR);
                    else me[q] = R(0);
                    const R cL = (R)cmL[q], cR = (R)cmR[q];
                    ca[q] = R(2) * cL * cR / (cL + cR);
                }
#pragma unroll
                for (int i = 0; i < ND; i++)
#pragma unroll
                    for (int j = 0; j < ND; j++) {
                        R v = R(0);
#pragma unroll
                        for (int q = 0; q < NC; q++)
                            v += ((j == q ? R(1) : R(0)) - ca[j]) * ((q == i ? R(1) : R(0)) - ca[i]) * me[q];
                        MB[i * ND + j] = v;
                    }
                const double *xL = selfL ? x : xn, *xR = selfL ? xn : x;
#pragma unroll
                for (int q = 0; q < ND; q++) dmu[q] = (R)(xR[c.dpo + q] - xL[c.dpo + q]);
                dmatvec<ND, R>(fd, MB, dmu);
#pragma unroll
                for (int q = 0; q < ND; q++) {
                    if (selfL) dlap[q].add(ff * fd[q] / volL);
                    else dlap[q].add(-(ff * fd[q] / volR));
                }
            }
            if (c.has_tcond) {
                const S *cS = a.cm + cell * (2 * NC + 1), *cN = a.cm + nb * (2 * NC + 1);
                const R tS = (R)cS[2 * NC], tN = (R)cN[2 * NC];
                const R tL = selfL ? tS : tN, tR = selfL ? tN : tS;
                if (fabs((double)tL) > 1e-32 && fabs((double)tR) > 1e-32) {
                    const double TL = selfL ? x[c.tmo] : xn[c.tmo], TR = selfL ? xn[c.tmo] : x[c.tmo];
                    const R ft = R(2) * tL * tR / (tL + tR) * (R)(TR - TL);
                    if (selfL) tlap.add(ff * ft / volL);
                    else tlap.add(-(ff * ft / volR));
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
                    R fd[NDV];
                    R ft = R(0);
                    if (bc.chem_type != CIPHER_BC_NONE) {
#pragma unroll
                        for (int q = 0; q < ND; q++) {
                            fd[q] = R(0);
                            if (bc.chem_mask & (1 << q)) {
                                if (bc.chem_type == CIPHER_BC_NEUMANN) fd[q] = (R)bc.chem_val[q] * dL;
                                else fd[q] = (R)a.cm[cell * (2 * NC + 1) + NC + q] * ((R)bc.chem_val[q] - (R)x[c.dpo + q]);
                            }
                        }
                    }
                    const R tc = c.has_tcond ? (R)a.cm[cell * (2 * NC + 1) + 2 * NC] : R(0);
                    if (bc.thermal_type == CIPHER_BC_NEUMANN) ft = tc * (R)bc.thermal_val * dL;
                    else if (bc.thermal_type == CIPHER_BC_DIRICHLET) ft = tc * ((R)bc.thermal_val - T);
                    const R cf = R(1) / dfastpow(dL, 2u);
                    if (bc.chem_type != CIPHER_BC_NONE) {
#pragma unroll
                        for (int q = 0; q < ND; q++) dlap[q].add(cf * fd[q]);
                    }
                    if (bc.thermal_type != CIPHER_BC_NONE) tlap.add(cf * ft);
                }
            }
        }
        const R plap = lap.s;

        // ---- cell (src/cipher.c:392-664) ------------------------------------
        R hh, hsum, h;
        interpolant<G, R>(c, n, phi, hh, hsum, h);
        if (!on) { h = R(0); hh = R(0); }
        const int mat = on ? c.phase_material[myid] : 0;
        const cipher_material &m = c.mat[mat];
        const int ns = on ? m.nsites : 0;

        // site fractions of the slot's phase (same arithmetic as k_prep)
        R y[2 * NC];
        R mui[NDV];
        R spim[2 * NDV];
#pragma unroll
        for (int k = 0; k < 2 * NC; k++) y[k] = R(0);
        if constexpr (NC > 1) {
            interface_mu<G, NC, R>(c, n, myid, h, mui);
#pragma unroll
            for (int s = 0; s < 2; s++)
#pragma unroll
                for (int k = 0; k < ND; k++)
                    spim[s * ND + k] = (s < ns) ? (R)m.stoichiometry[s] * ((R)x[c.dpo + k] - mui[k])
                                                  - (R)x[c.exo + slot * c.SP + s * ND + k]
                                                : R(0);
            if (on) d_sitefrac<NC, R>(c, m, y, spim, T);
        } else {
            if (on) {
                R dummy[2] = {R(0), R(0)};
                d_sitefrac<NC, R>(c, m, y, dummy, T);
            }
        }

        R pfrhs = R(0);
        bool act = false;
        if (n > 1) {
            // capillary driving force (src/cipher.c:474-497), replayed per slot
            R cap = R(0);
            for (int gk = 0; gk < n; gk++) {
                const R phk = gsh<G>(phi, gk), lapk = gsh<G>(plap, gk);
                const int idk = gsh<G>(myid, gk);
                for (int gj = gk + 1; gj < n; gj++) {
                    const R phj = gsh<G>(phi, gj), lapj = gsh<G>(plap, gj);
                    const int idj = gsh<G>(myid, gj);
                    const int it = c.ilist[idk * c.npf + idj];
                    const R e = itf_energy<R>(c, it, T);
                    const R w = (R)c.itf[it].width;
                    if (slot == gk) cap -= e * (phj / w + lapj * w / (R)CPI / (R)CPI);
                    if (slot == gj) cap -= e * (phk / w + lapk * w / (R)CPI / (R)CPI);
                    for (int gi = gj + 1; gi < n; gi++) {
                        const R phi_i = gsh<G>(phi, gi);
                        const int idi = gsh<G>(myid, gi);
                        const int it2 = c.ilist[idk * c.npf + idi];
                        const int it3 = c.ilist[idj * c.npf + idi];
                        R tj = R(0);
                        const R r1 = e / w;
                        tj = tj > r1 ? tj : r1;
                        const R r2 = itf_energy<R>(c, it2, T) / (R)c.itf[it2].width;
                        tj = tj > r2 ? tj : r2;
                        const R r3 = itf_energy<R>(c, it3, T) / (R)c.itf[it3].width;
                        tj = tj > r3 ? tj : r3;
                        tj *= (R)c.junctionpenalty;
                        if (slot == gk) cap -= tj * phj * phi_i;
                        if (slot == gj) cap -= tj * phk * phi_i;
                        if (slot == gi) cap -= tj * phk * phj;
                    }
                }
                if (slot == gk) {
                    cap *= R(8);
                    cap += lapk;
                }
            }
            // chemical driving force (src/cipher.c:499-522; Chemenergy, src/material.c:480-490)
            R chs = R(0);
            if (on) {
                chs = d_chemenergy<NC, R>(c, m, y, T);
                chs /= (R)m.molarvolume;
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    if (s >= ns) break;
#pragma unroll
                    for (int k = 0; k < ND; k++) chs -= (R)x[c.dpo + k] * (R)m.stoichiometry[s] * y[s * NC + k];
                }
            }
            if constexpr (NC > 1) {
                if (c.has_potential) {
                    R wdp[NDV];
#pragma unroll
                    for (int k = 0; k < ND; k++) wdp[k] = R(0);
                    for (int g = 0; g < n; g++) {
                        const int nsg = gsh<G>(ns, g);
#pragma unroll
                        for (int s = 0; s < 2; s++) {
                            const R st = (s < ns) ? (R)m.stoichiometry[s] : R(0);
#pragma unroll
                            for (int k = 0; k < ND; k++) {
                                const R v = gsh<G>(h * st * y[s * NC + k], g);
                                if (s < nsg) wdp[k] += v;
                            }
                        }
                    }
                    for (int gk = 0; gk < n; gk++) {
                        const int idk = gsh<G>(myid, gk);
                        const R hk = gsh<G>(h, gk);
                        for (int gj = gk + 1; gj < n; gj++) {
                            const int idj = gsh<G>(myid, gj);
                            const R hj = gsh<G>(h, gj);
                            const cipher_interface &f = c.itf[c.ilist[idk * c.npf + idj]];
                            if (f.has_potential) {
                                R rv = R(0);
#pragma unroll
                                for (int k = 0; k < ND; k++) rv += (R)f.potential[k] * wdp[k];
                                if (slot == gk) chs += rv * hj;
                                if (slot == gj) chs += rv * hk;
                            }
                        }
                    }
                }
            }
            chs = interp_deriv<G, R>(c, n, phi, hh, chs);
            // unconstrained rate and active flags (src/cipher.c:524-539)
            R ru = R(0);
            R nact = R(0);
            for (int gk = 0; gk < n; gk++) {
                const R capk = gsh<G>(cap, gk), chk = gsh<G>(chs, gk), hk = gsh<G>(h, gk);
                const int idk = gsh<G>(myid, gk);
                for (int gj = gk + 1; gj < n; gj++) {
                    const R capj = gsh<G>(cap, gj), chj = gsh<G>(chs, gj), hj = gsh<G>(h, gj);
                    const int idj = gsh<G>(myid, gj);
                    const R mkj = itf_mobility<R>(c, c.ilist[idk * c.npf + idj], T);
                    const R rv = mkj * ((capk - capj) - (chk - chj) * R(8) * sqrt(hk * hj) / (R)CPI);
                    if (slot == gk) ru += rv;
                    if (slot == gj) ru -= rv;
                }
                if (slot == gk)
                    act = (phid > CIPHER_TOL && phid < 1.0 - CIPHER_TOL) || (phid < CIPHER_TOL && ru > R(0))
                          || (phid > 1.0 - CIPHER_TOL && ru < R(0));
            }
            for (int g = 0; g < n; g++) nact += gsh<G>((int)act, g) ? R(1) : R(0);
            // constrained rate (src/cipher.c:541-555)
            for (int gk = 0; gk < n; gk++) {
                const int ak = gsh<G>((int)act, gk);
                if (!ak) continue;
                const R capk = gsh<G>(cap, gk), chk = gsh<G>(chs, gk), hk = gsh<G>(h, gk);
                const int idk = gsh<G>(myid, gk);
                for (int gj = gk + 1; gj < n; gj++) {
                    const int aj = gsh<G>((int)act, gj);
                    const R capj = gsh<G>(cap, gj), chj = gsh<G>(chs, gj), hj = gsh<G>(h, gj);
                    const int idj = gsh<G>(myid, gj);
                    if (!aj) continue;
                    const R mkj = itf_mobility<R>(c, c.ilist[idk * c.npf + idj], T);
                    const R rv = mkj * ((capk - capj) - (chk - chj) * R(8) * sqrt(hk * hj) / (R)CPI);
                    if (slot == gk) pfrhs += rv;
                    if (slot == gj) pfrhs -= rv;
                }
                if (slot == gk) pfrhs /= nact;
            }
        }
        if (a.active) {
            const unsigned long long bal = __ballot(act);
            const int gbase = (threadIdx.x & 63) & ~(G - 1);
            if (slot == 0) a.active[cell] = (uint32_t)((bal >> gbase) & ((G == 64) ? ~0ull : ((1ull << G) - 1)));
        }

        // temperature rate (src/cipher.c:558-568)
        R tmrhs = tlap.s;
        if (c.has_theat || c.has_trate) {
            R cpv = R(0);
            const R cpt = (on && c.ts[m.specific_heat].n) ? h * dtseries<R>(&c.ts[m.specific_heat], T) : R(0);
            const R lht = (on && c.ts[m.latent_heat].n) ? pfrhs * dtseries<R>(&c.ts[m.latent_heat], T) : R(0);
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

        R dprhs[NDV];
#pragma unroll
        for (int k = 0; k < NDV; k++) dprhs[k] = R(0);
        R cprhs[2 * NDV];
#pragma unroll
        for (int k = 0; k < 2 * NDV; k++) cprhs[k] = R(0);
        if constexpr (NC > 1) {
            // explicit potential rate (src/cipher.c:572-581)
            R spw[2 * NDV];
            if (on) {
                d_sitepot_explicit<NC, R>(c, m, spw, y, T);
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    if (s >= ns) break;
#pragma unroll
                    for (int k = 0; k < ND; k++)
                        cprhs[s * ND + k] = (R)m.chempot_ex_kineticcoeff
                                            * (spw[s * ND + k] - (R)x[c.exo + slot * c.SP + s * ND + k]);
                }
            }
            // composition rate (src/cipher.c:584-662)
            R wcp[NDV];
#pragma unroll
            for (int k = 0; k < ND; k++) wcp[k] = R(0);
            if (c.has_potential) {
                for (int gk = 0; gk < n; gk++) {
                    const R hk = gsh<G>(h, gk);
                    const int idk = gsh<G>(myid, gk);
                    for (int gj = gk + 1; gj < n; gj++) {
                        const R hj = gsh<G>(h, gj);
                        const int idj = gsh<G>(myid, gj);
                        const cipher_interface &f = c.itf[c.ilist[idk * c.npf + idj]];
                        if (f.has_potential) {
#pragma unroll
                            for (int k = 0; k < ND; k++) {
                                if (slot == gk) wcp[k] += (R)f.potential[k] * hj;
                                if (slot == gj) wcp[k] += (R)f.potential[k] * hk;
                            }
                        }
                    }
                }
            }
            R tpot[2 * ND * ND], tT[2 * ND];
            R exs = R(0);
#pragma unroll
            for (int k = 0; k < 2 * ND * ND; k++) tpot[k] = R(0);
#pragma unroll
            for (int k = 0; k < 2 * ND; k++) tT[k] = R(0);
            if (on) d_sitefrac_tangent<NC, R>(c, m, tpot, tT, exs, y, spim, T);
            R dcdm[ND * ND];
#pragma unroll
            for (int q = 0; q < ND * ND; q++) dcdm[q] = R(0);
            for (int g = 0; g < n; g++) {
                const int nsg = gsh<G>(ns, g);
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    const R st = (s < ns) ? (R)m.stoichiometry[s] : R(0);
#pragma unroll
                    for (int q = 0; q < ND * ND; q++) {
                        const R v = gsh<G>(h * st * st * tpot[s * ND * ND + q], g);
                        if (s < nsg) dcdm[q] += v;
                    }
                }
            }
            {
                R tmp[ND];
#pragma unroll
                for (int k = 0; k < ND; k++) tmp[k] = wcp[k];
                dmatvec<ND, R>(wcp, dcdm, tmp);
            }
            R wdp[ND];
#pragma unroll
            for (int k = 0; k < ND; k++) wdp[k] = R(0);
            if (n > 1) {
#pragma unroll
                for (int q = 0; q < ND; q++) {
                    R wpf = R(0);
#pragma unroll
                    for (int s = 0; s < 2; s++)
                        if (s < ns) wpf += (R)m.stoichiometry[s] * y[s * NC + q];
                    wpf -= wcp[q];
                    if (!on) wpf = R(0);
                    wpf = interp_deriv<G, R>(c, n, phi, hh, wpf);
                    for (int g = 0; g < n; g++) wdp[q] -= gsh<G>(pfrhs * wpf, g);
                }
            }
            // Chemsource (src/material.c:321-351), called with the interpolants
            R csrc[ND];
#pragma unroll
            for (int k = 0; k < ND; k++) csrc[k] = R(0);
            if (c.has_sources) {
                const S *cmc = a.cm + cell * (2 * NC + 1);
                for (int gk = 0; gk < n; gk++) {
                    const int idk = gsh<G>(myid, gk);
                    const R hk = gsh<G>(h, gk);
                    const cipher_material &mk = c.mat[c.phase_material[idk]];
                    for (int s = 0; s < mk.nsources; s++) {
                        const cipher_source &src = c.src[mk.source + s];
                        if (src.model != CIPHER_SOURCE_SINK) continue;
#pragma unroll
                        for (int q = 0; q < ND; q++)
                            csrc[q] += hk * hk * ((R)src.rate[q] * ((R)src.ceq[q] - (R)cmc[q]));
                    }
                    for (int gj = gk + 1; gj < n; gj++) {
                        const int idj = gsh<G>(myid, gj);
                        const R hj = gsh<G>(h, gj);
                        const cipher_interface &f = c.itf[c.ilist[idk * c.npf + idj]];
                        for (int s = 0; s < f.nsources; s++) {
                            const cipher_source &src = c.src[f.source + s];
                            if (src.model != CIPHER_SOURCE_SINK) continue;
#pragma unroll
                            for (int q = 0; q < ND; q++)
                                csrc[q] += R(2) * hk * hj * ((R)src.rate[q] * ((R)src.ceq[q] - (R)cmc[q]));
                        }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < ND; q++) wdp[q] += (dlap[q].s + csrc[q]);
            // sum_g sum_s h nu (T_ex . EXdot + T_T * Tdot)
            for (int g = 0; g < n; g++) {
                const int nsg = gsh<G>(ns, g);
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    R tv[ND];
                    {
                        R tex[ND * ND];
#pragma unroll
                        for (int q = 0; q < ND * ND; q++) tex[q] = exs == R(0) ? R(0) : -tpot[s * ND * ND + q];
                        dmatvec<ND, R>(tv, tex, &cprhs[s * ND]);
                    }
                    const R st = (s < ns) ? (R)m.stoichiometry[s] : R(0);
#pragma unroll
                    for (int q = 0; q < ND; q++) {
                        const R v = gsh<G>(h * st * (tv[q] + tT[s * ND + q] * tmrhs), g);
                        if (s < nsg) wdp[q] -= v;
                    }
                }
            }
            R dmdc[ND * ND];
            dinvert<ND, R>(dmdc, dcdm);
            dmatvec<ND, R>(dprhs, dmdc, wdp);
        }

        // ---- assemble F in LDS ------------------------------------------------
        if (on) fr[c.pfo + slot] = (double)pfrhs;
        if (slot == 0) {
            fr[c.tmo] = (double)tmrhs;
#pragma unroll
            for (int q = 0; q < ND; q++) fr[c.dpo + q] = (double)dprhs[q];
        }
        if constexpr (NC > 1) {
            if (on) {
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    if (s >= ns) break;
#pragma unroll
                    for (int q = 0; q < ND; q++) fr[c.exo + slot * c.SP + s * ND + q] = (double)cprhs[s * ND + q];
                }
            }
        }
    }
    __syncthreads();
    stage_out(a.F, ldsF, c0, nblk, c.D);
}

// ---------------------------------------------------------------------------
template <int G, int NC, class R, class S>
static cipher_status launch_G(const RhsLaunch &L, hipStream_t st) {
    constexpr int CPB = 256 / G;
    const long long nb = (L.ncells + CPB - 1) / CPB;
    if (nb == 0) return CIPHER_OK;
    const DevCfg &c = *L.cfg;
    const bool need_prep = (NC > 1) || c.has_tcond;
    if (need_prep) {
        const size_t sh = (size_t)CPB * c.D * sizeof(double);
        hipLaunchKernelGGL((k_prep<G, NC, R, S>), dim3((unsigned)nb), dim3(256), sh, st, L.X, (S *)L.cm, L.ncells, c);
    }
    RhsArgs<S> a;
    a.X = L.X; a.F = L.F; a.active = L.active; a.cellgeom = L.cellgeom;
    a.cell_off = L.cell_off; a.adj = L.adj; a.bcell_off = L.bcell_off; a.bface_id = L.bface_id;
    a.cm = (const S *)L.cm; a.ncells = L.ncells;
    const size_t sh = 2 * (size_t)CPB * c.D * sizeof(double);
    hipLaunchKernelGGL((k_rhs<G, NC, R, S>), dim3((unsigned)nb), dim3(256), sh, st, a, c);
    return hipGetLastError() == hipSuccess ? CIPHER_OK : CIPHER_ERR_HIP;
}

template <int NC, class R, class S>
static cipher_status launch_NC(const RhsLaunch &L, hipStream_t st) {
    const int P = L.cfg->P;
    if (P <= 4) return launch_G<4, NC, R, S>(L, st);
    if (P <= 8) return launch_G<8, NC, R, S>(L, st);
    if (P <= 16) return launch_G<16, NC, R, S>(L, st);
    return CIPHER_ERR_UNSUPPORTED;
}

template <class R, class S>
static cipher_status launch_R(const RhsLaunch &L, hipStream_t st) {
    switch (L.cfg->ncp) {
    case 1: return launch_NC<1, R, S>(L, st);
    case 2: return launch_NC<2, R, S>(L, st);
    case 3: return launch_NC<3, R, S>(L, st);
    case 4: return launch_NC<4, R, S>(L, st);
    case 5: return launch_NC<5, R, S>(L, st);
    default: return CIPHER_ERR_UNSUPPORTED;
    }
}

cipher_status launch_rhs(const RhsLaunch &L, int precision, hipStream_t st) {
    if (precision == CIPHER_FP32_MIXED) return launch_R<float, float>(L, st);
    return launch_R<double, double>(L, st);
}

size_t prep_record_bytes(int ncp, int precision) {
    return (size_t)(2 * ncp + 1) * (precision == CIPHER_FP32_MIXED ? sizeof(float) : sizeof(double));
}

}  // namespace cipher
