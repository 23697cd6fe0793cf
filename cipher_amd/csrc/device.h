// device.h - device-side configuration view and CALPHAD/utility functions.
//
// Restates the per-phase constitutive functions of micmog/CIPHER
// (src/material.c, src/utility.c) for gfx950, templated on the number of
// components NC (compile-time loop bounds keep every per-phase array in
// VGPRs) and on the arithmetic type R (double, or float in the mixed mode).
// Operation order follows the reference so that FP64 results agree to
// rounding of the transcendental functions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cipher_amd.h"

#define CPI 3.14159265358979323846264338327950288419716939937510582
#define CRG CIPHER_R_GAS_CONST

struct DevCfg {
    const cipher_tseries *ts;
    const cipher_rk *rk;
    const int32_t *rkindex;
    const cipher_mobility *mob;
    const cipher_material *mat;
    const cipher_interface *itf;
    const cipher_source *src;
    const cipher_bc *bc;
    const double *aniso;
    const int32_t *phase_material;
    const int32_t *ilist;
    const double *itab;       // [e0, width, m0] per interface type
    int has_grad;             // gradient_calculation: anisotropic interfaces
    const double *gw;         // least-squares gradient weights, dim per face-list entry
    double junctionpenalty;
    double trate;
    int has_trate;
    int dim, npf, ncp, nd, nmat, nf, nbcs, interp;
    int P, D, aso, pfo, dpo, exo, tmo, SP, SF;
    int has_tcond;      // some material has a thermal conductivity series
    int has_theat;      // some material has specific or latent heat series
    int has_sources;    // some material/interface has a sink source
    int has_potential;  // some interface has a segregation potential
    int has_imobc;      // some interface has a solute mobility factor
    int has_iact;       // some interface has an energy/mobility activation
    int ntseries;
    // T-series values folded at one temperature: tsv[i] = SumTSeries(tsv[ntseries], ts[i])
    const double *tsv;
    // bit r set: Redlich-Kister entry rk[r] has terms (n > 0); two zero words
    // of padding after the last entry
    const uint64_t *rkbits;
};

// bits [off, off + 64) of rkbits: which of the (at most 64) entries of an RK
// block starting at rk[off] have terms, so the loops over a block skip empty
// entries without loading them (they add nothing to any sum)
__device__ __forceinline__ uint64_t rk_window(const DevCfg &c, int off) {
    const int w = off >> 6, b = off & 63;
    const uint64_t lo = c.rkbits[w];
    return b ? (lo >> b) | (c.rkbits[w + 1] << (64 - b)) : lo;
}

template <class R> __device__ __forceinline__ R dexp(R x);
template <> __device__ __forceinline__ double dexp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float dexp<float>(float x) { return __expf(x); }
template <class R> __device__ __forceinline__ R dlog(R x);
template <> __device__ __forceinline__ double dlog<double>(double x) { return log(x); }
template <> __device__ __forceinline__ float dlog<float>(float x) { return __logf(x); }

// FastPow, src/utility.c:112-124
template <class R>
__device__ __forceinline__ R dfastpow(R base, unsigned e) {
    R r = R(1);
    while (e) {
        if (e & 1u) r *= base;
        e >>= 1;
        base *= base;
    }
    return r;
}

// SumTSeries, src/utility.c:48-62 (out of line: cells at the folding
// temperature never call it, and one copy serves every call site)
template <class R>
__device__ __noinline__ R dtseries(const cipher_tseries *t, R T) {
    R r = R(0);
    const int n = t->n;
    for (int i = 0; i < n; i++) {
        const int e = t->exp[i];
        const R c = (R)t->coeff[i];
        if (e > 0) r += c * dfastpow(T, (unsigned)e);
        else if (e < 0) r += c / dfastpow(T, (unsigned)(-e));
        else r += c;
    }
    const double lc = t->logcoeff;
    if (fabs(lc) > 0.0) r += (R)lc * T * dlog(T);
    return r;
}

// SumTSeries of table entry i; a cell at the folding temperature reads the
// value the fold kernel computed with the same code (identical bits).  ISO:
// the caller has checked that every lane of the wave is at the folding
// temperature, so the table value is returned without a per-lane test.
template <class R, bool ISO = false>
__device__ __forceinline__ R dts(const DevCfg &c, int i, R T) {
    if constexpr (ISO) {
        return (R)c.tsv[i];
    } else {
        if (c.tsv && (double)T == c.tsv[c.ntseries]) return (R)c.tsv[i];
        return dtseries<R>(&c.ts[i], T);
    }
}

// SumTSeries_derivative, src/utility.c:67-79
template <class R>
__device__ __noinline__ R dtseries_deriv(const cipher_tseries *t, R T) {
    R r = R(0);
    for (int i = 0; i < t->n; i++) {
        const int e = t->exp[i];
        if (e > 0) r += (R)e * (R)t->coeff[i] * dfastpow(T, (unsigned)(e - 1));
        else if (e < 0) r += (R)e * (R)t->coeff[i] / dfastpow(T, (unsigned)(-e + 1));
    }
    if (fabs(t->logcoeff) > 0.0) r += (R)t->logcoeff * (R(1) + dlog(T));
    return r;
}

template <class R>
__device__ __forceinline__ R dts_deriv(const DevCfg &c, int i, R T) {
    return dtseries_deriv<R>(&c.ts[i], T);
}

// active-set count/id read from a double in the state, clamped so that a
// corrupt state can never index outside the tables or loop without bound
__device__ __forceinline__ int clamp_count(double v, int P) {
    const double r = round(v);
    return r < 0.0 ? 0 : (r > (double)P ? P : (int)r);
}
__device__ __forceinline__ int clamp_id(double v, int npf) {
    const double r = round(v);
    return r < 0.0 ? 0 : (r > (double)(npf - 1) ? npf - 1 : (int)r);
}

template <int N> struct Cnt { static constexpr int v = N > 0 ? N : 1; };

// Invertmatrix for n <= 4 (src/utility.c:232-405): adjugate of a matrix taken
// to be symmetric, with the reference's cofactor term order.
template <int ND, class R>
__host__ __device__ inline void dinvert(R *d, const R *s) {
    R det;
    if constexpr (ND == 1) {
        d[0] = R(1) / s[0];
    } else if constexpr (ND == 2) {
        d[0] = +s[3];
        d[2] = -s[2];
        d[3] = +s[0];
        det = s[0] * d[0] + s[1] * d[2];
        det = R(1) / det;
        d[0] *= det; d[2] *= det; d[3] *= det;
        d[1] = d[2];
    } else if constexpr (ND == 3) {
        d[0] = +s[4] * s[8] - s[5] * s[7];
        d[3] = -s[3] * s[8] + s[5] * s[6];
        d[4] = +s[0] * s[8] - s[2] * s[6];
        d[6] = +s[3] * s[7] - s[4] * s[6];
        d[7] = -s[0] * s[7] + s[1] * s[6];
        d[8] = +s[0] * s[4] - s[1] * s[3];
        det = s[0] * d[0] + s[1] * d[3] + s[2] * d[6];
        det = R(1) / det;
        d[0] *= det; d[3] *= det; d[4] *= det; d[6] *= det; d[7] *= det; d[8] *= det;
        d[1] = d[3]; d[2] = d[6]; d[5] = d[7];
    } else if constexpr (ND == 4) {
        d[0] = +s[5] * s[10] * s[15] - s[5] * s[11] * s[14] - s[9] * s[6] * s[15]
               + s[9] * s[7] * s[14] + s[13] * s[6] * s[11] - s[13] * s[7] * s[10];
        d[4] = -s[4] * s[10] * s[15] + s[4] * s[11] * s[14] + s[8] * s[6] * s[15]
               - s[8] * s[7] * s[14] - s[12] * s[6] * s[11] + s[12] * s[7] * s[10];
        d[5] = +s[0] * s[10] * s[15] - s[0] * s[11] * s[14] - s[8] * s[2] * s[15]
               + s[8] * s[3] * s[14] + s[12] * s[2] * s[11] - s[12] * s[3] * s[10];
        d[8] = +s[4] * s[9] * s[15] - s[4] * s[11] * s[13] - s[8] * s[5] * s[15]
               + s[8] * s[7] * s[13] + s[12] * s[5] * s[11] - s[12] * s[7] * s[9];
        d[9] = -s[0] * s[9] * s[15] + s[0] * s[11] * s[13] + s[8] * s[1] * s[15]
               - s[8] * s[3] * s[13] - s[12] * s[1] * s[11] + s[12] * s[3] * s[9];
        d[10] = +s[0] * s[5] * s[15] - s[0] * s[7] * s[13] - s[4] * s[1] * s[15]
                + s[4] * s[3] * s[13] + s[12] * s[1] * s[7] - s[12] * s[3] * s[5];
        d[12] = -s[4] * s[9] * s[14] + s[4] * s[10] * s[13] + s[8] * s[5] * s[14]
                - s[8] * s[6] * s[13] - s[12] * s[5] * s[10] + s[12] * s[6] * s[9];
        d[13] = +s[0] * s[9] * s[14] - s[0] * s[10] * s[13] - s[8] * s[1] * s[14]
                + s[8] * s[2] * s[13] + s[12] * s[1] * s[10] - s[12] * s[2] * s[9];
        d[14] = -s[0] * s[5] * s[14] + s[0] * s[6] * s[13] + s[4] * s[1] * s[14]
                - s[4] * s[2] * s[13] - s[12] * s[1] * s[6] + s[12] * s[2] * s[5];
        d[15] = +s[0] * s[5] * s[10] - s[0] * s[6] * s[9] - s[4] * s[1] * s[10]
                + s[4] * s[2] * s[9] + s[8] * s[1] * s[6] - s[8] * s[2] * s[5];
        det = +s[0] * d[0] + s[1] * d[4] + s[2] * d[8] + s[3] * d[12];
        det = R(1) / det;
        d[0] *= det; d[4] *= det; d[5] *= det; d[8] *= det; d[9] *= det;
        d[10] *= det; d[12] *= det; d[13] *= det; d[14] *= det; d[15] *= det;
        d[1] = d[4]; d[2] = d[8]; d[3] = d[12]; d[6] = d[9]; d[7] = d[13]; d[11] = d[14];
    }
}

// MatVecMult_CIPHER, src/utility.c:456-465
template <int N, class R>
__device__ __forceinline__ void dmatvec(R *dst, const R *a, const R *b) {
#pragma unroll
    for (int k = 0; k < N; k++) {
        R s = R(0);
#pragma unroll
        for (int j = 0; j < N; j++) s += a[k * N + j] * b[j];
        dst[k] = s;
    }
}

// element ix of a compile-time sized register array without dynamic indexing
// (a sum of masked terms; the compiler turns a plain select chain back into
// an indexed private-memory load)
template <int N, class R>
__device__ __forceinline__ R dsel(const R *a, int ix) {
    R v = R(0);
#pragma unroll
    for (int c = 0; c < N; c++) {
        const R m = (c == ix) ? R(1) : R(0);
        v += m * a[c];
    }
    return v;
}
template <int N, class R>
__device__ __forceinline__ void daddat(R *a, int ix, R v) {
#pragma unroll
    for (int c = 0; c < N; c++) {
        const R m = (c == ix) ? R(1) : R(0);
        a[c] += m * v;
    }
}

// ---------------------------------------------------------------------------
// Sitefrac, src/material.c:732-822.  spim: nsites*(NC-1) implicit site
// potentials; y: nsites*NC site fractions.  The closed form is evaluated in
// FP64 in both precisions (exp, normalisation and the 1-sum remainder are
// the end-member sensitive part).
template <int NC, class R, bool ISO = false>
__device__ inline void d_sitefrac(const DevCfg &c, const cipher_material &m, R *y, const R *spim, R T) {
    constexpr int ND = NC - 1;
    if (m.model == CIPHER_CHEM_CALPHADDIS || m.model == CIPHER_CHEM_CALPHAD2SL) {
        const int ns = m.model == CIPHER_CHEM_CALPHAD2SL ? 2 : 1;
        const double Td = (double)T;
        const double RT = CRG * Td;
#pragma unroll
        for (int s = 0; s < 2; s++) {
            if (s >= ns) break;
            double yy[Cnt<NC>::v];
            double sum = 1.0;
            const double st = m.stoichiometry[s];
#pragma unroll
            for (int k = 0; k < ND; k++) {
                const double sc = m.molarvolume * (double)spim[s * ND + k];
                yy[k] = (ns == 2) ? exp(sc / st / RT) : exp(sc / RT);
                sum += yy[k];
            }
            yy[ND] = 1.0;
#pragma unroll
            for (int k = 0; k < ND; k++) {
                yy[k] /= sum;
                yy[ND] -= yy[k];
            }
#pragma unroll
            for (int k = 0; k < NC; k++) y[s * NC + k] = (R)yy[k];
        }
    } else if (m.model == CIPHER_CHEM_QUADRATIC) {
        R delta[Cnt<ND>::v], ceq[NC], un[NC], bi[NC];
        R mt[Cnt<ND * ND>::v], mi[Cnt<ND * ND>::v];
#pragma unroll
        for (int k = 0; k < NC; k++) {
            ceq[k] = dts<R, ISO>(c, m.quad_ceq + k, T);
            un[k] = dts<R, ISO>(c, m.unary + k, T);
            bi[k] = dts<R, ISO>(c, m.quad_binary + k, T);
        }
#pragma unroll
        for (int k = 0; k < ND * ND; k++) { mt[k] = R(0); mi[k] = R(0); }
#pragma unroll
        for (int j = 0; j < ND; j++) {
            const R sc = (R)m.molarvolume * spim[j];
            delta[j] = R(0.5) * sc + bi[ND] - R(0.5) * (un[j] - un[ND]) + (bi[j] * ceq[j] - bi[ND] * ceq[ND]);
            mt[j * ND + j] += bi[j];
#pragma unroll
            for (int i = 0; i < ND; i++) mt[j * ND + i] += bi[ND];
        }
        dinvert<ND, R>(mi, mt);
        dmatvec<ND, R>(y, mi, delta);
        y[ND] = R(1);
#pragma unroll
        for (int k = 0; k < ND; k++) y[ND] -= y[k];
    } else {
#pragma unroll
        for (int k = 0; k < NC; k++) y[k] = R(0);
        y[ND] = R(1);
    }
}

// CompositionMobility{,Component}, src/material.c:1023-1120 (times Vm)
template <int NC, class R, bool ISO = false>
__device__ inline void d_mobility(const DevCfg &c, const cipher_material &m, R *mob, const R *y, R T) {
    if (m.model == CIPHER_CHEM_NONE) {
#pragma unroll
        for (int k = 0; k < NC; k++) mob[k] = R(0);
    } else {
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const cipher_mobility mb = c.mob[m.mobility + k];
            const uint64_t nz = rk_window(c, mb.binary);
            int r = 0;
            R mig = R(0);
#pragma unroll
            for (int j = 0; j < NC; j++) {
                mig += y[j] * dts<R, ISO>(c, mb.unary + j, T);
#pragma unroll
                for (int i = j + 1; i < NC; i++, r++) {
                    if (!((nz >> r) & 1)) continue;
                    const cipher_rk bin = c.rk[mb.binary + r];
                    for (int o = 0; o < bin.n; o++)
                        mig += y[j] * y[i] * dts<R, ISO>(c, bin.ts + o, T) * dfastpow(y[j] - y[i], (unsigned)o);
                }
            }
            mob[k] = y[k] * (R)mb.m0 / (R)CRG / T * dexp<R>(mig / (R)CRG / T);
        }
    }
#pragma unroll
    for (int k = 0; k < NC; k++) mob[k] *= (R)m.molarvolume;
}

// Chemenergy_{calphad2sl,calphaddis,quad,none}, src/material.c:356-475
template <int NC, class R, bool ISO = false>
__device__ inline R d_chemenergy(const DevCfg &c, const cipher_material &m, const R *y, R T) {
    if (m.model == CIPHER_CHEM_CALPHAD2SL) {
        const R *yp = y, *yq = y + NC;
        const R p = (R)m.stoichiometry[0], q = (R)m.stoichiometry[1];
        R e = R(0);
#pragma unroll
        for (int k = 0; k < NC; k++) e += (R)CRG * T * (p * yp[k] * dlog(yp[k]) + q * yq[k] * dlog(yq[k]));
#pragma unroll
        for (int cp = 0; cp < NC; cp++)
#pragma unroll
            for (int cq = 0; cq < NC; cq++) e += dts<R, ISO>(c, m.unary + cp * NC + cq, T) * yp[cp] * yq[cq];
        {
            const uint64_t nz = rk_window(c, m.binary);
            int r = 0;
#pragma unroll
            for (int i = 0; i < NC; i++)
#pragma unroll
                for (int j = i + 1; j < NC; j++)
#pragma unroll
                    for (int cq = 0; cq < NC; cq++, r++) {
                        if (!((nz >> r) & 1)) continue;
                        const cipher_rk rv = c.rk[m.binary + r];
                        for (int o = 0; o < rv.n; o++)
                            e += dts<R, ISO>(c, rv.ts + o, T) * yp[i] * yp[j] * yq[cq] * dfastpow(yp[i] - yp[j], (unsigned)o);
                    }
        }
        {
            const uint64_t nz = rk_window(c, m.binaryq);
            int r = 0;
#pragma unroll
            for (int cp = 0; cp < NC; cp++)
#pragma unroll
                for (int i = 0; i < NC; i++)
#pragma unroll
                    for (int j = i + 1; j < NC; j++, r++) {
                        if (!((nz >> r) & 1)) continue;
                        const cipher_rk rv = c.rk[m.binaryq + r];
                        for (int o = 0; o < rv.n; o++)
                            e += dts<R, ISO>(c, rv.ts + o, T) * yp[cp] * yq[i] * yq[j] * dfastpow(yq[i] - yq[j], (unsigned)o);
                    }
        }
        {
            const uint64_t nz = rk_window(c, m.ternary);
            int r = 0;
#pragma unroll
            for (int i = 0; i < NC; i++)
#pragma unroll
                for (int j = i + 1; j < NC; j++)
#pragma unroll
                    for (int k = j + 1; k < NC; k++)
#pragma unroll
                        for (int cq = 0; cq < NC; cq++, r++) {
                            if (!((nz >> r) & 1)) continue;
                            const cipher_rk rv = c.rk[m.ternary + r];
                            for (int o = 0; o < rv.n; o++)
                                e += dts<R, ISO>(c, rv.ts + o, T) * yp[i] * yp[j] * yp[k] * yq[cq];
                        }
        }
        {
            const uint64_t nz = rk_window(c, m.ternaryq);
            int r = 0;
#pragma unroll
            for (int cp = 0; cp < NC; cp++)
#pragma unroll
                for (int i = 0; i < NC; i++)
#pragma unroll
                    for (int j = i + 1; j < NC; j++)
#pragma unroll
                        for (int k = j + 1; k < NC; k++, r++) {
                            if (!((nz >> r) & 1)) continue;
                            const cipher_rk rv = c.rk[m.ternaryq + r];
                            for (int o = 0; o < rv.n; o++)
                                e += dts<R, ISO>(c, rv.ts + o, T) * yq[i] * yq[j] * yq[k] * yp[cp];
                        }
        }
        return e;
    } else if (m.model == CIPHER_CHEM_CALPHADDIS) {
        const uint64_t nzb = rk_window(c, m.binary), nzt = rk_window(c, m.ternary);
        int rb = 0, rt = 0;
        R e = dts<R, ISO>(c, m.ref, T);
#pragma unroll
        for (int k = 0; k < NC; k++) {
            e += dts<R, ISO>(c, m.unary + k, T) * y[k] + (R)CRG * T * y[k] * dlog(y[k]);
#pragma unroll
            for (int j = k + 1; j < NC; j++, rb++) {
                if ((nzb >> rb) & 1) {
                    const cipher_rk bin = c.rk[m.binary + rb];
                    for (int o = 0; o < bin.n; o++)
                        e += dts<R, ISO>(c, bin.ts + o, T) * y[k] * y[j] * dfastpow(y[k] - y[j], (unsigned)o);
                }
#pragma unroll
                for (int i = j + 1; i < NC; i++, rt++) {
                    if (!((nzt >> rt) & 1)) continue;
                    const cipher_rk ter = c.rk[m.ternary + rt];
                    for (int o = 0; o < ter.n; o++)
                        e += dts<R, ISO>(c, ter.ts + o, T) * y[k] * y[j] * y[i]
                             * (dsel<NC, R>(y, c.rkindex[ter.idx + o]) + (R(1) - y[k] - y[j] - y[i]) / R(3));
                }
            }
        }
        return e;
    } else if (m.model == CIPHER_CHEM_QUADRATIC) {
        R e = dts<R, ISO>(c, m.ref, T);
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const R ceq = dts<R, ISO>(c, m.quad_ceq + k, T);
            e += dts<R, ISO>(c, m.unary + k, T) * (y[k] - ceq)
                 + dts<R, ISO>(c, m.quad_binary + k, T) * (y[k] - ceq) * (y[k] - ceq);
        }
        return e;
    }
    return dts<R, ISO>(c, m.ref, T);
}

// SitepotentialExplicit, src/material.c:495-650: out nsites*(NC-1)
template <int NC, class R, bool ISO = false>
__device__ inline void d_sitepot_explicit(const DevCfg &c, const cipher_material &m, R *sp, const R *y, R T) {
    constexpr int ND = NC - 1;
    R full[2 * NC];
#pragma unroll
    for (int k = 0; k < 2 * NC; k++) full[k] = R(0);
    if (m.model == CIPHER_CHEM_CALPHAD2SL) {
        const R *yp = y, *yq = y + NC;
        R *pp = full, *pq = full + NC;
        R a, b, cc;
#pragma unroll
        for (int cp = 0; cp < NC; cp++)
#pragma unroll
            for (int cq = 0; cq < NC; cq++) {
                a = dts<R, ISO>(c, m.unary + cp * NC + cq, T);
                pp[cp] += a * yq[cq];
                pq[cq] += a * yp[cp];
            }
        {
            const uint64_t nz = rk_window(c, m.binary);
            int r = 0;
#pragma unroll
            for (int i = 0; i < NC; i++)
#pragma unroll
                for (int j = i + 1; j < NC; j++)
#pragma unroll
                    for (int cq = 0; cq < NC; cq++, r++) {
                        if (!((nz >> r) & 1)) continue;
                        const cipher_rk rv = c.rk[m.binary + r];
                        for (int o = 0; o < rv.n; o++) {
                            a = dts<R, ISO>(c, rv.ts + o, T);
                            b = dfastpow(yp[i] - yp[j], (unsigned)o);
                            cc = (o == 0) ? R(0) : ((R)o) * dfastpow(yp[i] - yp[j], (unsigned)(o - 1));
                            pp[i] += a * yp[j] * yq[cq] * (b + cc * yp[i]);
                            pp[j] += a * yp[i] * yq[cq] * (b - cc * yp[j]);
                            pq[cq] += a * yp[i] * b * yp[j];
                        }
                    }
        }
        {
            const uint64_t nz = rk_window(c, m.binaryq);
            int r = 0;
#pragma unroll
            for (int cp = 0; cp < NC; cp++)
#pragma unroll
                for (int i = 0; i < NC; i++)
#pragma unroll
                    for (int j = i + 1; j < NC; j++, r++) {
                        if (!((nz >> r) & 1)) continue;
                        const cipher_rk rv = c.rk[m.binaryq + r];
                        for (int o = 0; o < rv.n; o++) {
                            a = dts<R, ISO>(c, rv.ts + o, T);
                            b = dfastpow(yq[i] - yq[j], (unsigned)o);
                            cc = (o == 0) ? R(0) : ((R)o) * dfastpow(yq[i] - yq[j], (unsigned)(o - 1));
                            pq[i] += a * yq[j] * yp[cp] * (b + cc * yq[i]);
                            pq[j] += a * yq[i] * yp[cp] * (b - cc * yq[j]);
                            pp[cp] += a * yq[i] * b * yq[j];
                        }
                    }
        }
        {
            const uint64_t nz = rk_window(c, m.ternary);
            int r = 0;
#pragma unroll
            for (int i = 0; i < NC; i++)
#pragma unroll
                for (int j = i + 1; j < NC; j++)
#pragma unroll
                    for (int k = j + 1; k < NC; k++)
#pragma unroll
                        for (int cq = 0; cq < NC; cq++, r++) {
                            if (!((nz >> r) & 1)) continue;
                            const cipher_rk rv = c.rk[m.ternary + r];
                            for (int o = 0; o < rv.n; o++) {
                                a = dts<R, ISO>(c, rv.ts + o, T);
                                pp[i] += a * yp[j] * yp[k] * yq[cq];
                                pp[j] += a * yp[k] * yp[i] * yq[cq];
                                pp[k] += a * yp[i] * yp[j] * yq[cq];
                                pq[cq] += a * yp[i] * yp[j] * yp[k];
                            }
                        }
        }
        {
            const uint64_t nz = rk_window(c, m.ternaryq);
            int r = 0;
#pragma unroll
            for (int cp = 0; cp < NC; cp++)
#pragma unroll
                for (int i = 0; i < NC; i++)
#pragma unroll
                    for (int j = i + 1; j < NC; j++)
#pragma unroll
                        for (int k = j + 1; k < NC; k++, r++) {
                            if (!((nz >> r) & 1)) continue;
                            const cipher_rk rv = c.rk[m.ternaryq + r];
                            for (int o = 0; o < rv.n; o++) {
                                a = dts<R, ISO>(c, rv.ts + o, T);
                                pq[i] += a * yq[j] * yq[k] * yp[cp];
                                pq[j] += a * yq[k] * yq[i] * yp[cp];
                                pq[k] += a * yq[i] * yq[j] * yp[cp];
                                pp[cp] += a * yq[i] * yq[j] * yq[k];
                            }
                        }
        }
    } else if (m.model == CIPHER_CHEM_CALPHADDIS) {
        R a, b, cc;
#pragma unroll
        for (int i = 0; i < NC; i++) full[i] = dts<R, ISO>(c, m.unary + i, T);
        {
            const uint64_t nz = rk_window(c, m.binary);
            int r = 0;
#pragma unroll
            for (int i = 0; i < NC; i++)
#pragma unroll
                for (int j = i + 1; j < NC; j++, r++) {
                    if (!((nz >> r) & 1)) continue;
                    const cipher_rk rv = c.rk[m.binary + r];
                    for (int o = 0; o < rv.n; o++) {
                        a = dts<R, ISO>(c, rv.ts + o, T);
                        b = dfastpow(y[i] - y[j], (unsigned)o);
                        cc = (o == 0) ? R(0) : ((R)o) * dfastpow(y[i] - y[j], (unsigned)(o - 1));
                        full[i] += a * y[j] * (b + cc * y[i]);
                        full[j] += a * y[i] * (b - cc * y[j]);
                    }
                }
        }
        const uint64_t nzt = rk_window(c, m.ternary);
        int rt = 0;
#pragma unroll
        for (int i = 0; i < NC; i++)
#pragma unroll
            for (int j = i + 1; j < NC; j++)
#pragma unroll
                for (int k = j + 1; k < NC; k++, rt++) {
                    if (!((nzt >> rt) & 1)) continue;
                    const cipher_rk rv = c.rk[m.ternary + rt];
                    for (int o = 0; o < rv.n; o++) {
                        const int ix = c.rkindex[rv.idx + o];
                        a = dts<R, ISO>(c, rv.ts + o, T);
                        b = dsel<NC, R>(y, ix) + (R(1) - y[k] - y[j] - y[i]) / R(3);
                        full[i] += a * y[j] * y[k] * (b - y[i] / R(3));
                        full[j] += a * y[k] * y[i] * (b - y[j] / R(3));
                        full[k] += a * y[i] * y[j] * (b - y[k] / R(3));
                        daddat<NC, R>(full, ix, a * y[i] * y[j] * y[k]);
                    }
                }
    }
    const int ns = m.nsites;
#pragma unroll
    for (int s = 0; s < 2; s++) {
        if (s >= ns) break;
#pragma unroll
        for (int k = 0; k < ND; k++) sp[s * ND + k] = (full[s * NC + k] - full[s * NC + ND]) / (R)m.molarvolume;
    }
}

// SitepotentialImplicit, src/material.c:655-711
template <int NC, class R, bool ISO = false>
__device__ inline void d_sitepot_implicit(const DevCfg &c, const cipher_material &m, R *sp, const R *y, R T) {
    constexpr int ND = NC - 1;
    R full[2 * NC];
#pragma unroll
    for (int k = 0; k < 2 * NC; k++) full[k] = R(0);
    if (m.model == CIPHER_CHEM_CALPHAD2SL) {
#pragma unroll
        for (int k = 0; k < NC; k++) {
            full[k] = (R)CRG * T * (R)m.stoichiometry[0] * dlog(y[k]);
            full[NC + k] = (R)CRG * T * (R)m.stoichiometry[1] * dlog(y[NC + k]);
        }
    } else if (m.model == CIPHER_CHEM_CALPHADDIS) {
#pragma unroll
        for (int k = 0; k < NC; k++) full[k] = (R)CRG * T * dlog(y[k]);
    } else if (m.model == CIPHER_CHEM_QUADRATIC) {
#pragma unroll
        for (int k = 0; k < NC; k++)
            full[k] = dts<R, ISO>(c, m.unary + k, T)
                      + R(2) * dts<R, ISO>(c, m.quad_binary + k, T) * (y[k] - dts<R, ISO>(c, m.quad_ceq + k, T));
    }
#pragma unroll
    for (int s = 0; s < 2; s++) {
        if (s >= m.nsites) break;
#pragma unroll
        for (int k = 0; k < ND; k++) sp[s * ND + k] = (full[s * NC + k] - full[s * NC + ND]) / (R)m.molarvolume;
    }
}

// SitefracTangent, src/material.c:827-1018: tpot (nsites*ND*ND) and tT
// (nsites*ND), both times Vm.  The ex tangent is -tpot for the CALPHAD
// models and 0 for quadratic/none (returned through *ex_sign).
template <int NC, class R, bool ISO = false>
__device__ inline void d_sitefrac_tangent(const DevCfg &c, const cipher_material &m, R *tpot, R *tT, R &ex_sign,
                                          const R *y, const R *spim, R T) {
    constexpr int ND = NC - 1, NN = ND * ND;
    const R RT = (R)CRG * T;
    if (m.model == CIPHER_CHEM_CALPHAD2SL || m.model == CIPHER_CHEM_CALPHADDIS) {
        const int ns = m.model == CIPHER_CHEM_CALPHAD2SL ? 2 : 1;
        ex_sign = R(-1);
#pragma unroll
        for (int s = 0; s < 2; s++) {
            if (s >= ns) break;
            const R *ys = y + s * NC;
            const R *ps = spim + s * ND;
            const R st = (R)m.stoichiometry[s];
            R t[Cnt<NN>::v];
#pragma unroll
            for (int k = 0; k < NN; k++) t[k] = R(0);
#pragma unroll
            for (int j = 0; j < ND; j++) {
                if (ns == 2) t[j * ND + j] += ys[j] / st / RT;
                else t[j * ND + j] += ys[j] / RT;
#pragma unroll
                for (int i = 0; i < ND; i++) {
                    if (ns == 2) t[j * ND + i] -= ys[j] * ys[i] / st / RT;
                    else t[j * ND + i] -= ys[j] * ys[i] / RT;
                }
            }
#pragma unroll
            for (int k = 0; k < NN; k++) tpot[s * NN + k] = t[k] * (R)m.molarvolume;
#pragma unroll
            for (int j = 0; j < ND; j++) {
                R v = R(0);
#pragma unroll
                for (int i = 0; i < ND; i++) v -= t[j * ND + i] * ps[i] / T;
                tT[s * ND + j] = v * (R)m.molarvolume;
            }
        }
    } else if (m.model == CIPHER_CHEM_QUADRATIC) {
        ex_sign = R(0);
        R delta[Cnt<ND>::v], ceqT[NC], unT[NC], biT[NC], ceq[NC], bi[NC];
        R mt[Cnt<NN>::v], mi[Cnt<NN>::v];
#pragma unroll
        for (int k = 0; k < NC; k++) {
            ceqT[k] = dts_deriv<R>(c, m.quad_ceq + k, T);
            unT[k] = dts_deriv<R>(c, m.unary + k, T);
            biT[k] = dts_deriv<R>(c, m.quad_binary + k, T);
            ceq[k] = dts<R, ISO>(c, m.quad_ceq + k, T);
            bi[k] = dts<R, ISO>(c, m.quad_binary + k, T);
        }
#pragma unroll
        for (int k = 0; k < NN; k++) { mt[k] = R(0); mi[k] = R(0); }
#pragma unroll
        for (int j = 0; j < ND; j++) {
            mt[j * ND + j] += bi[j];
#pragma unroll
            for (int i = 0; i < ND; i++) mt[j * ND + i] += bi[ND];
        }
        dinvert<ND, R>(mi, mt);
#pragma unroll
        for (int k = 0; k < NN; k++) tpot[k] = mi[k] * R(0.5) * (R)m.molarvolume;
#pragma unroll
        for (int j = 0; j < ND; j++)
            delta[j] = biT[ND] - R(0.5) * (unT[j] - unT[ND])
                       + (biT[j] * ceq[j] + bi[j] * ceqT[j] - biT[ND] * ceq[ND] - bi[ND] * ceqT[ND]);
        R tt[Cnt<ND>::v];
        dmatvec<ND, R>(tt, mi, delta);
#pragma unroll
        for (int j = 0; j < ND; j++) tT[j] = tt[j] * (R)m.molarvolume;
    } else {
        ex_sign = R(0);
#pragma unroll
        for (int k = 0; k < NN; k++) tpot[k] = R(0);
#pragma unroll
        for (int j = 0; j < ND; j++) tT[j] = R(0);
    }
}
