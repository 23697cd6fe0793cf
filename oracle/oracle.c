/*
 * oracle.c - CPU restatement of CIPHER's phase-field right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (see oracle.h).
 * FP64 throughout, in the reference's operation order, so that on the same
 * inputs it reproduces the reference's arithmetic.  Citations are
 * micmog/CIPHER file:line.
 *
 * Deliberate differences from the reference (each is undefined behaviour or a
 * PETSc-internal choice there):
 *  - the no-chemistry branch of the precompute reads the cell's own
 *    temperature (src/cipher.c:230 dereferences an unset pointer);
 *  - neighbours for the active-set union are the cells sharing an interior
 *    face of the flattened mesh (src/cipher.c:936-950 walks DMPlex cones).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_D 3.14159265358979323846264338327950288419716939937510582
#define RG CIPHER_R_GAS_CONST

/* ------------------------------------------------------------------ */
/* layout, src/init.c:1330-1343                                        */
/* ------------------------------------------------------------------ */
typedef struct {
    int maxsites, sp, sf, as, pf, dp, ex, tm;
    int aso, pfo, dpo, exo, tmo, ndof;
} olayout;

static olayout orc_layout(const cipher_config *cfg)
{
    olayout L;
    int mpn = cfg->max_phase_neighbours > 0 ? cfg->max_phase_neighbours : CIPHER_MAXAP;
    int ndp = cfg->ncp - 1;
    L.maxsites = 0;
    for (int m = 0; m < cfg->nmat; m++)
        if (cfg->material[m].nsites > L.maxsites) L.maxsites = cfg->material[m].nsites;
    L.sf = L.maxsites * cfg->ncp;
    L.sp = L.maxsites * ndp;
    L.aso = 0;
    L.as = (mpn < cfg->npf ? mpn : cfg->npf) + 1;
    L.pfo = L.aso + L.as;
    L.pf = (mpn < cfg->npf ? mpn : cfg->npf);
    L.dpo = L.pfo + L.pf;
    L.dp = ndp;
    L.exo = L.dpo + L.dp;
    L.ex = L.pf * L.sp;
    L.tmo = L.exo + L.ex;
    L.tm = 1;
    L.ndof = L.tmo + L.tm;
    return L;
}

/* ------------------------------------------------------------------ */
/* utility.c                                                           */
/* ------------------------------------------------------------------ */

/* F2IFUNC, src/utility.c:24-29 */
static void f2i(uint16_t *ia, const double *fa)
{
    ia[0] = (uint16_t)round(fa[0]);
    for (int k = 1; k <= ia[0]; ++k) ia[k] = (uint16_t)round(fa[k]);
}

/* I2FFUNC, src/utility.c:30-35 */
static void i2f(double *fa, const uint16_t *ia)
{
    fa[0] = (double)ia[0];
    for (int k = 1; k <= ia[0]; ++k) fa[k] = (double)ia[k];
}

/* FastPow, src/utility.c:112-124: binary exponentiation */
double orc_fastpow(double base, unsigned expn)
{
    double r = 1.0;
    while (expn) {
        if (expn & 1u) r *= base;
        expn >>= 1;
        base *= base;
    }
    return r;
}

/* SumTSeries, src/utility.c:48-62 */
double orc_sum_tseries(double T, const cipher_tseries *ts)
{
    double r = 0.0;
    for (int i = 0; i < ts->n; i++) {
        int e = ts->exp[i];
        if (e > 0)      r += ts->coeff[i] * orc_fastpow(T, (unsigned)e);
        else if (e < 0) r += ts->coeff[i] / orc_fastpow(T, (unsigned)(-e));
        else            r += ts->coeff[i];
    }
    if (fabs(ts->logcoeff) > 0.0) r += ts->logcoeff * T * log(T);
    return r;
}

/* SumTSeries_derivative, src/utility.c:67-79 */
double orc_sum_tseries_derivative(double T, const cipher_tseries *ts)
{
    double r = 0.0;
    for (int i = 0; i < ts->n; i++) {
        int e = ts->exp[i];
        if (e > 0)      r += e * ts->coeff[i] * orc_fastpow(T, (unsigned)(e - 1));
        else if (e < 0) r += e * ts->coeff[i] / orc_fastpow(T, (unsigned)(-e + 1));
    }
    if (fabs(ts->logcoeff) > 0.0) r += ts->logcoeff * (1.0 + log(T));
    return r;
}

/* SetIntersection, src/utility.c:168-190: two-pointer merge of sorted lists
 * (element 0 = count) with injection maps into A and B. */
void orc_set_intersection(uint16_t *outc, uint16_t *cina, uint16_t *cinb,
                          const uint16_t *ina, const uint16_t *inb)
{
    int i = 0, j = 0, k = 0;
    while (i < ina[0] && j < inb[0]) {
        uint16_t a = ina[i + 1], b = inb[j + 1];
        if (a == b) {
            cina[k] = (uint16_t)i;
            cinb[k] = (uint16_t)j;
            outc[++k] = a;
            i++; j++;
        } else if (a < b) {
            i++;
        } else {
            j++;
        }
    }
    outc[0] = (uint16_t)k;
}

/* SetUnion, src/utility.c:195-230: sorted merge with injection maps. */
void orc_set_union(uint16_t *outc, uint16_t *ainc, uint16_t *binc,
                   const uint16_t *ina, const uint16_t *inb)
{
    int i = 0, j = 0, k = 0;
    while (i < ina[0] && j < inb[0]) {
        uint16_t a = ina[i + 1], b = inb[j + 1];
        if (a == b) {
            ainc[i++] = (uint16_t)k; binc[j++] = (uint16_t)k; outc[++k] = a;
        } else if (a < b) {
            ainc[i++] = (uint16_t)k; outc[++k] = a;
        } else {
            binc[j++] = (uint16_t)k; outc[++k] = b;
        }
    }
    while (i < ina[0]) { ainc[i] = (uint16_t)k; outc[++k] = ina[i + 1]; i++; }
    while (j < inb[0]) { binc[j] = (uint16_t)k; outc[++k] = inb[j + 1]; j++; }
    outc[0] = (uint16_t)k;
}

/* Invertmatrix, src/utility.c:232-441.  n<=4: adjugate of a matrix taken to
 * be symmetric (the upper part is copied from the lower, :260, :290-292,
 * :399-404); n>=5: Gauss-Jordan without pivoting, overwriting src.  The
 * cofactor expressions keep the reference's term order. */
void orc_invert_matrix(double *d, double *s, int n)
{
    double det;
    if (n == 1) {
        d[0] = 1.0 / s[0];
    } else if (n == 2) {
        d[0] = +s[3];
        d[2] = -s[2];
        d[3] = +s[0];
        det = s[0] * d[0] + s[1] * d[2];
        det = 1.0 / det;
        d[0] *= det; d[2] *= det; d[3] *= det;
        d[1] = d[2];
    } else if (n == 3) {
        d[0] = +s[4] * s[8] - s[5] * s[7];
        d[3] = -s[3] * s[8] + s[5] * s[6];
        d[4] = +s[0] * s[8] - s[2] * s[6];
        d[6] = +s[3] * s[7] - s[4] * s[6];
        d[7] = -s[0] * s[7] + s[1] * s[6];
        d[8] = +s[0] * s[4] - s[1] * s[3];
        det = s[0] * d[0] + s[1] * d[3] + s[2] * d[6];
        det = 1.0 / det;
        d[0] *= det; d[3] *= det; d[4] *= det; d[6] *= det; d[7] *= det; d[8] *= det;
        d[1] = d[3]; d[2] = d[6]; d[5] = d[7];
    } else if (n == 4) {
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
        det = 1.0 / det;
        d[0] *= det; d[4] *= det; d[5] *= det; d[8] *= det; d[9] *= det;
        d[10] *= det; d[12] *= det; d[13] *= det; d[14] *= det; d[15] *= det;
        d[1] = d[4]; d[2] = d[8]; d[3] = d[12]; d[6] = d[9]; d[7] = d[13]; d[11] = d[14];
    } else {
        memset(d, 0, (size_t)n * n * sizeof(double));
        for (int i = 0; i < n; i++) d[i * n + i] = 1.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++)
                if (i != j) {
                    double ratio = s[j * n + i] / s[i * n + i];
                    for (int k = 0; k < n; k++) s[j * n + k] -= ratio * s[i * n + k];
                    for (int k = 0; k < n; k++) d[j * n + k] -= ratio * d[i * n + k];
                }
        for (int i = 0; i < n; i++) {
            double a = s[i * n + i];
            for (int j = 0; j < n; j++) d[i * n + j] /= a;
        }
    }
}

/* MatVecMult_CIPHER, src/utility.c:456-465 */
static void matvec(double *dst, const double *a, const double *b, int n)
{
    for (int k = 0; k < n; k++) {
        dst[k] = 0.0;
        for (int j = 0; j < n; j++) dst[k] += a[k * n + j] * b[j];
    }
}

/* Interpolant_{linear,quad,cubic}, src/utility.c:470-535 */
void orc_interpolant(int interp, double *h, const double *phi, int n)
{
    if (interp == CIPHER_INTERP_LINEAR) {
        memcpy(h, phi, (size_t)n * sizeof(double));
        return;
    }
    if (n == 1) { h[0] = 1.0; return; }
    double sum = 0.0;
    for (int g = 0; g < n; g++) {
        if (interp == CIPHER_INTERP_QUADRATIC) h[g] = phi[g] * phi[g];
        else h[g] = phi[g] * phi[g] * (3.0 - 2.0 * phi[g]);
        sum += h[g];
    }
    for (int g = 0; g < n; g++) h[g] /= sum;
}

/* InterpolantDerivative_{linear,quad,cubic}, src/utility.c:478-557:
 * in <- dh/dphi^T in (linear leaves in unchanged). */
void orc_interpolant_derivative(int interp, double *in, const double *phi, int n)
{
    if (interp == CIPHER_INTERP_LINEAR) return;
    if (n == 1) { in[0] = 0.0; return; }
    double out[64];
    double avg = 0.0, sum = 0.0;
    for (int g = 0; g < n; g++) {
        double hh = (interp == CIPHER_INTERP_QUADRATIC) ? phi[g] * phi[g]
                                                         : phi[g] * phi[g] * (3.0 - 2.0 * phi[g]);
        sum += hh;
        avg += hh * in[g];
    }
    for (int g = 0; g < n; g++) {
        if (interp == CIPHER_INTERP_QUADRATIC)
            out[g] = 2.0 * phi[g] * (in[g] - avg / sum) / sum;
        else
            out[g] = 6.0 * phi[g] * (1.0 - phi[g]) * (in[g] - avg / sum) / sum;
    }
    memcpy(in, out, (size_t)n * sizeof(double));
}

static int cmp_desc(const void *x, const void *y)
{
    double a = *(const double *)x, b = *(const double *)y;
    return (b > a) ? 1 : ((b < a) ? -1 : 0);
}

/* SimplexProjection, src/utility.c:587-603: Euclidean projection onto the
 * Gibbs simplex by sorting (Held-Wolfe-Crowder threshold). */
void orc_simplex_projection(double *out, const double *in, int size)
{
    double s[64], tmax = 0.0, tsum = 0.0;
    int found = 0;
    memcpy(s, in, (size_t)size * sizeof(double));
    qsort(s, (size_t)size, sizeof(double), cmp_desc);
    for (int i = 0; i < size - 1; i++) {
        tsum += s[i];
        tmax = (tsum - 1.0) / ((double)(i + 1));
        if (tmax >= s[i + 1]) { found = 1; break; }
    }
    if (!found) tmax = (tsum + s[size - 1] - 1.0) / ((double)size);
    for (int i = 0; i < size; i++) out[i] = in[i] - tmax > 0.0 ? in[i] - tmax : 0.0;
}

/* ------------------------------------------------------------------ */
/* material.c                                                          */
/* ------------------------------------------------------------------ */
#define TS(i) (&cfg->tseries[(i)])
#define RKP(i) (&cfg->rk[(i)])

static double rk_term(const cipher_config *cfg, const cipher_rk *rk, int o, double T)
{
    return orc_sum_tseries(T, TS(rk->ts + o));
}

/* Chemenergy_calphad2sl, src/material.c:356-425 */
static double chemenergy_2sl(const cipher_config *cfg, const cipher_material *m, const double *y, double T)
{
    int nc = cfg->ncp;
    const double *yp = y, *yq = y + nc;
    double p = m->stoichiometry[0], q = m->stoichiometry[1];
    double e = 0.0;
    for (int c = 0; c < nc; c++)
        e += RG * T * (p * yp[c] * log(yp[c]) + q * yq[c] * log(yq[c]));
    for (int cp = 0; cp < nc; cp++)
        for (int cq = 0; cq < nc; cq++)
            e += orc_sum_tseries(T, TS(m->unary + cp * nc + cq)) * yp[cp] * yq[cq];
    const cipher_rk *rk = RKP(m->binary);
    for (int i = 0; i < nc; i++)
        for (int j = i + 1; j < nc; j++)
            for (int cq = 0; cq < nc; cq++, rk++)
                for (int o = 0; o < rk->n; o++)
                    e += rk_term(cfg, rk, o, T) * yp[i] * yp[j] * yq[cq] * orc_fastpow(yp[i] - yp[j], (unsigned)o);
    rk = RKP(m->binaryq);
    for (int cp = 0; cp < nc; cp++)
        for (int i = 0; i < nc; i++)
            for (int j = i + 1; j < nc; j++, rk++)
                for (int o = 0; o < rk->n; o++)
                    e += rk_term(cfg, rk, o, T) * yp[cp] * yq[i] * yq[j] * orc_fastpow(yq[i] - yq[j], (unsigned)o);
    rk = RKP(m->ternary);
    for (int i = 0; i < nc; i++)
        for (int j = i + 1; j < nc; j++)
            for (int k = j + 1; k < nc; k++)
                for (int cq = 0; cq < nc; cq++, rk++)
                    for (int o = 0; o < rk->n; o++)
                        e += rk_term(cfg, rk, o, T) * yp[i] * yp[j] * yp[k] * yq[cq];
    rk = RKP(m->ternaryq);
    for (int cp = 0; cp < nc; cp++)
        for (int i = 0; i < nc; i++)
            for (int j = i + 1; j < nc; j++)
                for (int k = j + 1; k < nc; k++, rk++)
                    for (int o = 0; o < rk->n; o++)
                        e += rk_term(cfg, rk, o, T) * yq[i] * yq[j] * yq[k] * yp[cp];
    return e;
}

/* Chemenergy_calphaddis, src/material.c:430-451 */
static double chemenergy_dis(const cipher_config *cfg, const cipher_material *m, const double *y, double T)
{
    int nc = cfg->ncp;
    const cipher_rk *bin = RKP(m->binary), *ter = RKP(m->ternary);
    double e = orc_sum_tseries(T, TS(m->ref));
    for (int k = 0; k < nc; k++) {
        e += orc_sum_tseries(T, TS(m->unary + k)) * y[k] + RG * T * y[k] * log(y[k]);
        for (int j = k + 1; j < nc; j++, bin++) {
            for (int o = 0; o < bin->n; o++)
                e += rk_term(cfg, bin, o, T) * y[k] * y[j] * orc_fastpow(y[k] - y[j], (unsigned)o);
            for (int i = j + 1; i < nc; i++, ter++)
                for (int o = 0; o < ter->n; o++)
                    e += rk_term(cfg, ter, o, T) * y[k] * y[j] * y[i]
                         * (y[cfg->rkindex[ter->idx + o]] + (1.0 - y[k] - y[j] - y[i]) / 3.0);
        }
    }
    return e;
}

/* Chemenergy_quad, src/material.c:456-466 */
static double chemenergy_quad(const cipher_config *cfg, const cipher_material *m, const double *y, double T)
{
    int nc = cfg->ncp;
    double e = orc_sum_tseries(T, TS(m->ref));
    for (int c = 0; c < nc; c++)
        e += orc_sum_tseries(T, TS(m->unary + c)) * (y[c] - orc_sum_tseries(T, TS(m->quad_ceq + c)))
             + orc_sum_tseries(T, TS(m->quad_binary + c)) * (y[c] - orc_sum_tseries(T, TS(m->quad_ceq + c)))
                   * (y[c] - orc_sum_tseries(T, TS(m->quad_ceq + c)));
    return e;
}

double orc_chemenergy_phase(const cipher_config *cfg, int mat, const double *y, double T)
{
    const cipher_material *m = &cfg->material[mat];
    switch (m->model) {
    case CIPHER_CHEM_CALPHAD2SL: return chemenergy_2sl(cfg, m, y, T);
    case CIPHER_CHEM_CALPHADDIS: return chemenergy_dis(cfg, m, y, T);
    case CIPHER_CHEM_QUADRATIC: return chemenergy_quad(cfg, m, y, T);
    default: return orc_sum_tseries(T, TS(m->ref)); /* Chemenergy_none, :471-475 */
    }
}

/* SitepotentialExplicit_calphad2sl, src/material.c:495-576 (full ncp per site) */
static void spex_2sl(const cipher_config *cfg, const cipher_material *m, double *sp, const double *y, double T)
{
    int nc = cfg->ncp;
    const double *yp = y, *yq = y + nc;
    double *pp = sp, *pq = sp + nc;
    double a, b, c;
    for (int k = 0; k < nc; k++) { pp[k] = 0.0; pq[k] = 0.0; }
    for (int cp = 0; cp < nc; cp++)
        for (int cq = 0; cq < nc; cq++) {
            a = orc_sum_tseries(T, TS(m->unary + cp * nc + cq));
            pp[cp] += a * yq[cq];
            pq[cq] += a * yp[cp];
        }
    const cipher_rk *rk = RKP(m->binary);
    for (int i = 0; i < nc; i++)
        for (int j = i + 1; j < nc; j++)
            for (int cq = 0; cq < nc; cq++, rk++)
                for (int o = 0; o < rk->n; o++) {
                    a = rk_term(cfg, rk, o, T);
                    b = orc_fastpow(yp[i] - yp[j], (unsigned)o);
                    c = (o == 0) ? 0.0 : ((double)o) * orc_fastpow(yp[i] - yp[j], (unsigned)(o - 1));
                    pp[i] += a * yp[j] * yq[cq] * (b + c * yp[i]);
                    pp[j] += a * yp[i] * yq[cq] * (b - c * yp[j]);
                    pq[cq] += a * yp[i] * b * yp[j];
                }
    rk = RKP(m->binaryq);
    for (int cp = 0; cp < nc; cp++)
        for (int i = 0; i < nc; i++)
            for (int j = i + 1; j < nc; j++, rk++)
                for (int o = 0; o < rk->n; o++) {
                    a = rk_term(cfg, rk, o, T);
                    b = orc_fastpow(yq[i] - yq[j], (unsigned)o);
                    c = (o == 0) ? 0.0 : ((double)o) * orc_fastpow(yq[i] - yq[j], (unsigned)(o - 1));
                    pq[i] += a * yq[j] * yp[cp] * (b + c * yq[i]);
                    pq[j] += a * yq[i] * yp[cp] * (b - c * yq[j]);
                    pp[cp] += a * yq[i] * b * yq[j];
                }
    rk = RKP(m->ternary);
    for (int i = 0; i < nc; i++)
        for (int j = i + 1; j < nc; j++)
            for (int k = j + 1; k < nc; k++)
                for (int cq = 0; cq < nc; cq++, rk++)
                    for (int o = 0; o < rk->n; o++) {
                        a = rk_term(cfg, rk, o, T);
                        pp[i] += a * yp[j] * yp[k] * yq[cq];
                        pp[j] += a * yp[k] * yp[i] * yq[cq];
                        pp[k] += a * yp[i] * yp[j] * yq[cq];
                        pq[cq] += a * yp[i] * yp[j] * yp[k];
                    }
    rk = RKP(m->ternaryq);
    for (int cp = 0; cp < nc; cp++)
        for (int i = 0; i < nc; i++)
            for (int j = i + 1; j < nc; j++)
                for (int k = j + 1; k < nc; k++, rk++)
                    for (int o = 0; o < rk->n; o++) {
                        a = rk_term(cfg, rk, o, T);
                        pq[i] += a * yq[j] * yq[k] * yp[cp];
                        pq[j] += a * yq[k] * yq[i] * yp[cp];
                        pq[k] += a * yq[i] * yq[j] * yp[cp];
                        pp[cp] += a * yq[i] * yq[j] * yq[k];
                    }
}

/* SitepotentialExplicit_calphaddis, src/material.c:581-619 */
static void spex_dis(const cipher_config *cfg, const cipher_material *m, double *sp, const double *y, double T)
{
    int nc = cfg->ncp;
    double a, b, c;
    for (int i = 0; i < nc; i++) sp[i] = orc_sum_tseries(T, TS(m->unary + i));
    const cipher_rk *rk = RKP(m->binary);
    for (int i = 0; i < nc; i++)
        for (int j = i + 1; j < nc; j++, rk++)
            for (int o = 0; o < rk->n; o++) {
                a = rk_term(cfg, rk, o, T);
                b = orc_fastpow(y[i] - y[j], (unsigned)o);
                c = (o == 0) ? 0.0 : ((double)o) * orc_fastpow(y[i] - y[j], (unsigned)(o - 1));
                sp[i] += a * y[j] * (b + c * y[i]);
                sp[j] += a * y[i] * (b - c * y[j]);
            }
    rk = RKP(m->ternary);
    for (int i = 0; i < nc; i++)
        for (int j = i + 1; j < nc; j++)
            for (int k = j + 1; k < nc; k++, rk++)
                for (int o = 0; o < rk->n; o++) {
                    int ix = cfg->rkindex[rk->idx + o];
                    a = rk_term(cfg, rk, o, T);
                    b = y[ix] + (1.0 - y[k] - y[j] - y[i]) / 3.0;
                    sp[i] += a * y[j] * y[k] * (b - y[i] / 3.0);
                    sp[j] += a * y[k] * y[i] * (b - y[j] / 3.0);
                    sp[k] += a * y[i] * y[j] * (b - y[k] / 3.0);
                    sp[ix] += a * y[i] * y[j] * y[k];
                }
}

/* SitepotentialExplicit, src/material.c:640-650 (quad/none: zero, :624-635) */
void orc_sitepot_explicit(const cipher_config *cfg, int mat, double *sp, const double *y, double T)
{
    const cipher_material *m = &cfg->material[mat];
    int nc = cfg->ncp, nd = nc - 1;
    double full[2 * CIPHER_MAXCP];
    memset(full, 0, sizeof(full));
    if (m->model == CIPHER_CHEM_CALPHAD2SL) spex_2sl(cfg, m, full, y, T);
    else if (m->model == CIPHER_CHEM_CALPHADDIS) spex_dis(cfg, m, full, y, T);
    for (int s = 0; s < m->nsites; s++)
        for (int c = 0; c < nd; c++)
            sp[s * nd + c] = (full[s * nc + c] - full[s * nc + nd]) / m->molarvolume;
}

/* SitepotentialImplicit, src/material.c:655-711 */
void orc_sitepot_implicit(const cipher_config *cfg, int mat, double *sp, const double *y, double T)
{
    const cipher_material *m = &cfg->material[mat];
    int nc = cfg->ncp, nd = nc - 1;
    double full[2 * CIPHER_MAXCP];
    memset(full, 0, sizeof(full));
    if (m->model == CIPHER_CHEM_CALPHAD2SL) {
        for (int c = 0; c < nc; c++) {
            full[c] = RG * T * m->stoichiometry[0] * log(y[c]);
            full[nc + c] = RG * T * m->stoichiometry[1] * log(y[nc + c]);
        }
    } else if (m->model == CIPHER_CHEM_CALPHADDIS) {
        for (int c = 0; c < nc; c++) full[c] = RG * T * log(y[c]);
    } else if (m->model == CIPHER_CHEM_QUADRATIC) {
        for (int c = 0; c < nc; c++)
            full[c] = orc_sum_tseries(T, TS(m->unary + c))
                      + 2.0 * orc_sum_tseries(T, TS(m->quad_binary + c))
                            * (y[c] - orc_sum_tseries(T, TS(m->quad_ceq + c)));
    }
    for (int s = 0; s < m->nsites; s++)
        for (int c = 0; c < nd; c++)
            sp[s * nd + c] = (full[s * nc + c] - full[s * nc + nd]) / m->molarvolume;
}

/* Sitefrac, src/material.c:732-822 (closed form; quad: linear solve) */
void orc_sitefrac(const cipher_config *cfg, int mat, double *y, const double *spim, double T)
{
    const cipher_material *m = &cfg->material[mat];
    int nc = cfg->ncp, nd = nc - 1;
    double sc[2 * CIPHER_MAXCP];
    for (int s = 0; s < m->nsites; s++)
        for (int c = 0; c < nd; c++) sc[s * nd + c] = m->molarvolume * spim[s * nd + c];
    if (m->model == CIPHER_CHEM_CALPHAD2SL) {
        double *yp = y, *yq = y + nc;
        const double *sp = sc, *sq = sc + nd;
        double RT = RG * T, sump = 1.0, sumq = 1.0;
        double p = m->stoichiometry[0], q = m->stoichiometry[1];
        for (int c = 0; c < nd; c++) {
            yp[c] = exp(sp[c] / p / RT); sump += yp[c];
            yq[c] = exp(sq[c] / q / RT); sumq += yq[c];
        }
        yp[nd] = 1.0; yq[nd] = 1.0;
        for (int c = 0; c < nd; c++) {
            yp[c] /= sump; yp[nd] -= yp[c];
            yq[c] /= sumq; yq[nd] -= yq[c];
        }
    } else if (m->model == CIPHER_CHEM_CALPHADDIS) {
        double RT = RG * T, sum = 1.0;
        for (int c = 0; c < nd; c++) { y[c] = exp(sc[c] / RT); sum += y[c]; }
        y[nd] = 1.0;
        for (int c = 0; c < nd; c++) { y[c] /= sum; y[nd] -= y[c]; }
    } else if (m->model == CIPHER_CHEM_QUADRATIC) {
        double delta[CIPHER_MAXCP], ceq[CIPHER_MAXCP], un[CIPHER_MAXCP], bi[CIPHER_MAXCP];
        double mt[CIPHER_MAXCP * CIPHER_MAXCP], mi[CIPHER_MAXCP * CIPHER_MAXCP];
        for (int c = 0; c < nc; c++) {
            ceq[c] = orc_sum_tseries(T, TS(m->quad_ceq + c));
            un[c] = orc_sum_tseries(T, TS(m->unary + c));
            bi[c] = orc_sum_tseries(T, TS(m->quad_binary + c));
        }
        memset(mt, 0, sizeof(mt)); memset(mi, 0, sizeof(mi));
        for (int j = 0; j < nd; j++) {
            delta[j] = 0.5 * sc[j] + bi[nd] - 0.5 * (un[j] - un[nd]) + (bi[j] * ceq[j] - bi[nd] * ceq[nd]);
            mt[j * nd + j] += bi[j];
            for (int i = 0; i < nd; i++) mt[j * nd + i] += bi[nd];
        }
        orc_invert_matrix(mi, mt, nd);
        matvec(y, mi, delta, nd);
        y[nd] = 1.0;
        for (int c = 0; c < nd; c++) y[nd] -= y[c];
    } else {
        for (int c = 0; c < nc; c++) y[c] = 0.0;   /* Sitefrac_none, :803-807 */
        y[nd] = 1.0;
    }
}

/* SitefracTangent, src/material.c:827-1018 */
void orc_sitefrac_tangent(const cipher_config *cfg, int mat, double *tpot, double *tex, double *tT,
                          const double *y, const double *spim, double T)
{
    const cipher_material *m = &cfg->material[mat];
    int nc = cfg->ncp, nd = nc - 1, nn = nd * nd;
    double RT = RG * T;
    if (m->model == CIPHER_CHEM_CALPHAD2SL || m->model == CIPHER_CHEM_CALPHADDIS) {
        int ns = m->model == CIPHER_CHEM_CALPHAD2SL ? 2 : 1;
        for (int s = 0; s < ns; s++) {
            double t[CIPHER_MAXCP * CIPHER_MAXCP];
            const double *ys = y + s * nc;
            const double *ps = spim + s * nd;
            double st = (ns == 2) ? m->stoichiometry[s] : 1.0;
            memset(t, 0, sizeof(t));
            for (int j = 0; j < nd; j++) {
                if (ns == 2) t[j * nd + j] += ys[j] / st / RT;
                else t[j * nd + j] += ys[j] / RT;
                for (int i = 0; i < nd; i++) {
                    if (ns == 2) t[j * nd + i] -= ys[j] * ys[i] / st / RT;
                    else t[j * nd + i] -= ys[j] * ys[i] / RT;
                }
            }
            if (tpot) for (int c = 0; c < nn; c++) tpot[s * nn + c] = t[c];
            if (tex) for (int c = 0; c < nn; c++) tex[s * nn + c] = -t[c];
            if (tT) {
                for (int j = 0; j < nd; j++) {
                    tT[s * nd + j] = 0.0;
                    for (int i = 0; i < nd; i++) tT[s * nd + j] -= t[j * nd + i] * ps[i] / T;
                }
            }
        }
    } else if (m->model == CIPHER_CHEM_QUADRATIC) {
        double delta[CIPHER_MAXCP], ceqT[CIPHER_MAXCP], unT[CIPHER_MAXCP], biT[CIPHER_MAXCP];
        double ceq[CIPHER_MAXCP], bi[CIPHER_MAXCP];
        double mt[CIPHER_MAXCP * CIPHER_MAXCP], mi[CIPHER_MAXCP * CIPHER_MAXCP];
        for (int c = 0; c < nc; c++) {
            ceqT[c] = orc_sum_tseries_derivative(T, TS(m->quad_ceq + c));
            unT[c] = orc_sum_tseries_derivative(T, TS(m->unary + c));
            biT[c] = orc_sum_tseries_derivative(T, TS(m->quad_binary + c));
            ceq[c] = orc_sum_tseries(T, TS(m->quad_ceq + c));
            bi[c] = orc_sum_tseries(T, TS(m->quad_binary + c));
        }
        memset(mt, 0, sizeof(mt)); memset(mi, 0, sizeof(mi));
        for (int j = 0; j < nd; j++) {
            mt[j * nd + j] += bi[j];
            for (int i = 0; i < nd; i++) mt[j * nd + i] += bi[nd];
        }
        orc_invert_matrix(mi, mt, nd);
        if (tpot) { for (int j = 0; j < nn; j++) tpot[j] = mi[j] * 0.5; }
        if (tex) { for (int j = 0; j < nn; j++) tex[j] = 0.0; }
        if (tT) {
            for (int j = 0; j < nd; j++)
                delta[j] = biT[nd] - 0.5 * (unT[j] - unT[nd])
                           + (biT[j] * ceq[j] + bi[j] * ceqT[j] - biT[nd] * ceq[nd] - bi[nd] * ceqT[nd]);
            matvec(tT, mi, delta, nd);
        }
    } else {
        if (tpot) for (int j = 0; j < nn; j++) tpot[j] = 0.0;
        if (tex) for (int j = 0; j < nn; j++) tex[j] = 0.0;
        if (tT) for (int j = 0; j < nd; j++) tT[j] = 0.0;
    }
    if (tpot) for (int c = 0; c < m->nsites * nn; c++) tpot[c] *= m->molarvolume;
    if (tex) for (int c = 0; c < m->nsites * nn; c++) tex[c] *= m->molarvolume;
    if (tT) for (int c = 0; c < m->nsites * nd; c++) tT[c] *= m->molarvolume;
}

/* CompositionMobility_{2sl,dis,quad,none}, src/material.c:1023-1110 and
 * CompositionMobilityComponent, :1115-1120 */
void orc_mobility_component(const cipher_config *cfg, int mat, double *mob, const double *y, double T)
{
    const cipher_material *m = &cfg->material[mat];
    int nc = cfg->ncp;
    if (m->model == CIPHER_CHEM_NONE) {
        for (int k = 0; k < nc; k++) mob[k] = 0.0;
    } else {
        for (int k = 0; k < nc; k++) {
            const cipher_mobility *mb = &cfg->mobility[m->mobility + k];
            const cipher_rk *bin = RKP(mb->binary);
            double mig = 0.0;
            for (int j = 0; j < nc; j++) {
                mig += y[j] * orc_sum_tseries(T, TS(mb->unary + j));
                for (int i = j + 1; i < nc; i++, bin++)
                    for (int o = 0; o < bin->n; o++)
                        mig += y[j] * y[i] * rk_term(cfg, bin, o, T) * orc_fastpow(y[j] - y[i], (unsigned)o);
            }
            mob[k] = y[k] * mb->m0 / RG / T * exp(mig / RG / T);
        }
    }
    for (int c = 0; c < nc; c++) mob[c] *= m->molarvolume;
}

/* CompositionMobilityVolumeRef, src/material.c:1125-1137 (Andersson-Agren) */
void orc_mobility_volumeref(const cipher_config *cfg, double *out, const double *mob, const double *comp)
{
    int nd = cfg->ncp - 1, nc = cfg->ncp;
    for (int i = 0; i < nd; i++)
        for (int j = 0; j < nd; j++) {
            out[i * nd + j] = 0.0;
            for (int k = 0; k < nc; k++)
                out[i * nd + j] += ((j == k ? 1.0 : 0.0) - comp[j]) * ((k == i ? 1.0 : 0.0) - comp[i]) * mob[k];
        }
}

/* Chemsource_sink, src/material.c:295-303 */
static void source_sink(double *out, const double *avgcomp, const cipher_source *s, int nc)
{
    for (int c = 0; c < nc; c++) out[c] = s->rate[c] * (s->ceq[c] - avgcomp[c]);
}

/* Chemsource, src/material.c:321-351.  The third argument receives the
 * interpolants at the call site (src/cipher.c:645).  Random sources draw from
 * rand() in the reference and contribute nothing here. */
static void chemsource(const cipher_config *cfg, double *out, const double *avgcomp,
                       const double *pf, const uint16_t *ids)
{
    int nc = cfg->ncp, npf = cfg->npf;
    double bulk[CIPHER_MAXCP], intf[CIPHER_MAXCP];
    for (int c = 0; c < nc; c++) out[c] = 0.0;
    for (int gk = 0; gk < ids[0]; gk++) {
        const cipher_material *m = &cfg->material[cfg->phase_material[ids[gk + 1]]];
        for (int s = 0; s < m->nsources; s++) {
            const cipher_source *src = &cfg->source[m->source + s];
            if (src->model != CIPHER_SOURCE_SINK) continue;
            source_sink(bulk, avgcomp, src, nc);
            for (int c = 0; c < nc; c++) out[c] += pf[gk] * pf[gk] * bulk[c];
        }
        for (int gj = gk + 1; gj < ids[0]; gj++) {
            const cipher_interface *it = &cfg->interface[cfg->interface_list[ids[gk + 1] * npf + ids[gj + 1]]];
            for (int s = 0; s < it->nsources; s++) {
                const cipher_source *src = &cfg->source[it->source + s];
                if (src->model != CIPHER_SOURCE_SINK) continue;
                source_sink(intf, avgcomp, src, nc);
                for (int c = 0; c < nc; c++) out[c] += 2.0 * pf[gk] * pf[gj] * intf[c];
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* face adjacency (both directions) used by the set unions              */
/* ------------------------------------------------------------------ */
static int build_adjacency(const cipher_mesh *mesh, int **off_out, int **nbr_out)
{
    int N = mesh->ncells;
    int *off = calloc((size_t)N + 1, sizeof(int));
    int *nbr = malloc((size_t)(2 * mesh->nfaces + 1) * sizeof(int));
    if (!off || !nbr) { free(off); free(nbr); return -1; }
    for (int f = 0; f < mesh->nfaces; f++) {
        off[mesh->face_cells[2 * f] + 1]++;
        off[mesh->face_cells[2 * f + 1] + 1]++;
    }
    for (int c = 0; c < N; c++) off[c + 1] += off[c];
    int *pos = malloc((size_t)N * sizeof(int));
    if (!pos) { free(off); free(nbr); return -1; }
    memcpy(pos, off, (size_t)N * sizeof(int));
    for (int f = 0; f < mesh->nfaces; f++) {
        int a = mesh->face_cells[2 * f], b = mesh->face_cells[2 * f + 1];
        nbr[pos[a]++] = b;
        nbr[pos[b]++] = a;
    }
    free(pos);
    *off_out = off;
    *nbr_out = nbr;
    return 0;
}

#define PMAX 16

/* Least-squares gradient weights, src/cipher.c:1233-1282: per cell, the
 * rows of A are the neighbours' centroid offsets (face neighbours in face
 * order; a periodic direction takes the nearest image), and the weights are
 * (A^T A)^-1 A^T, stored per adjacency entry: g[k*dim + d]. */
static int lsq_setup(const cipher_config *cfg, const cipher_mesh *mesh, const int *off, const int *nbr, double **out)
{
    const int N = mesh->ncells, dim = cfg->dim;
    if (!mesh->centroid) return CIPHER_ERR_ARG;
    double *g = calloc((size_t)(off[N] > 0 ? off[N] : 1) * dim, sizeof(double));
    if (!g) return CIPHER_ERR_ARG;
    for (int cell = 0; cell < N; cell++) {
        const int nls = off[cell + 1] - off[cell];
        double a[24 * 3], ata[9], atainv[9], gm[24 * 3];
        if (nls > 24) { free(g); return CIPHER_ERR_ARG; }
        for (int j = 0; j < nls; j++) {
            const int nb = nbr[off[cell] + j];
            for (int d = 0; d < dim; d++) {
                double v = mesh->centroid[(size_t)nb * dim + d];
                v -= mesh->centroid[(size_t)cell * dim + d];
                const double L = mesh->period[d];
                if (L > 0.0) {
                    if (v > 0.5 * L) v -= L;
                    else if (v < -0.5 * L) v += L;
                }
                a[j * dim + d] = v;
            }
        }
        for (int di = 0; di < dim; di++)
            for (int dj = di; dj < dim; dj++) {
                double val = 0.0;
                for (int j = 0; j < nls; j++) val += a[j * dim + di] * a[j * dim + dj];
                ata[di * dim + dj] = val;
                ata[dj * dim + di] = val;
            }
        orc_invert_matrix(atainv, ata, dim);
        memset(gm, 0, sizeof(gm));
        for (int di = 0; di < dim; di++)
            for (int j = 0; j < nls; j++)
                for (int d = 0; d < dim; d++) gm[di * nls + j] += atainv[di * dim + d] * a[j * dim + d];
        for (int j = 0; j < nls; j++)
            for (int d = 0; d < dim; d++) g[(size_t)(off[cell] + j) * dim + d] = gm[d * nls + j];
    }
    *out = g;
    return CIPHER_OK;
}

/* Phase-field gradients of every cell's active phases, src/cipher.c:82-146:
 * grad[(cell*PMAX + g)*dim + d]. */
static void lsq_gradients(const cipher_config *cfg, const cipher_mesh *mesh, const int *off, const int *nbr,
                          const double *gw, const double *X, int D, int aso, int pfo, double *grad)
{
    const int N = mesh->ncells, dim = cfg->dim;
    uint16_t sl[PMAX + 1], nl[PMAX + 1], is[PMAX + 1], iL[PMAX + 1], iR[PMAX + 1];
    for (int cell = 0; cell < N; cell++) {
        const double *x = X + (size_t)cell * D;
        f2i(sl, x + aso);
        const double *pc = x + pfo;
        const int nls = off[cell + 1] - off[cell];
        double *gc = grad + (size_t)cell * PMAX * dim;
        for (int g = 0; g < sl[0]; g++)
            for (int d = 0; d < dim; d++) {
                gc[g * dim + d] = 0.0;
                for (int j = 0; j < nls; j++) gc[g * dim + d] -= gw[(size_t)(off[cell] + j) * dim + d] * pc[g];
            }
        for (int j = 0; j < nls; j++) {
            const double *xn = X + (size_t)nbr[off[cell] + j] * D;
            f2i(nl, xn + aso);
            orc_set_intersection(is, iL, iR, sl, nl);
            for (int g = 0; g < is[0]; g++)
                for (int d = 0; d < dim; d++)
                    gc[iL[g] * dim + d] += gw[(size_t)(off[cell] + j) * dim + d] * xn[pfo + iR[g]];
        }
    }
}

/* Gradients of every cell's active phase fractions (test access):
 * grad[(cell*16 + g)*dim + d]. */
int orc_gradients(const cipher_config *cfg, const cipher_mesh *mesh, const double *X, double *grad)
{
    olayout L = orc_layout(cfg);
    int *off, *nbr;
    double *gw;
    if (build_adjacency(mesh, &off, &nbr)) return CIPHER_ERR_ARG;
    int rc = lsq_setup(cfg, mesh, off, nbr, &gw);
    if (rc == CIPHER_OK) {
        lsq_gradients(cfg, mesh, off, nbr, gw, X, L.ndof, L.aso, L.pfo, grad);
        free(gw);
    }
    free(off); free(nbr);
    return rc;
}

/* ------------------------------------------------------------------ */
/* RHSFunction, src/cipher.c:27-681                                    */
/* ------------------------------------------------------------------ */

int orc_rhs(const cipher_config *cfg, const cipher_mesh *mesh, double t,
            const double *X, double *F, uint32_t *activeout)
{
    (void)t;
    olayout L = orc_layout(cfg);
    const int N = mesh->ncells, D = L.ndof, P = L.pf;
    const int nc = cfg->ncp, nd = nc - 1, npf = cfg->npf, dim = cfg->dim;
    const int SF = L.sf, SP = L.sp, DPS = L.dp;
    if (P > PMAX) return CIPHER_ERR_ARG;
    double *grad = NULL;
    if (cfg->gradient_calculation) {
        int *goff, *gnbr;
        double *gw;
        if (build_adjacency(mesh, &goff, &gnbr)) return CIPHER_ERR_ARG;
        int rc = lsq_setup(cfg, mesh, goff, gnbr, &gw);
        if (rc == CIPHER_OK) {
            grad = calloc((size_t)N * PMAX * dim, sizeof(double));
            if (grad) lsq_gradients(cfg, mesh, goff, gnbr, gw, X, D, L.aso, L.pfo, grad);
            else rc = CIPHER_ERR_ARG;
            free(gw);
        }
        free(goff); free(gnbr);
        if (rc != CIPHER_OK) return rc;
    }

    double *lap = calloc((size_t)N * D, sizeof(double));
    double *sfg = calloc((size_t)N * P * (SF > 0 ? SF : 1), sizeof(double));
    double *compg = calloc((size_t)N * nc, sizeof(double));
    double *mobg = calloc((size_t)N * nc, sizeof(double));
    double *tcond = calloc((size_t)N, sizeof(double));
    if (!lap || !sfg || !compg || !mobg || !tcond) {
        free(lap); free(sfg); free(compg); free(mobg); free(tcond); free(grad);
        return CIPHER_ERR_ARG;
    }
    memset(F, 0, (size_t)N * D * sizeof(double));
    if (activeout) memset(activeout, 0, (size_t)N * sizeof(uint32_t));

    uint16_t sl[PMAX + 1], sL[PMAX + 1], sR[PMAX + 1], isect[PMAX + 1], injL[PMAX + 1], injR[PMAX + 1];
    double h[PMAX];

    /* per-cell precompute, src/cipher.c:149-233 */
    for (int cell = 0; cell < N; ++cell) {
        const double *x = X + (size_t)cell * D;
        f2i(sl, x + L.aso);
        const double *pc = x + L.pfo, *mu = x + L.dpo, *ex = x + L.exo;
        double T = x[L.tmo];
        orc_interpolant(cfg->interpolation, h, pc, sl[0]);
        tcond[cell] = 0.0;
        if (nd) {
            double *sf = sfg + (size_t)cell * P * SF;
            double *comp = compg + (size_t)cell * nc, *mob = mobg + (size_t)cell * nc;
            double mui[CIPHER_MAXCP], spim[2 * CIPHER_MAXCP], me[CIPHER_MAXCP];
            for (int c = 0; c < nc; c++) { mob[c] = 0.0; comp[c] = 0.0; }
            for (int c = 0; c < nd; c++) mui[c] = 0.0;
            for (int gk = 0; gk < sl[0]; gk++)
                for (int gj = gk + 1; gj < sl[0]; gj++) {
                    const cipher_interface *it = &cfg->interface[cfg->interface_list[sl[gk + 1] * npf + sl[gj + 1]]];
                    if (it->has_potential)
                        for (int c = 0; c < nd; c++) mui[c] += h[gk] * h[gj] * it->potential[c];
                }
            for (int g = 0; g < sl[0]; g++) {
                int mat = cfg->phase_material[sl[g + 1]];
                const cipher_material *m = &cfg->material[mat];
                for (int s = 0; s < m->nsites; s++)
                    for (int c = 0; c < nd; c++)
                        spim[s * nd + c] = m->stoichiometry[s] * (mu[c] - mui[c]) - ex[g * SP + s * nd + c];
                orc_sitefrac(cfg, mat, sf + g * SF, spim, T);
                orc_mobility_component(cfg, mat, me, sf + g * SF, T);
                for (int c = 0; c < nc; c++) mob[c] += h[g] * me[c];
                for (int s = 0; s < m->nsites; s++)
                    for (int c = 0; c < nc; c++) comp[c] += h[g] * m->stoichiometry[s] * sf[g * SF + s * nc + c];
                if (cfg->tseries[m->tconductivity].n)
                    tcond[cell] += h[g] * orc_sum_tseries(T, TS(m->tconductivity));
            }
            for (int gk = 0; gk < sl[0]; gk++)
                for (int gj = gk + 1; gj < sl[0]; gj++) {
                    const cipher_interface *it = &cfg->interface[cfg->interface_list[sl[gk + 1] * npf + sl[gj + 1]]];
                    if (it->has_mobilityc)
                        for (int c = 0; c < nc; c++) mob[c] *= exp(it->mobilityc[c] / RG / T);
                }
        } else {
            for (int g = 0; g < sl[0]; g++) {
                const cipher_material *m = &cfg->material[cfg->phase_material[sl[g + 1]]];
                if (cfg->tseries[m->tconductivity].n)
                    tcond[cell] += h[g] * orc_sum_tseries(T, TS(m->tconductivity));
            }
        }
    }

    /* interior faces, src/cipher.c:236-317 */
    for (int f = 0; f < mesh->nfaces; ++f) {
        int cL = mesh->face_cells[2 * f], cR = mesh->face_cells[2 * f + 1];
        if (cL >= N || cR >= N) continue;
        const double *xL = X + (size_t)cL * D, *xR = X + (size_t)cR * D;
        double *lL = lap + (size_t)cL * D, *lR = lap + (size_t)cR * D;
        f2i(sL, xL + L.aso);
        f2i(sR, xR + L.aso);
        const double *pL = xL + L.pfo, *pR = xR + L.pfo;
        double dL = mesh->cellgeom[cL], dR = mesh->cellgeom[cR];
        double vL = orc_fastpow(dL, (unsigned)dim), vR = orc_fastpow(dR, (unsigned)dim);
        double ff = 2.0 * (dL < dR ? vL / dL : vR / dR) / (dL + dR);
        if (sL[0] > 1 || sL[1] > 1) {
            orc_set_intersection(isect, injL, injR, sL, sR);
            for (int k = 0; k < isect[0]; k++) {
                double fl = pR[injR[k]] - pL[injL[k]];
                lL[L.pfo + injL[k]] += ff * fl / vL;
                lR[L.pfo + injR[k]] -= ff * fl / vR;
            }
        }
        if (nd) {
            const double *mL = mobg + (size_t)cL * nc, *mR = mobg + (size_t)cR * nc;
            const double *cmL = compg + (size_t)cL * nc, *cmR = compg + (size_t)cR * nc;
            double me[CIPHER_MAXCP], ca[CIPHER_MAXCP], MB[CIPHER_MAXCP * CIPHER_MAXCP];
            double dmu[CIPHER_MAXCP], fd[CIPHER_MAXCP];
            for (int c = 0; c < nc; c++) {
                if (fabs(mL[c]) > 1e-32 && fabs(mR[c]) > 1e-32) me[c] = 2.0 * mL[c] * mR[c] / (mL[c] + mR[c]);
                else me[c] = 0.0;
                ca[c] = 2.0 * cmL[c] * cmR[c] / (cmL[c] + cmR[c]);
            }
            orc_mobility_volumeref(cfg, MB, me, ca);
            for (int c = 0; c < nd; c++) dmu[c] = xR[L.dpo + c] - xL[L.dpo + c];
            matvec(fd, MB, dmu, nd);
            for (int c = 0; c < nd; c++) {
                lL[L.dpo + c] += ff * fd[c] / vL;
                lR[L.dpo + c] -= ff * fd[c] / vR;
            }
        }
        if (fabs(tcond[cL]) > 1e-32 && fabs(tcond[cR]) > 1e-32) {
            double ft = 2.0 * tcond[cL] * tcond[cR] / (tcond[cL] + tcond[cR]) * (xR[L.tmo] - xL[L.tmo]);
            lL[L.tmo] += ff * ft / vL;
            lR[L.tmo] -= ff * ft / vR;
        }
    }

    /* boundary faces, src/cipher.c:320-389 (per condition, faces in order) */
    for (int b = 0; b < cfg->nbcs; b++) {
        const cipher_bc *bc = &cfg->bc[b];
        for (int f = 0; f < mesh->nbfaces; f++) {
            if (mesh->bface_id[f] != bc->boundary_id) continue;
            int cL = mesh->bface_cell[f];
            const double *xL = X + (size_t)cL * D;
            double *lL = lap + (size_t)cL * D;
            const double *mL = mobg + (size_t)cL * nc;
            double dL = mesh->cellgeom[cL];
            double fd[CIPHER_MAXCP], ft = 0.0;
            if (bc->chem_type != CIPHER_BC_NONE) {
                for (int c = 0; c < nd; c++) fd[c] = 0.0;
                if (bc->chem_type == CIPHER_BC_NEUMANN) {
                    for (int c = 0; c < nd; c++) if (bc->chem_mask & (1 << c)) fd[c] = bc->chem_val[c] * dL;
                } else {
                    for (int c = 0; c < nd; c++)
                        if (bc->chem_mask & (1 << c)) fd[c] = mL[c] * (bc->chem_val[c] - xL[L.dpo + c]);
                }
            }
            if (bc->thermal_type == CIPHER_BC_NEUMANN) ft = tcond[cL] * bc->thermal_val * dL;
            else if (bc->thermal_type == CIPHER_BC_DIRICHLET) ft = tcond[cL] * (bc->thermal_val - xL[L.tmo]);
            double cf = 1.0 / orc_fastpow(dL, 2);
            if (bc->chem_type != CIPHER_BC_NONE)
                for (int c = 0; c < nd; c++) lL[L.dpo + c] += cf * fd[c];
            if (bc->thermal_type != CIPHER_BC_NONE) lL[L.tmo] += cf * ft;
        }
    }

    /* cells, src/cipher.c:392-664 */
    for (int cell = 0; cell < N; ++cell) {
        const double *x = X + (size_t)cell * D;
        const double *lp = lap + (size_t)cell * D;
        double *fr = F + (size_t)cell * D;
        f2i(sl, x + L.aso);
        int n = sl[0];
        const double *pc = x + L.pfo, *mu = x + L.dpo, *ex = x + L.exo;
        double T = x[L.tmo];
        const double *sf = sfg + (size_t)cell * P * SF;
        const double *comp = compg + (size_t)cell * nc;
        const double *plapl = lp + L.pfo, *dlapl = lp + L.dpo, tlapl = lp[L.tmo];
        double *pfrhs = fr + L.pfo, *dprhs = fr + L.dpo, *cprhs = fr + L.exo, *tmrhs = fr + L.tmo;
        orc_interpolant(cfg->interpolation, h, pc, n);
        if (n > 1) {
            double ie[PMAX * PMAX], im[PMAX * PMAX], iw[PMAX * PMAX];
            double cap[PMAX], chs[PMAX], ru[PMAX];
            int act[PMAX];
            double wdp[CIPHER_MAXCP], rv, tj, nact;
            for (int gk = 0; gk < n; gk++)
                for (int gj = gk + 1; gj < n; gj++) {
                    const cipher_interface *it = &cfg->interface[cfg->interface_list[sl[gk + 1] * npf + sl[gj + 1]]];
                    int kj = gk * n + gj;
                    ie[kj] = 1.0;
                    im[kj] = 1.0;
                    if (grad) {
                        /* anisotropy, src/cipher.c:429-457 */
                        const double *gk_ = grad + ((size_t)cell * PMAX + gk) * dim;
                        const double *gj_ = grad + ((size_t)cell * PMAX + gj) * dim;
                        double nv[3], rn = 0.0, dot;
                        for (int d = 0; d < dim; d++) {
                            nv[d] = gk_[d] - gj_[d];
                            rn += nv[d] * nv[d];
                        }
                        rn = sqrt(rn);
                        if (rn > 1e-16) for (int d = 0; d < dim; d++) nv[d] /= rn;
                        for (int r = 0; r < it->e_naniso; r++) {
                            dot = 0.0;
                            for (int d = 0; d < dim; d++) dot += nv[d] * cfg->aniso[it->e_aniso + dim * r + d];
                            ie[kj] += orc_fastpow(dot, 4) * cfg->aniso[it->e_aniso + dim * it->e_naniso + r];
                        }
                        for (int r = 0; r < it->m_naniso; r++) {
                            dot = 0.0;
                            for (int d = 0; d < dim; d++) dot += nv[d] * cfg->aniso[it->m_aniso + dim * r + d];
                            im[kj] += orc_fastpow(dot, 4) * cfg->aniso[it->m_aniso + dim * it->m_naniso + r];
                        }
                    }
                    ie[kj] *= it->e0;
                    if (cfg->tseries[it->e_act].n) ie[kj] *= exp(orc_sum_tseries(T, TS(it->e_act)) / RG / T);
                    im[kj] *= it->m0;
                    if (cfg->tseries[it->m_act].n) im[kj] *= exp(orc_sum_tseries(T, TS(it->m_act)) / RG / T);
                    iw[kj] = it->width;
                }
            for (int g = 0; g < n; g++) cap[g] = 0.0;
            for (int gk = 0; gk < n; gk++) {
                for (int gj = gk + 1; gj < n; gj++) {
                    int kj = gk * n + gj;
                    cap[gk] -= ie[kj] * (pc[gj] / iw[kj] + plapl[gj] * iw[kj] / PI_D / PI_D);
                    cap[gj] -= ie[kj] * (pc[gk] / iw[kj] + plapl[gk] * iw[kj] / PI_D / PI_D);
                    for (int gi = gj + 1; gi < n; gi++) {
                        tj = 0.0;
                        tj = tj > ie[gk * n + gj] / iw[gk * n + gj] ? tj : ie[gk * n + gj] / iw[gk * n + gj];
                        tj = tj > ie[gk * n + gi] / iw[gk * n + gi] ? tj : ie[gk * n + gi] / iw[gk * n + gi];
                        tj = tj > ie[gj * n + gi] / iw[gj * n + gi] ? tj : ie[gj * n + gi] / iw[gj * n + gi];
                        tj *= cfg->junctionpenalty;
                        cap[gk] -= tj * pc[gj] * pc[gi];
                        cap[gj] -= tj * pc[gk] * pc[gi];
                        cap[gi] -= tj * pc[gk] * pc[gj];
                    }
                }
                cap[gk] *= 8.0;
                cap[gk] += plapl[gk];
            }
            /* Chemenergy, src/material.c:480-490 */
            for (int g = 0; g < n; g++) {
                int mat = cfg->phase_material[sl[g + 1]];
                const cipher_material *m = &cfg->material[mat];
                chs[g] = orc_chemenergy_phase(cfg, mat, sf + g * SF, T);
                chs[g] /= m->molarvolume;
                for (int s = 0; s < m->nsites; s++)
                    for (int c = 0; c < nd; c++) chs[g] -= mu[c] * m->stoichiometry[s] * sf[g * SF + s * nc + c];
            }
            for (int c = 0; c < DPS; c++) wdp[c] = 0.0;
            for (int g = 0; g < n; g++) {
                const cipher_material *m = &cfg->material[cfg->phase_material[sl[g + 1]]];
                for (int s = 0; s < m->nsites; s++)
                    for (int c = 0; c < nd; c++) wdp[c] += h[g] * m->stoichiometry[s] * sf[g * SF + s * nc + c];
            }
            for (int gk = 0; gk < n; gk++)
                for (int gj = gk + 1; gj < n; gj++) {
                    const cipher_interface *it = &cfg->interface[cfg->interface_list[sl[gk + 1] * npf + sl[gj + 1]]];
                    if (it->has_potential) {
                        rv = 0.0;
                        for (int c = 0; c < nd; c++) rv += it->potential[c] * wdp[c];
                        chs[gk] += rv * h[gj];
                        chs[gj] += rv * h[gk];
                    }
                }
            orc_interpolant_derivative(cfg->interpolation, chs, pc, n);
            nact = 0.0;
            for (int g = 0; g < n; g++) ru[g] = 0.0;
            for (int gk = 0; gk < n; gk++) {
                for (int gj = gk + 1; gj < n; gj++) {
                    int kj = gk * n + gj;
                    rv = im[kj] * ((cap[gk] - cap[gj]) - (chs[gk] - chs[gj]) * 8.0 * sqrt(h[gk] * h[gj]) / PI_D);
                    ru[gk] += rv;
                    ru[gj] -= rv;
                }
                act[gk] = (pc[gk] > CIPHER_TOL && pc[gk] < 1.0 - CIPHER_TOL)
                          || (pc[gk] < CIPHER_TOL && ru[gk] > 0.0)
                          || (pc[gk] > 1.0 - CIPHER_TOL && ru[gk] < 0.0);
                if (act[gk]) nact += 1.0;
            }
            if (activeout) {
                uint32_t bits = 0;
                for (int g = 0; g < n; g++) if (act[g]) bits |= 1u << g;
                activeout[cell] = bits;
            }
            for (int gk = 0; gk < n; gk++) {
                if (act[gk]) {
                    for (int gj = gk + 1; gj < n; gj++) {
                        if (act[gj]) {
                            int kj = gk * n + gj;
                            rv = im[kj] * ((cap[gk] - cap[gj]) - (chs[gk] - chs[gj]) * 8.0 * sqrt(h[gk] * h[gj]) / PI_D);
                            pfrhs[gk] += rv;
                            pfrhs[gj] -= rv;
                        }
                    }
                    pfrhs[gk] /= nact;
                }
            }
        }

        /* temperature rate, src/cipher.c:558-568 */
        {
            double cp_ = 0.0;
            *tmrhs = tlapl;
            for (int g = 0; g < n; g++) {
                const cipher_material *m = &cfg->material[cfg->phase_material[sl[g + 1]]];
                if (cfg->tseries[m->specific_heat].n) cp_ += h[g] * orc_sum_tseries(T, TS(m->specific_heat));
                if (cfg->tseries[m->latent_heat].n) *tmrhs += pfrhs[g] * orc_sum_tseries(T, TS(m->latent_heat));
            }
            if (cp_ != 0.0) *tmrhs /= cp_;
            if (cfg->has_temperature_rate) *tmrhs += cfg->temperature_rate;
        }

        if (nd) {
            double spw[2 * CIPHER_MAXCP], mui[CIPHER_MAXCP], spim[2 * CIPHER_MAXCP];
            double wcp[PMAX * CIPHER_MAXCP], wdp[CIPHER_MAXCP], wdp2[CIPHER_MAXCP], wpf[PMAX];
            double ctpot[PMAX * 2 * CIPHER_MAXCP * CIPHER_MAXCP], ctex[PMAX * 2 * CIPHER_MAXCP * CIPHER_MAXCP];
            double ctt[PMAX * 2 * CIPHER_MAXCP];
            double dcdm[CIPHER_MAXCP * CIPHER_MAXCP], dmdc[CIPHER_MAXCP * CIPHER_MAXCP], csrc[CIPHER_MAXCP];
            const int SPD = SP * DPS;   /* per-phase tangent stride, src/cipher.c:610 */
            /* explicit potential rate, :572-581 */
            for (int g = 0; g < n; g++) {
                int mat = cfg->phase_material[sl[g + 1]];
                const cipher_material *m = &cfg->material[mat];
                orc_sitepot_explicit(cfg, mat, spw, sf + g * SF, T);
                for (int s = 0; s < m->nsites; s++)
                    for (int c = 0; c < nd; c++)
                        cprhs[g * SP + s * nd + c] = m->chempot_ex_kineticcoeff * (spw[s * nd + c] - ex[g * SP + s * nd + c]);
            }
            /* composition rate, :584-662 */
            for (int c = 0; c < nd; c++) mui[c] = 0.0;
            for (int c = 0; c < P * DPS; c++) wcp[c] = 0.0;
            for (int gk = 0; gk < n; gk++)
                for (int gj = gk + 1; gj < n; gj++) {
                    const cipher_interface *it = &cfg->interface[cfg->interface_list[sl[gk + 1] * npf + sl[gj + 1]]];
                    if (it->has_potential)
                        for (int c = 0; c < nd; c++) {
                            mui[c] += h[gk] * h[gj] * it->potential[c];
                            wcp[gk * DPS + c] += it->potential[c] * h[gj];
                            wcp[gj * DPS + c] += it->potential[c] * h[gk];
                        }
                }
            for (int c = 0; c < DPS * DPS; c++) dcdm[c] = 0.0;
            for (int g = 0; g < n; g++) {
                int mat = cfg->phase_material[sl[g + 1]];
                const cipher_material *m = &cfg->material[mat];
                for (int s = 0; s < m->nsites; s++)
                    for (int c = 0; c < nd; c++)
                        spim[s * nd + c] = m->stoichiometry[s] * (mu[c] - mui[c]) - ex[g * SP + s * nd + c];
                orc_sitefrac_tangent(cfg, mat, ctpot + g * SPD, ctex + g * SPD, ctt + g * SP, sf + g * SF, spim, T);
                for (int s = 0; s < m->nsites; s++)
                    for (int c = 0; c < nd * nd; c++)
                        dcdm[c] += h[g] * m->stoichiometry[s] * m->stoichiometry[s] * ctpot[g * SPD + s * DPS * DPS + c];
            }
            for (int g = 0; g < n; g++) {
                memcpy(wdp2, wcp + g * DPS, DPS * sizeof(double));
                matvec(wcp + g * DPS, dcdm, wdp2, DPS);
            }
            for (int c = 0; c < DPS; c++) wdp[c] = 0.0;
            if (n > 1) {
                for (int c = 0; c < nd; c++) {
                    for (int g = 0; g < n; g++) {
                        const cipher_material *m = &cfg->material[cfg->phase_material[sl[g + 1]]];
                        wpf[g] = 0.0;
                        for (int s = 0; s < m->nsites; s++) wpf[g] += m->stoichiometry[s] * sf[g * SF + s * nc + c];
                        wpf[g] -= wcp[g * DPS + c];
                    }
                    orc_interpolant_derivative(cfg->interpolation, wpf, pc, n);
                    for (int g = 0; g < n; g++) wdp[c] -= pfrhs[g] * wpf[g];
                }
            }
            chemsource(cfg, csrc, comp, h, sl);
            for (int c = 0; c < nd; c++) wdp[c] += (dlapl[c] + csrc[c]);
            for (int g = 0; g < n; g++) {
                const cipher_material *m = &cfg->material[cfg->phase_material[sl[g + 1]]];
                for (int s = 0; s < m->nsites; s++) {
                    matvec(wcp + g * DPS, ctex + g * SPD + s * DPS * DPS, cprhs + g * SP + s * DPS, DPS);
                    for (int c = 0; c < nd; c++)
                        wdp[c] -= h[g] * m->stoichiometry[s] * (wcp[g * DPS + c] + ctt[g * SP + s * DPS + c] * (*tmrhs));
                }
            }
            orc_invert_matrix(dmdc, dcdm, DPS);
            matvec(dprhs, dmdc, wdp, DPS);
        }
    }
    free(lap); free(sfg); free(compg); free(mobg); free(tcond); free(grad);
    return CIPHER_OK;
}

/* PostStep active-set block, src/cipher.c:869-974 */
int orc_poststep(const cipher_config *cfg, const cipher_mesh *mesh, double *X)
{
    olayout L = orc_layout(cfg);
    const int N = mesh->ncells, D = L.ndof, SP = L.sp;
    int *off, *nbr, rc = CIPHER_OK;
    if (build_adjacency(mesh, &off, &nbr)) return CIPHER_ERR_ARG;
    double *gX = malloc((size_t)N * D * sizeof(double));
    if (!gX) { free(off); free(nbr); return CIPHER_ERR_ARG; }
    memcpy(gX, X, (size_t)N * D * sizeof(double));
    uint16_t sl[2 * PMAX + 2], gl[2 * PMAX + 2];
    /* projection and pruning, :875-909 */
    for (int cell = 0; cell < N; ++cell) {
        double *x = X + (size_t)cell * D;
        double *gx = gX + (size_t)cell * D;
        f2i(sl, x + L.aso);
        double *pc = x + L.pfo;
        if (sl[0] > 1) {
            double up[PMAX];
            memcpy(up, pc, sl[0] * sizeof(double));
            orc_simplex_projection(pc, up, sl[0]);
            gl[0] = 0;
            for (int g = 0; g < sl[0]; g++)
                if (pc[g] > CIPHER_TOL) gl[++gl[0]] = sl[g + 1];
            i2f(gx + L.aso, gl);
        } else {
            pc[0] = 1.0;
            gl[0] = 1; gl[1] = sl[1];
            i2f(gx + L.aso, gl);
        }
    }
    /* superset with neighbours, reorder, :920-969 */
    for (int cell = 0; cell < N && rc == CIPHER_OK; ++cell) {
        double *x = X + (size_t)cell * D;
        uint16_t ss[2 * PMAX + 2], un[2 * PMAX + 2], nl[2 * PMAX + 2], is[2 * PMAX + 2];
        uint16_t iL[2 * PMAX + 2], iR[2 * PMAX + 2];
        f2i(sl, x + L.aso);
        double *pc = x + L.pfo, *cc = x + L.exo;
        double T = x[L.tmo];
        f2i(ss, gX + (size_t)cell * D + L.aso);
        for (int k = off[cell]; k < off[cell + 1]; k++) {
            f2i(nl, gX + (size_t)nbr[k] * D + L.aso);
            if (ss[0] + nl[0] > 2 * PMAX) { rc = CIPHER_ERR_CAPACITY; break; }
            orc_set_union(un, iL, iR, ss, nl);
            memcpy(ss, un, (un[0] + 1) * sizeof(uint16_t));
        }
        if (rc != CIPHER_OK) break;
        if (ss[0] >= L.as) { rc = CIPHER_ERR_CAPACITY; break; }
        /* zeroed: a one-sublattice material fills only the first nd entries of
         * its slot; the rest are never read, and stay deterministic */
        double phiss[PMAX], spss[PMAX * 2 * CIPHER_MAXCP] = {0};
        for (int g = 0; g < ss[0]; g++) {
            phiss[g] = 0.0;
            int mat = cfg->phase_material[ss[g + 1]];
            orc_sitepot_explicit(cfg, mat, spss + g * SP, cfg->material[mat].c0, T);
        }
        orc_set_intersection(is, iL, iR, sl, ss);
        for (int g = 0; g < is[0]; g++) {
            phiss[iR[g]] = pc[iL[g]];
            memcpy(spss + iR[g] * SP, cc + iL[g] * SP, SP * sizeof(double));
        }
        memcpy(pc, phiss, ss[0] * sizeof(double));
        memcpy(cc, spss, (size_t)ss[0] * SP * sizeof(double));
        i2f(x + L.aso, ss);
    }
    free(gX); free(off); free(nbr);
    return rc;
}

/* SetUpProblem, src/init.c:1721-1830 */
int orc_setup_problem(const cipher_config *cfg, const cipher_mesh *mesh, const int32_t *phase, double *X)
{
    olayout L = orc_layout(cfg);
    const int N = mesh->ncells, D = L.ndof, SP = L.sp, nd = cfg->ncp - 1;
    int *off, *nbr;
    if (build_adjacency(mesh, &off, &nbr)) return CIPHER_ERR_ARG;
    memset(X, 0, (size_t)N * D * sizeof(double));
    for (int cell = 0; cell < N; ++cell) {
        X[(size_t)cell * D + L.aso] = 1.0;
        X[(size_t)cell * D + L.aso + 1] = (double)phase[cell];
    }
    uint16_t *lists = malloc((size_t)N * (2 * PMAX + 2) * sizeof(uint16_t));
    if (!lists) { free(off); free(nbr); return CIPHER_ERR_ARG; }
    int rc = CIPHER_OK;
    for (int cell = 0; cell < N; ++cell) {
        uint16_t gs[2 * PMAX + 2], un[2 * PMAX + 2], iL[2 * PMAX + 2], iR[2 * PMAX + 2];
        gs[0] = 1; gs[1] = (uint16_t)phase[cell];
        for (int k = off[cell]; k < off[cell + 1]; k++) {
            uint16_t nl[2] = {1, (uint16_t)phase[nbr[k]]};
            if (gs[0] + 1 > 2 * PMAX) { rc = CIPHER_ERR_CAPACITY; break; }
            orc_set_union(un, iL, iR, gs, nl);
            memcpy(gs, un, (un[0] + 1) * sizeof(uint16_t));
        }
        if (gs[0] >= L.as) rc = CIPHER_ERR_CAPACITY;
        if (rc != CIPHER_OK) break;
        memcpy(lists + (size_t)cell * (2 * PMAX + 2), gs, sizeof(gs));
    }
    for (int cell = 0; cell < N && rc == CIPHER_OK; ++cell) {
        double *x = X + (size_t)cell * D;
        const uint16_t *gs = lists + (size_t)cell * (2 * PMAX + 2);
        i2f(x + L.aso, gs);
        double *pc = x + L.pfo, *dc = x + L.dpo, *cc = x + L.exo, *tc = x + L.tmo;
        int ph = phase[cell];
        for (int g = 0; g < gs[0]; g++)
            if (gs[g + 1] == ph) *tc = cfg->material[cfg->phase_material[gs[g + 1]]].temperature0;
        for (int g = 0; g < gs[0]; g++) {
            int mat = cfg->phase_material[gs[g + 1]];
            const cipher_material *m = &cfg->material[mat];
            orc_sitepot_explicit(cfg, mat, cc + g * SP, m->c0, *tc);
            if (gs[g + 1] == ph) {
                double spi[2 * CIPHER_MAXCP], spe[2 * CIPHER_MAXCP];
                pc[g] = 1.0;
                orc_sitepot_implicit(cfg, mat, spi, m->c0, *tc);
                orc_sitepot_explicit(cfg, mat, spe, m->c0, *tc);
                for (int c = 0; c < nd; c++) dc[c] = 0.0;
                for (int s = 0; s < m->nsites; s++)
                    for (int c = 0; c < nd; c++) dc[c] += spi[s * nd + c] + spe[s * nd + c];
            } else {
                pc[g] = 0.0;
            }
        }
    }
    free(lists); free(off); free(nbr);
    return rc;
}

/* TSErrorWeightedNorm 2-norm form (PETSc): sqrt(mean(((|u-y|)/tol)^2)),
 * tol = atol + rtol*max(|u|,|y|). */
/* Output fields, src/cipher.c:1005-1080: the interpolants of the cell's
 * active phases, the phase with the largest one (first on ties, `>` against a
 * running maximum from -LARGE, src/typedef.h:25), its material, the
 * interface type of the last pair visited, mu, phi of a phase id, T, and the
 * composition sum_g h_g sum_s nu_s y_{g,s,c} with the site fractions at the
 * site potentials nu_s (mu_c - mu_int_c) - EX_{g,s,c} (src/cipher.c:1038-1066,
 * the same arithmetic as the RHS precompute, src/cipher.c:170-200). */
int orc_output(const cipher_config *cfg, const cipher_mesh *mesh, const double *X, int nfields,
               const cipher_output_field *fields, double *out)
{
    olayout L = orc_layout(cfg);
    const int N = mesh->ncells, D = L.ndof, nc = cfg->ncp, nd = nc - 1, npf = cfg->npf;
    const int SF = L.sf > 0 ? L.sf : 1, SP = L.sp;
    if (L.pf > PMAX || nfields < 0) return CIPHER_ERR_ARG;
    uint16_t sl[PMAX + 1];
    double h[PMAX], sf[PMAX * 2 * CIPHER_MAXCP];
    for (int cell = 0; cell < N; ++cell) {
        const double *x = X + (size_t)cell * D;
        f2i(sl, x + L.aso);
        const double *pc = x + L.pfo, *mu = x + L.dpo, *ex = x + L.exo;
        const double T = x[L.tmo];
        orc_interpolant(cfg->interpolation, h, pc, sl[0]);
        for (int f = 0; f < nfields; f++) {
            double *o = out + (size_t)cell * nfields + f;
            const int idx = fields[f].index;
            double v = 0.0;
            switch (fields[f].kind) {
            case CIPHER_OUT_PHASEID:
            case CIPHER_OUT_MATID: {
                double mx = -1.0e+6;
                for (int g = 0; g < sl[0]; g++)
                    if (h[g] > mx) {
                        mx = h[g];
                        v = fields[f].kind == CIPHER_OUT_PHASEID ? (double)sl[g + 1]
                                                                 : (double)cfg->phase_material[sl[g + 1]];
                    }
                break;
            }
            case CIPHER_OUT_INTERFACEID:
                v = -1.0;
                for (int gk = 0; gk < sl[0]; gk++)
                    for (int gj = gk + 1; gj < sl[0]; gj++)
                        v = (double)cfg->interface_list[sl[gk + 1] * npf + sl[gj + 1]];
                break;
            case CIPHER_OUT_CHEMPOT:
                if (idx >= 0 && idx < nd) v = mu[idx];
                break;
            case CIPHER_OUT_COMP: {
                if (idx < 0 || idx >= nc) break;
                double mui[CIPHER_MAXCP], spim[2 * CIPHER_MAXCP], avg[CIPHER_MAXCP];
                for (int c = 0; c < nd; c++) mui[c] = 0.0;
                for (int gk = 0; gk < sl[0]; gk++)
                    for (int gj = gk + 1; gj < sl[0]; gj++) {
                        const cipher_interface *it = &cfg->interface[cfg->interface_list[sl[gk + 1] * npf + sl[gj + 1]]];
                        if (it->has_potential)
                            for (int c = 0; c < nd; c++) mui[c] += h[gk] * h[gj] * it->potential[c];
                    }
                for (int c = 0; c < nc; c++) avg[c] = 0.0;
                for (int g = 0; g < sl[0]; g++) {
                    const int mat = cfg->phase_material[sl[g + 1]];
                    const cipher_material *m = &cfg->material[mat];
                    for (int s = 0; s < m->nsites; s++)
                        for (int c = 0; c < nd; c++)
                            spim[s * nd + c] = m->stoichiometry[s] * (mu[c] - mui[c]) - ex[g * SP + s * nd + c];
                    orc_sitefrac(cfg, mat, sf + g * SF, spim, T);
                    for (int s = 0; s < m->nsites; s++)
                        for (int c = 0; c < nc; c++) avg[c] += h[g] * m->stoichiometry[s] * sf[g * SF + s * nc + c];
                }
                v = avg[idx];
                break;
            }
            case CIPHER_OUT_PHI:
                for (int g = 0; g < sl[0]; g++)
                    if (sl[g + 1] == idx) v = pc[g];
                break;
            case CIPHER_OUT_TEMPERATURE:
                v = T;
                break;
            default:
                return CIPHER_ERR_ARG;
            }
            *o = v;
        }
    }
    return CIPHER_OK;
}

double orc_wrms_norm(const double *U, const double *Y, int64_t n, double atol, double rtol)
{
    double sum = 0.0;
    for (int64_t i = 0; i < n; i++) {
        double au = fabs(U[i]), ay = fabs(Y[i]);
        double tol = atol + rtol * (au > ay ? au : ay);
        double e = fabs(U[i] - Y[i]) / tol;
        sum += e * e;
    }
    return n > 0 ? sqrt(sum / (double)n) : 0.0;
}
