/*
 * oracle.h - CPU restatement of CIPHER's phase-field right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; nothing in the product library links or calls it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Plain C99, FP64, the reference's operation order.  Each function cites the
 * reference file:line it restates (paths relative to micmog/CIPHER).
 */
#ifndef CIPHER_ORACLE_H
#define CIPHER_ORACLE_H

#include <stdint.h>
#include "../include/cipher_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* utility.c */
double orc_fastpow(double base, unsigned expn);
double orc_sum_tseries(double T, const cipher_tseries *ts);
double orc_sum_tseries_derivative(double T, const cipher_tseries *ts);
void orc_set_intersection(uint16_t *outc, uint16_t *cina, uint16_t *cinb,
                          const uint16_t *ina, const uint16_t *inb);
void orc_set_union(uint16_t *outc, uint16_t *ainc, uint16_t *binc,
                   const uint16_t *ina, const uint16_t *inb);
void orc_invert_matrix(double *dst, double *src, int n);
void orc_interpolant(int interp, double *h, const double *phi, int n);
void orc_interpolant_derivative(int interp, double *in, const double *phi, int n);
void orc_simplex_projection(double *out, const double *in, int size);

/* material.c: phase-level entry points on material index mat */
void orc_sitefrac(const cipher_config *cfg, int mat, double *y, const double *sitepot_im, double T);
void orc_sitefrac_tangent(const cipher_config *cfg, int mat, double *tpot, double *tex, double *tT,
                          const double *y, const double *sitepot_im, double T);
void orc_mobility_component(const cipher_config *cfg, int mat, double *mob, const double *y, double T);
double orc_chemenergy_phase(const cipher_config *cfg, int mat, const double *y, double T);
void orc_sitepot_explicit(const cipher_config *cfg, int mat, double *sp, const double *y, double T);
void orc_sitepot_implicit(const cipher_config *cfg, int mat, double *sp, const double *y, double T);
void orc_mobility_volumeref(const cipher_config *cfg, double *out, const double *mob, const double *comp);

/* RHSFunction, src/cipher.c:27-681.  X, F: ncells*ndof doubles.  active
 * (optional): bit g of active[cell] = active[g]. */
/* least-squares phase-field gradients, src/cipher.c:82-146 and 1233-1282:
 * grad[(cell*16 + g)*dim + d] for the cell's active slots g */
int orc_gradients(const cipher_config *cfg, const cipher_mesh *mesh, const double *X, double *grad);

int orc_rhs(const cipher_config *cfg, const cipher_mesh *mesh, double t,
            const double *X, double *F, uint32_t *active);

/* PostStep active-set block, src/cipher.c:869-974, in place. Returns
 * CIPHER_ERR_CAPACITY where the reference asserts. */
int orc_poststep(const cipher_config *cfg, const cipher_mesh *mesh, double *X);

/* SetUpProblem, src/init.c:1721-1830. X is zeroed first. */
int orc_setup_problem(const cipher_config *cfg, const cipher_mesh *mesh,
                      const int32_t *phase, double *X);

/* Output fields of src/cipher.c:1005-1080: out[cell*nfields + f]. */
int orc_output(const cipher_config *cfg, const cipher_mesh *mesh, const double *X, int nfields,
               const cipher_output_field *fields, double *out);

/* WRMS error norm of PETSc's TSErrorWeightedNorm (2-norm form). */
double orc_wrms_norm(const double *U, const double *Y, int64_t n, double atol, double rtol);

#ifdef __cplusplus
}
#endif
#endif
