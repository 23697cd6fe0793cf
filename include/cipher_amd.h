/*
 * cipher_amd.h - C ABI of the MI355X phase-field right-hand side.
 *
 * This library is a drop-in for the per-stage right-hand side of CIPHER
 * (Calphad Integrated PHase-field solvER).  Every entry point takes plain
 * pointers and sizes; state vectors use CIPHER's per-cell layout
 *
 *     [ AS | PF | DP | EX | TM ]   (doubles, D = AS+PF+DP+EX+TM per cell)
 *
 *   AS[0]      number n of active phases, AS[1..n] ascending phase ids
 *              (integers stored as doubles, rounded on read)
 *   PF[g]      phase fraction of phase AS[g+1]
 *   DP[c]      chemical potential of component c < ncp-1
 *   EX[g*SP + s*ndp + c]  explicit (non-ideal) site potential
 *   TM         temperature
 *
 * with the sizes of src/init.c:1330-1343.  Pointers named *_dev are device
 * pointers (HBM); every other pointer is host memory.  Calls on one context
 * are ordered on that context's stream.  Every function returns a
 * cipher_status; the library never aborts the caller.
 *
 * Reference interfaces replaced (file:line in micmog/CIPHER):
 *   cipher_rhs            RHSFunction(TS,PetscReal,Vec,Vec,void*)   src/cipher.c:20,27-681
 *   cipher_poststep       PostStep(TS), active-set block            src/cipher.c:21,869-974
 *   cipher_setup_problem  SetUpProblem(Vec,AppCtx*)                 src/init.c:1721-1830
 *   cipher_wrms_norm      TSErrorWeightedNorm (PETSc, configured at src/cipher.c:1112)
 *   cipher_rk_*           TSRK stage arithmetic (PETSc, configured at src/cipher.c:1104)
 *   cipher_config         AppCtx material/interface tables          src/typedef.h:145-319
 *   cipher_mesh           flattened cells/faces/cellgeom            src/cipher.c:1187-1231
 */
#ifndef CIPHER_AMD_H
#define CIPHER_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Constants of src/typedef.h:25-30. */
#define CIPHER_LARGE 1.0e+6
#define CIPHER_TOL 1.0e-6
#define CIPHER_MAXCP 5
#define CIPHER_MAXAP 15
#define CIPHER_R_GAS_CONST 8.314462
#define CIPHER_MAXTSER 10

typedef enum {
    CIPHER_OK = 0,
    CIPHER_ERR_ARG = 1,         /* bad argument / unsupported combination */
    CIPHER_ERR_HIP = 2,         /* HIP runtime error */
    CIPHER_ERR_CAPACITY = 3,    /* active-set union exceeded capacity (src/cipher.c:968) */
    CIPHER_ERR_NONFINITE = 4,   /* NaN/Inf in an error norm */
    CIPHER_ERR_UNSUPPORTED = 5, /* feature not available on this path */
    CIPHER_ERR_STATE = 6        /* call out of order (e.g. no mesh uploaded) */
} cipher_status;

/* chemfe_model_t, src/typedef.h:54-59 (same numbering) */
enum { CIPHER_CHEM_QUADRATIC = 0, CIPHER_CHEM_CALPHADDIS = 1,
       CIPHER_CHEM_CALPHAD2SL = 2, CIPHER_CHEM_NONE = 3 };
/* interpolation_t, src/typedef.h:69-74 */
enum { CIPHER_INTERP_LINEAR = 0, CIPHER_INTERP_QUADRATIC = 1,
       CIPHER_INTERP_CUBIC = 2 };
/* boundary_t, src/typedef.h:62-66 */
enum { CIPHER_BC_NEUMANN = 0, CIPHER_BC_DIRICHLET = 1, CIPHER_BC_NONE = 2 };
/* source_model_t, src/typedef.h:40-43 */
enum { CIPHER_SOURCE_SINK = 0, CIPHER_SOURCE_RANDOM = 1 };
/* arithmetic of the right-hand side */
enum { CIPHER_FP64 = 0, CIPHER_FP32_MIXED = 1 };

/* TSeries, src/typedef.h:145-150: sum_i coeff[i]*T^exp[i] + logcoeff*T*ln(T) */
typedef struct cipher_tseries {
    double coeff[CIPHER_MAXTSER];
    double logcoeff;
    int32_t exp[CIPHER_MAXTSER];
    int32_t n;
    int32_t pad_;
} cipher_tseries;

/* RK, src/typedef.h:153-156: n Redlich-Kister orders, enthalpy series
 * tseries[ts .. ts+n-1], ternary component index rkindex[idx .. idx+n-1]. */
typedef struct cipher_rk {
    int32_t n;
    int32_t ts;
    int32_t idx;
    int32_t pad_;
} cipher_rk;

/* TACTIVATIONPROP of a component mobility, src/typedef.h:159-163:
 * unary: ncp series from tseries[unary]; binary: ncp*(ncp-1)/2 RK from rk[binary]. */
typedef struct cipher_mobility {
    double m0;
    int32_t unary;
    int32_t binary;
} cipher_mobility;

/* SOURCE, src/typedef.h:77-103 */
typedef struct cipher_source {
    double rate[CIPHER_MAXCP];
    double ceq[CIPHER_MAXCP];
    double amplitude[CIPHER_MAXCP];
    int32_t model;
    int32_t pad_;
} cipher_source;

/* MATERIAL, src/typedef.h:210-218 with the CHEMFE union of :166-201 flattened
 * into table offsets.  Table order is the reference's loop order. */
typedef struct cipher_material {
    double molarvolume;
    double chempot_ex_kineticcoeff;
    double c0[2 * CIPHER_MAXCP];        /* calphad2sl: c0_p then c0_q */
    double stoichiometry[2];            /* site multiplicities (p, q) */
    double temperature0;
    int32_t model;                      /* CIPHER_CHEM_* */
    int32_t nsites;                     /* 0 none, 1 dis/quad, 2 calphad2sl */
    int32_t ref;                        /* tseries: ref enthalpy (none/dis/quad) */
    int32_t unary;                      /* tseries: ncp (dis/quad) or ncp*ncp (2sl, [cp*ncp+cq]) */
    int32_t binary;                     /* rk: dis binary ncp(ncp-1)/2; 2sl binaryp ncp*ncp(ncp-1)/2 */
    int32_t ternary;                    /* rk: dis ternary C(ncp,3); 2sl ternaryp C(ncp,3)*ncp */
    int32_t binaryq;                    /* rk: 2sl binaryq ncp*ncp(ncp-1)/2 */
    int32_t ternaryq;                   /* rk: 2sl ternaryq ncp*C(ncp,3) */
    int32_t quad_binary;                /* tseries: quadratic binary, ncp */
    int32_t quad_ceq;                   /* tseries: quadratic ceq, ncp */
    int32_t mobility;                   /* first of ncp cipher_mobility */
    int32_t specific_heat;              /* tseries (n==0: absent) */
    int32_t latent_heat;                /* tseries (n==0: absent) */
    int32_t tconductivity;              /* tseries (n==0: absent) */
    int32_t nsources;
    int32_t source;                     /* first cipher_source */
} cipher_material;

/* INTERFACE, src/typedef.h:221-240 */
typedef struct cipher_interface {
    double width;
    double e0;                          /* ienergy->e->m0 */
    double m0;                          /* imobility->m->m0 */
    double potential[CIPHER_MAXCP];     /* segregation potential (ndp) */
    double mobilityc[CIPHER_MAXCP];     /* interface solute mobility (ncp) */
    int32_t e_act;                      /* tseries of the energy activation */
    int32_t m_act;                      /* tseries of the mobility activation */
    int32_t e_naniso, e_aniso;          /* anisotropy: aniso[e_aniso..] = dim*n dirs then n vals */
    int32_t m_naniso, m_aniso;
    int32_t has_potential;
    int32_t has_mobilityc;
    int32_t nsources;
    int32_t source;
} cipher_interface;

/* BOUNDARYCONDITIONS, src/typedef.h:243-248 */
typedef struct cipher_bc {
    double chem_val[CIPHER_MAXCP];
    double thermal_val;
    int32_t boundary_id;                /* 1..2*dim, src/cipher.c:1214-1220 */
    int32_t chem_type;                  /* CIPHER_BC_* */
    int32_t thermal_type;
    int32_t chem_mask;                  /* bit c set: component c constrained */
} cipher_bc;

/* The AppCtx fields the right-hand side reads (src/typedef.h:277-319). */
typedef struct cipher_config {
    const cipher_tseries *tseries;
    const cipher_rk *rk;
    const int32_t *rkindex;
    const cipher_mobility *mobility;
    const cipher_material *material;
    const cipher_interface *interface;
    const cipher_source *source;
    const cipher_bc *bc;
    const double *aniso;
    const int32_t *phase_material;      /* npf */
    const int32_t *interface_list;      /* npf*npf, upper triangle used */
    double junctionpenalty;
    double temperature_rate;            /* current load case */
    int32_t has_temperature_rate;
    int32_t dim, npf, ncp, nmat, nf, nbcs;
    int32_t ntseries, nrk, nrkindex, nmobility, nsource, naniso;
    int32_t max_phase_neighbours;       /* default CIPHER_MAXAP */
    int32_t interpolation;              /* CIPHER_INTERP_* */
    int32_t gradient_calculation;       /* must be 0 on the GPU path */
} cipher_config;

/* Field sizes/offsets of src/init.c:1330-1343. */
typedef struct cipher_layout {
    int32_t maxsites, sp_size, sf_size;
    int32_t as_size, pf_size, dp_size, ex_size, tm_size;
    int32_t as_offset, pf_offset, dp_offset, ex_offset, tm_offset;
    int32_t ndof;                       /* dofs per cell */
} cipher_layout;

/* Flattened leaf mesh after an adapt (src/cipher.c:1187-1231).  Cells are the
 * owned leaves; faces are the interior faces with two supports and no tree
 * children, in the order the right-hand side visits them; (L,R) is the support
 * order.  Boundary faces carry the label id of src/cipher.c:1214-1220. */
typedef struct cipher_mesh {
    const int32_t *face_cells;          /* 2*nfaces */
    const double *cellgeom;             /* ncells: pow(volume, 1/dim) */
    const int32_t *bface_cell;          /* nbfaces */
    const int32_t *bface_id;            /* nbfaces */
    int32_t ncells;
    int32_t nfaces;
    int32_t nbfaces;
    int32_t pad_;
    /* Cell centroids, ncells*dim (x fastest per cell), for the least-squares
     * gradients of gradient_calculation (src/cipher.c:1233-1282); may be NULL
     * otherwise.  period[d] > 0: the domain length of a periodic direction,
     * a neighbour's offset is taken as the nearest periodic image. */
    const double *centroid;
    double period[3];
} cipher_mesh;

typedef struct cipher_ctx cipher_ctx;

/* Layout of a config (no device work). */
int cipher_layout_of(const cipher_config *cfg, cipher_layout *out);

/* Create a context on the current HIP device: copies the config tables to HBM.
 * precision is CIPHER_FP64 or CIPHER_FP32_MIXED. */
int cipher_ctx_create(const cipher_config *cfg, int precision, cipher_ctx **out);
int cipher_ctx_destroy(cipher_ctx *ctx);
int cipher_ctx_layout(const cipher_ctx *ctx, cipher_layout *out);
int cipher_set_precision(cipher_ctx *ctx, int precision);
/* hipStream_t passed as void*; NULL selects the default stream. */
int cipher_set_stream(cipher_ctx *ctx, void *stream);
/* Update the load-case temperature rate (src/cipher.c:568). */
int cipher_set_temperature_rate(cipher_ctx *ctx, int has_rate, double rate);

/* Upload a flattened mesh; builds the per-cell face lists (and, with
 * gradient_calculation, each cell's least-squares gradient weights over its
 * face neighbours) and preallocates every scratch buffer so that later calls
 * allocate nothing. */
int cipher_upload_mesh(cipher_ctx *ctx, const cipher_mesh *mesh);

/* F = RHS(t, X): RHSFunction, src/cipher.c:27-681.  X_dev and F_dev hold
 * ncells*ndof doubles; F is fully defined (AS and inactive slots are 0). */
int cipher_rhs(cipher_ctx *ctx, double t, const double *X_dev, double *F_dev);

/* Same as cipher_rhs and also writes, per cell, bit g of active_dev[cell]
 * = active[g] of src/cipher.c:535-537 (0 for single-phase cells). */
int cipher_rhs_active(cipher_ctx *ctx, double t, const double *X_dev,
                      double *F_dev, uint32_t *active_dev);

/* Active-set update of PostStep, src/cipher.c:869-974, in place on X_dev.
 * Returns CIPHER_ERR_CAPACITY if a union exceeded PF_SIZE phases (the
 * reference asserts, src/cipher.c:968); X_dev is then unspecified. */
int cipher_poststep(cipher_ctx *ctx, double *X_dev);

/* Initial state of SetUpProblem, src/init.c:1721-1830: phase_dev[cell] is the
 * cell's phase label (0-based). */
int cipher_setup_problem(cipher_ctx *ctx, const int32_t *phase_dev, double *X_dev);

/* Weighted RMS error norm sqrt(1/N sum ((|u-y|)/(atol+rtol*max(|u|,|y|)))^2)
 * over n doubles, written to *norm_out (host). */
int cipher_wrms_norm(cipher_ctx *ctx, const double *U_dev, const double *Y_dev,
                     int64_t n, double atol, double rtol, double *norm_out);

/* No counterpart in the reference.  cipher_wrms_norm without a context, for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): same kernels, partition and
 * result bits, with the partial sums in a static device array of the library
 * (1 056 doubles per GPU, held until the process ends; the fixed partition
 * uses at most 1 024 partial sums, so every n fits).  Allocates nothing.
 * stream: a hipStream_t as void* (NULL: the default stream); synchronises
 * it.  Calls that overlap on one GPU share the array and race. */
int cipher_wrms_norm_noalloc(const double *U_dev, const double *Y_dev, int64_t n, double atol,
                             double rtol, void *stream, double *norm_out);

/* Y = X + sum_j a[j]*K_j (j < nk), the RK stage / update combination, FP64.
 * K_dev points to nk device vectors of n doubles each. */
int cipher_axpy_n(cipher_ctx *ctx, int64_t n, double *Y_dev, const double *X_dev,
                  int nk, const double *a, const double *const *K_dev);

/* ---- adaptive explicit Runge-Kutta integration on device vectors ----------
 * The integrator CIPHER configures in PETSc (TSRK, default type 3bs =
 * Bogacki-Shampine 3(2); TSADAPTDSP; TSSetTolerances; TSAdaptSetStepLimits;
 * src/cipher.c:1097-1115) and its time loop with an exact final time
 * (TS_EXACTFINALTIME_MATCHSTEP, src/cipher.c:1589-1606), with PostStep's
 * active-set update after every accepted step.  The stage values use
 * cipher_rhs, the error is cipher_wrms_norm of the third- against the
 * embedded second-order solution. */
typedef struct cipher_ts_params {
    double atol, rtol;          /* TSSetTolerances, src/cipher.c:1112 (abstol/reltol, src/init.c:1286-1291) */
    double dt_min, dt_max;      /* TSAdaptSetStepLimits, src/cipher.c:1111 (timestepmin/max) */
    double safety;              /* step-ratio safety factor (0.9) */
    double reject_safety;       /* extra factor after consecutive rejections (0.5) */
    double clip_lo, clip_hi;    /* bounds of the step ratio (0.1, 10) */
    int32_t use_graph;          /* replay each attempt from a captured HIP graph */
    int32_t method;             /* CIPHER_RK_3BS (0) or CIPHER_RK_2A: TSRK type, src/cipher.c:1104
                                   (petscoptions -ts_rk_type) */
} cipher_ts_params;

/* TSRK types (PETSc's names): Bogacki-Shampine 3(2), four stages with the
 * last one evaluated at the new solution; Heun's method 2(1), two stages,
 * the error estimate against the explicit Euler step of the first stage. */
enum { CIPHER_RK_3BS = 0, CIPHER_RK_2A = 1 };

/* Controller memory carried between steps (zero-initialise before a run). */
typedef struct cipher_ts_ctrl {
    double err_prev;            /* error norm of the last accepted step, 0 if none */
    double safety;              /* current safety factor, 0 = params.safety */
    int32_t rejected_last;      /* the previous attempt was rejected */
    int32_t pad_;
} cipher_ts_ctrl;

typedef struct cipher_ts_stats {
    int64_t steps;              /* accepted steps */
    int64_t rejected;           /* rejected attempts */
    int64_t rhs_evals;
    double t;                   /* time reached */
    double dt;                  /* next step size */
    double err;                 /* error norm of the last attempt */
    int64_t host_syncs;         /* times the host waited for the device */
} cipher_ts_stats;

/* Host-only step-size decision for an attempt of size dt with error norm err
 * (no device work): *accept = err <= 1 or dt at dt_min; *dt_next from the
 * PI controller with an arctan limiter, clipped and clamped to the limits.
 * CIPHER_ERR_NONFINITE for a non-finite err at the minimum step.  The
 * device-driven solves take the same decisions on the device, in the same
 * IEEE operations (identical bits). */
int cipher_ts_adapt(const cipher_ts_params *p, cipher_ts_ctrl *st, double dt, double err,
                    int *accept, double *dt_next);

/* Allocate the stage vectors for the uploaded mesh and reset the controller.
 * Call again after cipher_upload_mesh. */
int cipher_ts_setup(cipher_ctx *ctx, const cipher_ts_params *p);

/* One attempt of size dt from X_dev (unchanged): four right-hand sides, the
 * new and embedded solutions, the error norm (host, synchronises). */
int cipher_ts_attempt(cipher_ctx *ctx, const double *X_dev, double dt, double *err_out);

/* Copy the last attempt's solution into X_dev and run cipher_poststep on it. */
int cipher_ts_accept(cipher_ctx *ctx, double *X_dev);

/* Integrate X_dev from t0 to t_end (exact final time) or for max_steps
 * accepted steps; *dt_inout: first step size in, next step size out.  The
 * step decisions run on the device (after each attempt a one-thread kernel
 * accepts or rejects it, updates the time and the step size; the copy and
 * PostStep of an accepted step are predicated on its decision), so the
 * host enqueues several attempts per synchronisation (up to 16, fewer
 * near the end time); attempts enqueued after the end do no work. */
int cipher_ts_solve(cipher_ctx *ctx, double *X_dev, double t0, double t_end, double *dt_inout,
                    int64_t max_steps, cipher_ts_stats *stats);

/* The load-case loop of CIPHER's main (src/cipher.c:1589-1606; cases from
 * solution_parameters/time and temperature_rate, src/init.c:1248-1256): for
 * load case k = 0..ncases-1, set the temperature rate rates[k] (rates NULL:
 * no temperature rate in any case, the reference's NULL array), restart from
 * the step size params.dt_min (timestepmin) and integrate with an exact final
 * time to T_k = (...((t0 + times[0]) + times[1]) ...) + times[k], the
 * reference's running sum.  The step controller's memory carries over from
 * case to case as it does between cipher_ts_solve calls.  stats[k] receives
 * case k's counts and the time reached.  Leaves the last case's temperature
 * rate set. */
int cipher_ts_loadcases(cipher_ctx *ctx, double *X_dev, int32_t ncases, const double *times,
                        const double *rates, double t0, cipher_ts_stats *stats);

/* ---- derived output fields ---------------------------------------------
 * The per-cell fields CIPHER writes at an output step (src/cipher.c:1005-1080;
 * field names of the YAML `outputs` list: "phaseid", "matid", "interfaceid",
 * "<component>_chempot", "<component>_c", "<phase id>_phi", "temperature"). */
typedef enum {
    CIPHER_OUT_PHASEID = 1,     /* id of the phase with the largest interpolant */
    CIPHER_OUT_MATID = 2,       /* its material */
    CIPHER_OUT_INTERFACEID = 3, /* interface type of the last active pair, -1 for one phase */
    CIPHER_OUT_CHEMPOT = 4,     /* chemical potential of component index (< ncp-1) */
    CIPHER_OUT_COMP = 5,        /* composition of component index (< ncp) */
    CIPHER_OUT_PHI = 6,         /* phase fraction of phase id index, 0 if inactive */
    CIPHER_OUT_TEMPERATURE = 7
} cipher_output_kind;

typedef struct cipher_output_field {
    int32_t kind, index;
} cipher_output_field;

#define CIPHER_MAX_OUTPUTS 32

/* out_dev[cell * nfields + f] = field f of cell (nfields <= CIPHER_MAX_OUTPUTS);
 * a component index out of range gives 0, as an unmatched name leaves the
 * reference's output untouched. */
int cipher_output(cipher_ctx *ctx, const double *X_dev, int32_t nfields, const cipher_output_field *fields,
                  double *out_dev);

/* ---- launch shapes ------------------------------------------------------- */
/* Shape of one of the five per-size cell lists (list q holds the cells with at
 * most 2^q active phases; all lists run in one launch of 256-thread blocks,
 * one block range per list): lanes per cell (0 = 2^q; else a power of two
 * between 2^q and the capacity's lane count, at most 16), the list's block
 * cap (its range strides beyond it), and map: bit 0 the block order (0 as
 * launched, 1 one contiguous part of the list per XCD), bit 1 (value 2; list
 * 0 of a configuration with chemistry, one lane per cell) the single-phase
 * cells in a launch of their own after the merged launch of the other lists,
 * compiled for more resident waves.  Shapes change timing only, never
 * results.  No counterpart in the reference. */
typedef struct cipher_list_shape {
    int32_t lanes, cap, map, pad_;
} cipher_list_shape;

int cipher_get_shape(const cipher_ctx *ctx, int32_t q, cipher_list_shape *out);
int cipher_set_shape(cipher_ctx *ctx, int32_t q, const cipher_list_shape *shape);

/* Pick every list's shape by timing candidates on the state X_dev (F_dev is
 * scratch of the same size): coordinate descent, each candidate timed as the
 * median of reps right-hand sides.  *ms_out (may be NULL): the best time of
 * one right-hand side in the current precision. */
int cipher_autotune(cipher_ctx *ctx, const double *X_dev, double *F_dev, int32_t reps, double *ms_out);

/* ---- batched ensembles ---------------------------------------------------
 * No counterpart in the reference, which integrates one microstructure per
 * process.  A batch runs many independent contexts (each with its own
 * configuration and uploaded mesh: different sizes, component counts and
 * capacities) through shared launches: members with the same component count
 * and lane capacity form a class, and each kernel is one launch per class and
 * cell list, its grid cut into one segment per member.  Each member's results
 * are those of its context run alone.  Batch state vectors hold the members'
 * records back to back; member i starts at offsets[i] doubles.  A batch
 * borrows its members' mesh and scratch buffers: while it exists, do not run
 * the members' own RHS/PostStep/integrator or re-upload their meshes. */
typedef struct cipher_batch cipher_batch;

int cipher_batch_create(cipher_ctx *const *members, int32_t nmembers, cipher_batch **out);
int cipher_batch_destroy(cipher_batch *b);
/* offsets[nmembers+1]: first double of each member in a batch vector;
 * cell_offsets[nmembers+1] (may be NULL): first cell, for per-cell arrays. */
int cipher_batch_offsets(const cipher_batch *b, int64_t *offsets, int64_t *cell_offsets);
int cipher_batch_set_precision(cipher_batch *b, int precision);
int cipher_batch_set_stream(cipher_batch *b, void *stream);
/* Member i's load-case temperature rate (cipher_set_temperature_rate on its
 * context, src/cipher.c:568, and in the batch's member table), for load
 * cases run through the batch.  Synchronises the batch's stream. */
int cipher_batch_set_temperature_rate(cipher_batch *b, int32_t member, int has_rate, double rate);

/* Every member's cipher_rhs (cipher_rhs_active) in one batched launch set;
 * active_dev is indexed by batch cell (cell_offsets). */
int cipher_batch_rhs(cipher_batch *b, double t, const double *X_dev, double *F_dev);
int cipher_batch_rhs_active(cipher_batch *b, double t, const double *X_dev, double *F_dev,
                            uint32_t *active_dev);

/* Every member's cipher_poststep; on CIPHER_ERR_CAPACITY *failed_member (may
 * be NULL) names the first member whose union overflowed, else -1. */
int cipher_batch_poststep(cipher_batch *b, double *X_dev, int32_t *failed_member);

/* norms[i] (host) = cipher_wrms_norm of member i's slices of U and Y. */
int cipher_batch_wrms_norm(cipher_batch *b, const double *U_dev, const double *Y_dev,
                           double atol, double rtol, double *norms);

/* Pick the launch shape (block size, grid cap) of every class's cell lists
 * by timing candidates on the state X_dev (F_dev is scratch of the same
 * size): coordinate descent, median of reps batched right-hand sides per
 * candidate.  Results never depend on the shape.  *ms_out (may be NULL): the
 * best time of one batched right-hand side. */
int cipher_batch_autotune(cipher_batch *b, const double *X_dev, double *F_dev, int32_t reps,
                          double *ms_out);
/* Report class k's component count and lane capacity and the shape of its
 * list q. */
int cipher_batch_shape(const cipher_batch *b, int32_t k, int32_t q, int32_t *ncp, int32_t *capacity_lanes,
                       cipher_list_shape *shape);

/* Per-member adaptive integration: cipher_ts_solve for every member at once,
 * each with its own step size, controller and end time; an attempt advances
 * all members that have not reached their end time (or max_steps accepted
 * steps) and is replayed from one captured HIP graph.  t0, t_end, dt_inout
 * and stats have nmembers entries.  Each member's controller memory carries
 * over from call to call, as a context's does; cipher_batch_ts_setup clears
 * it. */
int cipher_batch_ts_setup(cipher_batch *b, const cipher_ts_params *p);
int cipher_batch_ts_solve(cipher_batch *b, double *X_dev, const double *t0, const double *t_end,
                          double *dt_inout, int64_t max_steps, cipher_ts_stats *stats);

/* Message of the last failing call on this thread. */
const char *cipher_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* CIPHER_AMD_H */
