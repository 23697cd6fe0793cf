// kernels.h - launch interfaces shared by the host API and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "device.h"

namespace cipher {

struct RhsArgs {
    const double *X;
    double *F;
    uint32_t *active;
    const double *cellgeom;
    const int32_t *cell_off;  // ncells+1: faces of each cell in face order
    const int32_t *adj;       // (neighbour << 1) | (self is R)
    const int32_t *bcell_off; // ncells+1: boundary faces of each cell
    const int32_t *bface_id;  // boundary label id per boundary-face entry
    const double *cm;         // per-cell [comp(ncp), mob(ncp), tcond]
    const int32_t *cells;     // cells to process and their (device) count
    const int32_t *cells_count;
    int32_t *redo;            // FP32-mixed: cells recomputed in FP64
    int32_t *redo_count;
    long long ncells;
};

struct RhsLaunch {
    const DevCfg *cfg;
    const double *X;
    double *F;
    uint32_t *active;
    const double *cellgeom;
    const int32_t *cell_off, *adj, *bcell_off, *bface_id;
    void *cm;
    int32_t *bins;            // 5 lists of ncells entries (cells by active-set size)
    int32_t *counts;          // 8 counters: bin sizes, then the recomputation list
    int32_t *redo;            // ncells entries
    hipStream_t aux[4];       // forked streams for the concurrent lists
    hipEvent_t ev_fork, ev_join[4];
    long long ncells;
};

cipher_status launch_rhs(const RhsLaunch &L, int precision, hipStream_t st);
size_t prep_record_bytes(int ncp);
cipher_status launch_tsfold(const double *X, long long ncells, const DevCfg &c, double *tsv, hipStream_t st);

struct PostLaunch {
    const DevCfg *cfg;
    double *X;
    const int32_t *cell_off, *adj;
    int32_t *glist;           // ncells*(P+1) scratch
    int32_t *err;             // device flag (0 ok, 1 capacity)
    const int32_t *phase;     // setup only
    long long ncells;
};
cipher_status launch_poststep(const PostLaunch &L, hipStream_t st);
cipher_status launch_setup(const PostLaunch &L, hipStream_t st);

cipher_status launch_wrms(const double *U, const double *Y, long long n, double atol, double rtol,
                          double *partial, int npart, double *out, hipStream_t st);
cipher_status launch_axpy_n(long long n, double *Y, const double *X, int nk, const double *a,
                            const double *const *K, hipStream_t st);
cipher_status launch_axpy_dt(long long n, double *Y, const double *X, const double *dt_dev, int nk, const double *A,
                             const double *const *K, hipStream_t st);

}  // namespace cipher
