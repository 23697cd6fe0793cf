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
    const double *fgeo;       // per adj entry: the face's flux factor 2*area/(hL + hR)
    const int32_t *bcell_off; // ncells+1: boundary faces of each cell
    const int32_t *bface_id;  // boundary label id per boundary-face entry
    const double *cm;         // per-cell [comp(ncp), mob(ncp), tcond]
    const int32_t *cells;     // cells to process and their (device) count
    const int32_t *cells_count;
    int32_t *redo;            // FP32-mixed: cells recomputed in FP64
    int32_t *redo_count;
    long long ncells;
    int32_t map;              // block order: 0 as launched, 1 one contiguous range per XCD
    int32_t pad_;
    const int32_t *pred;      // non-null: the launch does nothing while *pred == 0
};

// Launch shape of one per-size cell list: lanes per cell (at least the
// list's active-set bound; more lanes take more faces per step), grid cap
// and block order.  Shapes change timing only, never results.
// map: bit 0 the block order (RhsArgs::map); SHAPE_OWN (list 0 of a mesh
// with chemistry, one lane per cell): the single-phase cells run in a launch
// of their own, compiled for more resident waves than the merged launch of
// the multi-phase lists allows
constexpr int SHAPE_OWN = 2;
struct ListShape {
    int lanes = 0;            // 0: 2^q for list q
    int cap = 16384;          // blocks of the list's range
    int map = 0;              // block order (bit 0) and SHAPE_OWN
};
// the single-phase list runs in its own launch
static inline bool own_launch(const ListShape &s0, int ncp) {
    return ncp > 1 && (s0.map & SHAPE_OWN) && s0.lanes <= 1;
}

// per-mesh list counters: 0..4 the per-size lists, 8..12 their FP32-mixed
// recomputation lists
constexpr int RHS_COUNTERS = 16;
// counter sets that eager right-hand sides take in turn between clears
constexpr int RHS_RING = 64;

// blocks per list of the FP32-mixed recomputation launch
constexpr int RHS_REDO_CAP = 512;

// cells per binning tile: 256 threads of BIN_CPT consecutive cells
constexpr int BIN_CPT = 8;
constexpr int BIN_TILE = 256 * BIN_CPT;

// LDS bytes of a k_rhs block of cpb cells of G lanes: the interface type of
// every slot pair, and energy, width and mobility of every pair gk < gj
static inline size_t rhs_lds_bytes(const DevCfg &c, int G, int cpb) {
    const int np = G > 1 ? G * (G - 1) / 2 : 1;
    return (size_t)cpb * ((size_t)G * G * sizeof(int) + 3 * (size_t)np * sizeof(double));
}

// Block ranges of one merged launch over the per-size lists: range r covers
// blocks [start[r], start[r+1]) and runs list list[r] with lanes[r] lanes per
// cell and block order map[r]; with redo set, the ranges run the lists'
// FP32-mixed recomputation lists instead.
struct ListRanges {
    int32_t start[6];
    int32_t list[5];
    int32_t lanes[5];
    int32_t map[5];
    int32_t n;
    int32_t redo;
};

struct RhsLaunch {
    const DevCfg *cfg;
    const DevCfg *dcfg;       // device copy of *cfg (read by the list and prep kernels)
    const double *X;
    double *F;
    uint32_t *active;
    const double *cellgeom;
    const int32_t *cell_off, *adj, *bcell_off, *bface_id;
    const double *fgeo;
    void *cm;
    int32_t *bins;            // 5 lists of ncells entries (cells by active-set size)
    int32_t *counts;          // RHS_COUNTERS counters
    int32_t *fill;            // cleared before binning: nfill ints from fill
    int nfill;
    int32_t *redo;            // 5 recomputation lists of ncells entries
    long long ncells;
    ListShape shape[5];
    // 0 or 1: bin the cells; 2 or 3: reuse the lists of the previous call,
    // whose state had the same active sets (stages of one attempt)
    int bins_mode;
    const int32_t *pred = nullptr;  // device flag: the launches do nothing while it is 0
};

cipher_status launch_rhs(const RhsLaunch &L, int precision, hipStream_t st);

// derived output fields (kind/index pairs of cipher_output_field)
struct OutFields {
    int32_t kind[CIPHER_MAX_OUTPUTS];
    int32_t index[CIPHER_MAX_OUTPUTS];
    int32_t n;
    int32_t need_comp;        // some field is a composition (the prep records are needed)
};
cipher_status launch_output(const DevCfg &c, const double *X, double *cm, long long ncells, const OutFields &f,
                            double *out, hipStream_t st);
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
    const int32_t *pred = nullptr;  // device flag: the update does nothing while it is 0
};
cipher_status launch_poststep(const PostLaunch &L, hipStream_t st);
cipher_status launch_setup(const PostLaunch &L, hipStream_t st);

// ---------------------------------------------------------------------------
// Batched ensembles: independent meshes of different sizes, component counts
// and active-set capacities processed by shared launches.  A launch's grid is
// cut into segments, one per member; every block finds its segment (a binary
// search over the segment offsets, uniform per block) and runs the
// single-mesh kernel body on that member's mesh, configuration and slice of
// the concatenated state vectors.
struct BatchMember {
    RhsArgs a;                // mesh pointers of the member (X, F, lists set per launch)
    DevCfg c;                 // the member's configuration tables
    long long xoff;           // first double of the member in the batch vectors
    long long celloff;        // first cell of the member in batch per-cell arrays
    int32_t *bins;            // 5 lists of a.ncells entries
    int32_t *counts;          // RHS_COUNTERS counters
    double *cm;               // per-cell prep records
    int32_t *glist;           // PostStep scratch, ncells*(P+1)
    int32_t *err;             // PostStep capacity flag of the member
    double *partial;          // error-norm partial sums of the member
    int npart;                // number of partial sums
    int fold;                 // fold the T-series at the member's first cell
};

// segment of block b: seg[k] <= b < seg[k+1] (uniform per block)
__device__ __forceinline__ int seg_find(const int32_t *__restrict__ seg, int nseg, int b) {
    int lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (seg[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

struct BatchSegs {
    const int32_t *seg = nullptr;     // nseg+1 block offsets
    const int32_t *segmem = nullptr;  // member index of each segment
    const int32_t *segq = nullptr;    // list of each segment (merged list launches)
    int nseg = 0;
    int nblocks = 0;
    size_t lds = 0;                   // dynamic LDS bytes (max over members)
};

// lanes per cell and block order of the five lists of a class
struct BatchLists {
    int32_t lanes[5];
    int32_t map[5];
};

// one class of members: same component count and lane capacity G
struct BatchLaunch {
    const BatchMember *mem;
    const int32_t *live;      // per-member flag (null: all members)
    int ncp, G, nbins;
    int do_bin;               // sort the cells into lists (0: reuse the previous call's lists)
    BatchSegs fold, prep, bin, lists, post, dofs;
    BatchSegs relists;        // the list launch's segments with at most RHS_REDO_CAP blocks per list
    // with the single-phase list in a launch of its own (own_launch): the
    // list launch's segments of lists 1.. and of list 0 (else empty)
    BatchSegs listsN, lists0;
    BatchLists shape;
    const double *X;
    double *F;
    uint32_t *active;
};

// Every launch of one class's right-hand side on stream st: prep records,
// per-size lists, all lists of all members in one launch (segments per list
// and member, largest active sets first), and in FP32-mixed the FP64
// recomputation lists in one more launch.
cipher_status launch_rhs_batch(const BatchLaunch &L, int precision, hipStream_t st);
cipher_status launch_fold_batch(const BatchLaunch &L, hipStream_t st);
cipher_status launch_post_batch(const BatchLaunch &L, double *X, hipStream_t st);
cipher_status launch_wrms_batch(const BatchLaunch &L, const double *U, const double *Y, double atol, double rtol,
                                double *norms, hipStream_t st);
cipher_status launch_axpy_dt_batch(const BatchLaunch &L, double *Y, const double *X, const double *dt_dev, int nk,
                                   const double *A, const double *const *K, hipStream_t st);
cipher_status launch_copy_batch(const BatchLaunch &L, double *Y, const double *X, const int32_t *mask,
                                hipStream_t st);
// partial sums of the error norm over n doubles (fixed for a given n)
int wrms_parts(long long n);

cipher_status launch_wrms(const double *U, const double *Y, long long n, double atol, double rtol,
                          double *partial, int npart, double *out, hipStream_t st);
cipher_status launch_axpy_n(long long n, double *Y, const double *X, int nk, const double *a,
                            const double *const *K, hipStream_t st);
cipher_status launch_axpy_dt(long long n, double *Y, const double *X, const double *dt_dev, int nk, const double *A,
                             const double *const *K, hipStream_t st);

}  // namespace cipher
