// api.hip - the C ABI of include/cipher_amd.h.
//
// Host side of the drop-in: validates and uploads the configuration tables
// (the AppCtx fields RHSFunction reads, src/typedef.h:145-319), builds the
// per-cell face lists from a flattened mesh (src/cipher.c:1187-1231), owns
// every scratch buffer, and launches the kernels on the context's stream.
// No entry point allocates after cipher_upload_mesh.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.h"

using namespace cipher;

static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
int cipher_fail(int code, const std::string &msg) { return fail(code, msg); }

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(CIPHER_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

static void free_mesh(cipher_ctx *c) {
    if (c->dc.gw) (void)hipFree((void *)c->dc.gw);
    c->dc.gw = nullptr;
    void *ptrs[] = {c->cell_off, c->adj, c->fgeo, c->bcell_off, c->bface_id, c->cellgeom, c->cm, c->glist, c->list, c->bins};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    c->cell_off = c->adj = c->bcell_off = c->bface_id = nullptr;
    c->cellgeom = c->fgeo = nullptr;
    c->cm = nullptr;
    c->glist = nullptr;
    c->list = nullptr;
    c->bins = nullptr;
    c->ncells = 0;
}

extern "C" const char *cipher_last_error(void) { return g_err.c_str(); }

extern "C" int cipher_layout_of(const cipher_config *cfg, cipher_layout *out) {
    if (!cfg || !out) return fail(CIPHER_ERR_ARG, "null argument");
    if (cfg->nmat <= 0 || !cfg->material) return fail(CIPHER_ERR_ARG, "no materials");
    int mpn = cfg->max_phase_neighbours > 0 ? cfg->max_phase_neighbours : CIPHER_MAXAP;
    cipher_layout L{};
    L.maxsites = 0;
    for (int m = 0; m < cfg->nmat; m++) L.maxsites = std::max(L.maxsites, (int32_t)cfg->material[m].nsites);
    L.sf_size = L.maxsites * cfg->ncp;
    L.sp_size = L.maxsites * (cfg->ncp - 1);
    L.as_offset = 0;
    L.as_size = std::min(mpn, (int)cfg->npf) + 1;
    L.pf_offset = L.as_offset + L.as_size;
    L.pf_size = std::min(mpn, (int)cfg->npf);
    L.dp_offset = L.pf_offset + L.pf_size;
    L.dp_size = cfg->ncp - 1;
    L.ex_offset = L.dp_offset + L.dp_size;
    L.ex_size = L.pf_size * L.sp_size;
    L.tm_offset = L.ex_offset + L.ex_size;
    L.tm_size = 1;
    L.ndof = L.tm_offset + L.tm_size;
    *out = L;
    return CIPHER_OK;
}

template <class T>
static int upload(cipher_ctx *c, const T *src, int n, const T **dst) {
    size_t bytes = sizeof(T) * (size_t)std::max(n, 1);
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, bytes));
    c->allocs.push_back(p);
    if (n > 0) HIPCHK(hipMemcpy(p, src, sizeof(T) * (size_t)n, hipMemcpyHostToDevice));
    else HIPCHK(hipMemset(p, 0, bytes));
    *dst = (const T *)p;
    return CIPHER_OK;
}

static int validate(const cipher_config *cfg) {
    if (cfg->dim < 1 || cfg->dim > 3) return fail(CIPHER_ERR_ARG, "dim must be 1..3");
    if (cfg->ncp < 1 || cfg->ncp > CIPHER_MAXCP) return fail(CIPHER_ERR_ARG, "ncp must be 1..5");
    if (cfg->npf < 1) return fail(CIPHER_ERR_ARG, "npf must be positive");
    if (cfg->npf > 65535) return fail(CIPHER_ERR_ARG, "phase ids are uint16 (npf <= 65535)");
    if (cfg->nf < 1 || !cfg->interface) return fail(CIPHER_ERR_ARG, "no interfaces");
    if (!cfg->phase_material || !cfg->interface_list) return fail(CIPHER_ERR_ARG, "missing mappings");
    for (int i = 0; i < cfg->nf; i++) {
        const cipher_interface &f = cfg->interface[i];
        const int ne = f.e_naniso > 0 ? f.e_aniso + (cfg->dim + 1) * f.e_naniso : 0;
        const int nm = f.m_naniso > 0 ? f.m_aniso + (cfg->dim + 1) * f.m_naniso : 0;
        if (f.e_naniso < 0 || f.m_naniso < 0 || ne > cfg->naniso || nm > cfg->naniso || ((ne || nm) && !cfg->aniso))
            return fail(CIPHER_ERR_ARG, "anisotropy table out of range");
    }
    if (cfg->interpolation < CIPHER_INTERP_LINEAR || cfg->interpolation > CIPHER_INTERP_CUBIC)
        return fail(CIPHER_ERR_ARG, "interpolation");
    if (cfg->ntseries < 1 || !cfg->tseries) return fail(CIPHER_ERR_ARG, "tseries pool must hold the empty series");
    for (int i = 0; i < cfg->ntseries; i++)
        if (cfg->tseries[i].n < 0 || cfg->tseries[i].n > CIPHER_MAXTSER) return fail(CIPHER_ERR_ARG, "tseries size");
    for (int i = 0; i < cfg->nrk; i++) {
        const cipher_rk &r = cfg->rk[i];
        if (r.n < 0 || (r.n > 0 && (r.ts < 0 || r.ts + r.n > cfg->ntseries)))
            return fail(CIPHER_ERR_ARG, "rk table out of range");
    }
    for (int p = 0; p < cfg->npf; p++)
        if (cfg->phase_material[p] < 0 || cfg->phase_material[p] >= cfg->nmat)
            return fail(CIPHER_ERR_ARG, "phase_material out of range");
    for (int i = 0; i < cfg->npf; i++)
        for (int j = i + 1; j < cfg->npf; j++) {
            int it = cfg->interface_list[(size_t)i * cfg->npf + j];
            if (it < 0 || it >= cfg->nf) return fail(CIPHER_ERR_ARG, "interface_list out of range");
        }
    auto ts_ok = [&](int i) { return i >= 0 && i < cfg->ntseries; };
    for (int m = 0; m < cfg->nmat; m++) {
        const cipher_material &mt = cfg->material[m];
        if (!(mt.model >= CIPHER_CHEM_QUADRATIC && mt.model <= CIPHER_CHEM_NONE))
            return fail(CIPHER_ERR_ARG, "material model");
        if (!ts_ok(mt.ref) || !ts_ok(mt.specific_heat) || !ts_ok(mt.latent_heat) || !ts_ok(mt.tconductivity))
            return fail(CIPHER_ERR_ARG, "material tseries index");
        if (mt.molarvolume == 0.0) return fail(CIPHER_ERR_ARG, "molarvolume must be nonzero");
        if (mt.model != CIPHER_CHEM_NONE && (mt.mobility < 0 || mt.mobility + cfg->ncp > cfg->nmobility))
            return fail(CIPHER_ERR_ARG, "material mobility index");
        // Redlich-Kister blocks the constitutive functions walk
        // (src/material.c:356-451, 495-619, 1023-1074)
        const int nc = cfg->ncp, p2 = nc * (nc - 1) / 2, p3 = nc * (nc - 1) * (nc - 2) / 6;
        auto rk_ok = [&](int off, int cnt) { return cnt == 0 || (off >= 0 && off + cnt <= cfg->nrk); };
        bool ok = true;
        if (mt.model == CIPHER_CHEM_CALPHADDIS) ok = rk_ok(mt.binary, p2) && rk_ok(mt.ternary, p3);
        if (mt.model == CIPHER_CHEM_CALPHAD2SL)
            ok = rk_ok(mt.binary, nc * p2) && rk_ok(mt.binaryq, nc * p2) && rk_ok(mt.ternary, nc * p3) &&
                 rk_ok(mt.ternaryq, nc * p3);
        if (mt.model != CIPHER_CHEM_NONE)
            for (int k = 0; k < nc && ok; k++) ok = rk_ok(cfg->mobility[mt.mobility + k].binary, p2);
        if (mt.model == CIPHER_CHEM_CALPHADDIS && ok)
            for (int t = 0; t < p3; t++) {
                const cipher_rk &rt = cfg->rk[mt.ternary + t];
                if (rt.n > 0 && (!cfg->rkindex || rt.idx < 0 || rt.idx + rt.n > cfg->nrkindex)) ok = false;
            }
        if (!ok) return fail(CIPHER_ERR_ARG, "material Redlich-Kister block out of range");
    }
    for (int i = 0; i < cfg->nf; i++) {
        const cipher_interface &f = cfg->interface[i];
        if (!ts_ok(f.e_act) || !ts_ok(f.m_act)) return fail(CIPHER_ERR_ARG, "interface tseries index");
        if (f.width == 0.0) return fail(CIPHER_ERR_ARG, "interface width must be nonzero");
    }
    return CIPHER_OK;
}

extern "C" int cipher_ctx_create(const cipher_config *cfg, int precision, cipher_ctx **out) {
    if (!cfg || !out) return fail(CIPHER_ERR_ARG, "null argument");
    if (precision != CIPHER_FP64 && precision != CIPHER_FP32_MIXED) return fail(CIPHER_ERR_ARG, "precision");
    int rc = validate(cfg);
    if (rc) return rc;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    cipher_ctx *c = new cipher_ctx();
    c->precision = precision;
    c->device = dev;
    cipher_layout_of(cfg, &c->L);
    if (c->L.pf_size > 16) {
        delete c;
        return fail(CIPHER_ERR_UNSUPPORTED, "active-set capacity above 16");
    }
    DevCfg &d = c->dc;
    int r = 0;
    r |= upload(c, cfg->tseries, cfg->ntseries, &d.ts);
    r |= upload(c, cfg->rk, cfg->nrk, &d.rk);
    {
        // bit r: rk[r] has terms; two zero words of padding for rk_window
        std::vector<uint64_t> bits((size_t)(cfg->nrk + 63) / 64 + 2, 0ull);
        for (int i = 0; i < cfg->nrk; i++)
            if (cfg->rk[i].n > 0) bits[(size_t)i >> 6] |= 1ull << (i & 63);
        r |= upload(c, bits.data(), (int)bits.size(), &d.rkbits);
    }
    r |= upload(c, cfg->rkindex, cfg->nrkindex, &d.rkindex);
    r |= upload(c, cfg->mobility, cfg->nmobility, &d.mob);
    r |= upload(c, cfg->material, cfg->nmat, &d.mat);
    r |= upload(c, cfg->interface, cfg->nf, &d.itf);
    r |= upload(c, cfg->source, cfg->nsource, &d.src);
    r |= upload(c, cfg->bc, cfg->nbcs, &d.bc);
    r |= upload(c, cfg->aniso, cfg->naniso, &d.aniso);
    r |= upload(c, cfg->phase_material, cfg->npf, &d.phase_material);
    r |= upload(c, cfg->interface_list, cfg->npf * cfg->npf, &d.ilist);
    {
        // compact [e0, width, m0] per interface type for the pair loops
        std::vector<double> it(3 * (size_t)std::max(cfg->nf, 1), 0.0);
        for (int i = 0; i < cfg->nf; i++) {
            it[3 * i] = cfg->interface[i].e0;
            it[3 * i + 1] = cfg->interface[i].width;
            it[3 * i + 2] = cfg->interface[i].m0;
        }
        r |= upload(c, it.data(), (int)it.size(), &d.itab);
    }
    if (r) {
        cipher_ctx_destroy(c);
        return CIPHER_ERR_HIP;
    }
    d.junctionpenalty = cfg->junctionpenalty;
    d.trate = cfg->temperature_rate;
    d.has_trate = cfg->has_temperature_rate;
    d.dim = cfg->dim;
    d.npf = cfg->npf;
    d.ncp = cfg->ncp;
    d.nd = cfg->ncp - 1;
    d.nmat = cfg->nmat;
    d.nf = cfg->nf;
    d.nbcs = cfg->nbcs;
    d.interp = cfg->interpolation;
    d.P = c->L.pf_size;
    d.D = c->L.ndof;
    d.aso = c->L.as_offset;
    d.pfo = c->L.pf_offset;
    d.dpo = c->L.dp_offset;
    d.exo = c->L.ex_offset;
    d.tmo = c->L.tm_offset;
    d.SP = c->L.sp_size;
    d.SF = c->L.sf_size;
    d.has_tcond = d.has_theat = d.has_sources = d.has_potential = d.has_imobc = d.has_iact = 0;
    d.has_grad = cfg->gradient_calculation ? 1 : 0;
    d.gw = nullptr;
    for (int m = 0; m < cfg->nmat; m++) {
        const cipher_material &mt = cfg->material[m];
        if (cfg->tseries[mt.tconductivity].n) d.has_tcond = 1;
        if (cfg->tseries[mt.specific_heat].n || cfg->tseries[mt.latent_heat].n) d.has_theat = 1;
        for (int s = 0; s < mt.nsources; s++)
            if (cfg->source[mt.source + s].model == CIPHER_SOURCE_SINK) d.has_sources = 1;
    }
    for (int i = 0; i < cfg->nf; i++) {
        const cipher_interface &f = cfg->interface[i];
        if (f.has_potential) d.has_potential = 1;
        if (f.has_mobilityc) d.has_imobc = 1;
        if (cfg->tseries[f.e_act].n || cfg->tseries[f.m_act].n) d.has_iact = 1;
        for (int s = 0; s < f.nsources; s++)
            if (cfg->source[f.source + s].model == CIPHER_SOURCE_SINK) d.has_sources = 1;
    }
    d.ntseries = cfg->ntseries;
    d.tsv = nullptr;
    c->fold = (cfg->ncp > 1) || d.has_tcond || d.has_theat || d.has_iact;
    if (hipMalloc(&c->tsv, sizeof(double) * (cfg->ntseries + 1)) != hipSuccess) {
        cipher_ctx_destroy(c);
        return fail(CIPHER_ERR_HIP, "scratch allocation failed");
    }
    {
        std::vector<double> nanv(cfg->ntseries + 1, std::nan(""));
        HIPCHK(hipMemcpy(c->tsv, nanv.data(), sizeof(double) * nanv.size(), hipMemcpyHostToDevice));
    }
    if (c->fold) d.tsv = c->tsv;
    if (hipMalloc(&c->d_cfg, sizeof(DevCfg)) != hipSuccess || cipher_sync_cfg(c) != CIPHER_OK ||
        hipMalloc(&c->err, sizeof(int32_t)) != hipSuccess || hipMalloc(&c->counts, (RHS_RING + 1) * RHS_COUNTERS * sizeof(int32_t)) != hipSuccess || hipMalloc(&c->partial, sizeof(double) * c->npart) != hipSuccess ||
        hipMalloc(&c->norm, sizeof(double)) != hipSuccess ||
        hipHostMalloc(&c->h_norm, sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&c->h_err, sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
        cipher_ctx_destroy(c);
        return fail(CIPHER_ERR_HIP, "scratch allocation failed");
    }
    *out = c;
    return CIPHER_OK;
}

int cipher_sync_cfg(cipher_ctx *c) {
    if (!c->d_cfg) return CIPHER_OK;
    if (c->stream) HIPCHK(hipStreamSynchronize(c->stream));
    else HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(c->d_cfg, &c->dc, sizeof(DevCfg), hipMemcpyHostToDevice));
    return CIPHER_OK;
}

extern "C" int cipher_ctx_destroy(cipher_ctx *c) {
    if (!c) return CIPHER_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    else (void)hipDeviceSynchronize();
    cipher_ts_free(c);
    free_mesh(c);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->err) (void)hipFree(c->err);
    if (c->counts) (void)hipFree(c->counts);
    if (c->tsv) (void)hipFree(c->tsv);
    if (c->partial) (void)hipFree(c->partial);
    if (c->norm) (void)hipFree(c->norm);
    if (c->h_norm) (void)hipHostFree(c->h_norm);
    if (c->h_err) (void)hipHostFree(c->h_err);
    if (c->d_cfg) (void)hipFree(c->d_cfg);
    delete c;
    return CIPHER_OK;
}

extern "C" int cipher_ctx_layout(const cipher_ctx *c, cipher_layout *out) {
    if (!c || !out) return fail(CIPHER_ERR_ARG, "null argument");
    *out = c->L;
    return CIPHER_OK;
}

extern "C" int cipher_set_precision(cipher_ctx *c, int precision) {
    if (!c) return fail(CIPHER_ERR_ARG, "null ctx");
    if (precision != CIPHER_FP64 && precision != CIPHER_FP32_MIXED) return fail(CIPHER_ERR_ARG, "precision");
    if (c->precision != precision) cipher_ts_invalidate(c);
    c->precision = precision;
    return CIPHER_OK;
}

extern "C" int cipher_set_stream(cipher_ctx *c, void *stream) {
    if (!c) return fail(CIPHER_ERR_ARG, "null ctx");
    if (c->stream != (hipStream_t)stream) cipher_ts_invalidate(c);
    c->stream = (hipStream_t)stream;
    return CIPHER_OK;
}

extern "C" int cipher_set_temperature_rate(cipher_ctx *c, int has_rate, double rate) {
    if (!c) return fail(CIPHER_ERR_ARG, "null ctx");
    cipher_ts_invalidate(c);
    c->dc.has_trate = has_rate ? 1 : 0;
    c->dc.trate = rate;
    return cipher_sync_cfg(c);
}

template <int DIM>
static void lsq_cell(const cipher_mesh *m, const std::vector<int32_t> &off, const std::vector<int32_t> &adj,
                     long long cell, double *gw) {
    const int k0 = off[cell], nls = off[cell + 1] - off[cell];
    double a[24 * DIM], ata[DIM * DIM], atainv[DIM * DIM], gm[24 * DIM];
    for (int j = 0; j < nls; j++) {
        const long long nb = adj[k0 + j] >> 1;
        for (int d = 0; d < DIM; d++) {
            double v = m->centroid[nb * DIM + d];
            v -= m->centroid[cell * DIM + d];
            const double L = m->period[d];
            if (L > 0.0) {
                if (v > 0.5 * L) v -= L;
                else if (v < -0.5 * L) v += L;
            }
            a[j * DIM + d] = v;
        }
    }
    for (int di = 0; di < DIM; di++)
        for (int dj = di; dj < DIM; dj++) {
            double val = 0.0;
            for (int j = 0; j < nls; j++) val += a[j * DIM + di] * a[j * DIM + dj];
            ata[di * DIM + dj] = val;
            ata[dj * DIM + di] = val;
        }
    dinvert<DIM, double>(atainv, ata);
    for (int q = 0; q < 24 * DIM; q++) gm[q] = 0.0;
    for (int di = 0; di < DIM; di++)
        for (int j = 0; j < nls; j++)
            for (int d = 0; d < DIM; d++) gm[di * nls + j] += atainv[di * DIM + d] * a[j * DIM + d];
    for (int j = 0; j < nls; j++)
        for (int d = 0; d < DIM; d++) gw[(size_t)(k0 + j) * DIM + d] = gm[d * nls + j];
}

static int lsq_weights(int dim, const cipher_mesh *m, const std::vector<int32_t> &off,
                       const std::vector<int32_t> &adj, std::vector<double> &gw) {
    if (!m->centroid) return fail(CIPHER_ERR_ARG, "gradient_calculation needs cell centroids");
    const long long N = m->ncells;
    gw.assign((size_t)std::max<long long>(off[N], 1) * dim, 0.0);
    for (long long cell = 0; cell < N; cell++) {
        if (off[cell + 1] - off[cell] > 24)
            return fail(CIPHER_ERR_UNSUPPORTED, "more than 24 face neighbours (src/cipher.c:1240)");
        if (dim == 1) lsq_cell<1>(m, off, adj, cell, gw.data());
        else if (dim == 2) lsq_cell<2>(m, off, adj, cell, gw.data());
        else lsq_cell<3>(m, off, adj, cell, gw.data());
    }
    return CIPHER_OK;
}

extern "C" int cipher_upload_mesh(cipher_ctx *c, const cipher_mesh *m) {
    if (!c || !m) return fail(CIPHER_ERR_ARG, "null argument");
    if (m->ncells < 0 || m->nfaces < 0 || m->nbfaces < 0) return fail(CIPHER_ERR_ARG, "negative sizes");
    if (m->ncells >= (1 << 30)) return fail(CIPHER_ERR_ARG, "more than 2^30 cells");
    if ((m->nfaces && !m->face_cells) || (m->ncells && !m->cellgeom) ||
        (m->nbfaces && (!m->bface_cell || !m->bface_id)))
        return fail(CIPHER_ERR_ARG, "null mesh array");
    const long long N = m->ncells;
    std::vector<int32_t> off(N + 1, 0), boff(N + 1, 0);
    for (int f = 0; f < m->nfaces; f++) {
        int a = m->face_cells[2 * f], b = m->face_cells[2 * f + 1];
        if (a < 0 || a >= N || b < 0 || b >= N) return fail(CIPHER_ERR_ARG, "face cell out of range");
        off[a + 1]++;
        off[b + 1]++;
    }
    for (long long i = 0; i < N; i++) off[i + 1] += off[i];
    std::vector<int32_t> adj(std::max<long long>(off[N], 1));
    std::vector<int32_t> pos(off.begin(), off.end() - 1);
    for (int f = 0; f < m->nfaces; f++) {
        int a = m->face_cells[2 * f], b = m->face_cells[2 * f + 1];
        adj[pos[a]++] = (b << 1) | 0;   // a is L
        adj[pos[b]++] = (a << 1) | 1;   // b is R
    }
    for (int k = 0; k < m->nbfaces; k++) {
        int a = m->bface_cell[k];
        if (a < 0 || a >= N) return fail(CIPHER_ERR_ARG, "boundary face cell out of range");
        boff[a + 1]++;
    }
    for (long long i = 0; i < N; i++) boff[i + 1] += boff[i];
    std::vector<int32_t> bid(std::max<long long>(boff[N], 1));
    std::vector<int32_t> bpos(boff.begin(), boff.end() - 1);
    for (int k = 0; k < m->nbfaces; k++) bid[bpos[m->bface_cell[k]]++] = m->bface_id[k];
    cipher_ts_free(c);                 // stage vectors are sized for the old mesh
    free_mesh(c);
    size_t cmrec = prep_record_bytes(c->dc.ncp);
    HIPCHK(hipMalloc(&c->cell_off, sizeof(int32_t) * (N + 1)));
    HIPCHK(hipMalloc(&c->adj, sizeof(int32_t) * adj.size()));
    HIPCHK(hipMalloc(&c->fgeo, sizeof(double) * adj.size()));
    HIPCHK(hipMalloc(&c->bcell_off, sizeof(int32_t) * (N + 1)));
    HIPCHK(hipMalloc(&c->bface_id, sizeof(int32_t) * bid.size()));
    HIPCHK(hipMalloc(&c->cellgeom, sizeof(double) * std::max<long long>(N, 1)));
    HIPCHK(hipMalloc(&c->cm, cmrec * std::max<long long>(N, 1)));
    HIPCHK(hipMalloc(&c->glist, sizeof(int32_t) * (size_t)(c->L.pf_size + 1) * std::max<long long>(N, 1)));
    HIPCHK(hipMalloc(&c->list, 5 * sizeof(int32_t) * std::max<long long>(N, 1)));
    HIPCHK(hipMalloc(&c->bins, 5 * sizeof(int32_t) * std::max<long long>(N, 1)));
    HIPCHK(hipMemcpy(c->cell_off, off.data(), sizeof(int32_t) * (N + 1), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->adj, adj.data(), sizeof(int32_t) * adj.size(), hipMemcpyHostToDevice));
    {
        // flux factor of every face entry, 2 * (the smaller cell's v/h) / (hL + hR)
        // with v = h^dim (src/cipher.c:248-262), by the same FP64 operations
        // the reference evaluates per face; the same for both cells of a face
        std::vector<double> fg(adj.size(), 0.0);
        const unsigned dim = (unsigned)c->dc.dim;
        auto vol = [dim](double b) {
            double r = 1.0;
            for (unsigned e = dim; e; e >>= 1) {
                if (e & 1u) r *= b;
                b *= b;
            }
            return r;
        };
        for (long long cell = 0; cell < N; cell++)
            for (int k = off[cell]; k < off[cell + 1]; k++) {
                const int nb = adj[k] >> 1;
                const bool selfL = (adj[k] & 1) == 0;
                const double dL = selfL ? m->cellgeom[cell] : m->cellgeom[nb];
                const double dR = selfL ? m->cellgeom[nb] : m->cellgeom[cell];
                const double vL = vol(dL), vR = vol(dR);
                fg[k] = 2.0 * (dL < dR ? vL / dL : vR / dR) / (dL + dR);
            }
        HIPCHK(hipMemcpy(c->fgeo, fg.data(), sizeof(double) * fg.size(), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(c->bcell_off, boff.data(), sizeof(int32_t) * (N + 1), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->bface_id, bid.data(), sizeof(int32_t) * bid.size(), hipMemcpyHostToDevice));
    if (N) HIPCHK(hipMemcpy(c->cellgeom, m->cellgeom, sizeof(double) * N, hipMemcpyHostToDevice));
    if (c->dc.has_grad) {
        // least-squares gradient weights (A^T A)^-1 A^T per cell over its face
        // neighbours in face order, src/cipher.c:1233-1282
        std::vector<double> gw;
        int rc = lsq_weights(c->dc.dim, m, off, adj, gw);
        if (rc) {
            free_mesh(c);
            return rc;
        }
        double *dgw = nullptr;
        HIPCHK(hipMalloc(&dgw, sizeof(double) * gw.size()));
        HIPCHK(hipMemcpy(dgw, gw.data(), sizeof(double) * gw.size(), hipMemcpyHostToDevice));
        c->dc.gw = dgw;
    }
    if (int rs = cipher_sync_cfg(c)) {
        free_mesh(c);
        return rs;
    }
    c->ncells = N;
    return CIPHER_OK;
}

int cipher_rhs_impl(cipher_ctx *c, const double *X, double *F, uint32_t *active, int bins_mode) {
    if (!c || !X || !F) return fail(CIPHER_ERR_ARG, "null argument");
    if (!c->cell_off) return fail(CIPHER_ERR_STATE, "no mesh uploaded");
    RhsLaunch L;
    L.cfg = &c->dc;
    L.dcfg = c->d_cfg;
    L.X = X;
    L.F = F;
    L.active = active;
    L.cellgeom = c->cellgeom;
    L.cell_off = c->cell_off;
    L.adj = c->adj;
    L.fgeo = c->fgeo;
    L.bcell_off = c->bcell_off;
    L.bface_id = c->bface_id;
    L.cm = c->cm;
    L.redo = c->list;
    L.bins = c->bins;
    // counters: an eager call takes the next set of the ring (the first call
    // of a round clears the whole ring); a captured call, or a stage of the
    // integrator, uses the last set and clears it itself, so a replayed
    // graph starts from zero every time
    hipStreamCaptureStatus cap = hipStreamCaptureStatusActive;
    if (bins_mode == 0 && hipStreamIsCapturing(c->stream, &cap) != hipSuccess) {
        (void)hipGetLastError();     // unknown: treat as captured
        cap = hipStreamCaptureStatusActive;
    }
    if (bins_mode == 0 && cap == hipStreamCaptureStatusNone) {
        if (c->ring_pos >= RHS_RING) {
            L.fill = c->counts;
            L.nfill = RHS_RING * RHS_COUNTERS;
            c->ring_pos = 0;
        } else {
            L.fill = nullptr;
            L.nfill = 0;
        }
        L.counts = c->counts + (size_t)c->ring_pos++ * RHS_COUNTERS;
    } else {
        L.counts = c->counts + (size_t)RHS_RING * RHS_COUNTERS;
        L.fill = L.counts;
        L.nfill = RHS_COUNTERS;
    }
    L.ncells = c->ncells;
    for (int q = 0; q < 5; q++) L.shape[q] = c->shape[q];
    L.bins_mode = bins_mode;
    L.pred = c->pred;
    if (c->fold) {
        cipher_status f = launch_tsfold(X, c->ncells, c->dc, c->tsv, c->stream);
        if (f != CIPHER_OK) return fail(f, "fold launch");
    }
    cipher_status s = launch_rhs(L, c->precision, c->stream);
    if (s != CIPHER_OK) return fail(s, std::string("rhs launch: ") + hipGetErrorString(hipGetLastError()));
    return CIPHER_OK;
}

extern "C" int cipher_rhs(cipher_ctx *c, double t, const double *X, double *F) {
    (void)t;
    return cipher_rhs_impl(c, X, F, nullptr);
}

extern "C" int cipher_rhs_active(cipher_ctx *c, double t, const double *X, double *F, uint32_t *active) {
    (void)t;
    if (!active) return fail(CIPHER_ERR_ARG, "null active buffer");
    return cipher_rhs_impl(c, X, F, active);
}

extern "C" int cipher_output(cipher_ctx *c, const double *X, int32_t nfields, const cipher_output_field *fields,
                             double *out) {
    if (!c || !X || !out || (nfields > 0 && !fields)) return fail(CIPHER_ERR_ARG, "null argument");
    if (nfields < 0 || nfields > CIPHER_MAX_OUTPUTS) return fail(CIPHER_ERR_ARG, "number of output fields");
    if (!c->cell_off) return fail(CIPHER_ERR_STATE, "no mesh uploaded");
    OutFields f{};
    f.n = nfields;
    for (int k = 0; k < nfields; k++) {
        if (fields[k].kind < CIPHER_OUT_PHASEID || fields[k].kind > CIPHER_OUT_TEMPERATURE)
            return fail(CIPHER_ERR_ARG, "unknown output field kind");
        f.kind[k] = fields[k].kind;
        f.index[k] = fields[k].index;
        if (fields[k].kind == CIPHER_OUT_COMP) f.need_comp = 1;
    }
    if (nfields == 0) return CIPHER_OK;
    if (c->fold && f.need_comp) {
        cipher_status s = launch_tsfold(X, c->ncells, c->dc, c->tsv, c->stream);
        if (s != CIPHER_OK) return fail(s, "fold launch");
    }
    cipher_status s = launch_output(c->dc, X, (double *)c->cm, c->ncells, f, out, c->stream);
    if (s != CIPHER_OK) return fail(s, std::string("output launch: ") + hipGetErrorString(hipGetLastError()));
    return CIPHER_OK;
}

static int lanes_cap(const cipher_ctx *c) {
    const int P = c->L.pf_size;
    return P <= 4 ? 4 : (P <= 8 ? 8 : 16);
}

extern "C" int cipher_get_shape(const cipher_ctx *c, int32_t q, cipher_list_shape *out) {
    if (!c || !out || q < 0 || q >= 5) return fail(CIPHER_ERR_ARG, "bad argument");
    const ListShape &s = c->shape[q];
    out->lanes = s.lanes > 0 ? s.lanes : (1 << q);
    out->cap = s.cap;
    out->map = s.map;
    out->pad_ = 0;
    return CIPHER_OK;
}

extern "C" int cipher_set_shape(cipher_ctx *c, int32_t q, const cipher_list_shape *shape) {
    if (!c || !shape || q < 0 || q >= 5) return fail(CIPHER_ERR_ARG, "bad argument");
    if (!shape_valid(*shape, q, lanes_cap(c), c->dc.ncp)) return fail(CIPHER_ERR_ARG, "list shape out of range");
    ListShape &s = c->shape[q];
    s.lanes = shape->lanes;
    s.cap = shape->cap;
    s.map = shape->map;
    cipher_ts_invalidate(c);
    return CIPHER_OK;
}

// Coordinate descent over each list's shape (lanes, block order, grid cap), every candidate timed as the median of reps right-hand sides
// on X in the context's precision; the fastest is kept.
extern "C" int cipher_autotune(cipher_ctx *c, const double *X, double *F, int32_t reps, double *ms_out) {
    if (!c || !X || !F) return fail(CIPHER_ERR_ARG, "null argument");
    if (!c->cell_off) return fail(CIPHER_ERR_STATE, "no mesh uploaded");
    if (reps < 1) reps = 1;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int rc = CIPHER_OK;
    auto time_once = [&](float &ms) -> int {
        std::vector<float> v;
        for (int r = 0; r < reps; r++) {
            HIPCHK(hipEventRecord(e0, c->stream));
            int q = cipher_rhs_impl(c, X, F, nullptr);
            if (q) return q;
            HIPCHK(hipEventRecord(e1, c->stream));
            HIPCHK(hipEventSynchronize(e1));
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, e0, e1));
            v.push_back(t);
        }
        std::sort(v.begin(), v.end());
        ms = v[v.size() / 2];
        return CIPHER_OK;
    };
    float best = 0.f;
    if ((rc = cipher_rhs_impl(c, X, F, nullptr)) == CIPHER_OK) rc = time_once(best);
    const int G = lanes_cap(c);
    const int nbins = G >= 16 ? 5 : (G >= 8 ? 4 : 3);
    for (int q = 0; q < nbins && rc == CIPHER_OK; q++) {
        for (const ShapeStep &step : shape_steps(q, G, c->dc.ncp)) {
            const ListShape keep = c->shape[q];
            const ListShape cand = apply_step(keep, step);
            const int kl = keep.lanes > 0 ? keep.lanes : (1 << q), cl = cand.lanes > 0 ? cand.lanes : (1 << q);
            if (kl == cl && cand.map == keep.map && cand.cap == keep.cap) continue;
            c->shape[q] = cand;
            float ms = 0.f;
            if ((rc = time_once(ms))) break;
            if (ms < best) best = ms;
            else c->shape[q] = keep;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    cipher_ts_invalidate(c);
    if (rc == CIPHER_OK && ms_out) *ms_out = best;
    return rc;
}

int cipher_post_common(cipher_ctx *c, double *X, const int32_t *phase, bool setup) {
    if (!c || !X) return fail(CIPHER_ERR_ARG, "null argument");
    if (!c->cell_off) return fail(CIPHER_ERR_STATE, "no mesh uploaded");
    PostLaunch L;
    L.cfg = &c->dc;
    L.X = X;
    L.cell_off = c->cell_off;
    L.adj = c->adj;
    L.glist = c->glist;
    L.err = c->err;
    L.phase = phase;
    L.ncells = c->ncells;
    HIPCHK(hipMemsetAsync(c->err, 0, sizeof(int32_t), c->stream));
    if (setup) HIPCHK(hipMemsetAsync(X, 0, sizeof(double) * (size_t)c->ncells * c->L.ndof, c->stream));
    cipher_status s = setup ? launch_setup(L, c->stream) : launch_poststep(L, c->stream);
    if (s != CIPHER_OK) return fail(s, std::string("poststep launch: ") + hipGetErrorString(hipGetLastError()));
    HIPCHK(hipMemcpyAsync(c->h_err, c->err, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (*c->h_err) return fail(CIPHER_ERR_CAPACITY, "active-set union exceeds the capacity PF_SIZE");
    return CIPHER_OK;
}

extern "C" int cipher_poststep(cipher_ctx *c, double *X) { return cipher_post_common(c, X, nullptr, false); }

extern "C" int cipher_setup_problem(cipher_ctx *c, const int32_t *phase, double *X) {
    if (!phase) return fail(CIPHER_ERR_ARG, "null phase labels");
    return cipher_post_common(c, X, phase, true);
}

extern "C" int cipher_wrms_norm(cipher_ctx *c, const double *U, const double *Y, int64_t n, double atol, double rtol,
                                double *norm_out) {
    if (!c || !U || !Y || !norm_out || n < 0) return fail(CIPHER_ERR_ARG, "bad argument");
    cipher_status s = launch_wrms(U, Y, n, atol, rtol, c->partial, wrms_parts(n), c->norm, c->stream);
    if (s != CIPHER_OK) return fail(s, "wrms launch");
    HIPCHK(hipMemcpyAsync(c->h_norm, c->norm, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *norm_out = *c->h_norm;
    if (!std::isfinite(*norm_out)) return fail(CIPHER_ERR_NONFINITE, "error norm is not finite");
    return CIPHER_OK;
}

extern "C" int cipher_axpy_n(cipher_ctx *c, int64_t n, double *Y, const double *X, int nk, const double *a,
                             const double *const *K) {
    if (!c || !Y || !X || n < 0 || nk < 0 || nk > 8 || (nk > 0 && (!a || !K)))
        return fail(CIPHER_ERR_ARG, "bad argument");
    cipher_status s = launch_axpy_n(n, Y, X, nk, a, K, c->stream);
    if (s != CIPHER_OK) return fail(s, "axpy launch");
    return CIPHER_OK;
}
