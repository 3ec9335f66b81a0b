// C ABI of libmeshoid_hip.so (declared in include/meshoid_hip.h): argument
// checks, context/scratch management and error plumbing around the kernels.
#include <stdio.h>
#include <string.h>

#include <new>

#include "../../include/meshoid_hip.h"
#include "mhip_internal.hpp"

namespace mhip {

int fail(mhip_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

int check_hip(mhip_ctx *ctx, hipError_t e, const char *what) {
    if (e == hipSuccess) return 0;
    return fail(ctx, MHIP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// grow-only per-context device scratch; callers use it within one call, on
// the context's stream, so a grow waits for earlier work before freeing
void *scratch(mhip_ctx *ctx, size_t bytes, int *rc) {
    *rc = 0;
    if (bytes <= ctx->scratch_bytes) return ctx->scratch;
    if (ctx->scratch) {
        hipStreamSynchronize(ctx->stream);
        hipFree(ctx->scratch);
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
    }
    size_t want = bytes + bytes / 4;
    hipError_t e = hipMalloc(&ctx->scratch, want);
    if (e != hipSuccess) {
        *rc = check_hip(ctx, e, "hipMalloc (scratch)");
        ctx->scratch = nullptr;
        return nullptr;
    }
    ctx->scratch_bytes = want;
    return ctx->scratch;
}

}  // namespace mhip

using namespace mhip;

#define MHIP_NEED(cond, msg) \
    do {                      \
        if (!(cond)) return fail(ctx, MHIP_ERR_VALUE, msg); \
    } while (0)

extern "C" {

int mhip_version(void) { return 10000; }

int mhip_ctx_create(int device, void *stream, mhip_ctx **out) {
    if (!out) return MHIP_ERR_VALUE;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MHIP_ERR_HIP;
    if (device < 0 || device >= ndev) return MHIP_ERR_VALUE;
    if (hipSetDevice(device) != hipSuccess) return MHIP_ERR_HIP;
    mhip_ctx *c = new (std::nothrow) mhip_ctx();
    if (!c) return MHIP_ERR_HIP;
    c->device = device;
    c->stream = (hipStream_t)stream;
    *out = c;
    return MHIP_OK;
}

int mhip_ctx_destroy(mhip_ctx *ctx) {
    if (!ctx) return MHIP_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->stats) (void)hipFree(ctx->stats);
    for (auto &r : ctx->recs) {
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->order_ev) (void)hipEventDestroy(ctx->order_ev);
    delete ctx;
    return MHIP_OK;
}

int mhip_prof_enable(mhip_ctx *ctx, int level) {
    if (!ctx) return MHIP_ERR_VALUE;
    ctx->prof = level < 0 ? 0 : level;
    if (ctx->prof >= 2 && !ctx->stats) {
        int rc = check_hip(ctx, hipMalloc((void **)&ctx->stats, 16 * sizeof(unsigned long long)),
                           "hipMalloc (stats)");
        if (rc) return rc;
        rc = check_hip(ctx, hipMemsetAsync(ctx->stats, 0, 16 * sizeof(unsigned long long), ctx->stream),
                       "stats clear");
        if (rc) return rc;
    }
    return MHIP_OK;
}

int mhip_prof_read(mhip_ctx *ctx, char *buf, int64_t buflen) {
    if (!ctx || !buf || buflen < 1) return MHIP_ERR_VALUE;
    int rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "profile synchronize");
    if (rc) return rc;
    // aggregate by region name, in first-seen order
    std::vector<std::string> names;
    std::vector<double> tot;
    std::vector<long> cnt;
    for (auto &r : ctx->recs) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, r.start, r.stop);
        size_t i = 0;
        while (i < names.size() && names[i] != r.name) i++;
        if (i == names.size()) {
            names.push_back(r.name);
            tot.push_back(0.0);
            cnt.push_back(0);
        }
        tot[i] += ms;
        cnt[i]++;
        ctx->event_pool.push_back(r.start);
        ctx->event_pool.push_back(r.stop);
    }
    ctx->recs.clear();
    std::string out;
    char line[160];
    for (size_t i = 0; i < names.size(); i++) {
        snprintf(line, sizeof line, "%s %ld %.6f\n", names[i].c_str(), cnt[i], tot[i]);
        out += line;
    }
    if (ctx->stats) {
        static const char *stat_names[STAT_COUNT] = {"stat_p1_leaves", "stat_p1_inserts", "stat_p1_nodes",
                                                     "stat_p2_leaves", "stat_p2_nodes", "stat_waves",
                                                     "stat_lane_need", "stat_maxlane_need",
                                                     "stat_parent_max", "stat_nn_waves", "stat_nn_nodes",
                                                     "stat_nn_leaves", "stat_nn_exact", "stat_newton",
                                                     "stat_newton_wmax", "stat_overflow_waves"};
        unsigned long long hs[16];
        rc = check_hip(ctx, hipMemcpy(hs, ctx->stats, sizeof hs, hipMemcpyDeviceToHost), "stats read");
        if (rc) return rc;
        for (int i = 0; i < STAT_COUNT; i++) {
            snprintf(line, sizeof line, "%s %llu 0\n", stat_names[i], hs[i]);
            out += line;
        }
        rc = check_hip(ctx, hipMemset(ctx->stats, 0, sizeof hs), "stats clear");
        if (rc) return rc;
    }
    if ((int64_t)out.size() + 1 > buflen) return fail(ctx, MHIP_ERR_VALUE, "profile buffer too small");
    memcpy(buf, out.c_str(), out.size() + 1);
    return MHIP_OK;
}

int mhip_ctx_set_stream(mhip_ctx *ctx, void *stream) {
    if (!ctx) return MHIP_ERR_VALUE;
    hipStream_t ns = (hipStream_t)stream;
    if (ns == ctx->stream) return MHIP_OK;
    // the context's scratch and neighbour-list buffers may still be in use by
    // work queued on the old stream: the new stream waits for it first
    hipSetDevice(ctx->device);
    if (!ctx->order_ev) {
        int rc = check_hip(ctx, hipEventCreateWithFlags(&ctx->order_ev, hipEventDisableTiming), "event create");
        if (rc) return rc;
    }
    int rc = check_hip(ctx, hipEventRecord(ctx->order_ev, ctx->stream), "event record");
    if (!rc) rc = check_hip(ctx, hipStreamWaitEvent(ns, ctx->order_ev, 0), "stream wait");
    if (rc) return rc;
    ctx->stream = ns;
    return MHIP_OK;
}

int mhip_ctx_synchronize(mhip_ctx *ctx) {
    if (!ctx) return MHIP_ERR_VALUE;
    return check_hip(ctx, hipStreamSynchronize(ctx->stream), "stream synchronize");
}

const char *mhip_last_error(const mhip_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "no context"; }

int mhip_index_create(mhip_index **out) {
    if (!out) return MHIP_ERR_VALUE;
    *out = new (std::nothrow) mhip_index();
    return *out ? MHIP_OK : MHIP_ERR_HIP;
}

int mhip_index_destroy(mhip_index *ix) {
    if (!ix) return MHIP_OK;
    void *arrs[] = {ix->sx, ix->sy, ix->sz, ix->lx, ix->ly, ix->lz, ix->perm, ix->iperm, ix->nodes};
    for (void *p : arrs)
        if (p) hipFree(p);
    delete ix;
    return MHIP_OK;
}

int mhip_index_build(mhip_ctx *ctx, mhip_index *ix, const void *pos, int pos_f64, int64_t n, int dim,
                     double boxsize) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(ix && pos, "null index or positions");
    hipSetDevice(ctx->device);
    return index_build(ctx, pos, pos_f64 ? 1 : 0, n, dim, boxsize > 0 ? boxsize : -1.0, ix);
}

int mhip_index_info(const mhip_index *ix, int64_t *n, int *dim, double *boxsize, float *bbox6) {
    if (!ix) return MHIP_ERR_VALUE;
    if (n) *n = ix->n;
    if (dim) *dim = ix->dim;
    if (boxsize) *boxsize = ix->box;
    if (bbox6)
        for (int c = 0; c < 6; c++) bbox6[c] = ix->bbox[c];
    return MHIP_OK;
}

int mhip_index_order(mhip_ctx *ctx, const mhip_index *ix, uint32_t *perm_out) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(ix && perm_out, "null argument");
    MHIP_NEED(ix->n > 0, "index is empty");
    hipSetDevice(ctx->device);
    return check_hip(ctx, hipMemcpyAsync(perm_out, ix->perm, ix->n * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                                         ctx->stream),
                     "index order copy");
}

int mhip_nlist_create(mhip_nlist **out) {
    if (!out) return MHIP_ERR_VALUE;
    *out = new (std::nothrow) mhip_nlist();
    return *out ? MHIP_OK : MHIP_ERR_HIP;
}

int mhip_nlist_destroy(mhip_nlist *nl) {
    if (!nl) return MHIP_OK;
    void *arrs[] = {nl->qsorted, nl->qrow, nl->rowrec, nl->idx, nl->d2, nl->tau, nl->nlt, nl->h, nl->h64, nl->status, nl->amb};
    for (void *p : arrs)
        if (p) hipFree(p);
    delete nl;
    return MHIP_OK;
}

int mhip_knn_hsml(mhip_ctx *ctx, const mhip_index *ix, const int64_t *qidx, int64_t nq, int k,
                  const double *m_rows, const float *seed_h, const float *seed_r, float *h_out,
                  float *rho_out, float *vol_out, int32_t *status_out, int flags, mhip_nlist *nl) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(ix && nl, "null index or neighbour list");
    MHIP_NEED(ix->n > 0, "index has not been built");
    (void)hipSetDevice(ctx->device);
    return knn_run(ctx, ix, qidx, nq, k, m_rows, seed_h, seed_r, h_out, rho_out, vol_out, status_out,
                   flags, nl);
}

int mhip_seed_features(mhip_ctx *ctx, const mhip_index *ix, float *F) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(ix && F, "null index or feature buffer");
    (void)hipSetDevice(ctx->device);
    return seed_features(ctx, ix, F);
}

int mhip_seed_predict(mhip_ctx *ctx, const mhip_index *ix, const void *blob, float *h_rows) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(ix && blob && h_rows, "null argument");
    (void)hipSetDevice(ctx->device);
    return seed_predict(ctx, ix, blob, h_rows);
}

int mhip_mlp_train_step(mhip_ctx *ctx, const float *F, const float *T, int64_t nb, const float *mu,
                        const float *sd, float *P, float *m, float *v, float *grad, float *loss, int step,
                        float lr) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(F && T && mu && sd && P && grad && loss, "null argument");
    MHIP_NEED(step == 0 || (m && v), "Adam state needed for step > 0");
    (void)hipSetDevice(ctx->device);
    return mlp_train_step(ctx, F, T, nb, mu, sd, P, m, v, grad, loss, step, lr);
}

int mhip_mlp_pack(mhip_ctx *ctx, const float *P, const float *mu, const float *sd, void *blob) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(P && mu && sd && blob, "null argument");
    (void)hipSetDevice(ctx->device);
    return mlp_pack(ctx, P, mu, sd, blob);
}

int64_t mhip_mlp_blob_bytes(void) { return (int64_t)mlp_blob_bytes(); }

int mhip_nlist_export(mhip_ctx *ctx, const mhip_nlist *nl, int32_t *ngb_out, float *dist_out) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(nl, "null neighbour list");
    hipSetDevice(ctx->device);
    return nlist_export(ctx, nl, ngb_out, dist_out);
}

int mhip_nlist_hsml(mhip_ctx *ctx, const mhip_nlist *nl, double *h_out) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(nl && h_out, "null argument");
    hipSetDevice(ctx->device);
    return nlist_hsml(ctx, nl, h_out);
}

int mhip_nlist_export_f64(mhip_ctx *ctx, const mhip_nlist *nl, int32_t *ngb_out, double *dist_out) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(nl, "null neighbour list");
    hipSetDevice(ctx->device);
    return nlist_export64(ctx, nl, ngb_out, dist_out);
}

int mhip_apply_D(mhip_ctx *ctx, const mhip_nlist *nl, const void *f, int f_f64, int64_t nf_rows, int K,
                 int order, int weighted, float *out_d1, float *out_d2, int32_t *status) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(nl && f, "null neighbour list or field");
    MHIP_NEED(nf_rows >= 0, "negative field length");
    hipSetDevice(ctx->device);
    return apply_ops(ctx, nl, f, f_f64 ? 1 : 0, nf_rows, K, order, weighted ? 1 : 0, out_d1, out_d2,
                     status);
}

int mhip_gradient_weights(mhip_ctx *ctx, const void *pos, int pos_f64, int64_t n, int dim, const int32_t *ngb,
                          int64_t nq, int k, const double *h, const int64_t *indices, double boxsize,
                          int weighted, int order, double *G, int32_t *status) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(pos && ngb && h && indices && G, "null argument");
    hipSetDevice(ctx->device);
    return weights_rows(ctx, pos, pos_f64 ? 1 : 0, n, dim, ngb, nq, k, h, indices,
                        boxsize > 0 ? boxsize : -1.0, weighted ? 1 : 0, order, G, status);
}

int mhip_hsml_iter(mhip_ctx *ctx, const void *dists, int dists_f64, int64_t n, int k, int dim,
                   double error_norm, double *h, int32_t *status) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(dists && h, "null argument");
    hipSetDevice(ctx->device);
    return hsml_dists(ctx, dists, dists_f64 ? 1 : 0, n, k, dim, error_norm, h, status);
}

int mhip_kernel(mhip_ctx *ctx, const void *q, int f64, int64_t n, void *out) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(q && out, "null argument");
    hipSetDevice(ctx->device);
    return kernel_eval(ctx, q, f64 ? 1 : 0, n, out);
}

int mhip_kernel_weights(mhip_ctx *ctx, const void *dist, int dist_f64, const void *h, int h_f64, int64_t nq, int k,
                        double *w) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(dist && h && w, "null argument");
    hipSetDevice(ctx->device);
    return kernel_weights(ctx, dist, dist_f64 ? 1 : 0, h, h_f64 ? 1 : 0, nq, k, w);
}

int mhip_kernel_average(mhip_ctx *ctx, const mhip_nlist *nl, const void *f, int f_f64, int64_t nf_rows,
                        double *out) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(nl && f && out, "null argument");
    MHIP_NEED(nf_rows >= 0, "negative field length");
    hipSetDevice(ctx->device);
    return kernel_stat(ctx, nl, f, f_f64 ? 1 : 0, nf_rows, 1, 0, out);
}

int mhip_kernel_variance(mhip_ctx *ctx, const mhip_nlist *nl, const void *f, int f_f64, int64_t nf_rows,
                         int K, double *out) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(nl && f && out, "null argument");
    MHIP_NEED(nf_rows >= 0, "negative field length");
    hipSetDevice(ctx->device);
    return kernel_stat(ctx, nl, f, f_f64 ? 1 : 0, nf_rows, K, 1, out);
}

int mhip_integrate(mhip_ctx *ctx, const void *vol, int vol_f64, const void *f, int f_f64, int64_t n, int K,
                   double *out) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(vol && f && out, "null argument");
    hipSetDevice(ctx->device);
    return integrate(ctx, vol, vol_f64 ? 1 : 0, f, f_f64 ? 1 : 0, n, K, out);
}

int mhip_nlist_subset(mhip_ctx *ctx, const mhip_nlist *full, const int64_t *qidx, int64_t nq, const double *m_rows,
                      float *h_out, float *rho_out, float *vol_out, int32_t *status_out, mhip_nlist *nl) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(full && qidx && nl, "null argument");
    hipSetDevice(ctx->device);
    return nlist_subset(ctx, full, qidx, nq, m_rows, h_out, rho_out, vol_out, status_out, nl);
}

int mhip_surface_density(mhip_ctx *ctx, const void *f, int f_f64, const void *pos, int pos_f64, int64_t n, int dim,
                         const void *h, int h_f64, double smooth_fac, double hmin, double cx, double cy, double size,
                         int res, int conservative, double *grid_out) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(grid_out && ((f && pos && h) || n == 0), "null argument");
    hipSetDevice(ctx->device);
    return surface_density(ctx, f, f_f64 ? 1 : 0, pos, pos_f64 ? 1 : 0, n, dim, h, h_f64 ? 1 : 0, smooth_fac, hmin,
                           cx, cy, size, res, conservative, grid_out);
}

int mhip_radtransfer(mhip_ctx *ctx, const double *j, const double *m, const double *kappa, const double *x,
                     const double *h, int64_t n, int nf, int gridres, double L, double cx, double cy, const double *i0,
                     double *image_out) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(image_out && ((j && m && kappa && x && h) || n == 0), "null argument");
    hipSetDevice(ctx->device);
    return radtransfer(ctx, j, m, kappa, x, h, n, nf, gridres, L, cx, cy, i0, image_out);
}

int mhip_grid_density(mhip_ctx *ctx, const void *f, int f_f64, const void *pos, int pos_f64, int64_t n, const void *h,
                      int h_f64, double hmin, double cx, double cy, double cz, double size, int res, int periodic,
                      double *grid_out) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(grid_out && ((f && pos && h) || n == 0), "null argument");
    hipSetDevice(ctx->device);
    return grid_density(ctx, f, f_f64 ? 1 : 0, pos, pos_f64 ? 1 : 0, n, h, h_f64 ? 1 : 0, hmin, cx, cy, cz, size, res,
                        periodic, grid_out);
}

int mhip_nearest(mhip_ctx *ctx, const mhip_index *ix, const double *points, int64_t nt, int dim,
                 int64_t *nbr_out) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(ix && nbr_out && (points || nt == 0), "null argument");
    hipSetDevice(ctx->device);
    return nearest_run(ctx, ix, points, nt, dim, nbr_out);
}

int mhip_unique_inverse(mhip_ctx *ctx, const int64_t *idx, int64_t nt, int64_t n, int64_t *uniq_out,
                        int64_t *inv_out, int64_t *nuniq) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED(nuniq && ((idx && uniq_out && inv_out) || nt == 0), "null argument");
    hipSetDevice(ctx->device);
    return unique_inverse(ctx, idx, nt, n, uniq_out, inv_out, nuniq);
}

int mhip_reconstruct(mhip_ctx *ctx, const double *points, int64_t nt, int dim, double boxsize, const void *pos,
                     int pos_f64, const int64_t *nbr, const int64_t *inv, const void *f, int f_f64, int K,
                     const float *d1, const float *d2, int order, void *out) {
    if (!ctx) return MHIP_ERR_VALUE;
    MHIP_NEED((points && pos && nbr && f && out && (inv || order == 0)) || nt == 0, "null argument");
    hipSetDevice(ctx->device);
    return reconstruct(ctx, points, nt, dim, boxsize, pos, pos_f64 ? 1 : 0, nbr, inv, f, f_f64 ? 1 : 0, K, d1, d2,
                       order, out);
}

}  // extern "C"
