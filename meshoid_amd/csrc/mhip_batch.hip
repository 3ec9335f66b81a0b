// Batched form: many independent particle sets ("snapshots") of uneven size,
// each with its own des_ngb, and K fields per particle, through one HIP graph.
//
// No counterpart in the reference, which handles one Meshoid at a time
// (src/meshoid/Meshoid.py:15-170 per snapshot, then D at :237-268).  Per
// snapshot the work is the same as the single-set path (index build, kNN +
// smoothing length + density, field permutation); the derivative contraction
// of all snapshots is one launch over a device table of per-snapshot
// arguments.
//
// Execution: snapshots are spread over nbranch streams (largest first, each
// to the branch with the least work so far); the main stream forks to the
// branches with events and joins them before the batched contraction.  A run
// with use_graph captures that fork/join sequence once (after one eager run
// has sized every buffer) and replays it with hipGraphLaunch; the graph is
// recaptured when the caller passes different buffers.  Index builds run in
// deferred-check mode (no host synchronisation inside the graph); their
// bbox / NaN records are validated after each run.
//
// The contraction kernel: with K >= 8 fields, k_apply_wide (a thread group
// per query, row-major fields; group width 8, 16 or 32 threads, each thread
// summing up to 4, 2 or 1 fields per pass) and, when K is 16, 32 or 64,
// k_apply_mfma (the same with the field contraction on the
// matrix cores); with fewer, the tile kernel (2, 4, 8 or 16 fields per
// thread).  The first eager run times the number of branch streams in use
// (1, 2 or 4, at most the nbranch given: the device runs 4 hardware queues
// per process) and then every contraction candidate, and keeps the fastest of
// each.
#include <algorithm>
#include <new>
#include <numeric>
#include <string>

#include "../../include/meshoid_hip.h"
#include "mhip_internal.hpp"

struct mhip_batch {
    int nsnap = 0, K = 0, nbranch = 0;  // nbranch: streams created
    int nb_use = 0;                     // branches in use (timed on the first run)
    int force_cand = -1, force_nb = 0;  // test hook: fixed choices (-1 / 0: timed)
    float branch_ms[3] = {0, 0, 0};     // per branch count 1, 2, 4 (first run)
    std::vector<int64_t> n;
    std::vector<int> k, branch_of;
    std::vector<mhip_index *> ix;
    std::vector<mhip_nlist *> nl;
    std::vector<mhip_ctx *> br;
    std::vector<hipStream_t> streams;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    float *recs = nullptr;              // [nsnap][12] index records
    float *fsorted = nullptr;           // per snapshot [K][npad] sorted-space fields
    std::vector<int64_t> foff;          // offsets into fsorted (floats)
    mhip::ApplyArgs *tab_d = nullptr;   // [nsnap]
    int wide = 0;                       // contraction kernel family (1: row-major fields)
    int64_t *blk0_d = nullptr;          // per candidate [nsnap + 1] block prefixes
    int64_t nblocks[4] = {0, 0, 0, 0};  // per candidate
    int choice = -1;                    // chosen candidate
    float tune_ms[4] = {0, 0, 0, 0};    // per candidate
    bool warmed = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<const void *> key;      // buffers the graph was captured with
};

namespace mhip {

// contraction candidates: fields per thread for the tile kernel (256
// queries per block), threads per query for the wide and matrix-core kernels
enum CandKind { CAND_TILE = 0, CAND_WIDE = 1, CAND_MFMA = 2 };
struct Cand {
    int kind, param, qpb;
};
static const Cand TILE_CAND[4] = {{CAND_TILE, 2, 256}, {CAND_TILE, 4, 256}, {CAND_TILE, 8, 256}, {CAND_TILE, 16, 256}};
static const Cand WIDE_CAND[3] = {{CAND_WIDE, 8, 16}, {CAND_WIDE, 16, 16}, {CAND_WIDE, 32, 8}};
static bool has_mfma(const mhip_batch *b) { return b->wide && (b->K == 16 || b->K == 32 || b->K == 64); }
static int ncand(const mhip_batch *b) { return b->wide ? (has_mfma(b) ? 4 : 3) : 4; }
static Cand cand(const mhip_batch *b, int c) {
    if (!b->wide) return TILE_CAND[c];
    if (c < 3) return WIDE_CAND[c];
    return Cand{CAND_MFMA, b->K, 256 / b->K};  // K == threads per query
}

static int launch_contraction(mhip_ctx *ctx, mhip_batch *b, int c) {
    const int64_t *blk0 = b->blk0_d + (int64_t)c * (b->nsnap + 1);
    const Cand cd = cand(b, c);
    if (cd.kind == CAND_MFMA) return apply_mfma_launch(ctx, b->tab_d, blk0, b->nsnap, b->nblocks[c], cd.param);
    if (cd.kind == CAND_WIDE) {
        const int kmax = *std::max_element(b->k.begin(), b->k.end());
        return apply_wide_launch(ctx, b->tab_d, blk0, b->nsnap, b->nblocks[c], cd.param, b->K, kmax);
    }
    return apply_batch_launch(ctx, b->tab_d, blk0, b->nsnap, b->nblocks[c], cd.param, b->K);
}

// snapshots to the first nb branches: largest first, each to the least
// loaded branch
static void assign_branches(mhip_batch *b, int nb) {
    std::vector<int> order(b->nsnap);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->n[x] > b->n[y]; });
    std::vector<double> load(nb, 0.0);
    b->branch_of.assign(b->nsnap, 0);
    for (int s : order) {
        const int i = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        b->branch_of[s] = i;
        load[i] += (double)b->n[s] * b->k[s];
    }
    b->nb_use = nb;
}

static int64_t npad_of(int64_t n) { return (n + LEAF - 1) / LEAF * LEAF; }

// per-snapshot work on its branch: index, kNN + h + rho, field permutation
static int enqueue_snapshot(mhip_batch *b, int s, const float *pos, const float *f, float *h, float *rho,
                            int32_t *status) {
    mhip_ctx *c = b->br[b->branch_of[s]];
    int rc = index_build(c, pos, 0, b->n[s], 3, -1.0, b->ix[s], b->recs + 12 * s);
    if (rc) return rc;
    rc = knn_run(c, b->ix[s], nullptr, b->n[s], b->k[s], nullptr, nullptr, nullptr, h, rho, nullptr, status,
                 0, b->nl[s]);
    if (rc) return rc;
    const int64_t np = b->ix[s]->npad;
    if (b->wide)
        hipLaunchKernelGGL(k_permute_fields_rows<float>, dim3((unsigned)((np * b->K + 255) / 256)), dim3(256), 0,
                           c->stream, f, b->n[s], b->K, b->ix[s]->perm, b->n[s], np, b->fsorted + b->foff[s]);
    else
        hipLaunchKernelGGL(k_permute_fields<float>, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->stream, f,
                           b->n[s], b->K, b->ix[s]->perm, b->n[s], np, b->fsorted + b->foff[s]);
    return check_hip(c, hipGetLastError(), "batch field permutation");
}

// fork to the branches, per-snapshot work, join (the neighbour lists and
// indices get their device buffers here on the first run)
static int enqueue_snapshots(mhip_ctx *ctx, mhip_batch *b, const float *const *pos, const float *const *f,
                             float *const *h, float *const *rho, int32_t *const *status) {
    int rc = check_hip(ctx, hipEventRecord(b->fork, ctx->stream), "batch fork");
    if (rc) return rc;
    for (int i = 0; i < b->nb_use; i++)
        if ((rc = check_hip(ctx, hipStreamWaitEvent(b->streams[i], b->fork, 0), "batch fork wait"))) return rc;
    for (int s = 0; s < b->nsnap; s++) {
        if ((rc = enqueue_snapshot(b, s, pos[s], f[s], h[s], rho[s], status[s]))) {
            ctx->last_error = b->br[b->branch_of[s]]->last_error;
            return rc;
        }
    }
    for (int i = 0; i < b->nb_use; i++) {
        if ((rc = check_hip(ctx, hipEventRecord(b->join[i], b->streams[i]), "batch join"))) return rc;
        if ((rc = check_hip(ctx, hipStreamWaitEvent(ctx->stream, b->join[i], 0), "batch join wait"))) return rc;
    }
    return 0;
}

static int upload_table(mhip_ctx *ctx, mhip_batch *b, float *const *grad, int32_t *const *status) {
    std::vector<ApplyArgs> tab(b->nsnap);
    const int nc = ncand(b);
    std::vector<int64_t> blk0((size_t)nc * (b->nsnap + 1), 0);
    for (int s = 0; s < b->nsnap; s++)
        apply_args_fill(b->nl[s], b->fsorted + b->foff[s], b->K, grad[s], status[s], 1, &tab[s]);
    for (int c = 0; c < nc; c++) {
        int64_t *bc = blk0.data() + (size_t)c * (b->nsnap + 1);
        const int64_t qpb = cand(b, c).qpb;
        for (int s = 0; s < b->nsnap; s++) bc[s + 1] = bc[s] + (b->n[s] + qpb - 1) / qpb;
        b->nblocks[c] = bc[b->nsnap];
    }
    int rc = check_hip(ctx, hipMemcpy(b->tab_d, tab.data(), tab.size() * sizeof(ApplyArgs), hipMemcpyHostToDevice),
                       "batch table upload");
    if (rc) return rc;
    return check_hip(ctx, hipMemcpy(b->blk0_d, blk0.data(), blk0.size() * sizeof(int64_t), hipMemcpyHostToDevice),
                     "batch table upload");
}

static int validate(mhip_ctx *ctx, mhip_batch *b) {
    std::vector<float> r((size_t)b->nsnap * 12);
    int rc = check_hip(ctx, hipMemcpyAsync(r.data(), b->recs, r.size() * sizeof(float), hipMemcpyDeviceToHost,
                                           ctx->stream),
                       "batch record readback");
    if (rc) return rc;
    if ((rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "batch run"))) return rc;
    for (int s = 0; s < b->nsnap; s++)
        if (r[12 * s + 6] != 0.f) return fail(ctx, MHIP_ERR_VALUE, "positions contain NaN (snapshot " + std::to_string(s) + ")");
    return 0;
}

static void drop_graph(mhip_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->key.clear();
}

}  // namespace mhip

using namespace mhip;

extern "C" {

int mhip_batch_create(mhip_ctx *ctx, int nsnap, const int64_t *n, const int *k, int K, int nbranch,
                      mhip_batch **out) {
    if (!ctx || !out) return MHIP_ERR_VALUE;
    *out = nullptr;
    if (nsnap < 1 || !n || !k) return fail(ctx, MHIP_ERR_VALUE, "batch needs at least one snapshot");
    if (K < 1) return fail(ctx, MHIP_ERR_VALUE, "no field components");
    if (nbranch < 1) nbranch = 1;
    nbranch = std::min(std::min(nbranch, nsnap), 4);
    for (int s = 0; s < nsnap; s++) {
        if (k[s] < 2 || k[s] > 128) return fail(ctx, MHIP_ERR_UNSUPPORTED, "des_ngb must be in [2, 128]");
        if (n[s] < k[s]) return fail(ctx, MHIP_ERR_VALUE, "des_ngb exceeds the number of particles");
    }
    hipSetDevice(ctx->device);
    mhip_batch *b = new (std::nothrow) mhip_batch();
    if (!b) return MHIP_ERR_HIP;
    b->nsnap = nsnap;
    b->K = K;
    b->nbranch = nbranch;
    b->wide = K >= 8;  // many fields: one query per thread group, coalesced field rows
    b->n.assign(n, n + nsnap);
    b->k.assign(k, k + nsnap);
    assign_branches(b, nbranch);
    int rc = 0;
    auto bail = [&](int code) {
        mhip_batch_destroy(b);
        return code;
    };
    for (int i = 0; i < nbranch; i++) {
        hipStream_t st;
        if ((rc = check_hip(ctx, hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "batch stream"))) return bail(rc);
        b->streams.push_back(st);
        mhip_ctx *c = nullptr;
        if (mhip_ctx_create(ctx->device, st, &c) != MHIP_OK) return bail(fail(ctx, MHIP_ERR_HIP, "batch context"));
        b->br.push_back(c);
        hipEvent_t e;
        if ((rc = check_hip(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming), "batch event"))) return bail(rc);
        b->join.push_back(e);
    }
    if ((rc = check_hip(ctx, hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "batch event"))) return bail(rc);
    b->foff.assign(nsnap + 1, 0);
    for (int s = 0; s < nsnap; s++) b->foff[s + 1] = b->foff[s] + (int64_t)K * npad_of(n[s]);
    if ((rc = check_hip(ctx, hipMalloc((void **)&b->fsorted, b->foff[nsnap] * sizeof(float)), "hipMalloc (batch fields)")))
        return bail(rc);
    if ((rc = check_hip(ctx, hipMalloc((void **)&b->recs, (size_t)nsnap * 12 * sizeof(float)), "hipMalloc (batch)")))
        return bail(rc);
    if ((rc = check_hip(ctx, hipMalloc((void **)&b->tab_d, (size_t)nsnap * sizeof(ApplyArgs)), "hipMalloc (batch)")))
        return bail(rc);
    if ((rc = check_hip(ctx, hipMalloc((void **)&b->blk0_d, (size_t)4 * (nsnap + 1) * sizeof(int64_t)), "hipMalloc (batch)")))
        return bail(rc);
    for (int s = 0; s < nsnap; s++) {
        mhip_index *x = nullptr;
        mhip_nlist *l = nullptr;
        if (mhip_index_create(&x) != MHIP_OK || mhip_nlist_create(&l) != MHIP_OK) {
            mhip_index_destroy(x);
            mhip_nlist_destroy(l);
            return bail(fail(ctx, MHIP_ERR_HIP, "batch index"));
        }
        b->ix.push_back(x);
        b->nl.push_back(l);
    }
    *out = b;
    return MHIP_OK;
}

int mhip_batch_run(mhip_ctx *ctx, mhip_batch *b, const float *const *pos, const float *const *f,
                   float *const *h, float *const *rho, float *const *grad, int32_t *const *status,
                   int use_graph) {
    if (!ctx || !b) return MHIP_ERR_VALUE;
    if (!pos || !f || !h || !rho || !grad || !status) return fail(ctx, MHIP_ERR_VALUE, "null argument");
    for (int s = 0; s < b->nsnap; s++)
        if (!pos[s] || !f[s] || !h[s] || !rho[s] || !grad[s] || !status[s])
            return fail(ctx, MHIP_ERR_VALUE, "null snapshot buffer");
    hipSetDevice(ctx->device);
    std::vector<const void *> key;
    for (int s = 0; s < b->nsnap; s++) {
        const void *p[6] = {pos[s], f[s], h[s], rho[s], grad[s], status[s]};
        key.insert(key.end(), p, p + 6);
    }
    int rc = 0;
    const int prof = ctx->prof;
    if (!b->warmed || !use_graph) {
        // eager run: sizes every buffer; the first one also picks the field
        // tile.  The argument table holds the lists' device pointers, which
        // exist once the snapshots are enqueued.
        if ((rc = enqueue_snapshots(ctx, b, pos, f, h, rho, status))) return rc;
        if (b->choice < 0) {
            // branch count: the snapshots' work timed with 1, 2 and 4
            // streams in use (the buffers exist after the run above)
            if ((rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "batch warm-up"))) return rc;
            int best_nb = b->nbranch;
            float best_t = 0.f;
            const int nbs[3] = {1, 2, 4};
            for (int t = 0; t < 3; t++) {
                b->branch_ms[t] = 0.f;
                if (nbs[t] > b->nbranch || (b->force_nb && nbs[t] != b->force_nb)) continue;
                assign_branches(b, nbs[t]);
                hipEvent_t t0, t1;
                (void)hipEventCreate(&t0);
                (void)hipEventCreate(&t1);
                (void)hipEventRecord(t0, ctx->stream);
                rc = enqueue_snapshots(ctx, b, pos, f, h, rho, status);
                (void)hipEventRecord(t1, ctx->stream);
                if (!rc) rc = check_hip(ctx, hipEventSynchronize(t1), "batch branch timing");
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, t0, t1);
                (void)hipEventDestroy(t0);
                (void)hipEventDestroy(t1);
                if (rc) return rc;
                b->branch_ms[t] = ms;
                if (best_t == 0.f || ms < best_t) {
                    best_t = ms;
                    best_nb = nbs[t];
                }
            }
            assign_branches(b, best_nb);
        }
        if ((rc = upload_table(ctx, b, grad, status))) return rc;
        if (b->choice < 0) {
            if ((rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "batch warm-up"))) return rc;
            hipEvent_t e0, e1;
            (void)hipEventCreate(&e0);
            (void)hipEventCreate(&e1);
            float best = 0.f;
            for (int c = 0; c < ncand(b); c++) {
                if (b->force_cand >= 0 && c != b->force_cand) continue;
                float ms = 0.f;
                for (int rep = 0; rep < 3; rep++) {
                    (void)hipEventRecord(e0, ctx->stream);
                    rc = launch_contraction(ctx, b, c);
                    (void)hipEventRecord(e1, ctx->stream);
                    if (rc) break;
                    (void)hipEventSynchronize(e1);
                    float t = 0.f;
                    (void)hipEventElapsedTime(&t, e0, e1);
                    if (rep > 0) ms += t / 2;
                }
                if (rc) break;
                b->tune_ms[c] = ms;
                if (b->choice < 0 || ms < best) {
                    b->choice = c;
                    best = ms;
                }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            if (rc) return rc;
        }
        if ((rc = launch_contraction(ctx, b, b->choice))) return rc;
        b->warmed = true;
        return validate(ctx, b);
    }
    if (!b->exec || key != b->key) {
        drop_graph(b);
        if ((rc = upload_table(ctx, b, grad, status))) return rc;
        ctx->prof = 0;  // no timing events inside the graph
        rc = check_hip(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "batch capture");
        if (!rc) {
            int rce = enqueue_snapshots(ctx, b, pos, f, h, rho, status);
            if (!rce) rce = launch_contraction(ctx, b, b->choice);
            hipGraph_t g = nullptr;
            const int rc2 = check_hip(ctx, hipStreamEndCapture(ctx->stream, &g), "batch capture end");
            rc = rce ? rce : rc2;
            if (!rc) {
                b->graph = g;
                rc = check_hip(ctx, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "graph instantiate");
            } else if (g) {
                (void)hipGraphDestroy(g);
            }
        }
        ctx->prof = prof;
        if (rc) {
            drop_graph(b);
            return rc;
        }
        b->key = key;
    }
    if ((rc = check_hip(ctx, hipGraphLaunch(b->exec, ctx->stream), "graph launch"))) return rc;
    return validate(ctx, b);
}

int mhip_batch_info(const mhip_batch *b, int *wide, int *choice, int *cand4, float *tune_ms4, int *nbranch) {
    if (!b) return MHIP_ERR_VALUE;
    if (wide) *wide = b->wide;
    if (choice) *choice = b->choice >= 0 ? cand(b, b->choice).param : 0;
    for (int c = 0; c < 4; c++) {
        const bool in = c < ncand(b);
        if (cand4) cand4[c] = in ? cand(b, c).param : 0;
        if (tune_ms4) tune_ms4[c] = in ? b->tune_ms[c] : 0.f;
    }
    if (nbranch) *nbranch = b->nb_use;
    return MHIP_OK;
}

int mhip_batch_tuning(const mhip_batch *b, int *ncand_out, int *kind4, int *param4, float *ms4, int *choice,
                      int *nb_use, float *branch_ms3) {
    if (!b) return MHIP_ERR_VALUE;
    const int nc = ncand(b);
    if (ncand_out) *ncand_out = nc;
    for (int c = 0; c < 4; c++) {
        const bool in = c < nc;
        if (kind4) kind4[c] = in ? cand(b, c).kind : -1;
        if (param4) param4[c] = in ? cand(b, c).param : 0;
        if (ms4) ms4[c] = in ? b->tune_ms[c] : 0.f;
    }
    if (choice) *choice = b->choice;
    if (nb_use) *nb_use = b->nb_use;
    for (int t = 0; t < 3; t++)
        if (branch_ms3) branch_ms3[t] = b->branch_ms[t];
    return MHIP_OK;
}

int mhip_batch_force(mhip_batch *b, int cand_index, int nbranch) {
    if (!b) return MHIP_ERR_VALUE;
    if (b->warmed) return MHIP_ERR_VALUE;  // choices are made on the first run
    if (cand_index >= ncand(b)) return MHIP_ERR_VALUE;
    if (nbranch != 0 && nbranch != 1 && nbranch != 2 && nbranch != 4) return MHIP_ERR_VALUE;
    if (nbranch > b->nbranch) return MHIP_ERR_VALUE;
    b->force_cand = cand_index;
    b->force_nb = nbranch;
    return MHIP_OK;
}

int mhip_batch_destroy(mhip_batch *b) {
    if (!b) return MHIP_OK;
    drop_graph(b);
    for (auto c : b->br) mhip_ctx_destroy(c);
    for (auto s : b->streams) (void)hipStreamDestroy(s);
    for (auto e : b->join) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (auto x : b->ix) mhip_index_destroy(x);
    for (auto l : b->nl) mhip_nlist_destroy(l);
    if (b->fsorted) (void)hipFree(b->fsorted);
    if (b->recs) (void)hipFree(b->recs);
    if (b->tab_d) (void)hipFree(b->tab_d);
    if (b->blk0_d) (void)hipFree(b->blk0_d);
    delete b;
    return MHIP_OK;
}

}  // extern "C"
