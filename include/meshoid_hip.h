/*
 * meshoid_hip.h -- C ABI of libmeshoid_hip.so, the MI355X (gfx950) backend of
 * meshoid's neighbour-weighted operator path.
 *
 * All array arguments are DEVICE pointers unless stated otherwise; sizes are
 * element counts.  Calls are ordered on the context's stream; functions that
 * must read a result on the host (index build: bounding box / range checks)
 * synchronise that stream themselves.  Every function returns MHIP_OK or a
 * negative MHIP_ERR_* code; mhip_last_error() gives the message.  Contexts are
 * not thread-safe: use one per host thread / stream.
 *
 * The reference (meshoid 1.46, pure Python + Numba) has no C interface; each
 * entry point below names the reference function or call it replaces.
 */
#ifndef MESHOID_HIP_H
#define MESHOID_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MHIP_OK 0
#define MHIP_ERR_VALUE -1       /* invalid argument / data      -> ValueError   */
#define MHIP_ERR_HIP -2         /* HIP runtime failure          -> RuntimeError */
#define MHIP_ERR_SINGULAR -3    /* singular moment matrix       -> LinAlgError  */
#define MHIP_ERR_UNSUPPORTED -4 /* outside the implemented range -> NotImplementedError */
#define MHIP_ERR_KEY -5         /* dim not in {1,2,3}           -> KeyError     */

/* per-row status bits (status arrays) */
#define MHIP_ST_AMBIGUOUS 1      /* fp32 could not order k-th/(k+1)-th: fp64 resolved */
#define MHIP_ST_BAND_OVERFLOW 2  /* more fp64 ties at the k-th distance than held     */
#define MHIP_ST_HSML_NOCONV 4    /* smoothing-length iteration hit its cap            */
#define MHIP_ST_SINGULAR 8       /* moment matrix exactly singular                    */
#define MHIP_ST_COUNT 16         /* neighbour collection did not return k members     */
#define MHIP_ST_SEED_MISS 32     /* learned seed rejected, bracket used               */
#define MHIP_ST_RANKDEF 64       /* moment matrix rank-deficient by rounding: row NaN */
#define MHIP_ST_RADIUS_MISS 128  /* search-radius seed too small: recomputed without  */

/* mhip_knn_hsml flags */
#define MHIP_KNN_KEEP_DIST 1 /* keep squared distances (needed by mhip_nlist_export) */
#define MHIP_KNN_SMALL_RECORD 2 /* test hook: cap each wavefront's leaf record at 8 entries,
                                  * so that every wavefront collects its members by the
                                  * overflow path (a second walk); results are unchanged */

typedef struct mhip_ctx mhip_ctx;
typedef struct mhip_index mhip_index;
typedef struct mhip_nlist mhip_nlist;
typedef struct mhip_batch mhip_batch;

/* library version, e.g. 10000 for 1.0.0 */
int mhip_version(void);

/* Context: device ordinal + stream (NULL = the default stream of that device). */
int mhip_ctx_create(int device, void *stream, mhip_ctx **out);
int mhip_ctx_destroy(mhip_ctx *ctx);
int mhip_ctx_set_stream(mhip_ctx *ctx, void *stream);
int mhip_ctx_synchronize(mhip_ctx *ctx);
const char *mhip_last_error(const mhip_ctx *ctx);

/*
 * Kernel timing (no counterpart in the reference, whose notebooks use %time).
 * level 1: every launch region records HIP events on the context stream;
 * level 2: also per-wavefront work counters of the kNN passes; 0: off.
 * mhip_prof_read synchronises, writes one line per region,
 * "name launches total_ms\n" (counters: "stat_name count 0\n"), into buf and
 * clears the records.
 */
int mhip_prof_enable(mhip_ctx *ctx, int level);
int mhip_prof_read(mhip_ctx *ctx, char *buf, int64_t buflen);

/*
 * Spatial index over n particles.  Replaces BuildTree / cKDTree(pos, boxsize)
 * (src/meshoid/Meshoid.py:142-145).  pos: (n, dim) row-major, float32
 * (pos_f64 = 0) or float64 (pos_f64 = 1); the buffer must stay alive while the
 * index is used (float64 neighbour-boundary resolution reads it).  boxsize > 0
 * makes the space periodic on [0, boxsize)^dim; data outside raises
 * MHIP_ERR_VALUE as cKDTree raises ValueError.  Rebuilding an index reuses its
 * buffers when they are large enough.
 */
int mhip_index_create(mhip_index **out);
int mhip_index_build(mhip_ctx *ctx, mhip_index *ix, const void *pos, int pos_f64, int64_t n,
                     int dim, double boxsize);
int mhip_index_destroy(mhip_index *ix);
/* host-side copy of n, dim, boxsize and the bounding box (lo xyz, hi xyz) */
int mhip_index_info(const mhip_index *ix, int64_t *n, int *dim, double *boxsize, float *bbox6);
/* the index's particle order: perm_out uint32[n] (device), the original row
 * of each sorted position (Hilbert order of the positions' cells in the cubic
 * frame of the bounding box or periodic box, 21 bits per axis, stable) */
int mhip_index_order(mhip_ctx *ctx, const mhip_index *ix, uint32_t *perm_out);

/*
 * Exact kNN (k = des_ngb, self included) + smoothing length + density for the
 * query rows.  Replaces TreeUpdate (src/meshoid/Meshoid.py:147-170):
 * tree.query(pos[mask], des_ngb) (:157), HsmlIter(ngbdist, error_norm=1e-13,
 * dim) (:165, src/meshoid/kernel_density.py:14-57) and
 * density = des_ngb*m/(volnorm*h^dim), vol = m/density (:169-170).
 *   qidx     int64[nq] particle index of each row (particle_mask); NULL = all
 *            particles in order (nq is then ignored and taken as n)
 *   m_rows   float64[nq] mass of each row; NULL = 1
 *   seed_h   float32[nq] initial smoothing-length guess per row for the Newton
 *            solve (learned seed); NULL = none.  A guess outside the bracket
 *            is ignored (MHIP_ST_SEED_MISS).
 *   seed_r   float32[nq] upper bound guess of the (k+1)-th neighbour distance
 *            per row; NULL = none.  The search starts with that radius; rows
 *            where it was too small are recomputed without it
 *            (MHIP_ST_RADIUS_MISS).  Results never depend on the seeds.
 *   h_out, rho_out, vol_out   float32[nq], any may be NULL
 *   status_out int32[nq] MHIP_ST_* bits, may be NULL
 *   nl       neighbour list filled for later mhip_nlist_export / mhip_apply_D
 * Supports 2 <= k <= 128 and k <= n.
 */
int mhip_nlist_create(mhip_nlist **out);
int mhip_nlist_destroy(mhip_nlist *nl);
int mhip_knn_hsml(mhip_ctx *ctx, const mhip_index *ix, const int64_t *qidx, int64_t nq, int k,
                  const double *m_rows, const float *seed_h, const float *seed_r, float *h_out,
                  float *rho_out, float *vol_out, int32_t *status_out, int flags, mhip_nlist *nl);

/*
 * Neighbour rows as cKDTree.query returns them (src/meshoid/Meshoid.py:157):
 * ngb_out int32[nq*k] original particle indices, dist_out float32[nq*k]
 * distances, each row ascending by distance (ties by index).  Either output
 * may be NULL.  Needs a list built with MHIP_KNN_KEEP_DIST.
 */
int mhip_nlist_export(mhip_ctx *ctx, const mhip_nlist *nl, int32_t *ngb_out, float *dist_out);

/*
 * The same rows with float64 distances, recomputed from the positions the
 * index was built from (the caller's buffer must still hold them): the
 * float64 distances cKDTree.query returns (src/meshoid/Meshoid.py:157), rows
 * ascending by (distance, index).  Works with or without MHIP_KNN_KEEP_DIST.
 */
int mhip_nlist_export_f64(mhip_ctx *ctx, const mhip_nlist *nl, int32_t *ngb_out, double *dist_out);

/*
 * The smoothing lengths of the list's rows in float64, h_out double[nq] in
 * row order: HsmlIter's root (src/meshoid/kernel_density.py:14-57, as
 * TreeUpdate stores it at src/meshoid/Meshoid.py:165), refined with one
 * float64 Newton step on the exact float64 distances of the final
 * neighbours (h_out of mhip_knn_hsml is this value rounded to float32).
 * The derivative and kernel-statistics entry points fit with it.
 */
int mhip_nlist_hsml(mhip_ctx *ctx, const mhip_nlist *nl, double *h_out);

/*
 * Least-squares derivative operators over the list's rows.  Replaces
 * Meshoid.D (src/meshoid/Meshoid.py:237-268), Meshoid.D2 (:270-295) with
 * ComputeDWeights (:110-140) and gradient_weights
 * (src/meshoid/derivatives.py:117-138), fused: the weights are never stored.
 *   f        (nf_rows, K) row-major, float32 (f_f64 = 0) or float64; row r is
 *            the value at original particle r (the reference's np.take(f, ngb))
 *   order    1: out_d1 = gradient (nq, K, dim)
 *            2: quadratic fit; out_d1 = first derivatives (nq, K, dim) as
 *               D(f, order=2), out_d2 = second derivatives (nq, K, nd2) with
 *               nd2 = 6 in 3D (xx,yy,zz,xy,yz,zx), 3 in 2D, 1 in 1D, as D2(f)
 *   weighted kernel-weighted (1) or unweighted (0) least squares
 *   status   int32[nq] per row (MHIP_ST_SINGULAR / MHIP_ST_RANKDEF, as for
 *            mhip_gradient_weights), may be NULL
 * Outputs are float32; either may be NULL.
 */
int mhip_apply_D(mhip_ctx *ctx, const mhip_nlist *nl, const void *f, int f_f64, int64_t nf_rows,
                 int K, int order, int weighted, float *out_d1, float *out_d2, int32_t *status);

/*
 * gradient_weights(pos, ngb, kernel_radius, indices, boxsize, weighted, order)
 * (src/meshoid/derivatives.py:117-138) on arbitrary row-major neighbour lists:
 * pos (n, dim) float32/float64, ngb int32[nq*k], h float64[nq],
 * indices int64[nq]; G float64[nq * nd * k] with nd = dim (order 1) or
 * {1:2, 2:5, 3:9}[dim] (order 2).  status int32[nq] may be NULL:
 * MHIP_ST_SINGULAR when the moment matrix has an exactly zero diagonal entry
 * (np.linalg.inv raises there; the Python layer raises LinAlgError),
 * MHIP_ST_RANKDEF when the LDL^T meets a zero pivot (the row is NaN; LU
 * returns unbounded values there).
 */
int mhip_gradient_weights(mhip_ctx *ctx, const void *pos, int pos_f64, int64_t n, int dim,
                          const int32_t *ngb, int64_t nq, int k, const double *h,
                          const int64_t *indices, double boxsize, int weighted, int order,
                          double *G, int32_t *status);

/*
 * HsmlIter(neighbor_dists, dim, error_norm) (src/meshoid/kernel_density.py:14-57)
 * on a given (n, k) distance array (float32/float64, rows ascending):
 * h float64[n]; status int32[n] (MHIP_ST_HSML_NOCONV) may be NULL.
 */
int mhip_hsml_iter(mhip_ctx *ctx, const void *dists, int dists_f64, int64_t n, int k, int dim,
                   double error_norm, double *h, int32_t *status);

/* Kernel(q) ufunc (src/meshoid/kernel_density.py:82-96), elementwise, same dtype */
int mhip_kernel(mhip_ctx *ctx, const void *q, int f64, int64_t n, void *out);

/*
 * Meshoid.get_kernel_weights (src/meshoid/Meshoid.py:172-177): W(d/h)
 * normalised per row, for a row-major (nq, k) distance array (the exported
 * ngbdist; float32, or float64 with dist_f64 = 1) and h[nq] (float32, or
 * float64 with h_f64 = 1); w float64[nq*k].  A row with a zero kernel sum is
 * NaN, as in the reference.
 */
int mhip_kernel_weights(mhip_ctx *ctx, const void *dist, int dist_f64, const void *h, int h_f64, int64_t nq, int k,
                        double *w);

/*
 * Meshoid.KernelAverage(f) (src/meshoid/Meshoid.py:369-381): per query row,
 * sum_j W(r_j/h) f[ngb_j] / sum_j W(r_j/h) over the list's neighbours, scalar
 * f (nf_rows) float32/float64 in original row order; out float64[nq] in row
 * order.  Pair distances are recomputed in float64.
 */
int mhip_kernel_average(mhip_ctx *ctx, const mhip_nlist *nl, const void *f, int f_f64, int64_t nf_rows,
                        double *out);

/*
 * Meshoid.KernelVariance(f) (src/meshoid/Meshoid.py:352-367): population
 * standard deviation np.std(f[ngb], axis=1) of each of the K components of
 * f (nf_rows, K); out float64[nq*K], rows in row order.
 */
int mhip_kernel_variance(mhip_ctx *ctx, const mhip_nlist *nl, const void *f, int f_f64, int64_t nf_rows,
                         int K, double *out);

/*
 * Meshoid.Integrate(f) (src/meshoid/Meshoid.py:333-350): sum_i vol_i f_i per
 * component, vol[n] float32 (vol_f64 = 0) or float64, f (n, K)
 * float32/float64 (rows already masked); out float64[K].  Deterministic
 * two-level float64 reduction.
 */
int mhip_integrate(mhip_ctx *ctx, const void *vol, int vol_f64, const void *f, int f_f64, int64_t n, int K,
                   double *out);

/*
 * mhip_integrate without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, same result bit for bit as mhip_integrate.  The partial
 * sums live in a library-owned static device array of
 * MHIP_NOALLOC_INTEGRATE_MAX_K x 1024 doubles (512 KiB, held by every device
 * the library runs on until the process ends); K above
 * MHIP_NOALLOC_INTEGRATE_MAX_K is rejected with MHIP_ERR_VALUE.  Calls that
 * overlap on one device share that array and race: order them on one stream.
 */
#define MHIP_NOALLOC_INTEGRATE_MAX_K 64
int mhip_integrate_noalloc(mhip_ctx *ctx, const void *vol, int vol_f64, const void *f, int f_f64, int64_t n, int K,
                           double *out);

/*
 * GridSurfaceDensity (src/meshoid/grid_deposition.py:123-295), behind
 * Meshoid.SurfaceDensity (src/meshoid/Meshoid.py:595-643): the surface
 * density of f on a res x res grid of sightlines along z, grid_out float64
 * [gx][gy] with points at center - size/2 + g * size/(res-1).  f (n)
 * float32/float64, pos (n, dim) with dim 2 or 3 (x, y used), h (n); the
 * smoothing length used is min(max(smooth_fac * h, hmin), 1e100)
 * (SurfaceDensity: hmin = 2 size / res; GridSurfaceDensity itself: 1, 0);
 * conservative = 1: each particle's weights normalised to f, grid / dx^2
 * (GridSurfaceDensity_conservative_core).  Accumulated in float64.
 * res in [2, 2^20].  Uses context scratch of about 28 n bytes.
 */
int mhip_surface_density(mhip_ctx *ctx, const void *f, int f_f64, const void *pos, int pos_f64, int64_t n, int dim,
                         const void *h, int h_f64, double smooth_fac, double hmin, double cx, double cy, double size,
                         int res, int conservative, double *grid_out);

/*
 * radtransfer (src/meshoid/radiation/radtransfer.py:11-177): intensity along
 * z through emitting, absorbing particles on a gridres x gridres grid of
 * sightlines at center + (-L/2 + g L/(gridres-1)) in x and y.  All arrays
 * float64: j (n, nf) emissivity per unit mass, m (n), kappa (n, nf) opacity,
 * x (n, 3), h (n); i0 (gridres, gridres, nf) background or NULL (0).  The
 * particles act farthest first (largest z first), as
 * radtransfer_singlethread orders them.  image_out float64 (gridres,
 * gridres, nf).  nf <= 64.
 */
int mhip_radtransfer(mhip_ctx *ctx, const double *j, const double *m, const double *kappa, const double *x,
                     const double *h, int64_t n, int nf, int gridres, double L, double cx, double cy, const double *i0,
                     double *image_out);

/*
 * GridDensity (src/meshoid/grid_deposition.py:534-603), behind
 * Meshoid.DepositToGrid (src/meshoid/Meshoid.py:557-593): the density of the
 * conserved quantity f on a res^3 grid, grid_out float64 [gx][gy][gz] with
 * points at center - size/2 + dx/2 + g dx, dx = size/(res-1).  Each particle
 * spreads f over its footprint with weights kernel3d(r/h) / (their sum); the
 * map is divided by dx^3.  pos (n, 3), h (n) clipped below at hmin
 * (DepositToGrid: size/(res-1)); periodic = 1 wraps indices (g mod res) as
 * the reference does for box_size >= 0, else they are clipped.
 */
int mhip_grid_density(mhip_ctx *ctx, const void *f, int f_f64, const void *pos, int pos_f64, int64_t n, const void *h,
                      int h_f64, double hmin, double cx, double cy, double cz, double size, int res, int periodic,
                      double *grid_out);

/*
 * TreeUpdate with a particle_mask (src/meshoid/Meshoid.py:147-170) served
 * from an unmasked list `full` of the same index and des_ngb: a particle's
 * neighbours, distances and smoothing length do not depend on the mask, so
 * the rows qidx int64[nq] of `full` are copied into `nl` (a list other than
 * `full`) and density/volume recomputed with m_rows float64[nq] (NULL = 1),
 * as mhip_knn_hsml with qidx would give them.  h_out/rho_out/vol_out float32
 * and status_out int32 [nq] in row order, each may be NULL.
 */
int mhip_nlist_subset(mhip_ctx *ctx, const mhip_nlist *full, const int64_t *qidx, int64_t nq, const double *m_rows,
                      float *h_out, float *rho_out, float *vol_out, int32_t *status_out, mhip_nlist *nl);

/*
 * Meshoid.Reconstruct (src/meshoid/Meshoid.py:383-440), in three steps.
 *
 * mhip_nearest: the nearest particle of each point, tree.query(points)[1]
 *   (Meshoid.py:409, k = 1).  points float64 (nt, dim), dim the index's;
 *   periodic indices wrap the points into [0, boxsize).  The pick is exact in
 *   float64 (the distance cKDTree computes); at an exact float64 tie the
 *   lowest particle index wins.  nbr_out int64[nt], original particle index,
 *   -1 for a point with non-finite coordinates.
 *
 * mhip_unique_inverse: np.unique(idx, return_inverse=True) (Meshoid.py:410)
 *   for idx int64[nt] in [0, n): uniq_out int64[n] (capacity) receives the
 *   sorted distinct indices, inv_out int64[nt] the position of each idx in
 *   uniq_out, *nuniq (host) their count.  Synchronises the stream.  An idx
 *   outside [0, n) is an error.
 *
 * mhip_reconstruct: f_target (Meshoid.py:413-440) for order 0 (f[nbr]), 1
 *   (+ dx . D) or 2 (3D only: + 0.5 dx_a^2 D2_aa + dx_a dx_(a+1) D2_a,a+1, the
 *   reference's term order).  pos: the particle positions (float32/float64,
 *   (n, dim)); dx = point - pos[nbr], nearest image when boxsize > 0.  f (n, K)
 *   float32/float64; d1 float32 (nuniq, K, dim) and d2 float32 (nuniq, K, 6)
 *   are D(f, order) and D2(f) on the rows uniq (NULL when the order does not
 *   use them); out (nt, K) in f's dtype, every partial sum rounded to it as
 *   the reference's in-place additions do.
 */
int mhip_nearest(mhip_ctx *ctx, const mhip_index *ix, const double *points, int64_t nt, int dim,
                 int64_t *nbr_out);
int mhip_unique_inverse(mhip_ctx *ctx, const int64_t *idx, int64_t nt, int64_t n, int64_t *uniq_out,
                        int64_t *inv_out, int64_t *nuniq);
int mhip_reconstruct(mhip_ctx *ctx, const double *points, int64_t nt, int dim, double boxsize, const void *pos,
                     int pos_f64, const int64_t *nbr, const int64_t *inv, const void *f, int f_f64, int K,
                     const float *d1, const float *d2, int order, void *out);

/*
 * Batched form (no counterpart in the reference, which runs one Meshoid at a
 * time: Meshoid.py:15-170 then D at :237-268 per snapshot).  nsnap particle
 * sets of n[s] float32 positions (n[s], 3) each, with des_ngb k[s] and K
 * float32 fields (n[s], K); per snapshot: h, rho float32[n[s]], the weighted
 * gradient of every field grad float32[n[s] * K * 3] (as mhip_apply_D, order
 * 1) and status int32[n[s]] (kNN bits OR derivative bits).  Snapshots are
 * spread over up to nbranch streams (at most 4); with use_graph the whole
 * batch (fork/join of the branches + one contraction launch over all
 * snapshots) is captured once into a HIP graph and replayed (recaptured when
 * the buffers change).  The first run is eager and times the number of
 * branches in use (1, 2 or 4, at most nbranch) and then the contraction's
 * candidates (K >= 8: a group of 8/16/32 threads per query over row-major
 * fields, each thread summing 4/2/1 fields per pass, wide = 1, plus, for
 * K = 16, 32 or 64, the matrix-core kernel;
 * fewer fields: 2/4/8/16 fields per thread); mhip_batch_info and
 * mhip_batch_tuning report them and the choices.  Non-periodic sets only.
 */
int mhip_batch_create(mhip_ctx *ctx, int nsnap, const int64_t *n, const int *k, int K, int nbranch,
                      mhip_batch **out);
int mhip_batch_run(mhip_ctx *ctx, mhip_batch *b, const float *const *pos, const float *const *f,
                   float *const *h, float *const *rho, float *const *grad, int32_t *const *status,
                   int use_graph);
int mhip_batch_info(const mhip_batch *b, int *wide, int *choice, int *candidates4, float *tune_ms4,
                    int *nbranch);
/* The first run's timings: ncand candidates (<= 4) with kind4 (0 tile, 1
 * wide, 2 matrix-core), param4 (fields per thread or threads per query) and
 * ms4; choice = index of the candidate in use; nb_use = branches in use;
 * branch_ms3 = the snapshots' work with 1, 2 and 4 branches (0: not timed).
 * Any output may be NULL. */
int mhip_batch_tuning(const mhip_batch *b, int *ncand, int *kind4, int *param4, float *ms4, int *choice,
                      int *nb_use, float *branch_ms3);
/* Test hook, before the first run: use candidate cand_index (-1: timed) and
 * nbranch branches (0: timed; else 1, 2 or 4, at most the batch's streams). */
int mhip_batch_force(mhip_batch *b, int cand_index, int nbranch);
int mhip_batch_destroy(mhip_batch *b);

/*
 * Learned smoothing-length seed (no counterpart in the reference, whose
 * HsmlIter always starts from its bracket, kernel_density.py:37-38).
 * Features per particle: log(1 + count) of the 3x3x3 coarse cells (1/128 of
 * the index frame) around it + log(cell size); network 28-64-64-1 SiLU with
 * bfloat16 weights and float32 accumulation, predicting log(h / cell).
 *   mhip_seed_features   F float32[n*28], raw features in original row order
 *   mhip_seed_predict    h_rows float32[n] predicted smoothing length per
 *                        original row (BF16 MFMA inference); blob from
 *                        mhip_mlp_pack, mhip_mlp_blob_bytes() bytes
 *   mhip_mlp_train_step  gradient of the mean squared error over nb samples
 *                        (F float32[nb*28], T float32[nb] targets log(h/cell))
 *                        accumulated into grad float32[6081] (zero on entry)
 *                        and loss float32[1]; step > 0 then applies Adam step
 *                        `step` with learning rate lr to P float32[6081]
 *                        (moments m, v) and clears grad.  Parameter order:
 *                        W1 (64x28), b1, W2 (64x64), b2, w3 (64), b3.
 */
int mhip_seed_features(mhip_ctx *ctx, const mhip_index *ix, float *F);
int mhip_seed_predict(mhip_ctx *ctx, const mhip_index *ix, const void *blob, float *h_rows);
int mhip_mlp_train_step(mhip_ctx *ctx, const float *F, const float *T, int64_t nb, const float *mu,
                        const float *sd, float *P, float *m, float *v, float *grad, float *loss,
                        int step, float lr);
int mhip_mlp_pack(mhip_ctx *ctx, const float *P, const float *mu, const float *sd, void *blob);
int64_t mhip_mlp_blob_bytes(void);

#ifdef __cplusplus
}
#endif
#endif /* MESHOID_HIP_H */
