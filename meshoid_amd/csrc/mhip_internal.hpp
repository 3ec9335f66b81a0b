// Internal (non-exported) launch helpers shared by the .hip translation units.
#pragma once
#include "mhip_common.hpp"
#include "mhip_sfc.hpp"

namespace mhip {

// error plumbing: every launcher returns 0 or a negative MHIP_ERR_* code and
// records a message in ctx->last_error
int fail(mhip_ctx *ctx, int code, const std::string &msg);
int check_hip(mhip_ctx *ctx, hipError_t e, const char *what);
void *scratch(mhip_ctx *ctx, size_t bytes, int *rc);

// Stable LSD radix sort of (u64 key, u32 value) pairs on bits [lo_bit, bits)
// of the key (lo_bit a multiple of 8).  Buffers: k0/v0 hold the input; the
// sorted result is returned in *kout/*vout (one of the two ping-pong
// buffers).  counts needs 256*ceil(n/4096)+256 u32.
int radix_sort_pairs(mhip_ctx *ctx, uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1,
                     uint32_t *counts, int64_t n, int bits, uint64_t **kout, uint32_t **vout, int lo_bit = 0);
size_t radix_counts_words(int64_t n);

// exclusive prefix sum of in[0, n) into out (uint32: the total must stay below
// 2^32); work holds scan_work_words(n) words.  Returns the device address of
// the total (inside work).
uint32_t *exclusive_scan_u32(mhip_ctx *ctx, const uint32_t *in, int64_t n, uint32_t *out, uint32_t *work);
size_t scan_work_words(int64_t n);

// spatial index build (bbox -> keys -> sort -> reorder -> boxes).  With rec
// (device, 12 floats: bbox lo/hi, NaN flag, pad, frame) the build neither
// synchronises nor checks: the record is left for the caller to validate
// (float32 positions only, since the float64 box-test slack needs the bbox).
int index_build(mhip_ctx *ctx, const void *pos, int pos_f64, int64_t n, int dim, double box,
                mhip_index *ix, float *rec = nullptr);

// query setup + pass1/pass2/resolve
int knn_run(mhip_ctx *ctx, const mhip_index *ix, const int64_t *qidx, int64_t nq, int k,
            const double *m_rows, const float *seed_h, const float *seed_r, float *h_out,
            float *rho_out, float *vol_out, int32_t *status_out, int flags, mhip_nlist *nl);

// learned smoothing-length seed (mhip_seed.hip)
int seed_features(mhip_ctx *ctx, const mhip_index *ix, float *F);
int seed_predict(mhip_ctx *ctx, const mhip_index *ix, const void *blob, float *h_rows);
int mlp_train_step(mhip_ctx *ctx, const float *F, const float *T, int64_t nb, const float *mu,
                   const float *sd, float *P, float *m, float *v, float *grad, float *loss, int step,
                   float lr);
int mlp_pack(mhip_ctx *ctx, const float *P, const float *mu, const float *sd, void *blob);
size_t mlp_blob_bytes();

int nlist_export(mhip_ctx *ctx, const mhip_nlist *nl, int32_t *ngb_out, float *dist_out);
int nlist_export64(mhip_ctx *ctx, const mhip_nlist *nl, int32_t *ngb_out, double *dist_out);
int nlist_hsml(mhip_ctx *ctx, const mhip_nlist *nl, double *h_out);
int nlist_subset(mhip_ctx *ctx, const mhip_nlist *full, const int64_t *qidx, int64_t nq, const double *m_rows,
                 float *h_out, float *rho_out, float *vol_out, int32_t *status_out, mhip_nlist *nl);

// nearest particle of arbitrary points (mhip_knn.hip) and the reconstruction
// around it (mhip_recon.hip)
int nearest_run(mhip_ctx *ctx, const mhip_index *ix, const double *targets, int64_t nt, int dim,
                int64_t *out);
int unique_inverse(mhip_ctx *ctx, const int64_t *nbr, int64_t nt, int64_t n, int64_t *uniq, int64_t *inv,
                   int64_t *nuniq);
int surface_density(mhip_ctx *ctx, const void *f, int f_f64, const void *pos, int pos_f64, int64_t n, int dim,
                    const void *h, int h_f64, double smooth, double hmin, double cx, double cy, double size, int res,
                    int conservative, double *grid);
int radtransfer(mhip_ctx *ctx, const double *j, const double *m, const double *kappa, const double *x,
                const double *h, int64_t n, int nf, int res, double L, double cx, double cy, const double *i0,
                double *out);
int grid_density(mhip_ctx *ctx, const void *f, int f_f64, const void *pos, int pos_f64, int64_t n, const void *h,
                 int h_f64, double hmin, double cx, double cy, double cz, double size, int res, int periodic,
                 double *grid);
int reconstruct(mhip_ctx *ctx, const double *targets, int64_t nt, int dim, double box, const void *pos,
                int pos_f64, const int64_t *nbr, const int64_t *inv, const void *f, int f_f64, int K,
                const float *d1, const float *d2, int order, void *out);

int apply_ops(mhip_ctx *ctx, const mhip_nlist *nl, const void *f, int f_f64, int64_t nf_rows,
              int K, int order, int weighted, float *out1, float *out2, int32_t *status);

int weights_rows(mhip_ctx *ctx, const void *pos, int pos_f64, int64_t n, int dim,
                 const int32_t *ngb, int64_t nq, int k, const double *h, const int64_t *indices,
                 double box, int weighted, int order, double *G, int32_t *status);

int hsml_dists(mhip_ctx *ctx, const void *d, int d_f64, int64_t n, int k, int dim,
               double error_norm, double *h, int32_t *status);

int kernel_eval(mhip_ctx *ctx, const void *q, int f64, int64_t n, void *out);

// arguments of the fused derivative operator over one neighbour list
struct ApplyArgs {
    const float *sx, *sy, *sz, *lx, *ly, *lz;
    float box, inv_box;
    int per, p64;
    const uint32_t *nl_idx;
    const uint32_t *qsorted, *qrow, *perm;
    const double *h;     // smoothing length per slot (float64-refined)
    int64_t nq, npad;
    int k;
    const void *fs;  // [K][npad] sorted-space fields
    const float4 *xyzf;  // null, or [npad] (x, y, z, field 0) in sorted space (float32 positions and fields)
    int K;
    int weighted;
    float *out1, *out2;  // [row][K][DIM], [row][K][ND-DIM]
    int32_t *status;     // per row, may be null
    int or_status;       // 1: OR the bits into status instead of storing them
};

// every snapshot of a batch in one launch (table and block prefix on the
// device); kc = fields per thread (2, 4, 8 or 16), K / kc field tiles in
// parallel
int apply_batch_launch(mhip_ctx *ctx, const ApplyArgs *tab_d, const int64_t *blk0_d, int nsnap,
                       int64_t nblocks, int kc, int K);
void apply_args_fill(const mhip_nlist *nl, const void *fs, int K, float *out1, int32_t *status,
                     int or_status, ApplyArgs *a);
// many fields at once (order 1, 3D, float32 positions, row-major fields
// [npad][K], des_ngb <= kmax <= 128): kw threads per query (8, 16 or 32), a
// block holds 256 / kw queries (16 for kw = 8); blk0 counts query blocks of
// that size
int apply_wide_launch(mhip_ctx *ctx, const ApplyArgs *tab_d, const int64_t *blk0_d, int nsnap,
                      int64_t nblocks, int kw, int K, int kmax);
// the same with the field contraction on the matrix cores (K == kw fields)
int apply_mfma_launch(mhip_ctx *ctx, const ApplyArgs *tab_d, const int64_t *blk0_d, int nsnap,
                      int64_t nblocks, int kw);

// field rows (original order, K components each) -> sorted-space rows
// [npad][K]; rows beyond nf_rows and the padding read NaN
template <typename FT>
__global__ __launch_bounds__(256) void k_permute_fields_rows(const FT *__restrict__ f, int64_t nf_rows, int K,
                                                            const uint32_t *__restrict__ perm, int64_t n,
                                                            int64_t npad, FT *__restrict__ fs) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= npad * K) return;
    const int64_t p = e / K;
    const int c = (int)(e - p * K);
    const uint32_t o = p < n ? perm[p] : 0xffffffffu;
    fs[e] = (o < nf_rows) ? f[(int64_t)o * K + c] : (FT)NAN;
}

// sorted-space pack (x, y, z, f) of float32 positions and the first field
// component, so that a neighbour's position and value are one 16-byte load;
// f0 is the sorted-space first field (k_permute_fields output)
__global__ __launch_bounds__(256) void k_pack_xyzf(const float *__restrict__ sx, const float *__restrict__ sy,
                                                  const float *__restrict__ sz, const float *__restrict__ f0,
                                                  int64_t npad, float4 *__restrict__ out);

// kernel-weighted neighbour statistics (mhip_kops.hip)
int kernel_weights(mhip_ctx *ctx, const void *dist, int dist_f64, const void *h, int h_f64, int64_t nq, int k,
                   double *w);
int kernel_stat(mhip_ctx *ctx, const mhip_nlist *nl, const void *f, int f_f64, int64_t nf_rows, int K,
                int which, double *out);
int integrate(mhip_ctx *ctx, const void *vol, int vol_f64, const void *f, int f_f64, int64_t n, int K, double *out);

// field rows (original order, K components each) -> sorted-space columns
// [K][npad]: the j-th neighbours of a wave's 64 consecutive queries lie close
// together in sorted order, so one field's gathers share cache lines; rows
// beyond nf_rows and the padding read NaN
template <typename FT>
__global__ __launch_bounds__(256) void k_permute_fields(const FT *__restrict__ f, int64_t nf_rows, int K,
                                                       const uint32_t *__restrict__ perm, int64_t n,
                                                       int64_t npad, FT *__restrict__ fs) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= npad) return;
    const uint32_t o = p < n ? perm[p] : 0xffffffffu;
    for (int c = 0; c < K; c++) {
        FT v = (o < nf_rows) ? f[(int64_t)o * K + c] : (FT)NAN;
        fs[(int64_t)c * npad + p] = v;
    }
}

}  // namespace mhip
