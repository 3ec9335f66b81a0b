// Reconstruction of particle fields at arbitrary points on the GPU.
//
// Replaces, in Meshoid.Reconstruct (src/meshoid/Meshoid.py:383-440),
//   np.unique(target_neighbors, return_inverse=True)      Meshoid.py:410
//   f_target = np.take(f, target_neighbors, axis=0)       Meshoid.py:413
//   dx = target_points - pos[target_neighbors] (+ image)  Meshoid.py:418-420
//   f_target += dx . D(f)[neighbor_idx]                   Meshoid.py:425-426
//   f_target += second-order terms of D2(f)               Meshoid.py:431-439
// The nearest-particle query itself is nearest_run (mhip_knn.hip).
//
// unique: a flag per particle, a two-level exclusive scan of the flags (ranks
// among the flagged particles, ascending particle index: np.unique's sorted
// order) and a compaction; the inverse of each point is its particle's rank.
#include "mhip_internal.hpp"

namespace mhip {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;  // flags per block

__global__ __launch_bounds__(256) void k_zero_u32(uint32_t *__restrict__ a, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = 0u;
}

// block-local exclusive ranks of the flags, and the block totals
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(const uint32_t *__restrict__ flags, int64_t n,
                                                             uint32_t *__restrict__ rank,
                                                             uint32_t *__restrict__ tile_sum) {
    __shared__ uint32_t sh[SCAN_THREADS];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS], run = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        v[j] = base + j < n ? flags[base + j] : 0u;
        const uint32_t t = v[j];
        v[j] = run;
        run += t;
    }
    sh[threadIdx.x] = run;
    __syncthreads();
    // Hillis-Steele inclusive scan of the thread totals
    for (int o = 1; o < SCAN_THREADS; o <<= 1) {
        const uint32_t add = (int)threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    const uint32_t off = sh[threadIdx.x] - run;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++)
        if (base + j < n) rank[base + j] = v[j] + off;
    if (threadIdx.x == SCAN_THREADS - 1) tile_sum[blockIdx.x] = sh[threadIdx.x];
}

// exclusive scan of the tile totals in place, one block: each thread scans
// SUMS_PER_THREAD consecutive totals, the block scans the thread totals
// (one pass up to 8192 tiles, i.e. 16.7 M flags; larger inputs loop with a
// carry); tile_sum[ntile] receives the grand total
constexpr int SUMS_PER_THREAD = 32;
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tile_sums(uint32_t *__restrict__ tile_sum, int64_t ntile) {
    __shared__ uint32_t sh[SCAN_THREADS];
    uint32_t carry = 0;
    for (int64_t c0 = 0; c0 < ntile; c0 += (int64_t)SCAN_THREADS * SUMS_PER_THREAD) {
        const int64_t b = c0 + (int64_t)threadIdx.x * SUMS_PER_THREAD;
        uint32_t v[SUMS_PER_THREAD], run = 0;
#pragma unroll
        for (int j = 0; j < SUMS_PER_THREAD; j++) {
            const uint32_t t = b + j < ntile ? tile_sum[b + j] : 0u;
            v[j] = run;
            run += t;
        }
        sh[threadIdx.x] = run;
        __syncthreads();
        for (int o = 1; o < SCAN_THREADS; o <<= 1) {
            const uint32_t add = (int)threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
            __syncthreads();
            sh[threadIdx.x] += add;
            __syncthreads();
        }
        const uint32_t off = carry + sh[threadIdx.x] - run;
#pragma unroll
        for (int j = 0; j < SUMS_PER_THREAD; j++)
            if (b + j < ntile) tile_sum[b + j] = v[j] + off;
        carry += sh[SCAN_THREADS - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) tile_sum[ntile] = carry;
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t *__restrict__ out, const uint32_t *__restrict__ tile_sum,
                                                  int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (i < n) out[i] += tile_sum[i / SCAN_TILE];
}

size_t scan_work_words(int64_t n) { return (size_t)((n + SCAN_TILE - 1) / SCAN_TILE) + 1; }

uint32_t *exclusive_scan_u32(mhip_ctx *ctx, const uint32_t *in, int64_t n, uint32_t *out, uint32_t *work) {
    const int64_t ntile = (n + SCAN_TILE - 1) / SCAN_TILE;
    hipStream_t s = ctx->stream;
    if (n > 0) {
        hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntile), dim3(SCAN_THREADS), 0, s, in, n, out, work);
        hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(SCAN_THREADS), 0, s, work, ntile);
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, work, n);
    } else {
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, s, work, (int64_t)1);
    }
    return work + ntile;
}

// unique / inverse through a bit per particle: particle j is marked in bit
// j % 32 of word j / 32; its rank among the marked ones is the exclusive
// prefix of the words' popcounts plus the popcount of the lower bits of its
// own word (ascending particle order, as np.unique)
__global__ __launch_bounds__(256) void k_mark_bits(const int64_t *__restrict__ nbr, int64_t nt, int64_t n,
                                                   uint32_t *__restrict__ bits, uint32_t *__restrict__ bad) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const int64_t j = nbr[i];
    if (j >= 0 && j < n)
        atomicOr(bits + (j >> 5), 1u << (j & 31));
    else
        atomicAdd(bad, 1u);
}

__global__ __launch_bounds__(256) void k_popc_words(const uint32_t *__restrict__ bits, int64_t nw,
                                                    uint32_t *__restrict__ cnt) {
    const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (w < nw) cnt[w] = (uint32_t)__popc(bits[w]);
}

// the marked particles of word w, in ascending order, from position wpre[w]
__global__ __launch_bounds__(256) void k_compact_bits(const uint32_t *__restrict__ bits,
                                                      const uint32_t *__restrict__ wpre, int64_t nw,
                                                      int64_t *__restrict__ uniq) {
    const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (w >= nw) return;
    uint32_t m = bits[w];
    int64_t pos = wpre[w];
    while (m) {
        uniq[pos++] = w * 32 + __builtin_ctz(m);
        m &= m - 1u;
    }
}

__global__ __launch_bounds__(256) void k_inverse_bits(const int64_t *__restrict__ nbr, int64_t nt, int64_t n,
                                                      const uint32_t *__restrict__ bits,
                                                      const uint32_t *__restrict__ wpre, int64_t *__restrict__ inv) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const int64_t j = nbr[i];
    if (j < 0 || j >= n) {
        inv[i] = -1;
        return;
    }
    const uint32_t below = bits[j >> 5] & ((1u << (j & 31)) - 1u);
    inv[i] = (int64_t)wpre[j >> 5] + __popc(below);
}

int unique_inverse(mhip_ctx *ctx, const int64_t *nbr, int64_t nt, int64_t n, int64_t *uniq, int64_t *inv,
                   int64_t *nuniq) {
    if (nt < 0 || n < 1) return fail(ctx, -1, "bad unique shape");
    if (n > 0xffffffffll) return fail(ctx, -4, "more than 2^32 particles");
    *nuniq = 0;
    if (nt == 0) return 0;
    const int64_t nw = (n + 31) / 32;
    int rc = 0;
    // bits, popcounts, their exclusive prefix, scan work, the bad-point count
    const size_t sw = scan_work_words(nw);
    uint32_t *bits = (uint32_t *)scratch(ctx, (size_t)(3 * nw + sw + 1) * 4, &rc);
    if (!bits) return rc;
    uint32_t *cnt = bits + nw, *wpre = cnt + nw, *work = wpre + nw, *bad = work + sw;
    hipStream_t s = ctx->stream;
    ProfRegion pr(ctx, "unique");
    const unsigned gz = (unsigned)std::min<int64_t>(4096, (nw + 255) / 256);
    hipLaunchKernelGGL(k_zero_u32, dim3(gz), dim3(256), 0, s, bits, nw);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, s, bad, (int64_t)1);
    const dim3 gw((unsigned)((nw + 255) / 256)), gt((unsigned)((nt + 255) / 256));
    hipLaunchKernelGGL(k_mark_bits, gt, dim3(256), 0, s, nbr, nt, n, bits, bad);
    hipLaunchKernelGGL(k_popc_words, gw, dim3(256), 0, s, bits, nw, cnt);
    const uint32_t *total = exclusive_scan_u32(ctx, cnt, nw, wpre, work);
    hipLaunchKernelGGL(k_compact_bits, gw, dim3(256), 0, s, bits, wpre, nw, uniq);
    hipLaunchKernelGGL(k_inverse_bits, gt, dim3(256), 0, s, nbr, nt, n, bits, wpre, inv);
    if ((rc = check_hip(ctx, hipGetLastError(), "unique launch"))) return rc;
    uint32_t host[2] = {0u, 0u};
    if ((rc = check_hip(ctx, hipMemcpyAsync(&host[0], total, 4, hipMemcpyDeviceToHost, s), "unique count")))
        return rc;
    if ((rc = check_hip(ctx, hipMemcpyAsync(&host[1], bad, 4, hipMemcpyDeviceToHost, s), "unique count")))
        return rc;
    if ((rc = check_hip(ctx, hipStreamSynchronize(s), "unique sync"))) return rc;
    if (host[1]) return fail(ctx, -1, "a point has no nearest particle (non-finite coordinates)");
    *nuniq = host[0];
    return 0;
}

// ---------------------------------------------------------------------------
// f_target per point and component.  The reference adds each term to an
// array of f's dtype in place, so for float32 f every partial sum is rounded
// to float32 (first order, then the pure second-order terms, then each mixed
// term); the terms themselves are float64.
// ---------------------------------------------------------------------------
struct ReconArgs {
    const double *tg;
    int64_t nt;
    int dim;
    double box;  // > 0: periodic
    const void *pos;
    const int64_t *nbr, *inv;
    const void *f;
    int K, order;
    const float *d1;  // (nu, K, dim)
    const float *d2;  // (nu, K, 6), 3D only
    void *out;        // (nt, K), f's dtype
};

template <typename FT, typename PT, bool PER>
__global__ __launch_bounds__(256) void k_reconstruct(ReconArgs a) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= a.nt) return;
    const int64_t j = a.nbr[i];
    const FT *f = (const FT *)a.f;
    FT *out = (FT *)a.out;
    if (j < 0) {  // a point without a nearest particle (non-finite coordinates)
        for (int c = 0; c < a.K; c++) out[i * a.K + c] = (FT)NAN;
        return;
    }
    if (a.order == 0) {
        for (int c = 0; c < a.K; c++) out[i * a.K + c] = f[j * a.K + c];
        return;
    }
    const PT *pos = (const PT *)a.pos;
    double dx[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < a.dim; d++) {
        double v = a.tg[i * a.dim + d] - (double)pos[j * a.dim + d];
        if (PER) v = wrap<double>(v, a.box);
        dx[d] = v;
    }
    const int64_t u = a.inv[i];
    for (int c = 0; c < a.K; c++) {
        FT v = f[j * a.K + c];
        const float *g = a.d1 + (u * a.K + c) * a.dim;
        double s = dx[0] * (double)g[0];
        for (int d = 1; d < a.dim; d++) s += dx[d] * (double)g[d];
        v = (FT)((double)v + s);
        if (a.order == 2) {
            const float *h = a.d2 + (u * a.K + c) * 6;
            const double s2 = (dx[0] * dx[0] * (double)h[0] + dx[1] * dx[1] * (double)h[1]) +
                              dx[2] * dx[2] * (double)h[2];
            v = (FT)((double)v + 0.5 * s2);
            for (int d = 0; d < 3; d++) v = (FT)((double)v + dx[d] * dx[(d + 1) % 3] * (double)h[3 + d]);
        }
        out[i * a.K + c] = v;
    }
}

int reconstruct(mhip_ctx *ctx, const double *targets, int64_t nt, int dim, double box, const void *pos,
                int pos_f64, const int64_t *nbr, const int64_t *inv, const void *f, int f_f64, int K,
                const float *d1, const float *d2, int order, void *out) {
    if (nt < 0 || K < 1 || dim < 1 || dim > 3) return fail(ctx, -1, "bad reconstruction shape");
    if (order < 0 || order > 2) return fail(ctx, -1, "reconstruction order must be 0, 1 or 2");
    if (order >= 1 && !d1) return fail(ctx, -1, "first derivatives are required for order >= 1");
    if (order == 2 && (dim != 3 || !d2)) return fail(ctx, -1, "order-2 reconstruction needs 3D second derivatives");
    if (nt == 0) return 0;
    ReconArgs a{targets, nt, dim, box > 0 ? box : 0.0, pos, nbr, inv, f, K, order, d1, d2, out};
    const dim3 g((unsigned)((nt + 255) / 256));
    hipStream_t s = ctx->stream;
    ProfRegion pr(ctx, "reconstruct");
    const bool per = box > 0;
#define MHIP_RECON(FT, PT)                                                                  \
    do {                                                                                    \
        if (per)                                                                            \
            hipLaunchKernelGGL((k_reconstruct<FT, PT, true>), g, dim3(256), 0, s, a);       \
        else                                                                                \
            hipLaunchKernelGGL((k_reconstruct<FT, PT, false>), g, dim3(256), 0, s, a);      \
    } while (0)
    if (f_f64) {
        if (pos_f64)
            MHIP_RECON(double, double);
        else
            MHIP_RECON(double, float);
    } else {
        if (pos_f64)
            MHIP_RECON(float, double);
        else
            MHIP_RECON(float, float);
    }
#undef MHIP_RECON
    return check_hip(ctx, hipGetLastError(), "reconstruct launch");
}

}  // namespace mhip
