// Spatial index: the GPU replacement of scipy's cKDTree build
// (src/meshoid/Meshoid.py:142-145).  Particles are sorted by the Hilbert index
// of their cell on a 2^21-per-axis grid (a project-written LSD radix sort) and
// reordered into float32
// structure-of-arrays; an implicit tree groups LEAF consecutive particles into
// leaves and FANOUT consecutive nodes into parents, each with its bounding box.
#include <algorithm>

#include "mhip_internal.hpp"

namespace mhip {

// ---------------------------------------------------------------------------
// bounding box
// ---------------------------------------------------------------------------
template <typename PT>
__global__ __launch_bounds__(256) void k_bbox_partial(const PT *__restrict__ pos, int64_t n, int dim,
                                                     float *__restrict__ part) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float nanflag = 0.f;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        for (int c = 0; c < dim; c++) {
            float v = (float)pos[i * dim + c];
            if (v != v) nanflag = 1.f;
            lo[c] = fminf(lo[c], v);
            hi[c] = fmaxf(hi[c], v);
        }
    }
    __shared__ float s[7][256];
    for (int c = 0; c < 3; c++) {
        s[c][threadIdx.x] = lo[c];
        s[3 + c][threadIdx.x] = hi[c];
    }
    s[6][threadIdx.x] = nanflag;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int c = 0; c < 3; c++) {
                s[c][threadIdx.x] = fminf(s[c][threadIdx.x], s[c][threadIdx.x + w]);
                s[3 + c][threadIdx.x] = fmaxf(s[3 + c][threadIdx.x], s[3 + c][threadIdx.x + w]);
            }
            s[6][threadIdx.x] = fmaxf(s[6][threadIdx.x], s[6][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 7) part[blockIdx.x * 8 + threadIdx.x] = s[threadIdx.x][0];
}

// The same for 3D float32 positions read as 16-byte vectors of the flat
// (n, 3) array: with a thread count divisible by 3, thread t always meets
// components (t, t + 1, t + 2, t) mod 3 in its vectors, so it keeps three
// rotated accumulators and maps them back to x, y, z at the end.  The at most
// 3 floats after the last whole vector go to thread 0.  pos must be 16-byte
// aligned.
__global__ __launch_bounds__(256) void k_bbox_partial_v3(const float *__restrict__ pos, int64_t n,
                                                        float *__restrict__ part) {
    const int64_t gt = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t nt = (int64_t)gridDim.x * blockDim.x;  // a multiple of 3
    const int64_t nv = 3 * n / 4;
    const float4 *v4 = reinterpret_cast<const float4 *>(pos);
    float rl[3] = {INFINITY, INFINITY, INFINITY}, rh[3] = {-INFINITY, -INFINITY, -INFINITY};
    float nanflag = 0.f;
    for (int64_t v = gt; v < nv; v += nt) {
        const float4 x = v4[v];
        if (x.x != x.x || x.y != x.y || x.z != x.z || x.w != x.w) nanflag = 1.f;
        rl[0] = fminf(rl[0], fminf(x.x, x.w));
        rh[0] = fmaxf(rh[0], fmaxf(x.x, x.w));
        rl[1] = fminf(rl[1], x.y);
        rh[1] = fmaxf(rh[1], x.y);
        rl[2] = fminf(rl[2], x.z);
        rh[2] = fmaxf(rh[2], x.z);
    }
    const int c0 = (int)(gt % 3);
    float lo[3], hi[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int r = (c - c0 + 3) % 3;  // accumulator of component c
        lo[c] = r == 0 ? rl[0] : (r == 1 ? rl[1] : rl[2]);
        hi[c] = r == 0 ? rh[0] : (r == 1 ? rh[1] : rh[2]);
    }
    if (gt == 0) {
        for (int64_t e = 4 * nv; e < 3 * n; e++) {
            const float x = pos[e];
            const int c = (int)(e % 3);
            if (x != x) nanflag = 1.f;
            lo[c] = fminf(lo[c], x);
            hi[c] = fmaxf(hi[c], x);
        }
    }
    __shared__ float s[7][256];
    for (int c = 0; c < 3; c++) {
        s[c][threadIdx.x] = lo[c];
        s[3 + c][threadIdx.x] = hi[c];
    }
    s[6][threadIdx.x] = nanflag;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int c = 0; c < 3; c++) {
                s[c][threadIdx.x] = fminf(s[c][threadIdx.x], s[c][threadIdx.x + w]);
                s[3 + c][threadIdx.x] = fmaxf(s[3 + c][threadIdx.x], s[3 + c][threadIdx.x + w]);
            }
            s[6][threadIdx.x] = fmaxf(s[6][threadIdx.x], s[6][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 7) part[blockIdx.x * 8 + threadIdx.x] = s[threadIdx.x][0];
}

// single block: reduce partials into out[0..6] (lo xyz, hi xyz, nan flag) and
// the quantisation frame frame[0..3] (origin xyz, scale)
__global__ __launch_bounds__(256) void k_bbox_final(const float *__restrict__ part, int nb, int dim,
                                                   float box, float *__restrict__ out) {
    __shared__ float s[7][256];
    float v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.f};
    for (int b = threadIdx.x; b < nb; b += 256) {
        for (int c = 0; c < 3; c++) {
            v[c] = fminf(v[c], part[b * 8 + c]);
            v[3 + c] = fmaxf(v[3 + c], part[b * 8 + 3 + c]);
        }
        v[6] = fmaxf(v[6], part[b * 8 + 6]);
    }
    for (int c = 0; c < 7; c++) s[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int c = 0; c < 3; c++) {
                s[c][threadIdx.x] = fminf(s[c][threadIdx.x], s[c][threadIdx.x + w]);
                s[3 + c][threadIdx.x] = fmaxf(s[3 + c][threadIdx.x], s[3 + c][threadIdx.x + w]);
            }
            s[6][threadIdx.x] = fmaxf(s[6][threadIdx.x], s[6][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int c = 0; c < 7; c++) out[c] = s[c][0];
        for (int c = dim; c < 3; c++) {
            out[c] = 0.f;
            out[3 + c] = 0.f;
        }
        // frame: cubic cells over the longest side (or the periodic box)
        float ext = 0.f;
        for (int c = 0; c < dim; c++) ext = fmaxf(ext, out[3 + c] - out[c]);
        float org[3] = {out[0], out[1], out[2]};
        if (box > 0.f) {
            ext = box;
            org[0] = org[1] = org[2] = 0.f;
        }
        for (int c = 0; c < 3; c++) out[8 + c] = org[c];
        out[11] = ext > 0.f ? 2097151.0f / ext : 0.f;
    }
}

// ---------------------------------------------------------------------------
// sort keys: Hilbert index of the particle's cell (21 bits per axis)
// ---------------------------------------------------------------------------
template <typename PT>
__global__ __launch_bounds__(256) void k_sfc_keys(const PT *__restrict__ pos, int64_t n, int dim,
                                                 const float *__restrict__ bb,
                                                 uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float scale = bb[11];
    uint32_t c[3] = {0u, 0u, 0u};
    for (int a = 0; a < dim; a++) c[a] = key_cell((float)pos[i * dim + a], bb[8 + a], scale);
    keys[i] = hilbert_key(c[0], c[1], c[2], dim);
    vals[i] = (uint32_t)i;
}

// ---------------------------------------------------------------------------
// LSD radix sort, 8-bit digits.  Per pass: block histograms (digit-major),
// per-digit scan across blocks, then a stable scatter in which each wavefront
// ranks its 64 keys by digit with ballot-based matching.
// ---------------------------------------------------------------------------
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;

size_t radix_counts_words(int64_t n) {
    int64_t nb = (n + RS_TILE - 1) / RS_TILE;
    return (size_t)(256 * std::max<int64_t>(nb, 1) + 256);
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t *__restrict__ keys, int64_t n,
                                                        int shift, uint32_t *__restrict__ counts,
                                                        int nblocks) {
    __shared__ uint32_t h[4][256];
    const int tid = threadIdx.x, w = tid >> 6;
    for (int i = tid; i < 1024; i += RS_THREADS) (&h[0][0])[i] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int r = 0; r < RS_ITEMS; r++) {
        int64_t idx = base + r * RS_THREADS + tid;
        if (idx < n) atomicAdd(&h[w][(uint32_t)(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[(int64_t)tid * nblocks + blockIdx.x] = h[0][tid] + h[1][tid] + h[2][tid] + h[3][tid];
}

// one block per digit: exclusive scan of counts[d][0..nblocks), total -> totals[d]
__global__ __launch_bounds__(256) void k_rs_scan(uint32_t *__restrict__ counts, int nblocks,
                                                 uint32_t *__restrict__ totals) {
    __shared__ uint32_t s[256];
    uint32_t *row = counts + (int64_t)blockIdx.x * nblocks;
    const int tid = threadIdx.x;
    const int per = (nblocks + 255) / 256;
    const int b0 = tid * per, b1 = min(b0 + per, nblocks);
    uint32_t sum = 0;
    for (int b = b0; b < b1; b++) sum += row[b];
    s[tid] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t v = tid >= off ? s[tid - off] : 0u;
        __syncthreads();
        s[tid] += v;
        __syncthreads();
    }
    uint32_t run = s[tid] - sum;
    for (int b = b0; b < b1; b++) {
        uint32_t c = row[b];
        row[b] = run;
        run += c;
    }
    if (tid == 255) totals[blockIdx.x] = s[255];
}

// Stable scatter of one tile: keys are ranked into tile-local digit order in
// LDS first, then written out in that order, so that the threads of a wave
// write runs of consecutive positions of one digit (coalesced) instead of 64
// scattered 12-byte writes.  Each wavefront ranks a contiguous quarter of the
// tile on its own: per round it matches its 64 keys' digits with 8 ballots
// and advances its own digit counters in LDS, in program order (no block
// barrier inside the rounds; an earlier form ranked block-wide rounds with
// three barriers each).  The quarters are combined afterwards: a key's
// tile-local position is its digit's base in the tile + the counts of that
// digit in the earlier quarters + its rank within its quarter, which keeps
// the order of equal digits (the sort is stable).
constexpr int RS_WAVES = RS_THREADS / WAVE;
constexpr int RS_QUARTER = RS_TILE / RS_WAVES;  // keys per wavefront
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint64_t *__restrict__ kin,
                                                           const uint32_t *__restrict__ vin, int64_t n,
                                                           int shift, const uint32_t *__restrict__ counts,
                                                           const uint32_t *__restrict__ totals,
                                                           int nblocks, uint64_t *__restrict__ kout,
                                                           uint32_t *__restrict__ vout) {
    __shared__ uint64_t s_key[RS_TILE];
    __shared__ uint32_t s_val[RS_TILE];
    __shared__ uint32_t s_wc[RS_WAVES][256];  // per-wave digit counts (running, then offsets)
    __shared__ uint32_t s_gb[256];            // global position of the tile's first key of each digit
    __shared__ uint32_t s_lb[256];            // tile-local position of the first key of each digit
    __shared__ uint32_t s_tmp[256];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    const int tn = (int)min<int64_t>(RS_TILE, n - base);
#pragma unroll
    for (int ww = 0; ww < RS_WAVES; ww++) s_wc[ww][tid] = 0;
    __syncthreads();
    // this wavefront's quarter: input indices w*QUARTER + r*64 + lane
    uint64_t key[RS_ITEMS];
    uint32_t val[RS_ITEMS], rk[RS_ITEMS];
    const uint64_t lt_mask = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const int i = w * RS_QUARTER + r * WAVE + lane;
        const bool valid = i < tn;
        key[r] = valid ? kin[base + i] : 0ull;
        val[r] = valid ? vin[base + i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const int i = w * RS_QUARTER + r * WAVE + lane;
        const bool valid = i < tn;
        const uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bb = 0; bb < 8; bb++) {
            const bool bit = (d >> bb) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = __popcll(peers & lt_mask);
        // every lane reads its digit's count before the lowest lane of each
        // digit group advances it (one wavefront: reads, then writes, in order)
        const uint32_t before = valid ? s_wc[w][d] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) s_wc[w][d] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rk[r] = before + rank;
    }
    __syncthreads();
    // per digit (thread tid = digit): offsets of the quarters, tile count
    uint32_t cnt = 0;
#pragma unroll
    for (int ww = 0; ww < RS_WAVES; ww++) {
        const uint32_t c = s_wc[ww][tid];
        s_wc[ww][tid] = cnt;
        cnt += c;
    }
    // exclusive scans over the digits: tile-local bases and global bases
    s_tmp[tid] = cnt;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t v = tid >= off ? s_tmp[tid - off] : 0u;
        __syncthreads();
        s_tmp[tid] += v;
        __syncthreads();
    }
    s_lb[tid] = s_tmp[tid] - cnt;
    __syncthreads();
    s_tmp[tid] = totals[tid];
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t v = tid >= off ? s_tmp[tid - off] : 0u;
        __syncthreads();
        s_tmp[tid] += v;
        __syncthreads();
    }
    s_gb[tid] = s_tmp[tid] - totals[tid] + counts[(int64_t)tid * nblocks + blockIdx.x];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const int i = w * RS_QUARTER + r * WAVE + lane;
        if (i < tn) {
            const uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
            const uint32_t pos = s_lb[d] + s_wc[w][d] + rk[r];
            s_key[pos] = key[r];
            s_val[pos] = val[r];
        }
    }
    __syncthreads();
    // write out in tile-local digit order: runs of one digit are contiguous
    for (int i = tid; i < tn; i += RS_THREADS) {
        const uint64_t k = s_key[i];
        const uint32_t d = (uint32_t)(k >> shift) & 255u;
        const uint32_t pos = s_gb[d] + (uint32_t)i - s_lb[d];
        kout[pos] = k;
        vout[pos] = s_val[i];
    }
}

int radix_sort_pairs(mhip_ctx *ctx, uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1,
                     uint32_t *counts, int64_t n, int bits, uint64_t **kout, uint32_t **vout, int lo_bit) {
    uint64_t *ka = k0, *kb = k1;
    uint32_t *va = v0, *vb = v1;
    if (n > 1) {
        const int nblocks = (int)((n + RS_TILE - 1) / RS_TILE);
        uint32_t *totals = counts + (size_t)256 * nblocks;
        for (int shift = lo_bit; shift < bits; shift += 8) {
            hipLaunchKernelGGL(k_rs_hist, dim3(nblocks), dim3(RS_THREADS), 0, ctx->stream, ka, n,
                               shift, counts, nblocks);
            hipLaunchKernelGGL(k_rs_scan, dim3(256), dim3(256), 0, ctx->stream, counts, nblocks,
                               totals);
            hipLaunchKernelGGL(k_rs_scatter, dim3(nblocks), dim3(RS_THREADS), 0, ctx->stream, ka,
                               va, n, shift, counts, totals, nblocks, kb, vb);
            std::swap(ka, kb);
            std::swap(va, vb);
        }
    }
    *kout = ka;
    *vout = va;
    return check_hip(ctx, hipGetLastError(), "radix sort");
}

// ---------------------------------------------------------------------------
// reorder into sorted space
// ---------------------------------------------------------------------------
template <typename PT>
__global__ __launch_bounds__(256) void k_gather_sorted(const PT *__restrict__ pos, int64_t n, int64_t npad,
                                                      int dim, const uint32_t *__restrict__ order,
                                                      float *__restrict__ sx, float *__restrict__ sy,
                                                      float *__restrict__ sz, float *__restrict__ lx,
                                                      float *__restrict__ ly, float *__restrict__ lz,
                                                      uint32_t *__restrict__ perm) {
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= npad) return;
    const float nan = __builtin_nanf("");
    float h[3] = {nan, nan, nan}, l[3] = {0.f, 0.f, 0.f};
    if (p < n) {
        uint32_t o = order[p];
        perm[p] = o;
        for (int c = 0; c < 3; c++) {
            if (c < dim) {
                PT v = pos[(int64_t)o * dim + c];
                h[c] = (float)v;
                l[c] = (float)(v - (PT)h[c]);
            } else {
                h[c] = 0.f;
            }
        }
    } else {
        perm[p] = 0xffffffffu;
    }
    sx[p] = h[0];
    sy[p] = h[1];
    sz[p] = h[2];
    if (lx) {
        lx[p] = l[0];
        ly[p] = l[1];
        lz[p] = l[2];
    }
}

// LEAF lanes per leaf: lane i of a group reads particle i of the leaf (a
// wavefront reads 64 consecutive particles of each coordinate array in one
// instruction) and the group's minimum and maximum are formed with shuffles.
// One thread per leaf looping over its particles read every cache line 16
// times, 64 lines per instruction, and fetched the positions from HBM ~6x.
static_assert(WAVE % LEAF == 0, "a leaf's lanes lie in one wavefront");
__global__ __launch_bounds__(256) void k_leaf_boxes(const float *__restrict__ sx, const float *__restrict__ sy,
                                                   const float *__restrict__ sz, int64_t n, int64_t nleaf,
                                                   int64_t nslots, float *__restrict__ leaves) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t L = t / LEAF;
    const int i = (int)(t % LEAF);
    const int64_t p = L * LEAF + i;
    const bool in = L < nleaf && p < n;
    float lo[3], hi[3];
    lo[0] = hi[0] = in ? sx[p] : NAN;
    lo[1] = hi[1] = in ? sy[p] : NAN;
    lo[2] = hi[2] = in ? sz[p] : NAN;
    // fminf / fmaxf ignore a NaN operand: lanes past the data drop out, and a
    // leaf with no particle stays NaN, replaced below by the empty box
#pragma unroll
    for (int o = LEAF / 2; o > 0; o >>= 1) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            lo[c] = fminf(lo[c], __shfl_xor(lo[c], o));
            hi[c] = fmaxf(hi[c], __shfl_xor(hi[c], o));
        }
    }
    if (L >= nslots || i != 0) return;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (lo[c] != lo[c]) lo[c] = INFINITY;  // empty leaf: lo = +inf, hi = -inf
        if (hi[c] != hi[c]) hi[c] = -INFINITY;
    }
    float4 *o = (float4 *)(leaves + L * 8);
    o[0] = make_float4(lo[0], lo[1], lo[2], hi[0]);
    o[1] = make_float4(hi[1], hi[2], 0.f, 0.f);
}

// parents x in [x0, x0 + count) of one depth: union of children 8x+1..8x+8
__global__ __launch_bounds__(256) void k_node_boxes(float *__restrict__ nodes, int64_t x0, int64_t count) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int64_t x = x0 + i;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int c = 1; c <= FANOUT; c++) {
        const float4 *s = (const float4 *)(nodes + (FANOUT * x + c) * 8);
        const float4 a = s[0], b = s[1];
        lo[0] = fminf(lo[0], a.x);
        lo[1] = fminf(lo[1], a.y);
        lo[2] = fminf(lo[2], a.z);
        hi[0] = fmaxf(hi[0], a.w);
        hi[1] = fmaxf(hi[1], b.x);
        hi[2] = fmaxf(hi[2], b.y);
    }
    float4 *o = (float4 *)(nodes + x * 8);
    o[0] = make_float4(lo[0], lo[1], lo[2], hi[0]);
    o[1] = make_float4(hi[1], hi[2], 0.f, 0.f);
}

static int grow(mhip_ctx *ctx, void **p, size_t bytes) {
    if (*p) {
        hipFree(*p);
        *p = nullptr;
    }
    if (bytes == 0) return 0;
    return check_hip(ctx, hipMalloc(p, bytes), "hipMalloc (index)");
}

int index_build(mhip_ctx *ctx, const void *pos, int pos_f64, int64_t n, int dim, double box,
                mhip_index *ix, float *rec) {
    if (n < 1) return fail(ctx, -1, "index needs at least one particle");
    if (n >= (int64_t)0xffffffff) return fail(ctx, -4, "more than 2^32-2 particles");
    if (dim < 1 || dim > 3) return fail(ctx, -5, "dim must be 1, 2 or 3");
    const int64_t npad = (n + LEAF - 1) / LEAF * LEAF;
    const int64_t nleaf = npad / LEAF;
    int rc = 0;
    // complete 8-ary tree over the leaves
    int depth = 0;
    int64_t nslots = 1;
    while (nslots < nleaf) {
        nslots *= FANOUT;
        depth++;
    }
    if (depth > MAX_DEPTH) return fail(ctx, -4, "tree too deep");
    const int64_t first_leaf = (nslots - 1) / (FANOUT - 1);
    const int64_t tot = first_leaf + nslots;
    if (ix->cap < npad || (pos_f64 && !ix->lx)) {
        void **arrs[6] = {(void **)&ix->sx, (void **)&ix->sy, (void **)&ix->sz,
                          (void **)&ix->lx, (void **)&ix->ly, (void **)&ix->lz};
        for (int a = 0; a < 6; a++)
            if ((rc = grow(ctx, arrs[a], (a < 3 || pos_f64) ? npad * sizeof(float) : 0))) return rc;
        if ((rc = grow(ctx, (void **)&ix->perm, npad * sizeof(uint32_t)))) return rc;
        if ((rc = grow(ctx, (void **)&ix->iperm, npad * sizeof(uint32_t)))) return rc;
        ix->cap = npad;
    }
    if (ix->node_cap < tot) {
        if ((rc = grow(ctx, (void **)&ix->nodes, tot * 8 * sizeof(float)))) return rc;
        ix->node_cap = tot;
    }
    ix->iperm_ok = 0;
    ix->n = n;
    ix->npad = npad;
    ix->dim = dim;
    ix->box = box;
    ix->pos_f64 = pos_f64;
    ix->pos_in = pos;
    ix->depth = depth;
    ix->nleaf = nleaf;
    ix->first_leaf = first_leaf;
    ix->nnodes = tot;
    // scratch layout: keys x2, vals x2, counts, bbox partials
    const size_t cw = radix_counts_words(n);
    const int nbb = 1024;
    size_t off_k0 = 0, off_k1 = off_k0 + n * 8, off_v0 = off_k1 + n * 8, off_v1 = off_v0 + n * 4,
           off_c = off_v1 + n * 4, off_bb = off_c + cw * 4, total = off_bb + (nbb + 2) * 8 * 4;
    char *s = (char *)scratch(ctx, total, &rc);
    if (!s) return rc;
    uint64_t *k0 = (uint64_t *)(s + off_k0), *k1 = (uint64_t *)(s + off_k1);
    uint32_t *v0 = (uint32_t *)(s + off_v0), *v1 = (uint32_t *)(s + off_v1);
    uint32_t *counts = (uint32_t *)(s + off_c);
    float *part = (float *)(s + off_bb), *bb = rec ? rec : part + nbb * 8;
    // 16-byte loads for aligned 3D float32 sets; 1020 blocks (a multiple of 3)
    const bool v3 = !pos_f64 && dim == 3 && n >= ((int64_t)1 << 18) && ((uintptr_t)pos & 15) == 0;
    const int nb = v3 ? 1020 : (int)std::min<int64_t>(nbb, (n + 255) / 256);
    const dim3 g1((unsigned)((n + 255) / 256)), b256(256);
    ProfRegion pr_bb(ctx, "index_bbox");
    if (pos_f64) {
        hipLaunchKernelGGL(k_bbox_partial<double>, dim3(nb), b256, 0, ctx->stream, (const double *)pos,
                           n, dim, part);
    } else if (v3) {
        hipLaunchKernelGGL(k_bbox_partial_v3, dim3(nb), b256, 0, ctx->stream, (const float *)pos, n, part);
    } else {
        hipLaunchKernelGGL(k_bbox_partial<float>, dim3(nb), b256, 0, ctx->stream, (const float *)pos,
                           n, dim, part);
    }
    hipLaunchKernelGGL(k_bbox_final, dim3(1), b256, 0, ctx->stream, part, nb, dim, (float)box, bb);
    pr_bb.end();
    ProfRegion pr_key(ctx, "index_key_sort");
    if (pos_f64)
        hipLaunchKernelGGL(k_sfc_keys<double>, g1, b256, 0, ctx->stream, (const double *)pos, n, dim, bb,
                           k0, v0);
    else
        hipLaunchKernelGGL(k_sfc_keys<float>, g1, b256, 0, ctx->stream, (const float *)pos, n, dim, bb,
                           k0, v0);
    uint64_t *ks;
    uint32_t *vs;
    if ((rc = radix_sort_pairs(ctx, k0, v0, k1, v1, counts, n, KEY_BITS_PER_AXIS * dim, &ks, &vs,
                               key_sort_lo_bit(dim))))
        return rc;
    pr_key.end();
    ProfRegion pr_g(ctx, "index_reorder_boxes");
    const dim3 gp((unsigned)((npad + 255) / 256));
    if (pos_f64)
        hipLaunchKernelGGL(k_gather_sorted<double>, gp, b256, 0, ctx->stream, (const double *)pos, n,
                           npad, dim, vs, ix->sx, ix->sy, ix->sz, ix->lx, ix->ly, ix->lz, ix->perm);
    else
        hipLaunchKernelGGL(k_gather_sorted<float>, gp, b256, 0, ctx->stream, (const float *)pos, n,
                           npad, dim, vs, ix->sx, ix->sy, ix->sz, (float *)nullptr, (float *)nullptr,
                           (float *)nullptr, ix->perm);
    hipLaunchKernelGGL(k_leaf_boxes, dim3((unsigned)((nslots * LEAF + 255) / 256)), b256, 0, ctx->stream,
                       ix->sx, ix->sy, ix->sz, n, nleaf, nslots, ix->nodes + first_leaf * 8);
    for (int d = depth - 1; d >= 0; d--) {
        int64_t x0 = 0, cnt = 1;
        for (int e = 0; e < d; e++) {
            x0 += cnt;
            cnt *= FANOUT;
        }
        hipLaunchKernelGGL(k_node_boxes, dim3((unsigned)((cnt + 255) / 256)), b256, 0, ctx->stream,
                           ix->nodes, x0, cnt);
    }
    pr_g.end();
    if ((rc = check_hip(ctx, hipGetLastError(), "index build launch"))) return rc;
    if (rec) {
        // deferred: the caller reads the record (bbox, NaN flag, frame) later;
        // the host copies are unknown until then
        for (int c = 0; c < 6; c++) ix->bbox[c] = NAN;
        for (int c = 0; c < 4; c++) ix->frame[c] = NAN;
        return 0;
    }
    // bbox + nan flag back to the host: periodic range and NaN checks
    float hb[12];
    if ((rc = check_hip(ctx, hipMemcpyAsync(hb, bb, 12 * sizeof(float), hipMemcpyDeviceToHost,
                                            ctx->stream), "bbox readback")))
        return rc;
    if ((rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "index build"))) return rc;
    for (int c = 0; c < 6; c++) ix->bbox[c] = hb[c];
    for (int c = 0; c < 4; c++) ix->frame[c] = hb[8 + c];
    if (hb[6] != 0.f) return fail(ctx, -1, "positions contain NaN");
    if (box > 0) {
        for (int c = 0; c < dim; c++)
            // float64 data just below box can round up to box in float32
            if (hb[c] < 0.f || hb[3 + c] > (float)box || (hb[3 + c] == (float)box && !pos_f64))
                return fail(ctx, -1,
                            "Some input data are greater than the size of the periodic box.");
    }
    return 0;
}

}  // namespace mhip
