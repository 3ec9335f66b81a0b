// Surface density of a particle quantity on a grid of sightlines.
//
// Replaces
//   Meshoid.SurfaceDensity                   src/meshoid/Meshoid.py:595-643
//   GridSurfaceDensity (parallel split/sum)  src/meshoid/grid_deposition.py:123-184
//   GridSurfaceDensity_core                  src/meshoid/grid_deposition.py:187-229
//   GridSurfaceDensity_conservative_core     src/meshoid/grid_deposition.py:232-295
//   kernel2d                                 src/meshoid/kernel_density.py:60-68
//   Meshoid.DepositToGrid                    src/meshoid/Meshoid.py:557-593
//   GridDensity                              src/meshoid/grid_deposition.py:534-603
//   kernel3d                                 src/meshoid/kernel_density.py:71-79
//
// Each particle adds kernel2d(r/h) f / h^2 to every grid point (spacing
// size/(res-1), point 0 at center - size/2) within h of its projected
// position; the conservative variant divides each particle's weights by
// their sum and scales the grid by 1/dx^2 at the end.  Accumulation is
// float64.  Up to res = 128 every block keeps a private copy of the whole
// map in LDS (float64 LDS atomics, 128 KB) for an interleaved share of the
// particles and adds it to the HBM map once at the end: the many particles
// of a clump that land on the same few cells contend in LDS, not in L2.
// Larger maps (and every GridDensity volume) are binned: particles are
// sorted by the brick (32^2 map cells, 8^3 volume cells) that holds the
// first cell of their footprint, and each block takes a run of the sorted
// particles; a brick's run of at least BIN_DENSE particles accumulates in an
// LDS window (the brick plus the footprint reach above it, 4096 cells) that
// is added to the HBM map once, shorter runs take float64 atomics in HBM.
// Particles whose footprint exceeds the window (or BIG_CELLS) go to a kernel
// that spreads one particle over a whole block.
#include "mhip_internal.hpp"

namespace mhip {

constexpr int DEP_THREADS = 256;
constexpr int64_t BIG_CELLS = 1024;           // footprint handled by one thread at most
constexpr int DEP_TILE = 128;                 // map side held whole in LDS (128 KB of float64)
constexpr double KERNEL2D_NORM = 1.8189136353359467;  // 40 / (7 pi)

// binned deposition
constexpr int BIN_CHUNK = 2048;   // sorted particles per block
constexpr int BIN_DENSE = 32;     // a brick's run of this many particles in a block uses the window
constexpr int WIN_CELLS = 4096;   // LDS window, float64 cells (32 KB)
constexpr int BRICK2 = 32, WIN2 = 64;  // maps: brick side, window side (footprint span <= 33)
constexpr int BRICK3 = 8, WIN3 = 16;   // volumes: brick side, window side (span <= 9)
static_assert(WIN2 * WIN2 == WIN_CELLS && WIN3 * WIN3 * WIN3 == WIN_CELLS, "window size");
static_assert(BIN_CHUNK <= 65535, "run offsets and lengths are packed in 16 bits");
// sort keys: brick << SIZE_BITS | (SIZE_MAX_RANK - footprint cells), so a
// brick's particles come largest footprint first and a wavefront's 64
// footprints are alike (its lanes finish together)
constexpr int SIZE_BITS = 11;
// conservative footprints up to this many cells keep their weights in
// registers between the total and the deposit (the window kernel is limited
// to 3 waves per SIMD by its LDS, so the registers cost no occupancy)
constexpr int WCACHE2 = 25, WCACHE3 = 27;
constexpr uint32_t SIZE_MAX_RANK = (1u << SIZE_BITS) - 1;

struct DepArgs {
    const void *f, *pos, *h;
    int f_f64, pos_f64, h_f64, dim;
    int64_t n;
    double smooth, hmin;  // h used = min(max(smooth h, hmin), 1e100)
    double ox, oy;        // center - size/2: the position of grid point 0
    double dx;            // size / (res - 1)
    int res, conservative;
    double *grid;           // (res, res), [gx][gy]
    uint32_t *big, *nbig;   // particles left to the block-per-particle kernel
    int nb;                 // bricks per axis (binned path)
};

MHIP_DEV double kernel2d(double q) {
    double k;
    if (q <= 0.5) {
        k = 1.0 - 6.0 * q * q * (1.0 - q);
    } else {
        const double a = 1.0 - q;
        k = 2.0 * a * a * a;
    }
    return k * KERNEL2D_NORM;
}

template <typename T>
MHIP_DEV double ld(const void *p, int64_t i) {
    return (double)((const T *)p)[i];
}

struct Foot {
    double xs0, xs1, hs, hinv, w;  // projected position, h, 1/h, per-weight factor
    int gx0, gx1, gy0, gy1;
};

// grid index range of int((x -/+ h) / dx (+ 1)) as the reference forms it
// (truncation toward zero, then clipped to the grid)
MHIP_DEV int lo_index(double a, int res) {
    if (!(a > 0.0)) return 0;
    return a >= (double)res ? res : (int)a;
}
MHIP_DEV int hi_index(double b, int res) {
    if (!(b > -1.0)) return -1;
    return b >= (double)(res - 1) ? res - 1 : (int)b;
}

// false: nothing to deposit (non-finite input or footprint off the grid)
MHIP_DEV bool footprint(const DepArgs &a, int64_t i, Foot &fp) {
    const double x0 = a.pos_f64 ? ld<double>(a.pos, i * a.dim) : ld<float>(a.pos, i * a.dim);
    const double x1 = a.pos_f64 ? ld<double>(a.pos, i * a.dim + 1) : ld<float>(a.pos, i * a.dim + 1);
    const double hv = a.h_f64 ? ld<double>(a.h, i) : ld<float>(a.h, i);
    const double fv = a.f_f64 ? ld<double>(a.f, i) : ld<float>(a.f, i);
    fp.xs0 = (x0 - a.ox);
    fp.xs1 = (x1 - a.oy);
    double hs = fmin(fmax(a.smooth * hv, a.hmin), 1e100);
    if (a.conservative) hs = fmax(hs, a.dx);
    if (!(hs > 0.0) || !isfinite(fp.xs0) || !isfinite(fp.xs1) || !isfinite(fv)) return false;
    fp.hs = hs;
    fp.hinv = 1.0 / hs;
    fp.w = a.conservative ? fv : fv * fp.hinv * fp.hinv;
    fp.gx0 = lo_index((fp.xs0 - hs) / a.dx + 1.0, a.res);
    fp.gx1 = hi_index((fp.xs0 + hs) / a.dx, a.res);
    fp.gy0 = lo_index((fp.xs1 - hs) / a.dx + 1.0, a.res);
    fp.gy1 = hi_index((fp.xs1 + hs) / a.dx, a.res);
    return fp.gx0 <= fp.gx1 && fp.gy0 <= fp.gy1;
}

// kernel weight at grid point (gx, gy), 0 when it lies beyond h
MHIP_DEV double cell_weight(const DepArgs &a, const Foot &fp, int gx, int gy) {
    double ddx = fp.xs0 - gx * a.dx;
    ddx *= ddx;
    double ddy = fp.xs1 - gy * a.dx;
    ddy *= ddy;
    const double r = sqrt(ddx + ddy);
    if (r > fp.hs) return 0.0;
    return kernel2d(r * fp.hinv);
}

// block b: map tile (b % ntile) (tw x tw cells held in LDS) and particle
// chunk (b / ntile), the chunk being every nchunk-th run of 256 particles
// (interleaved, so that a clump's particles, stored together, spread over
// all chunks instead of piling onto one block's LDS atomics); it deposits the
// cells of its tile only, and adds the tile to the HBM map once at the end
__global__ __launch_bounds__(DEP_THREADS) void k_deposit_tiles(DepArgs a, int tw, int ntx, int nchunk) {
    extern __shared__ double sg[];
    const int ntile = ntx * ntx;
    const int tile = blockIdx.x % ntile, chunk = blockIdx.x / ntile;
    const int tx0 = (tile / ntx) * tw, ty0 = (tile % ntx) * tw;
    const int tx1 = min(tx0 + tw, a.res) - 1, ty1 = min(ty0 + tw, a.res) - 1;
    for (int c = threadIdx.x; c < tw * tw; c += DEP_THREADS) sg[c] = 0.0;
    __syncthreads();
    for (int64_t i = (int64_t)chunk * DEP_THREADS + threadIdx.x; i < a.n; i += (int64_t)nchunk * DEP_THREADS) {
        Foot fp;
        if (!footprint(a, i, fp)) continue;
        const int64_t cells = (int64_t)(fp.gx1 - fp.gx0 + 1) * (fp.gy1 - fp.gy0 + 1);
        if (cells > BIG_CELLS) {
            // listed once, by the tile that holds its first cell; the list
            // holds n entries
            if (fp.gx0 >= tx0 && fp.gx0 <= tx1 && fp.gy0 >= ty0 && fp.gy0 <= ty1)
                a.big[atomicAdd(a.nbig, 1u)] = (uint32_t)i;
            continue;
        }
        if (a.conservative && ntx == 1 && cells <= WCACHE2) {
            // one tile holds the whole map (and footprint): small footprints
            // keep their weights in registers between the total (summed in
            // the order of the loop below) and the deposit
            double wc[WCACHE2];
            double tot = 0.0;
            int gx = fp.gx0, gy = fp.gy0;
#pragma unroll
            for (int c = 0; c < WCACHE2; c++) {
                wc[c] = 0.0;
                if (c < cells) {
                    wc[c] = cell_weight(a, fp, gx, gy);
                    tot += wc[c];
                    if (++gy > fp.gy1) {
                        gy = fp.gy0;
                        gx++;
                    }
                }
            }
            if (tot == 0.0) continue;
            gx = fp.gx0;
            gy = fp.gy0;
#pragma unroll
            for (int c = 0; c < WCACHE2; c++) {
                if (c < cells) {
                    if (wc[c] != 0.0) atomicAdd(&sg[(gx - tx0) * tw + (gy - ty0)], wc[c] * fp.w / tot);
                    if (++gy > fp.gy1) {
                        gy = fp.gy0;
                        gx++;
                    }
                }
            }
            continue;
        }
        const int gx0 = max(fp.gx0, tx0), gx1 = min(fp.gx1, tx1);
        const int gy0 = max(fp.gy0, ty0), gy1 = min(fp.gy1, ty1);
        if (gx0 > gx1 || gy0 > gy1) continue;
        // conservative: each weight over the particle's total weight (the
        // whole footprint, not only this tile's part)
        double tot = 1.0;
        if (a.conservative) {
            tot = 0.0;
            for (int gx = fp.gx0; gx <= fp.gx1; gx++)
                for (int gy = fp.gy0; gy <= fp.gy1; gy++) tot += cell_weight(a, fp, gx, gy);
            if (tot == 0.0) continue;
        }
        for (int gx = gx0; gx <= gx1; gx++)
            for (int gy = gy0; gy <= gy1; gy++) {
                const double k = cell_weight(a, fp, gx, gy);
                if (k != 0.0) atomicAdd(&sg[(gx - tx0) * tw + (gy - ty0)], a.conservative ? k * fp.w / tot : k * fp.w);
            }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < tw * tw; c += DEP_THREADS) {
        const double v = sg[c];
        if (v != 0.0) atomicAdd(&a.grid[(int64_t)(tx0 + c / tw) * a.res + (ty0 + c % tw)], v);
    }
}

// conservative variant: the particle's total weight over its (clipped)
// footprint, by which each weight is divided; 1 for the plain variant
MHIP_DEV double foot_total(const DepArgs &a, const Foot &fp) {
    if (!a.conservative) return 1.0;
    double tot = 0.0;
    for (int gx = fp.gx0; gx <= fp.gx1; gx++)
        for (int gy = fp.gy0; gy <= fp.gy1; gy++) tot += cell_weight(a, fp, gx, gy);
    return tot;
}

// ---------------------------------------------------------------------------
// Binned deposition.  A policy P gives, for its kind of grid:
//   key(a, i, cells)     the brick of particle i's first footprint cell (and
//                        the footprint's cells), or skip(a) when it has
//                        nothing to deposit here (no footprint, or listed for
//                        the block-per-particle kernel)
//   direct(a, i)         deposit particle i with float64 atomics in HBM
//   window(a, i, k, w)   deposit particle i (brick k) into the LDS window w
//   flush(a, k, c, v)    add window cell c of brick k (value v) to the map
// ---------------------------------------------------------------------------
struct Map2 {
    using Args = DepArgs;
    static MHIP_DEV uint32_t skip(const DepArgs &a) { return (uint32_t)a.nb * (uint32_t)a.nb; }
    static MHIP_DEV uint32_t key(const DepArgs &a, int64_t i, uint32_t &cells) {
        Foot fp;
        cells = 0;
        if (!footprint(a, i, fp)) return skip(a);
        const int sx = fp.gx1 - fp.gx0 + 1, sy = fp.gy1 - fp.gy0 + 1;
        if (sx > WIN2 - BRICK2 + 1 || sy > WIN2 - BRICK2 + 1 || (int64_t)sx * sy > BIG_CELLS) {
            a.big[atomicAdd(a.nbig, 1u)] = (uint32_t)i;
            return skip(a);
        }
        cells = (uint32_t)(sx * sy);
        return (uint32_t)(fp.gx0 / BRICK2) * (uint32_t)a.nb + (uint32_t)(fp.gy0 / BRICK2);
    }
    static MHIP_DEV void direct(const DepArgs &a, uint32_t i) {
        Foot fp;
        footprint(a, i, fp);  // keyed, so it has a footprint
        const double tot = foot_total(a, fp);
        if (tot == 0.0) return;
        for (int gx = fp.gx0; gx <= fp.gx1; gx++)
            for (int gy = fp.gy0; gy <= fp.gy1; gy++) {
                const double k = cell_weight(a, fp, gx, gy);
                if (k != 0.0) atomicAdd(&a.grid[(int64_t)gx * a.res + gy], a.conservative ? k * fp.w / tot : k * fp.w);
            }
    }
    // the window holds cells [ox, ox + 64) x [oy, oy + 64) of the brick's
    // origin (ox, oy); a keyed footprint starts inside the brick and spans at
    // most 33 cells per axis, so it stays inside
    static MHIP_DEV void window(const DepArgs &a, uint32_t i, uint32_t key, double *win) {
        Foot fp;
        footprint(a, i, fp);
        const int ox = (int)(key / (uint32_t)a.nb) * BRICK2, oy = (int)(key % (uint32_t)a.nb) * BRICK2;
        const int wy = fp.gy1 - fp.gy0 + 1, ncell = (fp.gx1 - fp.gx0 + 1) * wy;
        if (a.conservative && ncell <= WCACHE2) {
            // small footprints: the weights stay in registers between the
            // total (summed in foot_total's order) and the deposit
            double wc[WCACHE2];
            double tot = 0.0;
            int gx = fp.gx0, gy = fp.gy0;
#pragma unroll
            for (int c = 0; c < WCACHE2; c++) {
                wc[c] = 0.0;
                if (c < ncell) {
                    wc[c] = cell_weight(a, fp, gx, gy);
                    tot += wc[c];
                    if (++gy > fp.gy1) {
                        gy = fp.gy0;
                        gx++;
                    }
                }
            }
            if (tot == 0.0) return;
            gx = fp.gx0;
            gy = fp.gy0;
#pragma unroll
            for (int c = 0; c < WCACHE2; c++) {
                if (c < ncell) {
                    if (wc[c] != 0.0) atomicAdd(&win[(gx - ox) * WIN2 + (gy - oy)], wc[c] * fp.w / tot);
                    if (++gy > fp.gy1) {
                        gy = fp.gy0;
                        gx++;
                    }
                }
            }
            return;
        }
        const double tot = foot_total(a, fp);
        if (tot == 0.0) return;
        for (int gx = fp.gx0; gx <= fp.gx1; gx++)
            for (int gy = fp.gy0; gy <= fp.gy1; gy++) {
                const double k = cell_weight(a, fp, gx, gy);
                if (k != 0.0) atomicAdd(&win[(gx - ox) * WIN2 + (gy - oy)], a.conservative ? k * fp.w / tot : k * fp.w);
            }
    }
    static MHIP_DEV void flush(const DepArgs &a, uint32_t key, int c, double v) {
        const int gx = (int)(key / (uint32_t)a.nb) * BRICK2 + c / WIN2;
        const int gy = (int)(key % (uint32_t)a.nb) * BRICK2 + c % WIN2;
        if (gx < a.res && gy < a.res) atomicAdd(&a.grid[(int64_t)gx * a.res + gy], v);
    }
};

template <class P>
__global__ __launch_bounds__(256) void k_bin_keys(typename P::Args a, uint64_t *__restrict__ keys,
                                                  uint32_t *__restrict__ vals) {
    const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (i >= a.n) return;
    uint32_t cells;
    const uint32_t b = P::key(a, i, cells);
    keys[i] = (uint64_t)b << SIZE_BITS | (SIZE_MAX_RANK - min(cells, SIZE_MAX_RANK));
    vals[i] = (uint32_t)i;
}

MHIP_DEV int first_not_below(const uint32_t *s, int m, uint32_t k) {
    int lo = 0, hi = m;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s[mid] < k)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

MHIP_DEV int first_above(const uint32_t *s, int m, uint32_t k) {
    int lo = 0, hi = m;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s[mid] <= k)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// block b: sorted particles [b BIN_CHUNK, (b + 1) BIN_CHUNK).  Their bricks'
// runs (lengths from binary searches of the block's keys in LDS) shorter
// than BIN_DENSE are deposited straight into HBM; each longer run is
// deposited into the LDS window by the whole block and added to the map
// once (cells still zero are skipped), the window cleared as it is read
template <class P>
__global__ __launch_bounds__(DEP_THREADS) void k_bin_deposit(typename P::Args a, const uint64_t *__restrict__ keys,
                                                             const uint32_t *__restrict__ vals) {
    __shared__ double win[WIN_CELLS];
    __shared__ uint32_t skey[BIN_CHUNK];
    __shared__ uint32_t runs[BIN_CHUNK / BIN_DENSE];  // offset | length << 16
    __shared__ uint32_t s_nrun;
    const int64_t cs = (int64_t)blockIdx.x * BIN_CHUNK;
    const int m = (int)min<int64_t>(BIN_CHUNK, a.n - cs);
    const uint32_t skip = P::skip(a);
    for (int c = threadIdx.x; c < WIN_CELLS; c += DEP_THREADS) win[c] = 0.0;
    for (int e = threadIdx.x; e < m; e += DEP_THREADS) skey[e] = (uint32_t)(keys[cs + e] >> SIZE_BITS);
    if (threadIdx.x == 0) s_nrun = 0u;
    __syncthreads();
    for (int e = threadIdx.x; e < m; e += DEP_THREADS) {
        const uint32_t k = skey[e];
        if (k == skip) continue;
        const int lo = first_not_below(skey, m, k), hi = first_above(skey, m, k);
        if (hi - lo >= BIN_DENSE) {
            // at most BIN_CHUNK / BIN_DENSE runs this long
            if (e == lo) runs[atomicAdd(&s_nrun, 1u)] = (uint32_t)lo | (uint32_t)(hi - lo) << 16;
        } else {
            P::direct(a, vals[cs + e]);
        }
    }
    __syncthreads();
    const uint32_t nrun = s_nrun;
    for (uint32_t r = 0; r < nrun; r++) {
        const int lo = (int)(runs[r] & 0xffffu), len = (int)(runs[r] >> 16);
        const uint32_t k = skey[lo];
        for (int e = threadIdx.x; e < len; e += DEP_THREADS) P::window(a, vals[cs + lo + e], k, win);
        __syncthreads();
        for (int c = threadIdx.x; c < WIN_CELLS; c += DEP_THREADS) {
            const double v = win[c];
            if (v != 0.0) {
                P::flush(a, k, c, v);
                win[c] = 0.0;
            }
        }
        __syncthreads();
    }
}

// bits of a sort key in [0, skip]
static int key_bits(uint32_t skip) { return 64 - __builtin_clzll((unsigned long long)skip | 1ull); }

// the binned path: keys, sort, runs; the block-per-particle list in big/nbig
// (cleared beforehand).  Scratch: keys 16 n, values 8 n, radix counts.
template <class P>
int bin_deposit(mhip_ctx *ctx, typename P::Args a, uint32_t skip, char *sc) {
    const int64_t n = a.n;
    if (n <= 0) return 0;
    uint64_t *k0 = (uint64_t *)sc, *k1 = k0 + n;
    uint32_t *v0 = (uint32_t *)(k1 + n), *v1 = v0 + n, *counts = v1 + n;
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(k_bin_keys<P>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, k0, v0);
    uint64_t *ks;
    uint32_t *vs;
    int rc = radix_sort_pairs(ctx, k0, v0, k1, v1, counts, n, key_bits(skip) + SIZE_BITS, &ks, &vs);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bin_deposit<P>, dim3((unsigned)((n + BIN_CHUNK - 1) / BIN_CHUNK)), dim3(DEP_THREADS), 0, s,
                       a, ks, vs);
    return 0;
}

// scratch bytes of bin_deposit plus the block-per-particle list (n + 1 words)
static size_t bin_scratch_bytes(int64_t n) {
    return (size_t)n * 24 + radix_counts_words(std::max<int64_t>(n, 1)) * 4 + ((size_t)n + 1) * 4 + 256;
}

// one block per listed particle (footprints above BIG_CELLS), HBM atomics
__global__ __launch_bounds__(DEP_THREADS) void k_deposit_big(DepArgs a) {
    __shared__ double red[DEP_THREADS];
    const uint32_t nb = *a.nbig;
    for (uint32_t e = blockIdx.x; e < nb; e += gridDim.x) {
        Foot fp;
        footprint(a, a.big[e], fp);  // listed, so it has a footprint
        const int wy = fp.gy1 - fp.gy0 + 1;
        const int64_t cells = (int64_t)(fp.gx1 - fp.gx0 + 1) * wy;
        double tot = 1.0;
        if (a.conservative) {
            double part = 0.0;
            for (int64_t c = threadIdx.x; c < cells; c += DEP_THREADS)
                part += cell_weight(a, fp, fp.gx0 + (int)(c / wy), fp.gy0 + (int)(c % wy));
            red[threadIdx.x] = part;
            __syncthreads();
            for (int o = DEP_THREADS / 2; o > 0; o >>= 1) {
                if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
                __syncthreads();
            }
            tot = red[0];
            __syncthreads();
            if (tot == 0.0) continue;
        }
        for (int64_t c = threadIdx.x; c < cells; c += DEP_THREADS) {
            const int gx = fp.gx0 + (int)(c / wy), gy = fp.gy0 + (int)(c % wy);
            const double k = cell_weight(a, fp, gx, gy);
            if (k != 0.0) atomicAdd(&a.grid[(int64_t)gx * a.res + gy], a.conservative ? k * fp.w / tot : k * fp.w);
        }
    }
}

__global__ __launch_bounds__(256) void k_grid_fill(double *__restrict__ g, int64_t n, double v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        g[i] = v;
}

__global__ __launch_bounds__(256) void k_grid_scale(double *__restrict__ g, int64_t n, double s) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        g[i] *= s;
}

__global__ void k_u32_zero(uint32_t *p) {
    if (threadIdx.x == 0) *p = 0u;
}

// ---------------------------------------------------------------------------
// GridDensity: 3-D conservative deposition.  Grid point g at
// center - size/2 + dx/2 + g dx (dx = size/(res-1)); each particle spreads f
// over the points of its cube footprint with weights kernel3d(r/h) / (their
// sum), and the map is divided by dx^3.  Periodic boxes wrap the indices
// (g mod res) instead of clipping them.  kernel3d is the compact-support
// spline: 0 beyond r = h (the reference leaves it unassigned there).
// ---------------------------------------------------------------------------
constexpr double KERNEL3D_NORM = 2.546479089470325;  // 8 / pi
constexpr int64_t BIG_CELLS_3D = 512;

struct Dep3Args {
    const void *f, *pos, *h;
    int f_f64, pos_f64, h_f64;
    int64_t n;
    double hmin;        // h used = min(max(h, hmin), 1e100)
    double ox, oy, oz;  // position of grid point 0
    double dx;
    int res, periodic;
    double *grid;  // (res, res, res)
    uint32_t *big, *nbig;
    int nb;  // bricks per axis
};

struct Foot3 {
    double x[3], hs, hinv, f;
    int g0[3], g1[3];
};

MHIP_DEV double kernel3d(double q) {
    if (q > 1.0) return 0.0;
    double k;
    if (q <= 0.5) {
        k = 1.0 - 6.0 * q * q * (1.0 - q);
    } else {
        const double a = 1.0 - q;
        k = 2.0 * a * a * a;
    }
    return k * KERNEL3D_NORM;
}

MHIP_DEV bool footprint3(const Dep3Args &a, int64_t i, Foot3 &fp) {
    const double o[3] = {a.ox, a.oy, a.oz};
    for (int c = 0; c < 3; c++)
        fp.x[c] = (a.pos_f64 ? ld<double>(a.pos, i * 3 + c) : ld<float>(a.pos, i * 3 + c)) - o[c];
    const double hv = a.h_f64 ? ld<double>(a.h, i) : ld<float>(a.h, i);
    fp.f = a.f_f64 ? ld<double>(a.f, i) : ld<float>(a.f, i);
    const double hs = fmin(fmax(hv, a.hmin), 1e100);
    if (!(hs > 0.0) || !isfinite(fp.x[0]) || !isfinite(fp.x[1]) || !isfinite(fp.x[2]) || !isfinite(fp.f))
        return false;
    fp.hs = hs;
    fp.hinv = 1.0 / hs;
    for (int c = 0; c < 3; c++) {
        const double lo = (fp.x[c] - hs) / a.dx + 1.0, hi = (fp.x[c] + hs) / a.dx;
        if (a.periodic) {
            // unclipped int() (truncation toward zero); the footprint spans at
            // most 2 h / dx + 1 points, h <= the map's extent in practice
            if (!(fabs(lo) < 1e9 && fabs(hi) < 1e9)) return false;
            fp.g0[c] = (int)lo;
            fp.g1[c] = (int)hi;
        } else {
            fp.g0[c] = lo_index(lo, a.res);
            fp.g1[c] = hi_index(hi, a.res);
        }
        if (fp.g0[c] > fp.g1[c]) return false;
    }
    return true;
}

MHIP_DEV double weight3(const Dep3Args &a, const Foot3 &fp, int gx, int gy, int gz) {
    double dx2 = fp.x[0] - gx * a.dx;
    dx2 *= dx2;
    double dy2 = fp.x[1] - gy * a.dx;
    dy2 *= dy2;
    double dz2 = fp.x[2] - gz * a.dx;
    dz2 *= dz2;
    return kernel3d(sqrt(dx2 + dy2 + dz2) * fp.hinv);
}

MHIP_DEV int64_t cell3(const Dep3Args &a, int gx, int gy, int gz) {
    if (a.periodic) {
        gx = ((gx % a.res) + a.res) % a.res;
        gy = ((gy % a.res) + a.res) % a.res;
        gz = ((gz % a.res) + a.res) % a.res;
    }
    return ((int64_t)gx * a.res + gy) * a.res + gz;
}

MHIP_DEV double foot3_total(const Dep3Args &a, const Foot3 &fp) {
    double tot = 0.0;
    for (int gx = fp.g0[0]; gx <= fp.g1[0]; gx++)
        for (int gy = fp.g0[1]; gy <= fp.g1[1]; gy++)
            for (int gz = fp.g0[2]; gz <= fp.g1[2]; gz++) tot += weight3(a, fp, gx, gy, gz);
    return tot;
}

// first footprint index on axis c, wrapped into the grid when periodic
MHIP_DEV int wrapped_first(const Dep3Args &a, const Foot3 &fp, int c) {
    return a.periodic ? ((fp.g0[c] % a.res) + a.res) % a.res : fp.g0[c];
}

// binned volumes: the brick of the (wrapped) first footprint cell; the
// window holds 16^3 cells from the brick's origin, a keyed footprint spans at
// most 9 cells per axis from inside the brick.  Periodic footprints keep
// their unwrapped extent in the window and wrap when it is added to the map.
struct Vol3 {
    using Args = Dep3Args;
    static MHIP_DEV uint32_t skip(const Dep3Args &a) { return (uint32_t)a.nb * (uint32_t)a.nb * (uint32_t)a.nb; }
    static MHIP_DEV uint32_t key(const Dep3Args &a, int64_t i, uint32_t &ncell) {
        Foot3 fp;
        ncell = 0;
        if (!footprint3(a, i, fp)) return skip(a);
        int64_t cells = 1;
        bool fits = true;
        for (int c = 0; c < 3; c++) {
            const int64_t sp = (int64_t)fp.g1[c] - fp.g0[c] + 1;
            fits = fits && sp <= WIN3 - BRICK3 + 1;
            cells *= fits ? sp : 1;
        }
        if (!fits || cells > BIG_CELLS_3D) {
            a.big[atomicAdd(a.nbig, 1u)] = (uint32_t)i;
            return skip(a);
        }
        ncell = (uint32_t)cells;
        const uint32_t nb = (uint32_t)a.nb;
        return ((uint32_t)(wrapped_first(a, fp, 0) / BRICK3) * nb + (uint32_t)(wrapped_first(a, fp, 1) / BRICK3)) * nb +
               (uint32_t)(wrapped_first(a, fp, 2) / BRICK3);
    }
    static MHIP_DEV void direct(const Dep3Args &a, uint32_t i) {
        Foot3 fp;
        footprint3(a, i, fp);  // keyed, so it has a footprint
        const double tot = foot3_total(a, fp);
        if (!(tot > 0.0)) return;
        for (int gx = fp.g0[0]; gx <= fp.g1[0]; gx++)
            for (int gy = fp.g0[1]; gy <= fp.g1[1]; gy++)
                for (int gz = fp.g0[2]; gz <= fp.g1[2]; gz++) {
                    const double k = weight3(a, fp, gx, gy, gz);
                    if (k != 0.0) atomicAdd(&a.grid[cell3(a, gx, gy, gz)], fp.f * k / tot);
                }
    }
    static MHIP_DEV void window(const Dep3Args &a, uint32_t i, uint32_t key, double *win) {
        Foot3 fp;
        footprint3(a, i, fp);
        // window index of the footprint's first cell on each axis (0..7)
        const uint32_t nb = (uint32_t)a.nb;
        const int b[3] = {(int)(key / (nb * nb)), (int)(key / nb % nb), (int)(key % nb)};
        int s[3];
        for (int c = 0; c < 3; c++) s[c] = wrapped_first(a, fp, c) - b[c] * BRICK3;
        const int wy = fp.g1[1] - fp.g0[1] + 1, wz = fp.g1[2] - fp.g0[2] + 1;
        const int ncell = (fp.g1[0] - fp.g0[0] + 1) * wy * wz;
        if (ncell <= WCACHE3) {
            // small footprints: the weights stay in registers between the
            // total (summed in foot3_total's order) and the deposit
            double wc[WCACHE3];
            double tot = 0.0;
            int gx = fp.g0[0], gy = fp.g0[1], gz = fp.g0[2];
#pragma unroll
            for (int c = 0; c < WCACHE3; c++) {
                wc[c] = 0.0;
                if (c < ncell) {
                    wc[c] = weight3(a, fp, gx, gy, gz);
                    tot += wc[c];
                    if (++gz > fp.g1[2]) {
                        gz = fp.g0[2];
                        if (++gy > fp.g1[1]) {
                            gy = fp.g0[1];
                            gx++;
                        }
                    }
                }
            }
            if (!(tot > 0.0)) return;
            gx = fp.g0[0];
            gy = fp.g0[1];
            gz = fp.g0[2];
#pragma unroll
            for (int c = 0; c < WCACHE3; c++) {
                if (c < ncell) {
                    if (wc[c] != 0.0) {
                        const int w = ((gx - fp.g0[0] + s[0]) * WIN3 + (gy - fp.g0[1] + s[1])) * WIN3 +
                                      (gz - fp.g0[2] + s[2]);
                        atomicAdd(&win[w], fp.f * wc[c] / tot);
                    }
                    if (++gz > fp.g1[2]) {
                        gz = fp.g0[2];
                        if (++gy > fp.g1[1]) {
                            gy = fp.g0[1];
                            gx++;
                        }
                    }
                }
            }
            return;
        }
        const double tot = foot3_total(a, fp);
        if (!(tot > 0.0)) return;
        for (int gx = fp.g0[0]; gx <= fp.g1[0]; gx++)
            for (int gy = fp.g0[1]; gy <= fp.g1[1]; gy++)
                for (int gz = fp.g0[2]; gz <= fp.g1[2]; gz++) {
                    const double k = weight3(a, fp, gx, gy, gz);
                    if (k != 0.0) {
                        const int c = ((gx - fp.g0[0] + s[0]) * WIN3 + (gy - fp.g0[1] + s[1])) * WIN3 +
                                      (gz - fp.g0[2] + s[2]);
                        atomicAdd(&win[c], fp.f * k / tot);
                    }
                }
    }
    static MHIP_DEV void flush(const Dep3Args &a, uint32_t key, int c, double v) {
        const uint32_t nb = (uint32_t)a.nb;
        const int gx = (int)(key / (nb * nb)) * BRICK3 + c / (WIN3 * WIN3);
        const int gy = (int)(key / nb % nb) * BRICK3 + c / WIN3 % WIN3;
        const int gz = (int)(key % nb) * BRICK3 + c % WIN3;
        if (!a.periodic && (gx >= a.res || gy >= a.res || gz >= a.res)) return;
        atomicAdd(&a.grid[cell3(a, gx, gy, gz)], v);
    }
};

// one block per listed particle (footprints above BIG_CELLS_3D)
__global__ __launch_bounds__(DEP_THREADS) void k_density3d_big(Dep3Args a) {
    __shared__ double red[DEP_THREADS];
    const uint32_t nb = *a.nbig;
    for (uint32_t e = blockIdx.x; e < nb; e += gridDim.x) {
        Foot3 fp;
        footprint3(a, a.big[e], fp);  // listed, so it has a footprint
        const int wy = fp.g1[1] - fp.g0[1] + 1, wz = fp.g1[2] - fp.g0[2] + 1;
        const int64_t cells = (int64_t)(fp.g1[0] - fp.g0[0] + 1) * wy * wz;
        double part = 0.0;
        for (int64_t c = threadIdx.x; c < cells; c += DEP_THREADS)
            part += weight3(a, fp, fp.g0[0] + (int)(c / ((int64_t)wy * wz)), fp.g0[1] + (int)((c / wz) % wy),
                            fp.g0[2] + (int)(c % wz));
        red[threadIdx.x] = part;
        __syncthreads();
        for (int o = DEP_THREADS / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        const double tot = red[0];
        __syncthreads();
        if (!(tot > 0.0)) continue;
        for (int64_t c = threadIdx.x; c < cells; c += DEP_THREADS) {
            const int gx = fp.g0[0] + (int)(c / ((int64_t)wy * wz)), gy = fp.g0[1] + (int)((c / wz) % wy),
                      gz = fp.g0[2] + (int)(c % wz);
            const double k = weight3(a, fp, gx, gy, gz);
            if (k != 0.0) atomicAdd(&a.grid[cell3(a, gx, gy, gz)], fp.f * k / tot);
        }
    }
}

int grid_density(mhip_ctx *ctx, const void *f, int f_f64, const void *pos, int pos_f64, int64_t n, const void *h,
                 int h_f64, double hmin, double cx, double cy, double cz, double size, int res, int periodic,
                 double *grid) {
    if (n < 0) return fail(ctx, -1, "negative particle count");
    if (res < 2 || res > 2048) return fail(ctx, -1, "res must be in [2, 2048]");
    if (!(size > 0.0)) return fail(ctx, -1, "size must be positive");
    if (n > 0xffffffffll) return fail(ctx, -4, "more than 2^32 particles");
    Dep3Args a;
    a.f = f;
    a.pos = pos;
    a.h = h;
    a.f_f64 = f_f64;
    a.pos_f64 = pos_f64;
    a.h_f64 = h_f64;
    a.n = n;
    a.hmin = hmin;
    a.dx = size / (res - 1);
    // x - center + size/2 - dx/2: point g at center - size/2 + dx/2 + g dx
    a.ox = cx - 0.5 * size + 0.5 * a.dx;
    a.oy = cy - 0.5 * size + 0.5 * a.dx;
    a.oz = cz - 0.5 * size + 0.5 * a.dx;
    a.res = res;
    a.periodic = periodic ? 1 : 0;
    a.grid = grid;
    a.nb = (res + BRICK3 - 1) / BRICK3;
    int rc = 0;
    char *sc = (char *)scratch(ctx, bin_scratch_bytes(n), &rc);
    if (!sc) return rc;
    // the sort buffers first (8-byte keys), the block-per-particle list last
    a.nbig = (uint32_t *)(sc + (size_t)n * 24 + radix_counts_words(std::max<int64_t>(n, 1)) * 4);
    a.big = a.nbig + 1;
    hipStream_t s = ctx->stream;
    const int64_t ncell = (int64_t)res * res * res;
    ProfRegion pr(ctx, "grid_density");
    hipLaunchKernelGGL(k_grid_fill, dim3((unsigned)std::min<int64_t>(4096, (ncell + 255) / 256)), dim3(256), 0, s,
                       grid, ncell, 0.0);
    hipLaunchKernelGGL(k_u32_zero, dim3(1), dim3(64), 0, s, a.nbig);
    if (n > 0) {
        if ((rc = bin_deposit<Vol3>(ctx, a, (uint32_t)(a.nb * a.nb * a.nb), sc))) return rc;
        hipLaunchKernelGGL(k_density3d_big, dim3(2048), dim3(DEP_THREADS), 0, s, a);
    }
    const double inv = 1.0 / (a.dx * a.dx * a.dx);
    hipLaunchKernelGGL(k_grid_scale, dim3((unsigned)std::min<int64_t>(4096, (ncell + 255) / 256)), dim3(256), 0, s,
                       grid, ncell, inv);
    return check_hip(ctx, hipGetLastError(), "grid density launch");
}

int surface_density(mhip_ctx *ctx, const void *f, int f_f64, const void *pos, int pos_f64, int64_t n, int dim,
                    const void *h, int h_f64, double smooth, double hmin, double cx, double cy, double size, int res,
                    int conservative, double *grid) {
    if (n < 0 || dim < 2 || dim > 3) return fail(ctx, -1, "surface density needs (n, 2) or (n, 3) positions");
    if (res < 2 || res > (1 << 20)) return fail(ctx, -1, "res must be in [2, 2^20]");
    if (!(size > 0.0)) return fail(ctx, -1, "size must be positive");
    if (n > 0xffffffffll) return fail(ctx, -4, "more than 2^32 particles");
    DepArgs a;
    a.f = f;
    a.pos = pos;
    a.h = h;
    a.f_f64 = f_f64;
    a.pos_f64 = pos_f64;
    a.h_f64 = h_f64;
    a.dim = dim;
    a.n = n;
    a.smooth = smooth;
    a.hmin = hmin;  // SurfaceDensity: np.clip(smooth_fac * h, 2 * size / res, 1e100)
    a.ox = cx - 0.5 * size;
    a.oy = cy - 0.5 * size;
    a.dx = size / (res - 1);
    a.res = res;
    a.conservative = conservative ? 1 : 0;
    a.grid = grid;
    a.nb = (res + BRICK2 - 1) / BRICK2;
    int rc = 0;
    char *sc = (char *)scratch(ctx, bin_scratch_bytes(n), &rc);
    if (!sc) return rc;
    a.nbig = (uint32_t *)(sc + (size_t)n * 24 + radix_counts_words(std::max<int64_t>(n, 1)) * 4);
    a.big = a.nbig + 1;
    hipStream_t s = ctx->stream;
    const int64_t ncell = (int64_t)res * res;
    ProfRegion pr(ctx, "surface_density");
    hipLaunchKernelGGL(k_grid_fill, dim3((unsigned)std::min<int64_t>(1024, (ncell + 255) / 256)), dim3(256), 0, s,
                       grid, ncell, 0.0);
    hipLaunchKernelGGL(k_u32_zero, dim3(1), dim3(64), 0, s, a.nbig);
    if (n > 0) {
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || ncu < 1)
            ncu = 256;
        if (res <= DEP_TILE) {
            // the whole map in LDS: one block per CU, particles interleaved
            // over the blocks, each block adds its map to HBM once
            const size_t lds = (size_t)res * res * sizeof(double);
            if (lds > 65536 && hipFuncSetAttribute((const void *)k_deposit_tiles,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
                return check_hip(ctx, hipGetLastError(), "deposit LDS size");
            const int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>(ncu, (n + 255) / 256));
            hipLaunchKernelGGL(k_deposit_tiles, dim3((unsigned)nchunk), dim3(DEP_THREADS), lds, s, a, res, 1,
                               (int)nchunk);
        } else {
            // binned by 32^2 bricks (per-tile LDS copies of the map would
            // re-read every particle per tile)
            if ((rc = bin_deposit<Map2>(ctx, a, (uint32_t)a.nb * (uint32_t)a.nb, sc))) return rc;
        }
        hipLaunchKernelGGL(k_deposit_big, dim3(2048), dim3(DEP_THREADS), 0, s, a);
    }
    if (conservative) {
        const double dx2inv = 1.0 / (a.dx * a.dx);
        hipLaunchKernelGGL(k_grid_scale, dim3((unsigned)std::min<int64_t>(1024, (ncell + 255) / 256)), dim3(256), 0,
                           s, grid, ncell, dx2inv);
    }
    return check_hip(ctx, hipGetLastError(), "surface density launch");
}

}  // namespace mhip
