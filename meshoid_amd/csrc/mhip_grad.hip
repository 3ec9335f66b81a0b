// Least-squares derivative operators on the GPU.
//
// Replaces, for the hot path,
//   gradient_weights(...)                       src/meshoid/derivatives.py:117-138
//     kernel_dx_and_weights                     src/meshoid/derivatives.py:26-52
//     polyfit_leastsq_matrices                  src/meshoid/derivatives.py:55-103
//   np.einsum("ikj,ij...->i...k", G, df)        src/meshoid/Meshoid.py:258,268,291,295
//
// The fused operator never materialises G = E^-1 [w a_j]: since
//   D f_i = sum_j G_ij (f_j - f_i) = E_i^-1 sum_j w_ij a_ij (f_j - f_i),
// each thread accumulates the moment matrix E (float64, from float32 pair
// offsets relative to the host particle) and the right-hand side of every
// field, then solves with an unpivoted float64 LDL^T (E is symmetric positive
// semi-definite; an exactly zero or NaN pivot is reported as singular, as
// np.linalg.inv raises on an exactly zero pivot).
#include <algorithm>
#include <type_traits>

#include "mhip_internal.hpp"
#include "mhip_linalg.hpp"

namespace mhip {

// ---------------------------------------------------------------------------
// fused operator over the neighbour list
// ---------------------------------------------------------------------------


// float64 offset of a neighbour (float32 position c, float32 residual cl of
// float64 inputs) from the host particle (qx, qy, qz): exact for float32
// positions, minimum image when periodic
MHIP_DEV void pair_offset(const ApplyArgs &a, const float (&c)[3], const float (&cl)[3], double qx, double qy,
                          double qz, double (&d)[3]) {
    d[0] = (double)c[0] - qx;
    d[1] = (double)c[1] - qy;
    d[2] = (double)c[2] - qz;
    if (a.p64) {
        d[0] += (double)cl[0];
        d[1] += (double)cl[1];
        d[2] += (double)cl[2];
    }
    if (a.per) {
        const double box = (double)a.box;
#pragma unroll
        for (int c2 = 0; c2 < 3; c2++) d[c2] = wrap<double>(d[c2], box);
    }
}

// neighbours per load batch of one lane (second-order fits): their indices
// are loaded first, then their positions and field values, then the sums run.
// A second-order lane serves all k neighbours of its query with few
// wavefronts per SIMD (~230 registers), so a loop that waits on one
// neighbour's dependent loads at a time stalls (measured 6.25 -> 5.24 ms at
// 2^24, K = 2, with batches of 4; 7.2 ms with 8).  First-order fits keep one
// neighbour per batch: their lanes and higher occupancy hide the latency
// already, and batching cost occupancy (with the packed loads and two lanes
// per query: 1.97 ms, against 2.31 with 2 and 2.42 with 4 per batch).
template <int ORDER>
constexpr int apply_nb() { return ORDER == 1 ? 1 : 4; }

// |d| from r2 = |d|^2 (float64): the hardware reciprocal square root and one
// Newton step on the root give ~1e-16 relative error without the correctly
// rounded expansion's scaling and fix-up instructions (apply_D1 2.27 ->
// 2.05 ms at 2^24 with four lanes per query).  Tiny r2 (float64 inputs at
// pathological scales, where the estimate would meet denormals) take sqrt.
MHIP_DEV double pair_dist(double r2) {
    if (r2 < 1e-280) return sqrt(r2);
    const double y = __builtin_amdgcn_rsq(r2);
    const double r = r2 * y;
    const double res = fma(-r, r, r2);
    return fma(0.5 * y, res, r);
}

__global__ __launch_bounds__(256) void k_pack_xyzf(const float *__restrict__ sx, const float *__restrict__ sy,
                                                  const float *__restrict__ sz, const float *__restrict__ f0,
                                                  int64_t npad, float4 *__restrict__ out) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p < npad) out[p] = make_float4(sx[p], sy[p], sz[p], f0[p]);
}

// the same pack straight from a one-component field in original row order
// (the field value through the sorted -> original map; NaN past nf_rows and
// in the padding, as k_permute_fields gives)
__global__ __launch_bounds__(256) void k_pack_xyzf_rows(const float *__restrict__ sx, const float *__restrict__ sy,
                                                       const float *__restrict__ sz, const float *__restrict__ f,
                                                       int64_t nf_rows, const uint32_t *__restrict__ perm,
                                                       int64_t n, int64_t npad, float4 *__restrict__ out) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= npad) return;
    const uint32_t o = p < n ? perm[p] : 0xffffffffu;
    out[p] = make_float4(sx[p], sy[p], sz[p], o < nf_rows ? f[o] : NAN);
}

// one query slot s of the list, fields [fb, fe): KC fields per pass over the
// neighbours.  With GW > 1 a group of GW consecutive lanes serves the query:
// lane gl takes neighbours gl, gl + GW, ..., the moment matrix and right-hand
// sides are summed over the group with shuffles, and lane 0 writes.  h is
// the list's float64-refined smoothing length (the kNN pass refines it over
// the final members, so every operator fits with the same h).
template <int DIM, int ORDER, typename FT, int KC, int GW = 1>
MHIP_DEV void apply_row(const ApplyArgs &a, int64_t s, int fb, int fe, int gl = 0) {
    constexpr int ND = NDer<DIM, ORDER>::ND, NE = NDer<DIM, ORDER>::NE;
    const uint32_t qp = a.qsorted ? a.qsorted[s] : (uint32_t)s;
    const uint32_t row = a.qrow ? a.qrow[s] : a.perm[s];
    // host particle in float64: hi + lo reproduces float64 inputs to 2^-48
    // and float32 inputs exactly, so the pair offsets below are exact for
    // float32 positions
    const double qx = (double)a.sx[qp] + (a.p64 ? (double)a.lx[qp] : 0.0);
    const double qy = (double)a.sy[qp] + (a.p64 ? (double)a.ly[qp] : 0.0);
    const double qz = (double)a.sz[qp] + (a.p64 ? (double)a.lz[qp] : 0.0);
    const double hd = a.h[s];
    const double ih = hd > 0.0 ? 1.0 / hd : (double)INFINITY;
    const FT *fs = (const FT *)a.fs;
    int32_t st = 0;
    for (int f0 = fb; f0 < fe; f0 += KC) {
        double E[NE];
        double B[KC][ND];
        FT fc[KC];
#pragma unroll
        for (int i = 0; i < NE; i++) E[i] = 0.0;
#pragma unroll
        for (int f = 0; f < KC; f++) {
#pragma unroll
            for (int i = 0; i < ND; i++) B[f][i] = 0.0;
            fc[f] = f0 + f < a.K ? fs[(int64_t)(f0 + f) * a.npad + qp] : FT(0);
        }
        constexpr int APPLY_NB = apply_nb<ORDER>();
        // field values of a batch are preloaded when they fit in registers
        constexpr bool PRE = KC * APPLY_NB <= 32;
        // packed (x, y, z, field 0): one 16-byte load per neighbour
        const bool pack = std::is_same<FT, float>::value && a.xyzf != nullptr;
        auto pair = [&](const double (&d)[3], const double (&df)[KC]) {
            double w = 1.0;
            if (a.weighted) w = kernel_w<double>(pair_dist(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * ih);
            double av[ND];
            basis<DIM, ORDER, double>(d, av);
            double wa[ND];
#pragma unroll
            for (int i = 0; i < ND; i++) wa[i] = w * av[i];
#pragma unroll
            for (int i = 0; i < ND; i++) {
#pragma unroll
                for (int jj = 0; jj <= i; jj++) E[tri(i, jj)] = fma(wa[i], av[jj], E[tri(i, jj)]);
            }
#pragma unroll
            for (int f = 0; f < KC; f++) {
                if (f0 + f < a.K) {
#pragma unroll
                    for (int i = 0; i < ND; i++) B[f][i] = fma(wa[i], df[f], B[f][i]);
                }
            }
        };
        for (int j0 = gl; j0 < a.k; j0 += GW * APPLY_NB) {
            uint32_t pj[APPLY_NB];
#pragma unroll
            for (int u = 0; u < APPLY_NB; u++) {
                const int j = j0 + u * GW;
                pj[u] = j < a.k ? a.nl_idx[(int64_t)j * a.nq + s] : qp;
            }
            float cx[APPLY_NB][3], cl[APPLY_NB][3], pw[APPLY_NB];
#pragma unroll
            for (int u = 0; u < APPLY_NB; u++) {
                if (pack) {
                    const float4 v = a.xyzf[pj[u]];
                    cx[u][0] = v.x;
                    cx[u][1] = v.y;
                    cx[u][2] = v.z;
                    pw[u] = v.w;
                } else {
                    cx[u][0] = a.sx[pj[u]];
                    cx[u][1] = a.sy[pj[u]];
                    cx[u][2] = a.sz[pj[u]];
                    pw[u] = 0.f;
                }
                cl[u][0] = a.p64 ? a.lx[pj[u]] : 0.f;
                cl[u][1] = a.p64 ? a.ly[pj[u]] : 0.f;
                cl[u][2] = a.p64 ? a.lz[pj[u]] : 0.f;
            }
            // field value f of neighbour u (field 0 from the pack when packed)
            auto fval = [&](int u, int f) -> FT {
                if (f0 + f >= a.K) return FT(0);
                if (pack && f0 + f == 0) return (FT)pw[u];
                return fs[(int64_t)(f0 + f) * a.npad + pj[u]];
            };
            FT fv[PRE ? APPLY_NB : 1][KC];
            if (PRE) {
#pragma unroll
                for (int u = 0; u < APPLY_NB; u++) {
#pragma unroll
                    for (int f = 0; f < KC; f++) fv[u][f] = fval(u, f);
                }
            }
#pragma unroll
            for (int u = 0; u < APPLY_NB; u++) {
                if (j0 + u * GW >= a.k) break;
                double d[3];
                pair_offset(a, cx[u], cl[u], qx, qy, qz, d);
                double df[KC];
#pragma unroll
                for (int f = 0; f < KC; f++) {
                    const FT v = PRE ? fv[PRE ? u : 0][f] : fval(u, f);
                    df[f] = (double)v - (double)fc[f];
                }
                pair(d, df);
            }
        }
        if (GW > 1) {
#pragma unroll
            for (int o = GW / 2; o > 0; o >>= 1) {
#pragma unroll
                for (int i = 0; i < NE; i++) E[i] += __shfl_xor(E[i], o);
#pragma unroll
                for (int f = 0; f < KC; f++) {
#pragma unroll
                    for (int i = 0; i < ND; i++) B[f][i] += __shfl_xor(B[f][i], o);
                }
            }
        }
        if (zero_diagonal<ND>(E)) st |= ST_SINGULAR;
        const bool ok = ldlt<ND>(E);
        if (!ok) st |= ST_RANKDEF;
#pragma unroll
        for (int f = 0; f < KC; f++) {
            if (f0 + f >= a.K || gl != 0) continue;
            ldlt_solve<ND>(E, B[f]);
            if (!ok) {
#pragma unroll
                for (int i = 0; i < ND; i++) B[f][i] = (double)NAN;
            }
            if (a.out1) {
                float *o = a.out1 + ((int64_t)row * a.K + f0 + f) * DIM;
#pragma unroll
                for (int i = 0; i < DIM; i++) o[i] = (float)B[f][i];
            }
            if (ORDER == 2 && a.out2) {
                float *o = a.out2 + ((int64_t)row * a.K + f0 + f) * (ND - DIM);
#pragma unroll
                for (int i = DIM; i < ND; i++) o[i - DIM] = (float)B[f][i];
            }
        }
    }
    if (a.status && gl == 0) a.status[row] = a.or_status ? (a.status[row] | st) : st;
}

// apply_row for the common case of one float32 field, 3D, first order,
// float32 positions and the packed (x, y, z, f) copy (a.K == 1, a.xyzf set,
// a.k >= 1): the same float64 terms as the generic loop (at the same lane
// count the results are bit-identical; with four lanes here against two there
// the partial sums are added in another order), with the loads
// software-pipelined.  The generic loop waits on a neighbour's
// index and then on its position before it can sum it, so every neighbour
// costs two memory round trips with nothing else in flight in the wavefront;
// here the position of the next neighbour and the index of the one after it
// are requested before the current neighbour is summed.  Measured at 2^24
// clustered, k = 32: 1.97 ms for the generic loop (two lanes per query);
// 2.06, 1.85, 1.78 and 2.18 ms with 1, 2, 4 and 8 lanes; prefetching two
// positions ahead instead of one, 1.83 ms (4 lanes).
template <int GW>
MHIP_DEV void apply_row_packed1(const ApplyArgs &a, int64_t s, int gl) {
    const uint32_t qp = a.qsorted ? a.qsorted[s] : (uint32_t)s;
    const uint32_t row = a.qrow ? a.qrow[s] : a.perm[s];
    const float4 qv = a.xyzf[qp];
    const double qx = (double)qv.x, qy = (double)qv.y, qz = (double)qv.z;
    const double fc = (double)qv.w;
    const double hd = a.h[s];
    const double ih = hd > 0.0 ? 1.0 / hd : (double)INFINITY;
    const int k = a.k;
    const int64_t nq = a.nq;
    const uint32_t *__restrict__ ip = a.nl_idx + s;  // entry j at ip[j * nq]
    double E[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double B[3] = {0.0, 0.0, 0.0};
    // list pointers advance by GW entries and stop at the row's last entry,
    // which keeps every prefetch inside the row
    const int64_t step = (int64_t)GW * nq;
    const uint32_t *const plast = ip + (int64_t)(k - 1) * nq;
    auto advance = [&](const uint32_t *p) { return p + step > plast ? plast : p + step; };
    const uint32_t *pnext = ip + (int64_t)min(gl, k - 1) * nq;
    float4 vn = a.xyzf[*pnext];
    pnext = advance(pnext);
    uint32_t pn = *pnext;
    // two neighbours per trip: the prefetch registers alternate instead of
    // being copied (6 waves per SIMD instead of 8, measured 1.76 -> 1.72 ms)
#pragma unroll 2
    for (int j = gl; j < k; j += GW) {
        const float4 v = vn;
        vn = a.xyzf[pn];
        pnext = advance(pnext);
        pn = *pnext;
        double d[3] = {(double)v.x - qx, (double)v.y - qy, (double)v.z - qz};
        if (a.per) {
            const double box = (double)a.box;
#pragma unroll
            for (int c = 0; c < 3; c++) d[c] = wrap<double>(d[c], box);
        }
        double w = 1.0;
        if (a.weighted) {
            // cubic spline with 1 - q clamped at 0 (zero past q = 1)
            const double q = pair_dist(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * ih;
            const double om = fmax(1.0 - q, 0.0);
            w = q <= 0.5 ? 1.0 + 6.0 * q * q * (q - 1.0) : 2.0 * om * om * om;
        }
        double wa[3];
#pragma unroll
        for (int i = 0; i < 3; i++) wa[i] = w * d[i];
#pragma unroll
        for (int i = 0; i < 3; i++) {
#pragma unroll
            for (int jj = 0; jj <= i; jj++) E[tri(i, jj)] = fma(wa[i], d[jj], E[tri(i, jj)]);
        }
        const double df = (double)v.w - fc;
#pragma unroll
        for (int i = 0; i < 3; i++) B[i] = fma(wa[i], df, B[i]);
    }
    if (GW > 1) {
#pragma unroll
        for (int o = GW / 2; o > 0; o >>= 1) {
#pragma unroll
            for (int i = 0; i < 6; i++) E[i] += __shfl_xor(E[i], o);
#pragma unroll
            for (int i = 0; i < 3; i++) B[i] += __shfl_xor(B[i], o);
        }
    }
    if (gl != 0) return;
    int32_t st = 0;
    if (zero_diagonal<3>(E)) st |= ST_SINGULAR;
    const bool ok = ldlt<3>(E);
    if (!ok) st |= ST_RANKDEF;
    ldlt_solve<3>(E, B);
    if (a.out1) {
        float *o = a.out1 + (int64_t)row * 3;
#pragma unroll
        for (int i = 0; i < 3; i++) o[i] = ok ? (float)B[i] : NAN;
    }
    if (a.status) a.status[row] = a.or_status ? (a.status[row] | st) : st;
}

// lanes per query: 2 for first-order fits (measured at 2^24, k = 32, with the
// packed (x, y, z, f) loads: 1.93 ms against 2.04 with 1 lane, 2.05 with 4
// and 2.74 with 8), 1 for the 9x9 second-order fits, whose 63 float64 sums
// would cost more to combine than they gain (9.5 vs 7.0 ms)
template <int ORDER>
constexpr int apply_gw() { return ORDER == 1 ? 2 : 1; }

template <int DIM, int ORDER, typename FT, int KC>
__global__ __launch_bounds__(256) void k_apply(ApplyArgs a) {
    constexpr int GW = apply_gw<ORDER>();
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t t = (int64_t)b * blockDim.x + threadIdx.x;
    const int64_t s = t / GW;
    if (s >= a.nq) return;
    apply_row<DIM, ORDER, FT, KC, GW>(a, s, 0, a.K, (int)(t % GW));
}

// apply_row_packed1 over the list (GW lanes per query)
template <int GW>
__global__ __launch_bounds__(256) void k_apply_packed1(ApplyArgs a) {
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t t = (int64_t)b * blockDim.x + threadIdx.x;
    const int64_t s = t / GW;
    if (s >= a.nq) return;
    apply_row_packed1<GW>(a, s, (int)(t % GW));
}

// every snapshot of a batch in one launch: blk0[i] is the first query block
// of snapshot i (blk0[nsnap] = query blocks); each block finds its
// snapshot's arguments in the device table.  The fields are split into ntile
// tiles of KC: the threads of a tile recompute the moment matrix, so tiles
// run in parallel instead of as passes of one thread (their status bits are
// identical).
template <int KC>
__global__ __launch_bounds__(256) void k_apply_batch(const ApplyArgs *__restrict__ tab,
                                                     const int64_t *__restrict__ blk0, int nsnap, int ntile) {
    // the tiles of one block of queries are consecutive logical blocks, so
    // they run together on one XCD and share its L2 for the neighbour
    // indices, positions and nearby field values
    const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t b = lb / (uint32_t)ntile;
    const int tile = (int)(lb % (uint32_t)ntile);
    int lo = 0, hi = nsnap;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (blk0[mid] <= b) lo = mid;
        else hi = mid;
    }
    const ApplyArgs a = tab[lo];
    const int64_t s = (b - blk0[lo]) * blockDim.x + threadIdx.x;
    if (s >= a.nq) return;
    const int fb = tile * KC;
    if (fb >= a.K) return;
    apply_row<3, 1, float, KC>(a, s, fb, min(fb + KC, a.K));
}

// Many fields at once (batched form, order 1, 3D, float32 positions and
// row-major fields [npad][K]): a block holds G = 256 / KW queries with KW
// threads each.  Phase 1 spreads a query's neighbours over its KW threads:
// each computes w_j a_j, the moment matrix is summed over the group with
// shuffles, every thread factorises it, and the owner of neighbour j stores
// g_j = E^-1 w_j a_j in LDS.  Phase 2 gives each thread fields f, f + KW, ...:
// out[f] = sum_j g_j (f_j - f_i) in float64, where the group's KW threads read
// one neighbour's KW consecutive field values (one cache line) together.
// Same operator as apply_row, summed in another order.
constexpr int WIDE_KMAX = 128;
// queries per block with KW threads each: 256 threads, or 16 queries of 8
constexpr int wide_g(int KW) { return KW >= 16 ? 256 / KW : 16; }

template <int KW>
struct WideQuery {
    int64_t s, sc;
    bool live;
    uint32_t qp, row;
    int32_t st;
    bool ok;
};

// the block's snapshot (binary search of the block prefix) and phase 1
template <int KW>
MHIP_DEV WideQuery<KW> wide_phase1(const ApplyArgs *__restrict__ tab, const int64_t *__restrict__ blk0, int nsnap,
                                   double *s_g, uint32_t *s_p, int kmax, const ApplyArgs *&ap) {
    constexpr int G = wide_g(KW);
    const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
    int lo = 0, hi = nsnap;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (blk0[mid] <= b) lo = mid;
        else hi = mid;
    }
    ap = tab + lo;
    const ApplyArgs &a = *ap;
    const int g = threadIdx.x / KW, f = threadIdx.x % KW;
    WideQuery<KW> w;
    w.s = (b - blk0[lo]) * G + g;
    w.live = w.s < a.nq;
    w.sc = w.live ? w.s : a.nq - 1;  // padding queries repeat the last one
    w.qp = a.qsorted ? a.qsorted[w.sc] : (uint32_t)w.sc;
    w.row = a.qrow ? a.qrow[w.sc] : a.perm[w.sc];
    const double qx = (double)a.sx[w.qp], qy = (double)a.sy[w.qp], qz = (double)a.sz[w.qp];
    const double hd = a.h[w.sc];
    const double ih = hd > 0.0 ? 1.0 / hd : (double)INFINITY;
    const int k = a.k;
    double E[6] = {0, 0, 0, 0, 0, 0};
    for (int j = f; j < k; j += KW) {
        const uint32_t p = a.nl_idx[(int64_t)j * a.nq + w.sc];
        const double d[3] = {(double)a.sx[p] - qx, (double)a.sy[p] - qy, (double)a.sz[p] - qz};
        const double wt = kernel_w<double>(pair_dist(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * ih);
        double wa[3];
#pragma unroll
        for (int i = 0; i < 3; i++) wa[i] = wt * d[i];
#pragma unroll
        for (int i = 0; i < 3; i++) {
#pragma unroll
            for (int jj = 0; jj <= i; jj++) E[tri(i, jj)] = fma(wa[i], d[jj], E[tri(i, jj)]);
        }
#pragma unroll
        for (int i = 0; i < 3; i++) s_g[((int64_t)g * kmax + j) * 3 + i] = wa[i];
        s_p[g * kmax + j] = p;
    }
#pragma unroll
    for (int o = KW / 2; o > 0; o >>= 1) {
#pragma unroll
        for (int e = 0; e < 6; e++) E[e] += __shfl_xor(E[e], o);
    }
    w.st = 0;
    if (zero_diagonal<3>(E)) w.st |= ST_SINGULAR;
    w.ok = ldlt<3>(E);
    if (!w.ok) w.st |= ST_RANKDEF;
    __syncthreads();
    for (int j = f; j < k; j += KW) {
        double *gj = s_g + ((int64_t)g * kmax + j) * 3;
        double v[3] = {gj[0], gj[1], gj[2]};
        ldlt_solve<3>(E, v);
#pragma unroll
        for (int i = 0; i < 3; i++) gj[i] = v[i];
    }
    __syncthreads();
    return w;
}

// Phase 2 with FPL fields per thread: thread f of a query sums FPL
// consecutive fields of each pass (one 4/8/16-byte load per neighbour when K
// is a multiple of FPL, else FPL single loads), so the neighbour's index and
// float64 weights are read from LDS once for FPL fields.  Every output is the
// same sequential float64 sum over j whatever KW and FPL are.  The LDS holds
// kmax neighbours per query (dynamic: the largest des_ngb of the batch).
template <int FPL>
struct FieldVec;
template <>
struct FieldVec<1> {
    typedef float T;
    static MHIP_DEV float at(const T &v, int) { return v; }
};
template <>
struct FieldVec<2> {
    typedef float2 T;
    static MHIP_DEV float at(const T &v, int i) { return i == 0 ? v.x : v.y; }
};
template <>
struct FieldVec<4> {
    typedef float4 T;
    static MHIP_DEV float at(const T &v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)); }
};

template <int FPL, bool VEC>
MHIP_DEV void wide_pass(const float *__restrict__ fs, int K, const double *sg, const uint32_t *sp, int k,
                        uint32_t qp, int c, double (&o)[FPL][3]) {
    // c: this thread's first field of the pass (c + FPL <= K when VEC)
    double fi[FPL];
    int col[FPL];
#pragma unroll
    for (int i = 0; i < FPL; i++) {
        col[i] = min(c + i, K - 1);
        fi[i] = (double)fs[(int64_t)qp * K + col[i]];
        o[i][0] = o[i][1] = o[i][2] = 0.0;
    }
    // unrolled so that several neighbours' field loads are in flight
    constexpr int UNR = 8 / FPL;
#pragma unroll UNR
    for (int j = 0; j < k; j++) {
        const float *row = fs + (int64_t)sp[j] * K;
        const double g0 = sg[3 * j], g1 = sg[3 * j + 1], g2 = sg[3 * j + 2];
        float v[FPL];
        if (VEC) {
            const typename FieldVec<FPL>::T x = *reinterpret_cast<const typename FieldVec<FPL>::T *>(row + c);
#pragma unroll
            for (int i = 0; i < FPL; i++) v[i] = FieldVec<FPL>::at(x, i);
        } else {
#pragma unroll
            for (int i = 0; i < FPL; i++) v[i] = row[col[i]];
        }
#pragma unroll
        for (int i = 0; i < FPL; i++) {
            const double df = (double)v[i] - fi[i];
            o[i][0] = fma(g0, df, o[i][0]);
            o[i][1] = fma(g1, df, o[i][1]);
            o[i][2] = fma(g2, df, o[i][2]);
        }
    }
}

template <int KW, int FPL>
__global__ __launch_bounds__(wide_g(KW) * KW) void k_apply_wide(const ApplyArgs *__restrict__ tab,
                                                                const int64_t *__restrict__ blk0, int nsnap,
                                                                int kmax) {
    constexpr int G = wide_g(KW);
    extern __shared__ double s_dyn[];
    double *s_g = s_dyn;
    uint32_t *s_p = reinterpret_cast<uint32_t *>(s_dyn + (size_t)G * kmax * 3);
    const ApplyArgs *ap;
    const WideQuery<KW> w = wide_phase1<KW>(tab, blk0, nsnap, s_g, s_p, kmax, ap);
    const ApplyArgs &a = *ap;
    const int g = threadIdx.x / KW, f = threadIdx.x % KW;
    const int k = a.k, K = a.K;
    const float *fs = (const float *)a.fs;
    const double *sg = s_g + (int64_t)g * kmax * 3;
    const uint32_t *sp = s_p + g * kmax;
    // vector loads need whole FPL-field groups (rows then stay aligned too)
    const bool vec = FPL > 1 && K % FPL == 0;
    for (int c0 = 0; c0 < K; c0 += KW * FPL) {
        const int c = c0 + f * FPL;
        double o[FPL][3];
        if (vec)
            wide_pass<FPL, true>(fs, K, sg, sp, k, w.qp, min(c, K - FPL), o);
        else
            wide_pass<FPL, false>(fs, K, sg, sp, k, w.qp, c, o);
#pragma unroll
        for (int i = 0; i < FPL; i++) {
            if (!w.live || c + i >= K) continue;
            if (!w.ok) o[i][0] = o[i][1] = o[i][2] = (double)NAN;
            float *out = a.out1 + ((int64_t)w.row * K + c + i) * 3;
            out[0] = (float)o[i][0];
            out[1] = (float)o[i][1];
            out[2] = (float)o[i][2];
        }
    }
    if (w.live && f == 0 && a.status) a.status[w.row] = a.or_status ? (a.status[w.row] | w.st) : w.st;
}

// The same operator with phase 2 on the matrix cores (K == KW fields, KW in
// {16, 32, 64}): v_mfma_f32_4x4x1_16b_f32 computes 16 independent 4 x 4
// blocks C_b += A_b B_b with A_b 4 x 1 and B_b 1 x 4.  Block b of a wavefront
// is (query q = b / (KW / 4) of the wavefront's 64 / KW queries, field group
// fg = b % (KW / 4)); its rows are the gradient components (g_j, row 3 zero)
// and its columns 4 fields (f_j - f_i), one neighbour j per instruction:
// C_b[c][t] = sum_j g_j[c] (f_j - f_i)[4 fg + t].  Lane l supplies A_b[l % 4]
// and B_b[l % 4] of block l / 4 and holds column l % 4 of C_b (rows in the 4
// result registers).  float32 products, float32 accumulation in four
// partial sums (neighbours j mod 4), added at the end.
template <int KW>
__global__ __launch_bounds__(256) void k_apply_mfma(const ApplyArgs *__restrict__ tab,
                                                    const int64_t *__restrict__ blk0, int nsnap) {
    constexpr int G = 256 / KW;
    constexpr int NGRP = KW / 4;  // field groups per query
    __shared__ double s_g[G][WIDE_KMAX][3];
    __shared__ uint32_t s_p[G][WIDE_KMAX];
    const ApplyArgs *ap;
    const WideQuery<KW> w = wide_phase1<KW>(tab, blk0, nsnap, &s_g[0][0][0], &s_p[0][0], WIDE_KMAX, ap);
    const ApplyArgs &a = *ap;
    const int g = threadIdx.x / KW, f = threadIdx.x % KW;
    // this lane's block: its query is g (a wavefront holds 64 / KW whole
    // queries, as in phase 1), its field group and column
    const int wl = threadIdx.x & 63;
    const int blk = wl >> 2, t = wl & 3;
    const int gq = (threadIdx.x >> 6) * (64 / KW) + blk / NGRP;  // == g
    const int field = 4 * (blk % NGRP) + t;
    const int k = a.k;
    const float *fs = (const float *)a.fs;
    // the query of this lane's block is the lane's own query (g == gq), so its
    // own position and row come from phase 1
    const float fi = fs[(int64_t)w.qp * a.K + field];
    typedef float v4f __attribute__((ext_vector_type(4)));
    // four accumulators (neighbours j mod 4): consecutive matrix instructions
    // do not wait on each other's results
    v4f acc[4];
#pragma unroll
    for (int u = 0; u < 4; u++) acc[u] = v4f{0.f, 0.f, 0.f, 0.f};
    int j = 0;
    for (; j + 4 <= k; j += 4) {
        float av[4], bv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            av[u] = t < 3 ? (float)s_g[gq][j + u][t] : 0.f;
            bv[u] = fs[(int64_t)s_p[gq][j + u] * a.K + field] - fi;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) acc[u] = __builtin_amdgcn_mfma_f32_4x4x1f32(av[u], bv[u], acc[u], 0, 0, 0);
    }
    for (; j < k; j++) {
        const float av = t < 3 ? (float)s_g[gq][j][t] : 0.f;
        const float bv = fs[(int64_t)s_p[gq][j] * a.K + field] - fi;
        acc[0] = __builtin_amdgcn_mfma_f32_4x4x1f32(av, bv, acc[0], 0, 0, 0);
    }
    const v4f res = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    if (w.live) {
        float *o = a.out1 + ((int64_t)w.row * a.K + field) * 3;
        o[0] = w.ok ? res[0] : NAN;
        o[1] = w.ok ? res[1] : NAN;
        o[2] = w.ok ? res[2] : NAN;
    }
    if (w.live && f == 0 && a.status) a.status[w.row] = a.or_status ? (a.status[w.row] | w.st) : w.st;
}

int apply_mfma_launch(mhip_ctx *ctx, const ApplyArgs *tab_d, const int64_t *blk0_d, int nsnap,
                      int64_t nblocks, int kw) {
    const dim3 g((unsigned)nblocks), b(256);
    switch (kw) {
        case 16: hipLaunchKernelGGL(k_apply_mfma<16>, g, b, 0, ctx->stream, tab_d, blk0_d, nsnap); break;
        case 32: hipLaunchKernelGGL(k_apply_mfma<32>, g, b, 0, ctx->stream, tab_d, blk0_d, nsnap); break;
        case 64: hipLaunchKernelGGL(k_apply_mfma<64>, g, b, 0, ctx->stream, tab_d, blk0_d, nsnap); break;
        default: return fail(ctx, -1, "unsupported field group");
    }
    return check_hip(ctx, hipGetLastError(), "mfma apply launch");
}

int apply_wide_launch(mhip_ctx *ctx, const ApplyArgs *tab_d, const int64_t *blk0_d, int nsnap,
                      int64_t nblocks, int kw, int K, int kmax) {
    if (kmax < 1 || kmax > WIDE_KMAX) return fail(ctx, -1, "des_ngb outside the wide kernel's range");
    const dim3 g((unsigned)nblocks);
    // fields per thread: enough for K (at most 32 per pass of the group)
    const int fpl = std::min(32 / kw, std::max(1, (K + kw - 1) / kw));
    const size_t lds = (size_t)wide_g(kw) * kmax * (3 * sizeof(double) + sizeof(uint32_t));
#define MHIP_WIDE(KW_, FPL_)                                                                                \
    hipLaunchKernelGGL((k_apply_wide<KW_, FPL_>), g, dim3(wide_g(KW_) * KW_), lds, ctx->stream, tab_d, blk0_d, \
                       nsnap, kmax)
    if (kw == 8 && fpl >= 3) MHIP_WIDE(8, 4);
    else if (kw == 8 && fpl == 2) MHIP_WIDE(8, 2);
    else if (kw == 8) MHIP_WIDE(8, 1);
    else if (kw == 16 && fpl >= 2) MHIP_WIDE(16, 2);
    else if (kw == 16) MHIP_WIDE(16, 1);
    else if (kw == 32) MHIP_WIDE(32, 1);
    else return fail(ctx, -1, "unsupported field group");
#undef MHIP_WIDE
    return check_hip(ctx, hipGetLastError(), "wide apply launch");
}

int apply_batch_launch(mhip_ctx *ctx, const ApplyArgs *tab_d, const int64_t *blk0_d, int nsnap,
                       int64_t nblocks, int kc, int K) {
    const int nt = (K + kc - 1) / kc;
    if (nblocks * nt >= ((int64_t)1 << 31)) return fail(ctx, -4, "batch too large for one launch");
    const dim3 g((unsigned)(nblocks * nt)), b(256);
    switch (kc) {
        case 2: hipLaunchKernelGGL(k_apply_batch<2>, g, b, 0, ctx->stream, tab_d, blk0_d, nsnap, nt); break;
        case 4: hipLaunchKernelGGL(k_apply_batch<4>, g, b, 0, ctx->stream, tab_d, blk0_d, nsnap, nt); break;
        case 8: hipLaunchKernelGGL(k_apply_batch<8>, g, b, 0, ctx->stream, tab_d, blk0_d, nsnap, nt); break;
        case 16: hipLaunchKernelGGL(k_apply_batch<16>, g, b, 0, ctx->stream, tab_d, blk0_d, nsnap, nt); break;
        default: return fail(ctx, -1, "unsupported field tile");
    }
    return check_hip(ctx, hipGetLastError(), "batched apply launch");
}

void apply_args_fill(const mhip_nlist *nl, const void *fs, int K, float *out1, int32_t *status,
                     int or_status, ApplyArgs *a) {
    const mhip_index *ix = nl->index;
    a->sx = ix->sx;
    a->sy = ix->sy;
    a->sz = ix->sz;
    a->lx = ix->lx;
    a->ly = ix->ly;
    a->lz = ix->lz;
    a->per = ix->box > 0;
    a->box = a->per ? (float)ix->box : 0.f;
    a->inv_box = a->per ? (float)(1.0 / ix->box) : 0.f;
    a->p64 = ix->pos_f64;
    a->nl_idx = nl->idx;
    a->qsorted = nl->masked ? nl->qsorted : nullptr;
    a->qrow = nl->masked ? nl->qrow : nullptr;
    a->perm = ix->perm;
    a->h = nl->h64;
    a->nq = nl->nq;
    a->npad = ix->npad;
    a->k = nl->k;
    a->fs = fs;
    a->xyzf = nullptr;
    a->K = K;
    a->weighted = 1;
    a->out1 = out1;
    a->out2 = nullptr;
    a->status = status;
    a->or_status = or_status;
}

template <int DIM, int ORDER, typename FT>
static void launch_apply(const ApplyArgs &a, dim3 g, hipStream_t s) {
    // one field: no unused right-hand sides in the registers and shuffles
    if (a.K == 1)
        hipLaunchKernelGGL((k_apply<DIM, ORDER, FT, 1>), g, dim3(256), 0, s, a);
    else if (ORDER == 1)
        hipLaunchKernelGGL((k_apply<DIM, ORDER, FT, 4>), g, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_apply<DIM, ORDER, FT, 2>), g, dim3(256), 0, s, a);
}

template <typename FT>
static void dispatch_apply(const ApplyArgs &a, int dim, int order, dim3 g, hipStream_t s) {
    if (order == 1) {
        if (dim == 1) launch_apply<1, 1, FT>(a, g, s);
        if (dim == 2) launch_apply<2, 1, FT>(a, g, s);
        if (dim == 3) launch_apply<3, 1, FT>(a, g, s);
    } else {
        if (dim == 1) launch_apply<1, 2, FT>(a, g, s);
        if (dim == 2) launch_apply<2, 2, FT>(a, g, s);
        if (dim == 3) launch_apply<3, 2, FT>(a, g, s);
    }
}

int apply_ops(mhip_ctx *ctx, const mhip_nlist *nl, const void *f, int f_f64, int64_t nf_rows, int K,
              int order, int weighted, float *out1, float *out2, int32_t *status) {
    const mhip_index *ix = nl->index;
    if (!ix || nl->nq < 1) return fail(ctx, -1, "neighbour list is empty");
    if (order != 1 && order != 2) return fail(ctx, -1, "order must be 1 or 2");
    if (K < 1) return fail(ctx, -1, "no field components");
    int rc = 0;
    const size_t fb = f_f64 ? 8 : 4;
    // float32 positions and fields: a packed (x, y, z, field 0) copy beside
    // the sorted-space fields, so a neighbour is one 16-byte load
    const bool pack = !f_f64 && !ix->pos_f64 && ix->dim == 3;
    // one float32 field, first order: the pipelined kernel reads only the pack
    const bool packed1 = pack && K == 1 && order == 1 && nl->k >= 1;
    const size_t fs_bytes = packed1 ? 0 : ((size_t)K * ix->npad * fb + 255) & ~(size_t)255;
    char *sc = (char *)scratch(ctx, fs_bytes + (pack ? (size_t)ix->npad * 16 : 0), &rc);
    if (!sc) return rc;
    void *fs = sc;
    float4 *xyzf = pack ? (float4 *)(sc + fs_bytes) : nullptr;
    const dim3 gp((unsigned)((ix->npad + 255) / 256));
    ProfRegion prp(ctx, "apply_permute_fields");
    if (packed1)
        hipLaunchKernelGGL(k_pack_xyzf_rows, gp, dim3(256), 0, ctx->stream, ix->sx, ix->sy, ix->sz,
                           (const float *)f, nf_rows, ix->perm, ix->n, ix->npad, xyzf);
    else if (f_f64)
        hipLaunchKernelGGL(k_permute_fields<double>, gp, dim3(256), 0, ctx->stream, (const double *)f,
                           nf_rows, K, ix->perm, ix->n, ix->npad, (double *)fs);
    else
        hipLaunchKernelGGL(k_permute_fields<float>, gp, dim3(256), 0, ctx->stream, (const float *)f,
                           nf_rows, K, ix->perm, ix->n, ix->npad, (float *)fs);
    if (pack && !packed1)
        hipLaunchKernelGGL(k_pack_xyzf, gp, dim3(256), 0, ctx->stream, ix->sx, ix->sy, ix->sz, (const float *)fs,
                           ix->npad, xyzf);
    ApplyArgs a;
    a.xyzf = xyzf;
    a.sx = ix->sx;
    a.sy = ix->sy;
    a.sz = ix->sz;
    a.lx = ix->lx;
    a.ly = ix->ly;
    a.lz = ix->lz;
    a.per = ix->box > 0;
    a.box = a.per ? (float)ix->box : 0.f;
    a.inv_box = a.per ? (float)(1.0 / ix->box) : 0.f;
    a.p64 = ix->pos_f64;
    a.nl_idx = nl->idx;
    a.qsorted = nl->masked ? nl->qsorted : nullptr;
    a.qrow = nl->masked ? nl->qrow : nullptr;
    a.perm = ix->perm;
    a.h = nl->h64;
    a.nq = nl->nq;
    a.npad = ix->npad;
    a.k = nl->k;
    a.fs = fs;
    a.K = K;
    a.weighted = weighted;
    a.out1 = out1;
    a.out2 = out2;
    a.status = status;
    a.or_status = 0;
    prp.end();
    ProfRegion pra(ctx, order == 1 ? "apply_D1" : "apply_D2");
    if (packed1) {
        // the pipelined packed-load kernel, four lanes per query
        constexpr int gwp = 4;
        const dim3 g((unsigned)((nl->nq * gwp + 255) / 256));
        hipLaunchKernelGGL(k_apply_packed1<gwp>, g, dim3(256), 0, ctx->stream, a);
        return check_hip(ctx, hipGetLastError(), "apply launch");
    }
    const int gw = order == 1 ? apply_gw<1>() : apply_gw<2>();
    const dim3 g((unsigned)((nl->nq * gw + 255) / 256));
    if (f_f64)
        dispatch_apply<double>(a, ix->dim, order, g, ctx->stream);
    else
        dispatch_apply<float>(a, ix->dim, order, g, ctx->stream);
    return check_hip(ctx, hipGetLastError(), "apply launch");
}

// ---------------------------------------------------------------------------
// materialised weights G (nq, ND, k) float64 for arbitrary row-major neighbour
// lists, gradient_weights (derivatives.py:117-138) signature
// ---------------------------------------------------------------------------
template <int DIM, int ORDER, typename PT>
__global__ __launch_bounds__(128) void k_weights_rows(const PT *__restrict__ pos, const int32_t *__restrict__ ngb,
                                                     int64_t nq, int k, const double *__restrict__ h,
                                                     const int64_t *__restrict__ indices, double box,
                                                     int weighted, double *__restrict__ G,
                                                     int32_t *__restrict__ status) {
    constexpr int ND = NDer<DIM, ORDER>::ND, NE = NDer<DIM, ORDER>::NE;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const int64_t c = indices[i];
    double xc[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < DIM; a++) xc[a] = (double)pos[c * DIM + a];
    const double ih = 1.0 / h[i];
    double E[NE];
#pragma unroll
    for (int e = 0; e < NE; e++) E[e] = 0.0;
    auto pair = [&](int j, double (&av)[ND], double &w) {
        const int64_t nb = ngb[i * k + j];
        double d[3] = {0, 0, 0};
        double r2 = 0.0;
#pragma unroll
        for (int a = 0; a < DIM; a++) {
            double v = (double)pos[nb * DIM + a] - xc[a];
            if (box > 0) v = wrap<double>(v, box);
            d[a] = v;
            r2 += v * v;
        }
        w = weighted ? kernel_w<double>(sqrt(r2) * ih) : 1.0;
        basis<DIM, ORDER, double>(d, av);
    };
    for (int j = 0; j < k; j++) {
        double av[ND], w;
        pair(j, av, w);
#pragma unroll
        for (int a = 0; a < ND; a++) {
#pragma unroll
            for (int b2 = 0; b2 <= a; b2++) E[tri(a, b2)] = fma(w * av[a], av[b2], E[tri(a, b2)]);
        }
    }
    const bool zero = zero_diagonal<ND>(E);
    const bool ok = ldlt<ND>(E);
    if (status) status[i] = (zero ? ST_SINGULAR : 0) | (ok ? 0 : ST_RANKDEF);
    double *out = G + i * (int64_t)ND * k;
    for (int j = 0; j < k; j++) {
        double av[ND], w;
        pair(j, av, w);
#pragma unroll
        for (int a = 0; a < ND; a++) av[a] *= w;
        ldlt_solve<ND>(E, av);
#pragma unroll
        for (int a = 0; a < ND; a++) out[(int64_t)a * k + j] = ok ? av[a] : NAN;
    }
}

template <typename PT>
static void dispatch_weights(const PT *pos, int dim, int order, const int32_t *ngb, int64_t nq, int k,
                             const double *h, const int64_t *indices, double box, int weighted,
                             double *G, int32_t *status, hipStream_t s) {
    const dim3 g((unsigned)((nq + 127) / 128)), b(128);
#define MHIP_W(D, O) \
    hipLaunchKernelGGL((k_weights_rows<D, O, PT>), g, b, 0, s, pos, ngb, nq, k, h, indices, box, weighted, G, status)
    if (order == 1) {
        if (dim == 1) MHIP_W(1, 1);
        if (dim == 2) MHIP_W(2, 1);
        if (dim == 3) MHIP_W(3, 1);
    } else {
        if (dim == 1) MHIP_W(1, 2);
        if (dim == 2) MHIP_W(2, 2);
        if (dim == 3) MHIP_W(3, 2);
    }
#undef MHIP_W
}

int weights_rows(mhip_ctx *ctx, const void *pos, int pos_f64, int64_t n, int dim, const int32_t *ngb,
                 int64_t nq, int k, const double *h, const int64_t *indices, double box, int weighted,
                 int order, double *G, int32_t *status) {
    (void)n;
    if (dim < 1 || dim > 3) return fail(ctx, -5, "dim must be 1, 2 or 3");
    if (order != 1 && order != 2) return fail(ctx, -1, "order must be 1 or 2");
    if (nq < 1 || k < 1) return 0;
    if (pos_f64)
        dispatch_weights<double>((const double *)pos, dim, order, ngb, nq, k, h, indices, box, weighted,
                                 G, status, ctx->stream);
    else
        dispatch_weights<float>((const float *)pos, dim, order, ngb, nq, k, h, indices, box, weighted, G,
                                status, ctx->stream);
    return check_hip(ctx, hipGetLastError(), "weights launch");
}

// ---------------------------------------------------------------------------
// HsmlIter on a given distance array (kernel_density.py:14-57): float64,
// bracket [d_1, c_d d_{k-1}], stop when |n - k| <= error_norm.  Safeguarded
// Newton replaces the plain bisection; the root is the same.
// ---------------------------------------------------------------------------
template <typename DT>
__global__ __launch_bounds__(128) void k_hsml_dists(const DT *__restrict__ d, int64_t n, int k, int dim,
                                                   double error_norm, double *__restrict__ hout,
                                                   int32_t *__restrict__ status) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DT *row = d + i * k;
    const double norm = hsml_norm(dim);
    double lo = (double)row[1], hi = (double)row[k - 1] * hsml_bound_coeff(dim);
    double h = 0.5 * (lo + hi);
    int32_t st = 0;
    int it = 0;
    for (; it < 200; it++) {
        const double ih = 1.0 / h;
        double s = 0.0, ds = 0.0;
        for (int j = 0; j < k; j++) {
            const double q = (double)row[j] * ih;
            double w, dw;
            kernel_w_dw<double>(q, w, dw);
            s += w;
            if (q <= 1.0) ds += dw * q;
        }
        const double f = norm * s - (double)k;
        if (fabs(f) <= error_norm) break;
        if (f > 0)
            hi = h;
        else
            lo = h;
        double hn = h - f / (-norm * ds * ih);
        if (!(hn > lo && hn < hi)) hn = 0.5 * (lo + hi);
        if (hn == h) break;
        h = hn;
    }
    if (it >= 200) st |= ST_HSML_NOCONV;
    hout[i] = h;
    if (status) status[i] = st;
}

int hsml_dists(mhip_ctx *ctx, const void *d, int d_f64, int64_t n, int k, int dim, double error_norm,
               double *h, int32_t *status) {
    if (k < 2) return fail(ctx, -1, "need at least 2 neighbour distances per row");
    if (dim < 1 || dim > 3) return fail(ctx, -5, "dim must be 1, 2 or 3");
    if (n < 1) return 0;
    const dim3 g((unsigned)((n + 127) / 128)), b(128);
    if (d_f64)
        hipLaunchKernelGGL(k_hsml_dists<double>, g, b, 0, ctx->stream, (const double *)d, n, k, dim,
                           error_norm, h, status);
    else
        hipLaunchKernelGGL(k_hsml_dists<float>, g, b, 0, ctx->stream, (const float *)d, n, k, dim,
                           error_norm, h, status);
    return check_hip(ctx, hipGetLastError(), "hsml launch");
}

// Kernel(q) ufunc (kernel_density.py:82-96)
template <typename T>
__global__ void k_kernel_eval(const T *__restrict__ q, int64_t n, T *__restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = kernel_w<T>(q[i]);
}

int kernel_eval(mhip_ctx *ctx, const void *q, int f64, int64_t n, void *out) {
    if (n < 1) return 0;
    const dim3 g((unsigned)((n + 255) / 256)), b(256);
    if (f64)
        hipLaunchKernelGGL(k_kernel_eval<double>, g, b, 0, ctx->stream, (const double *)q, n, (double *)out);
    else
        hipLaunchKernelGGL(k_kernel_eval<float>, g, b, 0, ctx->stream, (const float *)q, n, (float *)out);
    return check_hip(ctx, hipGetLastError(), "kernel launch");
}

}  // namespace mhip
