// Emission-absorption radiative transfer along z through the particles.
//
// Replaces
//   radtransfer_singlethread      src/meshoid/radiation/radtransfer.py:79-177
//   radtransfer (slab split/sum)  src/meshoid/radiation/radtransfer.py:11-76
//
// Sightlines: a res x res grid of points center + (-L/2 + g L/(res-1)) in x
// and y.  The particles are taken farthest first (largest z first, as
// radtransfer_singlethread orders them); particle i changes every pixel of
// its footprint (r <= max(h, dx), kernel2d weight) in every band b by
//   tau = kappa_ib kernel2d(r/h) m_i / h^2,
//   I <- exp(-tau) I + (-expm1(-tau)) j_ib / (4 pi kappa_ib + 1e-100),
//   tau_total += tau.
// Each step is affine in I, so the ordered product splits exactly into
// segments of consecutive particles: a segment computed from I = 0 gives
// (B_s, tau_s) and acts as I <- exp(-tau_s) I + B_s.  The image is cut into
// tiles; every (particle, tile it touches) pair is a record, the records are
// stably sorted by tile (so each tile's records keep the far-first order),
// and the sorted records are cut into chunks of equal length, one wavefront
// per chunk.  A wavefront holds one tile in LDS at a time and applies the
// chunk's records of that tile in order (one lane per pixel of a footprint,
// so its LDS updates stay in program order), writing one segment per (chunk,
// tile).  A tile's particles thus spread over as many wavefronts as their
// number asks for, however clustered they are.  A last kernel composes each
// pixel's segments farthest first from the background i0.
//
// The reference's multi-threaded driver composes its chunks with the largest
// z nearest the observer while each chunk is swept with the largest z
// farthest, and adds i0 once per chunk, so its result depends on the thread
// count; this follows the single-chunk order throughout.
#include "mhip_internal.hpp"

namespace mhip {

constexpr double FOUR_PI = 12.566370614359172;
constexpr int64_t RT_CHUNK = 1024;  // records per wavefront, at least
constexpr double RT_KERNEL2D_NORM = 1.8189136353359467;  // 40 / (7 pi)

MHIP_DEV double rt_kernel2d(double q) {
    double k;
    if (q <= 0.5) {
        k = 1.0 - 6.0 * q * q * (1.0 - q);
    } else {
        const double a = 1.0 - q;
        k = 2.0 * a * a * a;
    }
    return k * RT_KERNEL2D_NORM;
}

// far-first key: descending z as ascending unsigned integers (stable sort
// keeps the input order among equal z)
__global__ __launch_bounds__(256) void k_rt_keys(const double *__restrict__ x, int64_t n, uint64_t *__restrict__ keys,
                                                 uint32_t *__restrict__ vals) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t b = (uint64_t)__double_as_longlong(x[i * 3 + 2]);
    b = (b >> 63) ? ~b : (b | 0x8000000000000000ull);  // ascending z -> ascending b
    keys[i] = ~b;
    vals[i] = (uint32_t)i;
}

// sorted structure-of-arrays: x, y (relative to the grid origin), h, m, then
// j / (4 pi kappa + 1e-100) and kappa per band
struct RtArgs {
    const double *sx, *sy, *sh, *sm, *ssrc, *skap;
    int64_t n;
    int nf, res, tw, ntx;
    double dx;
};

// footprint of a particle at (xs0, xs1) relative to the grid origin, h, m:
// the index ranges of the sightlines within max(h, dx), clipped to the grid
struct RtFoot {
    double xs0, xs1, hs;
    int gx0, gx1, gy0, gy1;
};

MHIP_DEV bool rt_footprint(double xs0, double xs1, double h, double m, double dx, int res, RtFoot &f) {
    f.xs0 = xs0;
    f.xs1 = xs1;
    f.hs = fmax(h, dx);
    const double l0 = (xs0 - f.hs) / dx + 1.0, h0 = (xs0 + f.hs) / dx;
    const double l1 = (xs1 - f.hs) / dx + 1.0, h1 = (xs1 + f.hs) / dx;
    if (!(f.hs != 0.0 && m != 0.0 && h0 > -1.0 && h1 > -1.0 && l0 < (double)res && l1 < (double)res)) return false;
    f.gx0 = l0 > 0.0 ? (int)l0 : 0;
    f.gx1 = h0 >= res - 1 ? res - 1 : (int)h0;
    f.gy0 = l1 > 0.0 ? (int)l1 : 0;
    f.gy1 = h1 >= res - 1 ? res - 1 : (int)h1;
    return f.gx0 <= f.gx1 && f.gy0 <= f.gy1;
}

// number of image tiles (tw x tw sightlines) a footprint touches
MHIP_DEV uint32_t rt_tiles(const RtFoot &f, int tw) {
    return (uint32_t)((f.gx1 / tw - f.gx0 / tw + 1) * (f.gy1 / tw - f.gy0 / tw + 1));
}

// total number of (particle, tile) records, from the caller's arrays: a
// grid-stride loop, one atomic per block
__global__ __launch_bounds__(256) void k_rt_count(const double *__restrict__ x, const double *__restrict__ h,
                                                  const double *__restrict__ m, int64_t n, double ox, double oy,
                                                  double dx, int res, int tw,
                                                  unsigned long long *__restrict__ total) {
    __shared__ unsigned long long part[4];
    unsigned long long c = 0;
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        RtFoot f;
        if (rt_footprint(x[i * 3] - ox, x[i * 3 + 1] - oy, h[i], m[i], dx, res, f)) c += rt_tiles(f, tw);
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long s = part[0] + part[1] + part[2] + part[3];
        if (s) atomicAdd(total, s);
    }
}

// the sorted structure-of-arrays and each sorted particle's record count
__global__ __launch_bounds__(256) void k_rt_gather(const double *__restrict__ x, const double *__restrict__ h,
                                                   const double *__restrict__ m, const double *__restrict__ j,
                                                   const double *__restrict__ kappa, const uint32_t *__restrict__ order,
                                                   RtArgs a, double ox, double oy, double *__restrict__ sx,
                                                   double *__restrict__ sy, double *__restrict__ sh,
                                                   double *__restrict__ sm, double *__restrict__ ssrc,
                                                   double *__restrict__ skap, uint32_t *__restrict__ cnt) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    const int64_t i = order[k];
    // x - center + L/2, as xs = x[i] + L/2 after x -= center
    const double xs0 = x[i * 3] - ox, xs1 = x[i * 3 + 1] - oy;
    sx[k] = xs0;
    sy[k] = xs1;
    sh[k] = h[i];
    sm[k] = m[i];
    for (int b = 0; b < a.nf; b++) {
        const double kap = kappa[i * a.nf + b];
        skap[k * a.nf + b] = kap;
        ssrc[k * a.nf + b] = j[i * a.nf + b] / (FOUR_PI * kap + 1e-100);
    }
    RtFoot f;
    cnt[k] = rt_footprint(xs0, xs1, h[i], m[i], a.dx, a.res, f) ? rt_tiles(f, a.tw) : 0u;
}

// records of sorted particle k at off[k]: key tile id, value k, tiles in
// ascending order
__global__ __launch_bounds__(256) void k_rt_records(RtArgs a, const uint32_t *__restrict__ off, int64_t nrec,
                                                    uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    const int64_t k = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (k >= a.n) return;
    RtFoot f;
    if (!rt_footprint(a.sx[k], a.sy[k], a.sh[k], a.sm[k], a.dx, a.res, f)) return;
    // the counts came from the same footprints, so o stays below nrec
    int64_t o = off[k];
    for (int tx = f.gx0 / a.tw; tx <= f.gx1 / a.tw; tx++)
        for (int ty = f.gy0 / a.tw; ty <= f.gy1 / a.tw && o < nrec; ty++) {
            keys[o] = (uint64_t)(tx * a.ntx + ty);
            vals[o] = (uint32_t)k;
            o++;
        }
}

// tstart[t]: first sorted record of tile t (tstart[ntile] = number of records)
__global__ __launch_bounds__(256) void k_rt_tile_bounds(const uint64_t *__restrict__ keys, int64_t nrec, int ntile,
                                                        uint32_t *__restrict__ tstart) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t > ntile) return;
    int64_t lo = 0, hi = nrec;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < (uint64_t)t)
            lo = mid + 1;
        else
            hi = mid;
    }
    tstart[t] = (uint32_t)lo;
}

struct RtChunkArgs {
    RtArgs p;
    const uint64_t *keys;    // sorted records: tile
    const uint32_t *vals;    //                 sorted particle
    const uint32_t *tstart;  // (ntile + 1)
    int64_t nrec, per;       // records, records per chunk
    double *B, *A;           // per segment slot: [slot][tw * tw][band]; A = exp(-tau_s)
};

// one wavefront per chunk of sorted records.  The chunk's records of tile t
// form the segment stored in slot (chunk + t): the records are sorted by tile,
// so a later chunk never holds a smaller tile and the slots are distinct.
__global__ __launch_bounds__(64) void k_rt_chunk(RtChunkArgs c) {
    extern __shared__ double lds[];
    const RtArgs &a = c.p;
    const int tcells = a.tw * a.tw * a.nf;
    double *I = lds, *T = lds + tcells;
    const int lane = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * c.per, r1 = min(c.nrec, r0 + c.per);
    const int t0 = (int)c.keys[r0], t1 = (int)c.keys[r1 - 1];
    for (int t = t0; t <= t1; t++) {
        const int64_t s0 = max<int64_t>(r0, c.tstart[t]), s1 = min<int64_t>(r1, c.tstart[t + 1]);
        if (s0 >= s1) continue;
        const int tx0 = (t / a.ntx) * a.tw, ty0 = (t % a.ntx) * a.tw;
        const int tx1 = min(tx0 + a.tw, a.res) - 1, ty1 = min(ty0 + a.tw, a.res) - 1;
        for (int e = lane; e < tcells; e += 64) {
            I[e] = 0.0;
            T[e] = 0.0;
        }
        __syncthreads();
        // 64 records at a time: each lane loads one particle and clips its
        // footprint to the tile; they are then applied one after another in
        // ascending order, their data broadcast from the loading lane
        for (int64_t base = s0; base < s1; base += 64) {
            const int64_t r = base + lane;
            bool hit = false;
            int64_t kk = 0;
            double mh2 = 0.0, hinv = 0.0, kap1 = 0.0, src1 = 0.0;
            RtFoot f;
            f.xs0 = f.xs1 = f.hs = 0.0;
            f.gx0 = f.gy0 = 0;
            f.gx1 = f.gy1 = -1;
            if (r < s1) {
                kk = c.vals[r];
                const double m = a.sm[kk];
                if (rt_footprint(a.sx[kk], a.sy[kk], a.sh[kk], m, a.dx, a.res, f)) {
                    f.gx0 = max(f.gx0, tx0);
                    f.gx1 = min(f.gx1, tx1);
                    f.gy0 = max(f.gy0, ty0);
                    f.gy1 = min(f.gy1, ty1);
                    hit = f.gx0 <= f.gx1 && f.gy0 <= f.gy1;
                    mh2 = m / (f.hs * f.hs);
                    hinv = 1.0 / f.hs;
                    if (a.nf == 1) {
                        // one band: kappa and the source term travel with
                        // the particle instead of being read per update
                        kap1 = a.skap[kk];
                        src1 = a.ssrc[kk];
                    }
                }
            }
            uint64_t todo = __ballot(hit);
            while (todo) {
                const int src = __ffsll((long long)todo) - 1;
                todo &= todo - 1;
                const int64_t kp = __shfl(kk, src);
                const double pxs0 = __shfl(f.xs0, src), pxs1 = __shfl(f.xs1, src), phs = __shfl(f.hs, src);
                const double pmh2 = __shfl(mh2, src), phinv = __shfl(hinv, src);
                const double pkap1 = __shfl(kap1, src), psrc1 = __shfl(src1, src);
                const int pgx0 = __shfl(f.gx0, src), pgx1 = __shfl(f.gx1, src), pgy0 = __shfl(f.gy0, src),
                          pgy1 = __shfl(f.gy1, src);
                const double hs2 = phs * phs;
                const int wy = pgy1 - pgy0 + 1, cells = (pgx1 - pgx0 + 1) * wy;
                for (int q = lane; q < cells; q += 64) {
                    const int gx = pgx0 + q / wy, gy = pgy0 + q % wy;
                    double ddx = pxs0 - gx * a.dx;
                    ddx *= ddx;
                    double ddy = pxs1 - gy * a.dx;
                    ddy *= ddy;
                    const double rr = ddx + ddy;
                    if (rr > hs2) continue;
                    const double kern = rt_kernel2d(sqrt(rr) * phinv);
                    const int p = ((gx - tx0) * a.tw + (gy - ty0)) * a.nf;
                    for (int b = 0; b < a.nf; b++) {
                        const double kap = a.nf == 1 ? pkap1 : a.skap[kp * a.nf + b];
                        const double sv = a.nf == 1 ? psrc1 : a.ssrc[kp * a.nf + b];
                        const double tau = kap * kern * pmh2;
                        // 1 - exp(-tau) = -expm1(-tau) stays accurate for small tau;
                        // exp(-tau) itself comes from 1 + expm1(-tau) only while that
                        // sum does not cancel (tau <= 1/2), from exp beyond
                        const double em1 = expm1(-tau), fac2 = -em1;
                        const double fac1 = tau <= 0.5 ? 1.0 + em1 : exp(-tau);
                        T[p + b] += tau;
                        I[p + b] = fac1 * I[p + b] + fac2 * sv;
                    }
                }
            }
        }
        __syncthreads();
        const int64_t o = ((int64_t)blockIdx.x + t) * tcells;
        for (int e = lane; e < tcells; e += 64) {
            c.B[o + e] = I[e];
            c.A[o + e] = exp(-T[e]);
        }
        __syncthreads();
    }
}

// I = i0, then the pixel's tile segments, chunk by chunk (farthest first):
// I <- A_s I + B_s, A_s = exp(-tau_s)
__global__ __launch_bounds__(256) void k_rt_compose(const double *__restrict__ B, const double *__restrict__ A,
                                                    const double *__restrict__ i0,
                                                    const uint32_t *__restrict__ tstart, int64_t per, int res,
                                                    int tw, int ntx, int nf, double *__restrict__ out) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= (int64_t)res * res * nf) return;
    const int b = (int)(p % nf);
    const int64_t pix = p / nf;
    const int gx = (int)(pix / res), gy = (int)(pix % res);
    const int t = (gx / tw) * ntx + gy / tw;
    const int64_t e = (int64_t)((gx % tw) * tw + gy % tw) * nf + b;
    const int64_t tcells = (int64_t)tw * tw * nf;
    double I = i0 ? i0[p] : 0.0;
    const int64_t s0 = tstart[t], s1 = tstart[t + 1];
    if (s0 < s1)
        for (int64_t ch = s0 / per; ch <= (s1 - 1) / per; ch++) {
            const int64_t o = (ch + t) * tcells + e;
            I = A[o] * I + B[o];
        }
    out[p] = I;
}

__global__ __launch_bounds__(256) void k_rt_zero(uint32_t *__restrict__ p, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// bits of a sort key in [0, v]
static int rt_key_bits(uint64_t v) { return 64 - __builtin_clzll(v | 1ull); }

int radtransfer(mhip_ctx *ctx, const double *j, const double *m, const double *kappa, const double *x,
                const double *h, int64_t n, int nf, int res, double L, double cx, double cy, const double *i0,
                double *out) {
    if (n < 0 || nf < 1 || res < 2) return fail(ctx, -1, "bad radiative-transfer shape");
    if (nf > 64) return fail(ctx, -4, "more than 64 frequency bands");
    if (!(L > 0.0)) return fail(ctx, -1, "L must be positive");
    if (n > 0xffffffffll) return fail(ctx, -4, "more than 2^32 particles");
    const int64_t npix = (int64_t)res * res;
    // tile side: LDS holds I and tau for tw^2 pixels and nf bands (<= 64 KB)
    int tw = 16;
    while (tw > 4 && (int64_t)tw * tw * nf * 16 > 65536) tw /= 2;
    const int ntx = (res + tw - 1) / tw, ntile = ntx * ntx;
    const double dx = L / (res - 1);
    const double ox = cx - 0.5 * L, oy = cy - 0.5 * L;
    hipStream_t s = ctx->stream;
    ProfRegion pr(ctx, "radtransfer");
    int rc = 0;
    // the number of records first (it sizes everything else)
    uint64_t nrec = 0;
    if (n > 0) {
        unsigned long long *tot = (unsigned long long *)scratch(ctx, 64, &rc);
        if (!tot) return rc;
        hipLaunchKernelGGL(k_rt_zero, dim3(1), dim3(64), 0, s, (uint32_t *)tot, (int64_t)2);
        hipLaunchKernelGGL(k_rt_count, dim3((unsigned)std::min<int64_t>(2048, (n + 255) / 256)), dim3(256), 0, s, x,
                           h, m, n, ox, oy, dx, res, tw, tot);
        if ((rc = check_hip(ctx, hipMemcpyAsync(&nrec, tot, 8, hipMemcpyDeviceToHost, s), "radtransfer count")))
            return rc;
        if ((rc = check_hip(ctx, hipStreamSynchronize(s), "radtransfer count"))) return rc;
        if (nrec > 0xffffffffull) return fail(ctx, -4, "more than 2^32 (particle, tile) pairs");
    }
    // records per chunk: at least RT_CHUNK, and the segment maps within 1 GB
    // (or twice one per tile)
    const int64_t tcells = (int64_t)tw * tw * nf;
    const int64_t slot_bytes = tcells * 16;
    const int64_t avail = std::max<int64_t>(2 * (int64_t)ntile, (int64_t)(1ll << 30) / slot_bytes);
    const int64_t per = std::max<int64_t>(RT_CHUNK, ((int64_t)nrec + (avail - ntile) - 1) / (avail - ntile));
    const int64_t nchunk = ((int64_t)nrec + per - 1) / per;
    const int64_t nslot = nchunk + ntile;
    const int64_t nr = (int64_t)nrec;
    // scratch: segment maps, sorted particles, particle sort, counts and
    // offsets, records and their sort, tile bounds
    const size_t cwn = radix_counts_words(std::max<int64_t>(n, 1)), cwr = radix_counts_words(std::max<int64_t>(nr, 1));
    const size_t bytes = (size_t)nslot * slot_bytes + (size_t)n * (4 + 2 * nf) * 8 + (size_t)n * 24 + cwn * 4 +
                         (size_t)n * 8 + scan_work_words(n) * 4 + (size_t)nr * 24 + cwr * 4 + ((size_t)ntile + 1) * 4 +
                         256;
    char *sc = (char *)scratch(ctx, bytes, &rc);
    if (!sc) return rc;
    double *B = (double *)sc, *A = B + (size_t)nslot * tcells;
    double *sx = A + (size_t)nslot * tcells, *sy = sx + n, *sh = sy + n, *sm = sh + n, *ssrc = sm + n,
           *skap = ssrc + n * nf;
    uint64_t *k0 = (uint64_t *)(skap + n * nf), *k1 = k0 + n, *rk0 = k1 + n, *rk1 = rk0 + nr;
    uint32_t *v0 = (uint32_t *)(rk1 + nr), *v1 = v0 + n, *rv0 = v1 + n, *rv1 = rv0 + nr, *cnt = rv1 + nr,
             *off = cnt + n, *work = off + n, *counts = work + scan_work_words(n), *tstart = counts + std::max(cwn, cwr);
    RtArgs a{sx, sy, sh, sm, ssrc, skap, n, nf, res, tw, ntx, dx};
    if (nr > 0) {
        const dim3 g((unsigned)((n + 255) / 256));
        hipLaunchKernelGGL(k_rt_keys, g, dim3(256), 0, s, x, n, k0, v0);
        uint64_t *ks;
        uint32_t *vs;
        if ((rc = radix_sort_pairs(ctx, k0, v0, k1, v1, counts, n, 64, &ks, &vs))) return rc;
        hipLaunchKernelGGL(k_rt_gather, g, dim3(256), 0, s, x, h, m, j, kappa, vs, a, ox, oy, sx, sy, sh, sm, ssrc,
                           skap, cnt);
        exclusive_scan_u32(ctx, cnt, n, off, work);
        hipLaunchKernelGGL(k_rt_records, g, dim3(256), 0, s, a, off, nr, rk0, rv0);
        uint64_t *rks;
        uint32_t *rvs;
        if ((rc = radix_sort_pairs(ctx, rk0, rv0, rk1, rv1, counts, nr, rt_key_bits((uint64_t)ntile), &rks, &rvs)))
            return rc;
        hipLaunchKernelGGL(k_rt_tile_bounds, dim3((unsigned)((ntile + 256) / 256)), dim3(256), 0, s, rks, nr, ntile,
                           tstart);
        RtChunkArgs c{a, rks, rvs, tstart, nr, per, B, A};
        hipLaunchKernelGGL(k_rt_chunk, dim3((unsigned)nchunk), dim3(64), (size_t)slot_bytes, s, c);
    } else {
        hipLaunchKernelGGL(k_rt_zero, dim3((unsigned)((ntile + 256) / 256)), dim3(256), 0, s, tstart,
                           (int64_t)ntile + 1);
    }
    hipLaunchKernelGGL(k_rt_compose, dim3((unsigned)((npix * nf + 255) / 256)), dim3(256), 0, s, B, A, i0, tstart, per,
                       res, tw, ntx, nf, out);
    return check_hip(ctx, hipGetLastError(), "radtransfer launch");
}

}  // namespace mhip
