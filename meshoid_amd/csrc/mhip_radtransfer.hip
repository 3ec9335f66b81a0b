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
// Each step is affine in I, so the ordered product splits exactly into slabs
// of consecutive particles: a slab computed from I = 0 gives (B_s, tau_s) and
// acts as I <- exp(-tau_s) I + B_s.  Slabs x image tiles run in parallel, one
// wavefront per (slab, tile) holding its tile in LDS and sweeping the slab's
// particles in order, 64 loaded at a time and only those touching the tile
// applied (one lane per pixel of a footprint, so a wavefront's LDS updates
// stay in program order); a last kernel composes the slabs per pixel,
// farthest first, from the background i0.
//
// The reference's multi-threaded driver composes its chunks with the largest
// z nearest the observer while each chunk is swept with the largest z
// farthest, and adds i0 once per chunk, so its result depends on the thread
// count; this follows the single-chunk order throughout.
#include "mhip_internal.hpp"

namespace mhip {

constexpr double FOUR_PI = 12.566370614359172;
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
__global__ __launch_bounds__(256) void k_rt_gather(const double *__restrict__ x, const double *__restrict__ h,
                                                   const double *__restrict__ m, const double *__restrict__ j,
                                                   const double *__restrict__ kappa, const uint32_t *__restrict__ order,
                                                   int64_t n, int nf, double ox, double oy, double *__restrict__ sx,
                                                   double *__restrict__ sy, double *__restrict__ sh,
                                                   double *__restrict__ sm, double *__restrict__ ssrc,
                                                   double *__restrict__ skap) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t i = order[k];
    // x - center + L/2, as xs = x[i] + L/2 after x -= center
    sx[k] = (x[i * 3] - ox);
    sy[k] = (x[i * 3 + 1] - oy);
    sh[k] = h[i];
    sm[k] = m[i];
    for (int b = 0; b < nf; b++) {
        const double kap = kappa[i * nf + b];
        skap[k * nf + b] = kap;
        ssrc[k * nf + b] = j[i * nf + b] / (FOUR_PI * kap + 1e-100);
    }
}

struct RtArgs {
    const double *sx, *sy, *sh, *sm, *ssrc, *skap;
    int64_t n, per;  // particles, particles per slab
    int nf, res, tw, ntx, nslab;
    double dx;
    double *B, *T;  // per slab: [slab][pixel][band]
};

// one wavefront per (slab, tile): the slab's particles in order, the tile in
// LDS (I and tau per pixel and band)
__global__ __launch_bounds__(64) void k_rt_slab(RtArgs a) {
    extern __shared__ double lds[];
    const int ntile = a.ntx * a.ntx;
    const int tile = blockIdx.x % ntile, slab = blockIdx.x / ntile;
    const int tx0 = (tile / a.ntx) * a.tw, ty0 = (tile % a.ntx) * a.tw;
    const int tx1 = min(tx0 + a.tw, a.res) - 1, ty1 = min(ty0 + a.tw, a.res) - 1;
    const int tcells = a.tw * a.tw * a.nf;
    double *I = lds, *T = lds + tcells;
    const int lane = threadIdx.x;
    for (int c = lane; c < tcells; c += 64) {
        I[c] = 0.0;
        T[c] = 0.0;
    }
    const int64_t k0 = (int64_t)slab * a.per, k1 = min(a.n, k0 + a.per);
    // 64 particles at a time: each lane loads one (coalesced) and tests it
    // against the tile; the particles that touch the tile are then applied
    // one after another in ascending order, their data broadcast from the
    // loading lane
    for (int64_t base = k0; base < k1; base += 64) {
        const int64_t k = base + lane;
        bool hit = false;
        double xs0 = 0.0, xs1 = 0.0, hs = 0.0, m = 0.0, mh2 = 0.0, hinv = 0.0;
        int gx0 = 0, gx1 = -1, gy0 = 0, gy1 = -1;
        if (k < k1) {
            m = a.sm[k];
            hs = fmax(a.sh[k], a.dx);
            xs0 = a.sx[k];
            xs1 = a.sy[k];
            // index range of the footprint, clipped to the grid and the tile
            const double l0 = (xs0 - hs) / a.dx + 1.0, h0 = (xs0 + hs) / a.dx;
            const double l1 = (xs1 - hs) / a.dx + 1.0, h1 = (xs1 + hs) / a.dx;
            if (hs != 0.0 && m != 0.0 && h0 > -1.0 && h1 > -1.0 && l0 < (double)a.res && l1 < (double)a.res) {
                gx0 = max(l0 > 0.0 ? (int)l0 : 0, tx0);
                gx1 = min(h0 >= a.res - 1 ? a.res - 1 : (int)h0, tx1);
                gy0 = max(l1 > 0.0 ? (int)l1 : 0, ty0);
                gy1 = min(h1 >= a.res - 1 ? a.res - 1 : (int)h1, ty1);
                hit = gx0 <= gx1 && gy0 <= gy1;
                mh2 = m / (hs * hs);
                hinv = 1.0 / hs;
            }
        }
        uint64_t todo = __ballot(hit);
        while (todo) {
            const int src = __ffsll((long long)todo) - 1;
            todo &= todo - 1;
            const int64_t kp = base + src;
            const double pxs0 = __shfl(xs0, src), pxs1 = __shfl(xs1, src), phs = __shfl(hs, src);
            const double pmh2 = __shfl(mh2, src), phinv = __shfl(hinv, src);
            const int pgx0 = __shfl(gx0, src), pgx1 = __shfl(gx1, src), pgy0 = __shfl(gy0, src),
                      pgy1 = __shfl(gy1, src);
            const double hs2 = phs * phs;
            const int wy = pgy1 - pgy0 + 1, cells = (pgx1 - pgx0 + 1) * wy;
            for (int c = lane; c < cells; c += 64) {
                const int gx = pgx0 + c / wy, gy = pgy0 + c % wy;
                double ddx = pxs0 - gx * a.dx;
                ddx *= ddx;
                double ddy = pxs1 - gy * a.dx;
                ddy *= ddy;
                const double r = ddx + ddy;
                if (r > hs2) continue;
                const double kern = rt_kernel2d(sqrt(r) * phinv);
                const int p = ((gx - tx0) * a.tw + (gy - ty0)) * a.nf;
                for (int b = 0; b < a.nf; b++) {
                    const double tau = a.skap[kp * a.nf + b] * kern * pmh2;
                    // exp(-tau) = 1 + expm1(-tau): one transcendental per update
                    const double em1 = expm1(-tau), fac1 = 1.0 + em1, fac2 = -em1;
                    T[p + b] += tau;
                    I[p + b] = fac1 * I[p + b] + fac2 * a.ssrc[kp * a.nf + b];
                }
            }
        }
    }
    // the slab's maps for this tile
    const int64_t npix = (int64_t)a.res * a.res;
    for (int c = lane; c < a.tw * a.tw; c += 64) {
        const int gx = tx0 + c / a.tw, gy = ty0 + c % a.tw;
        if (gx > tx1 || gy > ty1) continue;
        const int64_t o = ((int64_t)slab * npix + (int64_t)gx * a.res + gy) * a.nf;
        for (int b = 0; b < a.nf; b++) {
            a.B[o + b] = I[c * a.nf + b];
            a.T[o + b] = T[c * a.nf + b];
        }
    }
}

// I = i0, then slab by slab farthest first: I <- exp(-tau_s) I + B_s
__global__ __launch_bounds__(256) void k_rt_compose(const double *__restrict__ B, const double *__restrict__ T,
                                                    const double *__restrict__ i0, int64_t npix_nf, int nslab,
                                                    double *__restrict__ out) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= npix_nf) return;
    double I = i0 ? i0[p] : 0.0;
    for (int s = 0; s < nslab; s++) {
        const int64_t o = (int64_t)s * npix_nf + p;
        I = exp(-T[o]) * I + B[o];
    }
    out[p] = I;
}

int radtransfer(mhip_ctx *ctx, const double *j, const double *m, const double *kappa, const double *x,
                const double *h, int64_t n, int nf, int res, double L, double cx, double cy, const double *i0,
                double *out) {
    if (n < 0 || nf < 1 || res < 2) return fail(ctx, -1, "bad radiative-transfer shape");
    if (nf > 64) return fail(ctx, -4, "more than 64 frequency bands");
    if (!(L > 0.0)) return fail(ctx, -1, "L must be positive");
    if (n > 0xffffffffll) return fail(ctx, -4, "more than 2^32 particles");
    const int64_t npix = (int64_t)res * res;
    // tile side: LDS holds I and tau for tw^2 pixels and nf bands (<= 64 KB);
    // small tiles spread a clump's pixels over many wavefronts
    int tw = 16;
    while (tw > 4 && (int64_t)tw * tw * nf * 16 > 65536) tw /= 2;
    const int ntx = (res + tw - 1) / tw, ntile = ntx * ntx;
    // slabs: about 16k (slab, tile) wavefronts, so that the sequential sweep
    // of one wavefront stays short; slab maps within 1 GB
    const int64_t map_bytes = npix * nf * 16;
    int64_t nslab = std::max<int64_t>(1, std::min<int64_t>(16384 / ntile + 1, (int64_t)(1ll << 30) / map_bytes));
    nslab = std::min<int64_t>(nslab, std::max<int64_t>(1, n));
    const int64_t per = n > 0 ? (n + nslab - 1) / nslab : 1;
    const size_t cw = radix_counts_words(std::max<int64_t>(n, 1));
    const size_t bytes = (size_t)n * 24 + cw * 4 + (size_t)n * (4 + 2 * nf) * 8 + (size_t)(2 * nslab * map_bytes) + 64;
    int rc = 0;
    char *sc = (char *)scratch(ctx, bytes, &rc);
    if (!sc) return rc;
    double *B = (double *)sc, *T = B + nslab * npix * nf;
    double *sx = T + nslab * npix * nf, *sy = sx + n, *sh = sy + n, *sm = sh + n, *ssrc = sm + n, *skap = ssrc + n * nf;
    uint64_t *k0 = (uint64_t *)(skap + n * nf), *k1 = k0 + n;
    uint32_t *v0 = (uint32_t *)(k1 + n), *v1 = v0 + n, *counts = v1 + n;
    hipStream_t s = ctx->stream;
    ProfRegion pr(ctx, "radtransfer");
    const double dx = L / (res - 1);
    if (n > 0) {
        const dim3 g((unsigned)((n + 255) / 256));
        hipLaunchKernelGGL(k_rt_keys, g, dim3(256), 0, s, x, n, k0, v0);
        uint64_t *ks;
        uint32_t *vs;
        if ((rc = radix_sort_pairs(ctx, k0, v0, k1, v1, counts, n, 64, &ks, &vs))) return rc;
        hipLaunchKernelGGL(k_rt_gather, g, dim3(256), 0, s, x, h, m, j, kappa, vs, n, nf, cx - 0.5 * L,
                           cy - 0.5 * L, sx, sy, sh, sm, ssrc, skap);
        RtArgs a{sx, sy, sh, sm, ssrc, skap, n, per, nf, res, tw, ntx, (int)nslab, dx, B, T};
        const size_t lds = (size_t)tw * tw * nf * 16;
        hipLaunchKernelGGL(k_rt_slab, dim3((unsigned)(nslab * ntile)), dim3(64), lds, s, a);
    } else {
        nslab = 0;
    }
    hipLaunchKernelGGL(k_rt_compose, dim3((unsigned)((npix * nf + 255) / 256)), dim3(256), 0, s, B, T, i0,
                       npix * nf, (int)nslab, out);
    return check_hip(ctx, hipGetLastError(), "radtransfer launch");
}

}  // namespace mhip
