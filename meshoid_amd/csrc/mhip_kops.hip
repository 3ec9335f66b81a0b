// Kernel-weighted neighbour statistics on the GPU.
//
// Replaces
//   Meshoid.get_kernel_weights                  src/meshoid/Meshoid.py:172-177
//   Meshoid.KernelAverage                       src/meshoid/Meshoid.py:369-381
//   Meshoid.KernelVariance                      src/meshoid/Meshoid.py:352-367
//   Meshoid.Integrate                           src/meshoid/Meshoid.py:333-350
//
// All reductions accumulate in float64.  The fused operators read the device
// neighbour list (column-major sorted-space indices) and recompute each pair
// distance in float64 from the hi + lo positions, which is exact for float32
// inputs, instead of reading back float32 distances.
#include <string>

#include "../../include/meshoid_hip.h"
#include "mhip_internal.hpp"

namespace mhip {

// W(r/h) / sum_j W(r_j/h) per row of a row-major (nq, k) distance array
template <typename DT, typename HT>
__global__ __launch_bounds__(256) void k_kernel_weights(const DT *__restrict__ dist, const HT *__restrict__ h,
                                                        int64_t nq, int k, double *__restrict__ w) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const double ih = 1.0 / (double)h[i];
    const DT *d = dist + i * k;
    double *o = w + i * k;
    double s = 0.0;
    for (int j = 0; j < k; j++) {
        const double v = kernel_w<double>((double)d[j] * ih);
        o[j] = v;
        s += v;
    }
    // a row whose kernel sum is zero gives 0/0 = NaN, as the reference does
    const double is = 1.0 / s;
    for (int j = 0; j < k; j++) o[j] *= is;
}

struct KopsArgs {
    const float *sx, *sy, *sz, *lx, *ly, *lz;
    double box;
    int per, p64;
    const uint32_t *nl_idx, *qsorted, *qrow, *perm;
    const double *h;  // float64-refined smoothing length per slot
    int64_t nq;
    int k;
    const void *fs;  // [K][npad] sorted-space field
    int64_t npad;
    int K;
    double *out;  // [row][K]
};

template <typename FT>
MHIP_DEV double field_at(const KopsArgs &a, int c, uint32_t p) {
    return (double)((const FT *)a.fs)[(int64_t)c * a.npad + p];
}

MHIP_DEV double pos_d(const float *hi, const float *lo, int p64, uint32_t p) {
    return (double)hi[p] + (p64 ? (double)lo[p] : 0.0);
}

// sum_j W(r_j/h) f_j / sum_j W(r_j/h) (scalar f, the einsum "ij,ij->i")
template <typename FT>
__global__ __launch_bounds__(256) void k_kernel_average(KopsArgs a) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= a.nq) return;
    const uint32_t qp = a.qsorted ? a.qsorted[s] : (uint32_t)s;
    const uint32_t row = a.qrow ? a.qrow[s] : a.perm[s];
    const double qx = pos_d(a.sx, a.lx, a.p64, qp), qy = pos_d(a.sy, a.ly, a.p64, qp),
                 qz = pos_d(a.sz, a.lz, a.p64, qp);
    const double ih = 1.0 / a.h[s];
    double sw = 0.0, swf = 0.0;
    for (int j = 0; j < a.k; j++) {
        const uint32_t p = a.nl_idx[(int64_t)j * a.nq + s];
        double dx = pos_d(a.sx, a.lx, a.p64, p) - qx;
        double dy = pos_d(a.sy, a.ly, a.p64, p) - qy;
        double dz = pos_d(a.sz, a.lz, a.p64, p) - qz;
        if (a.per) {
            dx = wrap<double>(dx, a.box);
            dy = wrap<double>(dy, a.box);
            dz = wrap<double>(dz, a.box);
        }
        const double w = kernel_w<double>(sqrt(dx * dx + dy * dy + dz * dz) * ih);
        sw += w;
        swf += w * field_at<FT>(a, 0, p);
    }
    a.out[row] = swf / sw;
}

// population standard deviation of f over the k neighbours, per component
// (np.std(f[ngb], axis=1)): two passes over the gathered values
template <typename FT>
__global__ __launch_bounds__(256) void k_kernel_variance(KopsArgs a) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= a.nq) return;
    const uint32_t row = a.qrow ? a.qrow[s] : a.perm[s];
    const double ik = 1.0 / (double)a.k;
    for (int c = 0; c < a.K; c++) {
        double m = 0.0;
        for (int j = 0; j < a.k; j++) m += field_at<FT>(a, c, a.nl_idx[(int64_t)j * a.nq + s]);
        m *= ik;
        double v = 0.0;
        for (int j = 0; j < a.k; j++) {
            const double d = field_at<FT>(a, c, a.nl_idx[(int64_t)j * a.nq + s]) - m;
            v += d * d;
        }
        a.out[(int64_t)row * a.K + c] = sqrt(v * ik);
    }
}

// sum_i vol_i f_i per component: per-block float64 partial sums, then one
// block folds the partials in a fixed order (deterministic)
constexpr int INT_BLOCK = 256;

template <typename VT, typename FT>
__global__ __launch_bounds__(INT_BLOCK) void k_integrate_partial(const VT *__restrict__ vol,
                                                                 const FT *__restrict__ f, int64_t n,
                                                                 int K, int c, double *__restrict__ part) {
    __shared__ double sh[INT_BLOCK];
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)INT_BLOCK + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * INT_BLOCK)
        acc += (double)vol[i] * (double)f[i * K + c];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = INT_BLOCK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(INT_BLOCK) void k_integrate_final(const double *__restrict__ part, int np,
                                                               double *__restrict__ out) {
    __shared__ double sh[INT_BLOCK];
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += INT_BLOCK) acc += part[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = INT_BLOCK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

int kernel_weights(mhip_ctx *ctx, const void *dist, int dist_f64, const void *h, int h_f64, int64_t nq, int k,
                   double *w) {
    if (nq < 0 || k < 1) return fail(ctx, -1, "bad kernel-weight shape");
    if (nq == 0) return 0;
    const dim3 g((unsigned)((nq + 255) / 256)), b(256);
#define MHIP_KW(DT, HT) \
    hipLaunchKernelGGL((k_kernel_weights<DT, HT>), g, b, 0, ctx->stream, (const DT *)dist, (const HT *)h, nq, k, w)
    if (dist_f64) {
        if (h_f64) MHIP_KW(double, double);
        else MHIP_KW(double, float);
    } else {
        if (h_f64) MHIP_KW(float, double);
        else MHIP_KW(float, float);
    }
#undef MHIP_KW
    return check_hip(ctx, hipGetLastError(), "kernel weights launch");
}

// which: 0 average (K must be 1), 1 standard deviation
int kernel_stat(mhip_ctx *ctx, const mhip_nlist *nl, const void *f, int f_f64, int64_t nf_rows, int K,
                int which, double *out) {
    const mhip_index *ix = nl->index;
    if (!ix || nl->nq < 1) return fail(ctx, -1, "neighbour list is empty");
    if (K < 1) return fail(ctx, -1, "no field components");
    if (which == 0 && K != 1) return fail(ctx, -1, "KernelAverage takes a scalar field");
    int rc = 0;
    const size_t fb = f_f64 ? 8 : 4;
    void *fs = scratch(ctx, (size_t)K * ix->npad * fb, &rc);
    if (!fs) return rc;
    const dim3 gp((unsigned)((ix->npad + 255) / 256));
    if (f_f64)
        hipLaunchKernelGGL(k_permute_fields<double>, gp, dim3(256), 0, ctx->stream, (const double *)f,
                           nf_rows, K, ix->perm, ix->n, ix->npad, (double *)fs);
    else
        hipLaunchKernelGGL(k_permute_fields<float>, gp, dim3(256), 0, ctx->stream, (const float *)f,
                           nf_rows, K, ix->perm, ix->n, ix->npad, (float *)fs);
    KopsArgs a;
    a.sx = ix->sx;
    a.sy = ix->sy;
    a.sz = ix->sz;
    a.lx = ix->lx;
    a.ly = ix->ly;
    a.lz = ix->lz;
    a.per = ix->box > 0;
    a.box = a.per ? ix->box : 0.0;
    a.p64 = ix->pos_f64;
    a.nl_idx = nl->idx;
    a.qsorted = nl->masked ? nl->qsorted : nullptr;
    a.qrow = nl->masked ? nl->qrow : nullptr;
    a.perm = ix->perm;
    a.h = nl->h64;
    a.nq = nl->nq;
    a.k = nl->k;
    a.fs = fs;
    a.npad = ix->npad;
    a.K = K;
    a.out = out;
    ProfRegion pr(ctx, which == 0 ? "kernel_average" : "kernel_variance");
    const dim3 g((unsigned)((nl->nq + 255) / 256));
    if (which == 0) {
        if (f_f64)
            hipLaunchKernelGGL(k_kernel_average<double>, g, dim3(256), 0, ctx->stream, a);
        else
            hipLaunchKernelGGL(k_kernel_average<float>, g, dim3(256), 0, ctx->stream, a);
    } else {
        if (f_f64)
            hipLaunchKernelGGL(k_kernel_variance<double>, g, dim3(256), 0, ctx->stream, a);
        else
            hipLaunchKernelGGL(k_kernel_variance<float>, g, dim3(256), 0, ctx->stream, a);
    }
    return check_hip(ctx, hipGetLastError(), "kernel statistic launch");
}

int integrate(mhip_ctx *ctx, const void *vol, int vol_f64, const void *f, int f_f64, int64_t n, int K, double *out) {
    if (n < 0 || K < 1) return fail(ctx, -1, "bad integrate shape");
    const int nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + INT_BLOCK - 1) / INT_BLOCK));
    int rc = 0;
    double *part = (double *)scratch(ctx, (size_t)nb * sizeof(double), &rc);
    if (!part) return rc;
    ProfRegion pr(ctx, "integrate");
    for (int c = 0; c < K; c++) {
#define MHIP_IP(VT, FT)                                                                                    \
    hipLaunchKernelGGL((k_integrate_partial<VT, FT>), dim3(nb), dim3(INT_BLOCK), 0, ctx->stream, (const VT *)vol, \
                       (const FT *)f, n, K, c, part)
        if (vol_f64) {
            if (f_f64) MHIP_IP(double, double);
            else MHIP_IP(double, float);
        } else {
            if (f_f64) MHIP_IP(float, double);
            else MHIP_IP(float, float);
        }
#undef MHIP_IP
        hipLaunchKernelGGL(k_integrate_final, dim3(1), dim3(INT_BLOCK), 0, ctx->stream, part, nb, out + c);
    }
    return check_hip(ctx, hipGetLastError(), "integrate launch");
}

}  // namespace mhip

// ---------------------------------------------------------------------------
// Integrate without allocation: the per-block partial sums of every
// component live in a file-scope device array instead of the context's
// scratch.  The array's host-side name is only a stand-in: its device address
// comes from hipGetSymbolAddress, looked up once per device (with that device
// current) and kept.  Calls on one device share the array, so calls that
// overlap on one device race; every partial is written before it is read, so
// nothing needs clearing between calls.
// ---------------------------------------------------------------------------
#include <mutex>

static __device__ __align__(256) double scratch[MHIP_NOALLOC_INTEGRATE_MAX_K * 1024];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

double *noalloc_scratch(mhip_ctx *ctx, int *rc) {
    *rc = 0;
    if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES) {
        *rc = mhip::fail(ctx, MHIP_ERR_VALUE, "device index beyond the no-allocation table");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    double *&p = g_scratch_ptr[ctx->device];
    if (!p) {
        void *dp = nullptr;
        hipError_t e = hipGetSymbolAddress(&dp, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !dp) {
            *rc = mhip::check_hip(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress (scratch)");
            return nullptr;
        }
        p = (double *)dp;
    }
    return p;
}
}  // namespace

extern "C" int mhip_integrate_noalloc(mhip_ctx *ctx, const void *vol, int vol_f64, const void *f, int f_f64,
                                      int64_t n, int K, double *out) {
    using namespace mhip;
    if (!ctx) return MHIP_ERR_VALUE;
    if (!vol || !f || !out) return fail(ctx, MHIP_ERR_VALUE, "null argument");
    if (n < 0 || K < 1) return fail(ctx, MHIP_ERR_VALUE, "bad integrate shape");
    if (K > MHIP_NOALLOC_INTEGRATE_MAX_K)
        return fail(ctx, MHIP_ERR_VALUE,
                    "mhip_integrate_noalloc: at most " + std::to_string(MHIP_NOALLOC_INTEGRATE_MAX_K) +
                        " components (K = " + std::to_string(K) + ")");
    int rc = check_hip(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    if (rc) return rc;
    double *part = noalloc_scratch(ctx, &rc);
    if (!part) return rc;
    // the partition of mhip_integrate: the same blocks and sums, bit for bit
    const int nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + INT_BLOCK - 1) / INT_BLOCK));
    ProfRegion pr(ctx, "integrate_noalloc");
    for (int c = 0; c < K; c++) {
        double *pc = part + (int64_t)c * 1024;
#define MHIP_IP(VT, FT)                                                                                    \
    hipLaunchKernelGGL((k_integrate_partial<VT, FT>), dim3(nb), dim3(INT_BLOCK), 0, ctx->stream, (const VT *)vol, \
                       (const FT *)f, n, K, c, pc)
        if (vol_f64) {
            if (f_f64) MHIP_IP(double, double);
            else MHIP_IP(double, float);
        } else {
            if (f_f64) MHIP_IP(float, double);
            else MHIP_IP(float, float);
        }
#undef MHIP_IP
        hipLaunchKernelGGL(k_integrate_final, dim3(1), dim3(INT_BLOCK), 0, ctx->stream, pc, nb, out + c);
    }
    return check_hip(ctx, hipGetLastError(), "integrate launch");
}
