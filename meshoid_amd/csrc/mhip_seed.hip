// Learned initial guess for the smoothing length (no counterpart in the
// reference, whose HsmlIter always starts from the bracket
// [r_1, c_d r_{k-1}] of src/meshoid/kernel_density.py:37-38).
//
//   features  28 per particle: log(1 + count) of the 3x3x3 block of coarse
//             cells around the particle's cell (cell = 1/128 of the sort-key
//             quantisation frame, so coarse cells are Hilbert-key prefixes and the
//             counts are run lengths of the sorted order) + log(cell size).
//   network   28 -> 64 -> 64 -> 1, SiLU; bfloat16 weights and activations,
//             float32 accumulation.  Output: log(h / cell).
//   inference v_mfma_f32_16x16x32_bf16, 16 particles per MFMA tile, hidden
//             activations re-laid out through LDS between layers.
//   training  per-sample forward + backward on the VALU (same bfloat16
//             rounding points, straight-through in the backward pass), weight
//             gradients reduced per wavefront in LDS and accumulated with
//             float32 atomics; Adam on float32 master weights.
//
// Parameter blob (float32, 6081): W1 (64x28 out-major), b1 (64), W2 (64x64),
// b2 (64), w3 (64), b3 (1).  Inference blob: W1 bf16 [64][32] (inputs 28..31
// zero), W2 bf16 [64][64], then float32 b1, b2, w3 (bf16-rounded), b3, pad x3,
// mu[28], sd[28].
#include "mhip_internal.hpp"

namespace mhip {

constexpr int SEED_G = 128;  // coarse cells per axis
constexpr int NIN = 28, NH = 64;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NH * NIN, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH, NPARAM = OFF_B3 + 1;
static_assert(NPARAM == 6081, "parameter count");
// inference blob offsets (bytes)
constexpr int IB_W1 = 0, IB_W2 = IB_W1 + NH * 32 * 2, IB_F = IB_W2 + NH * NH * 2;
constexpr int IF_B1 = 0, IF_B2 = 64, IF_W3 = 128, IF_B3 = 192, IF_MU = 196, IF_SD = 224, IF_N = 252;
constexpr size_t IB_BYTES = IB_F + IF_N * 4;

typedef __bf16 v8bf16 __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

MHIP_DEV float bf16r(float x) { return (float)(__bf16)x; }  // round to nearest even
MHIP_DEV float silu(float z) { return z / (1.0f + __expf(-z)); }
MHIP_DEV float dsilu(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

struct SeedView {
    const float *sx, *sy, *sz;
    const uint32_t *perm;
    int64_t n;
    int dim, per;
    float org[3], scale, logcell;
    const uint32_t *counts;  // [G^3] particles per coarse cell
};

// coarse cell of sorted particle p: the same arithmetic as the sort keys
MHIP_DEV void coarse_cell(const SeedView &v, int64_t p, int (&c)[3]) {
    const float x[3] = {v.sx[p], v.sy[p], v.sz[p]};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (a < v.dim) {
            float t = (x[a] - v.org[a]) * v.scale;
            t = fminf(fmaxf(t, 0.f), 2097151.0f);
            c[a] = (int)((uint32_t)t >> 14);
        } else {
            c[a] = 0;
        }
    }
}

MHIP_DEV uint32_t cell_id(int cx, int cy, int cz) {
    return ((uint32_t)cx * SEED_G + (uint32_t)cy) * SEED_G + (uint32_t)cz;
}

// raw feature j of particle p (j < 27: log1p of a neighbour cell count, j ==
// 27: log cell size, j > 27: 0)
MHIP_DEV float feature(const SeedView &v, const int (&c)[3], int j) {
    if (j == 27) return v.logcell;
    if (j > 27) return 0.f;
    const int d[3] = {j / 9 - 1, (j / 3) % 3 - 1, j % 3 - 1};
    int q[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        q[a] = c[a] + d[a];
        if (a >= v.dim) {
            if (q[a] != 0) return 0.f;
        } else if (v.per) {
            q[a] = (q[a] + SEED_G) % SEED_G;
        } else if (q[a] < 0 || q[a] >= SEED_G) {
            return 0.f;
        }
    }
    return log1pf((float)v.counts[cell_id(q[0], q[1], q[2])]);
}

__global__ void k_fill_u32(uint32_t *__restrict__ a, int64_t n, uint32_t val) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) a[i] = val;
}

// particles per coarse cell: a cell is a Hilbert-key prefix, so its particles are a
// contiguous run of the sorted order; each wavefront adds its runs' lengths
__global__ __launch_bounds__(256) void k_seed_hist(SeedView v, uint32_t *__restrict__ counts) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (WAVE - 1);
    const bool valid = p < v.n;
    uint32_t id = 0xffffffffu;
    if (valid) {
        int c[3];
        coarse_cell(v, p, c);
        id = cell_id(c[0], c[1], c[2]);
    }
    const uint32_t prev = __shfl_up(id, 1);
    const bool head = valid && (lane == 0 || prev != id);
    const uint64_t heads = __ballot(head);
    const uint64_t valids = __ballot(valid);
    if (head) {
        const uint64_t above = heads & ~((2ull << lane) - 1ull);  // heads after this lane
        const int end = above ? __ffsll((long long)above) - 1 : 64 - __clzll((long long)valids);
        atomicAdd(counts + id, (uint32_t)(end - lane));
    }
}

// features of every particle, in original row order: F[row][28] (raw, not
// normalised), for building training sets
__global__ __launch_bounds__(256) void k_seed_features(SeedView v, float *__restrict__ F) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= v.n) return;
    int c[3];
    coarse_cell(v, p, c);
    float *o = F + (int64_t)v.perm[p] * NIN;
    for (int j = 0; j < NIN; j++) o[j] = feature(v, c, j);
}

// inference: each wavefront predicts 64 consecutive sorted particles as four
// 16-particle MFMA tiles.  MFMA 16x16x32 bf16 operand layout (lane l):
//   A[i][k]: i = l % 16, k = 8 (l / 16) + 0..7;  B[k][j]: k likewise, j = l % 16;
//   D[i][j]: i = 4 (l / 16) + 0..3, j = l % 16.
__global__ __launch_bounds__(256) void k_seed_infer(SeedView v, const unsigned char *__restrict__ blob,
                                                   float *__restrict__ h_rows) {
    __shared__ __bf16 s_a1[4][16][NH];
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x >> 6;
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t base = ((int64_t)b * 4 + w) * WAVE;
    if (base >= v.n) return;
    const __bf16 *W1b = (const __bf16 *)(blob + IB_W1);
    const __bf16 *W2b = (const __bf16 *)(blob + IB_W2);
    const float *fb = (const float *)(blob + IB_F);
    const int col = lane & 15, grp = lane >> 4;
    v8bf16 W1B[4], W2B[4][2];
    float b1v[4], b2v[4], w3v[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        W1B[t] = *(const v8bf16 *)(W1b + (16 * t + col) * 32 + 8 * grp);
#pragma unroll
        for (int kb = 0; kb < 2; kb++) W2B[t][kb] = *(const v8bf16 *)(W2b + (16 * t + col) * NH + 32 * kb + 8 * grp);
        b1v[t] = fb[IF_B1 + 16 * t + col];
        b2v[t] = fb[IF_B2 + 16 * t + col];
        w3v[t] = fb[IF_W3 + 16 * t + col];
    }
    const float b3 = fb[IF_B3];
    float mu[8], isd[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int j = 8 * grp + e;
        mu[e] = j < NIN ? fb[IF_MU + j] : 0.f;
        isd[e] = j < NIN ? 1.0f / fb[IF_SD + j] : 0.f;
    }
    for (int tile = 0; tile < 4; tile++) {
        const int64_t t0 = base + 16 * tile;
        const int64_t p = min<int64_t>(t0 + col, v.n - 1);
        int c[3];
        coarse_cell(v, p, c);
        v8bf16 A1;
#pragma unroll
        for (int e = 0; e < 8; e++) A1[e] = (__bf16)((feature(v, c, 8 * grp + e) - mu[e]) * isd[e]);
        // layer 1
#pragma unroll
        for (int t = 0; t < 4; t++) {
            v4f z = {0.f, 0.f, 0.f, 0.f};
            z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, W1B[t], z, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; r++) s_a1[w][4 * grp + r][16 * t + col] = (__bf16)silu(z[r] + b1v[t]);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        v8bf16 A2[2];
#pragma unroll
        for (int kb = 0; kb < 2; kb++) A2[kb] = *(const v8bf16 *)(&s_a1[w][col][32 * kb + 8 * grp]);
        __builtin_amdgcn_wave_barrier();
        // layer 2 + output layer
        float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < 4; t++) {
            v4f z = {0.f, 0.f, 0.f, 0.f};
            z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A2[0], W2B[t][0], z, 0, 0, 0);
            z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A2[1], W2B[t][1], z, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; r++) part[r] += bf16r(silu(z[r] + b2v[t])) * w3v[t];
        }
        // sum over the 16 lanes of each row group (the j index)
#pragma unroll
        for (int r = 0; r < 4; r++) {
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) part[r] += __shfl_xor(part[r], o);
        }
        if (col == 0) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t q = t0 + 4 * grp + r;
                if (q < v.n) h_rows[v.perm[q]] = __expf(part[r] + b3 + v.logcell);
            }
        }
    }
}

// training forward + backward for 64 samples per block (one wavefront);
// gradients of the mean squared error over a batch of nb_total samples
__global__ __launch_bounds__(64) void k_mlp_train(const float *__restrict__ F, const float *__restrict__ T,
                                                 int64_t nb, int64_t nb_total, const float *__restrict__ P,
                                                 const float *__restrict__ mu, const float *__restrict__ sd,
                                                 float *__restrict__ grad, float *__restrict__ loss) {
    __shared__ float sW1[NH][NIN], sW2[NH][NH], sb1[NH], sb2[NH], sw3[NH];
    __shared__ float sx[WAVE][NIN + 1], sa1[WAVE][NH + 1], sdz2[WAVE][NH + 1], sdz1[WAVE][NH + 1];
    const int s = threadIdx.x;
    for (int i = s; i < NH * NIN; i += WAVE) (&sW1[0][0])[i] = bf16r(P[OFF_W1 + i]);
    for (int i = s; i < NH * NH; i += WAVE) (&sW2[0][0])[i] = bf16r(P[OFF_W2 + i]);
    sb1[s] = P[OFF_B1 + s];
    sb2[s] = P[OFF_B2 + s];
    sw3[s] = bf16r(P[OFF_W3 + s]);
    const float b3 = P[OFF_B3];
    __syncthreads();
    const int64_t i0 = (int64_t)blockIdx.x * WAVE + s;
    const bool valid = i0 < nb;
    const int64_t ii = valid ? i0 : nb - 1;
    // phase 1: thread = sample
    float x[NIN];
#pragma unroll
    for (int j = 0; j < NIN; j++) {
        x[j] = bf16r((F[ii * NIN + j] - mu[j]) / sd[j]);
        sx[s][j] = x[j];
    }
    float a1[NH], ds1[NH];
#pragma unroll 4
    for (int o = 0; o < NH; o++) {
        float z = sb1[o];
#pragma unroll
        for (int j = 0; j < NIN; j++) z = fmaf(sW1[o][j], x[j], z);
        a1[o] = bf16r(silu(z));
        ds1[o] = dsilu(z);
        sa1[s][o] = a1[o];
    }
    float y = b3;
    float a2[NH], ds2[NH];
#pragma unroll 4
    for (int o = 0; o < NH; o++) {
        float z = sb2[o];
#pragma unroll
        for (int i = 0; i < NH; i++) z = fmaf(sW2[o][i], a1[i], z);
        a2[o] = bf16r(silu(z));
        ds2[o] = dsilu(z);
        y = fmaf(sw3[o], a2[o], y);
    }
    const float r = valid ? y - T[ii] : 0.f;
    const float dy = 2.0f * r / (float)nb_total;
    float l2 = r * r / (float)nb_total;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) l2 += __shfl_xor(l2, o);
    if (s == 0) atomicAdd(loss, l2);
    // output layer gradients (reduced over the 64 samples)
    float gb3 = dy;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) gb3 += __shfl_xor(gb3, o);
    if (s == 0) atomicAdd(grad + OFF_B3, gb3);
    for (int o = 0; o < NH; o++) {
        float g = dy * a2[o];
#pragma unroll
        for (int m = 32; m > 0; m >>= 1) g += __shfl_xor(g, m);
        if (s == 0) atomicAdd(grad + OFF_W3 + o, g);
    }
    float dz2[NH];
#pragma unroll
    for (int o = 0; o < NH; o++) {
        dz2[o] = dy * sw3[o] * ds2[o];
        sdz2[s][o] = dz2[o];
    }
#pragma unroll 4
    for (int i = 0; i < NH; i++) {
        float da = 0.f;
#pragma unroll
        for (int o = 0; o < NH; o++) da = fmaf(sW2[o][i], dz2[o], da);
        sdz1[s][i] = da * ds1[i];
    }
    __syncthreads();
    // phase 2: thread = output unit o
    const int o = s;
    float gb2 = 0.f, gb1 = 0.f;
    float gW2[NH], gW1[NIN];
#pragma unroll
    for (int i = 0; i < NH; i++) gW2[i] = 0.f;
#pragma unroll
    for (int j = 0; j < NIN; j++) gW1[j] = 0.f;
    for (int m = 0; m < WAVE; m++) {
        const float d2 = sdz2[m][o], d1 = sdz1[m][o];
        gb2 += d2;
        gb1 += d1;
#pragma unroll
        for (int i = 0; i < NH; i++) gW2[i] = fmaf(d2, sa1[m][i], gW2[i]);
#pragma unroll
        for (int j = 0; j < NIN; j++) gW1[j] = fmaf(d1, sx[m][j], gW1[j]);
    }
    atomicAdd(grad + OFF_B2 + o, gb2);
    atomicAdd(grad + OFF_B1 + o, gb1);
    for (int i = 0; i < NH; i++) atomicAdd(grad + OFF_W2 + o * NH + i, gW2[i]);
    for (int j = 0; j < NIN; j++) atomicAdd(grad + OFF_W1 + o * NIN + j, gW1[j]);
}

// Adam on the float32 master weights; clears the gradient for the next step
__global__ void k_mlp_adam(float *__restrict__ P, float *__restrict__ m, float *__restrict__ v,
                           float *__restrict__ grad, int step, float lr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - powf(b1, (float)step));
    const float vh = vi / (1.0f - powf(b2, (float)step));
    P[i] -= lr * mh / (sqrtf(vh) + eps);
    grad[i] = 0.f;
}

// inference blob from the float32 parameters and the input normalisation
__global__ void k_mlp_pack(const float *__restrict__ P, const float *__restrict__ mu,
                           const float *__restrict__ sd, unsigned char *__restrict__ blob) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    __bf16 *W1b = (__bf16 *)(blob + IB_W1);
    __bf16 *W2b = (__bf16 *)(blob + IB_W2);
    float *fb = (float *)(blob + IB_F);
    if (i < NH * 32) {
        const int o = i / 32, j = i % 32;
        W1b[i] = (__bf16)(j < NIN ? P[OFF_W1 + o * NIN + j] : 0.f);
    }
    if (i < NH * NH) W2b[i] = (__bf16)P[OFF_W2 + i];
    if (i < NH) {
        fb[IF_B1 + i] = P[OFF_B1 + i];
        fb[IF_B2 + i] = P[OFF_B2 + i];
        fb[IF_W3 + i] = bf16r(P[OFF_W3 + i]);
    }
    if (i == 0) fb[IF_B3] = P[OFF_B3];
    if (i < NIN) {
        fb[IF_MU + i] = mu[i];
        fb[IF_SD + i] = sd[i];
    }
}

static SeedView seed_view(const mhip_index *ix, const uint32_t *counts) {
    SeedView v;
    v.sx = ix->sx;
    v.sy = ix->sy;
    v.sz = ix->sz;
    v.perm = ix->perm;
    v.n = ix->n;
    v.dim = ix->dim;
    v.per = ix->box > 0;
    for (int c = 0; c < 3; c++) v.org[c] = ix->frame[c];
    v.scale = ix->frame[3];
    const double ext = ix->frame[3] > 0 ? 2097151.0 / ix->frame[3] : 1.0;
    v.logcell = (float)log(ext / SEED_G);
    v.counts = counts;
    return v;
}

static int seed_counts(mhip_ctx *ctx, const mhip_index *ix, uint32_t **counts) {
    int rc = 0;
    const int64_t ncell = (int64_t)SEED_G * SEED_G * SEED_G;
    uint32_t *c = (uint32_t *)scratch(ctx, ncell * 4, &rc);
    if (!c) return rc;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, ctx->stream, c,
                       ncell, 0u);
    SeedView v = seed_view(ix, c);
    hipLaunchKernelGGL(k_seed_hist, dim3((unsigned)((ix->n + 255) / 256)), dim3(256), 0, ctx->stream, v, c);
    *counts = c;
    return check_hip(ctx, hipGetLastError(), "seed histogram");
}

int seed_features(mhip_ctx *ctx, const mhip_index *ix, float *F) {
    if (!ix || ix->n < 1) return fail(ctx, -1, "index has not been built");
    ProfRegion pr(ctx, "seed_features");
    uint32_t *c;
    int rc = seed_counts(ctx, ix, &c);
    if (rc) return rc;
    SeedView v = seed_view(ix, c);
    hipLaunchKernelGGL(k_seed_features, dim3((unsigned)((ix->n + 255) / 256)), dim3(256), 0, ctx->stream, v, F);
    return check_hip(ctx, hipGetLastError(), "seed features");
}

int seed_predict(mhip_ctx *ctx, const mhip_index *ix, const void *blob, float *h_rows) {
    if (!ix || ix->n < 1) return fail(ctx, -1, "index has not been built");
    ProfRegion pr(ctx, "seed_predict");
    uint32_t *c;
    int rc = seed_counts(ctx, ix, &c);
    if (rc) return rc;
    SeedView v = seed_view(ix, c);
    const int64_t nw = (ix->n + WAVE - 1) / WAVE;
    hipLaunchKernelGGL(k_seed_infer, dim3((unsigned)((nw + 3) / 4)), dim3(256), 0, ctx->stream, v,
                       (const unsigned char *)blob, h_rows);
    return check_hip(ctx, hipGetLastError(), "seed inference");
}

int mlp_train_step(mhip_ctx *ctx, const float *F, const float *T, int64_t nb, const float *mu,
                   const float *sd, float *P, float *m, float *v, float *grad, float *loss, int step,
                   float lr) {
    if (nb < 1) return fail(ctx, -1, "empty batch");
    ProfRegion pr(ctx, "mlp_train_step");
    hipLaunchKernelGGL(k_mlp_train, dim3((unsigned)((nb + WAVE - 1) / WAVE)), dim3(WAVE), 0, ctx->stream, F,
                       T, nb, nb, P, mu, sd, grad, loss);
    if (step > 0)
        hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream, P, m, v,
                           grad, step, lr);
    return check_hip(ctx, hipGetLastError(), "mlp train step");
}

int mlp_pack(mhip_ctx *ctx, const float *P, const float *mu, const float *sd, void *blob) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((NH * NH + 255) / 256), dim3(256), 0, ctx->stream, P, mu, sd,
                       (unsigned char *)blob);
    return check_hip(ctx, hipGetLastError(), "mlp pack");
}

size_t mlp_blob_bytes() { return IB_BYTES; }

}  // namespace mhip
