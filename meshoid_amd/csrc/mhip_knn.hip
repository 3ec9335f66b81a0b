// Exact k-nearest-neighbour search, smoothing length and density on the GPU.
//
// Replaces, for the hot path, the reference's
//   cKDTree.query(pos[mask], des_ngb)           src/meshoid/Meshoid.py:157
//   HsmlIter(ngbdist, error_norm=1e-13, dim)     src/meshoid/Meshoid.py:165,
//                                                src/meshoid/kernel_density.py:14-57
//   density = des_ngb*m/(volnorm*h^dim)          src/meshoid/Meshoid.py:169-170
//
// Design (one lane per query; 64 consecutive queries in Hilbert order per
// wavefront; the tree is a complete 8-ary tree in heap numbering, root 0,
// children of x at 8x+1..8x+8, leaves of 16 particles on the last level):
//   main    Every lane keeps the k+1 smallest float32 squared distances of
//           its query sorted in VGPRs; an insertion is one in-place v_med3
//           per slot.  (1) The leaves holding the wave's own queries are
//           scanned first.  (2) A wave-uniform depth-first walk without a
//           stack: expanding a node tests its 8 children (one contiguous
//           block of boxes, scalar loads) against every lane's bound and
//           keeps the survivors as an 8-bit mask per level in SGPRs; the
//           child nearest (in index) to the path of the wave's middle query
//           goes first.  (3) A leaf's 16 candidates (scalar loads) go to each
//           lane's LDS queue row qn; qn advances when the candidate is under
//           the lane's (k+1)-th distance.  Queues drain into the sorted lists
//           in uniform counted loops.  Leaves where some lane queued a
//           candidate are recorded.  (4) Epilogue: smoothing length
//           (safeguarded float32 Newton, then one float64 Newton step over
//           the final members), density, volume,
//           and the member set {d2 <= tau}: each lane tests the recorded
//           leaves against its own final radius and scans only those, one
//           leaf per lane per step (per-lane loads); the indices are staged
//           in LDS and written as whole rows.  Queries whose k-th and
//           (k+1)-th distances are closer than float32 rounding can order
//           are flagged.  A wavefront whose leaf record overflows walks
//           the tree again with every lane's final radius and collects the
//           members in batches of recorded leaves.
//   resolve for flagged queries only: members in the rounding band around tau
//           are re-ranked with float64 distances from the caller's positions,
//           so neighbour sets equal the float64 kNN except for float64 ties.
#include "../../include/meshoid_hip.h"
#include "mhip_internal.hpp"
#include "mhip_sortnet.hpp"

namespace mhip {

constexpr int KNN_BLOCK = 256;
constexpr int KNN_WAVES = KNN_BLOCK / WAVE;
constexpr int QCAP = 32;                     // per-lane pending-candidate queue
constexpr int WLCAP = 384;                   // per-wave record of candidate leaves
constexpr float U32 = 5.9604645e-8f;         // 2^-24
constexpr float AMBIG_REL = 12.0f * U32;     // fp32 vs fp64 ordering ambiguity of d^2
constexpr float BAND_REL = 16.0f * U32;      // resolver band half-width
constexpr float PRUNE_REL = 1.0f + 1e-5f;    // box-test safety factor
constexpr int BAND_CAP = 32;
constexpr int64_t RESOLVE_WAVES = 8192;      // resolver grid (one wavefront per ambiguous slot)

struct TreeView {
    const float *sx, *sy, *sz, *lx, *ly, *lz;
    const float *nodes;
    int64_t first_leaf, nleaf;
    float box, half;  // periodic box length and half of it (0: open)
    float slack;      // absolute box-test slack: 0 for float32 inputs, covers
                      // the float32 residuals of float64 inputs otherwise
    int depth;        // level of the leaves (root = level 0)
};

struct QPos {
    float x, y, z, lx, ly, lz;
};

template <bool P64>
MHIP_DEV QPos load_q(const TreeView &t, uint32_t p) {
    QPos q;
    q.x = t.sx[p];
    q.y = t.sy[p];
    q.z = t.sz[p];
    q.lx = P64 ? t.lx[p] : 0.f;
    q.ly = P64 ? t.ly[p] : 0.f;
    q.lz = P64 ? t.lz[p] : 0.f;
    return q;
}

// Offset c - q along one axis.  Periodic: the nearest image is formed by
// moving whichever coordinate lies in the far half by one box length; that
// subtraction (x - box with box/2 <= x < box) is exact (Sterbenz), so the
// wrapped offset carries a single rounding, like an unwrapped one.
template <bool PER>
MHIP_DEV float axis_off(float c, float q, float box, float half) {
    const float d = c - q;
    if (!PER) return d;
    const float up = (c - box) - q;  // image of c one box lower (c in the far half)
    const float dn = c - (q - box);  // image of c one box higher (q in the far half)
    return d > half ? up : (d < -half ? dn : d);
}

// squared distance, float32, fixed operation order (shared by every pass so
// that the same pair always yields the same bits)
template <bool PER, bool P64>
MHIP_DEV float dist2(const TreeView &t, const QPos &q, float cx, float cy, float cz, float clx,
                     float cly, float clz) {
    float dx = axis_off<PER>(cx, q.x, t.box, t.half);
    float dy = axis_off<PER>(cy, q.y, t.box, t.half);
    float dz = axis_off<PER>(cz, q.z, t.box, t.half);
    if (P64) {
        dx = dx + (clx - q.lx);
        dy = dy + (cly - q.ly);
        dz = dz + (clz - q.lz);
    }
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

template <bool PER>
MHIP_DEV float axis_gap(float q, float lo, float hi, float box, float slack) {
    float g = fmaxf(lo - q, q - hi);
    if (PER) {
        // the box one period lower / higher
        g = fminf(g, fmaxf((lo - box) - q, q - (hi - box)));
        g = fminf(g, fmaxf(lo - (q - box), (q - box) - hi));
    }
    return fmaxf(g - slack, 0.f);
}

// lower bound of the squared distance from q to any particle under node x
template <bool PER>
MHIP_DEV float node_d2(const TreeView &t, const QPos &q, uint32_t x) {
    const float4 *nb = (const float4 *)(t.nodes + (int64_t)x * 8);
    const float4 a = nb[0], b = nb[1];  // lo xyz, hi x | hi yz
    const float dx = axis_gap<PER>(q.x, a.x, a.w, t.box, t.slack);
    const float dy = axis_gap<PER>(q.y, a.y, b.x, t.box, t.slack);
    const float dz = axis_gap<PER>(q.z, a.z, b.y, t.box, t.slack);
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// next node in depth-first order after the subtree of x is finished; false
// at the end of the traversal (children of p are 8p+1..8p+8)
MHIP_DEV bool next_after(uint32_t &x) {
    while (x != 0 && (x & (FANOUT - 1)) == 0) x = (x - 1) >> 3;
    if (x == 0) return false;
    x++;
    return true;
}

// one leaf's candidates of this lane (per-lane gathers, 64-byte aligned)
template <bool P64>
MHIP_DEV void load_leaf(const TreeView &t, int64_t leaf, float (&cx)[LEAF], float (&cy)[LEAF],
                        float (&cz)[LEAF], float (&clx)[LEAF], float (&cly)[LEAF], float (&clz)[LEAF]) {
    const int64_t base = leaf * LEAF;
#pragma unroll
    for (int v = 0; v < LEAF / 4; v++) {
        const float4 X = ((const float4 *)(t.sx + base))[v];
        const float4 Y = ((const float4 *)(t.sy + base))[v];
        const float4 Z = ((const float4 *)(t.sz + base))[v];
        cx[4 * v] = X.x, cx[4 * v + 1] = X.y, cx[4 * v + 2] = X.z, cx[4 * v + 3] = X.w;
        cy[4 * v] = Y.x, cy[4 * v + 1] = Y.y, cy[4 * v + 2] = Y.z, cy[4 * v + 3] = Y.w;
        cz[4 * v] = Z.x, cz[4 * v + 1] = Z.y, cz[4 * v + 2] = Z.z, cz[4 * v + 3] = Z.w;
        if (P64) {
            const float4 LX = ((const float4 *)(t.lx + base))[v];
            const float4 LY = ((const float4 *)(t.ly + base))[v];
            const float4 LZ = ((const float4 *)(t.lz + base))[v];
            clx[4 * v] = LX.x, clx[4 * v + 1] = LX.y, clx[4 * v + 2] = LX.z, clx[4 * v + 3] = LX.w;
            cly[4 * v] = LY.x, cly[4 * v + 1] = LY.y, cly[4 * v + 2] = LY.z, cly[4 * v + 3] = LY.w;
            clz[4 * v] = LZ.x, clz[4 * v + 1] = LZ.y, clz[4 * v + 2] = LZ.z, clz[4 * v + 3] = LZ.w;
        } else {
            clx[4 * v] = clx[4 * v + 1] = clx[4 * v + 2] = clx[4 * v + 3] = 0.f;
            cly[4 * v] = cly[4 * v + 1] = cly[4 * v + 2] = cly[4 * v + 3] = 0.f;
            clz[4 * v] = clz[4 * v + 1] = clz[4 * v + 2] = clz[4 * v + 3] = 0.f;
        }
    }
}

// Inserts x into the ascending list A, dropping the largest entry: slot j
// becomes med3(A[j-1], x, A[j]).  The slots are updated in place from the top
// down (each instruction's destination is its own old value), so the list
// keeps its registers across loop iterations instead of alternating between
// two register sets joined by copies.  x is never NaN.
template <int KCAP>
MHIP_DEV void insert_sorted(float (&A)[KCAP], float x) {
#pragma unroll
    for (int j = KCAP - 1; j > 0; --j)
        asm volatile("v_med3_f32 %0, %1, %2, %0" : "+v"(A[j]) : "v"(A[j - 1]), "v"(x));
    asm volatile("v_min_f32 %0, %1, %0" : "+v"(A[0]) : "v"(x));
}

MHIP_DEV int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return (int)uniform((uint32_t)v);
}

MHIP_DEV uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

struct KnnArgs {
    TreeView t;
    const uint32_t *qsorted;  // null: slot q is sorted particle q
    const uint32_t *qrow;     // null: row = perm[q]
    const uint32_t *perm;
    int64_t nq, n;
    int k, dim;
    const float *seed_h;   // Newton seed per output row, null = none
    const float *seed_r;   // search-radius bound per output row, null = none
    int redo;              // 1: second launch, redo waves with a radius-seed miss
    int32_t *status_out;
    // per slot state
    float *tau;
    int32_t *nlt;
    float *h_slot;
    double *h64_slot;
    int32_t *st_slot;
    uint32_t *nl_idx;
    float *nl_d2;
    unsigned long long *stats;  // work counters, may be null
    uint32_t *amb;              // [0] count, [1..] ambiguous slots
    uint32_t wlcap;             // leaf-record capacity used (<= WLCAP; smaller only as a test hook)
};

MHIP_DEV int64_t slot_of_lane(int64_t *first) {
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t wave0 = ((int64_t)b * KNN_WAVES + (threadIdx.x >> 6)) * WAVE;
    *first = wave0;
    return wave0 + (threadIdx.x & (WAVE - 1));
}

MHIP_DEV void pad_list(const KnnArgs &a, int cnt, uint32_t qp, int64_t slot) {
    // never leave undefined indices behind: pad with the query itself
    for (int j = cnt; j < a.k; j++) {
        a.nl_idx[(int64_t)j * a.nq + slot] = qp;
        if (a.nl_d2) a.nl_d2[(int64_t)j * a.nq + slot] = INFINITY;
    }
}

// One float64 Newton step on the smoothing length over the final member
// list, member(j) giving the sorted index of member j < k.  The float32
// solve below is at the float32 noise floor (relative error up to ~2.4e-7,
// from the float32 distances); the derivative fits are sensitive to h through
// the weights of the outermost neighbours (d ln W / d ln h = 3q / (1 - q) as
// q -> 1), so a 2.4e-7 error of h moves second-order derivatives by up to
// ~6e-5, and first-order derivatives of rough fields by ~4e-5, on
// well-conditioned rows.  One step on the exact float64 pair distances
// (float32 positions are exact in float64; float64 inputs are hi + lo) takes
// h to ~1e-13 of the root of norm_d sum_j W(r_j / h) = k
// (kernel_density.py:36-56).  A step beyond 1e-4 h means the float32 solve
// did not converge; h is then kept.
template <bool PER, bool P64, typename F>
MHIP_DEV double refine_h64(const KnnArgs &a, const QPos &q, double hd, F member) {
    const double qx = (double)q.x + (double)q.lx, qy = (double)q.y + (double)q.ly, qz = (double)q.z + (double)q.lz;
    const double box = (double)a.t.box, ih = 1.0 / hd;
    double sw = 0.0, sd = 0.0;
    for (int j = 0; j < a.k; j++) {
        const uint32_t p = member(j);
        double d[3] = {(double)a.t.sx[p] - qx, (double)a.t.sy[p] - qy, (double)a.t.sz[p] - qz};
        if (P64) {
            d[0] += (double)a.t.lx[p];
            d[1] += (double)a.t.ly[p];
            d[2] += (double)a.t.lz[p];
        }
        if (PER) {
#pragma unroll
            for (int c = 0; c < 3; c++) d[c] = wrap<double>(d[c], box);
        }
        const double qq = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * ih;
        double w, dw;
        kernel_w_dw<double>(qq, w, dw);
        sw += w;
        sd += qq <= 1.0 ? dw * qq : 0.0;
    }
    const double norm = hsml_norm(a.dim);
    // n(h) grows with h: dn/dh = -norm sum_j W'(q_j) q_j / h > 0
    const double fn = norm * sw - (double)a.k, dndh = -norm * sd * ih;
    const double hn = hd - fn / dndh;
    return (dndh > 0.0 && fabs(hn - hd) <= 1e-4 * hd) ? hn : hd;
}

// smoothing length of one query slot (float32 copy and float64); the
// row-order h, density and volume are formed from it by k_rows_out
MHIP_DEV void write_h(const KnnArgs &a, int64_t slot, double h) {
    a.h_slot[slot] = (float)h;
    a.h64_slot[slot] = h;
}

// ---------------------------------------------------------------------------
// smoothing length on the k nearest distances (kernel_density.py:36-56 root):
// norm * sum_j W(r_j/h) = k, bracket [r_1, c_d r_{k-1}].  Safeguarded Newton
// with compensated float32 sums; the bracket is the fallback.  On return
// A[0..KCAP-2] hold the distances (sentinel slots +inf), not their squares:
// the list is not needed afterwards, and reusing its registers keeps the
// kernel's register count at that of the traversal.
// ---------------------------------------------------------------------------
template <int KCAP>
MHIP_DEV float hsml_solve(float (&A)[KCAP], int k, int dim, float seed, int32_t &st, int &iters) {
    const int e = KCAP - 1 - k;  // first real slot
    const float norm = (float)hsml_norm(dim);
    float(&r)[KCAP] = A;
    float lo = 0.f, rmax = 0.f;
#pragma unroll
    for (int j = 0; j < KCAP - 1; j++) {
        r[j] = A[j] >= 0.f ? __builtin_sqrtf(A[j]) : INFINITY;
        if (j == e + 1) lo = r[j];
        if (j == KCAP - 2) rmax = r[j];
    }
    float hi = (float)hsml_bound_coeff(dim) * rmax;
    iters = 0;
    if (!(hi > 0.f)) return 0.f;  // all k at distance 0: the reference returns 0
    float h = rmax;
    if (seed > 0.f) {
        if (seed > lo && seed < hi)
            h = seed;
        else
            st |= ST_SEED_MISS;
    }
    const float kf = (float)k;
    int it = 0;
    for (; it < 40; it++) {
        const float ih = 1.0f / h;
        float s = 0.f, c = 0.f, ds = 0.f;
#pragma unroll
        for (int j = 0; j < KCAP - 1; j++) {
            float q = r[j] * ih, w, dw;
            kernel_w_dw(q, w, dw);
            float y = w - c, tt = s + y;
            c = (tt - s) - y;
            s = tt;
            ds += q <= 1.0f ? dw * q : 0.0f;
            // keep the terms in order: evaluating all k kernel values ahead
            // of the sums would hold them all in registers at once
            asm volatile("" : "+v"(s), "+v"(c), "+v"(ds));
        }
        const float f = norm * s - kf;
        if (f == 0.f) break;
        if (f > 0.f)
            hi = h;
        else
            lo = h;
        const float dndh = -norm * ds * ih;
        const float step = f / dndh;
        // n(h) is C^2 (W is C^2 at q = 1/2 and q = 1), so Newton converges
        // quadratically: after a step below 3e-4 h the error of h - step is
        // at the float32 noise floor (~3e-7 h with the compensated sums), and
        // the float64 Newton step over the final members that follows
        // (refine_h64) takes it to ~2e-13 h.  Such a step ends the
        // solve even when it rounds onto the bracket end that h itself has
        // just become (otherwise a converged lane would bisect the whole
        // bracket again and hold its wavefront for ~15 more iterations).  A
        // collapsed bracket also ends the solve.
        if (fabsf(step) <= 3e-4f * h) {
            h = fminf(fmaxf(h - step, lo), hi);
            break;
        }
        float hn = h - step;
        if (!(hn > lo && hn < hi)) hn = 0.5f * (lo + hi);
        if (hi - lo <= 2.5e-7f * h) {
            h = hn;
            break;
        }
        h = hn;
    }
    if (it >= 40) st |= ST_HSML_NOCONV;
    iters = min(it + 1, 40);
    return h;
}

// Read-only data read at wave-uniform addresses goes through the constant
// address space, which the compiler serves with scalar (SMEM) loads into
// SGPRs: one load per wavefront instead of 64 lanes, and no VGPRs.
typedef const float __attribute__((address_space(4))) cfloat;
MHIP_DEV float ldc(const float *p, int64_t i) { return ((cfloat *)p)[i]; }

// one leaf's candidates, wave-uniform address (scalar loads)
template <bool P64>
MHIP_DEV void load_leaf_uniform(const float *sx, const float *sy, const float *sz, const float *lx,
                                const float *ly, const float *lz, int64_t leaf, float (&cx)[LEAF],
                                float (&cy)[LEAF], float (&cz)[LEAF], float (&clx)[LEAF],
                                float (&cly)[LEAF], float (&clz)[LEAF]) {
    const int64_t base = leaf * LEAF;
#pragma unroll
    for (int i = 0; i < LEAF; i++) {
        cx[i] = ldc(sx, base + i);
        cy[i] = ldc(sy, base + i);
        cz[i] = ldc(sz, base + i);
        clx[i] = P64 ? ldc(lx, base + i) : 0.f;
        cly[i] = P64 ? ldc(ly, base + i) : 0.f;
        clz[i] = P64 ? ldc(lz, base + i) : 0.f;
    }
}


// squared gap from a lane's query to a box read at wave-uniform addresses
// (scalar loads); the query interval [ql, qh] per axis is the position widened
// by the float64-residual slack (ql = qh for float32 inputs); zero inside
template <bool PER>
MHIP_DEV float axis_gap_u(float ql, float qh, float lo, float hi, float box) {
    float g = fmaxf(lo - qh, ql - hi);
    if (PER) {
        g = fminf(g, fmaxf((lo - box) - qh, ql - (hi - box)));  // the box one period lower
        g = fminf(g, fmaxf(lo - (qh - box), (ql - box) - hi));  // the box one period higher
    }
    return fmaxf(g, 0.f);
}

struct QBox {
    float lx, ly, lz, hx, hy, hz;
};

template <bool PER>
MHIP_DEV float box_gap2_u(const TreeView &t, const QBox &b, uint32_t x) {
    const int64_t o = (int64_t)x * 8;
    const float gx = axis_gap_u<PER>(b.lx, b.hx, ldc(t.nodes, o), ldc(t.nodes, o + 3), t.box);
    const float gy = axis_gap_u<PER>(b.ly, b.hy, ldc(t.nodes, o + 1), ldc(t.nodes, o + 4), t.box);
    const float gz = axis_gap_u<PER>(b.lz, b.hz, ldc(t.nodes, o + 2), ldc(t.nodes, o + 5), t.box);
    return fmaf(gz, gz, fmaf(gy, gy, gx * gx));
}

// bit c set: child 8x+1+c lies within bound of some lane's query
template <bool PER>
MHIP_DEV uint32_t expand_mask(const TreeView &t, const QBox &b, uint32_t x, float bound) {
    uint32_t m = 0;
    const uint32_t c0 = FANOUT * x + 1;
#pragma unroll
    for (int c = 0; c < FANOUT; c++)
        if (__any(box_gap2_u<PER>(t, b, c0 + c) <= bound)) m |= 1u << c;
    return m;
}

// the set bit of m nearest to position o (m != 0); ties go upwards
MHIP_DEV uint32_t pick_nearest(uint32_t m, uint32_t o) {
    const uint32_t up = m >> o;
    const uint32_t down = m & ((1u << o) - 1u);
    if (!down) return o + __builtin_ctz(up);
    const uint32_t cd = 31u - __builtin_clz(down);
    if (!up) return cd;
    const uint32_t cu = o + __builtin_ctz(up);
    return cu - o <= o - cd ? cu : cd;
}

// 8-bit child masks of the traversal path, levels 0..7 and 8..15
struct MaskStack {
    uint64_t lo = 0, hi = 0;
    MHIP_DEV uint32_t get(int l) const {
        return l < 8 ? (uint32_t)(lo >> (8 * l)) & 0xffu : (uint32_t)(hi >> (8 * (l - 8))) & 0xffu;
    }
    MHIP_DEV void put(int l, uint32_t m) {
        if (l < 8)
            lo |= (uint64_t)m << (8 * l);
        else
            hi |= (uint64_t)m << (8 * (l - 8));
    }
    MHIP_DEV void clear(int l, uint32_t c) {
        if (l < 8)
            lo &= ~(1ull << (8 * l + c));
        else
            hi &= ~(1ull << (8 * (l - 8) + c));
    }
};

constexpr int LROWS = 33;                          // staged member rows (k <= 32)
constexpr int QROWS = QCAP > LROWS ? QCAP : LROWS;  // per-lane LDS rows

template <int KCAP, bool PER, bool P64>
__global__ __launch_bounds__(KNN_BLOCK) void k_knn_main(KnnArgs a) {
    // members are staged in LDS and written as whole rows when they fit
    constexpr bool STAGE = KCAP <= LROWS;
    __shared__ uint32_t s_buf[KNN_WAVES][QROWS][WAVE];
    __shared__ uint32_t s_wl[KNN_WAVES][WLCAP];
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x >> 6;
    int64_t wave0;
    const int64_t slot = slot_of_lane(&wave0);
    if (wave0 >= a.nq) return;
    const bool live = slot < a.nq;
    const int64_t qs_slot = min<int64_t>(slot, a.nq - 1);
    // redo launch: only waves where some lane missed its radius seed
    const bool missed = a.redo && live && (a.st_slot[qs_slot] & ST_RADIUS_MISS);
    if (a.redo && !__any(missed)) return;
    const uint32_t qp = a.qsorted ? a.qsorted[qs_slot] : (uint32_t)qs_slot;
    const uint32_t row = a.qrow ? a.qrow[qs_slot] : a.perm[qs_slot];
    // a radius seed starts the k+1 slots at (seed radius)^2 instead of +inf: a
    // lane whose (k+1)-th slot still holds that bound at the end missed
    const float sr = (a.seed_r && !a.redo) ? a.seed_r[row] : 0.f;
    const float tau0 = sr > 0.f && sr < 3e18f ? sr * sr : INFINITY;
    const TreeView &t = a.t;
    const float *__restrict__ sx = t.sx;
    const float *__restrict__ sy = t.sy;
    const float *__restrict__ sz = t.sz;
    const float *__restrict__ lx = t.lx;
    const float *__restrict__ ly = t.ly;
    const float *__restrict__ lz = t.lz;
    const QPos q = load_q<P64>(t, qp);
    QBox qb;
    qb.lx = P64 ? q.x - t.slack : q.x;
    qb.ly = P64 ? q.y - t.slack : q.y;
    qb.lz = P64 ? q.z - t.slack : q.z;
    qb.hx = P64 ? q.x + t.slack : q.x;
    qb.hy = P64 ? q.y + t.slack : q.y;
    qb.hz = P64 ? q.z + t.slack : q.z;

    // sorted k+1 smallest squared distances; slots below e are -1 sentinels
    // when k + 1 < KCAP, so A[KCAP-2] is the k-th and A[KCAP-1] the (k+1)-th
    float A[KCAP];
    const int e = KCAP - 1 - a.k;
#pragma unroll
    for (int j = 0; j < KCAP; j++) A[j] = j < e ? -1.0f : tau0;
    float cx[LEAF], cy[LEAF], cz[LEAF], clx[LEAF], cly[LEAF], clz[LEAF];
    uint32_t n_ins = 0, n_leaf = 0, n_nodes = 0;
    uint32_t need_lane = 0, parent_max = 0;  // work counters (stats only)

    // (1) the leaves holding this wave's own queries first (they fill every
    // list with a nearby bound), then (2) a wave-uniform depth-first traversal
    // without a stack: expanding a node tests all 8 children against every
    // lane's bound at once and keeps the survivors as an 8-bit mask per level;
    // children nearest (in index) to the path of the wave's middle query are
    // taken first.  Every candidate of a leaf is written to its lane's LDS queue
    // row qn, and qn advances when the candidate is under the lane's current
    // (k+1)-th distance; the queues drain into the sorted lists in batches.
    // Leaves where some lane queued a candidate are recorded for (3).
    const uint32_t p_first = uniform(__shfl(qp, 0));
    const uint32_t p_last = uniform(__shfl(qp, WAVE - 1));
    const int64_t l0 = p_first / LEAF;
    const int64_t l1 = min<int64_t>(p_last / LEAF, l0 + 7);
    float t1 = A[KCAP - 1];
    float bound = t1 * PRUNE_REL;
    int qn = 0;
    uint32_t nwl = 0;  // recorded leaves (uniform)
    const uint32_t first_leaf = (uint32_t)t.first_leaf;
    const int depth = t.depth;
    const uint32_t lc = uniform(__shfl(qp, WAVE / 2 - 1)) / LEAF;
    MaskStack ms;
    int lvl = 0;
    uint32_t x = 0, off = 0;  // node whose children (level lvl + 1) are pending; its offset in its level
    int64_t L = l0;           // next own leaf
    bool walking = false, done = false;
    if (KCAP == 33 && e == 0) {
        // k = 32: the own leaves fill the lists through sorting networks, one
        // leaf at a time (its 16 distances sorted in registers, then merged
        // into the 33-slot list; the merge's other 16 inputs are +inf),
        // instead of one queued insertion step per candidate (64 steps of 34
        // instructions for four own leaves)
        for (int64_t li = l0; li <= l1; li++) {
            load_leaf_uniform<P64>(sx, sy, sz, lx, ly, lz, li, cx, cy, cz, clx, cly, clz);
            float S[32];
#pragma unroll
            for (int i = 0; i < LEAF; i++) {
                const float d2 = dist2<PER, P64>(t, q, cx[i], cy[i], cz[i], clx[i], cly[i], clz[i]);
                S[i] = fminf(d2, INFINITY);  // padding (NaN) -> +inf
            }
            oddeven_sort<LEAF>(*reinterpret_cast<float(*)[LEAF]>(&S[0]));
#pragma unroll
            for (int i = LEAF; i < 32; i++) S[i] = INFINITY;
            merge_into<32>(*reinterpret_cast<float(*)[33]>(&A[0]), S);
            if (nwl < a.wlcap && lane == 0) s_wl[w][nwl] = (uint32_t)li;
            nwl++;
            n_leaf++;
        }
        L = l1 + 1;
        t1 = A[KCAP - 1];
        bound = t1 * PRUNE_REL;
    }
    // batches: scan leaves until some queue is nearly full (the sorted lists
    // are not touched there), then drain every queue into its list
    while (true) {
        while (true) {
            int64_t li = -1;
            if (L <= l1) {
                li = L++;
            } else if (depth > 0) {
                if (!walking) {
                    walking = true;
                    ms.put(0, expand_mask<PER>(t, qb, 0u, bound));
                    n_nodes += FANOUT;
                }
                while (true) {
                    const uint32_t m = ms.get(lvl);
                    if (m == 0u) {
                        if (lvl == 0) break;
                        lvl--;
                        x = (x - 1) >> 3;
                        off >>= 3;
                        continue;
                    }
                    // the child on (or nearest to) the middle query's path first
                    const uint32_t anc = lc >> (3 * (depth - 1 - lvl));
                    const uint32_t base = FANOUT * off;
                    const uint32_t o = anc < base ? 0u : (anc > base + 7 ? 7u : anc - base);
                    const uint32_t c = pick_nearest(m, o);
                    ms.clear(lvl, c);
                    const uint32_t y = FANOUT * x + 1 + c;
                    if (lvl + 1 < depth) {
                        lvl++;
                        x = y;
                        off = base + c;
                        ms.put(lvl, expand_mask<PER>(t, qb, y, bound));
                        n_nodes += FANOUT;
                        if (a.stats && lvl + 1 == depth) {
                            // work counters: leaves this lane itself needs
                            int pc = 0;
                            for (int cc = 0; cc < FANOUT; cc++) {
                                const int64_t lf = (int64_t)(FANOUT * y + 1 + cc - first_leaf);
                                pc += box_gap2_u<PER>(t, qb, FANOUT * y + 1 + cc) <= bound &&
                                      !(lf >= l0 && lf <= l1) && lf < t.nleaf;
                            }
                            need_lane += pc;
                            parent_max += wave_max(pc);
                        }
                        continue;
                    }
                    const int64_t cand = (int64_t)(y - first_leaf);
                    // own leaves are done; empty leaves pass the box test
                    // only while some list is not full
                    if ((cand >= l0 && cand <= l1) || cand >= t.nleaf) continue;
                    li = cand;
                    break;
                }
            }
            if (li < 0) {
                done = true;
                break;
            }
            n_leaf++;
            load_leaf_uniform<P64>(sx, sy, sz, lx, ly, lz, li, cx, cy, cz, clx, cly, clz);
            const int before = qn;
#pragma unroll
            for (int i = 0; i < LEAF; i++) {
                const float d2 = dist2<PER, P64>(t, q, cx[i], cy[i], cz[i], clx[i], cly[i], clz[i]);
                // qn <= QCAP - LEAF when a leaf starts, so the row stays in range
                s_buf[w][qn][lane] = __float_as_uint(d2);
                qn += d2 < t1;
            }
            if (__any(qn > before)) {
                if (nwl < a.wlcap && lane == 0) s_wl[w][nwl] = (uint32_t)li;
                nwl++;
            }
            if (__any(qn > QCAP - LEAF)) break;
        }
        const int nd = wave_max(qn);
        for (int i = 0; i < nd; i++) {
            // rows below QCAP are always in range; rows at or past qn are stale
            const float r = __uint_as_float(s_buf[w][i][lane]);
            const float v = i < qn ? r : INFINITY;
            // rows that no lane's list takes any more are skipped
            if (__any(v < A[KCAP - 1])) {
                insert_sorted<KCAP>(A, v);
                n_ins++;
            }
        }
        qn = 0;
        t1 = A[KCAP - 1];
        bound = t1 * PRUNE_REL;
        if (done) break;
    }
    __builtin_amdgcn_wave_barrier();
    if (a.stats && lane == 0) {
        atomicAdd(a.stats + STAT_P1_LEAVES, (unsigned long long)n_leaf);
        atomicAdd(a.stats + STAT_P1_NODES, (unsigned long long)n_nodes);
        atomicAdd(a.stats + STAT_P1_INSERTS, (unsigned long long)n_ins);
        atomicAdd(a.stats + STAT_P2_LEAVES, (unsigned long long)nwl);
        atomicAdd(a.stats + STAT_WAVES, 1ull);
        atomicAdd(a.stats + STAT_PARENT_MAX, (unsigned long long)parent_max);
    }
    if (a.stats) {
        const uint32_t tot = wave_sum(need_lane), mx = (uint32_t)wave_max((int)need_lane);
        if (lane == 0) {
            atomicAdd(a.stats + STAT_LANE_NEED, (unsigned long long)tot);
            atomicAdd(a.stats + STAT_MAXLANE_NEED, (unsigned long long)mx);
        }
    }

    // (3) epilogue
    int32_t st = missed ? ST_RADIUS_MISS : 0;
    const float tau = A[KCAP - 2], tau1 = A[KCAP - 1];
    if (tau0 < INFINITY && tau1 >= tau0) {
        // radius seed missed: the redo launch recomputes this wave plainly
        if (live) a.st_slot[slot] = ST_RADIUS_MISS;
        return;
    }
    if (tau1 <= tau * (1.0f + AMBIG_REL)) st |= ST_AMBIGUOUS;
    int nlt = 0;
#pragma unroll
    for (int j = 0; j < KCAP - 1; j++) nlt += (A[j] >= 0.f && A[j] < tau);
    const float seed = a.seed_h ? a.seed_h[row] : 0.f;
    int n_newton = 0;
    const float h = hsml_solve<KCAP>(A, a.k, a.dim, seed, st, n_newton);
    if (a.stats) {
        // a redo launch counts only the lanes it recomputes for a miss
        const bool cnt = live && (!a.redo || missed);
        const uint32_t tot = wave_sum(cnt ? (uint32_t)n_newton : 0u);
        const uint32_t mx = (uint32_t)wave_max(cnt ? n_newton : 0);
        if (lane == 0) {
            atomicAdd(a.stats + STAT_NEWTON, (unsigned long long)tot);
            atomicAdd(a.stats + STAT_NEWTON_WMAX, (unsigned long long)mx);
        }
    }
    if (live) {
        a.tau[slot] = tau;
        a.nlt[slot] = nlt;
    }
    double h64 = (double)h;
    // members: a candidate at or below tau passed its lane's queue filter
    // when its leaf was scanned, so it lies in a recorded leaf; leaves no
    // final ball reaches are skipped
    {
        // Unless the row is ambiguous, exactly k particles have d2 <= tau
        // (the (k+1)-th distance is strictly larger), so "d2 <= tau" selects
        // the members.  Ambiguous rows (float32 ties at tau included) are
        // rebuilt by the float64 resolver; for them the count is only capped.
        // Each lane scans only the recorded leaves its own final ball reaches,
        // 64 recorded leaves at a time: a step scans one leaf per lane (its
        // particles loaded by the lane itself), so the wave pays for the
        // busiest lane, not for every leaf some lane needs.
        int cnt = 0;
        uint32_t nscan = 0;
        const float tb = tau * PRUNE_REL;
        auto collect = [&](uint32_t nrec) {
        for (uint32_t r0 = 0; r0 < nrec; r0 += 64) {
            const uint32_t nr = min(64u, nrec - r0);
            uint32_t nlo = 0, nhi = 0;
            for (uint32_t r = 0; r < nr; r++) {
                const uint32_t li = uniform(s_wl[w][r0 + r]);
                if (box_gap2_u<PER>(t, qb, first_leaf + li) <= tb) {
                    if (r < 32)
                        nlo |= 1u << r;
                    else
                        nhi |= 1u << (r - 32);
                }
            }
            while (__any((nlo | nhi) != 0u)) {
                const bool act = (nlo | nhi) != 0u;
                const uint32_t r = nlo ? (uint32_t)__builtin_ctz(nlo) : (nhi ? 32u + (uint32_t)__builtin_ctz(nhi) : 0u);
                if (nlo)
                    nlo &= nlo - 1u;
                else
                    nhi &= nhi - 1u;
                const uint32_t li = s_wl[w][r0 + r];
                nscan++;
                load_leaf<P64>(t, li, cx, cy, cz, clx, cly, clz);
#pragma unroll
                for (int i = 0; i < LEAF; i++) {
                    const float d2 = dist2<PER, P64>(t, q, cx[i], cy[i], cz[i], clx[i], cly[i], clz[i]);
                    const bool take = act & (d2 <= tau) & (cnt < a.k);
                    if (a.nl_d2 && take && live) a.nl_d2[(int64_t)cnt * a.nq + slot] = d2;
                    if (STAGE) {
                        // row cnt <= k < LROWS; the row is kept only on a hit
                        s_buf[w][cnt][lane] = li * LEAF + i;
                    } else if (take && live) {
                        a.nl_idx[(int64_t)cnt * a.nq + slot] = li * LEAF + i;
                    }
                    cnt += take;
                }
            }
        }
        };
        if (nwl <= a.wlcap) {
            collect(nwl);
        } else {
            // the leaf record overflowed (a wavefront whose queries spread
            // over a large region): walk the tree again, pruned by every
            // lane's final radius, and collect in batches of wlcap leaves
            MaskStack ws;
            int wl = 0;
            uint32_t wx = 0, nrec = 0;
            ws.put(0, expand_mask<PER>(t, qb, 0u, tb));
            while (true) {
                const uint32_t m = ws.get(wl);
                if (m == 0u) {
                    if (wl == 0) break;
                    wl--;
                    wx = (wx - 1) >> 3;
                    continue;
                }
                const uint32_t c = (uint32_t)__builtin_ctz(m);
                ws.clear(wl, c);
                const uint32_t y = FANOUT * wx + 1 + c;
                if (wl + 1 < depth) {
                    wl++;
                    wx = y;
                    ws.put(wl, expand_mask<PER>(t, qb, y, tb));
                    continue;
                }
                const int64_t cand = (int64_t)(y - first_leaf);
                if (cand >= t.nleaf) continue;
                if (lane == 0) s_wl[w][nrec] = (uint32_t)cand;
                if (++nrec == a.wlcap) {
                    __builtin_amdgcn_wave_barrier();
                    collect(nrec);
                    nrec = 0;
                    __builtin_amdgcn_wave_barrier();
                }
            }
            __builtin_amdgcn_wave_barrier();
            collect(nrec);
            if (a.stats && lane == 0) atomicAdd(a.stats + STAT_OVERFLOW_WAVES, 1ull);
        }
        if (a.stats && lane == 0) atomicAdd(a.stats + STAT_P2_NODES, (unsigned long long)nscan);
        if (STAGE && live) {
            for (int j = 0; j < a.k; j++) a.nl_idx[(int64_t)j * a.nq + slot] = j < cnt ? s_buf[w][j][lane] : qp;
        }
        if (live && cnt != a.k) {
            st |= ST_COUNT;
            if (!STAGE) pad_list(a, cnt, qp, slot);
            else if (a.nl_d2)
                for (int j = cnt; j < a.k; j++) a.nl_d2[(int64_t)j * a.nq + slot] = INFINITY;
        }
        // float64 refinement over the members just collected; ambiguous rows
        // (whose capped collection may hold the wrong ties) are refined by
        // the resolver once it has rebuilt them
        if (live && cnt == a.k && h > 0.f && !(st & (ST_HSML_NOCONV | ST_AMBIGUOUS))) {
            if (STAGE)
                h64 = refine_h64<PER, P64>(a, q, h64, [&](int j) { return s_buf[w][j][lane]; });
            else
                h64 = refine_h64<PER, P64>(a, q, h64, [&](int j) { return a.nl_idx[(int64_t)j * a.nq + slot]; });
        }
    }
    if (live) {
        write_h(a, slot, h64);
        a.st_slot[slot] = st;
    }
    // ambiguous queries for the resolver, one vector atomic per wavefront by
    // its lowest active lane (lanes that missed a radius seed left earlier;
    // in the redo launch only the recomputed lanes are new)
    const bool amb = live && (st & ST_AMBIGUOUS) && (!a.redo || missed);
    const uint64_t am = __ballot(amb);
    if (am) {
        const int leader = __ffsll((long long)__ballot(1)) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(a.amb, (uint32_t)__popcll(am));
        base = __shfl(base, leader);
        if (amb) a.amb[1 + base + __popcll(am & ((1ull << lane) - 1ull))] = (uint32_t)slot;
    }
}

// float64 squared distance between original particles i and j
template <typename PT, bool PER>
MHIP_DEV double exact_d2(const PT *pos, int dim, uint32_t i, uint32_t j, double box) {
    double s = 0.0;
    for (int c = 0; c < dim; c++) {
        double d = (double)pos[(int64_t)j * dim + c] - (double)pos[(int64_t)i * dim + c];
        if (PER) d = wrap<double>(d, box);
        s = __dadd_rn(s, __dmul_rn(d, d));
    }
    return s;
}

// one ambiguous slot (from the main kernel's list), served by a whole
// wavefront.  The walk is the depth-first one of a single lane (children in
// index order; a node is entered when its box lies within the band's outer
// radius), with a node's 8 children tested by 8 lanes at once and a leaf's
// particles classified by 16 lanes at once: members and band entries are
// taken in the order a lone lane would take them.  The band (at most BAND_CAP
// entries, in LDS) is then ordered and completed by lane 0.
template <bool PER, bool P64, typename PT>
MHIP_DEV void resolve_slot(const KnnArgs &a, const PT *__restrict__ pos_in, int64_t slot, double *s_bd,
                           uint32_t *s_bi) {
    const int lane = threadIdx.x;  // one wavefront per block
    const TreeView &t = a.t;
    const uint32_t qp = a.qsorted ? a.qsorted[slot] : (uint32_t)slot;
    const QPos q = load_q<P64>(t, qp);
    const float tau = a.tau[slot];
    const float band_lo = tau * (1.0f - BAND_REL), band_hi = tau * (1.0f + BAND_REL);
    const float thr = band_hi * PRUNE_REL;
    const uint32_t qo = a.perm[qp];
    const int64_t nq = a.nq;
    const uint64_t below = (1ull << lane) - 1ull;
    int nb = 0, cnt = 0;
    int32_t st = a.st_slot[slot];
    const uint32_t first_leaf = (uint32_t)t.first_leaf;
    // children of internal node x whose boxes pass, as a bit mask
    auto expand = [&](uint32_t x) -> uint32_t {
        const bool pass = lane < FANOUT && node_d2<PER>(t, q, FANOUT * x + 1 + lane) <= thr;
        return (uint32_t)__ballot(pass) & 0xffu;
    };
    auto leaf = [&](uint32_t x) {
        const int64_t li = (int64_t)(x - first_leaf);
        if (li >= t.nleaf) return;
        const int64_t p = li * LEAF + lane;
        float d2 = INFINITY;
        if (lane < LEAF)
            d2 = dist2<PER, P64>(t, q, t.sx[p], t.sy[p], t.sz[p], P64 ? t.lx[p] : 0.f, P64 ? t.ly[p] : 0.f,
                                 P64 ? t.lz[p] : 0.f);
        const bool mem = lane < LEAF && d2 < band_lo;
        const bool inb = lane < LEAF && !mem && d2 <= band_hi && p < a.n;
        const uint64_t bm = __ballot(mem);
        if (mem) {
            const int pos = cnt + __popcll(bm & below);
            if (pos < a.k) {
                a.nl_idx[(int64_t)pos * nq + slot] = (uint32_t)p;
                if (a.nl_d2) a.nl_d2[(int64_t)pos * nq + slot] = d2;
            }
        }
        cnt += __popcll(bm);
        const uint64_t bb = __ballot(inb);
        if (!bb) return;
        const double e = inb ? exact_d2<PT, PER>(pos_in, a.dim, qo, a.perm[p], (double)t.box) : 0.0;
        const int nbb = __popcll(bb);
        if (nb + nbb <= BAND_CAP) {
            if (inb) {
                const int i = nb + __popcll(bb & below);
                s_bd[i] = e;
                s_bi[i] = (uint32_t)p;
            }
            nb += nbb;
            __syncthreads();
            return;
        }
        // entry by entry: appended while there is room, then replacing the
        // largest entry when smaller
        __syncthreads();
        uint64_t rem = bb;
        while (rem) {
            const int l = __ffsll((long long)rem) - 1;
            rem &= rem - 1ull;
            const double el = __shfl(e, l);
            const uint32_t pl = (uint32_t)__shfl((int)p, l);
            if (lane == 0) {
                if (nb < BAND_CAP) {
                    s_bd[nb] = el;
                    s_bi[nb] = pl;
                } else {
                    int wmax = 0;
                    for (int m = 1; m < BAND_CAP; m++)
                        if (s_bd[m] > s_bd[wmax]) wmax = m;
                    if (el < s_bd[wmax]) {
                        s_bd[wmax] = el;
                        s_bi[wmax] = pl;
                    }
                }
            }
            if (nb < BAND_CAP)
                nb++;
            else
                st |= ST_BAND_OVERFLOW;
            __syncthreads();
        }
    };
    if (node_d2<PER>(t, q, 0) <= thr) {
        if (first_leaf == 0) {
            leaf(0);
        } else {
            MaskStack ms;
            ms.put(0, expand(0));
            int lvl = 0;
            uint32_t x = 0;
            while (true) {
                const uint32_t m = ms.get(lvl);
                if (m == 0u) {
                    if (lvl == 0) break;
                    lvl--;
                    x = (x - 1) >> 3;
                    continue;
                }
                const uint32_t c = (uint32_t)__builtin_ctz(m);
                ms.clear(lvl, c);
                const uint32_t y = FANOUT * x + 1 + c;
                if (y >= first_leaf) {
                    leaf(y);
                    continue;
                }
                lvl++;
                x = y;
                ms.put(lvl, expand(y));
            }
        }
    }
    __syncthreads();
    if (lane != 0) return;
    // order the band by (exact d^2, original index) and take what is missing
    for (int i = 1; i < nb; i++) {
        const double v = s_bd[i];
        const uint32_t vi = s_bi[i];
        int j = i - 1;
        while (j >= 0 && (s_bd[j] > v || (s_bd[j] == v && a.perm[s_bi[j]] > a.perm[vi]))) {
            s_bd[j + 1] = s_bd[j];
            s_bi[j + 1] = s_bi[j];
            j--;
        }
        s_bd[j + 1] = v;
        s_bi[j + 1] = vi;
    }
    for (int i = 0; i < nb && cnt < a.k; i++, cnt++) {
        a.nl_idx[(int64_t)cnt * nq + slot] = s_bi[i];
        if (a.nl_d2) a.nl_d2[(int64_t)cnt * nq + slot] = (float)s_bd[i];
    }
    if (cnt != a.k) {
        st |= ST_COUNT;
        pad_list(a, min(cnt, a.k), qp, slot);
    } else {
        st &= ~ST_COUNT;
        if (a.h_slot[slot] > 0.f && !(st & ST_HSML_NOCONV)) {
            const double h64 = refine_h64<PER, P64>(a, q, (double)a.h_slot[slot],
                                                    [&](int j) { return a.nl_idx[(int64_t)j * nq + slot]; });
            write_h(a, slot, h64);
        }
    }
    a.st_slot[slot] = st;
}

// empties the list of ambiguous slots (a kernel, so graph capture records no
// memset node)
__global__ void k_amb_reset(uint32_t *__restrict__ amb) {
    if (threadIdx.x == 0) amb[0] = 0u;
}

// the list of ambiguous slots, one wavefront per slot (a grid-stride loop
// over the list, whose length only the device knows)
template <bool PER, bool P64, typename PT>
__global__ __launch_bounds__(64) void k_knn_resolve(KnnArgs a, const PT *__restrict__ pos_in) {
    __shared__ double s_bd[BAND_CAP];
    __shared__ uint32_t s_bi[BAND_CAP];
    const uint32_t count = a.amb[0];
    for (int64_t i = blockIdx.x; i < count; i += gridDim.x) {
        const int64_t slot = a.amb[1 + i];
        if (slot >= a.nq || !(a.st_slot[slot] & ST_AMBIGUOUS)) continue;
        resolve_slot<PER, P64, PT>(a, pos_in, slot, s_bd, s_bi);
        __syncthreads();  // the next slot reuses the band arrays
    }
}

__global__ void k_qsetup_mask(const int64_t *__restrict__ qidx, int64_t nq,
                              const uint32_t *__restrict__ iperm, uint64_t *__restrict__ keys,
                              uint32_t *__restrict__ vals) {
    int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= nq) return;
    keys[r] = iperm[qidx[r]];
    vals[r] = (uint32_t)r;
}

__global__ void k_qsetup_finish(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ vals,
                                int64_t nq, uint32_t *__restrict__ qsorted, uint32_t *__restrict__ qrow) {
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= nq) return;
    qsorted[s] = (uint32_t)keys[s];
    qrow[s] = vals[s];
}

// inverse of the index's sorted -> original map (built on first use: only
// queries on a subset of the particles need it)
__global__ __launch_bounds__(256) void k_iperm(const uint32_t *__restrict__ perm, int64_t n,
                                              uint32_t *__restrict__ iperm) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p < n) iperm[perm[p]] = (uint32_t)p;
}

static int ensure_iperm(mhip_ctx *ctx, const mhip_index *ix) {
    if (ix->iperm_ok) return 0;
    hipLaunchKernelGGL(k_iperm, dim3((unsigned)((ix->n + 255) / 256)), dim3(256), 0, ctx->stream, ix->perm, ix->n,
                       ix->iperm);
    ix->iperm_ok = 1;
    return check_hip(ctx, hipGetLastError(), "inverse permutation");
}

// Row-order outputs from the slot-ordered results.  One thread per query
// slot forms h, density = k m / (V_d h^d), volume = m / density
// (Meshoid.py:165-170) and the status bits, and writes them as one 16-byte
// record to its output row (row = srow[slot]: one scattered write per row);
// a second pass copies the records, in row order, into the four output
// arrays (coalesced).  The slot arrays are read in order, so no inverse of
// the sort order is needed.
__global__ __launch_bounds__(256) void k_rows_rec(const uint32_t *__restrict__ srow, int64_t nq, int k, int dim,
                                                 const double *__restrict__ h64, const int32_t *__restrict__ st,
                                                 const double *__restrict__ m_rows, RowRec *__restrict__ rec) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= nq) return;
    const uint32_t r = srow[s];
    const double h = h64[s];
    const double vn = dim == 3 ? 4.18879020478639098 : (dim == 2 ? 3.14159265358979324 : 2.0);
    const double m = m_rows ? m_rows[r] : 1.0;
    const double hp = dim == 3 ? h * h * h : (dim == 2 ? h * h : h);
    const double rho = (double)k * m / (vn * hp);
    rec[r] = RowRec{(float)h, (float)rho, (float)(m / rho), st[s]};
}

__global__ __launch_bounds__(256) void k_rows_split(const RowRec *__restrict__ rec, int64_t nq,
                                                   float *__restrict__ h_out, float *__restrict__ rho_out,
                                                   float *__restrict__ vol_out, int32_t *__restrict__ status_out) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= nq) return;
    const RowRec v = rec[r];
    if (h_out) h_out[r] = v.h;
    if (rho_out) rho_out[r] = v.rho;
    if (vol_out) vol_out[r] = v.vol;
    if (status_out) status_out[r] = v.status;
}

static int nl_grow(mhip_ctx *ctx, void **p, size_t bytes) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    if (!bytes) return 0;
    return check_hip(ctx, hipMalloc(p, bytes), "hipMalloc (neighbour list)");
}

template <int KCAP>
static void launch_main(const KnnArgs &a, bool per, bool p64, dim3 g, hipStream_t s) {
    if (per) {
        if (p64)
            hipLaunchKernelGGL((k_knn_main<KCAP, true, true>), g, dim3(KNN_BLOCK), 0, s, a);
        else
            hipLaunchKernelGGL((k_knn_main<KCAP, true, false>), g, dim3(KNN_BLOCK), 0, s, a);
    } else {
        if (p64)
            hipLaunchKernelGGL((k_knn_main<KCAP, false, true>), g, dim3(KNN_BLOCK), 0, s, a);
        else
            hipLaunchKernelGGL((k_knn_main<KCAP, false, false>), g, dim3(KNN_BLOCK), 0, s, a);
    }
}

static int nl_reserve(mhip_ctx *ctx, mhip_nlist *nl, int64_t nq, int k) {
    int rc = 0;
    if (nl->cap < nq || nl->k < k) {
        void **arr[] = {(void **)&nl->qsorted, (void **)&nl->qrow, (void **)&nl->tau,
                        (void **)&nl->nlt,     (void **)&nl->h,    (void **)&nl->status};
        for (auto p : arr)
            if ((rc = nl_grow(ctx, p, nq * 4))) return rc;
        if ((rc = nl_grow(ctx, (void **)&nl->rowrec, nq * 16))) return rc;
        if ((rc = nl_grow(ctx, (void **)&nl->amb, (nq + 1) * 4))) return rc;
        if ((rc = nl_grow(ctx, (void **)&nl->h64, nq * 8))) return rc;
        if ((rc = nl_grow(ctx, (void **)&nl->idx, (size_t)nq * k * 4))) return rc;
        if ((rc = nl_grow(ctx, (void **)&nl->d2, (size_t)nq * k * 4))) return rc;
        nl->cap = nq;
        nl->k = k;
    }
    return 0;
}

// slots of the query rows qidx in sorted (Hilbert) order: qsorted[s] is the
// sorted-space particle of slot s, qrow[s] its row in qidx
static int query_setup(mhip_ctx *ctx, const mhip_index *ix, const int64_t *qidx, int64_t nq, mhip_nlist *nl) {
    ProfRegion pr(ctx, "knn_query_setup");
    int rc = 0;
    size_t cw = radix_counts_words(nq);
    char *s = (char *)scratch(ctx, nq * 24 + cw * 4, &rc);
    if (!s) return rc;
    uint64_t *k0 = (uint64_t *)s, *k1 = k0 + nq;
    uint32_t *v0 = (uint32_t *)(k1 + nq), *v1 = v0 + nq, *counts = v1 + nq;
    if ((rc = ensure_iperm(ctx, ix))) return rc;
    hipLaunchKernelGGL(k_qsetup_mask, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, ctx->stream, qidx, nq,
                       ix->iperm, k0, v0);
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) < ix->n) bits++;
    uint64_t *ks;
    uint32_t *vs;
    if ((rc = radix_sort_pairs(ctx, k0, v0, k1, v1, counts, nq, bits, &ks, &vs))) return rc;
    hipLaunchKernelGGL(k_qsetup_finish, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, ctx->stream, ks, vs,
                       nq, nl->qsorted, nl->qrow);
    return 0;
}

static TreeView tree_view(const mhip_index *ix) {
    TreeView t;
    t.sx = ix->sx;
    t.sy = ix->sy;
    t.sz = ix->sz;
    t.lx = ix->lx;
    t.ly = ix->ly;
    t.lz = ix->lz;
    t.nodes = ix->nodes;
    t.first_leaf = ix->first_leaf;
    t.nleaf = ix->nleaf;
    t.depth = ix->depth;
    const bool per = ix->box > 0;
    t.box = per ? (float)ix->box : 0.f;
    t.half = per ? 0.5f * (float)ix->box : 0.f;
    float m = per ? (float)ix->box : 0.f;
    for (int c = 0; c < 6; c++) m = fmaxf(m, fabsf(ix->bbox[c]));
    t.slack = ix->pos_f64 ? 2.4e-7f * m : 0.f;  // 4 ulp of the largest |coordinate|
    return t;
}

int knn_run(mhip_ctx *ctx, const mhip_index *ix, const int64_t *qidx, int64_t nq, int k,
            const double *m_rows, const float *seed_h, const float *seed_r, float *h_out,
            float *rho_out, float *vol_out, int32_t *status_out, int flags, mhip_nlist *nl) {
    if (k < 2) return fail(ctx, -1, "des_ngb must be at least 2");
    if (k > 128) return fail(ctx, -4, "des_ngb above 128 is not supported by the GPU kNN");
    if (k > ix->n) return fail(ctx, -1, "des_ngb exceeds the number of particles");
    if (!qidx) nq = ix->n;
    if (nq < 1) return fail(ctx, -1, "no query rows");
    int rc = 0;
    if ((rc = nl_reserve(ctx, nl, nq, k))) return rc;
    nl->nq = nq;
    nl->k = k;
    nl->index = ix;
    nl->has_d2 = (flags & 1) != 0;
    // query slots in Hilbert order
    const bool masked = qidx != nullptr;
    nl->masked = masked;
    if (masked && (rc = query_setup(ctx, ix, qidx, nq, nl))) return rc;
    KnnArgs a;
    a.t = tree_view(ix);
    const bool per = ix->box > 0;
    a.qsorted = masked ? nl->qsorted : nullptr;
    a.qrow = masked ? nl->qrow : nullptr;
    a.perm = ix->perm;
    a.nq = nq;
    a.n = ix->n;
    a.k = k;
    a.dim = ix->dim;
    a.seed_h = seed_h;
    a.seed_r = seed_r;
    a.redo = 0;
    a.status_out = status_out;
    a.tau = nl->tau;
    a.nlt = nl->nlt;
    a.h_slot = nl->h;
    a.h64_slot = nl->h64;
    a.st_slot = nl->status;
    a.nl_idx = nl->idx;
    a.nl_d2 = nl->has_d2 ? nl->d2 : nullptr;
    a.stats = ctx->prof >= 2 ? ctx->stats : nullptr;
    a.amb = nl->amb;
    a.wlcap = (flags & MHIP_KNN_SMALL_RECORD) ? 8u : (uint32_t)WLCAP;
    const bool p64 = ix->pos_f64 != 0;
    const dim3 g((unsigned)((nq + KNN_BLOCK - 1) / KNN_BLOCK));
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(k_amb_reset, dim3(1), dim3(64), 0, s, nl->amb);
    const int need = k + 1;
    ProfRegion pr1(ctx, "knn_main");
    if (need <= 5)
        launch_main<5>(a, per, p64, g, s);
    else if (need <= 9)
        launch_main<9>(a, per, p64, g, s);
    else if (need <= 17)
        launch_main<17>(a, per, p64, g, s);
    else if (need <= 21)
        launch_main<21>(a, per, p64, g, s);
    else if (need <= 33)
        launch_main<33>(a, per, p64, g, s);
    else if (need <= 65)
        launch_main<65>(a, per, p64, g, s);
    else if (need <= 101)
        launch_main<101>(a, per, p64, g, s);
    else
        launch_main<129>(a, per, p64, g, s);
    if (seed_r) {
        // second launch: waves with a radius-seed miss are recomputed plainly
        KnnArgs r = a;
        r.redo = 1;
        if (need <= 5)
            launch_main<5>(r, per, p64, g, s);
        else if (need <= 9)
            launch_main<9>(r, per, p64, g, s);
        else if (need <= 17)
            launch_main<17>(r, per, p64, g, s);
        else if (need <= 21)
            launch_main<21>(r, per, p64, g, s);
        else if (need <= 33)
            launch_main<33>(r, per, p64, g, s);
        else if (need <= 65)
            launch_main<65>(r, per, p64, g, s);
        else if (need <= 101)
            launch_main<101>(r, per, p64, g, s);
        else
            launch_main<129>(r, per, p64, g, s);
    }
    pr1.end();
    ProfRegion pr3(ctx, "knn_resolve");
    const dim3 gr((unsigned)std::min<int64_t>(RESOLVE_WAVES, (nq + 63) / 64));
    if (per) {
        if (p64)
            hipLaunchKernelGGL((k_knn_resolve<true, true, double>), gr, dim3(64), 0, s, a,
                               (const double *)ix->pos_in);
        else
            hipLaunchKernelGGL((k_knn_resolve<true, false, float>), gr, dim3(64), 0, s, a,
                               (const float *)ix->pos_in);
    } else {
        if (p64)
            hipLaunchKernelGGL((k_knn_resolve<false, true, double>), gr, dim3(64), 0, s, a,
                               (const double *)ix->pos_in);
        else
            hipLaunchKernelGGL((k_knn_resolve<false, false, float>), gr, dim3(64), 0, s, a,
                               (const float *)ix->pos_in);
    }
    pr3.end();
    if (status_out || h_out || rho_out || vol_out) {
        ProfRegion pr4(ctx, "knn_rows_out");
        const dim3 gq((unsigned)((nq + 255) / 256));
        hipLaunchKernelGGL(k_rows_rec, gq, dim3(256), 0, s, masked ? nl->qrow : ix->perm, nq, k, ix->dim, nl->h64,
                           nl->status, m_rows, nl->rowrec);
        hipLaunchKernelGGL(k_rows_split, gq, dim3(256), 0, s, nl->rowrec, nq, h_out, rho_out, vol_out, status_out);
    }
    return check_hip(ctx, hipGetLastError(), "knn launch");
}

// ---------------------------------------------------------------------------
// export: rows in output-row order, ascending (distance, original index),
// original particle indices -- the layout cKDTree.query returns
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_export(const uint32_t *__restrict__ idx, const float *__restrict__ d2,
                                               const uint32_t *__restrict__ qrow,
                                               const uint32_t *__restrict__ perm, int64_t nq, int k,
                                               int32_t *__restrict__ ngb_out, float *__restrict__ dist_out) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= nq) return;
    const int64_t row = qrow ? qrow[s] : perm[s];
    for (int j = 0; j < k; j++) {
        const float dj = d2[(int64_t)j * nq + s];
        const uint32_t oj = perm[idx[(int64_t)j * nq + s]];
        int rank = 0;
        for (int l = 0; l < k; l++) {
            const float dl = d2[(int64_t)l * nq + s];
            const uint32_t ol = perm[idx[(int64_t)l * nq + s]];
            rank += (dl < dj) || (dl == dj && ol < oj);
        }
        if (ngb_out) ngb_out[row * k + rank] = (int32_t)oj;
        if (dist_out) dist_out[row * k + rank] = __builtin_sqrtf(dj);
    }
}

// ---------------------------------------------------------------------------
// rows of a full-set list for a subset of query rows.  A particle's k nearest
// neighbours, their distances and its smoothing length do not depend on
// which other particles are queried, so TreeUpdate with a particle_mask
// (src/meshoid/Meshoid.py:147-170) equals these rows of the unmasked list;
// density and volume use the subset's masses (Meshoid.py:169-170).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_nl_subset(
    const uint32_t *__restrict__ fidx, const float *__restrict__ fd2, const float *__restrict__ ftau,
    const int32_t *__restrict__ fnlt, const float *__restrict__ fh, const double *__restrict__ fh64,
    const int32_t *__restrict__ fst, int64_t nfull, const uint32_t *__restrict__ qsorted,
    const uint32_t *__restrict__ qrow, int64_t nq, int k, int dim, const double *__restrict__ m_rows,
    uint32_t *__restrict__ idx, float *__restrict__ d2, float *__restrict__ tau, int32_t *__restrict__ nlt,
    float *__restrict__ hs, double *__restrict__ hs64, int32_t *__restrict__ st, float *__restrict__ h_out,
    float *__restrict__ rho_out, float *__restrict__ vol_out, int32_t *__restrict__ status_out) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= nq) return;
    const uint32_t f = qsorted[s], row = qrow[s];
    for (int j = 0; j < k; j++) {
        idx[(int64_t)j * nq + s] = fidx[(int64_t)j * nfull + f];
        if (d2) d2[(int64_t)j * nq + s] = fd2[(int64_t)j * nfull + f];
    }
    const double h = fh64[f];
    tau[s] = ftau[f];
    nlt[s] = fnlt[f];
    hs[s] = fh[f];
    hs64[s] = h;
    st[s] = fst[f];
    // the main kernel's density and volume arithmetic
    const double vn = dim == 3 ? 4.18879020478639098 : (dim == 2 ? 3.14159265358979324 : 2.0);
    const double m = m_rows ? m_rows[row] : 1.0;
    const double hp = dim == 3 ? h * h * h : (dim == 2 ? h * h : h);
    const double rho = (double)k * m / (vn * hp);
    if (h_out) h_out[row] = (float)h;
    if (rho_out) rho_out[row] = (float)rho;
    if (vol_out) vol_out[row] = (float)(m / rho);
    if (status_out) status_out[row] = fst[f];
}

int nlist_subset(mhip_ctx *ctx, const mhip_nlist *full, const int64_t *qidx, int64_t nq, const double *m_rows,
                 float *h_out, float *rho_out, float *vol_out, int32_t *status_out, mhip_nlist *nl) {
    const mhip_index *ix = full->index;
    if (!ix || full->nq < 1) return fail(ctx, -1, "neighbour list is empty");
    if (full->masked || full->nq != ix->n) return fail(ctx, -1, "the source list must cover every particle");
    if (nl == full) return fail(ctx, -1, "the subset needs a list of its own");
    if (nq < 1) return fail(ctx, -1, "no query rows");
    const int k = full->k;
    int rc = 0;
    if ((rc = nl_reserve(ctx, nl, nq, k))) return rc;
    nl->nq = nq;
    nl->k = k;
    nl->index = ix;
    nl->has_d2 = full->has_d2;
    nl->masked = true;
    if ((rc = query_setup(ctx, ix, qidx, nq, nl))) return rc;
    ProfRegion pr(ctx, "nlist_subset");
    hipLaunchKernelGGL(k_nl_subset, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, ctx->stream, full->idx,
                       full->has_d2 ? full->d2 : nullptr, full->tau, full->nlt, full->h, full->h64, full->status,
                       full->nq, nl->qsorted, nl->qrow, nq, k, ix->dim, m_rows, nl->idx,
                       full->has_d2 ? nl->d2 : nullptr, nl->tau, nl->nlt, nl->h, nl->h64, nl->status, h_out, rho_out,
                       vol_out, status_out);
    return check_hip(ctx, hipGetLastError(), "subset launch");
}

// The same rows with float64 distances recomputed from the caller's
// positions (original indices), the distance cKDTree.query returns
// (src/meshoid/Meshoid.py:157): d = sqrt((dx^2 + dy^2) + dz^2) of the float64
// offsets, minimum image when periodic; rows ascending by (distance, index).
template <typename PT, bool PER>
__global__ __launch_bounds__(256) void k_export64(const uint32_t *__restrict__ idx, const uint32_t *__restrict__ qsorted,
                                                 const uint32_t *__restrict__ qrow, const uint32_t *__restrict__ perm,
                                                 const PT *__restrict__ pos, int dim, double box, int64_t nq, int k,
                                                 int32_t *__restrict__ ngb_out, double *__restrict__ dist_out) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= nq) return;
    const int64_t row = qrow ? qrow[s] : perm[s];
    const uint32_t qo = perm[qsorted ? qsorted[s] : (uint32_t)s];
    auto d2_of = [&](uint32_t o) {
        double r = 0.0;
        for (int c = 0; c < dim; c++) {
            double d = (double)pos[(int64_t)o * dim + c] - (double)pos[(int64_t)qo * dim + c];
            if (PER) d = wrap<double>(d, box);
            r = __dadd_rn(r, __dmul_rn(d, d));
        }
        return r;
    };
    for (int j = 0; j < k; j++) {
        const uint32_t oj = perm[idx[(int64_t)j * nq + s]];
        const double dj = d2_of(oj);
        int rank = 0;
        for (int l = 0; l < k; l++) {
            if (l == j) continue;
            const uint32_t ol = perm[idx[(int64_t)l * nq + s]];
            const double dl = d2_of(ol);
            rank += (dl < dj) || (dl == dj && ol < oj);
        }
        if (ngb_out) ngb_out[row * k + rank] = (int32_t)oj;
        if (dist_out) dist_out[row * k + rank] = sqrt(dj);
    }
}

int nlist_export64(mhip_ctx *ctx, const mhip_nlist *nl, int32_t *ngb_out, double *dist_out) {
    const mhip_index *ix = nl->index;
    if (!ix || nl->nq < 1) return fail(ctx, -1, "neighbour list is empty");
    if (!ix->pos_in) return fail(ctx, -1, "the index holds no positions");
    ProfRegion pr(ctx, "nlist_export64");
    const dim3 g((unsigned)((nl->nq + 255) / 256)), b(256);
    const uint32_t *qs = nl->masked ? nl->qsorted : nullptr, *qr = nl->masked ? nl->qrow : nullptr;
    const bool per = ix->box > 0;
#define MHIP_EXP(PT, P)                                                                                        \
    hipLaunchKernelGGL((k_export64<PT, P>), g, b, 0, ctx->stream, nl->idx, qs, qr, ix->perm, (const PT *)ix->pos_in, \
                       ix->dim, ix->box, nl->nq, nl->k, ngb_out, dist_out)
    if (ix->pos_f64) {
        if (per) MHIP_EXP(double, true);
        else MHIP_EXP(double, false);
    } else {
        if (per) MHIP_EXP(float, true);
        else MHIP_EXP(float, false);
    }
#undef MHIP_EXP
    return check_hip(ctx, hipGetLastError(), "export launch");
}

// the list's float64 smoothing lengths in output-row order
__global__ __launch_bounds__(256) void k_rows_h64(const double *__restrict__ h64, const uint32_t *__restrict__ qrow,
                                                 const uint32_t *__restrict__ perm, int64_t nq,
                                                 double *__restrict__ out) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= nq) return;
    out[qrow ? qrow[s] : perm[s]] = h64[s];
}

int nlist_hsml(mhip_ctx *ctx, const mhip_nlist *nl, double *h_out) {
    if (!nl->index || nl->nq < 1) return fail(ctx, -1, "neighbour list is empty");
    hipLaunchKernelGGL(k_rows_h64, dim3((unsigned)((nl->nq + 255) / 256)), dim3(256), 0, ctx->stream, nl->h64,
                       nl->masked ? nl->qrow : nullptr, nl->index->perm, nl->nq, h_out);
    return check_hip(ctx, hipGetLastError(), "hsml copy launch");
}

int nlist_export(mhip_ctx *ctx, const mhip_nlist *nl, int32_t *ngb_out, float *dist_out) {
    const mhip_index *ix = nl->index;
    if (!ix || nl->nq < 1) return fail(ctx, -1, "neighbour list is empty");
    if (!nl->has_d2) return fail(ctx, -1, "neighbour list was built without distances");
    ProfRegion pr(ctx, "nlist_export");
    hipLaunchKernelGGL(k_export, dim3((unsigned)((nl->nq + 255) / 256)), dim3(256), 0, ctx->stream,
                       nl->idx, nl->d2, nl->masked ? nl->qrow : nullptr, ix->perm, nl->nq, nl->k,
                       ngb_out, dist_out);
    return check_hip(ctx, hipGetLastError(), "export launch");
}

// ---------------------------------------------------------------------------
// nearest particle of arbitrary float64 points: cKDTree.query(x)[1], k = 1
// (src/meshoid/Meshoid.py:409).  The points are first put in the index's
// Hilbert order, so that a wavefront's 64 lanes serve a compact block of
// points.  (1) Each lane finds, by binary search over the leaves' first
// sort keys, the leaf where its point's key falls, and scans it and its
// two neighbours (in Hilbert order these usually hold the nearest particle):
// every lane starts with a tight bound.  (2) A wave-uniform walk, inside out
// from the middle lane's leaf (its siblings, then the rest of the
// grandparent's subtree, ... up to the root), tests each node once for all
// lanes and scans the leaves some lane's bound (widened by the float32
// rounding band) reaches; it keeps the nearest and the second-nearest
// float32 d2.  (3) Where those two lie closer than float32 rounding can
// order, the band is re-ranked with float64 distances from the caller's
// positions (the distance cKDTree computes), ties to the lowest original
// index.  The point is split into float32 hi + lo parts, so its float32
// distances are as accurate as those between two particles.
// ---------------------------------------------------------------------------
struct NearestArgs {
    TreeView t;
    const uint32_t *order;  // point of each thread (Hilbert order), null = identity
    const uint32_t *perm;
    const void *pos_in;
    int64_t n;
    int dim;
    double box;        // > 0: periodic in [0, box)
    float org[3], scale;  // the index's key quantisation frame
    const double *tg;  // (nt, dim) row-major
    int64_t nt;
    int64_t *out;      // original particle index per point, -1 for a non-finite point
    unsigned long long *stats;  // work counters, may be null
    const uint64_t *leaf_key;   // first key of each leaf, on the bits the index is sorted on
};

template <typename PT, bool PER>
MHIP_DEV double exact_d2_point(const PT *pos, int dim, uint32_t j, const double (&tp)[3], double box) {
    double s = 0.0;
    for (int c = 0; c < dim; c++) {
        double d = (double)pos[(int64_t)j * dim + c] - tp[c];
        if (PER) d = wrap<double>(d, box);
        s = __dadd_rn(s, __dmul_rn(d, d));
    }
    return s;
}

// sort key of each point in the index's quantisation frame (as the
// particles' keys are formed), periodic points wrapped into the box first
__global__ __launch_bounds__(256) void k_point_keys(const double *__restrict__ tg, int64_t nt, int dim,
                                                    double box, float ox, float oy, float oz, float scale,
                                                    uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const float org[3] = {ox, oy, oz};
    uint32_t cc[3] = {0u, 0u, 0u};
    for (int c = 0; c < dim; c++) {
        double v = tg[i * dim + c];
        if (box > 0.0 && (v < 0.0 || v >= box)) v -= floor(v / box) * box;
        cc[c] = key_cell((float)v, org[c], scale);
    }
    keys[i] = hilbert_key(cc[0], cc[1], cc[2], dim);
    vals[i] = (uint32_t)i;
}

// sort key of sorted particle p, formed as the index build forms it
MHIP_DEV uint64_t particle_key(const TreeView &t, const float (&org)[3], float scale, int dim, int64_t p) {
    const float v[3] = {t.sx[p], t.sy[p], t.sz[p]};
    uint32_t cc[3] = {0u, 0u, 0u};
    for (int c = 0; c < dim; c++) cc[c] = key_cell(v[c], org[c], scale);
    return hilbert_key(cc[0], cc[1], cc[2], dim);
}

// first sort key of every leaf (on the bits the index is sorted on), for the
// points' first-leaf search
__global__ __launch_bounds__(256) void k_leaf_keys(TreeView t, float ox, float oy, float oz, float scale, int dim,
                                                  uint64_t *__restrict__ out) {
    const int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (l >= t.nleaf) return;
    const float org[3] = {ox, oy, oz};
    out[l] = particle_key(t, org, scale, dim, l * LEAF) >> key_sort_lo_bit(dim);
}

template <bool PER, bool P64, typename PT>
__global__ __launch_bounds__(256) void k_nearest(NearestArgs a) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t wave0 = tid - lane;
    if (wave0 >= a.nt) return;
    const bool live = tid < a.nt;
    // lanes past the end take the wave's first point (never written)
    const int64_t slot = live ? tid : wave0;
    const int64_t i = a.order ? (int64_t)a.order[slot] : slot;
    double tp[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < a.dim; c++) {
        double v = a.tg[i * a.dim + c];
        if (PER && (v < 0.0 || v >= a.box)) {
            v -= floor(v / a.box) * a.box;
            if (!(v >= 0.0 && v < a.box)) v = 0.0;  // rounded onto the upper edge
        }
        tp[c] = v;
    }
    QPos q;
    q.x = (float)tp[0];
    q.y = (float)tp[1];
    q.z = (float)tp[2];
    q.lx = (float)(tp[0] - (double)q.x);
    q.ly = (float)(tp[1] - (double)q.y);
    q.lz = (float)(tp[2] - (double)q.z);
    // box tests use the hi parts: the slack covers the point's lo parts and
    // (float64 particles) the particles' lo parts
    TreeView t = a.t;
    t.slack = a.t.slack + fmaxf(fabsf(q.lx), fmaxf(fabsf(q.ly), fabsf(q.lz)));
    QBox qb;
    qb.lx = q.x - t.slack;
    qb.ly = q.y - t.slack;
    qb.lz = q.z - t.slack;
    qb.hx = q.x + t.slack;
    qb.hy = q.y + t.slack;
    qb.hz = q.z + t.slack;
    const uint32_t first_leaf = (uint32_t)t.first_leaf;
    // (1) this lane's leaf: the last leaf whose first key is at most the
    // point's key (on the bits the index is sorted on); it and its neighbours
    // give the lane a first bound
    const int kshift = key_sort_lo_bit(a.dim);
    uint64_t kq = 0;
    {
        uint32_t cc[3] = {0u, 0u, 0u};
        for (int c = 0; c < a.dim; c++) cc[c] = key_cell((float)tp[c], a.org[c], a.scale);
        kq = hilbert_key(cc[0], cc[1], cc[2], a.dim) >> kshift;
    }
    uint32_t lo = 0, hi = (uint32_t)t.nleaf;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.leaf_key[mid] <= kq)
            lo = mid;
        else
            hi = mid;
    }
    const uint32_t own_lo = lo > 0 ? lo - 1 : 0, own_hi = min(lo + 1, (uint32_t)t.nleaf - 1);
    float best = INFINITY, second = INFINITY;
    uint32_t bp = 0xffffffffu;
    float cx[LEAF], cy[LEAF], cz[LEAF], clx[LEAF], cly[LEAF], clz[LEAF];
    // candidates of one leaf; skip: this lane has the leaf already
    auto take = [&](int64_t leaf, bool skip) {
#pragma unroll
        for (int j = 0; j < LEAF; j++) {
            // padding particles carry NaN coordinates: never below a bound
            const float d2 = dist2<PER, true>(t, q, cx[j], cy[j], cz[j], clx[j], cly[j], clz[j]);
            const bool nb = !skip && d2 < best;
            second = nb ? best : (skip ? second : fminf(second, d2));
            best = nb ? d2 : best;
            bp = nb ? (uint32_t)(leaf * LEAF + j) : bp;
        }
    };
    auto scan = [&](uint32_t node) {
        const int64_t leaf = node - first_leaf;
        load_leaf_uniform<P64>(t.sx, t.sy, t.sz, t.lx, t.ly, t.lz, leaf, cx, cy, cz, clx, cly, clz);
        take(leaf, leaf >= own_lo && leaf <= own_hi);
    };
    for (uint32_t l = own_lo; l <= own_hi; l++) {
        load_leaf<P64>(t, l, cx, cy, cz, clx, cly, clz);
        take(l, false);
    }
    // the middle lane's leaf, for every lane
    const uint32_t x = first_leaf + uniform(__shfl(lo, WAVE / 2));
    uint32_t n_nodes = 0, n_leaves = 1;
    scan(x);
    // (2) wave-uniform walk, inside out: the subtree of each ancestor of the
    // middle lane's leaf in turn (parent, grandparent, ... root), minus the
    // subtree already done, depth first (children in index order) and pruned
    // by every lane's bound.  A node's 8 children are tested together (their
    // boxes loaded at once) into a mask; a leaf is tested again with the
    // bound of the moment before it is scanned.  A leaf inside a child that
    // the bound of the moment would have pruned lies in that child's box, so
    // the leaf test prunes it too: the leaves scanned, and their order, are
    // those of a walk that tests each node just before entering it.
    const float F = (1.0f + BAND_REL) * PRUNE_REL;
    uint32_t done = x;  // scanned
    while (done != 0) {
        const uint32_t top = (done - 1) >> 3;
        MaskStack ms;
        ms.put(0, expand_mask<PER>(t, qb, top, best * F) & ~(1u << (done - (FANOUT * top + 1))));
        n_nodes += FANOUT;
        int lvl = 0;
        uint32_t cur = top;
        while (true) {
            const uint32_t m = ms.get(lvl);
            if (m == 0u) {
                if (lvl == 0) break;
                lvl--;
                cur = (cur - 1) >> 3;
                continue;
            }
            const uint32_t c = (uint32_t)__builtin_ctz(m);
            ms.clear(lvl, c);
            const uint32_t y = FANOUT * cur + 1 + c;
            if (y >= first_leaf) {
                if (y - first_leaf < t.nleaf && __any(box_gap2_u<PER>(t, qb, y) <= best * F)) {
                    scan(y);
                    n_leaves++;
                }
                continue;
            }
            lvl++;
            cur = y;
            ms.put(lvl, expand_mask<PER>(t, qb, y, best * F));
            n_nodes += FANOUT;
        }
        done = top;
    }
    if (a.stats) {
        const bool ex = live && bp != 0xffffffffu && second <= best * (1.0f + AMBIG_REL);
        const uint32_t nex = (uint32_t)__popcll(__ballot(ex));
        if (lane == 0) {
            atomicAdd(a.stats + STAT_NN_WAVES, 1ull);
            atomicAdd(a.stats + STAT_NN_NODES, (unsigned long long)n_nodes);
            atomicAdd(a.stats + STAT_NN_LEAVES, (unsigned long long)n_leaves);
            atomicAdd(a.stats + STAT_NN_EXACT, (unsigned long long)nex);
        }
    }
    if (!live) return;
    int64_t res = -1;
    if (bp != 0xffffffffu) {
        res = a.perm[bp];
        if (second <= best * (1.0f + AMBIG_REL)) {
            // (3) float64 re-rank of this point's rounding band (rare; per lane)
            const PT *pos = (const PT *)a.pos_in;
            const float lim = best * (1.0f + BAND_REL);
            double be = INFINITY;
            uint32_t bo = 0xffffffffu;
            uint32_t y = 0;
            bool more = true;
            while (more) {
                if (node_d2<PER>(t, q, y) <= lim * PRUNE_REL) {
                    if (y < first_leaf) {
                        y = FANOUT * y + 1;
                        continue;
                    }
                    if (y - first_leaf < t.nleaf) {
                        const int64_t leaf = y - first_leaf;
                        load_leaf<P64>(t, leaf, cx, cy, cz, clx, cly, clz);
                        for (int j = 0; j < LEAF; j++) {
                            const int64_t p = leaf * LEAF + j;
                            const float d2 = dist2<PER, true>(t, q, cx[j], cy[j], cz[j], clx[j], cly[j], clz[j]);
                            if (d2 <= lim && p < a.n) {
                                const uint32_t o = a.perm[p];
                                const double e = exact_d2_point<PT, PER>(pos, a.dim, o, tp, a.box);
                                if (e < be || (e == be && o < bo)) {
                                    be = e;
                                    bo = o;
                                }
                            }
                        }
                    }
                }
                more = next_after(y);
            }
            res = bo;
        }
    }
    a.out[i] = res;
}

int nearest_run(mhip_ctx *ctx, const mhip_index *ix, const double *targets, int64_t nt, int dim,
                int64_t *out) {
    if (!ix || ix->n < 1) return fail(ctx, -1, "the index holds no particles");
    if (dim != ix->dim) return fail(ctx, -1, "point dimension differs from the particles'");
    if (nt < 0) return fail(ctx, -1, "negative point count");
    if (nt == 0) return 0;
    NearestArgs a;
    a.t = tree_view(ix);
    a.order = nullptr;
    a.perm = ix->perm;
    a.pos_in = ix->pos_in;
    a.n = ix->n;
    a.dim = ix->dim;
    a.box = ix->box > 0 ? ix->box : 0.0;
    for (int c = 0; c < 3; c++) a.org[c] = ix->frame[c];
    a.scale = ix->frame[3];
    a.tg = targets;
    a.nt = nt;
    a.out = out;
    a.stats = ctx->prof >= 2 ? ctx->stats : nullptr;
    const bool per = ix->box > 0, p64 = ix->pos_f64 != 0;
    const dim3 g((unsigned)((nt + 255) / 256));
    hipStream_t s = ctx->stream;
    // scratch: the points' sort (large sets), then the leaves' first keys
    const bool sort_pts = nt >= 4096 && nt <= 0xffffffffll;
    const size_t cw = sort_pts ? radix_counts_words(nt) : 0;
    const size_t pts_bytes = sort_pts ? (((size_t)nt * 24 + cw * 4 + 255) & ~(size_t)255) : 0;
    int rc = 0;
    char *sc = (char *)scratch(ctx, pts_bytes + (size_t)ix->nleaf * sizeof(uint64_t), &rc);
    if (!sc) return rc;
    uint64_t *lkey = (uint64_t *)(sc + pts_bytes);
    a.leaf_key = lkey;
    if (sort_pts) {
        // points in the index's Hilbert order
        ProfRegion pk(ctx, "nearest_sort");
        uint64_t *k0 = (uint64_t *)sc, *k1 = k0 + nt;
        uint32_t *v0 = (uint32_t *)(k1 + nt), *v1 = v0 + nt, *counts = v1 + nt;
        hipLaunchKernelGGL(k_point_keys, g, dim3(256), 0, s, targets, nt, dim, a.box, ix->frame[0], ix->frame[1],
                           ix->frame[2], ix->frame[3], k0, v0);
        uint64_t *ks;
        uint32_t *vs;
        if ((rc = radix_sort_pairs(ctx, k0, v0, k1, v1, counts, nt, KEY_BITS_PER_AXIS * dim, &ks, &vs,
                                   key_sort_lo_bit(dim))))
            return rc;
        a.order = vs;
    }
    ProfRegion pr(ctx, "nearest");
    hipLaunchKernelGGL(k_leaf_keys, dim3((unsigned)((ix->nleaf + 255) / 256)), dim3(256), 0, s, a.t, a.org[0],
                       a.org[1], a.org[2], a.scale, a.dim, lkey);
    if (per) {
        if (p64)
            hipLaunchKernelGGL((k_nearest<true, true, double>), g, dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((k_nearest<true, false, float>), g, dim3(256), 0, s, a);
    } else {
        if (p64)
            hipLaunchKernelGGL((k_nearest<false, true, double>), g, dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((k_nearest<false, false, float>), g, dim3(256), 0, s, a);
    }
    return check_hip(ctx, hipGetLastError(), "nearest launch");
}

}  // namespace mhip
