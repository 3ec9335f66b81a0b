// Shared device helpers and internal host structures of libmeshoid_hip.
//
// Conventions used by every kernel in this library:
//   * particles are reordered once into Hilbert order ("sorted space"); positions
//     live there as float32 structure-of-arrays sx/sy/sz padded to a multiple of
//     LEAF with NaN coordinates (every distance comparison against a padding
//     slot is false, also after a periodic wrap).  Float64 inputs keep a float32 residual
//     (lx/ly/lz) so that x = hi + lo to ~2^-48.
//   * dim < 3 is handled by zero coordinates in the unused axes.
//   * one wavefront (64 lanes) serves 64 consecutive queries in sorted order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#define MHIP_DEV __device__ __forceinline__

namespace mhip {

constexpr int WAVE = 64;
constexpr int LEAF = 16;    // particles per leaf of the implicit tree
constexpr int FANOUT = 8;   // children per internal node
constexpr int MAX_DEPTH = 10;  // complete 8-ary tree: up to 8^10 leaves

// status bits written per query row
enum : int32_t {
    ST_AMBIGUOUS = 1,      // k-th / (k+1)-th distances within fp32 rounding: resolved in fp64
    ST_BAND_OVERFLOW = 2,  // more fp64 ties at the k-th distance than the resolver holds
    ST_HSML_NOCONV = 4,    // smoothing-length iteration hit its cap
    ST_SINGULAR = 8,       // exactly singular moment matrix (zero pivot)
    ST_COUNT = 16,         // neighbour collection did not return exactly k members
    ST_SEED_MISS = 32,     // learned seed rejected, bracket used instead
    ST_RANKDEF = 64,       // moment matrix rank-deficient at rounding level: row is NaN
    ST_RADIUS_MISS = 128,  // search-radius seed too small: row recomputed without it
};

// ---------------------------------------------------------------------------
// cubic-spline kernel W(q) (unnormalised), src/meshoid/kernel_density.py:82-96
//   1 - 6q^2 + 6q^3   q <= 1/2
//   2 (1-q)^3         1/2 < q <= 1
//   0                 q > 1
// ---------------------------------------------------------------------------
template <typename T>
MHIP_DEV T kernel_w(T q) {
    T a = T(1) - q;
    T inner = T(1) + T(6) * q * q * (q - T(1));
    T outer = T(2) * a * a * a;
    T w = q <= T(0.5) ? inner : outer;
    return q <= T(1) ? w : T(0);
}

// W and dW/dq together
template <typename T>
MHIP_DEV void kernel_w_dw(T q, T &w, T &dw) {
    T a = T(1) - q;
    bool in = q <= T(0.5);
    T wi = T(1) + T(6) * q * q * (q - T(1));
    T dwi = q * (T(18) * q - T(12));
    T wo = T(2) * a * a * a;
    T dwo = T(-6) * a * a;
    w = in ? wi : wo;
    dw = in ? dwi : dwo;
    if (!(q <= T(1))) {
        w = T(0);
        dw = T(0);
    }
}

// norm_d of kernel_density.py:26-31 and the bracket factor of :35
MHIP_DEV double hsml_norm(int dim) { return dim == 3 ? 32.0 / 3 : (dim == 2 ? 40.0 / 7 : 8.0 / 3); }
MHIP_DEV double hsml_bound_coeff(int dim) {
    return dim == 3 ? 1.5638742363463305 : (dim == 2 ? 1.7984058958344455 : 2.33839974837265);
}

// nearest_image of src/meshoid/derivatives.py:8-14 (box > 0)
template <typename T>
MHIP_DEV T wrap(T d, T box) {
    T ad = d < T(0) ? -d : d;
    if (ad > box * T(0.5)) return d > T(0) ? -(box - ad) : (box - ad);
    return d;
}

MHIP_DEV float med3(float a, float b, float c) { return __builtin_amdgcn_fmed3f(a, b, c); }

MHIP_DEV uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// XCD-aware block remap: consecutive logical blocks land on the same XCD
// (the hardware dispatches physical block b to XCD b % 8), so neighbouring
// query groups share that XCD's L2.
MHIP_DEV uint32_t xcd_remap(uint32_t b, uint32_t nblocks) {
    constexpr uint32_t NX = 8;
    uint32_t per = nblocks / NX, rem = nblocks % NX;
    uint32_t x = b % NX, i = b / NX;
    // XCDs 0..rem-1 own per+1 logical blocks, the rest own per
    uint32_t start = x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
    return start + i;
}

}  // namespace mhip

// ---------------------------------------------------------------------------
// host-side objects behind the opaque C handles
// ---------------------------------------------------------------------------
struct mhip_prof_rec {
    const char *name;
    hipEvent_t start, stop;
};

struct mhip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    hipEvent_t order_ev = nullptr;  // orders work across a stream change
    // kernel timing: HIP events recorded on the context stream around each
    // launch region while prof is on (mhip_prof_enable / mhip_prof_read)
    int prof = 0;
    std::vector<mhip_prof_rec> recs;
    std::vector<hipEvent_t> event_pool;
    // work counters (prof >= 2): device u64[16], summed per wavefront
    unsigned long long *stats = nullptr;
};

namespace mhip {
enum : int {
    STAT_P1_LEAVES = 0,   // leaves scanned by the kNN main pass (sum over lanes)
    STAT_P1_INSERTS,      // insertion steps executed per wavefront (sum over waves)
    STAT_P1_NODES,        // tree nodes tested (sum over lanes)
    STAT_P2_LEAVES,       // leaves scanned in pass 2
    STAT_P2_NODES,        // recorded leaves re-scanned for members (main pass)
    STAT_WAVES,           // wavefronts of pass 1
    STAT_LANE_NEED,       // leaves under a leaf-parent that a lane itself needs (sum over lanes)
    STAT_MAXLANE_NEED,    // per wavefront, the largest such count of one lane (sum over waves)
    STAT_PARENT_MAX,      // per leaf-parent expansion, the largest count of one lane (sum)
    STAT_NN_WAVES,        // wavefronts of the nearest-point query
    STAT_NN_NODES,        // its wave-uniform node tests (sum over waves)
    STAT_NN_LEAVES,       // its wave-uniform leaf scans (sum over waves)
    STAT_NN_EXACT,        // points re-ranked in float64
    STAT_NEWTON,          // smoothing-length iterations (sum over query lanes)
    STAT_NEWTON_WMAX,     // per wavefront, the largest iteration count of a lane (sum)
    STAT_OVERFLOW_WAVES,  // wavefronts whose leaf record overflowed (members by a second walk)
    STAT_COUNT
};
static_assert(STAT_COUNT <= 16, "the work-counter buffer holds 16 entries");
}

namespace mhip {
// records start/stop events around a launch region when profiling is on
struct ProfRegion {
    mhip_ctx *ctx;
    int slot = -1;
    ProfRegion(mhip_ctx *c, const char *name) : ctx(c) {
        if (!ctx->prof) return;
        mhip_prof_rec r;
        r.name = name;
        r.start = take(), r.stop = take();
        (void)hipEventRecord(r.start, ctx->stream);
        ctx->recs.push_back(r);
        slot = (int)ctx->recs.size() - 1;
    }
    void end() {
        if (slot >= 0) (void)hipEventRecord(ctx->recs[slot].stop, ctx->stream);
        slot = -1;
    }
    ~ProfRegion() { end(); }
    hipEvent_t take() {
        hipEvent_t e;
        if (!ctx->event_pool.empty()) {
            e = ctx->event_pool.back();
            ctx->event_pool.pop_back();
        } else {
            (void)hipEventCreate(&e);
        }
        return e;
    }
};
}  // namespace mhip

namespace mhip {
// one output row of the kNN pass (16 bytes: one write per row)
struct __align__(16) RowRec {
    float h, rho, vol;
    int32_t status;
};
}  // namespace mhip

struct mhip_index {
    int64_t n = 0, npad = 0, cap = 0;
    int dim = 3;
    double box = -1.0;  // > 0: periodic in [0, box)^dim
    int pos_f64 = 0;
    // sorted space
    float *sx = nullptr, *sy = nullptr, *sz = nullptr;  // hi parts, npad entries
    float *lx = nullptr, *ly = nullptr, *lz = nullptr;  // lo parts (f64 inputs only)
    uint32_t *perm = nullptr;                           // sorted -> original
    uint32_t *iperm = nullptr;                          // original -> sorted (built on first use)
    mutable int iperm_ok = 0;                           // iperm holds this build's inverse
    const void *pos_in = nullptr;                       // caller's positions (for fp64 resolution)
    // complete 8-ary tree in heap numbering: node x has children 8x+1..8x+8,
    // leaves are nodes first_leaf .. first_leaf + 8^depth - 1, leaf j holds
    // sorted particles [LEAF*j, LEAF*j + LEAF); leaves j >= nleaf are empty
    // (box lo = +inf, hi = -inf).  Boxes: 8 floats (lo xyz, hi xyz, pad).
    int depth = 0;
    int64_t nleaf = 0, first_leaf = 0, nnodes = 0;
    float *nodes = nullptr;
    int64_t node_cap = 0;
    float bbox[6] = {};
    float frame[4] = {};  // sort-key quantisation: origin xyz, scale (cells per unit, 2^21 - 1 over the extent)
};

struct mhip_nlist {
    int64_t nq = 0, cap = 0;
    int k = 0;
    uint32_t *qsorted = nullptr;  // sorted-space particle index of each query slot
    uint32_t *qrow = nullptr;     // output row of each query slot
    mhip::RowRec *rowrec = nullptr;  // per output row: h, rho, vol, status, staged for the row outputs
    uint32_t *idx = nullptr;      // column-major [k][nq] sorted-space neighbour indices
    float *d2 = nullptr;          // column-major [k][nq] squared distances
    float *tau = nullptr;         // k-th squared distance per slot
    int32_t *nlt = nullptr;       // count strictly below tau per slot
    float *h = nullptr;           // smoothing length per slot (float32 copy of h64)
    double *h64 = nullptr;        // smoothing length per slot, refined in float64
    int32_t *status = nullptr;    // per slot
    uint32_t *amb = nullptr;      // [0] count, [1..] slots flagged ambiguous (for the resolver)
    int masked = 0;               // qsorted/qrow valid (query rows are a subset/reordering)
    int has_d2 = 0;               // d2 filled (needed for export)
    const mhip_index *index = nullptr;
};
