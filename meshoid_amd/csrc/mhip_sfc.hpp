// Sort keys of the spatial index: the Hilbert index of a particle's cell on
// the 2^21-per-axis grid of the quantisation frame.  Header-only and usable
// from host code (a CPU unit test compiles it with g++).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define MHIP_HD __host__ __device__ __forceinline__
#else
#define MHIP_HD inline
#endif

namespace mhip {

// the 21 low bits of x spread to every third bit
MHIP_HD uint64_t spread21(uint64_t x) {
    x &= 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

// the 21 low bits of x spread to every second bit
MHIP_HD uint64_t spread21_2(uint64_t x) {
    x &= 0x1fffffull;
    x = (x | x << 16) & 0x0000ffff0000ffffull;
    x = (x | x << 8) & 0x00ff00ff00ff00ffull;
    x = (x | x << 4) & 0x0f0f0f0f0f0f0f0full;
    x = (x | x << 2) & 0x3333333333333333ull;
    x = (x | x << 1) & 0x5555555555555555ull;
    return x;
}

constexpr int KEY_BITS_PER_AXIS = 21;

// The index sorts its keys on bits [key_sort_lo_bit(dim), dim * 21): in 3D
// the lowest 8-bit digit is left out (7 radix passes instead of 8), so
// particles order along the curve on a 2^(55/3)-per-axis grid (18.3 bits) and
// those sharing a cell of it keep their input order.  Measured at 2^24
// clustered particles: the kNN time is unchanged with 55 or 47 sorted bits
// and 13% longer with 39 (cells then hold several particles in the densest
// clumps), so 55 bits leave 256 times that density before ties cost anything.
// 1D and 2D keys are sorted whole.
MHIP_HD constexpr int key_sort_lo_bit(int dim) { return dim == 3 ? 8 : 0; }

// Position of the cell (c0, c1, c2) (each < 2^21) along the Hilbert curve
// through the 2^21 x 2^21 (x 2^21) grid: a dim*21-bit key (dim 1: the
// coordinate itself).  Consecutive keys are face-adjacent cells, so a run of
// sorted particles covers a compact region (a Z-order jumps across octant
// boundaries).  Method: J. Skilling, "Programming the Hilbert curve",
// AIP Conf. Proc. 707, 381 (2004): the axes are turned in place into the
// "transposed" index (bit b of axis i is key bit dim*b + dim-1-i) by one
// sweep from the top bit down that inverts or exchanges low bits, then a Gray
// decode along the axes.
MHIP_HD uint64_t hilbert_key(uint32_t c0, uint32_t c1, uint32_t c2, int dim) {
    if (dim == 1) return c0;
    uint32_t X[3] = {c0, c1, dim == 3 ? c2 : 0u};
    constexpr uint32_t TOP = 1u << (KEY_BITS_PER_AXIS - 1);
    for (uint32_t q = TOP; q > 1u; q >>= 1) {
        const uint32_t p = q - 1u;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (i >= dim) break;
            // bit q of axis i set: invert the low bits of axis 0; else swap
            // the low bits of axes 0 and i
            const uint32_t sw = (X[0] ^ X[i]) & p;
            const bool set = (X[i] & q) != 0u;
            X[0] ^= set ? p : sw;
            if (i > 0) X[i] ^= set ? 0u : sw;
        }
    }
    X[1] ^= X[0];
    if (dim == 3) X[2] ^= X[1];
    uint32_t t = 0u;
    const uint32_t last = X[dim - 1];
    for (uint32_t q = TOP; q > 1u; q >>= 1)
        if (last & q) t ^= q - 1u;
    X[0] ^= t;
    X[1] ^= t;
    X[2] ^= t;
    if (dim == 2) return spread21_2(X[0]) << 1 | spread21_2(X[1]);
    return spread21(X[0]) << 2 | spread21(X[1]) << 1 | spread21(X[2]);
}

// cell of coordinate v in the index's quantisation frame (origin o, scale s
// cells per unit, 2^21 - 1 over the longest side); NaN -> 0
MHIP_HD uint32_t key_cell(float v, float o, float s) {
    float t = (v - o) * s;
    t = t > 0.f ? (t < 2097151.0f ? t : 2097151.0f) : 0.f;
    return (uint32_t)t;
}

}  // namespace mhip
