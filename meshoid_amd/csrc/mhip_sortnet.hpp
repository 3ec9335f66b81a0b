// In-register sorting networks for the kNN's list fill: Batcher's odd-even
// merge sort of 32 values and the merge of a sorted 32 into a sorted k+1 = 33
// list.  Index arithmetic is compile-time, so the arrays stay in registers.
// Header-only and usable from host code (a CPU unit test compiles it with g++).
// Values must not be NaN (callers map NaN to +inf first).
#pragma once

#if defined(__HIPCC__)
#define MHIP_HD __host__ __device__ __forceinline__
#else
#define MHIP_HD inline
#endif

namespace mhip {

MHIP_HD void cmpx(float &a, float &b) {
    const float lo = a < b ? a : b;
    const float hi = a < b ? b : a;
    a = lo;
    b = hi;
}

// Batcher's odd-even merge sort, ascending (K. E. Batcher, "Sorting networks
// and their applications", AFIPS 1968): for N = 32, 191 compare-exchanges.
template <int N>
MHIP_HD void oddeven_sort(float (&v)[N]) {
    static_assert((N & (N - 1)) == 0, "power of two");
#pragma unroll
    for (int p = 1; p < N; p <<= 1) {
#pragma unroll
        for (int k = p; k >= 1; k >>= 1) {
#pragma unroll
            for (int j = k % p; j + k < N; j += 2 * k) {
#pragma unroll
                for (int i = 0; i < k; i++) {
                    if (i + j + k < N && (i + j) / (2 * p) == (i + j + k) / (2 * p)) cmpx(v[i + j], v[i + j + k]);
                }
            }
        }
    }
}

// A[0..N] ascending (N + 1 slots), S[0..N-1] ascending: afterwards A holds
// the N + 1 smallest of A and S, ascending.  The N smallest of A[0..N-1] and S
// are min(A[i], S[N-1-i]) (a bitonic sequence, sorted by a bitonic merge);
// the (N+1)-th is the smallest of A[N] and the larger halves max(A[i], S[N-1-i]).
template <int N>
MHIP_HD void merge_into(float (&A)[N + 1], const float (&S)[N]) {
    static_assert((N & (N - 1)) == 0, "power of two");
    // in place: A[i] <- min(A[i], S[N-1-i]); the larger of each pair can only
    // be the (N+1)-th smallest
    float big = A[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        const float a = A[i], s = S[N - 1 - i];
        A[i] = a < s ? a : s;
        const float mx = a < s ? s : a;
        big = mx < big ? mx : big;
    }
    // bitonic merge of A[0..N-1]: half-cleaners of width N/2, N/4, ..., 1
#pragma unroll
    for (int w = N / 2; w >= 1; w >>= 1) {
#pragma unroll
        for (int i = 0; i < N; i++) {
            if ((i & w) == 0) cmpx(A[i], A[i + w]);
        }
    }
    A[N] = big;
}

}  // namespace mhip
