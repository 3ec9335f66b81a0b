// Small dense linear algebra shared by the derivative kernels: the
// polynomial basis of the least-squares fit and an in-register LDL^T of the
// symmetric moment matrix.  Header-only and usable from host code (the
// CPU unit test compiles it with g++).
#pragma once

#if defined(__HIPCC__)
#define MHIP_HD __host__ __device__ __forceinline__
#else
#define MHIP_HD inline
#endif

namespace mhip {

template <int DIM, int ORDER>
struct NDer {
    static constexpr int ND = ORDER == 1 ? DIM : (DIM == 1 ? 2 : (DIM == 2 ? 5 : 9));
    static constexpr int NE = ND * (ND + 1) / 2;
};

MHIP_HD constexpr int tri(int i, int j) { return i * (i + 1) / 2 + j; }

// polynomial basis of polyfit_leastsq_matrices (derivatives.py:83-92):
// [dx.., 0.5 dx_d^2.., dx_d dx_{(d+1)%dim}..]
template <int DIM, int ORDER, typename T>
MHIP_HD void basis(const T (&d)[3], T (&a)[NDer<DIM, ORDER>::ND]) {
    constexpr int ND = NDer<DIM, ORDER>::ND;
#pragma unroll
    for (int i = 0; i < ND; i++) {
        if (i < DIM)
            a[i] = d[i];
        else if (i < 2 * DIM)
            a[i] = T(0.5) * d[i - DIM] * d[i - DIM];
        else {
            const int c = i - 2 * DIM;
            a[i] = d[c % DIM] * d[(c + 1) % DIM];
        }
    }
}

// true when a diagonal entry is exactly zero: a whole basis direction is
// absent, so every LU factorisation (np.linalg.inv) meets a zero pivot
template <int M>
MHIP_HD bool zero_diagonal(const double (&E)[M * (M + 1) / 2]) {
    bool z = false;
#pragma unroll
    for (int j = 0; j < M; j++) z = z || E[tri(j, j)] == 0.0;
    return z;
}

// in-place LDL^T of the lower triangle; returns false on a zero/NaN pivot
template <int M>
MHIP_HD bool ldlt(double (&E)[M * (M + 1) / 2]) {
    bool ok = true;
#pragma unroll
    for (int j = 0; j < M; j++) {
        double d = E[tri(j, j)];
#pragma unroll
        for (int k = 0; k < j; k++) d -= E[tri(j, k)] * E[tri(j, k)] * E[tri(k, k)];
        E[tri(j, j)] = d;
        if (d == 0.0 || d != d) ok = false;
        const double id = 1.0 / d;
#pragma unroll
        for (int i = j + 1; i < M; i++) {
            double v = E[tri(i, j)];
#pragma unroll
            for (int k = 0; k < j; k++) v -= E[tri(i, k)] * E[tri(j, k)] * E[tri(k, k)];
            E[tri(i, j)] = v * id;
        }
    }
    return ok;
}

template <int M>
MHIP_HD void ldlt_solve(const double (&E)[M * (M + 1) / 2], double (&x)[M]) {
#pragma unroll
    for (int i = 0; i < M; i++) {
#pragma unroll
        for (int k = 0; k < i; k++) x[i] -= E[tri(i, k)] * x[k];
    }
#pragma unroll
    for (int i = 0; i < M; i++) x[i] /= E[tri(i, i)];
#pragma unroll
    for (int i = M - 1; i >= 0; i--) {
#pragma unroll
        for (int k = i + 1; k < M; k++) x[i] -= E[tri(k, i)] * x[k];
    }
}

}  // namespace mhip
