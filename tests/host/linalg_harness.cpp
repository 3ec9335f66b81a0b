// Host harness for the device linear algebra in mhip_linalg.hpp (compiled
// with g++ by tests/test_linalg_host.py).  fit(): weighted least-squares
// derivatives of one particle from its neighbour offsets, exactly as the
// kernels assemble them (E = sum w a a^T, b = sum w a df, LDL^T solve).
#include "../../meshoid_amd/csrc/mhip_linalg.hpp"

using namespace mhip;

template <int DIM, int ORDER>
static int fit_t(int k, const double *dx, const double *w, const double *df, double *out) {
    constexpr int ND = NDer<DIM, ORDER>::ND, NE = NDer<DIM, ORDER>::NE;
    double E[NE] = {}, B[ND] = {};
    for (int j = 0; j < k; j++) {
        double d[3] = {0, 0, 0};
        for (int c = 0; c < DIM; c++) d[c] = dx[j * DIM + c];
        double a[ND];
        basis<DIM, ORDER, double>(d, a);
        for (int i = 0; i < ND; i++) {
            for (int jj = 0; jj <= i; jj++) E[tri(i, jj)] += w[j] * a[i] * a[jj];
            B[i] += w[j] * a[i] * df[j];
        }
    }
    int ok = ldlt<ND>(E);
    ldlt_solve<ND>(E, B);
    for (int i = 0; i < ND; i++) out[i] = B[i];
    return ok ? 0 : 1;
}

extern "C" int fit(int dim, int order, int k, const double *dx, const double *w, const double *df,
                   double *out) {
    if (order == 1) {
        if (dim == 1) return fit_t<1, 1>(k, dx, w, df, out);
        if (dim == 2) return fit_t<2, 1>(k, dx, w, df, out);
        return fit_t<3, 1>(k, dx, w, df, out);
    }
    if (dim == 1) return fit_t<1, 2>(k, dx, w, df, out);
    if (dim == 2) return fit_t<2, 2>(k, dx, w, df, out);
    return fit_t<3, 2>(k, dx, w, df, out);
}
