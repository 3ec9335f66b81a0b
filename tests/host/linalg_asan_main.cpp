// Sanitizer driver for the device linear algebra header (test
// infrastructure; built with -fsanitize=address,undefined by
// tests/test_sanitizers.py): weighted least-squares fits of every
// (dim, order) through the host harness, on seeded offsets, including a
// rank-deficient row.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

extern "C" int fit(int dim, int order, int k, const double *dx, const double *w, const double *df, double *out);

int main() {
    std::mt19937_64 g(7);
    std::normal_distribution<double> nd(0.0, 1.0);
    int bad = 0;
    for (int dim = 1; dim <= 3; dim++)
        for (int order = 1; order <= 2; order++)
            for (int rep = 0; rep < 50; rep++) {
                const int k = 32;
                std::vector<double> dx(k * dim), w(k), df(k), out(9);
                for (auto &v : dx) v = 1e-3 * nd(g);
                for (auto &v : w) v = std::fabs(nd(g));
                for (auto &v : df) v = nd(g);
                if (rep == 0)
                    for (int j = 0; j < k; j++) dx[j * dim] = 0.0;  // one basis direction absent
                (void)fit(dim, order, k, dx.data(), w.data(), df.data(), out.data());
                if (rep > 0)
                    for (int i = 0; i < (order == 1 ? dim : 9) && i < 9; i++)
                        if (dim == 3 && !std::isfinite(out[i])) bad = 1;
            }
    if (!bad) std::printf("linalg sanitizer run: ok\n");
    return bad;
}
