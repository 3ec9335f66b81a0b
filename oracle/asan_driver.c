/*
 * CPU sanitizer driver for the FP64 oracle (test infrastructure only; built
 * by `make -C oracle asan` with -fsanitize=address,undefined and run by
 * tests/test_sanitizers.py).  Exercises every exported oracle routine on
 * small seeded inputs -- kNN (kd-tree and brute force, open and periodic,
 * 1D-3D, k > n), HsmlIter, gradient_weights of orders 1 and 2 (weighted,
 * unweighted, periodic, singular rows) and the weight contraction -- and
 * checks that kd-tree and brute force agree, so a memory error or undefined
 * behaviour anywhere on those paths fails the run.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int orc_knn(const double *pos, int64_t n, int dim, double boxsize, const double *queries, int64_t nq, int k,
            int64_t *ngb, double *dist);
int orc_knn_brute(const double *pos, int64_t n, int dim, double boxsize, const double *queries, int64_t nq,
                  int k, int64_t *ngb, double *dist);
void orc_hsml_iter(const double *dists, int64_t n, int k, int dim, double error_norm, int max_iter,
                   double *hsml, int32_t *iters);
int orc_num_derivs(int dim, int order);
int orc_gradient_weights(const double *pos, int64_t n, int dim, const int64_t *ngb, int64_t nq, int k,
                         const double *h, const int64_t *indices, double boxsize, int weighted, int order,
                         double *result, int32_t *status);
void orc_apply_weights(const double *G, int64_t nq, int nd, int k, int row0, int nrows, const int64_t *ngb,
                       const int64_t *indices, const double *fld, int K, double *out);

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;
static double urand(void) {
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}

static int run_case(int64_t n, int dim, int k, double box, int planar) {
    double *pos = malloc(sizeof(double) * n * dim);
    for (int64_t i = 0; i < n * dim; i++) pos[i] = urand() * (box > 0 ? box : 1.0);
    if (planar && dim == 3)
        for (int64_t i = 0; i < n; i++) pos[i * 3 + 2] = 0.25;
    const int kk = k < n ? k : (int)n;
    int64_t *ngb = malloc(sizeof(int64_t) * n * k), *ngb2 = malloc(sizeof(int64_t) * n * k);
    double *d = malloc(sizeof(double) * n * k), *d2 = malloc(sizeof(double) * n * k);
    if (orc_knn(pos, n, dim, box, pos, n, k, ngb, d) || orc_knn_brute(pos, n, dim, box, pos, n, k, ngb2, d2)) {
        fprintf(stderr, "knn failed\n");
        return 1;
    }
    for (int64_t i = 0; i < n * k; i++)
        if (d[i] != d2[i] && !(isinf(d[i]) && isinf(d2[i]))) {
            fprintf(stderr, "kd-tree and brute force disagree (n=%lld dim=%d k=%d)\n", (long long)n, dim, k);
            return 1;
        }
    int rc = 0;
    if (kk == k) {
        double *h = malloc(sizeof(double) * n);
        int32_t *it = malloc(sizeof(int32_t) * n), *st = malloc(sizeof(int32_t) * n);
        int64_t *idx = malloc(sizeof(int64_t) * n);
        orc_hsml_iter(d, n, k, dim, 1e-13, 200, h, it);
        for (int64_t i = 0; i < n; i++) idx[i] = i;
        double *fld = malloc(sizeof(double) * n * 2);
        for (int64_t i = 0; i < n * 2; i++) fld[i] = urand();
        for (int order = 1; order <= 2; order++)
            for (int weighted = 0; weighted <= 1; weighted++) {
                const int nd = orc_num_derivs(dim, order);
                double *G = malloc(sizeof(double) * n * nd * k);
                double *out = malloc(sizeof(double) * n * 2 * nd);
                (void)orc_gradient_weights(pos, n, dim, ngb, n, k, h, idx, box, weighted, order, G, st);
                orc_apply_weights(G, n, nd, k, 0, nd, ngb, idx, fld, 2, out);
                free(G);
                free(out);
            }
        free(h);
        free(it);
        free(st);
        free(idx);
        free(fld);
    }
    free(pos);
    free(ngb);
    free(ngb2);
    free(d);
    free(d2);
    return rc;
}

int main(void) {
    int bad = 0;
    bad |= run_case(700, 3, 32, -1.0, 0);
    bad |= run_case(700, 3, 32, 1.0, 0);
    bad |= run_case(500, 2, 20, -1.0, 0);
    bad |= run_case(300, 1, 4, -1.0, 0);
    bad |= run_case(300, 3, 16, -1.0, 1); /* exactly singular moment matrices */
    bad |= run_case(5, 3, 8, -1.0, 0);    /* k > n: missing neighbours */
    if (!bad) printf("oracle sanitizer run: ok\n");
    return bad;
}
