/*
 * meshoid_oracle.c -- strict-FP64 CPU restatement of meshoid's neighbour-weighted
 * operator path.  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and
 * the CPU-baseline leg of the benchmark link this file as the checker.  The
 * product path (meshoid_amd + libmeshoid_hip.so) never calls it.
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off -fPIC -shared (see oracle/Makefile).
 * -ffp-contract=off keeps every multiply and add separately rounded, so the
 * distance arithmetic is the IEEE sequence ((0+dx*dx)+dy*dy)+dz*dz.
 *
 * Functions and the reference code each one restates:
 *   orc_knn                  scipy.spatial.cKDTree(pos, boxsize).query(x, k)
 *                            (call sites src/meshoid/Meshoid.py:145 and :157):
 *                            exact Euclidean kNN, self included, ascending,
 *                            min-image distances when boxsize > 0.
 *   orc_hsml_iter            HsmlIter, src/meshoid/kernel_density.py:14-57
 *                            (bisection on the kernel-weighted neighbour number).
 *   orc_kernel               Kernel, src/meshoid/kernel_density.py:82-96.
 *   orc_gradient_weights     gradient_weights, src/meshoid/derivatives.py:117-138,
 *                            with kernel_dx_and_weights (:26-52), nearest_image
 *                            (:8-14), polyfit_leastsq_matrices (:55-103) and
 *                            get_num_derivs (:106-114); np.linalg.inv is restated
 *                            as Gauss-Jordan elimination with partial pivoting.
 *   orc_apply_weights        np.einsum("ikj,ij...->i...k", G, f[ngb]-f[mask,None]),
 *                            src/meshoid/Meshoid.py:258,268 (D) and :291,295 (D2).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* cubic-spline kernel, kernel_density.py:82-96                             */
/* ------------------------------------------------------------------------ */
double orc_kernel(double q) {
    if (q <= 0.5) return 1 - 6 * q * q + 6 * q * q * q;
    if (q <= 1.0) return 2 * (1 - q) * (1 - q) * (1 - q);
    return 0.0;
}

static double norm_for_dim(int dim) {
    /* kernel_density.py:26-31 */
    if (dim == 3) return 32.0 / 3;
    if (dim == 2) return 40.0 / 7;
    return 8.0 / 3;
}

/* ------------------------------------------------------------------------ */
/* HsmlIter restatement, kernel_density.py:36-56.  max_iter caps the while   */
/* loop (the reference has no cap); iterations used are returned per row.    */
/* ------------------------------------------------------------------------ */
void orc_hsml_iter(const double *dists, int64_t n, int k, int dim, double error_norm,
                   int max_iter, double *hsml, int32_t *iters) {
    const double norm = norm_for_dim(dim);
    const double bound_coeff = 1.0 / (1 - pow(2 * norm, -1.0 / 3));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        const double *d = dists + i * (int64_t)k;
        double upper = d[k - 1] * bound_coeff;
        double lower = d[1];
        double error = 1e100, h = 0.0;
        int it = 0;
        while (error > error_norm && it < max_iter) {
            h = (upper + lower) / 2;
            double n_ngb = 0.0;
            for (int j = 0; j < k; j++) {
                double q = d[j] / h;
                if (q <= 0.5)
                    n_ngb += 1 - 6 * q * q + 6 * q * q * q;
                else if (q <= 1.0)
                    n_ngb += 2 * (1 - q) * (1 - q) * (1 - q);
            }
            n_ngb *= norm;
            if (n_ngb > k)
                upper = h;
            else
                lower = h;
            error = fabs(n_ngb - k);
            it++;
        }
        hsml[i] = h;
        if (iters) iters[i] = it;
    }
}

/* ------------------------------------------------------------------------ */
/* exact kNN: kd-tree with bounding boxes, min-image metric when periodic   */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t start, end; /* range in the permutation */
    int64_t left, right; /* children, -1 for leaf */
    double lo[3], hi[3];
} kd_node;

typedef struct {
    const double *pos;
    int dim;
    int64_t *perm;
    kd_node *nodes;
    int64_t n_nodes, cap;
} kd_tree;

static double coord(const kd_tree *t, int64_t p, int c) { return t->pos[p * t->dim + c]; }

static int64_t kd_build_rec(kd_tree *t, int64_t start, int64_t end) {
    if (t->n_nodes == t->cap) {
        t->cap *= 2;
        t->nodes = (kd_node *)realloc(t->nodes, sizeof(kd_node) * t->cap);
    }
    int64_t id = t->n_nodes++;
    kd_node nd;
    nd.start = start;
    nd.end = end;
    nd.left = nd.right = -1;
    for (int c = 0; c < 3; c++) {
        nd.lo[c] = 0;
        nd.hi[c] = 0;
    }
    for (int c = 0; c < t->dim; c++) {
        double lo = INFINITY, hi = -INFINITY;
        for (int64_t i = start; i < end; i++) {
            double v = coord(t, t->perm[i], c);
            if (v < lo) lo = v;
            if (v > hi) hi = v;
        }
        nd.lo[c] = lo;
        nd.hi[c] = hi;
    }
    if (end - start > 16) {
        int ax = 0;
        double best = -1;
        for (int c = 0; c < t->dim; c++)
            if (nd.hi[c] - nd.lo[c] > best) {
                best = nd.hi[c] - nd.lo[c];
                ax = c;
            }
        if (best > 0) {
            /* quickselect the median along ax */
            int64_t m = start + (end - start) / 2;
            int64_t l = start, r = end - 1;
            while (l < r) {
                double pv = coord(t, t->perm[(l + r) / 2], ax);
                int64_t i = l, j = r;
                while (i <= j) {
                    while (coord(t, t->perm[i], ax) < pv) i++;
                    while (coord(t, t->perm[j], ax) > pv) j--;
                    if (i <= j) {
                        int64_t tmp = t->perm[i];
                        t->perm[i] = t->perm[j];
                        t->perm[j] = tmp;
                        i++;
                        j--;
                    }
                }
                if (m <= j)
                    r = j;
                else if (m >= i)
                    l = i;
                else
                    break;
            }
            t->nodes[id] = nd;
            int64_t left = kd_build_rec(t, start, m);
            int64_t right = kd_build_rec(t, m, end);
            t->nodes[id].left = left;
            t->nodes[id].right = right;
            return id;
        }
    }
    t->nodes[id] = nd;
    return id;
}

static double wrap1(double dx, double box) {
    /* nearest_image, derivatives.py:8-14 */
    if (box > 0 && fabs(dx) > box / 2) return -copysign(box - fabs(dx), dx);
    return dx;
}

static double point_d2(const kd_tree *t, const double *q, int64_t p, double box) {
    double s = 0.0;
    for (int c = 0; c < t->dim; c++) {
        double dx = wrap1(coord(t, p, c) - q[c], box);
        s = s + dx * dx;
    }
    return s;
}

static double box_d2(const kd_tree *t, const kd_node *nd, const double *q, double box) {
    double s = 0.0;
    for (int c = 0; c < t->dim; c++) {
        double d;
        if (box > 0) {
            /* distance from q to the interval [lo,hi] on a circle of length box */
            if (q[c] >= nd->lo[c] && q[c] <= nd->hi[c])
                d = 0;
            else {
                double a = fabs(nd->lo[c] - q[c]), b = fabs(q[c] - nd->hi[c]);
                a = fmin(a, box - a);
                b = fmin(b, box - b);
                d = fmin(a, b);
                if (nd->hi[c] - nd->lo[c] >= box / 2) d = 0; /* conservative */
            }
        } else {
            d = 0;
            if (q[c] < nd->lo[c]) d = nd->lo[c] - q[c];
            if (q[c] > nd->hi[c]) d = q[c] - nd->hi[c];
        }
        s += d * d;
    }
    return s;
}

/* bounded sorted list keyed by (d2, index) */
typedef struct {
    double *d2;
    int64_t *idx;
    int k, n;
} klist;

static void klist_push(klist *L, double d2, int64_t idx) {
    if (L->n == L->k) {
        double wd = L->d2[L->k - 1];
        int64_t wi = L->idx[L->k - 1];
        if (d2 > wd || (d2 == wd && idx > wi)) return;
        L->n--;
    }
    int j = L->n;
    while (j > 0 && (L->d2[j - 1] > d2 || (L->d2[j - 1] == d2 && L->idx[j - 1] > idx))) {
        L->d2[j] = L->d2[j - 1];
        L->idx[j] = L->idx[j - 1];
        j--;
    }
    L->d2[j] = d2;
    L->idx[j] = idx;
    L->n++;
}

static void kd_query_rec(const kd_tree *t, int64_t node, const double *q, double box, klist *L) {
    const kd_node *nd = &t->nodes[node];
    if (L->n == L->k && box_d2(t, nd, q, box) > L->d2[L->k - 1]) return;
    if (nd->left < 0) {
        for (int64_t i = nd->start; i < nd->end; i++) {
            int64_t p = t->perm[i];
            klist_push(L, point_d2(t, q, p, box), p);
        }
        return;
    }
    const kd_node *a = &t->nodes[nd->left], *b = &t->nodes[nd->right];
    double da = box_d2(t, a, q, box), db = box_d2(t, b, q, box);
    if (da <= db) {
        kd_query_rec(t, nd->left, q, box, L);
        kd_query_rec(t, nd->right, q, box, L);
    } else {
        kd_query_rec(t, nd->right, q, box, L);
        kd_query_rec(t, nd->left, q, box, L);
    }
}

/*
 * Exact kNN.  pos: (n, dim) f64.  queries: (nq, dim) f64 query coordinates.
 * Output rows ascending by distance (ties by index), self included when the
 * query is a particle.  Missing neighbours (k > n) get index n and inf, as
 * cKDTree returns them.
 */
int orc_knn(const double *pos, int64_t n, int dim, double boxsize, const double *queries,
            int64_t nq, int k, int64_t *ngb, double *dist) {
    if (dim < 1 || dim > 3 || k < 1) return -1;
    kd_tree t;
    t.pos = pos;
    t.dim = dim;
    t.perm = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) t.perm[i] = i;
    t.cap = 1024;
    t.n_nodes = 0;
    t.nodes = (kd_node *)malloc(sizeof(kd_node) * t.cap);
    if (n > 0) kd_build_rec(&t, 0, n);
#pragma omp parallel
    {
        double *bd = (double *)malloc(sizeof(double) * k);
        int64_t *bi = (int64_t *)malloc(sizeof(int64_t) * k);
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < nq; i++) {
            klist L = {bd, bi, k, 0};
            double q[3] = {0, 0, 0};
            for (int c = 0; c < dim; c++) q[c] = queries[i * dim + c];
            if (n > 0) kd_query_rec(&t, 0, q, boxsize, &L);
            for (int j = 0; j < k; j++) {
                if (j < L.n) {
                    ngb[i * k + j] = L.idx[j];
                    dist[i * k + j] = sqrt(L.d2[j]);
                } else {
                    ngb[i * k + j] = n;
                    dist[i * k + j] = INFINITY;
                }
            }
        }
        free(bd);
        free(bi);
    }
    free(t.perm);
    free(t.nodes);
    return 0;
}

/* brute force kNN (O(n*nq)), used only to pin orc_knn on small inputs */
int orc_knn_brute(const double *pos, int64_t n, int dim, double boxsize, const double *queries,
                  int64_t nq, int k, int64_t *ngb, double *dist) {
    kd_tree t;
    t.pos = pos;
    t.dim = dim;
#pragma omp parallel
    {
        double *bd = (double *)malloc(sizeof(double) * k);
        int64_t *bi = (int64_t *)malloc(sizeof(int64_t) * k);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < nq; i++) {
            klist L = {bd, bi, k, 0};
            double q[3] = {0, 0, 0};
            for (int c = 0; c < dim; c++) q[c] = queries[i * dim + c];
            for (int64_t p = 0; p < n; p++) klist_push(&L, point_d2(&t, q, p, boxsize), p);
            for (int j = 0; j < k; j++) {
                ngb[i * k + j] = j < L.n ? L.idx[j] : n;
                dist[i * k + j] = j < L.n ? sqrt(L.d2[j]) : INFINITY;
            }
        }
        free(bd);
        free(bi);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* least-squares gradient weights, derivatives.py:26-138                     */
/* ------------------------------------------------------------------------ */
int orc_num_derivs(int dim, int order) {
    /* get_num_derivs, derivatives.py:106-114 */
    if (order == 1) return dim;
    return dim == 1 ? 2 : (dim == 2 ? 5 : 9);
}

/* Gauss-Jordan inverse with partial pivoting; returns 0, or 1 if a pivot is 0 */
static int invert(double *a, double *inv, int m) {
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) inv[i * m + j] = (i == j);
    for (int c = 0; c < m; c++) {
        int p = c;
        for (int r = c + 1; r < m; r++)
            if (fabs(a[r * m + c]) > fabs(a[p * m + c])) p = r;
        if (a[p * m + c] == 0.0) return 1;
        if (p != c)
            for (int j = 0; j < m; j++) {
                double t = a[c * m + j];
                a[c * m + j] = a[p * m + j];
                a[p * m + j] = t;
                t = inv[c * m + j];
                inv[c * m + j] = inv[p * m + j];
                inv[p * m + j] = t;
            }
        double piv = a[c * m + c];
        for (int j = 0; j < m; j++) {
            a[c * m + j] /= piv;
            inv[c * m + j] /= piv;
        }
        for (int r = 0; r < m; r++) {
            if (r == c) continue;
            double f = a[r * m + c];
            if (f == 0.0) continue;
            for (int j = 0; j < m; j++) {
                a[r * m + j] -= f * a[c * m + j];
                inv[r * m + j] -= f * inv[c * m + j];
            }
        }
    }
    return 0;
}

/*
 * pos (n, dim) f64; ngb (nq, k) int64 indices into pos; h (nq,); indices (nq,)
 * the particle index of each row.  result (nq, nd, k) f64.  status[i] = 1 when
 * the moment matrix of row i is exactly singular (np.linalg.inv would raise).
 */
int orc_gradient_weights(const double *pos, int64_t n, int dim, const int64_t *ngb, int64_t nq,
                         int k, const double *h, const int64_t *indices, double boxsize,
                         int weighted, int order, double *result, int32_t *status) {
    (void)n;
    const int nd = orc_num_derivs(dim, order);
    int bad = 0;
#pragma omp parallel
    {
        double *dx = (double *)malloc(sizeof(double) * k * dim);
        double *w = (double *)malloc(sizeof(double) * k);
        double *rhs = (double *)malloc(sizeof(double) * nd * k);
        double lhs[81], inv[81];
#pragma omp for schedule(static) reduction(| : bad)
        for (int64_t i = 0; i < nq; i++) {
            int64_t index = indices[i];
            /* kernel_dx_and_weights, derivatives.py:42-51 */
            for (int j = 0; j < k; j++) {
                double r = 0;
                int64_t nb = ngb[i * k + j];
                for (int c = 0; c < dim; c++) {
                    double d = pos[nb * dim + c] - pos[index * dim + c];
                    if (boxsize > 0) d = wrap1(d, boxsize);
                    dx[j * dim + c] = d;
                    r += d * d;
                }
                w[j] = weighted ? orc_kernel(sqrt(r) / h[i]) : 1.0;
            }
            /* polyfit_leastsq_matrices, derivatives.py:69-101 */
            memset(lhs, 0, sizeof(lhs));
            if (order == 1) {
                for (int j = 0; j < k; j++)
                    for (int a = 0; a < dim; a++) rhs[a * k + j] = w[j] * dx[j * dim + a];
                for (int j = 0; j < k; j++)
                    for (int a = 0; a < dim; a++)
                        for (int b = 0; b < dim; b++)
                            lhs[a * nd + b] += rhs[a * k + j] * dx[j * dim + b];
            } else {
                for (int j = 0; j < k; j++)
                    for (int a = 0; a < nd; a++) {
                        double v;
                        if (a < dim)
                            v = dx[j * dim + a];
                        else if (a < 2 * dim) {
                            int d = a - dim;
                            v = 0.5 * dx[j * dim + d] * dx[j * dim + d];
                        } else {
                            int d = a - 2 * dim;
                            v = dx[j * dim + d % dim] * dx[j * dim + (d + 1) % dim];
                        }
                        rhs[a * k + j] = v;
                    }
                for (int j = 0; j < k; j++)
                    for (int a = 0; a < nd; a++)
                        for (int b = 0; b < nd; b++)
                            lhs[a * nd + b] += rhs[a * k + j] * rhs[b * k + j] * w[j];
                for (int j = 0; j < k; j++)
                    for (int a = 0; a < nd; a++) rhs[a * k + j] *= w[j];
            }
            int sing = invert(lhs, inv, nd);
            if (status) status[i] = sing;
            bad |= sing;
            /* result = inv @ rhs, derivatives.py:133-136 */
            double *out = result + i * (int64_t)nd * k;
            for (int a = 0; a < nd; a++)
                for (int j = 0; j < k; j++) {
                    double s = 0;
                    for (int b = 0; b < nd; b++) s += inv[a * nd + b] * rhs[b * k + j];
                    out[a * k + j] = sing ? NAN : s;
                }
        }
        free(dx);
        free(w);
        free(rhs);
    }
    return bad;
}

/*
 * out[i, f, a] = sum_j G[i, row0 + a, j] * (fld[ngb[i,j], f] - fld[indices[i], f])
 * for a in [0, nrows): the einsum of Meshoid.D / Meshoid.D2.
 * G: (nq, nd, k); fld: (nf_rows, K) f64; out (nq, K, nrows).
 */
void orc_apply_weights(const double *G, int64_t nq, int nd, int k, int row0, int nrows,
                       const int64_t *ngb, const int64_t *indices, const double *fld, int K,
                       double *out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nq; i++) {
        const double *g = G + i * (int64_t)nd * k;
        for (int f = 0; f < K; f++) {
            double f0 = fld[indices[i] * K + f];
            for (int a = 0; a < nrows; a++) {
                double s = 0;
                for (int j = 0; j < k; j++)
                    s += g[(row0 + a) * k + j] * (fld[ngb[i * k + j] * K + f] - f0);
                out[(i * K + f) * nrows + a] = s;
            }
        }
    }
}

int orc_num_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
