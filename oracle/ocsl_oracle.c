/*
 * ocsl_oracle.c -- CPU restatement of the reference's sparse-Krylov hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (tests, smoke) and the
 * timed CPU baseline (bench.py).  Never part of the product path.
 *
 * Each function cites the reference file:line whose mathematics it restates.
 * Loop structure follows the reference (row-serial CSR SpMV, separate BLAS-1
 * passes, OpenMP reductions) so that the timed baseline is representative.
 * With one thread every reduction runs left to right.
 */
#include "ocsl_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;

void or_set_threads(int t) { g_threads = t > 0 ? t : 1; }
int or_get_threads(void) { return g_threads; }

#define OMP_FOR _Pragma("omp parallel for num_threads(g_threads) schedule(static)")

static double* dvec(int64_t n) { return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

/* cmatrix_solver.cpp:233-241 */
double or_norm(int64_t n, const double* a, const double* b) {
    double s = 0.0;
    if (g_threads == 1) {
        for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
        return s;
    }
#pragma omp parallel for num_threads(g_threads) schedule(static) reduction(+ : s)
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* cmatrix_solver.cpp:244-252 */
double or_sum(int64_t n, const double* a) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i];
    return s;
}

/* cmatrix_solver.cpp:255-263 (abs of a double) */
double or_abssum(int64_t n, const double* a) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += fabs(a[i]);
    return s;
}

/* cmatrix_solver.cpp:284-335: out = a*x + b*y (the +-1 special cases give the
 * same IEEE result as the general form). */
void or_axpy(int64_t n, double a, const double* x, double b, const double* y, double* out) {
    OMP_FOR
    for (int64_t i = 0; i < n; ++i) out[i] = a * x[i] + b * y[i];
}

/* cmatrix_solver.cpp:338-453: out_i = a * (sum_j A_ij x_j) + b * z_i, the row
 * sum taken left to right; z may be NULL (b ignored).  Sentinel columns are
 * skipped: the reference multiplies their 0.0 by an out-of-range load. */
void or_gemv(int64_t n, const uint64_t* rp, const uint64_t* col, const double* val, double a,
             const double* x, double b, const double* z, double* out) {
    OMP_FOR
    for (int64_t i = 0; i < n; ++i) {
        double dum = 0.0;
        for (uint64_t p = rp[i]; p < rp[i + 1]; ++p)
            if (col[p] < (uint64_t)n) dum += val[p] * x[col[p]];
        double o = (a == 1.0) ? dum : a * dum;
        if (z) o = (b == 1.0) ? o + z[i] : (b == -1.0 ? o - z[i] : o + b * z[i]);
        out[i] = o;
    }
}

/* cmatrix_solver.cpp:455-469 */
void or_diag_index(int64_t n, const uint64_t* rp, const uint64_t* col, int64_t* diag) {
    for (int64_t i = 0; i < n; ++i) {
        diag[i] = -1;
        for (uint64_t p = rp[i]; p < rp[i + 1]; ++p)
            if (col[p] == (uint64_t)i) { diag[i] = (int64_t)p; break; }
    }
}

/* CCSR::get_element, cmatrix_solver.cpp:47-55: linear scan, 0 when absent. */
static double get_el(const uint64_t* rp, const uint64_t* col, const double* v, uint64_t i,
                     uint64_t j) {
    for (uint64_t p = rp[i]; p < rp[i + 1]; ++p)
        if (col[p] == j) return v[p];
    return 0.0;
}

/* cmatrix_solver.cpp:519-551.  Each sweep reads the previous iterate `old`
 * and writes `nw`; lower entries (col < i) are divided by the previous
 * diagonal of row col. */
void or_ilu_decompose(int64_t n, const uint64_t* rp, const uint64_t* col, const double* a_val,
                      const int64_t* diag, int sweeps, double* m_val) {
    const uint64_t nnz = rp[n];
    double* old = (double*)malloc(sizeof(double) * (nnz ? nnz : 1));
    memcpy(m_val, a_val, sizeof(double) * nnz);
    for (int s = 0; s < sweeps; ++s) {
        memcpy(old, m_val, sizeof(double) * nnz);
        OMP_FOR
        for (int64_t i = 0; i < n; ++i) {
            for (uint64_t j = rp[i]; j < rp[i + 1]; ++j) {
                const uint64_t cj = col[j];
                if ((uint64_t)i > cj) {
                    double dum = 0.0;
                    for (uint64_t k = rp[i]; k < j; ++k) dum += old[k] * get_el(rp, col, old, col[k], cj);
                    m_val[j] = (a_val[j] - dum) / old[diag[cj]];
                } else {
                    double dum = 0.0;
                    for (uint64_t k = rp[i]; (int64_t)k < diag[i]; ++k)
                        dum += old[k] * get_el(rp, col, old, col[k], cj);
                    m_val[j] = a_val[j] - dum;
                }
            }
        }
    }
    free(old);
}

/* Exact IKJ ILU(0), cmatrix_solver.cpp:474-518 (commented-out there). */
void or_ilu_exact(int64_t n, const uint64_t* rp, const uint64_t* col, const double* a_val,
                  const int64_t* diag, double* m_val) {
    const uint64_t nnz = rp[n];
    memcpy(m_val, a_val, sizeof(double) * nnz);
    int64_t* iw = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) iw[i] = -1;
    for (int64_t k = 0; k < n; ++k) {
        for (uint64_t p = rp[k]; p < rp[k + 1]; ++p)
            if (col[p] < (uint64_t)n) iw[col[p]] = (int64_t)p;
        for (uint64_t p = rp[k]; p < rp[k + 1]; ++p) {
            const uint64_t jr = col[p];
            if (jr >= (uint64_t)k) break;
            const double tl = m_val[p] / m_val[diag[jr]];
            m_val[p] = tl;
            for (uint64_t q = (uint64_t)diag[jr] + 1; q < rp[jr + 1]; ++q) {
                if (col[q] >= (uint64_t)n) continue;
                const int64_t w = iw[col[q]];
                if (w >= 0) m_val[w] -= tl * m_val[q];
            }
        }
        for (uint64_t p = rp[k]; p < rp[k + 1]; ++p)
            if (col[p] < (uint64_t)n) iw[col[p]] = -1;
    }
    free(iw);
}

/* cmatrix_solver.cpp:555-571 */
void or_forsub(int64_t n, const uint64_t* rp, const uint64_t* col, const double* m,
               const int64_t* diag, const double* b, double* y) {
    y[0] = b[0];
    for (int64_t i = 1; i < n; ++i) {
        double s = 0.0;
        for (uint64_t p = rp[i]; (int64_t)p < diag[i]; ++p) s += m[p] * y[col[p]];
        y[i] = b[i] - s;
    }
}

/* cmatrix_solver.cpp:595-610 */
void or_backsub(int64_t n, const uint64_t* rp, const uint64_t* col, const double* m,
                const int64_t* diag, const double* y, double* x) {
    x[n - 1] = y[n - 1] / m[diag[n - 1]];
    for (int64_t i = n - 1; i >= 0; --i) {
        double s = 0.0;
        for (uint64_t p = (uint64_t)diag[i] + 1; p < rp[i + 1]; ++p)
            if (col[p] < (uint64_t)n) s += m[p] * x[col[p]];
        x[i] = (y[i] - s) / m[diag[i]];
    }
}

/* cmatrix_solver.cpp:573-592 (commented-out damped Jacobi forward sweeps) */
void or_forsub_jacobi(int64_t n, const uint64_t* rp, const uint64_t* col, const double* m,
                      const int64_t* diag, const double* b, int sweeps, double w, double* y) {
    double* yn = dvec(n);
    for (int s = 0; s < sweeps; ++s) {
        OMP_FOR
        for (int64_t i = 0; i < n; ++i) {
            double acc = 0.0;
            for (uint64_t p = rp[i]; (int64_t)p < diag[i]; ++p) acc += m[p] * y[col[p]];
            yn[i] = w * (b[i] - acc) + (1.0 - w) * y[i];
        }
        memcpy(y, yn, sizeof(double) * (size_t)n);
    }
    free(yn);
}

/* cmatrix_solver.cpp:612-631 (commented-out damped Jacobi backward sweeps) */
void or_backsub_jacobi(int64_t n, const uint64_t* rp, const uint64_t* col, const double* m,
                       const int64_t* diag, const double* y, int sweeps, double w, double* x) {
    double* xn = dvec(n);
    for (int s = 0; s < sweeps; ++s) {
        OMP_FOR
        for (int64_t i = 0; i < n; ++i) {
            double acc = 0.0;
            const double inv = 1.0 / m[diag[i]];
            for (uint64_t p = (uint64_t)diag[i] + 1; p < rp[i + 1]; ++p)
                if (col[p] < (uint64_t)n) acc += m[p] * x[col[p]];
            xn[i] = w * inv * (y[i] - acc) + (1.0 - w) * x[i];
        }
        memcpy(x, xn, sizeof(double) * (size_t)n);
    }
    free(xn);
}

/* CMatrix_solver::CG_solve, cmatrix_solver.cpp:635-680: at most 4 iterations,
 * two SpMVs per iteration, exit when the absolute ||r'||^2 < 1e-6. */
double or_cg_reference(int64_t n, const uint64_t* rp, const uint64_t* col, const double* val,
                       const double* b, double* x, double* hist, int* n_hist) {
    const int n_iter = 4;
    const double res_limit = 1.0e-6;
    double res = 0.0;
    double *r = dvec(n), *rn = dvec(n), *p = dvec(n), *pn = dvec(n), *q = dvec(n);
    int nh = 0;
    or_gemv(n, rp, col, val, -1.0, x, 1.0, b, r);
    memcpy(p, r, sizeof(double) * (size_t)n);
    for (int it = 0; it < n_iter; ++it) {
        or_gemv(n, rp, col, val, 1.0, p, 0.0, NULL, q);
        const double a = or_norm(n, r, r) / (or_norm(n, p, q) + 1.0e-40);
        or_axpy(n, 1.0, x, a, p, x);
        or_gemv(n, rp, col, val, -1.0 * a, p, 1.0, r, rn);
        res = or_norm(n, rn, rn);
        if (hist) hist[nh] = res;
        ++nh;
        if (res < res_limit) break;
        const double beta = or_norm(n, rn, rn) / (or_norm(n, r, r) + 1.0e-40);
        or_axpy(n, 1.0, rn, beta, p, pn);
        memcpy(r, rn, sizeof(double) * (size_t)n);
        memcpy(p, pn, sizeof(double) * (size_t)n);
    }
    if (n_hist) *n_hist = nh;
    free(r); free(rn); free(p); free(pn); free(q);
    return res;
}

/* CMatrix_solver::ILU_BICGSTAB_solve, cmatrix_solver.cpp:682-762, with its
 * quirks: rho is never updated and beta is unused (:689, :715), the p update
 * uses omega (:719-720), omega comes from forward-substitution-only vectors
 * (:746-750), the early exit returns before x = h (:731-735). */
double or_ilu_bicgstab_reference(int64_t n, const uint64_t* rp, const uint64_t* col,
                                 const double* val, const double* b, double* x, double* hist,
                                 int* n_hist) {
    const int n_iter = 3;
    const double res_limit = 1.0e-20;
    const uint64_t nnz = rp[n];
    double res = 0.0;
    double alpha = 1.0, omega = 1.0, rhon;
    double *r = dvec(n), *p = dvec(n), *v = dvec(n), *r0 = dvec(n), *s = dvec(n), *dum = dvec(n);
    double *h = dvec(n), *t = dvec(n), *y = dvec(n), *z = dvec(n), *dum2 = dvec(n);
    double* m = (double*)malloc(sizeof(double) * (nnz ? nnz : 1));
    int64_t* diag = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int nh = 0;
    or_diag_index(n, rp, col, diag);
    or_ilu_decompose(n, rp, col, val, diag, 3, m);
    or_gemv(n, rp, col, val, -1.0, x, 1.0, b, r);
    memcpy(r0, r, sizeof(double) * (size_t)n);
    for (int it = 0; it < n_iter; ++it) {
        rhon = or_norm(n, r0, r);
        or_axpy(n, 1.0, p, -1.0 * omega, v, dum);
        or_axpy(n, 1.0, r, omega, dum, p);
        or_forsub(n, rp, col, m, diag, p, dum);
        or_backsub(n, rp, col, m, diag, dum, y);
        or_gemv(n, rp, col, val, 1.0, y, 0.0, NULL, v);
        const double dum1 = or_norm(n, r0, v);
        alpha = rhon / (dum1 + DBL_MIN);
        or_axpy(n, 1.0, x, alpha, y, h);
        res = or_norm(n, r, r) / or_norm(n, b, b) + DBL_MIN;
        if (hist) hist[nh] = res;
        ++nh;
        if (res < res_limit) break;
        or_axpy(n, -1.0 * alpha, v, 1.0, r, s);
        or_forsub(n, rp, col, m, diag, s, dum);
        or_backsub(n, rp, col, m, diag, dum, z);
        or_gemv(n, rp, col, val, 1.0, z, 0.0, NULL, t);
        or_forsub(n, rp, col, m, diag, t, dum);
        or_forsub(n, rp, col, m, diag, s, dum2);
        omega = or_norm(n, dum, dum2) / (or_norm(n, dum, dum) + 1.0e-40);
        or_axpy(n, 1.0, h, omega, z, x);
        or_axpy(n, -1.0 * omega, t, 1.0, s, r);
    }
    if (n_hist) *n_hist = nh;
    free(r); free(p); free(v); free(r0); free(s); free(dum);
    free(h); free(t); free(y); free(z); free(dum2); free(m); free(diag);
    return res;
}

static double true_relres(int64_t n, const uint64_t* rp, const uint64_t* col, const double* val,
                          const double* b, const double* x) {
    double* r = dvec(n);
    or_gemv(n, rp, col, val, -1.0, x, 1.0, b, r);
    const double rr = or_norm(n, r, r), bb = or_norm(n, b, b);
    free(r);
    return bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
}

/* Textbook (Hestenes-Stiefel) CG; the reference's CG_solve without its
 * 4-iteration cap, one SpMV per iteration, relative-residual stop. */
int or_cg_tol(int64_t n, const uint64_t* rp, const uint64_t* col, const double* val,
              const double* b, double* x, double tol, int maxit, double* hist,
              double* true_rel) {
    double *r = dvec(n), *p = dvec(n), *q = dvec(n);
    or_gemv(n, rp, col, val, -1.0, x, 1.0, b, r);
    memcpy(p, r, sizeof(double) * (size_t)n);
    const double bb = or_norm(n, b, b);
    const double scale = bb > 0.0 ? bb : 1.0;
    double rr = or_norm(n, r, r);
    int k = 0;
    if (hist) hist[0] = sqrt(rr / scale);
    while (k < maxit && sqrt(rr / scale) > tol) {
        or_gemv(n, rp, col, val, 1.0, p, 0.0, NULL, q);
        const double alpha = rr / or_norm(n, p, q);
        or_axpy(n, 1.0, x, alpha, p, x);
        or_axpy(n, 1.0, r, -alpha, q, r);
        const double rr_new = or_norm(n, r, r);
        const double beta = rr_new / rr;
        or_axpy(n, 1.0, r, beta, p, p);
        rr = rr_new;
        ++k;
        if (hist) hist[k] = sqrt(rr / scale);
    }
    if (true_rel) *true_rel = true_relres(n, rp, col, val, b, x);
    free(r); free(p); free(q);
    return k;
}

typedef struct {
    int64_t n;
    const uint64_t* rp;
    const uint64_t* col;
    const double* m;
    const int64_t* diag;
    int tri_sweeps;
    double tri_w;
    double* tmp;
} precond_t;

static void apply_precond(const precond_t* P, const double* in, double* out) {
    if (P->tri_sweeps == 0) {
        or_forsub(P->n, P->rp, P->col, P->m, P->diag, in, P->tmp);
        or_backsub(P->n, P->rp, P->col, P->m, P->diag, P->tmp, out);
    } else {
        memset(P->tmp, 0, sizeof(double) * (size_t)P->n);
        memset(out, 0, sizeof(double) * (size_t)P->n);
        or_forsub_jacobi(P->n, P->rp, P->col, P->m, P->diag, in, P->tri_sweeps, P->tri_w, P->tmp);
        or_backsub_jacobi(P->n, P->rp, P->col, P->m, P->diag, P->tmp, P->tri_sweeps, P->tri_w, out);
    }
}

/* Right-preconditioned BiCGSTAB (van der Vorst 1992) with ILU(0) factors as
 * built by the reference's fine-grained sweeps (cmatrix_solver.cpp:519-551). */
int or_bicgstab_tol(int64_t n, const uint64_t* rp, const uint64_t* col, const double* val,
                    const double* b, double* x, double tol, int maxit, int ilu_sweeps,
                    int tri_sweeps, double tri_w, double* hist, double* true_rel) {
    const uint64_t nnz = rp[n];
    double *r = dvec(n), *rh = dvec(n), *p = dvec(n), *v = dvec(n), *s = dvec(n), *t = dvec(n);
    double *yh = dvec(n), *zh = dvec(n), *tmp = dvec(n);
    double* m = (double*)malloc(sizeof(double) * (nnz ? nnz : 1));
    int64_t* diag = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    or_diag_index(n, rp, col, diag);
    if (ilu_sweeps > 0) or_ilu_decompose(n, rp, col, val, diag, ilu_sweeps, m);
    else or_ilu_exact(n, rp, col, val, diag, m);
    precond_t P = {n, rp, col, m, diag, tri_sweeps, tri_w, tmp};

    or_gemv(n, rp, col, val, -1.0, x, 1.0, b, r);
    memcpy(rh, r, sizeof(double) * (size_t)n);
    const double bb = or_norm(n, b, b);
    const double scale = bb > 0.0 ? bb : 1.0;
    double rr = or_norm(n, r, r);
    double rho_old = 1.0, alpha = 1.0, omega = 1.0;
    int k = 0;
    if (hist) hist[0] = sqrt(rr / scale);
    while (k < maxit && sqrt(rr / scale) > tol) {
        const double rho = or_norm(n, rh, r);
        const double beta = (rho / rho_old) * (alpha / omega);
        for (int64_t i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        apply_precond(&P, p, yh);
        or_gemv(n, rp, col, val, 1.0, yh, 0.0, NULL, v);
        alpha = rho / or_norm(n, rh, v);
        or_axpy(n, 1.0, r, -alpha, v, s);
        const double ss = or_norm(n, s, s);
        if (sqrt(ss / scale) <= tol) {
            or_axpy(n, 1.0, x, alpha, yh, x);
            memcpy(r, s, sizeof(double) * (size_t)n);
            rr = ss;
            ++k;
            if (hist) hist[k] = sqrt(rr / scale);
            break;
        }
        apply_precond(&P, s, zh);
        or_gemv(n, rp, col, val, 1.0, zh, 0.0, NULL, t);
        omega = or_norm(n, t, s) / or_norm(n, t, t);
        for (int64_t i = 0; i < n; ++i) x[i] += alpha * yh[i] + omega * zh[i];
        or_axpy(n, 1.0, s, -omega, t, r);
        rr = or_norm(n, r, r);
        rho_old = rho;
        ++k;
        if (hist) hist[k] = sqrt(rr / scale);
    }
    if (true_rel) *true_rel = true_relres(n, rp, col, val, b, x);
    free(r); free(rh); free(p); free(v); free(s); free(t); free(yh); free(zh); free(tmp);
    free(m); free(diag);
    return k;
}

int64_t or_transpose(int64_t n, const uint64_t* rp, const uint64_t* col, const double* val,
                     uint64_t* at_rp, uint64_t* at_col, double* at_val) {
    memset(at_rp, 0, sizeof(uint64_t) * (size_t)(n + 1));
    for (int64_t i = 0; i < n; ++i)
        for (uint64_t p = rp[i]; p < rp[i + 1]; ++p)
            if (col[p] < (uint64_t)n) at_rp[col[p] + 1]++;
    for (int64_t i = 0; i < n; ++i) at_rp[i + 1] += at_rp[i];
    uint64_t* fill = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) fill[i] = at_rp[i];
    for (int64_t i = 0; i < n; ++i)
        for (uint64_t p = rp[i]; p < rp[i + 1]; ++p) {
            const uint64_t c = col[p];
            if (c >= (uint64_t)n) continue;
            at_col[fill[c]] = (uint64_t)i;
            at_val[fill[c]] = val[p];
            fill[c]++;
        }
    free(fill);
    return (int64_t)at_rp[n];
}

void or_sampled_outer(int64_t n, const uint64_t* rp, const uint64_t* col, const double* lambda,
                      const double* x, double* dA) {
    OMP_FOR
    for (int64_t i = 0; i < n; ++i)
        for (uint64_t p = rp[i]; p < rp[i + 1]; ++p)
            dA[p] = col[p] < (uint64_t)n ? -lambda[i] * x[col[p]] : 0.0;
}
