/*
 * ocsl_oracle.h -- CPU oracle for the OpenCSL sparse-Krylov hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP solver
 * path is compared against (and, timed, the CPU baseline).  The product path
 * never links, loads or calls it.
 *
 * Matrices use the reference's CSR layout (CMatrix_solver::CCSR,
 * cmatrix_solver.h:33-51): uint64 row_ptr[n+1], uint64 col[nnz] ascending per
 * row, double val[nnz].  Columns >= n are the reference's (size_t)-1 boundary
 * sentinels (cfvm.cpp:90-93); they hold 0.0 and contribute nothing.
 */
#ifndef OCSL_ORACLE_H
#define OCSL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void or_set_threads(int t);
int or_get_threads(void);

/* BLAS-1 / SpMV with the reference's definitions (cmatrix_solver.cpp:233-453). */
double or_norm(int64_t n, const double* a, const double* b);
double or_sum(int64_t n, const double* a);
double or_abssum(int64_t n, const double* a);
void or_axpy(int64_t n, double a, const double* x, double b, const double* y, double* out);
void or_gemv(int64_t n, const uint64_t* rp, const uint64_t* col, const double* val, double a,
             const double* x, double b, const double* z, double* out);

/* diag[i] = slot of (i,i); -1 if absent (cmatrix_solver.cpp:455-469). */
void or_diag_index(int64_t n, const uint64_t* rp, const uint64_t* col, int64_t* diag);

/* Jacobi sweeps of the fine-grained ILU(0) starting from M = A
 * (cmatrix_solver.cpp:519-551; the reference runs 3). */
void or_ilu_decompose(int64_t n, const uint64_t* rp, const uint64_t* col, const double* a_val,
                      const int64_t* diag, int sweeps, double* m_val);
/* Exact sequential ILU(0) (IKJ), the reference's commented-out variant
 * (cmatrix_solver.cpp:474-518); used only to check sweep convergence. */
void or_ilu_exact(int64_t n, const uint64_t* rp, const uint64_t* col, const double* a_val,
                  const int64_t* diag, double* m_val);

/* Exact unit-lower / upper solves with the ILU factors (cmatrix_solver.cpp:555-571, 595-610). */
void or_forsub(int64_t n, const uint64_t* rp, const uint64_t* col, const double* m,
               const int64_t* diag, const double* b, double* y);
void or_backsub(int64_t n, const uint64_t* rp, const uint64_t* col, const double* m,
                const int64_t* diag, const double* y, double* x);
/* Damped Jacobi sweeps of the same solves, the reference's commented-out
 * variants (cmatrix_solver.cpp:573-592, 612-631): out <- w*(...) + (1-w)*out,
 * starting from the content of the output buffer. */
void or_forsub_jacobi(int64_t n, const uint64_t* rp, const uint64_t* col, const double* m,
                      const int64_t* diag, const double* b, int sweeps, double w, double* y);
void or_backsub_jacobi(int64_t n, const uint64_t* rp, const uint64_t* col, const double* m,
                       const int64_t* diag, const double* y, int sweeps, double w, double* x);

/* REFERENCE mode: the reference solvers with their caps and thresholds.
 * hist[it] receives the residual the reference computes in iteration it
 * (at most 4 for CG, 3 for BiCGSTAB); *n_hist the number written.
 * Return value = what CG_solve / ILU_BICGSTAB_solve return. */
double or_cg_reference(int64_t n, const uint64_t* rp, const uint64_t* col, const double* val,
                       const double* b, double* x, double* hist, int* n_hist);
double or_ilu_bicgstab_reference(int64_t n, const uint64_t* rp, const uint64_t* col,
                                 const double* val, const double* b, double* x, double* hist,
                                 int* n_hist);

/* TOLERANCE mode.  Stop when ||r_k||/||b|| <= tol (recurrence residual) or
 * after maxit iterations.  hist[k] = ||r_k||/||b||, k = 0..iters.
 * Returns the iteration count; *true_relres = ||b - A x||/||b|| in FP64. */
int or_cg_tol(int64_t n, const uint64_t* rp, const uint64_t* col, const double* val,
              const double* b, double* x, double tol, int maxit, double* hist,
              double* true_relres);
/* Right-preconditioned ILU(0)-BiCGSTAB (van der Vorst 1992).  The ILU factors
 * are `ilu_sweeps` fine-grained sweeps (0 = exact IKJ ILU(0)); the triangular
 * solves are exact when tri_sweeps == 0, else tri_sweeps damped-Jacobi sweeps
 * with weight tri_w starting from zero. */
int or_bicgstab_tol(int64_t n, const uint64_t* rp, const uint64_t* col, const double* val,
                    const double* b, double* x, double tol, int maxit, int ilu_sweeps,
                    int tri_sweeps, double tri_w, double* hist, double* true_relres);

/* CSR transpose (sentinels dropped): at_rp[n+1], at_col/at_val[nnz_real]. */
int64_t or_transpose(int64_t n, const uint64_t* rp, const uint64_t* col, const double* val,
                     uint64_t* at_rp, uint64_t* at_col, double* at_val);

/* Sampled outer product dA[p] = -lambda[i] * x[col[p]] on the CSR pattern
 * (0 on sentinels). */
void or_sampled_outer(int64_t n, const uint64_t* rp, const uint64_t* col, const double* lambda,
                      const double* x, double* dA);

#ifdef __cplusplus
}
#endif

#endif /* OCSL_ORACLE_H */
