/*
 * ocsl.h -- C ABI of the OpenCSL MI355X sparse-Krylov hot path (libocsl_hip.so).
 *
 * Drop-in device replacement for the linear-solver half of the reference's
 * CMatrix_solver (cmatrix_solver.h:27-95): the CSR system it owns
 * (matrix / b_vector / x_vector), its two solvers CG_solve and
 * ILU_BICGSTAB_solve, the dispatching solve(), and its BLAS helpers.  Plain C
 * types only; every entry point returns an ocsl_status and never exits or
 * throws.  One handle = one CSR structure + one HIP stream; calls on a handle
 * are serialised by the caller, distinct handles are independent.
 *
 * CSR arrays use the reference's layout (CMatrix_solver::CCSR,
 * cmatrix_solver.h:33-51): uint64 row_ptr[n+1], uint64 column_index[nnz]
 * ascending per row, double value[nnz].  Column (uint64_t)-1, the sentinel
 * CFVM::define_matrix_structure puts in boundary rows (cfvm.cpp:90-93), is
 * accepted and ignored, as is any column >= n.
 */
#ifndef OCSL_H
#define OCSL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ocsl_solver* ocsl_handle;

typedef enum {
    OCSL_OK = 0,
    /* > 0: the solve ran; warning flags (may be OR-ed) */
    OCSL_WARN_MAXITER = 1,     /* tolerance not reached within max_iter */
    OCSL_WARN_BREAKDOWN = 2,   /* a Krylov denominator was zero */
    OCSL_WARN_NONFINITE = 4,   /* residual became NaN/Inf */
    /* < 0: hard errors, nothing was solved */
    OCSL_ERR_ARG = -1,
    OCSL_ERR_STRUCTURE = -2,   /* missing diagonal, unsorted row, bad row_ptr */
    OCSL_ERR_NOMEM = -3,
    OCSL_ERR_HIP = -4,
    OCSL_ERR_STATE = -5,       /* values / rhs not set */
    OCSL_ERR_NODEVICE = -6
} ocsl_status;

typedef enum { OCSL_CG = 0, OCSL_ILU_BICGSTAB = 1 } ocsl_method;

/* REFERENCE: the reference's fixed-iteration semantics (CG: <= 4 iterations,
 * stop on absolute ||r||^2 < 1e-6; BiCGSTAB: <= 3 iterations, stop on
 * ||r||^2/||b||^2 < 1e-20; 3-sweep ILU, exact triangular solves).
 * TOLERANCE: textbook CG / right-preconditioned ILU(0)-BiCGSTAB to
 * ||b - A x|| / ||b|| <= tol. */
typedef enum { OCSL_MODE_REFERENCE = 0, OCSL_MODE_TOLERANCE = 1 } ocsl_mode;

/* FP64: all arithmetic in double.  MIXED: FP32 matrix, vectors and inner
 * iterations; FP64 solution accumulator, true residual b - A x, refinement
 * update and dot-product accumulation (TOLERANCE mode only). */
typedef enum { OCSL_FP64 = 0, OCSL_MIXED = 1 } ocsl_precision;

typedef enum { OCSL_HOST = 0, OCSL_DEVICE = 1 } ocsl_where;

typedef enum { OCSL_TRI_EXACT = 0, OCSL_TRI_JACOBI = 1 } ocsl_tri_mode;

typedef struct {
    double tol;          /* relative residual target (TOLERANCE); default 1e-8 */
    int max_iter;        /* default 20000 */
    int ilu_sweeps;      /* fine-grained ILU(0) sweeps (>= 1); default 3 */
    int tri_mode;        /* ocsl_tri_mode for TOLERANCE BiCGSTAB; default exact */
    int tri_sweeps;      /* Jacobi sweeps per triangular solve; default 3 */
    double tri_weight;   /* Jacobi damping; default 0.9 */
    double replace_ratio;/* MIXED: FP64 residual replacement when ||r|| drops by this factor; default 0.1 */
    int check_every;     /* iterations per device chunk between host checks; default 32 */
    int reorder;         /* 1: reverse Cuthill-McKee row order for CG (default); 0: input order */
    int use_graph;       /* 1: replay iteration chunks as hipGraphs (default) */
    int block_size;      /* threads per workgroup of the CG kernels, 64..1024; 0 = default (1024) */
    int grid_blocks;     /* workgroups of the CG SpMV kernels; 0 = default (512) */
    int vec_blocks;      /* workgroups of the CG vector-update kernels; 0 = as grid_blocks */
    int rows_per_lane;   /* rows per lane of the CG SELL-P layout, 1, 2 or 4; 0 = tuned, else 2 */
    int cg_kernels;      /* tolerance-CG iteration form: 1 = one fused kernel per iteration (MIXED:
                            with each block's row window of the direction staged in LDS);
                            2 = split (SpMV + update for MIXED, three kernels for FP64);
                            3 = one kernel, gathering from memory; 0 = tuned, else 2 */
    int x_period;        /* MIXED split CG and MIXED batch: iterations per update of the FP64
                            solution, 2, 4 or 8 (the batch also takes 1); 0 = 8 */
    int bicgstab_order;  /* ocsl_bicgstab_order: row order of the TOLERANCE ILU(0)-BiCGSTAB; default input */
} ocsl_options;

/* Row order in which TOLERANCE-mode ILU(0)-BiCGSTAB factors and solves
 * (REFERENCE mode always keeps the input order, as the reference does).
 * INPUT: the caller's numbering (the factor CMatrix_solver::ILU_decompose
 * builds).  COLOUR: rows renumbered by a greedy colouring of the
 * symmetrised row graph (reverse Cuthill-McKee order within a colour): no two
 * rows of a colour are coupled, so each triangular solve has one level per
 * colour and the SpMV gathers stay local; the ILU(0) factor is that of the
 * renumbered matrix, so iterates differ from INPUT's while the residual
 * target is the same.  b, x and the results stay in the caller's numbering. */
typedef enum { OCSL_BICGSTAB_ORDER_INPUT = 0, OCSL_BICGSTAB_ORDER_COLOUR = 1 } ocsl_bicgstab_order;

typedef struct {
    int status;          /* ocsl_status of the solve */
    int iterations;      /* Krylov iterations performed */
    int replacements;    /* MIXED: FP64 residual replacements */
    int history_len;     /* entries available from ocsl_get_history */
    double residual;     /* REFERENCE: the value CG_solve / ILU_BICGSTAB_solve return;
                            TOLERANCE: final ||b - A x||/||b|| in FP64 */
    double true_relres;  /* ||b - A x||/||b|| in FP64 (both modes) */
    double solve_ms;     /* device time of the solve (HIP events), excluding uploads */
    double bytes;        /* algorithmic HBM bytes of the Krylov iterations */
} ocsl_report;

/* Fills *opt with the defaults listed above. */
void ocsl_default_options(ocsl_options* opt);

/* Number of HIP devices (0 when none); never fails. */
int ocsl_device_count(void);

/* Builds, on the device, the CSR structure CFVM::define_matrix_structure
 * makes (cfvm.cpp:74-98, each element inserted through CCSR::insert_element,
 * cmatrix_solver.cpp:57-100): row i holds, ascending and each once, i, the
 * cells across its faces and, when it owns a boundary face, the sentinel
 * column (uint64_t)-1 (the -1 of its neighbour list, cgeometry.cpp:602-612).
 * Input: n_cells, the n_faces interior faces (face_left[f], face_right[f]:
 * the two cells) and the cells of the n_bc boundary faces.  Output: row_ptr
 * (n_cells + 1), column_index (column_cap entries), *nnz; optionally
 * face_slot (2 n_faces: the slot of column face_right[f] in row face_left[f],
 * then of face_left[f] in row face_right[f]) and diag_slot (n_cells: the slot
 * of each diagonal).  With column_index NULL only *nnz and (if given)
 * row_ptr are produced: the size query.  All arrays are host or device
 * memory per `where`.  Cell indices out of range, or a face with one cell on
 * both sides, return OCSL_ERR_STRUCTURE before anything is filled. */
ocsl_status ocsl_build_structure(int64_t n_cells, int64_t n_faces, const int32_t* face_left,
                                 const int32_t* face_right, int64_t n_bc, const int32_t* bc_left,
                                 ocsl_where where, uint64_t* row_ptr, uint64_t* column_index,
                                 int64_t column_cap, int64_t* nnz, int64_t* face_slot, int64_t* diag_slot);

/* Creates a handle for an n x n CSR structure (the reference's CCSR layout).
 * Replaces CFVM::define_matrix_structure's CMatrix_solver construction
 * (cfvm.cpp:74-98: CMatrix_solver(n, n, nnz), M = matrix,
 * extract_diag_indexes()).  Sentinel columns are stripped; every row must hold
 * its diagonal.  Builds the device structure, reorderings and ILU tables. */
ocsl_status ocsl_create(ocsl_handle* h, int64_t n, const uint64_t* row_ptr,
                        const uint64_t* column_index);
ocsl_status ocsl_destroy(ocsl_handle h);

/* Row count and stored-entry count (including sentinels) of the handle. */
ocsl_status ocsl_get_size(ocsl_handle h, int64_t* n, int64_t* nnz);

/* Values of matrix / b_vector / x_vector (cmatrix_solver.h:87-89), in the
 * reference's order, from host or device memory.  A device pointer is read
 * during ocsl_solve, not copied here: keep it alive until the solve returns. */
ocsl_status ocsl_set_values(ocsl_handle h, const double* value, ocsl_where where);
ocsl_status ocsl_set_rhs(ocsl_handle h, const double* b, ocsl_where where);
ocsl_status ocsl_set_x(ocsl_handle h, const double* x0, ocsl_where where);
/* Copies x_vector out after a solve. */
ocsl_status ocsl_get_x(ocsl_handle h, double* x, ocsl_where where);

/* One linear solve.  method/mode/precision replace the global
 * ematrix_solver_type that CMatrix_solver::solve() reads
 * (cmatrix_solver.cpp:764-780).  opt may be NULL (defaults); rep may be NULL. */
ocsl_status ocsl_solve(ocsl_handle h, int method, int mode, int precision,
                       const ocsl_options* opt, ocsl_report* rep);

/* k right-hand sides sharing the handle's matrix (and ILU factors), solved
 * together by TOLERANCE-mode ILU(0)-BiCGSTAB: every matrix / factor pass
 * serves all k.  Replaces the three momentum solves per iteration of
 * CFlow_solver::iterate (cflow_solver.cpp:259-301), whose matrices coincide
 * unless no-slip viscous walls couple the components
 * (cflow_solver_cell_face_loops.cpp:94-113).  B and X hold k column blocks of
 * n values (B[c*n + i]); X holds the initial guesses on entry and the
 * solutions on return.  1 <= k <= 4; reps (optional) receives k reports. */
ocsl_status ocsl_solve_multi(ocsl_handle h, int method, int precision, int k, const double* B,
                             double* X, ocsl_where where, const ocsl_options* opt,
                             ocsl_report* reps);

/* CMatrix_solver::CG_solve (cmatrix_solver.cpp:635-680) and
 * CMatrix_solver::ILU_BICGSTAB_solve (cmatrix_solver.cpp:682-762) in
 * REFERENCE mode; *res receives their return value. */
ocsl_status ocsl_CG_solve(ocsl_handle h, double* res);
ocsl_status ocsl_ILU_BICGSTAB_solve(ocsl_handle h, double* res);

/* CMatrix_solver::ILU_BICGSTAB_solve (cmatrix_solver.cpp:682-762) for k
 * right-hand sides of the handle's matrix at once, in REFERENCE mode: each
 * right-hand side's solution, residual (res[c]) and history are bitwise
 * those of ocsl_ILU_BICGSTAB_solve with that right-hand side and guess,
 * while the factor is computed once and every matrix, factor and vector
 * pass serves all k.  Replaces consecutive calls on one matrix, such as the
 * three momentum solves of CFlow_solver::iterate (cflow_solver.cpp:259-301)
 * when no no-slip viscous wall couples the components.  B and X hold k
 * column blocks of n values (B[c*n + i]); X holds the guesses on entry and
 * the solutions on return.  1 <= k <= 4; res may be NULL.  The history
 * (ocsl_get_history) is the k histories one after another. */
ocsl_status ocsl_ILU_BICGSTAB_solve_multi(ocsl_handle h, int k, const double* B, double* X,
                                          ocsl_where where, double* res);

/* Residual history of the last solve: REFERENCE mode, the value each
 * iteration computes (||r'||^2 for CG, ||r||^2/||b||^2 + DBL_MIN for
 * BiCGSTAB); TOLERANCE mode, ||r_k||/||b|| for k = 0..iterations. */
ocsl_status ocsl_get_history(ocsl_handle h, double* out, int cap, int* len);

/* CMatrix_solver::ABSsum / sum / norm (cmatrix_solver.cpp:233-263) of
 * n-vectors, on the handle's device and stream. */
ocsl_status ocsl_ABSsum(ocsl_handle h, const double* v, ocsl_where where, double* out);
ocsl_status ocsl_sum(ocsl_handle h, const double* v, ocsl_where where, double* out);
ocsl_status ocsl_norm(ocsl_handle h, const double* a, const double* b, ocsl_where where,
                      double* out);

/* CMatrix_solver::gemv (cmatrix_solver.cpp:338-453): out = a*(A x) + b*z
 * with the handle's current values; z may be NULL. */
ocsl_status ocsl_gemv(ocsl_handle h, double a, const double* x, double b, const double* z,
                      double* out, ocsl_where where);

/* CMatrix_solver::ILU_decompose (cmatrix_solver.cpp:472-552): the ILU factor
 * values M in the reference's slot order after `sweeps` sweeps. */
ocsl_status ocsl_ILU_decompose(ocsl_handle h, int sweeps, double* m_out, ocsl_where where);

/* CMatrix_solver::ILU_decompose(A, M, uptr) as called on an M that need not
 * hold A's values (cmatrix_solver.cpp:519-551: the sweeps start from M's
 * current values): m0 is that M^0 in the handle's slot order, A the handle's
 * current values.  Sentinel slots of m_out are 0. */
ocsl_status ocsl_ILU_decompose_from(ocsl_handle h, int sweeps, const double* m0, double* m_out,
                                    ocsl_where where);

/* CMatrix_solver::ILU_forsub (cmatrix_solver.cpp:555-571): y = L^-1 b with
 * the unit-lower part of the factor m (the handle's slot order);
 * CMatrix_solver::ILU_backsub (cmatrix_solver.cpp:595-610): x = U^-1 y with
 * its upper part.  Exact level-scheduled solves in the reference's order of
 * terms per row; the diagonal positions are the handle's own (what
 * extract_diag_indexes gives, cmatrix_solver.cpp:455-469). */
ocsl_status ocsl_ILU_forsub(ocsl_handle h, const double* m, const double* b, double* y, ocsl_where where);
ocsl_status ocsl_ILU_backsub(ocsl_handle h, const double* m, const double* y, double* x, ocsl_where where);

/* CMatrix_solver::norm, the static sum and ABSsum (cmatrix_solver.cpp:233-263),
 * product (:265-272) and both axpy forms (:275-335: out = a x + b y, or a x
 * when y is NULL), on n-vectors, without a handle: the reference calls sum
 * and ABSsum as static members (CMatrix_solver::ABSsum(b_vector),
 * cflow_solver.cpp:346).  They run on a stream the library keeps for the
 * current device (created at the first call, held until the process ends;
 * calls are serialised per device).  Device operands must be complete when
 * the call is made.  Reductions add in a fixed order; each product of axpy is
 * rounded before the sum, as the reference's loops compute it. */
ocsl_status ocsl_vec_norm(int64_t n, const double* a, const double* b, ocsl_where where, double* out);
ocsl_status ocsl_vec_sum(int64_t n, const double* a, ocsl_where where, double* out);
ocsl_status ocsl_vec_ABSsum(int64_t n, const double* a, ocsl_where where, double* out);
ocsl_status ocsl_vec_product(int64_t n, const double* x, const double* y, double* out, ocsl_where where);
ocsl_status ocsl_vec_axpy(int64_t n, double a, const double* x, double b, const double* y, double* out,
                          ocsl_where where);

/* ocsl_vec_ABSsum without any allocation (no counterpart in the reference:
 * for callers that cannot allocate -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  a: n doubles in device
 * memory; out: one double in device memory, written when `stream` (a
 * hipStream_t; NULL: the null stream) reaches the call; nothing is copied to
 * the host and nothing waits.  The partial sums live in a fixed device array
 * of the library (16 KiB per device, held until the process ends) shared by
 * every call on that device: calls that overlap on one device race.  n from 0
 * to 2^31 - 1; the same result, bit for bit, as ocsl_vec_ABSsum. */
ocsl_status ocsl_vec_ABSsum_noalloc(int64_t n, const double* a, double* out, void* stream);

/* Measurement hook (no counterpart in the reference): launches the tolerance-
 * mode CG SpMV kernel (q = A p with its fused dot products) `reps` times on
 * the handle's stream between two HIP events and returns its average duration
 * and the algorithmic bytes of one launch: off-diagonals * (sv + cb) + n*sv
 * (diagonal) + 4*(slices+1) + 2*n*sv (p gathered once, q written), plus n*sv
 * (r read) for MIXED; sv = 8 (FP64) or 4 (MIXED), cb = 2 with packed 16-bit
 * column offsets, else 4.  Values must be set. */
ocsl_status ocsl_time_cg_spmv(ocsl_handle h, int precision, int reorder, int rows_per_lane, int reps,
                              double* avg_ms, double* bytes_per_launch);

/* Measurement hook (no counterpart in the reference): the one-kernel CG
 * iteration (cg_kernels 1, or 3 when tuned so) of the handle's tuned launch.  After a true
 * residual from x = 0 and the first q = A p, runs `pairs` (even, odd) pairs of
 * iteration kernels with a HIP event between every two launches; avg_ms[0] /
 * avg_ms[1] receive the average duration of the even / odd kernel and bytes[0]
 * / bytes[1] the algorithmic bytes of one launch: the matrix as for
 * ocsl_time_cg_spmv, r, q, p read once and written once (6*n*sv), and for the
 * odd kernel x read and written and the previous direction read (16*n +
 * n*sv).  Values and b must be set; the handle's x is left unchanged. */
ocsl_status ocsl_time_cg_iteration(ocsl_handle h, int precision, int pairs, double* avg_ms, double* bytes);

/* Autotuner (no counterpart in the reference): runs `iterations` iterations
 * of the TOLERANCE-mode CG solve of the handle's system for every launch
 * candidate -- iteration form {1, 2, 3} x rows per lane {1, 2, 4} x threads
 * per workgroup {256, 512, 1024} x SpMV workgroups {256 .. 2048} x (split
 * form) vector-update workgroups -- timed with HIP events; the six fastest
 * run again over four times the iterations and the fastest of those is kept
 * in the handle for later solves of that precision whose options leave those
 * fields at 0.  best (optional) receives it and best_us_per_iter its refined
 * time; table (optional) every candidate's time in the sweep.  x is
 * preserved. */
typedef struct {
    int rows_per_lane, block_size, grid_blocks, vec_blocks, cg_kernels;
    double us_per_iteration;
} ocsl_tune_entry;
ocsl_status ocsl_tune_cg(ocsl_handle h, int precision, int iterations, ocsl_options* best,
                         double* best_us_per_iter, ocsl_tune_entry* table, int table_cap, int* table_len);

/* Sets the handle's tuned launch of a precision (e.g. from a cache of
 * ocsl_tune_cg results); all-zero fields clear it. */
ocsl_status ocsl_set_tuned_cg(ocsl_handle h, int precision, const ocsl_options* tuned);

/* REFERENCE-mode solves (ocsl_CG_solve, ocsl_ILU_BICGSTAB_solve and
 * ocsl_solve in REFERENCE mode) replay their whole device work as one hipGraph
 * once a solve with the same right-hand-side address has run (on: 1, the
 * default).  0 issues every launch each time; results are bitwise the same. */
ocsl_status ocsl_set_reference_graphs(ocsl_handle h, int on);

/* Form of the exact triangular solves (ILU_forsub / ILU_backsub and the
 * BiCGSTAB preconditioner; no counterpart in the reference, whose solves are
 * sequential loops): 0 (default) chosen from the level schedule; 1 one
 * workgroup, a barrier per level; 2 one launch per level; 3 serial, one
 * thread over the rows staged in LDS (systems whose level-ordered factor fits
 * in 64 KB; otherwise as 2).  Every form adds the same terms in the same
 * order: results are bitwise the same. */
ocsl_status ocsl_set_tri_form(ocsl_handle h, int form);

/* Measurement hook (no counterpart in the reference): the exact lower
 * (lower = 1) or upper (0) solve of the handle's 3-sweep ILU(0) factor
 * (ILU_forsub / ILU_backsub, cmatrix_solver.cpp:553-633, level-scheduled),
 * `reps` times between two HIP events on the handle's stream; *avg_ms
 * receives the average per solve.  Values must be set. */
ocsl_status ocsl_time_tri_solve(ocsl_handle h, int lower, int reps, double* avg_ms);
/* The same for one form of the solve (the forms of ocsl_set_tri_form; form
 * 3 is an argument error when the system does not fit it). */
ocsl_status ocsl_time_tri_solve_form(ocsl_handle h, int lower, int form, int reps, double* avg_ms);

/* The row order of OCSL_BICGSTAB_ORDER_COLOUR solves (no counterpart in the
 * reference): order[i] = the input row placed at position i (n entries;
 * NULL to query only), *n_colours = the colours, i.e. the levels of each
 * triangular solve.  Built on first use and kept by the handle. */
ocsl_status ocsl_colour_order(ocsl_handle h, int* order, int* n_colours);

/* HIP stream of the handle (a hipStream_t), for callers that pass device
 * pointers produced on another stream. */
void* ocsl_get_stream(ocsl_handle h);

/* ---- batches of independent systems (no counterpart in the reference, which
 * solves one system at a time through CMatrix_solver::solve) ----
 * m systems with their own CSR structures (the reference layout), values and
 * right-hand sides, solved by TOLERANCE-mode CG in the same launches: every
 * system keeps its own scalars and convergence, blocks of converged systems
 * return at once, iteration chunks replay as hipGraphs.  MIXED: FP32
 * iterations with FP64 solution and per-system FP64 residual replacement. */
typedef struct ocsl_batch_s* ocsl_batch;
ocsl_status ocsl_batch_create(ocsl_batch* bt, int m, const int64_t* n_rows,
                              const uint64_t* const* row_ptr, const uint64_t* const* column_index);
ocsl_status ocsl_batch_destroy(ocsl_batch bt);
ocsl_status ocsl_batch_size(ocsl_batch bt, int* m, int64_t* rows_total);
ocsl_status ocsl_batch_set_values(ocsl_batch bt, const double* const* values, ocsl_where where);
ocsl_status ocsl_batch_set_rhs(ocsl_batch bt, const double* const* b, ocsl_where where);
/* x[k]: initial guess on entry, solution on return.  reps (optional): m
 * reports; latency_ms (optional): per system, the wall time from the start
 * of the solve until its convergence was observed (at chunk granularity). */
ocsl_status ocsl_batch_solve(ocsl_batch bt, int precision, const ocsl_options* opt, double* const* x,
                             ocsl_where where, ocsl_report* reps, double* latency_ms);
/* Launch shape of the batch: rows per lane of the concatenated SELL-P layout
 * (1, 2 or 4: slices of 64, 128 or 256 rows, one per wave) and threads per
 * SpMV workgroup (256, 512 or 1024); values and right-hand side already set
 * follow into the new layout.  Default 1 and 1024. */
ocsl_status ocsl_batch_set_launch(ocsl_batch bt, int rows_per_lane, int block_size);
ocsl_status ocsl_batch_get_launch(ocsl_batch bt, int* rows_per_lane, int* block_size);
/* Average device time of one CG iteration of the whole batch in the current
 * launch shape (every system restarted from x = 0 and kept running), over
 * `iterations` iterations replayed as one hipGraph. */
ocsl_status ocsl_batch_time_iterations(ocsl_batch bt, int precision, int iterations, double* us_per_iter);
/* Autotuner of the batched launch: times every rows-per-lane x workgroup-size
 * shape with ocsl_batch_time_iterations and keeps the fastest.  table
 * (optional, table_cap entries) receives each measurement with grid_blocks =
 * workgroups per launch, vec_blocks = threads per update workgroup and
 * cg_kernels = 2 (an iteration is two kernels). */
ocsl_status ocsl_batch_tune(ocsl_batch bt, int precision, int iterations, int* best_rows_per_lane,
                            int* best_block_size, double* best_us_per_iter, ocsl_tune_entry* table, int table_cap,
                            int* table_len);

/* ---- differentiation (no counterpart in the reference, which has no
 * derivative path; for gradient-based callers of CMatrix_solver::solve,
 * cmatrix_solver.cpp:764-780) ----
 * Adjoint of the handle's last solve x = A^{-1} b for a loss L(x) with
 * g = dL/dx: solves A^T lambda = g to opt->tol (TOLERANCE mode; method CG
 * for symmetric positive definite A, else ILU(0)-BiCGSTAB; FP64 or MIXED),
 * then dL/db = lambda and dL/dA_ij = -lambda_i x_j on every stored entry,
 * in the caller's slot order (0 in sentinel slots).  x is the handle's
 * current solution (ocsl_get_x).  dLdb or dLdA may be NULL. */
ocsl_status ocsl_adjoint(ocsl_handle h, int method, int precision, const double* g, double* dLdb,
                         double* dLdA, ocsl_where where, const ocsl_options* opt, ocsl_report* rep);

/* out[p] = alpha * u[i] * v[j] for every stored entry p = (i, j), in the
 * caller's slot order (0 in sentinel slots). */
ocsl_status ocsl_sampled_outer(ocsl_handle h, double alpha, const double* u, const double* v,
                               double* out, ocsl_where where);

/* Measurement hook: average duration of the sampled outer-product kernel
 * over `reps` launches, and its algorithmic bytes nnz*(8+4) + 4*(n+1) + 2*n*8
 * (values written, columns read, row offsets, u and v once). */
ocsl_status ocsl_time_sampled_outer(ocsl_handle h, int reps, double* avg_ms, double* bytes_per_launch);

/* Measurement hook (no counterpart in the reference): STREAM triad
 * a = b + 3 c over three device arrays of n doubles (n even), 16-byte
 * accesses; avg_ms over `reps` launches after a warm-up, bytes_per_launch =
 * 24 n.  The bandwidth a plain streaming kernel reaches on the device, for
 * comparison with the published HBM peak. */
ocsl_status ocsl_time_stream_triad(int64_t n, int reps, double* avg_ms, double* bytes_per_launch);

/* Last error message of the calling thread ("" if none). */
const char* ocsl_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* OCSL_H */
