// ref_driver.cpp -- C entry points into the reference solver, compiled from the
// reference's own sources where they lie (see oracle/Makefile).  Output goes to
// oracle/_ref/ only.  TEST INFRASTRUCTURE: used to pin the oracle, never shipped
// and never part of the product path.
//
// Every function builds a CMatrix_solver (cmatrix_solver.h:27-95) from the
// caller's CSR arrays, exactly as CFVM::define_matrix_structure leaves it
// (cfvm.cpp:74-98: M = matrix, extract_diag_indexes()), runs one reference
// routine with one OpenMP thread and copies the results out.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <vector>

#include "cmatrix_solver.h"

namespace {

CMatrix_solver make_solver(int64_t n, const uint64_t* rp, const uint64_t* col, const double* val,
                           const double* b, const double* x) {
    omp_set_dynamic(0);
    omp_set_num_threads(1);
    const size_t nnz = rp[n];
    CMatrix_solver s((size_t)n, (size_t)n, nnz);
    s.matrix.value.assign(val, val + nnz);
    s.matrix.row_ptr.assign(rp, rp + n + 1);
    s.matrix.column_index.assign(col, col + nnz);
    s.matrix.n_nnz = nnz;
    if (b) s.b_vector.value.assign(b, b + n);
    if (x) s.x_vector.value.assign(x, x + n);
    s.M = s.matrix;
    s.extract_diag_indexes();
    return s;
}

}  // namespace

extern "C" {

double ref_cg_solve(int64_t n, const uint64_t* rp, const uint64_t* col, const double* val,
                    const double* b, double* x) {
    CMatrix_solver s = make_solver(n, rp, col, val, b, x);
    const double res = s.CG_solve();
    std::memcpy(x, s.x_vector.value.data(), sizeof(double) * n);
    return res;
}

double ref_ilu_bicgstab_solve(int64_t n, const uint64_t* rp, const uint64_t* col,
                              const double* val, const double* b, double* x) {
    CMatrix_solver s = make_solver(n, rp, col, val, b, x);
    const double res = s.ILU_BICGSTAB_solve();
    std::memcpy(x, s.x_vector.value.data(), sizeof(double) * n);
    return res;
}

void ref_ilu_decompose(int64_t n, const uint64_t* rp, const uint64_t* col, const double* val,
                       double* m_out) {
    CMatrix_solver s = make_solver(n, rp, col, val, nullptr, nullptr);
    s.M.copy_same_structure(s.matrix);
    s.ILU_decompose(s.matrix, s.M, s.diag_index);
    std::memcpy(m_out, s.M.value.data(), sizeof(double) * rp[n]);
}

void ref_forsub(int64_t n, const uint64_t* rp, const uint64_t* col, const double* m,
                const double* b, double* y) {
    CMatrix_solver s = make_solver(n, rp, col, m, b, nullptr);
    CMatrix_solver::Cvector out((int)n);
    s.ILU_forsub(s.matrix, s.diag_index, s.b_vector, out);
    std::memcpy(y, out.value.data(), sizeof(double) * n);
}

void ref_backsub(int64_t n, const uint64_t* rp, const uint64_t* col, const double* m,
                 const double* yv, double* x) {
    CMatrix_solver s = make_solver(n, rp, col, m, yv, nullptr);
    CMatrix_solver::Cvector out((int)n);
    s.ILU_backsub(s.matrix, s.diag_index, s.b_vector, out);
    std::memcpy(x, out.value.data(), sizeof(double) * n);
}

void ref_gemv(int64_t n, const uint64_t* rp, const uint64_t* col, const double* val, double a,
              const double* xv, double b, const double* z, double* out) {
    CMatrix_solver s = make_solver(n, rp, col, val, z, xv);
    CMatrix_solver::Cvector o((int)n);
    if (z) s.gemv(a, s.matrix, s.x_vector, b, s.b_vector, o);
    else s.gemv(a, s.matrix, s.x_vector, o);
    std::memcpy(out, o.value.data(), sizeof(double) * n);
}

double ref_norm(int64_t n, const double* a, const double* b) {
    omp_set_num_threads(1);
    CMatrix_solver s((size_t)n, (size_t)n, 0);
    CMatrix_solver::Cvector va((int)n), vb((int)n);
    va.value.assign(a, a + n);
    vb.value.assign(b, b + n);
    return s.norm(va, vb);
}

double ref_abssum(int64_t n, const double* a) {
    omp_set_num_threads(1);
    CMatrix_solver::Cvector va((int)n);
    va.value.assign(a, a + n);
    return CMatrix_solver::ABSsum(va);
}

// Times the reference's CG_solve (cmatrix_solver.cpp:635-680) as a caller
// sees it, with `threads` OpenMP threads: `calls` calls on one system, each the
// reference's fixed 4 iterations from the caller's x (x reset before each call,
// so no call stops early), and as many runs of the call's own set-up alone
// (its five work vectors, r = b - A x, p = r).  Medians in seconds; one
// iteration costs (call - setup) / 4.
void ref_cg_timing(int64_t n, const uint64_t* rp, const uint64_t* col, const double* val,
                   const double* b, const double* x0, int threads, int calls, double* call_s,
                   double* setup_s) {
    using clk = std::chrono::steady_clock;
    CMatrix_solver s = make_solver(n, rp, col, val, b, x0);
    omp_set_num_threads(threads > 0 ? threads : 1);
    std::vector<double> tc, ts;
    for (int k = 0; k < calls; ++k) {
        s.x_vector.value.assign(x0, x0 + n);
        const auto t0 = clk::now();
        (void)s.CG_solve();
        tc.push_back(std::chrono::duration<double>(clk::now() - t0).count());
        const auto t1 = clk::now();
        {
            CMatrix_solver::Cvector r((int)n), r_next((int)n), p((int)n), p_next((int)n), dum((int)n);
            s.gemv(-1.0, s.matrix, s.x_vector, 1.0, s.b_vector, r);
            p = r;
        }
        ts.push_back(std::chrono::duration<double>(clk::now() - t1).count());
    }
    auto median = [](std::vector<double> v) {
        std::sort(v.begin(), v.end());
        return v.empty() ? 0.0 : v[v.size() / 2];
    };
    *call_s = median(tc);
    *setup_s = median(ts);
    omp_set_num_threads(1);
}

}  // extern "C"
