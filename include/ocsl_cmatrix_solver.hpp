// C++17 header-only mirror of the reference's CMatrix_solver public surface
// (cmatrix_solver.h:27-95) over the C ABI of ocsl.h.  A caller that fills
// matrix.value / b_vector.value / x_vector.value and calls solve(),
// CG_solve() or ILU_BICGSTAB_solve() keeps doing exactly that; the solves run
// on the GPU.  Link with libocsl_hip.so.
//
//   ocsl::CMatrix_solver S(n_rows, n_cols, nnz);        // cmatrix_solver.cpp:30-34
//   S.matrix.row_ptr / column_index                     // structure (u64, sentinels allowed)
//   S.matrix.value, S.b_vector.value, S.x_vector.value  // filled by the caller
//   S.solve();                                          // method from `method`
//
// The reference picks the method from its global configuration
// (cmatrix_solver.cpp:764-780); here it is the member `method`.  The
// structure is read at the first solve (the reference fixes it after
// CFVM::define_matrix_structure); later solves re-upload values, b and x.
#ifndef OCSL_CMATRIX_SOLVER_HPP
#define OCSL_CMATRIX_SOLVER_HPP

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "ocsl.h"

namespace ocsl {

class Error : public std::runtime_error {
public:
    Error(ocsl_status st, const std::string& what) : std::runtime_error(what), status(st) {}
    ocsl_status status;
};

inline void check(ocsl_status st, const char* what) {
    if (st < 0) throw Error(st, std::string(what) + ": " + ocsl_last_error());
}

class CMatrix_solver {
public:
    // CMatrix_solver::CCSR (cmatrix_solver.h:33-51)
    struct CCSR {
        std::vector<double> value;
        std::vector<std::size_t> row_ptr;
        std::vector<std::size_t> column_index;
        std::size_t n_rows = 0, n_cols = 0;
        double get_element(std::size_t i, std::size_t j) const {
            for (std::size_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p)
                if (column_index[p] == j) return value[p];
            return 0.0;
        }
    };
    // Cvector (cmatrix_solver.h:53-66)
    struct Cvector {
        std::vector<double> value;
        double get_element(std::size_t i) const { return value[i]; }
        void insert_element(std::size_t i, double v) { value[i] = v; }
        void add_insert_element(std::size_t i, double v) { value[i] += v; }
        void clear() { std::fill(value.begin(), value.end(), 0.0); }
    };

    CCSR matrix;
    Cvector b_vector, x_vector;
    ocsl_method method = OCSL_ILU_BICGSTAB;
    ocsl_mode mode = OCSL_MODE_REFERENCE;       // the reference's capped solvers
    ocsl_precision precision = OCSL_FP64;       // TOLERANCE mode only
    ocsl_options options{};
    ocsl_report last_report{};
    double last_residual = 0.0;                 // what CG_solve / ILU_BICGSTAB_solve returned

    CMatrix_solver(std::size_t n_rows, std::size_t n_cols, std::size_t nnz) {
        matrix.n_rows = n_rows;
        matrix.n_cols = n_cols;
        matrix.value.assign(nnz, 0.0);
        matrix.row_ptr.assign(n_rows + 1, 0);
        matrix.column_index.assign(nnz, 0);
        b_vector.value.assign(n_rows, 0.0);
        x_vector.value.assign(n_rows, 0.0);
        ocsl_default_options(&options);
    }
    ~CMatrix_solver() {
        if (h_) ocsl_destroy(h_);
    }
    CMatrix_solver(const CMatrix_solver&) = delete;
    CMatrix_solver& operator=(const CMatrix_solver&) = delete;

    // cmatrix_solver.cpp:764-780 (method from the member, not a global)
    void solve() {
        if (mode == OCSL_MODE_REFERENCE) {
            if (method == OCSL_CG) CG_solve();
            else ILU_BICGSTAB_solve();
            return;
        }
        upload();
        check(ocsl_solve(h_, method, OCSL_MODE_TOLERANCE, precision, &options, &last_report), "ocsl_solve");
        last_residual = last_report.residual;
        download();
    }
    // cmatrix_solver.cpp:635-680; returns the final <r', r'>
    double CG_solve() {
        upload();
        check(ocsl_CG_solve(h_, &last_residual), "ocsl_CG_solve");
        download();
        return last_residual;
    }
    // cmatrix_solver.cpp:682-762; returns <r, r> / <b, b> + DBL_MIN
    double ILU_BICGSTAB_solve() {
        upload();
        check(ocsl_ILU_BICGSTAB_solve(h_, &last_residual), "ocsl_ILU_BICGSTAB_solve");
        download();
        return last_residual;
    }
    // cmatrix_solver.cpp:472-552: M after `sweeps` sweeps, in the slot order of `matrix`
    std::vector<double> ILU_decompose(int sweeps = 3) {
        upload();
        std::vector<double> m(matrix.value.size());
        check(ocsl_ILU_decompose(h_, sweeps, m.data(), OCSL_HOST), "ocsl_ILU_decompose");
        return m;
    }
    // cmatrix_solver.cpp:233-263 (device reductions)
    double norm(const Cvector& a, const Cvector& b) {
        ensure_handle();
        double out = 0.0;
        check(ocsl_norm(h_, a.value.data(), b.value.data(), OCSL_HOST, &out), "ocsl_norm");
        return out;
    }
    double sum(const Cvector& a) {
        ensure_handle();
        double out = 0.0;
        check(ocsl_sum(h_, a.value.data(), OCSL_HOST, &out), "ocsl_sum");
        return out;
    }
    double ABSsum(const Cvector& a) {
        ensure_handle();
        double out = 0.0;
        check(ocsl_ABSsum(h_, a.value.data(), OCSL_HOST, &out), "ocsl_ABSsum");
        return out;
    }
    // cmatrix_solver.cpp:338-453: out = a (A x) + b z
    void gemv(double a, const Cvector& x, double b, const Cvector* z, Cvector& out) {
        upload_values();
        out.value.resize(matrix.n_rows);
        check(ocsl_gemv(h_, a, x.value.data(), b, z ? z->value.data() : nullptr, out.value.data(), OCSL_HOST),
              "ocsl_gemv");
    }

    ocsl_handle handle() {
        ensure_handle();
        return h_;
    }

private:
    ocsl_handle h_ = nullptr;

    void ensure_handle() {
        if (h_) return;
        static_assert(sizeof(std::size_t) == sizeof(uint64_t), "64-bit size_t expected");
        check(ocsl_create(&h_, (int64_t)matrix.n_rows, reinterpret_cast<const uint64_t*>(matrix.row_ptr.data()),
                          reinterpret_cast<const uint64_t*>(matrix.column_index.data())),
              "ocsl_create");
    }
    void upload_values() {
        ensure_handle();
        check(ocsl_set_values(h_, matrix.value.data(), OCSL_HOST), "ocsl_set_values");
    }
    void upload() {
        upload_values();
        check(ocsl_set_rhs(h_, b_vector.value.data(), OCSL_HOST), "ocsl_set_rhs");
        check(ocsl_set_x(h_, x_vector.value.data(), OCSL_HOST), "ocsl_set_x");
    }
    void download() { check(ocsl_get_x(h_, x_vector.value.data(), OCSL_HOST), "ocsl_get_x"); }
};

}  // namespace ocsl

#endif  // OCSL_CMATRIX_SOLVER_HPP
