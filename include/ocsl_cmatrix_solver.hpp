// C++17 header-only drop-in for the reference's CMatrix_solver
// (cmatrix_solver.h:27-95) over the C ABI of ocsl.h: the same class name,
// nested types, members, constructors and member functions with the same
// signatures and meaning, so that a caller written against the reference --
// CFVM::define_matrix_structure (cfvm.cpp:84-97), the assembly loops
// (cflow_solver_cell_face_loops.cpp), CFlow_solver::iterate
// (cflow_solver.cpp:259-263, 346-349) -- compiles unchanged against this
// header and runs its solves and vector operations on the GPU.  Link with
// libocsl_hip.so.
//
// What stays on the host, as in the reference: the CSR and vector containers
// (CCSR, Cvector: std::vector storage the caller reads and writes directly),
// their element access and structural inserts, and extract_diag_indexes()
// (setup bookkeeping).  What runs on the device: CG_solve,
// ILU_BICGSTAB_solve, solve, ILU_decompose, ILU_forsub, ILU_backsub, gemv,
// norm, sum, ABSsum, product and axpy.  Each of those uploads the operands it
// reads and downloads what it writes (the containers are host memory);
// callers that keep their data in device memory use the C ABI directly.
//
// Beyond the reference (which reads one global configuration,
// cmatrix_solver.cpp:764-780): the members `method`, `mode`, `precision` and
// `options` choose what solve() runs; `last_report` holds the last solve's
// report.  Device errors throw ocsl::SolverError (the reference has no error
// path); warnings (iteration cap, breakdown) are left in last_report.status.
//
// Device handles are created lazily for the structure of `matrix` (and
// reused for M and any CCSR with the same row_ptr / column_index), and
// recreated when that structure grows: CCSR::insert_element and
// add_insert_element only ever add entries, so a structural change always
// changes the entry count.  Writing row_ptr / column_index directly to a
// different pattern of the same size is not detected.
//
// CMatrix_solver is declared in namespace ocsl and, unless
// OCSL_NO_GLOBAL_NAMES is defined, also made visible in the global namespace
// as the reference's callers name it.
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

class SolverError : public std::runtime_error {
public:
    SolverError(ocsl_status st, const std::string& what) : std::runtime_error(what), status(st) {}
    ocsl_status status;
};

inline void check(ocsl_status st, const char* what) {
    if (st < 0) throw SolverError(st, std::string(what) + ": " + ocsl_last_error());
}

class CMatrix_solver {
public:
    // CMatrix_solver::CCSR (cmatrix_solver.h:33-51, cmatrix_solver.cpp:37-184)
    class CCSR {
    public:
        std::size_t n_rows = 0;     // Number of rows
        std::size_t n_columns = 0;  // Number of columns
        std::size_t n_nnz = 0;      // entries reserved at construction (not updated by inserts)
        std::vector<double> value;
        std::vector<std::size_t> row_ptr;
        std::vector<std::size_t> column_index;

        CCSR() {}
        // empty pattern: row_ptr of n_rows + 1 zeros, room for n_nnz entries
        CCSR(std::size_t n_rows_, std::size_t n_columns_, std::size_t n_nnz_)
            : n_rows(n_rows_), n_columns(n_columns_), n_nnz(n_nnz_) {
            value.reserve(n_nnz_);
            row_ptr.assign(n_rows_ + 1, 0);
            column_index.reserve(n_nnz_);
        }

        // value at (i, j), 0 when (i, j) is not stored
        double get_element(const std::size_t i, const std::size_t j) const {
            const std::size_t p = find(i, j);
            return p == npos ? 0.0 : value[p];
        }
        // overwrites (i, j), or inserts it keeping the row's columns ascending
        bool insert_element(const std::size_t i, const std::size_t j, const double value_inp) {
            const std::size_t p = find(i, j);
            if (p != npos) value[p] = value_inp;
            else insert_new(i, j, value_inp);
            return true;
        }
        // adds to (i, j), or inserts it with that value
        bool add_insert_element(const std::size_t i, const std::size_t j, const double value_inp) {
            const std::size_t p = find(i, j);
            if (p != npos) value[p] += value_inp;
            else insert_new(i, j, value_inp);
            return true;
        }
        void clear() { std::fill(value.begin(), value.end(), 0.0); }
        // values of a matrix with the same pattern
        void copy_same_structure(const CCSR& original) {
            std::copy(original.value.begin(), original.value.begin() + value.size(), value.begin());
        }

    private:
        static constexpr std::size_t npos = ~(std::size_t)0;
        std::size_t find(std::size_t i, std::size_t j) const {
            for (std::size_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p)
                if (column_index[p] == j) return p;
            return npos;
        }
        // position: before the first larger column of row i, else at its end
        void insert_new(std::size_t i, std::size_t j, double v) {
            std::size_t at = row_ptr[i + 1];
            for (std::size_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p)
                if (j < column_index[p]) {
                    at = p;
                    break;
                }
            value.insert(value.begin() + at, v);
            column_index.insert(column_index.begin() + at, j);
            for (std::size_t r = i + 1; r < row_ptr.size(); ++r) ++row_ptr[r];
        }
    };

    // CMatrix_solver::Cvector (cmatrix_solver.h:53-66, cmatrix_solver.cpp:187-230)
    class Cvector {
    public:
        std::size_t n_values = 0;
        std::vector<double> value;

        Cvector() {}
        Cvector(const int n) : n_values((std::size_t)n), value((std::size_t)n, 0.0) {}

        double get_element(const std::size_t i) const { return value[i]; }
        void insert_element(const std::size_t i, const double value_inp) { value[i] = value_inp; }
        void add_insert_element(std::size_t i, const double value_inp) { value[i] += value_inp; }
        void clear() { std::fill(value.begin(), value.end(), 0.0); }
    };

    CMatrix_solver() {}
    // cmatrix_solver.cpp:30-34
    CMatrix_solver(std::size_t n_rows, std::size_t n_columns, std::size_t n_nnz)
        : matrix(n_rows, n_columns, n_nnz), b_vector((int)n_rows), x_vector((int)n_rows),
          M(n_rows, n_columns, n_nnz) {
        diag_index.resize(n_rows);
        ocsl_default_options(&options);
    }
    ~CMatrix_solver() { release(); }
    // copies take the data, not the device handle
    CMatrix_solver(const CMatrix_solver& o)
        : matrix(o.matrix), b_vector(o.b_vector), x_vector(o.x_vector), diag_index(o.diag_index), M(o.M),
          method(o.method), mode(o.mode), precision(o.precision), options(o.options), last_report(o.last_report) {}
    CMatrix_solver& operator=(const CMatrix_solver& o) {
        if (this != &o) {
            release();
            matrix = o.matrix;
            b_vector = o.b_vector;
            x_vector = o.x_vector;
            diag_index = o.diag_index;
            M = o.M;
            method = o.method;
            mode = o.mode;
            precision = o.precision;
            options = o.options;
            last_report = o.last_report;
        }
        return *this;
    }

    // cmatrix_solver.cpp:233-241: sum of inp1_i inp2_i
    double norm(const Cvector& inp1, const Cvector& inp2) {
        double out = 0.0;
        check(ocsl_vec_norm((int64_t)inp1.n_values, inp1.value.data(), inp2.value.data(), OCSL_HOST, &out),
              "ocsl_vec_norm");
        return out;
    }
    // cmatrix_solver.cpp:244-263
    static double sum(const Cvector& inp) {
        double out = 0.0;
        check(ocsl_vec_sum((int64_t)inp.n_values, inp.value.data(), OCSL_HOST, &out), "ocsl_vec_sum");
        return out;
    }
    static double ABSsum(const Cvector& inp) {
        double out = 0.0;
        check(ocsl_vec_ABSsum((int64_t)inp.n_values, inp.value.data(), OCSL_HOST, &out), "ocsl_vec_ABSsum");
        return out;
    }
    // cmatrix_solver.cpp:265-272: out = inp1 * inp2 elementwise
    void product(Cvector& inp1, Cvector& inp2, Cvector& out) {
        fit(out, inp1.n_values);
        check(ocsl_vec_product((int64_t)inp1.n_values, inp1.value.data(), inp2.value.data(), out.value.data(),
                               OCSL_HOST),
              "ocsl_vec_product");
    }
    // cmatrix_solver.cpp:275-282: out = a inp1
    void axpy(double a, const Cvector& inp1, Cvector& out) {
        fit(out, inp1.n_values);
        check(ocsl_vec_axpy((int64_t)inp1.n_values, a, inp1.value.data(), 0.0, nullptr, out.value.data(), OCSL_HOST),
              "ocsl_vec_axpy");
    }
    // cmatrix_solver.cpp:284-335: out = a inp1 + b inp2
    void axpy(const double a, const Cvector& inp1, const double b, const Cvector& inp2, Cvector& out) {
        fit(out, inp1.n_values);
        check(ocsl_vec_axpy((int64_t)inp1.n_values, a, inp1.value.data(), b, inp2.value.data(), out.value.data(),
                            OCSL_HOST),
              "ocsl_vec_axpy");
    }
    // cmatrix_solver.cpp:338-356: out = a (inp1 inp2)
    void gemv(const double a, const CCSR& inp1, const Cvector& inp2, Cvector& out) {
        Bound hb = handle_for(inp1);
        check(ocsl_set_values(hb.h, inp1.value.data(), OCSL_HOST), "ocsl_set_values");
        fit(out, inp1.n_rows);
        check(ocsl_gemv(hb.h, a, inp2.value.data(), 0.0, nullptr, out.value.data(), OCSL_HOST), "ocsl_gemv");
    }
    // cmatrix_solver.cpp:358-453: out = a (inp1 inp2) + b inp3
    void gemv(const double a, const CCSR& inp1, const Cvector& inp2, const double b, const Cvector& inp3,
              Cvector& out) {
        Bound hb = handle_for(inp1);
        check(ocsl_set_values(hb.h, inp1.value.data(), OCSL_HOST), "ocsl_set_values");
        fit(out, inp1.n_rows);
        check(ocsl_gemv(hb.h, a, inp2.value.data(), b, inp3.value.data(), out.value.data(), OCSL_HOST), "ocsl_gemv");
    }

    // cmatrix_solver.cpp:455-469: diag_index[i] = position of (i, i) in M
    void extract_diag_indexes() {
        if (diag_index.size() < M.n_rows) diag_index.resize(M.n_rows);
        for (std::size_t i = 0; i < M.n_rows; ++i)
            for (std::size_t p = M.row_ptr[i]; p < M.row_ptr[i + 1]; ++p)
                if (M.column_index[p] == i) {
                    diag_index[i] = (int)p;
                    break;
                }
    }
    // cmatrix_solver.cpp:472-552: three fine-grained ILU(0) sweeps from M's
    // current values; M must have A's pattern.  uptr is the diagonal
    // positions (the reference reads its member diag_index; the device keeps
    // its own, which extract_diag_indexes gives).
    void ILU_decompose(const CCSR& A, CCSR& M_, const std::vector<int>& uptr) {
        (void)uptr;
        if (M_.value.size() != A.value.size())
            throw SolverError(OCSL_ERR_ARG, "ILU_decompose: M must have A's pattern");
        Bound hb = handle_for(A);
        check(ocsl_set_values(hb.h, A.value.data(), OCSL_HOST), "ocsl_set_values");
        std::vector<double> out(A.value.size());
        check(ocsl_ILU_decompose_from(hb.h, 3, M_.value.data(), out.data(), OCSL_HOST), "ocsl_ILU_decompose_from");
        M_.value.swap(out);
    }
    // cmatrix_solver.cpp:555-571: y = L^-1 b (unit lower part of M_)
    void ILU_forsub(const CCSR& M_, const std::vector<int>& diag, const Cvector& b, Cvector& y) {
        (void)diag;
        Bound hb = handle_for(M_);
        fit(y, M_.n_rows);
        check(ocsl_ILU_forsub(hb.h, M_.value.data(), b.value.data(), y.value.data(), OCSL_HOST), "ocsl_ILU_forsub");
    }
    // cmatrix_solver.cpp:595-610: x = U^-1 y (upper part of M_, diagonal included)
    void ILU_backsub(const CCSR& M_, const std::vector<int>& uptr, const Cvector& y, Cvector& x) {
        (void)uptr;
        Bound hb = handle_for(M_);
        fit(x, M_.n_rows);
        check(ocsl_ILU_backsub(hb.h, M_.value.data(), y.value.data(), x.value.data(), OCSL_HOST), "ocsl_ILU_backsub");
    }

    // cmatrix_solver.cpp:635-680 (REFERENCE mode); returns the final <r', r'>
    double CG_solve() {
        ocsl_handle h = upload_system();
        double res = 0.0;
        check(ocsl_CG_solve(h, &res), "ocsl_CG_solve");
        download_x(h);
        last_report.residual = res;
        return res;
    }
    // cmatrix_solver.cpp:682-762 (REFERENCE mode); returns <r, r>/<b, b> +
    // DBL_MIN.  As in the reference, M holds the ILU factor of `matrix`
    // afterwards.
    double ILU_BICGSTAB_solve() {
        ocsl_handle h = upload_system();
        double res = 0.0;
        check(ocsl_ILU_BICGSTAB_solve(h, &res), "ocsl_ILU_BICGSTAB_solve");
        download_x(h);
        last_report.residual = res;
        factor_into_M(h);
        return res;
    }
    // cmatrix_solver.cpp:764-780, the method taken from `method` (and
    // `mode`, `precision`, `options`) instead of the global configuration
    void solve() {
        if (mode == OCSL_MODE_REFERENCE) {
            if (method == OCSL_CG) CG_solve();
            else ILU_BICGSTAB_solve();
            return;
        }
        ocsl_handle h = upload_system();
        const ocsl_status st = ocsl_solve(h, method, OCSL_MODE_TOLERANCE, precision, &options, &last_report);
        check(st, "ocsl_solve");
        download_x(h);
    }

    // Linear equation system memory (cmatrix_solver.h:86-89)
    CCSR matrix;
    Cvector b_vector;
    Cvector x_vector;
    // ILU related memory (cmatrix_solver.h:91-93)
    std::vector<int> diag_index;
    CCSR M;

    // beyond the reference: what solve() runs, and its last report
    ocsl_method method = OCSL_ILU_BICGSTAB;
    ocsl_mode mode = OCSL_MODE_REFERENCE;  // the reference's capped solvers
    ocsl_precision precision = OCSL_FP64;  // TOLERANCE mode only
    ocsl_options options = default_options();
    ocsl_report last_report{};

    // the device handle of `matrix`'s structure (created if needed)
    ocsl_handle handle() { return handle_for(matrix).h; }

private:
    // the main handle: its structure, and the CCSRs known to share it
    ocsl_handle h_ = nullptr;
    std::size_t h_rows_ = 0, h_nnz_ = 0;
    std::vector<std::size_t> h_rp_, h_col_;
    struct Known {
        const CCSR* who;
        std::size_t rows, nnz;
    };
    std::vector<Known> known_;
    // a handle for a pattern other than the main one, for one call
    ocsl_handle tmp_ = nullptr;

    struct Bound {
        ocsl_handle h;
    };

    static ocsl_options default_options() {
        ocsl_options o;
        ocsl_default_options(&o);
        return o;
    }
    static void fit(Cvector& v, std::size_t n) {
        if (v.value.size() < n) v.value.resize(n, 0.0);
        if (v.n_values < n) v.n_values = n;
    }
    void release() {
        if (h_) ocsl_destroy(h_);
        if (tmp_) ocsl_destroy(tmp_);
        h_ = tmp_ = nullptr;
        known_.clear();
        h_rp_.clear();
        h_col_.clear();
    }
    static ocsl_handle create(const CCSR& A) {
        static_assert(sizeof(std::size_t) == sizeof(uint64_t), "64-bit size_t expected");
        if (A.row_ptr.size() != A.n_rows + 1 || A.row_ptr[A.n_rows] != A.column_index.size())
            throw SolverError(OCSL_ERR_STRUCTURE, "CCSR: row_ptr does not match column_index");
        ocsl_handle h = nullptr;
        check(ocsl_create(&h, (int64_t)A.n_rows, reinterpret_cast<const uint64_t*>(A.row_ptr.data()),
                          reinterpret_cast<const uint64_t*>(A.column_index.data())),
              "ocsl_create");
        return h;
    }
    bool same_as_main(const CCSR& A) const {
        return h_ && A.n_rows == h_rows_ && A.column_index.size() == h_nnz_ && A.row_ptr == h_rp_ &&
               A.column_index == h_col_;
    }
    // the handle whose structure is A's
    Bound handle_for(const CCSR& A) {
        for (const Known& k : known_)
            if (k.who == &A && k.rows == A.n_rows && k.nnz == A.column_index.size() && h_) return {h_};
        if (same_as_main(A)) {
            known_.push_back({&A, A.n_rows, A.column_index.size()});
            return {h_};
        }
        if (&A == &matrix || !h_) {
            release();
            h_ = create(A);
            h_rows_ = A.n_rows;
            h_nnz_ = A.column_index.size();
            h_rp_ = A.row_ptr;
            h_col_ = A.column_index;
            known_.push_back({&A, A.n_rows, A.column_index.size()});
            return {h_};
        }
        if (tmp_) ocsl_destroy(tmp_);
        tmp_ = nullptr;
        tmp_ = create(A);
        return {tmp_};
    }
    ocsl_handle upload_system() {
        Bound hb = handle_for(matrix);
        check(ocsl_set_values(hb.h, matrix.value.data(), OCSL_HOST), "ocsl_set_values");
        if (b_vector.value.size() < matrix.n_rows || x_vector.value.size() < matrix.n_rows)
            throw SolverError(OCSL_ERR_ARG, "b_vector / x_vector shorter than the matrix");
        check(ocsl_set_rhs(hb.h, b_vector.value.data(), OCSL_HOST), "ocsl_set_rhs");
        check(ocsl_set_x(hb.h, x_vector.value.data(), OCSL_HOST), "ocsl_set_x");
        return hb.h;
    }
    void download_x(ocsl_handle h) { check(ocsl_get_x(h, x_vector.value.data(), OCSL_HOST), "ocsl_get_x"); }
    // M = matrix's pattern with the factor ILU_BICGSTAB_solve computed
    // (M.copy_same_structure(matrix); ILU_decompose(matrix, M, diag_index),
    // cmatrix_solver.cpp:702-704)
    void factor_into_M(ocsl_handle h) {
        if (M.value.size() != matrix.value.size()) return;  // M not set up by the caller
        check(ocsl_ILU_decompose(h, 3, M.value.data(), OCSL_HOST), "ocsl_ILU_decompose");
    }
};

}  // namespace ocsl

#ifndef OCSL_NO_GLOBAL_NAMES
using ocsl::CMatrix_solver;
#endif

#endif  // OCSL_CMATRIX_SOLVER_HPP
