// A caller written against the reference's CMatrix_solver, compiled against
// include/ocsl_cmatrix_solver.hpp instead.  Every use below is one of the
// reference's own call forms:
//   * structure: CFVM::define_matrix_structure (cfvm.cpp:84-97) -- the
//     shared_ptr construction with (cells, cells, nnz), insert_element of the
//     diagonal and of every neighbour (-1 for a boundary face, converted to
//     the size_t column), M = matrix, extract_diag_indexes();
//   * assembly: CFlow_solver::iterate's clears (cflow_solver.cpp:261-263) and
//     the loops' add_insert_element on matrix and b_vector, get_element and
//     insert_element of the diagonal for under-relaxation, insert_element of
//     x_vector (cflow_solver_cell_face_loops.cpp:102, 130, 153, 163-169);
//   * the pressure residual CMatrix_solver::ABSsum(b_vector) and solve()
//     (cflow_solver.cpp:346-349);
//   * the remaining public members of cmatrix_solver.h:27-95 (norm, sum,
//     product, both axpy and gemv forms, ILU_decompose, ILU_forsub,
//     ILU_backsub, CG_solve, ILU_BICGSTAB_solve).
// The system is a convection-diffusion problem on a randomly numbered
// nx x ny x nz hex box.  Output: named arrays, one per line ("name count v0
// v1 ...", %.17g), for tests/test_cpp_mirror.py to check against the oracle.
//
//   cmatrix_solver_dropin [structure]   ("structure": host-only, no device call)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "ocsl_cmatrix_solver.hpp"

namespace {

struct Box {
    int n = 0;
    std::vector<std::vector<int>> neighbours;  // per cell, in face order; -1: boundary face
    std::vector<int> face_left, face_right, bc_left;
    std::vector<double> face_d, face_mdot, bc_d, volume;
};

// deterministic 64-bit LCG (Knuth's MMIX constants)
struct Lcg {
    uint64_t s;
    double next() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(s >> 11) * (1.0 / 9007199254740992.0);
    }
};

Box make_box(int nx, int ny, int nz, uint64_t seed) {
    Box m;
    m.n = nx * ny * nz;
    Lcg rng{seed};
    std::vector<int> perm(m.n);
    for (int i = 0; i < m.n; ++i) perm[i] = i;
    for (int i = m.n - 1; i > 0; --i) {
        const int j = (int)(rng.next() * (i + 1));
        std::swap(perm[i], perm[j]);
    }
    auto id = [&](int i, int j, int k) { return perm[i + nx * (j + ny * k)]; };
    m.neighbours.assign(m.n, {});
    m.volume.assign(m.n, 0.0);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int c = id(i, j, k);
                m.volume[c] = 1.0 + 0.25 * rng.next();
                const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                                      {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
                for (auto& q : nb) {
                    const bool in = q[0] >= 0 && q[0] < nx && q[1] >= 0 && q[1] < ny && q[2] >= 0 && q[2] < nz;
                    const int o = in ? id(q[0], q[1], q[2]) : -1;
                    m.neighbours[c].push_back(o);
                    if (!in) {
                        m.bc_left.push_back(c);
                        m.bc_d.push_back(0.5 + rng.next());
                    } else if (c < o) {
                        m.face_left.push_back(c);
                        m.face_right.push_back(o);
                        m.face_d.push_back(0.2 + rng.next());
                        m.face_mdot.push_back(rng.next() - 0.5);
                    }
                }
            }
    return m;
}

void put(const char* name, const std::vector<double>& v) {
    std::printf("%s %zu", name, v.size());
    for (double x : v) std::printf(" %.17g", x);
    std::printf("\n");
}
void put_u(const char* name, const std::vector<std::size_t>& v) {
    std::printf("%s %zu", name, v.size());
    for (std::size_t x : v) std::printf(" %zu", x);
    std::printf("\n");
}
void put_i(const char* name, const std::vector<int>& v) {
    std::printf("%s %zu", name, v.size());
    for (int x : v) std::printf(" %d", x);
    std::printf("\n");
}
void put1(const char* name, double v) { std::printf("%s 1 %.17g\n", name, v); }

// cfvm.cpp:74-98
void define_matrix_structure(const Box& m, std::shared_ptr<CMatrix_solver>& Matrix_solver) {
    std::size_t nnz = 0;
    for (int i = 0; i < m.n; i++) nnz += m.neighbours[i].size();
    Matrix_solver = std::shared_ptr<CMatrix_solver>(new CMatrix_solver(m.n, m.n, nnz));
    for (std::size_t i_cell = 0; i_cell < (std::size_t)m.n; i_cell++) {
        Matrix_solver->matrix.insert_element(i_cell, i_cell, 0.0);
        for (std::size_t i_neig = 0; i_neig < m.neighbours[i_cell].size(); i_neig++) {
            int neigh_index = m.neighbours[i_cell][i_neig];
            Matrix_solver->matrix.insert_element(i_cell, neigh_index, 0.0);
        }
    }
    Matrix_solver->M = Matrix_solver->matrix;
    Matrix_solver->extract_diag_indexes();
}

// an upwind convection-diffusion system assembled face by face and cell by
// cell with the reference's scatter calls
void assemble(const Box& m, std::shared_ptr<CMatrix_solver>& Matrix_solver, double urf) {
    Matrix_solver->matrix.clear();
    Matrix_solver->b_vector.clear();
    Matrix_solver->x_vector.clear();
    for (std::size_t f = 0; f < m.face_left.size(); ++f) {
        const int l = m.face_left[f], r = m.face_right[f];
        const double D = m.face_d[f], F = m.face_mdot[f];
        Matrix_solver->matrix.add_insert_element(l, l, D + std::fmax(F, 0.0));
        Matrix_solver->matrix.add_insert_element(l, r, -D - std::fmax(-F, 0.0));
        Matrix_solver->matrix.add_insert_element(r, r, D + std::fmax(-F, 0.0));
        Matrix_solver->matrix.add_insert_element(r, l, -D - std::fmax(F, 0.0));
    }
    for (std::size_t f = 0; f < m.bc_left.size(); ++f) {
        const int c = m.bc_left[f];
        Matrix_solver->matrix.add_insert_element(c, c, m.bc_d[f]);
        Matrix_solver->b_vector.add_insert_element(c, m.bc_d[f] * 0.3);
    }
    for (int c = 0; c < m.n; ++c) {
        const double t = m.volume[c] / 0.5;
        Matrix_solver->matrix.add_insert_element(c, c, t);
        Matrix_solver->b_vector.add_insert_element(c, t * std::sin(0.1 * c));
        // implicit under-relaxation of the diagonal (cflow_solver_cell_face_loops.cpp:163-169)
        double dum = Matrix_solver->matrix.get_element(c, c);
        Matrix_solver->matrix.insert_element(c, c, dum / urf);
        dum = (1.0 - urf) * dum / urf * std::cos(0.05 * c);
        Matrix_solver->b_vector.add_insert_element(c, dum);
        Matrix_solver->x_vector.insert_element(c, std::cos(0.05 * c));
    }
}

}  // namespace

int main(int argc, char** argv) {
    const bool structure_only = argc > 1 && std::strcmp(argv[1], "structure") == 0;
    const Box m = make_box(8, 7, 6, 12345);
    std::shared_ptr<CMatrix_solver> Matrix_solver;
    try {
        define_matrix_structure(m, Matrix_solver);
        put_i("face_left", m.face_left);
        put_i("face_right", m.face_right);
        put_i("bc_left", m.bc_left);
        put_u("row_ptr", Matrix_solver->matrix.row_ptr);
        put_u("col", Matrix_solver->matrix.column_index);
        put_i("diag_index", Matrix_solver->diag_index);
        std::printf("n_rows %zu %zu %zu\n", Matrix_solver->matrix.n_rows, Matrix_solver->matrix.n_columns,
                    Matrix_solver->matrix.n_nnz);
        assemble(m, Matrix_solver, 0.7);
        put("val", Matrix_solver->matrix.value);
        put("b", Matrix_solver->b_vector.value);
        put("x0", Matrix_solver->x_vector.value);
        if (structure_only) return 0;

        CMatrix_solver& S = *Matrix_solver;
        const CMatrix_solver::Cvector b0 = S.b_vector, x0 = S.x_vector;
        // cflow_solver.cpp:346
        put1("abssum_b", CMatrix_solver::ABSsum(Matrix_solver->b_vector));
        put1("sum_b", CMatrix_solver::sum(S.b_vector));
        put1("norm_bx", S.norm(S.b_vector, S.x_vector));
        CMatrix_solver::Cvector out(m.n), out2(m.n), tmp(m.n);
        S.gemv(1.0, S.matrix, x0, out);
        put("gemv_ax", out.value);
        S.gemv(-1.0, S.matrix, x0, 1.0, b0, out2);
        put("gemv_r", out2.value);
        S.axpy(2.5, b0, tmp);
        put("axpy1", tmp.value);
        S.axpy(0.5, b0, -1.0, x0, tmp);
        put("axpy2", tmp.value);
        CMatrix_solver::Cvector bb = b0, xx = x0;
        S.product(bb, xx, tmp);
        put("product", tmp.value);

        // the factor and the triangular solves as ILU_BICGSTAB_solve uses them
        S.M.copy_same_structure(S.matrix);
        S.ILU_decompose(S.matrix, S.M, S.diag_index);
        put("ilu_m", S.M.value);
        CMatrix_solver::Cvector y(m.n), z(m.n);
        S.ILU_forsub(S.M, S.diag_index, b0, y);
        put("forsub_y", y.value);
        S.ILU_backsub(S.M, S.diag_index, y, z);
        put("backsub_x", z.value);

        // the two REFERENCE solvers, then solve() dispatching on the member
        const double r_cg = S.CG_solve();
        put1("cg_res", r_cg);
        put("cg_x", S.x_vector.value);
        S.x_vector = x0;
        const double r_bi = S.ILU_BICGSTAB_solve();
        put1("bicg_res", r_bi);
        put("bicg_x", S.x_vector.value);
        put("bicg_M", S.M.value);
        assemble(m, Matrix_solver, 0.7);
        Matrix_solver->method = OCSL_CG;
        Matrix_solver->solve();  // cflow_solver.cpp:349
        put("solve_cg_x", Matrix_solver->x_vector.value);
        assemble(m, Matrix_solver, 0.7);
        Matrix_solver->method = OCSL_ILU_BICGSTAB;
        Matrix_solver->mode = OCSL_MODE_TOLERANCE;
        Matrix_solver->options.tol = 1e-12;
        Matrix_solver->solve();
        put("tol_x", Matrix_solver->x_vector.value);
        put1("tol_status", Matrix_solver->last_report.status);
        put1("tol_relres", Matrix_solver->last_report.true_relres);
    } catch (const ocsl::SolverError& e) {
        std::fprintf(stderr, "error %d: %s\n", (int)e.status, e.what());
        return 2;
    }
    return 0;
}
