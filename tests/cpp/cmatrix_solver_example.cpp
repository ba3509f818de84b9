// A reference-style caller of the CMatrix_solver mirror: builds the 10x10
// system of the reference's CG unit test (cmatrix_solver_supply.h:29-82:
// diagonal 5, off-diagonals 1, b = 2, x0 = 1), calls CG_solve() and
// ILU_BICGSTAB_solve() as that test does, and prints the returned residuals and
// solutions (%.17g), one per line, for the test to compare.
#include <cstdio>

#include "ocsl_cmatrix_solver.hpp"

static void fill(ocsl::CMatrix_solver& S, int n) {
    std::size_t p = 0;
    for (int i = 0; i < n; ++i) {
        S.matrix.row_ptr[i] = p;
        for (int j = i - 1; j <= i + 1; ++j)
            if (j >= 0 && j < n) {
                S.matrix.column_index[p] = (std::size_t)j;
                S.matrix.value[p] = (i == j) ? 5.0 : 1.0;
                ++p;
            }
    }
    S.matrix.row_ptr[n] = p;
    for (int i = 0; i < n; ++i) {
        S.b_vector.value[i] = 2.0;
        S.x_vector.value[i] = 1.0;
    }
}

int main() {
    const int n = 10, nnz = 3 * n - 2;
    try {
        ocsl::CMatrix_solver S(n, n, nnz);
        fill(S, n);
        const double r_cg = S.CG_solve();
        std::printf("cg %.17g\n", r_cg);
        for (int i = 0; i < n; ++i) std::printf("x %.17g\n", S.x_vector.value[i]);
        for (int i = 0; i < n; ++i) S.x_vector.value[i] = 1.0;
        const double r_bi = S.ILU_BICGSTAB_solve();
        std::printf("bicgstab %.17g\n", r_bi);
        for (int i = 0; i < n; ++i) std::printf("x %.17g\n", S.x_vector.value[i]);
        ocsl::CMatrix_solver::Cvector ones;
        ones.value.assign(n, 1.0);
        std::printf("abssum %.17g\n", S.ABSsum(S.b_vector));
        std::printf("norm %.17g\n", S.norm(ones, S.b_vector));
    } catch (const ocsl::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", (int)e.status, e.what());
        return 2;
    }
    return 0;
}
