/*
 * ocsl_gen.h -- synthetic linear systems for the OpenCSL MI355X solver path.
 *
 * These generators have no counterpart in the reference solver (it builds its
 * systems from a mesh inside CFVM::define_matrix_structure, cfvm.cpp:74-98, and
 * the assembly loops of cflow_solver_cell_face_loops.cpp).  They produce the
 * benchmark and parity systems from seeds so that the same bytes are obtained
 * on any machine:
 *
 *   engine      std::mt19937_64(seed)  (output sequence fixed by the C++ standard)
 *   U[0,1)      (e >> 11) * 2^-53
 *   N(0,1)      Box-Muller, cosine branch: sqrt(-2 ln(1-U1)) * cos(2 pi U2)
 *   lognormal   exp(sigma * N)
 *   permutation Fisher-Yates from the top, j = floor(U * (i+1)); cell with
 *               natural id c = i + n*(j + n*k) gets the new id perm[c]
 *
 * Matrices are emitted in the reference's CSR layout: uint64 row_ptr[n+1],
 * uint64 column_index[nnz] sorted ascending per row, double value[nnz].  With
 * `sentinels` != 0 every boundary row additionally ends with the column
 * (uint64_t)-1 holding 0.0, which is what CFVM::define_matrix_structure
 * produces for cells that own a boundary face (cfvm.cpp:90-93).
 *
 * Every function returns 0 on success and a negative value on bad arguments.
 * Nothing here touches a GPU.
 */
#ifndef OCSL_GEN_H
#define OCSL_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Sentinel column written into boundary rows (the reference's (size_t)-1). */
#define OCSL_SENTINEL_COL ((uint64_t)-1)

/* Row and entry counts of the n^3 seven-point hex-mesh systems below. */
int ocsl_gen_hex_sizes(int n, int sentinels, int64_t* n_rows, int64_t* nnz);

/*
 * SPD pressure-correction system on an n^3 hex mesh of the unit cube.
 * Face coefficient a_f = rho_f * A_f / d_f = rho_f with rho_f ~ lognormal(0, sigma);
 * A_ij = -a_f, A_ii = sum_f a_f + leak * [cell owns a boundary face].
 * Stream of `seed_mesh`: the permutation draws (if permute), then rho_f for
 * x-faces, y-faces, z-faces, each in natural order (i fastest).
 * `perm_out` (optional, length n^3) receives perm[c].
 * `rhs_out` (optional, length n^3) receives b ~ U[-1,1] drawn in new-id
 * order from mt19937_64(seed_rhs), or, when seed_rhs == UINT64_MAX, from the
 * continuation of the mesh stream.
 */
int ocsl_gen_poisson3d(int n, uint64_t seed_mesh, uint64_t seed_rhs, double sigma,
                       double leak, int permute, int sentinels,
                       uint64_t* row_ptr, uint64_t* col, double* val,
                       double* rhs_out, int64_t* perm_out);

/*
 * Nonsymmetric momentum-like convection-diffusion system on an n^3 hex mesh of
 * [0, 2 pi]^3.  Face mass flux mdot_f = rho * u_f . S_f with the Taylor-Green
 * field u = sin x cos y cos z, v = -cos x sin y cos z, w = 0 at face centres.
 * First-order upwind convection and central diffusion (D = Gamma * A_f / d_f,
 * Gamma = rho * h / peclet so that the cell Peclet number with |u|max = 1 is
 * `peclet`), assembled with the reference's upwind/diffusive stencil rules,
 * plus the implicit time term rho * V / dt on the diagonal.  The box faces carry
 * no flux (the Taylor-Green normal velocity vanishes there).
 * Stream of `seed_perm`: the permutation draws.
 */
int ocsl_gen_convdiff3d(int n, uint64_t seed_perm, double rho, double dt,
                        double peclet, int permute, int sentinels,
                        uint64_t* row_ptr, uint64_t* col, double* val,
                        int64_t* perm_out);

/* out[i] = lo + (hi - lo) * U_i from mt19937_64(seed), i = 0..n-1. */
int ocsl_gen_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out);

/* out[i] = N_i(0,1) from mt19937_64(seed) (two uniforms per value). */
int ocsl_gen_normal(uint64_t seed, int64_t n, double* out);

/*
 * Sizes and seeds of m independent systems: from mt19937_64(seed), first
 * n_k = round(exp(ln lo + U_k (ln hi - ln lo))) for k = 0..m-1, then two raw
 * 64-bit outputs per system (seed_mesh_k, seed_rhs_k).
 */
int ocsl_gen_batch_plan(uint64_t seed, int m, int lo, int hi, int* n_out,
                        uint64_t* seed_mesh_out, uint64_t* seed_rhs_out);

/*
 * Seeded random sample of `count` distinct CSR slots of real (non-sentinel)
 * entries, drawn by rejection from mt19937_64(seed); slots ascending.
 */
int ocsl_gen_sample_slots(uint64_t seed, int64_t n_rows, const uint64_t* row_ptr,
                          const uint64_t* col, int count, int64_t* slots_out);

/* 64-bit FNV-1a hash over the bytes of (row_ptr, col, val, b); b may be NULL. */
uint64_t ocsl_gen_hash(int64_t n_rows, const uint64_t* row_ptr, const uint64_t* col,
                       const double* val, const double* b);

#ifdef __cplusplus
}
#endif

#endif /* OCSL_GEN_H */
