// ref_sod_driver.cpp -- runs the reference flow solver on the 1D shock tube and
// records every linear solve.  Compiled with the reference sources where they
// lie (oracle/Makefile); output to oracle/_ref/ only.  TEST INFRASTRUCTURE.
//
// The hook: the library is linked with --wrap for CMatrix_solver::solve(), so
// the calls from CFlow_solver::iterate (cflow_solver.cpp:289, 349, 460) land in
// __wrap__ZN14CMatrix_solver5solveEv below.  It records the system, dispatches
// to CG_solve / ILU_BICGSTAB_solve as solve() does (cmatrix_solver.cpp:764-780)
// and keeps the residual those return, which solve() discards.
#include <omp.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "cconfiguration.h"
#include "cflow_solver.h"
#include "cmatrix_solver.h"
#include "croot.h"
#include "enums.h"

extern Croot root;

namespace {

struct Recorder {
    int solver_mode = 0;  // 0: configured type for all; 1: CG for pressure, BiCGSTAB otherwise
    int64_t count = 0;
    int64_t hist_cap = 0;
    double* res_hist = nullptr;
    int max_dump = 0;
    int64_t n = 0, nnz_cap = 0;
    double *dump_val = nullptr, *dump_b = nullptr, *dump_x0 = nullptr, *dump_x = nullptr;
    int* dump_method = nullptr;
};
Recorder g_rec;

}  // namespace

extern "C" void __real__ZN14CMatrix_solver5solveEv(CMatrix_solver* self);

extern "C" void __wrap__ZN14CMatrix_solver5solveEv(CMatrix_solver* self) {
    Recorder& R = g_rec;
    const int64_t k = R.count;
    bool use_cg = root.configuration->ematrix_solver_type == enums::Ematrix_solver_type::CG;
    if (R.solver_mode == 1) use_cg = (k % 5) == 3;  // order per iteration: u, v, w, p', h
    const int64_t nnz = (int64_t)self->matrix.value.size();
    const bool dump = k < R.max_dump && nnz <= R.nnz_cap;
    if (dump) {
        std::memcpy(R.dump_val + k * R.nnz_cap, self->matrix.value.data(), sizeof(double) * nnz);
        std::memcpy(R.dump_b + k * R.n, self->b_vector.value.data(), sizeof(double) * R.n);
        std::memcpy(R.dump_x0 + k * R.n, self->x_vector.value.data(), sizeof(double) * R.n);
        R.dump_method[k] = use_cg ? 0 : 1;
    }
    const double res = use_cg ? self->CG_solve() : self->ILU_BICGSTAB_solve();
    if (dump) std::memcpy(R.dump_x + k * R.n, self->x_vector.value.data(), sizeof(double) * R.n);
    if (k < R.hist_cap) R.res_hist[k] = res;
    R.count++;
}

extern "C" {

// Returns the number of solves performed, or a negative value when the
// buffers are too small.  Field outputs are per cell in the reference's cell
// order.  The reference writes solution.vtk / solution_BC.vtk into the
// current directory at the end of the run (croot.cpp:239-240).
int64_t ref_sod_run(int n_cells, double lx, double ly, double lz, int steps, int inner, double dt,
                    int solver_type, int solver_mode, int conv_type, int visc_type, double p_left,
                    double t_left, double p_right, double t_right, int64_t hist_cap,
                    double* res_hist, int max_dump, int64_t nnz_cap, uint64_t* rp_out,
                    uint64_t* col_out, int64_t* nnz_out, double* dump_val, double* dump_b,
                    double* dump_x0, double* dump_x, int* dump_method, double* p, double* u,
                    double* T, double* rho, double* h) {
    Cconfiguration c;
    c.openmp_number_of_threads = 1;
    c.eproblem_type = enums::Eproblem_type::Fluid_flow;
    c.eflow_type_visc = visc_type ? enums::Eflow_type_visc::Inviscid : enums::Eflow_type_visc::Viscous;
    c.etimestepping_type = enums::Etimestepping_type::Transient;
    c.ematrix_solver_type = solver_type ? enums::Ematrix_solver_type::ILU_BICSTAB : enums::Ematrix_solver_type::CG;
    c.econv_flux_type = conv_type ? enums::Econv_flux_type::Blended_central : enums::Econv_flux_type::First_order_upwind;
    c.emesh_source = enums::Emesh_source::generate_1D;
    c.geo_parameters_1D[0] = lx;
    c.geo_parameters_1D[1] = ly;
    c.geo_parameters_1D[2] = lz;
    c.geo_parameters_1D[3] = n_cells;
    c.global_delta_t = dt;
    c.inital_condition->einitial_condition_type = enums::Einitial_condition_type::Shock_tube;
    c.inital_condition->Pressure = p_left;
    c.inital_condition->Temperature = t_left;
    c.inital_condition->Pressure2 = p_right;
    c.inital_condition->Temperature2 = t_right;
    c.inital_condition->Velocity.val[0] = 0.0;
    c.inital_condition->Velocity.val[1] = 0.0;
    c.inital_condition->Velocity.val[2] = 0.0;
    c.set_BC_number(3);
    c.boundary_conditions[0].eflow_BC_type = enums::Eflow_BC_type::wall;
    c.boundary_conditions[0].ewall_BC_type = enums::Ewall_BC_type::slip;
    c.boundary_conditions[1].eflow_BC_type = enums::Eflow_BC_type::empty;
    c.boundary_conditions[2].eflow_BC_type = enums::Eflow_BC_type::wall;
    c.boundary_conditions[2].ewall_BC_type = enums::Ewall_BC_type::slip;
    c.max_iteration = steps;
    c.max_inner_iteration = inner;

    g_rec = Recorder();
    g_rec.solver_mode = solver_mode;
    g_rec.hist_cap = hist_cap;
    g_rec.res_hist = res_hist;
    g_rec.max_dump = max_dump;
    g_rec.n = n_cells;
    g_rec.nnz_cap = nnz_cap;
    g_rec.dump_val = dump_val;
    g_rec.dump_b = dump_b;
    g_rec.dump_x0 = dump_x0;
    g_rec.dump_x = dump_x;
    g_rec.dump_method = dump_method;

    root.run_config(c);

    const CMatrix_solver& S = *root.Matrix_solver;
    const int64_t nnz = (int64_t)S.matrix.value.size();
    *nnz_out = nnz;
    if (nnz > nnz_cap) return -1;
    std::memcpy(rp_out, S.matrix.row_ptr.data(), sizeof(uint64_t) * (n_cells + 1));
    std::memcpy(col_out, S.matrix.column_index.data(), sizeof(uint64_t) * nnz);
    const auto& cells = root.Flow_solver->cells;
    for (int i = 0; i < n_cells; ++i) {
        p[i] = cells[i].Pressure;
        for (int d = 0; d < 3; ++d) u[3 * i + d] = cells[i].Velocity.val[d];
        T[i] = cells[i].Temperature;
        rho[i] = cells[i].Density;
        h[i] = cells[i].Enthalpy;
    }
    return g_rec.count;
}

}  // extern "C"
