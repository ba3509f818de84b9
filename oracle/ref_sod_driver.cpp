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
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "cconfiguration.h"
#include "cflow_solver.h"
#include "cfvm.h"
#include "cgeometry.h"
#include "cgeometry_supply.h"
#include "cmatrix_solver.h"
#include "croot.h"
#include "enums.h"

extern Croot root;

// Flow state recorded at a solve call, i.e. right after the assembly of the
// system being solved.  Per-cell arrays hold n values (3n for vectors, x y z
// interleaved), per-face arrays F, per-boundary-face arrays B; each points to
// max_state consecutive records.
struct RefFlowState {
    double *u, *u_old, *rho, *rho_old, *p, *p_old, *p_corr, *grad_p, *T, *h, *h_old, *ap, *ap_vec, *dt;
    double *f_mdot, *f_rho, *f_qy, *f_stress;
    double *b_T, *b_p, *b_rho, *b_mdot, *b_h, *b_qy, *b_u;
    int* b_super;
};

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
    // flow state at the first max_state solves (see RefFlowState)
    int max_state = 0;
    struct RefFlowState* state = nullptr;
};
Recorder g_rec;

// OpenMP threads of the shock-tube runs (1 for the fixtures: the reference's
// reductions and atomics then run in a fixed order) and the reference's own
// per-inner-iteration clock of the last run (Croot::start_solver
// accumulates each inner iteration's wall time in total_elapsed_iteration_time).
int g_threads = 1;
double g_last_iter_us = 0.0;

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
    if (k < R.max_state && R.state && root.Flow_solver) {
        const auto& fs = *root.Flow_solver;
        RefFlowState& st = *R.state;
        const int64_t n = (int64_t)fs.cells.size(), F = (int64_t)fs.faces.size(), B = (int64_t)fs.BC_faces.size();
        for (int64_t i = 0; i < n; ++i) {
            const auto& c = fs.cells[i];
            for (int d = 0; d < 3; ++d) {
                st.u[(k * n + i) * 3 + d] = c.Velocity.val[d];
                st.u_old[(k * n + i) * 3 + d] = c.Velocity_last_t.val[d];
                st.grad_p[(k * n + i) * 3 + d] = c.Pressure_grad.val[d];
                st.ap_vec[(k * n + i) * 3 + d] = c.ap_vec.val[d];
            }
            st.rho[k * n + i] = c.Density;
            st.rho_old[k * n + i] = c.Density_last_t;
            st.p[k * n + i] = c.Pressure;
            st.p_old[k * n + i] = c.Pressure_last_t;
            st.p_corr[k * n + i] = c.Pressure_corr;
            st.T[k * n + i] = c.Temperature;
            st.h[k * n + i] = c.Enthalpy;
            st.h_old[k * n + i] = c.Enthalpy_last_t;
            st.ap[k * n + i] = c.ap;
            st.dt[k * n + i] = c.delta_t;
        }
        for (int64_t f = 0; f < F; ++f) {
            const auto& fc = fs.faces[f];
            st.f_mdot[k * F + f] = fc.Mdot;
            st.f_rho[k * F + f] = fc.Density;
            st.f_qy[k * F + f] = fc.Qy;
            for (int d = 0; d < 3; ++d) st.f_stress[(k * F + f) * 3 + d] = fc.stress_explicit_part.val[d];
        }
        for (int64_t b = 0; b < B; ++b) {
            const auto& bc = fs.BC_faces[b];
            st.b_T[k * B + b] = bc.Temperature;
            st.b_p[k * B + b] = bc.Pressure;
            st.b_rho[k * B + b] = bc.Density;
            st.b_mdot[k * B + b] = bc.Mdot;
            st.b_h[k * B + b] = bc.Enthalpy;
            st.b_qy[k * B + b] = bc.Qy;
            st.b_super[k * B + b] = bc.supersonic_flag ? 1 : 0;
            for (int d = 0; d < 3; ++d) st.b_u[(k * B + b) * 3 + d] = bc.Velocity.val[d];
        }
    }
    const double res = use_cg ? self->CG_solve() : self->ILU_BICGSTAB_solve();
    if (dump) std::memcpy(R.dump_x + k * R.n, self->x_vector.value.data(), sizeof(double) * R.n);
    if (k < R.hist_cap) R.res_hist[k] = res;
    R.count++;
}

extern "C" {

// OpenMP threads of the next ref_sod_run (default 1) and the average inner
// iteration of the last one by the reference's own clock, in microseconds.
void ref_set_threads(int threads) { g_threads = threads > 0 ? threads : 1; }
double ref_last_iteration_us() { return g_last_iter_us; }

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
    c.openmp_number_of_threads = g_threads;
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

    const unsigned long long before = root.total_elapsed_iteration_time;
    root.run_config(c);
    g_last_iter_us = (double)(root.total_elapsed_iteration_time - before) / ((double)steps * inner);

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

// General 1D-mesh flow run: configurable physics, boundary conditions and
// initial condition; records the FVM mesh and the flow state at the first
// solves.  bc_* arrays have 3 entries (markers 0: x = 0 end, 1: side faces,
// 2: x = lx end), bc_type in enums::Eflow_BC_type order.
struct RefFlowConfig {
    int n_cells;
    double lx, ly, lz;
    int steps, inner;
    double dt;
    int steady;
    double implicit_underrelax, implicit_underrelax_max;
    int ramp_iterations;
    int solver_type, solver_mode, conv_type, visc_type;
    int init_shock_tube;
    double p0, T0, u0[3], p2, T2;
    int bc_type[3], bc_no_slip[3], bc_const_T[3];
    double bc_T[3], bc_p[3], bc_u[3][3];
};

struct RefFlowMesh {
    double* cell_volume;
    int *face_left, *face_right;
    double *face_area_vector, *face_unit, *face_alpha, *face_area, *face_dist, *face_ksi, *face_geo_aver;
    int *bc_left, *bc_marker;
    double *bc_area_vector, *bc_unit, *bc_alpha, *bc_area, *bc_dist, *bc_ksi, *bc_perp;
};

int ref_flow_sizes(int n_cells, double lx, double ly, double lz, int64_t* n_faces, int64_t* n_bc, int64_t* nnz) {
    Cgeometry g = generate_1d_mesh(lx, ly, lz, n_cells);
    int64_t F = 0, B = 0;
    for (auto& f : g.faces) (f.is_boundary() ? B : F)++;
    *n_faces = F;
    *n_bc = B;
    *nnz = 0;
    for (auto& c : g.cells) {
        std::vector<int> nb(c.i_neighbours.begin(), c.i_neighbours.end());
        nb.push_back((int)(&c - &g.cells[0]));
        std::sort(nb.begin(), nb.end());
        *nnz += (int64_t)(std::unique(nb.begin(), nb.end()) - nb.begin());
    }
    return 0;
}

// The run settings of a RefFlowConfig common to every mesh source.
static void flow_settings(const RefFlowConfig* cf, Cconfiguration& c) {
    c.openmp_number_of_threads = 1;
    c.eproblem_type = enums::Eproblem_type::Fluid_flow;
    c.eflow_type_visc = cf->visc_type ? enums::Eflow_type_visc::Inviscid : enums::Eflow_type_visc::Viscous;
    c.etimestepping_type = cf->steady ? enums::Etimestepping_type::Steady : enums::Etimestepping_type::Transient;
    c.ematrix_solver_type = cf->solver_type ? enums::Ematrix_solver_type::ILU_BICSTAB : enums::Ematrix_solver_type::CG;
    c.econv_flux_type = cf->conv_type ? enums::Econv_flux_type::Blended_central : enums::Econv_flux_type::First_order_upwind;
    c.implicit_underrelax = cf->implicit_underrelax;
    c.implicit_underrelax_max = cf->implicit_underrelax_max;
    c.underrelax_ramp_iteration = cf->ramp_iterations;
    c.global_delta_t = cf->dt;
    c.inital_condition->einitial_condition_type =
        cf->init_shock_tube ? enums::Einitial_condition_type::Shock_tube : enums::Einitial_condition_type::Uniform;
    c.inital_condition->Pressure = cf->p0;
    c.inital_condition->Temperature = cf->T0;
    c.inital_condition->Pressure2 = cf->p2;
    c.inital_condition->Temperature2 = cf->T2;
    for (int d = 0; d < 3; ++d) c.inital_condition->Velocity.val[d] = cf->u0[d];
    c.max_iteration = cf->steps;
    c.max_inner_iteration = cf->inner;
}

// One boundary condition of the configuration.
static void flow_bc(Cconfiguration& c, int m, int type, int no_slip, int const_T, double T, double p, const double* u) {
    auto& bc = c.boundary_conditions[m];
    bc.eflow_BC_type = (enums::Eflow_BC_type)type;
    bc.ewall_BC_type = no_slip ? enums::Ewall_BC_type::no_slip : enums::Ewall_BC_type::slip;
    bc.ethermal_BC_type = const_T ? enums::Ethermal_BC_type::constant_temperature : enums::Ethermal_BC_type::adiabatic;
    bc.Temperature = T;
    bc.Pressure = p;
    for (int d = 0; d < 3; ++d) bc.Velocity.val[d] = u[d];
}

int64_t ref_flow_run(const RefFlowConfig* cf, int64_t hist_cap, double* res_hist, int max_dump, int64_t nnz_cap,
                     uint64_t* rp_out, uint64_t* col_out, int64_t* nnz_out, double* dump_val, double* dump_b,
                     double* dump_x0, double* dump_x, int* dump_method, int max_state, RefFlowState* state,
                     RefFlowMesh* mesh, double* p, double* u, double* T, double* rho, double* h) {
    Cconfiguration c;
    c.openmp_number_of_threads = 1;
    c.eproblem_type = enums::Eproblem_type::Fluid_flow;
    c.eflow_type_visc = cf->visc_type ? enums::Eflow_type_visc::Inviscid : enums::Eflow_type_visc::Viscous;
    c.etimestepping_type = cf->steady ? enums::Etimestepping_type::Steady : enums::Etimestepping_type::Transient;
    c.ematrix_solver_type = cf->solver_type ? enums::Ematrix_solver_type::ILU_BICSTAB : enums::Ematrix_solver_type::CG;
    c.econv_flux_type = cf->conv_type ? enums::Econv_flux_type::Blended_central : enums::Econv_flux_type::First_order_upwind;
    c.implicit_underrelax = cf->implicit_underrelax;
    c.implicit_underrelax_max = cf->implicit_underrelax_max;
    c.underrelax_ramp_iteration = cf->ramp_iterations;
    c.emesh_source = enums::Emesh_source::generate_1D;
    c.geo_parameters_1D[0] = cf->lx;
    c.geo_parameters_1D[1] = cf->ly;
    c.geo_parameters_1D[2] = cf->lz;
    c.geo_parameters_1D[3] = cf->n_cells;
    c.global_delta_t = cf->dt;
    c.inital_condition->einitial_condition_type =
        cf->init_shock_tube ? enums::Einitial_condition_type::Shock_tube : enums::Einitial_condition_type::Uniform;
    c.inital_condition->Pressure = cf->p0;
    c.inital_condition->Temperature = cf->T0;
    c.inital_condition->Pressure2 = cf->p2;
    c.inital_condition->Temperature2 = cf->T2;
    for (int d = 0; d < 3; ++d) c.inital_condition->Velocity.val[d] = cf->u0[d];
    c.set_BC_number(3);
    for (int m = 0; m < 3; ++m) {
        auto& bc = c.boundary_conditions[m];
        bc.eflow_BC_type = (enums::Eflow_BC_type)cf->bc_type[m];
        bc.ewall_BC_type = cf->bc_no_slip[m] ? enums::Ewall_BC_type::no_slip : enums::Ewall_BC_type::slip;
        bc.ethermal_BC_type = cf->bc_const_T[m] ? enums::Ethermal_BC_type::constant_temperature
                                                 : enums::Ethermal_BC_type::adiabatic;
        bc.Temperature = cf->bc_T[m];
        bc.Pressure = cf->bc_p[m];
        for (int d = 0; d < 3; ++d) bc.Velocity.val[d] = cf->bc_u[m][d];
    }
    c.max_iteration = cf->steps;
    c.max_inner_iteration = cf->inner;

    g_rec = Recorder();
    g_rec.solver_mode = cf->solver_mode;
    g_rec.hist_cap = hist_cap;
    g_rec.res_hist = res_hist;
    g_rec.max_dump = max_dump;
    g_rec.n = cf->n_cells;
    g_rec.nnz_cap = nnz_cap;
    g_rec.dump_val = dump_val;
    g_rec.dump_b = dump_b;
    g_rec.dump_x0 = dump_x0;
    g_rec.dump_x = dump_x;
    g_rec.dump_method = dump_method;
    g_rec.max_state = max_state;
    g_rec.state = state;

    root.run_config(c);

    const CMatrix_solver& S = *root.Matrix_solver;
    const int64_t nnz = (int64_t)S.matrix.value.size();
    *nnz_out = nnz;
    if (nnz > nnz_cap) return -1;
    std::memcpy(rp_out, S.matrix.row_ptr.data(), sizeof(uint64_t) * (cf->n_cells + 1));
    std::memcpy(col_out, S.matrix.column_index.data(), sizeof(uint64_t) * nnz);
    const auto& cells = root.Flow_solver->cells;
    for (int i = 0; i < cf->n_cells; ++i) {
        p[i] = cells[i].Pressure;
        for (int d = 0; d < 3; ++d) u[3 * i + d] = cells[i].Velocity.val[d];
        T[i] = cells[i].Temperature;
        rho[i] = cells[i].Density;
        h[i] = cells[i].Enthalpy;
    }
    if (mesh) {
        const CFVM& fvm = *root.FVM;
        for (size_t i = 0; i < fvm.cells.size(); ++i) mesh->cell_volume[i] = fvm.cells[i].volume;
        for (size_t f = 0; f < fvm.faces.size(); ++f) {
            const auto& fc = fvm.faces[f];
            mesh->face_left[f] = fc.left_cell_index;
            mesh->face_right[f] = fc.right_cell_index;
            for (int d = 0; d < 3; ++d) {
                mesh->face_area_vector[3 * f + d] = fc.area_vector.val[d];
                mesh->face_unit[3 * f + d] = fc.cell_center_unit_vector.val[d];
                mesh->face_alpha[3 * f + d] = fc.alpha.val[d];
            }
            mesh->face_area[f] = fc.area;
            mesh->face_dist[f] = fc.cell_center_distance;
            mesh->face_ksi[f] = fc.diffusion_ksi_A;
            mesh->face_geo_aver[f] = fc.geo_aver;
        }
        for (size_t b = 0; b < fvm.BC_faces.size(); ++b) {
            const auto& bc = fvm.BC_faces[b];
            mesh->bc_left[b] = bc.left_cell_index;
            mesh->bc_marker[b] = root.geometry->faces[bc.global_face_index].boundary_marker;
            for (int d = 0; d < 3; ++d) {
                mesh->bc_area_vector[3 * b + d] = bc.area_vector.val[d];
                mesh->bc_unit[3 * b + d] = bc.cell_center_unit_vector.val[d];
                mesh->bc_alpha[3 * b + d] = bc.alpha.val[d];
            }
            mesh->bc_area[b] = bc.area;
            mesh->bc_dist[b] = bc.cell_center_distance;
            mesh->bc_ksi[b] = bc.diffusion_ksi_A;
            mesh->bc_perp[b] = bc.perp_dist;
        }
    }
    return g_rec.count;
}

// The same run on a mesh read from an SU2 file (Emesh_source::from_file,
// croot.cpp:72-76) with n_bc boundary conditions, one per marker in file
// order; cf's 1D geometry and its three boundary conditions are not used.
// Outputs: the per-solve residual history, the matrix structure and the
// final fields of the n_cells cells.
int64_t ref_flow_run_file(const RefFlowConfig* cf, const char* path, int n_bc, const int* bc_type,
                          const int* bc_no_slip, const int* bc_const_T, const double* bc_T, const double* bc_p,
                          const double* bc_u, int64_t n_cells, int64_t hist_cap, double* res_hist, int64_t nnz_cap,
                          uint64_t* rp_out, uint64_t* col_out, int64_t* nnz_out, double* p, double* u, double* T,
                          double* rho, double* h) {
    Cconfiguration c;
    flow_settings(cf, c);
    c.emesh_source = enums::Emesh_source::from_file;
    c.mesh_filename = path;
    c.set_BC_number(n_bc);
    for (int m = 0; m < n_bc; ++m)
        flow_bc(c, m, bc_type[m], bc_no_slip[m], bc_const_T[m], bc_T[m], bc_p[m], bc_u + 3 * m);
    g_rec = Recorder();
    g_rec.solver_mode = cf->solver_mode;
    g_rec.hist_cap = hist_cap;
    g_rec.res_hist = res_hist;
    g_rec.n = (int)n_cells;
    g_rec.nnz_cap = nnz_cap;
    root.run_config(c);
    const auto& cells = root.Flow_solver->cells;
    if ((int64_t)cells.size() != n_cells) return -2;
    const CMatrix_solver& S = *root.Matrix_solver;
    const int64_t nnz = (int64_t)S.matrix.value.size();
    *nnz_out = nnz;
    if (nnz > nnz_cap) return -1;
    std::memcpy(rp_out, S.matrix.row_ptr.data(), sizeof(uint64_t) * (n_cells + 1));
    std::memcpy(col_out, S.matrix.column_index.data(), sizeof(uint64_t) * nnz);
    for (int64_t i = 0; i < n_cells; ++i) {
        p[i] = cells[i].Pressure;
        for (int d = 0; d < 3; ++d) u[3 * i + d] = cells[i].Velocity.val[d];
        T[i] = cells[i].Temperature;
        rho[i] = cells[i].Density;
        h[i] = cells[i].Enthalpy;
    }
    return g_rec.count;
}

// The same run on `threads` OpenMP threads, timed by the reference itself:
// Croot::start_solver accumulates each inner iteration's wall time in
// total_elapsed_iteration_time (croot.cpp:126-219; the time-step update and
// the VTK output are outside it).  iter_us receives the average per inner
// iteration, total_s the whole run (mesh reading, FVM, matrix structure,
// initialisation, iterations, VTK files).  Outputs as ref_flow_run_file:
// the final fields of the n_cells cells.  Returns the number of solves.
int64_t ref_flow_time_file(const RefFlowConfig* cf, const char* path, int n_bc, const int* bc_type,
                           const int* bc_no_slip, const int* bc_const_T, const double* bc_T, const double* bc_p,
                           const double* bc_u, int threads, int64_t n_cells, double* iter_us, double* total_s,
                           double* p, double* u, double* T, double* rho, double* h) {
    Cconfiguration c;
    flow_settings(cf, c);
    c.openmp_number_of_threads = threads > 0 ? threads : 1;
    c.emesh_source = enums::Emesh_source::from_file;
    c.mesh_filename = path;
    c.set_BC_number(n_bc);
    for (int m = 0; m < n_bc; ++m)
        flow_bc(c, m, bc_type[m], bc_no_slip[m], bc_const_T[m], bc_T[m], bc_p[m], bc_u + 3 * m);
    g_rec = Recorder();
    g_rec.solver_mode = cf->solver_mode;
    const unsigned long long before = root.total_elapsed_iteration_time;
    const auto t0 = std::chrono::steady_clock::now();
    root.run_config(c);
    *total_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const int inner = c.etimestepping_type == enums::Etimestepping_type::Steady ? 1 : cf->inner;
    *iter_us = (double)(root.total_elapsed_iteration_time - before) / ((double)cf->steps * inner);
    const auto& cells = root.Flow_solver->cells;
    if ((int64_t)cells.size() != n_cells) return -2;
    for (int64_t i = 0; i < n_cells; ++i) {
        p[i] = cells[i].Pressure;
        for (int d = 0; d < 3; ++d) u[3 * i + d] = cells[i].Velocity.val[d];
        T[i] = cells[i].Temperature;
        rho[i] = cells[i].Density;
        h[i] = cells[i].Enthalpy;
    }
    return g_rec.count;
}

}  // extern "C"
