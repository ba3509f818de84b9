// C++17 header-only bridge between the reference's flow-solver objects and
// the GPU assembly of ocsl_flow.h, for a CFlow_solver::iterate
// (cflow_solver.cpp:229-476) that keeps its own objects and swaps only its
// face / boundary-face / cell loops (cflow_solver_cell_face_loops.cpp) for
// the device kernels.  Templates over the reference's types, using their
// member names:
//   * the mesh, CFVM (cfvm.h:33-93): cells[i].volume; faces[f].left_cell_index,
//     right_cell_index, area_vector, cell_center_unit_vector, alpha, area,
//     cell_center_distance, diffusion_ksi_A, geo_aver; BC_faces[b]
//     .left_cell_index, BC_index and the same factors plus perp_dist;
//   * the flow state, CFlow_solver (cflow_solver.h:34-98): cells[i].Velocity,
//     Velocity_last_t, Density, Density_last_t, Pressure, Pressure_last_t,
//     Pressure_corr, Pressure_grad, Pressure_corr_grad, Temperature,
//     Enthalpy, Enthalpy_last_t, ap, ap_vec, delta_t, Mach; faces[f].Mdot,
//     Density, Qy, stress_explicit_part; BC_faces[b].Temperature, Pressure,
//     Out_Pressure, Density, Mdot, Enthalpy, Qy, Velocity, supersonic_flag;
//     Svector values as .val[3];
//   * the system, CMatrix_solver (cmatrix_solver.h:27-95, or the drop-in
//     ocsl_cmatrix_solver.hpp): matrix.value, b_vector.value, x_vector.value.
//
// In iterate, the block "clear; face loop; BC-face loop; cell loop" of an
// equation becomes one call:
//
//   ocsl::FlowAssembly gpu(*FVM, Matrix_solver->matrix, params);   // once
//   ...
//   gpu.assemble(OCSL_EQ_MOMENTUM_X + i_dimention, *this, *Matrix_solver);
//   Matrix_solver->solve();
//
// assemble() copies the whole state to the device, assembles there, and
// copies back the system and, for momentum, ap and ap_vec (which the
// reference's cell loop updates).  That is two PCIe transfers of the state
// per equation: the step that keeps the state resident on the device is
// ocsl_flow_iterate.
#ifndef OCSL_CFLOW_ADAPTER_HPP
#define OCSL_CFLOW_ADAPTER_HPP

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "ocsl_flow.h"

namespace ocsl {

class FlowError : public std::runtime_error {
public:
    FlowError(ocsl_status st, const std::string& what) : std::runtime_error(what), status(st) {}
    ocsl_status status;
};

inline void flow_check(ocsl_status st, const char* what) {
    if (st < 0) throw FlowError(st, std::string(what) + ": " + ocsl_last_error());
}

namespace detail {
template <class V>
inline void put3(std::vector<double>& a, std::size_t i, const V& v) {
    a[3 * i] = v.val[0];
    a[3 * i + 1] = v.val[1];
    a[3 * i + 2] = v.val[2];
}
template <class V>
inline void get3(const std::vector<double>& a, std::size_t i, V& v) {
    v.val[0] = a[3 * i];
    v.val[1] = a[3 * i + 1];
    v.val[2] = a[3 * i + 2];
}
}  // namespace detail

// CFVM's faces and cells as the flat arrays ocsl_mesh points into.
struct FlowMeshArrays {
    int64_t n_cells = 0, n_faces = 0, n_bc = 0;
    std::vector<double> cell_volume;
    std::vector<int32_t> face_left, face_right, bc_left, bc_marker;
    std::vector<double> face_area_vector, face_unit, face_alpha, face_area, face_dist, face_ksi, face_geo_aver;
    std::vector<double> bc_area_vector, bc_unit, bc_alpha, bc_area, bc_dist, bc_ksi, bc_perp;

    ocsl_mesh view() const {
        return {n_cells,          n_faces,         n_bc,           cell_volume.data(),    face_left.data(),
                face_right.data(), face_area_vector.data(), face_unit.data(), face_alpha.data(), face_area.data(),
                face_dist.data(),  face_ksi.data(), face_geo_aver.data(), bc_left.data(),  bc_marker.data(),
                bc_area_vector.data(), bc_unit.data(), bc_alpha.data(), bc_area.data(), bc_dist.data(),
                bc_ksi.data(),     bc_perp.data()};
    }
};

template <class FVM>
FlowMeshArrays mesh_arrays(const FVM& fvm) {
    FlowMeshArrays m;
    const std::size_t n = fvm.cells.size(), F = fvm.faces.size(), B = fvm.BC_faces.size();
    m.n_cells = (int64_t)n;
    m.n_faces = (int64_t)F;
    m.n_bc = (int64_t)B;
    m.cell_volume.resize(n);
    for (std::size_t i = 0; i < n; ++i) m.cell_volume[i] = fvm.cells[i].volume;
    m.face_left.resize(F);
    m.face_right.resize(F);
    m.face_area_vector.resize(3 * F);
    m.face_unit.resize(3 * F);
    m.face_alpha.resize(3 * F);
    m.face_area.resize(F);
    m.face_dist.resize(F);
    m.face_ksi.resize(F);
    m.face_geo_aver.resize(F);
    for (std::size_t f = 0; f < F; ++f) {
        const auto& fc = fvm.faces[f];
        m.face_left[f] = (int32_t)fc.left_cell_index;
        m.face_right[f] = (int32_t)fc.right_cell_index;
        detail::put3(m.face_area_vector, f, fc.area_vector);
        detail::put3(m.face_unit, f, fc.cell_center_unit_vector);
        detail::put3(m.face_alpha, f, fc.alpha);
        m.face_area[f] = fc.area;
        m.face_dist[f] = fc.cell_center_distance;
        m.face_ksi[f] = fc.diffusion_ksi_A;
        m.face_geo_aver[f] = fc.geo_aver;
    }
    m.bc_left.resize(B);
    m.bc_marker.resize(B);
    m.bc_area_vector.resize(3 * B);
    m.bc_unit.resize(3 * B);
    m.bc_alpha.resize(3 * B);
    m.bc_area.resize(B);
    m.bc_dist.resize(B);
    m.bc_ksi.resize(B);
    m.bc_perp.resize(B);
    for (std::size_t b = 0; b < B; ++b) {
        const auto& bf = fvm.BC_faces[b];
        m.bc_left[b] = (int32_t)bf.left_cell_index;
        m.bc_marker[b] = (int32_t)bf.BC_index;
        detail::put3(m.bc_area_vector, b, bf.area_vector);
        detail::put3(m.bc_unit, b, bf.cell_center_unit_vector);
        detail::put3(m.bc_alpha, b, bf.alpha);
        m.bc_area[b] = bf.area;
        m.bc_dist[b] = bf.cell_center_distance;
        m.bc_ksi[b] = bf.diffusion_ksi_A;
        m.bc_perp[b] = bf.perp_dist;
    }
    return m;
}

// CFlow_solver's cell, face and boundary-face fields as the flat arrays
// ocsl_flow_state points into.
struct FlowStateArrays {
    std::size_t n = 0, F = 0, B = 0;
    std::vector<double> u, u_old, rho, rho_old, p, p_old, p_corr, grad_p, grad_pc, T, h, h_old, ap, ap_vec, dt,
        mach;
    std::vector<double> f_mdot, f_rho, f_qy, f_stress;
    std::vector<double> b_T, b_p, b_p_out, b_rho, b_mdot, b_h, b_qy, b_u;
    std::vector<int> b_supersonic;

    void resize(std::size_t n_, std::size_t F_, std::size_t B_) {
        n = n_;
        F = F_;
        B = B_;
        for (auto* v : {&u, &u_old, &grad_p, &grad_pc, &ap_vec}) v->assign(3 * n, 0.0);
        for (auto* v : {&rho, &rho_old, &p, &p_old, &p_corr, &T, &h, &h_old, &ap, &dt, &mach}) v->assign(n, 0.0);
        for (auto* v : {&f_mdot, &f_rho, &f_qy}) v->assign(F, 0.0);
        f_stress.assign(3 * F, 0.0);
        for (auto* v : {&b_T, &b_p, &b_p_out, &b_rho, &b_mdot, &b_h, &b_qy}) v->assign(B, 0.0);
        b_u.assign(3 * B, 0.0);
        b_supersonic.assign(B, 0);
    }
    ocsl_flow_state view() {
        ocsl_flow_state s;
        s.u = u.data();
        s.u_old = u_old.data();
        s.rho = rho.data();
        s.rho_old = rho_old.data();
        s.p = p.data();
        s.p_old = p_old.data();
        s.p_corr = p_corr.data();
        s.grad_p = grad_p.data();
        s.grad_pc = grad_pc.data();
        s.T = T.data();
        s.h = h.data();
        s.h_old = h_old.data();
        s.ap = ap.data();
        s.ap_vec = ap_vec.data();
        s.dt = dt.data();
        s.mach = mach.data();
        s.f_mdot = f_mdot.data();
        s.f_rho = f_rho.data();
        s.f_qy = f_qy.data();
        s.f_stress = f_stress.data();
        s.b_T = b_T.data();
        s.b_p = b_p.data();
        s.b_p_out = b_p_out.data();
        s.b_rho = b_rho.data();
        s.b_mdot = b_mdot.data();
        s.b_h = b_h.data();
        s.b_qy = b_qy.data();
        s.b_u = b_u.data();
        s.b_supersonic = b_supersonic.data();
        return s;
    }
};

template <class Flow>
void gather_state(const Flow& fl, FlowStateArrays& s) {
    s.resize(fl.cells.size(), fl.faces.size(), fl.BC_faces.size());
    for (std::size_t i = 0; i < s.n; ++i) {
        const auto& c = fl.cells[i];
        detail::put3(s.u, i, c.Velocity);
        detail::put3(s.u_old, i, c.Velocity_last_t);
        s.rho[i] = c.Density;
        s.rho_old[i] = c.Density_last_t;
        s.p[i] = c.Pressure;
        s.p_old[i] = c.Pressure_last_t;
        s.p_corr[i] = c.Pressure_corr;
        detail::put3(s.grad_p, i, c.Pressure_grad);
        detail::put3(s.grad_pc, i, c.Pressure_corr_grad);
        s.T[i] = c.Temperature;
        s.h[i] = c.Enthalpy;
        s.h_old[i] = c.Enthalpy_last_t;
        s.ap[i] = c.ap;
        detail::put3(s.ap_vec, i, c.ap_vec);
        s.dt[i] = c.delta_t;
        s.mach[i] = c.Mach;
    }
    for (std::size_t f = 0; f < s.F; ++f) {
        const auto& fc = fl.faces[f];
        s.f_mdot[f] = fc.Mdot;
        s.f_rho[f] = fc.Density;
        s.f_qy[f] = fc.Qy;
        detail::put3(s.f_stress, f, fc.stress_explicit_part);
    }
    for (std::size_t b = 0; b < s.B; ++b) {
        const auto& bf = fl.BC_faces[b];
        s.b_T[b] = bf.Temperature;
        s.b_p[b] = bf.Pressure;
        s.b_p_out[b] = bf.Out_Pressure;
        s.b_rho[b] = bf.Density;
        s.b_mdot[b] = bf.Mdot;
        s.b_h[b] = bf.Enthalpy;
        s.b_qy[b] = bf.Qy;
        detail::put3(s.b_u, b, bf.Velocity);
        s.b_supersonic[b] = bf.supersonic_flag ? 1 : 0;
    }
}

template <class Flow>
void scatter_state(const FlowStateArrays& s, Flow& fl) {
    for (std::size_t i = 0; i < s.n; ++i) {
        auto& c = fl.cells[i];
        detail::get3(s.u, i, c.Velocity);
        detail::get3(s.u_old, i, c.Velocity_last_t);
        c.Density = s.rho[i];
        c.Density_last_t = s.rho_old[i];
        c.Pressure = s.p[i];
        c.Pressure_last_t = s.p_old[i];
        c.Pressure_corr = s.p_corr[i];
        detail::get3(s.grad_p, i, c.Pressure_grad);
        detail::get3(s.grad_pc, i, c.Pressure_corr_grad);
        c.Temperature = s.T[i];
        c.Enthalpy = s.h[i];
        c.Enthalpy_last_t = s.h_old[i];
        c.ap = s.ap[i];
        detail::get3(s.ap_vec, i, c.ap_vec);
        c.delta_t = s.dt[i];
        c.Mach = s.mach[i];
    }
    for (std::size_t f = 0; f < s.F; ++f) {
        auto& fc = fl.faces[f];
        fc.Mdot = s.f_mdot[f];
        fc.Density = s.f_rho[f];
        fc.Qy = s.f_qy[f];
        detail::get3(s.f_stress, f, fc.stress_explicit_part);
    }
    for (std::size_t b = 0; b < s.B; ++b) {
        auto& bf = fl.BC_faces[b];
        bf.Temperature = s.b_T[b];
        bf.Pressure = s.b_p[b];
        bf.Out_Pressure = s.b_p_out[b];
        bf.Density = s.b_rho[b];
        bf.Mdot = s.b_mdot[b];
        bf.Enthalpy = s.b_h[b];
        bf.Qy = s.b_qy[b];
        detail::get3(s.b_u, b, bf.Velocity);
        bf.supersonic_flag = s.b_supersonic[b] != 0;
    }
}

// The GPU assembly of one mesh and matrix structure.
class FlowAssembly {
public:
    // fvm: the CFVM; matrix: the CCSR CFVM::define_matrix_structure built
    // (row_ptr, column_index, n_rows); prm: the configuration's choices.
    template <class FVM, class CCSR>
    FlowAssembly(const FVM& fvm, const CCSR& matrix, const ocsl_flow_params& prm) {
        static_assert(sizeof(matrix.row_ptr[0]) == sizeof(uint64_t), "64-bit CSR indices expected");
        const FlowMeshArrays m = mesh_arrays(fvm);
        const ocsl_mesh mv = m.view();
        flow_check(ocsl_flow_create(&f_, &mv, reinterpret_cast<const uint64_t*>(matrix.row_ptr.data()),
                                    reinterpret_cast<const uint64_t*>(matrix.column_index.data()), &prm),
                   "ocsl_flow_create");
        n_ = (std::size_t)m.n_cells;
        nnz_ = matrix.column_index.size();
    }
    ~FlowAssembly() {
        if (f_) ocsl_flow_destroy(f_);
    }
    FlowAssembly(const FlowAssembly&) = delete;
    FlowAssembly& operator=(const FlowAssembly&) = delete;

    // One equation's clear + face, boundary-face and cell loops
    // (cflow_solver.cpp:261-286, 319-343, 431-453): S.matrix.value,
    // S.b_vector.value and S.x_vector.value as those loops leave them; for
    // momentum, fl.cells[i].ap and .ap_vec as the cell loop leaves them.
    template <class Flow, class Solver>
    void assemble(int equation, Flow& fl, Solver& S) {
        gather_state(fl, st_);
        ocsl_flow_state sv = st_.view();
        flow_check(ocsl_flow_set_state(f_, &sv), "ocsl_flow_set_state");
        if (S.matrix.value.size() != nnz_ || S.b_vector.value.size() < n_ || S.x_vector.value.size() < n_)
            throw FlowError(OCSL_ERR_ARG, "FlowAssembly: system sizes differ from the structure");
        flow_check(ocsl_flow_assemble(f_, equation, S.matrix.value.data(), S.b_vector.value.data(),
                                      S.x_vector.value.data(), OCSL_HOST),
                   "ocsl_flow_assemble");
        if (equation <= OCSL_EQ_MOMENTUM_Z) {
            ocsl_flow_state out{};
            out.ap = st_.ap.data();
            out.ap_vec = st_.ap_vec.data();
            flow_check(ocsl_flow_get_state(f_, &out), "ocsl_flow_get_state");
            for (std::size_t i = 0; i < n_; ++i) {
                fl.cells[i].ap = st_.ap[i];
                detail::get3(st_.ap_vec, i, fl.cells[i].ap_vec);
            }
        }
    }

    ocsl_flow get() const { return f_; }

private:
    ocsl_flow f_ = nullptr;
    std::size_t n_ = 0, nnz_ = 0;
    FlowStateArrays st_;
};

}  // namespace ocsl

#endif  // OCSL_CFLOW_ADAPTER_HPP
