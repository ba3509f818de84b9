// ref_su2_driver.cpp -- C entry points into the reference's mesh set-up,
// compiled with the reference's own sources where they lie (oracle/Makefile).
// TEST INFRASTRUCTURE: pins libocsl_mesh.so; never shipped, never on the
// product path.
//
// ref_su2_read runs Read_SU2_mesh (cgeometry_supply.h:201-433),
// CFVM::create_FVM_from_geometry (cfvm.cpp:100-170) and
// CFVM::define_matrix_structure (cfvm.cpp:74-98) on one file and keeps the
// result; ref_su2_copy copies its arrays out in the reference's order.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>

#include "cfvm.h"
#include "cgeometry_supply.h"
#include "cmatrix_solver.h"
#include "cflow_solver.h"
#include "cfluid_properties.h"
#include "croot.h"

namespace {
std::shared_ptr<CFVM> g_fvm;
std::shared_ptr<CMatrix_solver> g_ms;
}  // namespace

extern "C" {

// sizes: n_cells, n_faces, n_bc, nnz, n_markers.  stage_s, when given,
// receives the seconds of the three stages: Read_SU2_mesh,
// create_FVM_from_geometry, define_matrix_structure.
int ref_su2_read_timed(const char* path, int64_t* sizes, double* stage_s) {
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    root.geometry = std::make_shared<Cgeometry>(Read_SU2_mesh(path));
    const auto t1 = clk::now();
    g_fvm = std::make_shared<CFVM>();
    root.FVM = g_fvm;
    g_fvm->create_FVM_from_geometry();
    const auto t2 = clk::now();
    g_fvm->define_matrix_structure(g_ms);
    const auto t3 = clk::now();
    if (stage_s) {
        stage_s[0] = std::chrono::duration<double>(t1 - t0).count();
        stage_s[1] = std::chrono::duration<double>(t2 - t1).count();
        stage_s[2] = std::chrono::duration<double>(t3 - t2).count();
    }
    sizes[0] = (int64_t)g_fvm->cells.size();
    sizes[1] = (int64_t)g_fvm->faces.size();
    sizes[2] = (int64_t)g_fvm->BC_faces.size();
    sizes[3] = (int64_t)g_ms->matrix.column_index.size();
    sizes[4] = (int64_t)root.geometry->n_boundary;
    return 0;
}

int ref_su2_read(const char* path, int64_t* sizes) { return ref_su2_read_timed(path, sizes, nullptr); }

static void put3(double* d, size_t i, Svector v) {
    for (int k = 0; k < 3; ++k) d[3 * i + k] = v[k];
}

void ref_su2_copy(double* vol, double* center, int32_t* fl, int32_t* fr, double* fS, double* fe, double* fdist,
                  double* falpha, double* fksi, double* fgeo, double* farea, int32_t* bl, int32_t* bm, double* bS,
                  double* be, double* bdist, double* balpha, double* bksi, double* bperp, double* barea,
                  uint64_t* row_ptr, uint64_t* col) {
    const CFVM& F = *g_fvm;
    for (size_t i = 0; i < F.cells.size(); ++i) {
        vol[i] = F.cells[i].volume;
        put3(center, i, root.geometry->cells[i].center);
    }
    for (size_t i = 0; i < F.faces.size(); ++i) {
        const CFVM::CFVM_face& f = F.faces[i];
        fl[i] = f.left_cell_index;
        fr[i] = f.right_cell_index;
        put3(fS, i, f.area_vector);
        put3(fe, i, f.cell_center_unit_vector);
        fdist[i] = f.cell_center_distance;
        put3(falpha, i, f.alpha);
        fksi[i] = f.diffusion_ksi_A;
        fgeo[i] = f.geo_aver;
        farea[i] = f.area;
    }
    for (size_t i = 0; i < F.BC_faces.size(); ++i) {
        const CFVM::CFVM_BC_face& f = F.BC_faces[i];
        bl[i] = f.left_cell_index;
        bm[i] = root.geometry->faces[(size_t)f.global_face_index].boundary_marker;
        put3(bS, i, f.area_vector);
        put3(be, i, f.cell_center_unit_vector);
        bdist[i] = f.cell_center_distance;
        put3(balpha, i, f.alpha);
        bksi[i] = f.diffusion_ksi_A;
        bperp[i] = f.perp_dist;
        barea[i] = f.area;
    }
    const auto& rp = g_ms->matrix.row_ptr;
    const auto& ci = g_ms->matrix.column_index;
    for (size_t i = 0; i < rp.size(); ++i) row_ptr[i] = rp[i];
    for (size_t i = 0; i < ci.size(); ++i) col[i] = ci[i];
}

// After a reference run (ref_sod_run): its mesh as nodes and SU2-ordered
// cells (the inverse of the reader's hexahedron reordering,
// cgeometry_supply.h:266-276), and the cell fields
// Cpost_processing::write_VTK_data writes (cpost_processing.cpp:26-127),
// viscosity from Cfluid_properties::get_viscosity.
void ref_last_run_sizes(int64_t* n_nodes, int64_t* n_cells) {
    *n_nodes = (int64_t)root.geometry->nodes.size();
    *n_cells = (int64_t)root.geometry->cells.size();
}

void ref_last_run_dump(double* nodes, int32_t* cell_type, int32_t* cell_nodes, double* rho, double* p, double* T,
                       double* mu, double* mach, double* u, double* gp) {
    const Cgeometry& g = *root.geometry;
    for (size_t i = 0; i < g.nodes.size(); ++i) put3(nodes, i, g.nodes[i]);
    const int hex_su2[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    for (size_t c = 0; c < g.cells.size(); ++c) {
        const auto& cell = g.cells[c];
        const int code = cell.ecell_type == enums::Ecell_type::tetra ? 10
                         : cell.ecell_type == enums::Ecell_type::wedge ? 13 : 12;
        cell_type[c] = code;
        for (int k = 0; k < 8; ++k) cell_nodes[8 * c + k] = -1;
        for (size_t k = 0; k < cell.i_nodes.size(); ++k)
            cell_nodes[8 * c + k] = cell.i_nodes[code == 12 ? hex_su2[k] : k];
    }
    const auto& cells = root.Flow_solver->cells;
    for (size_t c = 0; c < cells.size(); ++c) {
        rho[c] = cells[c].Density;
        p[c] = cells[c].Pressure;
        T[c] = cells[c].Temperature;
        mu[c] = root.Fluid_properties->get_viscosity(cells[c].Temperature);
        mach[c] = cells[c].Mach;
        put3(u, c, cells[c].Velocity);
        put3(gp, c, cells[c].Pressure_grad);
    }
}

// The boundary faces of the last run as Cpost_processing::write_boundary_VTK_file
// and write_boundary_VTK_data see them (cpost_processing.cpp:129-193,
// 335-401): per face of BC_face_global_indexes, its node count, nodes in the
// stored order and marker; per boundary face of the flow solver, the Mach
// number of its cell and CFlow_solver::calc_wall_stress
// (cflow_solver_supply.cpp:223-239).
void ref_last_run_bc_size(int64_t* n_bc) { *n_bc = (int64_t)root.geometry->BC_face_global_indexes.size(); }

void ref_last_run_bc_dump(int32_t* face_nn, int32_t* face_nodes, int32_t* marker, double* mach, double* stress) {
    const Cgeometry& g = *root.geometry;
    for (size_t q = 0; q < g.BC_face_global_indexes.size(); ++q) {
        const auto& f = g.faces[(size_t)g.BC_face_global_indexes[q]];
        face_nn[q] = (int32_t)f.i_nodes.size();
        for (int k = 0; k < 4; ++k) face_nodes[4 * q + k] = k < (int)f.i_nodes.size() ? f.i_nodes[(size_t)k] : -1;
        marker[q] = f.boundary_marker;
    }
    auto& fl = *root.Flow_solver;
    for (size_t q = 0; q < fl.BC_faces.size(); ++q) {
        mach[q] = fl.BC_faces[q].left_cell_ptr->Mach;
        const Svector w = fl.calc_wall_stress(root.FVM->BC_faces[q], fl.BC_faces[q], root.Fluid_properties);
        put3(stress, q, w);
    }
}

}  // extern "C"
