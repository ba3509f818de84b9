// Exercises include/ocsl_cflow_adapter.hpp with objects shaped like the
// reference's: CFVM with CFVM_cell / CFVM_face / CFVM_BC_face (cfvm.h:33-93),
// CFlow_solver with CFlow_cell / CFlow_face / CFlow_BC_face
// (cflow_solver.h:34-98), Svector as tensor.h:29 declares it, and the
// CMatrix_solver drop-in.  The classes below are stand-ins that carry the
// reference's member names and types (the adapter reads nothing else).
//
// Input (argv[1]): a text file of named arrays ("name count v0 v1 ...") with
// a reference run's mesh, matrix structure, flow parameters (raw bytes of
// ocsl_flow_params, "params_hex") and recorded flow states; written by
// tests/test_cflow_adapter.py from a golden fixture.
//   cflow_adapter_check FILE roundtrip   host only: AoS -> arrays -> AoS
//   cflow_adapter_check FILE assemble    each recorded state through
//                                        FlowAssembly::assemble on the GPU
// Output: named arrays, one per line (%.17g).
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "ocsl_cflow_adapter.hpp"
#include "ocsl_cmatrix_solver.hpp"

struct Svector {
    double val[3];
    Svector() { val[0] = val[1] = val[2] = 0.0; }
};

class CFVM {
public:
    class CFVM_cell {
    public:
        size_t index = 0;
        double volume = 0.0;
    };
    class CFVM_face {
    public:
        Svector area_vector, cell_center_unit_vector, alpha;
        double area = 0.0, cell_center_distance = 0.0, diffusion_ksi_A = 0.0, geo_aver = 0.0;
        CFVM_cell *left_cell_ptr = nullptr, *right_cell_ptr = nullptr;
        int global_face_index = -1, left_cell_index = -1, right_cell_index = -1;
    };
    class CFVM_BC_face {
    public:
        Svector area_vector, cell_center_unit_vector, alpha;
        double area = 0, cell_center_distance = 0, diffusion_ksi_A = 0, perp_dist = 0.0;
        CFVM_cell* left_cell_ptr = nullptr;
        int left_cell_index = -1, global_face_index = -1, BC_index = 0;
    };
    std::vector<CFVM_cell> cells;
    std::vector<CFVM_face> faces;
    std::vector<CFVM_BC_face> BC_faces;
};

class CFlow_solver {
public:
    class CFlow_cell {
    public:
        Svector Velocity, Velocity_last_t, Velocity_grad[3];
        double Density = 0.0, Density_last_t = 0.0;
        Svector Density_grad;
        double Pressure = 0.0, Pressure_last_t = 0.0;
        Svector Pressure_grad;
        double Pressure_corr = 0.0;
        Svector Pressure_corr_grad;
        double Enthalpy = 0.0, Enthalpy_last_t = 0.0;
        Svector Enthalpy_grad;
        double ap = 1.0, delta_t = 0.001;
        Svector ap_vec;
        double Temperature = 0.0, total_mdot = 0.0, Mach = 0.0;
    };
    class CFlow_face {
    public:
        Svector stress_explicit_part;
        double Mdot = 0.0, Density = 0.0, Qy = 1.0;
    };
    class CFlow_BC_face {
    public:
        double Temperature = 0.0, Pressure = 0.0, Out_Pressure = 0.0, Density = 0.0, Mdot = 0.0, Enthalpy = 0.0,
               Qy = 1.0;
        Svector Velocity;
        bool supersonic_flag = false;
    };
    std::vector<CFlow_cell> cells;
    std::vector<CFlow_face> faces;
    std::vector<CFlow_BC_face> BC_faces;
};

namespace {

std::map<std::string, std::vector<double>> arr;
std::string params_hex;

void read_file(const char* path) {
    std::ifstream in(path);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string name;
        size_t cnt = 0;
        ls >> name >> cnt;
        if (name == "params_hex") {
            ls >> params_hex;
            continue;
        }
        std::vector<double>& v = arr[name];
        v.resize(cnt);
        for (size_t i = 0; i < cnt; ++i) ls >> v[i];
    }
}

const std::vector<double>& A(const std::string& k) {
    auto it = arr.find(k);
    if (it == arr.end()) throw std::runtime_error("missing array " + k);
    return it->second;
}

void set3(Svector& s, const std::vector<double>& a, size_t i) {
    for (int d = 0; d < 3; ++d) s.val[d] = a[3 * i + d];
}

CFVM make_fvm() {
    CFVM m;
    const size_t n = A("mesh_cell_volume").size(), F = A("mesh_face_left").size(), B = A("mesh_bc_left").size();
    m.cells.resize(n);
    for (size_t i = 0; i < n; ++i) {
        m.cells[i].index = i;
        m.cells[i].volume = A("mesh_cell_volume")[i];
    }
    m.faces.resize(F);
    for (size_t f = 0; f < F; ++f) {
        auto& fc = m.faces[f];
        fc.left_cell_index = (int)A("mesh_face_left")[f];
        fc.right_cell_index = (int)A("mesh_face_right")[f];
        set3(fc.area_vector, A("mesh_face_area_vector"), f);
        set3(fc.cell_center_unit_vector, A("mesh_face_unit"), f);
        set3(fc.alpha, A("mesh_face_alpha"), f);
        fc.area = A("mesh_face_area")[f];
        fc.cell_center_distance = A("mesh_face_dist")[f];
        fc.diffusion_ksi_A = A("mesh_face_ksi")[f];
        fc.geo_aver = A("mesh_face_geo_aver")[f];
    }
    m.BC_faces.resize(B);
    for (size_t b = 0; b < B; ++b) {
        auto& bf = m.BC_faces[b];
        bf.left_cell_index = (int)A("mesh_bc_left")[b];
        bf.BC_index = (int)A("mesh_bc_marker")[b];
        set3(bf.area_vector, A("mesh_bc_area_vector"), b);
        set3(bf.cell_center_unit_vector, A("mesh_bc_unit"), b);
        set3(bf.alpha, A("mesh_bc_alpha"), b);
        bf.area = A("mesh_bc_area")[b];
        bf.cell_center_distance = A("mesh_bc_dist")[b];
        bf.diffusion_ksi_A = A("mesh_bc_ksi")[b];
        bf.perp_dist = A("mesh_bc_perp")[b];
    }
    return m;
}

// recorded state k into the AoS objects
void load_state(CFlow_solver& fl, int k, size_t n, size_t F, size_t B) {
    const std::string p = "state" + std::to_string(k) + "_";
    auto S = [&](const char* name) -> const std::vector<double>& { return A(p + name); };
    for (size_t i = 0; i < n; ++i) {
        auto& c = fl.cells[i];
        set3(c.Velocity, S("u"), i);
        set3(c.Velocity_last_t, S("u_old"), i);
        c.Density = S("rho")[i];
        c.Density_last_t = S("rho_old")[i];
        c.Pressure = S("p")[i];
        c.Pressure_last_t = S("p_old")[i];
        c.Pressure_corr = S("p_corr")[i];
        set3(c.Pressure_grad, S("grad_p"), i);
        c.Temperature = S("T")[i];
        c.Enthalpy = S("h")[i];
        c.Enthalpy_last_t = S("h_old")[i];
        c.ap = S("ap")[i];
        set3(c.ap_vec, S("ap_vec"), i);
        c.delta_t = S("dt")[i];
    }
    for (size_t f = 0; f < F; ++f) {
        auto& fc = fl.faces[f];
        fc.Mdot = S("f_mdot")[f];
        fc.Density = S("f_rho")[f];
        fc.Qy = S("f_qy")[f];
        set3(fc.stress_explicit_part, S("f_stress"), f);
    }
    for (size_t b = 0; b < B; ++b) {
        auto& bf = fl.BC_faces[b];
        bf.Temperature = S("b_T")[b];
        bf.Pressure = S("b_p")[b];
        bf.Density = S("b_rho")[b];
        bf.Mdot = S("b_mdot")[b];
        bf.Enthalpy = S("b_h")[b];
        bf.Qy = S("b_qy")[b];
        set3(bf.Velocity, S("b_u"), b);
        bf.supersonic_flag = S("b_super")[b] != 0.0;
    }
}

void put(const std::string& name, const std::vector<double>& v) {
    std::printf("%s %zu", name.c_str(), v.size());
    for (double x : v) std::printf(" %.17g", x);
    std::printf("\n");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) return 1;
    try {
        read_file(argv[1]);
        const CFVM fvm = make_fvm();
        const size_t n = fvm.cells.size(), F = fvm.faces.size(), B = fvm.BC_faces.size();
        CFlow_solver fl;
        fl.cells.resize(n);
        fl.faces.resize(F);
        fl.BC_faces.resize(B);
        const int records = (int)A("records")[0];
        if (std::strcmp(argv[2], "roundtrip") == 0) {
            // mesh and every recorded state through the flat arrays and back
            const ocsl::FlowMeshArrays m = ocsl::mesh_arrays(fvm);
            put("rt_face_unit", m.face_unit);
            put("rt_bc_perp", m.bc_perp);
            std::vector<double> marker(m.bc_marker.begin(), m.bc_marker.end());
            put("rt_bc_marker", marker);
            for (int k = 0; k < records; ++k) {
                load_state(fl, k, n, F, B);
                ocsl::FlowStateArrays s;
                ocsl::gather_state(fl, s);
                CFlow_solver back;
                back.cells.resize(n);
                back.faces.resize(F);
                back.BC_faces.resize(B);
                ocsl::scatter_state(s, back);
                ocsl::FlowStateArrays s2;
                ocsl::gather_state(back, s2);
                const std::string p = "rt" + std::to_string(k) + "_";
                put(p + "u", s2.u);
                put(p + "h", s2.h);
                put(p + "ap_vec", s2.ap_vec);
                put(p + "f_stress", s2.f_stress);
                put(p + "b_u", s2.b_u);
                std::vector<double> sup(s2.b_supersonic.begin(), s2.b_supersonic.end());
                put(p + "b_super", sup);
            }
            return 0;
        }
        // the matrix structure of CFVM::define_matrix_structure, in a drop-in CMatrix_solver
        const std::vector<double>& rp = A("row_ptr");
        const std::vector<double>& col = A("col");
        std::shared_ptr<CMatrix_solver> Matrix_solver(new CMatrix_solver(n, n, col.size()));
        Matrix_solver->matrix.row_ptr.assign(rp.size(), 0);
        for (size_t i = 0; i < rp.size(); ++i) Matrix_solver->matrix.row_ptr[i] = (size_t)rp[i];
        Matrix_solver->matrix.column_index.resize(col.size());
        for (size_t i = 0; i < col.size(); ++i)
            Matrix_solver->matrix.column_index[i] = col[i] < 0 ? ~(size_t)0 : (size_t)col[i];
        Matrix_solver->matrix.value.assign(col.size(), 0.0);
        Matrix_solver->M = Matrix_solver->matrix;
        Matrix_solver->extract_diag_indexes();
        ocsl_flow_params prm;
        if (params_hex.size() != 2 * sizeof(prm)) throw std::runtime_error("params size");
        for (size_t i = 0; i < sizeof(prm); ++i)
            reinterpret_cast<unsigned char*>(&prm)[i] = (unsigned char)std::stoi(params_hex.substr(2 * i, 2), nullptr, 16);
        ocsl::FlowAssembly gpu(fvm, Matrix_solver->matrix, prm);
        for (int k = 0; k < records; ++k) {
            load_state(fl, k, n, F, B);
            const int eq = k % 5;
            gpu.assemble(eq, fl, *Matrix_solver);
            const std::string p = "k" + std::to_string(k) + "_";
            put(p + "val", Matrix_solver->matrix.value);
            put(p + "b", Matrix_solver->b_vector.value);
            put(p + "x0", Matrix_solver->x_vector.value);
            std::vector<double> ap(n), apv(3 * n);
            for (size_t i = 0; i < n; ++i) {
                ap[i] = fl.cells[i].ap;
                for (int d = 0; d < 3; ++d) apv[3 * i + d] = fl.cells[i].ap_vec.val[d];
            }
            put(p + "ap", ap);
            put(p + "ap_vec", apv);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 2;
    }
    return 0;
}
