/* Pressure-based flow solver around the GPU assembly kernels: the face/cell
 * coefficient loops of CFlow_solver (cflow_solver_cell_face_loops.cpp with the
 * stencil helpers of solvers_common.h) run as HIP kernels that write the
 * system's CSR values, right-hand side and initial guess on the device, in the
 * caller's slot layout, where the solver handle reads them.  The rest of the
 * SIMPLE-like step (CFlow_solver::iterate, cflow_solver.cpp:229-476) runs as
 * device kernels over a state kept in device memory (OCSL_SUPPORT_DEVICE, the
 * default), or as a host restatement around the assembly kernels
 * (OCSL_SUPPORT_HOST).
 *
 * Arrays of 3-vectors are x, y, z interleaved.  Enumerations follow the
 * reference's (enums.h).  All calls return ocsl_status as in ocsl.h. */
#ifndef OCSL_FLOW_H
#define OCSL_FLOW_H

#include "ocsl.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ocsl_flow_s* ocsl_flow;

/* Eflow_BC_type (enums.h:38) */
typedef enum {
    OCSL_BC_EMPTY = 0,
    OCSL_BC_SYMMETRY = 1,
    OCSL_BC_WALL = 2,
    OCSL_BC_VELOCITY_INLET = 3,
    OCSL_BC_SUPERSONIC_INLET = 4,
    OCSL_BC_PRESSURE_OUTLET = 5
} ocsl_bc_type;

/* Equations, in the order CFlow_solver::iterate solves them. */
typedef enum {
    OCSL_EQ_MOMENTUM_X = 0,
    OCSL_EQ_MOMENTUM_Y = 1,
    OCSL_EQ_MOMENTUM_Z = 2,
    OCSL_EQ_PRESSURE = 3,
    OCSL_EQ_ENTHALPY = 4
} ocsl_equation;

/* Finite-volume mesh (CFVM, cfvm.h:33-93), host arrays, copied at create.
 * Interior faces: left/right cells, area vector, cell-centre unit vector and
 * distance, alpha = S / (S . d), diffusion factor ksi_A = alpha . S, the
 * geometric weight geo_aver.  Boundary faces: owner cell, boundary marker
 * (index into ocsl_flow_params.bc), the same factors and perp_dist. */
typedef struct {
    int64_t n_cells, n_faces, n_bc;
    const double* cell_volume;
    const int32_t* face_left;
    const int32_t* face_right;
    const double* face_area_vector;
    const double* face_unit;
    const double* face_alpha;
    const double* face_area;
    const double* face_dist;
    const double* face_ksi;
    const double* face_geo_aver;
    const int32_t* bc_left;
    const int32_t* bc_marker;
    const double* bc_area_vector;
    const double* bc_unit;
    const double* bc_alpha;
    const double* bc_area;
    const double* bc_dist;
    const double* bc_ksi;
    const double* bc_perp;
} ocsl_mesh;

/* One boundary condition (Cconfiguration::Cboundary_condition). */
typedef struct {
    int type;          /* ocsl_bc_type */
    int no_slip;       /* walls: 1 no-slip, 0 slip */
    int const_temperature;
    double temperature, pressure;
    double velocity[3];
} ocsl_bc;

#define OCSL_FLOW_MAX_BC 16

typedef struct {
    int viscous;       /* Eflow_type_visc: 1 viscous, 0 inviscid */
    int blended;       /* Econv_flux_type: 1 blended central (alpha 0.7), 0 first-order upwind */
    int transient;     /* Etimestepping_type: 1 transient, 0 steady */
    double dt;
    double implicit_underrelax, implicit_underrelax_max;
    int ramp_iterations;
    /* fluid (Cfluid_properties, get_AIR_properties) */
    double gamma, molar_mass, t_ref, h_formation, sutherland[2], prandtl;
    /* initial condition (Cinitial_condition) */
    int shock_tube;
    double p0, t0, u0[3], p2, t2;
    int n_bc;
    ocsl_bc bc[OCSL_FLOW_MAX_BC];
    /* linear solves: method (ocsl_method) per equation group and mode */
    int method_momentum, method_pressure, method_enthalpy;
    int mode;          /* ocsl_mode; REFERENCE is the drop-in */
    int precision;     /* ocsl_precision (TOLERANCE mode only) */
    int support;       /* ocsl_flow_support: where the SIMPLE step's support loops run */
    ocsl_options solve_options;
    int cell_order;    /* ocsl_flow_cell_order: the order the device keeps cells in */
    /* 1 (default): in REFERENCE mode with ILU_BICGSTAB for momentum, the three
     * momentum systems are solved by one ocsl_ILU_BICGSTAB_solve_multi call
     * when they share their matrix and none reads another's solution, that is
     * unless a viscous flow has a no-slip wall
     * (cflow_solver_cell_face_loops.cpp:94-113); the results are bitwise
     * those of the three consecutive solves of CFlow_solver::iterate
     * (cflow_solver.cpp:259-301).  0: always three solves.  Device support
     * loops only. */
    int momentum_together;
} ocsl_flow_params;

/* The order in which the device stores cells (and interior faces, by their
 * left cell) for the kernels that gather over a cell's neighbours.  The
 * caller's numbering, rows, right-hand side and solution order are unchanged
 * whatever the choice, and so are all results, bit for bit: each cell's sums
 * run over its faces in the caller's face order either way.  LOCALITY: the
 * reverse Cuthill-McKee order of the matrix graph.  CURVE: the cells sorted
 * along a 3-D Hilbert curve through their centres (rebuilt from the faces'
 * centre distance vectors), which keeps most face neighbours of a workgroup's
 * cells inside it.  AUTO (default): CURVE when the caller's numbering puts
 * face neighbours more than twice as far apart on average, else the
 * caller's.  CALLER forces the caller's. */
typedef enum {
    OCSL_CELL_ORDER_AUTO = 0,
    OCSL_CELL_ORDER_CALLER = 1,
    OCSL_CELL_ORDER_LOCALITY = 2,
    OCSL_CELL_ORDER_CURVE = 3
} ocsl_flow_cell_order;

/* The device cell order chosen at create: order[k] = the caller's cell stored
 * k-th (n_cells entries). */
ocsl_status ocsl_flow_cell_order_get(ocsl_flow f, int32_t* order);

/* Where the support loops of the SIMPLE step run (boundary updates, face mass
 * fluxes, pressure and pressure-correction gradients, mass-flux correction,
 * the field updates after each solve: cflow_solver.cpp:192-227, 286-476,
 * cflow_solver_supply.cpp:25-214).  DEVICE (default): HIP kernels, the flow
 * state stays resident in device memory between steps and the solutions are
 * read where the solver leaves them.  HOST: a host restatement that uploads
 * each equation's inputs, kept as an independent cross-check. */
typedef enum { OCSL_SUPPORT_DEVICE = 0, OCSL_SUPPORT_HOST = 1 } ocsl_flow_support;

/* Defaults: air (gamma 1.4, M 28.9651159e-3 kg/mol, Tref 298.15 K, Sutherland
 * 1.458e-6 / 110.4, Pr 0.713), transient, blended, inviscid, REFERENCE-mode
 * ILU(0)-BiCGSTAB for every equation, no boundary conditions, support loops on
 * the device. */
void ocsl_flow_default_params(ocsl_flow_params* prm);

/* Flow state (CFlow_solver::CFlow_cell / CFlow_face / CFlow_BC_face).  Any
 * pointer may be NULL in get/set (skipped). */
typedef struct {
    double *u, *u_old, *rho, *rho_old, *p, *p_old, *p_corr, *grad_p, *grad_pc, *T, *h, *h_old, *ap,
        *ap_vec, *dt, *mach;                                  /* cells: n (3n for vectors) */
    double *f_mdot, *f_rho, *f_qy, *f_stress;                 /* faces: F (3F) */
    double *b_T, *b_p, *b_p_out, *b_rho, *b_mdot, *b_h, *b_qy, *b_u; /* boundary faces: B (3B) */
    int* b_supersonic;
} ocsl_flow_state;

/* row_ptr / column_index: the CSR structure of CFVM::define_matrix_structure
 * (cfvm.cpp:74-98), sentinel columns included. */
ocsl_status ocsl_flow_create(ocsl_flow* f, const ocsl_mesh* mesh, const uint64_t* row_ptr,
                             const uint64_t* column_index, const ocsl_flow_params* prm);
ocsl_status ocsl_flow_destroy(ocsl_flow f);

/* The solver handle that holds the flow's systems (owned by the flow). */
ocsl_handle ocsl_flow_solver(ocsl_flow f);

/* CFlow_solver::initialize (cflow_solver.cpp:34-190). */
ocsl_status ocsl_flow_initialize(ocsl_flow f);

ocsl_status ocsl_flow_set_state(ocsl_flow f, const ocsl_flow_state* st);
ocsl_status ocsl_flow_get_state(ocsl_flow f, ocsl_flow_state* st);

/* Assembles one equation's system from the current state on the device
 * (face, boundary-face and cell loops: cflow_solver_cell_face_loops.cpp), into
 * the flow's solver handle.  For momentum, ap += a_ii / 3 and ap_vec[d] = a_ii
 * are written to the state as the reference's cell loop does (ap is not
 * cleared here).  val / b / x0 (optional, `where`) receive copies of the
 * system: values in the caller's slot layout, right-hand side, initial guess. */
ocsl_status ocsl_flow_assemble(ocsl_flow f, int equation, double* val, double* b, double* x0,
                               ocsl_where where);

/* Measurement hook: launches one equation's assembly kernel `reps` times on
 * the solver handle's stream between two HIP events (state as last set);
 * returns the average duration and the algorithmic bytes of one launch (every
 * array the kernel needs, read or written once).  Leaves the state as set. */
ocsl_status ocsl_flow_time_assemble(ocsl_flow f, int equation, int reps, double* avg_ms, double* bytes);

/* Explicit viscous stress of every interior face (calc_flux_explicit_visc_part,
 * cflow_solver_cell_face_loops.cpp:25-35) from the current state. */
ocsl_status ocsl_flow_viscous_stress(ocsl_flow f);

/* One CFlow_solver::iterate; then, when end_of_step, update_timestep
 * (cflow_solver.cpp:497-519). */
ocsl_status ocsl_flow_iterate(ocsl_flow f, int end_of_step);

/* Croot::start_solver's loop (croot.cpp:120-230): steps x inner iterations
 * (one inner iteration when steady). */
ocsl_status ocsl_flow_run(ocsl_flow f, int steps, int inner);

/* Residuals returned by the linear solves so far (reference definition), in
 * solve order; *len receives the count, at most cap are copied. */
ocsl_status ocsl_flow_solve_history(ocsl_flow f, double* out, int64_t cap, int64_t* len);

#ifdef __cplusplus
}
#endif

#endif /* OCSL_FLOW_H */
