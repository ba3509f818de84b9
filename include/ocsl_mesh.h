/* ocsl_mesh.h -- SU2 mesh reader and finite-volume geometry (host C++,
 * libocsl_mesh.so).
 *
 * Replaces the reference's mesh set-up path:
 *   Read_SU2_mesh                       cgeometry_supply.h:201-433
 *   Cgeometry::calc_cell_face_connectivity, add_BC_face_tri / _quad,
 *   calc_face_area_and_normal, calc_cell_volume_and_center
 *                                       cgeometry.cpp:239-680
 *   CFVM::create_FVM_from_geometry      cfvm.cpp:100-170
 * and hands over the arrays ocsl_flow_create takes (ocsl_flow.h), in the
 * reference's face order: interior faces in the order the connectivity pass
 * discovers them, boundary faces marker by marker in file order.
 *
 * Supported: 3-D SU2 files (NDIME= 3) with tetrahedra (10), prisms (13) and
 * hexahedra (12); boundary triangles (5) and quadrilaterals (9).  Element
 * lines carry a trailing index; point lines carry x y z and an index.
 */
#ifndef OCSL_MESH_H
#define OCSL_MESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ocsl_mesh_s* ocsl_mesh;

/* Reads and processes an SU2 file.  Returns 0, or a negative status (-1 bad
 * argument, -2 unreadable file, -3 malformed or unsupported content, -4 out of
 * memory); ocsl_mesh_last_error() says what.
 * Reference: Read_SU2_mesh (cgeometry_supply.h:201-433) followed by
 * CFVM::create_FVM_from_geometry (cfvm.cpp:100-170). */
int ocsl_mesh_read_su2(const char* path, ocsl_mesh* out);

/* The same from SU2 text held in memory (len bytes). */
int ocsl_mesh_parse_su2(const char* text, int64_t len, ocsl_mesh* out);

void ocsl_mesh_destroy(ocsl_mesh m);

/* Thread-local message of the last failed call. */
const char* ocsl_mesh_last_error(void);

/* Counts: nodes, cells, interior faces, boundary faces, boundary markers. */
int ocsl_mesh_sizes(ocsl_mesh m, int64_t* n_nodes, int64_t* n_cells, int64_t* n_faces, int64_t* n_bc,
                    int32_t* n_markers);

/* Marker i's tag (MARKER_TAG, first blank removed as the reference does),
 * NUL-terminated into buf of cap bytes; returns its length or a negative
 * status. */
int ocsl_mesh_marker_tag(ocsl_mesh m, int32_t i, char* buf, int64_t cap);

/* Per cell: volume and centroid [n][3] (Cgeo_cell::volume / center).  Either
 * pointer may be NULL. */
int ocsl_mesh_cells(ocsl_mesh m, double* volume, double* center);

/* Per interior face (CFVM_face, cfvm.cpp:114-139): left and right cell, area
 * vector [F][3], unit vector between the cell centres [F][3], that distance,
 * alpha = S / (S . d) [F][3], diffusion_ksi_A, geo_aver, area.  NULL skips an
 * array. */
int ocsl_mesh_faces(ocsl_mesh m, int32_t* left, int32_t* right, double* area_vector, double* unit, double* dist,
                    double* alpha, double* ksi, double* geo_aver, double* area);

/* Per boundary face (CFVM_BC_face, cfvm.cpp:143-167): owner cell, marker
 * index, area vector, unit vector from the cell centre to the face centre,
 * that distance, alpha, diffusion_ksi_A, perp_dist, area. */
int ocsl_mesh_bc_faces(ocsl_mesh m, int32_t* left, int32_t* marker, double* area_vector, double* unit,
                       double* dist, double* alpha, double* ksi, double* perp, double* area);

/* Cell fields for ocsl_mesh_write_vtk, one value (vectors: three) per cell;
 * a NULL array is left out of the file. */
typedef struct ocsl_vtk_fields {
    const double* density;
    const double* pressure;
    const double* temperature;
    const double* viscosity;
    const double* mach;
    const double* velocity;      /* [n][3] */
    const double* pressure_grad; /* [n][3] */
} ocsl_vtk_fields;

/* Legacy-VTK unstructured grid of the mesh: nodes, cells (VTK node order and
 * types 10 / 13 / 12), then cell data Volume[m3], Center and the given fields
 * with the reference's names and order.  Numbers are written %.{precision}g:
 * 6 reproduces the reference's stream output, 17 round-trips every double.
 * Reference: Cpost_processing::write_VTK_file / write_VTK_data
 * (cpost_processing.cpp:26-127, 260-333). */
int ocsl_mesh_write_vtk(ocsl_mesh m, const char* path, const ocsl_vtk_fields* fields, int precision);

/* Boundary-face fields for ocsl_mesh_write_boundary_vtk, one value (vectors:
 * three) per boundary face in ocsl_mesh_bc_faces order; NULL leaves a field
 * out. */
typedef struct ocsl_vtk_bc_fields {
    const double* mach;        /* the Mach number of the face's cell */
    const double* wall_stress; /* [B][3]: CFlow_solver::calc_wall_stress
                                  (cflow_solver_supply.cpp:223-239) */
} ocsl_vtk_bc_fields;

/* Legacy-VTK unstructured grid of the mesh's boundary faces: all nodes, the
 * boundary faces as cells (types 5 / 9), then per face Area[m2], normal,
 * Marker, area_vector, center and the given fields, with the reference's
 * names, keywords and order; numbers %.{precision}g (6: the reference's text).
 * Reference: Cpost_processing::write_boundary_VTK_file /
 * write_boundary_VTK_data (cpost_processing.cpp:335-401, 129-193). */
int ocsl_mesh_write_boundary_vtk(ocsl_mesh m, const char* path, const ocsl_vtk_bc_fields* fields, int precision);

#ifdef __cplusplus
}
#endif

#endif /* OCSL_MESH_H */
