"""SU2 meshes: the reader (libocsl_mesh.so, include/ocsl_mesh.h) and a box
generator that writes SU2 text.

read_su2 / parse_su2 return the finite-volume arrays that ``flow.Flow`` takes
(the keys of ``gen.hex_mesh``), in the reference's order: interior faces as
Cgeometry::calc_cell_face_connectivity discovers them (cgeometry.cpp:620-660),
boundary faces marker by marker in file order; values as
CFVM::create_FVM_from_geometry computes them (cfvm.cpp:100-170).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._native import mesh_lib


class MeshError(RuntimeError):
    pass


def _arr(shape, dt):
    return np.empty(shape, dt)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _collect(lib, h) -> dict:
    c = ctypes
    nn, nc, nf, nb = (c.c_int64() for _ in range(4))
    nm = c.c_int32()
    lib.ocsl_mesh_sizes(h, c.byref(nn), c.byref(nc), c.byref(nf), c.byref(nb), c.byref(nm))
    N, F, B = nc.value, nf.value, nb.value
    tags = []
    for i in range(nm.value):
        buf = c.create_string_buffer(4096)
        lib.ocsl_mesh_marker_tag(h, i, buf, 4096)
        tags.append(buf.value.decode())
    vol, ctr = _arr(N, np.float64), _arr((N, 3), np.float64)
    lib.ocsl_mesh_cells(h, _ptr(vol), _ptr(ctr))
    f = {k: _arr(F, np.int32) for k in ("face_left", "face_right")}
    for k in ("face_area_vector", "face_unit", "face_alpha"):
        f[k] = _arr((F, 3), np.float64)
    for k in ("face_dist", "face_ksi", "face_geo_aver", "face_area"):
        f[k] = _arr(F, np.float64)
    lib.ocsl_mesh_faces(h, *(_ptr(f[k]) for k in ("face_left", "face_right", "face_area_vector", "face_unit",
                                                   "face_dist", "face_alpha", "face_ksi", "face_geo_aver",
                                                   "face_area")))
    b = {k: _arr(B, np.int32) for k in ("bc_left", "bc_marker")}
    for k in ("bc_area_vector", "bc_unit", "bc_alpha"):
        b[k] = _arr((B, 3), np.float64)
    for k in ("bc_dist", "bc_ksi", "bc_perp", "bc_area"):
        b[k] = _arr(B, np.float64)
    lib.ocsl_mesh_bc_faces(h, *(_ptr(b[k]) for k in ("bc_left", "bc_marker", "bc_area_vector", "bc_unit",
                                                      "bc_dist", "bc_alpha", "bc_ksi", "bc_perp", "bc_area")))
    out = {"cell_volume": vol, "cell_center": ctr, **f, **b, "n_nodes": nn.value, "tags": tags}
    return out


def _finish(lib, st, h) -> dict:
    if st != 0:
        raise MeshError(f"SU2 mesh: {lib.ocsl_mesh_last_error().decode()} (status {st})")
    try:
        return _collect(lib, h)
    finally:
        lib.ocsl_mesh_destroy(h)


def read_su2(path: str) -> dict:
    """Reads an SU2 file (Read_SU2_mesh, cgeometry_supply.h:201-433) into the
    finite-volume arrays, plus cell_center, n_nodes and the marker tags."""
    lib = mesh_lib()
    h = ctypes.c_void_p()
    st = lib.ocsl_mesh_read_su2(str(path).encode(), ctypes.byref(h))
    return _finish(lib, st, h)


def parse_su2(text: str | bytes) -> dict:
    """The same from SU2 text in memory."""
    lib = mesh_lib()
    data = text.encode() if isinstance(text, str) else bytes(text)
    h = ctypes.c_void_p()
    st = lib.ocsl_mesh_parse_su2(data, len(data), ctypes.byref(h))
    return _finish(lib, st, h)


# ------------------------------------------------------------------ VTK output
VTK_FIELDS = ("density", "pressure", "temperature", "viscosity", "mach", "velocity", "pressure_grad")


class _VtkFields(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in VTK_FIELDS]


def vtk_fields(state: dict, sutherland) -> dict:
    """The reference's output fields from a flow state (``Flow.get_state()``):
    density, pressure, temperature, Sutherland viscosity
    s0 T^1.5 / (T + s1) (Cfluid_properties::get_viscosity,
    cfluid_properties.cpp:123-128), Mach, velocity, pressure gradient."""
    T = np.asarray(state["T"], np.float64)
    mu = sutherland[0] * np.power(T, 1.5)
    mu = mu / (T + sutherland[1])
    return {"density": state["rho"], "pressure": state["p"], "temperature": T, "viscosity": mu,
            "mach": state["mach"], "velocity": state["u"], "pressure_grad": state["grad_p"]}


def write_vtk(su2, path: str, fields: dict | None = None, precision: int = 6) -> None:
    """Writes the mesh of `su2` (SU2 text, or a path to an SU2 file) with cell
    fields as a legacy-VTK file, as Cpost_processing::write_VTK_file does
    (cpost_processing.cpp:26-127, 260-333): fields by the keys of VTK_FIELDS
    (velocity and pressure_grad [n, 3]); precision 6 gives the reference's
    text, 17 round-trips every value."""
    lib = mesh_lib()
    h = ctypes.c_void_p()
    if isinstance(su2, (bytes, str)) and ("\n" in (su2 if isinstance(su2, str) else su2.decode(errors="ignore"))):
        data = su2.encode() if isinstance(su2, str) else bytes(su2)
        st = lib.ocsl_mesh_parse_su2(data, len(data), ctypes.byref(h))
    else:
        st = lib.ocsl_mesh_read_su2(str(su2).encode(), ctypes.byref(h))
    if st != 0:
        raise MeshError(f"SU2 mesh: {lib.ocsl_mesh_last_error().decode()} (status {st})")
    try:
        n = ctypes.c_int64()
        lib.ocsl_mesh_sizes(h, None, ctypes.byref(n), None, None, None)
        keep, f = [], _VtkFields()
        for k, v in (fields or {}).items():
            if k not in VTK_FIELDS:
                raise ValueError(f"unknown VTK field {k!r}")
            a = np.ascontiguousarray(v, dtype=np.float64)
            if a.size != n.value * (3 if k in ("velocity", "pressure_grad") else 1):
                raise ValueError(f"field {k!r} has {a.size} values for {n.value} cells")
            keep.append(a)
            setattr(f, k, a.ctypes.data)
        st = lib.ocsl_mesh_write_vtk(h, str(path).encode(), ctypes.byref(f) if fields else None, int(precision))
        if st != 0:
            raise MeshError(f"VTK: {lib.ocsl_mesh_last_error().decode()} (status {st})")
    finally:
        lib.ocsl_mesh_destroy(h)


class _VtkBcFields(ctypes.Structure):
    _fields_ = [("mach", ctypes.c_void_p), ("wall_stress", ctypes.c_void_p)]


def wall_stress(fvm: dict, state: dict, bcs, viscous: bool, sutherland) -> np.ndarray:
    """Wall shear stress per boundary face [B, 3], as
    CFlow_solver::calc_wall_stress computes it (cflow_solver_supply.cpp:223-239):
    on a no-slip wall of a viscous run, mu(T_L) / perp * (u_L - (u_L . n) n)
    with n = S / |S| and u_L, T_L the face's cell; zero on every other face.
    `fvm` holds bc_left, bc_marker, bc_area_vector, bc_area, bc_perp (the keys
    read_su2 returns); `bcs` one dict per marker with "type" and "no_slip"
    (the flow module's boundary conditions)."""
    from .flow import WALL

    left = np.asarray(fvm["bc_left"], np.int64)
    marker = np.asarray(fvm["bc_marker"], np.int64)
    out = np.zeros((left.size, 3))
    if not viscous:
        return out
    on = np.array([bc.get("type") == WALL and bool(bc.get("no_slip", 0)) for bc in bcs], bool)[marker]
    if not on.any():
        return out
    u = np.asarray(state["u"], np.float64).reshape(-1, 3)[left[on]]
    T = np.asarray(state["T"], np.float64)[left[on]]
    n = (1.0 / np.asarray(fvm["bc_area"], np.float64)[on])[:, None] * np.asarray(fvm["bc_area_vector"])[on]
    un = (u[:, 0] * n[:, 0] + u[:, 1] * n[:, 1]) + u[:, 2] * n[:, 2]
    par = u - un[:, None] * n
    mu = sutherland[0] * np.power(T, 1.5)
    mu = mu / (T + sutherland[1])
    out[on] = (mu / np.asarray(fvm["bc_perp"], np.float64)[on])[:, None] * par
    return out


def write_boundary_vtk(su2, path: str, fields: dict | None = None, precision: int = 6) -> None:
    """Writes the boundary faces of `su2` (SU2 text, or a path to an SU2 file)
    as a legacy-VTK file, as Cpost_processing::write_boundary_VTK_file does
    (cpost_processing.cpp:129-193, 335-401): one VTK cell per boundary face,
    marker by marker in file order, with area, unit normal, marker index, area
    vector and centre; `fields` may add "mach" [B] (the Mach number of each
    face's cell) and "wall_stress" [B, 3] (``wall_stress``), which a flow run
    writes."""
    lib = mesh_lib()
    h = ctypes.c_void_p()
    if isinstance(su2, (bytes, str)) and ("\n" in (su2 if isinstance(su2, str) else su2.decode(errors="ignore"))):
        data = su2.encode() if isinstance(su2, str) else bytes(su2)
        st = lib.ocsl_mesh_parse_su2(data, len(data), ctypes.byref(h))
    else:
        st = lib.ocsl_mesh_read_su2(str(su2).encode(), ctypes.byref(h))
    if st != 0:
        raise MeshError(f"SU2 mesh: {lib.ocsl_mesh_last_error().decode()} (status {st})")
    try:
        nb = ctypes.c_int64()
        lib.ocsl_mesh_sizes(h, None, None, None, ctypes.byref(nb), None)
        keep, f = [], _VtkBcFields()
        for k, v in (fields or {}).items():
            if k not in ("mach", "wall_stress"):
                raise ValueError(f"unknown boundary VTK field {k!r}")
            a = np.ascontiguousarray(v, dtype=np.float64)
            if a.size != nb.value * (3 if k == "wall_stress" else 1):
                raise ValueError(f"field {k!r} has {a.size} values for {nb.value} boundary faces")
            keep.append(a)
            setattr(f, k, a.ctypes.data)
        st = lib.ocsl_mesh_write_boundary_vtk(h, str(path).encode(), ctypes.byref(f) if fields else None,
                                              int(precision))
        if st != 0:
            raise MeshError(f"VTK: {lib.ocsl_mesh_last_error().decode()} (status {st})")
    finally:
        lib.ocsl_mesh_destroy(h)


# ------------------------------------------------------------------ writer
SU2_TET, SU2_HEX, SU2_PRISM, SU2_TRI, SU2_QUAD = 10, 12, 13, 5, 9


def write_su2_text(nodes, elements, markers) -> str:
    """SU2 text: nodes [n, 3]; elements a list of (type code, node list);
    markers a list of (tag, [(type code, node list), ...])."""
    out = ["NDIME= 3", f"NELEM= {len(elements)}"]
    out += [f"{t} " + " ".join(str(int(v)) for v in ns) + f" {i}" for i, (t, ns) in enumerate(elements)]
    out.append(f"NPOIN= {len(nodes)}")
    out += [f"{x!r} {y!r} {z!r} {i}" for i, (x, y, z) in enumerate(np.asarray(nodes, np.float64).tolist())]
    out.append(f"NMARK= {len(markers)}")
    for tag, elems in markers:
        out += [f"MARKER_TAG= {tag}", f"MARKER_ELEMS= {len(elems)}"]
        out += [f"{t} " + " ".join(str(int(v)) for v in ns) for t, ns in elems]
    return "\n".join(out) + "\n"


# Kuhn subdivision of a cube into six tetrahedra along the diagonal 000-111:
# one per order of the three axes; corners as (x, y, z) bits.
_KUHN = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def box_su2(n, kind: str = "hex", h: float = 1.0, seed: int | None = None) -> str:
    """An n^3 box of cubes of side h as SU2 text, each cube one hexahedron,
    two prisms (split along its 100-010 diagonal) or six tetrahedra (Kuhn
    subdivision); positively oriented elements; six markers xmin, xmax, ymin,
    ymax, zmin, zmax.  With a seed, nodes and elements are numbered in a
    random order (the connectivity order depends on the numbering)."""
    n = int(n)
    m = n + 1
    rng = np.random.default_rng(seed) if seed is not None else None
    nid = np.arange(m ** 3) if rng is None else rng.permutation(m ** 3)

    def node(i, j, k):
        return int(nid[i + m * (j + m * k)])

    def corner(i, j, k, b):
        return node(i + (b & 1), j + ((b >> 1) & 1), k + ((b >> 2) & 1))

    coords = np.empty((m ** 3, 3))
    g = np.arange(m) * h
    kk, jj, ii = np.meshgrid(g, g, g, indexing="ij")
    coords[nid] = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1)

    def vol6(ns):
        p = coords[list(ns)]
        return float(np.dot(p[1] - p[0], np.cross(p[2] - p[0], p[3] - p[0])))

    elems = []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                c = [corner(i, j, k, b) for b in range(8)]
                if kind == "hex":
                    elems.append((SU2_HEX, [c[0], c[1], c[3], c[2], c[4], c[5], c[7], c[6]]))
                elif kind == "prism":
                    elems.append((SU2_PRISM, [c[0], c[1], c[2], c[4], c[5], c[6]]))
                    elems.append((SU2_PRISM, [c[1], c[3], c[2], c[5], c[7], c[6]]))
                elif kind == "tet":
                    for order in _KUHN:
                        b, path = 0, [c[0]]
                        for ax in order:
                            b |= 1 << ax
                            path.append(c[b])
                        if vol6(path) < 0:
                            path[2], path[3] = path[3], path[2]
                        elems.append((SU2_TET, path))
                else:
                    raise ValueError("kind must be hex, prism or tet")
    markers = []
    for axis, name in enumerate("xyz"):
        for side, tag in ((0, name + "min"), (1, name + "max")):
            faces = []
            a, b = [d for d in range(3) if d != axis]
            for u in range(n):
                for v in range(n):
                    def pt(du, dv):
                        idx = [0, 0, 0]
                        idx[axis] = side * n
                        idx[a] = u + du
                        idx[b] = v + dv
                        return node(*idx)
                    q = [pt(0, 0), pt(1, 0), pt(1, 1), pt(0, 1)]
                    split = kind == "tet" or (kind == "prism" and axis == 2)
                    if not split:
                        faces.append((SU2_QUAD, q))
                    elif kind == "prism":
                        # the prisms' 100-010 diagonal on the z sides
                        p00, p10, p11, p01 = q
                        faces.append((SU2_TRI, [p00, p10, p01]))
                        faces.append((SU2_TRI, [p10, p11, p01]))
                    else:
                        # Kuhn faces split along the diagonal from the low corner
                        p00, p10, p11, p01 = q
                        faces.append((SU2_TRI, [p00, p10, p11]))
                        faces.append((SU2_TRI, [p00, p11, p01]))
            markers.append((tag, faces))
    if rng is not None:
        order = rng.permutation(len(elems))
        elems = [elems[q] for q in order]
    return write_su2_text(coords, elems, markers)
