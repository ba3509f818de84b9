// ocsl_su2.cpp -- SU2 mesh reader, cell/face connectivity, geometry and the
// finite-volume face factors, on the host (libocsl_mesh.so, include/ocsl_mesh.h).
//
// The output is the reference's, in the reference's order:
//   * parsing: Read_SU2_mesh (cgeometry_supply.h:201-433) -- NELEM, NPOIN and
//     NMARK are found by scanning lines for "KEY=", the entries after them are
//     read as whitespace-separated numbers;
//   * interior faces in the order Cgeometry::calc_cell_face_connectivity
//     (cgeometry.cpp:620-660) discovers them: nodes ascending, ordered pairs of
//     cells sharing the node, the left cell's faces in local order;
//   * boundary faces marker by marker in file order (add_BC_face_tri/_quad,
//     cgeometry.cpp:239-336), placed on the one cell holding all their nodes;
//   * metrics with the reference's operation order (cgeometry.cpp:338-555,
//     tensor.h's vector helpers) and the CFVM factors (cfvm.cpp:100-170).
// The connectivity keeps per-cell face tables and per-node cell lists in flat
// arrays; nothing is searched through std::vector copies.
#include "ocsl_mesh.h"

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct MeshError {
    int status;
    std::string msg;
};

[[noreturn]] void fail(int st, const std::string& msg) { throw MeshError{st, msg}; }

// ---------------------------------------------------------------- vectors
// Component order and summation order of tensor.h (dot sums from 0.0, mag is
// sqrt of that dot, the cross product written out per component).
struct V3 {
    double v[3] = {0.0, 0.0, 0.0};
};
inline V3 operator+(const V3& a, const V3& b) { return {{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]}}; }
inline V3 operator-(const V3& a, const V3& b) { return {{a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]}}; }
inline V3 operator*(double s, const V3& a) { return {{s * a.v[0], s * a.v[1], s * a.v[2]}}; }
inline V3 cross(const V3& a, const V3& b) {
    return {{a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2], a.v[0] * b.v[1] - a.v[1] * b.v[0]}};
}
inline double dot(const V3& a, const V3& b) {
    double r = 0.0;
    for (int k = 0; k < 3; ++k) r += a.v[k] * b.v[k];
    return r;
}
inline double mag(const V3& a) { return std::sqrt(dot(a, a)); }

// ---------------------------------------------------------------- cell types
// Local faces of each cell type as local node indices (the face node order
// of Cgeometry::return_face, cgeometry.cpp:80-228); -1 pads triangles.
enum CellType : int8_t { kTet = 0, kPrism = 1, kHex = 2 };
constexpr int kNodes[3] = {4, 6, 8};
constexpr int kFaces[3] = {4, 5, 6};
constexpr int kFaceNodes[3][6][4] = {
    {{1, 0, 2, -1}, {0, 1, 3, -1}, {1, 2, 3, -1}, {2, 0, 3, -1}, {}, {}},
    {{0, 1, 4, 3}, {1, 2, 5, 4}, {2, 0, 3, 5}, {0, 2, 1, -1}, {3, 4, 5, -1}, {}},
    {{0, 1, 5, 4}, {1, 3, 7, 5}, {3, 2, 6, 7}, {2, 0, 4, 6}, {1, 0, 2, 3}, {4, 5, 7, 6}},
};
inline int face_size(int type, int lf) { return kFaceNodes[type][lf][3] < 0 ? 3 : 4; }

// For each cell type and local node, the local faces through it, ascending.
struct NodeFaces {
    int n[3][8];
    int f[3][8][6];
    NodeFaces() {
        for (int t = 0; t < 3; ++t)
            for (int k = 0; k < 8; ++k) {
                n[t][k] = 0;
                for (int lf = 0; lf < kFaces[t]; ++lf)
                    for (int q = 0; q < face_size(t, lf); ++q)
                        if (kFaceNodes[t][lf][q] == k) f[t][k][n[t][k]++] = lf;
            }
    }
};
const NodeFaces kNodeFaces;

struct Cell {
    CellType type;
    int node[8];
    int face[6];
    int nb[6];  // the cell across each local face, -1 until found
};

struct Face {
    int node[4];
    int nn;             // 3 or 4
    int left, right;    // right = -1 on boundary faces
    int marker;         // -1 on interior faces
    V3 normal, center;  // calc_face_area_and_normal
    double area;
};

// ---------------------------------------------------------------- parsing
// A cursor over the file text that reads the way the reference's stream
// does: key lines with getline + split at '=', numbers with operator>>.
struct Text {
    const char* p;
    const char* end;

    // Scans lines from the cursor for one whose text before '=' (the whole
    // line when it has none) is `key`; returns the text after '='.
    bool seek(const char* key, std::string& value) {
        const size_t klen = std::strlen(key);
        while (p < end) {
            const char* eol = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
            const char* le = eol ? eol : end;
            const char* eq = static_cast<const char*>(std::memchr(p, '=', (size_t)(le - p)));
            const char* ke = eq ? eq : le;
            const bool hit = (size_t)(ke - p) == klen && std::memcmp(p, key, klen) == 0;
            if (hit) value.assign(eq ? eq + 1 : le, le);
            p = eol ? eol + 1 : end;
            if (hit) return true;
        }
        return false;
    }
    // the token copied into a scratch buffer ended there, not at its size (a
    // NUL byte is not a separator: strchr would match the terminator)
    bool at_end(size_t n) const { return p + n >= end || (p[n] != '\0' && std::strchr(" \t\n\r\v\f", p[n])); }
    void skip_ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
    }
    long long integer(const char* what) {
        skip_ws();
        char buf[64];
        size_t n = 0;
        while (p + n < end && n < sizeof(buf) - 1 && !std::strchr(" \t\n\r\v\f", p[n])) buf[n] = p[n], ++n;
        buf[n] = 0;
        char* q = nullptr;
        errno = 0;
        const long long v = std::strtoll(buf, &q, 10);
        if (n == 0 || q != buf + n || !at_end(n) || errno == ERANGE)
            fail(-3, std::string("expected an integer (") + what + ")");
        p += n;
        return v;
    }
    // a node number: 0 .. INT_MAX, range-checked before it is narrowed
    int index(const char* what) {
        const long long v = integer(what);
        if (v < 0 || v > INT_MAX) fail(-3, std::string(what) + " out of range");
        return (int)v;
    }
    double real(const char* what) {
        skip_ws();
        char buf[128];
        size_t n = 0;
        while (p + n < end && n < sizeof(buf) - 1 && !std::strchr(" \t\n\r\v\f", p[n])) buf[n] = p[n], ++n;
        buf[n] = 0;
        char* q = nullptr;
        errno = 0;
        const double v = std::strtod(buf, &q);
        if (n == 0 || q != buf + n || !at_end(n)) fail(-3, std::string("expected a number (") + what + ")");
        // overflow and inf / nan spellings; an underflow to a subnormal or zero is a value
        if (!std::isfinite(v) || (errno == ERANGE && std::fabs(v) > 1.0))
            fail(-3, std::string("non-finite number (") + what + ")");
        p += n;
        return v;
    }
};

// The count after "KEY=" (the reference converts it with atof).
long long count_of(const std::string& v, const char* key) {
    char* q = nullptr;
    const double d = std::strtod(v.c_str(), &q);
    if (q == v.c_str() || !(d >= 0.0) || d > 2.0e9) fail(-3, std::string("bad count after ") + key + "=");
    return (long long)d;
}

}  // namespace

struct ocsl_mesh_s {
    std::vector<V3> nodes;
    std::vector<Cell> cells;
    std::vector<Face> faces;  // geometry faces: interior (discovery order), then boundary
    std::vector<std::string> tags;
    std::vector<double> vol;
    std::vector<V3> center;
    std::vector<int> interior, boundary;  // indexes into faces, in order
};

namespace {

// ---------------------------------------------------------------- connectivity
struct NodeCells {
    std::vector<int> ptr, cell;  // cells holding each node, ascending (one entry per occurrence)
};

NodeCells node_cells(const ocsl_mesh_s& m) {
    NodeCells nc;
    const size_t N = m.nodes.size();
    nc.ptr.assign(N + 1, 0);
    for (const Cell& c : m.cells)
        for (int k = 0; k < kNodes[c.type]; ++k) nc.ptr[(size_t)c.node[k] + 1]++;
    for (size_t i = 0; i < N; ++i) nc.ptr[i + 1] += nc.ptr[i];
    nc.cell.resize((size_t)nc.ptr[N]);
    std::vector<int> cur(nc.ptr.begin(), nc.ptr.end() - 1);
    for (size_t ci = 0; ci < m.cells.size(); ++ci) {
        const Cell& c = m.cells[ci];
        for (int k = 0; k < kNodes[c.type]; ++k) nc.cell[(size_t)cur[c.node[k]]++] = (int)ci;
    }
    return nc;
}

inline void face_nodes(const Cell& c, int lf, int (&g)[4], int& nn) {
    nn = face_size(c.type, lf);
    for (int k = 0; k < nn; ++k) g[k] = c.node[kFaceNodes[c.type][lf][k]];
}

inline bool has(const int* a, int n, int x) {
    for (int k = 0; k < n; ++k)
        if (a[k] == x) return true;
    return false;
}

// Interior faces in the reference's discovery order: for each node, for each
// ordered pair (l, r) of distinct positions in its cell list, each face of l
// through the node against each face of r through the node; a face whose
// nodes all appear in r's face (same kind) becomes a new face with l on its
// left, unless l's side of it is already taken (cgeometry.cpp:575-660).
void connect(ocsl_mesh_s& m, const NodeCells& nc) {
    m.faces.reserve(m.cells.size() * 3);
    auto local = [](const Cell& c, int node) {
        for (int k = 0; k < kNodes[c.type]; ++k)
            if (c.node[k] == node) return k;
        return -1;
    };
    const size_t N = m.nodes.size();
    for (size_t node = 0; node < N; ++node) {
        const int b = nc.ptr[node], e = nc.ptr[node + 1];
        for (int pl = b; pl < e; ++pl)
            for (int pr = b; pr < e; ++pr) {
                if (pl == pr) continue;
                const int il = nc.cell[(size_t)pl], ir = nc.cell[(size_t)pr];
                Cell& L = m.cells[(size_t)il];
                Cell& R = m.cells[(size_t)ir];
                // only faces through the node can match; visited in ascending
                // local order, as the full scan would meet them.  A face of L
                // whose side is taken creates nothing, so it is skipped.
                const int kl = local(L, (int)node), kr = local(R, (int)node);
                for (int a = 0; a < kNodeFaces.n[L.type][kl]; ++a) {
                    const int fl = kNodeFaces.f[L.type][kl][a];
                    if (L.nb[fl] != -1) continue;
                    int gl[4], nl;
                    face_nodes(L, fl, gl, nl);
                    for (int c = 0; c < kNodeFaces.n[R.type][kr]; ++c) {
                        const int fr = kNodeFaces.f[R.type][kr][c];
                        int gr[4], nr;
                        face_nodes(R, fr, gr, nr);
                        if (nl != nr) continue;
                        bool same = true;
                        for (int k = 0; k < nl && same; ++k) same = has(gr, nr, gl[k]);
                        if (!same) continue;
                        Face f{};
                        for (int k = 0; k < 4; ++k) f.node[k] = k < nl ? gl[k] : -1;
                        f.nn = nl;
                        f.left = il;
                        f.right = ir;
                        f.marker = -1;
                        const int id = (int)m.faces.size();
                        m.faces.push_back(f);
                        L.face[fl] = id;
                        R.face[fr] = id;
                        L.nb[fl] = ir;
                        R.nb[fr] = il;
                        break;  // L's side of fl is taken now
                    }
                }
            }
    }
}

// A boundary face: the first cell holding all its nodes (the candidates of
// its first node, kept in order while the others filter them), and that
// cell's first local face containing all of them (cgeometry.cpp:239-336).
void add_bc_face(ocsl_mesh_s& m, const NodeCells& nc, const int* nodes, int nn, int marker) {
    const size_t N = m.nodes.size();
    for (int k = 0; k < nn; ++k)
        if (nodes[k] < 0 || (size_t)nodes[k] >= N) fail(-3, "boundary face node out of range");
    int cell = -1;
    for (int q = nc.ptr[(size_t)nodes[0]]; q < nc.ptr[(size_t)nodes[0] + 1] && cell < 0; ++q) {
        const int c = nc.cell[(size_t)q];
        bool all = true;
        for (int k = 1; k < nn && all; ++k) {
            bool in = false;
            for (int s = nc.ptr[(size_t)nodes[k]]; s < nc.ptr[(size_t)nodes[k] + 1] && !in; ++s)
                in = nc.cell[(size_t)s] == c;
            all = in;
        }
        if (all) cell = c;
    }
    if (cell < 0) fail(-3, "boundary face whose nodes share no cell");
    Cell& C = m.cells[(size_t)cell];
    for (int lf = 0; lf < kFaces[C.type]; ++lf) {
        int g[4], n;
        face_nodes(C, lf, g, n);
        bool all = true;
        for (int k = 0; k < nn && all; ++k) all = has(g, n, nodes[k]);
        if (!all) continue;
        Face f{};
        for (int k = 0; k < 4; ++k) f.node[k] = k < n ? g[k] : -1;
        f.nn = n;
        f.left = cell;
        f.right = -1;
        f.marker = marker;
        C.face[lf] = (int)m.faces.size();
        m.faces.push_back(f);
        return;
    }
    fail(-3, "boundary face is not a face of the cell holding its nodes");
}

// ---------------------------------------------------------------- metrics
// calc_face_area_and_normal (cgeometry.cpp:338-381): the normal of the
// first triangle (nodes 0, 1, 2); a quadrilateral adds the triangle (0, 2, 3)
// to the area and area-weights the two triangle centroids.
void face_metrics(const ocsl_mesh_s& m, Face& f) {
    const V3 &n0 = m.nodes[(size_t)f.node[0]], &n1 = m.nodes[(size_t)f.node[1]], &n2 = m.nodes[(size_t)f.node[2]];
    V3 nrm = cross(n1 - n0, n2 - n0);
    double len = mag(nrm);
    f.normal = (1.0 / len) * nrm;
    f.area = 0.5 * len;
    f.center = (1.0 / 3.0) * (n0 + n1 + n2);
    if (f.nn == 4) {
        const V3& n3 = m.nodes[(size_t)f.node[3]];
        f.center = f.area * f.center;
        len = mag(cross(n2 - n0, n3 - n0));
        f.area = f.area + (0.5 * len);
        f.center = f.center + ((0.5 * len / 3.0) * (n0 + n2 + n3));
        f.center = (1.0 / f.area) * f.center;
    }
}

inline double tet_volume(const V3& a, const V3& b, const V3& c, const V3& d) {
    return (1.0 / 6.0) * dot(b - a, cross(c - a, d - a));
}

// calc_cell_volume_and_center (cgeometry.cpp:383-555): a tetrahedron's own
// volume and vertex mean; prisms (3) and hexahedra (6) as sums of
// tetrahedra, the centroid their volume-weighted vertex means.
constexpr int kTets[3][6][4] = {
    {{0, 1, 2, 3}},
    {{0, 4, 5, 3}, {0, 4, 1, 5}, {0, 1, 2, 5}},
    {{0, 2, 7, 3}, {0, 2, 6, 7}, {0, 7, 1, 3}, {0, 7, 5, 1}, {0, 7, 4, 5}, {0, 7, 6, 4}},
};
constexpr int kNTets[3] = {1, 3, 6};

void cell_metrics(const ocsl_mesh_s& m, const Cell& c, double& vol, V3& ctr) {
    auto P = [&](int t, int k) -> const V3& { return m.nodes[(size_t)c.node[kTets[c.type][t][k]]]; };
    vol = tet_volume(P(0, 0), P(0, 1), P(0, 2), P(0, 3));
    ctr = P(0, 0) + P(0, 1) + P(0, 2) + P(0, 3);
    if (c.type == kTet) {
        ctr = (1.0 / 4.0) * ctr;
        return;
    }
    ctr = (vol / 4.0) * ctr;
    for (int t = 1; t < kNTets[c.type]; ++t) {
        const double dv = tet_volume(P(t, 0), P(t, 1), P(t, 2), P(t, 3));
        vol += dv;
        const V3 cd = (dv / 4.0) * (P(t, 0) + P(t, 1) + P(t, 2) + P(t, 3));
        ctr = ctr + cd;
    }
    ctr = (1.0 / vol) * ctr;
}

void parse(ocsl_mesh_s& m, const char* text, size_t len) {
    Text t{text, text + len};
    std::string v;
    if (!t.seek("NELEM", v)) fail(-3, "no NELEM= line");
    const long long n_cell = count_of(v, "NELEM");
    m.cells.resize((size_t)n_cell);
    std::vector<int> raw;
    for (long long ci = 0; ci < n_cell; ++ci) {
        const long long code = t.integer("element type");
        Cell& c = m.cells[(size_t)ci];
        for (int k = 0; k < 6; ++k) c.face[k] = c.nb[k] = -1;
        for (int k = 0; k < 8; ++k) c.node[k] = -1;
        int nn;
        if (code == 10) c.type = kTet, nn = 4;
        else if (code == 13) c.type = kPrism, nn = 6;
        else if (code == 12) c.type = kHex, nn = 8;
        else fail(-3, "unsupported element type " + std::to_string(code) + " (tetrahedra, prisms, hexahedra only)");
        int in[8];
        for (int k = 0; k < nn; ++k) in[k] = t.index("element node");
        if (c.type == kHex) {
            // SU2 lists a hexahedron's bottom and top faces around; the cell
            // tables take nodes 2/3 and 6/7 swapped (cgeometry_supply.h:266-276)
            const int map[8] = {0, 1, 3, 2, 4, 5, 7, 6};
            for (int k = 0; k < 8; ++k) c.node[map[k]] = in[k];
        } else {
            for (int k = 0; k < nn; ++k) c.node[k] = in[k];
        }
        (void)t.integer("element index");
    }
    if (!t.seek("NPOIN", v)) fail(-3, "no NPOIN= line");
    const long long n_node = count_of(v, "NPOIN");
    m.nodes.resize((size_t)n_node);
    for (long long i = 0; i < n_node; ++i) {
        for (int k = 0; k < 3; ++k) m.nodes[(size_t)i].v[k] = t.real("point coordinate");
        (void)t.integer("point index");
    }
    for (const Cell& c : m.cells)
        for (int k = 0; k < kNodes[c.type]; ++k)
            if (c.node[k] < 0 || c.node[k] >= n_node) fail(-3, "element node out of range");
    const NodeCells nc = node_cells(m);
    connect(m, nc);
    if (!t.seek("NMARK", v)) fail(-3, "no NMARK= line");
    const long long n_mark = count_of(v, "NMARK");
    m.tags.resize((size_t)n_mark);
    for (long long mk = 0; mk < n_mark; ++mk) {
        if (!t.seek("MARKER_TAG", v)) fail(-3, "missing MARKER_TAG= line");
        const size_t sp = v.find(' ');
        if (sp != std::string::npos) v.erase(sp, 1);
        m.tags[(size_t)mk] = v;
        if (!t.seek("MARKER_ELEMS", v)) fail(-3, "missing MARKER_ELEMS= line");
        const long long ne = count_of(v, "MARKER_ELEMS");
        for (long long q = 0; q < ne; ++q) {
            const long long code = t.integer("boundary element type");
            int nodes[4];
            if (code == 5) {
                for (int k = 0; k < 3; ++k) nodes[k] = t.index("boundary node");
                add_bc_face(m, nc, nodes, 3, (int)mk);
            } else if (code == 9) {
                // quadrilateral read as 0, 1, 3, 2 (cgeometry_supply.h:416-421)
                const int map[4] = {0, 1, 3, 2};
                for (int k = 0; k < 4; ++k) nodes[map[k]] = t.index("boundary node");
                add_bc_face(m, nc, nodes, 4, (int)mk);
            } else {
                fail(-3, "unsupported boundary element type " + std::to_string(code));
            }
        }
    }
    for (Face& f : m.faces) face_metrics(m, f);
    m.vol.resize(m.cells.size());
    m.center.resize(m.cells.size());
    for (size_t c = 0; c < m.cells.size(); ++c) cell_metrics(m, m.cells[c], m.vol[c], m.center[c]);
    for (size_t f = 0; f < m.faces.size(); ++f) (m.faces[f].marker < 0 ? m.interior : m.boundary).push_back((int)f);
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const MeshError& e) {
        g_err = e.msg;
        return e.status;
    } catch (const std::bad_alloc&) {
        g_err = "out of memory";
        return -4;
    }
}

inline void put3(double* dst, size_t i, const V3& a) {
    for (int k = 0; k < 3; ++k) dst[3 * i + k] = a.v[k];
}

}  // namespace

extern "C" {

int ocsl_mesh_parse_su2(const char* text, int64_t len, ocsl_mesh* out) {
    if (!out || (!text && len) || len < 0) {
        g_err = "bad argument";
        return -1;
    }
    *out = nullptr;
    ocsl_mesh_s* m = nullptr;
    const int st = guarded([&] {
        m = new ocsl_mesh_s;
        parse(*m, text, (size_t)len);
    });
    if (st) {
        delete m;
        return st;
    }
    *out = m;
    return 0;
}

int ocsl_mesh_read_su2(const char* path, ocsl_mesh* out) {
    if (!path || !out) {
        g_err = "bad argument";
        return -1;
    }
    *out = nullptr;
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        g_err = std::string("cannot open ") + path;
        return -2;
    }
    std::string text;
    const int st = guarded([&] {
        std::ostringstream ss;
        ss << in.rdbuf();
        text = ss.str();
    });
    if (st) return st;
    return ocsl_mesh_parse_su2(text.data(), (int64_t)text.size(), out);
}

void ocsl_mesh_destroy(ocsl_mesh m) { delete m; }

const char* ocsl_mesh_last_error(void) { return g_err.c_str(); }

int ocsl_mesh_sizes(ocsl_mesh m, int64_t* n_nodes, int64_t* n_cells, int64_t* n_faces, int64_t* n_bc,
                    int32_t* n_markers) {
    if (!m) return -1;
    if (n_nodes) *n_nodes = (int64_t)m->nodes.size();
    if (n_cells) *n_cells = (int64_t)m->cells.size();
    if (n_faces) *n_faces = (int64_t)m->interior.size();
    if (n_bc) *n_bc = (int64_t)m->boundary.size();
    if (n_markers) *n_markers = (int32_t)m->tags.size();
    return 0;
}

int ocsl_mesh_marker_tag(ocsl_mesh m, int32_t i, char* buf, int64_t cap) {
    if (!m || i < 0 || (size_t)i >= m->tags.size() || !buf || cap <= 0) return -1;
    const std::string& s = m->tags[(size_t)i];
    const size_t n = std::min((size_t)cap - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
    return (int)s.size();
}

int ocsl_mesh_cells(ocsl_mesh m, double* volume, double* center) {
    if (!m) return -1;
    for (size_t c = 0; c < m->cells.size(); ++c) {
        if (volume) volume[c] = m->vol[c];
        if (center) put3(center, c, m->center[c]);
    }
    return 0;
}

// CFVM::create_FVM_from_geometry, interior faces (cfvm.cpp:114-139).
int ocsl_mesh_faces(ocsl_mesh m, int32_t* left, int32_t* right, double* area_vector, double* unit, double* dist,
                    double* alpha, double* ksi, double* geo_aver, double* area) {
    if (!m) return -1;
    for (size_t q = 0; q < m->interior.size(); ++q) {
        const Face& f = m->faces[(size_t)m->interior[q]];
        const V3 &cl = m->center[(size_t)f.left], &cr = m->center[(size_t)f.right];
        const V3 S = f.area * f.normal;
        V3 d = cr - cl;
        const V3 al = (1.0 / dot(S, d)) * S;
        const double len = mag(d);
        d = (1.0 / len) * d;
        if (left) left[q] = f.left;
        if (right) right[q] = f.right;
        if (area_vector) put3(area_vector, q, S);
        if (unit) put3(unit, q, d);
        if (dist) dist[q] = len;
        if (alpha) put3(alpha, q, al);
        if (ksi) ksi[q] = dot(al, S);
        if (geo_aver) geo_aver[q] = mag(cr - f.center) / len;
        if (area) area[q] = f.area;
    }
    return 0;
}

// CFVM::create_FVM_from_geometry, boundary faces (cfvm.cpp:143-167).
int ocsl_mesh_bc_faces(ocsl_mesh m, int32_t* left, int32_t* marker, double* area_vector, double* unit,
                       double* dist, double* alpha, double* ksi, double* perp, double* area) {
    if (!m) return -1;
    for (size_t q = 0; q < m->boundary.size(); ++q) {
        const Face& f = m->faces[(size_t)m->boundary[q]];
        const V3& cl = m->center[(size_t)f.left];
        const V3 S = f.area * f.normal;
        V3 d = f.center - cl;
        const double len = mag(f.center - cl);
        const V3 al = (1.0 / dot(S, f.center - cl)) * S;
        d = (1.0 / len) * d;
        if (left) left[q] = f.left;
        if (marker) marker[q] = f.marker;
        if (area_vector) put3(area_vector, q, S);
        if (unit) put3(unit, q, d);
        if (dist) dist[q] = len;
        if (alpha) put3(alpha, q, al);
        if (ksi) ksi[q] = dot(al, S);
        if (perp) perp[q] = dot(len * d, (1.0 / f.area) * S);
        if (area) area[q] = f.area;
    }
    return 0;
}

// Cpost_processing::write_VTK_file (cpost_processing.cpp:260-333) and
// write_VTK_data (:26-127): VTK node order per cell type (prism 0 2 1 3 5 4,
// hexahedron back to the file's order), then the cell data.
int ocsl_mesh_write_vtk(ocsl_mesh m, const char* path, const ocsl_vtk_fields* fields, int precision) {
    if (!m || !path || precision < 1 || precision > 17) {
        g_err = "bad argument";
        return -1;
    }
    FILE* f = std::fopen(path, "w");
    if (!f) {
        g_err = std::string("cannot open ") + path + " for writing";
        return -2;
    }
    const int p = precision;
    const size_t n = m->cells.size();
    std::fprintf(f, "# vtk DataFile Version 1.0\nSolver Data File\nASCII\nDATASET UNSTRUCTURED_GRID\nPOINTS %zu double\n",
                 m->nodes.size());
    for (const V3& x : m->nodes) std::fprintf(f, "%.*g %.*g %.*g\n", p, x.v[0], p, x.v[1], p, x.v[2]);
    size_t list = 0;
    for (const Cell& c : m->cells) list += 1 + (size_t)kNodes[c.type];
    std::fprintf(f, "CELLS %zu %zu\n", n, list);
    static const int kVtkOrder[3][8] = {{0, 1, 2, 3}, {0, 2, 1, 3, 5, 4}, {0, 1, 3, 2, 4, 5, 7, 6}};
    for (const Cell& c : m->cells) {
        std::fprintf(f, "%d", kNodes[c.type]);
        for (int k = 0; k < kNodes[c.type]; ++k) std::fprintf(f, " %d", c.node[kVtkOrder[c.type][k]]);
        std::fputc('\n', f);
    }
    static const int kVtkType[3] = {10, 13, 12};
    std::fprintf(f, "CELL_TYPES %zu\n", n);
    for (const Cell& c : m->cells) std::fprintf(f, "%d\n", kVtkType[c.type]);
    std::fprintf(f, "CELL_DATA %zu\n", n);
    auto scalars = [&](const char* name, const double* v) {
        if (!v) return;
        std::fprintf(f, "SCALARS %s double 1\nLOOKUP_TABLE default\n", name);
        for (size_t c = 0; c < n; ++c) std::fprintf(f, "%.*g\n", p, v[c]);
    };
    auto vectors = [&](const char* name, const double* v) {
        if (!v) return;
        std::fprintf(f, "VECTORS %s double\n", name);
        for (size_t c = 0; c < n; ++c) std::fprintf(f, "%.*g %.*g %.*g\n", p, v[3 * c], p, v[3 * c + 1], p, v[3 * c + 2]);
    };
    scalars("Volume[m3]", m->vol.data());
    std::vector<double> ctr(3 * n);
    for (size_t c = 0; c < n; ++c) put3(ctr.data(), c, m->center[c]);
    vectors("Center", ctr.data());
    if (fields) {
        scalars("Density[Kg/m^3]", fields->density);
        scalars("Pressure[Pa]", fields->pressure);
        scalars("Temperature[K]", fields->temperature);
        scalars("viscosity", fields->viscosity);
        scalars("MACH", fields->mach);
        vectors("Velocity[m/s]", fields->velocity);
        vectors("Pres_grad", fields->pressure_grad);
    }
    const bool ok = !std::ferror(f);
    if (std::fclose(f) != 0 || !ok) {
        g_err = std::string("write to ") + path + " failed";
        return -2;
    }
    return 0;
}

// Cpost_processing::write_boundary_VTK_file (cpost_processing.cpp:335-401)
// and write_boundary_VTK_data (:129-193): the mesh's nodes, its boundary
// faces as VTK cells (nodes in the stored order, types 5 / 9), then per face
// its area, unit normal, marker, area vector and centre, and the flow
// fields.
int ocsl_mesh_write_boundary_vtk(ocsl_mesh m, const char* path, const ocsl_vtk_bc_fields* fields, int precision) {
    if (!m || !path || precision < 1 || precision > 17) {
        g_err = "bad argument";
        return -1;
    }
    FILE* f = std::fopen(path, "w");
    if (!f) {
        g_err = std::string("cannot open ") + path + " for writing";
        return -2;
    }
    const int p = precision;
    const size_t B = m->boundary.size();
    std::fprintf(f, "# vtk DataFile Version 1.0\nSolver Data File\nASCII\nDATASET UNSTRUCTURED_GRID\nPOINTS %zu double\n",
                 m->nodes.size());
    for (const V3& x : m->nodes) std::fprintf(f, "%.*g %.*g %.*g\n", p, x.v[0], p, x.v[1], p, x.v[2]);
    size_t list = 0;
    for (int q : m->boundary) list += 1 + (size_t)m->faces[(size_t)q].nn;
    std::fprintf(f, "CELLS %zu %zu\n", B, list);
    for (int q : m->boundary) {
        const Face& fc = m->faces[(size_t)q];
        std::fprintf(f, "%d", fc.nn);
        for (int k = 0; k < fc.nn; ++k) std::fprintf(f, " %d", fc.node[k]);
        std::fputc('\n', f);
    }
    std::fprintf(f, "CELL_TYPES %zu\n", B);
    for (int q : m->boundary) std::fprintf(f, "%d\n", m->faces[(size_t)q].nn == 3 ? 5 : 9);
    std::fprintf(f, "CELL_DATA %zu\n", B);
    auto vec3 = [&](const V3& a) { std::fprintf(f, "%.*g %.*g %.*g\n", p, a.v[0], p, a.v[1], p, a.v[2]); };
    std::fprintf(f, "SCALARS Area[m2] double 1\nLOOKUP_TABLE default\n");
    for (int q : m->boundary) std::fprintf(f, "%.*g\n", p, m->faces[(size_t)q].area);
    std::fprintf(f, "VECTORS normal double\n");
    for (int q : m->boundary) vec3(m->faces[(size_t)q].normal);
    std::fprintf(f, "SCALARS Marker double 1\nLOOKUP_TABLE default\n");
    for (int q : m->boundary) std::fprintf(f, "%d\n", m->faces[(size_t)q].marker);
    std::fprintf(f, "Vectors area_vector double\n");
    for (int q : m->boundary) {
        const Face& fc = m->faces[(size_t)q];
        vec3(fc.area * fc.normal);
    }
    std::fprintf(f, "Vectors center double\n");
    for (int q : m->boundary) vec3(m->faces[(size_t)q].center);
    if (fields && fields->mach) {
        std::fprintf(f, "SCALARS MACH double 1\nLOOKUP_TABLE default\n");
        for (size_t b = 0; b < B; ++b) std::fprintf(f, "%.*g\n", p, fields->mach[b]);
    }
    if (fields && fields->wall_stress) {
        std::fprintf(f, "Vectors Stress_vector double\n");
        const double* w = fields->wall_stress;
        for (size_t b = 0; b < B; ++b) std::fprintf(f, "%.*g %.*g %.*g\n", p, w[3 * b], p, w[3 * b + 1], p, w[3 * b + 2]);
    }
    const bool ok = !std::ferror(f);
    if (std::fclose(f) != 0 || !ok) {
        g_err = std::string("write to ") + path + " failed";
        return -2;
    }
    return 0;
}

}  // extern "C"
