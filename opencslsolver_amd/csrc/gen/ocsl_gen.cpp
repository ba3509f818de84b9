// Synthetic seven-point hex-mesh systems for the OpenCSL MI355X solver path.
// The contract (engine, transforms, draw order) is stated in include/ocsl_gen.h.
#include "ocsl_gen.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <set>
#include <utility>
#include <vector>

namespace {

struct Stream {
    std::mt19937_64 eng;
    explicit Stream(uint64_t seed) : eng(seed) {}
    uint64_t raw() { return eng(); }
    double uniform() { return (double)(eng() >> 11) * 0x1.0p-53; }
    double normal() {
        const double u1 = 1.0 - uniform();  // (0, 1]
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

// perm[c] = new id of natural cell c; Fisher-Yates from the top.
std::vector<int64_t> make_perm(Stream& s, int64_t n_cells, bool permute) {
    std::vector<int64_t> perm(n_cells);
    for (int64_t c = 0; c < n_cells; ++c) perm[c] = c;
    if (!permute) return perm;
    for (int64_t i = n_cells - 1; i >= 1; --i) {
        const int64_t j = (int64_t)(s.uniform() * (double)(i + 1));
        std::swap(perm[i], perm[j]);
    }
    return perm;
}

// Per natural cell: 6 face slots (-x,+x,-y,+y,-z,+z), -1 neighbour = boundary.
struct HexMesh {
    int n;
    int64_t cells;
    explicit HexMesh(int n_) : n(n_), cells((int64_t)n_ * n_ * n_) {}
    int64_t id(int i, int j, int k) const { return i + (int64_t)n * (j + (int64_t)n * k); }
    int64_t nb(int64_t c, int f) const {
        const int i = (int)(c % n), j = (int)((c / n) % n), k = (int)(c / ((int64_t)n * n));
        switch (f) {
            case 0: return i > 0 ? c - 1 : -1;
            case 1: return i < n - 1 ? c + 1 : -1;
            case 2: return j > 0 ? c - n : -1;
            case 3: return j < n - 1 ? c + n : -1;
            case 4: return k > 0 ? c - (int64_t)n * n : -1;
            default: return k < n - 1 ? c + (int64_t)n * n : -1;
        }
    }
    bool boundary(int64_t c) const {
        for (int f = 0; f < 6; ++f)
            if (nb(c, f) < 0) return true;
        return false;
    }
};

// Writes the permuted CSR.  coef(c, f) gives the off-diagonal coefficient of
// row c towards neighbour f, diag[c] the diagonal of row c (natural ids).
template <class Coef>
void emit_csr(const HexMesh& m, const std::vector<int64_t>& perm, const std::vector<double>& diag,
              Coef coef, int sentinels, uint64_t* row_ptr, uint64_t* col, double* val) {
    std::vector<int64_t> inv(m.cells);
    for (int64_t c = 0; c < m.cells; ++c) inv[perm[c]] = c;
    row_ptr[0] = 0;
    for (int64_t r = 0; r < m.cells; ++r) {
        const int64_t c = inv[r];
        int len = 1;
        bool bnd = false;
        for (int f = 0; f < 6; ++f) {
            if (m.nb(c, f) >= 0) ++len; else bnd = true;
        }
        if (bnd && sentinels) ++len;
        row_ptr[r + 1] = row_ptr[r] + (uint64_t)len;
    }
    for (int64_t r = 0; r < m.cells; ++r) {
        const int64_t c = inv[r];
        std::pair<uint64_t, double> e[8];
        int len = 0;
        bool bnd = false;
        e[len++] = {(uint64_t)r, diag[c]};
        for (int f = 0; f < 6; ++f) {
            const int64_t d = m.nb(c, f);
            if (d < 0) { bnd = true; continue; }
            e[len++] = {(uint64_t)perm[d], coef(c, f)};
        }
        std::sort(e, e + len, [](const auto& a, const auto& b) { return a.first < b.first; });
        if (bnd && sentinels) e[len++] = {OCSL_SENTINEL_COL, 0.0};
        uint64_t p = row_ptr[r];
        for (int q = 0; q < len; ++q, ++p) {
            col[p] = e[q].first;
            val[p] = e[q].second;
        }
    }
}

}  // namespace

extern "C" int ocsl_gen_hex_sizes(int n, int sentinels, int64_t* n_rows, int64_t* nnz) {
    if (n < 2 || !n_rows || !nnz) return -1;
    const int64_t N = (int64_t)n * n * n;
    const int64_t faces = 3 * (int64_t)(n - 1) * n * n;
    int64_t nz = N + 2 * faces;
    if (sentinels) nz += N - (int64_t)(n - 2) * (n - 2) * (n - 2);
    *n_rows = N;
    *nnz = nz;
    return 0;
}

extern "C" int ocsl_gen_poisson3d(int n, uint64_t seed_mesh, uint64_t seed_rhs, double sigma,
                                  double leak, int permute, int sentinels, uint64_t* row_ptr,
                                  uint64_t* col, double* val, double* rhs_out, int64_t* perm_out) {
    if (n < 2 || !row_ptr || !col || !val) return -1;
    HexMesh m(n);
    Stream s(seed_mesh);
    std::vector<int64_t> perm = make_perm(s, m.cells, permute != 0);
    // a[3*c + dir] = coefficient of the face between c and its + neighbour in dir.
    std::vector<double> a(3 * (size_t)m.cells, 0.0);
    for (int dir = 0; dir < 3; ++dir)
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    const int lim[3] = {i, j, k};
                    if (lim[dir] == n - 1) continue;
                    a[3 * (size_t)m.id(i, j, k) + dir] = std::exp(sigma * s.normal());
                }
    auto face = [&](int64_t c, int f) -> double {
        const int dir = f / 2;
        const int64_t owner = (f & 1) ? c : m.nb(c, f);
        return a[3 * (size_t)owner + dir];
    };
    std::vector<double> diag(m.cells);
    for (int64_t c = 0; c < m.cells; ++c) {
        double d = 0.0;
        bool bnd = false;
        for (int f = 0; f < 6; ++f) {
            if (m.nb(c, f) >= 0) d += face(c, f); else bnd = true;
        }
        diag[c] = d + (bnd ? leak : 0.0);
    }
    emit_csr(m, perm, diag, [&](int64_t c, int f) { return -face(c, f); }, sentinels, row_ptr, col, val);
    if (rhs_out) {
        if (seed_rhs == UINT64_MAX) {
            for (int64_t r = 0; r < m.cells; ++r) rhs_out[r] = -1.0 + 2.0 * s.uniform();
        } else {
            Stream sb(seed_rhs);
            for (int64_t r = 0; r < m.cells; ++r) rhs_out[r] = -1.0 + 2.0 * sb.uniform();
        }
    }
    if (perm_out) std::memcpy(perm_out, perm.data(), sizeof(int64_t) * m.cells);
    return 0;
}

extern "C" int ocsl_gen_convdiff3d(int n, uint64_t seed_perm, double rho, double dt, double peclet,
                                   int permute, int sentinels, uint64_t* row_ptr, uint64_t* col,
                                   double* val, int64_t* perm_out) {
    if (n < 2 || !row_ptr || !col || !val || !(dt > 0.0) || !(peclet > 0.0)) return -1;
    HexMesh m(n);
    Stream s(seed_perm);
    std::vector<int64_t> perm = make_perm(s, m.cells, permute != 0);
    const double h = 2.0 * M_PI / n;
    const double area = h * h, vol = h * h * h;
    const double gamma = rho * h / peclet;
    const double D = gamma * area / h;
    // Mass flux through the face between c and its + neighbour in dir.
    auto mdot = [&](int64_t c, int dir) -> double {
        const int i = (int)(c % n), j = (int)((c / n) % n), k = (int)(c / ((int64_t)n * n));
        double x = (i + 0.5) * h, y = (j + 0.5) * h, z = (k + 0.5) * h;
        if (dir == 0) { x = (i + 1) * h; return rho * std::sin(x) * std::cos(y) * std::cos(z) * area; }
        if (dir == 1) { y = (j + 1) * h; return -rho * std::cos(x) * std::sin(y) * std::cos(z) * area; }
        return 0.0;
    };
    // Row accumulators in natural ids: diag and the 6 neighbour slots.
    std::vector<double> diag(m.cells, rho * vol / dt);
    std::vector<double> off(6 * (size_t)m.cells, 0.0);
    for (int dir = 0; dir < 3; ++dir)
        for (int64_t L = 0; L < m.cells; ++L) {
            const int64_t R = m.nb(L, 2 * dir + 1);
            if (R < 0) continue;
            const double md = mdot(L, dir);
            // Convective flux, first-order upwind (left owns the +face).
            if (md >= 0.0) {
                diag[L] += md;
                off[6 * (size_t)R + 2 * dir] += -md;
            } else {
                off[6 * (size_t)L + 2 * dir + 1] += md;
                diag[R] += -md;
            }
            // Central diffusion.
            diag[L] += D;
            off[6 * (size_t)L + 2 * dir + 1] += -D;
            diag[R] += D;
            off[6 * (size_t)R + 2 * dir] += -D;
        }
    emit_csr(m, perm, diag, [&](int64_t c, int f) { return off[6 * (size_t)c + f]; }, sentinels,
             row_ptr, col, val);
    if (perm_out) std::memcpy(perm_out, perm.data(), sizeof(int64_t) * m.cells);
    return 0;
}

extern "C" int ocsl_gen_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out) {
    if (n < 0 || (n > 0 && !out)) return -1;
    Stream s(seed);
    for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * s.uniform();
    return 0;
}

extern "C" int ocsl_gen_normal(uint64_t seed, int64_t n, double* out) {
    if (n < 0 || (n > 0 && !out)) return -1;
    Stream s(seed);
    for (int64_t i = 0; i < n; ++i) out[i] = s.normal();
    return 0;
}

extern "C" int ocsl_gen_batch_plan(uint64_t seed, int m, int lo, int hi, int* n_out,
                                   uint64_t* seed_mesh_out, uint64_t* seed_rhs_out) {
    if (m <= 0 || lo < 2 || hi < lo || !n_out || !seed_mesh_out || !seed_rhs_out) return -1;
    Stream s(seed);
    const double a = std::log((double)lo), b = std::log((double)hi);
    for (int k = 0; k < m; ++k) n_out[k] = (int)std::lround(std::exp(a + (b - a) * s.uniform()));
    for (int k = 0; k < m; ++k) {
        seed_mesh_out[k] = s.raw();
        seed_rhs_out[k] = s.raw();
    }
    return 0;
}

extern "C" int ocsl_gen_sample_slots(uint64_t seed, int64_t n_rows, const uint64_t* row_ptr,
                                     const uint64_t* col, int count, int64_t* slots_out) {
    if (n_rows <= 0 || !row_ptr || !col || count < 0 || (count > 0 && !slots_out)) return -1;
    const uint64_t nnz = row_ptr[n_rows];
    uint64_t real = 0;
    for (uint64_t p = 0; p < nnz; ++p) real += col[p] < (uint64_t)n_rows;
    if ((uint64_t)count > real) return -2;
    Stream s(seed);
    std::set<int64_t> got;
    while ((int)got.size() < count) {
        const int64_t p = (int64_t)(s.uniform() * (double)nnz);
        if (col[p] < (uint64_t)n_rows) got.insert(p);
    }
    int q = 0;
    for (int64_t p : got) slots_out[q++] = p;
    return 0;
}

extern "C" uint64_t ocsl_gen_hash(int64_t n_rows, const uint64_t* row_ptr, const uint64_t* col,
                                  const double* val, const double* b) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t bytes) {
        const unsigned char* c = (const unsigned char*)p;
        for (size_t i = 0; i < bytes; ++i) {
            h ^= c[i];
            h *= 1099511628211ull;
        }
    };
    const uint64_t nnz = row_ptr[n_rows];
    mix(row_ptr, sizeof(uint64_t) * (size_t)(n_rows + 1));
    mix(col, sizeof(uint64_t) * nnz);
    mix(val, sizeof(double) * nnz);
    if (b) mix(b, sizeof(double) * (size_t)n_rows);
    return h;
}
