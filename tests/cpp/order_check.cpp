// Host-only check of the cell order of ocsl_order.h (the header the library
// is built from):
//  * hilbert3 is a bijection of the 2^b cube onto [0, 8^b) whose consecutive
//    indices are face-adjacent points, for b = 1..6 (the defining property of
//    the Hilbert curve; a wrong rotation or Gray step breaks adjacency);
//  * cell_centres rebuilds the centres of a randomly numbered box mesh, with
//    randomly ordered and oriented faces, to within rounding of the true ones
//    (up to the translation that puts the first cell at the origin);
//  * curve_order on a randomly numbered 32^3 box stores every cell next to a
//    face neighbour of the cell stored before it (the whole curve), is a
//    permutation on a 20 x 13 x 7 box, and keeps most face neighbours of a
//    256-cell workgroup inside it.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "ocsl_order.h"

using namespace ocsl;

static int fails = 0;
#define CHECK(c, ...)                                  \
    do {                                               \
        if (!(c)) {                                    \
            std::fprintf(stderr, "FAIL %s:%d ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);         \
            std::fprintf(stderr, "\n");                \
            if (++fails > 20) std::exit(1);            \
        }                                              \
    } while (0)

struct Box {
    int n;
    std::vector<int> fL, fR, num;       // faces by new cell numbers; num[lex] = new number
    std::vector<double> fe, fd, centre;  // centre by new number
};

// nx x ny x nz box with spacing (hx, hy, hz), cells numbered by a random
// permutation, faces in random order, each face oriented at random.
static Box make_box(int nx, int ny, int nz, double hx, double hy, double hz, unsigned seed) {
    Box b;
    const int n = nx * ny * nz;
    b.n = n;
    std::mt19937 rng(seed);
    b.num.resize(n);
    for (int i = 0; i < n; ++i) b.num[i] = i;
    std::shuffle(b.num.begin(), b.num.end(), rng);
    b.centre.resize(3 * (size_t)n);
    auto lex = [&](int x, int y, int z) { return (z * ny + y) * nx + x; };
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const int c = b.num[lex(x, y, z)];
                b.centre[3 * c] = 0.1 + x * hx;
                b.centre[3 * c + 1] = -2.0 + y * hy;
                b.centre[3 * c + 2] = 7.0 + z * hz;
            }
    std::vector<std::pair<int, int>> faces;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                if (x + 1 < nx) faces.push_back({lex(x, y, z), lex(x + 1, y, z)});
                if (y + 1 < ny) faces.push_back({lex(x, y, z), lex(x, y + 1, z)});
                if (z + 1 < nz) faces.push_back({lex(x, y, z), lex(x, y, z + 1)});
            }
    std::shuffle(faces.begin(), faces.end(), rng);
    for (auto& f : faces) {
        int l = b.num[f.first], r = b.num[f.second];
        if (rng() & 1) std::swap(l, r);
        double d[3], m = 0.0;
        for (int k = 0; k < 3; ++k) {
            d[k] = b.centre[3 * r + k] - b.centre[3 * l + k];
            m += d[k] * d[k];
        }
        m = std::sqrt(m);
        b.fL.push_back(l);
        b.fR.push_back(r);
        b.fd.push_back(m);
        for (int k = 0; k < 3; ++k) b.fe.push_back(d[k] / m);
    }
    return b;
}

int main() {
    // the curve itself
    for (int bits = 1; bits <= 6; ++bits) {
        const uint32_t side = 1u << bits;
        const uint64_t total = (uint64_t)side * side * side;
        std::vector<int> at(total, -1);
        for (uint32_t z = 0; z < side; ++z)
            for (uint32_t y = 0; y < side; ++y)
                for (uint32_t x = 0; x < side; ++x) {
                    const uint64_t k = hilbert3(x, y, z, bits);
                    CHECK(k < total, "bits %d key %llu out of range", bits, (unsigned long long)k);
                    if (k >= total) continue;
                    CHECK(at[k] < 0, "bits %d key %llu taken twice", bits, (unsigned long long)k);
                    at[k] = (int)((z * side + y) * side + x);
                }
        for (uint64_t k = 1; k < total; ++k) {
            const int a = at[k - 1], b = at[k];
            const int dx = std::abs(a % (int)side - b % (int)side);
            const int dy = std::abs(a / (int)side % (int)side - b / (int)side % (int)side);
            const int dz = std::abs(a / (int)(side * side) - b / (int)(side * side));
            CHECK(dx + dy + dz == 1, "bits %d: keys %llu, %llu not adjacent", bits, (unsigned long long)k - 1,
                  (unsigned long long)k);
        }
        CHECK(hilbert3(0, 0, 0, bits) == 0, "curve starts at the origin");
    }
    // the top of the range stays below 2^63
    CHECK(hilbert3((1u << 21) - 1, (1u << 21) - 1, (1u << 21) - 1, 21) < (1ull << 63), "21-bit key range");

    // centres rebuilt from the faces
    {
        const Box b = make_box(9, 6, 5, 0.25, 0.5, 0.125, 11);
        const std::vector<double> c = cell_centres(b.n, (int)b.fL.size(), b.fL.data(), b.fR.data(), b.fe.data(),
                                                   b.fd.data());
        double worst = 0.0;
        for (int i = 0; i < b.n; ++i)
            for (int k = 0; k < 3; ++k)
                worst = std::max(worst, std::abs(c[3 * i + k] - (b.centre[3 * i + k] - b.centre[k])));
        CHECK(worst < 1e-12, "centres off by %g", worst);
    }
    // the whole curve through a 32^3 box
    {
        const Box b = make_box(32, 32, 32, 1e-3, 1e-3, 1e-3, 5);
        const std::vector<int> o = curve_order(b.n, (int)b.fL.size(), b.fL.data(), b.fR.data(), b.fe.data(),
                                               b.fd.data());
        std::vector<int> lexof(b.n);
        for (int i = 0; i < b.n; ++i) lexof[b.num[i]] = i;
        std::vector<char> hit(b.n, 0);
        for (int k = 0; k < b.n; ++k) {
            CHECK(o[k] >= 0 && o[k] < b.n && !hit[o[k]], "not a permutation at %d", k);
            if (o[k] >= 0 && o[k] < b.n) hit[o[k]] = 1;
        }
        int apart = 0;
        for (int k = 1; k < b.n; ++k) {
            const int a = lexof[o[k - 1]], c = lexof[o[k]];
            const int d = std::abs(a % 32 - c % 32) + std::abs(a / 32 % 32 - c / 32 % 32) + std::abs(a / 1024 - c / 1024);
            apart += d != 1;
        }
        CHECK(apart == 0, "%d consecutive stored cells are not face neighbours", apart);
    }
    // an uneven box: a permutation with most neighbours in the workgroup
    {
        const Box b = make_box(20, 13, 7, 0.3, 0.2, 0.7, 9);
        const int F = (int)b.fL.size();
        const std::vector<int> o = curve_order(b.n, F, b.fL.data(), b.fR.data(), b.fe.data(), b.fd.data());
        std::vector<int> pos(b.n, -1);
        for (int k = 0; k < b.n; ++k) {
            CHECK(o[k] >= 0 && o[k] < b.n && pos[o[k]] < 0, "not a permutation at %d", k);
            if (o[k] >= 0 && o[k] < b.n) pos[o[k]] = k;
        }
        int same = 0, same_in = 0;
        for (int f = 0; f < F; ++f) {
            same += pos[b.fL[f]] / 256 == pos[b.fR[f]] / 256;
            same_in += b.fL[f] / 256 == b.fR[f] / 256;
        }
        CHECK(same > 0.7 * F, "only %d of %d faces inside a 256-cell block", same, F);
        std::printf("uneven box: %.3f of faces inside a 256-cell block (input numbering %.3f)\n", (double)same / F,
                    (double)same_in / F);
    }
    if (fails) return 1;
    std::printf("order ok\n");
    return 0;
}
