// Space-filling-curve order of a finite-volume mesh's cells: plain host code
// with no HIP dependency, compiled by the library (ocsl_flow_create) and by a
// host-only checker (tests/cpp/order_check.cpp).
//
// The flow kernels gather each cell's face neighbours.  An order in which
// most neighbours of a workgroup's cells fall in the same workgroup, or in
// one processed shortly before, lets those gathers hit the XCD's L2.  The
// cells are sorted along a 3-D Hilbert curve through their centres.  The
// mesh interface carries no cell centres, so they are rebuilt from the
// faces: the centre of a face's right cell lies at the left cell's centre
// plus cell_center_distance * cell_center_unit_vector (cfvm.cpp:129-133), so
// a breadth-first walk over the faces places every cell of a connected mesh
// relative to its first cell.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

namespace ocsl {

// Index of the lattice point (x, y, z), each coordinate below 2^bits
// (1 <= bits <= 21), along the 3-D Hilbert curve of that cube.  The method
// is J. Skilling's ("Programming the Hilbert curve", AIP Conf. Proc. 707,
// 381 (2004)): the coordinates are turned into the curve's "transposed"
// index by undoing the excess work of the rotations top bit down, then Gray
// encoding; the index interleaves the transposed words' bits, most
// significant first.  Consecutive indices are face-adjacent lattice points.
inline uint64_t hilbert3(uint32_t x, uint32_t y, uint32_t z, int bits) {
    uint32_t X[3] = {x, y, z};
    const uint32_t M = 1u << (bits - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    X[1] ^= X[0];
    X[2] ^= X[1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    for (int i = 0; i < 3; ++i) X[i] ^= t;
    uint64_t key = 0;
    for (int b = bits - 1; b >= 0; --b)
        for (int i = 0; i < 3; ++i) key = (key << 1) | ((X[i] >> b) & 1u);
    return key;
}

// Cell centres rebuilt from the interior faces, up to one translation per
// connected part of the mesh (each part's first cell at the origin).
// fe: unit vectors (3 per face, left to right), fdist: centre distances.
inline std::vector<double> cell_centres(int n, int F, const int* fL, const int* fR, const double* fe,
                                        const double* fdist) {
    std::vector<int> ptr(n + 1, 0), adj(2 * (size_t)F);
    for (int f = 0; f < F; ++f) {
        ptr[fL[f] + 1]++;
        ptr[fR[f] + 1]++;
    }
    for (int i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int f = 0; f < F; ++f) {
        adj[fill[fL[f]]++] = f;
        adj[fill[fR[f]]++] = f;
    }
    std::vector<double> c(3 * (size_t)n, 0.0);
    std::vector<char> seen(n, 0);
    std::vector<int> queue;
    queue.reserve(n);
    for (int root = 0; root < n; ++root) {
        if (seen[root]) continue;
        seen[root] = 1;
        queue.push_back(root);
        for (size_t head = queue.size() - 1; head < queue.size(); ++head) {
            const int i = queue[head];
            for (int q = ptr[i]; q < ptr[i + 1]; ++q) {
                const int f = adj[q];
                const bool left = fL[f] == i;
                const int j = left ? fR[f] : fL[f];
                if (seen[j]) continue;
                seen[j] = 1;
                const double sgn = left ? fdist[f] : -fdist[f];
                for (int k = 0; k < 3; ++k) c[3 * (size_t)j + k] = c[3 * (size_t)i + k] + sgn * fe[3 * (size_t)f + k];
                queue.push_back(j);
            }
        }
    }
    return c;
}

// Cells sorted along the Hilbert curve through their centres: order[k] =
// the cell stored k-th.  The centres are placed on a lattice of half the
// median centre distance (so face neighbours land on distinct points) and
// shrunk to at most 2^21 points per axis; cells on one lattice point keep
// their input order.
inline std::vector<int> curve_order(int n, int F, const int* fL, const int* fR, const double* fe,
                                    const double* fdist) {
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    if (n < 2 || F < 1) return order;
    const std::vector<double> c = cell_centres(n, F, fL, fR, fe, fdist);
    double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {0.0, 0.0, 0.0};  // non-finite centres (bad faces) count as 0
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < n; ++i) {
            const double v = c[3 * (size_t)i + k];
            if (!std::isfinite(v)) continue;
            lo[k] = std::min(lo[k], v);
            hi[k] = std::max(hi[k], v);
        }
    std::vector<double> d(fdist, fdist + F);
    std::nth_element(d.begin(), d.begin() + F / 2, d.end());
    double h = 0.5 * d[F / 2];
    const double span = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));  // finite
    const double top = (double)((1u << 21) - 1);
    if (!(h > 0.0) || !std::isfinite(h) || span / h > top) h = span > 0.0 ? span / top : 1.0;
    int bits = 1;
    while (bits < 21 && ((double)(1u << bits) - 1.0) * h < span) ++bits;
    std::vector<uint64_t> key(n);
    for (int i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int k = 0; k < 3; ++k) {
            double v = std::floor((c[3 * (size_t)i + k] - lo[k]) / h + 0.5);
            if (!(v >= 0.0)) v = 0.0;
            v = std::min(v, (double)((1u << bits) - 1));
            q[k] = (uint32_t)v;
        }
        key[i] = hilbert3(q[0], q[1], q[2], bits);
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
    return order;
}

}  // namespace ocsl
