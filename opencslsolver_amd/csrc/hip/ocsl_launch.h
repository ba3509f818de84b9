// Launch geometry of the tolerance-CG kernels and the key of their cached
// hipGraphs: plain host arithmetic with no HIP dependency, so that the same
// code the library runs is also compiled by a host-only checker
// (tests/cpp/launch_plan_check.cpp) and swept there over every system size
// and option the library accepts.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>

#include "ocsl.h"

#if defined(__HIPCC__)
#define OCSL_HD __host__ __device__
#else
#define OCSL_HD
#endif

namespace ocsl {

// What every entry point turns into a status and ocsl_last_error().
struct Error {
    ocsl_status status;
    std::string what;
};

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kBlock = 256;        // threads per workgroup of the generic kernels (4 waves)
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kXcds = 8;           // MI355X: 8 XCDs, workgroups dealt round-robin
constexpr int kMaxRedBlocks = 2048;
constexpr int kMaxRed = 16;        // reductions finalised by one kernel
constexpr int kCgMaxBlock = 1024;  // CG kernels run with 64..1024 threads per block

// Logical block id so that each XCD processes one contiguous range of blocks:
// hardware block b runs on XCD b % 8.  A bijection of [0, g) for every g.
OCSL_HD inline int xcd_map(int b, int g) {
    if (g % kXcds) return b;
    return (b % kXcds) * (g / kXcds) + b / kXcds;
}

// Slices [s0, s1) of logical block lb of g over n_slices slices: equal
// contiguous runs, the last ones possibly short or empty (s0 == s1), never
// past n_slices.
struct SliceRange {
    int s0, s1;
};
OCSL_HD inline SliceRange block_slices(int lb, int g, int n_slices) {
    const int spb = (n_slices + g - 1) / g;
    const int b = lb * spb;
    const int s0 = b < n_slices ? b : n_slices;
    const int e = s0 + spb;
    return {s0, e < n_slices ? e : n_slices};
}

struct CgLaunch {
    int grid, block;  // SpMV and residual kernels
    int vgrid;        // vector-update kernels
};

// A whole multiple of the XCD count when there are that many blocks.
inline int round_xcds(int g) { return g >= kXcds ? (g / kXcds) * kXcds : g; }

// Throws unless the launch is one every tolerance-CG kernel accepts: at
// least one block, at most kMaxRedBlocks blocks (one partial sum per block
// and reduction slot), whole wavefronts, at most kCgMaxBlock threads (the
// kernels' launch bound).
inline void check_cg_launch(const CgLaunch& L) {
    auto bad = [](const std::string& m) { throw Error{OCSL_ERR_ARG, "CG launch: " + m}; };
    if (L.grid < 1 || L.grid > kMaxRedBlocks) bad("grid " + std::to_string(L.grid));
    if (L.vgrid < 1 || L.vgrid > kMaxRedBlocks) bad("vector grid " + std::to_string(L.vgrid));
    if (L.block < kWave || L.block > kCgMaxBlock || L.block % kWave) bad("block " + std::to_string(L.block));
}

// Default launch: 512 blocks of 1024 threads (2 blocks, 8 waves per SIMD, on
// each of the 256 CUs).  Every block of a consuming kernel re-adds one
// partial per producing block, so fewer, larger blocks keep that cheap.  Small
// systems get one wave per slice.  All kernels of a solve share the launch:
// each re-adds partial sums the others stored.
inline CgLaunch cg_launch(int n_slices, int n, const ocsl_options& o) {
    int block = o.block_size > 0 ? o.block_size : 1024;
    block = std::max(kWave, std::min(kCgMaxBlock, block / kWave * kWave));
    const int waves = block / kWave;
    int cap = o.grid_blocks > 0 ? o.grid_blocks : 512;
    cap = std::min(cap, kMaxRedBlocks);
    const int grid = round_xcds(std::max(1, std::min(cap, (n_slices + waves - 1) / waves)));
    int vcap = o.vec_blocks > 0 ? o.vec_blocks : grid;
    vcap = std::min(vcap, kMaxRedBlocks);
    // one 16-byte vector per thread at most
    const int vneed = (int)std::max<int64_t>(1, ((int64_t)n / 2 + block - 1) / block);
    const CgLaunch L{grid, block, round_xcds(std::max(1, std::min(vcap, vneed)))};
    check_cg_launch(L);
    return L;
}

// The stand-alone FP64 residual of a finished solve: 16 waves of 1024-thread
// blocks, one slice each.
inline CgLaunch residual_launch(int n_slices) {
    const CgLaunch L{std::max(1, std::min(kMaxRedBlocks, (n_slices + 15) / 16)), kCgMaxBlock, 1};
    check_cg_launch(L);
    return L;
}

// Slices per wave of the one-kernel launch over a SELL-P layout.
inline int slices_per_wave(int n_slices, CgLaunch L) {
    const int spb = (n_slices + L.grid - 1) / L.grid;
    const int waves = L.block / kWave;
    return (spb + waves - 1) / waves;
}

// One chunk of tolerance-CG iterations between host checks: `chunk` launches
// of the iteration (whole periods of `per`: the MIXED x-update period, or
// the pairs of the one-kernel form), the last `chunk - kstop` of which find
// the solve stopped; replayed as a graph when `graph`.  `left` (>= 1) is the
// iteration budget still open.
struct CgChunk {
    int chunk, kstop;
    bool graph;
};

inline CgChunk cg_chunk(int check_every, int per, int left, bool use_graph) {
    if (check_every < 1 || per < 1 || left < 1) throw Error{OCSL_ERR_ARG, "CG chunk: bad arguments"};
    const int every = (check_every + per - 1) / per * per;
    const bool graph = use_graph && left > every;
    int chunk = graph ? every : std::max(1, std::min(every, left));
    if (per > 1) {
        // the kernel after the last allowed iteration records its residual
        // and stops; iterations past max_iter stop on their own
        if (!graph && chunk == left) ++chunk;
        chunk = (chunk + per - 1) / per * per;
    }
    return {chunk, std::min(chunk, left), graph};
}

// Key of a cached iteration-chunk graph.  Everything a captured launch bakes
// into its arguments that can change between solves on one handle must be in
// the key (buffer addresses cannot: a reallocation drops every graph).
struct GraphKey {
    int64_t a;       // solver family, iteration form, launch geometry
    int b;           // chunk length and rows per lane
    uint64_t c = 0;  // scalar options captured as kernel arguments (bit patterns)
    bool operator<(const GraphKey& o) const {
        if (a != o.a) return a < o.a;
        if (b != o.b) return b < o.b;
        return c < o.c;
    }
    bool operator==(const GraphKey& o) const { return a == o.a && b == o.b && c == o.c; }
};

// Tolerance CG: the iteration form `kind` (< 256), the launch, the chunk
// length and the rows per lane.  Distinct arguments give distinct keys (bit
// 48 set: no key of the BiCGSTAB family below); fields out of range throw
// instead of aliasing.
inline GraphKey cg_graph_key(int kind, const CgLaunch& L, int chunk, int rpl) {
    check_cg_launch(L);
    auto bad = [](const std::string& m) { throw Error{OCSL_ERR_ARG, "graph key: " + m}; };
    if (kind < 0 || kind >= 256) bad("kind " + std::to_string(kind));
    if (chunk < 1 || chunk > (1 << 26)) bad("chunk " + std::to_string(chunk));
    if (rpl < 1 || rpl > 7) bad("rows per lane " + std::to_string(rpl));
    // kind: bits 40..47; grid, vgrid (<= 2048): bits 24..35, 8..19; waves (<= 16): bits 0..4
    const int64_t a = (int64_t)1 << 48 | (int64_t)kind << 40 | (int64_t)L.grid << 24 | (int64_t)L.vgrid << 8 |
                      (L.block / kWave);
    return {a, chunk * 8 + rpl, 0};
}

// REFERENCE-mode solves (family 1: ILU_BICGSTAB_solve, 2: CG_solve): the
// whole solve is one graph; it bakes in the right-hand side's address, the
// one caller buffer a solve reads that can change between solves (bit 50
// set: no key of the other families).
// Family 3, the REFERENCE BiCGSTAB of k right-hand sides together, also
// keys on k (1..4).
inline GraphKey reference_graph_key(int family, const void* b, int k = 0) {
    if (family < 1 || family > 3) throw Error{OCSL_ERR_ARG, "graph key: reference family"};
    if (family == 3 ? (k < 1 || k > 4) : k != 0) throw Error{OCSL_ERR_ARG, "graph key: right-hand sides"};
    return {(int64_t)1 << 50 | family, k, (uint64_t)(uintptr_t)b};
}

// Tolerance BiCGSTAB: its iteration form (< 2^40), the chunk length and the
// damped-Jacobi weight of its triangular sweeps, which the captured sweeps
// take as an argument.
inline GraphKey bicgstab_graph_key(int64_t kind, int chunk, double tri_weight) {
    if (kind < 0 || kind >= ((int64_t)1 << 40)) throw Error{OCSL_ERR_ARG, "graph key: kind"};
    if (chunk < 1) throw Error{OCSL_ERR_ARG, "graph key: chunk " + std::to_string(chunk)};
    uint64_t w;
    static_assert(sizeof(w) == sizeof(tri_weight), "double is 64-bit");
    std::memcpy(&w, &tri_weight, sizeof(w));
    return {kind, chunk, w};
}

}  // namespace ocsl
