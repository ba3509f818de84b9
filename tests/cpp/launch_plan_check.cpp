// Host-only check of the tolerance-CG launch planning in
// opencslsolver_amd/csrc/hip/ocsl_launch.h: the very functions the library
// calls (cg_launch, residual_launch, block_slices, xcd_map, cg_chunk and the
// graph keys), swept over every system size up to 4096 rows, every n^3 hex
// system up to 130^3 and every launch option class the library accepts.
//
// For each launch it checks what the kernels assume of it:
//  * the SpMV-type kernels (k_cg_residual, k_cg_spmv, k_cgm_spmv, the
//    one-kernel forms): every slice is processed by exactly one wave of one
//    block, and no row they touch lies past the vectors' allocation
//    (whole slices + 4 elements);
//  * the update kernels (k_cg_update_*, k_cgm_update: one 16-byte vector per
//    thread, grid-stride, first vector loaded at an index clamped to the
//    last full vector, the remainder done by the last block): every vector
//    index is visited once, the clamped load stays in the allocation and the
//    remainder fits one block;
//  * every block writes its partial sums inside its slot (grid <= 2048);
//  * the iteration chunks hold whole periods and make progress;
//  * distinct graph-cache arguments give distinct keys.
// Prints "launch plan ok" and exits 0, or names the first violation.
#include <cstdio>
#include <cstdlib>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "ocsl_launch.h"

using namespace ocsl;

namespace {

int failures = 0;

void fail(const std::string& what) {
    if (failures < 20) std::fprintf(stderr, "FAIL %s\n", what.c_str());
    ++failures;
}

std::string ctx(int n, int rpl, const ocsl_options& o) {
    return "n=" + std::to_string(n) + " rpl=" + std::to_string(rpl) + " block_size=" +
           std::to_string(o.block_size) + " grid_blocks=" + std::to_string(o.grid_blocks) +
           " vec_blocks=" + std::to_string(o.vec_blocks);
}

// SpMV-type kernel: block b is logical block xcd_map(b, g); its waves take
// slices s0 + w, s0 + w + waves, ... below s1; wave lane l of slice s touches
// rows s * 64 R + l R + j, j < R.
void check_spmv(int n, int rpl, const CgLaunch& L, const std::string& where) {
    const int rows_per_slice = kWave * rpl;
    const int n_slices = (n + rows_per_slice - 1) / rows_per_slice;
    const long alloc = (long)n_slices * rows_per_slice + 4;
    std::vector<int> seen(n_slices, 0);
    const int waves = L.block / kWave;
    for (int b = 0; b < L.grid; ++b) {
        const int lb = xcd_map(b, L.grid);
        const SliceRange r = block_slices(lb, L.grid, n_slices);
        if (r.s0 < 0 || r.s1 > n_slices || r.s0 > r.s1) fail(where + ": slice range");
        for (int w = 0; w < waves; ++w)
            for (int s = r.s0 + w; s < r.s1; s += waves) {
                ++seen[s];
                const long last_row = (long)s * rows_per_slice + (kWave - 1) * rpl + (rpl - 1);
                if (last_row >= alloc) fail(where + ": row past allocation");
            }
    }
    for (int s = 0; s < n_slices; ++s)
        if (seen[s] != 1) {
            fail(where + ": slice " + std::to_string(s) + " processed " + std::to_string(seen[s]) + " times");
            break;
        }
}

// Update kernel with N elements per 16-byte vector.
void check_update(int n, int rpl, int N, const CgLaunch& L, const std::string& where) {
    const int rows_per_slice = kWave * rpl;
    const long alloc = (long)((n + rows_per_slice - 1) / rows_per_slice) * rows_per_slice + 4;
    const int nv = n / N;
    const long stride = (long)L.vgrid * L.block;
    std::vector<unsigned char> seen(nv, 0);
    // threads t >= nv visit nothing and load at the clamped index nv
    if ((long)nv * N + N - 1 >= alloc) fail(where + ": clamped load past allocation");
    for (long t = 0; t < stride && t < nv; ++t)
        for (long v = t; v < nv; v += stride)
            if (seen[v]++) fail(where + ": vector visited twice");
    for (int v = 0; v < nv; ++v)
        if (!seen[v]) {
            fail(where + ": vector " + std::to_string(v) + " not visited");
            break;
        }
    if (n - nv * N >= L.block) fail(where + ": remainder larger than a block");
}

void check_launch(int n, int rpl, const ocsl_options& o) {
    const std::string where = ctx(n, rpl, o);
    const int n_slices = (n + kWave * rpl - 1) / (kWave * rpl);
    CgLaunch L;
    try {
        L = cg_launch(n_slices, n, o);
    } catch (const Error& e) {
        fail(where + ": cg_launch threw " + e.what);
        return;
    }
    if (L.grid > kMaxRedBlocks || L.vgrid > kMaxRedBlocks) fail(where + ": partial-sum slot overflow");
    check_spmv(n, rpl, L, where);
    check_update(n, rpl, 4, L, where + " (FP32 vectors)");
    check_update(n, rpl, 2, L, where + " (FP64 vectors)");
    const CgLaunch R = residual_launch(n_slices);
    check_spmv(n, rpl, R, where + " (final residual)");
    if (slices_per_wave(n_slices, L) < 1) fail(where + ": slices per wave");
}

void check_xcd_map() {
    for (int g = 1; g <= kMaxRedBlocks; ++g) {
        std::vector<unsigned char> hit(g, 0);
        for (int b = 0; b < g; ++b) {
            const int lb = xcd_map(b, g);
            if (lb < 0 || lb >= g || hit[lb]++) {
                fail("xcd_map is not a bijection for g=" + std::to_string(g));
                break;
            }
        }
    }
}

void check_chunks() {
    for (int per : {1, 2, 4, 8})
        for (int every = 1; every <= 70; ++every)
            for (int left = 1; left <= 200; ++left)
                for (int ug = 0; ug < 2; ++ug) {
                    const CgChunk c = cg_chunk(every, per, left, ug != 0);
                    const std::string w = "chunk per=" + std::to_string(per) + " every=" + std::to_string(every) +
                                          " left=" + std::to_string(left);
                    if (c.chunk < 1 || c.chunk % per) fail(w + ": not whole periods");
                    if (c.kstop < 1 || c.kstop > c.chunk || c.kstop > left) fail(w + ": kstop");
                    if (c.graph && !(ug && c.chunk < left)) fail(w + ": graph chunk reaches the budget");
                    if (per > 1 && !c.graph && c.chunk <= left && c.chunk < ((every + per - 1) / per) * per)
                        fail(w + ": short chunk without reaching the budget");
                }
    bool threw = false;
    try {
        cg_chunk(0, 1, 5, true);
    } catch (const Error&) {
        threw = true;
    }
    if (!threw) fail("cg_chunk accepted check_every 0");
}

void check_keys() {
    std::set<GraphKey> keys;
    long made = 0;
    const int grids[] = {1, 2, 7, 8, 16, 512, 1000, 2040, 2048};
    for (int kind = 0; kind < 256; kind += 5)
        for (int g : grids)
            for (int vg : grids)
                for (int block = kWave; block <= kCgMaxBlock; block += 3 * kWave)
                    for (int chunk : {1, 2, 8, 32, 33, 64, 1 << 20})
                        for (int rpl : {1, 2, 4}) {
                            const GraphKey k = cg_graph_key(kind, CgLaunch{g, block, vg}, chunk, rpl);
                            ++made;
                            if (!keys.insert(k).second) fail("CG graph keys collide");
                            if (k.a < ((int64_t)1 << 48)) fail("CG graph key outside its family");
                        }
    if ((long)keys.size() != made) fail("CG graph keys not distinct");
    for (int64_t kind : {(int64_t)0, (int64_t)16, (int64_t)1 << 20, ((int64_t)1 << 40) - 1})
        for (double w : {0.9, 0.8, 1.0})
            if (!keys.insert(bicgstab_graph_key(kind, 32, w)).second) fail("BiCGSTAB key collides");
    static double b0[4], b1[4];
    for (int fam : {1, 2})
        for (const void* b : {(const void*)b0, (const void*)b1, (const void*)nullptr})
            if (!keys.insert(reference_graph_key(fam, b)).second) fail("REFERENCE key collides");
    for (int k = 1; k <= 4; ++k)
        for (const void* b : {(const void*)b0, (const void*)b1})
            if (!keys.insert(reference_graph_key(3, b, k)).second) fail("REFERENCE k-RHS key collides");
    for (auto fk : {std::pair<int, int>{3, 0}, {3, 5}, {1, 2}, {4, 1}}) {
        bool threw = false;
        try {
            reference_graph_key(fk.first, b0, fk.second);
        } catch (const Error&) {
            threw = true;
        }
        if (!threw) fail("reference_graph_key accepted family " + std::to_string(fk.first));
    }
    for (int kind : {-1, 256}) {
        bool threw = false;
        try {
            cg_graph_key(kind, CgLaunch{8, 256, 8}, 32, 2);
        } catch (const Error&) {
            threw = true;
        }
        if (!threw) fail("cg_graph_key accepted kind " + std::to_string(kind));
    }
    bool threw = false;
    try {
        cg_graph_key(0, CgLaunch{0, 256, 8}, 32, 2);
    } catch (const Error&) {
        threw = true;
    }
    if (!threw) fail("cg_graph_key accepted an empty grid");
}

}  // namespace

int main() {
    check_xcd_map();
    check_chunks();
    check_keys();
    std::vector<int> sizes;
    for (int n = 1; n <= 4096; ++n) sizes.push_back(n);
    for (int k = 17; k <= 130; ++k) sizes.push_back(k * k * k);
    std::vector<ocsl_options> opts;
    for (int bs : {0, 64, 100, 256, 512, 1024, 4000})
        for (int gb : {0, 1, 7, 512, 5000})
            for (int vb : {0, 13, 2048, 9000}) {
                ocsl_options o{};
                o.block_size = bs;
                o.grid_blocks = gb;
                o.vec_blocks = vb;
                opts.push_back(o);
            }
    long launches = 0;
    for (int n : sizes) {
        // every option class on the small sizes and the hex systems up to
        // 81^3 (the batch's range), the default launch on the rest
        const bool all = n <= 300 || n % 1000 == 0 || (n > 4096 && n <= 81 * 81 * 81);
        for (int rpl : {1, 2, 4})
            for (size_t i = 0; i < (all ? opts.size() : 1); ++i) {
                check_launch(n, rpl, opts[i]);
                ++launches;
            }
    }
    if (failures) {
        std::fprintf(stderr, "%d violations\n", failures);
        return 1;
    }
    std::printf("launch plan ok: %ld launches checked\n", launches);
    return 0;
}
