// The solver handle: one CSR structure on one device with one stream.
#pragma once

#include <map>
#include <memory>
#include <type_traits>

#include "ocsl_common.h"

namespace ocsl {

// Host-side analysis of the caller's CSR structure (setup only).
struct HostCSR {
    int n = 0;
    int64_t nnz_in = 0;       // entries in the caller's arrays, sentinels included
    std::vector<int> rp, col;  // real entries only, int32
    std::vector<int> src;      // real entry -> slot in the caller's arrays
    std::vector<int> diag;     // slot (real CSR) of (i,i)
};

// A SELL-64 copy of the matrix in some row order (identity or RCM).
struct Sell {
    int n = 0, n_slices = 0, max_w = 0;
    int64_t total = 0;         // padded entries
    DBuf<int> slice_ptr, col, src, perm, iperm;  // perm: new -> old (empty for identity)
    DBuf<double> val64;
    DBuf<float> val32;
    bool have64 = false, have32 = false;
    bool identity = true;
    template <class V>
    SellView<V> view() const;
};
template <>
inline SellView<double> Sell::view<double>() const {
    return {n, n_slices, slice_ptr.p, col.p, val64.p, max_w};
}
template <>
inline SellView<float> Sell::view<float>() const {
    return {n, n_slices, slice_ptr.p, col.p, val32.p, max_w};
}

// SELL-P copy of the matrix (tolerance-mode CG), RCM or input row order,
// rpl rows per lane (slices of 64 rpl rows).
struct SellP {
    int n = 0, n_slices = 0, rpl = 1;
    uint64_t epoch64 = ~0ull, epoch32 = ~0ull;  // value epochs of val64 / val32
    int64_t total = 0;                      // padded off-diagonal entries
    DBuf<int> slice_ptr, col, src, diag_src, perm;  // perm: new -> old (empty for identity)
    DBuf<int> win_lo, win_hi;               // per slice: row window of its SpMV (own rows and columns)
    DBuf<unsigned> colw;                    // packed 16-bit column offsets (compact)
    bool compact = false;
    DBuf<double> val64, diag64;
    DBuf<float> val32, diag32;              // diag32: row sums (difference form)
    bool identity = true;
    // rows of whole slices: the allocation size of every vector the SELL-P
    // kernels read by row groups without a tail guard
    size_t padded() const { return (size_t)n_slices * 64 * rpl; }
    template <class V>
    SellPView<V> view() const;
};
template <>
inline SellPView<double> SellP::view<double>() const {
    return {n, n_slices, slice_ptr.p, compact ? nullptr : col.p, compact ? colw.p : nullptr, val64.p, diag64.p,
            win_lo.p, win_hi.p};
}
template <>
inline SellPView<float> SellP::view<float>() const {
    return {n, n_slices, slice_ptr.p, compact ? nullptr : col.p, compact ? colw.p : nullptr, val32.p, diag32.p,
            win_lo.p, win_hi.p};
}

// Level schedule of a triangular solve: rows grouped by dependency depth.
// Rows with at most this many entries on one side of the diagonal keep their
// level-ordered entries in padded columns (Levels::ell_w).
constexpr int kTriEll = 8;

struct Levels {
    int n_levels = 0;
    int max_width = 0;
    // Padded-column copy of the entries when every row has at most kTriEll
    // (ell_w = that maximum, else 0): entry e of position k at e * n + k,
    // cnt[k] of them used; esw the factor slots (the diagonal's in padding).
    int ell_w = 0;
    DBuf<int> cnt, ecw, esw;
    std::vector<int> ptr_host;  // n_levels + 1
    DBuf<int> rows, ptr;
    // The direction's entries in level order: position k (row rows[k]) owns
    // entries eptr[k] .. eptr[k+1] (the row's strictly lower or strictly
    // upper part, in the row's own order), with their columns ecol and their
    // slots esrc in the factor; dslot[k] is the slot of the row's diagonal.
    DBuf<int> eptr, ecol, esrc, dslot;
    // Factor values gathered into that order, and the factor (address and
    // decomposition epoch) they were gathered from.
    DBuf<double> eval64, dval64;
    DBuf<float> eval32, dval32;
    const void* from64 = nullptr;
    const void* from32 = nullptr;
    int64_t epoch64 = -1, epoch32 = -1;
};

// Fine-grained ILU(0) sweep tables (cmatrix_solver.cpp:519-551): for each
// real entry p the (k, q) slot pairs whose products enter its sum, in the
// reference's k order, and the slot of the dividing diagonal (-1 for upper).
struct IluTables {
    DBuf<int> pair_ptr, pair_k, pair_q, den;
    int64_t n_pairs = 0;
};

struct Work {
    DBuf<double> d[16];
    DBuf<float> f[12];
};

}  // namespace ocsl

// Tuned tolerance-CG launch of a handle (0: not tuned).
struct ocsl_solver_tuned {
    int rows_per_lane = 0, block_size = 0, grid_blocks = 0, vec_blocks = 0, cg_kernels = 0;
};

struct ocsl_solver {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ocsl::HostCSR host;

    // device CSR of the real entries (input order)
    ocsl::DBuf<int> rp, col, src, diag;
    ocsl::DBuf<double> csr_val;  // gathered values (input order, real entries)
    ocsl::DBuf<float> csr_val32;
    bool csr_val_valid = false, csr_val32_valid = false;

    // caller data
    ocsl::DBuf<double> val_in_own, b_own, x_own;
    const double* val_in = nullptr;  // device pointer to the caller's value array
    const double* b_dev = nullptr;   // device pointer to b (input order)
    bool have_vals = false, have_b = false;
    uint64_t val_epoch = 0;          // bumped on every ocsl_set_values

    ocsl::Sell sell_id, sell_rcm;
    ocsl::SellP sellp[2][3];  // [reorder][rows per lane 1, 2, 4]
    ocsl_solver_tuned tuned[2];  // launch picked by ocsl_tune_cg per precision (zeros: none)
    uint64_t sell_id_epoch64 = ~0ull, sell_id_epoch32 = ~0ull;
    uint64_t sell_rcm_epoch64 = ~0ull, sell_rcm_epoch32 = ~0ull;

    ocsl::Levels lev_fwd, lev_bwd;
    bool levels_built = false;
    // form of the exact triangular solves: 0 chosen from the level schedule,
    // 1 one workgroup for the whole solve, 2 one launch per level, 3 serial
    // over the rows staged in LDS (ocsl_set_tri_form; the timing hook sets it
    // for the duration of a measurement)
    int tri_form = 0;
    int lds_max = 65536;  // LDS bytes one workgroup may use on this device
    ocsl::IluTables ilu;
    bool ilu_built = false;
    ocsl::DBuf<double> ilu_m, ilu_old;
    int64_t ilu_epoch = 0;  // bumped by every decomposition (ilu_m, ilu_m32 rewritten)
    ocsl::DBuf<float> ilu_m32;

    ocsl::Work work;
    ocsl::DBuf<ocsl::Scalars> S;
    ocsl::DBuf<ocsl::Scalars> S_multi;  // per right-hand side: REFERENCE BiCGSTAB of k together
    ocsl::Scalars* S_host = nullptr;  // pinned mirror
    ocsl::DBuf<double> red_partial;
    ocsl::DBuf<double> dscal;          // small scalar outputs of k_dots
    double* dscal_host = nullptr;      // pinned mirror
    ocsl::DBuf<double> hist;
    std::vector<double> hist_host;
    int hist_len = 0;

    // adjoint: handle of the transposed pattern, slot map into the caller's values
    ocsl_solver* adj = nullptr;
    ocsl::DBuf<int> adj_map;
    uint64_t adj_epoch = ~0ull;
    ocsl::DBuf<double> grad_own;

    // colour-ordered ILU(0)-BiCGSTAB: handle of the renumbered matrix, its
    // value map into the caller's slots, new -> old rows, staged vectors
    ocsl_solver* col_h = nullptr;
    ocsl::DBuf<int> col_map, col_perm;
    uint64_t col_epoch = ~0ull;
    int n_colours = 0;
    ocsl::DBuf<double> col_b, col_x;

    // hipGraph cache of iteration chunks (cg_graph_key, bicgstab_graph_key)
    // and of whole REFERENCE-mode solves (reference_graph_key)
    std::map<ocsl::GraphKey, hipGraphExec_t> graphs;
    bool ref_graphs = true;  // REFERENCE solves replay a graph after their first run

    ocsl::RedScratch red() const { return {red_partial.p}; }
};

namespace ocsl {

// setup (ocsl_setup.hip)
void analyse_structure(int64_t n, const uint64_t* rp, const uint64_t* col, HostCSR& out);
std::vector<int> rcm_order(const HostCSR& a);
void build_sell(const HostCSR& a, const std::vector<int>* order, Sell& s, hipStream_t st);
void build_sellp(const HostCSR& a, const std::vector<int>* order, int rpl, SellP& s, hipStream_t st);
void build_levels(const HostCSR& a, bool lower, Levels& L, hipStream_t st);
void build_ilu_tables(const HostCSR& a, IluTables& t, hipStream_t st);

// helpers shared by drivers (ocsl_api.hip)
void ensure_csr_values(ocsl_solver* h, bool need32);
void ensure_sell_values(ocsl_solver* h, Sell& s, bool need64, bool need32);
Sell& sell_for(ocsl_solver* h, bool reorder);
SellP& sellp_for(ocsl_solver* h, bool reorder, int rpl);
void ensure_sellp_values(ocsl_solver* h, SellP& s, bool need64, bool need32);
void sync(ocsl_solver* h);
void drop_graphs(ocsl_solver* h);
double device_dot(ocsl_solver* h, const double* a, const double* b, int n);
void device_dots(ocsl_solver* h, int n, int nr, const double* const* a, const double* const* b,
                 const int* kind, double* out);

// Work buffer of at least `count` elements; a reallocation invalidates the
// cached hipGraphs, whose kernel arguments hold the old addresses.
template <class T>
inline T* work_buf(ocsl_solver* h, DBuf<T>& b, size_t count) {
    if (!b.p || count > b.n) {
        drop_graphs(h);
        b.alloc(count);
    }
    return b.p;
}

// A REFERENCE-mode solve's device work (`body` issues it on the handle's
// stream): run eagerly the first time a right-hand-side address is seen (the
// run allocates whatever the solve allocates lazily), and captured then as a
// graph that later solves with the same address replay.  Returns after the
// work is issued.
template <class F>
void run_reference(ocsl_solver* h, int family, const void* b, F&& body, int k = 0) {
    if (!h->ref_graphs) {
        body();
        return;
    }
    const GraphKey key = reference_graph_key(family, b, k);
    auto it = h->graphs.find(key);
    if (it != h->graphs.end()) {
        OCSL_HIP(hipGraphLaunch(it->second, h->stream));
        return;
    }
    body();
    // at most a few right-hand-side buffers per family are kept
    int same = 0;
    for (const auto& kv : h->graphs) same += kv.first.a == key.a;
    if (same >= 4) drop_graphs(h);
    hipGraph_t gr;
    OCSL_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    body();
    OCSL_HIP(hipStreamEndCapture(h->stream, &gr));
    hipGraphExec_t ex;
    OCSL_HIP(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
    OCSL_HIP(hipGraphDestroy(gr));
    h->graphs.emplace(key, ex);
}

// Calls f(std::integral_constant<int, R>) for the SELL-P rows per lane R.
template <class F>
inline void with_rpl(int rpl, F&& f) {
    switch (rpl) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        default: throw Error{OCSL_ERR_ARG, "rows_per_lane must be 0, 1, 2 or 4"};
    }
}

inline int grid_for(int64_t n, int per_block = kBlock, int cap = 4096) {
    int64_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// solvers (ocsl_cg.hip, ocsl_bicgstab.hip)
void cg_reference(ocsl_solver* h, ocsl_report* rep);
void cg_tolerance(ocsl_solver* h, bool mixed, const ocsl_options& o, ocsl_report* rep);
void bicgstab_reference(ocsl_solver* h, ocsl_report* rep);
void bicgstab_reference_multi(ocsl_solver* h, int k, const double* B_dev, double* X_dev, ocsl_report* reps);
void bicgstab_tolerance(ocsl_solver* h, bool mixed, const ocsl_options& o, ocsl_report* rep);
// The reference's ILU sweeps from M^0 = m0 (device, real-entry order), or
// from A when m0 is null; the result lands in h->ilu_m.
void ilu_decompose_device(ocsl_solver* h, int sweeps, bool want32, const double* m0 = nullptr);
void colour_order_get(ocsl_solver* h, int* order, int* n_colours);
void bicgstab_multi(ocsl_solver* h, bool mixed, const ocsl_options& o, int k, const double* B_dev,
                    double* X_dev, ocsl_report* reps);

}  // namespace ocsl
