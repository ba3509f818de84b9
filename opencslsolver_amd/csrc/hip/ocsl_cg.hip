// CG drivers.
//   cg_reference : CMatrix_solver::CG_solve (cmatrix_solver.cpp:635-680) with
//                  its 4-iteration cap, absolute 1e-6 stop and two SpMVs per
//                  iteration, FP64, input row order.
//   cg_tolerance : textbook CG to a relative residual, three fused kernels per
//                  iteration with device-resident scalars, RCM row order,
//                  iteration chunks replayed as hipGraphs.  MIXED runs the
//                  iterations in FP32 (matrix in difference form, r, p, q)
//                  with FP64 dot accumulation, the solution accumulated in
//                  FP64 every iteration and the FP64 true residual b - A x
//                  replacing the FP32 recurrence residual whenever the latter
//                  has dropped by `replace_ratio` (reliable updates).
#include <cmath>
#include <cstdio>

#include "ocsl_handle.h"
#include "ocsl_kernels.h"

namespace ocsl {

static void launch_axpby(ocsl_solver* h, double a, const double* x, double b, const double* y,
                         double* out, int n) {
    k_axpby<<<grid_for(n), kBlock, 0, h->stream>>>(n, a, x, b, y, out);
    OCSL_HIP(hipGetLastError());
}

static void launch_spmv(ocsl_solver* h, const Sell& s, double a, const double* x, double b,
                        const double* z, double* out) {
    k_sell_axpby<double><<<grid_for((int64_t)s.n_slices * kWave), kBlock, 0, h->stream>>>(
        s.view<double>(), a, x, b, z, out);
    OCSL_HIP(hipGetLastError());
}

// ||b - A x|| / ||b|| over SELL-P: the residual kernel writing into scratch.
static double true_relres_p(ocsl_solver* h, const SellP& s, const double* b, const double* x) {
    const int n = s.n;
    double* r = work_buf(h, h->work.d[11], n);
    const int g = std::min(kMaxRedBlocks, (s.n_slices + 15) / 16 * 1);
    k_cg_residual<false><<<std::max(1, g), 1024, 0, h->stream>>>(s.view<double>(), b, x, r, nullptr, nullptr,
                                                                 nullptr, h->red(), 0);
    k_finalize<2, FinStore<2>><<<1, kBlock, 0, h->stream>>>(h->red(), std::max(1, g), FinStore<2>{h->dscal.p},
                                                            kSlotRR);
    OCSL_HIP(hipGetLastError());
    OCSL_HIP(hipMemcpyAsync(h->dscal_host, h->dscal.p, 2 * sizeof(double), hipMemcpyDeviceToHost,
                            h->stream));
    sync(h);
    const double rr = h->dscal_host[0], bb = h->dscal_host[1];
    return bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
}

static double true_relres(ocsl_solver* h, const Sell& s, const double* b, const double* x) {
    const int g = red_blocks(kBlock * 4, s.n);
    k_true_residual<double><<<g, kBlock, 0, h->stream>>>(s.view<double>(), b, x, h->red());
    k_finalize<2, FinStore<2>><<<1, kBlock, 0, h->stream>>>(h->red(), g, FinStore<2>{h->dscal.p});
    OCSL_HIP(hipGetLastError());
    OCSL_HIP(hipMemcpyAsync(h->dscal_host, h->dscal.p, 2 * sizeof(double), hipMemcpyDeviceToHost,
                            h->stream));
    sync(h);
    const double rr = h->dscal_host[0], bb = h->dscal_host[1];
    return bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
}

void cg_reference(ocsl_solver* h, ocsl_report* rep) {
    const int n = h->host.n;
    Sell& s = sell_for(h, false);
    ensure_sell_values(h, s, true, false);
    Work& w = h->work;
    for (int i = 0; i < 5; ++i) work_buf(h, w.d[i], n);
    double *r = w.d[0].p, *rn = w.d[1].p, *p = w.d[2].p, *pn = w.d[3].p, *q = w.d[4].p;
    double* x = h->x_own.p;
    const double* b = h->b_dev;
    std::vector<double> hist;
    OCSL_HIP(hipEventRecord(h->ev0, h->stream));
    // r = b - A x ; p = r
    launch_spmv(h, s, -1.0, x, 1.0, b, r);
    OCSL_HIP(hipMemcpyAsync(p, r, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
    double res = 0.0;
    int it = 0;
    for (; it < 4; ++it) {
        launch_spmv(h, s, 1.0, p, 0.0, nullptr, q);
        const double* da[2] = {r, p};
        const double* db[2] = {r, q};
        const int kind[2] = {0, 0};
        double d[2];
        device_dots(h, n, 2, da, db, kind, d);
        const double a = d[0] / (d[1] + 1.0e-40);
        launch_axpby(h, 1.0, x, a, p, x, n);
        launch_spmv(h, s, -1.0 * a, p, 1.0, r, rn);
        res = device_dot(h, rn, rn, n);
        hist.push_back(res);
        if (res < 1.0e-6) { ++it; break; }
        const double beta = res / (d[0] + 1.0e-40);
        launch_axpby(h, 1.0, rn, beta, p, pn, n);
        std::swap(r, rn);
        std::swap(p, pn);
    }
    OCSL_HIP(hipEventRecord(h->ev1, h->stream));
    sync(h);
    float ms = 0.f;
    OCSL_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->hist_host = hist;
    h->hist_len = (int)hist.size();
    rep->status = OCSL_OK;
    if (!std::isfinite(res)) rep->status = OCSL_WARN_NONFINITE;
    rep->iterations = it;
    rep->residual = res;
    rep->true_relres = true_relres(h, s, b, x);
    rep->history_len = h->hist_len;
    rep->solve_ms = ms;
    const double nnz = (double)h->host.col.size();
    rep->bytes = it * 2.0 * (nnz * 12.0 + 16.0 * n);
}

namespace {

struct CgLaunch {
    int grid, block;  // SpMV and residual kernels
    int vgrid;        // vector-update kernels
};

int round8(int g) { return g >= kXcds ? (g / kXcds) * kXcds : g; }

// Default launch: 512 blocks of 1024 threads (2 blocks, 8 waves per SIMD, on
// each of the 256 CUs).  Every block of a consuming kernel re-adds one
// partial per producing block, so fewer, larger blocks keep that cheap.  Small
// systems get one wave per slice.  All kernels of a solve share the launch:
// each re-adds partial sums the others stored.
CgLaunch cg_launch(int n_slices, int n, const ocsl_options& o) {
    int block = o.block_size > 0 ? o.block_size : 1024;
    block = std::max(kWave, std::min(kCgMaxBlock, block / kWave * kWave));
    const int waves = block / kWave;
    int cap = o.grid_blocks > 0 ? o.grid_blocks : 512;
    cap = std::min(cap, kMaxRedBlocks);
    const int grid = round8(std::max(1, std::min(cap, (n_slices + waves - 1) / waves)));
    int vcap = o.vec_blocks > 0 ? o.vec_blocks : grid;
    vcap = std::min(vcap, kMaxRedBlocks);
    // one 16-byte vector per thread at most
    const int vneed = (int)std::max<int64_t>(1, ((int64_t)n / 2 + block - 1) / block);
    return {grid, block, round8(std::max(1, std::min(vcap, vneed)))};
}

// FP64: the textbook three-kernel iteration; MIXED: the two-kernel iteration
// with merged reductions.
template <class V, class X>
void launch_iter(ocsl_solver* h, const CgArgs<V, X>& c, int n, CgLaunch L) {
    if constexpr (sizeof(X) == 4) {
        k_cgm_spmv<V, X><<<L.grid, L.block, 0, h->stream>>>(c);
        k_cgm_update<V, X><<<L.vgrid, L.block, 0, h->stream>>>(c, n);
    } else {
        k_cg_spmv<V, X><<<L.grid, L.block, 0, h->stream>>>(c);
        k_cg_update_r<V, X><<<L.vgrid, L.block, 0, h->stream>>>(c, n);
        k_cg_update_xp<V, X><<<L.vgrid, L.block, 0, h->stream>>>(c, n);
    }
}

// `chunk` iterations, replayed as one hipGraph when use_graph.  Kernel arguments are fixed per handle: every buffer
// involved is allocated once at the handle's size (work_buf drops the cached
// graphs if one is reallocated).
template <class V, class X>
void run_chunk(ocsl_solver* h, const CgArgs<V, X>& c, int n, int chunk, bool use_graph, int kind,
               CgLaunch L) {
    if (!use_graph) {
        for (int k = 0; k < chunk; ++k) launch_iter(h, c, n, L);
        OCSL_HIP(hipGetLastError());
        return;
    }
    auto key = std::make_pair(kind + 4 * L.grid + (1 << 13) * L.vgrid + (1 << 25) * (L.block / kWave), chunk);
    auto it = h->graphs.find(key);
    if (it == h->graphs.end()) {
        hipGraph_t gr;
        OCSL_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < chunk; ++k) launch_iter(h, c, n, L);
        OCSL_HIP(hipStreamEndCapture(h->stream, &gr));
        hipGraphExec_t ex;
        OCSL_HIP(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
        OCSL_HIP(hipGraphDestroy(gr));
        it = h->graphs.emplace(key, ex).first;
    }
    OCSL_HIP(hipGraphLaunch(it->second, h->stream));
}

void read_scalars(ocsl_solver* h) {
    OCSL_HIP(hipMemcpyAsync(h->S_host, h->S.p, sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
    sync(h);
}

}  // namespace

void cg_tolerance(ocsl_solver* h, bool mixed, const ocsl_options& o, ocsl_report* rep) {
    SellP& s = sellp_for(h, o.reorder != 0);
    ensure_sellp_values(h, s, true, mixed);
    const int n = s.n;
    const int* perm = s.identity ? nullptr : s.perm.p;
    Work& w = h->work;
    // b, x in FP64 (d5, d6); r, p, q in the Krylov precision
    double* b = work_buf(h, w.d[5], n);
    double* x = work_buf(h, w.d[6], n);
    CgArgs<double, double> c64{};
    CgArgs<float, float> c32{};
    if (mixed) {
        c32 = {s.view<float>(), work_buf(h, w.f[0], n), work_buf(h, w.f[1], n), work_buf(h, w.f[2], n), x,
               h->S.p, nullptr, h->red()};
    } else {
        c64 = {s.view<double>(), work_buf(h, w.d[7], n), work_buf(h, w.d[8], n), work_buf(h, w.d[9], n), x,
               h->S.p, nullptr, h->red()};
    }
    work_buf(h, h->hist, (size_t)o.max_iter + 2);
    c32.hist = c64.hist = h->hist.p;

    Scalars init{};
    init.tol2 = o.tol * o.tol;
    init.max_iter = o.max_iter;
    init.replace_ratio = mixed ? o.replace_ratio : 0.0;
    init.hist_cap = o.max_iter + 2;
    OCSL_HIP(hipMemcpyAsync(h->S.p, &init, sizeof(Scalars), hipMemcpyHostToDevice, h->stream));

    const int gv = grid_for(n);
    k_permute_in<double><<<gv, kBlock, 0, h->stream>>>(n, perm, h->b_dev, b);
    k_permute_in<double><<<gv, kBlock, 0, h->stream>>>(n, perm, h->x_own.p, x);
    OCSL_HIP(hipGetLastError());

    OCSL_HIP(hipEventRecord(h->ev0, h->stream));
    // the iteration kernels re-add the partial sums of one another and of the
    // residual kernel: every launch of the solve uses this geometry
    const CgLaunch L = cg_launch(s.n_slices, n, o);
    c32.g_spmv = c64.g_spmv = L.grid;
    c32.g_vec = c64.g_vec = L.vgrid;
    auto residual = [&](int first) {
        if (mixed)
            k_cg_residual<true><<<L.grid, L.block, 0, h->stream>>>(s.view<double>(), b, x, nullptr, c32.r,
                                                                  nullptr, c32.p, h->red(), first);
        else
            k_cg_residual<false><<<L.grid, L.block, 0, h->stream>>>(s.view<double>(), b, x, c64.r, nullptr,
                                                                   c64.p, nullptr, h->red(), first);
        k_finalize<2, FinCgResidual><<<1, kBlock, 0, h->stream>>>(
            h->red(), L.grid, FinCgResidual{h->S.p, h->hist.p, first, L.grid}, kSlotRR);
        OCSL_HIP(hipGetLastError());
    };
    residual(1);
    read_scalars(h);
    int replacements = 0;
    const int kind = (mixed ? 1 : 0) + (s.identity ? 0 : 2);
    while (!h->S_host->done) {
        const int left = o.max_iter - h->S_host->iter;
        const bool graph = o.use_graph && left >= o.check_every;
        const int chunk = graph ? o.check_every : std::max(1, std::min(o.check_every, left));
        if (mixed) run_chunk(h, c32, n, chunk, graph, kind, L);
        else run_chunk(h, c64, n, chunk, graph, kind, L);
        read_scalars(h);
        const Scalars& S = *h->S_host;
        if (S.nonfinite || S.breakdown) break;
        if (mixed && (S.need_update || S.done)) {
            residual(0);
            ++replacements;
            read_scalars(h);
        }
        if (h->S_host->iter >= o.max_iter) break;
    }
    OCSL_HIP(hipEventRecord(h->ev1, h->stream));
    k_permute_out<double><<<gv, kBlock, 0, h->stream>>>(n, perm, x, h->x_own.p);
    OCSL_HIP(hipGetLastError());
    read_scalars(h);
    float ms = 0.f;
    OCSL_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    const Scalars S = *h->S_host;
    const int iters = S.iter;
    h->hist_len = std::min(iters + 1, o.max_iter + 2);
    h->hist_host.assign(h->hist_len, 0.0);
    OCSL_HIP(hipMemcpy(h->hist_host.data(), h->hist.p, sizeof(double) * h->hist_len, hipMemcpyDeviceToHost));
    const double tr = true_relres_p(h, s, b, x);
    rep->iterations = iters;
    rep->replacements = replacements;
    rep->history_len = h->hist_len;
    rep->true_relres = tr;
    rep->residual = tr;
    rep->solve_ms = ms;
    int st = OCSL_OK;
    if (S.breakdown) st |= OCSL_WARN_BREAKDOWN;
    if (S.nonfinite || !std::isfinite(tr)) st |= OCSL_WARN_NONFINITE;
    const double scale = S.bb > 0.0 ? S.bb : 1.0;
    const bool converged = mixed ? (tr <= o.tol) : (S.rr_last <= S.tol2 * scale);
    if (!converged && st == OCSL_OK) st |= OCSL_WARN_MAXITER;
    rep->status = st;
    // algorithmic bytes of an iteration: the SpMV (off-diagonal values and
    // columns -- 2 or 4 bytes each --, the diagonal, slice pointers, p
    // gathered, q written) and the vector updates.  FP64 (three kernels):
    // r, q read, r written; x read/written, p, r read, p written.  MIXED (two
    // kernels): r read in the SpMV; x (FP64), r, q, p read, x, r, p written.
    const double off = (double)h->host.col.size() - n;
    const double sv = mixed ? 4.0 : 8.0;
    const double cb = s.compact ? 2.0 : 4.0;
    const double spmv_bytes = off * (sv + cb) + n * sv + 4.0 * (s.n_slices + 1) + 2.0 * n * sv;
    const double it_bytes = mixed ? spmv_bytes + n * sv + (16.0 + 3.0 * sv) * n + (16.0 + 2.0 * sv) * n
                                  : spmv_bytes + 3.0 * n * sv + (16.0 + 3.0 * sv) * n;
    const double repl_bytes = off * (8.0 + cb) + 8.0 * n + 4.0 * (s.n_slices + 1) + n * (8.0 + 8.0 + sv);
    rep->bytes = iters * it_bytes + replacements * repl_bytes;
}

template <class V, class X>
static double time_spmv(ocsl_solver* h, SellP& s, X* p, X* q, int reps, const ocsl_options& o) {
    const CgLaunch L = cg_launch(s.n_slices, s.n, o);
    CgArgs<V, X> c{s.view<V>(), p, p, q, nullptr, h->S.p, nullptr, h->red(), L.grid, L.vgrid};
    Scalars init{};
    init.max_iter = 1 << 30;
    init.tol2 = -1.0;
    init.rr_blocks = 1;
    OCSL_HIP(hipMemcpyAsync(h->S.p, &init, sizeof(Scalars), hipMemcpyHostToDevice, h->stream));
    auto launch = [&] {
        if constexpr (sizeof(X) == 4) k_cgm_spmv<V, X><<<L.grid, L.block, 0, h->stream>>>(c);
        else k_cg_spmv<V, X><<<L.grid, L.block, 0, h->stream>>>(c);
    };
    launch();  // warm
    OCSL_HIP(hipEventRecord(h->ev0, h->stream));
    for (int i = 0; i < reps; ++i) launch();
    OCSL_HIP(hipEventRecord(h->ev1, h->stream));
    OCSL_HIP(hipGetLastError());
    sync(h);
    float ms = 0.f;
    OCSL_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    return ms / reps;
}

}  // namespace ocsl

extern "C" ocsl_status ocsl_time_cg_spmv(ocsl_handle h, int precision, int reorder, int reps,
                                         double* avg_ms, double* bytes_per_launch) {
    using namespace ocsl;
    if (!h || !avg_ms || !bytes_per_launch || reps <= 0) return OCSL_ERR_ARG;
    try {
        ocsl_options o;
        ocsl_default_options(&o);
        const bool mixed = precision == OCSL_MIXED;
        SellP& s = sellp_for(h, reorder != 0);
        ensure_sellp_values(h, s, true, mixed);
        const int n = s.n;
        DBuf<double> p64, q64;
        DBuf<float> p32, q32;
        double ms;
        if (mixed) {
            p32.alloc(n);
            q32.alloc(n);
            k_fill<float><<<grid_for(n), kBlock, 0, h->stream>>>(n, 1.0f, p32.p);
            ms = time_spmv<float, float>(h, s, p32.p, q32.p, reps, o);
        } else {
            p64.alloc(n);
            q64.alloc(n);
            k_fill<double><<<grid_for(n), kBlock, 0, h->stream>>>(n, 1.0, p64.p);
            ms = time_spmv<double, double>(h, s, p64.p, q64.p, reps, o);
        }
        const double sv = mixed ? 4.0 : 8.0;
        *avg_ms = ms;
        // off-diagonal values and columns (2 bytes when packed), the diagonal
        // (no column index), slice pointers, p gathered once, q written; the
        // mixed kernel also reads r
        const double off = (double)h->host.col.size() - n;
        const double cb = s.compact ? 2.0 : 4.0;
        *bytes_per_launch = off * (sv + cb) + n * sv + 4.0 * (s.n_slices + 1) + 2.0 * n * sv + (mixed ? n * sv : 0.0);
        return OCSL_OK;
    } catch (const Error& e) {
        set_last_error(e.what);
        return e.status;
    }
}
