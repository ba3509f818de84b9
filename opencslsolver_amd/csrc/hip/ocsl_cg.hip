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

static double true_relres(ocsl_solver* h, const Sell& s, const double* b, const double* x) {
    k_true_residual<double><<<red_blocks(kBlock * 4, s.n), kBlock, 0, h->stream>>>(
        s.view<double>(), b, x, h->red(), h->dscal.p);
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
    for (int i = 0; i < 5; ++i) w.d[i].alloc(n);
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

struct CgBuffers {
    double *b, *x64, *r64, *p64, *q64;
    float *r32, *p32, *q32;
};

template <class V, class X>
void launch_cg_iter(ocsl_solver* h, const CgArgs<V, X>& c, int n, int g_spmv, int g_vec) {
    k_cg_spmv<V, X><<<g_spmv, kBlock, 0, h->stream>>>(c);
    k_cg_update_r<V, X><<<g_vec, kBlock, 0, h->stream>>>(c, n);
    k_cg_update_xp<V, X><<<g_vec, kBlock, 0, h->stream>>>(c, n);
}

int round8(int g) { return g >= kXcds ? (g / kXcds) * kXcds : g; }

template <class V, class X>
void run_chunk(ocsl_solver* h, const CgArgs<V, X>& c, int n, int chunk, bool use_graph, int kind) {
    const int g_spmv = round8(std::min(kMaxRedBlocks, (c.A.n_slices + kWavesPerBlock - 1) / kWavesPerBlock));
    const int g_vec = round8(red_blocks(kBlock * 4, n));
    if (!use_graph) {
        for (int k = 0; k < chunk; ++k) launch_cg_iter(h, c, n, g_spmv, g_vec);
        OCSL_HIP(hipGetLastError());
        return;
    }
    // Graphs are keyed by (kind, chunk); kernel arguments are fixed per handle
    // because every buffer involved is allocated once at the handle's size.
    auto key = std::make_pair(kind, chunk);
    auto it = h->graphs.find(key);
    if (it == h->graphs.end()) {
        hipGraph_t g;
        OCSL_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < chunk; ++k) launch_cg_iter(h, c, n, g_spmv, g_vec);
        OCSL_HIP(hipStreamEndCapture(h->stream, &g));
        hipGraphExec_t ex;
        OCSL_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
        OCSL_HIP(hipGraphDestroy(g));
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
    Sell& s = sell_for(h, o.reorder != 0);
    ensure_sell_values(h, s, true, mixed);
    const int n = s.n;
    const int* perm = s.identity ? nullptr : s.perm.p;
    Work& w = h->work;
    // FP64: d[5..9]; MIXED: d[5..6] + f[0..2]
    for (int i = 5; i < 10; ++i) w.d[i].alloc(n);
    if (mixed)
        for (int i = 0; i < 3; ++i) w.f[i].alloc(n);
    CgBuffers B{w.d[5].p, w.d[6].p, w.d[7].p, w.d[8].p, w.d[9].p,
                mixed ? w.f[0].p : nullptr, mixed ? w.f[1].p : nullptr,
                mixed ? w.f[2].p : nullptr};
    const double* hist_before = h->hist.p;
    h->hist.alloc((size_t)o.max_iter + 2);
    if (h->hist.p != hist_before) drop_graphs(h);

    Scalars init{};
    init.tol2 = o.tol * o.tol;
    init.max_iter = o.max_iter;
    init.replace_ratio = mixed ? o.replace_ratio : 0.0;
    init.hist_cap = o.max_iter + 2;
    OCSL_HIP(hipMemcpyAsync(h->S.p, &init, sizeof(Scalars), hipMemcpyHostToDevice, h->stream));

    const int gv = grid_for(n);
    k_permute_in<double><<<gv, kBlock, 0, h->stream>>>(n, perm, h->b_dev, B.b);
    k_permute_in<double><<<gv, kBlock, 0, h->stream>>>(n, perm, h->x_own.p, B.x64);
    OCSL_HIP(hipGetLastError());

    OCSL_HIP(hipEventRecord(h->ev0, h->stream));
    const int g_res = round8(std::min(kMaxRedBlocks, (s.n_slices + kWavesPerBlock - 1) / kWavesPerBlock));
    auto residual = [&](int first) {
        if (mixed)
            k_cg_residual<true><<<g_res, kBlock, 0, h->stream>>>(s.view<double>(), B.b, B.x64, nullptr,
                                                                B.r32, B.p32, nullptr, h->S.p,
                                                                h->hist.p, h->red(), first);
        else
            k_cg_residual<false><<<g_res, kBlock, 0, h->stream>>>(s.view<double>(), B.b, B.x64, B.r64,
                                                                 nullptr, nullptr, B.p64, h->S.p,
                                                                 h->hist.p, h->red(), first);
        OCSL_HIP(hipGetLastError());
    };
    residual(1);
    read_scalars(h);
    int replacements = 0;
    const int kind = (mixed ? 1 : 0) + (s.identity ? 0 : 2);
    if (!h->S_host->done) {
        while (true) {
            const int left = o.max_iter - h->S_host->iter;
            const int chunk = std::min(o.check_every, std::max(left, 1));
            if (mixed) {
                CgArgs<float, float> c{s.view<float>(), B.r32, B.p32, B.q32, B.x64, h->S.p, h->hist.p, h->red()};
                run_chunk(h, c, n, chunk, o.use_graph && chunk == o.check_every, kind);
            } else {
                CgArgs<double, double> c{s.view<double>(), B.r64, B.p64, B.q64, B.x64, h->S.p, h->hist.p, h->red()};
                run_chunk(h, c, n, chunk, o.use_graph && chunk == o.check_every, kind);
            }
            read_scalars(h);
            const Scalars& S = *h->S_host;
            if (S.nonfinite || S.breakdown) break;
            if (mixed && (S.need_update || S.done)) {
                residual(0);
                ++replacements;
                read_scalars(h);
                if (h->S_host->done || h->S_host->nonfinite) break;
            } else if (!mixed && S.done) {
                break;
            }
            if (h->S_host->iter >= o.max_iter) break;
        }
    }
    OCSL_HIP(hipEventRecord(h->ev1, h->stream));
    k_permute_out<double><<<gv, kBlock, 0, h->stream>>>(n, perm, B.x64, h->x_own.p);
    OCSL_HIP(hipGetLastError());
    read_scalars(h);
    float ms = 0.f;
    OCSL_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    const Scalars S = *h->S_host;
    const int iters = S.iter;
    h->hist_len = std::min(iters + 1, o.max_iter + 2);
    h->hist_host.assign(h->hist_len, 0.0);
    OCSL_HIP(hipMemcpy(h->hist_host.data(), h->hist.p, sizeof(double) * h->hist_len, hipMemcpyDeviceToHost));
    const double tr = true_relres(h, s, B.b, B.x64);
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
    const bool converged = mixed ? (tr <= o.tol) : (S.rr <= S.tol2 * scale);
    if (!converged && st == OCSL_OK) st |= OCSL_WARN_MAXITER;
    rep->status = st;
    // algorithmic bytes of the iterations: SpMV (values + columns + slice
    // pointers + p gather + q write) + r update (r, q read, r write) + x/p
    // update (x read/write in FP64, p, r read, p write)
    const double nnz = (double)h->host.col.size();
    const double sv = mixed ? 4.0 : 8.0;
    const double it_bytes = nnz * (sv + 4.0) + 4.0 * (s.n_slices + 1) + 2.0 * n * sv +
                            3.0 * n * sv + (16.0 + 3.0 * sv) * n;
    // FP64 residual replacement: FP64 SpMV of x, b read, r32 write
    const double repl_bytes = nnz * 12.0 + 4.0 * (s.n_slices + 1) + n * (8.0 + 8.0 + 4.0);
    rep->bytes = iters * it_bytes + replacements * repl_bytes;
}

template <class V, class X>
static double time_spmv(ocsl_solver* h, Sell& s, X* p, X* q, int reps) {
    CgArgs<V, X> c{s.view<V>(), nullptr, p, q, nullptr, h->S.p, nullptr, h->red()};
    const int g = round8(std::min(kMaxRedBlocks, (s.n_slices + kWavesPerBlock - 1) / kWavesPerBlock));
    Scalars init{};
    init.rr = 1.0;
    OCSL_HIP(hipMemcpyAsync(h->S.p, &init, sizeof(Scalars), hipMemcpyHostToDevice, h->stream));
    k_cg_spmv<V, X><<<g, kBlock, 0, h->stream>>>(c);  // warm
    OCSL_HIP(hipEventRecord(h->ev0, h->stream));
    for (int i = 0; i < reps; ++i) k_cg_spmv<V, X><<<g, kBlock, 0, h->stream>>>(c);
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
        const bool mixed = precision == OCSL_MIXED;
        Sell& s = sell_for(h, reorder != 0);
        ensure_sell_values(h, s, true, mixed);
        const int n = s.n;
        DBuf<double> p64, q64;
        DBuf<float> p32, q32;
        double ms;
        if (mixed) {
            p32.alloc(n);
            q32.alloc(n);
            OCSL_HIP(hipMemsetAsync(p32.p, 0, sizeof(float) * n, h->stream));
            ms = time_spmv<float, float>(h, s, p32.p, q32.p, reps);
        } else {
            p64.alloc(n);
            q64.alloc(n);
            OCSL_HIP(hipMemsetAsync(p64.p, 0, sizeof(double) * n, h->stream));
            ms = time_spmv<double, double>(h, s, p64.p, q64.p, reps);
        }
        const double sv = mixed ? 4.0 : 8.0;
        *avg_ms = ms;
        *bytes_per_launch = (double)h->host.col.size() * (sv + 4.0) + 4.0 * (s.n_slices + 1) + 2.0 * n * sv;
        return OCSL_OK;
    } catch (const Error& e) {
        set_last_error(e.what);
        return e.status;
    }
}
