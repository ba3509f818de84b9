// C ABI of include/ocsl.h: handle life cycle, data upload, dispatch.
#include <cmath>
#include <cstdio>

#include "ocsl_handle.h"
#include "ocsl_kernels.h"

namespace ocsl {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

void sync(ocsl_solver* h) { OCSL_HIP(hipStreamSynchronize(h->stream)); }

void drop_graphs(ocsl_solver* h) {
    for (auto& kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
    h->graphs.clear();
}

// Real-entry CSR values gathered from the caller's array (input order).
void ensure_csr_values(ocsl_solver* h, bool need32) {
    OCSL_REQUIRE(h->have_vals, OCSL_ERR_STATE, "values not set");
    const int64_t nnz = (int64_t)h->host.col.size();
    if (!h->csr_val_valid) {
        h->csr_val.alloc(nnz);
        k_gather_vals<double><<<grid_for(nnz), kBlock, 0, h->stream>>>(nnz, h->src.p, h->val_in,
                                                                       h->csr_val.p);
        OCSL_HIP(hipGetLastError());
        h->csr_val_valid = true;
    }
    if (need32 && !h->csr_val32_valid) {
        h->csr_val32.alloc(nnz);
        k_gather_vals<float><<<grid_for(nnz), kBlock, 0, h->stream>>>(nnz, h->src.p, h->val_in,
                                                                      h->csr_val32.p);
        OCSL_HIP(hipGetLastError());
        h->csr_val32_valid = true;
    }
}

Sell& sell_for(ocsl_solver* h, bool reorder) {
    Sell& s = reorder ? h->sell_rcm : h->sell_id;
    if (s.n == 0) {
        drop_graphs(h);
        if (reorder) {
            std::vector<int> order = rcm_order(h->host);
            build_sell(h->host, &order, s, h->stream);
        } else {
            build_sell(h->host, nullptr, s, h->stream);
        }
    }
    return s;
}

void ensure_sell_values(ocsl_solver* h, Sell& s, bool need64, bool need32) {
    OCSL_REQUIRE(h->have_vals, OCSL_ERR_STATE, "values not set");
    uint64_t& e64 = (&s == &h->sell_rcm) ? h->sell_rcm_epoch64 : h->sell_id_epoch64;
    uint64_t& e32 = (&s == &h->sell_rcm) ? h->sell_rcm_epoch32 : h->sell_id_epoch32;
    const bool want64 = need64 && e64 != h->val_epoch;
    const bool want32 = need32 && e32 != h->val_epoch;
    if (!want64 && !want32) return;
    if (need64 && !s.val64.p) { drop_graphs(h); s.val64.alloc(s.total); }
    if (need32 && !s.val32.p) { drop_graphs(h); s.val32.alloc(s.total); }
    const int g = grid_for(s.total);
    if (want64 && want32) {
        k_gather_vals2<<<g, kBlock, 0, h->stream>>>(s.total, s.src.p, h->val_in, s.val64.p, s.val32.p);
    } else if (want64) {
        k_gather_vals<double><<<g, kBlock, 0, h->stream>>>(s.total, s.src.p, h->val_in, s.val64.p);
    } else {
        k_gather_vals<float><<<g, kBlock, 0, h->stream>>>(s.total, s.src.p, h->val_in, s.val32.p);
    }
    OCSL_HIP(hipGetLastError());
    if (want32) {
        // the FP32 copy is evaluated in difference form (sell_row_diff)
        OCSL_REQUIRE(s.val64.p && (want64 || e64 == h->val_epoch), OCSL_ERR_STATE,
                     "FP32 values need the FP64 copy");
        k_sell_rowsum_diag<<<grid_for((int64_t)s.n_slices * kWave), kBlock, 0, h->stream>>>(
            s.view<double>(), s.val32.p);
        OCSL_HIP(hipGetLastError());
    }
    if (want64) e64 = h->val_epoch;
    if (want32) e32 = h->val_epoch;
}

SellP& sellp_for(ocsl_solver* h, bool reorder, int rpl) {
    OCSL_REQUIRE(rpl == 1 || rpl == 2 || rpl == 4, OCSL_ERR_ARG, "rows_per_lane must be 0, 1, 2 or 4");
    SellP& s = h->sellp[reorder ? 1 : 0][rpl == 4 ? 2 : rpl - 1];
    if (s.n == 0) {
        drop_graphs(h);
        if (reorder) {
            std::vector<int> order = rcm_order(h->host);
            build_sellp(h->host, &order, rpl, s, h->stream);
        } else {
            build_sellp(h->host, nullptr, rpl, s, h->stream);
        }
    }
    return s;
}

// FP64 values (off-diagonals + diagonal) and, for the FP32 copy, the
// difference form: off-diagonals rounded, diagonal replaced by the FP64 row
// sum rounded once.
void ensure_sellp_values(ocsl_solver* h, SellP& s, bool need64, bool need32) {
    OCSL_REQUIRE(h->have_vals, OCSL_ERR_STATE, "values not set");
    uint64_t& e64 = s.epoch64;
    uint64_t& e32 = s.epoch32;
    const bool want64 = (need64 || need32) && e64 != h->val_epoch;
    const bool want32 = need32 && e32 != h->val_epoch;
    if (!want64 && !want32) return;
    if (!s.val64.p || !s.diag64.p) {
        drop_graphs(h);
        s.val64.alloc(s.total);
        s.diag64.alloc(s.padded());
    }
    if (need32 && (!s.val32.p || !s.diag32.p)) {
        drop_graphs(h);
        s.val32.alloc(s.total);
        s.diag32.alloc(s.padded());
    }
    const int g = grid_for(s.total);
    if (want64) {
        k_gather_vals<double><<<g, kBlock, 0, h->stream>>>(s.total, s.src.p, h->val_in, s.val64.p);
        k_gather_vals<double><<<grid_for(s.n), kBlock, 0, h->stream>>>(s.n, s.diag_src.p, h->val_in, s.diag64.p);
        e64 = h->val_epoch;
    }
    if (want32) {
        k_gather_vals<float><<<g, kBlock, 0, h->stream>>>(s.total, s.src.p, h->val_in, s.val32.p);
        const int gr = grid_for((int64_t)s.n_slices * kWave);
        with_rpl(s.rpl, [&](auto RC) {
            k_sellp_rowsum<decltype(RC)::value><<<gr, kBlock, 0, h->stream>>>(s.view<double>(), s.diag32.p);
        });
        e32 = h->val_epoch;
    }
    OCSL_HIP(hipGetLastError());
}

void device_dots(ocsl_solver* h, int n, int nr, const double* const* a, const double* const* b,
                 const int* kind, double* out) {
    const int g = red_blocks(kBlock * 4, n);
    if (nr == 1) {
        DotArgs<1> d{{a[0]}, {b[0]}, {kind[0]}};
        k_dots<1><<<g, kBlock, 0, h->stream>>>(n, d, h->red());
        k_finalize<1, FinStore<1>><<<1, kBlock, 0, h->stream>>>(h->red(), g, FinStore<1>{h->dscal.p});
    } else if (nr == 2) {
        DotArgs<2> d{{a[0], a[1]}, {b[0], b[1]}, {kind[0], kind[1]}};
        k_dots<2><<<g, kBlock, 0, h->stream>>>(n, d, h->red());
        k_finalize<2, FinStore<2>><<<1, kBlock, 0, h->stream>>>(h->red(), g, FinStore<2>{h->dscal.p});
    } else {
        DotArgs<3> d{{a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {kind[0], kind[1], kind[2]}};
        k_dots<3><<<g, kBlock, 0, h->stream>>>(n, d, h->red());
        k_finalize<3, FinStore<3>><<<1, kBlock, 0, h->stream>>>(h->red(), g, FinStore<3>{h->dscal.p});
    }
    OCSL_HIP(hipGetLastError());
    OCSL_HIP(hipMemcpyAsync(h->dscal_host, h->dscal.p, sizeof(double) * nr, hipMemcpyDeviceToHost,
                            h->stream));
    sync(h);
    for (int r = 0; r < nr; ++r) out[r] = h->dscal_host[r];
}

double device_dot(ocsl_solver* h, const double* a, const double* b, int n) {
    const double* aa[1] = {a};
    const double* bb[1] = {b};
    const int kind[1] = {0};
    double out;
    device_dots(h, n, 1, aa, bb, kind, &out);
    return out;
}

}  // namespace ocsl

using namespace ocsl;

namespace {

template <class F>
ocsl_status guarded(F&& f) {
    try {
        f();
        return OCSL_OK;
    } catch (const Error& e) {
        set_last_error(e.what);
        return e.status;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return OCSL_ERR_NOMEM;
    } catch (...) {
        set_last_error("unexpected exception");
        return OCSL_ERR_ARG;
    }
}

const double* stage(ocsl_solver* h, DBuf<double>& own, const double* p, int64_t count, int where) {
    OCSL_REQUIRE(p, OCSL_ERR_ARG, "null data pointer");
    OCSL_REQUIRE(where == OCSL_HOST || where == OCSL_DEVICE, OCSL_ERR_ARG, "bad ocsl_where");
    if (where == OCSL_DEVICE) return p;
    own.upload(p, (size_t)count, h->stream);
    return own.p;
}

}  // namespace

extern "C" {

void ocsl_default_options(ocsl_options* o) {
    if (!o) return;
    o->tol = 1e-8;
    o->max_iter = 20000;
    o->ilu_sweeps = 3;
    o->tri_mode = OCSL_TRI_EXACT;
    o->tri_sweeps = 3;
    o->tri_weight = 0.9;
    o->replace_ratio = 0.1;
    o->check_every = 32;
    o->reorder = 1;
    o->use_graph = 1;
    o->block_size = 0;
    o->grid_blocks = 0;
    o->vec_blocks = 0;
    o->rows_per_lane = 0;
    o->cg_kernels = 0;
    o->x_period = 0;
    o->bicgstab_order = OCSL_BICGSTAB_ORDER_INPUT;
}

int ocsl_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

const char* ocsl_last_error(void) { return g_last_error.c_str(); }

ocsl_status ocsl_create(ocsl_handle* out, int64_t n, const uint64_t* row_ptr,
                        const uint64_t* column_index) {
    if (!out) return OCSL_ERR_ARG;
    *out = nullptr;
    if (ocsl_device_count() <= 0) {
        set_last_error("no HIP device");
        return OCSL_ERR_NODEVICE;
    }
    ocsl_solver* h = new (std::nothrow) ocsl_solver();
    if (!h) return OCSL_ERR_NOMEM;
    ocsl_status st = guarded([&] {
        OCSL_HIP(hipGetDevice(&h->device));
        OCSL_HIP(hipDeviceGetAttribute(&h->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
        OCSL_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        OCSL_HIP(hipEventCreate(&h->ev0));
        OCSL_HIP(hipEventCreate(&h->ev1));
        analyse_structure(n, row_ptr, column_index, h->host);
        h->rp.upload(h->host.rp.data(), h->host.rp.size(), h->stream);
        h->col.upload(h->host.col.data(), h->host.col.size(), h->stream);
        h->src.upload(h->host.src.data(), h->host.src.size(), h->stream);
        h->diag.upload(h->host.diag.data(), h->host.diag.size(), h->stream);
        h->S.alloc(1);
        OCSL_HIP(hipHostMalloc((void**)&h->S_host, sizeof(Scalars), hipHostMallocDefault));
        h->red_partial.alloc(kMaxRed * kMaxRedBlocks);
        h->dscal.alloc(8);
        OCSL_HIP(hipHostMalloc((void**)&h->dscal_host, sizeof(double) * 8, hipHostMallocDefault));
        sync(h);
    });
    if (st != OCSL_OK) {
        ocsl_destroy(h);
        return st;
    }
    *out = h;
    return OCSL_OK;
}

ocsl_status ocsl_destroy(ocsl_handle h) {
    if (!h) return OCSL_ERR_ARG;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graphs(h);
    if (h->adj) ocsl_destroy(h->adj);
    if (h->col_h) ocsl_destroy(h->col_h);
    if (h->S_host) (void)hipHostFree(h->S_host);
    if (h->dscal_host) (void)hipHostFree(h->dscal_host);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    hipStream_t s = h->stream;
    delete h;  // frees device buffers
    if (s) (void)hipStreamDestroy(s);
    return OCSL_OK;
}

ocsl_status ocsl_get_size(ocsl_handle h, int64_t* n, int64_t* nnz) {
    if (!h || !n || !nnz) return OCSL_ERR_ARG;
    *n = h->host.n;
    *nnz = h->host.nnz_in;
    return OCSL_OK;
}

ocsl_status ocsl_set_values(ocsl_handle h, const double* value, ocsl_where where) {
    if (!h) return OCSL_ERR_ARG;
    return guarded([&] {
        h->val_in = stage(h, h->val_in_own, value, h->host.nnz_in, where);
        h->have_vals = true;
        h->val_epoch++;
        h->csr_val_valid = h->csr_val32_valid = false;
    });
}

ocsl_status ocsl_set_rhs(ocsl_handle h, const double* b, ocsl_where where) {
    if (!h) return OCSL_ERR_ARG;
    return guarded([&] {
        h->b_dev = stage(h, h->b_own, b, h->host.n, where);
        h->have_b = true;
    });
}

ocsl_status ocsl_set_x(ocsl_handle h, const double* x0, ocsl_where where) {
    if (!h) return OCSL_ERR_ARG;
    return guarded([&] {
        OCSL_REQUIRE(x0, OCSL_ERR_ARG, "null x0");
        h->x_own.alloc(h->host.n);
        OCSL_HIP(hipMemcpyAsync(h->x_own.p, x0, sizeof(double) * h->host.n,
                                where == OCSL_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                                h->stream));
        if (where == OCSL_HOST) sync(h);
    });
}

ocsl_status ocsl_get_x(ocsl_handle h, double* x, ocsl_where where) {
    if (!h || !x) return OCSL_ERR_ARG;
    return guarded([&] {
        OCSL_REQUIRE(h->x_own.p, OCSL_ERR_STATE, "x not set");
        OCSL_HIP(hipMemcpyAsync(x, h->x_own.p, sizeof(double) * h->host.n,
                                where == OCSL_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                                h->stream));
        sync(h);
    });
}

ocsl_status ocsl_solve(ocsl_handle h, int method, int mode, int precision, const ocsl_options* opt,
                       ocsl_report* rep) {
    if (!h) return OCSL_ERR_ARG;
    ocsl_report local{};
    ocsl_report* r = rep ? rep : &local;
    std::memset(r, 0, sizeof(*r));
    ocsl_options o;
    ocsl_default_options(&o);
    if (opt) o = *opt;
    ocsl_status st = guarded([&] {
        OCSL_REQUIRE(method == OCSL_CG || method == OCSL_ILU_BICGSTAB, OCSL_ERR_ARG, "bad method");
        OCSL_REQUIRE(mode == OCSL_MODE_REFERENCE || mode == OCSL_MODE_TOLERANCE, OCSL_ERR_ARG, "bad mode");
        OCSL_REQUIRE(precision == OCSL_FP64 || precision == OCSL_MIXED, OCSL_ERR_ARG, "bad precision");
        OCSL_REQUIRE(!(mode == OCSL_MODE_REFERENCE && precision == OCSL_MIXED), OCSL_ERR_ARG,
                     "REFERENCE mode is FP64 only");
        OCSL_REQUIRE(o.max_iter > 0 && o.check_every > 0 && o.tol >= 0.0, OCSL_ERR_ARG, "bad options");
        OCSL_REQUIRE(h->have_vals && h->have_b, OCSL_ERR_STATE, "values or rhs not set");
        if (!h->x_own.p) {
            h->x_own.alloc(h->host.n);
            OCSL_HIP(hipMemsetAsync(h->x_own.p, 0, sizeof(double) * h->host.n, h->stream));
        }
        const bool mixed = precision == OCSL_MIXED;
        if (method == OCSL_CG) {
            if (mode == OCSL_MODE_REFERENCE) cg_reference(h, r);
            else cg_tolerance(h, mixed, o, r);
        } else {
            if (mode == OCSL_MODE_REFERENCE) bicgstab_reference(h, r);
            else bicgstab_tolerance(h, mixed, o, r);
        }
    });
    if (st != OCSL_OK) return st;
    return (ocsl_status)r->status;
}

ocsl_status ocsl_solve_multi(ocsl_handle h, int method, int precision, int k, const double* B,
                             double* X, ocsl_where where, const ocsl_options* opt,
                             ocsl_report* reps) {
    if (!h || !B || !X || k < 1 || k > 4) return OCSL_ERR_ARG;
    std::vector<ocsl_report> local(k);
    ocsl_report* r = reps ? reps : local.data();
    std::memset(r, 0, sizeof(ocsl_report) * k);
    ocsl_options o;
    ocsl_default_options(&o);
    if (opt) o = *opt;
    ocsl_status st = guarded([&] {
        OCSL_REQUIRE(method == OCSL_ILU_BICGSTAB, OCSL_ERR_ARG, "multi-RHS solves use ILU_BICGSTAB");
        OCSL_REQUIRE(precision == OCSL_FP64 || precision == OCSL_MIXED, OCSL_ERR_ARG, "bad precision");
        OCSL_REQUIRE(o.max_iter > 0 && o.check_every > 0 && o.tol >= 0.0, OCSL_ERR_ARG, "bad options");
        OCSL_REQUIRE(h->have_vals, OCSL_ERR_STATE, "values not set");
        const size_t cnt = (size_t)h->host.n * k;
        DBuf<double> bd, xd;
        const double* Bd = B;
        double* Xd = X;
        if (where == OCSL_HOST) {
            bd.upload(B, cnt, h->stream);
            xd.upload(X, cnt, h->stream);
            Bd = bd.p;
            Xd = xd.p;
        }
        bicgstab_multi(h, precision == OCSL_MIXED, o, k, Bd, Xd, r);
        if (where == OCSL_HOST) {
            OCSL_HIP(hipMemcpyAsync(X, Xd, sizeof(double) * cnt, hipMemcpyDeviceToHost, h->stream));
            sync(h);
        }
    });
    if (st != OCSL_OK) return st;
    int worst = 0;
    for (int c = 0; c < k; ++c) worst |= r[c].status;
    return (ocsl_status)worst;
}

ocsl_status ocsl_CG_solve(ocsl_handle h, double* res) {
    ocsl_report rep;
    ocsl_status st = ocsl_solve(h, OCSL_CG, OCSL_MODE_REFERENCE, OCSL_FP64, nullptr, &rep);
    if (st >= 0 && res) *res = rep.residual;
    return st;
}

ocsl_status ocsl_ILU_BICGSTAB_solve(ocsl_handle h, double* res) {
    ocsl_report rep;
    ocsl_status st = ocsl_solve(h, OCSL_ILU_BICGSTAB, OCSL_MODE_REFERENCE, OCSL_FP64, nullptr, &rep);
    if (st >= 0 && res) *res = rep.residual;
    return st;
}

ocsl_status ocsl_ILU_BICGSTAB_solve_multi(ocsl_handle h, int k, const double* B, double* X, ocsl_where where,
                                          double* res) {
    if (!h || !B || !X || k < 1 || k > 4) return OCSL_ERR_ARG;
    std::vector<ocsl_report> r(k);
    std::memset(r.data(), 0, sizeof(ocsl_report) * k);
    ocsl_status st = guarded([&] {
        OCSL_REQUIRE(h->have_vals, OCSL_ERR_STATE, "values not set");
        const size_t cnt = (size_t)h->host.n * k;
        DBuf<double> bd, xd;
        const double* Bd = B;
        double* Xd = X;
        if (where == OCSL_HOST) {
            bd.upload(B, cnt, h->stream);
            xd.upload(X, cnt, h->stream);
            Bd = bd.p;
            Xd = xd.p;
        }
        bicgstab_reference_multi(h, k, Bd, Xd, r.data());
        if (where == OCSL_HOST) OCSL_HIP(hipMemcpyAsync(X, Xd, sizeof(double) * cnt, hipMemcpyDeviceToHost, h->stream));
        sync(h);
    });
    if (st != OCSL_OK) return st;
    int worst = 0;
    for (int c = 0; c < k; ++c) {
        worst |= r[c].status;
        if (res) res[c] = r[c].residual;
    }
    return (ocsl_status)worst;
}

ocsl_status ocsl_get_history(ocsl_handle h, double* out, int cap, int* len) {
    if (!h || !len || (cap > 0 && !out)) return OCSL_ERR_ARG;
    *len = h->hist_len;
    const int m = std::min(cap, h->hist_len);
    for (int i = 0; i < m; ++i) out[i] = h->hist_host[i];
    return OCSL_OK;
}

static ocsl_status reduce_api(ocsl_handle h, const double* a, const double* b, int kind,
                              ocsl_where where, double* out) {
    if (!h || !out) return OCSL_ERR_ARG;
    return guarded([&] {
        DBuf<double> ta, tb;
        const double* da = stage(h, ta, a, h->host.n, where);
        const double* db = b ? stage(h, tb, b, h->host.n, where) : da;
        const double* aa[1] = {da};
        const double* bb[1] = {db};
        const int k[1] = {kind};
        device_dots(h, h->host.n, 1, aa, bb, k, out);
    });
}

ocsl_status ocsl_ABSsum(ocsl_handle h, const double* v, ocsl_where where, double* out) {
    return reduce_api(h, v, nullptr, 2, where, out);
}

ocsl_status ocsl_sum(ocsl_handle h, const double* v, ocsl_where where, double* out) {
    return reduce_api(h, v, nullptr, 1, where, out);
}

ocsl_status ocsl_norm(ocsl_handle h, const double* a, const double* b, ocsl_where where,
                      double* out) {
    if (!b) return OCSL_ERR_ARG;
    return reduce_api(h, a, b, 0, where, out);
}

ocsl_status ocsl_gemv(ocsl_handle h, double a, const double* x, double b, const double* z,
                      double* out, ocsl_where where) {
    if (!h || !out) return OCSL_ERR_ARG;
    return guarded([&] {
        const int n = h->host.n;
        DBuf<double> tx, tz, to;
        const double* dx = stage(h, tx, x, n, where);
        const double* dz = z ? stage(h, tz, z, n, where) : nullptr;
        double* dout = out;
        if (where == OCSL_HOST) {
            to.alloc(n);
            dout = to.p;
        }
        Sell& s = sell_for(h, false);
        ensure_sell_values(h, s, true, false);
        k_sell_axpby<double><<<grid_for((int64_t)s.n_slices * kWave), kBlock, 0, h->stream>>>(
            s.view<double>(), a, nullptr, dx, b, dz, dout);
        OCSL_HIP(hipGetLastError());
        if (where == OCSL_HOST)
            OCSL_HIP(hipMemcpyAsync(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
        sync(h);
    });
}

ocsl_status ocsl_ILU_decompose(ocsl_handle h, int sweeps, double* m_out, ocsl_where where) {
    if (!h || !m_out || sweeps < 0) return OCSL_ERR_ARG;
    return guarded([&] {
        ensure_csr_values(h, false);
        ilu_decompose_device(h, sweeps, false);
        const int64_t nnz_in = h->host.nnz_in;
        DBuf<double> full;
        full.alloc(nnz_in);
        OCSL_HIP(hipMemsetAsync(full.p, 0, sizeof(double) * nnz_in, h->stream));
        k_permute_out<double><<<grid_for(h->host.col.size()), kBlock, 0, h->stream>>>(
            (int)h->host.col.size(), h->src.p, h->ilu_m.p, full.p);
        OCSL_HIP(hipGetLastError());
        OCSL_HIP(hipMemcpyAsync(m_out, full.p, sizeof(double) * nnz_in,
                                where == OCSL_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                                h->stream));
        sync(h);
    });
}

void* ocsl_get_stream(ocsl_handle h) { return h ? (void*)h->stream : nullptr; }

ocsl_status ocsl_colour_order(ocsl_handle h, int* order, int* n_colours) {
    if (!h) return OCSL_ERR_ARG;
    return guarded([&] { colour_order_get(h, order, n_colours); });
}

ocsl_status ocsl_set_reference_graphs(ocsl_handle h, int on) {
    if (!h || (on != 0 && on != 1)) return OCSL_ERR_ARG;
    return guarded([&] {
        sync(h);
        h->ref_graphs = on != 0;
    });
}

ocsl_status ocsl_set_tri_form(ocsl_handle h, int form) {
    if (!h || form < 0 || form > 3) return OCSL_ERR_ARG;
    return guarded([&] {
        sync(h);
        if (form != h->tri_form) drop_graphs(h);  // captured solves hold the form they were captured with
        h->tri_form = form;
    });
}

}  // extern "C"
