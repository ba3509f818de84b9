// Adjoint of a solve x = A^{-1} b for a scalar loss L(x).
//
// Given g = dL/dx, the transpose system A^T lambda = g is solved on the
// transposed pattern, held by an internal second handle whose values are
// gathered from the caller's value array through a slot map (no host copy
// of the values).  Then dL/db = lambda and dL/dA_ij = -lambda_i x_j on every
// stored entry: a sampled outer product over the CSR pattern, written in the
// caller's slot layout.
//
// The outer-product kernel gives each workgroup kBlock consecutive rows.
// Their slots form one contiguous range; the workgroup stages the rows'
// offsets in LDS and walks the range with consecutive threads on consecutive
// slots (coalesced column reads and gradient writes), finding each slot's
// row by a binary search in LDS.
#include <cstring>

#include "ocsl_handle.h"
#include "ocsl_kernels.h"

namespace ocsl {

template <bool kMap>
__global__ void __launch_bounds__(kBlock)
    k_sampled_outer(int n, const int* __restrict__ rp, const int* __restrict__ col, const int* __restrict__ src,
                    double alpha, const double* __restrict__ u, const double* __restrict__ v,
                    double* __restrict__ out) {
    __shared__ int srp[kBlock + 1];
    const int r0 = blockIdx.x * kBlock;
    const int rows = min(kBlock, n - r0);
    for (int t = threadIdx.x; t <= rows; t += kBlock) srp[t] = rp[r0 + t];
    __syncthreads();
    const int s0 = srp[0], s1 = srp[rows];
    for (int s = s0 + (int)threadIdx.x; s < s1; s += kBlock) {
        int lo = 0, hi = rows - 1;  // largest r with srp[r] <= s
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (srp[mid] <= s) lo = mid;
            else hi = mid - 1;
        }
        const double val = alpha * u[r0 + lo] * v[col[s]];
        out[kMap ? src[s] : s] = val;
    }
}

namespace {

// Transposed structure of the real entries: row j of A^T holds (j, i) for
// every stored A_ij, in ascending i.  map[t] = A's caller slot of entry t.
ocsl_solver* transposed(ocsl_solver* h) {
    if (h->adj) return h->adj;
    const HostCSR& a = h->host;
    const int n = a.n;
    const size_t nnz = a.col.size();
    std::vector<uint64_t> rpt(n + 1, 0), colt(nnz);
    std::vector<int> map(nnz);
    for (size_t p = 0; p < nnz; ++p) rpt[a.col[p] + 1]++;
    for (int i = 0; i < n; ++i) rpt[i + 1] += rpt[i];
    std::vector<uint64_t> fill(rpt.begin(), rpt.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int p = a.rp[i]; p < a.rp[i + 1]; ++p) {
            const uint64_t t = fill[a.col[p]]++;
            colt[t] = (uint64_t)i;
            map[t] = a.src[p];
        }
    ocsl_handle t = nullptr;
    const ocsl_status st = ocsl_create(&t, n, rpt.data(), colt.data());
    if (st != OCSL_OK) throw Error{st, std::string("transpose handle: ") + ocsl_last_error()};
    h->adj = t;
    h->adj_map.upload(map.data(), map.size(), h->stream);
    h->adj_epoch = ~0ull;
    sync(h);
    return t;
}

void sampled_outer(ocsl_solver* h, double alpha, const double* u, const double* v, double* out) {
    const HostCSR& a = h->host;
    const int n = a.n;
    const bool mapped = (int64_t)a.col.size() != a.nnz_in;  // sentinel slots present
    if (mapped) k_fill<double><<<grid_for(a.nnz_in), kBlock, 0, h->stream>>>((int)a.nnz_in, 0.0, out);
    const int grid = (n + kBlock - 1) / kBlock;
    if (grid > 0) {
        if (mapped)
            k_sampled_outer<true><<<grid, kBlock, 0, h->stream>>>(n, h->rp.p, h->col.p, h->src.p, alpha, u, v, out);
        else
            k_sampled_outer<false><<<grid, kBlock, 0, h->stream>>>(n, h->rp.p, h->col.p, nullptr, alpha, u, v, out);
    }
    OCSL_HIP(hipGetLastError());
}

double outer_bytes(const ocsl_solver* h) {
    // values written and columns read per stored entry, row offsets, u and v once
    const double nnz = (double)h->host.col.size(), n = h->host.n;
    return nnz * (8.0 + 4.0) + 4.0 * (n + 1) + 2.0 * n * 8.0;
}

template <class F>
ocsl_status guarded_a(F&& f) {
    try {
        f();
        return OCSL_OK;
    } catch (const Error& e) {
        set_last_error(e.what);
        return e.status;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return OCSL_ERR_NOMEM;
    }
}

}  // namespace
}  // namespace ocsl

using namespace ocsl;

extern "C" {

ocsl_status ocsl_adjoint(ocsl_handle h, int method, int precision, const double* g, double* dLdb, double* dLdA,
                         ocsl_where where, const ocsl_options* opt, ocsl_report* rep) {
    if (!h || !g) return OCSL_ERR_ARG;
    ocsl_report local{};
    ocsl_report* r = rep ? rep : &local;
    std::memset(r, 0, sizeof(*r));
    ocsl_options o;
    ocsl_default_options(&o);
    if (opt) o = *opt;
    ocsl_status st = guarded_a([&] {
        OCSL_REQUIRE(method == OCSL_CG || method == OCSL_ILU_BICGSTAB, OCSL_ERR_ARG, "bad method");
        OCSL_REQUIRE(precision == OCSL_FP64 || precision == OCSL_MIXED, OCSL_ERR_ARG, "bad precision");
        OCSL_REQUIRE(where == OCSL_HOST || where == OCSL_DEVICE, OCSL_ERR_ARG, "bad ocsl_where");
        OCSL_REQUIRE(o.max_iter > 0 && o.check_every > 0 && o.tol >= 0.0, OCSL_ERR_ARG, "bad options");
        OCSL_REQUIRE(h->have_vals, OCSL_ERR_STATE, "values not set");
        OCSL_REQUIRE(h->x_own.p, OCSL_ERR_STATE, "no solution to differentiate (solve or set x first)");
        ocsl_solver* t = transposed(h);
        const int n = h->host.n;
        const int64_t nnz = (int64_t)h->host.col.size();
        sync(h);  // the caller's values and x are complete
        if (h->adj_epoch != h->val_epoch) {
            t->val_in_own.alloc(nnz);
            k_gather_vals<double><<<grid_for(nnz), kBlock, 0, t->stream>>>(nnz, h->adj_map.p, h->val_in,
                                                                          t->val_in_own.p);
            OCSL_HIP(hipGetLastError());
            t->val_in = t->val_in_own.p;
            t->have_vals = true;
            t->val_epoch++;
            t->csr_val_valid = t->csr_val32_valid = false;
            h->adj_epoch = h->val_epoch;
        }
        if (where == OCSL_DEVICE) {
            t->b_dev = g;
        } else {
            t->b_own.upload(g, n, t->stream);
            t->b_dev = t->b_own.p;
        }
        t->have_b = true;
        t->x_own.alloc(n);
        OCSL_HIP(hipMemsetAsync(t->x_own.p, 0, sizeof(double) * n, t->stream));
        if (method == OCSL_CG) cg_tolerance(t, precision == OCSL_MIXED, o, r);
        else bicgstab_tolerance(t, precision == OCSL_MIXED, o, r);
        sync(t);
        const auto kind = where == OCSL_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
        if (dLdb) OCSL_HIP(hipMemcpyAsync(dLdb, t->x_own.p, sizeof(double) * n, kind, h->stream));
        if (dLdA) {
            double* dst = dLdA;
            if (where == OCSL_HOST) {
                h->grad_own.alloc(h->host.nnz_in);
                dst = h->grad_own.p;
            }
            sampled_outer(h, -1.0, t->x_own.p, h->x_own.p, dst);
            if (where == OCSL_HOST)
                OCSL_HIP(hipMemcpyAsync(dLdA, dst, sizeof(double) * h->host.nnz_in, hipMemcpyDeviceToHost,
                                        h->stream));
        }
        sync(h);
    });
    if (st != OCSL_OK) return st;
    return (ocsl_status)r->status;
}

ocsl_status ocsl_sampled_outer(ocsl_handle h, double alpha, const double* u, const double* v, double* out,
                               ocsl_where where) {
    if (!h || !u || !v || !out) return OCSL_ERR_ARG;
    return guarded_a([&] {
        OCSL_REQUIRE(where == OCSL_HOST || where == OCSL_DEVICE, OCSL_ERR_ARG, "bad ocsl_where");
        const int n = h->host.n;
        if (where == OCSL_DEVICE) {
            sampled_outer(h, alpha, u, v, out);
        } else {
            DBuf<double> du, dv, dout;
            du.upload(u, n, h->stream);
            dv.upload(v, n, h->stream);
            dout.alloc(h->host.nnz_in);
            sampled_outer(h, alpha, du.p, dv.p, dout.p);
            OCSL_HIP(hipMemcpyAsync(out, dout.p, sizeof(double) * h->host.nnz_in, hipMemcpyDeviceToHost,
                                    h->stream));
        }
        sync(h);
    });
}

ocsl_status ocsl_time_sampled_outer(ocsl_handle h, int reps, double* avg_ms, double* bytes_per_launch) {
    if (!h || !avg_ms || !bytes_per_launch || reps <= 0) return OCSL_ERR_ARG;
    return guarded_a([&] {
        const int n = h->host.n;
        DBuf<double> u, v, out;
        u.alloc(n);
        v.alloc(n);
        out.alloc(h->host.nnz_in);
        k_fill<double><<<grid_for(n), kBlock, 0, h->stream>>>(n, 0.5, u.p);
        k_fill<double><<<grid_for(n), kBlock, 0, h->stream>>>(n, 2.0, v.p);
        sampled_outer(h, -1.0, u.p, v.p, out.p);  // warm
        OCSL_HIP(hipEventRecord(h->ev0, h->stream));
        for (int i = 0; i < reps; ++i) sampled_outer(h, -1.0, u.p, v.p, out.p);
        OCSL_HIP(hipEventRecord(h->ev1, h->stream));
        sync(h);
        float ms = 0.f;
        OCSL_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *avg_ms = ms / reps;
        *bytes_per_launch = outer_bytes(h);
    });
}

}  // extern "C"
