// Vector helpers of CMatrix_solver that take no matrix -- norm, the static
// sum and ABSsum, product and the two axpy forms (cmatrix_solver.cpp:233-335)
// -- on the current device without a solver handle.  The reference's sum and
// ABSsum are static members (cmatrix_solver.h:47-48), called as
// CMatrix_solver::ABSsum(b_vector) (cflow_solver.cpp:346), so these entry
// points own a stream and scratch of their own: one context per device,
// created at its first use and kept until the process ends; calls are
// serialised per device.
#include <map>
#include <mutex>

#include "ocsl_handle.h"
#include "ocsl_kernels.h"

namespace ocsl {
namespace {

struct VecCtx {
    hipStream_t stream = nullptr;
    DBuf<double> a, b, out;  // staging of host operands
    DBuf<double> partial;    // kMaxRed x kMaxRedBlocks partial sums
    DBuf<double> res;
    double* res_host = nullptr;
    std::mutex mu;
};

std::mutex g_ctx_mu;
std::map<int, VecCtx*> g_ctx;  // one per device, never freed

VecCtx& vec_ctx() {
    int dev = 0;
    OCSL_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    auto it = g_ctx.find(dev);
    if (it != g_ctx.end()) return *it->second;
    VecCtx* c = new VecCtx();
    try {
        OCSL_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->partial.alloc((size_t)kMaxRed * kMaxRedBlocks);
        c->res.alloc(1);
        OCSL_HIP(hipHostMalloc((void**)&c->res_host, sizeof(double), hipHostMallocDefault));
    } catch (...) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        throw;
    }
    g_ctx.emplace(dev, c);
    return *c;
}

const double* stage_in(VecCtx& c, DBuf<double>& buf, const double* p, int64_t n, int where) {
    if (where == OCSL_DEVICE) return p;
    buf.upload(p, (size_t)n, c.stream);
    return buf.p;
}

template <class F>
ocsl_status vec_guard(int64_t n, int where, F&& f) {
    if (n < 0 || n > 0x7fffffff || (where != OCSL_HOST && where != OCSL_DEVICE)) return OCSL_ERR_ARG;
    try {
        if (ocsl_device_count() <= 0) throw Error{OCSL_ERR_NODEVICE, "no HIP device"};
        VecCtx& c = vec_ctx();
        std::lock_guard<std::mutex> lk(c.mu);
        f(c);
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

// sum over i of a_i b_i (kind 0), a_i (1) or |a_i| (2), in a fixed order.
ocsl_status vec_reduce(int64_t n, const double* a, const double* b, int kind, int where, double* out) {
    if (!out || (n > 0 && !a) || (kind == 0 && n > 0 && !b)) return OCSL_ERR_ARG;
    return vec_guard(n, where, [&](VecCtx& c) {
        if (n == 0) {
            *out = 0.0;
            return;
        }
        const double* da = stage_in(c, c.a, a, n, where);
        const double* db = kind == 0 ? stage_in(c, c.b, b, n, where) : da;
        const int g = red_blocks(kBlock * 4, n);
        DotArgs<1> d{{da}, {db}, {kind}};
        const RedScratch red{c.partial.p};
        k_dots<1><<<g, kBlock, 0, c.stream>>>((int)n, d, red);
        k_finalize<1, FinStore<1>><<<1, kBlock, 0, c.stream>>>(red, g, FinStore<1>{c.res.p});
        OCSL_HIP(hipGetLastError());
        OCSL_HIP(hipMemcpyAsync(c.res_host, c.res.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
        OCSL_HIP(hipStreamSynchronize(c.stream));
        *out = *c.res_host;
    });
}

// out = a x + b y (y == nullptr: out = a x) with each product rounded, as the
// reference's loops compute it (cmatrix_solver.cpp:275-335); mul == 1:
// out = x * y (cmatrix_solver.cpp:265-272).
__global__ void k_vec_axpby(int n, double a, const double* __restrict__ x, double b, const double* __restrict__ y,
                            double* __restrict__ out, int mul) {
#pragma clang fp contract(off)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (mul) out[i] = x[i] * y[i];
        else if (y) out[i] = a * x[i] + b * y[i];
        else out[i] = a * x[i];
    }
}

ocsl_status vec_elementwise(int64_t n, double a, const double* x, double b, const double* y, double* out,
                            int where, int mul) {
    if (!out || (n > 0 && !x) || (mul && n > 0 && !y)) return OCSL_ERR_ARG;
    return vec_guard(n, where, [&](VecCtx& c) {
        if (n == 0) return;
        const double* dx = stage_in(c, c.a, x, n, where);
        const double* dy = y ? stage_in(c, c.b, y, n, where) : nullptr;
        double* dout = out;
        if (where == OCSL_HOST) {
            c.out.alloc((size_t)n);
            dout = c.out.p;
        }
        k_vec_axpby<<<grid_for(n), kBlock, 0, c.stream>>>((int)n, a, dx, b, dy, dout, mul);
        OCSL_HIP(hipGetLastError());
        if (where == OCSL_HOST)
            OCSL_HIP(hipMemcpyAsync(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
        OCSL_HIP(hipStreamSynchronize(c.stream));
    });
}

// ---------------------------------------------------------------- no allocation
// Partial sums of ocsl_vec_ABSsum_noalloc: one per reduction block (at most
// kMaxRedBlocks), in a file-scope device array instead of an allocation.
// Each device holds its own copy at its own address, looked up once per
// device with hipGetSymbolAddress; host code uses only those addresses.
constexpr int kNoAllocPartials = kMaxRedBlocks;  // 16 KiB
static __device__ __align__(256) double scratch[kNoAllocPartials];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

double* scratch_for_current_device() {
    int dev = 0;
    OCSL_HIP(hipGetDevice(&dev));
    OCSL_REQUIRE(dev >= 0 && dev < kMaxDevices, OCSL_ERR_ARG, "device index beyond the supported 64");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        double* ptr = nullptr;
        OCSL_HIP(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
        OCSL_REQUIRE(ptr, OCSL_ERR_HIP, "no device address for the reduction scratch");
        g_scratch[dev] = ptr;
    }
    return g_scratch[dev];
}

}  // namespace
}  // namespace ocsl

using namespace ocsl;

extern "C" {

ocsl_status ocsl_vec_norm(int64_t n, const double* a, const double* b, ocsl_where where, double* out) {
    return vec_reduce(n, a, b, 0, where, out);
}

ocsl_status ocsl_vec_sum(int64_t n, const double* a, ocsl_where where, double* out) {
    return vec_reduce(n, a, nullptr, 1, where, out);
}

ocsl_status ocsl_vec_ABSsum(int64_t n, const double* a, ocsl_where where, double* out) {
    return vec_reduce(n, a, nullptr, 2, where, out);
}

ocsl_status ocsl_vec_ABSsum_noalloc(int64_t n, const double* a, double* out, void* stream) {
    if (n < 0 || n > 0x7fffffff) {
        set_last_error("ocsl_vec_ABSsum_noalloc: n must be 0 .. 2^31 - 1");
        return OCSL_ERR_ARG;
    }
    if (!out || (n > 0 && !a)) {
        set_last_error("ocsl_vec_ABSsum_noalloc: null pointer");
        return OCSL_ERR_ARG;
    }
    try {
        double* part = scratch_for_current_device();
        hipStream_t st = (hipStream_t)stream;
        // the same launch as ocsl_vec_ABSsum: the same partial sums, added in
        // the same order; every partial read is written first by this call
        const int g = n > 0 ? red_blocks(kBlock * 4, n) : 1;
        OCSL_REQUIRE(g <= kNoAllocPartials, OCSL_ERR_ARG, "ocsl_vec_ABSsum_noalloc: too many reduction blocks");
        DotArgs<1> d{{a}, {a}, {2}};
        const RedScratch red{part};
        k_dots<1><<<g, kBlock, 0, st>>>((int)n, d, red);
        k_finalize<1, FinStore<1>><<<1, kBlock, 0, st>>>(red, g, FinStore<1>{out});
        OCSL_HIP(hipGetLastError());
        return OCSL_OK;
    } catch (const Error& e) {
        set_last_error(e.what);
        return e.status;
    } catch (...) {
        set_last_error("unexpected exception");
        return OCSL_ERR_ARG;
    }
}

ocsl_status ocsl_vec_axpy(int64_t n, double a, const double* x, double b, const double* y, double* out,
                          ocsl_where where) {
    return vec_elementwise(n, a, x, b, y, out, where, 0);
}

ocsl_status ocsl_vec_product(int64_t n, const double* x, const double* y, double* out, ocsl_where where) {
    return vec_elementwise(n, 1.0, x, 0.0, y, out, where, 1);
}

}  // extern "C"
