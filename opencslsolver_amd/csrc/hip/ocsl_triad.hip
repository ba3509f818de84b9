// STREAM triad a = b + s c in FP64 over device arrays of n elements: the
// bandwidth a plain streaming kernel reaches on this device, reported beside
// the published HBM peak as the practical ceiling of the bandwidth-bound
// kernels.  One 16-byte vector per thread over a full grid of 256-thread
// workgroups: the fastest of the shapes measured on MI355X (grid-stride
// loops with 4096-65536 workgroups of 256-1024 threads, and 2, 4 or 8
// vectors per thread, reach 4.3-5.7 TB/s on 1.5 GB; this one 5.73).
#include "ocsl_handle.h"

namespace ocsl {
namespace {

__global__ void __launch_bounds__(kBlock) k_triad(int64_t nv, double2* __restrict__ a, const double2* __restrict__ b,
                                                  const double2* __restrict__ c, double s) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nv) {
        const double2 bv = b[v], cv = c[v];
        a[v] = make_double2(bv.x + s * cv.x, bv.y + s * cv.y);
    }
}

__global__ void k_triad_init(int64_t n, double* __restrict__ b, double* __restrict__ c) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        b[i] = 1.0 + 1e-9 * (double)(i & 1023);
        c[i] = 2.0 - 1e-9 * (double)(i & 511);
    }
}

}  // namespace
}  // namespace ocsl

using namespace ocsl;

extern "C" ocsl_status ocsl_time_stream_triad(int64_t n, int reps, double* avg_ms, double* bytes_per_launch) {
    if (n <= 0 || (n & 1) || reps <= 0 || !avg_ms || !bytes_per_launch) return OCSL_ERR_ARG;
    if (ocsl_device_count() <= 0) {
        set_last_error("no HIP device");
        return OCSL_ERR_NODEVICE;
    }
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ocsl_status st = OCSL_OK;
    try {
        OCSL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        OCSL_HIP(hipEventCreate(&e0));
        OCSL_HIP(hipEventCreate(&e1));
        DBuf<double> a, b, c;
        a.alloc(n);
        b.alloc(n);
        c.alloc(n);
        k_triad_init<<<grid_for(n), kBlock, 0, s>>>(n, b.p, c.p);
        const int64_t nv = n / 2;
        OCSL_REQUIRE((nv + kBlock - 1) / kBlock < ((int64_t)1 << 31), OCSL_ERR_ARG, "triad too long");
        const int grid = (int)((nv + kBlock - 1) / kBlock);
        auto launch = [&] {
            k_triad<<<grid, kBlock, 0, s>>>(nv, reinterpret_cast<double2*>(a.p), reinterpret_cast<const double2*>(b.p),
                                            reinterpret_cast<const double2*>(c.p), 3.0);
        };
        launch();  // warm
        OCSL_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) launch();
        OCSL_HIP(hipEventRecord(e1, s));
        OCSL_HIP(hipGetLastError());
        OCSL_HIP(hipStreamSynchronize(s));
        float ms = 0.f;
        OCSL_HIP(hipEventElapsedTime(&ms, e0, e1));
        *avg_ms = (double)ms / reps;
        *bytes_per_launch = 24.0 * (double)n;
    } catch (const Error& e) {
        set_last_error(e.what);
        st = e.status;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        st = OCSL_ERR_NOMEM;
    }
    if (s) (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    return st;
}
