// Shared internals of libocsl_hip.so: handle layout, device buffers, launch
// helpers and the device building blocks used by every solver kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "ocsl.h"

namespace ocsl {

constexpr int kWave = 64;         // CDNA wavefront width
constexpr int kBlock = 256;       // threads per workgroup (4 waves)
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kXcds = 8;          // MI355X: 8 XCDs, workgroups dealt round-robin
constexpr int kMaxRedBlocks = 2048;
constexpr int kMaxRed = 4;        // reductions finalised by one kernel

// ---------------------------------------------------------------- errors
struct Error {
    ocsl_status status;
    std::string what;
};

void set_last_error(const std::string& msg);

#define OCSL_HIP(x)                                                                       \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess)                                                             \
            throw ::ocsl::Error{OCSL_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)}; \
    } while (0)

#define OCSL_REQUIRE(cond, st, msg)                          \
    do {                                                     \
        if (!(cond)) throw ::ocsl::Error{(st), (msg)};       \
    } while (0)

// ---------------------------------------------------------------- buffers
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count <= n && p) return;
        release();
        const size_t bytes = sizeof(T) * (count ? count : 1);
        hipError_t e = hipMalloc((void**)&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            throw Error{OCSL_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed"};
        }
        n = count;
    }
    void upload(const T* host, size_t count, hipStream_t s) {
        alloc(count);
        if (count) OCSL_HIP(hipMemcpyAsync(p, host, sizeof(T) * count, hipMemcpyHostToDevice, s));
    }
};

// Device-side scalar state of one solve.  Written by the last block of each
// reducing kernel; read by later kernels (kernel boundaries order them).
struct Scalars {
    double rr, rr0, bb, pq, alpha, beta, omega, rho, rho_old, ts, tt, ss, r_ref, tol2;
    double rrk[4], alphak[4], omegak[4], rhok[4], rho_oldk[4], betak[4], bbk[4], ssk[4];
    double aux[8];
    int iter, max_iter, done, halt, early, need_update, breakdown, nonfinite;
    int donek[4], earlyk[4];
    double replace_ratio;
    int hist_cap;
    int pad[3];
};

// Scratch of the deterministic grid reductions: partial sums per block per
// reduction, and an arrival counter that the last block resets.
struct RedScratch {
    double* partial;    // [kMaxRed][kMaxRedBlocks]
    unsigned* counter;  // 1
};

// ---------------------------------------------------------------- device helpers
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Logical block id so that each XCD processes one contiguous range of blocks:
// hardware block b runs on XCD b % 8.
__device__ __forceinline__ int xcd_block(int b, int g) {
    if (g % kXcds) return b;
    return (b % kXcds) * (g / kXcds) + b / kXcds;
}

// Grid-wide deterministic sum of NR per-thread values.  Every block stores
// its partials; the last block to arrive adds them in block order and calls
// fin(totals) on thread 0.  Returns true in the finalising block.
template <int NR, class Fin>
__device__ __forceinline__ void grid_reduce(const double (&v)[NR], RedScratch red, Fin fin) {
    __shared__ double sh[NR][kWavesPerBlock];
    __shared__ int is_last;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const double s = wave_sum(v[r]);
        if (lane == 0) sh[r][wid] = s;
    }
    __syncthreads();
    const int G = gridDim.x;
    if (tid == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            double s = 0.0;
            for (int w = 0; w < kWavesPerBlock; ++w) s += sh[r][w];
            __hip_atomic_store(&red.partial[r * kMaxRedBlocks + blockIdx.x], s, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        __threadfence();
        const unsigned t = atomicAdd(red.counter, 1u);
        is_last = (t == (unsigned)(G - 1));
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double tot[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        double s = 0.0;
        for (int i = tid; i < G; i += kBlock)
            s += __hip_atomic_load(&red.partial[r * kMaxRedBlocks + i], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        s = wave_sum(s);
        if (lane == 0) sh[r][wid] = s;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            double s = 0.0;
            for (int w = 0; w < kWavesPerBlock; ++w) s += sh[r][w];
            tot[r] = s;
        }
        *red.counter = 0u;
        fin(tot);
    }
}

// Number of reduction blocks for n items handled kBlock at a time.
inline int red_blocks(int64_t items_per_block_unit, int64_t n) {
    int64_t g = (n + items_per_block_unit - 1) / items_per_block_unit;
    if (g < 1) g = 1;
    if (g > kMaxRedBlocks) g = kMaxRedBlocks;
    return (int)g;
}

// ---------------------------------------------------------------- SELL-64
// Sliced ELLPACK with slice height = wavefront width: slice s holds rows
// 64s..64s+63, stored column-major (entry k of lane l at base + 64k + l), so
// every load of a wavefront is one contiguous 256/512-byte segment.  Padding
// entries carry value 0 and column = their own row (an in-window gather).
template <class V>
struct SellView {
    int n;
    int n_slices;
    const int* slice_ptr;  // n_slices + 1, element offsets (multiples of 64)
    const int* col;
    const V* val;
};

// Row sum of slice-row `row` (lane of slice s) against x, in stored order.
template <class V, class X, class Acc>
__device__ __forceinline__ Acc sell_row(const SellView<V>& A, int s, int lane, const X* __restrict__ x) {
    const int base = A.slice_ptr[s];
    const int w = (A.slice_ptr[s + 1] - base) >> 6;
    Acc acc = 0;
    const int* cp = A.col + base + lane;
    const V* vp = A.val + base + lane;
    for (int k = 0; k < w; ++k) {
        const int c = __builtin_nontemporal_load(cp + k * kWave);
        const V v = __builtin_nontemporal_load(vp + k * kWave);
        acc += (Acc)v * (Acc)x[c];
    }
    return acc;
}

// Row sum in difference form: s_i x_i + sum_{j != i} a_ij (x_j - x_i), where the
// diagonal slot (the first slot whose column is the row itself) holds the
// row sum s_i = sum_j a_ij instead of a_ii.  For matrices whose rows nearly
// sum to zero (diffusion operators with a weak Dirichlet leak) this keeps the
// low-precision product accurate relative to its result: no term of size
// |a_ii x_i| has to cancel.  Storage may be FP32; the arithmetic is FP64
// (free in a bandwidth-bound kernel), so the only rounding is that of the
// stored operands.  Padding slots (value 0, column = row) add 0.
template <class V, class X>
__device__ __forceinline__ double sell_row_diff(const SellView<V>& A, int s, int lane, int row,
                                                const X* __restrict__ x) {
    const int base = A.slice_ptr[s];
    const int w = (A.slice_ptr[s + 1] - base) >> 6;
    const double xi = row < A.n ? (double)x[row] : 0.0;
    double acc = 0.0;
    const int* cp = A.col + base + lane;
    const V* vp = A.val + base + lane;
    for (int k = 0; k < w; ++k) {
        const int c = __builtin_nontemporal_load(cp + k * kWave);
        const V v = __builtin_nontemporal_load(vp + k * kWave);
        acc += (double)v * (c == row ? xi : (double)x[c] - xi);
    }
    return acc;
}

}  // namespace ocsl
