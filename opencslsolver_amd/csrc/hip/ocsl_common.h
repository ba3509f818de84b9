// Shared internals of libocsl_hip.so: handle layout, device buffers, launch
// helpers and the device building blocks used by every solver kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "ocsl.h"
#include "ocsl_launch.h"

namespace ocsl {


// ---------------------------------------------------------------- errors

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
    double rr, rr0, rr_last, bb, pq, alpha, beta, omega, rho, rho_old, ts, tt, ss, r_ref, tol2;
    double rrk[4], alphak[4], omegak[4], rhok[4], rho_oldk[4], betak[4], bbk[4], ssk[4], rrefk[4];
    double aux[8];
    int iter, max_iter, done, halt, early, need_update, breakdown, nonfinite;
    int rr_blocks;  // producing grid of the r.r partials the next CG SpMV re-adds
    int donek[4], earlyk[4];
    double replace_ratio;
    int hist_cap;
    int lag;            // MIXED CG: x lags steps of the current period (or pair)
    double alpha_prev;  // MIXED CG: alpha of that even iteration
    int fstop[2];       // one-kernel CG: stop flag written by the kernels of each parity
    double alpha_q[8];  // MIXED CG: alpha of the iterations of the current x-update period
};

// Scratch of the deterministic grid reductions: one partial sum per block
// per reduction.
struct RedScratch {
    double* partial;  // [kMaxRed][kMaxRedBlocks]
};

// Partial sums of batched kernels: slot r, block b at partial[r * stride + b].
struct RedScratchBig {
    double* partial;
    size_t stride;
};

// ---------------------------------------------------------------- device helpers
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Logical block id so that each XCD processes one contiguous range of blocks:
// hardware block b runs on XCD b % 8.
__device__ __forceinline__ int xcd_block(int b, int g) { return xcd_map(b, g); }

// Deterministic grid-wide reductions in two launches.  The producing kernel
// stores one partial sum per block and reduction (block_partials, plain
// stores: no fences, no atomics); a one-block k_finalize launched after it on
// the same stream adds the partials in block order and hands the totals to a
// finaliser functor on thread 0.  The kernel boundary between the two makes
// every partial visible, across XCDs too; an in-kernel "last block" scheme
// would need an agent-scope fence per block, which on CDNA writes back the
// XCD's L2 each time.
template <int NR>
__device__ __forceinline__ void block_partials(const double (&v)[NR], RedScratch red, int slot0 = 0) {
    static_assert(NR <= kMaxRed, "reduction scratch holds kMaxRed partial sums per block");
    constexpr int kMaxWaves = 16;
    __shared__ double sh[NR][kMaxWaves];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int nw = blockDim.x / kWave;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const double s = wave_sum(v[r]);
        if (lane == 0) sh[r][wid] = s;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += sh[r][w];
            red.partial[(slot0 + r) * kMaxRedBlocks + blockIdx.x] = s;
        }
    }
}

template <int NR, class Fin>
__global__ void __launch_bounds__(kBlock) k_finalize(RedScratch red, int G, Fin fin, int slot0 = 0) {
    __shared__ double sh[NR][kWavesPerBlock];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    // every slot's partials in one pass (each thread's sum of a slot still
    // runs over i ascending), loads in flight before the finaliser's own
    // state is read
    double s[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) s[r] = 0.0;
    for (int i = tid; i < G; i += kBlock)
#pragma unroll
        for (int r = 0; r < NR; ++r) s[r] += red.partial[(slot0 + r) * kMaxRedBlocks + i];
    if (!fin.active()) return;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const double t = wave_sum(s[r]);
        if (lane == 0) sh[r][wid] = t;
    }
    __syncthreads();
    if (tid == 0) {
        double tot[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            double s = 0.0;
            for (int w = 0; w < kWavesPerBlock; ++w) s += sh[r][w];
            tot[r] = s;
        }
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
    int max_w;             // widest slice (entries per row)
};

// Entries of a SELL row are fetched kBatch at a time: all column indices and
// values of a batch are loaded before any x gather, so each wavefront keeps
// up to 2*kBatch streaming loads and then kBatch gathers in flight.  The batch
// predicates depend only on the slice width and are uniform across the wave.
constexpr int kBatch = 8;

// Row sum of slice-row `row` (lane of slice s) against x, in stored order.
template <class V, class X, class Acc>
__device__ __forceinline__ Acc sell_row(const SellView<V>& A, int s, int lane, const X* __restrict__ x) {
    const int base = A.slice_ptr[s];
    const int w = (A.slice_ptr[s + 1] - base) >> 6;
    Acc acc = 0;
    const int* cp = A.col + base + lane;
    const V* vp = A.val + base + lane;
    for (int k0 = 0; k0 < w; k0 += kBatch) {
        int c[kBatch];
        V v[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j)
            if (k0 + j < w) {
                c[j] = __builtin_nontemporal_load(cp + (k0 + j) * kWave);
                v[j] = __builtin_nontemporal_load(vp + (k0 + j) * kWave);
            }
#pragma unroll
        for (int j = 0; j < kBatch; ++j)
            if (k0 + j < w) acc += (Acc)v[j] * (Acc)x[c[j]];
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
    for (int k0 = 0; k0 < w; k0 += kBatch) {
        int c[kBatch];
        V v[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j)
            if (k0 + j < w) {
                c[j] = __builtin_nontemporal_load(cp + (k0 + j) * kWave);
                v[j] = __builtin_nontemporal_load(vp + (k0 + j) * kWave);
            }
#pragma unroll
        for (int j = 0; j < kBatch; ++j)
            if (k0 + j < w) acc += (double)v[j] * (c[j] == row ? xi : (double)x[c[j]] - xi);
    }
    return acc;
}

// SELL-P: the layout of the tolerance-mode CG kernels.  The diagonal is a
// dense vector of its own (no column index); the off-diagonal entries of each
// 64-row slice are stored in pairs, pair t of lane l at base + 128 t + 2 l, so
// a lane fetches two values (float2 / double2) per load: 512 or 1024
// contiguous bytes per wavefront instruction.  Columns are either int2 pairs
// at the same positions, or -- when every |column - row| fits in 16 bits, as
// it does in a bandwidth-reducing order -- one 32-bit word per pair holding
// the two signed 16-bit offsets, word t of lane l at base/2 + 64 t + l.
// Slices are padded to an even width with value 0 and column = the row.
template <class V>
struct SellPView {
    int n;
    int n_slices;
    const int* slice_ptr;   // n_slices + 1, element offsets (multiples of 128)
    const int* col;         // int32 columns (null when colw is used)
    const unsigned* colw;   // packed 16-bit column offsets, or null
    const V* val;
    const V* diag;          // a_ii, or the row sum in the FP32 difference form
    const int* win_lo;      // per slice: smallest row the slice's rows touch (own rows, stored columns)
    const int* win_hi;      // per slice: largest such row
};

template <class V>
struct Pair2;
template <>
struct Pair2<float> {
    using T = float2;
};
template <>
struct Pair2<double> {
    using T = double2;
};

// K consecutive elements of T moved as one aligned vector access.
template <class T, int K>
struct alignas(sizeof(T) * K) Pack {
    T v[K];
};
template <class T, int K>
__device__ __forceinline__ Pack<T, K> ldp(const T* __restrict__ p) {
    return *reinterpret_cast<const Pack<T, K>*>(p);
}

// R consecutive rows of a length-n vector as doubles (0 past the end); one
// vector load unless the group straddles n.
template <int R, class T>
__device__ __forceinline__ void load_rows(const T* __restrict__ a, int row0, int n, double (&out)[R]) {
    if (row0 + R <= n) {
        const Pack<T, R> v = ldp<T, R>(a + row0);
#pragma unroll
        for (int j = 0; j < R; ++j) out[j] = (double)v.v[j];
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j) out[j] = row0 + j < n ? (double)a[row0 + j] : 0.0;
    }
}
template <int R, class T>
__device__ __forceinline__ void load_rows_t(const T* __restrict__ a, int row0, int n, T (&out)[R]) {
    if (row0 + R <= n) {
        const Pack<T, R> v = ldp<T, R>(a + row0);
#pragma unroll
        for (int j = 0; j < R; ++j) out[j] = v.v[j];
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j) out[j] = row0 + j < n ? a[row0 + j] : T(0);
    }
}
// The same without the tail guard, for vectors allocated to the layout's
// padded row count (whole slices): one vector load, nothing to branch on, so
// the compiler keeps the load in flight past the code that follows.  Values
// of padding rows are unspecified and must not reach a result.
template <int R, class T>
__device__ __forceinline__ void load_rows_p(const T* __restrict__ a, int row0, double (&out)[R]) {
    const Pack<T, R> v = ldp<T, R>(a + row0);
#pragma unroll
    for (int j = 0; j < R; ++j) out[j] = (double)v.v[j];
}
template <int R, class T>
__device__ __forceinline__ void load_rows_tp(const T* __restrict__ a, int row0, T (&out)[R]) {
    const Pack<T, R> v = ldp<T, R>(a + row0);
#pragma unroll
    for (int j = 0; j < R; ++j) out[j] = v.v[j];
}
template <int R, class T>
__device__ __forceinline__ void store_rows(T* __restrict__ a, int row0, int n, const T (&v)[R]) {
    if (row0 + R <= n) {
        Pack<T, R> pk;
#pragma unroll
        for (int j = 0; j < R; ++j) pk.v[j] = v[j];
        *reinterpret_cast<Pack<T, R>*>(a + row0) = pk;
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j)
            if (row0 + j < n) a[row0 + j] = v[j];
    }
}

// R rows of A x over SELL-P with R rows per lane: slices of 64 R rows, lane
// l owning rows s*64R + l*R + j.  Pair t of row j of lane l is at element
// base + 128 R t + 2 R l + 2 j, so a lane fetches 2R values (and R packed
// column words) per access; R = 1 is the one-row-per-lane layout.  kDiff:
// difference form, diag = row sum s_i: s_i x_i + sum_off a_ij (x_j - x_i);
// otherwise a_ii x_i + sum_off a_ij x_j.  Arithmetic in FP64 whatever the
// storage; per row the terms are added in the same order for every R.
// gather(c) returns the vector's element c as a double (a plain load, or a
// value formed from several vectors).  kPairs pairs are loaded per round: 4
// at R = 1, 2 at R = 2, and 1 at R = 4, where 2 would take the MIXED SpMV
// from 78 to 96 registers (6 to 5 waves per SIMD) and measured 21.3 against
// 20.4 us at 128^3.
// A slice's start, its pairs per row and the lane's diagonals: the loads that
// open sellp_rows_g, separable so that a kernel can issue them early.
template <class V, int R>
struct SellpHead {
    int base, np;
    V dg[R];  // as stored (fewer registers while held)
};
template <class V, int R>
__device__ __forceinline__ SellpHead<V, R> sellp_head(const SellPView<V>& A, int s, int row0) {
    SellpHead<V, R> h;
    h.base = A.slice_ptr[s];
    h.np = (A.slice_ptr[s + 1] - h.base) / (128 * R);  // pairs per row
    load_rows_tp<R>(A.diag, row0, h.dg);               // diagonals are allocated to whole slices
    return h;
}

template <class V, bool kDiff, int R, class G, int kPairs = (R == 1 ? 4 : (R == 4 ? 1 : 2))>
__device__ __forceinline__ void sellp_rows_gh(const SellPView<V>& A, const SellpHead<V, R>& hd, int lane, int row0,
                                              const G& gather, const double (&xi)[R], double (&acc)[R]) {
    const int base = hd.base, np = hd.np;
#pragma unroll
    for (int j = 0; j < R; ++j) acc[j] = (double)hd.dg[j] * xi[j];
    for (int t0 = 0; t0 < np; t0 += kPairs) {
        int c[kPairs][2 * R];
        V v[kPairs][2 * R];
#pragma unroll
        for (int jj = 0; jj < kPairs; ++jj) {
            const int t = t0 + jj;
            if (t < np) {
                if (A.colw) {
                    const Pack<unsigned, R> w = ldp<unsigned, R>(A.colw + (base >> 1) + t * 64 * R + lane * R);
#pragma unroll
                    for (int j = 0; j < R; ++j) {
                        c[jj][2 * j] = row0 + j + (int)(short)(w.v[j] & 0xffffu);
                        c[jj][2 * j + 1] = row0 + j + (int)(short)(w.v[j] >> 16);
                    }
                } else {
                    const Pack<int, 2 * R> cc = ldp<int, 2 * R>(A.col + base + t * 128 * R + lane * 2 * R);
#pragma unroll
                    for (int e = 0; e < 2 * R; ++e) c[jj][e] = cc.v[e];
                }
            }
        }
#pragma unroll
        for (int jj = 0; jj < kPairs; ++jj) {
            const int t = t0 + jj;
            if (t < np) {
                const Pack<V, 2 * R> vv = ldp<V, 2 * R>(A.val + base + t * 128 * R + lane * 2 * R);
#pragma unroll
                for (int e = 0; e < 2 * R; ++e) v[jj][e] = vv.v[e];
            }
        }
#pragma unroll
        for (int jj = 0; jj < kPairs; ++jj)
            if (t0 + jj < np) {
#pragma unroll
                for (int j = 0; j < R; ++j)
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const double a = (double)v[jj][2 * j + e];
                        const double xj = gather(c[jj][2 * j + e]);
                        if constexpr (kDiff) acc[j] += a * (xj - xi[j]);
                        else acc[j] += a * xj;
                    }
            }
    }
}

template <class V, bool kDiff, int R, class G, int kPairs = (R == 1 ? 4 : (R == 4 ? 1 : 2))>
__device__ __forceinline__ void sellp_rows_g(const SellPView<V>& A, int s, int lane, int row0, const G& gather,
                                             const double (&xi)[R], double (&acc)[R]) {
    sellp_rows_gh<V, kDiff, R, G, kPairs>(A, sellp_head<V, R>(A, s, row0), lane, row0, gather, xi, acc);
}

template <class V, class X, bool kDiff, int R>
__device__ __forceinline__ void sellp_rows(const SellPView<V>& A, int s, int lane, int row0,
                                           const X* __restrict__ x, const double (&xi)[R], double (&acc)[R]) {
    sellp_rows_g<V, kDiff, R>(A, s, lane, row0, [x](int c) { return (double)x[c]; }, xi, acc);
}

// One row per lane (R = 1).
template <class V, class X, bool kDiff>
__device__ __forceinline__ double sellp_row(const SellPView<V>& A, int s, int lane, int row,
                                            const X* __restrict__ x, double xi) {
    const double xr[1] = {xi};
    double acc[1];
    sellp_rows<V, X, kDiff, 1>(A, s, lane, row, x, xr, acc);
    return acc[0];
}

// NR interleaved vectors (x[row * NR + k]) against one SELL row: one pass
// over the matrix entries serves all right-hand sides.  FP32 storage uses
// the difference form of sell_row_diff; FP64 the plain row sum.
template <class V, class X, int NR>
__device__ __forceinline__ void sell_rows_nr(const SellView<V>& A, int s, int lane, int row,
                                             const X* __restrict__ x, double (&out)[NR]) {
    const int base = A.slice_ptr[s];
    const int w = (A.slice_ptr[s + 1] - base) >> 6;
    const int* cp = A.col + base + lane;
    const V* vp = A.val + base + lane;
#pragma unroll
    for (int k = 0; k < NR; ++k) out[k] = 0.0;
    if constexpr (sizeof(V) == 4) {
        double xi[NR];
#pragma unroll
        for (int k = 0; k < NR; ++k) xi[k] = row < A.n ? (double)x[(size_t)row * NR + k] : 0.0;
        for (int j = 0; j < w; ++j) {
            const int c = __builtin_nontemporal_load(cp + j * kWave);
            const double v = (double)__builtin_nontemporal_load(vp + j * kWave);
#pragma unroll
            for (int k = 0; k < NR; ++k)
                out[k] += v * (c == row ? xi[k] : (double)x[(size_t)c * NR + k] - xi[k]);
        }
    } else {
        for (int j = 0; j < w; ++j) {
            const int c = __builtin_nontemporal_load(cp + j * kWave);
            const double v = (double)__builtin_nontemporal_load(vp + j * kWave);
#pragma unroll
            for (int k = 0; k < NR; ++k) out[k] += v * (double)x[(size_t)c * NR + k];
        }
    }
}

}  // namespace ocsl
