// Device kernels shared by the solver drivers: value/vector gathers, SELL-64
// SpMV with fused epilogues, BLAS-1 and the fused CG iteration kernels.
#pragma once

#include "ocsl_common.h"

namespace ocsl {

// ------------------------------------------------------------------ gathers
template <class V>
__global__ void k_gather_vals(int64_t m, const int* __restrict__ src, const double* __restrict__ in,
                              V* __restrict__ out) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < m;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int s = src[q];
        out[q] = s >= 0 ? (V)in[s] : (V)0;
    }
}

// Two precisions from one read of the source values.
[[maybe_unused]] static __global__ void k_gather_vals2(int64_t m, const int* __restrict__ src, const double* __restrict__ in,
                               double* __restrict__ out64, float* __restrict__ out32) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < m;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int s = src[q];
        const double v = s >= 0 ? in[s] : 0.0;
        out64[q] = v;
        out32[q] = (float)v;
    }
}

// out[i] = in[perm[i]] (perm == nullptr: copy)
template <class T>
__global__ void k_permute_in(int n, const int* __restrict__ perm, const T* __restrict__ in,
                             T* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = in[perm ? perm[i] : i];
}

// out[perm[i]] = in[i]
template <class T>
__global__ void k_permute_out(int n, const int* __restrict__ perm, const T* __restrict__ in,
                              T* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[perm ? perm[i] : i] = in[i];
}

// *S = v, from a kernel: a solve that may be replayed as a graph sets its
// scalars without a host-to-device copy.
[[maybe_unused]] static __global__ void k_set_scalars(Scalars* S, Scalars v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *S = v;
}

template <class T>
__global__ void k_fill(int n, T v, T* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = v;
}

// ------------------------------------------------------------------ BLAS-1
// out = a*x + b*y (cmatrix_solver.cpp:284-335)
[[maybe_unused]] static __global__ void k_axpby(int n, double a, const double* __restrict__ x, double b,
                        const double* __restrict__ y, double* out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = a * x[i] + b * y[i];
}

// out = (ca * a) x + (cb * b) y with a = *pa and b = *pb where given (device
// scalars; else 1); nothing once S->halt is set.
[[maybe_unused]] static __global__ void k_axpby_s(int n, double ca, const double* pa, const double* __restrict__ x,
                                                  double cb, const double* pb, const double* __restrict__ y,
                                                  double* out, const Scalars* S) {
    if (S->halt) return;
    const double a = pa ? ca * *pa : ca;
    const double b = pb ? cb * *pb : cb;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = a * x[i] + b * y[i];
}

// Up to 3 dot products / sums in one pass; results land in out[0..NR-1].
// kind: 0 = a.b, 1 = sum(a), 2 = sum|a|.
template <int NR>
struct DotArgs {
    const double* a[NR];
    const double* b[NR];
    int kind[NR];
};

template <int NR>
__global__ void __launch_bounds__(kBlock) k_dots(int n, DotArgs<NR> args, RedScratch red) {
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const double av = args.a[r][i];
            acc[r] += args.kind[r] == 0 ? av * args.b[r][i] : (args.kind[r] == 1 ? av : fabs(av));
        }
    }
    block_partials<NR>(acc, red);
}

// Copies the totals out (k_dots, k_true_residual).
template <int NR>
struct FinStore {
    double* out;
    __device__ bool active() const { return true; }
    __device__ void operator()(const double* t) const {
        for (int r = 0; r < NR; ++r) out[r] = t[r];
    }
};

// ------------------------------------------------------------------ SpMV
// out = a*(A x) + b*z over a SELL matrix whose row order is the vector order
// (cmatrix_solver.cpp:338-453).  z may be null.
template <class V>
__global__ void __launch_bounds__(kBlock) k_sell_axpby(SellView<V> A, double a, const double* a_dev,
                                                       const double* __restrict__ x, double b,
                                                       const double* __restrict__ z,
                                                       double* __restrict__ out) {
    // a_dev (optional): a device scalar the coefficient a multiplies
    if (a_dev) a *= *a_dev;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * kWavesPerBlock;
    for (int s = wave; s < A.n_slices; s += nwaves) {
        const int row = s * kWave + lane;
        const double ax = sell_row<V, double, double>(A, s, lane, x);
        if (row < A.n) {
            double o = (a == 1.0) ? ax : a * ax;
            if (z) o += b * z[row];
            out[row] = o;
        }
    }
}

// ------------------------------------------------------------------ CG (tolerance mode)
// Vectors live in the SELL row order.  X is the Krylov vector type (double or
// float); dot products always accumulate in double.
//
// One iteration is three kernels, with no host round trip and no separate
// finalisation launch: each kernel stores per-block partial sums, and the
// kernels after it re-add those partials themselves, every block in the same
// fixed order, so all blocks derive bitwise identical scalars.
//   K1 k_cg_spmv      q = A p, partials of p.q   (block 0: S->rr = sum r.r)
//   K2 k_cg_update_r  alpha = rr / pq; r -= alpha q, partials of r.r
//   K3 k_cg_update_xp alpha, beta = rr_new / rr; x += alpha p; p = r + beta p
//                     (block 0: iteration count, history, convergence flags)
// Flags: `done` (written by K3 block 0, or K2 block 0 on breakdown) stops K1
// and K2; K1 turns it into `halt`, which stops K3 -- K3 cannot test `done`
// because its own block 0 writes it.  Only block 0 writes shared scalars, and
// never one that other blocks of the same kernel read.
constexpr int kSlotPQ = 0;  // p.q of K1
constexpr int kSlotRR = 1;  // r.r of K2 and of the residual kernel
constexpr int kSlotBB = 2;  // b.b of the residual kernel
constexpr int kSlotRQ = 3;  // r.q of the merged K1 (mixed precision)
constexpr int kSlotQQ = 4;  // q.q of the merged K1

template <class V, class X>
struct CgArgs {
    SellPView<V> A;
    X* r;
    X* p;
    X* q;
    X* p_alt;       // MIXED: the other direction buffer (p is double-buffered)
    double* x;      // solution, FP64 in both precisions
    Scalars* S;
    double* hist;
    RedScratch red;
    int g_spmv;     // blocks of the SpMV kernels (and of the residual kernel)
    int g_vec;      // blocks of the vector-update kernels
    X* pb[8];       // MIXED: the direction buffers of one x-update period
};

// Sum of the G partials of one slot, identical in every block (fixed order,
// butterfly wave sums give every lane the same bits); all threads get it.
__device__ __forceinline__ double block_sum_partials(RedScratch red, int slot, int G) {
    __shared__ double sh[16];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int nw = blockDim.x / kWave;
    double s = 0.0;
    for (int i = tid; i < G; i += blockDim.x) s += red.partial[slot * kMaxRedBlocks + i];
    s = wave_sum(s);
    if (lane == 0) sh[wid] = s;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += sh[w];
    __syncthreads();
    return t;
}

// Several slots at once: one pass, one pair of barriers.

// The same sums in two phases, read by the block's first wave only, with
// 16-byte loads.  issue() starts the wave's loads of the first 512 partials
// of every slot (raw loads at clamped indices, nothing computed from them, so
// that nothing waits for them yet); a kernel calls it before its other global
// loads, because
// vector loads complete in issue order: waiting for the partials then waits
// for nothing issued after them.  finish() adds any further partials, reduces
// across the wave and hands the totals to the whole block through LDS with a
// single barrier -- so each instantiation's finish() may run once per kernel.
// Every block runs the same code, so all get the same bits.
template <int K>
struct PartialSums {
    static constexpr int kRounds = 4;  // double2 loads per lane and slot issued up front
    double2 v[K][kRounds];
    __device__ __forceinline__ void issue(RedScratch red, const int (&slot)[K], const int (&G)[K]) {
        if (threadIdx.x >= kWave) return;
        const int lane = threadIdx.x;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double2* src = reinterpret_cast<const double2*>(red.partial + slot[k] * kMaxRedBlocks);
#pragma unroll
            for (int r = 0; r < kRounds; ++r) v[k][r] = src[min(lane + kWave * r, kMaxRedBlocks / 2 - 1)];
        }
    }
    __device__ __forceinline__ void finish(RedScratch red, const int (&slot)[K], const int (&G)[K],
                                           double (&out)[K]) {
        __shared__ double sh[K];
        if (threadIdx.x < kWave) {
            const int lane = threadIdx.x;
            double a[K];
            int gmax = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                a[k] = 0.0;
#pragma unroll
                for (int r = 0; r < kRounds; ++r) {
                    const int i2 = lane + kWave * r;  // partials past G[k] (stale or clamped) add 0
                    a[k] += (2 * i2 < G[k] ? v[k][r].x : 0.0) + (2 * i2 + 1 < G[k] ? v[k][r].y : 0.0);
                }
                gmax = max(gmax, G[k]);
            }
            for (int i2 = lane + kWave * kRounds; 2 * i2 < gmax; i2 += kWave)
#pragma unroll
                for (int k = 0; k < K; ++k)
                    if (2 * i2 < G[k]) {
                        const double2 t = reinterpret_cast<const double2*>(red.partial + slot[k] * kMaxRedBlocks)[i2];
                        a[k] += t.x;
                        if (2 * i2 + 1 < G[k]) a[k] += t.y;
                    }
#pragma unroll
            for (int k = 0; k < K; ++k) a[k] = wave_sum(a[k]);
            if (lane == 0)
#pragma unroll
                for (int k = 0; k < K; ++k) sh[k] = a[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] = sh[k];
    }
};

template <int K>
__device__ __forceinline__ void block_sum_partials_k(RedScratch red, const int (&slot)[K], int G,
                                                     double (&out)[K]) {
    int gs[K];
#pragma unroll
    for (int k = 0; k < K; ++k) gs[k] = G;
    PartialSums<K> ps;
    ps.issue(red, slot, gs);
    ps.finish(red, slot, gs, out);
}

// 16-byte vector of X: each lane of the vector kernels moves kVec<X>
// consecutive elements per load or store.
template <class X>
struct Vec16;
template <>
struct Vec16<float> {
    static constexpr int N = 4;
    using T = float4;
};
template <>
struct Vec16<double> {
    static constexpr int N = 2;
    using T = double2;
};

template <class X>
__device__ __forceinline__ void vload(const X* __restrict__ p, int v, X (&e)[Vec16<X>::N]) {
    const typename Vec16<X>::T t = reinterpret_cast<const typename Vec16<X>::T*>(p)[v];
    __builtin_memcpy(e, &t, sizeof(t));
}
template <class X>
__device__ __forceinline__ void vstore(X* __restrict__ p, int v, const X (&e)[Vec16<X>::N]) {
    typename Vec16<X>::T t;
    __builtin_memcpy(&t, e, sizeof(t));
    reinterpret_cast<typename Vec16<X>::T*>(p)[v] = t;
}
// N doubles starting at element v*N (N = 2 or 4), as double2 loads / stores
template <int N>
__device__ __forceinline__ void dload(const double* __restrict__ p, int v, double (&e)[N]) {
#pragma unroll
    for (int h = 0; h < N / 2; ++h) {
        const double2 t = reinterpret_cast<const double2*>(p)[v * (N / 2) + h];
        e[2 * h] = t.x;
        e[2 * h + 1] = t.y;
    }
}
template <int N>
__device__ __forceinline__ void dstore(double* __restrict__ p, int v, const double (&e)[N]) {
#pragma unroll
    for (int h = 0; h < N / 2; ++h) reinterpret_cast<double2*>(p)[v * (N / 2) + h] = make_double2(e[2 * h], e[2 * h + 1]);
}

// r = b - A x (FP64), stored as r64 or r32; p = r on the first call; partials
// of r.r (kSlotRR) and b.b (kSlotBB).
template <bool kMixed, int R>
__global__ void __launch_bounds__(kCgMaxBlock) k_cg_residual(SellPView<double> A, const double* __restrict__ b,
                                                            const double* __restrict__ x,
                                                            double* __restrict__ r64, float* __restrict__ r32,
                                                            double* __restrict__ p64, float* __restrict__ p32,
                                                            RedScratch red, int first) {
    const int lane = threadIdx.x & (kWave - 1);
    const int g = gridDim.x;
    const int lb = xcd_block(blockIdx.x, g);
    const SliceRange sr = block_slices(lb, g, A.n_slices);
    const int s0 = sr.s0, s1 = sr.s1;
    double acc[2] = {0.0, 0.0};
    for (int s = s0 + (int)(threadIdx.x / kWave); s < s1; s += (int)(blockDim.x / kWave)) {
        const int row0 = s * kWave * R + lane * R;
        double xi[R], ax[R], bv[R];
        load_rows_p<R>(x, row0, xi);
        sellp_rows<double, double, false, R>(A, s, lane, row0, x, xi, ax);
        load_rows_p<R>(b, row0, bv);
        double rv[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            rv[j] = bv[j] - ax[j];
            if (row0 + j < A.n) {
                acc[0] += rv[j] * rv[j];
                acc[1] += bv[j] * bv[j];
            }
        }
        if constexpr (kMixed) {
            float rf[R];
#pragma unroll
            for (int j = 0; j < R; ++j) rf[j] = (float)rv[j];
            store_rows<R>(r32, row0, A.n, rf);
            if (first) store_rows<R>(p32, row0, A.n, rf);
        } else {
            store_rows<R>(r64, row0, A.n, rv);
            if (first) store_rows<R>(p64, row0, A.n, rv);
        }
    }
    block_partials<2>(acc, red, kSlotRR);
}

// After the residual kernel: bb and history entry 0 on the first call;
// convergence judged on this (true) residual; r_ref for the replacement test.
struct FinCgResidual {
    Scalars* S;
    double* hist;
    int first;
    int blocks;  // grid of the residual kernel: the partial count of its r.r
    __device__ bool active() const { return true; }
    __device__ void operator()(const double* t) const {
        const double rr = t[0], bb = t[1];
        S->rr_blocks = blocks;
        S->rr = rr;
        S->rr_last = rr;
        S->r_ref = sqrt(rr);
        if (first) {
            S->bb = bb;
            S->rr0 = rr;
        }
        const double scale = S->bb > 0.0 ? S->bb : 1.0;
        if (first && hist && S->hist_cap > 0) hist[0] = sqrt(rr / scale);
        S->done = (rr <= S->tol2 * scale) ? 1 : 0;
        S->halt = S->done;
        S->need_update = 0;
        S->fstop[0] = S->fstop[1] = 0;
        if (!isfinite(rr)) { S->nonfinite = 1; S->done = 1; S->halt = 1; }
    }
};

// K1: q = A p, partials of p.q.  FP32 storage uses the difference form.
template <class V, class X, int R>
__global__ void __launch_bounds__(kCgMaxBlock) k_cg_spmv(CgArgs<V, X> c) {
    Scalars* S = c.S;
    // the wave's first slice opens with its loads issued beside the stop
    // flag's, not after it; each later slice's at the end of the one before
    const int done = S->done;
    const int g = gridDim.x;
    const int lane = threadIdx.x & (kWave - 1);
    const int lb = xcd_block(blockIdx.x, g);
    const SliceRange sr = block_slices(lb, g, c.A.n_slices);
    const int s1 = sr.s1, sf = sr.s0 + (int)(threadIdx.x / kWave), step = (int)(blockDim.x / kWave);
    X pf[R] = {};
    SellpHead<V, R> hf{};
    if (sf < s1) {
        const int row0 = sf * kWave * R + lane * R;
        load_rows_tp<R>(c.p, row0, pf);
        hf = sellp_head<V, R>(c.A, sf, row0);
    }
    if (done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) S->halt = 1;
        return;
    }
    if (blockIdx.x == 0) {
        const double rr = block_sum_partials(c.red, kSlotRR, S->rr_blocks);
        if (threadIdx.x == 0) S->rr = rr;
    }
    double acc[1] = {0.0};
    for (int s = sf; s < s1; s += step) {
        const int row0 = s * kWave * R + lane * R;
        double pi[R], qd[R];
#pragma unroll
        for (int j = 0; j < R; ++j) pi[j] = (double)pf[j];
        const X* px = c.p;
        sellp_rows_gh<V, sizeof(V) == 4, R>(c.A, hf, lane, row0, [px](int cc) { return (double)px[cc]; }, pi, qd);
        X qv[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            qv[j] = (X)qd[j];
            if (row0 + j < c.A.n) acc[0] += pi[j] * (double)qv[j];
        }
        store_rows<R>(c.q, row0, c.A.n, qv);
        if (s + step < s1) {
            const int row1 = (s + step) * kWave * R + lane * R;
            load_rows_tp<R>(c.p, row1, pf);
            hf = sellp_head<V, R>(c.A, s + step, row1);
        }
    }
    block_partials<1>(acc, c.red, kSlotPQ);
}

__device__ __forceinline__ bool cg_pq_ok(double pq) { return pq != 0.0 && isfinite(pq); }

// Each thread's first kPre vectors are loaded before the partial sums are
// re-added, so the two memory latencies overlap.  Elements past the last full
// vector are handled by the last block.
constexpr int kPre = 2;

// K2: alpha = rr / pq ; r -= alpha q ; partials of r.r
template <class V, class X>
__global__ void __launch_bounds__(kCgMaxBlock) k_cg_update_r(CgArgs<V, X> c, int n) {
    constexpr int N = Vec16<X>::N;
    Scalars* S = c.S;
    const int done = S->done;  // tested once the first loads are in flight
    const int nv = n / N;
    const int stride = gridDim.x * blockDim.x;
    const int v0 = blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int kSl[1] = {kSlotPQ};
    const int gs[1] = {c.g_spmv};
    double pqs[1];
    PartialSums<1> ps;
    ps.issue(c.red, kSl, gs);
    // unconditional loads at a clamped index (the vectors hold n + 4 elements)
    X rv[kPre][N], qv[kPre][N];
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const int v = min(v0 + k * stride, nv);
        vload(c.r, v, rv[k]);
        vload(c.q, v, qv[k]);
    }
    if (done) return;
    ps.finish(c.red, kSl, gs, pqs);
    const double pq = pqs[0];
    if (!cg_pq_ok(pq)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            S->breakdown = (pq == 0.0);
            S->nonfinite = !isfinite(pq);
            S->done = 1;
        }
        return;
    }
    const double alpha = S->rr / pq;
    double acc[1] = {0.0};
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const int v = v0 + k * stride;
        if (v < nv) {
            X out[N];
#pragma unroll
            for (int e = 0; e < N; ++e) {
                out[e] = (X)((double)rv[k][e] - alpha * (double)qv[k][e]);
                acc[0] += (double)out[e] * (double)out[e];
            }
            vstore(c.r, v, out);
        }
    }
    for (int v = v0 + kPre * stride; v < nv; v += stride) {
        X a[N], b[N], out[N];
        vload(c.r, v, a);
        vload(c.q, v, b);
#pragma unroll
        for (int e = 0; e < N; ++e) {
            out[e] = (X)((double)a[e] - alpha * (double)b[e]);
            acc[0] += (double)out[e] * (double)out[e];
        }
        vstore(c.r, v, out);
    }
    if (blockIdx.x == gridDim.x - 1) {
        const int i = nv * N + (int)threadIdx.x;
        if (i < n) {
            const X r2 = (X)((double)c.r[i] - alpha * (double)c.q[i]);
            c.r[i] = r2;
            acc[0] += (double)r2 * (double)r2;
        }
    }
    block_partials<1>(acc, c.red, kSlotRR);
    if (blockIdx.x == 0 && threadIdx.x == 0) S->rr_blocks = gridDim.x;
}

// K3: x += alpha p (in FP64 for both precisions) ; p = r + beta p ; block 0
// records the iteration.
template <class V, class X>
__global__ void __launch_bounds__(kCgMaxBlock) k_cg_update_xp(CgArgs<V, X> c, int n) {
    constexpr int N = Vec16<X>::N;
    Scalars* S = c.S;
    const int halt = S->halt;  // tested once the first loads are in flight
    const int g = gridDim.x;
    const int nv = n / N;
    const int stride = g * blockDim.x;
    const int v0 = blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int kSl[2] = {kSlotPQ, kSlotRR};
    const int gs[2] = {c.g_spmv, g};
    double sums[2];
    PartialSums<2> ps;
    ps.issue(c.red, kSl, gs);
    // unconditional loads at a clamped index (the vectors hold n + 4 elements)
    double xv[kPre][N];
    X pv[kPre][N], rv[kPre][N];
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const int v = min(v0 + k * stride, nv);
        dload<N>(c.x, v, xv[k]);
        vload(c.p, v, pv[k]);
        vload(c.r, v, rv[k]);
    }
    if (halt) return;
    ps.finish(c.red, kSl, gs, sums);
    const double pq = sums[0];
    if (!cg_pq_ok(pq)) return;
    const double rr_new = sums[1];
    const double rr = S->rr;
    const double alpha = rr / pq, beta = rr_new / rr;
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const int v = v0 + k * stride;
        if (v < nv) {
            X pn[N];
#pragma unroll
            for (int e = 0; e < N; ++e) {
                const double p0 = (double)pv[k][e];
                xv[k][e] += alpha * p0;
                pn[e] = (X)((double)rv[k][e] + beta * p0);
            }
            dstore<N>(c.x, v, xv[k]);
            vstore(c.p, v, pn);
        }
    }
    for (int v = v0 + kPre * stride; v < nv; v += stride) {
        double xe[N];
        X pe[N], re[N], pn[N];
        dload<N>(c.x, v, xe);
        vload(c.p, v, pe);
        vload(c.r, v, re);
#pragma unroll
        for (int e = 0; e < N; ++e) {
            const double p0 = (double)pe[e];
            xe[e] += alpha * p0;
            pn[e] = (X)((double)re[e] + beta * p0);
        }
        dstore<N>(c.x, v, xe);
        vstore(c.p, v, pn);
    }
    if (blockIdx.x == g - 1) {
        const int i = nv * N + (int)threadIdx.x;
        if (i < n) {
            const double p0 = (double)c.p[i];
            c.x[i] += alpha * p0;
            c.p[i] = (X)((double)c.r[i] + beta * p0);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S->alpha = alpha;
        S->beta = beta;
        S->rr_last = rr_new;
        const int it = ++S->iter;
        const double scale = S->bb > 0.0 ? S->bb : 1.0;
        if (c.hist && it < S->hist_cap) c.hist[it] = sqrt(rr_new / scale);
        if (rr_new <= S->tol2 * scale || it >= S->max_iter) S->done = 1;
        if (!isfinite(rr_new)) { S->nonfinite = 1; S->done = 1; }
        if (sqrt(rr_new) < S->replace_ratio * S->r_ref) S->need_update = 1;
    }
}

// ------------------------------------------------------------------ CG, merged reductions
// The mixed-precision iteration in two kernels (kernel fusion with merged
// reductions).  K1 also forms r.q and q.q, which give
//   ||r_{k+1}||^2 = rr_k - 2 alpha (r.q) + alpha^2 (q.q),
// so K2 can compute beta = ||r_{k+1}||^2 / rr_k and update x, r and p in one
// pass.  The true r.r of the stored r_{k+1} (partials of K2) still drives
// alpha of the next iteration, the convergence test and the FP64 residual
// replacement trigger; only beta comes from the recurrence.
//   K1m k_cgm_spmv    block 0: rr = sum r.r, convergence, history;
//                     q = A p; partials of p.q, r.q, q.q
//   K2m k_cgm_update  alpha, beta; r -= alpha q; p' = r + beta p; partials of r.r
// x is updated once per period of M iterations (M = 2, 4 or 8): iteration j
// of a period (j = 0..M-1) reads its direction from buffer j and writes the
// next one to buffer (j + 1) mod M, so the period's directions stay at hand;
// the last update of the period adds alpha_0 P_0 + ... + alpha_{M-1} P_{M-1}
// to the FP64 x, one addition per iteration in iteration order -- the same
// sums as an update every iteration.  x is read and written once per period
// (for FP32 vectors: 36 bytes per row and iteration when x is updated every
// iteration, 5 sv + (16 + (M - 1) sv) / M in general: 30, 27, 25.5).
// When a period is cut by a stop after its update j > 0, the update of the
// stopping iteration adds the j pending steps (S->lag marks a period with
// pending steps: set by update 0, cleared by the SpMV of the next period and
// by any SpMV that finds the solve stopped, so it is read and written in
// different kernels).
template <class V, class X, int R>
__global__ void __launch_bounds__(kCgMaxBlock) k_cgm_spmv(CgArgs<V, X> c, int first) {
    Scalars* S = c.S;
    // the wave's first slice opens with its loads issued beside the stop
    // flag's, not after it
    const int done = S->done;
    const int lane = threadIdx.x & (kWave - 1);
    const int g = gridDim.x;
    const int lb = xcd_block(blockIdx.x, g);
    const SliceRange sr = block_slices(lb, g, c.A.n_slices);
    const int s0 = sr.s0, s1 = sr.s1;
    const int sf = s0 + (int)(threadIdx.x / kWave);
    X pf[R] = {};
    SellpHead<V, R> hf{};
    if (sf < s1) {
        const int row0 = sf * kWave * R + lane * R;
        load_rows_tp<R>(c.p, row0, pf);
        hf = sellp_head<V, R>(c.A, sf, row0);
    }
    if (done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            S->halt = 1;
            S->lag = 0;
        }
        return;
    }
    if (blockIdx.x == 0) {
        const double rr = block_sum_partials(c.red, kSlotRR, S->rr_blocks);
        if (threadIdx.x == 0) {
            if (first) S->lag = 0;
            S->rr = rr;
            S->rr_last = rr;
            const int it = S->iter;
            const double scale = S->bb > 0.0 ? S->bb : 1.0;
            if (c.hist && it < S->hist_cap) c.hist[it] = sqrt(rr / scale);
            if (!(rr > S->tol2 * scale) || it >= S->max_iter) S->done = 1;
            if (!isfinite(rr)) { S->nonfinite = 1; S->done = 1; }
            // the FP64 residual replacement follows its trigger at once: the
            // solve stops here (as on a convergence claim) and the host
            // replaces r before the next iteration.  Iterations run on past
            // the trigger can let the FP32 recurrence drift beyond recovery
            // on systems with condition near 1 / eps32.
            if (sqrt(rr) < S->replace_ratio * S->r_ref) {
                S->need_update = 1;
                S->done = 1;
            }
        }
    }
    double acc[3] = {0.0, 0.0, 0.0};
    const int step = (int)(blockDim.x / kWave);
    for (int s = sf; s < s1; s += step) {
        const int row0 = s * kWave * R + lane * R;
        double pi[R], qd[R], rv[R];
#pragma unroll
        for (int j = 0; j < R; ++j) pi[j] = (double)pf[j];
        const X* px = c.p;
        sellp_rows_gh<V, sizeof(V) == 4, R>(c.A, hf, lane, row0, [px](int cc) { return (double)px[cc]; }, pi, qd);
        load_rows_p<R>(c.r, row0, rv);
        X qv[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            qv[j] = (X)qd[j];
            if (row0 + j < c.A.n) {
                const double qj = (double)qv[j];
                acc[0] += pi[j] * qj;
                acc[1] += rv[j] * qj;
                acc[2] += qj * qj;
            }
        }
        store_rows<R>(c.q, row0, c.A.n, qv);
        // the next slice's opening loads
        if (s + step < s1) {
            const int row1 = (s + step) * kWave * R + lane * R;
            load_rows_tp<R>(c.p, row1, pf);
            hf = sellp_head<V, R>(c.A, s + step, row1);
        }
    }
    const double pq1[1] = {acc[0]};
    block_partials<1>(pq1, c.red, kSlotPQ);
    // r.q and q.q go to their own slots (kSlotRQ, kSlotQQ follow kSlotBB)
    const double rq_qq[2] = {acc[1], acc[2]};
    block_partials<2>(rq_qq, c.red, kSlotRQ);
}

// x += alpha_0 P_0 + ... + alpha_{J-1} P_{J-1} over the whole vector: the
// pending steps of a period cut before its last update.
template <class V, class X, int J>
__device__ __forceinline__ void cgm_flush_pending(const CgArgs<V, X>& c, const Scalars* S, int n) {
    if constexpr (J > 0) {
        constexpr int N = Vec16<X>::N;
        const int g = gridDim.x, nv = n / N, stride = g * blockDim.x;
        double al[J];
#pragma unroll
        for (int i = 0; i < J; ++i) al[i] = S->alpha_q[i];
        for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += stride) {
            double xe[N];
            dload<N>(c.x, v, xe);
#pragma unroll
            for (int i = 0; i < J; ++i) {
                X pp[N];
                vload(c.pb[i], v, pp);
#pragma unroll
                for (int e = 0; e < N; ++e) xe[e] += al[i] * (double)pp[e];
            }
            dstore<N>(c.x, v, xe);
        }
        if (blockIdx.x == g - 1) {
            const int i0 = nv * N + (int)threadIdx.x;
            if (i0 < n) {
                double xi = c.x[i0];
#pragma unroll
                for (int i = 0; i < J; ++i) xi += al[i] * (double)c.pb[i][i0];
                c.x[i0] = xi;
            }
        }
    }
}

// Update j = kJ of a period of kM iterations: c.p is P_j, c.p_alt is
// P_{(j+1) mod kM}; the last update (kJ = kM - 1) also adds the period's
// steps to x, reading P_0 .. P_{kM-2} (P_0 before p' replaces it).
template <class V, class X, int kJ, int kM>
__global__ void __launch_bounds__(kCgMaxBlock) k_cgm_update(CgArgs<V, X> c, int n) {
    constexpr int N = Vec16<X>::N;
    constexpr bool kLast = kJ == kM - 1;
    constexpr int kOld = kLast ? kM - 1 : 0;  // earlier directions read for x
    Scalars* S = c.S;
    const int g = gridDim.x;
    const int nv = n / N;
    const int stride = g * blockDim.x;
    const int v0 = blockIdx.x * blockDim.x + threadIdx.x;
    // the stop flag, the partial sums and the first vector are all in flight
    // before the flag is tested
    const int done = S->done;
    constexpr int kSlots[3] = {kSlotPQ, kSlotRQ, kSlotQQ};
    const int gs[3] = {c.g_spmv, c.g_spmv, c.g_spmv};
    double d3[3];
    PartialSums<3> ps;
    ps.issue(c.red, kSlots, gs);
    // first vector in flight while the partial sums are re-added:
    // unconditional loads at a clamped index (the vectors hold n + 4 elements)
    double xv[N];
    X pv[N], rv[N], qv[N], ppv[kOld > 0 ? kOld : 1][N];
    const bool have = v0 < nv;
    {
        const int vs = have ? v0 : nv;
        if constexpr (kLast) {
            dload<N>(c.x, vs, xv);
#pragma unroll
            for (int i = 0; i < kOld; ++i) vload(c.pb[i], vs, ppv[i]);
        }
        vload(c.p, vs, pv);
        vload(c.r, vs, rv);
        vload(c.q, vs, qv);
    }
    if (done) {
        // a period cut after its update kJ - 1: add the pending steps
        if (kJ > 0 && S->lag) cgm_flush_pending<V, X, kJ>(c, S, n);
        return;
    }
    ps.finish(c.red, kSlots, gs, d3);
    const double pq = d3[0], rq = d3[1], qq = d3[2];
    if (!cg_pq_ok(pq)) {
        cgm_flush_pending<V, X, kJ>(c, S, n);  // keep the period's earlier steps
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            S->breakdown = (pq == 0.0);
            S->nonfinite = !isfinite(pq);
            S->done = 1;
            S->lag = 0;
        }
        return;
    }
    const double rr = S->rr;
    const double alpha = rr / pq;
    double rr_next = rr - 2.0 * alpha * rq + alpha * alpha * qq;
    if (!(rr_next > 0.0)) rr_next = 0.0;  // rounding: restart the direction
    const double beta = rr_next / rr;
    double al[kOld > 0 ? kOld : 1];
#pragma unroll
    for (int i = 0; i < kOld; ++i) al[i] = S->alpha_q[i];
    double acc[1] = {0.0};
    auto body = [&](int v, double (&xe)[N], const X (&pe)[N], const X (&re)[N], const X (&qe)[N],
                    const X (&ppe)[kOld > 0 ? kOld : 1][N]) {
        X rn[N], pn[N];
#pragma unroll
        for (int e = 0; e < N; ++e) {
            const double p0 = (double)pe[e];
            if constexpr (kLast) {
#pragma unroll
                for (int i = 0; i < kOld; ++i) xe[e] += al[i] * (double)ppe[i][e];
                xe[e] += alpha * p0;
            }
            rn[e] = (X)((double)re[e] - alpha * (double)qe[e]);
            pn[e] = (X)((double)rn[e] + beta * p0);
            acc[0] += (double)rn[e] * (double)rn[e];
        }
        if constexpr (kLast) dstore<N>(c.x, v, xe);
        vstore(c.r, v, rn);
        vstore(c.p_alt, v, pn);
    };
    if (have) body(v0, xv, pv, rv, qv, ppv);
    for (int v = v0 + stride; v < nv; v += stride) {
        double xe[N];
        X pe[N], re[N], qe[N], ppe[kOld > 0 ? kOld : 1][N];
        if constexpr (kLast) {
            dload<N>(c.x, v, xe);
#pragma unroll
            for (int i = 0; i < kOld; ++i) vload(c.pb[i], v, ppe[i]);
        }
        vload(c.p, v, pe);
        vload(c.r, v, re);
        vload(c.q, v, qe);
        body(v, xe, pe, re, qe, ppe);
    }
    if (blockIdx.x == g - 1) {
        const int i = nv * N + (int)threadIdx.x;
        if (i < n) {
            const double p0 = (double)c.p[i];
            if constexpr (kLast) {
                double xi = c.x[i];
#pragma unroll
                for (int o = 0; o < kOld; ++o) xi += al[o] * (double)c.pb[o][i];
                xi += alpha * p0;
                c.x[i] = xi;
            }
            const X r2 = (X)((double)c.r[i] - alpha * (double)c.q[i]);
            c.r[i] = r2;
            c.p_alt[i] = (X)((double)r2 + beta * p0);
            acc[0] += (double)r2 * (double)r2;
        }
    }
    block_partials<1>(acc, c.red, kSlotRR);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S->alpha = alpha;
        S->beta = beta;
        S->pq = pq;
        S->rr_blocks = gridDim.x;
        if constexpr (!kLast) S->alpha_q[kJ] = alpha;
        if constexpr (kJ == 0) S->lag = 1;
        ++S->iter;
    }
}

// ------------------------------------------------------------------ CG, one kernel per iteration
// The whole iteration k -> k+1 in one launch.  Every block re-adds the four
// partial sums of the previous launch (p_k.q_k, r_k.q_k, q_k.q_k, r_k.r_k) in
// the same fixed order, so all blocks hold bitwise identical alpha_k, beta_k
// (beta from the merged recurrence as in the two-kernel form) and the same
// stop decision.  Each block then, for its own rows,
//   r_{k+1} = r_k - alpha q_k,  p_{k+1} = r_{k+1} + beta p_k,
//   q_{k+1} = A p_{k+1},
// where the SpMV forms every gathered p_{k+1}[j] from r_k[j], q_k[j] and
// p_k[j] with the same two roundings as the owner of row j, so a gathered
// element and the stored one are the same value.  r, q and p are
// double-buffered (the kernels of parity pi read buffer pi and write buffer
// 1 - pi); the partial sums likewise use the slot set of the parity.  x
// follows in pairs as in the two-kernel form: the odd kernel adds
// alpha_{k-1} p_{k-1} + alpha_k p_k, p_{k-1} being its own rows of the buffer
// it is about to overwrite.
// Stop (convergence on r_k.r_k, breakdown, non-finite, or k reaching kstop):
// every block sees it at once; the odd kernel first adds the pending
// alpha_{k-1} p_{k-1}.  Block 0 raises fstop[pi]; a kernel that finds the
// flag of the previous parity raised only passes it on.  A chunk holds an
// even number of kernels, so it starts and ends with everything in buffer 0.
// k_cgf_first (after every true-residual computation) forms q = A p and the
// four sums of buffer 0.
constexpr int kSetF = 5;  // slot set pi: p.q, r.q, q.q, r.r at kSetF + 4 pi + 0..3
static_assert(kSetF + 8 <= kMaxRed, "two slot sets of four");

template <class V, class X>
struct CgfArgs {
    SellPView<V> A;
    X* r[2];
    X* q[2];
    X* p[2];
    double* x;
    Scalars* S;
    double* hist;
    RedScratch red;
};

template <class V, class X, int R>
__global__ void __launch_bounds__(kCgMaxBlock) k_cgf_first(CgfArgs<V, X> c) {
    const int lane = threadIdx.x & (kWave - 1);
    const int g = gridDim.x;
    const int lb = xcd_block(blockIdx.x, g);
    const SliceRange sr = block_slices(lb, g, c.A.n_slices);
    const int s0 = sr.s0, s1 = sr.s1;
    const X* __restrict__ p = c.p[0];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int s = s0 + (int)(threadIdx.x / kWave); s < s1; s += (int)(blockDim.x / kWave)) {
        const int row0 = s * kWave * R + lane * R;
        double pi[R], qd[R], rv[R];
        load_rows_p<R>(p, row0, pi);
        sellp_rows<V, X, sizeof(V) == 4, R>(c.A, s, lane, row0, p, pi, qd);
        load_rows_p<R>(c.r[0], row0, rv);
        X qv[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            qv[j] = (X)qd[j];
            if (row0 + j < c.A.n) {
                const double qj = (double)qv[j];
                acc[0] += pi[j] * qj;
                acc[1] += rv[j] * qj;
                acc[2] += qj * qj;
                acc[3] += rv[j] * rv[j];
            }
        }
        store_rows<R>(c.q[0], row0, c.A.n, qv);
    }
    block_partials<4>(acc, c.red, kSetF);
}

// The two roundings of the direction update, shared by the owner of a row and
// by every gather of it.
template <class X>
struct CgfStep {
    double alpha, beta;
    __device__ __forceinline__ X r_next(double r, double q) const { return (X)fma(-alpha, q, r); }
    __device__ __forceinline__ X p_next(X rn, double p) const { return (X)fma(beta, p, (double)rn); }
};

// The stop test and bookkeeping shared by the one-kernel iterations: every
// block evaluates the same sums; block 0 records the history entry, the
// replacement trigger and, on a stop, the flags.
__device__ __forceinline__ bool cgf_stop(Scalars* S, double* hist, double pq, double rr, int k, int kstop, int odd) {
    const double scale = S->bb > 0.0 ? S->bb : 1.0;
    const bool conv = !(rr > S->tol2 * scale);
    const bool fin = isfinite(rr);
    const bool pq_ok = cg_pq_ok(pq);
    // a due FP64 residual replacement stops the solve too: the host replaces
    // r before the next iteration
    const bool replace = sqrt(rr) < S->replace_ratio * S->r_ref;
    const bool stop = conv || !fin || !pq_ok || replace || k >= kstop;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const int it = S->iter;
        if (hist && it < S->hist_cap) hist[it] = sqrt(rr / scale);
        S->rr = rr;
        S->rr_last = rr;
        if (replace) S->need_update = 1;
        if (stop) {
            S->fstop[odd] = 1;
            S->done = 1;
            if (!conv) {
                if (!fin) S->nonfinite = 1;
                else if (!pq_ok) {
                    S->breakdown = (pq == 0.0);
                    S->nonfinite = !isfinite(pq);
                }
            }
        }
    }
    return stop;
}

__device__ __forceinline__ void cgf_record(Scalars* S, double alpha, double beta, double pq, int odd) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S->alpha = alpha;
        S->beta = beta;
        S->pq = pq;
        if (!odd) S->alpha_prev = alpha;
        ++S->iter;
    }
}

template <class V, class X, int R, int kOdd>
__global__ void __launch_bounds__(kCgMaxBlock) k_cgf_iter(CgfArgs<V, X> c, int k, int kstop) {
    Scalars* S = c.S;
    if (S->fstop[1 - kOdd]) {
        if (blockIdx.x == 0 && threadIdx.x == 0) S->fstop[kOdd] = 1;
        return;
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int g = gridDim.x;
    const int lb = xcd_block(blockIdx.x, g);
    const SliceRange sr = block_slices(lb, g, c.A.n_slices);
    const int s0 = sr.s0, s1 = sr.s1;
    const int n = c.A.n;
    const X* __restrict__ rin = c.r[kOdd];
    const X* __restrict__ qin = c.q[kOdd];
    const X* __restrict__ pin = c.p[kOdd];
    X* __restrict__ rout = c.r[1 - kOdd];
    X* __restrict__ qout = c.q[1 - kOdd];
    X* __restrict__ pout = c.p[1 - kOdd];
    const int sw = s0 + (int)(threadIdx.x / kWave);  // this wave's first slice
    const int sstep = (int)(blockDim.x / kWave);
    constexpr int kIn[4] = {kSetF + 4 * kOdd, kSetF + 4 * kOdd + 1, kSetF + 4 * kOdd + 2, kSetF + 4 * kOdd + 3};
    const int gs[4] = {g, g, g, g};
    double d4[4];
    PartialSums<4> ps;
    ps.issue(c.red, kIn, gs);
    // the first slice's own rows are in flight while the sums are re-added
    X rv[R], qv[R], pv[R];
    if (sw < s1) {
        const int row0 = sw * kWave * R + lane * R;
        load_rows_tp<R>(rin, row0, rv);
        load_rows_tp<R>(qin, row0, qv);
        load_rows_tp<R>(pin, row0, pv);
    }
    ps.finish(c.red, kIn, gs, d4);
    const double pq = d4[0], rq = d4[1], qq = d4[2], rr = d4[3];
    if (cgf_stop(S, c.hist, pq, rr, k, kstop, kOdd)) {
        if constexpr (kOdd) {  // the pending alpha_{k-1} p_{k-1} of the pair
            const double ap = S->alpha_prev;
            for (int s = sw; s < s1; s += sstep) {
                const int row0 = s * kWave * R + lane * R;
                double xv[R], pp[R];
                load_rows_p<R>(c.x, row0, xv);
                load_rows_p<R>(pout, row0, pp);
#pragma unroll
                for (int j = 0; j < R; ++j) xv[j] += ap * pp[j];
                store_rows<R>(c.x, row0, n, xv);
            }
        }
        return;
    }
    const double alpha = rr / pq;
    double rr_next = rr - 2.0 * alpha * rq + alpha * alpha * qq;
    if (!(rr_next > 0.0)) rr_next = 0.0;  // rounding: restart the direction
    const CgfStep<X> st{alpha, rr_next / rr};
    const double ap = kOdd ? S->alpha_prev : 0.0;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    auto pnext = [&](int j) -> double {
        return (double)st.p_next(st.r_next((double)rin[j], (double)qin[j]), (double)pin[j]);
    };
    for (int s = sw; s < s1; s += sstep) {
        const int row0 = s * kWave * R + lane * R;
        if (s != sw) {
            load_rows_tp<R>(rin, row0, rv);
            load_rows_tp<R>(qin, row0, qv);
            load_rows_tp<R>(pin, row0, pv);
        }
        X rn[R], pn[R];
        double pnd[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            rn[j] = st.r_next((double)rv[j], (double)qv[j]);
            pn[j] = st.p_next(rn[j], (double)pv[j]);
            pnd[j] = (double)pn[j];
        }
        if constexpr (kOdd) {
            double xv[R], pp[R];
            load_rows_p<R>(c.x, row0, xv);
            load_rows_p<R>(pout, row0, pp);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                xv[j] += ap * pp[j];
                xv[j] += alpha * (double)pv[j];
            }
            store_rows<R>(c.x, row0, n, xv);
        }
        double qd[R];
        sellp_rows_g<V, sizeof(V) == 4, R, decltype(pnext), R == 4 ? 1 : 2>(c.A, s, lane, row0, pnext, pnd, qd);
        X qn[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            qn[j] = (X)qd[j];
            if (row0 + j < n) {
                const double qj = (double)qn[j], rj = (double)rn[j];
                acc[0] += pnd[j] * qj;
                acc[1] += rj * qj;
                acc[2] += qj * qj;
                acc[3] += rj * rj;
            }
        }
        store_rows<R>(rout, row0, n, rn);
        store_rows<R>(pout, row0, n, pn);
        store_rows<R>(qout, row0, n, qn);
    }
    block_partials<4>(acc, c.red, kSetF + 4 * (1 - kOdd));
    cgf_record(S, alpha, st.beta, pq, kOdd);
}

// ------------------------------------------------------------------ CG, one kernel, LDS window (FP32)
// The one-kernel iteration with the direction staged in LDS.  A block (one
// per CU: it takes 158 KB of the 160 KB LDS of a gfx950 CU) owns a contiguous
// range of rows; its row window is the union of the per-slice windows of the
// layout (own rows and every column they touch -- with RCM on a 128^3 mesh a
// block of 8192 rows spans at most ~33K rows).  The block
//   1. loads its own rows' r_k, q_k, p_k (each wave up to kSpw slices, kept
//      in registers) and, in float4 groups with coalesced loads, the rest of
//      the window; the partial sums of the previous launch are re-added while
//      these loads are in flight;
//   2. forms r_{k+1}, p_{k+1} of its own rows (global stores, x update) and
//      p_{k+1} over the window into LDS;
//   3. after one barrier, runs the SpMV of its own rows with every gather
//      served from LDS.
// A block whose window exceeds the LDS gathers from memory as k_cgf_iter does
// (same values).  The launch guarantees at most kSpw slices per wave.
constexpr int kLdsFloats = 40448;
constexpr int kHaloBatch = 2;  // float4 window groups per thread in flight

template <int R, int kOdd, int kSpw>
__global__ void __launch_bounds__(kCgMaxBlock) k_cgl_iter(CgfArgs<float, float> c, int k, int kstop) {
    using X = float;
    Scalars* S = c.S;
    if (S->fstop[1 - kOdd]) {
        if (blockIdx.x == 0 && threadIdx.x == 0) S->fstop[kOdd] = 1;
        return;
    }
    __shared__ float pw[kLdsFloats];  // p_{k+1} of window row j at pw[j - base]
    __shared__ int wb[2];
    const int lane = threadIdx.x & (kWave - 1);
    const int g = gridDim.x;
    const int lb = xcd_block(blockIdx.x, g);
    const SliceRange sr = block_slices(lb, g, c.A.n_slices);
    const int s0 = sr.s0, s1 = sr.s1;
    const int n = c.A.n;
    const X* __restrict__ rin = c.r[kOdd];
    const X* __restrict__ qin = c.q[kOdd];
    const X* __restrict__ pin = c.p[kOdd];
    X* __restrict__ rout = c.r[1 - kOdd];
    X* __restrict__ qout = c.q[1 - kOdd];
    X* __restrict__ pout = c.p[1 - kOdd];
    const int sw = s0 + (int)(threadIdx.x / kWave);
    const int sstep = (int)(blockDim.x / kWave);
    constexpr int kIn[4] = {kSetF + 4 * kOdd, kSetF + 4 * kOdd + 1, kSetF + 4 * kOdd + 2, kSetF + 4 * kOdd + 3};
    const int gs[4] = {g, g, g, g};
    double d4[4];
    PartialSums<4> ps;
    ps.issue(c.red, kIn, gs);
    // row window of the block (slice windows loaded before the own rows)
    int wlo = 0x7fffffff, whi = -1;
    for (int q = s0 + (int)threadIdx.x; q < s1; q += (int)blockDim.x) {
        wlo = min(wlo, c.A.win_lo[q]);
        whi = max(whi, c.A.win_hi[q]);
    }
    // own rows of this wave's slices
    X rv[kSpw][R], qv[kSpw][R], pv[kSpw][R];
#pragma unroll
    for (int t = 0; t < kSpw; ++t) {
        const int s = sw + t * sstep;
        if (s < s1) {
            const int row0 = s * kWave * R + lane * R;
            load_rows_tp<R>(rin, row0, rv[t]);
            load_rows_tp<R>(qin, row0, qv[t]);
            load_rows_tp<R>(pin, row0, pv[t]);
        }
    }
    if (threadIdx.x == 0) {
        wb[0] = 0x7fffffff;
        wb[1] = -1;
    }
    __syncthreads();
    if (whi >= 0) {
        atomicMin(&wb[0], wlo);
        atomicMax(&wb[1], whi);
    }
    __syncthreads();
    const int ra = s0 * kWave * R, rb = min(s1 * kWave * R, n);  // own rows [ra, rb)
    const int base = wb[0] & ~3;
    const int nv4 = wb[1] >= base ? (wb[1] - base) / 4 + 1 : 0;   // float4 groups of the window
    const bool fits = s0 < s1 && 4 * nv4 <= kLdsFloats;
    const int vo0 = fits ? (ra - base) / 4 : 0;                     // own groups [vo0, vo1)
    const int vo1 = fits ? vo0 + ((rb - ra + 3) >> 2) : 0;
    const int nhalo = fits ? vo0 + (nv4 - vo1) : 0;
    auto halo_group = [&](int h) { return h < vo0 ? h : h - vo0 + vo1; };
    // the first halo groups in flight with the sums
    Pack<float, 4> hr[kHaloBatch], hq[kHaloBatch], hp[kHaloBatch];
    auto load_group = [&](int v, Pack<float, 4>& a, Pack<float, 4>& b, Pack<float, 4>& p) {
        const int j0 = base + 4 * v;
        if (j0 + 4 <= n) {
            a = ldp<float, 4>(rin + j0);
            b = ldp<float, 4>(qin + j0);
            p = ldp<float, 4>(pin + j0);
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const bool in = j0 + e < n;
                a.v[e] = in ? rin[j0 + e] : 0.f;
                b.v[e] = in ? qin[j0 + e] : 0.f;
                p.v[e] = in ? pin[j0 + e] : 0.f;
            }
        }
    };
#pragma unroll
    for (int i = 0; i < kHaloBatch; ++i) {
        const int h = (int)threadIdx.x + i * (int)blockDim.x;
        if (h < nhalo) load_group(halo_group(h), hr[i], hq[i], hp[i]);
    }
    ps.finish(c.red, kIn, gs, d4);
    const double pq = d4[0], rq = d4[1], qq = d4[2], rr = d4[3];
    if (cgf_stop(S, c.hist, pq, rr, k, kstop, kOdd)) {
        if constexpr (kOdd) {  // the pending alpha_{k-1} p_{k-1} of the pair
            const double ap = S->alpha_prev;
#pragma unroll
            for (int t = 0; t < kSpw; ++t) {
                const int s = sw + t * sstep;
                if (s < s1) {
                    const int row0 = s * kWave * R + lane * R;
                    double xe[R], pe[R];
                    load_rows_p<R>(c.x, row0, xe);
                    load_rows_p<R>(pout, row0, pe);
#pragma unroll
                    for (int j = 0; j < R; ++j) xe[j] += ap * pe[j];
                    store_rows<R>(c.x, row0, n, xe);
                }
            }
        }
        return;
    }
    const double alpha = rr / pq;
    double rr_next = rr - 2.0 * alpha * rq + alpha * alpha * qq;
    if (!(rr_next > 0.0)) rr_next = 0.0;  // rounding: restart the direction
    const CgfStep<X> st{alpha, rr_next / rr};
    const double ap = kOdd ? S->alpha_prev : 0.0;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    // own rows: r_{k+1}, p_{k+1} (odd: x, which needs p_{k-1} from pout
    // before p_{k+1} replaces it); p_{k+1} into the window
    X rn[kSpw][R], pn[kSpw][R];
#pragma unroll
    for (int t = 0; t < kSpw; ++t) {
        const int s = sw + t * sstep;
        if (s < s1) {
            const int row0 = s * kWave * R + lane * R;
#pragma unroll
            for (int j = 0; j < R; ++j) {
                rn[t][j] = st.r_next((double)rv[t][j], (double)qv[t][j]);
                pn[t][j] = st.p_next(rn[t][j], (double)pv[t][j]);
                if (row0 + j < n) acc[3] += (double)rn[t][j] * (double)rn[t][j];
            }
            if constexpr (kOdd) {
                double xe[R], pe[R];
                load_rows_p<R>(c.x, row0, xe);
                load_rows_p<R>(pout, row0, pe);
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    xe[j] += ap * pe[j];
                    xe[j] += alpha * (double)pv[t][j];
                }
                store_rows<R>(c.x, row0, n, xe);
            }
            store_rows<R>(rout, row0, n, rn[t]);
            store_rows<R>(pout, row0, n, pn[t]);
            if (fits) {
                if (row0 + R <= n) {
                    Pack<float, R> pk;
#pragma unroll
                    for (int j = 0; j < R; ++j) pk.v[j] = pn[t][j];
                    *reinterpret_cast<Pack<float, R>*>(pw + (row0 - base)) = pk;
                } else {
#pragma unroll
                    for (int j = 0; j < R; ++j)
                        if (row0 + j < n) pw[row0 + j - base] = pn[t][j];
                }
            }
        }
    }
    if (fits) {
        // the rest of the window
        auto put_group = [&](int v, const Pack<float, 4>& a, const Pack<float, 4>& b, const Pack<float, 4>& p) {
            Pack<float, 4> o;
#pragma unroll
            for (int e = 0; e < 4; ++e) o.v[e] = st.p_next(st.r_next((double)a.v[e], (double)b.v[e]), (double)p.v[e]);
            *reinterpret_cast<Pack<float, 4>*>(pw + 4 * v) = o;
        };
#pragma unroll
        for (int i = 0; i < kHaloBatch; ++i) {
            const int h = (int)threadIdx.x + i * (int)blockDim.x;
            if (h < nhalo) put_group(halo_group(h), hr[i], hq[i], hp[i]);
        }
        for (int h0 = (int)threadIdx.x + kHaloBatch * (int)blockDim.x; h0 < nhalo;
             h0 += kHaloBatch * (int)blockDim.x) {
#pragma unroll
            for (int i = 0; i < kHaloBatch; ++i) {
                const int h = h0 + i * (int)blockDim.x;
                if (h < nhalo) load_group(halo_group(h), hr[i], hq[i], hp[i]);
            }
#pragma unroll
            for (int i = 0; i < kHaloBatch; ++i) {
                const int h = h0 + i * (int)blockDim.x;
                if (h < nhalo) put_group(halo_group(h), hr[i], hq[i], hp[i]);
            }
        }
    }
    __syncthreads();
    auto pl = [&](int j) -> double { return (double)pw[j - base]; };
    auto pnext = [&](int j) -> double {
        return (double)st.p_next(st.r_next((double)rin[j], (double)qin[j]), (double)pin[j]);
    };
#pragma unroll
    for (int t = 0; t < kSpw; ++t) {
        const int s = sw + t * sstep;
        if (s < s1) {
            const int row0 = s * kWave * R + lane * R;
            double pnd[R], qd[R];
#pragma unroll
            for (int j = 0; j < R; ++j) pnd[j] = (double)pn[t][j];
            if (fits) sellp_rows_g<float, true, R, decltype(pl), R == 1 ? 4 : 2>(c.A, s, lane, row0, pl, pnd, qd);
            else sellp_rows_g<float, true, R, decltype(pnext), R == 4 ? 1 : 2>(c.A, s, lane, row0, pnext, pnd, qd);
            X qn[R];
#pragma unroll
            for (int j = 0; j < R; ++j) {
                qn[j] = (X)qd[j];
                if (row0 + j < n) {
                    const double qj = (double)qn[j];
                    acc[0] += pnd[j] * qj;
                    acc[1] += (double)rn[t][j] * qj;
                    acc[2] += qj * qj;
                }
            }
            store_rows<R>(qout, row0, n, qn);
        }
    }
    block_partials<4>(acc, c.red, kSetF + 4 * (1 - kOdd));
    cgf_record(S, alpha, st.beta, pq, kOdd);
}

// SELL-P FP32 difference form: diag32 = row sum of the FP64 values, the
// off-diagonals added in their storage order (the same for every R).
template <int R>
__global__ void __launch_bounds__(kBlock) k_sellp_rowsum(SellPView<double> A, float* __restrict__ diag32) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * kWavesPerBlock;
    for (int s = wave; s < A.n_slices; s += nwaves) {
        const int base = A.slice_ptr[s];
        const int w = (A.slice_ptr[s + 1] - base) / (kWave * R);
        for (int j = 0; j < R; ++j) {
            const int row = s * kWave * R + lane * R + j;
            if (row >= A.n) continue;
            double sum = A.diag[row];
            for (int k = 0; k < w; ++k) sum += A.val[base + (k / 2) * 128 * R + lane * 2 * R + 2 * j + (k & 1)];
            diag32[row] = (float)sum;
        }
    }
}

// FP32 difference-form values: the diagonal slot of each SELL row receives the
// row sum computed in FP64 from the FP64 values (padding slots sum to 0).
[[maybe_unused]] static __global__ void __launch_bounds__(kBlock) k_sell_rowsum_diag(SellView<double> A, float* __restrict__ val32) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * kWavesPerBlock;
    for (int s = wave; s < A.n_slices; s += nwaves) {
        const int row = s * kWave + lane;
        if (row >= A.n) continue;
        const int base = A.slice_ptr[s];
        const int w = (A.slice_ptr[s + 1] - base) >> 6;
        double sum = 0.0;
        int dslot = -1;
        for (int k = 0; k < w; ++k) {
            const int q = base + k * kWave + lane;
            sum += A.val[q];
            if (dslot < 0 && A.col[q] == row) dslot = q;
        }
        if (dslot >= 0) val32[dslot] = (float)sum;
    }
}

// ||b - A x||^2 and ||b||^2 without writing anything (true residual).
template <class V>
__global__ void __launch_bounds__(kBlock) k_true_residual(SellView<V> A, const double* __restrict__ b,
                                                          const double* __restrict__ x, RedScratch red) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * kWavesPerBlock;
    double acc[2] = {0.0, 0.0};
    for (int s = wave; s < A.n_slices; s += nwaves) {
        const int row = s * kWave + lane;
        const double ax = sell_row<V, double, double>(A, s, lane, x);
        if (row < A.n) {
            const double rv = b[row] - ax;
            acc[0] += rv * rv;
            acc[1] += b[row] * b[row];
        }
    }
    block_partials<2>(acc, red);
}

}  // namespace ocsl
