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
__global__ void __launch_bounds__(kBlock) k_sell_axpby(SellView<V> A, double a,
                                                       const double* __restrict__ x, double b,
                                                       const double* __restrict__ z,
                                                       double* __restrict__ out) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * kWavesPerBlock;
    for (int s = wave; s < A.n_slices; s += nwaves) {
        const int row = s * kWave + lane;
        const double dum = sell_row<V, double, double>(A, s, lane, x);
        if (row < A.n) {
            double o = (a == 1.0) ? dum : a * dum;
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
constexpr int kCgMaxBlock = 1024;  // CG kernels run with 64..1024 threads per block
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
template <int K>
__device__ __forceinline__ void block_sum_partials_k(RedScratch red, const int (&slot)[K], int G,
                                                     double (&out)[K]) {
    __shared__ double sh[K][16];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int nw = blockDim.x / kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (int i = tid; i < G; i += blockDim.x) s += red.partial[slot[k] * kMaxRedBlocks + i];
        s = wave_sum(s);
        if (lane == 0) sh[k][wid] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double t = 0.0;
        for (int w = 0; w < nw; ++w) t += sh[k][w];
        out[k] = t;
    }
    __syncthreads();
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
    const int spb = (A.n_slices + g - 1) / g;
    const int s0 = lb * spb, s1 = min(A.n_slices, s0 + spb);
    double acc[2] = {0.0, 0.0};
    for (int s = s0 + (int)(threadIdx.x / kWave); s < s1; s += (int)(blockDim.x / kWave)) {
        const int row0 = s * kWave * R + lane * R;
        double xi[R], ax[R], bv[R];
        load_rows<R>(x, row0, A.n, xi);
        sellp_rows<double, double, false, R>(A, s, lane, row0, x, xi, ax);
        load_rows<R>(b, row0, A.n, bv);
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
        if (!isfinite(rr)) { S->nonfinite = 1; S->done = 1; S->halt = 1; }
    }
};

// K1: q = A p, partials of p.q.  FP32 storage uses the difference form.
template <class V, class X, int R>
__global__ void __launch_bounds__(kCgMaxBlock) k_cg_spmv(CgArgs<V, X> c) {
    Scalars* S = c.S;
    if (S->done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) S->halt = 1;
        return;
    }
    const int g = gridDim.x;
    if (blockIdx.x == 0) {
        const double rr = block_sum_partials(c.red, kSlotRR, S->rr_blocks);
        if (threadIdx.x == 0) S->rr = rr;
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int lb = xcd_block(blockIdx.x, g);
    const int spb = (c.A.n_slices + g - 1) / g;
    const int s0 = lb * spb, s1 = min(c.A.n_slices, s0 + spb);
    double acc[1] = {0.0};
    for (int s = s0 + (int)(threadIdx.x / kWave); s < s1; s += (int)(blockDim.x / kWave)) {
        const int row0 = s * kWave * R + lane * R;
        double pi[R], qd[R];
        load_rows<R>(c.p, row0, c.A.n, pi);
        sellp_rows<V, X, sizeof(V) == 4, R>(c.A, s, lane, row0, c.p, pi, qd);
        X qv[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            qv[j] = (X)qd[j];
            if (row0 + j < c.A.n) acc[0] += pi[j] * (double)qv[j];
        }
        store_rows<R>(c.q, row0, c.A.n, qv);
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
    if (S->done) return;
    const int nv = n / N;
    const int stride = gridDim.x * blockDim.x;
    const int v0 = blockIdx.x * blockDim.x + threadIdx.x;
    X rv[kPre][N], qv[kPre][N];
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const int v = v0 + k * stride;
        if (v < nv) {
            vload(c.r, v, rv[k]);
            vload(c.q, v, qv[k]);
        }
    }
    const double pq = block_sum_partials(c.red, kSlotPQ, c.g_spmv);
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
    if (S->halt) return;
    const int g = gridDim.x;
    const int nv = n / N;
    const int stride = g * blockDim.x;
    const int v0 = blockIdx.x * blockDim.x + threadIdx.x;
    double xv[kPre][N];
    X pv[kPre][N], rv[kPre][N];
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const int v = v0 + k * stride;
        if (v < nv) {
            dload<N>(c.x, v, xv[k]);
            vload(c.p, v, pv[k]);
            vload(c.r, v, rv[k]);
        }
    }
    const double pq = block_sum_partials(c.red, kSlotPQ, c.g_spmv);
    if (!cg_pq_ok(pq)) return;
    const double rr_new = block_sum_partials(c.red, kSlotRR, g);
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
// Iterations run in pairs with p double-buffered (p -> p_alt -> p): the even
// update writes p' to the other buffer and leaves x alone; the odd one, with
// both directions at hand, adds alpha_prev p_prev + alpha p to x.  The FP64 x
// is then read and written every other iteration (30 instead of 36 bytes per
// row and iteration); the sums are the same two FP64 additions in the same
// order.  When a pair is cut by convergence after its even update, the odd
// update still adds the pending alpha_prev p_prev (S->lag, cleared by the even
// SpMV of the next pair, so it is read and written in different kernels).
template <class V, class X, int R>
__global__ void __launch_bounds__(kCgMaxBlock) k_cgm_spmv(CgArgs<V, X> c, int even) {
    Scalars* S = c.S;
    if (S->done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            S->halt = 1;
            if (even) S->lag = 0;
        }
        return;
    }
    const int g = gridDim.x;
    if (blockIdx.x == 0) {
        const double rr = block_sum_partials(c.red, kSlotRR, S->rr_blocks);
        if (threadIdx.x == 0) {
            if (even) S->lag = 0;
            S->rr = rr;
            S->rr_last = rr;
            const int it = S->iter;
            const double scale = S->bb > 0.0 ? S->bb : 1.0;
            if (c.hist && it < S->hist_cap) c.hist[it] = sqrt(rr / scale);
            if (!(rr > S->tol2 * scale) || it >= S->max_iter) S->done = 1;
            if (!isfinite(rr)) { S->nonfinite = 1; S->done = 1; }
            if (sqrt(rr) < S->replace_ratio * S->r_ref) S->need_update = 1;
        }
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int lb = xcd_block(blockIdx.x, g);
    const int spb = (c.A.n_slices + g - 1) / g;
    const int s0 = lb * spb, s1 = min(c.A.n_slices, s0 + spb);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int s = s0 + (int)(threadIdx.x / kWave); s < s1; s += (int)(blockDim.x / kWave)) {
        const int row0 = s * kWave * R + lane * R;
        double pi[R], qd[R], rv[R];
        load_rows<R>(c.p, row0, c.A.n, pi);
        sellp_rows<V, X, sizeof(V) == 4, R>(c.A, s, lane, row0, c.p, pi, qd);
        load_rows<R>(c.r, row0, c.A.n, rv);
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
    }
    const double pq1[1] = {acc[0]};
    block_partials<1>(pq1, c.red, kSlotPQ);
    // r.q and q.q go to their own slots (kSlotRQ, kSlotQQ follow kSlotBB)
    const double rq_qq[2] = {acc[1], acc[2]};
    block_partials<2>(rq_qq, c.red, kSlotRQ);
}

template <class V, class X, int kOdd>
__global__ void __launch_bounds__(kCgMaxBlock) k_cgm_update(CgArgs<V, X> c, int n) {
    constexpr int N = Vec16<X>::N;
    Scalars* S = c.S;
    const int g = gridDim.x;
    const int nv = n / N;
    const int stride = g * blockDim.x;
    const int v0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (S->done) {
        // a pair cut after its even update: add the pending alpha_prev p_prev
        if (kOdd && S->lag) {
            const double ap = S->alpha_prev;
            for (int v = v0; v < nv; v += stride) {
                double xe[N];
                X pp[N];
                dload<N>(c.x, v, xe);
                vload(c.p_alt, v, pp);
#pragma unroll
                for (int e = 0; e < N; ++e) xe[e] += ap * (double)pp[e];
                dstore<N>(c.x, v, xe);
            }
            if (blockIdx.x == g - 1) {
                const int i = nv * N + (int)threadIdx.x;
                if (i < n) c.x[i] += ap * (double)c.p_alt[i];
            }
        }
        return;
    }
    double xv[N];
    X pv[N], rv[N], qv[N], ppv[N];
    const bool have = v0 < nv;
    if (have) {  // first vector loaded before the partial sums are re-added
        if (kOdd) {
            dload<N>(c.x, v0, xv);
            vload(c.p_alt, v0, ppv);
        }
        vload(c.p, v0, pv);
        vload(c.r, v0, rv);
        vload(c.q, v0, qv);
    }
    constexpr int kSlots[3] = {kSlotPQ, kSlotRQ, kSlotQQ};
    double d3[3];
    block_sum_partials_k<3>(c.red, kSlots, c.g_spmv, d3);
    const double pq = d3[0], rq = d3[1], qq = d3[2];
    if (!cg_pq_ok(pq)) {
        if (kOdd) {  // keep the even iteration's step
            const double ap = S->alpha_prev;
            for (int v = v0; v < nv; v += stride) {
                double xe[N];
                X pp[N];
                dload<N>(c.x, v, xe);
                vload(c.p_alt, v, pp);
#pragma unroll
                for (int e = 0; e < N; ++e) xe[e] += ap * (double)pp[e];
                dstore<N>(c.x, v, xe);
            }
            if (blockIdx.x == g - 1) {
                const int i = nv * N + (int)threadIdx.x;
                if (i < n) c.x[i] += ap * (double)c.p_alt[i];
            }
        }
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
    const double ap = kOdd ? S->alpha_prev : 0.0;
    double acc[1] = {0.0};
    auto body = [&](int v, double (&xe)[N], const X (&pe)[N], const X (&re)[N], const X (&qe)[N],
                    const X (&ppe)[N]) {
        X rn[N], pn[N];
#pragma unroll
        for (int e = 0; e < N; ++e) {
            const double p0 = (double)pe[e];
            if (kOdd) {
                xe[e] += ap * (double)ppe[e];
                xe[e] += alpha * p0;
            }
            rn[e] = (X)((double)re[e] - alpha * (double)qe[e]);
            pn[e] = (X)((double)rn[e] + beta * p0);
            acc[0] += (double)rn[e] * (double)rn[e];
        }
        if (kOdd) dstore<N>(c.x, v, xe);
        vstore(c.r, v, rn);
        vstore(c.p_alt, v, pn);
    };
    if (have) body(v0, xv, pv, rv, qv, ppv);
    for (int v = v0 + stride; v < nv; v += stride) {
        double xe[N];
        X pe[N], re[N], qe[N], ppe[N];
        if (kOdd) {
            dload<N>(c.x, v, xe);
            vload(c.p_alt, v, ppe);
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
            if (kOdd) {
                double xi = c.x[i];
                xi += ap * (double)c.p_alt[i];
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
        if (!kOdd) {
            S->alpha_prev = alpha;
            S->lag = 1;
        }
        ++S->iter;
    }
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
