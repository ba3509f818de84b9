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
__global__ void __launch_bounds__(kBlock) k_dots(int n, DotArgs<NR> args, RedScratch red,
                                                 double* out) {
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
    grid_reduce<NR>(acc, red, [&](const double* t) {
#pragma unroll
        for (int r = 0; r < NR; ++r) out[r] = t[r];
    });
}

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
template <class V, class X>
struct CgArgs {
    SellView<V> A;
    X* r;
    X* p;
    X* q;
    double* x;      // solution, FP64 in both precisions
    Scalars* S;
    double* hist;
    RedScratch red;
};

// r = b - A x (FP64), stored as r64 or r32; p = r on the first call; rr, bb.
// Finaliser (re)starts the recurrence: rr, bb, r_ref, history entry 0 when
// `first`, convergence flag.
template <bool kMixed>
__global__ void __launch_bounds__(kBlock) k_cg_residual(SellView<double> A, const double* __restrict__ b,
                                                       const double* __restrict__ x,
                                                       double* __restrict__ r64, float* __restrict__ r32,
                                                       float* __restrict__ p32,
                                                       double* __restrict__ p64, Scalars* S, double* hist,
                                                       RedScratch red, int first) {
    const int lane = threadIdx.x & (kWave - 1);
    const int g = gridDim.x;
    const int lb = xcd_block(blockIdx.x, g);
    const int spb = (A.n_slices + g - 1) / g;
    const int s0 = lb * spb, s1 = min(A.n_slices, s0 + spb);
    double acc[2] = {0.0, 0.0};
    for (int s = s0 + threadIdx.x / kWave; s < s1; s += kWavesPerBlock) {
        const int row = s * kWave + lane;
        const double ax = sell_row<double, double, double>(A, s, lane, x);
        if (row < A.n) {
            const double bv = b[row];
            const double rv = bv - ax;
            acc[0] += rv * rv;
            acc[1] += bv * bv;
            if (kMixed) {
                r32[row] = (float)rv;
                if (first) p32[row] = (float)rv;
            } else {
                r64[row] = rv;
                if (first) p64[row] = rv;
            }
        }
    }
    grid_reduce<2>(acc, red, [&](const double* t) {
        const double rr = t[0], bb = t[1];
        S->rr = rr;
        S->r_ref = sqrt(rr);
        if (first) {
            S->bb = bb;
            S->rr0 = rr;
            const double scale = bb > 0.0 ? bb : 1.0;
            if (hist && S->hist_cap > 0) hist[0] = sqrt(rr / scale);
            S->done = (rr <= S->tol2 * scale) ? 1 : 0;
        } else {
            const double scale = S->bb > 0.0 ? S->bb : 1.0;
            S->done = (rr <= S->tol2 * scale) ? 1 : 0;
        }
        S->halt = S->done;
        S->need_update = 0;
        if (!isfinite(rr)) { S->nonfinite = 1; S->done = 1; S->halt = 1; }
    });
}

// q = A p ; alpha = rr / (p.q)
template <class V, class X>
__global__ void __launch_bounds__(kBlock) k_cg_spmv(CgArgs<V, X> c) {
    Scalars* S = c.S;
    if (S->done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) S->halt = 1;
        return;
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int g = gridDim.x;
    const int lb = xcd_block(blockIdx.x, g);
    const int spb = (c.A.n_slices + g - 1) / g;
    const int s0 = lb * spb, s1 = min(c.A.n_slices, s0 + spb);
    double acc[1] = {0.0};
    for (int s = s0 + threadIdx.x / kWave; s < s1; s += kWavesPerBlock) {
        const int row = s * kWave + lane;
        X qv;
        if constexpr (sizeof(V) == 4)
            qv = (X)sell_row_diff<V, X>(c.A, s, lane, row, c.p);
        else
            qv = sell_row<V, X, X>(c.A, s, lane, c.p);
        if (row < c.A.n) {
            c.q[row] = qv;
            acc[0] += (double)c.p[row] * (double)qv;
        }
    }
    grid_reduce<1>(acc, c.red, [&](const double* t) {
        const double pq = t[0];
        if (pq == 0.0 || !isfinite(pq)) {
            S->breakdown = (pq == 0.0);
            S->nonfinite = !isfinite(pq);
            S->done = 1;
            S->alpha = 0.0;
        } else {
            S->alpha = S->rr / pq;
        }
        S->pq = pq;
    });
}

// r -= alpha q ; rr_new = r.r ; beta = rr_new / rr ; convergence tests
template <class V, class X>
__global__ void __launch_bounds__(kBlock) k_cg_update_r(CgArgs<V, X> c, int n) {
    Scalars* S = c.S;
    if (S->done) return;
    const double alpha = S->alpha;
    double acc[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const X rv = (X)((double)c.r[i] - alpha * (double)c.q[i]);
        c.r[i] = rv;
        acc[0] += (double)rv * (double)rv;
    }
    grid_reduce<1>(acc, c.red, [&](const double* t) {
        const double rr_new = t[0];
        S->beta = rr_new / S->rr;
        S->rr = rr_new;
        const int it = ++S->iter;
        const double scale = S->bb > 0.0 ? S->bb : 1.0;
        if (c.hist && it < S->hist_cap) c.hist[it] = sqrt(rr_new / scale);
        if (rr_new <= S->tol2 * scale || it >= S->max_iter) S->done = 1;
        if (!isfinite(rr_new)) { S->nonfinite = 1; S->done = 1; }
        if (sqrt(rr_new) < S->replace_ratio * S->r_ref) S->need_update = 1;
    });
}

// x += alpha p (in FP64 for both precisions) ; p = r + beta p
template <class V, class X>
__global__ void __launch_bounds__(kBlock) k_cg_update_xp(CgArgs<V, X> c, int n) {
    Scalars* S = c.S;
    if (S->halt) return;
    const double alpha = S->alpha, beta = S->beta;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double pv = (double)c.p[i];
        c.x[i] += alpha * pv;
        c.p[i] = (X)((double)c.r[i] + beta * pv);
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
                                                          const double* __restrict__ x, RedScratch red,
                                                          double* out) {
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
    grid_reduce<2>(acc, red, [&](const double* t) {
        out[0] = t[0];
        out[1] = t[1];
    });
}

}  // namespace ocsl
