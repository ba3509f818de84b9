// ILU(0)-BiCGSTAB.
//   ilu_decompose_device : the reference's fine-grained ILU(0) sweeps
//                          (cmatrix_solver.cpp:519-551), one thread per entry
//                          over precomputed (k, q) slot pairs.
//   triangular solves    : exact, level-scheduled (cmatrix_solver.cpp:555-571,
//                          595-610), or damped Jacobi sweeps (the reference's
//                          commented-out variants, :573-592, :612-631).
//   bicgstab_reference   : CMatrix_solver::ILU_BICGSTAB_solve
//                          (cmatrix_solver.cpp:682-762) with its quirks.
//   bicgstab_tolerance   : right-preconditioned BiCGSTAB (van der Vorst 1992)
//                          to a relative residual, device-resident scalars,
//                          1 or 3 right-hand sides sharing A and M (one matrix
//                          pass serves all of them), FP64 or MIXED (FP32
//                          storage, FP64 arithmetic, FP64 solution and
//                          residual replacement).
#include <algorithm>
#include <cfloat>
#include <cstddef>
#include <cmath>
#include <type_traits>

#include "ocsl_handle.h"
#include "ocsl_kernels.h"

namespace ocsl {

// ------------------------------------------------------------------ ILU sweeps
__global__ void k_ilu_sweep(int nnz, const double* __restrict__ a, const double* __restrict__ old,
                            double* __restrict__ nw, const int* __restrict__ pptr,
                            const int* __restrict__ pk, const int* __restrict__ pq,
                            const int* __restrict__ den) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += gridDim.x * blockDim.x) {
        double tmp = 0.0;
        for (int t = pptr[p]; t < pptr[p + 1]; ++t) tmp += old[pk[t]] * old[pq[t]];
        const int d = den[p];
        nw[p] = d >= 0 ? (a[p] - tmp) / old[d] : a[p] - tmp;
    }
}

__global__ void k_to_float(int n, const double* __restrict__ in, float* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

void ilu_decompose_device(ocsl_solver* h, int sweeps, bool want32, const double* m0) {
    OCSL_REQUIRE(sweeps >= 0, OCSL_ERR_ARG, "ILU sweeps must be >= 0");
    if (!h->ilu_built) {
        build_ilu_tables(h->host, h->ilu, h->stream);
        h->ilu_built = true;
    }
    const int nnz = (int)h->host.col.size();
    h->ilu_m.alloc(nnz);
    h->ilu_old.alloc(nnz);
    // M^0 = A (copy_same_structure, cmatrix_solver.cpp:702), or the
    // caller's M (ILU_decompose starts from whatever M holds, :520)
    double* cur = h->ilu_m.p;
    double* prev = h->ilu_old.p;
    if (sweeps % 2 == 1) std::swap(cur, prev);  // the last sweep lands in ilu_m
    k_permute_in<double><<<grid_for(nnz), kBlock, 0, h->stream>>>(nnz, nullptr, m0 ? m0 : h->csr_val.p, cur);
    ++h->ilu_epoch;  // the level-ordered copies of the factor are stale
    const int g = grid_for(nnz);
    for (int s = 0; s < sweeps; ++s) {
        std::swap(cur, prev);
        k_ilu_sweep<<<g, kBlock, 0, h->stream>>>(nnz, h->csr_val.p, prev, cur, h->ilu.pair_ptr.p,
                                                 h->ilu.pair_k.p, h->ilu.pair_q.p, h->ilu.den.p);
    }
    OCSL_HIP(hipGetLastError());
    if (want32) {
        h->ilu_m32.alloc(nnz);
        k_to_float<<<g, kBlock, 0, h->stream>>>(nnz, h->ilu_m.p, h->ilu_m32.p);
        OCSL_HIP(hipGetLastError());
    }
}

// ------------------------------------------------------------------ triangular solves
struct TriCSR {
    const int* rp;
    const int* col;
    const int* diag;
};

// One row of a triangular solve, at position k of the direction's level
// order: lower: out_i = in_i - sum_{col < i} M x_col (unit diagonal); upper:
// out_i = (in_i - sum_{col > i} M x_col) / M_ii.  The sum is formed first and
// subtracted once, as the reference does, over the row's entries in their
// order.  The row index, its entry range and its diagonal are read by
// position from level-ordered copies, so a row costs three rounds of
// dependent loads (position data, entries, gathered values) rather than five
// (row, row pointers and diagonal slot, entries, values, gathered values).
struct LevelTri {
    const int* rows;
    const int* eptr;
    const int* ecol;
};
template <class VM, class X, int NR>
__device__ __forceinline__ void tri_row_level(int k, LevelTri t, const VM* __restrict__ ev,
                                              const VM* __restrict__ dv, const X* in, X* out, bool lower) {
    const int i = t.rows[k];
    const int p0 = t.eptr[k], p1 = t.eptr[k + 1];
    const double dd = lower ? 1.0 : (double)dv[k];
    double sum[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) sum[c] = 0.0;
    for (int p = p0; p < p1; ++p) {
        const double mv = (double)ev[p];
        const int j = t.ecol[p];
#pragma unroll
        for (int c = 0; c < NR; ++c) sum[c] += mv * (double)out[(size_t)j * NR + c];
    }
#pragma unroll
    for (int c = 0; c < NR; ++c) {
        const double v = (double)in[(size_t)i * NR + c] - sum[c];
        out[(size_t)i * NR + c] = (X)(lower ? v : v / dd);
    }
}

// tri_row_level in two parts, so that a level kernel can load its first
// row's position data together with the halt flag instead of after it.
struct TriPos {
    int i, p0, p1;
    double dd;
};
template <class VM>
__device__ __forceinline__ TriPos tri_pos(int k, LevelTri t, const VM* __restrict__ dv, bool lower) {
    return {t.rows[k], t.eptr[k], t.eptr[k + 1], lower ? 1.0 : (double)dv[k]};
}
template <class VM, class X, int NR>
__device__ __forceinline__ void tri_row_at(const TriPos& q, LevelTri t, const VM* __restrict__ ev, const X* in,
                                           X* out, bool lower) {
    double sum[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) sum[c] = 0.0;
    for (int p = q.p0; p < q.p1; ++p) {
        const double mv = (double)ev[p];
        const int j = t.ecol[p];
#pragma unroll
        for (int c = 0; c < NR; ++c) sum[c] += mv * (double)out[(size_t)j * NR + c];
    }
#pragma unroll
    for (int c = 0; c < NR; ++c) {
        const double v = (double)in[(size_t)q.i * NR + c] - sum[c];
        out[(size_t)q.i * NR + c] = (X)(lower ? v : v / q.dd);
    }
}

template <class VM, class X, int NR>
__global__ void k_tri_level(LevelTri t, int begin, int count, const VM* __restrict__ ev, const VM* __restrict__ dv,
                            const X* in, X* out, int lower, const Scalars* S) {
    const int k0 = blockIdx.x * blockDim.x + threadIdx.x;
    TriPos q{0, 0, 0, 1.0};
    if (k0 < count) q = tri_pos<VM>(begin + k0, t, dv, lower != 0);
    if (S && S->halt) return;
    if (k0 < count) tri_row_at<VM, X, NR>(q, t, ev, in, out, lower != 0);
    for (int k = k0 + gridDim.x * blockDim.x; k < count; k += gridDim.x * blockDim.x)
        tri_row_level<VM, X, NR>(begin + k, t, ev, dv, in, out, lower != 0);
}

// Two independent solves with the same factor and direction in one pass
// over the levels (each vector's rows computed exactly as alone).
template <class VM>
__global__ void k_tri_level_pair(LevelTri t, int begin, int count, const VM* __restrict__ ev,
                                 const VM* __restrict__ dv, const double* in0, double* out0, const double* in1,
                                 double* out1, int lower, const Scalars* S) {
    const int k0 = blockIdx.x * blockDim.x + threadIdx.x;
    TriPos q{0, 0, 0, 1.0};
    if (k0 < count) q = tri_pos<VM>(begin + k0, t, dv, lower != 0);
    if (S && S->halt) return;
    if (k0 < count) {
        tri_row_at<VM, double, 1>(q, t, ev, in0, out0, lower != 0);
        tri_row_at<VM, double, 1>(q, t, ev, in1, out1, lower != 0);
    }
    for (int k = k0 + gridDim.x * blockDim.x; k < count; k += gridDim.x * blockDim.x) {
        tri_row_level<VM, double, 1>(begin + k, t, ev, dv, in0, out0, lower != 0);
        tri_row_level<VM, double, 1>(begin + k, t, ev, dv, in1, out1, lower != 0);
    }
}

// All levels in one workgroup; a barrier separates consecutive levels.
constexpr int kTriBlock = 1024;
template <class VM, class X, int NR>
__global__ void __launch_bounds__(kTriBlock) k_tri_block(LevelTri t, const int* __restrict__ lptr, int n_levels,
                                                         const VM* __restrict__ ev, const VM* __restrict__ dv,
                                                         const X* in, X* out, int lower, const Scalars* S) {
    if (S && S->halt) return;
    for (int l = 0; l < n_levels; ++l) {
        const int b = lptr[l], e = lptr[l + 1];
        for (int k = b + threadIdx.x; k < e; k += kTriBlock) tri_row_level<VM, X, NR>(k, t, ev, dv, in, out, lower != 0);
        __syncthreads();
    }
}

// The same solves from the padded-column copy (Levels::ell_w): a row's
// columns and factor values sit at fixed offsets from its position, so they
// are loaded in the same round as its row index, and only the solution
// values remain to gather.  Terms added in the same order as tri_row_level.
struct LevelEll {
    const int* rows;
    const int* cnt;
    const int* ecol;  // w * n
    int n, w;
};
struct EllRow {
    int i, c;
    double dd;
    int j[kTriEll];
    double v[kTriEll];
};
template <class VM>
__device__ __forceinline__ void ell_load(int k, const LevelEll& t, const VM* __restrict__ ev,
                                         const VM* __restrict__ dv, bool lower, EllRow& r) {
    r.i = t.rows[k];
    r.c = t.cnt[k];
    r.dd = lower ? 1.0 : (double)dv[k];
#pragma unroll
    for (int e = 0; e < kTriEll; ++e) {
        r.j[e] = 0;
        r.v[e] = 0.0;
        if (e < t.w) {
            r.j[e] = t.ecol[(size_t)e * t.n + k];
            r.v[e] = (double)ev[(size_t)e * t.n + k];
        }
    }
}
template <class X, int NR>
__device__ __forceinline__ void ell_apply(const EllRow& r, const X* in, X* out, bool lower) {
    double sum[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) sum[c] = 0.0;
#pragma unroll
    for (int e = 0; e < kTriEll; ++e)
        if (e < r.c) {
#pragma unroll
            for (int c = 0; c < NR; ++c) sum[c] += r.v[e] * (double)out[(size_t)r.j[e] * NR + c];
        }
#pragma unroll
    for (int c = 0; c < NR; ++c) {
        const double v = (double)in[(size_t)r.i * NR + c] - sum[c];
        out[(size_t)r.i * NR + c] = (X)(lower ? v : v / r.dd);
    }
}

template <class VM, class X, int NR>
__global__ void k_tri_level_ell(LevelEll t, int begin, int count, const VM* __restrict__ ev,
                                const VM* __restrict__ dv, const X* in, X* out, int lower, const Scalars* S) {
    const int k0 = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    EllRow r;
    if (k0 < count) ell_load<VM>(begin + k0, t, ev, dv, lower != 0, r);
    if (S && S->halt) return;
    if (k0 < count) ell_apply<X, NR>(r, in, out, lower != 0);
    for (int k = k0 + stride; k < count; k += stride) {
        ell_load<VM>(begin + k, t, ev, dv, lower != 0, r);
        ell_apply<X, NR>(r, in, out, lower != 0);
    }
}

template <class VM>
__global__ void k_tri_level_pair_ell(LevelEll t, int begin, int count, const VM* __restrict__ ev,
                                     const VM* __restrict__ dv, const double* in0, double* out0, const double* in1,
                                     double* out1, int lower, const Scalars* S) {
    const int k0 = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    EllRow r;
    if (k0 < count) ell_load<VM>(begin + k0, t, ev, dv, lower != 0, r);
    if (S && S->halt) return;
    for (int k = k0; k < count; k += stride) {
        if (k != k0) ell_load<VM>(begin + k, t, ev, dv, lower != 0, r);
        ell_apply<double, 1>(r, in0, out0, lower != 0);
        ell_apply<double, 1>(r, in1, out1, lower != 0);
    }
}

template <class VM, class X, int NR>
__global__ void __launch_bounds__(kTriBlock) k_tri_block_ell(LevelEll t, const int* __restrict__ lptr, int n_levels,
                                                             const VM* __restrict__ ev, const VM* __restrict__ dv,
                                                             const X* in, X* out, int lower, const Scalars* S) {
    if (S && S->halt) return;
    EllRow r;
    for (int l = 0; l < n_levels; ++l) {
        const int b = lptr[l], e = lptr[l + 1];
        for (int k = b + threadIdx.x; k < e; k += kTriBlock) {
            ell_load<VM>(k, t, ev, dv, lower != 0, r);
            ell_apply<X, NR>(r, in, out, lower != 0);
        }
        __syncthreads();
    }
}

// Deep, narrow schedules (levels of a few rows: the chains of 1D meshes):
// the workgroup stages the level-ordered factor and the input in LDS, then
// one thread walks the positions in order -- every row's dependencies sit at
// earlier positions -- and keeps the row it just solved in registers, so a
// chain costs one dependent multiply-add per row instead of a workgroup
// barrier and a memory round trip per level.  The terms are added in the
// order of ell_apply and each result is rounded to X before it is read
// again, so results are bitwise those of the level kernels.
constexpr int kTriSerialBlock = 1024;  // threads that stage the arrays
constexpr int kTriSerialLds = 65536;   // LDS bytes of a serial solve, at most
inline size_t tri_serial_lds(int n, int w, int nr) {
    return (size_t)n * (16 * (size_t)nr + 12 * (size_t)w + 16);
}
// Vectors of kS interleaved columns (v[i kS + c]); workgroup b solves
// columns (b mod (kS / NR)) NR .. + NR - 1 of in[b / (kS / NR)] into the
// same of out[...] (two vector pairs: the paired solve), so independent
// right-hand sides run on separate compute units at once.
template <class X>
struct SerialIO {
    const X* in[2];
    X* out[2];
};

template <class VM, class X, int NR, int kS, bool kLower, int kW>
__global__ void __launch_bounds__(kTriSerialBlock) k_tri_serial_ell(LevelEll t, const VM* __restrict__ ev,
                                                                    const VM* __restrict__ dv, SerialIO<X> io,
                                                                    const Scalars* S) {
    extern __shared__ double lds[];
    if (S && S->halt) return;
    const int n = t.n, w = t.w;
    constexpr int kG = kS / NR;  // workgroups per vector
    const int pair = blockIdx.x / kG;
    const int col0 = (blockIdx.x % kG) * NR;
    const X* __restrict__ in = io.in[pair];
    X* __restrict__ out = io.out[pair];
    double* xs = lds;               // n NR: the solution, by row
    double* xin = xs + n * NR;      // n NR: the input, by position
    double* vv = xin + n * NR;      // w n: factor values, entry e of position k at e n + k
    double* vd = vv + w * n;        // n: diagonal by position (upper)
    int* ri = (int*)(vd + n);       // n: row of each position
    int* rc = ri + n;               // n: entries of each position
    int* rj = rc + n;               // w n: columns, as vv
    for (int q = threadIdx.x; q < w * n; q += blockDim.x) {
        vv[q] = (double)ev[q];
        rj[q] = t.ecol[q];
    }
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int i = t.rows[k];
        ri[k] = i;
        rc[k] = t.cnt[k];
        vd[k] = kLower ? 1.0 : (double)dv[k];
#pragma unroll
        for (int c = 0; c < NR; ++c) xin[k * NR + c] = (double)in[(size_t)i * kS + col0 + c];
    }
    __syncthreads();
    // A chain: position 0 has no entry, every later position one, on the row
    // of the position before it (the factors of 1D meshes).
    int chain = 1;
    if constexpr (kW == 1) {
        for (int k = threadIdx.x; k < n; k += blockDim.x)
            if (rc[k] != (k > 0 ? 1 : 0) || (k > 0 && rj[k] != ri[k - 1])) chain = 0;
        chain = __syncthreads_and(chain);
    } else {
        chain = 0;
    }
    if (chain && threadIdx.x == 0) {
        // y_k = x_k - m_k y_{k-1} (upper: then / d_k), one dependent
        // multiply-add per row and right-hand side (the NR recurrences are
        // independent, so they overlap); the terms as in the general loop
        constexpr int kC = NR == 1 ? 8 : 4;
        double pv[NR];
#pragma unroll
        for (int c = 0; c < NR; ++c) {
            const double v = xin[c];
            const X r0 = (X)(kLower ? v : v / vd[0]);
            pv[c] = (double)r0;
            xs[ri[0] * NR + c] = pv[c];
        }
        auto step = [&](double x, double m, double d, int c) {
            double sum = 0.0;
            sum += m * pv[c];
            const double v = x - sum;
            const X r = (X)(kLower ? v : v / d);
            pv[c] = (double)r;
        };
        int k0 = 1;
        for (; k0 + kC <= n; k0 += kC) {
            double xv[kC][NR], mv[kC], dd[kC];
            int rr[kC];
#pragma unroll
            for (int u = 0; u < kC; ++u) {
#pragma unroll
                for (int c = 0; c < NR; ++c) xv[u][c] = xin[(k0 + u) * NR + c];
                mv[u] = vv[k0 + u];
                dd[u] = kLower ? 1.0 : vd[k0 + u];
                rr[u] = ri[k0 + u];
            }
#pragma unroll
            for (int u = 0; u < kC; ++u) {
#pragma unroll
                for (int c = 0; c < NR; ++c) {
                    step(xv[u][c], mv[u], dd[u], c);
                    xs[rr[u] * NR + c] = pv[c];
                }
            }
        }
        for (; k0 < n; ++k0) {
#pragma unroll
            for (int c = 0; c < NR; ++c) {
                step(xin[k0 * NR + c], vv[k0], kLower ? 1.0 : vd[k0], c);
                xs[ri[k0] * NR + c] = pv[c];
            }
        }
    }
    if (!chain && threadIdx.x == 0) {
        // Rows in chunks: a chunk's position data is read from LDS at once,
        // so one LDS round trip serves kChunk rows, each of which then waits
        // only for the multiply-adds (and, upper, the division) of the row
        // before it.
        constexpr int kChunk = (kW == 1 && NR <= 2) ? 8 : 2;
        struct Row {
            int i, c;
            double d, x[NR];
            int j[kW];
            double v[kW];
        };
        auto load = [&](int k, Row& r) {
            r.i = ri[k];
            r.c = rc[k];
            r.d = kLower ? 1.0 : vd[k];
#pragma unroll
            for (int c = 0; c < NR; ++c) r.x[c] = xin[k * NR + c];
#pragma unroll
            for (int e = 0; e < kW; ++e) {
                const bool in_w = kW == 1 || e < w;
                r.j[e] = in_w ? rj[e * n + k] : 0;
                r.v[e] = in_w ? vv[e * n + k] : 0.0;
            }
        };
        int prev = -1;
        double pv[NR];
#pragma unroll
        for (int c = 0; c < NR; ++c) pv[c] = 0.0;
        auto solve = [&](const Row& r) {
            double sum[NR];
#pragma unroll
            for (int c = 0; c < NR; ++c) sum[c] = 0.0;
#pragma unroll
            for (int e = 0; e < kW; ++e)
                if (e < r.c) {
                    const int j = r.j[e];
                    if (j == prev) {
#pragma unroll
                        for (int c = 0; c < NR; ++c) sum[c] += r.v[e] * pv[c];
                    } else {
#pragma unroll
                        for (int c = 0; c < NR; ++c) sum[c] += r.v[e] * xs[j * NR + c];
                    }
                }
#pragma unroll
            for (int c = 0; c < NR; ++c) {
                const double v = r.x[c] - sum[c];
                const X res = (X)(kLower ? v : v / r.d);
                pv[c] = (double)res;
                xs[r.i * NR + c] = (double)res;
            }
            prev = r.i;
        };
        for (int k0 = 0; k0 < n; k0 += kChunk) {
            Row r[kChunk];
#pragma unroll
            for (int u = 0; u < kChunk; ++u)
                if (k0 + u < n) load(k0 + u, r[u]);
#pragma unroll
            for (int u = 0; u < kChunk; ++u)
                if (k0 + u < n) solve(r[u]);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < n * NR; q += blockDim.x) out[(size_t)(q / NR) * kS + col0 + q % NR] = (X)xs[q];
}

// Launches the serial form: lower or upper, the one-entry chain or up to
// kTriEll entries per row; NR interleaved right-hand sides take NR
// workgroups of one each (a serial solve is bound by its dependent chain,
// so separate compute units run them in the time of one).
// A second vector (in1, out1: not null) is solved in the same launch.
template <class VM, class X, int NR>
void launch_tri_serial(const LevelEll& te, const VM* ev, const VM* dv, const X* in, X* out, bool lower,
                       const Scalars* S, hipStream_t st, const X* in1 = nullptr, X* out1 = nullptr) {
    const size_t lds = tri_serial_lds(te.n, te.w, 1);
    const SerialIO<X> io{{in, in1 ? in1 : in}, {out, out1 ? out1 : out}};
    const int grid = in1 ? 2 * NR : NR;
    auto go = [&](auto L, auto W) {
        k_tri_serial_ell<VM, X, 1, NR, decltype(L)::value, decltype(W)::value>
            <<<grid, kTriSerialBlock, lds, st>>>(te, ev, dv, io, S);
    };
    using T = std::true_type;
    using F = std::false_type;
    using W1 = std::integral_constant<int, 1>;
    using W8 = std::integral_constant<int, kTriEll>;
    if (lower) {
        if (te.w == 1) go(T{}, W1{});
        else go(T{}, W8{});
    } else {
        if (te.w == 1) go(F{}, W1{});
        else go(F{}, W8{});
    }
}

// Factor values into level order: ev[e] = m[esrc[e]], dv[k] = m[dslot[k]].
template <class VM>
__global__ void k_level_gather(int ne, const int* __restrict__ esrc, int n, const int* __restrict__ dslot,
                               const VM* __restrict__ m, VM* __restrict__ ev, VM* __restrict__ dv) {
    const int stride = gridDim.x * blockDim.x;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += stride) ev[e] = m[esrc[e]];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) dv[k] = m[dslot[k]];
}

// One damped Jacobi sweep: out = w * (in - off(M) old) [/ M_ii] + (1 - w) old
template <class VM, class X, int NR>
__global__ void k_tri_jacobi(int n, TriCSR t, const VM* __restrict__ m, const X* __restrict__ in,
                             const X* __restrict__ old, X* __restrict__ out, int lower, double w,
                             int first, const Scalars* S) {
    if (S && S->halt) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double sum[NR];
#pragma unroll
        for (int c = 0; c < NR; ++c) sum[c] = 0.0;
        const int d = t.diag[i];
        if (!first) {
            const int p0 = lower ? t.rp[i] : d + 1;
            const int p1 = lower ? d : t.rp[i + 1];
            for (int p = p0; p < p1; ++p) {
                const double mv = (double)m[p];
                const int j = t.col[p];
#pragma unroll
                for (int c = 0; c < NR; ++c) sum[c] += mv * (double)old[(size_t)j * NR + c];
            }
        }
        const double inv = lower ? 1.0 : 1.0 / (double)m[d];
#pragma unroll
        for (int c = 0; c < NR; ++c) {
            const double o = first ? 0.0 : (double)old[(size_t)i * NR + c];
            out[(size_t)i * NR + c] = (X)(w * inv * ((double)in[(size_t)i * NR + c] - sum[c]) + (1.0 - w) * o);
        }
    }
}

namespace {

TriCSR tri_csr(ocsl_solver* h) { return {h->rp.p, h->col.p, h->diag.p}; }

void ensure_levels(ocsl_solver* h) {
    if (h->levels_built) return;
    build_levels(h->host, true, h->lev_fwd, h->stream);
    build_levels(h->host, false, h->lev_bwd, h->stream);
    h->levels_built = true;
}

// One workgroup for the whole solve when its levels average fewer than 1024
// rows, else one launch per level.  Measured (lower solves of the 3-sweep
// factor of the convection-diffusion systems): a launch costs ~4 us per
// level whatever its size, a workgroup level ~3 ns per row; randomly
// numbered 32^3 (about 2,000 rows per level) 62 us per level launch against
// 98 us in one workgroup, lexicographic 96^3 (3,100) 1.07 against 4.8 ms.
bool use_block_tri(const ocsl_solver* h, const Levels& L, int n) {
    if (h->tri_form != 0) return h->tri_form == 1;
    return (int64_t)L.n_levels * 1024 > n;
}

// The serial form (k_tri_serial_ell): padded-column entries, the staged
// arrays within kTriSerialLds, and levels of 8 rows or fewer on average
// (each serial row costs ~50 cycles of LDS and multiply-add latency, a
// workgroup level ~0.35 us).
bool use_serial_tri(const ocsl_solver* h, const Levels& L, int n) {
    const size_t cap = std::min((size_t)kTriSerialLds, (size_t)h->lds_max);
    if (!L.ell_w || tri_serial_lds(n, L.ell_w, 1) > cap) return false;
    if (h->tri_form != 0) return h->tri_form == 3;
    return (int64_t)L.n_levels * 8 >= n;
}

// The factor m's values in L's level order, gathered once per factor
// (address) and decomposition.
template <class VM>
void level_values(ocsl_solver* h, Levels& L, const VM* m, const VM** ev, const VM** dv) {
    DBuf<VM>* e;
    DBuf<VM>* d;
    const void** from;
    int64_t* epoch;
    if constexpr (sizeof(VM) == 8) {
        e = &L.eval64;
        d = &L.dval64;
        from = &L.from64;
        epoch = &L.epoch64;
    } else {
        e = &L.eval32;
        d = &L.dval32;
        from = &L.from32;
        epoch = &L.epoch32;
    }
    const int n = h->host.n;
    // the padded-column copy when there is one, else the row-packed entries
    const int ne = L.ell_w ? L.ell_w * n : (int)L.esrc.n;
    const int* esrc = L.ell_w ? L.esw.p : L.esrc.p;
    if (*from != (const void*)m || *epoch != h->ilu_epoch || !e->p) {
        e->alloc(std::max(ne, 1));
        d->alloc(n);
        k_level_gather<VM><<<grid_for(std::max(ne, n)), kBlock, 0, h->stream>>>(ne, esrc, n, L.dslot.p, m, e->p,
                                                                                 d->p);
        OCSL_HIP(hipGetLastError());
        *from = (const void*)m;
        *epoch = h->ilu_epoch;
    }
    *ev = e->p;
    *dv = d->p;
}

// Exact solve (lower or upper) of NR interleaved vectors.
template <class VM, class X, int NR>
void tri_exact(ocsl_solver* h, const VM* m, const X* in, X* out, bool lower, const Scalars* S) {
    Levels& L = lower ? h->lev_fwd : h->lev_bwd;
    const int n = h->host.n;
    const VM *ev, *dv;
    level_values<VM>(h, L, m, &ev, &dv);
    if (L.ell_w) {
        const LevelEll te{L.rows.p, L.cnt.p, L.ecw.p, n, L.ell_w};
        if (use_serial_tri(h, L, n)) {
            launch_tri_serial<VM, X, NR>(te, ev, dv, in, out, lower, S, h->stream);
        } else if (use_block_tri(h, L, n)) {
            k_tri_block_ell<VM, X, NR><<<1, kTriBlock, 0, h->stream>>>(te, L.ptr.p, L.n_levels, ev, dv, in, out, lower,
                                                                       S);
        } else {
            for (int l = 0; l < L.n_levels; ++l) {
                const int b = L.ptr_host[l], cnt = L.ptr_host[l + 1] - b;
                k_tri_level_ell<VM, X, NR><<<grid_for(cnt), kBlock, 0, h->stream>>>(te, b, cnt, ev, dv, in, out, lower,
                                                                                     S);
            }
        }
        return;
    }
    const LevelTri t{L.rows.p, L.eptr.p, L.ecol.p};
    if (use_block_tri(h, L, n)) {
        k_tri_block<VM, X, NR><<<1, kTriBlock, 0, h->stream>>>(t, L.ptr.p, L.n_levels, ev, dv, in, out, lower, S);
    } else {
        for (int l = 0; l < L.n_levels; ++l) {
            const int b = L.ptr_host[l], cnt = L.ptr_host[l + 1] - b;
            k_tri_level<VM, X, NR><<<grid_for(cnt), kBlock, 0, h->stream>>>(t, b, cnt, ev, dv, in, out, lower, S);
        }
    }
}

// Two exact solves (same factor, same direction) of independent vectors:
// out0 = T^-1 in0 and out1 = T^-1 in1, one launch per level for both.
template <class VM>
void tri_exact_pair(ocsl_solver* h, const VM* m, const double* in0, double* out0, const double* in1, double* out1,
                    bool lower, const Scalars* S) {
    Levels& L = lower ? h->lev_fwd : h->lev_bwd;
    const int n = h->host.n;
    if (use_serial_tri(h, L, n)) {
        const VM *ev, *dv;
        level_values<VM>(h, L, m, &ev, &dv);
        const LevelEll te{L.rows.p, L.cnt.p, L.ecw.p, n, L.ell_w};
        launch_tri_serial<VM, double, 1>(te, ev, dv, in0, out0, lower, S, h->stream, in1, out1);
        return;
    }
    if (use_block_tri(h, L, n)) {
        tri_exact<VM, double, 1>(h, m, in0, out0, lower, S);
        tri_exact<VM, double, 1>(h, m, in1, out1, lower, S);
        return;
    }
    const VM *ev, *dv;
    level_values<VM>(h, L, m, &ev, &dv);
    if (L.ell_w) {
        const LevelEll te{L.rows.p, L.cnt.p, L.ecw.p, n, L.ell_w};
        for (int l = 0; l < L.n_levels; ++l) {
            const int b = L.ptr_host[l], cnt = L.ptr_host[l + 1] - b;
            k_tri_level_pair_ell<VM><<<grid_for(cnt), kBlock, 0, h->stream>>>(te, b, cnt, ev, dv, in0, out0, in1, out1,
                                                                              lower, S);
        }
        return;
    }
    const LevelTri t{L.rows.p, L.eptr.p, L.ecol.p};
    for (int l = 0; l < L.n_levels; ++l) {
        const int b = L.ptr_host[l], cnt = L.ptr_host[l + 1] - b;
        k_tri_level_pair<VM><<<grid_for(cnt), kBlock, 0, h->stream>>>(t, b, cnt, ev, dv, in0, out0, in1, out1, lower,
                                                                      S);
    }
}

// Jacobi sweeps into out, tmp as the ping-pong partner.
template <class VM, class X, int NR>
void tri_jacobi(ocsl_solver* h, const VM* m, const X* in, X* out, X* tmp, bool lower, int sweeps, double w,
                const Scalars* S) {
    const int n = h->host.n;
    X* bufs[2] = {out, tmp};
    int cur = (sweeps % 2 == 1) ? 0 : 1;  // the last sweep writes `out`
    for (int s = 0; s < sweeps; ++s) {
        X* dst = bufs[cur];
        const X* src = bufs[cur ^ 1];
        k_tri_jacobi<VM, X, NR><<<grid_for(n), kBlock, 0, h->stream>>>(n, tri_csr(h), m, in, src, dst, lower, w,
                                                                       s == 0, S);
        cur ^= 1;
    }
}

}  // namespace

// M^-1 in = U^-1 L^-1 in, through tmp (and tmp2 for Jacobi).
template <class VM, class X, int NR>
static void apply_precond(ocsl_solver* h, const VM* m, const X* in, X* tmp, X* tmp2, X* out,
                          const ocsl_options& o, const Scalars* S) {
    if (o.tri_mode == OCSL_TRI_JACOBI) {
        tri_jacobi<VM, X, NR>(h, m, in, tmp, tmp2, true, o.tri_sweeps, o.tri_weight, S);
        tri_jacobi<VM, X, NR>(h, m, tmp, out, tmp2, false, o.tri_sweeps, o.tri_weight, S);
    } else {
        tri_exact<VM, X, NR>(h, m, in, tmp, true, S);
        tri_exact<VM, X, NR>(h, m, tmp, out, false, S);
    }
}

// ------------------------------------------------------------------ reference mode
// Device-resident scalars of the REFERENCE BiCGSTAB: the dot products land
// in S->aux (0: r0.r, 1: r0.v, 2: r.r, 3: b.b, 4: (L^-1 t).(L^-1 s), 5: (L^-1 t).(L^-1 t)), the
// last stage of the reduction that completes a step's inputs forms alpha,
// the residual or omega with the reference's formulas (bref_step), and the
// vector updates read their coefficients from S.  The
// reference's early exit (residual below 1e-20, before x = h) sets S->halt,
// after which the updates and the triangular solves do nothing.
__device__ __forceinline__ void bref_step(Scalars* S, double* hist, int step, int it) {
    if (step < 0 || S->halt) return;  // step < 0: no step
    if (step == 0) {
        S->alpha = S->aux[0] / (S->aux[1] + DBL_MIN);
    } else if (step == 1) {
        const double res = S->aux[2] / S->aux[3] + DBL_MIN;
        S->rr = res;
        hist[it] = res;
        S->iter = it + 1;
        if (res < 1.0e-20) S->halt = 1;
    } else {
        S->omega = S->aux[4] / (S->aux[5] + 1.0e-40);
    }
}

// The totals of the dot products a step reads (aux[slot ..]), then the
// step: one launch for the reduction's last stage and the scalar.
template <int NR>
struct FinBref {
    Scalars* S;
    double* hist;
    int slot, step, it;
    __device__ bool active() const { return true; }
    __device__ void operator()(const double* t) const {
        for (int r = 0; r < NR; ++r) S->aux[slot + r] = t[r];
        bref_step(S, hist, step, it);
    }
};

void bicgstab_reference(ocsl_solver* h, ocsl_report* rep) {
    const int n = h->host.n;
    ensure_csr_values(h, false);
    ensure_levels(h);
    Sell& s = sell_for(h, false);
    ensure_sell_values(h, s, true, false);
    Work& w = h->work;
    for (int i = 0; i < 11; ++i) work_buf(h, w.d[i], n);
    double *r = w.d[0].p, *p = w.d[1].p, *v = w.d[2].p, *r0 = w.d[3].p, *sv = w.d[4].p, *tmp = w.d[5].p;
    double *hh = w.d[6].p, *t = w.d[7].p, *y = w.d[8].p, *z = w.d[9].p, *lts = w.d[10].p;
    double* x = h->x_own.p;
    const double* b = h->b_dev;
    const int gv = grid_for(n);
    auto spmv = [&](double a, const double* xin, double bz, const double* zin, double* out) {
        k_sell_axpby<double><<<grid_for((int64_t)s.n_slices * kWave), kBlock, 0, h->stream>>>(
            s.view<double>(), a, nullptr, xin, bz, zin, out);
    };
    Scalars* S = h->S.p;
    // device addresses of fields of S (S is never dereferenced on the host)
    auto field = [S](size_t off) { return reinterpret_cast<double*>(reinterpret_cast<char*>(S) + off); };
    double* aux = field(offsetof(Scalars, aux));
    double* hist_d = work_buf(h, h->hist, 4);
    const int gd = red_blocks(kBlock * 4, n);
    // dot products into S->aux[slot], in the reduction device_dots uses
    auto dot1 = [&](const double* a1, const double* b1, int slot) {
        DotArgs<1> d{{a1}, {b1}, {0}};
        k_dots<1><<<gd, kBlock, 0, h->stream>>>(n, d, h->red());
        k_finalize<1, FinStore<1>><<<1, kBlock, 0, h->stream>>>(h->red(), gd, FinStore<1>{aux + slot});
    };
    // dot products whose totals a scalar step reads, with that step
    auto dot1s = [&](const double* a1, const double* b1, int slot, int step, int it) {
        DotArgs<1> d{{a1}, {b1}, {0}};
        k_dots<1><<<gd, kBlock, 0, h->stream>>>(n, d, h->red());
        k_finalize<1, FinBref<1>><<<1, kBlock, 0, h->stream>>>(h->red(), gd, FinBref<1>{S, hist_d, slot, step, it});
    };
    auto dot2s = [&](const double* a1, const double* b1, const double* a2, const double* b2, int slot, int step,
                     int it) {
        DotArgs<2> d{{a1, a2}, {b1, b2}, {0, 0}};
        k_dots<2><<<gd, kBlock, 0, h->stream>>>(n, d, h->red());
        k_finalize<2, FinBref<2>><<<1, kBlock, 0, h->stream>>>(h->red(), gd, FinBref<2>{S, hist_d, slot, step, it});
    };
    auto upd = [&](double ca, const double* pa, const double* xa, double cb, const double* pb, const double* ya,
                   double* out) {
        k_axpby_s<<<gv, kBlock, 0, h->stream>>>(n, ca, pa, xa, cb, pb, ya, out, S);
    };
    Scalars init{};
    init.alpha = 1.0;
    init.omega = 1.0;
    const double* alpha = field(offsetof(Scalars, alpha));
    const double* omega = field(offsetof(Scalars, omega));
    OCSL_HIP(hipEventRecord(h->ev0, h->stream));
    // the solve's device work: replayed as one graph after its first run
    // (no host copies or memsets inside: kernels set S and clear p, v)
    auto body = [&] {
    k_set_scalars<<<1, 64, 0, h->stream>>>(S, init);
    ilu_decompose_device(h, 3, false);  // M.copy_same_structure + ILU_decompose
    const double* m = h->ilu_m.p;       // allocated by the call above on first use
    k_fill<double><<<gv, kBlock, 0, h->stream>>>(n, 0.0, p);
    k_fill<double><<<gv, kBlock, 0, h->stream>>>(n, 0.0, v);
    spmv(-1.0, x, 1.0, b, r);
    k_permute_in<double><<<gv, kBlock, 0, h->stream>>>(n, nullptr, r, r0);
    for (int it = 0; it < 3; ++it) {
        dot1(r0, r, 0);                                                // rhon
        upd(1.0, nullptr, p, -1.0, omega, v, tmp);                     // p - omega v
        upd(1.0, nullptr, r, 1.0, omega, tmp, p);                      // p = r + omega (p - omega v)
        tri_exact<double, double, 1>(h, m, p, tmp, true, S);
        tri_exact<double, double, 1>(h, m, tmp, y, false, S);
        spmv(1.0, y, 0.0, nullptr, v);
        dot1s(r0, v, 1, 0, it);                                        // alpha
        upd(1.0, nullptr, x, 1.0, alpha, y, hh);                       // h = x + alpha y
        dot2s(r, r, b, b, 2, 1, it);                                   // residual, early exit
        upd(-1.0, alpha, v, 1.0, nullptr, r, sv);                      // s = r - alpha v
        tri_exact<double, double, 1>(h, m, sv, tmp, true, S);
        tri_exact<double, double, 1>(h, m, tmp, z, false, S);
        spmv(1.0, z, 0.0, nullptr, t);
        tri_exact_pair<double>(h, m, t, tmp, sv, lts, true, S);       // L^-1 t, L^-1 s
        dot2s(tmp, lts, tmp, tmp, 4, 2, it);                           // omega
        upd(1.0, nullptr, hh, 1.0, omega, z, x);                       // x = h + omega z
        upd(-1.0, omega, t, 1.0, nullptr, sv, r);                      // r = s - omega t
    }
    OCSL_HIP(hipGetLastError());
    };
    run_reference(h, 1, b, body);
    ++h->ilu_epoch;  // a replay recomputed the factor
    OCSL_HIP(hipEventRecord(h->ev1, h->stream));
    // the one read-back of the solve: iterations, residual, history
    OCSL_HIP(hipMemcpyAsync(h->S_host, S, sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
    sync(h);
    const int it = h->S_host->iter;
    const double res = h->S_host->rr;
    std::vector<double> hist(it);
    if (it) {
        OCSL_HIP(hipMemcpyAsync(hist.data(), hist_d, sizeof(double) * it, hipMemcpyDeviceToHost, h->stream));
        sync(h);
    }
    float ms = 0.f;
    OCSL_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->hist_host = hist;
    h->hist_len = (int)hist.size();
    rep->status = std::isfinite(res) ? OCSL_OK : OCSL_WARN_NONFINITE;
    rep->iterations = it;
    rep->residual = res;
    {
        const int g = red_blocks(kBlock * 4, n);
        k_true_residual<double><<<g, kBlock, 0, h->stream>>>(s.view<double>(), b, x, h->red());
        k_finalize<2, FinStore<2>><<<1, kBlock, 0, h->stream>>>(h->red(), g, FinStore<2>{h->dscal.p});
        OCSL_HIP(hipMemcpyAsync(h->dscal_host, h->dscal.p, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                h->stream));
        sync(h);
        const double rr = h->dscal_host[0], bb = h->dscal_host[1];
        rep->true_relres = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    }
    rep->history_len = h->hist_len;
    rep->solve_ms = ms;
    const double nnz = (double)h->host.col.size();
    const double b_spmv = nnz * 12.0 + 16.0 * n;
    const double b_tri = nnz * 12.0 + 8.0 * n + 16.0 * n;
    rep->bytes = it * (2.0 * b_spmv + 3.0 * b_tri + 20.0 * 8.0 * n);
}

// ------------------------------------------------------------------ reference mode, k right-hand sides
// ILU_BICGSTAB_solve (cmatrix_solver.cpp:682-762) for k right-hand sides of
// one matrix at once: the factor is computed once (every call of the
// reference computes the same one from the same values), and every matrix,
// factor and vector pass serves all k, held interleaved (v[i k + c]).  Each
// right-hand side keeps its own scalars (S[c]) and takes exactly the steps,
// terms and reduction orders of the one-vector solve above, so its results
// are bitwise those of that solve.  The triangular solves run for every
// right-hand side; one whose solve has halted computes values that nothing
// reads (its x and scalars are frozen).
namespace {

constexpr size_t kNoCoef = ~(size_t)0;

__global__ void k_set_scalars_nr(Scalars* S, Scalars v, int k) {
    if (blockIdx.x == 0 && (int)threadIdx.x < k) S[threadIdx.x] = v;
}


__device__ __forceinline__ double scalar_at(const Scalars* s, size_t off) {
    return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(s) + off);
}

// k_axpby_s per right-hand side: out = (ca a_c) x + (cb b_c) y, a_c and b_c
// the fields at byte offsets oa, ob of S[c] (kNoCoef: 1).
template <int NR>
__global__ void k_axpby_nr(int n, double ca, size_t oa, const double* __restrict__ x, double cb, size_t ob,
                           const double* __restrict__ y, double* out, const Scalars* S) {
    double a[NR], b[NR];
    bool on[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) {
        on[c] = !S[c].halt;
        a[c] = oa != kNoCoef ? ca * scalar_at(S + c, oa) : ca;
        b[c] = ob != kNoCoef ? cb * scalar_at(S + c, ob) : cb;
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < NR; ++c)
            if (on[c]) {
                const size_t q = (size_t)i * NR + c;
                out[q] = a[c] * x[q] + b[c] * y[q];
            }
}

// k_dots with K products per right-hand side (kind 0, a.b) of interleaved
// vectors: the same block and thread decomposition, sum c K + j in slot
// c K + j.
template <int NR, int K>
__global__ void __launch_bounds__(kBlock) k_dots_nr(int n, DotArgs<K> args, RedScratch red) {
    double acc[NR * K];
#pragma unroll
    for (int r = 0; r < NR * K; ++r) acc[r] = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
#pragma unroll
        for (int c = 0; c < NR; ++c)
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const size_t q = (size_t)i * NR + c;
                const double av = args.a[j][q];
                acc[c * K + j] += args.kind[j] == 0 ? av * args.b[j][q] : (args.kind[j] == 1 ? av : fabs(av));
            }
    }
    block_partials<NR * K>(acc, red);
}

// Totals c K + j to S[c].aux[slot + j], then bref_step of each right-hand
// side (history entry it of c at hist[3 c + it]).
template <int NR, int K>
struct FinAux {
    Scalars* S;
    double* hist;
    int slot, step, it;
    __device__ bool active() const { return true; }
    __device__ void operator()(const double* t) const {
        for (int c = 0; c < NR; ++c) {
            for (int j = 0; j < K; ++j) S[c].aux[slot + j] = t[c * K + j];
            bref_step(S + c, hist + 3 * c, step, it);
        }
    }
};

// k_sell_axpby of interleaved vectors: out = a (A x) + b z, z may be null.
template <int NR>
__global__ void __launch_bounds__(kBlock) k_sell_axpby_nr(SellView<double> A, double a, const double* __restrict__ x,
                                                          double b, const double* __restrict__ z,
                                                          double* __restrict__ out) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * kWavesPerBlock;
    for (int s = wave; s < A.n_slices; s += nwaves) {
        const int row = s * kWave + lane;
        double ax[NR];
        sell_rows_nr<double, double, NR>(A, s, lane, row, x, ax);
        if (row < A.n)
#pragma unroll
            for (int c = 0; c < NR; ++c) {
                const size_t q = (size_t)row * NR + c;
                double o = (a == 1.0) ? ax[c] : a * ax[c];
                if (z) o += b * z[q];
                out[q] = o;
            }
    }
}

template <int NR>
void bicgstab_reference_nr(ocsl_solver* h, const double* b, double* x, ocsl_report* reps) {
    const int n = h->host.n;
    ensure_csr_values(h, false);
    ensure_levels(h);
    Sell& s = sell_for(h, false);
    ensure_sell_values(h, s, true, false);
    Work& w = h->work;
    const size_t nk = (size_t)n * NR;
    for (int i = 0; i < 11; ++i) work_buf(h, w.d[i], nk);
    double *r = w.d[0].p, *p = w.d[1].p, *v = w.d[2].p, *r0 = w.d[3].p, *sv = w.d[4].p, *tmp = w.d[5].p;
    double *hh = w.d[6].p, *t = w.d[7].p, *y = w.d[8].p, *z = w.d[9].p, *lts = w.d[10].p;
    Scalars* S = work_buf(h, h->S_multi, NR);
    double* hist_d = work_buf(h, h->hist, 3 * NR);
    const int gv = grid_for(n), gk = grid_for((int64_t)nk);
    auto spmv = [&](double a, const double* xin, double bz, const double* zin, double* out) {
        k_sell_axpby_nr<NR><<<grid_for((int64_t)s.n_slices * kWave), kBlock, 0, h->stream>>>(s.view<double>(), a,
                                                                                             xin, bz, zin, out);
    };
    const int gd = red_blocks(kBlock * 4, n);
    // step < 0: the totals only; else with the scalar step that reads them
    auto dot1 = [&](const double* a1, const double* b1, int slot, int step = -1, int it = 0) {
        DotArgs<1> d{{a1}, {b1}, {0}};
        k_dots_nr<NR, 1><<<gd, kBlock, 0, h->stream>>>(n, d, h->red());
        k_finalize<NR, FinAux<NR, 1>><<<1, kBlock, 0, h->stream>>>(h->red(), gd,
                                                                    FinAux<NR, 1>{S, hist_d, slot, step, it});
    };
    auto dot2 = [&](const double* a1, const double* b1, const double* a2, const double* b2, int slot, int step,
                    int it) {
        DotArgs<2> d{{a1, a2}, {b1, b2}, {0, 0}};
        k_dots_nr<NR, 2><<<gd, kBlock, 0, h->stream>>>(n, d, h->red());
        k_finalize<2 * NR, FinAux<NR, 2>><<<1, kBlock, 0, h->stream>>>(h->red(), gd,
                                                                        FinAux<NR, 2>{S, hist_d, slot, step, it});
    };
    auto upd = [&](double ca, size_t oa, const double* xa, double cb, size_t ob, const double* ya, double* out) {
        k_axpby_nr<NR><<<gv, kBlock, 0, h->stream>>>(n, ca, oa, xa, cb, ob, ya, out, S);
    };
    const size_t alpha = offsetof(Scalars, alpha), omega = offsetof(Scalars, omega);
    auto lower = [&](const double* in, double* out) { tri_exact<double, double, NR>(h, h->ilu_m.p, in, out, true, nullptr); };
    auto upper = [&](const double* in, double* out) {
        tri_exact<double, double, NR>(h, h->ilu_m.p, in, out, false, nullptr);
    };
    // two independent lower solves: one launch in the serial form
    auto lower_pair = [&](const double* in0, double* out0, const double* in1, double* out1) {
        Levels& L = h->lev_fwd;
        if (use_serial_tri(h, L, n)) {
            const double *ev, *dv;
            level_values<double>(h, L, h->ilu_m.p, &ev, &dv);
            const LevelEll te{L.rows.p, L.cnt.p, L.ecw.p, n, L.ell_w};
            launch_tri_serial<double, double, NR>(te, ev, dv, in0, out0, true, nullptr, h->stream, in1, out1);
        } else {
            lower(in0, out0);
            lower(in1, out1);
        }
    };
    Scalars init{};
    init.alpha = 1.0;
    init.omega = 1.0;
    OCSL_HIP(hipEventRecord(h->ev0, h->stream));
    auto body = [&] {
        k_set_scalars_nr<<<1, 64, 0, h->stream>>>(S, init, NR);
        ilu_decompose_device(h, 3, false);
        k_fill<double><<<gk, kBlock, 0, h->stream>>>((int)nk, 0.0, p);
        k_fill<double><<<gk, kBlock, 0, h->stream>>>((int)nk, 0.0, v);
        spmv(-1.0, x, 1.0, b, r);
        k_permute_in<double><<<gk, kBlock, 0, h->stream>>>((int)nk, nullptr, r, r0);
        for (int it = 0; it < 3; ++it) {
            dot1(r0, r, 0);
            upd(1.0, kNoCoef, p, -1.0, omega, v, tmp);
            upd(1.0, kNoCoef, r, 1.0, omega, tmp, p);
            lower(p, tmp);
            upper(tmp, y);
            spmv(1.0, y, 0.0, nullptr, v);
            dot1(r0, v, 1, 0, it);
            upd(1.0, kNoCoef, x, 1.0, alpha, y, hh);
            dot2(r, r, b, b, 2, 1, it);
            upd(-1.0, alpha, v, 1.0, kNoCoef, r, sv);
            lower(sv, tmp);
            upper(tmp, z);
            spmv(1.0, z, 0.0, nullptr, t);
            lower_pair(t, tmp, sv, lts);
            dot2(tmp, lts, tmp, tmp, 4, 2, it);
            upd(1.0, kNoCoef, hh, 1.0, omega, z, x);
            upd(-1.0, omega, t, 1.0, kNoCoef, sv, r);
        }
        OCSL_HIP(hipGetLastError());
    };
    run_reference(h, 3, b, body, NR);
    ++h->ilu_epoch;
    OCSL_HIP(hipEventRecord(h->ev1, h->stream));
    std::vector<Scalars> Sh(NR);
    std::vector<double> hist(3 * NR);
    OCSL_HIP(hipMemcpyAsync(Sh.data(), S, sizeof(Scalars) * NR, hipMemcpyDeviceToHost, h->stream));
    OCSL_HIP(hipMemcpyAsync(hist.data(), hist_d, sizeof(double) * 3 * NR, hipMemcpyDeviceToHost, h->stream));
    sync(h);
    float ms = 0.f;
    OCSL_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    const double nnz = (double)h->host.col.size();
    const double b_spmv = nnz * 12.0 + 16.0 * n * NR;
    const double b_tri = nnz * 12.0 + 8.0 * n + 16.0 * n * NR;
    // the history of the last call: right-hand side after right-hand side
    h->hist_host.clear();
    for (int c = 0; c < NR; ++c) {
        const int it = Sh[c].iter;
        const double res = Sh[c].rr;
        ocsl_report& rep = reps[c];
        rep.status = std::isfinite(res) ? OCSL_OK : OCSL_WARN_NONFINITE;
        rep.iterations = it;
        rep.residual = res;
        rep.history_len = it;
        rep.solve_ms = ms;
        rep.true_relres = std::nan("");  // not computed: the call returns the reference's residual
        // each right-hand side's share of the passes its iterations ran in
        rep.bytes = it * (2.0 * b_spmv + 4.0 * b_tri + 20.0 * 8.0 * n * NR) / NR;
        for (int i = 0; i < it; ++i) h->hist_host.push_back(hist[3 * c + i]);
    }
    h->hist_len = (int)h->hist_host.size();
}

}  // namespace

template <class T>
__global__ void k_interleave(int n, int k, const T* __restrict__ in, T* __restrict__ out, int to_rows);

// B_dev, X_dev: k column blocks of n (X: guesses in, solutions out).
void bicgstab_reference_multi(ocsl_solver* h, int k, const double* B_dev, double* X_dev, ocsl_report* reps) {
    OCSL_REQUIRE(k >= 1 && k <= 4, OCSL_ERR_ARG, "1 to 4 right-hand sides");
    const int n = h->host.n;
    // interleaved copies at fixed addresses (the solve's graph reads them)
    double* b_il = work_buf(h, h->work.d[14], (size_t)n * k);
    double* x_il = work_buf(h, h->work.d[15], (size_t)n * k);
    const int g = grid_for(n);
    k_interleave<double><<<g, kBlock, 0, h->stream>>>(n, k, B_dev, b_il, 1);
    k_interleave<double><<<g, kBlock, 0, h->stream>>>(n, k, X_dev, x_il, 1);
    OCSL_HIP(hipGetLastError());
    switch (k) {
        case 1: bicgstab_reference_nr<1>(h, b_il, x_il, reps); break;
        case 2: bicgstab_reference_nr<2>(h, b_il, x_il, reps); break;
        case 3: bicgstab_reference_nr<3>(h, b_il, x_il, reps); break;
        default: bicgstab_reference_nr<4>(h, b_il, x_il, reps); break;
    }
    k_interleave<double><<<g, kBlock, 0, h->stream>>>(n, k, x_il, X_dev, 0);
    OCSL_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ tolerance mode
template <class V, class X, int NR>
struct BiArgs {
    SellView<V> A;
    const X* rh;  // shadow residual r^
    X *r, *p, *v, *s, *t, *yh, *zh;
    double* x;
    Scalars* S;
    double* hist;
    RedScratch red;
    int n;
};

// p = r + beta (p - omega v); on the first iteration beta = 0, p = v = 0.
template <class V, class X, int NR>
__global__ void __launch_bounds__(kBlock) k_bi_p(BiArgs<V, X, NR> c) {
    Scalars* S = c.S;
    if (S->done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) S->halt = 1;
        return;
    }
    double beta[NR], omega[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        beta[k] = S->donek[k] ? 0.0 : S->betak[k];
        omega[k] = S->omegak[k];
    }
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < c.n; i += gridDim.x * kBlock) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const size_t q = (size_t)i * NR + k;
            c.p[q] = (X)((double)c.r[q] + beta[k] * ((double)c.p[q] - omega[k] * (double)c.v[q]));
        }
    }
}

// v = A yh ; alpha = rho / (r^ . v)
template <class V, class X, int NR>
__global__ void __launch_bounds__(kBlock) k_bi_spmv_v(BiArgs<V, X, NR> c) {
    Scalars* S = c.S;
    if (S->halt) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int g = gridDim.x;
    const int lb = xcd_block(blockIdx.x, g);
    const SliceRange sr = block_slices(lb, g, c.A.n_slices);
    const int s0 = sr.s0, s1 = sr.s1;
    double acc[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) acc[k] = 0.0;
    for (int sl = s0 + threadIdx.x / kWave; sl < s1; sl += kWavesPerBlock) {
        const int row = sl * kWave + lane;
        double out[NR];
        sell_rows_nr<V, X, NR>(c.A, sl, lane, row, c.yh, out);
        if (row < c.A.n) {
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const X vv = (X)out[k];
                c.v[(size_t)row * NR + k] = vv;
                acc[k] += (double)c.rh[(size_t)row * NR + k] * (double)vv;
            }
        }
    }
    block_partials<NR>(acc, c.red);
}

// alpha = rho / (r^ . v) per right-hand side
template <int NR>
struct FinBiSpmvV {
    Scalars* S;
    __device__ bool active() const { return !S->halt; }
    __device__ void operator()(const double* t) const {
        for (int k = 0; k < NR; ++k) {
            if (S->donek[k]) { S->alphak[k] = 0.0; continue; }
            if (t[k] == 0.0 || !isfinite(t[k])) {
                S->breakdown |= (t[k] == 0.0);
                S->nonfinite |= !isfinite(t[k]);
                S->donek[k] = 1;
                S->alphak[k] = 0.0;
            } else {
                S->alphak[k] = S->rhok[k] / t[k];
            }
        }
    }
};

// s = r - alpha v ; early exit test on ||s||
template <class V, class X, int NR>
__global__ void __launch_bounds__(kBlock) k_bi_s(BiArgs<V, X, NR> c) {
    Scalars* S = c.S;
    if (S->halt) return;
    double alpha[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) alpha[k] = S->alphak[k];
    double acc[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) acc[k] = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < c.n; i += gridDim.x * kBlock) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const size_t q = (size_t)i * NR + k;
            const X sv = (X)((double)c.r[q] - alpha[k] * (double)c.v[q]);
            c.s[q] = sv;
            acc[k] += (double)sv * (double)sv;
        }
    }
    block_partials<NR>(acc, c.red);
}

// ||s||: right-hand sides already converged at s exit early (x += alpha yh)
template <int NR>
struct FinBiS {
    Scalars* S;
    __device__ bool active() const { return !S->halt; }
    __device__ void operator()(const double* t) const {
        for (int k = 0; k < NR; ++k) {
            S->ssk[k] = t[k];
            const double scale = S->bbk[k] > 0.0 ? S->bbk[k] : 1.0;
            S->earlyk[k] = (!S->donek[k] && t[k] <= S->tol2 * scale) ? 1 : 0;
        }
    }
};

// t = A zh ; omega = (t . s) / (t . t)
template <class V, class X, int NR>
__global__ void __launch_bounds__(kBlock) k_bi_spmv_t(BiArgs<V, X, NR> c) {
    Scalars* S = c.S;
    if (S->halt) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int g = gridDim.x;
    const int lb = xcd_block(blockIdx.x, g);
    const SliceRange sr = block_slices(lb, g, c.A.n_slices);
    const int s0 = sr.s0, s1 = sr.s1;
    double acc[2 * NR];
#pragma unroll
    for (int k = 0; k < 2 * NR; ++k) acc[k] = 0.0;
    for (int sl = s0 + threadIdx.x / kWave; sl < s1; sl += kWavesPerBlock) {
        const int row = sl * kWave + lane;
        double out[NR];
        sell_rows_nr<V, X, NR>(c.A, sl, lane, row, c.zh, out);
        if (row < c.A.n) {
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const X tv = (X)out[k];
                c.t[(size_t)row * NR + k] = tv;
                acc[k] += (double)tv * (double)c.s[(size_t)row * NR + k];
                acc[NR + k] += (double)tv * (double)tv;
            }
        }
    }
    block_partials<2 * NR>(acc, c.red);
}

// omega = (t . s) / (t . t) per right-hand side
template <int NR>
struct FinBiSpmvT {
    Scalars* S;
    __device__ bool active() const { return !S->halt; }
    __device__ void operator()(const double* t) const {
        for (int k = 0; k < NR; ++k) {
            if (S->donek[k] || S->earlyk[k]) { S->omegak[k] = 0.0; continue; }
            if (t[NR + k] == 0.0 || !isfinite(t[NR + k])) {
                S->breakdown |= (t[NR + k] == 0.0);
                S->nonfinite |= !isfinite(t[NR + k]);
                S->omegak[k] = 0.0;
                S->donek[k] = 1;
            } else {
                S->omegak[k] = t[k] / t[NR + k];
            }
        }
    }
};

// x += alpha yh + omega zh ; r = s - omega t ; rr, rho = r^ . r, beta
template <class V, class X, int NR>
__global__ void __launch_bounds__(kBlock) k_bi_xr(BiArgs<V, X, NR> c) {
    Scalars* S = c.S;
    if (S->halt) return;
    double alpha[NR], omega[NR];
    bool skip[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        alpha[k] = S->alphak[k];
        omega[k] = S->omegak[k];
        skip[k] = S->donek[k] && !S->earlyk[k];
    }
    double acc[2 * NR];
#pragma unroll
    for (int k = 0; k < 2 * NR; ++k) acc[k] = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < c.n; i += gridDim.x * kBlock) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const size_t q = (size_t)i * NR + k;
            double rv = (double)c.r[q];
            if (!skip[k]) {
                c.x[q] += alpha[k] * (double)c.yh[q] + omega[k] * (double)c.zh[q];
                rv = (double)c.s[q] - omega[k] * (double)c.t[q];
                c.r[q] = (X)rv;
                rv = (double)(X)rv;
            }
            acc[k] += rv * rv;
            acc[NR + k] += (double)c.rh[q] * rv;
        }
    }
    block_partials<2 * NR>(acc, c.red);
}

// rho_new = r^ . r, beta, iteration count, history, convergence
template <int NR>
struct FinBiXr {
    Scalars* S;
    double* hist;
    __device__ bool active() const { return !S->halt; }
    __device__ void operator()(const double* t) const {
        const int it = ++S->iter;
        bool all = true;
        for (int k = 0; k < NR; ++k) {
            const double scale = S->bbk[k] > 0.0 ? S->bbk[k] : 1.0;
            if (!S->donek[k]) {
                const double rho_new = t[NR + k];
                S->betak[k] = (rho_new / S->rhok[k]) * (S->alphak[k] / S->omegak[k]);
                S->rhok[k] = rho_new;
                S->rrk[k] = t[k];
                if (S->earlyk[k] || t[k] <= S->tol2 * scale) S->donek[k] = 1;
                if (!isfinite(t[k])) { S->nonfinite = 1; S->donek[k] = 1; }
                if (rho_new == 0.0 && !S->donek[k]) { S->breakdown = 1; S->donek[k] = 1; }
                if (sqrt(t[k]) < S->replace_ratio * S->rrefk[k]) S->need_update = 1;
            }
            S->earlyk[k] = 0;
            if (hist && (it + 1) * NR <= S->hist_cap) hist[(size_t)it * NR + k] = sqrt(S->rrk[k] / scale);
            all = all && S->donek[k];
        }
        if (all || it >= S->max_iter) S->done = 1;
    }
};

// r = b - A x in FP64 (standard form), stored in X; on the first call also
// r^ = r, p = v = 0 and the scalar state; later calls (residual replacement)
// recompute rho = r^ . r.
template <class X, int NR>
__global__ void __launch_bounds__(kBlock) k_bi_residual(SellView<double> A, const double* __restrict__ b,
                                                       const double* __restrict__ x, X* r, X* rh, X* p,
                                                       X* v, RedScratch red, int first) {
    const int lane = threadIdx.x & (kWave - 1);
    const int g = gridDim.x;
    const int lb = xcd_block(blockIdx.x, g);
    const SliceRange sr = block_slices(lb, g, A.n_slices);
    const int s0 = sr.s0, s1 = sr.s1;
    double acc[3 * NR];
#pragma unroll
    for (int k = 0; k < 3 * NR; ++k) acc[k] = 0.0;
    for (int sl = s0 + threadIdx.x / kWave; sl < s1; sl += kWavesPerBlock) {
        const int row = sl * kWave + lane;
        double ax[NR];
        sell_rows_nr<double, double, NR>(A, sl, lane, row, x, ax);
        if (row < A.n) {
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const size_t q = (size_t)row * NR + k;
                const double bv = b[q];
                const X rv = (X)(bv - ax[k]);
                r[q] = rv;
                if (first) {
                    rh[q] = rv;
                    p[q] = (X)0;
                    v[q] = (X)0;
                }
                acc[k] += (double)rv * (double)rv;
                acc[NR + k] += bv * bv;
                acc[2 * NR + k] += (double)rh[q] * (double)rv;
            }
        }
    }
    block_partials<3 * NR>(acc, red);
}

// (Re)starts from the residual b - A x: bb and the initial scalars on the
// first call; convergence judged on this true residual, also for right-hand
// sides the recurrence had already declared converged.
template <int NR>
struct FinBiResidual {
    Scalars* S;
    double* hist;
    int first;
    __device__ bool active() const { return true; }
    __device__ void operator()(const double* t) const {
        bool all = true;
        for (int k = 0; k < NR; ++k) {
            if (first) {
                S->bbk[k] = t[NR + k];
                S->alphak[k] = 1.0;
                S->omegak[k] = 1.0;
                S->betak[k] = 0.0;
                S->donek[k] = 0;
                S->earlyk[k] = 0;
            }
            const double scale = S->bbk[k] > 0.0 ? S->bbk[k] : 1.0;
            S->rrk[k] = t[k];
            S->rrefk[k] = sqrt(t[k]);
            if (first && hist) hist[k] = sqrt(t[k] / scale);
            const int was_done = S->donek[k];
            S->donek[k] = (t[k] <= S->tol2 * scale) ? 1 : 0;
            if (!isfinite(t[k])) { S->nonfinite = 1; S->donek[k] = 1; }
            if (!S->donek[k]) {
                S->rhok[k] = t[2 * NR + k];
                if (was_done) { S->betak[k] = 0.0; S->alphak[k] = 1.0; S->omegak[k] = 1.0; }
            }
            all = all && S->donek[k];
        }
        S->done = all ? 1 : 0;
        S->halt = S->done;
        S->need_update = 0;
    }
};

namespace {

template <class V, class X, int NR>
void bi_iteration(ocsl_solver* h, const BiArgs<V, X, NR>& c, const void* m, X* tmp, X* tmp2,
                  const ocsl_options& o, int g_spmv, int g_vec) {
    k_bi_p<V, X, NR><<<g_vec, kBlock, 0, h->stream>>>(c);
    apply_precond<V, X, NR>(h, (const V*)m, c.p, tmp, tmp2, c.yh, o, c.S);
    k_bi_spmv_v<V, X, NR><<<g_spmv, kBlock, 0, h->stream>>>(c);
    k_finalize<NR, FinBiSpmvV<NR>><<<1, kBlock, 0, h->stream>>>(c.red, g_spmv, FinBiSpmvV<NR>{c.S});
    k_bi_s<V, X, NR><<<g_vec, kBlock, 0, h->stream>>>(c);
    k_finalize<NR, FinBiS<NR>><<<1, kBlock, 0, h->stream>>>(c.red, g_vec, FinBiS<NR>{c.S});
    apply_precond<V, X, NR>(h, (const V*)m, c.s, tmp, tmp2, c.zh, o, c.S);
    k_bi_spmv_t<V, X, NR><<<g_spmv, kBlock, 0, h->stream>>>(c);
    k_finalize<2 * NR, FinBiSpmvT<NR>><<<1, kBlock, 0, h->stream>>>(c.red, g_spmv, FinBiSpmvT<NR>{c.S});
    k_bi_xr<V, X, NR><<<g_vec, kBlock, 0, h->stream>>>(c);
    k_finalize<2 * NR, FinBiXr<NR>><<<1, kBlock, 0, h->stream>>>(c.red, g_vec, FinBiXr<NR>{c.S, c.hist});
}

int round8b(int g) { return g >= kXcds ? (g / kXcds) * kXcds : g; }

template <class V, class X, int NR>
void bicgstab_run(ocsl_solver* h, bool mixed, const ocsl_options& o, const double* b_il, double* x_il,
                  ocsl_report* reps) {
    const int n = h->host.n;
    Sell& s = sell_for(h, false);
    ensure_sell_values(h, s, true, mixed);
    ensure_csr_values(h, false);
    if (o.tri_mode == OCSL_TRI_EXACT) ensure_levels(h);
    Work& w = h->work;
    // X vectors: r, rh, p, v, s, t, yh, zh, tmp, tmp2 -> 10 x n*NR
    X* vec[10];
    for (int i = 0; i < 10; ++i) {
        if (sizeof(X) == 8) vec[i] = (X*)work_buf(h, w.d[i], (size_t)n * NR);
        else vec[i] = (X*)work_buf(h, w.f[i], (size_t)n * NR);
    }
    work_buf(h, h->hist, (size_t)(o.max_iter + 2) * NR);

    Scalars init{};
    init.tol2 = o.tol * o.tol;
    init.max_iter = o.max_iter;
    init.replace_ratio = mixed ? o.replace_ratio : 0.0;
    init.hist_cap = (o.max_iter + 2) * NR;
    OCSL_HIP(hipMemcpyAsync(h->S.p, &init, sizeof(Scalars), hipMemcpyHostToDevice, h->stream));

    OCSL_HIP(hipEventRecord(h->ev0, h->stream));
    ilu_decompose_device(h, o.ilu_sweeps, mixed);
    const void* m = mixed ? (const void*)h->ilu_m32.p : (const void*)h->ilu_m.p;
    BiArgs<V, X, NR> c{s.view<V>(), vec[1], vec[0], vec[2], vec[3], vec[4], vec[5], vec[6], vec[7],
                       x_il, h->S.p, h->hist.p, h->red(), n};
    const int g_spmv = round8b(std::min(kMaxRedBlocks, (s.n_slices + kWavesPerBlock - 1) / kWavesPerBlock));
    const int g_vec = round8b(red_blocks(kBlock * 4, n));
    auto residual = [&](int first) {
        k_bi_residual<X, NR><<<g_spmv, kBlock, 0, h->stream>>>(s.view<double>(), b_il, x_il, c.r, (X*)c.rh, c.p,
                                                               c.v, h->red(), first);
        k_finalize<3 * NR, FinBiResidual<NR>><<<1, kBlock, 0, h->stream>>>(
            h->red(), g_spmv, FinBiResidual<NR>{h->S.p, h->hist.p, first});
        OCSL_HIP(hipGetLastError());
    };
    auto read_S = [&] {
        OCSL_HIP(hipMemcpyAsync(h->S_host, h->S.p, sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
        sync(h);
    };
    residual(1);
    read_S();
    int replacements = 0;
    const int kind = 16 + (mixed ? 1 : 0) + 2 * NR + 64 * o.tri_mode + 256 * o.tri_sweeps;
    int done_iters = 0;
    while (!h->S_host->done) {
        const int left = o.max_iter - h->S_host->iter;
        if (left <= 0) break;
        // geometric chunk growth: few wasted (early-exiting) launches on fast
        // convergence, few host checks on slow convergence
        int chunk = std::max(1, std::min({o.check_every, left, std::max(1, done_iters)}));
        const bool graph = o.use_graph && chunk == o.check_every;
        if (graph) {
            const GraphKey key = bicgstab_graph_key(kind, chunk, o.tri_weight);
            auto it = h->graphs.find(key);
            if (it == h->graphs.end()) {
                hipGraph_t gr;
                OCSL_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                for (int k = 0; k < chunk; ++k) bi_iteration<V, X, NR>(h, c, m, vec[8], vec[9], o, g_spmv, g_vec);
                OCSL_HIP(hipStreamEndCapture(h->stream, &gr));
                hipGraphExec_t ex;
                OCSL_HIP(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
                OCSL_HIP(hipGraphDestroy(gr));
                it = h->graphs.emplace(key, ex).first;
            }
            OCSL_HIP(hipGraphLaunch(it->second, h->stream));
        } else {
            for (int k = 0; k < chunk; ++k) bi_iteration<V, X, NR>(h, c, m, vec[8], vec[9], o, g_spmv, g_vec);
            OCSL_HIP(hipGetLastError());
        }
        read_S();
        done_iters = h->S_host->iter;
        const Scalars& S = *h->S_host;
        if (S.nonfinite) break;
        if (mixed && (S.need_update || S.done)) {
            residual(0);
            ++replacements;
            read_S();
        }
    }
    OCSL_HIP(hipEventRecord(h->ev1, h->stream));
    read_S();
    float ms = 0.f;
    OCSL_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    const Scalars S = *h->S_host;
    const int iters = S.iter;
    h->hist_len = std::min(iters + 1, o.max_iter + 2) * NR;
    h->hist_host.assign(h->hist_len, 0.0);
    OCSL_HIP(hipMemcpy(h->hist_host.data(), h->hist.p, sizeof(double) * h->hist_len, hipMemcpyDeviceToHost));
    // true residual of each right-hand side
    residual(0);
    read_S();
    const double nnz = (double)h->host.col.size();
    const double sv = mixed ? 4.0 : 8.0;
    const double b_spmv = nnz * (sv + 4.0) + 4.0 * (s.n_slices + 1) + (1.0 + NR * 3.0) * n * sv;
    const double b_tri = nnz * (sv + 4.0) + 4.0 * (n + 1) + 4.0 * n + 2.0 * NR * n * sv;
    const double b_vec = NR * n * (sv * 4.0 + sv * 3.0 + sv * 6.0 + 16.0 + 2.0 * sv);
    for (int k = 0; k < NR; ++k) {
        ocsl_report* r = reps + k;
        const double scale = h->S_host->bbk[k] > 0.0 ? h->S_host->bbk[k] : 1.0;
        const double tr = std::sqrt(h->S_host->rrk[k] / scale);
        r->iterations = iters;
        r->replacements = replacements;
        r->history_len = h->hist_len;
        r->true_relres = tr;
        r->residual = tr;
        r->solve_ms = ms;
        r->bytes = iters * (2.0 * b_spmv + 4.0 * b_tri + b_vec) / NR;
        int st = OCSL_OK;
        if (S.breakdown) st |= OCSL_WARN_BREAKDOWN;
        if (S.nonfinite || !std::isfinite(tr)) st |= OCSL_WARN_NONFINITE;
        if (!(tr <= o.tol * 1.0000001) && !(S.donek[k] && !mixed && st == OCSL_OK)) st |= OCSL_WARN_MAXITER;
        r->status = st;
    }
}

}  // namespace

// Interleaves k vectors of n (column blocks) into n x k rows.
template <class T>
__global__ void k_interleave(int n, int k, const T* __restrict__ in, T* __restrict__ out, int to_rows) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int c = 0; c < k; ++c) {
            if (to_rows) out[(size_t)i * k + c] = in[(size_t)c * n + i];
            else out[(size_t)c * n + i] = in[(size_t)i * k + c];
        }
}

// k column blocks of n between two row orders: gather out[c n + i] =
// in[c n + perm[i]], or scatter out[c n + perm[i]] = in[c n + i].
__global__ void k_perm_cols(int n, int k, const int* __restrict__ perm, const double* __restrict__ in,
                            double* __restrict__ out, int gather) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int j = perm[i];
        for (int c = 0; c < k; ++c) {
            if (gather) out[(size_t)c * n + i] = in[(size_t)c * n + j];
            else out[(size_t)c * n + j] = in[(size_t)c * n + i];
        }
    }
}

namespace {

// Greedy colouring of the symmetrised row graph (A and A^T), rows visited in
// reverse Cuthill-McKee order, each taking the smallest colour none of its
// coloured neighbours has; rows renumbered colour by colour, in visiting
// order within a colour.  Returns new -> old.
std::vector<int> colour_order(const HostCSR& a, int* n_colours) {
    const int n = a.n;
    std::vector<int> ptr(n + 1, 0);
    for (int i = 0; i < n; ++i)
        for (int p = a.rp[i]; p < a.rp[i + 1]; ++p)
            if (a.col[p] != i) {
                ++ptr[i + 1];
                ++ptr[a.col[p] + 1];
            }
    for (int i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
    std::vector<int> adj(ptr[n]), fill(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int p = a.rp[i]; p < a.rp[i + 1]; ++p) {
            const int j = a.col[p];
            if (j != i) {
                adj[fill[i]++] = j;
                adj[fill[j]++] = i;
            }
        }
    const std::vector<int> visit = rcm_order(a);
    std::vector<int> colour(n, -1), seen;
    int nc = 0;
    for (int v : visit) {
        // colours of v's neighbours are < its degree + 1
        seen.assign((size_t)(ptr[v + 1] - ptr[v]) + 1, 0);
        for (int q = ptr[v]; q < ptr[v + 1]; ++q) {
            const int c = colour[adj[q]];
            if (c >= 0 && c < (int)seen.size()) seen[c] = 1;
        }
        int c = 0;
        while (seen[c]) ++c;
        colour[v] = c;
        nc = std::max(nc, c + 1);
    }
    std::vector<int> start(nc + 1, 0), order(n);
    for (int v = 0; v < n; ++v) ++start[colour[v] + 1];
    for (int c = 0; c < nc; ++c) start[c + 1] += start[c];
    for (int v : visit) order[start[colour[v]]++] = v;
    *n_colours = nc;
    return order;
}

// The handle of the colour-renumbered matrix P A P^T (built once): its rows
// in the new order, each row's entries by ascending new column; the value
// map points every entry at its slot in the caller's value array.
ocsl_solver* coloured(ocsl_solver* h) {
    if (h->col_h) return h->col_h;
    const HostCSR& a = h->host;
    const int n = a.n;
    int nc = 0;
    const std::vector<int> order = colour_order(a, &nc);
    std::vector<int> inv(n);
    for (int i = 0; i < n; ++i) inv[order[i]] = i;
    std::vector<uint64_t> rp(n + 1, 0), col;
    std::vector<int> map;
    col.reserve(a.col.size());
    map.reserve(a.col.size());
    std::vector<std::pair<int, int>> row;
    for (int i = 0; i < n; ++i) {
        const int o = order[i];
        row.clear();
        for (int p = a.rp[o]; p < a.rp[o + 1]; ++p) row.emplace_back(inv[a.col[p]], a.src[p]);
        std::sort(row.begin(), row.end());
        for (const auto& e : row) {
            col.push_back((uint64_t)e.first);
            map.push_back(e.second);
        }
        rp[i + 1] = col.size();
    }
    ocsl_handle t = nullptr;
    const ocsl_status st = ocsl_create(&t, n, rp.data(), col.data());
    if (st != OCSL_OK) throw Error{st, std::string("colour-ordered handle: ") + ocsl_last_error()};
    h->col_h = t;
    h->n_colours = nc;
    h->col_map.upload(map.data(), map.size(), h->stream);
    h->col_perm.upload(order.data(), order.size(), h->stream);
    h->col_epoch = ~0ull;
    sync(h);
    return t;
}

}  // namespace

void bicgstab_multi(ocsl_solver* h, bool mixed, const ocsl_options& o, int k, const double* B_dev,
                    double* X_dev, ocsl_report* reps);

// The colour order and its colour count (ocsl_colour_order).
void colour_order_get(ocsl_solver* h, int* order, int* n_colours) {
    (void)coloured(h);
    if (order) {
        OCSL_HIP(hipMemcpyAsync(order, h->col_perm.p, sizeof(int) * h->host.n, hipMemcpyDeviceToHost, h->stream));
        sync(h);
    }
    if (n_colours) *n_colours = h->n_colours;
}

// TOLERANCE ILU(0)-BiCGSTAB in the colour order: values gathered into the
// renumbered handle (when they changed), b and x permuted in, the solve run
// there, x permuted back.  Reports and history are the inner solve's.
static void bicgstab_coloured(ocsl_solver* h, bool mixed, const ocsl_options& o, int k, const double* B_dev,
                              double* X_dev, ocsl_report* reps) {
    ocsl_solver* t = coloured(h);
    const int n = h->host.n;
    const int64_t nr = (int64_t)h->host.col.size();
    sync(h);  // the caller's values, b and x are complete
    if (h->col_epoch != h->val_epoch) {
        t->val_in_own.alloc(nr);
        k_gather_vals<double><<<grid_for(nr), kBlock, 0, t->stream>>>(nr, h->col_map.p, h->val_in, t->val_in_own.p);
        OCSL_HIP(hipGetLastError());
        t->val_in = t->val_in_own.p;
        t->have_vals = true;
        t->val_epoch++;
        t->csr_val_valid = t->csr_val32_valid = false;
        h->col_epoch = h->val_epoch;
    }
    const size_t cnt = (size_t)n * k;
    h->col_b.alloc(cnt);
    h->col_x.alloc(cnt);
    k_perm_cols<<<grid_for(n), kBlock, 0, t->stream>>>(n, k, h->col_perm.p, B_dev, h->col_b.p, 1);
    k_perm_cols<<<grid_for(n), kBlock, 0, t->stream>>>(n, k, h->col_perm.p, X_dev, h->col_x.p, 1);
    OCSL_HIP(hipGetLastError());
    ocsl_options oi = o;
    oi.bicgstab_order = OCSL_BICGSTAB_ORDER_INPUT;
    bicgstab_multi(t, mixed, oi, k, h->col_b.p, h->col_x.p, reps);
    k_perm_cols<<<grid_for(n), kBlock, 0, t->stream>>>(n, k, h->col_perm.p, h->col_x.p, X_dev, 0);
    OCSL_HIP(hipGetLastError());
    sync(t);
    h->hist_host = t->hist_host;
    h->hist_len = t->hist_len;
}

void bicgstab_multi(ocsl_solver* h, bool mixed, const ocsl_options& o, int k, const double* B_dev,
                    double* X_dev, ocsl_report* reps) {
    OCSL_REQUIRE(o.ilu_sweeps >= 1, OCSL_ERR_ARG, "ilu_sweeps must be >= 1");
    OCSL_REQUIRE(o.bicgstab_order == OCSL_BICGSTAB_ORDER_INPUT || o.bicgstab_order == OCSL_BICGSTAB_ORDER_COLOUR,
                 OCSL_ERR_ARG, "bad bicgstab_order");
    if (o.bicgstab_order == OCSL_BICGSTAB_ORDER_COLOUR) {
        bicgstab_coloured(h, mixed, o, k, B_dev, X_dev, reps);
        return;
    }
    const int n = h->host.n;
    Work& w = h->work;
    double* b_il = work_buf(h, w.d[14], (size_t)n * k);
    double* x_il = work_buf(h, w.d[15], (size_t)n * k);
    const int gv = grid_for(n);
    if (k == 1) {
        OCSL_HIP(hipMemcpyAsync(b_il, B_dev, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
        OCSL_HIP(hipMemcpyAsync(x_il, X_dev, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
    } else {
        k_interleave<double><<<gv, kBlock, 0, h->stream>>>(n, k, B_dev, b_il, 1);
        k_interleave<double><<<gv, kBlock, 0, h->stream>>>(n, k, X_dev, x_il, 1);
    }
    OCSL_HIP(hipGetLastError());
    if (k == 1) {
        if (mixed) bicgstab_run<float, float, 1>(h, true, o, b_il, x_il, reps);
        else bicgstab_run<double, double, 1>(h, false, o, b_il, x_il, reps);
    } else if (k == 2) {
        if (mixed) bicgstab_run<float, float, 2>(h, true, o, b_il, x_il, reps);
        else bicgstab_run<double, double, 2>(h, false, o, b_il, x_il, reps);
    } else if (k == 3) {
        if (mixed) bicgstab_run<float, float, 3>(h, true, o, b_il, x_il, reps);
        else bicgstab_run<double, double, 3>(h, false, o, b_il, x_il, reps);
    } else {
        if (mixed) bicgstab_run<float, float, 4>(h, true, o, b_il, x_il, reps);
        else bicgstab_run<double, double, 4>(h, false, o, b_il, x_il, reps);
    }
    if (k == 1)
        OCSL_HIP(hipMemcpyAsync(X_dev, x_il, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
    else
        k_interleave<double><<<gv, kBlock, 0, h->stream>>>(n, k, x_il, X_dev, 0);
    OCSL_HIP(hipGetLastError());
    sync(h);
}

void bicgstab_tolerance(ocsl_solver* h, bool mixed, const ocsl_options& o, ocsl_report* rep) {
    bicgstab_multi(h, mixed, o, 1, h->b_dev, h->x_own.p, rep);
    h->hist_len = std::min<int>(h->hist_len, (int)h->hist_host.size());
}

}  // namespace ocsl

// ------------------------------------------------------------------ factor-level entry points
namespace ocsl {
namespace {

template <class F>
ocsl_status tri_guard(F&& f) {
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

// The caller's factor (its slot order, sentinels included) in the handle's
// real-entry order.
void gather_factor(ocsl_solver* h, const double* m, int where, DBuf<double>& stage, DBuf<double>& real) {
    const double* md = m;
    if (where == OCSL_HOST) {
        stage.upload(m, (size_t)h->host.nnz_in, h->stream);
        md = stage.p;
    }
    const int64_t nr = (int64_t)h->host.col.size();
    real.alloc((size_t)nr);
    k_gather_vals<double><<<grid_for(nr), kBlock, 0, h->stream>>>(nr, h->src.p, md, real.p);
    OCSL_HIP(hipGetLastError());
}

// y = L^-1 b (lower) or x = U^-1 y (upper) with the caller's factor m.
ocsl_status tri_api(ocsl_handle h, const double* m, const double* in, double* out, int where, bool lower) {
    if (!h || !m || !in || !out || (where != OCSL_HOST && where != OCSL_DEVICE)) return OCSL_ERR_ARG;
    return tri_guard([&] {
        const int n = h->host.n;
        DBuf<double> mstage, mreal, ib, ob;
        gather_factor(h, m, where, mstage, mreal);
        ensure_levels(h);
        ++h->ilu_epoch;  // the level-ordered copies follow this factor
        const double* din = in;
        double* dout = out;
        if (where == OCSL_HOST) {
            ib.upload(in, (size_t)n, h->stream);
            ob.alloc((size_t)n);
            din = ib.p;
            dout = ob.p;
        }
        tri_exact<double, double, 1>(h, mreal.p, din, dout, lower, nullptr);
        OCSL_HIP(hipGetLastError());
        if (where == OCSL_HOST)
            OCSL_HIP(hipMemcpyAsync(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
        sync(h);
    });
}

}  // namespace
}  // namespace ocsl

using namespace ocsl;

extern "C" ocsl_status ocsl_ILU_forsub(ocsl_handle h, const double* m, const double* b, double* y, ocsl_where where) {
    return tri_api(h, m, b, y, where, true);
}

extern "C" ocsl_status ocsl_ILU_backsub(ocsl_handle h, const double* m, const double* y, double* x,
                                       ocsl_where where) {
    return tri_api(h, m, y, x, where, false);
}

extern "C" ocsl_status ocsl_ILU_decompose_from(ocsl_handle h, int sweeps, const double* m0, double* m_out,
                                              ocsl_where where) {
    if (!h || !m0 || !m_out || sweeps < 0 || (where != OCSL_HOST && where != OCSL_DEVICE)) return OCSL_ERR_ARG;
    return tri_guard([&] {
        ensure_csr_values(h, false);
        DBuf<double> mstage, mreal, full;
        gather_factor(h, m0, where, mstage, mreal);
        ilu_decompose_device(h, sweeps, false, mreal.p);
        const int64_t nnz_in = h->host.nnz_in;
        full.alloc((size_t)nnz_in);
        k_fill<double><<<grid_for(nnz_in), kBlock, 0, h->stream>>>((int)nnz_in, 0.0, full.p);
        k_permute_out<double><<<grid_for((int64_t)h->host.col.size()), kBlock, 0, h->stream>>>(
            (int)h->host.col.size(), h->src.p, h->ilu_m.p, full.p);
        OCSL_HIP(hipGetLastError());
        OCSL_HIP(hipMemcpyAsync(m_out, full.p, sizeof(double) * nnz_in,
                                where == OCSL_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, h->stream));
        sync(h);
    });
}

extern "C" ocsl_status ocsl_time_tri_solve(ocsl_handle h, int lower, int reps, double* avg_ms) {
    return ocsl_time_tri_solve_form(h, lower, 0, reps, avg_ms);
}

extern "C" ocsl_status ocsl_time_tri_solve_form(ocsl_handle h, int lower, int form, int reps, double* avg_ms) {
    if (!h || !avg_ms || reps <= 0 || (lower != 0 && lower != 1) || form < 0 || form > 3) return OCSL_ERR_ARG;
    struct Restore {
        ocsl_solver* h;
        int form;
        ~Restore() { h->tri_form = form; }
    } restore{h, h->tri_form};
    h->tri_form = form;
    return tri_guard([&] {
        OCSL_REQUIRE(h->have_vals, OCSL_ERR_STATE, "values not set");
        const int n = h->host.n;
        ensure_csr_values(h, false);
        ensure_levels(h);
        OCSL_REQUIRE(form != 3 || use_serial_tri(h, lower ? h->lev_fwd : h->lev_bwd, n), OCSL_ERR_ARG,
                     "the serial form does not fit this system");
        ilu_decompose_device(h, 3, false);
        DBuf<double> in, out;
        in.alloc((size_t)n);
        out.alloc((size_t)n);
        k_fill<double><<<grid_for(n), kBlock, 0, h->stream>>>(n, 1.0, in.p);
        OCSL_HIP(hipGetLastError());
        tri_exact<double, double, 1>(h, h->ilu_m.p, in.p, out.p, lower != 0, nullptr);  // warm: level values
        OCSL_HIP(hipEventRecord(h->ev0, h->stream));
        for (int r = 0; r < reps; ++r) tri_exact<double, double, 1>(h, h->ilu_m.p, in.p, out.p, lower != 0, nullptr);
        OCSL_HIP(hipEventRecord(h->ev1, h->stream));
        sync(h);
        float ms = 0.f;
        OCSL_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *avg_ms = ms / reps;
    });
}
