// Batched solves of many independent CSR systems in the same launches.
//
// The systems' SELL-P structures (reverse Cuthill-McKee order, 16-bit column
// offsets, R = 1, 2 or 4 rows per lane) are concatenated, each padded to
// whole slices of 64 R rows so no slice and no workgroup straddles two
// systems.  One block map -- workgroup ->
// (system, slice range) -- serves every kernel, so the partial sums of system
// k always sit in the contiguous block range bptr[k]..bptr[k+1] and a
// consuming block re-adds only its own system's partials.  Every system has
// its own scalars (rr, alpha, beta, flags): the solves are independent, they
// merely share the launches.  An iteration is two kernels (the merged-
// reduction CG of the single-system mixed solver); blocks of converged
// systems return at once, and chunks of iterations replay as one hipGraph.
// MIXED: when a system's FP32 recurrence has dropped by replace_ratio since
// its last replacement, the system's next iteration is an FP64 residual
// replacement (r = b - A x) inside the same launches, so the replacement
// comes one iteration after its trigger whatever the chunk length; a claim
// of convergence is confirmed on the true FP64 residual between chunks.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <tuple>

#include "ocsl_handle.h"
#include "ocsl_kernels.h"

namespace ocsl {

// Most threads per workgroup of the update kernel, which covers the rows of
// one SpMV workgroup with one 16-byte vector per thread and round.
constexpr int kBtUpdateBlock = 256;

// Most iterations per FP64 update of x (the period of the MIXED solves).
constexpr int kBtMaxPeriod = 8;

struct SysScalars {
    double rr, bb, r_ref, tol2, alpha, beta, replace_ratio;
    int iter, max_iter, done, halt, need_update, breakdown, nonfinite, replacements;
    int verified;     // convergence confirmed on the true FP64 residual (MIXED)
    int replace_now;  // the next SpMV kernel replaces r instead (set by the update kernel)
    int replaced;     // this iteration's SpMV kernel replaced r (set by the SpMV kernel)
    // x period (see k_bt_update): the steps of the current period, whether
    // position j's update left x holding every step up to j, the direction
    // buffer of the iteration at which the system stopped, and the stop of
    // this iteration (its update adds the pending steps)
    double alpha_q[kBtMaxPeriod];
    int added[kBtMaxPeriod];
    int pdir, flush_now;
};

// partials of the true r.r of an in-iteration replacement (the batch's
// partial scratch holds 8 slots per block)
constexpr int kBtSlotTR = 5;

struct BatchMap {
    const int* bsys;    // block -> system
    const int* bs0;     // block -> first slice (global)
    const int* bs1;     // block -> one past its last slice
    const int* bptr;    // system -> first block (m + 1)
    const int* active;  // launched workgroup -> block (null: all blocks, in order)
    __device__ int block() const { return active ? active[blockIdx.x] : (int)blockIdx.x; }
};

template <class V, class X>
struct BtArgs {
    SellPView<V> A;  // concatenated; n = padded total rows
    BatchMap map;
    X* r;
    X* p;      // this iteration's direction: P_j
    X* q;
    double* x;
    SysScalars* sys;
    RedScratchBig red;
    int slice_rows;        // 64 R
    SellPView<double> A64; // FP64 values (the residual replacement)
    const double* b;
    X* pb;          // direction buffers P_0 .. P_{M-1}, pstride elements apart
    int64_t pstride;
    int M;          // iterations per FP64 update of x
    int j;          // this iteration's position in the period
    X* p_alt;       // the next direction: P_{(j + 1) mod M}
};

// Sum of system k's partials of one slot; identical in all its blocks.
__device__ __forceinline__ double sys_sum(const RedScratchBig& red, int slot, int b0, int b1) {
    __shared__ double sh[16];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int nw = blockDim.x / kWave;
    double s = 0.0;
    for (int i = b0 + tid; i < b1; i += blockDim.x) s += red.partial[(size_t)slot * red.stride + i];
    s = wave_sum(s);
    if (lane == 0) sh[wid] = s;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += sh[w];
    __syncthreads();
    return t;
}

// K slots of system k's partials (blocks [b0, b1)) in two phases, as the
// single-system solver does it: issue() starts the first wave's loads of up
// to 512 partials per slot (raw loads at clamped indices, before the kernel's
// vector loads, which complete after them); finish() adds the rest, reduces
// across the wave and hands the totals to the block through LDS with one
// barrier (once per kernel).  Identical in all of the system's blocks.
template <int K>
struct SysSums {
    static constexpr int kRounds = 8;
    double v[K][kRounds];
    __device__ __forceinline__ void issue(const RedScratchBig& red, const int (&slot)[K], int b0, int b1) {
        if (threadIdx.x >= kWave) return;
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int r = 0; r < kRounds; ++r)
                v[k][r] = red.partial[(size_t)slot[k] * red.stride + min(b0 + (int)threadIdx.x + kWave * r, b1 - 1)];
    }
    __device__ __forceinline__ void finish(const RedScratchBig& red, const int (&slot)[K], int b0, int b1,
                                           double (&out)[K]) {
        __shared__ double sh[K];
        if (threadIdx.x < kWave) {
            const int lane = threadIdx.x;
            double a[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                a[k] = 0.0;
#pragma unroll
                for (int r = 0; r < kRounds; ++r) a[k] += b0 + lane + kWave * r < b1 ? v[k][r] : 0.0;
            }
            for (int i = b0 + lane + kWave * kRounds; i < b1; i += kWave)
#pragma unroll
                for (int k = 0; k < K; ++k) a[k] += red.partial[(size_t)slot[k] * red.stride + i];
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

template <int NR>
__device__ __forceinline__ void bt_partials(const double (&v)[NR], const RedScratchBig& red, int slot0, int gb) {
    __shared__ double sh[NR][16];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int nw = blockDim.x / kWave;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const double s = wave_sum(v[r]);
        if (lane == 0) sh[r][wid] = s;
    }
    __syncthreads();
    if (tid == 0)
        for (int r = 0; r < NR; ++r) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += sh[r][w];
            red.partial[(size_t)(slot0 + r) * red.stride + gb] = s;
        }
}

// A system takes the residual pass between chunks when it claims convergence
// not yet confirmed on the true residual.
__device__ __host__ __forceinline__ bool bt_wants_residual(const SysScalars& S) {
    return S.done && !S.verified && !S.nonfinite && !S.breakdown;
}

// r = b - A x (FP64) for the systems being (re)started; p = r (in P_0) on
// the first call; partials of r.r and b.b.  One wave per slice of 64 R rows;
// every row of the concatenation exists (systems are padded to whole slices,
// padding rows are zero rows with zero right-hand side).  A system that goes
// on iterating after this pass resumes at period position phi (>= 0): its
// direction moves from the buffer it stopped at (pdir) to P_phi.
template <class X, int R>
__global__ void __launch_bounds__(1024) k_bt_residual(SellPView<double> A, BatchMap map, const double* __restrict__ b,
                                                    const double* __restrict__ x, X* r, X* pb, int64_t pstride,
                                                    int phi, SysScalars* sys, RedScratchBig red, int first) {
    const int gb = map.block();
    const int k = map.bsys[gb];
    if (!first && !bt_wants_residual(sys[k])) return;
    const int lane = threadIdx.x & (kWave - 1);
    X* p = pb;  // P_0
    const int pdir = sys[k].pdir;
    const bool move = !first && phi >= 0 && pdir != phi;
    const X* p_from = pb + (int64_t)pdir * pstride;
    X* p_to = pb + (int64_t)(phi > 0 ? phi : 0) * pstride;
    double acc[2] = {0.0, 0.0};
    for (int s = map.bs0[gb] + (int)(threadIdx.x / kWave); s < map.bs1[gb];
         s += (int)(blockDim.x / kWave)) {
        const int row0 = s * kWave * R + lane * R;
        double xi[R], ax[R], bv[R];
        load_rows_p<R>(x, row0, xi);
        load_rows_p<R>(b, row0, bv);
        sellp_rows<double, double, false, R>(A, s, lane, row0, x, xi, ax);
        X rv[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const double rj = bv[j] - ax[j];
            acc[0] += rj * rj;
            acc[1] += bv[j] * bv[j];
            rv[j] = (X)rj;
        }
        store_rows<R>(r, row0, A.n, rv);
        if (first) store_rows<R>(p, row0, A.n, rv);
        if (move) {
            X pv[R];
            load_rows_tp<R>(p_from, row0, pv);
            store_rows<R>(p_to, row0, A.n, pv);
        }
    }
    bt_partials<2>(acc, red, kSlotRR, gb);
}

// One block per system: bb, rr, the convergence test on the true residual.
// A system that goes on iterating resumes at period position phi: the
// positions before it hold no pending step.
__global__ void k_bt_fin_residual(BatchMap map, SysScalars* sys, RedScratchBig red, int first, int m, int phi) {
    const int k = blockIdx.x;
    if (k >= m) return;
    SysScalars& S = sys[k];
    if (!first && !bt_wants_residual(S)) return;
    const double rr = sys_sum(red, kSlotRR, map.bptr[k], map.bptr[k + 1]);
    const double bb = sys_sum(red, kSlotBB, map.bptr[k], map.bptr[k + 1]);
    if (threadIdx.x == 0) {
        if (first) S.bb = bb;
        else S.replacements++;
        S.rr = rr;
        S.r_ref = sqrt(rr);
        const double scale = S.bb > 0.0 ? S.bb : 1.0;
        const int conv = (rr <= S.tol2 * scale) ? 1 : 0;
        // at the iteration cap the system stays finished, converged or not
        S.done = conv || S.iter >= S.max_iter;
        S.verified = S.done;
        if (!isfinite(rr)) { S.nonfinite = 1; S.done = 1; S.verified = 1; }
        S.halt = S.done;
        S.need_update = 0;
        S.replace_now = 0;
        S.replaced = 0;
        S.flush_now = 0;
        for (int i = 0; i < kBtMaxPeriod; ++i) S.added[i] = i < phi ? 1 : 0;
    }
}

// K1: per-system rr (first block of the system), q = A p, partials p.q, r.q,
// q.q; or, for a system whose replacement is due, r = b - A x in FP64 with
// partials of r.r (the iteration's update kernel then adopts it).
template <class V, class X, int R>
__global__ void __launch_bounds__(1024) k_bt_spmv(BtArgs<V, X> c) {
    const int gb = c.map.block();
    const int k = c.map.bsys[gb];
    SysScalars* S = c.sys + k;
    // the wave's first slice opens with its loads (direction rows, slice
    // start, diagonals) issued beside those of the system's flags, not after
    // them; each later slice's at the end of the one before
    const int lane = threadIdx.x & (kWave - 1);
    const int s1 = c.map.bs1[gb], step = (int)(blockDim.x / kWave);
    const int sf = c.map.bs0[gb] + (int)(threadIdx.x / kWave);
    X pf[R] = {};
    SellpHead<V, R> hf{};
    if (sf < s1) {
        const int row0 = sf * kWave * R + lane * R;
        load_rows_tp<R>(c.p, row0, pf);
        hf = sellp_head<V, R>(c.A, sf, row0);
    }
    if (S->done || S->breakdown || S->nonfinite) {
        if (gb == c.map.bptr[k] && threadIdx.x == 0) {
            S->halt = 1;
            S->done = 1;       // a breakdown found by the last update stops the system here
            S->flush_now = 0;  // the update after the stop has added the pending steps
        }
        return;
    }
    if (S->replace_now) {
        double acc[1] = {0.0};
        for (int s = c.map.bs0[gb] + (int)(threadIdx.x / kWave); s < c.map.bs1[gb];
             s += (int)(blockDim.x / kWave)) {
            const int row0 = s * kWave * R + lane * R;
            double xi[R], ax[R], bv[R];
            load_rows_p<R>(c.x, row0, xi);
            load_rows_p<R>(c.b, row0, bv);
            sellp_rows<double, double, false, R>(c.A64, s, lane, row0, c.x, xi, ax);
            X rv[R];
#pragma unroll
            for (int j = 0; j < R; ++j) {
                rv[j] = (X)(bv[j] - ax[j]);
                acc[0] += (double)rv[j] * (double)rv[j];  // r.r of the stored r
            }
            store_rows<R>(c.r, row0, c.A.n, rv);
        }
        bt_partials<1>(acc, c.red, kBtSlotTR, gb);
        if (gb == c.map.bptr[k] && threadIdx.x == 0) S->replaced = 1;
        return;
    }
    if (gb == c.map.bptr[k]) {
        const double rr = sys_sum(c.red, kSlotRR, c.map.bptr[k], c.map.bptr[k + 1]);
        if (threadIdx.x == 0) {
            S->rr = rr;
            const double scale = S->bb > 0.0 ? S->bb : 1.0;
            if (!(rr > S->tol2 * scale) || S->iter >= S->max_iter) S->done = 1;
            if (!isfinite(rr)) { S->nonfinite = 1; S->done = 1; }
            if (S->done) {
                // this iteration's update adds the period's pending steps;
                // the direction stays in P_j
                S->flush_now = 1;
                S->pdir = c.j;
            }
            S->need_update = sqrt(rr) < S->replace_ratio * S->r_ref;
            S->replaced = 0;
        }
    }
    double acc[3] = {0.0, 0.0, 0.0};
    for (int s = sf; s < s1; s += step) {
        const int row0 = s * kWave * R + lane * R;
        double pi[R], rv[R], qd[R];
#pragma unroll
        for (int j = 0; j < R; ++j) pi[j] = (double)pf[j];
        const X* px = c.p;
        sellp_rows_gh<V, sizeof(V) == 4, R>(c.A, hf, lane, row0, [px](int cc) { return (double)px[cc]; }, pi, qd);
        load_rows_p<R>(c.r, row0, rv);
        X qv[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            qv[j] = (X)qd[j];
            acc[0] += pi[j] * (double)qv[j];
            acc[1] += rv[j] * (double)qv[j];
            acc[2] += (double)qv[j] * (double)qv[j];
        }
        store_rows<R>(c.q, row0, c.A.n, qv);
        if (s + step < s1) {
            const int row1 = (s + step) * kWave * R + lane * R;
            load_rows_tp<R>(c.p, row1, pf);
            hf = sellp_head<V, R>(c.A, s + step, row1);
        }
    }
    // p.q, then r.q and q.q in slots of their own: the r.r slot is being read
    // by the first block of each system during this kernel
    const double pq1[1] = {acc[0]};
    bt_partials<1>(pq1, c.red, kSlotPQ, gb);
    const double rq_qq[2] = {acc[1], acc[2]};
    bt_partials<2>(rq_qq, c.red, kSlotRQ, gb);
}

// x += alpha_q[lo] P_lo + ... + alpha_q[hi-1] P_{hi-1} over the block's rows
// (16-byte vectors v0..v1 of X): the pending steps of a period cut short.
template <class V, class X>
__device__ __forceinline__ void bt_add_pending(const BtArgs<V, X>& c, const SysScalars* S, int v0, int v1, int lo,
                                               int hi) {
    constexpr int N = Vec16<X>::N;
    if (lo >= hi) return;
    for (int v = v0 + (int)threadIdx.x; v < v1; v += (int)blockDim.x) {
        double xe[N];
        dload<N>(c.x, v, xe);
        for (int i = lo; i < hi; ++i) {
            const double ai = S->alpha_q[i];
            X pp[N];
            vload(c.pb + (int64_t)i * c.pstride, v, pp);
#pragma unroll
            for (int e = 0; e < N; ++e) xe[e] += ai * (double)pp[e];
        }
        dstore<N>(c.x, v, xe);
    }
}

// K2: alpha, beta per system; r -= alpha q; p' = r + beta p (into P_{j+1});
// partials of r.r.
// x period: the FP64 x takes the steps only at the period's last position
// (j = M - 1), before an FP64 residual replacement (which reads x), and when
// the system stops (a breakdown here, or a stop found by the next SpMV,
// after which this kernel adds the steps).  Every step is one FP64 addition
// in iteration order, so x is bitwise the x of an update every iteration;
// x moves 16 bytes per row once per period instead of every iteration.
// Position j's update records its step (alpha_q[j]) and whether x then holds
// every step up to j (added[j]); the pending steps at position j are those
// after the last position i < j with added[i].  A position writes only its
// own entries and reads only earlier ones, so no kernel reads what it writes.
template <class V, class X>
__global__ void __launch_bounds__(1024) k_bt_update(BtArgs<V, X> c) {
    constexpr int N = Vec16<X>::N;
    const int gb = c.map.block();
    const int k = c.map.bsys[gb];
    SysScalars* S = c.sys + k;
    const int j = c.j;
    // the block's rows are whole slices: 16-byte vectors v0..v1
    const int v0 = c.map.bs0[gb] * (c.slice_rows / N), v1 = c.map.bs1[gb] * (c.slice_rows / N);
    const int b0 = c.map.bptr[k], b1 = c.map.bptr[k + 1];
    // the SpMV's partial sums, then the first vector, in flight before the
    // system's flags are tested (vector loads complete in issue order, so the
    // re-add waits for the partials only)
    constexpr int kSl[3] = {kSlotPQ, kSlotRQ, kSlotQQ};
    SysSums<3> ps;
    ps.issue(c.red, kSl, b0, b1);
    const int va = v0 + (int)threadIdx.x;
    double xa[N];
    X pa[N], ra[N], qa[N];
    if (va < v1) {
        vload(c.p, va, pa);
        vload(c.r, va, ra);
        vload(c.q, va, qa);
    }
    int lo = 0;  // first pending position of the period
    for (int i = j - 1; i >= 0; --i)
        if (S->added[i]) {
            lo = i + 1;
            break;
        }
    // (done is set only by other kernels; the breakdown flags this kernel
    // may set stop the system at the next SpMV)
    if (S->done) {
        if (S->flush_now) bt_add_pending(c, S, v0, v1, lo, j);
        return;
    }
    if (S->replaced) {
        // no step at this position (x already holds every earlier step): the
        // direction moves on to P_{j+1}; the replaced residual's r.r partials
        // become the next rr, its norm the reference of the next replacement
        // test
        if (c.p_alt != c.p)
            for (int v = v0 + (int)threadIdx.x; v < v1; v += (int)blockDim.x) {
                X pe[N];
                vload(c.p, v, pe);
                vstore(c.p_alt, v, pe);
            }
        if (threadIdx.x == 0)
            c.red.partial[(size_t)kSlotRR * c.red.stride + gb] = c.red.partial[(size_t)kBtSlotTR * c.red.stride + gb];
        if (gb == b0) {
            const double tr = sys_sum(c.red, kBtSlotTR, b0, b1);
            if (threadIdx.x == 0) {
                S->r_ref = sqrt(tr);
                S->replacements++;
                S->replace_now = 0;
                S->need_update = 0;
                S->added[j] = 1;
                if (!isfinite(tr)) S->nonfinite = 1;
            }
        }
        return;
    }
    // x takes the steps at the period's end and before a replacement
    const bool add = j == c.M - 1 || S->need_update;
    if (add && va < v1) dload<N>(c.x, va, xa);
    double d3[3];
    ps.finish(c.red, kSl, b0, b1, d3);
    const double pq = d3[0], rq = d3[1], qq = d3[2];
    if (!(pq != 0.0 && isfinite(pq))) {
        // x keeps the period's earlier steps; the next SpMV stops the system
        bt_add_pending(c, S, v0, v1, lo, j);
        if (gb == b0 && threadIdx.x == 0) {
            S->breakdown = (pq == 0.0);
            S->nonfinite = !isfinite(pq);
        }
        return;
    }
    const double rr = S->rr;
    const double alpha = rr / pq;
    double rr_next = rr - 2.0 * alpha * rq + alpha * alpha * qq;
    if (!(rr_next > 0.0)) rr_next = 0.0;
    const double beta = rr_next / rr;
    // the pending steps' sizes, in registers (static indices)
    double al[kBtMaxPeriod];
#pragma unroll
    for (int i = 0; i < kBtMaxPeriod; ++i) al[i] = (add && i >= lo && i < j) ? S->alpha_q[i] : 0.0;
    double acc[1] = {0.0};
    auto body = [&](int v, double (&xe)[N], const X (&pe)[N], const X (&re)[N], const X (&qe)[N]) {
        X rn[N], pn[N];
#pragma unroll
        for (int e = 0; e < N; ++e) {
            rn[e] = (X)((double)re[e] - alpha * (double)qe[e]);
            pn[e] = (X)((double)rn[e] + beta * (double)pe[e]);
            acc[0] += (double)rn[e] * (double)rn[e];
        }
        if (add) {
#pragma unroll
            for (int i = 0; i < kBtMaxPeriod; ++i)
                if (i >= lo && i < j) {
                    X pp[N];
                    vload(c.pb + (int64_t)i * c.pstride, v, pp);
#pragma unroll
                    for (int e = 0; e < N; ++e) xe[e] += al[i] * (double)pp[e];
                }
#pragma unroll
            for (int e = 0; e < N; ++e) xe[e] += alpha * (double)pe[e];
            dstore<N>(c.x, v, xe);
        }
        vstore(c.r, v, rn);
        vstore(c.p_alt, v, pn);
    };
    if (va < v1) body(va, xa, pa, ra, qa);
    for (int v = va + (int)blockDim.x; v < v1; v += (int)blockDim.x) {
        double xe[N];
        X pe[N], re[N], qe[N];
        if (add) dload<N>(c.x, v, xe);
        vload(c.p, v, pe);
        vload(c.r, v, re);
        vload(c.q, v, qe);
        body(v, xe, pe, re, qe);
    }
    bt_partials<1>(acc, c.red, kSlotRR, gb);
    if (gb == b0 && threadIdx.x == 0) {
        S->alpha = alpha;
        S->beta = beta;
        ++S->iter;
        S->replace_now = S->need_update;
        S->alpha_q[j] = alpha;
        S->added[j] = add ? 1 : 0;
    }
}

}  // namespace ocsl

using namespace ocsl;

struct ocsl_batch_s {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr;
    int m = 0;
    std::vector<int64_t> n_sys;          // rows per system
    std::vector<int64_t> nnz_in;         // caller entries per system
    std::vector<int> row_off;            // padded row offset per system (m + 1)
    int n_tot = 0, ns_tot = 0, nblocks = 0;
    int rpl = 1;                         // rows per lane: slices of 64 rpl rows
    int block = 1024;                    // threads per SpMV/residual workgroup (one slice per wave)
    std::vector<int> bptr_host;
    std::vector<HostCSR> csr;            // the systems' structures (real entries)
    std::vector<std::vector<int>> order; // their reverse Cuthill-McKee orders
    std::vector<int> ns_sys;             // slices per system in the current layout
    // concatenated SELL-P (global numbering)
    DBuf<int> slice_ptr, src, diag_src, bsys, bs0, bs1, bptr, perm;  // perm: padded row -> (flat input index)
    DBuf<int> active;                    // compacted list of the blocks of unfinished systems
    DBuf<unsigned> colw;
    DBuf<int> col;
    bool compact = true;
    int64_t total = 0;
    DBuf<double> val_in, val64, diag64, b_in, b_glob, x_glob;
    DBuf<double> x_flat;                 // the solutions in the callers' order (staging of x0 and x)
    DBuf<float> val32, diag32;
    std::vector<int64_t> val_off;        // offset of system k's values in val_in
    int64_t n_flat = 0;                  // sum of the systems' rows
    bool have_vals = false, have_b = false;
    DBuf<double> d[4];
    DBuf<float> f[4];
    DBuf<float> pb32;                    // MIXED direction buffers P_0 .. P_{M-1}, n_tot apart
    DBuf<double> red_partial;
    DBuf<SysScalars> sys;
    SysScalars* sys_host = nullptr;
    // iteration chunks by (precision, chunk, launch size, compacted, x period, first position)
    std::map<std::tuple<int, int, int, int, int, int>, hipGraphExec_t> graphs;
    ~ocsl_batch_s() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        if (sys_host) (void)hipHostFree(sys_host);
        if (ev0) (void)hipEventDestroy(ev0);
        if (stream) (void)hipStreamDestroy(stream);
    }
    RedScratchBig red() const { return {red_partial.p, (size_t)nblocks}; }
    // threads of the update kernel: one 16-byte vector each over a block's rows
    int upd_block(bool mixed) const {
        const int vecs = (block / kWave) * kWave * rpl / (mixed ? 4 : 2);
        return std::min(kBtUpdateBlock, std::max(kWave, vecs));
    }
    std::vector<int64_t> flat_off() const {
        std::vector<int64_t> o(m + 1, 0);
        for (int k = 0; k < m; ++k) o[k + 1] = o[k] + n_sys[k];
        return o;
    }
    SellPView<double> view64() const {
        return {n_tot, ns_tot, slice_ptr.p, compact ? nullptr : col.p, compact ? colw.p : nullptr, val64.p, diag64.p,
                nullptr, nullptr};
    }
    SellPView<float> view32() const {
        return {n_tot, ns_tot, slice_ptr.p, compact ? nullptr : col.p, compact ? colw.p : nullptr, val32.p, diag32.p,
                nullptr, nullptr};
    }
    BatchMap map(bool compact = false) const { return {bsys.p, bs0.p, bs1.p, bptr.p, compact ? active.p : nullptr}; }
};

namespace {

template <class F>
ocsl_status guarded_b(F&& f) {
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

// out[i] = in[perm[i]] for padded rows (perm -1: 0)
__global__ void k_bt_gather(int n, const int* __restrict__ perm, const double* __restrict__ in, double* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int p = perm[i];
        out[i] = p >= 0 ? in[p] : 0.0;
    }
}

__global__ void k_bt_scatter(int n, const int* __restrict__ perm, const double* __restrict__ in, double* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int p = perm[i];
        if (p >= 0) out[p] = in[i];
    }
}

void bt_drop_graphs(ocsl_batch bt) {
    for (auto& kv : bt->graphs) (void)hipGraphExecDestroy(kv.second);
    bt->graphs.clear();
}

// Concatenated SELL-P of the batch with rpl rows per lane.  Row r of system
// k's reverse Cuthill-McKee order is row row_off[k] + r; in slice sl lane l
// owns rows sl*64R + l*R + j, and off-diagonal kk of such a row sits at
// base + (kk/2)*128R + 2R l + 2j + (kk&1) -- the single-system SELL-P
// layout, so the same row routines serve.  Padding rows of a system are
// zero rows whose columns are themselves.
void bt_build_layout(ocsl_batch bt, int rpl) {
    OCSL_REQUIRE(rpl == 1 || rpl == 2 || rpl == 4, OCSL_ERR_ARG, "rows_per_lane must be 1, 2 or 4");
    const int m = bt->m, H = kWave * rpl;
    std::vector<int> sp_all{0}, colg, srcg, dsrc, perm, row_off{0}, ns_sys(m);
    bool compact = true;
    int64_t flat_base = 0;
    for (int k = 0; k < m; ++k) {
        const HostCSR& a = bt->csr[k];
        const std::vector<int>& order = bt->order[k];
        const int n = a.n, ns = (n + H - 1) / H;
        OCSL_REQUIRE((int64_t)row_off.back() + (int64_t)ns * H < ((int64_t)1 << 30), OCSL_ERR_ARG, "batch too large");
        std::vector<int> iperm(n);
        for (int i = 0; i < n; ++i) iperm[order[i]] = i;
        const int r0 = row_off.back();
        const int64_t vb = bt->val_off[k];
        for (int sl = 0; sl < ns; ++sl) {
            int w = 0;
            for (int q = 0; q < H && sl * H + q < n; ++q) {
                const int o = order[sl * H + q];
                w = std::max(w, a.rp[o + 1] - a.rp[o] - 1);
            }
            w += w & 1;
            const int base = sp_all.back();
            OCSL_REQUIRE((int64_t)base + (int64_t)w * H < ((int64_t)1 << 31), OCSL_ERR_ARG, "batch too large");
            sp_all.push_back(base + w * H);
            colg.resize(base + (size_t)w * H);
            srcg.resize(base + (size_t)w * H);
            for (int l = 0; l < kWave; ++l)
                for (int j = 0; j < rpl; ++j) {
                    const int r = sl * H + l * rpl + j;  // local row (new order)
                    auto slot = [&](int kk) {
                        return (int64_t)base + (int64_t)(kk / 2) * 2 * H + 2 * rpl * l + 2 * j + (kk & 1);
                    };
                    int kk = 0;
                    if (r < n) {
                        const int o = order[r];
                        for (int p = a.rp[o]; p < a.rp[o + 1]; ++p) {
                            if (p == a.diag[o]) continue;
                            const int64_t q = slot(kk++);
                            colg[q] = r0 + iperm[a.col[p]];
                            srcg[q] = (int)(vb + a.src[p]);
                            const int d = colg[q] - (r0 + r);
                            if (d < -32768 || d > 32767) compact = false;
                        }
                    }
                    for (; kk < w; ++kk) {
                        const int64_t q = slot(kk);
                        colg[q] = r0 + r;
                        srcg[q] = -1;
                    }
                }
        }
        for (int r = 0; r < ns * H; ++r) {
            perm.push_back(r < n ? (int)(flat_base + order[r]) : -1);
            dsrc.push_back(r < n ? (int)(vb + a.src[a.diag[order[r]]]) : -1);
        }
        ns_sys[k] = ns;
        row_off.push_back(r0 + ns * H);
        flat_base += n;
    }
    bt_drop_graphs(bt);
    hipStream_t s = bt->stream;
    OCSL_HIP(hipStreamSynchronize(s));  // earlier uploads may still read host buffers of an earlier build
    bt->rpl = rpl;
    bt->row_off = row_off;
    bt->ns_sys = ns_sys;
    bt->n_tot = row_off.back();
    bt->ns_tot = (int)(sp_all.size() - 1);
    bt->total = sp_all.back();
    bt->compact = compact;
    bt->slice_ptr.upload(sp_all.data(), sp_all.size(), s);
    if (compact) {
        // word of pair t of row j of lane l: base / 2 + t H + l R + j
        std::vector<unsigned> colw(bt->total / 2);
        for (int sl = 0; sl < bt->ns_tot; ++sl) {
            const int np = (sp_all[sl + 1] - sp_all[sl]) / (2 * H);
            for (int t = 0; t < np; ++t)
                for (int l = 0; l < kWave; ++l)
                    for (int j = 0; j < rpl; ++j) {
                        const int64_t q = sp_all[sl] + (int64_t)t * 2 * H + 2 * rpl * l + 2 * j;
                        const int row = sl * H + l * rpl + j;
                        const unsigned lo = (unsigned)(uint16_t)(int16_t)(colg[q] - row);
                        const unsigned hi = (unsigned)(uint16_t)(int16_t)(colg[q + 1] - row);
                        colw[sp_all[sl] / 2 + (int64_t)t * H + l * rpl + j] = lo | (hi << 16);
                    }
        }
        bt->colw.upload(colw.data(), colw.size(), s);
        OCSL_HIP(hipStreamSynchronize(s));
    } else {
        bt->col.upload(colg.data(), colg.size(), s);
    }
    bt->src.upload(srcg.data(), srcg.size(), s);
    bt->diag_src.upload(dsrc.data(), dsrc.size(), s);
    bt->perm.upload(perm.data(), perm.size(), s);
    OCSL_HIP(hipStreamSynchronize(s));
}

// Workgroup -> (system, slice range) map for SpMV workgroups of `block`
// threads, one slice per wave; the update kernel covers the same rows.
void bt_build_map(ocsl_batch bt, int block) {
    OCSL_REQUIRE(block == 256 || block == 512 || block == 1024, OCSL_ERR_ARG, "block_size must be 256, 512 or 1024");
    const int spb = block / kWave;
    std::vector<int> bsys, bs0, bs1, bptr{0};
    int sb = 0;
    for (int k = 0; k < bt->m; ++k) {
        const int ns = bt->ns_sys[k];
        for (int s0 = 0; s0 < ns; s0 += spb) {
            bsys.push_back(k);
            bs0.push_back(sb + s0);
            bs1.push_back(sb + std::min(ns, s0 + spb));
        }
        bptr.push_back((int)bsys.size());
        sb += ns;
    }
    bt_drop_graphs(bt);
    hipStream_t s = bt->stream;
    OCSL_HIP(hipStreamSynchronize(s));
    bt->block = block;
    bt->nblocks = (int)bsys.size();
    bt->bptr_host = bptr;
    bt->bsys.upload(bsys.data(), bsys.size(), s);
    bt->bs0.upload(bs0.data(), bs0.size(), s);
    bt->bs1.upload(bs1.data(), bs1.size(), s);
    bt->bptr.upload(bptr.data(), bptr.size(), s);
    bt->red_partial.alloc((size_t)bt->nblocks * 8);
    bt->active.alloc(bt->nblocks);
    OCSL_HIP(hipStreamSynchronize(s));
}

// Matrix values of the current layout from the caller's values (val_in):
// FP64 entries and diagonal, FP32 entries and the FP64 row sums as the FP32
// diagonal (difference form).
void bt_gather_values(ocsl_batch bt) {
    hipStream_t s = bt->stream;
    const void* before[4] = {bt->val64.p, bt->diag64.p, bt->val32.p, bt->diag32.p};
    bt->val64.alloc(bt->total);
    bt->diag64.alloc(bt->n_tot);
    bt->val32.alloc(bt->total);
    bt->diag32.alloc(bt->n_tot);
    const void* after[4] = {bt->val64.p, bt->diag64.p, bt->val32.p, bt->diag32.p};
    if (!std::equal(before, before + 4, after)) bt_drop_graphs(bt);  // graphs hold the arrays' addresses
    k_gather_vals<double><<<grid_for(bt->total), kBlock, 0, s>>>(bt->total, bt->src.p, bt->val_in.p, bt->val64.p);
    k_gather_vals<double><<<grid_for(bt->n_tot), kBlock, 0, s>>>(bt->n_tot, bt->diag_src.p, bt->val_in.p,
                                                                bt->diag64.p);
    k_gather_vals<float><<<grid_for(bt->total), kBlock, 0, s>>>(bt->total, bt->src.p, bt->val_in.p, bt->val32.p);
    with_rpl(bt->rpl, [&](auto RC) {
        k_sellp_rowsum<decltype(RC)::value>
            <<<grid_for((int64_t)bt->ns_tot * kWave), kBlock, 0, s>>>(bt->view64(), bt->diag32.p);
    });
    OCSL_HIP(hipGetLastError());
}

// Layout with rpl rows per lane and workgroups of `block` threads; values
// and right-hand side, when set, follow into the new layout.
void bt_set_launch(ocsl_batch bt, int rpl, int block) {
    if (rpl != bt->rpl) {
        bt_build_layout(bt, rpl);
        if (bt->have_vals) bt_gather_values(bt);
        if (bt->have_b) {
            bt->b_glob.alloc(bt->n_tot);
            k_bt_gather<<<grid_for(bt->n_tot), kBlock, 0, bt->stream>>>(bt->n_tot, bt->perm.p, bt->b_in.p,
                                                                       bt->b_glob.p);
            OCSL_HIP(hipGetLastError());
        }
        bt_build_map(bt, block);
    } else if (block != bt->block) {
        bt_build_map(bt, block);
    }
    OCSL_HIP(hipStreamSynchronize(bt->stream));
}

// Iterations per FP64 update of x: the MIXED option's 1, 2, 4 or 8 (other
// values: 8); FP64 updates x every iteration.
int bt_period(bool mixed, const ocsl_options& o) {
    if (!mixed) return 1;
    return o.x_period == 1 || o.x_period == 2 || o.x_period == 4 ? o.x_period : kBtMaxPeriod;
}

// Krylov vectors of the precision (zeroed: padding rows stay zero; MIXED:
// M direction buffers) and fresh per-system scalars for a solve to tol
// within max_iter iterations.
void bt_prepare(ocsl_batch bt, bool mixed, double tol, int max_iter, double replace_ratio, int M) {
    hipStream_t s = bt->stream;
    const int N = bt->n_tot, m = bt->m;
    const void* before[4] = {bt->f[0].p, bt->f[2].p, bt->pb32.p, bt->d[0].p};
    for (int i = 0; i < 3; ++i) {
        if (mixed) {
            if (i == 1) continue;  // the directions live in pb32
            bt->f[i].alloc(N);
            OCSL_HIP(hipMemsetAsync(bt->f[i].p, 0, sizeof(float) * N, s));
        } else {
            bt->d[i].alloc(N);
            OCSL_HIP(hipMemsetAsync(bt->d[i].p, 0, sizeof(double) * N, s));
        }
    }
    if (mixed) {
        bt->pb32.alloc((size_t)M * N);
        OCSL_HIP(hipMemsetAsync(bt->pb32.p, 0, sizeof(float) * (size_t)M * N, s));
    }
    const void* after[4] = {bt->f[0].p, bt->f[2].p, bt->pb32.p, bt->d[0].p};
    if (!std::equal(before, before + 4, after)) bt_drop_graphs(bt);  // graphs hold the buffers' addresses
    std::vector<SysScalars> init(m);
    for (auto& S : init) {
        S = SysScalars{};
        S.tol2 = tol * tol;
        S.max_iter = max_iter;
        S.replace_ratio = mixed ? replace_ratio : 0.0;
    }
    OCSL_HIP(hipMemcpyAsync(bt->sys.p, init.data(), sizeof(SysScalars) * m, hipMemcpyHostToDevice, s));
    OCSL_HIP(hipStreamSynchronize(s));
}

// The batch's kernel launches in one precision over the current layout.
// The batch's kernel launches in one precision over the current layout,
// with x updated once per M iterations (bt_prepare allocates the M direction
// buffers first).
struct BtLaunch {
    ocsl_batch bt;
    bool mixed;
    int UB;
    int M;
    BtArgs<float, float> c32;
    BtArgs<double, double> c64;
    BtLaunch(ocsl_batch b, bool mx, int period)
        : bt(b), mixed(mx), UB(b->upd_block(mx)), M(period),
          c32{b->view32(), b->map(), b->f[0].p, b->pb32.p, b->f[2].p, b->x_glob.p, b->sys.p, b->red(), kWave * b->rpl,
              b->view64(), b->b_glob.p, b->pb32.p, b->n_tot, period, 0, b->pb32.p},
          c64{b->view64(), b->map(), b->d[0].p, b->d[1].p, b->d[2].p, b->x_glob.p, b->sys.p, b->red(),
              kWave * b->rpl, b->view64(), b->b_glob.p, b->d[1].p, b->n_tot, 1, 0, b->d[1].p} {
        OCSL_REQUIRE(mixed ? (M >= 1 && M <= kBtMaxPeriod && b->pb32.n >= (size_t)M * b->n_tot) : M == 1,
                     OCSL_ERR_STATE, "batch direction buffers do not match the x period");
    }
    void set_map(const BatchMap& mp) { c32.map = c64.map = mp; }
    // r = b - A x (and p = r when first) for the systems of the map's blocks;
    // a system that goes on iterating resumes at period position phi
    // (phi < 0: none does)
    void residual(int first, const BatchMap& mp, int grid, int phi) {
        hipStream_t s = bt->stream;
        with_rpl(bt->rpl, [&](auto RC) {
            constexpr int R = decltype(RC)::value;
            if (mixed)
                k_bt_residual<float, R><<<grid, bt->block, 0, s>>>(bt->view64(), mp, bt->b_glob.p, bt->x_glob.p,
                                                                   c32.r, c32.pb, c32.pstride, phi, bt->sys.p,
                                                                   bt->red(), first);
            else
                k_bt_residual<double, R><<<grid, bt->block, 0, s>>>(bt->view64(), mp, bt->b_glob.p, bt->x_glob.p,
                                                                    c64.r, c64.pb, c64.pstride, phi, bt->sys.p,
                                                                    bt->red(), first);
        });
        k_bt_fin_residual<<<bt->m, 256, 0, s>>>(bt->map(), bt->sys.p, bt->red(), first, bt->m, phi);
        OCSL_HIP(hipGetLastError());
    }
    // one CG iteration, at period position j, of every unfinished system of
    // the map's blocks
    void iter(int grid, int j) {
        hipStream_t s = bt->stream;
        auto at = [&](auto c) {
            c.j = j;
            c.p = c.pb + (int64_t)j * c.pstride;
            c.p_alt = c.pb + (int64_t)((j + 1) % c.M) * c.pstride;
            return c;
        };
        OCSL_REQUIRE(j >= 0 && j < M, OCSL_ERR_ARG, "period position out of range");
        with_rpl(bt->rpl, [&](auto RC) {
            constexpr int R = decltype(RC)::value;
            if (mixed) {
                const BtArgs<float, float> c = at(c32);
                k_bt_spmv<float, float, R><<<grid, bt->block, 0, s>>>(c);
                k_bt_update<float, float><<<grid, UB, 0, s>>>(c);
            } else {
                const BtArgs<double, double> c = at(c64);
                k_bt_spmv<double, double, R><<<grid, bt->block, 0, s>>>(c);
                k_bt_update<double, double><<<grid, UB, 0, s>>>(c);
            }
        });
    }
};

// Average time of one iteration of the whole batch: every system restarted
// from x = 0 with tolerance 0 (none finishes), `iters` iterations captured
// as one graph, the second of two replays timed with events.
double bt_time_iterations(ocsl_batch bt, bool mixed, int iters) {
    hipStream_t s = bt->stream;
    bt->x_glob.alloc(bt->n_tot);
    hipGraphExec_t ex = nullptr;
    hipEvent_t e1 = nullptr;
    float ms = 0.0f;
    try {
        OCSL_HIP(hipEventCreate(&e1));
        for (int rep = 0; rep < 2; ++rep) {
            OCSL_HIP(hipMemsetAsync(bt->x_glob.p, 0, sizeof(double) * bt->n_tot, s));
            const int M = mixed ? kBtMaxPeriod : 1;  // the solve's default period
            bt_prepare(bt, mixed, 0.0, 1 << 30, 0.0, M);
            BtLaunch L(bt, mixed, M);
            L.residual(1, bt->map(), bt->nblocks, 0);
            if (!ex) {
                hipGraph_t gr;
                OCSL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                for (int i = 0; i < iters; ++i) L.iter(bt->nblocks, i % M);
                OCSL_HIP(hipStreamEndCapture(s, &gr));
                const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                OCSL_HIP(e);
            }
            OCSL_HIP(hipEventRecord(bt->ev0, s));
            OCSL_HIP(hipGraphLaunch(ex, s));
            OCSL_HIP(hipEventRecord(e1, s));
            OCSL_HIP(hipStreamSynchronize(s));
            OCSL_HIP(hipEventElapsedTime(&ms, bt->ev0, e1));
        }
    } catch (...) {
        if (ex) (void)hipGraphExecDestroy(ex);
        if (e1) (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipGraphExecDestroy(ex);
    (void)hipEventDestroy(e1);
    return 1e3 * (double)ms / iters;
}

}  // namespace

extern "C" {

ocsl_status ocsl_batch_create(ocsl_batch* out, int m, const int64_t* n_rows, const uint64_t* const* row_ptr,
                              const uint64_t* const* column_index) {
    if (!out || m <= 0 || !n_rows || !row_ptr || !column_index) return OCSL_ERR_ARG;
    *out = nullptr;
    if (ocsl_device_count() <= 0) {
        set_last_error("no HIP device");
        return OCSL_ERR_NODEVICE;
    }
    ocsl_batch_s* bt = new (std::nothrow) ocsl_batch_s();
    if (!bt) return OCSL_ERR_NOMEM;
    ocsl_status st = guarded_b([&] {
        OCSL_HIP(hipGetDevice(&bt->device));
        OCSL_HIP(hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking));
        OCSL_HIP(hipEventCreate(&bt->ev0));
        bt->m = m;
        bt->n_sys.assign(n_rows, n_rows + m);
        bt->csr.resize(m);
        bt->order.resize(m);
        bt->val_off.assign(m + 1, 0);
        int64_t flat = 0;
        for (int k = 0; k < m; ++k) {
            HostCSR& a = bt->csr[k];
            analyse_structure(n_rows[k], row_ptr[k], column_index[k], a);
            bt->order[k] = rcm_order(a);
            bt->nnz_in.push_back(a.nnz_in);
            bt->val_off[k + 1] = bt->val_off[k] + a.nnz_in;
            flat += a.n;
        }
        OCSL_REQUIRE(bt->val_off[m] < ((int64_t)1 << 31) && flat < ((int64_t)1 << 30), OCSL_ERR_ARG, "batch too large");
        bt->n_flat = flat;
        bt->rpl = 0;  // forces the first build
        bt_set_launch(bt, 1, 1024);
        bt->sys.alloc(m);
        OCSL_HIP(hipHostMalloc((void**)&bt->sys_host, sizeof(SysScalars) * m, hipHostMallocDefault));
    });
    if (st != OCSL_OK) {
        delete bt;
        return st;
    }
    *out = bt;
    return OCSL_OK;
}

ocsl_status ocsl_batch_destroy(ocsl_batch bt) {
    if (!bt) return OCSL_ERR_ARG;
    if (bt->stream) (void)hipStreamSynchronize(bt->stream);
    delete bt;
    return OCSL_OK;
}

ocsl_status ocsl_batch_size(ocsl_batch bt, int* m, int64_t* rows_total) {
    if (!bt || !m || !rows_total) return OCSL_ERR_ARG;
    *m = bt->m;
    *rows_total = bt->n_flat;
    return OCSL_OK;
}

// values[k]: the value array of system k in its caller layout.
ocsl_status ocsl_batch_set_values(ocsl_batch bt, const double* const* values, ocsl_where where) {
    if (!bt || !values) return OCSL_ERR_ARG;
    return guarded_b([&] {
        hipStream_t s = bt->stream;
        const int64_t tot = bt->val_off[bt->m];
        bt->val_in.alloc(tot);
        for (int k = 0; k < bt->m; ++k) {
            OCSL_REQUIRE(values[k], OCSL_ERR_ARG, "null value array");
            OCSL_HIP(hipMemcpyAsync(bt->val_in.p + bt->val_off[k], values[k], sizeof(double) * bt->nnz_in[k],
                                    where == OCSL_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s));
        }
        bt_gather_values(bt);
        OCSL_HIP(hipStreamSynchronize(s));
        bt->have_vals = true;
    });
}

static void bt_flat_in(ocsl_batch bt, const double* const* v, ocsl_where where, DBuf<double>& flat_tmp,
                       double* out_padded) {
    hipStream_t s = bt->stream;
    const std::vector<int64_t> fo = bt->flat_off();
    flat_tmp.alloc(bt->n_flat);
    for (int k = 0; k < bt->m; ++k) {
        OCSL_REQUIRE(v[k], OCSL_ERR_ARG, "null vector");
        OCSL_HIP(hipMemcpyAsync(flat_tmp.p + fo[k], v[k], sizeof(double) * bt->n_sys[k],
                                where == OCSL_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s));
    }
    k_bt_gather<<<grid_for(bt->n_tot), kBlock, 0, s>>>(bt->n_tot, bt->perm.p, flat_tmp.p, out_padded);
    OCSL_HIP(hipGetLastError());
}

ocsl_status ocsl_batch_set_rhs(ocsl_batch bt, const double* const* b, ocsl_where where) {
    if (!bt || !b) return OCSL_ERR_ARG;
    return guarded_b([&] {
        const double* before = bt->b_glob.p;
        bt->b_glob.alloc(bt->n_tot);
        if (before != bt->b_glob.p) bt_drop_graphs(bt);  // graphs hold b's address
        bt_flat_in(bt, b, where, bt->b_in, bt->b_glob.p);
        OCSL_HIP(hipStreamSynchronize(bt->stream));
        bt->have_b = true;
    });
}

ocsl_status ocsl_batch_set_launch(ocsl_batch bt, int rows_per_lane, int block_size) {
    if (!bt) return OCSL_ERR_ARG;
    return guarded_b([&] { bt_set_launch(bt, rows_per_lane, block_size); });
}

ocsl_status ocsl_batch_get_launch(ocsl_batch bt, int* rows_per_lane, int* block_size) {
    if (!bt || !rows_per_lane || !block_size) return OCSL_ERR_ARG;
    *rows_per_lane = bt->rpl;
    *block_size = bt->block;
    return OCSL_OK;
}

ocsl_status ocsl_batch_time_iterations(ocsl_batch bt, int precision, int iterations, double* us_per_iter) {
    if (!bt || !us_per_iter || iterations <= 0) return OCSL_ERR_ARG;
    return guarded_b([&] {
        OCSL_REQUIRE(precision == OCSL_FP64 || precision == OCSL_MIXED, OCSL_ERR_ARG, "bad precision");
        OCSL_REQUIRE(bt->have_vals && bt->have_b, OCSL_ERR_STATE, "values or rhs not set");
        *us_per_iter = bt_time_iterations(bt, precision == OCSL_MIXED, iterations);
    });
}

ocsl_status ocsl_batch_tune(ocsl_batch bt, int precision, int iterations, int* best_rpl, int* best_block,
                            double* best_us, ocsl_tune_entry* table, int table_cap, int* table_len) {
    if (!bt || iterations <= 0 || (table && table_cap < 0)) return OCSL_ERR_ARG;
    if (table_len) *table_len = 0;
    return guarded_b([&] {
        OCSL_REQUIRE(precision == OCSL_FP64 || precision == OCSL_MIXED, OCSL_ERR_ARG, "bad precision");
        OCSL_REQUIRE(bt->have_vals && bt->have_b, OCSL_ERR_STATE, "values or rhs not set");
        const bool mixed = precision == OCSL_MIXED;
        int br = bt->rpl, bb = bt->block, len = 0;
        double bu = 0.0;
        for (int R : {1, 2, 4})
            for (int B : {256, 512, 1024}) {
                bt_set_launch(bt, R, B);
                const double us = bt_time_iterations(bt, mixed, iterations);
                if (table && len < table_cap) {
                    ocsl_tune_entry& e = table[len];
                    e.rows_per_lane = R;
                    e.block_size = B;
                    e.grid_blocks = bt->nblocks;
                    e.vec_blocks = bt->upd_block(mixed);
                    e.cg_kernels = 2;
                    e.us_per_iteration = us;
                }
                ++len;
                if (bu == 0.0 || us < bu) {
                    bu = us;
                    br = R;
                    bb = B;
                }
            }
        bt_set_launch(bt, br, bb);
        if (best_rpl) *best_rpl = br;
        if (best_block) *best_block = bb;
        if (best_us) *best_us = bu;
        if (table_len) *table_len = table ? std::min(len, table_cap) : len;
    });
}

ocsl_status ocsl_batch_solve(ocsl_batch bt, int precision, const ocsl_options* opt, double* const* x,
                             ocsl_where where, ocsl_report* reps, double* latency_ms) {
    if (!bt || !x) return OCSL_ERR_ARG;
    ocsl_options o;
    ocsl_default_options(&o);
    if (opt) o = *opt;
    int worst = 0;
    ocsl_status st = guarded_b([&] {
        OCSL_REQUIRE(precision == OCSL_FP64 || precision == OCSL_MIXED, OCSL_ERR_ARG, "bad precision");
        OCSL_REQUIRE(bt->have_vals && bt->have_b, OCSL_ERR_STATE, "values or rhs not set");
        OCSL_REQUIRE(o.max_iter > 0 && o.check_every > 0 && o.tol >= 0.0, OCSL_ERR_ARG, "bad options");
        const bool mixed = precision == OCSL_MIXED;
        hipStream_t s = bt->stream;
        const int m = bt->m, N = bt->n_tot, G = bt->nblocks;
        DBuf<double>& flat = bt->x_flat;
        const double* x_before = bt->x_glob.p;
        bt->x_glob.alloc(N);
        if (x_before != bt->x_glob.p) bt_drop_graphs(bt);  // graphs hold x's address
        bt_flat_in(bt, x, where, flat, bt->x_glob.p);
        const int M = bt_period(mixed, o);
        bt_prepare(bt, mixed, o.tol, o.max_iter, o.replace_ratio, M);
        // launched blocks: all, or (compacted) the blocks of unfinished systems
        int G_cur = G;
        bool comp = false;
        BtLaunch L(bt, mixed, M);
        // iterations launched so far: the next one sits at period position
        // launched mod M
        int64_t launched = 0;
        auto residual = [&](int first, bool all, int phi) {
            const bool c = comp && !all;
            L.residual(first, bt->map(c), c ? G_cur : G, phi);
        };
        auto read_sys = [&] {
            OCSL_HIP(hipMemcpyAsync(bt->sys_host, bt->sys.p, sizeof(SysScalars) * m, hipMemcpyDeviceToHost, s));
            OCSL_HIP(hipStreamSynchronize(s));
        };
        std::vector<double> lat(m, -1.0);
        auto t0 = std::chrono::steady_clock::now();
        auto finished = [&](const SysScalars& S) {
            if (S.nonfinite || S.breakdown) return true;
            return mixed ? (S.done && S.verified) : (bool)S.done;
        };
        auto stamp = [&] {
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            for (int k = 0; k < m; ++k)
                if (lat[k] < 0.0 && finished(bt->sys_host[k])) lat[k] = ms;
        };
        // Compaction: once the blocks of unfinished systems are at most 3/4
        // of those launched, launch only them (a workgroup reads its block
        // from the active list; partial sums stay indexed by block).
        std::vector<int> act;
        auto compact = [&] {
            act.clear();
            for (int k = 0; k < m; ++k)
                if (!finished(bt->sys_host[k]))
                    for (int q = bt->bptr_host[k]; q < bt->bptr_host[k + 1]; ++q) act.push_back(q);
            if (act.empty() || 4 * (int64_t)act.size() > 3 * (int64_t)G_cur) return;
            OCSL_HIP(hipMemcpyAsync(bt->active.p, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice, s));
            OCSL_HIP(hipStreamSynchronize(s));
            G_cur = (int)act.size();
            comp = true;
            L.set_map(bt->map(true));
        };
        OCSL_HIP(hipEventRecord(bt->ev0, s));
        residual(1, true, 0);
        read_sys();
        stamp();
        const int chunk = o.check_every;
        if (bt->graphs.size() > 64) bt_drop_graphs(bt);
        while (true) {
            bool all = true;
            for (int k = 0; k < m; ++k) all = all && finished(bt->sys_host[k]);
            // an iteration is a CG step or (MIXED) a residual replacement,
            // never two replacements in a row: 2 max_iter launches bound it
            if (all || launched >= 2 * (int64_t)o.max_iter) break;
            compact();
            // graphs hold the launch size and map; the active list is read
            // through its (fixed) device address when the graph runs
            const int j0 = (int)(launched % M);
            const auto key = std::make_tuple(mixed ? 1 : 0, chunk, G_cur, comp ? 1 : 0, M, j0);
            auto chunk_iters = [&] {
                for (int i = 0; i < chunk; ++i) L.iter(G_cur, (int)((j0 + i) % M));
            };
            if (o.use_graph) {
                auto it = bt->graphs.find(key);
                if (it == bt->graphs.end()) {
                    hipGraph_t gr;
                    OCSL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    chunk_iters();
                    OCSL_HIP(hipStreamEndCapture(s, &gr));
                    hipGraphExec_t ex;
                    OCSL_HIP(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
                    OCSL_HIP(hipGraphDestroy(gr));
                    it = bt->graphs.emplace(key, ex).first;
                }
                OCSL_HIP(hipGraphLaunch(it->second, s));
            } else {
                chunk_iters();
                OCSL_HIP(hipGetLastError());
            }
            launched += chunk;
            read_sys();
            if (mixed) {
                bool any = false;
                for (int k = 0; k < m; ++k) any = any || bt_wants_residual(bt->sys_host[k]);
                if (any) {
                    residual(0, false, (int)(launched % M));
                    read_sys();
                }
            }
            stamp();
        }
        OCSL_HIP(hipStreamSynchronize(s));
        const double total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        // solutions back to the caller's arrays
        k_bt_scatter<<<grid_for(N), kBlock, 0, s>>>(N, bt->perm.p, bt->x_glob.p, flat.p);
        OCSL_HIP(hipGetLastError());
        const std::vector<int64_t> fo = bt->flat_off();
        for (int k = 0; k < m; ++k)
            OCSL_HIP(hipMemcpyAsync(x[k], flat.p + fo[k], sizeof(double) * bt->n_sys[k],
                                    where == OCSL_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s));
        // true residuals: one more residual pass for every system
        std::vector<SysScalars> before(bt->sys_host, bt->sys_host + m);
        for (int k = 0; k < m; ++k) {
            bt->sys_host[k].done = 1;
            bt->sys_host[k].verified = 0;
            bt->sys_host[k].nonfinite = 0;
            bt->sys_host[k].breakdown = 0;
        }
        OCSL_HIP(hipMemcpyAsync(bt->sys.p, bt->sys_host, sizeof(SysScalars) * m, hipMemcpyHostToDevice, s));
        residual(0, true, -1);
        read_sys();
        for (int k = 0; k < m; ++k) {
            const SysScalars& S = bt->sys_host[k];
            const double scale = S.bb > 0.0 ? S.bb : 1.0;
            const double tr = std::sqrt(S.rr / scale);
            int stk = OCSL_OK;
            if (before[k].breakdown) stk |= OCSL_WARN_BREAKDOWN;
            if (before[k].nonfinite || !std::isfinite(tr)) stk |= OCSL_WARN_NONFINITE;
            if (!(tr <= o.tol * (mixed ? 1.0 : 1.5)) && stk == OCSL_OK) stk |= OCSL_WARN_MAXITER;
            worst |= stk;
            if (reps) {
                ocsl_report& r = reps[k];
                std::memset(&r, 0, sizeof(r));
                r.status = stk;
                r.iterations = before[k].iter;
                r.replacements = before[k].replacements;
                r.residual = tr;
                r.true_relres = tr;
                r.solve_ms = total_ms;
                // algorithmic bytes, as the single-system split CG counts
                // them: SpMV (off-diagonal values and columns, diagonal,
                // slice pointers, p, r read, q written), update (p, r, q
                // read, r, p written, and x read and written once per period
                // of M iterations with the period's M - 1 earlier
                // directions); a replacement reads the FP64 matrix, x, b and
                // writes r
                const double n = (double)bt->n_sys[k], sv = mixed ? 4.0 : 8.0, cb = bt->compact ? 2.0 : 4.0;
                const double off = (double)bt->csr[k].col.size() - n, sl = 4.0 * (bt->ns_sys[k] + 1);
                const double xb = (16.0 + (M - 1) * sv) / M;
                const double it = off * (sv + cb) + sl + n * (9.0 * sv + xb);
                const double res = off * (8.0 + cb) + sl + n * (8.0 + 8.0 + 8.0 + sv);
                r.bytes = (double)r.iterations * it + (double)r.replacements * res;
            }
            if (latency_ms) latency_ms[k] = lat[k] >= 0.0 ? lat[k] : total_ms;
        }
        OCSL_HIP(hipStreamSynchronize(s));
    });
    if (st != OCSL_OK) return st;
    return (ocsl_status)worst;
}

}  // extern "C"
