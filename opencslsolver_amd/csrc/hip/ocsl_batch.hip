// Batched solves of many independent CSR systems in the same launches.
//
// The systems' SELL-P structures (reverse Cuthill-McKee order, 16-bit column
// offsets) are concatenated, each padded to whole 64-row slices so no slice
// and no workgroup straddles two systems.  One block map -- workgroup ->
// (system, slice range) -- serves every kernel, so the partial sums of system
// k always sit in the contiguous block range bptr[k]..bptr[k+1] and a
// consuming block re-adds only its own system's partials.  Every system has
// its own scalars (rr, alpha, beta, flags): the solves are independent, they
// merely share the launches.  An iteration is two kernels (the merged-
// reduction CG of the single-system mixed solver); blocks of converged
// systems return at once, and chunks of iterations replay as one hipGraph.
// FP64 residual replacement (MIXED) is done per system when its FP32
// recurrence has dropped by replace_ratio or claims convergence.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>

#include "ocsl_handle.h"
#include "ocsl_kernels.h"

namespace ocsl {

// Threads per workgroup of the update kernel: one 16-byte FP32 vector per
// thread over a block's 16 slices of 64 rows.
constexpr int kBtUpdateBlock = 256;

struct SysScalars {
    double rr, bb, r_ref, tol2, alpha, beta, replace_ratio;
    int iter, max_iter, done, halt, need_update, breakdown, nonfinite, replacements;
    int verified;  // convergence confirmed on the true FP64 residual (MIXED)
    int pad;
};

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
    X* p;
    X* q;
    double* x;
    SysScalars* sys;
    RedScratchBig red;
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

// A system takes the residual pass when its recurrence asks for a
// replacement, or claims convergence not yet confirmed on the true residual.
__device__ __forceinline__ bool bt_wants_residual(const SysScalars& S) {
    return S.need_update || (S.done && !S.verified && !S.nonfinite && !S.breakdown);
}

// r = b - A x (FP64) for the systems being (re)started; p = r on the first
// call; partials of r.r and b.b.
template <class X>
__global__ void __launch_bounds__(1024) k_bt_residual(SellPView<double> A, BatchMap map, const double* __restrict__ b,
                                                    const double* __restrict__ x, X* r, X* p, SysScalars* sys,
                                                    RedScratchBig red, int first) {
    const int gb = map.block();
    const int k = map.bsys[gb];
    if (!first && !bt_wants_residual(sys[k])) return;
    const int lane = threadIdx.x & (kWave - 1);
    double acc[2] = {0.0, 0.0};
    for (int s = map.bs0[gb] + (int)(threadIdx.x / kWave); s < map.bs1[gb];
         s += (int)(blockDim.x / kWave)) {
        const int row = s * kWave + lane;
        const double ax = sellp_row<double, double, false>(A, s, lane, row, x, row < A.n ? x[row] : 0.0);
        if (row < A.n) {
            const double bv = b[row];
            const double rv = bv - ax;
            acc[0] += rv * rv;
            acc[1] += bv * bv;
            r[row] = (X)rv;
            if (first) p[row] = (X)rv;
        }
    }
    bt_partials<2>(acc, red, kSlotRR, gb);
}

// One block per system: bb, rr, the convergence test on the true residual.
__global__ void k_bt_fin_residual(BatchMap map, SysScalars* sys, RedScratchBig red, int first, int m) {
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
    }
}

// K1: per-system rr (first block of the system), q = A p, partials p.q, r.q, q.q
template <class V, class X>
__global__ void __launch_bounds__(1024) k_bt_spmv(BtArgs<V, X> c) {
    const int gb = c.map.block();
    const int k = c.map.bsys[gb];
    SysScalars* S = c.sys + k;
    if (S->done) {
        if (gb == c.map.bptr[k] && threadIdx.x == 0) S->halt = 1;
        return;
    }
    if (gb == c.map.bptr[k]) {
        const double rr = sys_sum(c.red, kSlotRR, c.map.bptr[k], c.map.bptr[k + 1]);
        if (threadIdx.x == 0) {
            S->rr = rr;
            const double scale = S->bb > 0.0 ? S->bb : 1.0;
            if (!(rr > S->tol2 * scale) || S->iter >= S->max_iter) S->done = 1;
            if (!isfinite(rr)) { S->nonfinite = 1; S->done = 1; }
            if (sqrt(rr) < S->replace_ratio * S->r_ref) S->need_update = 1;
        }
    }
    const int lane = threadIdx.x & (kWave - 1);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int s = c.map.bs0[gb] + (int)(threadIdx.x / kWave); s < c.map.bs1[gb]; s += (int)(blockDim.x / kWave)) {
        const int row = s * kWave + lane;
        const X pi = row < c.A.n ? c.p[row] : (X)0;
        const X qv = (X)sellp_row<V, X, sizeof(V) == 4>(c.A, s, lane, row, c.p, (double)pi);
        if (row < c.A.n) {
            const double rv = (double)c.r[row];
            c.q[row] = qv;
            acc[0] += (double)pi * (double)qv;
            acc[1] += rv * (double)qv;
            acc[2] += (double)qv * (double)qv;
        }
    }
    // p.q, then r.q and q.q in slots of their own: the r.r slot is being read
    // by the first block of each system during this kernel
    const double pq1[1] = {acc[0]};
    bt_partials<1>(pq1, c.red, kSlotPQ, gb);
    const double rq_qq[2] = {acc[1], acc[2]};
    bt_partials<2>(rq_qq, c.red, kSlotRQ, gb);
}

// K2: alpha, beta per system; x += alpha p; r -= alpha q; p = r + beta p; r.r
template <class V, class X>
__global__ void __launch_bounds__(1024) k_bt_update(BtArgs<V, X> c) {
    constexpr int N = Vec16<X>::N;
    const int gb = c.map.block();
    const int k = c.map.bsys[gb];
    SysScalars* S = c.sys + k;
    if (S->done) return;
    const int b0 = c.map.bptr[k], b1 = c.map.bptr[k + 1];
    constexpr int kSl[3] = {kSlotPQ, kSlotRQ, kSlotQQ};
    SysSums<3> ps;
    ps.issue(c.red, kSl, b0, b1);
    // the block's rows are whole 64-row slices: 16-byte vectors, the first
    // one in flight while the partial sums are re-added
    const int v0 = c.map.bs0[gb] * (kWave / N), v1 = c.map.bs1[gb] * (kWave / N);
    const int va = v0 + (int)threadIdx.x;
    double xa[N];
    X pa[N], ra[N], qa[N];
    if (va < v1) {
        dload<N>(c.x, va, xa);
        vload(c.p, va, pa);
        vload(c.r, va, ra);
        vload(c.q, va, qa);
    }
    double d3[3];
    ps.finish(c.red, kSl, b0, b1, d3);
    const double pq = d3[0], rq = d3[1], qq = d3[2];
    if (!(pq != 0.0 && isfinite(pq))) {
        if (gb == b0 && threadIdx.x == 0) {
            S->breakdown = (pq == 0.0);
            S->nonfinite = !isfinite(pq);
            S->done = 1;
        }
        return;
    }
    const double rr = S->rr;
    const double alpha = rr / pq;
    double rr_next = rr - 2.0 * alpha * rq + alpha * alpha * qq;
    if (!(rr_next > 0.0)) rr_next = 0.0;
    const double beta = rr_next / rr;
    double acc[1] = {0.0};
    auto body = [&](int v, double (&xe)[N], const X (&pe)[N], const X (&re)[N], const X (&qe)[N]) {
        X rn[N], pn[N];
#pragma unroll
        for (int e = 0; e < N; ++e) {
            const double p0 = (double)pe[e];
            xe[e] += alpha * p0;
            rn[e] = (X)((double)re[e] - alpha * (double)qe[e]);
            pn[e] = (X)((double)rn[e] + beta * p0);
            acc[0] += (double)rn[e] * (double)rn[e];
        }
        dstore<N>(c.x, v, xe);
        vstore(c.r, v, rn);
        vstore(c.p, v, pn);
    };
    if (va < v1) body(va, xa, pa, ra, qa);
    for (int v = va + (int)blockDim.x; v < v1; v += (int)blockDim.x) {
        double xe[N];
        X pe[N], re[N], qe[N];
        dload<N>(c.x, v, xe);
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
    int n_tot = 0, ns_tot = 0, nblocks = 0, block = 1024;
    std::vector<int> bptr_host;
    // concatenated SELL-P (global numbering)
    DBuf<int> slice_ptr, src, diag_src, bsys, bs0, bs1, bptr, perm;  // perm: padded row -> (flat input index)
    DBuf<int> active;                    // compacted list of the blocks of unfinished systems
    DBuf<unsigned> colw;
    DBuf<int> col;
    bool compact = true;
    int64_t total = 0;
    DBuf<double> val_in, val64, diag64, b_in, b_glob, x_glob;
    DBuf<float> val32, diag32;
    std::vector<int64_t> val_off;        // offset of system k's values in val_in
    int64_t n_flat = 0;                  // sum of the systems' rows
    bool have_vals = false, have_b = false;
    DBuf<double> d[4];
    DBuf<float> f[4];
    DBuf<double> red_partial;
    DBuf<SysScalars> sys;
    SysScalars* sys_host = nullptr;
    std::map<int64_t, hipGraphExec_t> graphs;  // iteration chunks by (precision, chunk, launch size, compacted)
    ~ocsl_batch_s() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        if (sys_host) (void)hipHostFree(sys_host);
        if (ev0) (void)hipEventDestroy(ev0);
        if (stream) (void)hipStreamDestroy(stream);
    }
    RedScratchBig red() const { return {red_partial.p, (size_t)nblocks}; }
    std::vector<int64_t> flat_off() const {
        std::vector<int64_t> o(m + 1, 0);
        for (int k = 0; k < m; ++k) o[k + 1] = o[k] + n_sys[k];
        return o;
    }
    SellPView<double> view64() const {
        return {n_tot, ns_tot, slice_ptr.p, compact ? nullptr : col.p, compact ? colw.p : nullptr, val64.p, diag64.p};
    }
    SellPView<float> view32() const {
        return {n_tot, ns_tot, slice_ptr.p, compact ? nullptr : col.p, compact ? colw.p : nullptr, val32.p, diag32.p};
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

// Per-slice-width conversion of the concatenated value arrays; the FP32
// diagonal is the FP64 row sum (difference form).
__global__ void __launch_bounds__(kBlock) k_bt_rowsum(SellPView<double> A, float* __restrict__ diag32) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * kWavesPerBlock;
    for (int s = wave; s < A.n_slices; s += nwaves) {
        const int row = s * kWave + lane;
        const int base = A.slice_ptr[s];
        const int w = (A.slice_ptr[s + 1] - base) >> 6;
        double sum = A.diag[row];
        for (int k = 0; k < w; ++k) sum += A.val[base + (k / 2) * 2 * kWave + 2 * lane + (k & 1)];
        diag32[row] = (float)sum;
    }
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
        std::vector<int> sp_all{0}, colg, srcg, dsrc, perm, bsys, bs0, bs1, bptr{0};
        std::vector<int> row_off{0};
        int64_t val_base = 0;
        bt->val_off.assign(m + 1, 0);
        bool compact = true;
        const int spb = bt->block / kWave;  // slices per block: one per wave
        int slice_base = 0;
        int64_t flat_base = 0;  // offset of system k in the flat (concatenated) vectors
        for (int k = 0; k < m; ++k) {
            HostCSR a;
            analyse_structure(n_rows[k], row_ptr[k], column_index[k], a);
            const int n = a.n, ns = (n + kWave - 1) / kWave;
            std::vector<int> order = rcm_order(a);
            std::vector<int> iperm(n);
            for (int i = 0; i < n; ++i) iperm[order[i]] = i;
            // slice widths
            const int r0 = row_off.back();
            for (int sl = 0; sl < ns; ++sl) {
                int w = 0;
                for (int l = 0; l < kWave && sl * kWave + l < n; ++l) {
                    const int o = order[sl * kWave + l];
                    w = std::max(w, a.rp[o + 1] - a.rp[o] - 1);
                }
                w += w & 1;
                const int base = sp_all.back();
                sp_all.push_back(base + w * kWave);
                colg.resize(base + (size_t)w * kWave);
                srcg.resize(base + (size_t)w * kWave);
                for (int l = 0; l < kWave; ++l) {
                    const int r = sl * kWave + l;  // local new row
                    int kk = 0;
                    if (r < n) {
                        const int o = order[r];
                        for (int p = a.rp[o]; p < a.rp[o + 1]; ++p) {
                            if (p == a.diag[o]) continue;
                            const int64_t q = base + (int64_t)(kk / 2) * 2 * kWave + 2 * l + (kk & 1);
                            colg[q] = r0 + iperm[a.col[p]];
                            srcg[q] = (int)(val_base + a.src[p]);
                            const int d = colg[q] - (r0 + r);
                            if (d < -32768 || d > 32767) compact = false;
                            ++kk;
                        }
                    }
                    for (; kk < w; ++kk) {
                        const int64_t q = base + (int64_t)(kk / 2) * 2 * kWave + 2 * l + (kk & 1);
                        colg[q] = r0 + r;  // padding rows of a system are its own (zero) rows
                        srcg[q] = -1;
                    }
                }
            }
            for (int r = 0; r < ns * kWave; ++r) {
                perm.push_back(r < n ? (int)(flat_base + order[r]) : -1);
                dsrc.push_back(r < n ? (int)(val_base + a.src[a.diag[order[r]]]) : -1);
            }
            // blocks of this system
            for (int s0 = 0; s0 < ns; s0 += spb) {
                bsys.push_back(k);
                bs0.push_back(slice_base + s0);
                bs1.push_back(slice_base + std::min(ns, s0 + spb));
            }
            bptr.push_back((int)bsys.size());
            slice_base += ns;
            row_off.push_back(r0 + ns * kWave);
            flat_base += n;
            val_base += a.nnz_in;
            bt->nnz_in.push_back(a.nnz_in);
            bt->val_off[k + 1] = val_base;
        }
        OCSL_REQUIRE(val_base < ((int64_t)1 << 31) && row_off.back() < (1 << 30), OCSL_ERR_ARG, "batch too large");
        bt->n_flat = flat_base;
        bt->row_off = row_off;
        bt->n_tot = row_off.back();
        bt->ns_tot = slice_base;
        bt->nblocks = (int)bsys.size();
        bt->bptr_host = bptr;
        bt->total = sp_all.back();
        bt->compact = compact;
        hipStream_t s = bt->stream;
        bt->slice_ptr.upload(sp_all.data(), sp_all.size(), s);
        if (compact) {
            std::vector<unsigned> colw(bt->total / 2);
            for (int sl = 0; sl < bt->ns_tot; ++sl) {
                const int np = (sp_all[sl + 1] - sp_all[sl]) / (2 * kWave);
                for (int t = 0; t < np; ++t)
                    for (int l = 0; l < kWave; ++l) {
                        const int64_t q = sp_all[sl] + (int64_t)t * 2 * kWave + 2 * l;
                        const int row = sl * kWave + l;
                        const unsigned lo = (unsigned)(uint16_t)(int16_t)(colg[q] - row);
                        const unsigned hi = (unsigned)(uint16_t)(int16_t)(colg[q + 1] - row);
                        colw[sp_all[sl] / 2 + (int64_t)t * kWave + l] = lo | (hi << 16);
                    }
            }
            bt->colw.upload(colw.data(), colw.size(), s);
        } else {
            bt->col.upload(colg.data(), colg.size(), s);
        }
        bt->src.upload(srcg.data(), srcg.size(), s);
        bt->diag_src.upload(dsrc.data(), dsrc.size(), s);
        bt->perm.upload(perm.data(), perm.size(), s);
        bt->bsys.upload(bsys.data(), bsys.size(), s);
        bt->bs0.upload(bs0.data(), bs0.size(), s);
        bt->bs1.upload(bs1.data(), bs1.size(), s);
        bt->bptr.upload(bptr.data(), bptr.size(), s);
        bt->red_partial.alloc((size_t)bt->nblocks * 8);
        bt->active.alloc(bt->nblocks);
        bt->sys.alloc(m);
        OCSL_HIP(hipHostMalloc((void**)&bt->sys_host, sizeof(SysScalars) * m, hipHostMallocDefault));
        OCSL_HIP(hipStreamSynchronize(s));
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
        bt->val64.alloc(bt->total);
        bt->diag64.alloc(bt->n_tot);
        bt->val32.alloc(bt->total);
        bt->diag32.alloc(bt->n_tot);
        k_gather_vals<double><<<grid_for(bt->total), kBlock, 0, s>>>(bt->total, bt->src.p, bt->val_in.p, bt->val64.p);
        k_gather_vals<double><<<grid_for(bt->n_tot), kBlock, 0, s>>>(bt->n_tot, bt->diag_src.p, bt->val_in.p,
                                                                    bt->diag64.p);
        k_gather_vals<float><<<grid_for(bt->total), kBlock, 0, s>>>(bt->total, bt->src.p, bt->val_in.p, bt->val32.p);
        k_bt_rowsum<<<grid_for((int64_t)bt->ns_tot * kWave), kBlock, 0, s>>>(bt->view64(), bt->diag32.p);
        OCSL_HIP(hipGetLastError());
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
        bt->b_glob.alloc(bt->n_tot);
        bt_flat_in(bt, b, where, bt->b_in, bt->b_glob.p);
        OCSL_HIP(hipStreamSynchronize(bt->stream));
        bt->have_b = true;
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
        const int m = bt->m, N = bt->n_tot, G = bt->nblocks, B = bt->block;
        DBuf<double> flat;
        bt->x_glob.alloc(N);
        bt_flat_in(bt, x, where, flat, bt->x_glob.p);
        // vectors in the Krylov precision
        const bool remap = mixed ? (!bt->f[0].p || bt->f[0].n < (size_t)N) : (!bt->d[0].p || bt->d[0].n < (size_t)N);
        if (remap) {
            for (auto& kv : bt->graphs) (void)hipGraphExecDestroy(kv.second);
            bt->graphs.clear();
        }
        for (int i = 0; i < 3; ++i) {
            if (mixed) bt->f[i].alloc(N);
            else bt->d[i].alloc(N);
        }
        std::vector<SysScalars> init(m);
        for (auto& S : init) {
            S = SysScalars{};
            S.tol2 = o.tol * o.tol;
            S.max_iter = o.max_iter;
            S.replace_ratio = mixed ? o.replace_ratio : 0.0;
        }
        OCSL_HIP(hipMemcpyAsync(bt->sys.p, init.data(), sizeof(SysScalars) * m, hipMemcpyHostToDevice, s));
        // padded rows must stay zero in every vector
        for (int i = 0; i < 3; ++i) {
            if (mixed) OCSL_HIP(hipMemsetAsync(bt->f[i].p, 0, sizeof(float) * N, s));
            else OCSL_HIP(hipMemsetAsync(bt->d[i].p, 0, sizeof(double) * N, s));
        }
        // launched blocks: all, or (compacted) the blocks of unfinished systems
        int G_cur = G;
        bool comp = false;
        BtArgs<float, float> c32{bt->view32(), bt->map(), bt->f[0].p, bt->f[1].p, bt->f[2].p, bt->x_glob.p,
                                 bt->sys.p, bt->red()};
        BtArgs<double, double> c64{bt->view64(), bt->map(), bt->d[0].p, bt->d[1].p, bt->d[2].p, bt->x_glob.p,
                                   bt->sys.p, bt->red()};
        auto residual = [&](int first, bool all) {
            const BatchMap mp = bt->map(comp && !all);
            const int gl = (comp && !all) ? G_cur : G;
            if (mixed)
                k_bt_residual<float><<<gl, B, 0, s>>>(bt->view64(), mp, bt->b_glob.p, bt->x_glob.p, c32.r, c32.p,
                                                      bt->sys.p, bt->red(), first);
            else
                k_bt_residual<double><<<gl, B, 0, s>>>(bt->view64(), mp, bt->b_glob.p, bt->x_glob.p, c64.r, c64.p,
                                                       bt->sys.p, bt->red(), first);
            k_bt_fin_residual<<<m, 256, 0, s>>>(bt->map(), bt->sys.p, bt->red(), first, m);
            OCSL_HIP(hipGetLastError());
        };
        auto iter = [&] {
            if (mixed) {
                k_bt_spmv<float, float><<<G_cur, B, 0, s>>>(c32);
                k_bt_update<float, float><<<G_cur, kBtUpdateBlock, 0, s>>>(c32);
            } else {
                k_bt_spmv<double, double><<<G_cur, B, 0, s>>>(c64);
                k_bt_update<double, double><<<G_cur, kBtUpdateBlock, 0, s>>>(c64);
            }
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
            c32.map = c64.map = bt->map(true);
        };
        OCSL_HIP(hipEventRecord(bt->ev0, s));
        residual(1, true);
        read_sys();
        stamp();
        const int chunk = o.check_every;
        int launched = 0;
        if (bt->graphs.size() > 64) {
            for (auto& kv : bt->graphs) (void)hipGraphExecDestroy(kv.second);
            bt->graphs.clear();
        }
        while (true) {
            bool all = true;
            for (int k = 0; k < m; ++k) all = all && finished(bt->sys_host[k]);
            if (all || launched >= o.max_iter) break;
            compact();
            // graphs hold the launch size and map; the active list is read
            // through its (fixed) device address when the graph runs
            const int64_t key = (mixed ? 1 : 0) + 2 * (int64_t)chunk + ((int64_t)G_cur << 24) + (comp ? (1ll << 56) : 0);
            if (o.use_graph) {
                auto it = bt->graphs.find(key);
                if (it == bt->graphs.end()) {
                    hipGraph_t gr;
                    OCSL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    for (int i = 0; i < chunk; ++i) iter();
                    OCSL_HIP(hipStreamEndCapture(s, &gr));
                    hipGraphExec_t ex;
                    OCSL_HIP(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
                    OCSL_HIP(hipGraphDestroy(gr));
                    it = bt->graphs.emplace(key, ex).first;
                }
                OCSL_HIP(hipGraphLaunch(it->second, s));
            } else {
                for (int i = 0; i < chunk; ++i) iter();
                OCSL_HIP(hipGetLastError());
            }
            launched += chunk;
            read_sys();
            if (mixed) {
                bool any = false;
                for (int k = 0; k < m; ++k) {
                    const SysScalars& S = bt->sys_host[k];
                    any = any || S.need_update || (S.done && !S.verified && !S.nonfinite && !S.breakdown);
                }
                if (any) {
                    residual(0, false);
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
            bt->sys_host[k].need_update = 1;
            bt->sys_host[k].nonfinite = 0;
            bt->sys_host[k].breakdown = 0;
        }
        OCSL_HIP(hipMemcpyAsync(bt->sys.p, bt->sys_host, sizeof(SysScalars) * m, hipMemcpyHostToDevice, s));
        residual(0, true);
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
            }
            if (latency_ms) latency_ms[k] = lat[k] >= 0.0 ? lat[k] : total_ms;
        }
        OCSL_HIP(hipStreamSynchronize(s));
    });
    if (st != OCSL_OK) return st;
    return (ocsl_status)worst;
}

}  // extern "C"
