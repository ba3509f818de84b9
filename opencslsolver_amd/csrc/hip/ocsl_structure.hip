// Matrix structure builder on the device: the CSR pattern of
// CFVM::define_matrix_structure (cfvm.cpp:74-98) and the slot maps an
// assembly needs, from the mesh's face lists.
//
// The reference inserts, for every cell i, the element (i, i) and then
// (i, neighbour) for each entry of the cell's neighbour list (cgeometry.cpp:
// 602-612: the cell across each face, -1 for a boundary face) through
// CCSR::insert_element (cmatrix_solver.cpp:57-100), which keeps each row
// sorted ascending and stores a repeated column once.  The -1 of a boundary
// face becomes the column (size_t)-1.  So row i holds, ascending, the set
// {i} + {cells across its faces} + {(uint64)-1 if it owns a boundary face}.
// Each insert shifts every later row pointer (O(n nnz) in all); here:
//   1. degree: 1 + faces per cell (atomics), boundary flag per cell;
//   2. scan of the degrees (int64, three phases);
//   3. fill: the diagonal, then each face's two columns at atomic cursors;
//   4. per row (one thread): insertion sort and removal of repeats, so the
//      order the atomics left does not matter -- the output is deterministic;
//   5. scan of the unique counts (+1 for the sentinel) = row_ptr;
//   6. emit the columns as uint64 and, optionally, the slot of each face's
//      two off-diagonal entries and of each diagonal.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "ocsl_handle.h"

namespace ocsl {
namespace {

constexpr int kScanBlock = 1024;
constexpr int kScanItems = 4;  // per thread
constexpr int kScanTile = kScanBlock * kScanItems;

__global__ void k_st_check(int n, int F, const int* __restrict__ fL, const int* __restrict__ fR, int B,
                           const int* __restrict__ bL, int* __restrict__ bad) {
    const int stride = gridDim.x * blockDim.x;
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < F; f += stride) {
        const int l = fL[f], r = fR[f];
        if (l < 0 || l >= n || r < 0 || r >= n || l == r) *bad = 1;
    }
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += stride) {
        const int c = bL[b];
        if (c < 0 || c >= n) *bad = 1;
    }
}

__global__ void k_st_init(int n, int* __restrict__ deg, int* __restrict__ bflag) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        deg[i] = 1;  // the diagonal
        bflag[i] = 0;
    }
}

__global__ void k_st_degree(int F, const int* __restrict__ fL, const int* __restrict__ fR, int B,
                            const int* __restrict__ bL, int* __restrict__ deg, int* __restrict__ bflag) {
    const int stride = gridDim.x * blockDim.x;
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < F; f += stride) {
        atomicAdd(&deg[fL[f]], 1);
        atomicAdd(&deg[fR[f]], 1);
    }
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += stride) bflag[bL[b]] = 1;
}

// Exclusive int64 scan of n ints in three phases: tiles of 4096 (block scan
// with wave shuffles), a one-block scan of the tile totals, the tile offsets
// added.  out[n] receives the total.
__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int64_t t = __shfl_up(v, off, kWave);
        if (lane >= off) v += t;
    }
    return v;
}

// Inclusive scan of one value per thread over the block; returns it, and the
// block total through *total.
__device__ __forceinline__ int64_t block_incl_scan(int64_t v, int64_t* total) {
    __shared__ int64_t sh[kScanBlock / kWave];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave, nw = blockDim.x / kWave;
    v = wave_incl_scan(v);
    if (lane == kWave - 1) sh[wid] = v;
    __syncthreads();
    if (wid == 0) {
        int64_t w = lane < nw ? sh[lane] : 0;
        w = wave_incl_scan(w);
        if (lane < nw) sh[lane] = w;
    }
    __syncthreads();
    const int64_t add = wid > 0 ? sh[wid - 1] : 0;
    *total = sh[nw - 1];
    __syncthreads();
    return v + add;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_tiles(int n, const int* __restrict__ in,
                                                           int64_t* __restrict__ out, int64_t* __restrict__ tile_sum) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int64_t v[kScanItems], s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        s += v[k];
    }
    int64_t total;
    const int64_t incl = block_incl_scan(s, &total);
    int64_t run = incl - s;  // exclusive prefix of this thread's items
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// One block: exclusive scan of the G tile totals in place; out[n] = total.
__global__ void __launch_bounds__(kScanBlock) k_scan_totals(int G, int64_t* __restrict__ tile_sum,
                                                            int64_t* __restrict__ out_total) {
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int g0 = 0; g0 < G; g0 += kScanBlock) {
        const int g = g0 + (int)threadIdx.x;
        const int64_t v = g < G ? tile_sum[g] : 0;
        int64_t total;
        const int64_t incl = block_incl_scan(v, &total);
        const int64_t c = carry;
        if (g < G) tile_sum[g] = c + incl - v;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + total;
        __syncthreads();
    }
    if (threadIdx.x == 0) *out_total = carry;
}

__global__ void k_scan_add(int n, int64_t* __restrict__ out, const int64_t* __restrict__ tile_sum) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] += tile_sum[i / kScanTile];
}

void scan_exclusive(int n, const int* in, int64_t* out, DBuf<int64_t>& tiles, hipStream_t s) {
    const int G = std::max(1, (n + kScanTile - 1) / kScanTile);
    tiles.alloc(G);
    k_scan_tiles<<<G, kScanBlock, 0, s>>>(n, in, out, tiles.p);
    k_scan_totals<<<1, kScanBlock, 0, s>>>(G, tiles.p, out + n);
    k_scan_add<<<grid_for(n), kBlock, 0, s>>>(n, out, tiles.p);
    OCSL_HIP(hipGetLastError());
}

__global__ void k_st_fill_diag(int n, const int64_t* __restrict__ tptr, int* __restrict__ tcol,
                               int* __restrict__ cursor) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        tcol[tptr[i]] = i;
        cursor[i] = 1;
    }
}

__global__ void k_st_fill(int F, const int* __restrict__ fL, const int* __restrict__ fR,
                          const int64_t* __restrict__ tptr, int* __restrict__ tcol, int* __restrict__ cursor) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < F; f += gridDim.x * blockDim.x) {
        const int l = fL[f], r = fR[f];
        tcol[tptr[l] + atomicAdd(&cursor[l], 1)] = r;
        tcol[tptr[r] + atomicAdd(&cursor[r], 1)] = l;
    }
}

// Sorts each row's columns ascending in place and drops repeats; ucount =
// distinct columns + 1 when the row takes the boundary sentinel.
__global__ void k_st_sort(int n, const int64_t* __restrict__ tptr, int* __restrict__ tcol,
                          const int* __restrict__ bflag, int* __restrict__ ucount) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int* c = tcol + tptr[i];
        const int d = (int)(tptr[i + 1] - tptr[i]);
        for (int a = 1; a < d; ++a) {
            const int v = c[a];
            int b = a - 1;
            while (b >= 0 && c[b] > v) {
                c[b + 1] = c[b];
                --b;
            }
            c[b + 1] = v;
        }
        int u = d > 0 ? 1 : 0;
        for (int a = 1; a < d; ++a)
            if (c[a] != c[u - 1]) c[u++] = c[a];
        ucount[i] = u + bflag[i];
    }
}

__global__ void k_st_emit(int n, const int64_t* __restrict__ tptr, const int* __restrict__ tcol,
                          const int64_t* __restrict__ rp, const int* __restrict__ bflag,
                          uint64_t* __restrict__ col_out, int64_t* __restrict__ dslot) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int64_t o = rp[i];
        const int u = (int)(rp[i + 1] - o) - bflag[i];
        const int* c = tcol + tptr[i];
        for (int a = 0; a < u; ++a) {
            col_out[o + a] = (uint64_t)c[a];
            if (dslot && c[a] == i) dslot[i] = o + a;
        }
        if (bflag[i]) col_out[o + u] = ~(uint64_t)0;
    }
}

__global__ void k_st_rowptr(int n, const int64_t* __restrict__ rp, uint64_t* __restrict__ rp_out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x) rp_out[i] = (uint64_t)rp[i];
}

// Slot of column `col` in row `row` (sorted, distinct): binary search.
__device__ __forceinline__ int64_t row_find(const int64_t* rp, const uint64_t* cols, int row, uint64_t col) {
    int64_t lo = rp[row], hi = rp[row + 1];
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (cols[mid] < col) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void k_st_face_slots(int F, const int* __restrict__ fL, const int* __restrict__ fR,
                                const int64_t* __restrict__ rp, const uint64_t* __restrict__ cols,
                                int64_t* __restrict__ fslot) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < F; f += gridDim.x * blockDim.x) {
        const int l = fL[f], r = fR[f];
        fslot[2 * (int64_t)f] = row_find(rp, cols, l, (uint64_t)r);
        fslot[2 * (int64_t)f + 1] = row_find(rp, cols, r, (uint64_t)l);
    }
}

template <class T>
const T* dev_in(const T* p, size_t count, ocsl_where where, DBuf<T>& tmp, hipStream_t s) {
    if (where == OCSL_DEVICE || count == 0) return p;
    tmp.upload(p, count, s);
    return tmp.p;
}

template <class T>
void copy_out(T* dst, const T* src, size_t count, ocsl_where where, hipStream_t s) {
    if (!dst || count == 0) return;
    OCSL_HIP(hipMemcpyAsync(dst, src, sizeof(T) * count,
                            where == OCSL_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s));
}

}  // namespace
}  // namespace ocsl

using namespace ocsl;

extern "C" ocsl_status ocsl_build_structure(int64_t n_cells, int64_t n_faces, const int32_t* face_left,
                                            const int32_t* face_right, int64_t n_bc, const int32_t* bc_left,
                                            ocsl_where where, uint64_t* row_ptr, uint64_t* column_index,
                                            int64_t column_cap, int64_t* nnz, int64_t* face_slot,
                                            int64_t* diag_slot) {
    if (n_cells <= 0 || n_cells >= (1ll << 30) || n_faces < 0 || n_faces >= (1ll << 30) || n_bc < 0 ||
        n_bc >= (1ll << 31) || !nnz || (n_faces && (!face_left || !face_right)) || (n_bc && !bc_left))
        return OCSL_ERR_ARG;
    if (where != OCSL_HOST && where != OCSL_DEVICE) return OCSL_ERR_ARG;
    if (ocsl_device_count() <= 0) {
        set_last_error("no HIP device");
        return OCSL_ERR_NODEVICE;
    }
    hipStream_t s = nullptr;
    try {
        OCSL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        const int n = (int)n_cells, F = (int)n_faces, B = (int)n_bc;
        DBuf<int> tmp_l, tmp_r, tmp_b, deg, bflag, cursor, tcol, ucount, bad;
        DBuf<int64_t> tptr, rp, tiles, fs, ds;
        DBuf<uint64_t> cols, rpo;
        const int* fL = dev_in(face_left, F, where, tmp_l, s);
        const int* fR = dev_in(face_right, F, where, tmp_r, s);
        const int* bL = dev_in(bc_left, B, where, tmp_b, s);
        // indices checked before any of them addresses memory
        bad.alloc(1);
        OCSL_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
        k_st_check<<<grid_for(std::max(F, B)), kBlock, 0, s>>>(n, F, fL, fR, B, bL, bad.p);
        int bad_h = 0;
        OCSL_HIP(hipMemcpyAsync(&bad_h, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
        OCSL_HIP(hipStreamSynchronize(s));
        OCSL_REQUIRE(!bad_h, OCSL_ERR_STRUCTURE, "face or boundary-face cell index out of range (or a face with one cell on both sides)");
        deg.alloc(n);
        bflag.alloc(n);
        k_st_init<<<grid_for(n), kBlock, 0, s>>>(n, deg.p, bflag.p);
        k_st_degree<<<grid_for(std::max(F, B)), kBlock, 0, s>>>(F, fL, fR, B, bL, deg.p, bflag.p);
        tptr.alloc((size_t)n + 1);
        scan_exclusive(n, deg.p, tptr.p, tiles, s);
        const int64_t tot = (int64_t)n + 2 * (int64_t)F;
        OCSL_REQUIRE(tot < (1ll << 31), OCSL_ERR_ARG, "structure too large");
        tcol.alloc((size_t)tot);
        cursor.alloc(n);
        k_st_fill_diag<<<grid_for(n), kBlock, 0, s>>>(n, tptr.p, tcol.p, cursor.p);
        k_st_fill<<<grid_for(F), kBlock, 0, s>>>(F, fL, fR, tptr.p, tcol.p, cursor.p);
        ucount.alloc(n);
        k_st_sort<<<grid_for(n), kBlock, 0, s>>>(n, tptr.p, tcol.p, bflag.p, ucount.p);
        rp.alloc((size_t)n + 1);
        scan_exclusive(n, ucount.p, rp.p, tiles, s);
        int64_t total = 0;
        OCSL_HIP(hipMemcpyAsync(&total, rp.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        OCSL_HIP(hipStreamSynchronize(s));
        *nnz = total;
        if (!column_index && row_ptr) {
            // size query: the row pointers only
            rpo.alloc((size_t)n + 1);
            k_st_rowptr<<<grid_for((int64_t)n + 1), kBlock, 0, s>>>(n, rp.p, rpo.p);
            OCSL_HIP(hipGetLastError());
            copy_out(row_ptr, rpo.p, (size_t)n + 1, where, s);
        }
        if (column_index) {
            OCSL_REQUIRE(column_cap >= total, OCSL_ERR_ARG, "column_index holds fewer entries than the structure");
            cols.alloc((size_t)std::max<int64_t>(total, 1));
            rpo.alloc((size_t)n + 1);
            if (diag_slot) ds.alloc(n);
            k_st_emit<<<grid_for(n), kBlock, 0, s>>>(n, tptr.p, tcol.p, rp.p, bflag.p, cols.p,
                                                     diag_slot ? ds.p : nullptr);
            k_st_rowptr<<<grid_for((int64_t)n + 1), kBlock, 0, s>>>(n, rp.p, rpo.p);
            if (face_slot && F) {
                fs.alloc(2 * (size_t)F);
                k_st_face_slots<<<grid_for(F), kBlock, 0, s>>>(F, fL, fR, rp.p, cols.p, fs.p);
            }
            OCSL_HIP(hipGetLastError());
            copy_out(column_index, cols.p, (size_t)total, where, s);
            copy_out(row_ptr, rpo.p, (size_t)n + 1, where, s);
            if (face_slot) copy_out(face_slot, fs.p, 2 * (size_t)F, where, s);
            if (diag_slot) copy_out(diag_slot, ds.p, (size_t)n, where, s);
        }
        OCSL_HIP(hipStreamSynchronize(s));
        OCSL_HIP(hipStreamDestroy(s));
        return OCSL_OK;
    } catch (const Error& e) {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        set_last_error(e.what);
        return e.status;
    } catch (const std::bad_alloc&) {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        set_last_error("host allocation failed");
        return OCSL_ERR_NOMEM;
    }
}
