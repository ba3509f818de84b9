// Host-side structure analysis (setup, once per handle): sentinel stripping,
// reverse Cuthill-McKee order, SELL-64 layout, triangular level schedules and
// the fine-grained ILU(0) sweep tables.
#include <algorithm>
#include <numeric>

#include "ocsl_handle.h"

namespace ocsl {

// Validates the reference CCSR layout (cmatrix_solver.h:33-51) and keeps the
// real entries: columns < n.  Sentinel columns (cfvm.cpp:90-93) and anything
// else >= n are dropped; they hold 0.0 in every reference code path.
void analyse_structure(int64_t n, const uint64_t* rp, const uint64_t* col, HostCSR& out) {
    OCSL_REQUIRE(n > 0 && n < (int64_t)1 << 30, OCSL_ERR_ARG, "n out of range");
    OCSL_REQUIRE(rp && col, OCSL_ERR_ARG, "null structure pointer");
    OCSL_REQUIRE(rp[0] == 0, OCSL_ERR_STRUCTURE, "row_ptr[0] != 0");
    const uint64_t nnz_in = rp[n];
    OCSL_REQUIRE(nnz_in < ((uint64_t)1 << 31), OCSL_ERR_ARG, "nnz exceeds int32 range");
    out.n = (int)n;
    out.nnz_in = (int64_t)nnz_in;
    out.rp.assign(n + 1, 0);
    out.col.clear();
    out.src.clear();
    out.diag.assign(n, -1);
    out.col.reserve(nnz_in);
    out.src.reserve(nnz_in);
    for (int64_t i = 0; i < n; ++i) {
        OCSL_REQUIRE(rp[i + 1] >= rp[i], OCSL_ERR_STRUCTURE, "row_ptr not monotone");
        int64_t prev = -1;
        for (uint64_t p = rp[i]; p < rp[i + 1]; ++p) {
            const uint64_t c = col[p];
            if (c >= (uint64_t)n) continue;
            OCSL_REQUIRE((int64_t)c > prev, OCSL_ERR_STRUCTURE,
                         "columns not strictly ascending in row " + std::to_string(i));
            prev = (int64_t)c;
            if ((int64_t)c == i) out.diag[i] = (int)out.col.size();
            out.col.push_back((int)c);
            out.src.push_back((int)p);
        }
        out.rp[i + 1] = (int)out.col.size();
        OCSL_REQUIRE(out.diag[i] >= 0, OCSL_ERR_STRUCTURE, "row " + std::to_string(i) + " has no diagonal");
    }
}

// Breadth-first levels from `root` within the unvisited part; returns the
// last level's nodes and the eccentricity.
static int bfs_levels(const HostCSR& a, int root, const std::vector<char>& done,
                      std::vector<int>& stamp, int tag, std::vector<int>& last) {
    std::vector<int> cur{root}, nxt;
    stamp[root] = tag;
    int depth = 0;
    while (true) {
        nxt.clear();
        for (int v : cur)
            for (int p = a.rp[v]; p < a.rp[v + 1]; ++p) {
                const int w = a.col[p];
                if (!done[w] && stamp[w] != tag) {
                    stamp[w] = tag;
                    nxt.push_back(w);
                }
            }
        if (nxt.empty()) break;
        cur.swap(nxt);
        ++depth;
    }
    last = cur;
    return depth;
}

// Reverse Cuthill-McKee on the row graph, one component at a time, started
// from a pseudo-peripheral node (George & Liu).  Returns new -> old.
std::vector<int> rcm_order(const HostCSR& a) {
    const int n = a.n;
    std::vector<int> deg(n);
    for (int i = 0; i < n; ++i) deg[i] = a.rp[i + 1] - a.rp[i];
    std::vector<char> done(n, 0);
    std::vector<int> stamp(n, -1), order;
    order.reserve(n);
    int tag = 0;
    std::vector<int> last, nb;
    for (int seed = 0; seed < n; ++seed) {
        if (done[seed]) continue;
        int root = seed;
        int ecc = bfs_levels(a, root, done, stamp, tag++, last);
        for (int rep = 0; rep < 6; ++rep) {
            int best = last[0];
            for (int v : last)
                if (deg[v] < deg[best]) best = v;
            const int e2 = bfs_levels(a, best, done, stamp, tag++, last);
            if (e2 <= ecc) break;
            ecc = e2;
            root = best;
        }
        const size_t start = order.size();
        order.push_back(root);
        done[root] = 1;
        for (size_t head = start; head < order.size(); ++head) {
            const int v = order[head];
            nb.clear();
            for (int p = a.rp[v]; p < a.rp[v + 1]; ++p) {
                const int w = a.col[p];
                if (!done[w]) {
                    done[w] = 1;
                    nb.push_back(w);
                }
            }
            std::sort(nb.begin(), nb.end(), [&](int x, int y) {
                return deg[x] != deg[y] ? deg[x] < deg[y] : x < y;
            });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

// SELL-64 in the row order `order` (new -> old; null = identity).  Columns
// are renumbered into the same order and kept in the reference's per-row
// ascending order of the original columns, so a row sum adds its terms in the
// order CMatrix_solver::gemv does.
void build_sell(const HostCSR& a, const std::vector<int>* order, Sell& s, hipStream_t st) {
    const int n = a.n;
    const int ns = (n + kWave - 1) / kWave;
    std::vector<int> iperm;
    if (order) {
        iperm.assign(n, 0);
        for (int i = 0; i < n; ++i) iperm[(*order)[i]] = i;
    }
    std::vector<int> sp(ns + 1, 0);
    for (int sl = 0; sl < ns; ++sl) {
        int w = 0;
        for (int l = 0; l < kWave; ++l) {
            const int r = sl * kWave + l;
            if (r >= n) break;
            const int old = order ? (*order)[r] : r;
            w = std::max(w, a.rp[old + 1] - a.rp[old]);
        }
        sp[sl + 1] = sp[sl] + w * kWave;
    }
    const int64_t total = sp[ns];
    std::vector<int> col(total), src(total);
    for (int sl = 0; sl < ns; ++sl) {
        const int w = (sp[sl + 1] - sp[sl]) / kWave;
        for (int l = 0; l < kWave; ++l) {
            const int r = sl * kWave + l;
            const int old = r < n ? (order ? (*order)[r] : r) : -1;
            const int len = old >= 0 ? a.rp[old + 1] - a.rp[old] : 0;
            for (int k = 0; k < w; ++k) {
                const int64_t q = sp[sl] + (int64_t)k * kWave + l;
                if (k < len) {
                    const int p = a.rp[old] + k;
                    col[q] = order ? iperm[a.col[p]] : a.col[p];
                    src[q] = a.src[p];
                } else {
                    col[q] = r < n ? r : 0;
                    src[q] = -1;
                }
            }
        }
    }
    s.n = n;
    s.n_slices = ns;
    s.total = total;
    s.max_w = 0;
    for (int sl = 0; sl < ns; ++sl) s.max_w = std::max(s.max_w, (sp[sl + 1] - sp[sl]) / kWave);
    s.identity = order == nullptr;
    s.slice_ptr.upload(sp.data(), sp.size(), st);
    s.col.upload(col.data(), col.size(), st);
    s.src.upload(src.data(), src.size(), st);
    if (order) {
        s.perm.upload(order->data(), n, st);
        s.iperm.upload(iperm.data(), n, st);
    }
    s.have64 = s.have32 = false;
    OCSL_HIP(hipStreamSynchronize(st));
}

// SELL-P (see SellPView): diagonal apart, off-diagonals of each row in the
// reference's ascending column order, slices padded to an even width.
void build_sellp(const HostCSR& a, const std::vector<int>* order, int rpl, SellP& s, hipStream_t st) {
    const int n = a.n;
    const int H = kWave * rpl;  // rows per slice
    const int ns = (n + H - 1) / H;
    std::vector<int> iperm;
    if (order) {
        iperm.assign(n, 0);
        for (int i = 0; i < n; ++i) iperm[(*order)[i]] = i;
    }
    auto old_of = [&](int r) { return order ? (*order)[r] : r; };
    // element of off-diagonal k of row j of lane l in a slice starting at base
    auto slot = [&](int64_t base, int l, int j, int k) {
        return base + (int64_t)(k / 2) * 2 * H + 2 * rpl * l + 2 * j + (k & 1);
    };
    std::vector<int> sp(ns + 1, 0);
    for (int sl = 0; sl < ns; ++sl) {
        int w = 0;
        for (int q = 0; q < H; ++q) {
            const int r = sl * H + q;
            if (r >= n) break;
            const int o = old_of(r);
            w = std::max(w, a.rp[o + 1] - a.rp[o] - 1);
        }
        w += w & 1;
        sp[sl + 1] = sp[sl] + w * H;
    }
    const int64_t total = sp[ns];
    std::vector<int> col(total), src(total), dsrc(n);
    for (int sl = 0; sl < ns; ++sl) {
        const int w = (sp[sl + 1] - sp[sl]) / H;
        for (int l = 0; l < kWave; ++l)
            for (int j = 0; j < rpl; ++j) {
                const int r = sl * H + l * rpl + j;
                int k = 0;
                if (r < n) {
                    const int o = old_of(r);
                    dsrc[r] = a.src[a.diag[o]];
                    for (int p = a.rp[o]; p < a.rp[o + 1]; ++p) {
                        if (p == a.diag[o]) continue;
                        const int64_t q = slot(sp[sl], l, j, k);
                        col[q] = order ? iperm[a.col[p]] : a.col[p];
                        src[q] = a.src[p];
                        ++k;
                    }
                }
                // padding: value 0; column = the row, or for rows past the
                // last one the last row (a valid index within 16 bits)
                for (; k < w; ++k) {
                    const int64_t q = slot(sp[sl], l, j, k);
                    col[q] = r < n ? r : n - 1;
                    src[q] = -1;
                }
            }
    }
    // row window of each slice: its real rows and every stored column
    std::vector<int> wlo(ns), whi(ns);
    for (int sl = 0; sl < ns; ++sl) {
        int lo = sl * H, hi = std::min(n, (sl + 1) * H) - 1;
        for (int64_t q = sp[sl]; q < sp[sl + 1]; ++q) {
            lo = std::min(lo, col[q]);
            hi = std::max(hi, col[q]);
        }
        wlo[sl] = lo;
        whi[sl] = hi;
    }
    s.win_lo.upload(wlo.data(), wlo.size(), st);
    s.win_hi.upload(whi.data(), whi.size(), st);
    // 16-bit column offsets when every entry allows it
    bool compact = true;
    for (int sl = 0; sl < ns && compact; ++sl) {
        const int w = (sp[sl + 1] - sp[sl]) / H;
        for (int l = 0; l < kWave && compact; ++l)
            for (int j = 0; j < rpl && compact; ++j)
                for (int k = 0; k < w; ++k) {
                    const int d = col[slot(sp[sl], l, j, k)] - (sl * H + l * rpl + j);
                    if (d < -32768 || d > 32767) { compact = false; break; }
                }
    }
    s.n = n;
    s.n_slices = ns;
    s.rpl = rpl;
    s.total = total;
    s.identity = order == nullptr;
    s.compact = compact;
    s.slice_ptr.upload(sp.data(), sp.size(), st);
    if (compact) {
        // word of pair t of row j of lane l: base / 2 + t H + l rpl + j
        std::vector<unsigned> colw(total / 2);
        for (int sl = 0; sl < ns; ++sl) {
            const int np = (sp[sl + 1] - sp[sl]) / (2 * H);
            for (int t = 0; t < np; ++t)
                for (int l = 0; l < kWave; ++l)
                    for (int j = 0; j < rpl; ++j) {
                        const int64_t q = slot(sp[sl], l, j, 2 * t);
                        const int row = sl * H + l * rpl + j;
                        const unsigned lo = (unsigned)(uint16_t)(int16_t)(col[q] - row);
                        const unsigned hi = (unsigned)(uint16_t)(int16_t)(col[q + 1] - row);
                        colw[sp[sl] / 2 + (int64_t)t * H + l * rpl + j] = lo | (hi << 16);
                    }
        }
        s.colw.upload(colw.data(), colw.size(), st);
    } else {
        s.col.upload(col.data(), col.size(), st);
    }
    s.src.upload(src.data(), src.size(), st);
    s.diag_src.upload(dsrc.data(), dsrc.size(), st);
    if (order) s.perm.upload(order->data(), n, st);
    OCSL_HIP(hipStreamSynchronize(st));
}

// Dependency levels of the unit-lower (lower = true) or upper solve with the
// ILU factors (cmatrix_solver.cpp:555-571, 595-610): a row's level is one more
// than the deepest row it reads.
void build_levels(const HostCSR& a, bool lower, Levels& L, hipStream_t st) {
    const int n = a.n;
    std::vector<int> lev(n, 0);
    int nl = 0;
    if (lower) {
        for (int i = 0; i < n; ++i) {
            int l = 0;
            for (int p = a.rp[i]; p < a.diag[i]; ++p) l = std::max(l, lev[a.col[p]] + 1);
            lev[i] = l;
            nl = std::max(nl, l + 1);
        }
    } else {
        for (int i = n - 1; i >= 0; --i) {
            int l = 0;
            for (int p = a.diag[i] + 1; p < a.rp[i + 1]; ++p) l = std::max(l, lev[a.col[p]] + 1);
            lev[i] = l;
            nl = std::max(nl, l + 1);
        }
    }
    std::vector<int> ptr(nl + 1, 0), rows(n);
    for (int i = 0; i < n; ++i) ptr[lev[i] + 1]++;
    for (int l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n; ++i) rows[fill[lev[i]]++] = i;
    L.n_levels = nl;
    L.max_width = 0;
    for (int l = 0; l < nl; ++l) L.max_width = std::max(L.max_width, ptr[l + 1] - ptr[l]);
    L.ptr_host = ptr;
    L.rows.upload(rows.data(), rows.size(), st);
    L.ptr.upload(ptr.data(), ptr.size(), st);
    // the entries in level order (one position per row)
    std::vector<int> eptr(n + 1, 0), ecol, esrc, dslot(n);
    ecol.reserve(a.col.size() / 2 + 1);
    esrc.reserve(a.col.size() / 2 + 1);
    for (int k = 0; k < n; ++k) {
        const int i = rows[k];
        const int p0 = lower ? a.rp[i] : a.diag[i] + 1;
        const int p1 = lower ? a.diag[i] : a.rp[i + 1];
        for (int q = p0; q < p1; ++q) {
            ecol.push_back(a.col[q]);
            esrc.push_back(q);
        }
        eptr[k + 1] = (int)ecol.size();
        dslot[k] = a.diag[i];
    }
    L.eptr.upload(eptr.data(), eptr.size(), st);
    L.ecol.upload(ecol.data(), ecol.size(), st);
    L.esrc.upload(esrc.data(), esrc.size(), st);
    L.dslot.upload(dslot.data(), dslot.size(), st);
    // padded columns, when every row is short
    int w = 0;
    for (int k = 0; k < n; ++k) w = std::max(w, eptr[k + 1] - eptr[k]);
    L.ell_w = (w >= 1 && w <= kTriEll) ? w : 0;
    if (L.ell_w) {
        std::vector<int> cnt(n), ecw((size_t)w * n, 0), esw((size_t)w * n);
        for (int k = 0; k < n; ++k) {
            cnt[k] = eptr[k + 1] - eptr[k];
            for (int e = 0; e < w; ++e) {
                const size_t at = (size_t)e * n + k;
                if (e < cnt[k]) {
                    ecw[at] = ecol[eptr[k] + e];
                    esw[at] = esrc[eptr[k] + e];
                } else {
                    esw[at] = dslot[k];
                }
            }
        }
        L.cnt.upload(cnt.data(), cnt.size(), st);
        L.ecw.upload(ecw.data(), ecw.size(), st);
        L.esw.upload(esw.data(), esw.size(), st);
    }
    L.from64 = L.from32 = nullptr;
    L.epoch64 = L.epoch32 = -1;
    OCSL_HIP(hipStreamSynchronize(st));
}

// For entry p = (i, j): lower (j < i) sums over the row-i slots k before p,
// upper over the row-i slots k before the diagonal; a pair (k, q) exists when
// row col[k] holds column j at slot q (CCSR::get_element returns 0 otherwise).
void build_ilu_tables(const HostCSR& a, IluTables& t, hipStream_t st) {
    const int n = a.n;
    const int nnz = (int)a.col.size();
    std::vector<int> ptr(nnz + 1, 0), pk, pq, den(nnz, -1);
    pk.reserve(4 * (size_t)nnz);
    pq.reserve(4 * (size_t)nnz);
    for (int i = 0; i < n; ++i) {
        for (int p = a.rp[i]; p < a.rp[i + 1]; ++p) {
            const int j = a.col[p];
            const int kend = j < i ? p : a.diag[i];
            if (j < i) den[p] = a.diag[j];
            for (int k = a.rp[i]; k < kend; ++k) {
                const int c = a.col[k];
                const int* b = a.col.data() + a.rp[c];
                const int* e = a.col.data() + a.rp[c + 1];
                const int* f = std::lower_bound(b, e, j);
                if (f != e && *f == j) {
                    pk.push_back(k);
                    pq.push_back((int)(f - a.col.data()));
                }
            }
            ptr[p + 1] = (int)pk.size();
        }
    }
    t.n_pairs = (int64_t)pk.size();
    t.pair_ptr.upload(ptr.data(), ptr.size(), st);
    t.pair_k.upload(pk.data(), pk.size(), st);
    t.pair_q.upload(pq.data(), pq.size(), st);
    t.den.upload(den.data(), den.size(), st);
    OCSL_HIP(hipStreamSynchronize(st));
}

}  // namespace ocsl
