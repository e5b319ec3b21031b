// Structural ILU(0) on the canonical CSR pattern, factorised on the device.
//
// The reference's "ilu0" is scipy's spilu (core/preconditioner.py:49-54);
// this is the textbook ILU(0) (Saad, Iterative Methods for Sparse Linear
// Systems, Alg. 10.4, IKJ variant) that can run on a GPU: unit-lower L and
// upper U stored on A's pattern.  Rows are level-scheduled: level(i) = 1 +
// max level of the rows it depends on (j < i in row i for L and for the
// factorisation, j > i for U).  The factorisation runs one launch per level;
// the triangular solves are sync-free kernels (bicgstab.hip) that walk the
// same level lists.
//
// Ordering.  With KCFD_ORDER_COLOR the factors are those of P A P^T, where P
// sorts the rows by a first-fit colouring of the graph of A + A^T (a row's
// colour is the smallest not used by an already coloured neighbour; rows in
// ascending order; within a colour rows keep their order).  No two rows of a
// colour are coupled, so every row of colour c depends only on colours < c
// (L) or > c (U): the solves have as many levels as colours.
//
// With KCFD_ORDER_SEGMENT the colouring is of segments instead of rows: runs
// of up to kSegLen consecutive rows of A, each coupled to the one before it
// (on a structured grid in its natural numbering, pieces of a grid line).
// Segments are coloured first-fit on the graph of A + A^T contracted to
// segments.  Within a colour, segments in the order of their first rows are
// taken 64 at a time; a group's segments are sorted longest first (stable)
// and its rows numbered step-major: the first row of every segment of the
// group, then the second rows, and so on.  A segment's rows keep their
// order.  The coupling inside a segment is kept by the factorisation, which
// costs the solves up to kSegLen levels per colour and buys fewer Krylov
// iterations than the row colouring on grids (the strong couplings along a
// line survive).
#include "common.hpp"

#include <algorithm>
#include <climits>
#include <cstdlib>

namespace kcfd {

// Factorise the rows of one level.  Row i: for each k < i in its pattern (in
// column order): l_ik = a_ik / u_kk; a_ij -= l_ik u_kj for j > k in both
// patterns.  Both rows are sorted, so the match is a merge.
__global__ __launch_bounds__(kBlock) void k_ilu0_level(const int *__restrict__ rows, int count,
                                                       const int *__restrict__ ip,
                                                       const int *__restrict__ ix,
                                                       const int *__restrict__ dpos,
                                                       double *__restrict__ lu) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int i = rows[t];
    const int rs = ip[i], re = ip[i + 1], di = dpos[i];
    for (int q = rs; q < di; ++q) {
        const int k = ix[q];
        const double lik = lu[q] / lu[dpos[k]];
        lu[q] = lik;
        int w = q + 1;  // merge pointer into row i (columns > k)
        for (int r = dpos[k] + 1; r < ip[k + 1]; ++r) {
            const int j = ix[r];
            while (w < re && ix[w] < j) ++w;
            if (w >= re) break;
            if (ix[w] == j) lu[w] -= lik * lu[r];
        }
    }
}

__global__ void k_f64_to_f32(size_t n, const double *__restrict__ a, float *__restrict__ b) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) b[i] = (float)a[i];
}

__global__ void k_pivot_check(int n, const int *__restrict__ dpos, const double *__restrict__ lu,
                              int *__restrict__ bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double u = lu[dpos[i]];
    if (u == 0.0 || !(u == u)) atomicMin(bad, i);
}

// level lists: rows sorted by level (stable: increasing row id inside a
// level for L, decreasing for U), with level offsets
static void level_schedule(const std::vector<int> &ip, const std::vector<int> &ix,
                           const std::vector<int> &dp, bool upper, std::vector<int> &rows,
                           std::vector<int> &ptr) {
    const int n = (int)ip.size() - 1;
    std::vector<int> lev(n, 0);
    int maxl = -1;
    if (!upper) {
        for (int i = 0; i < n; ++i) {
            int l = 0;
            for (int q = ip[i]; q < dp[i]; ++q) l = std::max(l, lev[ix[q]] + 1);
            lev[i] = l;
            maxl = std::max(maxl, l);
        }
    } else {
        for (int i = n - 1; i >= 0; --i) {
            int l = 0;
            for (int q = dp[i] + 1; q < ip[i + 1]; ++q) l = std::max(l, lev[ix[q]] + 1);
            lev[i] = l;
            maxl = std::max(maxl, l);
        }
    }
    const int L = maxl + 1;
    ptr.assign(L + 1, 0);
    for (int i = 0; i < n; ++i) ptr[lev[i] + 1]++;
    for (int l = 0; l < L; ++l) ptr[l + 1] += ptr[l];
    rows.assign(n, 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    if (!upper) {
        for (int i = 0; i < n; ++i) rows[fill[lev[i]]++] = i;
    } else {
        for (int i = n - 1; i >= 0; --i) rows[fill[lev[i]]++] = i;
    }
}

__global__ void k_gather_vals(int nnz, const int *__restrict__ map, const double *__restrict__ a,
                              double *__restrict__ b) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nnz) b[q] = a[map[q]];
}

// first-fit colouring of A + A^T, rows visited in ascending order
static int color_rows(const std::vector<int> &ip, const std::vector<int> &ix, std::vector<int> &col) {
    const int n = (int)ip.size() - 1;
    // transpose pattern (rows j listing the i with a_ij != 0)
    std::vector<int> tp(n + 1, 0), tx(ix.size());
    for (int c : ix) tp[c + 1]++;
    for (int i = 0; i < n; ++i) tp[i + 1] += tp[i];
    {
        std::vector<int> f(tp.begin(), tp.end() - 1);
        for (int i = 0; i < n; ++i)
            for (int q = ip[i]; q < ip[i + 1]; ++q) tx[f[ix[q]]++] = i;
    }
    col.assign(n, -1);
    std::vector<int> mark;   // mark[c] == i: colour c is taken by a neighbour of i
    int ncol = 0;
    for (int i = 0; i < n; ++i) {
        auto take = [&](int j) {
            if (j == i) return;
            const int c = col[j];
            if (c >= 0) mark[c] = i;
        };
        for (int q = ip[i]; q < ip[i + 1]; ++q) take(ix[q]);
        for (int q = tp[i]; q < tp[i + 1]; ++q) take(tx[q]);
        int c = 0;
        while (c < (int)mark.size() && mark[c] == i) ++c;
        if (c == (int)mark.size()) mark.push_back(-1);
        col[i] = c;
        ncol = std::max(ncol, c + 1);
    }
    return ncol;
}

// First-fit colouring of segments (see the ordering note above): col[i] =
// colour of row i's segment, seg[i] its segment; returns the number of
// colours.
static int color_segments(const std::vector<int> &ip, const std::vector<int> &ix, std::vector<int> &col,
                          std::vector<int> &seg) {
    const int n = (int)ip.size() - 1;
    auto has = [&](int i, int j) {   // a_ij in the (sorted) pattern
        return std::binary_search(ix.begin() + ip[i], ix.begin() + ip[i + 1], j);
    };
    std::vector<int> start;
    seg.assign(n, 0);
    for (int i = 0, len = 0; i < n; ++i) {
        if (i == 0 || len == kSegLen || !(has(i, i - 1) || has(i - 1, i))) {
            start.push_back(i);
            len = 0;
        }
        seg[i] = (int)start.size() - 1;
        ++len;
    }
    const int ns = (int)start.size();
    start.push_back(n);
    // transpose pattern, for the A^T couplings
    std::vector<int> tp(n + 1, 0), tx(ix.size());
    for (int c : ix) tp[c + 1]++;
    for (int i = 0; i < n; ++i) tp[i + 1] += tp[i];
    {
        std::vector<int> f(tp.begin(), tp.end() - 1);
        for (int i = 0; i < n; ++i)
            for (int q = ip[i]; q < ip[i + 1]; ++q) tx[f[ix[q]]++] = i;
    }
    std::vector<int> scol(ns, -1), mark;
    int ncol = 0;
    for (int g = 0; g < ns; ++g) {
        auto take = [&](int j) {
            const int h = seg[j];
            if (h != g && scol[h] >= 0) mark[scol[h]] = g;
        };
        for (int i = start[g]; i < start[g + 1]; ++i) {
            for (int q = ip[i]; q < ip[i + 1]; ++q) take(ix[q]);
            for (int q = tp[i]; q < tp[i + 1]; ++q) take(tx[q]);
        }
        int c = 0;
        while (c < (int)mark.size() && mark[c] == g) ++c;
        if (c == (int)mark.size()) mark.push_back(-1);
        scol[g] = c;
        ncol = std::max(ncol, c + 1);
    }
    col.resize(n);
    for (int i = 0; i < n; ++i) col[i] = scol[seg[i]];
    return ncol;
}

void ilu_free_order(kcfd_mat *m) {
    if (m->perm) {
        hipFree(m->perm);
        hipFree(m->iperm);
        hipFree(m->f_vmap);
        hipFree(m->f_indptr);
        hipFree(m->f_indices);
        hipFree(m->f_diagpos);
        hipFree(m->f_uix);
    }
    m->perm = m->iperm = m->f_vmap = m->f_indptr = m->f_indices = m->f_diagpos = m->f_uix = nullptr;
    if (m->sell_permuted) {   // the SELL copy was in the dropped order: rebuilt at the next set-up
        hipFree(m->six);
        m->six = nullptr;
    }
    m->h_perm.clear();
    m->hf_indptr.clear();
    m->hf_indices.clear();
    m->hf_diagpos.clear();
    m->ncolors = 0;
    m->ilu_built_order = -1;
}

// factor structure for m->ilu_order (host copies + device arrays)
static int build_order(kcfd_mat *m) {
    ilu_free_order(m);
    const int n = m->n, nnz = m->nnz;
    if (m->ilu_order == KCFD_ORDER_NATURAL || n == 0) {
        m->f_indptr = m->indptr;
        m->f_indices = m->indices;
        m->f_diagpos = m->diagpos;
        m->f_uix = m->indices;
        m->hf_indptr = m->h_indptr;
        m->hf_indices = m->h_indices;
        m->hf_diagpos = m->h_diagpos;
        m->ilu_built_order = m->ilu_order;
        return KCFD_OK;
    }
    std::vector<int> col, seg;
    m->ncolors = m->ilu_order == KCFD_ORDER_SEGMENT ? color_segments(m->h_indptr, m->h_indices, col, seg)
                                                     : color_rows(m->h_indptr, m->h_indices, col);
    // counting sort by colour, stable in the row id
    std::vector<int> cp(m->ncolors + 1, 0);
    for (int c : col) cp[c + 1]++;
    for (int c = 0; c < m->ncolors; ++c) cp[c + 1] += cp[c];
    m->ncolour0 = cp[1];
    std::vector<int> &perm = m->h_perm, iperm(n);
    perm.assign(n, 0);
    for (int i = 0; i < n; ++i) perm[cp[col[i]]++] = i;
    for (int k = 0; k < n; ++k) iperm[perm[k]] = k;
    // Segment ordering: groups of 64 consecutive segments of one colour,
    // longest first within a group, numbered step-major (row k of every
    // segment of the group together), so that the lanes of the segment
    // solves read consecutive rows.  h_sg: per segment, in group order,
    // {its row at step 0, length}.
    m->h_sg.clear();
    if (m->ilu_order == KCFD_ORDER_SEGMENT) {
        std::vector<int2> segs;   // {first row of A, length}, colour-major, natural order
        for (int k = 0; k < n; ++k) {
            if (k == 0 || seg[perm[k]] != seg[perm[k - 1]]) segs.push_back(make_int2(perm[k], 0));
            segs.back().y++;
        }
        std::vector<int> np;
        np.reserve(n);
        const int nseg = (int)segs.size();
        for (int s0 = 0; s0 < nseg;) {
            const int c = col[segs[s0].x];
            int s1 = s0;
            while (s1 < nseg && s1 - s0 < 64 && col[segs[s1].x] == c) ++s1;
            std::stable_sort(segs.begin() + s0, segs.begin() + s1,
                             [](const int2 &a, const int2 &b) { return a.y > b.y; });
            const int base = (int)np.size();
            for (int k = 0; k < segs[s0].y; ++k)
                for (int t = s0; t < s1 && segs[t].y > k; ++t) np.push_back(segs[t].x + k);
            for (int t = s0; t < s1; ++t) m->h_sg.push_back(make_int2(base + (t - s0), segs[t].y));
            s0 = s1;
        }
        perm = np;
        for (int k = 0; k < n; ++k) iperm[perm[k]] = k;
    }
    // P A P^T: row k = row perm[k] of A, columns renumbered and re-sorted
    std::vector<int> &fp = m->hf_indptr, &fx = m->hf_indices, &fd = m->hf_diagpos;
    std::vector<int> vmap(nnz), uix(nnz);
    fp.assign(n + 1, 0);
    fx.assign(nnz, 0);
    fd.assign(n, -1);
    std::vector<std::pair<int, int>> row;
    for (int k = 0; k < n; ++k) {
        const int r = perm[k];
        row.clear();
        for (int q = m->h_indptr[r]; q < m->h_indptr[r + 1]; ++q) row.emplace_back(iperm[m->h_indices[q]], q);
        std::sort(row.begin(), row.end());
        fp[k + 1] = fp[k] + (int)row.size();
        for (size_t t = 0; t < row.size(); ++t) {
            const int q = fp[k] + (int)t;
            fx[q] = row[t].first;
            vmap[q] = row[t].second;
            uix[q] = m->h_indices[row[t].second];
            if (row[t].first == k) fd[k] = q;
        }
    }
    KCFD_HIP(hipMalloc(&m->perm, sizeof(int) * n));
    KCFD_HIP(hipMalloc(&m->iperm, sizeof(int) * n));
    KCFD_HIP(hipMalloc(&m->f_vmap, sizeof(int) * std::max(nnz, 1)));
    KCFD_HIP(hipMalloc(&m->f_indptr, sizeof(int) * (n + 1)));
    KCFD_HIP(hipMalloc(&m->f_indices, sizeof(int) * std::max(nnz, 1)));
    KCFD_HIP(hipMalloc(&m->f_diagpos, sizeof(int) * n));
    KCFD_HIP(hipMalloc(&m->f_uix, sizeof(int) * std::max(nnz, 1)));
    // on the library stream (non-blocking: the null stream is not ordered
    // with the kernels that read these); the host vectors are locals, so the
    // copies finish before return
    hipStream_t s = m->ctx->stream;
    KCFD_HIP(hipMemcpyAsync(m->perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    KCFD_HIP(hipMemcpyAsync(m->iperm, iperm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    KCFD_HIP(hipMemcpyAsync(m->f_indptr, fp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    KCFD_HIP(hipMemcpyAsync(m->f_diagpos, fd.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    if (nnz) {
        KCFD_HIP(hipMemcpyAsync(m->f_vmap, vmap.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, s));
        KCFD_HIP(hipMemcpyAsync(m->f_indices, fx.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, s));
        KCFD_HIP(hipMemcpyAsync(m->f_uix, uix.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, s));
    }
    KCFD_HIP(hipStreamSynchronize(s));
    m->ilu_built_order = m->ilu_order;
    return KCFD_OK;
}

// Chunked column-major layout of the L and U parts of the factor rows (see
// kcfd_mat::chl_off).  Chunks follow the level lists: per level, runs of up
// to 64 consecutive list entries.
static void chunk_layout(const kcfd_mat *m, const std::vector<int> &rows, const std::vector<int> &lvl,
                         bool upper, std::vector<int> &ch_ptr, std::vector<int> &off,
                         std::vector<int> &col, std::vector<int> &map) {
    const std::vector<int> &ip = m->hf_indptr, &ix = m->hf_indices, &dp = m->hf_diagpos;
    const int L = (int)lvl.size() - 1;
    ch_ptr.assign(1, 0);
    off.assign(1, 0);
    col.clear();
    map.clear();
    for (int l = 0; l < L; ++l) {
        for (int s0 = lvl[l]; s0 < lvl[l + 1]; s0 += 64) {
            const int cnt = std::min(64, lvl[l + 1] - s0);
            int w = 0;
            for (int t = 0; t < cnt; ++t) {
                const int i = rows[s0 + t];
                w = std::max(w, upper ? ip[i + 1] - dp[i] - 1 : dp[i] - ip[i]);
            }
            const size_t base = col.size();
            col.resize(base + (size_t)64 * w, -1);
            map.resize(base + (size_t)64 * w, -1);
            for (int t = 0; t < cnt; ++t) {
                const int i = rows[s0 + t];
                const int q0 = upper ? dp[i] + 1 : ip[i], q1 = upper ? ip[i + 1] : dp[i];
                for (int q = q0; q < q1; ++q) {
                    const size_t e = base + (size_t)64 * (q - q0) + t;
                    col[e] = ix[q];
                    map[e] = q;
                }
            }
            off.push_back((int)col.size());
        }
        ch_ptr.push_back((int)off.size() - 1);
    }
}

// Segment layout of a two-colour segment ordering (kcfd_mat::seg2): the
// in-segment entries of each row (sgl: the row before, sgu: the row after)
// and, per group of up to 64 segments of one colour and per step k, the
// entries of row k of each segment outside its segment -- colour 1: its L
// entries (loff/lcol/lmap), colour 0: its U entries (uoff/ucol/umap) -- in
// CSR order, column-major over the group's lanes.  False when some row has
// an entry outside its segment to a row of its own colour, which the
// segment solves cannot take (the level layout serves such a matrix).
static bool seg_layout(kcfd_mat *m, std::vector<int> &loff, std::vector<int> &lcol, std::vector<int> &lmap,
                       std::vector<int> &uoff, std::vector<int> &ucol, std::vector<int> &umap,
                       std::vector<int> &sgl, std::vector<int> &sgu) {
    const std::vector<int> &ip = m->hf_indptr, &ix = m->hf_indices, &dp = m->hf_diagpos;
    const int n = m->n, n0 = m->ncolour0;
    const std::vector<int2> &sg = m->h_sg;
    sgl.assign(n, -1);
    sgu.assign(n, -1);
    int nseg0 = 0;
    for (const int2 &g : sg) nseg0 += g.x < n0;
    m->h_sg_steps.clear();
    loff.assign(1, 0);
    uoff.assign(1, 0);
    lcol.clear(); lmap.clear(); ucol.clear(); umap.clear();
    for (int c = 0; c < 2; ++c) {
        std::vector<int> &grp = m->h_sgg[c], &off = c ? loff : uoff, &col = c ? lcol : ucol,
                         &map = c ? lmap : umap;
        grp.clear();
        const int s0 = c ? nseg0 : 0, s1 = c ? (int)sg.size() : nseg0;
        for (int f = s0; f < s1; f += 64) {
            const int cnt = std::min(64, s1 - f);
            const int maxlen = sg[f].y;   // the longest segment leads its group
            // row offsets of the steps: row k of lane t is sg[f + t].x + roff[k]
            std::vector<int> roff(maxlen + 1, 0);
            for (int k = 0; k < maxlen; ++k) {
                int ck = 0;
                for (int t = 0; t < cnt; ++t) ck += sg[f + t].y > k;
                roff[k + 1] = roff[k] + ck;
            }
            // structure: inside a segment only the rows before and after,
            // outside it only rows of the other colour
            for (int t = 0; t < cnt; ++t) {
                const int2 g = sg[f + t];
                if (g.y < 1 || g.y > kSegLen || g.y > maxlen || (g.x >= n0) != (c == 1)) return false;
                for (int k = 0; k < g.y; ++k) {
                    const int i = g.x + roff[k];
                    if ((i >= n0) != (c == 1)) return false;
                    const int prev = k > 0 ? g.x + roff[k - 1] : -1, next = k + 1 < g.y ? g.x + roff[k + 1] : -1;
                    for (int q = ip[i]; q < ip[i + 1]; ++q) {
                        const int j = ix[q];
                        if (j == i) continue;
                        if (j == prev) { sgl[i] = q; continue; }
                        if (j == next) { sgu[i] = q; continue; }
                        if ((j >= n0) == (c == 1)) return false;   // outside: the other colour only
                    }
                }
            }
            grp.insert(grp.end(), {f, cnt, maxlen, (int)m->h_sg_steps.size()});
            for (int k = 0; k < maxlen; ++k) {
                // colour 1: the L part without the row before; colour 0: the
                // U part without the row after
                auto outside = [&](int i, int q) { return q != sgl[i] && q != sgu[i]; };
                int w = 0;
                for (int t = 0; t < cnt; ++t) {
                    if (k >= sg[f + t].y) continue;
                    const int i = sg[f + t].x + roff[k];
                    const int q0 = c ? ip[i] : dp[i] + 1, q1 = c ? dp[i] : ip[i + 1];
                    int e = 0;
                    for (int q = q0; q < q1; ++q) e += outside(i, q);
                    w = std::max(w, e);
                }
                if (w > kSegW) return false;   // the segment kernel holds at most kSegW per row
                const size_t base = col.size();
                col.resize(base + (size_t)64 * w, -1);
                map.resize(base + (size_t)64 * w, -1);
                for (int t = 0; t < cnt; ++t) {
                    if (k >= sg[f + t].y) continue;
                    const int i = sg[f + t].x + roff[k];
                    const int q0 = c ? ip[i] : dp[i] + 1, q1 = c ? dp[i] : ip[i + 1];
                    int e = 0;
                    for (int q = q0; q < q1; ++q)
                        if (outside(i, q)) {
                            col[base + (size_t)64 * e + t] = ix[q];
                            map[base + (size_t)64 * e + t] = q;
                            ++e;
                        }
                }
                m->h_sg_steps.push_back(make_int4((int)base, w, roff[k], 0));
                off.push_back((int)col.size());
            }
        }
    }
    return true;
}

static int upload_ints(const std::vector<int> &h, int **d, hipStream_t s) {
    KCFD_HIP(hipMalloc(d, sizeof(int) * std::max<size_t>(h.size(), 1)));
    if (!h.empty()) KCFD_HIP(hipMemcpyAsync(*d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, s));
    return KCFD_OK;
}

template <typename T>
static int upload_vec(const std::vector<T> &h, T **d, hipStream_t s) {
    KCFD_HIP(hipMalloc(d, sizeof(T) * std::max<size_t>(h.size(), 1)));
    if (!h.empty()) KCFD_HIP(hipMemcpyAsync(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
    return KCFD_OK;
}

// Narrow-index form of a chunk layout: per chunk its smallest column as the
// base and each entry's column minus it, padding (column -1) as kSellDiag.
// False when a chunk's columns span more than kSellSpan.
static bool chunk_offsets16(const std::vector<int> &off, const std::vector<int> &col,
                            std::vector<unsigned short> &o16, std::vector<int> &base) {
    const size_t nch = off.size() - 1;
    o16.assign(col.size(), (unsigned short)kSellDiag);
    base.assign(nch, 0);
    for (size_t c = 0; c < nch; ++c) {
        int lo = INT_MAX, hi = INT_MIN;
        for (int e = off[c]; e < off[c + 1]; ++e)
            if (col[e] >= 0) {
                lo = std::min(lo, col[e]);
                hi = std::max(hi, col[e]);
            }
        if (lo == INT_MAX) continue;
        if (hi - lo > kSellSpan) return false;
        base[c] = lo;
        for (int e = off[c]; e < off[c + 1]; ++e)
            if (col[e] >= 0) o16[e] = (unsigned short)(col[e] - lo);
    }
    return true;
}

// The solvers' chunk records carry a chunk's width in 13 bits.
static int check_chunk_widths(const std::vector<int> &off) {
    for (size_t c = 0; c + 1 < off.size(); ++c)
        if (((off[c + 1] - off[c]) >> 6) > 8191) {
            set_error("ILU(0): a row has more than 8191 entries in its L or U part");
            return KCFD_EARG;
        }
    return KCFD_OK;
}

// chunk-layout values: FP64, and (MIXED) the packed {column, FP32 value}
// entries; with col null the FP32 copy goes to v32 (u_ii by row)
__global__ void k_chunk_vals(size_t ne, const int *__restrict__ map, const int *__restrict__ col,
                             const double *__restrict__ lu, double *__restrict__ v64,
                             int2 *__restrict__ pk32, float *__restrict__ v32) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int q = map[e];
    const double v = q >= 0 ? lu[q] : 0.0;
    v64[e] = v;
    if (pk32) pk32[e] = make_int2(col[e], __float_as_int((float)v));
    if (v32) v32[e] = (float)v;
}

// Colour-0 pivots of the folded mixed BiCGSTAB product, rows i < n0 of
// P A P^T (two colours: these rows have no L entries, so their factor row
// is A's own row): piv_i = a_ii + sum_{j != i} (a_ij - fl32(a_ij)).  The FP32
// U entries of the row are fl32(a_ij), so the row [piv_i, fl32(a_ij)] sums to
// A's row sum up to FP64 rounding, as the difference-form SELL copy does.
__global__ void k_colour0_pivots(int n0, const int *__restrict__ ip, const int *__restrict__ dpos,
                                 const double *__restrict__ lu, double *__restrict__ piv) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n0) return;
    double d = lu[dpos[i]];
    for (int q = ip[i]; q < ip[i + 1]; ++q)
        if (q != dpos[i]) d += lu[q] - (double)(float)lu[q];
    piv[i] = d;
}

// Ordering, level schedule and chunked solve layout: they depend on the
// structure only and are built once per ordering.
int ilu0_structure(kcfd_mat *m) {
    const int n = m->n;
    hipStream_t s = m->ctx->stream;
    for (int i = 0; i < n; ++i) {
        if (m->h_diagpos[i] < 0) {
            set_error("ILU(0): row %d has no diagonal entry", i);
            return KCFD_EPIVOT;
        }
    }
    if (!m->lvl_rows_l || m->ilu_built_order != m->ilu_order) {
        mat_free_ilu_structure(m);
        KCFD_CHECK(build_order(m));
        std::vector<int> rl, ru;
        level_schedule(m->hf_indptr, m->hf_indices, m->hf_diagpos, false, rl, m->h_lvl_ptr_l);
        level_schedule(m->hf_indptr, m->hf_indices, m->hf_diagpos, true, ru, m->h_lvl_ptr_u);
        m->nlev_l = (int)m->h_lvl_ptr_l.size() - 1;
        m->nlev_u = (int)m->h_lvl_ptr_u.size() - 1;
        KCFD_HIP(hipMalloc(&m->lvl_rows_l, sizeof(int) * std::max(n, 1)));
        KCFD_HIP(hipMalloc(&m->lvl_rows_u, sizeof(int) * std::max(n, 1)));
        if (n) {
            KCFD_HIP(hipMemcpyAsync(m->lvl_rows_l, rl.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
            KCFD_HIP(hipMemcpyAsync(m->lvl_rows_u, ru.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
            KCFD_HIP(hipStreamSynchronize(s));   // rl, ru are rewritten below
        }
        // chunked L / U layout of the solves: per level of the level lists,
        // or, with two colours, the colour-1 rows (L) and colour-0 rows (U)
        m->two_colour = m->perm && m->ncolors == 2 && m->ilu_order == KCFD_ORDER_COLOR;
        // segment ordering with two colours: the segment layout, when the
        // factor structure allows it
        std::vector<int> sl_off, sl_col, sl_map, su_off, su_col, su_map, sgl, sgu;
        m->seg2 = m->ilu_order == KCFD_ORDER_SEGMENT && m->perm && m->ncolors == 2 &&
                  seg_layout(m, sl_off, sl_col, sl_map, su_off, su_col, su_map, sgl, sgu);
        if (m->two_colour) {
            const int n0 = m->ncolour0;
            m->nnz_colour0 = m->hf_indptr[n0] - m->hf_indptr[0];
            rl.resize(n - n0);
            for (int k = n0; k < n; ++k) rl[k - n0] = k;
            ru.resize(n0);
            for (int k = 0; k < n0; ++k) ru[k] = k;
        }
        const std::vector<int> lvl_l = m->two_colour ? std::vector<int>{0, n - m->ncolour0} : m->h_lvl_ptr_l;
        const std::vector<int> lvl_u = m->two_colour ? std::vector<int>{0, m->ncolour0} : m->h_lvl_ptr_u;
        {
            // the two-colour layouts also get their narrow-index form (for
            // the mixed solver) when both fit (KCFD_SELL16=0: never)
            const char *env = std::getenv("KCFD_SELL16");
            bool narrow = m->two_colour && !(env && env[0] == '0');
            std::vector<int> off, col, map, lbase, ubase;
            std::vector<unsigned short> lo16, uo16;
            if (m->seg2) {
                off = sl_off; col = sl_col; map = sl_map;
                m->h_chl_ptr = {0, (int)off.size() - 1};
            } else {
                chunk_layout(m, rl, lvl_l, false, m->h_chl_ptr, off, col, map);
            }
            KCFD_CHECK(check_chunk_widths(off));
            m->nch_l = (int)off.size() - 1;
            m->nent_l = (int)col.size();
            m->h_chl_off = off;
            if (narrow) narrow = chunk_offsets16(off, col, lo16, lbase);
            KCFD_CHECK(upload_ints(off, &m->chl_off, s));
            KCFD_CHECK(upload_ints(col, &m->chl_col, s));
            KCFD_CHECK(upload_ints(map, &m->chl_map, s));
            KCFD_HIP(hipStreamSynchronize(s));
            if (m->seg2) {
                off = su_off; col = su_col; map = su_map;
                m->h_chu_ptr = {0, (int)off.size() - 1};
            } else {
                chunk_layout(m, ru, lvl_u, true, m->h_chu_ptr, off, col, map);
            }
            KCFD_CHECK(check_chunk_widths(off));
            m->nch_u = (int)off.size() - 1;
            m->nent_u = (int)col.size();
            m->h_chu_off = off;
            if (narrow) narrow = chunk_offsets16(off, col, uo16, ubase);
            KCFD_CHECK(upload_ints(off, &m->chu_off, s));
            KCFD_CHECK(upload_ints(col, &m->chu_col, s));
            KCFD_CHECK(upload_ints(map, &m->chu_map, s));
            if (narrow) {
                m->h_chl_base = lbase;
                m->h_chu_base = ubase;
                KCFD_CHECK(upload_vec(lo16, &m->chl_o16, s));
                KCFD_CHECK(upload_vec(uo16, &m->chu_o16, s));
                KCFD_CHECK(upload_vec(lbase, &m->chl_base, s));
                KCFD_CHECK(upload_vec(ubase, &m->chu_base, s));
            }
            m->ch16 = narrow;
            if (m->seg2) {
                std::vector<int> sgi(2 * m->h_sg.size()), sti(4 * m->h_sg_steps.size());
                for (size_t g = 0; g < m->h_sg.size(); ++g) { sgi[2 * g] = m->h_sg[g].x; sgi[2 * g + 1] = m->h_sg[g].y; }
                for (size_t t = 0; t < m->h_sg_steps.size(); ++t) {
                    sti[4 * t] = m->h_sg_steps[t].x;
                    sti[4 * t + 1] = m->h_sg_steps[t].y;
                    sti[4 * t + 2] = m->h_sg_steps[t].z;
                    sti[4 * t + 3] = 0;
                }
                KCFD_CHECK(upload_ints(sgi, (int **)&m->sg, s));
                KCFD_CHECK(upload_ints(sti, (int **)&m->sg_steps, s));
                KCFD_CHECK(upload_ints(sgl, &m->sgl_map, s));
                KCFD_CHECK(upload_ints(sgu, &m->sgu_map, s));
            }
            KCFD_HIP(hipStreamSynchronize(s));   // the host vectors are locals
        }
    }
    return KCFD_OK;
}

// Factor values (after ilu0_structure): ILU(0) of P A P^T level by level,
// FP32 copies, the chunked solve layout's values and a pivot check.
int ilu0_setup(kcfd_mat *m, bool want32) {
    const int n = m->n;
    hipStream_t s = m->ctx->stream;
    KCFD_CHECK(ilu0_structure(m));
    const std::vector<int> &pl = m->h_lvl_ptr_l;
    if (!m->lu) KCFD_HIP(hipMalloc(&m->lu, sizeof(double) * std::max(m->nnz, 1)));
    if (m->nnz && !m->perm)
        KCFD_HIP(hipMemcpyAsync(m->lu, m->val, sizeof(double) * m->nnz, hipMemcpyDeviceToDevice, s));
    else if (m->nnz)
        hipLaunchKernelGGL(k_gather_vals, dim3((m->nnz + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                           m->nnz, m->f_vmap, m->val, m->lu);
    for (int l = 0; l < m->nlev_l; ++l) {
        const int cnt = pl[l + 1] - pl[l];
        hipLaunchKernelGGL(k_ilu0_level, dim3((cnt + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                           m->lvl_rows_l + pl[l], cnt, m->f_indptr, m->f_indices, m->f_diagpos, m->lu);
    }
    KCFD_HIP(hipGetLastError());
    if (want32 && m->nnz) {
        if (!m->lu32) KCFD_HIP(hipMalloc(&m->lu32, sizeof(float) * m->nnz));
        size_t nb = ((size_t)m->nnz + kBlock - 1) / kBlock;
        if (nb > 65536) nb = 65536;
        hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)nb), dim3(kBlock), 0, s, (size_t)m->nnz,
                           m->lu, m->lu32);
        KCFD_HIP(hipGetLastError());
    }
    // factor values in the chunked solve layout, u_ii by row
    {
        const size_t nl = (size_t)m->nent_l, nu = (size_t)m->nent_u;
        if (!m->chl_v64) {
            KCFD_HIP(hipMalloc(&m->chl_v64, sizeof(double) * std::max<size_t>(nl, 1)));
            KCFD_HIP(hipMalloc(&m->chu_v64, sizeof(double) * std::max<size_t>(nu, 1)));
            KCFD_HIP(hipMalloc(&m->ud64, sizeof(double) * std::max(n, 1)));
        }
        const bool narrow = want32 && m->ch16;   // FP32 values beside the 16-bit offsets
        if (want32 && !m->ud32) {
            if (narrow) {
                KCFD_HIP(hipMalloc(&m->chl_v32, sizeof(float) * std::max<size_t>(nl, 1)));
                KCFD_HIP(hipMalloc(&m->chu_v32, sizeof(float) * std::max<size_t>(nu, 1)));
            } else {
                KCFD_HIP(hipMalloc(&m->chl_pk32, sizeof(int2) * std::max<size_t>(nl, 1)));
                KCFD_HIP(hipMalloc(&m->chu_pk32, sizeof(int2) * std::max<size_t>(nu, 1)));
            }
            KCFD_HIP(hipMalloc(&m->ud32, sizeof(float) * std::max(n, 1)));
        }
        int2 *l32 = want32 && !narrow ? m->chl_pk32 : nullptr, *u32 = want32 && !narrow ? m->chu_pk32 : nullptr;
        float *lv32 = narrow ? m->chl_v32 : nullptr, *uv32 = narrow ? m->chu_v32 : nullptr;
        if (nl)
            hipLaunchKernelGGL(k_chunk_vals, dim3((unsigned)((nl + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                               nl, m->chl_map, m->chl_col, m->lu, m->chl_v64, l32, lv32);
        if (nu)
            hipLaunchKernelGGL(k_chunk_vals, dim3((unsigned)((nu + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                               nu, m->chu_map, m->chu_col, m->lu, m->chu_v64, u32, uv32);
        if (n)
            hipLaunchKernelGGL(k_chunk_vals, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, (size_t)n,
                               m->f_diagpos, nullptr, m->lu, m->ud64, nullptr, want32 ? m->ud32 : nullptr);
        if (m->seg2 && n) {   // the in-segment entries by row
            if (!m->sgl64) {
                KCFD_HIP(hipMalloc(&m->sgl64, sizeof(double) * n));
                KCFD_HIP(hipMalloc(&m->sgu64, sizeof(double) * n));
            }
            if (want32 && !m->sgl32) {
                KCFD_HIP(hipMalloc(&m->sgl32, sizeof(float) * n));
                KCFD_HIP(hipMalloc(&m->sgu32, sizeof(float) * n));
            }
            hipLaunchKernelGGL(k_chunk_vals, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, (size_t)n,
                               m->sgl_map, nullptr, m->lu, m->sgl64, nullptr, want32 ? m->sgl32 : nullptr);
            hipLaunchKernelGGL(k_chunk_vals, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, (size_t)n,
                               m->sgu_map, nullptr, m->lu, m->sgu64, nullptr, want32 ? m->sgu32 : nullptr);
        }
        if (want32 && m->two_colour && m->ncolour0 > 0) {
            if (!m->c0piv) KCFD_HIP(hipMalloc(&m->c0piv, sizeof(double) * m->ncolour0));
            hipLaunchKernelGGL(k_colour0_pivots, dim3((m->ncolour0 + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                               m->ncolour0, m->f_indptr, m->f_diagpos, m->lu, m->c0piv);
        }
        KCFD_HIP(hipGetLastError());
    }
    // pivot check: first row with a zero or NaN u_ii (or n if none)
    KCFD_CHECK(ensure_scratch(m->ctx, sizeof(int)));
    int *bad = (int *)m->ctx->scratch;
    const int init = n;
    KCFD_HIP(hipMemcpyAsync(bad, &init, sizeof(int), hipMemcpyHostToDevice, s));
    if (n)
        hipLaunchKernelGGL(k_pivot_check, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n,
                           m->f_diagpos, m->lu, bad);
    int first = n;
    KCFD_HIP(hipMemcpyAsync(&first, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    KCFD_HIP(hipStreamSynchronize(s));
    if (first < n) {
        set_error("ILU(0): zero or NaN pivot at row %d", m->perm ? m->h_perm[first] : first);
        return KCFD_EPIVOT;
    }
    return KCFD_OK;
}

// One level of the ILU(0) apply, rows i (positions in P A P^T) of that level.
//   lower: t_i = z_{perm i} - sum_{j<i} l_ij t_j          (t in P-order)
//   upper: w_{perm i} = (t_i - sum_{j>i} u_ij w_{perm j}) / u_ii  (w in A's order)
// perm null is the identity; uix[q] = perm[ix[q]] (with perm null and
// uix = ix, z and w are both in P-order).  The launch is a no-op once
// *skip < col (a caller's stop flag; skip null: never).
__global__ __launch_bounds__(kBlock) void k_trsv_level(const int *__restrict__ rows, int count,
                                                       const int *__restrict__ ip,
                                                       const int *__restrict__ ix,
                                                       const int *__restrict__ uix,
                                                       const int *__restrict__ dpos,
                                                       const int *__restrict__ perm,
                                                       const double *__restrict__ lu,
                                                       const double *z, double *t, double *w,
                                                       int upper, const int *skip, int col) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count || (skip && *skip < col)) return;
    const int i = rows[k];
    const int r = perm ? perm[i] : i;
    const int dp = dpos[i];
    if (!upper) {
        double acc = z[r];
        for (int q = ip[i]; q < dp; ++q) acc -= lu[q] * t[ix[q]];
        t[i] = acc;
    } else {
        double acc = t[i];
        for (int q = dp + 1; q < ip[i + 1]; ++q) acc -= lu[q] * w[uix[q]];
        w[r] = acc / lu[dp];
    }
}

__global__ void k_jacobi_apply(int n, const double *__restrict__ dinv, const double *__restrict__ x,
                               double *__restrict__ y, const int *skip, int col) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(skip && *skip < col)) y[i] = dinv[i] * x[i];
}

__global__ void k_copy(int n, const double *__restrict__ x, double *__restrict__ y, const int *skip,
                       int col) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(skip && *skip < col)) y[i] = x[i];
}

// y = M^{-1} x on the library stream, no host synchronisation.  porder: x
// and y are in the ILU(0) row order (positions of P A P^T) rather than A's;
// t: n doubles of scratch (ILU(0)).
int precond_apply_async(kcfd_mat *m, const double *x, double *y, double *t, bool porder,
                        const int *skip, int col) {
    hipStream_t s = m->ctx->stream;
    const int n = m->n;
    if (!n) return KCFD_OK;
    const unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
    if (m->pc_kind == KCFD_PC_JACOBI) {
        hipLaunchKernelGGL(k_jacobi_apply, dim3(nb), dim3(kBlock), 0, s, n, m->dinv, x, y, skip, col);
    } else if (m->pc_kind == KCFD_PC_ILU0) {
        // level by level: t = L^{-1} P x, then y = P^T U^{-1} t
        const int *perm = porder ? nullptr : m->perm;
        const int *uix = porder ? m->f_indices : m->f_uix;
        for (int l = 0; l < m->nlev_l; ++l) {
            const int c = m->h_lvl_ptr_l[l + 1] - m->h_lvl_ptr_l[l];
            hipLaunchKernelGGL(k_trsv_level, dim3((c + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                               m->lvl_rows_l + m->h_lvl_ptr_l[l], c, m->f_indptr, m->f_indices, uix,
                               m->f_diagpos, perm, m->lu, x, t, y, 0, skip, col);
        }
        for (int l = 0; l < m->nlev_u; ++l) {
            const int c = m->h_lvl_ptr_u[l + 1] - m->h_lvl_ptr_u[l];
            hipLaunchKernelGGL(k_trsv_level, dim3((c + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                               m->lvl_rows_u + m->h_lvl_ptr_u[l], c, m->f_indptr, m->f_indices, uix,
                               m->f_diagpos, perm, m->lu, x, t, y, 1, skip, col);
        }
    } else {
        hipLaunchKernelGGL(k_copy, dim3(nb), dim3(kBlock), 0, s, n, x, y, skip, col);
    }
    KCFD_HIP(hipGetLastError());
    return KCFD_OK;
}

}  // namespace kcfd

using namespace kcfd;

extern "C" int kcfd_precond_apply(kcfd_mat *m, const double *x, double *y) {
    KCFD_ARG(m && (m->n == 0 || (x && y)), "null argument");
    KCFD_ARG(x != y, "x and y must not alias");
    std::lock_guard<std::recursive_mutex> g(m->ctx->mu);
    KCFD_HIP(hipSetDevice(m->ctx->device));
    const int n = m->n;
    if (!n) return KCFD_OK;
    double *t = nullptr;
    if (m->pc_kind == KCFD_PC_ILU0) {
        KCFD_CHECK(ensure_scratch(m->ctx, sizeof(double) * n));
        t = (double *)m->ctx->scratch;
    }
    KCFD_CHECK(precond_apply_async(m, x, y, t, false, nullptr, 0));
    KCFD_HIP(hipStreamSynchronize(m->ctx->stream));
    return KCFD_OK;
}

extern "C" int kcfd_ilu0_get(kcfd_mat *m, double *lu, int *levels_l, int *levels_u) {
    KCFD_ARG(m != nullptr, "null matrix");
    std::lock_guard<std::recursive_mutex> g(m->ctx->mu);
    if (!m->lu) {
        set_error("ILU(0) factors not set up");
        return KCFD_ESTATE;
    }
    if (lu && m->nnz) {
        KCFD_HIP(hipMemcpyAsync(lu, m->lu, sizeof(double) * m->nnz, hipMemcpyDeviceToHost,
                                m->ctx->stream));
        KCFD_HIP(hipStreamSynchronize(m->ctx->stream));
    }
    if (levels_l) *levels_l = m->nlev_l;
    if (levels_u) *levels_u = m->nlev_u;
    return KCFD_OK;
}

extern "C" int kcfd_precond_set_order(kcfd_mat *m, int order) {
    KCFD_ARG(m != nullptr, "null matrix");
    KCFD_ARG(order == KCFD_ORDER_NATURAL || order == KCFD_ORDER_COLOR || order == KCFD_ORDER_SEGMENT,
             "unknown ILU(0) ordering");
    std::lock_guard<std::recursive_mutex> g(m->ctx->mu);
    m->ilu_order = order;
    return KCFD_OK;
}

// device copy of n ints (null stays null)
static int dup_ints(const int *src, size_t n, int **dst, hipStream_t s) {
    *dst = nullptr;
    if (!src) return KCFD_OK;
    KCFD_HIP(hipMalloc(dst, sizeof(int) * std::max<size_t>(n, 1)));
    if (n) KCFD_HIP(hipMemcpyAsync(*dst, src, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
    return KCFD_OK;
}

extern "C" int kcfd_mat_copy_order(kcfd_mat *dst, const kcfd_mat *src) {
    KCFD_ARG(dst && src, "null matrix");
    KCFD_ARG(dst->ctx == src->ctx, "matrices of different contexts");
    KCFD_ARG(src->lvl_rows_l != nullptr, "source has no ILU(0) structure (set it up first)");
    KCFD_ARG(dst->n == src->n && dst->nnz == src->nnz && dst->h_indptr == src->h_indptr &&
                 dst->h_indices == src->h_indices,
             "matrices differ in structure");
    std::lock_guard<std::recursive_mutex> g(dst->ctx->mu);
    KCFD_HIP(hipSetDevice(dst->ctx->device));
    KCFD_HIP(hipStreamSynchronize(dst->ctx->stream));
    if (dst->solver) {
        kcfd_batch_destroy(dst->solver);
        dst->solver = nullptr;
    }
    mat_free_ilu_structure(dst);
    const size_t n = (size_t)dst->n, nnz = (size_t)dst->nnz;
    if (src->perm) {
        KCFD_CHECK(dup_ints(src->perm, n, &dst->perm, dst->ctx->stream));
        KCFD_CHECK(dup_ints(src->iperm, n, &dst->iperm, dst->ctx->stream));
        KCFD_CHECK(dup_ints(src->f_vmap, nnz, &dst->f_vmap, dst->ctx->stream));
        KCFD_CHECK(dup_ints(src->f_indptr, n + 1, &dst->f_indptr, dst->ctx->stream));
        KCFD_CHECK(dup_ints(src->f_indices, nnz, &dst->f_indices, dst->ctx->stream));
        KCFD_CHECK(dup_ints(src->f_diagpos, n, &dst->f_diagpos, dst->ctx->stream));
        KCFD_CHECK(dup_ints(src->f_uix, nnz, &dst->f_uix, dst->ctx->stream));
        dst->h_perm = src->h_perm;
    } else {
        dst->f_indptr = dst->indptr;
        dst->f_indices = dst->indices;
        dst->f_diagpos = dst->diagpos;
        dst->f_uix = dst->indices;
    }
    KCFD_CHECK(dup_ints(src->lvl_rows_l, n, &dst->lvl_rows_l, dst->ctx->stream));
    KCFD_CHECK(dup_ints(src->lvl_rows_u, n, &dst->lvl_rows_u, dst->ctx->stream));
    KCFD_CHECK(dup_ints(src->chl_off, (size_t)src->nch_l + 1, &dst->chl_off, dst->ctx->stream));
    KCFD_CHECK(dup_ints(src->chu_off, (size_t)src->nch_u + 1, &dst->chu_off, dst->ctx->stream));
    KCFD_CHECK(dup_ints(src->chl_col, (size_t)src->nent_l, &dst->chl_col, dst->ctx->stream));
    KCFD_CHECK(dup_ints(src->chu_col, (size_t)src->nent_u, &dst->chu_col, dst->ctx->stream));
    KCFD_CHECK(dup_ints(src->chl_map, (size_t)src->nent_l, &dst->chl_map, dst->ctx->stream));
    dst->two_colour = src->two_colour;
    dst->seg2 = src->seg2;
    if (src->seg2) {
        KCFD_CHECK(dup_ints((const int *)src->sg, 2 * src->h_sg.size(), (int **)&dst->sg, dst->ctx->stream));
        KCFD_CHECK(dup_ints((const int *)src->sg_steps, 4 * src->h_sg_steps.size(), (int **)&dst->sg_steps,
                            dst->ctx->stream));
        KCFD_CHECK(dup_ints(src->sgl_map, n, &dst->sgl_map, dst->ctx->stream));
        KCFD_CHECK(dup_ints(src->sgu_map, n, &dst->sgu_map, dst->ctx->stream));
        dst->h_sg = src->h_sg;
        dst->h_sg_steps = src->h_sg_steps;
        dst->h_sgg[0] = src->h_sgg[0];
        dst->h_sgg[1] = src->h_sgg[1];
    }
    dst->ncolour0 = src->ncolour0;
    dst->nnz_colour0 = src->nnz_colour0;
    if (src->ch16) {
        KCFD_HIP(hipMalloc(&dst->chl_o16, sizeof(unsigned short) * std::max(src->nent_l, 1)));
        KCFD_HIP(hipMalloc(&dst->chu_o16, sizeof(unsigned short) * std::max(src->nent_u, 1)));
        if (src->nent_l)
            KCFD_HIP(hipMemcpyAsync(dst->chl_o16, src->chl_o16, sizeof(unsigned short) * src->nent_l,
                                    hipMemcpyDeviceToDevice, dst->ctx->stream));
        if (src->nent_u)
            KCFD_HIP(hipMemcpyAsync(dst->chu_o16, src->chu_o16, sizeof(unsigned short) * src->nent_u,
                                    hipMemcpyDeviceToDevice, dst->ctx->stream));
        KCFD_CHECK(dup_ints(src->chl_base, (size_t)src->nch_l, &dst->chl_base, dst->ctx->stream));
        KCFD_CHECK(dup_ints(src->chu_base, (size_t)src->nch_u, &dst->chu_base, dst->ctx->stream));
    }
    dst->ch16 = src->ch16;
    KCFD_CHECK(dup_ints(src->chu_map, (size_t)src->nent_u, &dst->chu_map, dst->ctx->stream));
    dst->h_chl_ptr = src->h_chl_ptr;
    dst->h_chu_ptr = src->h_chu_ptr;
    dst->h_chl_off = src->h_chl_off;
    dst->h_chu_off = src->h_chu_off;
    dst->h_chl_base = src->h_chl_base;
    dst->h_chu_base = src->h_chu_base;
    dst->nch_l = src->nch_l;
    dst->nch_u = src->nch_u;
    dst->nent_l = src->nent_l;
    dst->nent_u = src->nent_u;
    dst->h_lvl_ptr_l = src->h_lvl_ptr_l;
    dst->h_lvl_ptr_u = src->h_lvl_ptr_u;
    dst->nlev_l = src->nlev_l;
    dst->nlev_u = src->nlev_u;
    dst->ncolors = src->ncolors;
    dst->ilu_order = src->ilu_built_order;
    dst->ilu_built_order = src->ilu_built_order;
    KCFD_HIP(hipStreamSynchronize(dst->ctx->stream));
    return KCFD_OK;
}

extern "C" int kcfd_ilu0_order(kcfd_mat *m, int *perm, int *ncolors) {
    KCFD_ARG(m != nullptr, "null matrix");
    std::lock_guard<std::recursive_mutex> g(m->ctx->mu);
    if (!m->lu) {
        set_error("ILU(0) factors not set up");
        return KCFD_ESTATE;
    }
    if (perm) {
        if (m->h_perm.empty())
            for (int i = 0; i < m->n; ++i) perm[i] = i;
        else
            std::copy(m->h_perm.begin(), m->h_perm.end(), perm);
    }
    if (ncolors) *ncolors = m->ncolors;
    return KCFD_OK;
}
