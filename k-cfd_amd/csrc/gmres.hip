// Restarted, left-preconditioned GMRES(m) on the device, FP64.
//
// Follows scipy.sparse.linalg.gmres (scipy 1.15, the solver the reference
// calls at core/linear_solver.py:64-68 as gmres(A, b, x0=x0, atol=tol,
// maxiter=maxiter, M=M) with restart 20): modified Gram-Schmidt Arnoldi on
// M^{-1} A, Givens rotations by LAPACK's xLARTG (the safe-scaling algorithm
// of Anderson, ACM TOMS 44(1) 2017, Algorithm 978), the inner stop at
// |S_{col+1}| <= ptol with the gh-8400 ptol adaptation between cycles, an
// upper-triangular back substitution that skips zero components, and the
// true residual after each cycle.  maxiter counts restart cycles.
//
// Layout.  One cycle is enqueued without a host synchronisation: the
// Hessenberg matrix, rotations, S and the inner tolerance live in a device
// GmState; every kernel of column col returns at once when the cycle stopped
// at an earlier column (stop_col < col), so a cycle that converges early
// costs only launches for its unused columns.  Every cycle thus enqueues the
// same launches; from the second cycle on they are replayed from a hipGraph
// captured once per solve (a cycle is ~15 launches per inner iteration, and
// replaying them removes the host launch gaps).  b and x0 are staged in the
// workspace so that the graph holds no caller pointer.  KCFD_GMRES_GRAPH=0
// in the environment launches every cycle directly instead (for tools that
// do not follow graph launches; the kernels and results are the same).  The
// host reads a few scalars once per cycle.  The
// reductions are two-stage and deterministic: each kernel writes one
// partial per workgroup, and every workgroup of the next kernel sums all
// partials in the same fixed order, so each gets the same scalar without a
// separate finalize launch.  Partials alternate between two buffers.
//
// With the colour-ordered ILU(0) the SELL copy of A is in the ILU(0) row
// order; the solve then runs in that order (b and x0 gathered, x scattered
// at the end), where every row sum still takes A's CSR entry order.

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>

#include "common.hpp"

namespace kcfd {

constexpr int kMaxRestart = 64;
constexpr int kGmBlock = 256;
constexpr int kMaxGmGrid = 2048;   // workgroups of the vector kernels (partials per reduction)
constexpr int kGmPerCu = 8;        // 8 workgroups of 4 wavefronts per CU: 8 wavefronts per SIMD

struct GmState {
    // read back by the host once per cycle
    double rnorm;        // ||b - A x|| after the last residual kernel
    double presid;       // |S_{col+1}| of the last column built
    double fin0, fin1;   // setup norms: ||b||, ||M^{-1} b||
    int inner;           // inner iterations, all cycles
    int ncol;            // columns built in the current cycle
    int breakdown;       // h_{col+1,col} <= eps h0 at the last column
    int stop_col;        // column at which the cycle stopped (INT_MAX while running)
    double atol;         // absolute stopping tolerance max(atol, rtol ||b||)
    double ptol;         // inner stopping tolerance of the current cycle
    double pmf;          // ptol_max_factor of the gh-8400 adaptation
    double h0;           // ||w|| of the current column before orthogonalisation
    double S[kMaxRestart + 1], y[kMaxRestart], gc[kMaxRestart], gs[kMaxRestart];
    double h[kMaxRestart][kMaxRestart + 1];   // h[col][k] as scipy's h[col, k]
};

struct GmSummary {
    double rnorm, presid, fin0, fin1;
    int inner, ncol, breakdown, stop_col;
};

// workspace cached on the matrix
struct GmresWs {
    size_t ldv = 0;      // vector stride (doubles)
    int cols = 0;        // restart + 1 vectors
    double *V = nullptr, *av = nullptr, *r = nullptr, *t = nullptr, *bp = nullptr, *xp = nullptr;
    double *part = nullptr;   // 2 buffers x kMaxGmGrid x 2
    GmState *st = nullptr;
    size_t bytes = 0;
};

// ------------------------------------------------------------ reductions

// Sum of the NV (<= 2) partials of all G workgroups, in a fixed order; the
// result is returned to every thread of the block.
template <int NV>
__device__ __forceinline__ void sum_parts(const double *__restrict__ part, int G, double (&out)[NV]) {
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = 0.0;
    for (int j = threadIdx.x; j < G; j += kGmBlock) {
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += part[2 * j + k];
    }
    block_sum<NV, kGmBlock>(v);
    __shared__ double bc[NV];
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) bc[k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = bc[k];
}

template <int NV>
__device__ __forceinline__ void put_parts(double (&v)[NV], double *__restrict__ part) {
    block_sum<NV, kGmBlock>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) part[2 * blockIdx.x + k] = v[k];
    }
}

__device__ __forceinline__ bool skipped(const GmState *st, int col) { return st->stop_col < col; }

// ------------------------------------------------------------ kernels

// y_i = sum_q a_q x_{col q} over the SELL-64 row (A's CSR order, padding
// last).  A wave's 64 rows are one slice, so the width is uniform: the
// common widths issue every load of the row before the in-order sum.
template <int W>
__device__ __forceinline__ double sell_fixed(int base, const int *__restrict__ six,
                                             const double *__restrict__ sa, const double *__restrict__ x) {
#pragma clang fp contract(off)
    int c[W];
    double a[W], xv[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        c[k] = gl(six)[base + 64 * k];
        a[k] = gl(sa)[base + 64 * k];
    }
#pragma unroll
    for (int k = 0; k < W; ++k) xv[k] = gl(x)[c[k]];
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < W; ++k) acc += a[k] * xv[k];
    return acc;
}

__device__ __forceinline__ double sell_row(int i, const int *__restrict__ sptr, const int *__restrict__ six,
                                           const double *__restrict__ sa, const double *__restrict__ x) {
#pragma clang fp contract(off)
    const int s = i >> 6, base = sptr[s] + (i & 63), w = uniform((sptr[s + 1] - sptr[s]) >> 6);
    if (w == 7) return sell_fixed<7>(base, six, sa, x);
    if (w == 5) return sell_fixed<5>(base, six, sa, x);
    double acc = 0.0;
    for (int k = 0; k < w; ++k) acc += sa[base + 64 * k] * x[six[base + 64 * k]];
    return acc;
}

__global__ __launch_bounds__(kGmBlock) void g_spmv(int n, const int *__restrict__ sptr,
                                                   const int *__restrict__ six,
                                                   const double *__restrict__ sa,
                                                   const double *__restrict__ x, double *__restrict__ y,
                                                   const GmState *st, int col) {
    if (skipped(st, col)) return;
    // whole waves (one SELL slice each, so one width); the lanes past n in
    // the last slice, whose entries are never written, compute row n - 1
    // (same slice) and store nothing
    const int nr = (n + 63) & ~63;
    for (int i = blockIdx.x * kGmBlock + threadIdx.x; i < nr; i += gridDim.x * kGmBlock) {
        const double v = sell_row(min(i, n - 1), sptr, six, sa, x);
        if (i < n) y[i] = v;
    }
}

// r = b - A x, partial sums of r.r
__global__ __launch_bounds__(kGmBlock) void g_resid(int n, const int *__restrict__ sptr,
                                                    const int *__restrict__ six,
                                                    const double *__restrict__ sa,
                                                    const double *__restrict__ x,
                                                    const double *__restrict__ b, double *__restrict__ r,
                                                    double *__restrict__ part) {
#pragma clang fp contract(off)
    double v[1] = {0.0};
    const int nr = (n + 63) & ~63;   // whole waves; lanes past n compute row n - 1
    for (int i = blockIdx.x * kGmBlock + threadIdx.x; i < nr; i += gridDim.x * kGmBlock) {
        const double ax = sell_row(min(i, n - 1), sptr, six, sa, x);
        if (i < n) {
            const double ri = b[i] - ax;
            r[i] = ri;
            v[0] += ri * ri;
        }
    }
    put_parts<1>(v, part);
}

// partial sums of a.a and (c ? c.c : 0)
__global__ __launch_bounds__(kGmBlock) void g_norm2_part(int n, const double *__restrict__ a,
                                                         const double *__restrict__ c,
                                                         double *__restrict__ part) {
#pragma clang fp contract(off)
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kGmBlock + threadIdx.x; i < n; i += gridDim.x * kGmBlock) {
        const double ai = a[i];
        v[0] += ai * ai;
        if (c) {
            const double ci = c[i];
            v[1] += ci * ci;
        }
    }
    put_parts<2>(v, part);
}

// Two-colour ILU(0) apply on P-order vectors, P A P^T = [[D0, B], [C, D1]]
// with the n0 colour-0 rows first.  Colour-0 rows have no L entries and
// colour-1 rows no U entries, so the sequential L then U solves reduce to
//   FIRST  (colour-1 rows): w_i = (z_i - sum_j l_ij z_j) / u_ii
//   SECOND (colour-0 rows): w_i = (z_i - sum_j u_ij w_j) / u_ii
// with the entries in CSR order: the same operations in the same order.
// The factors are read in the chunked column-major layout (entry k of the
// row in lane l of chunk c at off[c] + 64k + l, padded with column -1), one
// wave per 64-row chunk.  Skipped once *skip < col (skip null: never).
template <int W, bool SECOND>
__device__ __forceinline__ double ilu_fixed(double acc, int base, const int *__restrict__ colv,
                                            const double *__restrict__ val, const double *__restrict__ src) {
#pragma clang fp contract(off)
    int c[W];
    double a[W], xv[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        c[k] = gl(colv)[base + 64 * k];
        a[k] = gl(val)[base + 64 * k];
    }
#pragma unroll
    for (int k = 0; k < W; ++k) xv[k] = c[k] >= 0 ? gl(src)[c[k]] : 0.0;
#pragma unroll
    for (int k = 0; k < W; ++k)
        if (c[k] >= 0) acc -= a[k] * xv[k];
    return acc;
}

template <bool SECOND>
__global__ __launch_bounds__(kGmBlock) void g_ilu2(int n0, int nrows, const int *__restrict__ off,
                                                   const int *__restrict__ colv,
                                                   const double *__restrict__ val,
                                                   const double *__restrict__ ud, const double *__restrict__ z,
                                                   double *__restrict__ w, const int *skip, int col) {
#pragma clang fp contract(off)
    if (skip && *skip < col) return;
    const int lane = threadIdx.x & 63;
    const int nch = (nrows + 63) >> 6;
    const int nw = (int)gridDim.x * (kGmBlock / 64);
    const double *src = SECOND ? w : z;
    for (int c = uniform((int)blockIdx.x * (kGmBlock / 64) + (int)(threadIdx.x >> 6)); c < nch; c += nw) {
        const int i = (SECOND ? 0 : n0) + 64 * c + lane;
        const int o0 = uniform(off[c]);
        const int wd = uniform((off[c + 1] - o0) >> 6);
        const bool live = 64 * c + lane < nrows;
        const int base = o0 + lane;
        double acc = live ? z[i] : 0.0;
        if (wd == 6) {
            acc = ilu_fixed<6, SECOND>(acc, base, colv, val, src);
        } else if (wd == 4) {
            acc = ilu_fixed<4, SECOND>(acc, base, colv, val, src);
        } else {
            for (int k = 0; k < wd; ++k) {
                const int j = colv[base + 64 * k];
                if (j >= 0) acc -= val[base + 64 * k] * src[j];
            }
        }
        if (live) w[i] = acc / ud[i];
    }
}

// one workgroup: mode 0 -> fin0 = sqrt(s0), fin1 = sqrt(s1); mode 1 -> rnorm = sqrt(s0)
__global__ __launch_bounds__(kGmBlock) void g_finalize(const double *__restrict__ part, int G, GmState *st,
                                                       int mode) {
    double s[2];
    sum_parts<2>(part, G, s);
    if (threadIdx.x != 0) return;
    if (mode == 0) {
        st->fin0 = sqrt(s[0]);
        st->fin1 = sqrt(s[1]);
    } else {
        st->rnorm = sqrt(s[0]);
    }
}

__global__ void g_state_init(GmState *st) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->inner = 0;
        st->ncol = 0;
        st->breakdown = 0;
        st->stop_col = INT_MAX;
        st->presid = 0.0;
    }
}

// v0 *= 1/||v0||; S = (||v0||, 0, ...); a new cycle
__global__ __launch_bounds__(kGmBlock) void g_cycle_start(int n, double *__restrict__ v0,
                                                          const double *__restrict__ part, int G,
                                                          GmState *st, int restart) {
#pragma clang fp contract(off)
    double s[1];
    sum_parts<1>(part, G, s);
    const double tmp = sqrt(s[0]);
    const double inv = 1.0 / tmp;
    for (int i = blockIdx.x * kGmBlock + threadIdx.x; i < n; i += gridDim.x * kGmBlock) v0[i] *= inv;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->S[0] = tmp;
        for (int k = 1; k <= restart; ++k) st->S[k] = 0.0;
        st->ncol = 0;
        st->breakdown = 0;
        st->stop_col = INT_MAX;
    }
}

// partial sums of w.w and v0.w (the first Gram-Schmidt coefficient)
__global__ __launch_bounds__(kGmBlock) void g_mgs_first(int n, const double *__restrict__ v0,
                                                        const double *__restrict__ w,
                                                        double *__restrict__ part, const GmState *st,
                                                        int col) {
#pragma clang fp contract(off)
    if (skipped(st, col)) return;
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kGmBlock + threadIdx.x; i < n; i += gridDim.x * kGmBlock) {
        const double wi = w[i];
        v[0] += wi * wi;
        v[1] += v0[i] * wi;
    }
    put_parts<2>(v, part);
}

// Gram-Schmidt step k of column col: h[col][k] = v_k.w (from the partials),
// w -= h[col][k] v_k, then the partials of v_{k+1}.w (k < col) or w.w.
__global__ __launch_bounds__(kGmBlock) void g_mgs(int n, const double *__restrict__ vk,
                                                  const double *__restrict__ vk1, double *__restrict__ w,
                                                  int k, int col, const double *__restrict__ pin, int G,
                                                  double *__restrict__ pout, GmState *st) {
#pragma clang fp contract(off)
    if (skipped(st, col)) return;
    double tmp;
    if (k == 0) {
        double s[2];
        sum_parts<2>(pin, G, s);
        tmp = s[1];
        if (blockIdx.x == 0 && threadIdx.x == 0) st->h0 = sqrt(s[0]);
    } else {
        double s[1];
        sum_parts<1>(pin, G, s);
        tmp = s[0];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) st->h[col][k] = tmp;
    const double *nx = k < col ? vk1 : nullptr;
    double v[1] = {0.0};
    for (int i = blockIdx.x * kGmBlock + threadIdx.x; i < n; i += gridDim.x * kGmBlock) {
        const double wi = w[i] - tmp * vk[i];
        w[i] = wi;
        v[0] += (nx ? nx[i] : wi) * wi;
    }
    put_parts<1>(v, pout);
}

// LAPACK xLARTG (safe scaling, Anderson 2017): c, s, r with
// [c s; -s c] [f; g] = [r; 0], c >= 0, r carrying the sign of f.
__device__ void lartg(double f, double g, double &c, double &s, double &r) {
#pragma clang fp contract(off)
    const double safmin = 2.2250738585072014e-308;   // smallest normal
    const double safmax = 1.0 / safmin;
    const double rtmin = sqrt(safmin);
    const double rtmax = sqrt(safmax / 2.0);
    const double f1 = fabs(f), g1 = fabs(g);
    if (g == 0.0) {
        c = 1.0;
        s = 0.0;
        r = f;
    } else if (f == 0.0) {
        c = 0.0;
        s = copysign(1.0, g);
        r = g1;
    } else if (f1 > rtmin && f1 < rtmax && g1 > rtmin && g1 < rtmax) {
        const double d = sqrt(f * f + g * g);
        c = f1 / d;
        r = copysign(d, f);
        s = g / r;
    } else {
        const double u = fmin(safmax, fmax(safmin, fmax(f1, g1)));
        const double fs = f / u, gs = g / u;
        const double d = sqrt(fs * fs + gs * gs);
        c = fabs(fs) / d;
        r = copysign(d, f);
        s = gs / r;
        r = r * u;
    }
}

// End of column col: h1 = ||w||, breakdown test, w /= h1; one thread then
// applies the earlier rotations to column col, forms the new one, updates S
// and decides whether the cycle stops here.
__global__ __launch_bounds__(kGmBlock) void g_mgs_last(int n, double *__restrict__ w, int col,
                                                       const double *__restrict__ pin, int G, GmState *st,
                                                       double eps) {
#pragma clang fp contract(off)
    if (skipped(st, col)) return;
    double s1[1];
    sum_parts<1>(pin, G, s1);
    const double h1 = sqrt(s1[0]);
    const bool bd = h1 <= eps * st->h0;
    if (!bd) {
        const double inv = 1.0 / h1;
        for (int i = blockIdx.x * kGmBlock + threadIdx.x; i < n; i += gridDim.x * kGmBlock) w[i] *= inv;
    }
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double *hc = st->h[col];
    hc[col + 1] = bd ? 0.0 : h1;
    for (int k = 0; k < col; ++k) {
        const double c = st->gc[k], s = st->gs[k];
        const double n0 = hc[k], n1 = hc[k + 1];
        hc[k] = c * n0 + s * n1;
        hc[k + 1] = -s * n0 + c * n1;
    }
    double c, s, mag;
    lartg(hc[col], hc[col + 1], c, s, mag);
    st->gc[col] = c;
    st->gs[col] = s;
    hc[col] = mag;
    hc[col + 1] = 0.0;
    const double tmp = -s * st->S[col];
    st->S[col] = c * st->S[col];
    st->S[col + 1] = tmp;
    st->presid = fabs(tmp);
    st->inner += 1;
    st->ncol = col + 1;
    st->breakdown = bd ? 1 : 0;
    if (st->presid <= st->ptol || bd) st->stop_col = col;
}

// y = H^{-1} S over the ncol built columns (zero components skipped; a zero
// last diagonal zeroes its S entry).  The block stages H and S in LDS, then
// one thread substitutes.
__global__ __launch_bounds__(kGmBlock) void g_backsub(GmState *st) {
#pragma clang fp contract(off)
    __shared__ double hs[kMaxRestart][kMaxRestart + 1];
    __shared__ double ss[kMaxRestart + 1];
    const int nc = uniform(st->ncol);
    if (nc <= 0) return;
    for (int e = threadIdx.x; e < nc * (kMaxRestart + 1); e += kGmBlock)
        hs[e / (kMaxRestart + 1)][e % (kMaxRestart + 1)] = st->h[e / (kMaxRestart + 1)][e % (kMaxRestart + 1)];
    for (int e = threadIdx.x; e < nc; e += kGmBlock) ss[e] = st->S[e];
    __syncthreads();
    if (threadIdx.x != 0) return;
    const int col = nc - 1;
    if (hs[col][col] == 0.0) {
        ss[col] = 0.0;
        st->S[col] = 0.0;
    }
    double *y = ss;   // solved in place
    for (int k = col; k > 0; --k) {
        if (y[k] != 0.0) {
            y[k] /= hs[k][k];
            const double tmp = y[k];
            for (int j = 0; j < k; ++j) y[j] -= tmp * hs[k][j];
        }
    }
    if (y[0] != 0.0) y[0] /= hs[0][0];
    for (int k = 0; k <= col; ++k) st->y[k] = y[k];
}

// x += sum_k y_k v_k
__global__ __launch_bounds__(kGmBlock) void g_xupdate(int n, const double *__restrict__ V, size_t ldv,
                                                      double *__restrict__ x, const GmState *st) {
#pragma clang fp contract(off)
    __shared__ double ys[kMaxRestart];
    const int nc = uniform(st->ncol);
    for (int k = threadIdx.x; k < nc; k += kGmBlock) ys[k] = st->y[k];
    __syncthreads();
    for (int i = blockIdx.x * kGmBlock + threadIdx.x; i < n; i += gridDim.x * kGmBlock) {
        double acc = 0.0;
        for (int k = 0; k < nc; ++k) acc += ys[k] * V[k * ldv + i];
        x[i] += acc;
    }
}

// End of a cycle, one workgroup: ||r|| from the residual partials, then the
// gh-8400 adaptation of the inner tolerance for the next cycle when this one
// neither converged nor broke down.
__global__ __launch_bounds__(kGmBlock) void g_cycle_end(const double *__restrict__ part, int G, GmState *st,
                                                        double eps) {
#pragma clang fp contract(off)
    double s[1];
    sum_parts<1>(part, G, s);
    if (threadIdx.x != 0) return;
    const double rnorm = sqrt(s[0]);
    st->rnorm = rnorm;
    if (rnorm <= st->atol || st->breakdown) return;
    if (st->presid <= st->ptol)
        st->pmf = fmax(eps, 0.25 * st->pmf);
    else
        st->pmf = fmin(1.0, 1.5 * st->pmf);
    st->ptol = st->presid * fmin(st->pmf, st->atol / rnorm);
}

__global__ void g_set_tol(GmState *st, double atol, double ptol) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->atol = atol;
        st->ptol = ptol;
        st->pmf = 1.0;
    }
}

__global__ void g_copy(int n, const double *__restrict__ src, double *__restrict__ dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}

__global__ void g_gather(int n, const int *__restrict__ perm, const double *__restrict__ src,
                         double *__restrict__ dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[perm[k]];
}

__global__ void g_scatter(int n, const int *__restrict__ perm, const double *__restrict__ src,
                          double *__restrict__ dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[perm[k]] = src[k];
}

// ------------------------------------------------------------ host side

// y = M^{-1} x for the solver: the two-colour ILU(0) through the chunked
// layout (two launches), every other preconditioner through the library's
// level-scheduled apply
static int gm_precond(kcfd_mat *m, const double *x, double *y, double *t, bool porder, const int *skip, int col,
                      int G) {
    if (m->pc_kind == KCFD_PC_ILU0 && porder && m->two_colour && m->chl_v64 && m->ud64) {
        hipStream_t s = m->ctx->stream;
        const int n0 = m->ncolour0, n1 = m->n - m->ncolour0;
        if (n1 > 0)
            hipLaunchKernelGGL(g_ilu2<false>, dim3(G), dim3(kGmBlock), 0, s, n0, n1, m->chl_off, m->chl_col,
                               m->chl_v64, m->ud64, x, y, skip, col);
        if (n0 > 0)
            hipLaunchKernelGGL(g_ilu2<true>, dim3(G), dim3(kGmBlock), 0, s, n0, n0, m->chu_off, m->chu_col,
                               m->chu_v64, m->ud64, x, y, skip, col);
        KCFD_HIP(hipGetLastError());
        return KCFD_OK;
    }
    return precond_apply_async(m, x, y, t, porder, skip, col);
}

size_t gmres_bytes(const kcfd_mat *m) { return m->gmres ? m->gmres->bytes + sizeof(GmState) : 0; }

void gmres_free(kcfd_mat *m) {
    GmresWs *w = m->gmres;
    if (!w) return;
    hipFree(w->V);
    hipFree(w->av);
    hipFree(w->r);
    hipFree(w->t);
    hipFree(w->bp);
    hipFree(w->xp);
    hipFree(w->part);
    hipFree(w->st);
    delete w;
    m->gmres = nullptr;
}

static int ensure_ws(kcfd_mat *m, int cols) {
    const size_t n = (size_t)m->n;
    const size_t ldv = (n + 63) & ~(size_t)63;
    GmresWs *w = m->gmres;
    if (w && w->cols >= cols && w->ldv == ldv) return KCFD_OK;
    gmres_free(m);
    w = new GmresWs();
    m->gmres = w;
    w->ldv = ldv;
    w->cols = cols;
    auto alloc = [&](double **p, size_t count) {
        w->bytes += sizeof(double) * count;
        return hipMalloc(p, sizeof(double) * std::max<size_t>(count, 1));
    };
    if (alloc(&w->V, ldv * cols) != hipSuccess || alloc(&w->av, ldv) != hipSuccess ||
        alloc(&w->r, ldv) != hipSuccess || alloc(&w->t, ldv) != hipSuccess ||
        alloc(&w->part, 2 * 2 * kMaxGmGrid) != hipSuccess || alloc(&w->bp, ldv) != hipSuccess ||
        alloc(&w->xp, ldv) != hipSuccess ||
        hipMalloc(&w->st, sizeof(GmState)) != hipSuccess) {
        gmres_free(m);
        set_error("GMRES workspace: hipMalloc failed (n=%zu, %d vectors)", n, cols);
        return KCFD_EOOM;
    }
    return KCFD_OK;
}

// algorithmic bytes of one SpMV and one preconditioner apply
static double spmv_bytes(const kcfd_mat *m) {
    return 12.0 * m->sell_nnz + 4.0 * (m->nslices + 1) + 16.0 * m->n;
}
static double precond_bytes(const kcfd_mat *m) {
    const double n = m->n;
    if (m->pc_kind == KCFD_PC_JACOBI) return 24.0 * n;
    if (m->pc_kind == KCFD_PC_ILU0) return 12.0 * m->nnz + 4.0 * (m->n + 1) + 12.0 * n + 32.0 * n;
    return 16.0 * n;
}

int gmres_solve(kcfd_mat *m, const double *b, double *x, const kcfd_solve_opts *o, kcfd_solve_result *res) {
    KCFD_ARG(o->precision == KCFD_FP64, "GMRES runs in FP64 only");
    KCFD_ARG(o->precond == KCFD_PC_NONE || o->precond == KCFD_PC_JACOBI || o->precond == KCFD_PC_ILU0,
             "unknown preconditioner");
    const int n = m->n;
    int restart = o->restart > 0 ? o->restart : 20;
    KCFD_ARG(restart <= kMaxRestart, "GMRES restart must be 1..64");
    restart = std::min(restart, n);
    if (m->pc_kind != o->precond) {
        set_error("preconditioner not set up as requested (have %d, want %d)", m->pc_kind, o->precond);
        return KCFD_ESTATE;
    }
    const bool porder = m->pc_kind == KCFD_PC_ILU0 && m->sell_permuted;
    if (m->sell_permuted != porder || (porder && !m->perm) ||
        (m->pc_kind == KCFD_PC_ILU0 && !porder && m->perm)) {
        set_error("SELL row order does not match the preconditioner's (set up the preconditioner again)");
        return KCFD_ESTATE;
    }
    const int maxiter = o->maxiter > 0 ? o->maxiter : 10 * n;
    KCFD_CHECK(ensure_ws(m, restart + 1));
    GmresWs *w = m->gmres;
    kcfd_ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    const int G = std::max(1, std::min((n + kGmBlock - 1) / kGmBlock, std::min(kMaxGmGrid, kGmPerCu * ctx->cus)));
    const unsigned nb = (unsigned)((n + kGmBlock - 1) / kGmBlock);
    const double eps = 2.220446049250313e-16;
    const size_t ldv = w->ldv;
    double *V = w->V, *p0 = w->part, *p1 = w->part + 2 * kMaxGmGrid;
    GmState *st = w->st;
    const int *stop_col = (const int *)((const char *)st + offsetof(GmState, stop_col));
    double *B = w->bp, *X = w->xp;   // b and x in the solver's row order
    GmSummary sm{};
    auto summary = [&]() -> int {
        KCFD_HIP(hipMemcpyAsync(&sm, st, sizeof(sm), hipMemcpyDeviceToHost, s));
        KCFD_HIP(hipStreamSynchronize(s));
        return KCFD_OK;
    };
    // one restart cycle: the same launches every cycle (all scalars it
    // depends on are in GmState)
    auto enqueue_cycle = [&]() -> int {
        KCFD_CHECK(gm_precond(m, w->r, V, w->t, porder, nullptr, 0, G));
        hipLaunchKernelGGL(g_norm2_part, dim3(G), dim3(kGmBlock), 0, s, n, V, (const double *)nullptr, p0);
        hipLaunchKernelGGL(g_cycle_start, dim3(G), dim3(kGmBlock), 0, s, n, V, p0, G, st, restart);
        for (int col = 0; col < restart; ++col) {
            double *vc = V + (size_t)col * ldv, *wv = V + (size_t)(col + 1) * ldv;
            hipLaunchKernelGGL(g_spmv, dim3(G), dim3(kGmBlock), 0, s, n, m->slice_ptr, m->six, m->sa64, vc,
                               w->av, st, col);
            KCFD_CHECK(gm_precond(m, w->av, wv, w->t, porder, stop_col, col, G));
            hipLaunchKernelGGL(g_mgs_first, dim3(G), dim3(kGmBlock), 0, s, n, V, wv, p0, st, col);
            for (int k = 0; k <= col; ++k) {
                const double *pin = (k & 1) ? p1 : p0;
                double *pout = (k & 1) ? p0 : p1;
                hipLaunchKernelGGL(g_mgs, dim3(G), dim3(kGmBlock), 0, s, n, V + (size_t)k * ldv,
                                   V + (size_t)(k + 1) * ldv, wv, k, col, pin, G, pout, st);
            }
            hipLaunchKernelGGL(g_mgs_last, dim3(G), dim3(kGmBlock), 0, s, n, wv, col,
                               ((col + 1) & 1) ? p1 : p0, G, st, eps);
        }
        hipLaunchKernelGGL(g_backsub, dim3(1), dim3(kGmBlock), 0, s, st);
        hipLaunchKernelGGL(g_xupdate, dim3(G), dim3(kGmBlock), 0, s, n, V, ldv, X, st);
        hipLaunchKernelGGL(g_resid, dim3(G), dim3(kGmBlock), 0, s, n, m->slice_ptr, m->six, m->sa64, X, B,
                           w->r, p0);
        hipLaunchKernelGGL(g_cycle_end, dim3(1), dim3(kGmBlock), 0, s, p0, G, st, eps);
        KCFD_HIP(hipGetLastError());
        return KCFD_OK;
    };
    double bytes = 0.0;
    const double vec = 8.0 * n;

    KCFD_HIP(hipEventRecord(ctx->ev0, s));
    if (porder) {
        hipLaunchKernelGGL(g_gather, dim3(nb), dim3(kGmBlock), 0, s, n, m->perm, b, B);
        hipLaunchKernelGGL(g_gather, dim3(nb), dim3(kGmBlock), 0, s, n, m->perm, x, X);
        bytes += 2 * (12.0 * n + 8.0 * n);
    } else {
        hipLaunchKernelGGL(g_copy, dim3(nb), dim3(kGmBlock), 0, s, n, b, B);
        hipLaunchKernelGGL(g_copy, dim3(nb), dim3(kGmBlock), 0, s, n, x, X);
        bytes += 2 * (16.0 * n);
    }
    hipLaunchKernelGGL(g_state_init, dim3(1), dim3(64), 0, s, st);
    // ||b||, ||M^{-1} b|| and the initial residual
    KCFD_CHECK(gm_precond(m, B, V, w->t, porder, nullptr, 0, G));
    hipLaunchKernelGGL(g_norm2_part, dim3(G), dim3(kGmBlock), 0, s, n, B, V, p0);
    hipLaunchKernelGGL(g_finalize, dim3(1), dim3(kGmBlock), 0, s, p0, G, st, 0);
    hipLaunchKernelGGL(g_resid, dim3(G), dim3(kGmBlock), 0, s, n, m->slice_ptr, m->six, m->sa64, X, B,
                       w->r, p1);
    hipLaunchKernelGGL(g_finalize, dim3(1), dim3(kGmBlock), 0, s, p1, G, st, 1);
    KCFD_HIP(hipGetLastError());
    bytes += precond_bytes(m) + 2 * vec + spmv_bytes(m) + 2 * vec;
    KCFD_CHECK(summary());

    const double bnrm2 = sm.fin0;
    int info = 0, cycles = 0;
    double rnorm = sm.rnorm;
    bool done = false;
    if (bnrm2 == 0.0) {   // scipy returns b itself (all zeros)
        rnorm = 0.0;
        done = true;
    }
    const double atol = std::max(o->atol, o->rtol * bnrm2);
    if (!done) {
        const double ptol = sm.fin1 * std::min(1.0, atol / bnrm2);
        hipLaunchKernelGGL(g_set_tol, dim3(1), dim3(64), 0, s, st, atol, ptol);
        KCFD_HIP(hipGetLastError());
        if (rnorm < atol) done = true;
    }
    hipGraphExec_t graph = nullptr;
    int rc = KCFD_OK;
    const char *genv = std::getenv("KCFD_GMRES_GRAPH");
    const bool use_graph = !(genv && genv[0] == '0');
    for (int it = 0; !done && it < maxiter; ++it) {
        ++cycles;
        if (it == 0 || !use_graph) {
            rc = enqueue_cycle();
        } else {
            if (!graph) {   // a second cycle: capture one and replay it from now on
                hipGraph_t g = nullptr;
                KCFD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                rc = enqueue_cycle();
                hipError_t e = hipStreamEndCapture(s, &g);
                if (rc == KCFD_OK && e != hipSuccess) {
                    set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
                    rc = KCFD_EHIP;
                }
                if (rc == KCFD_OK) {
                    e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
                    if (e != hipSuccess) {
                        set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
                        rc = KCFD_EHIP;
                        graph = nullptr;
                    }
                }
                if (g) hipGraphDestroy(g);
            }
            if (rc == KCFD_OK && hipGraphLaunch(graph, s) != hipSuccess) {
                set_error("hipGraphLaunch failed");
                rc = KCFD_EHIP;
            }
        }
        if (rc == KCFD_OK) rc = summary();
        if (rc != KCFD_OK) break;
        const int ncol = sm.ncol;
        // bytes: preconditioned residual, v0 norm and scaling, then per column
        // SpMV + apply + Gram-Schmidt (w and v_0..v_col read, w written
        // col + 2 times), then the update and the true residual
        bytes += precond_bytes(m) + 3 * vec;
        for (int c = 0; c < ncol; ++c)
            bytes += spmv_bytes(m) + precond_bytes(m) + 2 * vec + (3.0 * c + 3) * vec + 2 * vec;
        bytes += (ncol + 2) * vec + spmv_bytes(m) + 2 * vec;
        rnorm = sm.rnorm;
        if (rnorm <= atol || sm.breakdown) break;   // otherwise g_cycle_end has set the next ptol
    }
    if (graph) hipGraphExecDestroy(graph);
    if (rc != KCFD_OK) return rc;
    if (!(rnorm <= atol)) info = maxiter;
    if (bnrm2 == 0.0) {
        hipLaunchKernelGGL(g_copy, dim3(nb), dim3(kGmBlock), 0, s, n, b, x);
    } else if (porder) {
        hipLaunchKernelGGL(g_scatter, dim3(nb), dim3(kGmBlock), 0, s, n, m->perm, X, x);
        bytes += 20.0 * n;
    } else {
        hipLaunchKernelGGL(g_copy, dim3(nb), dim3(kGmBlock), 0, s, n, X, x);
        bytes += 16.0 * n;
    }
    KCFD_HIP(hipGetLastError());
    KCFD_HIP(hipEventRecord(ctx->ev1, s));
    KCFD_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    KCFD_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms = ms;
    if (res) {
        std::memset(res, 0, sizeof(*res));
        res->info = info;
        res->iters = sm.inner;
        res->refinements = cycles;
        res->rel_resid = bnrm2 > 0 ? rnorm / bnrm2 : 0.0;
        res->bytes = bytes;
        res->ms = ms;
    }
    return KCFD_OK;
}

}  // namespace kcfd
