// Batched, device-resident BiCGSTAB (Jacobi / ILU(0) / none) for gfx950.
//
// One launch per phase covers the row tiles of every active system of a
// batch (persistent, XCD-aware tile walk); each system carries its own scalar
// state machine on the device (status, rho, alpha, omega, beta, counts), so
// scalars never leave the device and converged systems drop out without the
// host.  The host replays a captured hipGraph of `check_every` iterations and
// reads the status words in between.
//
// Recurrences: scipy's right-preconditioned BiCGSTAB
// (scipy/sparse/linalg/_isolve/iterative.py:157-302, the solver behind
// core/linear_solver.py:59): same stopping and breakdown tests (rhotol =
// omegatol = eps^2, rv == 0), half-step exit on ||s||, iteration count =
// callback count, and the same operation order in every vector update
// (products and sums not fused).  The SpMV reads a SELL-64 copy of A whose
// rows keep CSR order, so A z is bit-identical to scipy's csr_matvec; FP64
// iterates differ from scipy's only through dot-product summation order.
//
// Per iteration (Jacobi folded into the gathers; ILU(0) adds two sync-free
// triangular solves after K1 and after K3):
//   K1 p = r + beta (p - omega v)        (p = r at iteration 0)
//   K2 v = A M p; rt.v           F2: alpha
//   K3 s = r - alpha v                   (with ILU(0): formed by the s apply)
//   K4 t = A M s; t.s, t.t, ||s||^2  F4: half-step exit on ||s|| (scipy's test,
//                                    made one finalize later), else omega
//   K5 x += alpha M p; x += omega M s; r = s - omega t; ||r||, rt.r
//      (half-step exit: x += alpha M p only)
//                                F5: convergence / breakdown / beta
// Folded products (two-colour ILU(0), kcfd_solve_opts.fold_colour0; on by
// default in MIXED): the colour-0 half of each apply also writes the
// colour-0 rows of v = A M p (t = A M s), and K2 / K4 stream the colour-1
// rows only.
// Reductions: wave64 xor-shuffle + LDS per tile, one partial record per tile,
// and a finalize kernel (one workgroup per system) that sums the partials in
// tile order -- deterministic, and free of in-kernel fences (on gfx950 an
// agent-scope fence writes back and invalidates the XCD's L2).
//
// KCFD_MIXED: the iteration runs on FP32 matrix values (difference form, see
// row_amz), FP32 preconditioner and FP32 vectors (products exact in FP64,
// sums and all dot products in FP64); the iterate x stays FP64.  Reliable
// updates (residual replacement: r = b - A x in FP64 on the FP64 matrix,
// r32 = r) are taken whenever the FP32 residual has dropped by `delta` since
// the last one, when it claims convergence, or at the iteration limit,
// without restarting the Krylov space (the shadow residual and p are kept;
// rho and beta are recomputed from the replaced residual).  Convergence is
// judged on the FP64 true residual; a stop without convergence returns the
// best iterate seen.
#include <chrono>
#include <cstdlib>

#include "common.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace kcfd {

enum : int {
    S_RUN = 0,      // iteration in progress
    S_START = 1,    // iteration 0 next: p = r
    S_HALF = 2,     // half-step exit: x += alpha M p pending
    S_REFINE = 3,   // MIXED: reliable update (x += d; r = b - A x) pending
    S_ZEROX = 4,    // ||b|| == 0: x = 0 pending
    S_RESUME = 5,   // MIXED: continue after a reliable update (p-update with new beta)
    S_NEW = 6,      // attached, initialisation (k_init / f_init) pending
    S_DONE = 16,    // terminal
    S_FREE = 32,    // empty slot of a dynamic batch
};

struct SysDesc {
    int n, tile0, ntiles;
    int polled;   // ILU(0) solved by the sync-free kernels (sentinel-flagged buffers)
    // Folded product (kcfd_solve_opts.fold_colour0, two-colour ILU(0)):
    // rows [0, nid) of the solver's order are colour 0; the colour-0 half
    // of each apply also writes those rows of A M z (k_ilu2 SECOND), so the
    // product kernels only read them.  c0piv: MIXED, their FP64 diagonal.
    // nid 0: no folding.
    int nid;
    const double *c0piv;
    // ILU(0) factor structure: CSR of P A P^T (A's own CSR in the natural
    // ordering); uix = columns as rows of A.  perm: the solver's row order,
    // row of A at each position (null: identity).  With a permuted ILU(0)
    // every vector and the SELL copy are in P-order; b and x are gathered
    // from the caller's buffers at the start and x is scattered back at the
    // end.
    const int *ip, *ix, *dpos, *uix, *perm;
    // chunked L / U solve layout (kcfd_mat::chl_off): offsets per chunk,
    // columns (rows of P A P^T), values, and u_ii by row
    const int *chl_off, *chl_col, *chu_off, *chu_col;
    const double *chl_v64, *chu_v64, *ud64;
    const int2 *chl_pk32, *chu_pk32;   // MIXED: packed {column, FP32 value} entries
    // MIXED, narrow-index two-colour layouts (kcfd_mat::ch16; null if packed):
    // FP32 values, 16-bit offsets from the chunk's base column, the bases
    const float *chl_v32, *chu_v32;
    const unsigned short *chl_o16, *chu_o16;
    const int *chl_base, *chu_base;
    const float *ud32;
    // segment ordering (kcfd_mat::seg2): {first row, length} per segment,
    // {entry offset, width} per step, in-segment L / U values by row
    const int2 *sg;
    const int4 *sg_steps;
    const double *sgl64, *sgu64;
    const float *sgl32, *sgu32;
    const int *sptr, *six;          // SELL-64 structure
    const double *sa64;
    const int2 *spk32;              // MIXED: packed {column, FP32 value} SELL entries
    const float *sv32;              // MIXED, narrow-index form (kcfd_mat::so16): values,
    const unsigned short *so16;     //   16-bit column offsets (kSellDiag: the row itself)
    const int *sbase;               //   and the column base of each slice
    const double *dinv64;
    const float *dinv32;
    const double *lu64;
    const float *lu32;
    const double *b;  // in the solver's row order
    double *x;        // the iterate, FP64 in both modes, in the solver's row order
    const double *b_user;   // the caller's b and x (natural order)
    double *x_user;
    double *xb;       // MIXED: the iterate with the lowest true residual so far
    // inner vectors (double in FP64 mode, float in MIXED mode)
    void *r, *rt, *p, *v, *s, *t, *ph, *sh, *y;
};

struct SysState {
    double rho, rho_prev, alpha, omega, beta;
    double atol_in, atol_out, bnrm;
    double rr, rr64, rmax, rbest;
    double true_rr, true_bb;
    int status, iters, info, half;
    int refinements, total_iters, maxiter, restart;
    int final_, stagnant, true_done, save;   // save: copy x to xb in the next K1
    int restore, best_iter, rollback, scatter;   // restore: x = xb before the true residual;
                                                 // rollback: x = xb before a residual replacement;
                                                 // scatter: x to the caller's order pending
};

struct BatchParams {
    double rtol, atol, delta, rhotol;
    int precision, precond, max_refine, nsys;
    int inner_maxiter, method;
    int poll_limit;   // sync-free triangular solves: polls of one operand before giving up
    int pad;
};

// counters[kTrsvFail] != 0: a sync-free triangular solve gave up waiting for
// an operand (the solve's result is invalid; the host reports KCFD_ESTATE)
constexpr int kTrsvFail = 6;
constexpr int kDefaultPollLimit = 1 << 20;
// ints of the kernels' counts table (kcfd_batch::d_counts)
constexpr int kCounts = 10;

// ILU trisolve work chunk: up to 64 consecutive rows of one system's level
// list, 16 bytes: the system (slot), row count and width packed in `meta`
// (12, 7 and 13 bits), the first row, the first entry in the matrix's chunk
// layout (kcfd_mat::chl_off / chu_off) and the base column of a
// narrow-index two-colour chunk.  Carrying the layout position here lets a
// wave's entry loads follow the chunk's own load, not another dependent
// load; keeping the record small keeps the per-round table uploads of a
// serving batch small.
constexpr int kChunkMaxSys = 1 << 12, kChunkMaxWidth = (1 << 13) - 1;
struct Chunk {
    int meta, start, o0, cb;
    __host__ __device__ int sys() const { return meta & (kChunkMaxSys - 1); }
    __host__ __device__ int count() const { return (meta >> 12) & 127; }
    __host__ __device__ int wd() const { return (int)((unsigned)meta >> 19); }
};
__host__ __device__ inline Chunk make_chunk(int sys, int start, int count, int o0, int wd, int cb) {
    return Chunk{(int)((unsigned)sys | ((unsigned)count << 12) | ((unsigned)wd << 19)), start, o0, cb};
}

// ----------------------------------------------------------- tile walk
// Blocks with blockIdx % 8 == x (the XCD they are dispatched to) cover the
// contiguous tile range [x*per, (x+1)*per) of that XCD, so that neighbouring
// rows share its L2.  walk 0: block j of the XCD (J blocks per XCD) takes
// tiles j, j + J, j + 2J, ...; walk 1: block j takes the contiguous run
// [j*c, (j+1)*c), c = ceil(per / J).  The walk mode is counts[5].
// gridDim.x is a multiple of 8; the trip count is uniform over a block.
// KCFD_TILE_LOOP_W takes the walk mode as an expression: the two product
// kernels walk contiguous runs when the batch has folded members
// (counts[6]), whose tile list alternates cheap colour-0 tiles with full
// colour-1 ones -- an interleaved walk with an even J would give each
// workgroup tiles of one colour only.
#define KCFD_TILE_LOOP_W(tile, ntiles, walk)                                                     \
    for (int per_ = ((ntiles) + kXcds - 1) / kXcds, J_ = (int)gridDim.x / kXcds,                  \
             j_ = (int)blockIdx.x / kXcds, x0_ = (int)(blockIdx.x % kXcds) * per_, w_ = (walk),   \
             c_ = w_ ? (per_ + J_ - 1) / J_ : 1, st_ = w_ ? 1 : J_,                                 \
             tile = x0_ + j_ * c_,                                                                 \
             e_ = min(min(x0_ + (w_ ? (j_ + 1) * c_ : per_), x0_ + per_), (ntiles));               \
         tile < e_; tile += st_)
#define KCFD_TILE_LOOP(tile, ntiles) KCFD_TILE_LOOP_W(tile, ntiles, counts[5])

// ----------------------------------------------------- element accessors

template <typename V> __device__ __forceinline__ const V *jac(const SysDesc &d);
template <> __device__ __forceinline__ const double *jac<double>(const SysDesc &d) { return d.dinv64; }
template <> __device__ __forceinline__ const float *jac<float>(const SysDesc &d) { return d.dinv32; }
template <typename V> __device__ __forceinline__ const V *luv(const SysDesc &d);
template <> __device__ __forceinline__ const double *luv<double>(const SysDesc &d) { return d.lu64; }
template <> __device__ __forceinline__ const float *luv<float>(const SysDesc &d) { return d.lu32; }

// Factor entries of the chunked L / U layouts: FP64 keeps columns and values
// in two arrays; MIXED reads one packed {column, FP32 value} pair per entry.
struct FacRef {
    const int *col;
    const double *v64;
    const int2 *pk32;
};
__device__ __forceinline__ FacRef fac_l(const SysDesc &d) { return FacRef{d.chl_col, d.chl_v64, d.chl_pk32}; }
__device__ __forceinline__ FacRef fac_u(const SysDesc &d) { return FacRef{d.chu_col, d.chu_v64, d.chu_pk32}; }
template <typename V>
__device__ __forceinline__ void fac_entry(const FacRef &f, int e, int &c, V &v) {
    if constexpr (sizeof(V) == sizeof(float)) {
        const kint2 q = gl(reinterpret_cast<const kint2 *>(f.pk32))[e];
        c = q.x;
        v = __int_as_float(q.y);
    } else {
        c = gl(f.col)[e];
        v = gl(f.v64)[e];
    }
}
// narrow-index chunk entries: an FP32 value and a 16-bit offset from the
// chunk's base column cb; padding (kSellDiag) decodes to column -1
struct FacRef16 {
    const float *v32;
    const unsigned short *o16;
    int cb;
};
template <typename V>
__device__ __forceinline__ void fac_get(const FacRef &f, int e, int &c, V &v) {
    fac_entry<V>(f, e, c, v);
}
template <typename V>
__device__ __forceinline__ void fac_get(const FacRef16 &f, int e, int &c, V &v) {
    const unsigned o = gl(f.o16)[e];
    v = gl(f.v32)[e];
    c = o == kSellDiag ? -1 : f.cb + (int)o;
}
// typed global view of a workspace vector (stored as void *)
template <typename V>
__device__ __forceinline__ KCFD_GLOBAL V *vec(void *p) {
    return (KCFD_GLOBAL V *)p;
}

template <typename V> __device__ __forceinline__ const V *udiag(const SysDesc &d);
template <> __device__ __forceinline__ const double *udiag<double>(const SysDesc &d) { return d.ud64; }
template <> __device__ __forceinline__ const float *udiag<float>(const SysDesc &d) { return d.ud32; }

template <typename V> struct Sentinel;
template <> struct Sentinel<double> {
    using U = unsigned long long;
    static constexpr U bits = 0x7FF4DEADBEEF0B0BULL;  // a signalling-NaN payload
};
template <> struct Sentinel<float> {
    using U = unsigned int;
    static constexpr U bits = 0x7FA5B0B5u;
};

template <typename V> __device__ __forceinline__ V sentinel_value() {
    return __builtin_bit_cast(V, Sentinel<V>::bits);
}

// preconditioned operand (M z)_j: Jacobi dinv_j z_j, ILU zh_j, none z_j
template <typename V, int PC>
__device__ __forceinline__ V precond_v(const SysDesc &d, const V *__restrict__ z,
                                       const V *__restrict__ zh, int j) {
#pragma clang fp contract(off)
    if (PC == KCFD_PC_JACOBI) return gl(jac<V>(d))[j] * gl(z)[j];
    if (PC == KCFD_PC_ILU0) return gl(zh)[j];
    return gl(z)[j];
}
template <typename V, int PC>
__device__ __forceinline__ double precond_at(const SysDesc &d, const V *__restrict__ z,
                                             const V *__restrict__ zh, int j) {
    return (double)precond_v<V, PC>(d, z, zh, j);
}

// (A M z)_i from the SELL-64 copy, entries in CSR order.  A wave's 64 rows
// are one slice, so the width w is wave-uniform.  The index and value loads
// of a group of entries, then its gathers, are all issued before the
// (in-order) accumulation.  The common widths (5- and 7-point stencils, tet
// rows) run as fixed-size groups without any guard; any other width goes in
// groups of kRowUnroll whose entries past the width re-load the last entry (the same
// cache line) and are not accumulated, so no load sits behind a branch.
constexpr int kRowUnroll = 4;

// entry e of the SELL copy of row i: column and value.  MIXED: one 8-byte
// {column, value} load, or (NARROW) a 4-byte value and a 2-byte offset from
// the slice's column base sb.  The layout is a template parameter so that
// each form is straight-line code (a per-entry branch would split the group
// of loads issued ahead of the accumulation).
template <typename V, bool NARROW>
__device__ __forceinline__ void sell_entry(const SysDesc &d, int e, int i, int sb, int &c, V &v) {
    if constexpr (sizeof(V) == sizeof(float) && NARROW) {
        const unsigned o = gl(d.so16)[e];
        v = gl(d.sv32)[e];
        c = o == kSellDiag ? i : sb + (int)o;
    } else if constexpr (sizeof(V) == sizeof(float)) {
        const kint2 q = gl(reinterpret_cast<const kint2 *>(d.spk32))[e];
        c = q.x;
        v = __int_as_float(q.y);
    } else {
        c = gl(d.six)[e];
        v = gl(d.sa64)[e];
    }
}

// G entries starting at entry k0 of the row at `base`, the first `valid`
// of which are accumulated into acc.  MIXED (V = float): the FP32 values are
// in difference form (row sum on the diagonal), so the row is
// sum_{j!=i} a_ij (z_j - z_i) + rowsum_i z_i with the differences in FP64.
template <typename V, int PC, int G, bool NARROW>
__device__ __forceinline__ void amz_group(const SysDesc &d, int i, const V *__restrict__ z,
                                          const V *__restrict__ zh, int base, int sb, int k0, int last,
                                          int valid, double zi, double &acc) {
#pragma clang fp contract(off)
    constexpr bool kDiff = sizeof(V) == sizeof(float);
    int col[G];
    V av[G], zv[G];
#pragma unroll
    for (int k = 0; k < G; ++k)
        sell_entry<V, NARROW>(d, base + 64 * min(k0 + k, last), i, sb, col[k], av[k]);
#pragma unroll
    for (int k = 0; k < G; ++k) zv[k] = precond_v<V, PC>(d, z, zh, col[k]);
#pragma unroll
    for (int k = 0; k < G; ++k)
        if (k < valid) {
            if (kDiff) acc += (double)av[k] * (col[k] == i ? zi : (double)zv[k] - zi);
            else acc += (double)av[k] * (double)zv[k];
        }
}

template <typename V, int PC, bool NARROW>
__device__ __forceinline__ double row_amz_layout(const SysDesc &d, int i, const V *__restrict__ z,
                                                 const V *__restrict__ zh, int s, int s0, int w) {
#pragma clang fp contract(off)
    constexpr bool kDiff = sizeof(V) == sizeof(float);
    const int base = s0 + (i & 63);
    const int sb = NARROW ? uniform(gl(d.sbase)[s]) : 0;
    const double zi = kDiff ? precond_at<V, PC>(d, z, zh, i) : 0.0;
    double acc = 0.0;
    if (w == 7) {
        amz_group<V, PC, 7, NARROW>(d, i, z, zh, base, sb, 0, 6, 7, zi, acc);
    } else if (w == 5) {
        amz_group<V, PC, 5, NARROW>(d, i, z, zh, base, sb, 0, 4, 5, zi, acc);
    } else {
        for (int k0 = 0; k0 < w; k0 += kRowUnroll)
            amz_group<V, PC, kRowUnroll, NARROW>(d, i, z, zh, base, sb, k0, w - 1, w - k0, zi, acc);
    }
    return acc;
}

template <typename V, int PC>
__device__ __forceinline__ double row_amz(const SysDesc &d, int i, const V *__restrict__ z,
                                          const V *__restrict__ zh) {
    const int s = uniform(i >> 6);
    const int s0 = uniform(gl(d.sptr)[s]), w = uniform((gl(d.sptr)[s + 1] - s0) >> 6);
    if constexpr (sizeof(V) == sizeof(float))
        if (d.so16) return row_amz_layout<V, PC, true>(d, i, z, zh, s, s0, w);
    return row_amz_layout<V, PC, false>(d, i, z, zh, s, s0, w);
}

// (A x)_i in FP64 on the FP64 values
template <int G>
__device__ __forceinline__ void ax_group(const SysDesc &d, const double *__restrict__ x, int base, int k0,
                                         int last, int valid, double &acc) {
#pragma clang fp contract(off)
    int col[G];
    double av[G], xv[G];
#pragma unroll
    for (int k = 0; k < G; ++k) sell_entry<double, false>(d, base + 64 * min(k0 + k, last), 0, 0, col[k], av[k]);
#pragma unroll
    for (int k = 0; k < G; ++k) xv[k] = gl(x)[col[k]];
#pragma unroll
    for (int k = 0; k < G; ++k)
        if (k < valid) acc += av[k] * xv[k];
}

__device__ __forceinline__ double row_ax64(const SysDesc &d, int i, const double *__restrict__ x) {
    const int s = uniform(i >> 6);
    const int s0 = uniform(gl(d.sptr)[s]), w = uniform((gl(d.sptr)[s + 1] - s0) >> 6);
    const int base = s0 + (i & 63);
    double acc = 0.0;
    if (w == 7) {
        ax_group<7>(d, x, base, 0, 6, 7, acc);
    } else if (w == 5) {
        ax_group<5>(d, x, base, 0, 4, 5, acc);
    } else {
        for (int k0 = 0; k0 < w; k0 += kRowUnroll) ax_group<kRowUnroll>(d, x, base, k0, w - 1, w - k0, acc);
    }
    return acc;
}

// Tile kernels: a tile is R kBlock consecutive rows of one system, and thread t
// of the workgroup takes rows t, t + kBlock, ... of it (R rows, each a
// coalesced wave access).  Every load of a thread's rows is issued before
// the first store, and the rows of the last tile past the system's end are
// loaded at the last row and discarded.
// R (rows per thread: 1, 2 or 4; a wave covers 64 R rows of the tile) is a
// template parameter of every tile kernel, chosen at run time per batch
// (kcfd_solve_opts.rows_per_thread, swept by the autotuner).
constexpr int kDefaultRows = 2;

template <int R>
struct TileRows {
    int ix[R];   // row index, clamped to n - 1
    bool ok[R];  // a row of the system
    __device__ __forceinline__ TileRows(int t0, int n) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = t0 + r * kBlock + (int)threadIdx.x;
            ok[r] = i < n;
            ix[r] = ok[r] ? i : n - 1;
        }
    }
};

template <int NV>
__device__ __forceinline__ void write_partials(double (&v)[NV], double *__restrict__ partials, int tile,
                                               int first = 0) {
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) partials[(size_t)tile * kMaxPartials + first + k] = v[k];
    }
}

// Sum of a system's tile partials (finalize kernels: one kFinBlock-thread
// workgroup per system).  Thread t adds tiles t, t + kFinBlock, ... in
// ascending order, then a fixed tree combines the threads: the order depends
// on the system's tile count only.  The loads of kFinUnroll tiles are all
// issued before the first add (the partials come from other XCDs' writes, so
// every load is a long trip).
constexpr int kFinUnroll = 8;
template <int NV>
__device__ __forceinline__ void sum_partials(const SysDesc &d, const double *__restrict__ partials,
                                             double (&tot)[NV]) {
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    const double *base = partials + (size_t)d.tile0 * kMaxPartials;
    for (int q0 = threadIdx.x; q0 < d.ntiles; q0 += kFinUnroll * kFinBlock) {
        double v[kFinUnroll][NV];
#pragma unroll
        for (int u = 0; u < kFinUnroll; ++u) {
            const int q = q0 + u * kFinBlock;
#pragma unroll
            for (int k = 0; k < NV; ++k) v[u][k] = q < d.ntiles ? base[(size_t)q * kMaxPartials + k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kFinUnroll; ++u)
#pragma unroll
            for (int k = 0; k < NV; ++k)
                if (q0 + u * kFinBlock < d.ntiles) acc[k] += v[u][k];
    }
    block_sum<NV, kFinBlock>(acc);
#pragma unroll
    for (int k = 0; k < NV; ++k) tot[k] = acc[k];
}

// ------------------------------------------------------------------ init
// FP64: r = b - A x, rt = r; sums ||b||^2, ||r||^2.  MIXED: ||b||^2.
// ILU: every triangular-solve buffer starts at the sentinel.
template <typename V, int PC, int R>
__global__ __launch_bounds__(kBlock) void k_init(const SysDesc *__restrict__ sys,
                                                 const SysState *__restrict__ sts,
                                                 const int2 *__restrict__ tiles, const int *__restrict__ counts,
                                                 double *__restrict__ partials, const BatchParams *__restrict__ bpp) {
#pragma clang fp contract(off)
    const BatchParams bp = *bpp;
    const int ntiles = counts[0];   // tiles of the current membership
    KCFD_TILE_LOOP(tile, ntiles) {
        const int2 ti = tiles[tile];
        const SysDesc &d = sys[ti.x];
        if (sts[ti.x].status != S_NEW || d.n <= 0) continue;
        const TileRows<R> tr(ti.y, d.n);
        double v[2] = {0.0, 0.0};
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!tr.ok[r]) continue;
            const int i = tr.ix[r];
            const double bi = gl(d.b)[i];
            v[0] += bi * bi;
            if (bp.precision == KCFD_FP64) {
                const double ri = bi - row_ax64(d, i, d.x);
                vec<double>(d.r)[i] = ri;
                vec<double>(d.rt)[i] = ri;
                v[1] += ri * ri;
            }
            if (PC == KCFD_PC_ILU0) {
                vec<V>(d.y)[i] = sentinel_value<V>();
                vec<V>(d.ph)[i] = sentinel_value<V>();
                vec<V>(d.sh)[i] = sentinel_value<V>();
            }
        }
        write_partials<2>(v, partials, tile);
    }
}

__global__ __launch_bounds__(kFinBlock) void f_init(const SysDesc *__restrict__ sys,
                                                 SysState *__restrict__ sts,
                                                 const double *__restrict__ partials, const BatchParams *__restrict__ bpp) {
    const BatchParams bp = *bpp;
    const SysDesc &d = sys[blockIdx.x];
    SysState &st = sts[blockIdx.x];
    if (st.status != S_NEW) return;
    double tot[2];
    sum_partials<2>(d, partials, tot);
    if (threadIdx.x != 0) return;
    const double bnrm = sqrt(tot[0]);
    st.bnrm = bnrm;
    const double thr = fmax(bp.atol, bp.rtol * bnrm);
    st.iters = st.total_iters = st.refinements = st.half = st.info = 0;
    st.stagnant = st.final_ = 0;
    if (bnrm == 0.0) {          // scipy returns x = b (= 0), info 0
        st.status = S_ZEROX;
        return;
    }
    if (bp.precision == KCFD_FP64) {
        st.atol_in = thr;
        st.rr = tot[1];
        if (st.maxiter <= 0 || sqrt(tot[1]) < thr) { st.status = S_DONE; return; }
        if (bp.method == KCFD_METHOD_CG) { st.status = S_START; return; }   // iteration 0: p = z
        st.rho = tot[1];           // rt = r, so rt.r = ||r||^2
        if (fabs(st.rho) < bp.rhotol) { st.info = KCFD_INFO_RHO_BREAKDOWN; st.status = S_DONE; return; }
        st.status = S_START;       // iteration 0 next: p = r
    } else {
        st.atol_out = thr;
        st.atol_in = thr;
        st.restart = 1;
        st.rbest = INFINITY;
        st.status = S_REFINE;
    }
}

// -------------------------------------------------------------------- K1
// p-update / p = r / x = 0 for a zero rhs; MIXED: xb = x after an improving
// reliable update, solve buffers reset before one.
// Block 0 also clears the chunk counters of the two p triangular solves.
template <typename V, int PC, int R>
__global__ __launch_bounds__(kBlock) void k_p(const SysDesc *__restrict__ sys,
                                              const SysState *__restrict__ sts,
                                              const int2 *__restrict__ tiles, const int *__restrict__ counts,
                                              unsigned int *__restrict__ counters) {
#pragma clang fp contract(off)
    if (blockIdx.x == 0 && threadIdx.x < 2) counters[threadIdx.x] = 0u;
    const int ntiles = counts[0];   // tiles of the current membership
    KCFD_TILE_LOOP(tile, ntiles) {
        const int2 ti = tiles[tile];
        const SysDesc &d = sys[ti.x];
        const SysState &st = sts[ti.x];
        const int status = st.status;
        if (d.n <= 0) continue;
        const TileRows<R> tr(ti.y, d.n);
        if (st.save) {                     // MIXED: the last update improved on the best
            double xv[R];
#pragma unroll
            for (int r = 0; r < R; ++r) xv[r] = glw(d.x)[tr.ix[r]];
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (tr.ok[r]) glw(d.xb)[tr.ix[r]] = xv[r];
        }
        if (status == S_RUN || status == S_RESUME) {
            KCFD_GLOBAL V *p = vec<V>(d.p);
            const KCFD_GLOBAL V *vv = vec<V>(d.v), *rr = vec<V>(d.r);
            const V om = (V)st.omega, be = (V)st.beta;
            V pv[R], vl[R], rv[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                pv[r] = p[tr.ix[r]];
                vl[r] = vv[tr.ix[r]];
                rv[r] = rr[tr.ix[r]];
            }
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (tr.ok[r]) {
                    V pi = pv[r] - om * vl[r];
                    pi = pi * be;
                    p[tr.ix[r]] = pi + rv[r];
                }
        } else if (status == S_START) {
            V rv[R];
#pragma unroll
            for (int r = 0; r < R; ++r) rv[r] = vec<V>(d.r)[tr.ix[r]];
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (tr.ok[r]) vec<V>(d.p)[tr.ix[r]] = rv[r];
        } else if (status == S_REFINE) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (!tr.ok[r]) continue;
                const int i = tr.ix[r];
                if (st.rollback) glw(d.x)[i] = glw(d.xb)[i];   // the inner iteration diverged
                if (PC == KCFD_PC_ILU0 && d.polled) {
                    vec<V>(d.ph)[i] = sentinel_value<V>();
                    vec<V>(d.sh)[i] = sentinel_value<V>();
                    vec<V>(d.y)[i] = sentinel_value<V>();
                }
            }
        } else if (status == S_ZEROX) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (tr.ok[r]) glw(d.x)[tr.ix[r]] = 0.0;
        }
    }
}

// -------------------------------------------------------------------- K2
// v = A M p; rt.v.   MIXED S_REFINE: r64 = b - A64 x; r = (V)r64 (and rt = r
// on a restart); sums ||r64||^2, rt.r, ||r||^2.
template <typename V, int PC, int R>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_spmv_p(const SysDesc *__restrict__ sys,
                                                   const SysState *__restrict__ sts,
                                                   const int2 *__restrict__ tiles, const int *__restrict__ counts,
                                                   double *__restrict__ partials) {
#pragma clang fp contract(off)
    const int ntiles = counts[0];   // tiles of the current membership
    KCFD_TILE_LOOP_W(tile, ntiles, counts[5] | counts[6]) {
        const int2 ti = tiles[tile];
        const SysDesc &d = sys[ti.x];
        const SysState &st = sts[ti.x];
        const int status = st.status;
        if (d.n <= 0) continue;
        const TileRows<R> tr(ti.y, d.n);
        if (status == S_RUN || status == S_START || status == S_RESUME) {
            // one row after the other (the store orders them): a row's loads
            // already fill the wave's memory pipeline, and two rows in flight
            // would cost occupancy
            double v[1] = {0.0};
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = tr.ix[r];
                if (PC == KCFD_PC_ILU0 && i < d.nid) {   // folded: the p apply wrote v_i
                    const double rtv = (double)vec<V>(d.rt)[i], vv = (double)vec<V>(d.v)[i];
                    if (tr.ok[r]) v[0] += rtv * vv;
                    continue;
                }
                const double a = row_amz<V, PC>(d, i, (const V *)d.p, (const V *)d.ph);
                const double rtv = (double)vec<V>(d.rt)[i];
                if (tr.ok[r]) {
                    const V vv = (V)a;
                    vec<V>(d.v)[i] = vv;
                    v[0] += rtv * (double)vv;
                }
            }
            write_partials<1>(v, partials, tile);
        } else if (status == S_REFINE) {
            double r64[R], rto[R];
#pragma unroll
            for (int r = 0; r < R; ++r) r64[r] = gl(d.b)[tr.ix[r]] - row_ax64(d, tr.ix[r], d.x);
#pragma unroll
            for (int r = 0; r < R; ++r) rto[r] = (double)vec<V>(d.rt)[tr.ix[r]];
            double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (tr.ok[r]) {
                    const V rv = (V)r64[r];
                    vec<V>(d.r)[tr.ix[r]] = rv;
                    if (st.restart) {
                        vec<V>(d.rt)[tr.ix[r]] = rv;
                        rto[r] = (double)rv;
                    }
                    v[0] += r64[r] * r64[r];
                    v[1] += rto[r] * (double)rv;
                    v[2] += (double)rv * (double)rv;
                }
            write_partials<3>(v, partials, tile);
        }
    }
}

__global__ __launch_bounds__(kFinBlock) void f_spmv_p(const SysDesc *__restrict__ sys,
                                                   SysState *__restrict__ sts,
                                                   const double *__restrict__ partials, const BatchParams *__restrict__ bpp) {
    const BatchParams bp = *bpp;
    const SysDesc &d = sys[blockIdx.x];
    SysState &st = sts[blockIdx.x];
    const int status = st.status;
    const bool mixed = bp.precision == KCFD_MIXED;
    if (status == S_RUN || status == S_START || status == S_RESUME) {
        double tot[1];
        sum_partials<1>(d, partials, tot);
        if (threadIdx.x != 0) return;
        st.save = 0;
        const double rv = tot[0];
        if (rv == 0.0) {
            st.info = KCFD_INFO_OMEGA_BREAKDOWN;
            if (!mixed || st.iters == 0) { st.status = S_DONE; return; }
            st.restart = 1;
            st.status = S_REFINE;
            return;
        }
        st.alpha = st.rho / rv;
        st.status = S_RUN;
    } else if (status == S_REFINE) {
        double tot[3];
        sum_partials<3>(d, partials, tot);
        if (threadIdx.x != 0) return;
        st.save = 0;
        st.rollback = 0;
        st.rr64 = tot[0];
        st.refinements += 1;
        const double rn = sqrt(tot[0]);
        if (rn < st.atol_out) { st.info = 0; st.status = S_DONE; return; }
        // A stop without convergence returns the best iterate seen (xb).
        if (st.final_) {
            st.info = st.total_iters > 0 ? st.total_iters : 1;
            st.restore = rn > st.rbest;
            st.status = S_DONE;
            return;
        }
        if (rn < st.rbest) {
            st.rbest = rn;
            st.best_iter = st.total_iters;
            st.save = 1;
        } else if (st.total_iters - st.best_iter > max(200, st.best_iter) ||
                   st.refinements >= bp.max_refine) {
            // stagnation: more iterations without a better true residual than
            // it took to reach the best one (at least 200)
            st.info = st.total_iters > 0 ? st.total_iters : 1;
            st.restore = 1;
            st.status = S_DONE;
            return;
        }
        st.rmax = sqrt(tot[2]);
        st.rr = tot[2];
        if (st.restart) {
            st.rho = tot[1];
            st.iters = 0;
            st.restart = 0;
            if (fabs(st.rho) < bp.rhotol) { st.info = KCFD_INFO_RHO_BREAKDOWN; st.status = S_DONE; return; }
            st.status = S_START;
        } else {
            const double rho = tot[1];
            if (fabs(rho) < bp.rhotol || fabs(st.omega) < bp.rhotol) {
                st.restart = 1;        // redo the replacement with a fresh shadow residual
                st.status = S_REFINE;
                return;
            }
            st.beta = (rho / st.rho) * (st.alpha / st.omega);
            st.rho_prev = st.rho;
            st.rho = rho;
            st.status = S_RESUME;
        }
    } else if (status == S_ZEROX) {
        if (threadIdx.x == 0) { st.info = 0; st.status = S_DONE; }
    }
}

// -------------------------------------------------------------------- K3
// s = r - alpha v (Jacobi / none; with ILU(0) the s apply forms s itself).
// ||s||^2 is summed by K4 over the same rows in the same order.
template <typename V, int PC, int R>
__global__ __launch_bounds__(kBlock) void k_s(const SysDesc *__restrict__ sys,
                                              const SysState *__restrict__ sts,
                                              const int2 *__restrict__ tiles, const int *__restrict__ counts) {
#pragma clang fp contract(off)
    const int ntiles = counts[0];   // tiles of the current membership
    KCFD_TILE_LOOP(tile, ntiles) {
        const int2 ti = tiles[tile];
        const SysDesc &d = sys[ti.x];
        const SysState &st = sts[ti.x];
        if (st.status != S_RUN || d.n <= 0) continue;
        const TileRows<R> tr(ti.y, d.n);
        const V al = (V)st.alpha;
        V rv[R], vl[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            rv[r] = vec<V>(d.r)[tr.ix[r]];
            vl[r] = vec<V>(d.v)[tr.ix[r]];
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (tr.ok[r]) vec<V>(d.s)[tr.ix[r]] = rv[r] - al * vl[r];
    }
}

// -------------------------------------------------------------------- K4
// t = A M s; t.s, t.t, ||s||^2.
template <typename V, int PC, int R>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_spmv_s(const SysDesc *__restrict__ sys,
                                                   const SysState *__restrict__ sts,
                                                   const int2 *__restrict__ tiles, const int *__restrict__ counts,
                                                   double *__restrict__ partials, const BatchParams *__restrict__ bpp) {
#pragma clang fp contract(off)
    const int ntiles = counts[0];   // tiles of the current membership
    KCFD_TILE_LOOP_W(tile, ntiles, counts[5] | counts[6]) {
        const int2 ti = tiles[tile];
        const SysDesc &d = sys[ti.x];
        const SysState &st = sts[ti.x];
        if (st.status != S_RUN || d.n <= 0) continue;
        const TileRows<R> tr(ti.y, d.n);
        double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int r = 0; r < R; ++r) {   // one row after the other, as in K2
            const int i = tr.ix[r];
            double tv;
            if (PC == KCFD_PC_ILU0 && i < d.nid) {   // folded: the s apply wrote t_i
                tv = (double)vec<V>(d.t)[i];
            } else {
                tv = (double)(V)row_amz<V, PC>(d, i, (const V *)d.s, (const V *)d.sh);
                if (tr.ok[r]) vec<V>(d.t)[i] = (V)tv;
            }
            const double sv = (double)vec<V>(d.s)[i];
            if (tr.ok[r]) {
                v[0] += tv * sv;
                v[1] += tv * tv;
                v[2] += sv * sv;
            }
        }
        write_partials<3>(v, partials, tile);
    }
}

__global__ __launch_bounds__(kFinBlock) void f_spmv_s(const SysDesc *__restrict__ sys,
                                                   SysState *__restrict__ sts,
                                                   const double *__restrict__ partials, const BatchParams *__restrict__ bpp) {
    const BatchParams bp = *bpp;
    const SysDesc &d = sys[blockIdx.x];
    SysState &st = sts[blockIdx.x];
    const int status = st.status;
    if (status == S_RUN) {
        double tot[3];   // t.s, t.t (K4), ||s||^2 (K3)
        sum_partials<3>(d, partials, tot);
        if (threadIdx.x != 0) return;
        if (sqrt(tot[2]) < st.atol_in) {   // scipy's half-step exit: x += alpha M p (K5)
            st.status = S_HALF;
            return;
        }
        st.omega = tot[0] / tot[1];
    }
}

// -------------------------------------------------------------------- K5
// x += alpha M p; x += omega M s; r = s - omega t; ||r||^2, rt.r
template <typename V, int PC, int R>
__global__ __launch_bounds__(kBlock) void k_xr(const SysDesc *__restrict__ sys,
                                               const SysState *__restrict__ sts,
                                               const int2 *__restrict__ tiles, const int *__restrict__ counts,
                                               double *__restrict__ partials, const BatchParams *__restrict__ bpp) {
#pragma clang fp contract(off)
    const BatchParams bp = *bpp;
    const int ntiles = counts[0];   // tiles of the current membership
    KCFD_TILE_LOOP(tile, ntiles) {
        const int2 ti = tiles[tile];
        const SysDesc &d = sys[ti.x];
        const SysState &st = sts[ti.x];
        if (d.n <= 0) continue;
        const TileRows<R> tr(ti.y, d.n);
        if (st.status == S_HALF) {                   // half-step exit: x += alpha M p
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (!tr.ok[r]) continue;
                const int i = tr.ix[r];
                const V al = (V)st.alpha;
                const V ph = (V)precond_at<V, PC>(d, (const V *)d.p, (const V *)d.ph, i);
                if (bp.precision == KCFD_FP64)       // scipy's operation, in FP64
                    vec<V>(d.x)[i] = vec<V>(d.x)[i] + al * ph;
                else                                 // MIXED: FP32 direction, FP64 iterate
                    glw(d.x)[i] = glw(d.x)[i] + st.alpha * (double)ph;
                if (PC == KCFD_PC_ILU0 && d.polled) {
                    vec<V>(d.ph)[i] = sentinel_value<V>();
                    vec<V>(d.sh)[i] = sentinel_value<V>();
                }
            }
            continue;
        }
        if (st.status != S_RUN) continue;
        const V al = (V)st.alpha, om = (V)st.omega;
        V ph[R], sh[R], sv[R], tv[R], rtv[R];
        double xv[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tr.ix[r];
            ph[r] = (V)precond_at<V, PC>(d, (const V *)d.p, (const V *)d.ph, i);
            sh[r] = (V)precond_at<V, PC>(d, (const V *)d.s, (const V *)d.sh, i);
            sv[r] = vec<V>(d.s)[i];
            tv[r] = vec<V>(d.t)[i];
            rtv[r] = vec<V>(d.rt)[i];
            xv[r] = glw(d.x)[i];
        }
        double v[2] = {0.0, 0.0};
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!tr.ok[r]) continue;
            const int i = tr.ix[r];
            if (bp.precision == KCFD_FP64) {         // scipy's operation order
                V xi = (V)xv[r] + al * ph[r];
                xi = xi + om * sh[r];
                glw(d.x)[i] = (double)xi;
            } else {   // MIXED: the iterate accumulates in FP64 (an FP32 correction
                       // would carry an error of eps_32 |x| into the residual)
                double xi = xv[r] + st.alpha * (double)ph[r];
                xi = xi + st.omega * (double)sh[r];
                glw(d.x)[i] = xi;
            }
            const V ri = sv[r] - om * tv[r];
            vec<V>(d.r)[i] = ri;
            v[0] += (double)ri * (double)ri;
            v[1] += (double)rtv[r] * (double)ri;
            if (PC == KCFD_PC_ILU0 && d.polled) {   // ready flags of the sync-free solves
                vec<V>(d.ph)[i] = sentinel_value<V>();
                vec<V>(d.sh)[i] = sentinel_value<V>();
            }
        }
        write_partials<2>(v, partials, tile);
    }
}

__global__ __launch_bounds__(kFinBlock) void f_xr(const SysDesc *__restrict__ sys, SysState *__restrict__ sts,
                                               const double *__restrict__ partials, const BatchParams *__restrict__ bpp) {
    const BatchParams bp = *bpp;
    const SysDesc &d = sys[blockIdx.x];
    SysState &st = sts[blockIdx.x];
    if (st.status == S_HALF) {    // K5 applied the half step
        if (threadIdx.x != 0) return;
        st.half = 1;
        st.info = 0;
        if (bp.precision == KCFD_MIXED) {
            st.restart = 1;
            st.status = S_REFINE;
        } else {
            st.status = S_DONE;
        }
        return;
    }
    if (st.status != S_RUN) return;
    double tot[2];
    sum_partials<2>(d, partials, tot);
    if (threadIdx.x != 0) return;
    st.iters += 1;
    st.total_iters += 1;
    st.rr = tot[0];
    const double rn = sqrt(tot[0]);
    const double rho = tot[1];
    if (bp.precision == KCFD_FP64) {
        if (st.total_iters >= st.maxiter) {          // scipy's for-loop exhausted
            st.info = st.total_iters;
            st.status = S_DONE;
            return;
        }
        if (rn < st.atol_in) { st.info = 0; st.status = S_DONE; return; }
        if (fabs(rho) < bp.rhotol) { st.info = KCFD_INFO_RHO_BREAKDOWN; st.status = S_DONE; return; }
        if (fabs(st.omega) < bp.rhotol) { st.info = KCFD_INFO_OMEGA_BREAKDOWN; st.status = S_DONE; return; }
    } else {
        if (st.total_iters >= st.maxiter) { st.final_ = 1; st.status = S_REFINE; return; }
        // reliable update when the FP32 residual claims convergence or has
        // dropped by delta since the last update; restart on divergence,
        // breakdown or the inner iteration cap
        // (BiCGSTAB's residual spikes by orders of magnitude on hard systems;
        // only an explosion counts as divergence -- slow drift is caught by
        // the stagnation window at the next replacement)
        const bool diverged = !(rn == rn) || rn > 1e6 * st.rmax;
        if (diverged || fabs(rho) < bp.rhotol || fabs(st.omega) < bp.rhotol ||
            st.iters >= bp.inner_maxiter) {
            st.restart = 1;
            st.rollback = diverged;   // back to the best iterate
            st.status = S_REFINE;
            return;
        }
        if (rn < st.atol_out || rn < bp.delta * st.rmax) { st.status = S_REFINE; return; }
        st.rmax = fmax(st.rmax, rn);
    }
    st.beta = (rho / st.rho) * (st.alpha / st.omega);
    st.rho_prev = st.rho;
    st.rho = rho;
}

// ------------------------------------------------ conjugate gradients
// scipy's cg (scipy/sparse/linalg/_isolve/iterative.py, the solver behind
// solve_pressure_correction, core/linear_solver.py:88-93), in FP64, with its
// operation order (p = p beta + z; x += alpha p; r -= alpha q):
//   ILU(0): z = M r by the triangular-solve kernels (into ph)
//   C1 z = M r (Jacobi / none inline); r.z        FC1: rho, beta = rho / rho_prev
//   C2 p = z (iteration 0) | p = p beta + z
//   C3 q = A p; p.q                               FC3: alpha = rho / p.q
//   C4 x += alpha p; r -= alpha q; ||r||^2         FC4: iteration count; stop at
//      maxiter (scipy's loop exhausted) or when ||r|| < atol (its next check)
// q lives in v.  A zero rhs: C1 sets x = 0, FC1 ends the solve (x = b = 0).
template <int PC>
__device__ __forceinline__ double cg_z(const SysDesc &d, int i) {
#pragma clang fp contract(off)
    if (PC == KCFD_PC_JACOBI) return gl(d.dinv64)[i] * vec<double>(d.r)[i];
    if (PC == KCFD_PC_ILU0) return vec<double>(d.ph)[i];
    return vec<double>(d.r)[i];
}

template <int PC, int R>
__global__ __launch_bounds__(kBlock) void k_cg_rz(const SysDesc *__restrict__ sys,
                                                  const SysState *__restrict__ sts,
                                                  const int2 *__restrict__ tiles, const int *__restrict__ counts,
                                                  double *__restrict__ partials) {
#pragma clang fp contract(off)
    const int ntiles = counts[0];   // tiles of the current membership
    KCFD_TILE_LOOP(tile, ntiles) {
        const int2 ti = tiles[tile];
        const SysDesc &d = sys[ti.x];
        const int status = sts[ti.x].status;
        if (d.n <= 0) continue;
        const TileRows<R> tr(ti.y, d.n);
        if (status == S_ZEROX) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (tr.ok[r]) glw(d.x)[tr.ix[r]] = 0.0;
            continue;
        }
        if (status != S_RUN && status != S_START) continue;
        double rv[R], zv[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            rv[r] = vec<double>(d.r)[tr.ix[r]];
            zv[r] = cg_z<PC>(d, tr.ix[r]);
        }
        double v[1] = {0.0};
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (tr.ok[r]) v[0] += rv[r] * zv[r];
        write_partials<1>(v, partials, tile);
    }
}

__global__ __launch_bounds__(kFinBlock) void f_cg_rho(const SysDesc *__restrict__ sys,
                                                    SysState *__restrict__ sts,
                                                    const double *__restrict__ partials) {
    const SysDesc &d = sys[blockIdx.x];
    SysState &st = sts[blockIdx.x];
    if (st.status == S_ZEROX) {
        if (threadIdx.x == 0) { st.info = 0; st.status = S_DONE; }
        return;
    }
    if (st.status != S_RUN && st.status != S_START) return;
    double tot[1];
    sum_partials<1>(d, partials, tot);
    if (threadIdx.x != 0) return;
    st.rho = tot[0];
    if (st.status == S_RUN) st.beta = st.rho / st.rho_prev;
}

template <int PC, int R>
__global__ __launch_bounds__(kBlock) void k_cg_p(const SysDesc *__restrict__ sys,
                                                 const SysState *__restrict__ sts,
                                                 const int2 *__restrict__ tiles, const int *__restrict__ counts) {
#pragma clang fp contract(off)
    const int ntiles = counts[0];   // tiles of the current membership
    KCFD_TILE_LOOP(tile, ntiles) {
        const int2 ti = tiles[tile];
        const SysDesc &d = sys[ti.x];
        const SysState &st = sts[ti.x];
        const int status = st.status;
        if ((status != S_RUN && status != S_START) || d.n <= 0) continue;
        const TileRows<R> tr(ti.y, d.n);
        double zv[R], pv[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            zv[r] = cg_z<PC>(d, tr.ix[r]);
            pv[r] = vec<double>(d.p)[tr.ix[r]];
        }
        const double beta = st.beta;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!tr.ok[r]) continue;
            const int i = tr.ix[r];
            double pi = zv[r];
            if (status == S_RUN) {
                pi = pv[r] * beta;
                pi = pi + zv[r];
            }
            vec<double>(d.p)[i] = pi;
            if (PC == KCFD_PC_ILU0 && d.polled) vec<double>(d.ph)[i] = sentinel_value<double>();
        }
    }
}

template <int PC, int R>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_cg_q(
    const SysDesc *__restrict__ sys, const SysState *__restrict__ sts, const int2 *__restrict__ tiles,
    const int *__restrict__ counts, double *__restrict__ partials) {
#pragma clang fp contract(off)
    const int ntiles = counts[0];   // tiles of the current membership
    KCFD_TILE_LOOP(tile, ntiles) {
        const int2 ti = tiles[tile];
        const SysDesc &d = sys[ti.x];
        const int status = sts[ti.x].status;
        if ((status != S_RUN && status != S_START) || d.n <= 0) continue;
        const TileRows<R> tr(ti.y, d.n);
        double v[1] = {0.0};
#pragma unroll
        for (int r = 0; r < R; ++r) {   // one row after the other, as in K2
            const double q = row_amz<double, KCFD_PC_NONE>(d, tr.ix[r], (const double *)d.p,
                                                           (const double *)d.p);
            const double pv = vec<double>(d.p)[tr.ix[r]];
            if (tr.ok[r]) {
                vec<double>(d.v)[tr.ix[r]] = q;
                v[0] += pv * q;
            }
        }
        write_partials<1>(v, partials, tile);
    }
}

__global__ __launch_bounds__(kFinBlock) void f_cg_alpha(const SysDesc *__restrict__ sys,
                                                      SysState *__restrict__ sts,
                                                      const double *__restrict__ partials) {
    const SysDesc &d = sys[blockIdx.x];
    SysState &st = sts[blockIdx.x];
    if (st.status != S_RUN && st.status != S_START) return;
    double tot[1];
    sum_partials<1>(d, partials, tot);
    if (threadIdx.x != 0) return;
    const double pq = tot[0];
    if (pq == 0.0 || !(pq == pq)) {   // scipy would divide by zero and carry NaN to maxiter
        st.info = KCFD_INFO_PQ_BREAKDOWN;
        st.status = S_DONE;
        return;
    }
    st.alpha = st.rho / pq;
}

template <int R>
__global__ __launch_bounds__(kBlock) void k_cg_xr(const SysDesc *__restrict__ sys,
                                                  const SysState *__restrict__ sts,
                                                  const int2 *__restrict__ tiles, const int *__restrict__ counts,
                                                  double *__restrict__ partials) {
#pragma clang fp contract(off)
    const int ntiles = counts[0];   // tiles of the current membership
    KCFD_TILE_LOOP(tile, ntiles) {
        const int2 ti = tiles[tile];
        const SysDesc &d = sys[ti.x];
        const SysState &st = sts[ti.x];
        if ((st.status != S_RUN && st.status != S_START) || d.n <= 0) continue;
        const TileRows<R> tr(ti.y, d.n);
        double xv[R], pv[R], rv[R], qv[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tr.ix[r];
            xv[r] = glw(d.x)[i];
            pv[r] = vec<double>(d.p)[i];
            rv[r] = vec<double>(d.r)[i];
            qv[r] = vec<double>(d.v)[i];
        }
        const double alpha = st.alpha;
        double v[1] = {0.0};
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!tr.ok[r]) continue;
            const int i = tr.ix[r];
            glw(d.x)[i] = xv[r] + alpha * pv[r];
            const double ri = rv[r] - alpha * qv[r];
            vec<double>(d.r)[i] = ri;
            v[0] += ri * ri;
        }
        write_partials<1>(v, partials, tile);
    }
}

__global__ __launch_bounds__(kFinBlock) void f_cg_check(const SysDesc *__restrict__ sys,
                                                      SysState *__restrict__ sts,
                                                      const double *__restrict__ partials) {
    const SysDesc &d = sys[blockIdx.x];
    SysState &st = sts[blockIdx.x];
    if (st.status != S_RUN && st.status != S_START) return;
    double tot[1];
    sum_partials<1>(d, partials, tot);
    if (threadIdx.x != 0) return;
    st.iters += 1;
    st.total_iters += 1;
    st.rho_prev = st.rho;
    st.rr = tot[0];
    if (st.total_iters >= st.maxiter) {           // scipy's for-loop exhausted
        st.info = st.total_iters;
        st.status = S_DONE;
        return;
    }
    if (sqrt(tot[0]) < st.atol_in) { st.info = 0; st.status = S_DONE; return; }
    st.status = S_RUN;
}

// ------------------------------------------------------- true residual
// ||b - A x||^2 and ||b||^2 in FP64 on the FP64 matrix, every system.
// Systems that are done and whose true residual is not yet computed.
__device__ __forceinline__ bool wants_true(const SysState &st) {
    return st.status == S_DONE && !st.true_done;
}

// MIXED stop without convergence: x = xb (the best iterate), in a launch of
// its own -- the true residual reads the neighbours' x.
template <int R>
__global__ __launch_bounds__(kBlock) void k_restore(const SysDesc *__restrict__ sys,
                                                    const SysState *__restrict__ sts,
                                                    const int2 *__restrict__ tiles,
                                                    const int *__restrict__ counts) {
    const int ntiles = counts[0];   // tiles of the current membership
    KCFD_TILE_LOOP(tile, ntiles) {
        const int2 ti = tiles[tile];
        const SysState &st = sts[ti.x];
        if (!wants_true(st) || !st.restore) continue;
        const SysDesc &d = sys[ti.x];
        if (d.n <= 0) continue;
        const TileRows<R> tr(ti.y, d.n);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (tr.ok[r]) glw(d.x)[tr.ix[r]] = glw(d.xb)[tr.ix[r]];
    }
}

template <int R>
__global__ __launch_bounds__(kBlock) void k_true_resid(const SysDesc *__restrict__ sys,
                                                       const SysState *__restrict__ sts,
                                                       const int2 *__restrict__ tiles, const int *__restrict__ counts,
                                                       double *__restrict__ partials) {
#pragma clang fp contract(off)
    const int ntiles = counts[0];   // tiles of the current membership
    KCFD_TILE_LOOP(tile, ntiles) {
        const int2 ti = tiles[tile];
        const SysDesc &d = sys[ti.x];
        if (!wants_true(sts[ti.x]) || d.n <= 0) continue;
        const TileRows<R> tr(ti.y, d.n);
        double v[2] = {0.0, 0.0};
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!tr.ok[r]) continue;
            const double bi = gl(d.b)[tr.ix[r]];
            const double ri = bi - row_ax64(d, tr.ix[r], d.x);
            v[0] += ri * ri;
            v[1] += bi * bi;
        }
        write_partials<2>(v, partials, tile);
    }
}

__global__ __launch_bounds__(kFinBlock) void f_true_resid(const SysDesc *__restrict__ sys,
                                                       SysState *__restrict__ sts,
                                                       const double *__restrict__ partials) {
    const SysDesc &d = sys[blockIdx.x];
    SysState &st = sts[blockIdx.x];
    if (!wants_true(st)) {
        if (threadIdx.x == 0) st.scatter = 0;   // scattered by an earlier k_perm_out
        return;
    }
    double tot[2];
    sum_partials<2>(d, partials, tot);
    if (threadIdx.x != 0) return;
    st.true_rr = tot[0];
    st.true_bb = tot[1];
    st.true_done = 1;
    st.scatter = 1;
}

// Permuted row order: b and the initial x gathered from the caller's
// buffers for systems about to start (before k_init, which reads x at
// neighbouring rows) ...
template <int R>
__global__ __launch_bounds__(kBlock) void k_perm_in(const SysDesc *__restrict__ sys,
                                                    const SysState *__restrict__ sts,
                                                    const int2 *__restrict__ tiles,
                                                    const int *__restrict__ counts) {
    const int ntiles = counts[0];   // tiles of the current membership
    KCFD_TILE_LOOP(tile, ntiles) {
        const int2 ti = tiles[tile];
        const SysDesc &d = sys[ti.x];
        if (!d.perm || sts[ti.x].status != S_NEW || d.n <= 0) continue;
        const TileRows<R> tr(ti.y, d.n);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!tr.ok[r]) continue;
            const int i = tr.ix[r], q = gl(d.perm)[i];
            glw(const_cast<double *>(d.b))[i] = gl(d.b_user)[q];
            glw(d.x)[i] = glw(d.x_user)[q];
        }
    }
}

// ... and the solution scattered back once its true residual is in.
template <int R>
__global__ __launch_bounds__(kBlock) void k_perm_out(const SysDesc *__restrict__ sys,
                                                     const SysState *__restrict__ sts,
                                                     const int2 *__restrict__ tiles,
                                                     const int *__restrict__ counts) {
    const int ntiles = counts[0];   // tiles of the current membership
    KCFD_TILE_LOOP(tile, ntiles) {
        const int2 ti = tiles[tile];
        const SysDesc &d = sys[ti.x];
        if (!d.perm || !sts[ti.x].scatter || d.n <= 0) continue;
        const TileRows<R> tr(ti.y, d.n);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (tr.ok[r]) glw(d.x_user)[gl(d.perm)[tr.ix[r]]] = glw(d.x)[tr.ix[r]];
    }
}

// Operands of the ILU(0) applies.  PlainSrc reads a vector.  The s apply
// forms s = r - alpha v where it reads it, with scipy's operations (the
// s-update of the BiCGSTAB loop), so s is not a pass of its own; the own
// row's value is stored by the apply.
template <typename V>
struct PlainSrc {
    const KCFD_GLOBAL V *__restrict__ a;
    __device__ __forceinline__ V operator()(int c) const { return a[c]; }
};
// s = r - alpha v
template <typename V>
struct SSrc {
    const KCFD_GLOBAL V *__restrict__ r, *__restrict__ v;
    V al;
    __device__ __forceinline__ V operator()(int c) const {
#pragma clang fp contract(off)
        return r[c] - al * v[c];
    }
};

// ---------------------------------------------------- ILU(0) trisolves
//
// Sync-free level-scheduled triangular solves: one launch per solve covers
// all systems; 64-row chunks (one per wave, never spanning a level) are
// listed in level-major order and dealt round-robin to the waves (wave w
// takes chunks w, w + W, w + 2W, ...).  The grid is small enough to be
// resident at once, so the lowest unfinished chunk always belongs to a
// running wave whose earlier chunks are done and whose inputs (lower
// levels, lower chunks) are ready: the solve cannot deadlock.  (A shared
// atomic ticket would serialise tens of thousands of same-address atomics
// on the wide levels of a colour ordering.)  The output
// vector doubles as the ready flag: it holds a NaN sentinel until the row is
// written (agent-scope atomic store) and waiters poll it with agent-scope
// atomic loads.  Row arithmetic follows the oracle's order exactly.
// Rows are positions i of P A P^T, the solver's row order (every vector is
// in that order):
//   LOWER: y_i = z_i - sum_{j<i} l_ij y_j                   (unit L)
//   UPPER: w_i = (y_i - sum_{j>i} u_ij w_j) / u_ii; y_i reset
template <typename V>
__device__ __forceinline__ void publish(V *p, V v) {
    using U = typename Sentinel<V>::U;
    __hip_atomic_store(reinterpret_cast<U *>(p), __builtin_bit_cast(U, v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// acc - sum_k a_k vec[col_k] over one row of a chunk (entries at base + 64k,
// k < w, in CSR order; column -1 marks padding), subtracting in k order.
// Coalesced: the 64 lanes of a wave read 64 consecutive entries.  Groups of
// 8 entries: their indices and values, then their (polled) operands are all
// requested before the first subtraction, so a row costs about one memory
// round trip per group instead of one per entry.
// The wait for an operand is bounded: after `limit` polls that still find
// the sentinel the lane stops waiting, raises the fail word (the host turns
// it into KCFD_ESTATE) and uses a quiet NaN, so that its own row is still
// published and every wave of the launch runs to its end.
template <typename V>
__device__ __forceinline__ double sub_ready(double acc, const FacRef &f, int base, int w, const V *vec, int limit,
                                            unsigned int *fail) {
#pragma clang fp contract(off)
    using U = typename Sentinel<V>::U;
    constexpr int G = 8;
    for (int kb = 0; kb < w; kb += G) {
        int c[G];
        V av[G];
        U bits[G];
#pragma unroll
        for (int k = 0; k < G; ++k) fac_entry<V>(f, base + 64 * min(kb + k, w - 1), c[k], av[k]);
#pragma unroll
        for (int k = 0; k < G; ++k)   // padding (column -1) polls entry 0 and ignores it
            bits[k] = __hip_atomic_load(reinterpret_cast<const U *>(vec + max(c[k], 0)), __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int k = 0; k < G; ++k)
            if (kb + k < w && c[k] >= 0) {
                int polls = 0;
                while (bits[k] == Sentinel<V>::bits) {
                    if (polls++ >= limit) {
                        __hip_atomic_store(fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        bits[k] = __builtin_bit_cast(U, (V)NAN);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                    bits[k] = __hip_atomic_load(reinterpret_cast<const U *>(vec + c[k]), __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
                }
                acc -= (double)av[k] * (double)__builtin_bit_cast(V, bits[k]);
            }
    }
    return acc;
}

template <typename V, bool UPPER, int WHICH>
__global__ __launch_bounds__(kBlock) void k_trsv(const SysDesc *__restrict__ sys,
                                                 const SysState *__restrict__ sts,
                                                 const Chunk *__restrict__ chunks,
                                                 const int *__restrict__ counts,
                                                 const int *const *__restrict__ lvl_rows,
                                                 const BatchParams *__restrict__ bpp,
                                                 unsigned int *__restrict__ counters) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int limit = bpp->poll_limit;
    unsigned int *fail = counters + kTrsvFail;
    // The chunk loop is wave-uniform (the chunk index depends on the wave
    // only); per-lane work sits inside an if.
    const int nchunks = counts[UPPER ? 2 : 1];
    const int nwaves = (int)(gridDim.x * (blockDim.x >> 6));
    const int wave = (int)__builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    for (int c = wave; c < nchunks; c += nwaves) {
        const Chunk ch = chunks[c];
        const SysDesc &d = sys[ch.sys()];
        const int status = sts[ch.sys()].status;
        // the p solves (after K1) serve systems in iteration 0 or later; the s
        // solves (after K3) only systems still running -- a system that the
        // refinement just restarted (S_START set inside F2) must not run them,
        // or stale values would be left in sh
        const bool want = WHICH == 0 ? (status == S_RUN || status == S_START || status == S_RESUME)
                                     : (WHICH == 1 ? status == S_RUN : (status == S_RUN || status == S_START));
        const bool live = want && lane < ch.count();
        if (live) {
            const int i = gl(lvl_rows[ch.sys()])[ch.start + lane];
            const int *off = UPPER ? d.chu_off : d.chl_off;
            const int o0 = ch.o0, w = ch.wd();
            if (!UPPER) {
                // the own operand: s is formed here (and stored), as in the
                // two-colour apply
                V zi;
                const SysState &st = sts[ch.sys()];
                if constexpr (WHICH == 0) {
                    zi = vec<V>(d.p)[i];
                } else if constexpr (WHICH == 1) {
                    zi = SSrc<V>{vec<V>(d.r), vec<V>(d.v), (V)st.alpha}(i);
                    vec<V>(d.s)[i] = zi;
                } else {
                    zi = vec<V>(d.r)[i];
                }
                V *y = (V *)d.y;
                const V acc = (V)sub_ready<V>((double)zi, fac_l(d), o0 + lane, w, y, limit, fail);
                publish<V>(y + i, acc);
            } else {
                V *y = (V *)d.y;
                V *wv = (V *)(WHICH == 1 ? d.sh : d.ph);
                const double acc = sub_ready<V>((double)glw(y)[i], fac_u(d), o0 + lane, w, wv, limit, fail);
                publish<V>(wv + i, (V)(acc / (double)gl(udiag<V>(d))[i]));
                glw(y)[i] = sentinel_value<V>();
            }
        }
    }
}

// acc - sum_k a_k vec[col_k] over one chunk row (layout as sub_ready), with
// plain loads: every operand was written by an earlier launch.  As in
// row_amz: fixed groups for the common widths, else groups of 8 whose loads
// past the width re-load the last entry; padding (column -1) gathers entry
// 0.  Neither is accumulated.
template <typename V, int G, typename F, typename S>
__device__ __forceinline__ void sub_group(double &acc, const F &f, int base, int k0, int last, int valid,
                                          const S &src) {
#pragma clang fp contract(off)
    int c[G];
    V av[G], xv[G];
#pragma unroll
    for (int k = 0; k < G; ++k) fac_get<V>(f, base + 64 * min(k0 + k, last), c[k], av[k]);
#pragma unroll
    for (int k = 0; k < G; ++k) xv[k] = src(max(c[k], 0));
#pragma unroll
    for (int k = 0; k < G; ++k)
        if (k < valid && c[k] >= 0) acc -= (double)av[k] * (double)xv[k];
}

template <typename V, typename F, typename S>
__device__ __forceinline__ double sub_plain(double acc, const F &f, int base, int w, const S &src) {
    if (w == 6) {
        sub_group<V, 6>(acc, f, base, 0, 5, 6, src);
    } else if (w == 4) {
        sub_group<V, 4>(acc, f, base, 0, 3, 4, src);
    } else {
        constexpr int G = 8;
        for (int kb = 0; kb < w; kb += G) sub_group<V, G>(acc, f, base, kb, w - 1, w - kb, src);
    }
    return acc;
}

// ILU(0) apply for a two-colour ordering, P A P^T = [[D0, B], [C, D1]] with
// colour 0 first (vectors in P-order).  Colour-0 rows have no L entries, so
// their y is z; colour-1 rows have no U entries.  Hence, exactly as the
// sequential L then U solves compute it (same operations, same order):
//   FIRST  (colour-1 rows i): y_i = z_i - sum_j l_ij z_j;  w_i = y_i / u_ii
//   SECOND (colour-0 rows i): w_i = (z_i - sum_j u_ij w_j) / u_ii
// Two plain launches per apply: no y vector, no ready flags.  Chunks are runs
// of 64 rows of one colour, one wave per chunk.  As in the tile walk, the
// chunk list is cut into one contiguous range per XCD (blockIdx % 8), whose
// waves take its chunks round-robin: a range's gathers stay in its XCD's
// L2.  Launched on the persistent grid (a multiple of 8 blocks).
// One row i of a two-colour apply.  The own operand z_i comes from src and
// is stored to zout (the p or s vector) when given; FIRST gathers the
// colour-0 operands from src, SECOND the colour-1 results from w.  The
// factor layout is chosen once per chunk (wave-uniform), so each form is
// straight-line code.
template <typename V, bool SECOND, typename S>
__device__ __forceinline__ void ilu2_row(const SysDesc &d, const Chunk &ch, int i, int o0, int wd, const S &src,
                                         KCFD_GLOBAL V *__restrict__ zout, KCFD_GLOBAL V *__restrict__ w) {
#pragma clang fp contract(off)
    const V zi = src(i);
    if (zout) zout[i] = zi;
    const PlainSrc<V> wsrc{w};
    double v;
    const unsigned short *o16 = nullptr;
    if constexpr (sizeof(V) == sizeof(float)) o16 = SECOND ? d.chu_o16 : d.chl_o16;
    if (o16) {
        const FacRef16 f{SECOND ? d.chu_v32 : d.chl_v32, o16, ch.cb};
        v = SECOND ? sub_plain<V>((double)zi, f, o0 + (i - ch.start), wd, wsrc)
                   : sub_plain<V>((double)zi, f, o0 + (i - ch.start), wd, src);
    } else {
        v = SECOND ? sub_plain<V>((double)zi, fac_u(d), o0 + (i - ch.start), wd, wsrc)
                   : sub_plain<V>((double)zi, fac_l(d), o0 + (i - ch.start), wd, src);
    }
    w[i] = (V)(v / (double)gl(udiag<V>(d))[i]);
}

// A colour-0 row of a folded apply (SECOND), which also writes the row's
// product (A M z)_i -- v for the p apply, t for the s apply.  Colour-0
// rows of U hold A's own off-diagonal entries (no L entries, so the
// factorisation leaves them alone), so the product needs nothing the apply
// does not already read: (A M z)_i = P_i + D_i w_i with P_i = sum_j u_ij w_j,
// summed here in FP64 beside the apply's own sum, w_i the apply's result
// and D_i the diagonal: a_ii (FP64), or in MIXED the FP64 pivot
// a_ii + sum_j (a_ij - fl32(a_ij)) (kcfd_mat::c0piv), for which the row
// [D_i, fl32(a_ij)] keeps A's row sum, as the difference-form copy of A
// does.  The apply's own arithmetic is unchanged.
template <typename V, int G, typename F, typename S>
__device__ __forceinline__ void sub_group_prod(double &acc, double &prod, const F &f, int base, int k0, int last,
                                               int valid, const S &src) {
#pragma clang fp contract(off)
    int c[G];
    V av[G], xv[G];
#pragma unroll
    for (int k = 0; k < G; ++k) fac_get<V>(f, base + 64 * min(k0 + k, last), c[k], av[k]);
#pragma unroll
    for (int k = 0; k < G; ++k) xv[k] = src(max(c[k], 0));
#pragma unroll
    for (int k = 0; k < G; ++k)
        if (k < valid && c[k] >= 0) {
            acc -= (double)av[k] * (double)xv[k];
            prod += (double)av[k] * (double)xv[k];
        }
}

template <typename V, typename F, typename S>
__device__ __forceinline__ double sub_plain_prod(double acc, double &prod, const F &f, int base, int w, const S &src) {
    if (w == 6) {
        sub_group_prod<V, 6>(acc, prod, f, base, 0, 5, 6, src);
    } else if (w == 4) {
        sub_group_prod<V, 4>(acc, prod, f, base, 0, 3, 4, src);
    } else {
        constexpr int G = 8;
        for (int kb = 0; kb < w; kb += G) sub_group_prod<V, G>(acc, prod, f, base, kb, w - 1, w - kb, src);
    }
    return acc;
}

template <typename V, typename S>
__device__ __forceinline__ void ilu2_row_fold(const SysDesc &d, const Chunk &ch, int i, int o0, int wd,
                                              const S &src, KCFD_GLOBAL V *__restrict__ zout,
                                              KCFD_GLOBAL V *__restrict__ w, KCFD_GLOBAL V *__restrict__ prod) {
#pragma clang fp contract(off)
    const V zi = src(i);
    if (zout) zout[i] = zi;
    const PlainSrc<V> wsrc{w};
    double pr = 0.0, v;
    const unsigned short *o16 = nullptr;
    if constexpr (sizeof(V) == sizeof(float)) o16 = d.chu_o16;
    if (o16) {
        const FacRef16 f{d.chu_v32, o16, ch.cb};
        v = sub_plain_prod<V>((double)zi, pr, f, o0 + (i - ch.start), wd, wsrc);
    } else {
        v = sub_plain_prod<V>((double)zi, pr, fac_u(d), o0 + (i - ch.start), wd, wsrc);
    }
    const V ud = gl(udiag<V>(d))[i];
    const V wi = (V)(v / (double)ud);
    w[i] = wi;
    const double dg = sizeof(V) == sizeof(float) ? gl(d.c0piv)[i] : (double)ud;
    prod[i] = (V)(pr + dg * (double)wi);
}

template <typename V, int WHICH>
__device__ __forceinline__ bool ilu2_wants(int status) {
    return WHICH == 0 ? (status == S_RUN || status == S_START || status == S_RESUME)
                      : (WHICH == 1 ? status == S_RUN : (status == S_RUN || status == S_START));
}

// One chunk alone (the general path): one row per lane.
template <typename V, bool SECOND, int WHICH>
__device__ __forceinline__ void ilu2_chunk(const SysDesc *__restrict__ sys, const SysState *__restrict__ sts,
                                           const Chunk &ch, int lane) {
#pragma clang fp contract(off)
    const SysDesc &d = sys[ch.sys()];
    const SysState &st = sts[ch.sys()];
    if (!ilu2_wants<V, WHICH>(st.status) || lane >= ch.count()) return;
    const int i = ch.start + lane;       // row of P A P^T
    KCFD_GLOBAL V *__restrict__ w = vec<V>(WHICH == 1 ? d.sh : d.ph);
    const int o0 = ch.o0, wd = ch.wd();
    if constexpr (SECOND && WHICH != 2) {
        if (d.nid > 0) {   // folded: the colour-0 rows of A M z as well
            if constexpr (WHICH == 0) {
                const PlainSrc<V> src{vec<V>(d.p)};
                ilu2_row_fold<V>(d, ch, i, o0, wd, src, (KCFD_GLOBAL V *)nullptr, w, vec<V>(d.v));
            } else {
                const SSrc<V> src{vec<V>(d.r), vec<V>(d.v), (V)st.alpha};
                ilu2_row_fold<V>(d, ch, i, o0, wd, src, vec<V>(d.s), w, vec<V>(d.t));
            }
            return;
        }
    }
    if constexpr (WHICH == 0) {
        const PlainSrc<V> src{vec<V>(d.p)};
        ilu2_row<V, SECOND>(d, ch, i, o0, wd, src, (KCFD_GLOBAL V *)nullptr, w);
    } else if constexpr (WHICH == 1) {
        const SSrc<V> src{vec<V>(d.r), vec<V>(d.v), (V)st.alpha};
        ilu2_row<V, SECOND>(d, ch, i, o0, wd, src, vec<V>(d.s), w);
    } else {
        const PlainSrc<V> src{vec<V>(d.r)};
        ilu2_row<V, SECOND>(d, ch, i, o0, wd, src, (KCFD_GLOBAL V *)nullptr, w);
    }
}

template <typename V, bool SECOND, int WHICH>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_ilu2(const SysDesc *__restrict__ sys,
                                                 const SysState *__restrict__ sts,
                                                 const Chunk *__restrict__ chunks,
                                                 const int *__restrict__ counts) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int nchunks = counts[SECOND ? 4 : 3];
    const int per = (nchunks + kXcds - 1) / kXcds;
    const int c0 = (int)(blockIdx.x % kXcds) * per, c1 = min(c0 + per, nchunks);
    const int wx = (int)(gridDim.x / kXcds) * kWaves;   // waves per XCD
    const int lw = uniform((int)(blockIdx.x / kXcds) * kWaves + (int)(threadIdx.x >> 6));
    // the next chunk's record is loaded while the current one is worked on
    Chunk next{};
    if (c0 + lw < c1) next = chunks[c0 + lw];
    for (int c = c0 + lw; c < c1; c += wx) {
        const Chunk cur = next;
        if (c + wx < c1) next = chunks[c + wx];
        ilu2_chunk<V, SECOND, WHICH>(sys, sts, cur, lane);
    }
}

// ILU(0) apply for a two-colour segment ordering (kcfd_mat::seg2).  A
// row's entries outside its segment are of the other colour (colour 0
// first), so one lane per segment can walk its rows, as the sequential L
// then U solves do (same operations, same order: the entries outside the
// segment come before the row before it in a colour-1 row's L part, after
// the row after it in a colour-0 row's U part):
//   PART 0, colour-0 segments: y_i = z_i - l_{i,i-} y_{i-}
//   PART 1, colour-1 segments: y_i = z_i - sum_j l_ij y_j - l_{i,i-} y_{i-},
//           then w_i = (y_i - u_{i,i+} w_{i+}) / u_ii, last row first
//   PART 2, colour-0 segments: w_i = (y_i - u_{i,i+} w_{i+} - sum_j u_ij w_j)
//           / u_ii, last row first
// (i- / i+: the row before / after i in its segment).  y lives in d.y.  A
// group record (Chunk) holds up to 64 segments numbered step-major, so
// that a step's rows are consecutive over the lanes: start = its first
// segment, count, wd = its longest segment, o0 = its first step {entry
// offset, width, row offset}.  Groups are walked like k_ilu2's chunks.  The
// own operand z and the formed s are as in k_ilu2 (WHICH 0: p, 1: s = r -
// alpha v, stored, 2: r).
template <typename V>
__device__ __forceinline__ const V *seg_l(const SysDesc &d);
template <> __device__ __forceinline__ const double *seg_l<double>(const SysDesc &d) { return d.sgl64; }
template <> __device__ __forceinline__ const float *seg_l<float>(const SysDesc &d) { return d.sgl32; }
template <typename V>
__device__ __forceinline__ const V *seg_u(const SysDesc &d);
template <> __device__ __forceinline__ const double *seg_u<double>(const SysDesc &d) { return d.sgu64; }
template <> __device__ __forceinline__ const float *seg_u<float>(const SysDesc &d) { return d.sgu32; }

// The products a_ij x_j of the entries outside the segment in step k's row,
// in CSR order (padding, column -1, comes last and is left out: nv).  A
// product is rounded as the sequential solve rounds it before subtracting.
template <typename V, typename F, typename S>
__device__ __forceinline__ void seg_products(const F &f, int base, int w, const S &src, double (&pr)[kSegW],
                                             int &nv) {
#pragma clang fp contract(off)
    int c[kSegW];
    V av[kSegW], xv[kSegW];
#pragma unroll
    for (int e = 0; e < kSegW; ++e) {
        c[e] = -1;
        av[e] = (V)0;
        if (e < w) fac_get<V>(f, base + 64 * e, c[e], av[e]);
    }
#pragma unroll
    for (int e = 0; e < kSegW; ++e) xv[e] = e < w ? src(max(c[e], 0)) : (V)0;
    nv = 0;
#pragma unroll
    for (int e = 0; e < kSegW; ++e) {
        pr[e] = (double)av[e] * (double)xv[e];
        nv += c[e] >= 0;
    }
}

// One segment (one lane).  Every load of the segment's rows is issued
// before its arithmetic, which then runs in the sequential solves' order,
// and the stores come last: the rows' loads do not wait on each other.
template <typename V, int PART, int WHICH, typename S>
__device__ __forceinline__ void seg_walk(const SysDesc &d, const Chunk &g, int lane, const S &src,
                                         KCFD_GLOBAL V *__restrict__ zout) {
#pragma clang fp contract(off)
    const kint2 sv = gl(reinterpret_cast<const kint2 *>(d.sg))[g.start + lane];
    const int r0 = sv.x, len = sv.y;
    const kint4 *__restrict__ steps = reinterpret_cast<const kint4 *>(d.sg_steps) + g.o0;
    KCFD_GLOBAL V *__restrict__ y = vec<V>(d.y);
    KCFD_GLOBAL V *__restrict__ w = vec<V>(WHICH == 1 ? d.sh : d.ph);
    const V *__restrict__ lv = seg_l<V>(d);
    const V *__restrict__ uv = seg_u<V>(d);
    const V *__restrict__ dg = udiag<V>(d);
    // the segment's rows (step-major numbering: row k at r0 + the step's row
    // offset) and its steps' entry layouts
    int ri[kSegLen];
    kint4 st[kSegLen];
#pragma unroll
    for (int k = 0; k < kSegLen; ++k) {
        st[k] = kint4{0, 0, 0, 0};
        if (k < len) st[k] = gl(steps)[k];
        ri[k] = r0 + st[k].z;
    }
    V out[kSegLen];
    if constexpr (PART == 0) {
        V z[kSegLen], l[kSegLen];
#pragma unroll
        for (int k = 0; k < kSegLen; ++k) {
            z[k] = l[k] = (V)0;
            if (k < len) {
                z[k] = src(ri[k]);
                l[k] = gl(lv)[ri[k]];
            }
        }
        double yp = 0.0;
#pragma unroll
        for (int k = 0; k < kSegLen; ++k) {
            double acc = (double)z[k];
            if (k > 0) acc -= (double)l[k] * yp;
            out[k] = (V)acc;
            yp = acc;
        }
#pragma unroll
        for (int k = 0; k < kSegLen; ++k)
            if (k < len) {
                if (zout) zout[ri[k]] = z[k];
                y[ri[k]] = out[k];
            }
    } else {
        double pr[kSegLen][kSegW];
        int nv[kSegLen];
        V z[kSegLen], l[kSegLen], u[kSegLen], dd[kSegLen];
        const PlainSrc<V> osrc{PART == 1 ? y : w};   // colour-0 y, or colour-1 w
#pragma unroll
        for (int k = 0; k < kSegLen; ++k) {
            z[k] = l[k] = u[k] = (V)0;
            dd[k] = (V)1;
            nv[k] = 0;
#pragma unroll
            for (int e = 0; e < kSegW; ++e) pr[k][e] = 0.0;
            if (k < len) {
                const int i = ri[k];
                z[k] = src(i);
                l[k] = gl(lv)[i];
                u[k] = gl(uv)[i];
                dd[k] = gl(dg)[i];
                if (PART == 1) seg_products<V>(fac_l(d), st[k].x + lane, st[k].y, osrc, pr[k], nv[k]);
                else seg_products<V>(fac_u(d), st[k].x + lane, st[k].y, osrc, pr[k], nv[k]);
            }
        }
        if constexpr (PART == 1) {
            double yk[kSegLen];
            double yp = 0.0;
#pragma unroll
            for (int k = 0; k < kSegLen; ++k) {
                double acc = (double)z[k];
#pragma unroll
                for (int e = 0; e < kSegW; ++e)
                    if (e < nv[k]) acc -= pr[k][e];
                if (k > 0) acc -= (double)l[k] * yp;
                yk[k] = acc;
                yp = acc;
            }
            double wn = 0.0;
#pragma unroll
            for (int k = kSegLen - 1; k >= 0; --k) {
                double acc = yk[k];
                if (k + 1 < len) acc -= (double)u[k] * wn;
                out[k] = (V)(acc / (double)dd[k]);
                wn = (double)out[k];
            }
        } else {
            // the colour-0 forward solve again, in registers: y is not read
            // back (in MIXED that keeps it unrounded; in FP64 it is the same
            // arithmetic as part 0's)
            double yk[kSegLen];
            double yp = 0.0;
#pragma unroll
            for (int k = 0; k < kSegLen; ++k) {
                double acc = (double)z[k];
                if (k > 0) acc -= (double)l[k] * yp;
                yk[k] = acc;
                yp = acc;
            }
            double wn = 0.0;
#pragma unroll
            for (int k = kSegLen - 1; k >= 0; --k) {
                double acc = yk[k];
                if (k + 1 < len) acc -= (double)u[k] * wn;
#pragma unroll
                for (int e = 0; e < kSegW; ++e)
                    if (e < nv[k]) acc -= pr[k][e];
                out[k] = (V)(acc / (double)dd[k]);
                wn = (double)out[k];
            }
        }
#pragma unroll
        for (int k = 0; k < kSegLen; ++k)
            if (k < len) {
                if (PART == 1 && zout) zout[ri[k]] = z[k];
                w[ri[k]] = out[k];
            }
    }
}

template <typename V, int PART, int WHICH>
__global__ __launch_bounds__(kBlock) void k_seg(const SysDesc *__restrict__ sys, const SysState *__restrict__ sts,
                                                const Chunk *__restrict__ groups,
                                                const int *__restrict__ counts) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int ng = counts[PART == 1 ? 9 : 8];
    const int per = (ng + kXcds - 1) / kXcds;
    const int c0 = (int)(blockIdx.x % kXcds) * per, c1 = min(c0 + per, ng);
    const int wx = (int)(gridDim.x / kXcds) * kWaves;   // waves per XCD
    const int lw = uniform((int)(blockIdx.x / kXcds) * kWaves + (int)(threadIdx.x >> 6));
    for (int c = c0 + lw; c < c1; c += wx) {
        const Chunk g = groups[c];
        const SysDesc &d = sys[g.sys()];
        const SysState &st = sts[g.sys()];
        if (!ilu2_wants<V, WHICH>(st.status) || lane >= g.count()) continue;
        if constexpr (WHICH == 0) {
            const PlainSrc<V> src{vec<V>(d.p)};
            seg_walk<V, PART, WHICH>(d, g, lane, src, (KCFD_GLOBAL V *)nullptr);
        } else if constexpr (PART == 2 && WHICH == 1) {   // s as part 0 stored it
            const PlainSrc<V> src{vec<V>(d.s)};
            seg_walk<V, PART, WHICH>(d, g, lane, src, (KCFD_GLOBAL V *)nullptr);
        } else if constexpr (WHICH == 1) {
            const SSrc<V> src{vec<V>(d.r), vec<V>(d.v), (V)st.alpha};
            seg_walk<V, PART, WHICH>(d, g, lane, src, vec<V>(d.s));
        } else {
            const PlainSrc<V> src{vec<V>(d.r)};
            seg_walk<V, PART, WHICH>(d, g, lane, src, (KCFD_GLOBAL V *)nullptr);
        }
    }
}

template <typename V>
__global__ void k_fill_sentinel(V *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = sentinel_value<V>();
}

}  // namespace kcfd

using namespace kcfd;

// The workspace of kcfd_solve_noalloc, one copy per device (see
// solve_noalloc below).  64 MiB: a FP64 Jacobi system of about 1.2M rows,
// a mixed colour-ILU(0) one of about 1.1M.
static constexpr size_t kNoAllocBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace kcfd {
// Device address of `scratch` on the current device; kcfd_init calls it once
// per context, with the context's device current.
int noalloc_resolve(kcfd_ctx *ctx) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        return KCFD_EHIP;
    }
    ctx->noalloc = p;
    return KCFD_OK;
}
}  // namespace kcfd

// ======================================================================= host

namespace kcfd {
struct SysTables {
    std::vector<int2> tiles;
    std::vector<std::vector<Chunk>> lvl_l, lvl_u;
    std::vector<Chunk> ca, cb;
    std::vector<Chunk> sg0, sg1;   // segment groups of colour 0 and colour 1
};
}  // namespace kcfd

struct kcfd_batch {
    kcfd_ctx *ctx = nullptr;
    int nsys = 0;                   // slots (the capacity of a dynamic batch)
    int ntiles = 0, grid = 0;
    int rows = kDefaultRows;        // rows per thread of the tile kernels (tile = rows * kBlock)
    bool dynamic = false;           // systems attach / detach between rounds
    kcfd_solve_opts opts{};
    BatchParams bp{};
    std::vector<kcfd_mat *> mats;   // per slot; null = free slot
    std::vector<SysDesc> h_desc;
    std::vector<int> sys_n;
    std::vector<void *> slot_mem;   // dynamic: vectors of each slot
    // dynamic: vector blocks of detached systems, {bytes, pointer}, reused by
    // attaches of the same size (a device free synchronises the device);
    // kcfd_batch_trim frees them
    std::vector<std::pair<size_t, void *>> spare;
    std::vector<char> fresh, reported;
    SysDesc *d_desc = nullptr;
    SysState *d_state = nullptr;
    std::vector<SysState> h_state;
    int2 *d_tiles = nullptr;
    int cap_tiles = 0;
    double *d_partials = nullptr;
    unsigned int *d_counters = nullptr;
    int *d_counts = nullptr;        // {ntiles, nchunks_l, nchunks_u, nchunks_a, nchunks_b, walk,
                                    // folded, 0, nsg0, nsg1}, read by the kernels
    void *slab = nullptr;           // static: all per-system vectors
    size_t slab_bytes = 0;
    // ILU chunk tables (level-major over systems) for the L and U solves
    Chunk *d_chunks_l = nullptr, *d_chunks_u = nullptr;
    int nchunks_l = 0, nchunks_u = 0, cap_chunks_l = 0, cap_chunks_u = 0, trsv_grid = 0;
    // two-colour members: colour-1 (a) and colour-0 (b) chunk runs
    Chunk *d_chunks_a = nullptr, *d_chunks_b = nullptr;
    // segment-ordered members: groups of colour 0 and colour 1
    Chunk *d_sg0 = nullptr, *d_sg1 = nullptr;
    int nsg0 = 0, nsg1 = 0, cap_sg0 = 0, cap_sg1 = 0;
    bool run_seg = false;
    int nchunks_a = 0, nchunks_b = 0, cap_chunks_a = 0, cap_chunks_b = 0;
    bool run_two = false, run_polled = false;   // which ILU kernels the graph launches
    const int **d_lvl_l = nullptr, **d_lvl_u = nullptr;
    BatchParams *d_params = nullptr;   // read by the kernels: the captured graph stays valid
    hipGraphExec_t graph = nullptr;
    int graph_iters = 0;
    bool dirty = false;             // membership changed: tables to rebuild
    bool no_graph = false;          // direct launches only (kcfd_solve_noalloc)
    bool fold = false;              // kcfd_solve_opts.fold_colour0 resolved for this batch
    // per-system table pieces (rebuild_tables); tables_dirty: the membership
    // or a member's structure changed since they were built; compacted: the
    // device tables hold only the systems that were still running
    std::vector<SysTables> sys_tables;
    bool tables_dirty = true, compacted = false;
    // kcfd_batch_profile: launch_iteration records ctx->ev0 / ev1 around the
    // kernel of this phase (0: none, as during graph capture)
    int timed_phase = 0;
};

namespace {

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// One kernel of an iteration.  The kernels of the phase being timed
// (kcfd_batch_profile) carry the profiling events in their own dispatch
// packets -- the start event on the phase's first kernel, the stop event on
// its last -- so the elapsed time is the kernels' own, without the launch
// gaps a stream-recorded event would add.
template <typename F, typename... Args>
void launch_k(const kcfd_batch *b, int phase, bool first, bool last, F kernel, int grid, int block,
              hipStream_t s, Args... args) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (phase != 0 && b->timed_phase == phase) {
        if (first) e0 = b->ctx->ev0;
        if (last) e1 = b->ctx->ev1;
    }
    if (e0 || e1)
        hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, s, e0, e1, 0, args...);
    else
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, s, args...);
}

template <typename V, int PC, int R>
void launch_iteration(kcfd_batch *b, hipStream_t s) {
    const int g = b->grid, ns = b->nsys, tg = b->trsv_grid;
    launch_k(b, 4, true, true, k_p<V, PC, R>, g, kBlock, s, b->d_desc, b->d_state, b->d_tiles, b->d_counts,
             b->d_counters);
    if (PC == KCFD_PC_ILU0) {
        // phase 3: the p apply's first launches (two-colour: the colour-1
        // rows; sync-free: the L solve); phase 5: the two-colour colour-0
        // rows, launched last (the two kinds of members are independent)
        // (phase 6: the whole apply, one event pair from its first launch to
        // its last)
        const bool whole = b->timed_phase == 6;
        if (b->run_two)
            launch_k(b, whole ? 6 : 3, true, !whole && !b->run_polled, k_ilu2<V, false, 0>, g, kBlock, s,
                     b->d_desc, b->d_state, b->d_chunks_a, b->d_counts);
        if (b->run_polled) {
            launch_k(b, whole ? 6 : 3, !b->run_two, !whole, k_trsv<V, false, 0>, tg, kBlock, s, b->d_desc,
                     b->d_state, b->d_chunks_l, b->d_counts, b->d_lvl_l, b->d_params, b->d_counters);
            launch_k(b, whole ? 6 : 0, false, whole && !b->run_two && !b->run_seg, k_trsv<V, true, 0>, tg, kBlock,
                     s, b->d_desc, b->d_state, b->d_chunks_u, b->d_counts, b->d_lvl_u, b->d_params, b->d_counters);
        }
        if (b->run_seg) {   // segment-ordered members (timed in phase 6 only)
            launch_k(b, whole ? 6 : 0, whole && !b->run_two && !b->run_polled, false, k_seg<V, 0, 0>, g, kBlock,
                     s, b->d_desc, b->d_state, b->d_sg0, b->d_counts);
            launch_k(b, 0, false, false, k_seg<V, 1, 0>, g, kBlock, s, b->d_desc, b->d_state, b->d_sg1,
                     b->d_counts);
            launch_k(b, whole ? 6 : 0, false, whole && !b->run_two, k_seg<V, 2, 0>, g, kBlock, s, b->d_desc,
                     b->d_state, b->d_sg0, b->d_counts);
        }
        if (b->run_two)
            launch_k(b, whole ? 6 : 5, !whole, true, k_ilu2<V, true, 0>, g, kBlock, s, b->d_desc, b->d_state,
                     b->d_chunks_b, b->d_counts);
    }
    launch_k(b, 1, true, true, k_spmv_p<V, PC, R>, g, kBlock, s, b->d_desc, b->d_state, b->d_tiles, b->d_counts,
             b->d_partials);
    launch_k(b, 0, false, false, f_spmv_p, ns, kFinBlock, s, b->d_desc, b->d_state, b->d_partials,
             b->d_params);
    if (PC != KCFD_PC_ILU0)
        launch_k(b, 0, false, false, k_s<V, PC, R>, g, kBlock, s, b->d_desc, b->d_state, b->d_tiles, b->d_counts);
    if (PC == KCFD_PC_ILU0) {
        if (b->run_two)
            launch_k(b, 0, false, false, k_ilu2<V, false, 1>, g, kBlock, s, b->d_desc, b->d_state,
                     b->d_chunks_a, b->d_counts);
        if (b->run_polled) {
            launch_k(b, 0, false, false, k_trsv<V, false, 1>, tg, kBlock, s, b->d_desc, b->d_state,
                     b->d_chunks_l, b->d_counts, b->d_lvl_l, b->d_params, b->d_counters);
            launch_k(b, 0, false, false, k_trsv<V, true, 1>, tg, kBlock, s, b->d_desc, b->d_state,
                     b->d_chunks_u, b->d_counts, b->d_lvl_u, b->d_params, b->d_counters);
        }
        if (b->run_seg) {
            launch_k(b, 0, false, false, k_seg<V, 0, 1>, g, kBlock, s, b->d_desc, b->d_state, b->d_sg0,
                     b->d_counts);
            launch_k(b, 0, false, false, k_seg<V, 1, 1>, g, kBlock, s, b->d_desc, b->d_state, b->d_sg1,
                     b->d_counts);
            launch_k(b, 0, false, false, k_seg<V, 2, 1>, g, kBlock, s, b->d_desc, b->d_state, b->d_sg0,
                     b->d_counts);
        }
        if (b->run_two)
            launch_k(b, 0, false, false, k_ilu2<V, true, 1>, g, kBlock, s, b->d_desc, b->d_state,
                     b->d_chunks_b, b->d_counts);
    }
    launch_k(b, 0, false, false, k_spmv_s<V, PC, R>, g, kBlock, s, b->d_desc, b->d_state, b->d_tiles,
             b->d_counts, b->d_partials, b->d_params);
    launch_k(b, 0, false, false, f_spmv_s, ns, kFinBlock, s, b->d_desc, b->d_state, b->d_partials,
             b->d_params);
    launch_k(b, 2, true, true, k_xr<V, PC, R>, g, kBlock, s, b->d_desc, b->d_state, b->d_tiles, b->d_counts,
             b->d_partials, b->d_params);
    launch_k(b, 0, false, false, f_xr, ns, kFinBlock, s, b->d_desc, b->d_state, b->d_partials, b->d_params);
}

// One CG iteration (FP64).  Profiled phases: 1 = q = A p (C3), 2 = x/r
// update (C4), 3 = the ILU(0) apply z = M r, 4 = p update (C2).
template <typename V, int PC, int R>
void launch_iteration_cg(kcfd_batch *b, hipStream_t s) {
    const int g = b->grid, ns = b->nsys, tg = b->trsv_grid;
    if (PC == KCFD_PC_ILU0) {
        if (b->run_two) {
            launch_k(b, 3, true, false, k_ilu2<V, false, 2>, g, kBlock, s, b->d_desc, b->d_state,
                     b->d_chunks_a, b->d_counts);
            launch_k(b, 3, false, !b->run_polled && !b->run_seg, k_ilu2<V, true, 2>, g, kBlock, s, b->d_desc,
                     b->d_state, b->d_chunks_b, b->d_counts);
        }
        if (b->run_polled) {
            launch_k(b, 3, !b->run_two, false, k_trsv<V, false, 2>, tg, kBlock, s, b->d_desc, b->d_state,
                     b->d_chunks_l, b->d_counts, b->d_lvl_l, b->d_params, b->d_counters);
            launch_k(b, 3, false, !b->run_seg, k_trsv<V, true, 2>, tg, kBlock, s, b->d_desc, b->d_state,
                     b->d_chunks_u, b->d_counts, b->d_lvl_u, b->d_params, b->d_counters);
        }
        if (b->run_seg) {
            launch_k(b, 3, !b->run_two && !b->run_polled, false, k_seg<V, 0, 2>, g, kBlock, s, b->d_desc,
                     b->d_state, b->d_sg0, b->d_counts);
            launch_k(b, 3, false, false, k_seg<V, 1, 2>, g, kBlock, s, b->d_desc, b->d_state, b->d_sg1,
                     b->d_counts);
            launch_k(b, 3, false, true, k_seg<V, 2, 2>, g, kBlock, s, b->d_desc, b->d_state, b->d_sg0,
                     b->d_counts);
        }
    }
    launch_k(b, 0, false, false, k_cg_rz<PC, R>, g, kBlock, s, b->d_desc, b->d_state, b->d_tiles, b->d_counts,
             b->d_partials);
    launch_k(b, 0, false, false, f_cg_rho, ns, kFinBlock, s, b->d_desc, b->d_state, b->d_partials);
    launch_k(b, 4, true, true, k_cg_p<PC, R>, g, kBlock, s, b->d_desc, b->d_state, b->d_tiles, b->d_counts);
    launch_k(b, 1, true, true, k_cg_q<PC, R>, g, kBlock, s, b->d_desc, b->d_state, b->d_tiles, b->d_counts,
             b->d_partials);
    launch_k(b, 0, false, false, f_cg_alpha, ns, kFinBlock, s, b->d_desc, b->d_state, b->d_partials);
    launch_k(b, 2, true, true, k_cg_xr<R>, g, kBlock, s, b->d_desc, b->d_state, b->d_tiles, b->d_counts,
             b->d_partials);
    launch_k(b, 0, false, false, f_cg_check, ns, kFinBlock, s, b->d_desc, b->d_state, b->d_partials);
}

template <typename V, int PC, int R>
void launch_init(kcfd_batch *b, hipStream_t s) {
    if (PC == KCFD_PC_ILU0)
        hipLaunchKernelGGL(k_perm_in<R>, dim3(b->grid), dim3(kBlock), 0, s, b->d_desc, b->d_state, b->d_tiles,
                           b->d_counts);
    hipLaunchKernelGGL((k_init<V, PC, R>), dim3(b->grid), dim3(kBlock), 0, s, b->d_desc, b->d_state,
                       b->d_tiles, b->d_counts, b->d_partials, b->d_params);
    hipLaunchKernelGGL(f_init, dim3(b->nsys), dim3(kFinBlock), 0, s, b->d_desc, b->d_state,
                       b->d_partials, b->d_params);
}

// after the iterations: best-iterate restore (mixed), FP64 true residual,
// scatter out of the ILU(0) row order
template <int R>
void launch_finish(kcfd_batch *b, hipStream_t s) {
    if (b->bp.precision == KCFD_MIXED)
        hipLaunchKernelGGL(k_restore<R>, dim3(b->grid), dim3(kBlock), 0, s, b->d_desc, b->d_state,
                           b->d_tiles, b->d_counts);
    hipLaunchKernelGGL(k_true_resid<R>, dim3(b->grid), dim3(kBlock), 0, s, b->d_desc, b->d_state,
                       b->d_tiles, b->d_counts, b->d_partials);
    hipLaunchKernelGGL(f_true_resid, dim3(b->nsys), dim3(kFinBlock), 0, s, b->d_desc, b->d_state,
                       b->d_partials);
    if (b->bp.precond == KCFD_PC_ILU0)
        hipLaunchKernelGGL(k_perm_out<R>, dim3(b->grid), dim3(kBlock), 0, s, b->d_desc, b->d_state,
                           b->d_tiles, b->d_counts);
}

using LaunchFn = void (*)(kcfd_batch *, hipStream_t);

template <int R>
LaunchFn pick_iter_r(int precision, int pc, int method) {
    if (method == KCFD_METHOD_CG) {   // FP64 only (check_opts)
        if (pc == KCFD_PC_JACOBI) return launch_iteration_cg<double, KCFD_PC_JACOBI, R>;
        if (pc == KCFD_PC_ILU0) return launch_iteration_cg<double, KCFD_PC_ILU0, R>;
        return launch_iteration_cg<double, KCFD_PC_NONE, R>;
    }
    if (precision == KCFD_FP64) {
        if (pc == KCFD_PC_JACOBI) return launch_iteration<double, KCFD_PC_JACOBI, R>;
        if (pc == KCFD_PC_ILU0) return launch_iteration<double, KCFD_PC_ILU0, R>;
        return launch_iteration<double, KCFD_PC_NONE, R>;
    }
    if (pc == KCFD_PC_JACOBI) return launch_iteration<float, KCFD_PC_JACOBI, R>;
    if (pc == KCFD_PC_ILU0) return launch_iteration<float, KCFD_PC_ILU0, R>;
    return launch_iteration<float, KCFD_PC_NONE, R>;
}

LaunchFn pick_iter(int precision, int pc, int method, int rows) {
    if (rows == 1) return pick_iter_r<1>(precision, pc, method);
    if (rows == 4) return pick_iter_r<4>(precision, pc, method);
    return pick_iter_r<2>(precision, pc, method);
}

template <int R>
LaunchFn pick_init_r(int precision, int pc) {
    if (precision == KCFD_FP64) {
        if (pc == KCFD_PC_JACOBI) return launch_init<double, KCFD_PC_JACOBI, R>;
        if (pc == KCFD_PC_ILU0) return launch_init<double, KCFD_PC_ILU0, R>;
        return launch_init<double, KCFD_PC_NONE, R>;
    }
    if (pc == KCFD_PC_JACOBI) return launch_init<float, KCFD_PC_JACOBI, R>;
    if (pc == KCFD_PC_ILU0) return launch_init<float, KCFD_PC_ILU0, R>;
    return launch_init<float, KCFD_PC_NONE, R>;
}

LaunchFn pick_init(int precision, int pc, int rows) {
    if (rows == 1) return pick_init_r<1>(precision, pc);
    if (rows == 4) return pick_init_r<4>(precision, pc);
    return pick_init_r<2>(precision, pc);
}

LaunchFn pick_finish(int rows) {
    if (rows == 1) return launch_finish<1>;
    if (rows == 4) return launch_finish<4>;
    return launch_finish<2>;
}

// algorithmic bytes (roofline model): CSR stream M_A, ILU factor stream M_F,
// vector streams of the fused schedule; MIXED adds one FP64 residual
// replacement (M_A at 8 bytes + 6 n*8) per reliable update
// folded product (nid > 0): the two products per iteration skip the
// colour-0 rows' entries and slice pointers (those rows only read v or t,
// and rt or s); the colour-0 half of each apply writes v or t (n0 s_v) and,
// in MIXED, reads the FP64 pivots (8 n0)
// Algorithmic bytes of one segment-ordered ILU(0) apply: the entries
// outside the segments (value + 4-byte column), the in-segment L and U
// values and u_ii by row, z read, w written twice over (colour 1 written,
// gathered once; colour 0 written), the segment and step records, and the
// colour-0 rows' y (written, read back).
double seg2_apply_bytes(const kcfd_mat *m, double sv) {
    const double n = m->n, n0 = m->ncolour0;
    return (double)(m->nent_l + m->nent_u) * (sv + 4) + 3 * n * sv + 3 * n * sv + 8.0 * m->h_sg.size() +
           16.0 * m->h_sg_steps.size() + 2 * n0 * sv;
}

double model_bytes(const kcfd_mat *m, int pc, int precision, const SysState &st,
                   int method = KCFD_METHOD_BICGSTAB, int nid = 0) {
    const double n = m->n, nnz = m->nnz;
    auto MA = [&](double sv) { return nnz * (sv + 4) + (n + 1) * 4; };
    auto MF = [&](double sv) { return nnz * (sv + 4) + 2 * (n + 1) * 4; };
    if (method == KCFD_METHOD_CG) {
        // C1..C4 per iteration: Jacobi M_A + 14 n s_v, ILU(0) M_A + M_F + 17 n s_v
        // (the apply's 4 streams included), none M_A + 12 n s_v
        const double per = (pc == KCFD_PC_JACOBI ? MA(8) + 14 * n * 8
                            : pc == KCFD_PC_ILU0 ? MA(8) + MF(8) + 17 * n * 8 : MA(8) + 12 * n * 8) +
                           (pc == KCFD_PC_ILU0 && m->seg2 ? 2 * m->ncolour0 * 8.0 : 0.0);
        return st.total_iters * per + MA(8) + 3 * n * 8;
    }
    auto per_iter = [&](double sv) {
        if (pc == KCFD_PC_JACOBI) return 2 * MA(sv) + 22 * n * sv;
        if (pc == KCFD_PC_ILU0)   // (segment ordering: the colour-0 rows' y, written and read, per apply)
            return 2 * MA(sv) + 2 * MF(sv) + 29 * n * sv + (m->seg2 ? 4 * m->ncolour0 * sv : 0.0);
        return 2 * MA(sv) + 19 * n * sv;
    };
    const double n0 = nid, nnz0 = nid > 0 ? (double)m->nnz_colour0 : 0.0;
    auto fold = [&](double sv) { return nid > 0 ? -2 * (nnz0 * (sv + 4) + n0 * 4) + 2 * n0 * sv : 0.0; };
    if (precision == KCFD_FP64)
        return (st.total_iters + 0.5 * st.half) * (per_iter(8) + fold(8)) + MA(8) + 3 * n * 8;
    // MIXED: the iterate x is FP64 (read + write, +8 n over an FP32 one);
    // a residual replacement reads A, x, b in FP64 and writes r (+ xb copy)
    const double fold64 = nid > 0 ? 2 * 8 * n0 : 0.0;   // the pivots, twice
    return st.total_iters * (per_iter(4) + 8 * n + fold(4) + fold64) +
           st.refinements * (MA(8) + 2 * n * 8 + n * 4);
}

void fill_desc(SysDesc &d, kcfd_mat *m) {
    d.ip = m->f_indptr;
    d.ix = m->f_indices;
    d.dpos = m->f_diagpos;
    d.uix = m->f_uix;
    d.perm = m->pc_kind == KCFD_PC_ILU0 && m->sell_permuted ? m->perm : nullptr;
    d.sptr = m->slice_ptr;
    d.six = m->six;
    d.sa64 = m->sa64;
    d.spk32 = m->spk32;
    d.sv32 = m->sv32;
    d.so16 = m->so16;
    d.sbase = m->sbase;
    d.dinv64 = m->dinv;
    d.dinv32 = m->dinv32;
    d.lu64 = m->lu;
    d.lu32 = m->lu32;
    d.polled = m->pc_kind == KCFD_PC_ILU0 && !m->two_colour && !m->seg2;
    d.sg = m->seg2 ? m->sg : nullptr;
    d.sg_steps = m->sg_steps;
    d.sgl64 = m->sgl64;
    d.sgu64 = m->sgu64;
    d.sgl32 = m->sgl32;
    d.sgu32 = m->sgu32;
    d.chl_off = m->chl_off;
    d.chl_col = m->chl_col;
    d.chu_off = m->chu_off;
    d.chu_col = m->chu_col;
    d.chl_v64 = m->chl_v64;
    d.chu_v64 = m->chu_v64;
    d.ud64 = m->ud64;
    d.chl_pk32 = m->chl_pk32;
    d.chu_pk32 = m->chu_pk32;
    const bool ch16 = m->ch16 && m->chl_v32 != nullptr;
    d.chl_v32 = ch16 ? m->chl_v32 : nullptr;
    d.chu_v32 = ch16 ? m->chu_v32 : nullptr;
    d.chl_o16 = ch16 ? m->chl_o16 : nullptr;
    d.chu_o16 = ch16 ? m->chu_o16 : nullptr;
    d.chl_base = ch16 ? m->chl_base : nullptr;
    d.chu_base = ch16 ? m->chu_base : nullptr;
    d.ud32 = m->ud32;
}

int check_opts(const kcfd_solve_opts *o) {
    const int pc = o->precond, prec = o->precision;
    KCFD_ARG(pc == KCFD_PC_NONE || pc == KCFD_PC_JACOBI || pc == KCFD_PC_ILU0,
             "unknown preconditioner");
    KCFD_ARG(prec == KCFD_FP64 || prec == KCFD_MIXED, "unknown precision");
    KCFD_ARG(o->grid_per_cu >= 0 && o->grid_per_cu <= 32, "grid_per_cu must be 0..32");
    KCFD_ARG(o->trsv_per_cu >= 0 && o->trsv_per_cu <= 4, "trsv_per_cu must be 0..4");
    KCFD_ARG(o->tile_walk >= 0 && o->tile_walk <= 2, "tile_walk must be 0..2");
    KCFD_ARG(o->rows_per_thread == 0 || o->rows_per_thread == 1 || o->rows_per_thread == 2 ||
                 o->rows_per_thread == 4,
             "rows_per_thread must be 0, 1, 2 or 4");
    KCFD_ARG(o->method == KCFD_METHOD_BICGSTAB || o->method == KCFD_METHOD_CG,
             o->method == KCFD_METHOD_GMRES ? "GMRES runs through kcfd_solve only (no batches)"
                                            : "unknown Krylov method");
    KCFD_ARG(o->method != KCFD_METHOD_CG || prec == KCFD_FP64, "CG runs in FP64 only");
    KCFD_ARG(o->fold_colour0 >= -1 && o->fold_colour0 <= 1, "fold_colour0 must be -1, 0 or 1");
    return KCFD_OK;
}

int check_mat(const kcfd_batch *b, const kcfd_mat *m, int k) {
    if (m->n <= 0) {   // a tile of no rows would leave its system's partials unwritten
        set_error("matrix %d is empty (n = 0); kcfd_solve handles it alone", k);
        return KCFD_EARG;
    }
    if (m->ctx != b->ctx) {
        set_error("matrix %d belongs to another context", k);
        return KCFD_EARG;
    }
    {
        char what[48];
        snprintf(what, sizeof(what), "matrix %d", k);
        KCFD_CHECK(check_current(m, what));
    }
    if (m->pc_kind != b->bp.precond) {
        set_error("matrix %d: preconditioner not set up as requested (have %d, want %d)", k,
                  m->pc_kind, b->bp.precond);
        return KCFD_ESTATE;
    }
    if (b->bp.precision == KCFD_MIXED && !m->have32) {
        set_error("matrix %d: no FP32 copy (set up the preconditioner with KCFD_MIXED)", k);
        return KCFD_ESTATE;
    }
    return KCFD_OK;
}

// common part of a batch: parameters and the small per-slot device arrays
// the host side of a batch: options, parameters and per-slot host arrays
void init_batch_host(kcfd_batch *b, kcfd_ctx *ctx, int slots, const kcfd_solve_opts *o) {
    b->ctx = ctx;
    b->nsys = slots;
    b->opts = *o;
    if (b->opts.check_every <= 0) b->opts.check_every = 8;
    b->rows = o->rows_per_thread > 0 ? o->rows_per_thread : kDefaultRows;
    b->bp.rtol = o->rtol;
    b->bp.atol = o->atol;
    b->bp.delta = o->inner_rtol > 0 ? o->inner_rtol : 0.1;
    const double eps = 2.220446049250313e-16;
    b->bp.rhotol = eps * eps;
    b->bp.precision = o->precision;
    b->bp.precond = o->precond;
    b->fold = o->fold_colour0 > 0 || (o->fold_colour0 == 0 && o->precision == KCFD_MIXED);
    b->bp.max_refine = o->max_refine > 0 ? o->max_refine : 1000;
    b->bp.nsys = slots;
    b->bp.inner_maxiter = 5000;
    b->bp.method = o->method;
    // polls of one operand in the sync-free triangular solves before they
    // give up (KCFD_TRSV_POLL_LIMIT overrides; at ~1 us a poll the default is
    // about a second, against ~1 ms for a whole 382-level solve)
    b->bp.poll_limit = kDefaultPollLimit;
    if (const char *e = std::getenv("KCFD_TRSV_POLL_LIMIT")) b->bp.poll_limit = std::max(0, std::atoi(e));
    b->mats.assign(slots, nullptr);
    b->h_desc.assign(slots, SysDesc{});
    b->sys_n.assign(slots, 0);
    b->slot_mem.assign(slots, nullptr);
    b->fresh.assign(slots, 0);
    b->reported.assign(slots, 0);
    SysState free_st{};
    free_st.status = S_FREE;
    b->h_state.assign(slots, free_st);
}

int new_batch(kcfd_ctx *ctx, int slots, const kcfd_solve_opts *o, kcfd_batch **out) {
    KCFD_CHECK(check_opts(o));
    KCFD_ARG(o->precond != KCFD_PC_ILU0 || slots <= kChunkMaxSys,
             "an ILU(0) batch holds at most 4096 systems (slots)");
    kcfd_batch *b = new kcfd_batch();
    init_batch_host(b, ctx, slots, o);
    *out = b;
    // Every host-issued copy and fill goes on the library stream (created
    // non-blocking, so the null stream is not ordered with it).
    hipStream_t s = ctx->stream;
    KCFD_HIP(hipMalloc(&b->d_desc, sizeof(SysDesc) * slots));
    KCFD_HIP(hipMalloc(&b->d_state, sizeof(SysState) * slots));
    KCFD_HIP(hipMemcpyAsync(b->d_state, b->h_state.data(), sizeof(SysState) * slots, hipMemcpyHostToDevice, s));
    KCFD_HIP(hipMalloc(&b->d_counters, sizeof(unsigned int) * 8));
    KCFD_HIP(hipMemsetAsync(b->d_counters, 0, sizeof(unsigned int) * 8, s));
    KCFD_HIP(hipMalloc(&b->d_counts, sizeof(int) * kCounts));
    KCFD_HIP(hipMemsetAsync(b->d_counts, 0, sizeof(int) * 8, s));
    KCFD_HIP(hipMalloc(&b->d_params, sizeof(BatchParams)));
    KCFD_HIP(hipMemcpyAsync(b->d_params, &b->bp, sizeof(BatchParams), hipMemcpyHostToDevice, s));
    if (o->precond == KCFD_PC_ILU0) {
        KCFD_HIP(hipMalloc((void **)&b->d_lvl_l, sizeof(int *) * slots));
        KCFD_HIP(hipMalloc((void **)&b->d_lvl_u, sizeof(int *) * slots));
    }
    KCFD_HIP(hipStreamSynchronize(s));
    return KCFD_OK;
}

// Per-system workspace: 6 Krylov vectors (9 with ILU) of the precision's
// value type; MIXED adds the FP64 best iterate xb; a permuted row order adds
// FP64 b and x in that order.
int vectors_per_system(const kcfd_batch *b) { return b->bp.precond == KCFD_PC_ILU0 ? 9 : 6; }
size_t value_size(const kcfd_batch *b) { return b->bp.precision == KCFD_FP64 ? sizeof(double) : sizeof(float); }
size_t vec_stride(const kcfd_batch *b, int n) { return align_up((size_t)std::max(n, 1) * value_size(b), 256); }
size_t f64_stride(int n) { return align_up((size_t)std::max(n, 1) * sizeof(double), 256); }
size_t slot_bytes(const kcfd_batch *b, const SysDesc &d) {
    size_t v = (size_t)vectors_per_system(b) * vec_stride(b, d.n);
    if (b->bp.precision == KCFD_MIXED) v += f64_stride(d.n);
    if (d.perm) v += 2 * f64_stride(d.n);
    return v;
}

// slot k's vectors at base.  In the natural order the iterate is the
// caller's FP64 x and b is read in place.
void place_vectors(kcfd_batch *b, int k, char *base) {
    SysDesc &d = b->h_desc[k];
    const size_t stride = vec_stride(b, d.n), s64 = f64_stride(d.n);
    void **vecs[9] = {&d.r, &d.rt, &d.p, &d.v, &d.s, &d.t, &d.ph, &d.sh, &d.y};
    const int nv = vectors_per_system(b);
    for (int q = 0; q < nv; ++q) *vecs[q] = base + q * stride;
    char *tail = base + nv * stride;
    d.xb = nullptr;
    if (b->bp.precision == KCFD_MIXED) {
        d.xb = (double *)tail;
        tail += s64;
    }
    if (d.perm) {
        d.b = (const double *)tail;
        d.x = (double *)(tail + s64);
    }
}

void describe_slot(kcfd_batch *b, int k, kcfd_mat *m, const double *bvec, double *x) {
    SysDesc &d = b->h_desc[k];
    std::memset(&d, 0, sizeof(d));
    d.n = m->n;
    fill_desc(d, m);
    // folded product: BiCGSTAB over a two-colour ILU(0); MIXED needs the
    // FP64 colour-0 pivots of the mixed set-up
    const bool mixed = b->bp.precision == KCFD_MIXED;
    if (b->fold && b->bp.method == KCFD_METHOD_BICGSTAB && b->bp.precond == KCFD_PC_ILU0 && m->two_colour &&
        m->ncolour0 > 0 && (!mixed || m->c0piv)) {
        d.nid = m->ncolour0;
        d.c0piv = mixed ? m->c0piv : nullptr;
    }
    d.b = d.b_user = bvec;
    d.x = d.x_user = x;
    b->mats[k] = m;
    b->sys_n[k] = m->n;
}

// grow a device table to hold `need` entries (contents are rewritten after)
template <typename T>
int ensure_table(T **p, int *cap, int need, bool *grew) {
    if (need <= *cap && *p) return KCFD_OK;
    int c = std::max(64, std::max(need, 2 * *cap));
    hipFree((void *)*p);
    *p = nullptr;
    KCFD_HIP(hipMalloc((void **)p, sizeof(T) * (size_t)c));
    *cap = c;
    *grew = true;
    return KCFD_OK;
}

// Per-system pieces of the tile and chunk tables: its tiles in the order
// they are listed, its sync-free ILU(0) chunks per level, its two-colour
// chunk runs.  Built when the membership changes; the device tables are
// concatenations of these, so dropping finished systems from the tables is
// a cheap copy.
void build_sys_tables(kcfd_batch *b, int k) {
    SysTables &t = b->sys_tables[k];
    t = SysTables{};
    const kcfd_mat *m = b->mats[k];
    if (!m) return;
    const SysDesc &d = b->h_desc[k];
    const int kTile = b->rows * kBlock;   // rows per tile
    const int nt = std::max(1, (d.n + kTile - 1) / kTile);
    if (d.perm && !m->h_perm.empty()) {
        // permuted row order: tiles listed by the row of A they start at, so
        // that the tiles of both colours over one region of A sit together in
        // the list and fall into the same XCD's range (its gathers then hit
        // values its own L2 wrote)
        std::vector<std::pair<int, int>> key(nt);
        for (int q = 0; q < nt; ++q) key[q] = {m->h_perm[(size_t)q * kTile], q};
        std::sort(key.begin(), key.end());
        for (const auto &kt : key) t.tiles.push_back(make_int2(k, kt.second * kTile));
    } else {
        for (int q = 0; q < nt; ++q) t.tiles.push_back(make_int2(k, q * kTile));
    }
    if (b->bp.precond != KCFD_PC_ILU0) return;
    if (m->seg2) {
        for (int c = 0; c < 2; ++c) {
            const std::vector<int> &grp = m->h_sgg[c];
            std::vector<Chunk> &out = c ? t.sg1 : t.sg0;
            for (size_t q = 0; q + 3 < grp.size(); q += 4)
                out.push_back(make_chunk(k, grp[q], grp[q + 1], grp[q + 3], grp[q + 2], 0));
        }
        return;
    }
    if (m->two_colour) {
        const int n0 = m->ncolour0, n1 = m->n - m->ncolour0;
        auto chunk = [&](int start, int count, int local, bool upper) {
            const std::vector<int> &off = upper ? m->h_chu_off : m->h_chl_off;
            const std::vector<int> &base = upper ? m->h_chu_base : m->h_chl_base;
            return make_chunk(k, start, count, off[local], (off[local + 1] - off[local]) >> 6,
                              m->ch16 ? base[local] : 0);
        };
        for (int s0 = 0; s0 < n1; s0 += 64) t.ca.push_back(chunk(n0 + s0, std::min(64, n1 - s0), s0 / 64, false));
        for (int s0 = 0; s0 < n0; s0 += 64) t.cb.push_back(chunk(s0, std::min(64, n0 - s0), s0 / 64, true));
        return;
    }
    for (int upper = 0; upper < 2; ++upper) {
        const std::vector<int> &lp = upper ? m->h_lvl_ptr_u : m->h_lvl_ptr_l;
        const std::vector<int> &cp = upper ? m->h_chu_ptr : m->h_chl_ptr;
        const int nl = upper ? m->nlev_u : m->nlev_l;
        std::vector<std::vector<Chunk>> &lv = upper ? t.lvl_u : t.lvl_l;
        lv.assign(nl, {});
        const std::vector<int> &off = upper ? m->h_chu_off : m->h_chl_off;
        for (int l = 0; l < nl; ++l)
            for (int s0 = lp[l]; s0 < lp[l + 1]; s0 += 64) {
                const int local = cp[l] + (s0 - lp[l]) / 64;
                lv[l].push_back(make_chunk(k, s0, std::min(64, lp[l + 1] - s0), off[local],
                                           (off[local + 1] - off[local]) >> 6, 0));
            }
    }
}

// Tile list (slot order), per-slot tile ranges, and the level-major ILU chunk
// tables of the systems in `only` (null: every member), uploaded on the
// solver stream.  A system's tiles stay contiguous and in their order, so
// its partial sums -- and its arithmetic -- do not depend on which other
// systems are in the tables.  A grown table invalidates the captured graph
// (its kernels hold the old pointers).
int rebuild_tables(kcfd_batch *b, const std::vector<char> *only = nullptr) {
    hipStream_t s = b->ctx->stream;
    if (b->tables_dirty || (int)b->sys_tables.size() != b->nsys) {
        b->sys_tables.assign(b->nsys, SysTables{});
        for (int k = 0; k < b->nsys; ++k) build_sys_tables(b, k);
        b->tables_dirty = false;
    }
    auto in = [&](int k) { return b->mats[k] && (!only || (*only)[k]); };
    std::vector<int2> tiles;
    for (int k = 0; k < b->nsys; ++k) {
        SysDesc &d = b->h_desc[k];
        d.tile0 = (int)tiles.size();
        d.ntiles = 0;
        if (!in(k)) continue;
        const SysTables &t = b->sys_tables[k];
        d.ntiles = (int)t.tiles.size();
        tiles.insert(tiles.end(), t.tiles.begin(), t.tiles.end());
    }
    b->ntiles = (int)tiles.size();
    bool grew = false;
    const int old_cap = b->cap_tiles;
    KCFD_CHECK(ensure_table(&b->d_tiles, &b->cap_tiles, b->ntiles, &grew));
    if (b->cap_tiles != old_cap || !b->d_partials) {
        hipFree(b->d_partials);
        b->d_partials = nullptr;
        KCFD_HIP(hipMalloc(&b->d_partials, sizeof(double) * kMaxPartials * (size_t)b->cap_tiles));
    }
    std::vector<Chunk> cl, cu, ca, cb, sg0, sg1;
    if (b->bp.precond == KCFD_PC_ILU0) {
        // sync-free chunks in level-major order over the systems (the solve's
        // progress argument needs every chunk of level l listed before any of
        // level l + 1)
        size_t maxlev_l = 0, maxlev_u = 0;
        for (int k = 0; k < b->nsys; ++k)
            if (in(k)) {
                maxlev_l = std::max(maxlev_l, b->sys_tables[k].lvl_l.size());
                maxlev_u = std::max(maxlev_u, b->sys_tables[k].lvl_u.size());
            }
        for (size_t l = 0; l < maxlev_l; ++l)
            for (int k = 0; k < b->nsys; ++k)
                if (in(k) && l < b->sys_tables[k].lvl_l.size())
                    cl.insert(cl.end(), b->sys_tables[k].lvl_l[l].begin(), b->sys_tables[k].lvl_l[l].end());
        for (size_t l = 0; l < maxlev_u; ++l)
            for (int k = 0; k < b->nsys; ++k)
                if (in(k) && l < b->sys_tables[k].lvl_u.size())
                    cu.insert(cu.end(), b->sys_tables[k].lvl_u[l].begin(), b->sys_tables[k].lvl_u[l].end());
        // two-colour members: runs of 64 rows of colour 1 (a) and colour 0 (b)
        for (int k = 0; k < b->nsys; ++k)
            if (in(k)) {
                ca.insert(ca.end(), b->sys_tables[k].ca.begin(), b->sys_tables[k].ca.end());
                cb.insert(cb.end(), b->sys_tables[k].cb.begin(), b->sys_tables[k].cb.end());
                sg0.insert(sg0.end(), b->sys_tables[k].sg0.begin(), b->sys_tables[k].sg0.end());
                sg1.insert(sg1.end(), b->sys_tables[k].sg1.begin(), b->sys_tables[k].sg1.end());
            }
        b->nsg0 = (int)sg0.size();
        b->nsg1 = (int)sg1.size();
        KCFD_CHECK(ensure_table(&b->d_sg0, &b->cap_sg0, b->nsg0, &grew));
        KCFD_CHECK(ensure_table(&b->d_sg1, &b->cap_sg1, b->nsg1, &grew));
        if (!sg0.empty())
            KCFD_HIP(hipMemcpyAsync(b->d_sg0, sg0.data(), sizeof(Chunk) * sg0.size(), hipMemcpyHostToDevice, s));
        if (!sg1.empty())
            KCFD_HIP(hipMemcpyAsync(b->d_sg1, sg1.data(), sizeof(Chunk) * sg1.size(), hipMemcpyHostToDevice, s));
        b->nchunks_l = (int)cl.size();
        b->nchunks_u = (int)cu.size();
        b->nchunks_a = (int)ca.size();
        b->nchunks_b = (int)cb.size();
        KCFD_CHECK(ensure_table(&b->d_chunks_l, &b->cap_chunks_l, b->nchunks_l, &grew));
        KCFD_CHECK(ensure_table(&b->d_chunks_u, &b->cap_chunks_u, b->nchunks_u, &grew));
        KCFD_CHECK(ensure_table(&b->d_chunks_a, &b->cap_chunks_a, b->nchunks_a, &grew));
        KCFD_CHECK(ensure_table(&b->d_chunks_b, &b->cap_chunks_b, b->nchunks_b, &grew));
        if (!ca.empty())
            KCFD_HIP(hipMemcpyAsync(b->d_chunks_a, ca.data(), sizeof(Chunk) * ca.size(), hipMemcpyHostToDevice, s));
        if (!cb.empty())
            KCFD_HIP(hipMemcpyAsync(b->d_chunks_b, cb.data(), sizeof(Chunk) * cb.size(), hipMemcpyHostToDevice, s));
        std::vector<const int *> pl(b->nsys, nullptr), pu(b->nsys, nullptr);
        for (int k = 0; k < b->nsys; ++k)
            if (b->mats[k]) {
                pl[k] = b->mats[k]->lvl_rows_l;
                pu[k] = b->mats[k]->lvl_rows_u;
            }
        KCFD_HIP(hipMemcpyAsync((void *)b->d_lvl_l, pl.data(), sizeof(int *) * b->nsys, hipMemcpyHostToDevice, s));
        KCFD_HIP(hipMemcpyAsync((void *)b->d_lvl_u, pu.data(), sizeof(int *) * b->nsys, hipMemcpyHostToDevice, s));
        if (!cl.empty())
            KCFD_HIP(hipMemcpyAsync(b->d_chunks_l, cl.data(), sizeof(Chunk) * cl.size(), hipMemcpyHostToDevice, s));
        if (!cu.empty())
            KCFD_HIP(hipMemcpyAsync(b->d_chunks_u, cu.data(), sizeof(Chunk) * cu.size(), hipMemcpyHostToDevice, s));
    }
    if (!tiles.empty())
        KCFD_HIP(hipMemcpyAsync(b->d_tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice, s));
    KCFD_HIP(hipMemcpyAsync(b->d_desc, b->h_desc.data(), sizeof(SysDesc) * b->nsys, hipMemcpyHostToDevice, s));
    const int walk = b->opts.tile_walk == 2 ? 1 : 0;
    int folded = 0;
    for (int k = 0; k < b->nsys; ++k) folded |= in(k) && b->h_desc[k].nid > 0;
    const int counts[kCounts] = {b->ntiles, b->nchunks_l, b->nchunks_u, b->nchunks_a, b->nchunks_b,
                                 walk, folded, 0, b->nsg0, b->nsg1};
    KCFD_HIP(hipMemcpyAsync(b->d_counts, counts, sizeof(counts), hipMemcpyHostToDevice, s));
    // the host arrays above are locals: finish the copies before they go
    KCFD_HIP(hipStreamSynchronize(s));
    if (grew && b->graph) {
        hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    b->dirty = false;
    b->compacted = only != nullptr;
    return KCFD_OK;
}

void set_launch_shape(kcfd_batch *b, const kcfd_solve_opts *o);

int build_batch(kcfd_ctx *ctx, int nsys, kcfd_mat *const *mats, const double *const *bs,
                double *const *xs, const kcfd_solve_opts *o, kcfd_batch **out) {
    kcfd_batch *b = nullptr;
    int rc = new_batch(ctx, nsys, o, &b);
    auto fail = [&](int code) {
        kcfd_batch_destroy(b);
        return code;
    };
    if (rc != KCFD_OK) return b ? fail(rc) : rc;
    size_t slab = 0;
    std::vector<size_t> off(nsys);
    for (int k = 0; k < nsys; ++k) {
        if ((rc = check_mat(b, mats[k], k)) != KCFD_OK) return fail(rc);
        describe_slot(b, k, mats[k], bs[k], xs[k]);
        off[k] = slab;
        slab += align_up(slot_bytes(b, b->h_desc[k]), 256);
    }
    if (hipMalloc(&b->slab, slab ? slab : 256) != hipSuccess) {
        set_error("out of device memory for the batch vectors (%zu bytes)", slab);
        return fail(KCFD_EOOM);
    }
    b->slab_bytes = slab;
    for (int k = 0; k < nsys; ++k) place_vectors(b, k, (char *)b->slab + off[k]);
    if ((rc = rebuild_tables(b)) != KCFD_OK) return fail(rc);
    set_launch_shape(b, o);
    *out = b;
    return KCFD_OK;
}

// which ILU(0) kernels run, the persistent grid and the triangular-solve grid
void set_launch_shape(kcfd_batch *b, const kcfd_solve_opts *o) {
    kcfd_ctx *ctx = b->ctx;
    b->run_two = b->nchunks_a + b->nchunks_b > 0;
    b->run_polled = b->nchunks_l + b->nchunks_u > 0;
    b->run_seg = b->nsg0 + b->nsg1 > 0;
    // persistent grid: at most grid_per_cu workgroups per CU, a multiple of
    // the XCD count
    const int gpc = o->grid_per_cu > 0 ? o->grid_per_cu : 8;
    const int cap = std::max(kXcds, (ctx->cus * gpc / kXcds) * kXcds);
    b->grid = std::min(xcd_grid(b->ntiles), cap);
    if (o->precond == KCFD_PC_ILU0) {
        // Waves beyond a few levels of look-ahead only poll (and load the
        // memory system with polls), so the grid covers about three of the
        // widest levels (summed over the batch).  At most 4 workgroups (16
        // waves) per CU: the kernel's occupancy limit is 32 waves per CU, so
        // the whole grid is resident.
        int widest = 0, maxlev = 0;
        for (auto *m : b->mats) maxlev = std::max(maxlev, std::max(m->nlev_l, m->nlev_u));
        for (int l = 0; l < maxlev; ++l) {
            int w = 0;
            for (auto *m : b->mats)
                if (l < m->nlev_l) w += (m->h_lvl_ptr_l[l + 1] - m->h_lvl_ptr_l[l] + 63) / 64;
            widest = std::max(widest, w);
        }
        const int want_waves = std::max(64, 3 * widest);
        const int tcap = ctx->cus * (o->trsv_per_cu > 0 ? o->trsv_per_cu : 4);
        b->trsv_grid = std::max(1, std::min((want_waves + kWaves - 1) / kWaves, tcap));
    }
}

int capture_graph(kcfd_batch *b, int iters) {
    hipStream_t s = b->ctx->stream;
    if (b->graph) {
        hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    LaunchFn fn = pick_iter(b->bp.precision, b->bp.precond, b->bp.method, b->rows);
    hipGraph_t g = nullptr;
    KCFD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < iters; ++k) fn(b, s);
    hipError_t e = hipStreamEndCapture(s, &g);
    if (e != hipSuccess) {
        set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
        return KCFD_EHIP;
    }
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
        return KCFD_EHIP;
    }
    b->graph_iters = iters;
    return KCFD_OK;
}

// The iterations between host checks run as one captured graph, or, with
// KCFD_BATCH_GRAPH=0 in the environment, as direct launches of the same
// kernels (same arguments: everything they read lives in device memory).
// The direct mode is for tools that trace kernels: rocprofv3's kernel trace
// has aborted the process inside hipGraphLaunch of the batch's graph.
bool batch_graph_enabled(const kcfd_batch *b) {
    if (b->no_graph) return false;
    const char *e = std::getenv("KCFD_BATCH_GRAPH");
    return !(e && e[0] == '0');
}

int replay_iterations(kcfd_batch *b, hipStream_t s) {
    if (!batch_graph_enabled(b)) {
        LaunchFn fn = pick_iter(b->bp.precision, b->bp.precond, b->bp.method, b->rows);
        for (int k = 0; k < b->opts.check_every; ++k) fn(b, s);
        KCFD_HIP(hipGetLastError());
        b->graph_iters = b->opts.check_every;
        return KCFD_OK;
    }
    if (!b->graph || b->graph_iters != b->opts.check_every)
        KCFD_CHECK(capture_graph(b, b->opts.check_every));
    KCFD_HIP(hipGraphLaunch(b->graph, s));
    return KCFD_OK;
}

// the fail word of the sync-free triangular solves (read after the state
// copy that the caller has synchronised)
int check_trsv(const kcfd_batch *b, unsigned int fail) {
    if (fail) {
        set_error("ILU(0) sync-free triangular solve: a wait for an operand exceeded %d polls; "
                  "the solve's result is invalid", b->bp.poll_limit);
        return KCFD_ESTATE;
    }
    return KCFD_OK;
}

int run_batch(kcfd_batch *b, kcfd_solve_result *res) {
    kcfd_ctx *ctx = b->ctx;
    hipStream_t s = ctx->stream;
    for (int k = 0; k < b->nsys; ++k) {
        SysState st{};
        const int n = b->sys_n[k];
        st.status = S_NEW;
        st.maxiter = b->opts.maxiter > 0 ? b->opts.maxiter : 10 * std::max(n, 1);
        b->h_state[k] = st;
    }
    KCFD_HIP(hipMemcpyAsync(b->d_state, b->h_state.data(), sizeof(SysState) * b->nsys,
                            hipMemcpyHostToDevice, s));
    KCFD_HIP(hipMemcpyAsync(b->d_params, &b->bp, sizeof(BatchParams), hipMemcpyHostToDevice, s));
    if (batch_graph_enabled(b) && (!b->graph || b->graph_iters != b->opts.check_every))
        KCFD_CHECK(capture_graph(b, b->opts.check_every));
    KCFD_HIP(hipMemsetAsync(b->d_counters, 0, sizeof(unsigned int) * 8, s));
    KCFD_HIP(hipEventRecord(ctx->ev0, s));
    pick_init(b->bp.precision, b->bp.precond, b->rows)(b, s);
    KCFD_HIP(hipGetLastError());
    unsigned int fail = 0;
    long long guard = 0;
    int maxit = 0;
    for (int k = 0; k < b->nsys; ++k) maxit = std::max(maxit, b->h_state[k].maxiter);
    // every graph iteration advances each running system by one iteration or
    // one reliable update
    const long long limit = 2LL * maxit + 2LL * b->bp.max_refine + 64;
    // KCFD_BATCH_TRACE=1: one stderr line per host check (iterations so far,
    // systems still running, tiles in the tables, host time since the start)
    const char *tr_env = std::getenv("KCFD_BATCH_TRACE");
    const bool trace = tr_env && tr_env[0] && tr_env[0] != '0';
    const auto t_start = std::chrono::steady_clock::now();
    int tr_tiles = b->ntiles;
    for (;;) {
        KCFD_CHECK(replay_iterations(b, s));
        guard += b->graph_iters;
        KCFD_HIP(hipMemcpyAsync(b->h_state.data(), b->d_state, sizeof(SysState) * b->nsys,
                                hipMemcpyDeviceToHost, s));
        if (b->run_polled)
            KCFD_HIP(hipMemcpyAsync(&fail, b->d_counters + kTrsvFail, sizeof(unsigned int),
                                    hipMemcpyDeviceToHost, s));
        KCFD_HIP(hipStreamSynchronize(s));
        KCFD_CHECK(check_trsv(b, fail));
        if (trace) {
            int running = 0;
            for (int k = 0; k < b->nsys; ++k) running += b->mats[k] && b->h_state[k].status != S_DONE;
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
                                  .count();
            std::fprintf(stderr, "[kcfd batch] iterations %lld running %d tiles %d at %.3f ms\n", guard, running,
                         tr_tiles, ms);
        }
        bool all = true;
        for (int k = 0; k < b->nsys; ++k)
            if (b->h_state[k].status != S_DONE) { all = false; break; }
        if (all) break;
        if (guard > limit) {
            set_error("solver did not reach a terminal state after %lld graph iterations", guard);
            return KCFD_ESTATE;
        }
        // Systems that finished leave the tables once they hold a fifth of
        // the tiles: every kernel walks the tile and chunk lists, and a
        // finished system's tiles would still cost a load chain each.
        std::vector<char> act(b->nsys, 0);
        int act_tiles = 0;
        for (int k = 0; k < b->nsys; ++k)
            if (b->mats[k] && b->h_state[k].status != S_DONE) {
                act[k] = 1;
                act_tiles += (int)b->sys_tables[k].tiles.size();
            }
        if (act_tiles < b->ntiles - b->ntiles / 5) {
            KCFD_CHECK(rebuild_tables(b, &act));
            tr_tiles = act_tiles;
        }
    }
    // the finishing kernels (true residual, restore, scatter) serve every
    // system: the full tables again
    if (b->compacted) KCFD_CHECK(rebuild_tables(b));
    KCFD_HIP(hipEventRecord(ctx->ev1, s));
    pick_finish(b->rows)(b, s);
    KCFD_HIP(hipGetLastError());
    KCFD_HIP(hipMemcpyAsync(b->h_state.data(), b->d_state, sizeof(SysState) * b->nsys,
                            hipMemcpyDeviceToHost, s));
    KCFD_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    KCFD_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms = ms;
    if (res) {
        for (int k = 0; k < b->nsys; ++k) {
            const SysState &st = b->h_state[k];
            kcfd_solve_result &r = res[k];
            r.info = st.info;
            r.iters = b->bp.precision == KCFD_FP64 ? st.iters : st.total_iters;
            r.half_step = st.half;
            r.refinements = st.refinements;
            r.rel_resid = st.true_bb > 0 ? sqrt(st.true_rr / st.true_bb) : sqrt(st.true_rr);
            r.bytes = model_bytes(b->mats[k], b->bp.precond, b->bp.precision, st, b->bp.method,
                                  b->h_desc[k].nid);
            r.ms = ms;
        }
    }
    return KCFD_OK;
}

// ------------------------------------------------- allocation-free solve
// kcfd_solve_noalloc keeps its whole workspace -- descriptor, state,
// parameters, counts, tile list, partials, ILU(0) chunk tables and the
// Krylov vectors -- in this array, one per device, held until the process
// ends.  Host code never uses the array's name as an address: kcfd_init
// resolves the device address with hipGetSymbolAddress (kcfd_ctx::noalloc)
// and everything below works from that pointer.
// one lock per device: calls on one GPU share its array
std::mutex g_noalloc_mu[64];

struct Carve {
    char *base;
    size_t used, cap;
    template <typename T>
    T *take(size_t count) {
        const size_t off = align_up(used, 256);
        used = off + align_up(std::max<size_t>(count, 1) * sizeof(T), 256);
        return used <= cap ? reinterpret_cast<T *>(base + off) : nullptr;
    }
};

int solve_noalloc(kcfd_mat *m, const double *bvec, double *x, const kcfd_solve_opts *o,
                  kcfd_solve_result *res) {
    kcfd_ctx *ctx = m->ctx;
    if (!ctx->noalloc) {
        set_error("kcfd_solve_noalloc: the context has no scratch address (kcfd_init failed to resolve it)");
        return KCFD_ESTATE;
    }
    kcfd_batch bb;
    kcfd_batch *b = &bb;
    init_batch_host(b, ctx, 1, o);
    b->no_graph = true;
    KCFD_CHECK(check_mat(b, m, 0));
    describe_slot(b, 0, m, bvec, x);
    // table sizes of this system
    const int ntiles = std::max(1, (m->n + b->rows * kBlock - 1) / (b->rows * kBlock));
    int nl = 0, nu = 0, na = 0, nb = 0, ns0 = 0, ns1 = 0;
    if (b->bp.precond == KCFD_PC_ILU0) {
        if (m->seg2) {
            ns0 = (int)m->h_sgg[0].size() / 4;
            ns1 = (int)m->h_sgg[1].size() / 4;
        } else if (m->two_colour) {
            na = (m->n - m->ncolour0 + 63) / 64;
            nb = (m->ncolour0 + 63) / 64;
        } else {
            for (int l = 0; l < m->nlev_l; ++l) nl += (m->h_lvl_ptr_l[l + 1] - m->h_lvl_ptr_l[l] + 63) / 64;
            for (int l = 0; l < m->nlev_u; ++l) nu += (m->h_lvl_ptr_u[l + 1] - m->h_lvl_ptr_u[l] + 63) / 64;
        }
    }
    Carve c{reinterpret_cast<char *>(ctx->noalloc), 0, kNoAllocBytes};
    b->d_desc = c.take<SysDesc>(1);
    b->d_state = c.take<SysState>(1);
    b->d_params = c.take<BatchParams>(1);
    b->d_counts = c.take<int>(kCounts);
    b->d_counters = c.take<unsigned int>(8);
    b->d_tiles = c.take<int2>(ntiles);
    b->d_partials = c.take<double>((size_t)kMaxPartials * ntiles);
    b->d_chunks_l = c.take<Chunk>(nl);
    b->d_chunks_u = c.take<Chunk>(nu);
    b->d_chunks_a = c.take<Chunk>(na);
    b->d_chunks_b = c.take<Chunk>(nb);
    b->d_sg0 = c.take<Chunk>(ns0);
    b->d_sg1 = c.take<Chunk>(ns1);
    b->d_lvl_l = c.take<const int *>(1);
    b->d_lvl_u = c.take<const int *>(1);
    char *vecs = c.take<char>(slot_bytes(b, b->h_desc[0]));
    if (c.used > c.cap) {
        set_error("kcfd_solve_noalloc: a system of %d rows needs %zu bytes of workspace, more than the "
                  "%zu-byte scratch array holds", m->n, c.used, c.cap);
        return KCFD_EARG;
    }
    b->cap_tiles = ntiles;
    b->cap_chunks_l = nl;
    b->cap_chunks_u = nu;
    b->cap_chunks_a = na;
    b->cap_chunks_b = nb;
    b->cap_sg0 = ns0;
    b->cap_sg1 = ns1;
    place_vectors(b, 0, vecs);
    // every table fits its slice, so rebuild_tables allocates nothing
    KCFD_CHECK(rebuild_tables(b));
    set_launch_shape(b, o);
    // the device arrays are slices of the scratch array: nothing to free
    return run_batch(b, res);
}

}  // namespace

extern "C" {

int kcfd_batch_create(kcfd_ctx *ctx, int nsys, kcfd_mat *const *mats, const double *const *bs,
                      double *const *xs, const kcfd_solve_opts *opts, kcfd_batch **out) {
    KCFD_ARG(ctx && mats && bs && xs && opts && out, "null argument");
    KCFD_ARG(nsys > 0, "empty batch");
    for (int k = 0; k < nsys; ++k) KCFD_ARG(mats[k] && bs[k] && xs[k], "null system entry");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    KCFD_HIP(hipSetDevice(ctx->device));
    return build_batch(ctx, nsys, mats, bs, xs, opts, out);
}

int kcfd_batch_run(kcfd_batch *b, kcfd_solve_result *results) {
    KCFD_ARG(b != nullptr, "null batch");
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    KCFD_HIP(hipSetDevice(b->ctx->device));
    return run_batch(b, results);
}

int kcfd_batch_destroy(kcfd_batch *b) {
    if (!b) return KCFD_OK;
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    if (b->graph) hipGraphExecDestroy(b->graph);
    hipFree(b->d_desc);
    hipFree(b->d_state);
    hipFree(b->d_tiles);
    hipFree(b->d_partials);
    hipFree(b->d_counters);
    hipFree(b->d_counts);
    hipFree(b->d_params);
    hipFree(b->slab);
    for (void *p : b->slot_mem) hipFree(p);
    for (auto &sp : b->spare) hipFree(sp.second);
    hipFree(b->d_chunks_l);
    hipFree(b->d_chunks_u);
    hipFree(b->d_chunks_a);
    hipFree(b->d_chunks_b);
    hipFree(b->d_sg0);
    hipFree(b->d_sg1);
    hipFree((void *)b->d_lvl_l);
    hipFree((void *)b->d_lvl_u);
    delete b;
    return KCFD_OK;
}

// ------------------------------------------------------- dynamic batches

int kcfd_batch_create_dynamic(kcfd_ctx *ctx, int capacity, const kcfd_solve_opts *opts,
                              kcfd_batch **out) {
    KCFD_ARG(ctx && opts && out, "null argument");
    KCFD_ARG(capacity > 0, "capacity must be positive");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    KCFD_HIP(hipSetDevice(ctx->device));
    kcfd_batch *b = nullptr;
    int rc = new_batch(ctx, capacity, opts, &b);
    if (rc != KCFD_OK) {
        if (b) kcfd_batch_destroy(b);
        return rc;
    }
    b->dynamic = true;
    b->run_two = b->run_polled = opts->precond == KCFD_PC_ILU0;   // any membership
    // the graph is captured once for any membership: full-size launches
    const int gpc = opts->grid_per_cu > 0 ? opts->grid_per_cu : 8;
    b->grid = std::max(kXcds, (ctx->cus * gpc / kXcds) * kXcds);
    b->trsv_grid = ctx->cus * (opts->trsv_per_cu > 0 ? opts->trsv_per_cu : 4);
    if ((rc = rebuild_tables(b)) != KCFD_OK) {
        kcfd_batch_destroy(b);
        return rc;
    }
    *out = b;
    return KCFD_OK;
}

int kcfd_batch_attach(kcfd_batch *b, kcfd_mat *m, const double *bvec, double *x, int *slot) {
    KCFD_ARG(b && m && slot && (m->n == 0 || (bvec && x)), "null argument");
    KCFD_ARG(b->dynamic, "not a dynamic batch");
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    KCFD_HIP(hipSetDevice(b->ctx->device));
    int k = 0;
    while (k < b->nsys && b->mats[k]) ++k;
    if (k == b->nsys) {
        set_error("dynamic batch full (%d slots)", b->nsys);
        return KCFD_EBUSY;
    }
    KCFD_CHECK(check_mat(b, m, k));
    if (b->bp.precond == KCFD_PC_ILU0 && m->seg2 && !b->run_seg) {
        // the first segment-ordered member: the iterations launch the
        // segment apply from now on (the graph is captured again)
        b->run_seg = true;
        if (b->graph) {
            KCFD_HIP(hipStreamSynchronize(b->ctx->stream));
            hipGraphExecDestroy(b->graph);
            b->graph = nullptr;
        }
    }
    describe_slot(b, k, m, bvec, x);
    void *mem = nullptr;
    const size_t bytes = slot_bytes(b, b->h_desc[k]);
    for (size_t q = 0; q < b->spare.size(); ++q)
        if (b->spare[q].first == bytes) {   // a detached system's block of this size
            mem = b->spare[q].second;
            b->spare[q] = b->spare.back();
            b->spare.pop_back();
            break;
        }
    if (!mem && hipMalloc(&mem, bytes) != hipSuccess) {
        (void)hipGetLastError();
        std::memset(&b->h_desc[k], 0, sizeof(SysDesc));
        b->mats[k] = nullptr;
        b->sys_n[k] = 0;
        set_error("out of device memory for a system's vectors (%zu bytes)", bytes);
        return KCFD_EOOM;
    }
    b->slot_mem[k] = mem;
    place_vectors(b, k, (char *)mem);
    SysState st{};
    st.status = S_NEW;
    st.maxiter = b->opts.maxiter > 0 ? b->opts.maxiter : 10 * std::max(m->n, 1);
    b->h_state[k] = st;
    b->fresh[k] = 1;
    b->reported[k] = 0;
    b->dirty = true;
    if ((int)b->sys_tables.size() == b->nsys) build_sys_tables(b, k);
    else b->tables_dirty = true;
    *slot = k;
    return KCFD_OK;
}

int kcfd_batch_step(kcfd_batch *b, int *finished, int *nfinished) {
    KCFD_ARG(b && nfinished, "null argument");
    KCFD_ARG(b->dynamic, "not a dynamic batch");
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    KCFD_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    *nfinished = 0;
    if (b->dirty) KCFD_CHECK(rebuild_tables(b));
    bool any_new = false, any = false;
    for (int k = 0; k < b->nsys; ++k) {
        if (!b->mats[k]) continue;
        any = true;
        if (b->fresh[k]) {
            KCFD_HIP(hipMemcpyAsync(b->d_state + k, &b->h_state[k], sizeof(SysState),
                                    hipMemcpyHostToDevice, s));
            any_new = true;
        }
    }
    if (!any) return KCFD_OK;
    KCFD_HIP(hipMemcpyAsync(b->d_params, &b->bp, sizeof(BatchParams), hipMemcpyHostToDevice, s));
    if (batch_graph_enabled(b) && (!b->graph || b->graph_iters != b->opts.check_every))
        KCFD_CHECK(capture_graph(b, b->opts.check_every));
    if (any_new) pick_init(b->bp.precision, b->bp.precond, b->rows)(b, s);
    KCFD_CHECK(replay_iterations(b, s));
    pick_finish(b->rows)(b, s);
    KCFD_HIP(hipGetLastError());
    KCFD_HIP(hipMemcpyAsync(b->h_state.data(), b->d_state, sizeof(SysState) * b->nsys,
                            hipMemcpyDeviceToHost, s));
    unsigned int fail = 0;
    if (b->run_polled)
        KCFD_HIP(hipMemcpyAsync(&fail, b->d_counters + kTrsvFail, sizeof(unsigned int), hipMemcpyDeviceToHost, s));
    KCFD_HIP(hipStreamSynchronize(s));
    KCFD_CHECK(check_trsv(b, fail));
    for (int k = 0; k < b->nsys; ++k) {
        b->fresh[k] = 0;
        if (b->mats[k] && !b->reported[k] && b->h_state[k].status == S_DONE && b->h_state[k].true_done) {
            b->reported[k] = 1;
            if (finished) finished[*nfinished] = k;
            ++*nfinished;
        }
    }
    return KCFD_OK;
}

int kcfd_batch_detach(kcfd_batch *b, int slot, kcfd_solve_result *res) {
    KCFD_ARG(b != nullptr, "null batch");
    KCFD_ARG(b->dynamic, "not a dynamic batch");
    KCFD_ARG(slot >= 0 && slot < b->nsys && b->mats[slot], "no system in this slot");
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    KCFD_HIP(hipSetDevice(b->ctx->device));
    const SysState &st = b->h_state[slot];
    if (res) {
        std::memset(res, 0, sizeof(*res));
        res->info = st.status == S_DONE ? st.info : KCFD_EBUSY;
        res->iters = b->bp.precision == KCFD_FP64 ? st.iters : st.total_iters;
        res->half_step = st.half;
        res->refinements = st.refinements;
        res->rel_resid = st.true_done ? (st.true_bb > 0 ? sqrt(st.true_rr / st.true_bb) : sqrt(st.true_rr))
                                      : NAN;
        res->bytes = model_bytes(b->mats[slot], b->bp.precond, b->bp.precision, st, b->bp.method,
                                 b->h_desc[slot].nid);
        res->ms = 0.0;
    }
    // the block is kept for a later attach: its kernels follow this batch's
    // on the same stream
    b->spare.emplace_back(slot_bytes(b, b->h_desc[slot]), b->slot_mem[slot]);
    b->slot_mem[slot] = nullptr;
    b->mats[slot] = nullptr;
    b->sys_n[slot] = 0;
    std::memset(&b->h_desc[slot], 0, sizeof(SysDesc));
    SysState fs{};
    fs.status = S_FREE;
    b->h_state[slot] = fs;
    KCFD_HIP(hipMemcpyAsync(b->d_state + slot, &b->h_state[slot], sizeof(SysState),
                            hipMemcpyHostToDevice, b->ctx->stream));
    KCFD_HIP(hipStreamSynchronize(b->ctx->stream));
    b->fresh[slot] = 0;
    b->reported[slot] = 0;
    b->dirty = true;
    if ((int)b->sys_tables.size() == b->nsys) build_sys_tables(b, slot);
    else b->tables_dirty = true;
    return KCFD_OK;
}

int kcfd_batch_trim(kcfd_batch *b) {
    KCFD_ARG(b != nullptr, "null batch");
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    if (b->spare.empty()) return KCFD_OK;
    KCFD_HIP(hipSetDevice(b->ctx->device));
    KCFD_HIP(hipStreamSynchronize(b->ctx->stream));
    for (auto &sp : b->spare) hipFree(sp.second);
    b->spare.clear();
    return KCFD_OK;
}

int kcfd_batch_members(kcfd_batch *b, int *members) {
    KCFD_ARG(b && members, "null argument");
    int n = 0;
    for (auto *m : b->mats) n += m != nullptr;
    *members = n;
    return KCFD_OK;
}

int kcfd_mat_solver(kcfd_mat *m, kcfd_batch **out) {
    KCFD_ARG(m && out, "null argument");
    *out = m->solver;
    if (!m->solver) {
        set_error("no cached solver: call kcfd_solve first");
        return KCFD_ESTATE;
    }
    return KCFD_OK;
}

int kcfd_batch_bytes(kcfd_batch *b, long long *bytes) {
    KCFD_ARG(b && bytes, "null argument");
    long long v = (long long)b->slab_bytes;
    for (int k = 0; k < b->nsys; ++k)
        if (b->slot_mem[k]) v += (long long)slot_bytes(b, b->h_desc[k]);
    for (auto &sp : b->spare) v += (long long)sp.first;
    *bytes = v + (long long)b->cap_tiles * (8 + 8 * kMaxPartials) +
             (long long)b->nsys * (long long)(sizeof(SysDesc) + sizeof(SysState) + 2 * sizeof(int *)) +
             (long long)(b->cap_chunks_l + b->cap_chunks_u) * (long long)sizeof(Chunk);
    return KCFD_OK;
}

// Time one phase kernel inside real iterations (see the loop below), with
// the events in the kernels' own dispatch packets (launch_k).
// phase: 1 = K2 (v = A M p + rt.v partials), 2 = K5 (x/r update), 3 = the
// ILU(0) apply to p (two-colour: both kernels; sync-free: the L solve),
// 4 = K1 (p update).
int kcfd_batch_profile(kcfd_batch *b, int phase, int reps, double *avg_ms) {
    KCFD_ARG(b && avg_ms && reps > 0, "bad argument");
    KCFD_ARG(phase >= 1 && phase <= 6, "unknown phase");
    KCFD_ARG(phase < 3 || phase == 4 || b->bp.precond == KCFD_PC_ILU0, "phases 3, 5 and 6 need ILU(0)");
    KCFD_ARG(phase != 6 || b->bp.method == KCFD_METHOD_BICGSTAB, "phase 6 needs BiCGSTAB");
    KCFD_ARG(phase != 5 || (b->run_two && b->bp.method == KCFD_METHOD_BICGSTAB),
             "phase 5 needs a two-colour ILU(0) member and BiCGSTAB");
    KCFD_ARG(phase != 3 || b->bp.method == KCFD_METHOD_CG || b->run_two || b->run_polled,
             "phase 3 of BiCGSTAB needs a two-colour or sync-free ILU(0) member (phase 6 times the whole apply)");
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    KCFD_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    KCFD_HIP(hipStreamSynchronize(s));
    // Full iterations, launched directly (not from the graph) so that events
    // can bracket the one kernel: it runs between the same kernels as in a
    // solve, with the caches in the state they leave.  Every system is put
    // back into S_RUN with its saved scalars before each iteration (a
    // converged system would otherwise skip the work); the vectors are
    // scratch afterwards and the states are restored at the end.
    std::vector<SysState> saved(b->nsys), run(b->nsys);
    KCFD_HIP(hipMemcpyAsync(saved.data(), b->d_state, sizeof(SysState) * b->nsys, hipMemcpyDeviceToHost, s));
    KCFD_HIP(hipStreamSynchronize(s));
    // No test may end the iteration early: convergence, half-step and
    // breakdown thresholds are disabled, so every kernel does its full work.
    for (int k = 0; k < b->nsys; ++k) {
        run[k] = saved[k];
        if (!b->mats[k]) continue;
        run[k].status = S_RUN;
        if (run[k].omega == 0.0) run[k].omega = 1.0;
        if (run[k].rho == 0.0) run[k].rho = 1.0;
        if (run[k].rho_prev == 0.0) run[k].rho_prev = 1.0;   // CG: beta = rho / rho_prev
        run[k].atol_in = run[k].atol_out = -1.0;
        run[k].maxiter = 1 << 30;
        run[k].iters = run[k].total_iters = 0;
        run[k].rmax = INFINITY;   // MIXED: no divergence restart
    }
    const LaunchFn it = pick_iter(b->bp.precision, b->bp.precond, b->bp.method, b->rows);
    BatchParams pp = b->bp;
    pp.rhotol = -1.0;     // no breakdown exit
    pp.delta = 0.0;       // MIXED: no residual replacement
    pp.inner_maxiter = 1 << 30;
    KCFD_HIP(hipMemcpyAsync(b->d_params, &pp, sizeof(BatchParams), hipMemcpyHostToDevice, s));
    double tot = 0.0;
    int rc = KCFD_OK;
    b->timed_phase = phase;
    for (int r = 0; r <= reps && rc == KCFD_OK; ++r) {   // r == 0 warms up
        // The state upload is ordered on the solver stream and the whole
        // iteration is waited for: a status changed under a running kernel
        // could make a triangular solve skip rows that other rows poll.
        hipError_t e = hipMemcpyAsync(b->d_state, run.data(), sizeof(SysState) * b->nsys,
                                      hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            it(b, s);
            e = hipStreamSynchronize(s);
        }
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, b->ctx->ev0, b->ctx->ev1);
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess) {
            set_error("kcfd_batch_profile: %s", hipGetErrorString(e));
            rc = KCFD_EHIP;
        }
        if (r > 0) tot += ms;
    }
    b->timed_phase = 0;
    KCFD_CHECK(rc);
    *avg_ms = tot / reps;
    KCFD_HIP(hipMemcpyAsync(b->d_params, &b->bp, sizeof(BatchParams), hipMemcpyHostToDevice, s));
    KCFD_HIP(hipMemcpyAsync(b->d_state, saved.data(), sizeof(SysState) * b->nsys, hipMemcpyHostToDevice, s));
    KCFD_HIP(hipStreamSynchronize(s));
    return KCFD_OK;
}

int kcfd_batch_kernel_bytes(kcfd_batch *b, int phase, double *bytes) {
    KCFD_ARG(b && bytes, "bad argument");
    const double sv = b->bp.precision == KCFD_FP64 ? 8.0 : 4.0;
    double tot = 0.0;
    for (size_t k = 0; k < b->mats.size(); ++k) {
        const kcfd_mat *m = b->mats[k];
        if (!m) continue;
        const double n = m->n, nnz = m->nnz;
        const double MA = nnz * (sv + 4) + (n + 1) * 4;
        // folded product: the product kernel's colour-0 rows read v (t) and
        // rt (s) only, the colour-0 half of the apply writes v (t) and, in
        // MIXED, reads the FP64 pivots
        const double n0 = b->h_desc[k].nid, nnz0 = n0 > 0 ? (double)m->nnz_colour0 : 0.0;
        const double fold = -(nnz0 * (sv + 4) + n0 * 4);
        const double x64 = b->bp.precision == KCFD_MIXED && n0 > 0 ? 1.0 : 0.0;
        const double MFh = 0.5 * nnz * (sv + 4) + (n + 1) * 4;  // one triangle + its level list
        const bool cg = b->bp.method == KCFD_METHOD_CG;
        if (cg && phase == 1) tot += MA + 2 * n * sv;          // q = A p: p gathered, q written
        else if (cg && phase == 2) tot += 6 * n * sv;          // x, p, r, q read; x, r written
        else if (cg && phase == 4) tot += (b->bp.precond == KCFD_PC_JACOBI ? 4 : 3) * n * sv;
        else if (phase == 1) tot += MA + (b->bp.precond == KCFD_PC_JACOBI ? 4 : 3) * n * sv + fold;
        else if (phase == 2) tot += 8 * n * sv + (b->bp.precision == KCFD_MIXED ? 8 * n : 0);  // FP64 x
        else if (phase == 6) {   // the whole apply: both halves, both sync-free solves, or the
                                 // three segment launches
            if (m->two_colour)
                tot += (double)(m->nent_l + m->nent_u) * (sv + 4) + 4 * n * sv + n0 * sv + x64 * 8 * n0;
            else if (m->seg2)    // entries, the in-segment values, z, w, u_ii, y of colour 0 (written, read)
                tot += seg2_apply_bytes(m, sv);
            else
                tot += 2 * (MFh + 2 * n * sv);
        }
        else if (phase == 3 || phase == 5) {
            // two-colour apply, colour-1 launch (3): its L entries, own z
            // read, w written, u_ii, the colour-0 operands gathered (once);
            // colour-0 launch (5): its U entries, own z, w, u_ii, the colour-1
            // results gathered, and folded: the product rows written and, in
            // MIXED, the FP64 pivots read
            const double n1 = n - m->ncolour0, nc0 = m->ncolour0;
            if (m->two_colour && cg)   // CG times both launches of its apply as phase 3
                tot += (double)(m->nent_l + m->nent_u) * (sv + 4) + 4 * n * sv;
            else if (m->seg2 && cg)
                tot += seg2_apply_bytes(m, sv);
            else if (m->two_colour && phase == 3)
                tot += (double)m->nent_l * (sv + 4) + (3 * n1 + nc0) * sv;
            else if (m->two_colour)
                tot += (double)m->nent_u * (sv + 4) + (3 * nc0 + n1) * sv + n0 * sv + x64 * 8 * n0;
            else if (phase == 3)   // the sync-free L solve (one triangle + its level list)
                tot += MFh + 2 * n * sv;
        }
        else if (phase == 4) tot += 4 * n * sv;
    }
    *bytes = tot;
    return KCFD_OK;
}

int kcfd_solve(kcfd_mat *m, const double *bvec, double *x, const kcfd_solve_opts *opts,
               kcfd_solve_result *res) {
    KCFD_ARG(m && opts, "null argument");
    KCFD_ARG(m->n == 0 || (bvec && x), "null vector");
    std::lock_guard<std::recursive_mutex> g(m->ctx->mu);
    KCFD_HIP(hipSetDevice(m->ctx->device));
    if (m->n == 0) {
        if (res) std::memset(res, 0, sizeof(*res));
        return KCFD_OK;
    }
    KCFD_CHECK(check_current(m, "kcfd_solve"));
    if (opts->method == KCFD_METHOD_GMRES) return gmres_solve(m, bvec, x, opts, res);
    // the cached single-system workspace is reused while the preconditioner,
    // precision, launch shape and the b / x buffers are unchanged
    bool reuse = m->solver != nullptr && m->solver->bp.precond == opts->precond &&
                 m->solver->bp.precision == opts->precision && m->solver->h_desc[0].b_user == bvec &&
                 m->solver->h_desc[0].x_user == x && m->solver->opts.grid_per_cu == opts->grid_per_cu &&
                 m->solver->opts.trsv_per_cu == opts->trsv_per_cu &&
                 m->solver->opts.tile_walk == opts->tile_walk && m->solver->opts.method == opts->method &&
                 m->solver->opts.rows_per_thread == opts->rows_per_thread &&
                 m->solver->opts.fold_colour0 == opts->fold_colour0;
    if (!reuse) {
        if (m->solver) kcfd_batch_destroy(m->solver);
        m->solver = nullptr;
        kcfd_mat *mm[1] = {m};
        const double *bb[1] = {bvec};
        double *xx[1] = {x};
        KCFD_CHECK(build_batch(m->ctx, 1, mm, bb, xx, opts, &m->solver));
    } else {
        kcfd_batch *b = m->solver;
        b->opts = *opts;
        if (b->opts.check_every <= 0) b->opts.check_every = 8;
        b->bp.rtol = opts->rtol;
        b->bp.atol = opts->atol;
        b->bp.delta = opts->inner_rtol > 0 ? opts->inner_rtol : 0.1;
        b->bp.max_refine = opts->max_refine > 0 ? opts->max_refine : 1000;
        // the parameters live in device memory (uploaded by run_batch), so the
        // captured graph is reused as is
    }
    return run_batch(m->solver, res);
}

int kcfd_solve_noalloc(kcfd_mat *m, const double *bvec, double *x, const kcfd_solve_opts *opts,
                       kcfd_solve_result *res) {
    KCFD_ARG(m && opts, "null argument");
    KCFD_ARG(m->n == 0 || (bvec && x), "null vector");
    KCFD_CHECK(check_opts(opts));
    KCFD_ARG(opts->method == KCFD_METHOD_BICGSTAB || opts->method == KCFD_METHOD_CG,
             "kcfd_solve_noalloc runs BiCGSTAB and CG");
    KCFD_ARG(m->ctx->device >= 0 && m->ctx->device < 64, "device index beyond the scratch table");
    std::lock_guard<std::recursive_mutex> g(m->ctx->mu);
    std::lock_guard<std::mutex> gd(g_noalloc_mu[m->ctx->device]);
    KCFD_HIP(hipSetDevice(m->ctx->device));
    if (m->n == 0) {
        if (res) std::memset(res, 0, sizeof(*res));
        return KCFD_OK;
    }
    KCFD_CHECK(check_current(m, "kcfd_solve_noalloc"));
    return solve_noalloc(m, bvec, x, opts, res);
}

}  // extern "C"
