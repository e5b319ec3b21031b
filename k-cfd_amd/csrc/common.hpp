// Internal definitions shared by the libkcfd translation units (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/kcfd.h"

namespace kcfd {

void set_error(const char *fmt, ...);

constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr int kMaxPartials = 4;      // doubles per block partial record
constexpr int kFinBlock = 1024;      // threads of a finalize workgroup (one per system)

// Number of XCDs on MI355X; workgroups are dispatched round-robin over them.
constexpr int kXcds = 8;
// rows per segment of the KCFD_ORDER_SEGMENT ILU(0) ordering (at most)
#ifndef KCFD_SEG_LEN
#define KCFD_SEG_LEN 4
#endif
constexpr int kSegLen = KCFD_SEG_LEN;
// entries outside its segment in one row of a segment-ordered factor (at
// most; a structure with more takes the level layout)
constexpr int kSegW = 6;

// narrow-index SELL copy: the offset that marks column = row, and the widest
// span of off-diagonal columns one 64-row slice may have
constexpr unsigned kSellDiag = 0xFFFFu;
constexpr int kSellSpan = 0xFFFE;

// Global-memory views of pointers.  Pointers read from a descriptor in
// memory are generic to the compiler, which then emits FLAT loads that wait
// on both the vector-memory and LDS counters; cast to the global address
// space, the loads are GLOBAL and can be kept in flight together.
#define KCFD_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const KCFD_GLOBAL T *gl(const T *p) {
    return (const KCFD_GLOBAL T *)p;
}
template <typename T>
__device__ __forceinline__ KCFD_GLOBAL T *glw(T *p) {
    return (KCFD_GLOBAL T *)p;
}
// clang's own 2 x int vector: loadable through a global-space pointer (HIP's
// int2 wrapper type cannot be copied out of one)
typedef int kint2 __attribute__((ext_vector_type(2)));
typedef int kint4 __attribute__((ext_vector_type(4)));

// a value the caller knows is the same in every lane of the wave, as such
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Remap a dispatch-order block id so that each XCD works on one contiguous
// range of tiles (neighbouring rows share that XCD's L2).  grid must be a
// multiple of kXcds; ids >= ntiles are idle.
__device__ __forceinline__ int xcd_tile(int bid, int grid) {
    int per = grid / kXcds;
    return (bid % kXcds) * per + bid / kXcds;
}

inline int xcd_grid(int ntiles) { return ((ntiles + kXcds - 1) / kXcds) * kXcds; }

// wave64 sum by xor-shuffle (fixed tree: deterministic)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block sum of NV values over NT threads; result valid in thread 0.  Fixed
// order.
template <int NV, int NT = kBlock>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    constexpr int kW = NT / 64;
    __shared__ double red[kW][NV > 0 ? NV : 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) red[w][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = red[0][k];
            for (int q = 1; q < kW; ++q) s += red[q][k];
            v[k] = s;
        }
    }
    __syncthreads();
}

}  // namespace kcfd

#define KCFD_HIP(expr)                                                                 \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            kcfd::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),      \
                            __FILE__, __LINE__);                                        \
            return e_ == hipErrorOutOfMemory ? KCFD_EOOM : KCFD_EHIP;                  \
        }                                                                              \
    } while (0)

#define KCFD_ARG(cond, msg)                                                            \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            kcfd::set_error("invalid argument: %s", msg);                              \
            return KCFD_EARG;                                                          \
        }                                                                              \
    } while (0)

#define KCFD_CHECK(call)              \
    do {                              \
        int rc_ = (call);             \
        if (rc_ != KCFD_OK) return rc_; \
    } while (0)

struct kcfd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::recursive_mutex mu;
    int cus = 0;
    double last_ms = 0.0;
    // staging scratch for small host<->device copies
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // device address of kcfd_solve_noalloc's static workspace array on this
    // context's device (hipGetSymbolAddress at kcfd_init)
    void *noalloc = nullptr;
};

struct kcfd_batch;
namespace kcfd {
struct GmresWs;
}

struct kcfd_mat {
    kcfd_ctx *ctx = nullptr;
    int n = 0, nnz = 0;
    int *indptr = nullptr, *indices = nullptr, *diagpos = nullptr;
    double *val = nullptr, *rhs = nullptr, *diagt = nullptr;
    std::vector<int> h_indptr, h_indices, h_diagpos;
    // SELL-64 copy of A used by every solver SpMV: slice s holds rows
    // 64s..64s+63, column-major, width = longest row of the slice; a row's
    // entries keep their CSR order and padding (column = row, value 0) comes
    // last, so its sum equals the sequential CSR row sum.
    int nslices = 0, sell_nnz = 0;
    bool sell_permuted = false;   // rows and columns in the ILU(0) order (perm)
    std::vector<int> h_slice_ptr;
    int *slice_ptr = nullptr, *six = nullptr;
    double *sa64 = nullptr;
    int2 *spk32 = nullptr;   // MIXED: {column, FP32 value bits} per SELL entry (one 8-byte load)
    // MIXED, narrow-index form (used instead of spk32 when every slice's
    // off-diagonal columns span at most kSellSpan): FP32 values, 16-bit
    // column offsets from the slice's smallest off-diagonal column, and that
    // base per slice; the diagonal and padding (column = row) as kSellDiag.
    // 6 bytes per entry instead of 8.
    float *sv32 = nullptr;
    unsigned short *so16 = nullptr;
    int *sbase = nullptr;
    // preconditioner
    int pc_kind = KCFD_PC_NONE;
    bool have32 = false;
    double *dinv = nullptr;
    float *dinv32 = nullptr;
    // ILU(0) ordering.  The factors are those of P A P^T; perm[k] is the row
    // of A at position k.  The factor structure f_* is P A P^T's canonical
    // CSR (aliases A's arrays in the natural ordering, where perm is null);
    // f_vmap[q] is the slot of A feeding factor slot q and f_uix[q] the
    // column as a row of A (what the U solve, writing in A's numbering, polls).
    int ilu_order = KCFD_ORDER_NATURAL, ilu_built_order = -1, ncolors = 0;
    int ncolour0 = 0;   // rows of colour 0 (they come first in P A P^T)
    int *perm = nullptr, *iperm = nullptr, *f_vmap = nullptr;
    int *f_indptr = nullptr, *f_indices = nullptr, *f_diagpos = nullptr, *f_uix = nullptr;
    std::vector<int> h_perm, hf_indptr, hf_indices, hf_diagpos;
    // ILU(0): factors on the (permuted) pattern, level schedule of the solves
    double *lu = nullptr;
    float *lu32 = nullptr;
    int nlev_l = 0, nlev_u = 0;
    int *lvl_rows_l = nullptr, *lvl_rows_u = nullptr;   // rows ordered by level
    std::vector<int> h_lvl_ptr_l, h_lvl_ptr_u;           // level offsets
    int max_lvl_rows_l = 0, max_lvl_rows_u = 0;
    // ILU(0) solve layout.  The level lists are cut into chunks of up to 64
    // rows (never across a level); chunk c keeps the L (resp. U) entries of
    // its rows column-major: entry k of the row in lane l at off[c] + 64k + l,
    // in CSR order, padded with column -1 / value 0.  Columns are rows of
    // P A P^T (the solver's row order).  h_ch*_ptr[l]: first chunk of level l;
    // ch*_map: factor slot of each entry (-1: padding); ud: u_ii by row of
    // P A P^T.
    int nch_l = 0, nch_u = 0, nent_l = 0, nent_u = 0;
    // Two colours: the L layout holds the colour-1 rows (positions
    // ncolour0..n-1) and the U layout the colour-0 rows, as single runs of
    // 64-row chunks; a colour-0 row's y is z itself.
    bool two_colour = false;
    long long nnz_colour0 = 0;   // entries of the colour-0 rows (diagonal included)
    // MIXED, two colours: FP64 pivots of the colour-0 rows for the folded
    // BiCGSTAB product (kcfd_solve_opts.fold_colour0): a_ii + sum_j (a_ij -
    // fl32(a_ij)), so that the row [pivot, fl32(a_ij)] has A's row sum
    double *c0piv = nullptr;
    // Segment ordering with two colours (KCFD_ORDER_SEGMENT): the solves
    // walk segments, one lane per segment.  A segment group is up to 64
    // consecutive segments of one colour (colour 0 first), longest first,
    // its rows numbered step-major: row k of lane t is h_sg[g].x + the
    // step's row offset, h_sg[g] = {row at step 0, length}.  For each step
    // the chunk layout above holds the rows' entries outside their segment
    // -- colour-1 rows: their L entries (colour-0 columns) in the L layout;
    // colour-0 rows: their U entries (colour-1 columns) in the U layout --
    // with entry e of lane l at off + 64e + l.  h_sg_steps[t] = {entry
    // offset, width, row offset, 0} of step t; a group's steps are
    // consecutive.  sgl / sgu map each row's in-segment L entry (the row
    // before) / U entry (the row after) to its factor slot (-1: none).
    bool seg2 = false;
    std::vector<int2> h_sg;
    std::vector<int4> h_sg_steps;
    // groups of colour 0 and colour 1, 4 ints each: first segment, segments,
    // longest segment, first step
    std::vector<int> h_sgg[2];
    int2 *sg = nullptr;
    int4 *sg_steps = nullptr;
    int *sgl_map = nullptr, *sgu_map = nullptr;
    double *sgl64 = nullptr, *sgu64 = nullptr;
    float *sgl32 = nullptr, *sgu32 = nullptr;
    std::vector<int> h_chl_ptr, h_chu_ptr;
    // host copies of the per-chunk entry offsets and (narrow-index form)
    // base columns, which the batch's chunk tables carry
    std::vector<int> h_chl_off, h_chu_off, h_chl_base, h_chu_base;
    int *chl_off = nullptr, *chu_off = nullptr, *chl_col = nullptr, *chu_col = nullptr;
    int *chl_map = nullptr, *chu_map = nullptr;
    double *chl_v64 = nullptr, *chu_v64 = nullptr, *ud64 = nullptr;
    // MIXED: {column, FP32 value bits} per chunk entry (one 8-byte load)
    int2 *chl_pk32 = nullptr, *chu_pk32 = nullptr;
    float *ud32 = nullptr;
    // MIXED, narrow-index form of the two-colour layouts (instead of the
    // packed pairs when every chunk's columns span at most kSellSpan):
    // 16-bit column offsets from each chunk's smallest column (structure,
    // padding as kSellDiag), the bases, and FP32 values
    bool ch16 = false;
    unsigned short *chl_o16 = nullptr, *chu_o16 = nullptr;
    int *chl_base = nullptr, *chu_base = nullptr;
    float *chl_v32 = nullptr, *chu_v32 = nullptr;
    // Values generation: bumped by every entry point that writes the values
    // (kcfd_mat_set_values, the assembly calls); kcfd_precond_setup records
    // it.  The solvers read the SELL copy and the preconditioner that the
    // set-up built, so they refuse a matrix whose values changed since.
    unsigned long long val_gen = 1, setup_gen = 0;
    // cached single-system solver workspaces (BiCGSTAB / CG batch, GMRES)
    kcfd_batch *solver = nullptr;
    kcfd::GmresWs *gmres = nullptr;
    int solver_key[4] = {-1, -1, -1, -1};
};

namespace kcfd {
// KCFD_ESTATE unless the preconditioner (and the SELL copy) was set up from
// the current values; `what` names the matrix in the message
int check_current(const kcfd_mat *m, const char *what);
int noalloc_resolve(kcfd_ctx *ctx);
int ensure_scratch(kcfd_ctx *ctx, size_t bytes);
int ilu0_structure(kcfd_mat *m);
int ilu0_setup(kcfd_mat *m, bool want32);
int ensure_sell(kcfd_mat *m, bool permuted);
void ilu_free_order(kcfd_mat *m);
int precond_apply_async(kcfd_mat *m, const double *x, double *y, double *t, bool porder,
                        const int *skip, int col);
int gmres_solve(kcfd_mat *m, const double *b, double *x, const kcfd_solve_opts *o,
                kcfd_solve_result *res);
void gmres_free(kcfd_mat *m);
size_t gmres_bytes(const kcfd_mat *m);   // device memory of the cached GMRES workspace
int mat_free_precond(kcfd_mat *m);
int mat_free_ilu_structure(kcfd_mat *m);
}  // namespace kcfd
