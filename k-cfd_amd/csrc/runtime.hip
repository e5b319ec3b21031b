// Context, error reporting, device buffers, CSR matrices and the Jacobi
// preconditioner set-up.
#include "common.hpp"

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>

namespace kcfd {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int check_current(const kcfd_mat *m, const char *what) {
    if (m->setup_gen != m->val_gen) {
        set_error("%s: the values changed after the last kcfd_precond_setup (or it was never called); "
                  "set the preconditioner up again before solving", what);
        return KCFD_ESTATE;
    }
    return KCFD_OK;
}

int ensure_scratch(kcfd_ctx *ctx, size_t bytes) {
    if (ctx->scratch_bytes >= bytes) return KCFD_OK;
    if (ctx->scratch) KCFD_HIP(hipFree(ctx->scratch));
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
    KCFD_HIP(hipMalloc(&ctx->scratch, bytes));
    ctx->scratch_bytes = bytes;
    return KCFD_OK;
}

// ---------------------------------------------------------------- kernels

__global__ void k_fill(double *p, double v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = v;
}

// Row-parallel CSR SpMV in CSR order, products and sums unfused (the
// sequential row sum of scipy's csr_matvec).
__global__ void k_spmv(int n, const int *__restrict__ ip, const int *__restrict__ ix,
                       const double *__restrict__ a, const double *__restrict__ x,
                       double *__restrict__ y) {
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int q = ip[i]; q < ip[i + 1]; ++q) s += a[q] * x[ix[q]];
    y[i] = s;
}

__global__ void k_diagpos(int n, const int *__restrict__ ip, const int *__restrict__ ix,
                          int *__restrict__ dp) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int p = -1;
    for (int q = ip[i]; q < ip[i + 1]; ++q)
        if (ix[q] == i) { p = q; break; }
    dp[i] = p;
}

// Jacobi: dinv = where(|a_ii| > 1e-30, 1/a_ii, 0) (core/preconditioner.py:43-47);
// a missing diagonal counts as 0.
__global__ void k_jacobi(int n, const int *__restrict__ dp, const double *__restrict__ a,
                         double *__restrict__ dinv, float *__restrict__ dinv32) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = dp[i] >= 0 ? a[dp[i]] : 0.0;
    double r = fabs(d) > 1e-30 ? 1.0 / d : 0.0;
    dinv[i] = r;
    if (dinv32) dinv32[i] = (float)r;
}

// SELL-64 column indices (structure) and values of one row per thread.
// Row k of the SELL is row rperm[k] of A (rperm null: k) with A's entries in
// A's CSR order and columns renumbered by icol (null: unchanged); the
// solver runs in that row order (P A P^T for a permuted ILU(0)), and every
// row sum keeps the sequential CSR order.
__global__ void k_sell_struct(int n, const int *__restrict__ ip, const int *__restrict__ ix,
                              const int *__restrict__ sptr, int *__restrict__ six,
                              const int *__restrict__ rperm, const int *__restrict__ icol) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int r = rperm ? rperm[k] : k;
    const int s = k >> 6, base = sptr[s] + (k & 63), w = (sptr[s + 1] - sptr[s]) >> 6;
    const int q0 = ip[r], len = ip[r + 1] - q0;
    for (int j = 0; j < w; ++j) {
        const int c = j < len ? ix[q0 + j] : -1;
        six[base + 64 * j] = c < 0 ? k : (icol ? icol[c] : c);
    }
}

// SELL values (row order as k_sell_struct).  The FP32 copy is kept in
// difference form for the mixed solver: the diagonal position holds the row
// sum sum_j a_ij (summed in FP64), so that
// (A z)_i = sum_{j!=i} a_ij (z_j - z_i) + rowsum_i z_i.  Rounding the
// off-diagonals to FP32 then leaves the row sums (and the near-null space of
// a diffusion operator) intact, where rounding a_ii separately would perturb
// the smallest eigenvalues by eps_32 |a_ii|.
// With so16 the FP32 copy goes to sv32 / so16 (narrow-index form, the base
// of each slice from k_sell_base) instead of spk32.
__global__ void k_sell_values(int n, const int *__restrict__ ip, const int *__restrict__ ix,
                              const int *__restrict__ sptr, const int *__restrict__ six,
                              const double *__restrict__ a, double *__restrict__ sa64,
                              int2 *__restrict__ spk32, float *__restrict__ sv32,
                              unsigned short *__restrict__ so16, const int *__restrict__ sbase,
                              const int *__restrict__ rperm) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int r = rperm ? rperm[k] : k;
    const int s = k >> 6, base = sptr[s] + (k & 63), w = (sptr[s + 1] - sptr[s]) >> 6;
    const int q0 = ip[r], len = ip[r + 1] - q0;
    double rowsum = 0.0;
    for (int j = 0; j < len; ++j) rowsum += a[q0 + j];
    for (int j = 0; j < w; ++j) {
        const double v = j < len ? a[q0 + j] : 0.0;
        sa64[base + 64 * j] = v;
        const float v32 = (float)(j < len && ix[q0 + j] == r ? rowsum : v);
        const int c = six[base + 64 * j];
        if (so16) {
            sv32[base + 64 * j] = v32;
            so16[base + 64 * j] = (unsigned short)(c == k ? kSellDiag : (unsigned)(c - sbase[s]));
        } else if (spk32) {
            spk32[base + 64 * j] = make_int2(c, __float_as_int(v32));
        }
    }
}

// Base of each slice for the narrow-index SELL copy: the smallest column
// other than the row itself over the slice's entries (0 when there is none),
// and *wide = 1 when a slice's such columns span more than kSellSpan.  One
// wave per slice (kBlock is a multiple of 64, so a wave's lanes are the rows
// of one slice); every lane takes part in the shuffles.
__global__ __launch_bounds__(kBlock) void k_sell_base(int n, const int *__restrict__ sptr,
                                                      const int *__restrict__ six, int *__restrict__ sbase,
                                                      int *__restrict__ wide) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    int lo = INT_MAX, hi = INT_MIN;
    if (k < n) {
        const int s = k >> 6, base = sptr[s] + (k & 63), w = (sptr[s + 1] - sptr[s]) >> 6;
        for (int j = 0; j < w; ++j) {
            const int c = six[base + 64 * j];
            if (c != k) {
                lo = min(lo, c);
                hi = max(hi, c);
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o));
        hi = max(hi, __shfl_xor(hi, o));
    }
    if (k < n && (k & 63) == 0) {
        sbase[k >> 6] = lo == INT_MAX ? 0 : lo;
        if (lo != INT_MAX && hi - lo > kSellSpan) *wide = 1;
    }
}

// Bandwidth ceilings.  Each workgroup streams one contiguous chunk of the
// array (whole 4 KB rows of 256 x 16-byte vectors), four rows in flight per
// thread, with non-temporal loads and stores (the data is touched once, so
// it should not displace anything in the caches).  copy: dst = src; read:
// every vector is loaded and folded into one value per thread.
constexpr int kCeilUnroll = 4;

__global__ __launch_bounds__(kBlock) void k_stream_copy(size_t n16, const double2 *__restrict__ src,
                                                        double2 *__restrict__ dst) {
    const size_t row = (size_t)kBlock * kCeilUnroll;
    const size_t per = (n16 + gridDim.x - 1) / gridDim.x;
    const size_t chunk = (per + row - 1) / row * row;
    const size_t b0 = (size_t)blockIdx.x * chunk, b1 = min(b0 + chunk, n16);
    for (size_t base = b0; base < b1; base += row) {
        double2 v[kCeilUnroll];
#pragma unroll
        for (int u = 0; u < kCeilUnroll; ++u) {
            const size_t i = base + (size_t)u * kBlock + threadIdx.x;
            if (i < b1) {
                v[u].x = __builtin_nontemporal_load(&src[i].x);
                v[u].y = __builtin_nontemporal_load(&src[i].y);
            }
        }
#pragma unroll
        for (int u = 0; u < kCeilUnroll; ++u) {
            const size_t i = base + (size_t)u * kBlock + threadIdx.x;
            if (i < b1) {
                __builtin_nontemporal_store(v[u].x, &dst[i].x);
                __builtin_nontemporal_store(v[u].y, &dst[i].y);
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_stream_read(size_t n16, const double2 *__restrict__ src,
                                                        double *__restrict__ out) {
    const size_t row = (size_t)kBlock * kCeilUnroll;
    const size_t per = (n16 + gridDim.x - 1) / gridDim.x;
    const size_t chunk = (per + row - 1) / row * row;
    const size_t b0 = (size_t)blockIdx.x * chunk, b1 = min(b0 + chunk, n16);
    double acc = 0.0;
    for (size_t base = b0; base < b1; base += row) {
        double2 v[kCeilUnroll];
#pragma unroll
        for (int u = 0; u < kCeilUnroll; ++u) {
            const size_t i = base + (size_t)u * kBlock + threadIdx.x;
            v[u] = make_double2(0.0, 0.0);
            if (i < b1) {
                v[u].x = __builtin_nontemporal_load(&src[i].x);
                v[u].y = __builtin_nontemporal_load(&src[i].y);
            }
        }
#pragma unroll
        for (int u = 0; u < kCeilUnroll; ++u) acc += v[u].x + v[u].y;
    }
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

// SELL-64 structure of A in the solver's row order: natural, or P A P^T
// (rows and columns renumbered by m->perm) when `permuted`.  Rebuilt only
// when the order changes; a rebuild frees the value buffers (and any cached
// solver, which points at them).
int ensure_sell(kcfd_mat *m, bool permuted) {
    if (m->six && m->sell_permuted == permuted) return KCFD_OK;
    const int n = m->n;
    hipStream_t st = m->ctx->stream;
    KCFD_HIP(hipStreamSynchronize(st));
    if (m->solver) {
        kcfd_batch_destroy(m->solver);
        m->solver = nullptr;
    }
    hipFree(m->slice_ptr); m->slice_ptr = nullptr;
    hipFree(m->six); m->six = nullptr;
    hipFree(m->sa64); m->sa64 = nullptr;
    hipFree(m->spk32); m->spk32 = nullptr;
    hipFree(m->sv32); m->sv32 = nullptr;
    hipFree(m->so16); m->so16 = nullptr;
    hipFree(m->sbase); m->sbase = nullptr;
    const std::vector<int> &ip = m->h_indptr;
    m->nslices = (n + 63) / 64;
    m->h_slice_ptr.assign(m->nslices + 1, 0);
    long long acc = 0;
    for (int sl = 0; sl < m->nslices; ++sl) {
        int w = 0;
        for (int k = sl * 64; k < std::min(n, sl * 64 + 64); ++k) {
            const int r = permuted ? m->h_perm[k] : k;
            w = std::max(w, ip[r + 1] - ip[r]);
        }
        acc += 64LL * w;
        if (acc >= (1LL << 31)) {
            set_error("matrix too large for 32-bit SELL offsets");
            return KCFD_EARG;
        }
        m->h_slice_ptr[sl + 1] = (int)acc;
    }
    m->sell_nnz = (int)acc;
    if (hipMalloc(&m->slice_ptr, sizeof(int) * (m->nslices + 1)) != hipSuccess ||
        hipMalloc(&m->six, sizeof(int) * std::max(m->sell_nnz, 1)) != hipSuccess ||
        hipMalloc(&m->sa64, sizeof(double) * std::max(m->sell_nnz, 1)) != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipMalloc failed for the SELL copy (n=%d, padded nnz=%d)", n, m->sell_nnz);
        return KCFD_EOOM;
    }
    KCFD_HIP(hipMemcpyAsync(m->slice_ptr, m->h_slice_ptr.data(), sizeof(int) * (m->nslices + 1),
                            hipMemcpyHostToDevice, st));
    if (n)
        hipLaunchKernelGGL(k_sell_struct, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, n,
                           m->indptr, m->indices, m->slice_ptr, m->six, permuted ? m->perm : nullptr,
                           permuted ? m->iperm : nullptr);
    KCFD_HIP(hipGetLastError());
    KCFD_HIP(hipStreamSynchronize(st));   // h_slice_ptr is read by the copy
    m->sell_permuted = permuted;
    return KCFD_OK;
}

// Buffers of the FP32 SELL copy: the narrow-index form when every slice's
// off-diagonal columns fit 16-bit offsets (KCFD_SELL16=0 forces the packed
// {column, value} form), else spk32.
int sell_mixed_alloc(kcfd_mat *m) {
    hipStream_t st = m->ctx->stream;
    const char *env = std::getenv("KCFD_SELL16");
    bool narrow = !(env && env[0] == '0') && m->n > 0;
    if (narrow) {
        int *wide_d = nullptr;
        int wide = 0;
        KCFD_HIP(hipMalloc(&m->sbase, sizeof(int) * m->nslices));
        KCFD_HIP(hipMalloc(&wide_d, sizeof(int)));
        // every step runs and the flag buffer is freed before an error returns
        hipError_t e = hipMemcpyAsync(wide_d, &wide, sizeof(int), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_sell_base, dim3((m->n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, m->n,
                               m->slice_ptr, m->six, m->sbase, wide_d);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(&wide, wide_d, sizeof(int), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        hipFree(wide_d);
        KCFD_HIP(e);
        narrow = wide == 0;
        if (!narrow) {
            hipFree(m->sbase);
            m->sbase = nullptr;
        }
    }
    if (narrow) {
        KCFD_HIP(hipMalloc(&m->sv32, sizeof(float) * std::max(m->sell_nnz, 1)));
        KCFD_HIP(hipMalloc(&m->so16, sizeof(unsigned short) * std::max(m->sell_nnz, 1)));
    } else {
        KCFD_HIP(hipMalloc(&m->spk32, sizeof(int2) * std::max(m->sell_nnz, 1)));
    }
    return KCFD_OK;
}

// Value buffers of the preconditioner.  The ILU(0) ordering and level
// schedule depend on the structure only and survive (mat_free_ilu_structure).
int mat_free_precond(kcfd_mat *m) {
    hipFree(m->spk32); m->spk32 = nullptr;
    hipFree(m->sv32); m->sv32 = nullptr;
    hipFree(m->so16); m->so16 = nullptr;
    hipFree(m->sbase); m->sbase = nullptr;
    hipFree(m->dinv); m->dinv = nullptr;
    hipFree(m->dinv32); m->dinv32 = nullptr;
    hipFree(m->lu); m->lu = nullptr;
    hipFree(m->lu32); m->lu32 = nullptr;
    hipFree(m->chl_v64); m->chl_v64 = nullptr;
    hipFree(m->chu_v64); m->chu_v64 = nullptr;
    hipFree(m->ud64); m->ud64 = nullptr;
    hipFree(m->chl_pk32); m->chl_pk32 = nullptr;
    hipFree(m->chu_pk32); m->chu_pk32 = nullptr;
    hipFree(m->chl_v32); m->chl_v32 = nullptr;
    hipFree(m->chu_v32); m->chu_v32 = nullptr;
    hipFree(m->ud32); m->ud32 = nullptr;
    hipFree(m->c0piv); m->c0piv = nullptr;
    hipFree(m->sgl64); m->sgl64 = nullptr;
    hipFree(m->sgu64); m->sgu64 = nullptr;
    hipFree(m->sgl32); m->sgl32 = nullptr;
    hipFree(m->sgu32); m->sgu32 = nullptr;
    m->have32 = false;
    m->pc_kind = KCFD_PC_NONE;
    return KCFD_OK;
}

int mat_free_ilu_structure(kcfd_mat *m) {
    hipFree(m->lvl_rows_l); m->lvl_rows_l = nullptr;
    hipFree(m->lvl_rows_u); m->lvl_rows_u = nullptr;
    int **arrs[] = {&m->chl_off, &m->chu_off, &m->chl_col, &m->chu_col, &m->chl_map, &m->chu_map};
    m->two_colour = false;
    for (int **a : arrs) {
        hipFree(*a);
        *a = nullptr;
    }
    m->h_chl_ptr.clear();
    m->h_chu_ptr.clear();
    m->h_chl_off.clear();
    m->h_chu_off.clear();
    m->h_chl_base.clear();
    m->h_chu_base.clear();
    m->nch_l = m->nch_u = m->nent_l = m->nent_u = 0;
    double **v64[] = {&m->chl_v64, &m->chu_v64, &m->ud64};
    for (double **a : v64) {
        hipFree(*a);
        *a = nullptr;
    }
    hipFree(m->chl_pk32); m->chl_pk32 = nullptr;
    hipFree(m->chu_pk32); m->chu_pk32 = nullptr;
    hipFree(m->chl_v32); m->chl_v32 = nullptr;
    hipFree(m->chu_v32); m->chu_v32 = nullptr;
    hipFree(m->ud32); m->ud32 = nullptr;
    hipFree(m->c0piv); m->c0piv = nullptr;
    m->nnz_colour0 = 0;
    hipFree(m->chl_o16); m->chl_o16 = nullptr;
    hipFree(m->chu_o16); m->chu_o16 = nullptr;
    hipFree(m->chl_base); m->chl_base = nullptr;
    hipFree(m->chu_base); m->chu_base = nullptr;
    m->ch16 = false;
    hipFree(m->sg); m->sg = nullptr;
    hipFree(m->sg_steps); m->sg_steps = nullptr;
    hipFree(m->sgl_map); m->sgl_map = nullptr;
    hipFree(m->sgu_map); m->sgu_map = nullptr;
    hipFree(m->sgl64); m->sgl64 = nullptr;
    hipFree(m->sgu64); m->sgu64 = nullptr;
    hipFree(m->sgl32); m->sgl32 = nullptr;
    hipFree(m->sgu32); m->sgu32 = nullptr;
    m->seg2 = false;
    m->h_sg.clear();
    m->h_sg_steps.clear();
    m->h_sgg[0].clear();
    m->h_sgg[1].clear();
    ilu_free_order(m);
    m->h_lvl_ptr_l.clear();
    m->h_lvl_ptr_u.clear();
    m->nlev_l = m->nlev_u = 0;
    return KCFD_OK;
}

}  // namespace kcfd

using namespace kcfd;

extern "C" int kcfd_batch_destroy(kcfd_batch *b);

static inline unsigned nblk(size_t n) {
    size_t b = (n + kBlock - 1) / kBlock;
    if (b == 0) b = 1;
    if (b > 65535u * 8u) b = 65535u * 8u;
    return (unsigned)b;
}

extern "C" {

const char *kcfd_last_error(void) { return g_err; }

int kcfd_init(int device, kcfd_ctx **out) {
    KCFD_ARG(out != nullptr, "out is null");
    int count = 0;
    KCFD_HIP(hipGetDeviceCount(&count));
    KCFD_ARG(device >= 0 && device < count, "device index out of range");
    KCFD_HIP(hipSetDevice(device));
    kcfd_ctx *c = new kcfd_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    KCFD_HIP(hipGetDeviceProperties(&prop, device));
    c->cus = prop.multiProcessorCount;
    KCFD_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    KCFD_HIP(hipEventCreate(&c->ev0));
    KCFD_HIP(hipEventCreate(&c->ev1));
    KCFD_CHECK(noalloc_resolve(c));
    *out = c;
    return KCFD_OK;
}

int kcfd_destroy(kcfd_ctx *ctx) {
    if (!ctx) return KCFD_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    if (ctx->scratch) hipFree(ctx->scratch);
    hipEventDestroy(ctx->ev0);
    hipEventDestroy(ctx->ev1);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return KCFD_OK;
}

int kcfd_device_info(kcfd_ctx *ctx, char *name, int len, int *cus, long long *total_bytes,
                     long long *free_bytes) {
    KCFD_ARG(ctx != nullptr, "ctx is null");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    KCFD_HIP(hipSetDevice(ctx->device));
    hipDeviceProp_t prop;
    KCFD_HIP(hipGetDeviceProperties(&prop, ctx->device));
    if (name && len > 0) {
        snprintf(name, (size_t)len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cus) *cus = prop.multiProcessorCount;
    size_t fr = 0, tot = 0;
    KCFD_HIP(hipMemGetInfo(&fr, &tot));
    if (total_bytes) *total_bytes = (long long)tot;
    if (free_bytes) *free_bytes = (long long)fr;
    return KCFD_OK;
}

int kcfd_sync(kcfd_ctx *ctx) {
    KCFD_ARG(ctx != nullptr, "ctx is null");
    KCFD_HIP(hipStreamSynchronize(ctx->stream));
    return KCFD_OK;
}

int kcfd_buf_alloc(kcfd_ctx *ctx, size_t bytes, void **dev) {
    KCFD_ARG(ctx && dev, "null argument");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    KCFD_HIP(hipSetDevice(ctx->device));
    *dev = nullptr;
    KCFD_HIP(hipMalloc(dev, bytes ? bytes : 8));
    return KCFD_OK;
}

int kcfd_buf_free(kcfd_ctx *ctx, void *dev) {
    KCFD_ARG(ctx != nullptr, "ctx is null");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (!dev) return KCFD_OK;
    KCFD_HIP(hipStreamSynchronize(ctx->stream));
    KCFD_HIP(hipFree(dev));
    return KCFD_OK;
}

int kcfd_buf_upload(kcfd_ctx *ctx, void *dev, const void *host, size_t bytes) {
    KCFD_ARG(ctx && (bytes == 0 || (dev && host)), "null argument");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (!bytes) return KCFD_OK;
    KCFD_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    KCFD_HIP(hipStreamSynchronize(ctx->stream));
    return KCFD_OK;
}

int kcfd_buf_download(kcfd_ctx *ctx, void *host, const void *dev, size_t bytes) {
    KCFD_ARG(ctx && (bytes == 0 || (dev && host)), "null argument");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (!bytes) return KCFD_OK;
    KCFD_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    KCFD_HIP(hipStreamSynchronize(ctx->stream));
    return KCFD_OK;
}

int kcfd_buf_fill(kcfd_ctx *ctx, double *dev, double value, size_t n) {
    KCFD_ARG(ctx && (n == 0 || dev), "null argument");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (!n) return KCFD_OK;
    hipLaunchKernelGGL(k_fill, dim3(nblk(n)), dim3(kBlock), 0, ctx->stream, dev, value, n);
    KCFD_HIP(hipGetLastError());
    return KCFD_OK;
}

// mean time of `reps` launches of a ceiling kernel over a `bytes` array (one
// untimed launch first); read_only: k_stream_read, else k_stream_copy
static int stream_ceiling(kcfd_ctx *ctx, size_t bytes, int reps, bool read_only, double *avg_ms) {
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    KCFD_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const size_t n16 = bytes / 16;
    const unsigned grid = (unsigned)(ctx->cus * 8);
    void *a = nullptr, *b = nullptr;
    const size_t bbytes = read_only ? (size_t)grid * kBlock * sizeof(double) : n16 * 16;
    KCFD_HIP(hipMalloc(&a, n16 * 16));
    if (hipMalloc(&b, bbytes) != hipSuccess) {
        hipFree(a);
        set_error("stream ceiling: hipMalloc of %zu bytes failed", bbytes);
        return KCFD_EOOM;
    }
    hipMemsetAsync(a, 0, n16 * 16, s);
    auto launch = [&]() {
        if (read_only)
            hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(kBlock), 0, s, n16, (const double2 *)a, (double *)b);
        else
            hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(kBlock), 0, s, n16, (const double2 *)a,
                               (double2 *)b);
    };
    launch();
    hipEventRecord(ctx->ev0, s);
    for (int r = 0; r < reps; ++r) launch();
    hipEventRecord(ctx->ev1, s);
    hipError_t e = hipEventSynchronize(ctx->ev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    hipFree(a);
    hipFree(b);
    if (e != hipSuccess) {
        set_error("stream ceiling: %s", hipGetErrorString(e));
        return KCFD_EHIP;
    }
    *avg_ms = ms / reps;
    return KCFD_OK;
}

int kcfd_stream_copy(kcfd_ctx *ctx, size_t bytes, int reps, double *avg_ms) {
    KCFD_ARG(ctx && avg_ms && reps > 0 && bytes >= 16, "bad argument");
    return stream_ceiling(ctx, bytes, reps, false, avg_ms);
}

int kcfd_stream_read(kcfd_ctx *ctx, size_t bytes, int reps, double *avg_ms) {
    KCFD_ARG(ctx && avg_ms && reps > 0 && bytes >= 16, "bad argument");
    return stream_ceiling(ctx, bytes, reps, true, avg_ms);
}

int kcfd_mat_create(kcfd_ctx *ctx, int n, int nnz, const int *indptr, const int *indices,
                    kcfd_mat **out) {
    KCFD_ARG(ctx && out, "null argument");
    KCFD_ARG(n >= 0 && nnz >= 0, "negative size");
    KCFD_ARG(n == 0 || (indptr && (nnz == 0 || indices)), "null structure");
    if (n > 0) {
        KCFD_ARG(indptr[0] == 0 && indptr[n] == nnz, "indptr inconsistent with nnz");
        for (int i = 0; i < n; ++i) {
            KCFD_ARG(indptr[i + 1] >= indptr[i], "indptr not monotone");
            for (int q = indptr[i]; q < indptr[i + 1]; ++q) {
                KCFD_ARG(indices[q] >= 0 && indices[q] < n, "column index out of range");
                KCFD_ARG(q == indptr[i] || indices[q] > indices[q - 1],
                         "columns not sorted/unique (canonical CSR required)");
            }
        }
    }
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    KCFD_HIP(hipSetDevice(ctx->device));
    kcfd_mat *m = new kcfd_mat();
    m->ctx = ctx;
    m->n = n;
    m->nnz = nnz;
    m->h_indptr.assign(indptr, indptr + n + 1);
    m->h_indices.assign(indices, indices + nnz);
    m->h_diagpos.assign(n, -1);
    for (int i = 0; i < n; ++i)
        for (int q = indptr[i]; q < indptr[i + 1]; ++q)
            if (indices[q] == i) m->h_diagpos[i] = q;
    auto fail = [&](int rc) { kcfd_mat_destroy(m); return rc; };
    if (hipMalloc(&m->indptr, sizeof(int) * (n + 1)) != hipSuccess ||
        hipMalloc(&m->indices, sizeof(int) * (nnz ? nnz : 1)) != hipSuccess ||
        hipMalloc(&m->diagpos, sizeof(int) * (n ? n : 1)) != hipSuccess ||
        hipMalloc(&m->val, sizeof(double) * (nnz ? nnz : 1)) != hipSuccess ||
        hipMalloc(&m->rhs, sizeof(double) * (n ? n : 1)) != hipSuccess ||
        hipMalloc(&m->diagt, sizeof(double) * (n ? n : 1)) != hipSuccess) {
        set_error("hipMalloc failed for a matrix of n=%d nnz=%d", n, nnz);
        return fail(KCFD_EOOM);
    }
    // uploads and fills on the library stream (created non-blocking: the
    // null stream is not ordered with the kernels that use these arrays)
    hipStream_t st = ctx->stream;
    if (hipMemcpyAsync(m->indptr, indptr, sizeof(int) * (n + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
        (nnz && hipMemcpyAsync(m->indices, indices, sizeof(int) * nnz, hipMemcpyHostToDevice, st) != hipSuccess) ||
        (n && hipMemcpyAsync(m->diagpos, m->h_diagpos.data(), sizeof(int) * n, hipMemcpyHostToDevice, st) != hipSuccess) ||
        hipMemsetAsync(m->val, 0, sizeof(double) * (nnz ? nnz : 1), st) != hipSuccess ||
        hipMemsetAsync(m->rhs, 0, sizeof(double) * (n ? n : 1), st) != hipSuccess ||
        hipMemsetAsync(m->diagt, 0, sizeof(double) * (n ? n : 1), st) != hipSuccess) {
        set_error("matrix upload failed");
        return fail(KCFD_EHIP);
    }
    {
        const int rc = ensure_sell(m, false);   // waits for the uploads above
        if (rc != KCFD_OK) return fail(rc);
    }
    *out = m;
    return KCFD_OK;
}

int kcfd_mat_destroy(kcfd_mat *m) {
    if (!m) return KCFD_OK;
    std::lock_guard<std::recursive_mutex> g(m->ctx->mu);
    hipSetDevice(m->ctx->device);
    hipStreamSynchronize(m->ctx->stream);
    if (m->solver) kcfd_batch_destroy(m->solver);
    gmres_free(m);
    mat_free_precond(m);
    mat_free_ilu_structure(m);
    hipFree(m->indptr);
    hipFree(m->indices);
    hipFree(m->diagpos);
    hipFree(m->val);
    hipFree(m->rhs);
    hipFree(m->diagt);
    hipFree(m->slice_ptr);
    hipFree(m->six);
    hipFree(m->sa64);
    delete m;
    return KCFD_OK;
}

int kcfd_mat_set_values(kcfd_mat *m, const double *vals) {
    KCFD_ARG(m && (m->nnz == 0 || vals), "null argument");
    std::lock_guard<std::recursive_mutex> g(m->ctx->mu);
    if (!m->nnz) return KCFD_OK;
    ++m->val_gen;
    KCFD_HIP(hipMemcpyAsync(m->val, vals, sizeof(double) * m->nnz, hipMemcpyHostToDevice,
                            m->ctx->stream));
    KCFD_HIP(hipStreamSynchronize(m->ctx->stream));
    return KCFD_OK;
}

int kcfd_mat_get_values(kcfd_mat *m, double *vals) {
    KCFD_ARG(m && (m->nnz == 0 || vals), "null argument");
    std::lock_guard<std::recursive_mutex> g(m->ctx->mu);
    if (!m->nnz) return KCFD_OK;
    KCFD_HIP(hipMemcpyAsync(vals, m->val, sizeof(double) * m->nnz, hipMemcpyDeviceToHost,
                            m->ctx->stream));
    KCFD_HIP(hipStreamSynchronize(m->ctx->stream));
    return KCFD_OK;
}

int kcfd_mat_dev_ptrs(kcfd_mat *m, double **vals, double **rhs, double **diag) {
    KCFD_ARG(m != nullptr, "null matrix");
    if (vals) *vals = m->val;
    if (rhs) *rhs = m->rhs;
    if (diag) *diag = m->diagt;
    return KCFD_OK;
}

int kcfd_mat_spmv(kcfd_mat *m, const double *x, double *y) {
    KCFD_ARG(m && (m->n == 0 || (x && y)), "null argument");
    std::lock_guard<std::recursive_mutex> g(m->ctx->mu);
    if (!m->n) return KCFD_OK;
    hipLaunchKernelGGL(k_spmv, dim3((m->n + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       m->ctx->stream, m->n, m->indptr, m->indices, m->val, x, y);
    KCFD_HIP(hipGetLastError());
    KCFD_HIP(hipStreamSynchronize(m->ctx->stream));
    return KCFD_OK;
}

int kcfd_mat_bytes(kcfd_mat *m, long long *bytes) {
    KCFD_ARG(m && bytes, "null argument");
    long long b = 4LL * (m->n + 1) + 4LL * m->nnz + 4LL * m->n + 8LL * m->nnz + 16LL * m->n;
    b += 4LL * (m->nslices + 1) + 12LL * m->sell_nnz;
    if (m->spk32) b += 8LL * m->sell_nnz;
    if (m->so16) b += 6LL * m->sell_nnz + 4LL * m->nslices;
    if (m->dinv) b += 8LL * m->n;
    if (m->dinv32) b += 4LL * m->n;
    if (m->lu) b += 8LL * m->nnz + 8LL * m->n;
    if (m->perm) b += 4LL * (m->n + 1 + m->n + m->n) + 12LL * m->nnz;
    if (m->lu32) b += 4LL * m->nnz;
    b += (long long)gmres_bytes(m);
    *bytes = b;
    return KCFD_OK;
}

int kcfd_precond_setup(kcfd_mat *m, int kind, int precision) {
    KCFD_ARG(m != nullptr, "null matrix");
    KCFD_ARG(kind == KCFD_PC_NONE || kind == KCFD_PC_JACOBI || kind == KCFD_PC_ILU0,
             "unknown preconditioner kind");
    KCFD_ARG(precision == KCFD_FP64 || precision == KCFD_MIXED, "unknown precision");
    std::lock_guard<std::recursive_mutex> g(m->ctx->mu);
    KCFD_HIP(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    const int n = m->n;
    const bool want32 = precision == KCFD_MIXED;
    if (want32) {
        // the FP32 copy holds each row's sum at its diagonal entry
        // (difference form): a row without a structural diagonal has no
        // place for it
        for (int i = 0; i < n; ++i)
            if (m->h_diagpos[i] < 0) {
                set_error("mixed precision needs a diagonal entry in every row (row %d has none in the "
                          "sparsity pattern); use KCFD_FP64", i);
                return KCFD_EARG;
            }
    }
    // Same kind and precision as last time: only the values changed, so the
    // buffers (and any solver workspace that points at them) are reused.
    const bool same = m->pc_kind == kind && m->have32 == want32 &&
                      (kind != KCFD_PC_ILU0 || (m->lu != nullptr && m->ilu_built_order == m->ilu_order)) &&
                      (kind != KCFD_PC_JACOBI || m->dinv != nullptr);
    if (!same) {
        KCFD_HIP(hipStreamSynchronize(s));
        if (m->solver) {
            kcfd_batch_destroy(m->solver);
            m->solver = nullptr;
        }
        mat_free_precond(m);
    }
    // the solver iterates in the ILU(0) ordering (P A P^T) when there is one:
    // ordering first, then the SELL copy in that row order
    if (kind == KCFD_PC_ILU0) KCFD_CHECK(ilu0_structure(m));
    KCFD_CHECK(ensure_sell(m, kind == KCFD_PC_ILU0 && m->perm != nullptr));
    if (want32 && !m->spk32 && !m->so16) KCFD_CHECK(sell_mixed_alloc(m));
    if (kind == KCFD_PC_JACOBI && !m->dinv) {
        KCFD_HIP(hipMalloc(&m->dinv, sizeof(double) * (n ? n : 1)));
        if (want32) KCFD_HIP(hipMalloc(&m->dinv32, sizeof(float) * (n ? n : 1)));
    }
    if (n) {
        hipLaunchKernelGGL(k_sell_values, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n,
                           m->indptr, m->indices, m->slice_ptr, m->six, m->val, m->sa64,
                           want32 ? m->spk32 : nullptr, want32 ? m->sv32 : nullptr,
                           want32 ? m->so16 : nullptr, m->sbase, m->sell_permuted ? m->perm : nullptr);
        KCFD_HIP(hipGetLastError());
    }
    if (kind == KCFD_PC_JACOBI && n) {
        hipLaunchKernelGGL(k_jacobi, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n,
                           m->diagpos, m->val, m->dinv, m->dinv32);
        KCFD_HIP(hipGetLastError());
    } else if (kind == KCFD_PC_ILU0) {
        KCFD_CHECK(ilu0_setup(m, want32));
    }
    m->pc_kind = kind;
    m->have32 = want32;
    KCFD_HIP(hipStreamSynchronize(s));
    m->setup_gen = m->val_gen;
    return KCFD_OK;
}

}  // extern "C"
