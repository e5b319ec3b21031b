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
#include "common.hpp"

#include <algorithm>

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

int ilu0_setup(kcfd_mat *m, bool want32) {
    const int n = m->n;
    hipStream_t s = m->ctx->stream;
    for (int i = 0; i < n; ++i) {
        if (m->h_diagpos[i] < 0) {
            set_error("ILU(0): row %d has no diagonal entry", i);
            return KCFD_EPIVOT;
        }
    }
    // the level schedule depends on the structure only: built once
    if (!m->lvl_rows_l) {
        std::vector<int> rl, ru;
        level_schedule(m->h_indptr, m->h_indices, m->h_diagpos, false, rl, m->h_lvl_ptr_l);
        level_schedule(m->h_indptr, m->h_indices, m->h_diagpos, true, ru, m->h_lvl_ptr_u);
        m->nlev_l = (int)m->h_lvl_ptr_l.size() - 1;
        m->nlev_u = (int)m->h_lvl_ptr_u.size() - 1;
        KCFD_HIP(hipMalloc(&m->lvl_rows_l, sizeof(int) * std::max(n, 1)));
        KCFD_HIP(hipMalloc(&m->lvl_rows_u, sizeof(int) * std::max(n, 1)));
        if (n) {
            KCFD_HIP(hipMemcpy(m->lvl_rows_l, rl.data(), sizeof(int) * n, hipMemcpyHostToDevice));
            KCFD_HIP(hipMemcpy(m->lvl_rows_u, ru.data(), sizeof(int) * n, hipMemcpyHostToDevice));
        }
    }
    const std::vector<int> &pl = m->h_lvl_ptr_l;
    if (!m->lu) KCFD_HIP(hipMalloc(&m->lu, sizeof(double) * std::max(m->nnz, 1)));
    if (m->nnz)
        KCFD_HIP(hipMemcpyAsync(m->lu, m->val, sizeof(double) * m->nnz, hipMemcpyDeviceToDevice, s));
    for (int l = 0; l < m->nlev_l; ++l) {
        const int cnt = pl[l + 1] - pl[l];
        hipLaunchKernelGGL(k_ilu0_level, dim3((cnt + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                           m->lvl_rows_l + pl[l], cnt, m->indptr, m->indices, m->diagpos, m->lu);
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
    // pivot check: first row with a zero or NaN u_ii (or n if none)
    KCFD_CHECK(ensure_scratch(m->ctx, sizeof(int)));
    int *bad = (int *)m->ctx->scratch;
    const int init = n;
    KCFD_HIP(hipMemcpyAsync(bad, &init, sizeof(int), hipMemcpyHostToDevice, s));
    if (n)
        hipLaunchKernelGGL(k_pivot_check, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n,
                           m->diagpos, m->lu, bad);
    int first = n;
    KCFD_HIP(hipMemcpyAsync(&first, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    KCFD_HIP(hipStreamSynchronize(s));
    if (first < n) {
        set_error("ILU(0): zero or NaN pivot at row %d", first);
        return KCFD_EPIVOT;
    }
    return KCFD_OK;
}

// one level of y = L^{-1} z (unit lower) or w = U^{-1} y, rows of that level
__global__ __launch_bounds__(kBlock) void k_trsv_level(const int *__restrict__ rows, int count,
                                                       const int *__restrict__ ip,
                                                       const int *__restrict__ ix,
                                                       const int *__restrict__ dpos,
                                                       const double *__restrict__ lu,
                                                       const double *z,  // may alias y (U solve in place)
                                                       double *y, int upper) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int i = rows[t];
    const int dp = dpos[i];
    double acc = z[i];
    if (!upper) {
        for (int q = ip[i]; q < dp; ++q) acc -= lu[q] * y[ix[q]];
        y[i] = acc;
    } else {
        for (int q = dp + 1; q < ip[i + 1]; ++q) acc -= lu[q] * y[ix[q]];
        y[i] = acc / lu[dp];
    }
}

__global__ void k_jacobi_apply(int n, const double *__restrict__ dinv, const double *__restrict__ x,
                               double *__restrict__ y) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = dinv[i] * x[i];
}

__global__ void k_copy(int n, const double *__restrict__ x, double *__restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i];
}

}  // namespace kcfd

using namespace kcfd;

extern "C" int kcfd_precond_apply(kcfd_mat *m, const double *x, double *y) {
    KCFD_ARG(m && (m->n == 0 || (x && y)), "null argument");
    KCFD_ARG(x != y, "x and y must not alias");
    std::lock_guard<std::recursive_mutex> g(m->ctx->mu);
    KCFD_HIP(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    const int n = m->n;
    if (!n) return KCFD_OK;
    const unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
    if (m->pc_kind == KCFD_PC_JACOBI) {
        hipLaunchKernelGGL(k_jacobi_apply, dim3(nb), dim3(kBlock), 0, s, n, m->dinv, x, y);
    } else if (m->pc_kind == KCFD_PC_ILU0) {
        // level by level: y = L^{-1} x, then y = U^{-1} y in place
        for (int l = 0; l < m->nlev_l; ++l) {
            const int c = m->h_lvl_ptr_l[l + 1] - m->h_lvl_ptr_l[l];
            hipLaunchKernelGGL(k_trsv_level, dim3((c + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                               m->lvl_rows_l + m->h_lvl_ptr_l[l], c, m->indptr, m->indices,
                               m->diagpos, m->lu, x, y, 0);
        }
        for (int l = 0; l < m->nlev_u; ++l) {
            const int c = m->h_lvl_ptr_u[l + 1] - m->h_lvl_ptr_u[l];
            hipLaunchKernelGGL(k_trsv_level, dim3((c + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                               m->lvl_rows_u + m->h_lvl_ptr_u[l], c, m->indptr, m->indices,
                               m->diagpos, m->lu, y, y, 1);
        }
    } else {
        hipLaunchKernelGGL(k_copy, dim3(nb), dim3(kBlock), 0, s, n, x, y);
    }
    KCFD_HIP(hipGetLastError());
    KCFD_HIP(hipStreamSynchronize(s));
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
