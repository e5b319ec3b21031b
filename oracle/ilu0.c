/*
 * Textbook structural ILU(0) on a canonical CSR pattern, and its triangular
 * solves.  TEST INFRASTRUCTURE ONLY: the oracle's preconditioner for the
 * GPU ILU(0) parity tests and the CPU baseline.
 *
 * Algorithm: Saad, "Iterative Methods for Sparse Linear Systems", 2nd ed.,
 * Alg. 10.4 (ILU(0), IKJ variant).  L is unit lower triangular and U upper
 * triangular; both are stored in place of A's values on A's pattern.
 *
 * The reference's "ilu0" preconditioner (core/preconditioner.py:49-54) is
 * scipy's spilu (SuperLU ILUTP); this structural ILU(0) is what the GPU path
 * implements, so it is the parity target for iteration counts.
 *
 * Requirements: sorted column indices, every diagonal entry present.
 * Returns 0, or -(row+1) for a missing or zero pivot.
 */
#include <stdlib.h>
#include <string.h>

int kcfd_oracle_ilu0_factor(int n, const int *indptr, const int *indices,
                            const double *a, double *lu, int *diagpos)
{
    int *iw = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    if (!iw) return -1000000000;
    for (int j = 0; j < n; ++j) iw[j] = -1;
    memcpy(lu, a, sizeof(double) * (size_t)indptr[n]);
    int status = 0;
    for (int i = 0; i < n; ++i) {
        int rs = indptr[i], re = indptr[i + 1];
        for (int q = rs; q < re; ++q) iw[indices[q]] = q;
        int dp = -1;
        for (int q = rs; q < re; ++q) {
            int k = indices[q];
            if (k >= i) { if (k == i) dp = q; break; }
            /* l_ik = a_ik / u_kk */
            double piv = lu[diagpos[k]];
            double lik = lu[q] / piv;
            lu[q] = lik;
            /* a_ij -= l_ik u_kj for j > k in pattern(row i) */
            for (int r = diagpos[k] + 1; r < indptr[k + 1]; ++r) {
                int pos = iw[indices[r]];
                if (pos >= 0) lu[pos] -= lik * lu[r];
            }
        }
        if (dp < 0) {
            for (int q = rs; q < re; ++q) if (indices[q] == i) dp = q;
        }
        diagpos[i] = dp;
        for (int q = rs; q < re; ++q) iw[indices[q]] = -1;
        if (dp < 0 || lu[dp] == 0.0) { status = -(i + 1); break; }
    }
    free(iw);
    return status;
}

/* x = U^{-1} L^{-1} b */
void kcfd_oracle_ilu0_solve(int n, const int *indptr, const int *indices,
                            const double *lu, const int *diagpos,
                            const double *b, double *x)
{
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        for (int q = indptr[i]; q < diagpos[i]; ++q) s -= lu[q] * x[indices[q]];
        x[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int q = diagpos[i] + 1; q < indptr[i + 1]; ++q) s -= lu[q] * x[indices[q]];
        x[i] = s / lu[diagpos[i]];
    }
}
