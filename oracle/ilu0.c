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

/*
 * First-fit colouring of the graph of A + A^T (the ILU(0) "color" ordering
 * of the GPU path).  Rows are visited in ascending order; a row takes the
 * smallest colour that no already coloured neighbour (row or column
 * neighbour) has.  perm receives the rows sorted by colour, ascending row id
 * within a colour; colors (may be NULL) the colour of each row.  Returns the
 * number of colours, or -1 on allocation failure.
 */
int kcfd_oracle_color_order(int n, const int *indptr, const int *indices, int *perm, int *colors)
{
    int nnz = n > 0 ? indptr[n] : 0;
    int *tp = (int *)calloc((size_t)n + 1, sizeof(int));
    int *tx = (int *)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    int *fill = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *col = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *seen = (int *)malloc(sizeof(int) * (size_t)(n + 1));  /* seen[c] == i: c is taken */
    int ncol = 0;
    if (!tp || !tx || !fill || !col || !seen) { ncol = -1; goto out; }
    for (int q = 0; q < nnz; ++q) tp[indices[q] + 1]++;
    for (int j = 0; j < n; ++j) tp[j + 1] += tp[j];
    for (int j = 0; j < n; ++j) fill[j] = tp[j];
    for (int i = 0; i < n; ++i)
        for (int q = indptr[i]; q < indptr[i + 1]; ++q) tx[fill[indices[q]]++] = i;
    for (int c = 0; c <= n; ++c) seen[c] = -1;
    for (int i = 0; i < n; ++i) col[i] = -1;
    for (int i = 0; i < n; ++i) {
        for (int q = indptr[i]; q < indptr[i + 1]; ++q) {
            int j = indices[q];
            if (j != i && col[j] >= 0) seen[col[j]] = i;
        }
        for (int q = tp[i]; q < tp[i + 1]; ++q) {
            int j = tx[q];
            if (j != i && col[j] >= 0) seen[col[j]] = i;
        }
        int c = 0;
        while (seen[c] == i) ++c;
        col[i] = c;
        if (c + 1 > ncol) ncol = c + 1;
    }
    {
        int k = 0;
        for (int c = 0; c < ncol; ++c)
            for (int i = 0; i < n; ++i)
                if (col[i] == c) perm[k++] = i;
    }
    if (colors) for (int i = 0; i < n; ++i) colors[i] = col[i];
out:
    free(tp); free(tx); free(fill); free(col); free(seen);
    return ncol;
}
