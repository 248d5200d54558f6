/* CPU restatement of pymatting's thresholded incomplete Cholesky.
 *
 * TEST INFRASTRUCTURE / CPU BASELINE ONLY (see oracle/__init__.py).
 *
 * Follows pymatting/preconditioner/ichol.py:6-145: a left-looking column
 * factorisation of a CSC matrix (lower triangle used), where the columns k
 * that contribute to column j are kept in per-row linked lists keyed by the
 * next unprocessed row of column k; off-diagonal entries whose magnitude is
 * at or below the discard threshold are dropped, optionally folding their
 * square into the diagonal. Returns nnz, -1 on a non-positive pivot or -2 on
 * overflow of max_nnz, like the reference's _ichol.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int cmp_i64(const void* a, const void* b) {
    int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return (x > y) - (x < y);
}

int64_t oracle_ichol(int64_t n, const double* Av, const int64_t* Ar, const int64_t* Ap,
                     double* Lv, int64_t* Lr, int64_t* Lp, double discard_threshold,
                     double shift, int64_t max_nnz, double relative_discard_threshold,
                     int diag_keep_discarded) {
    int64_t* next_row = (int64_t*)calloc((size_t)n, sizeof(int64_t)); /* next entry of column k */
    int64_t* first_sub = (int64_t*)calloc((size_t)n, sizeof(int64_t)); /* first sub-diag idx of A col */
    int64_t* head = (int64_t*)malloc((size_t)n * sizeof(int64_t));     /* linked lists by row */
    double* acc = (double*)calloc((size_t)n, sizeof(double));          /* dense column accumulator */
    double* colabs = (double*)calloc((size_t)n, sizeof(double));
    unsigned char* used = (unsigned char*)calloc((size_t)n, 1);
    int64_t* rows = (int64_t*)malloc((size_t)n * sizeof(int64_t));
    double* diag = (double*)malloc((size_t)n * sizeof(double));
    int64_t nnz = 0, nrows = 0, result = 0;

    for (int64_t j = 0; j < n; ++j) {
        head[j] = -1;
        diag[j] = shift;
    }
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t p = Ap[j]; p < Ap[j + 1]; ++p) {
            int64_t i = Ar[p];
            if (i == j) {
                diag[j] += Av[p];
                first_sub[j] = p + 1;
            }
            if (i >= j) colabs[j] += fabs(Av[p]);
        }
    }
    Lp[0] = 0;
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t p = first_sub[j]; p < Ap[j + 1]; ++p) {
            int64_t i = Ar[p];
            double v = Av[p];
            if (v != 0.0 && i > j) {
                acc[i] += v;
                if (!used[i]) {
                    used[i] = 1;
                    rows[nrows++] = i;
                }
            }
        }
        int64_t k = head[j];
        while (k != -1) {
            int64_t k0 = next_row[k];
            int64_t k1 = Lp[k + 1];
            int64_t k_next = head[k];
            double L_jk = Lv[k0];
            ++k0;
            if (k0 < k1) {
                next_row[k] = k0;
                int64_t i0 = Lr[k0];
                head[k] = head[i0];
                head[i0] = k;
                for (int64_t p = k0; p < k1; ++p) {
                    int64_t i = Lr[p];
                    acc[i] -= Lv[p] * L_jk;
                    if (!used[i]) {
                        used[i] = 1;
                        rows[nrows++] = i;
                    }
                }
            }
            k = k_next;
        }
        if (diag[j] <= 0.0) { result = -1; goto done; }
        if (nnz + 1 + nrows > max_nnz) { result = -2; goto done; }
        diag[j] = sqrt(diag[j]);
        Lv[nnz] = diag[j];
        Lr[nnz] = j;
        ++nnz;
        next_row[j] = nnz;
        qsort(rows, (size_t)nrows, sizeof(int64_t), cmp_i64);
        for (int64_t t = 0; t < nrows; ++t) {
            int64_t i = rows[t];
            double L_ij = acc[i] / diag[j];
            if (diag_keep_discarded) diag[i] -= L_ij * L_ij;
            double rel = relative_discard_threshold * colabs[j];
            if (fabs(L_ij) > discard_threshold && fabs(acc[i]) > rel) {
                if (!diag_keep_discarded) diag[i] -= L_ij * L_ij;
                Lv[nnz] = L_ij;
                Lr[nnz] = i;
                ++nnz;
            }
            acc[i] = 0.0;
            used[i] = 0;
        }
        nrows = 0;
        Lp[j + 1] = nnz;
        if (Lp[j] + 1 < Lp[j + 1]) {
            int64_t i = Lr[Lp[j] + 1];
            head[j] = head[i];
            head[i] = j;
        }
    }
    result = nnz;
done:
    free(next_row); free(first_sub); free(head); free(acc); free(colabs);
    free(used); free(rows); free(diag);
    return result;
}

/* Forward substitution L y = x in place (ichol.py:116-129). */
void oracle_backsub_L(const double* data, const int64_t* indices, const int64_t* indptr,
                      double* x, int64_t n) {
    for (int64_t j = 0; j < n; ++j) {
        double t = x[j] / data[indptr[j]];
        x[j] = t;
        for (int64_t k = indptr[j] + 1; k < indptr[j + 1]; ++k) x[indices[k]] -= data[k] * t;
    }
}

/* Back substitution L^T x = y in place (ichol.py:132-145). */
void oracle_backsub_LT(const double* data, const int64_t* indices, const int64_t* indptr,
                       double* x, int64_t n) {
    for (int64_t i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int64_t k = indptr[i] + 1; k < indptr[i + 1]; ++k) s -= data[k] * x[indices[k]];
        x[i] = s / data[indptr[i]];
    }
}
