"""Thresholded incomplete Cholesky preconditioner (CPU; test infrastructure).

Restates pymatting/preconditioner/ichol.py:148-301 around the C factorisation
in ``ichol.c``: CSR input is transposed to CSC, the factorisation is retried
over increasing diagonal shifts when a pivot is not positive, and the result
applies L^-T L^-1 by forward and back substitution.
"""
import numpy as np
import scipy.sparse

from oracle._native import lib, ptr


class CholeskyDecomposition:
    """ichol.py:148-181: applies the factor by two triangular solves."""

    def __init__(self, Ltuple):
        self.Ltuple = Ltuple

    @property
    def L(self):
        Lv, Lr, Lp = self.Ltuple
        n = len(Lp) - 1
        return scipy.sparse.csc_matrix(self.Ltuple, (n, n))

    def __call__(self, b):
        Lv, Lr, Lp = self.Ltuple
        x = np.array(b, dtype=np.float64, copy=True)
        n = len(x)
        lib().oracle_backsub_L(ptr(Lv), ptr(Lr), ptr(Lp), ptr(x), n)
        lib().oracle_backsub_LT(ptr(Lv), ptr(Lr), ptr(Lp), ptr(x), n)
        return x


def ichol(A, discard_threshold=1e-4,
          shifts=[0.0, 1e-4, 1e-3, 1e-2, 0.1, 0.5, 1.0, 10.0, 100, 1e3, 1e4, 1e5],
          max_nnz=int(4e9 / 16), relative_discard_threshold=0.0, diag_keep_discarded=True):
    """ichol.py:184-301."""
    if isinstance(A, scipy.sparse.csr_matrix):
        A = A.T
    if not isinstance(A, scipy.sparse.csc_matrix):
        raise ValueError("Matrix A must be a scipy.sparse.csc_matrix")
    if not A.has_canonical_format:
        A.sum_duplicates()
    m, n = A.shape
    assert m == n
    Lv = np.empty(max_nnz, dtype=np.float64)
    Lr = np.empty(max_nnz, dtype=np.int64)
    Lp = np.zeros(n + 1, dtype=np.int64)
    Av = np.ascontiguousarray(A.data, dtype=np.float64)
    Ar = np.ascontiguousarray(A.indices, dtype=np.int64)
    Ap = np.ascontiguousarray(A.indptr, dtype=np.int64)
    nnz = -1
    for shift in shifts:
        nnz = lib().oracle_ichol(n, ptr(Av), ptr(Ar), ptr(Ap), ptr(Lv), ptr(Lr), ptr(Lp),
                                 discard_threshold, shift, max_nnz,
                                 relative_discard_threshold, int(bool(diag_keep_discarded)))
        if nnz >= 0:
            break
        if nnz == -1:
            print("PERFORMANCE WARNING:")
            print("Thresholded incomplete Cholesky decomposition failed due to insufficient "
                  "positive-definiteness of matrix A with parameters:")
            print("    discard_threshold = %e" % discard_threshold)
            print("    shift = %e" % shift)
            print("Try decreasing discard_threshold or start with a larger shift")
            print("")
        if nnz == -2:
            raise ValueError("Thresholded incomplete Cholesky decomposition failed because more "
                             "than max_nnz non-zero elements were created. Try increasing max_nnz "
                             "or discard_threshold.")
    if nnz < 0:
        raise ValueError("Thresholded incomplete Cholesky decomposition failed due to insufficient "
                         "positive-definiteness of matrix A and diagonal shifts did not help.")
    return CholeskyDecomposition((Lv[:nnz].copy(), Lr[:nnz].copy(), Lp))
