"""The library's incomplete Cholesky factorisation (hm_ichol_factor, host
code, no device call) against the oracle's restatement of
pymatting/preconditioner/ichol.py:6-113: the same factor, bit for bit, on the
reference's own test matrices (tests/test_ichol.py: random sparse SPD,
no dropping), on the lemur lambda-system at scale 0.125 with the default
threshold (tests/test_preconditioners.py), with relative thresholds and
both diagonal rules, and the reference's return codes for a pivot that is
not positive (-1) and for max_nnz (-2)."""
import ctypes
import os

import numpy as np
import pytest
import scipy.sparse

from conftest import ROOT, load_golden

LIB = os.path.join(ROOT, "pymatting_amd", "libpymatting_amd.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")


def _lib():
    import torch  # noqa: F401  (loads the HIP runtime the library links)
    lib = ctypes.CDLL(LIB)
    P, i64, f64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int
    lib.hm_ichol_factor.argtypes = [i64, P, P, P, f64, f64, i64, f64, i32,
                                    ctypes.POINTER(i64), ctypes.POINTER(P)]
    lib.hm_ichol_copy_factor.argtypes = [P, P, P, P]
    lib.hm_ichol_nnz.restype = i64
    lib.hm_ichol_nnz.argtypes = [P]
    lib.hm_ichol_destroy.argtypes = [P]
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def factor(A, discard=1e-4, shift=0.0, max_nnz=10 ** 8, rel=0.0, keep=True):
    lib = _lib()
    A = scipy.sparse.csc_matrix(A)
    A.sum_duplicates()
    Ap = A.indptr.astype(np.int64)
    Ar = A.indices.astype(np.int64)
    Av = A.data.astype(np.float64)
    res = ctypes.c_int64()
    h = ctypes.c_void_p()
    assert lib.hm_ichol_factor(A.shape[0], _p(Ap), _p(Ar), _p(Av), discard, shift, max_nnz, rel,
                               int(keep), ctypes.byref(res), ctypes.byref(h)) == 0
    if res.value < 0:
        assert not h.value
        return res.value, None
    nnz = lib.hm_ichol_nnz(h)
    Lv, Lr, Lp = np.empty(nnz), np.empty(nnz, np.int64), np.empty(A.shape[0] + 1, np.int64)
    assert lib.hm_ichol_copy_factor(h, _p(Lv), _p(Lr), _p(Lp)) == 0
    lib.hm_ichol_destroy(h)
    return res.value, (Lv, Lr, Lp)


def oracle_factor(A, discard=1e-4, shift=0.0, max_nnz=10 ** 8, rel=0.0, keep=True):
    from oracle._native import lib, ptr
    A = scipy.sparse.csc_matrix(A)
    A.sum_duplicates()
    n = A.shape[0]
    Lv = np.empty(max(A.nnz * 50, 16))
    Lr = np.empty(len(Lv), np.int64)
    Lp = np.zeros(n + 1, np.int64)
    Av = np.ascontiguousarray(A.data, np.float64)
    Ar = np.ascontiguousarray(A.indices, np.int64)
    Ap = np.ascontiguousarray(A.indptr, np.int64)
    nnz = lib().oracle_ichol(n, ptr(Av), ptr(Ar), ptr(Ap), ptr(Lv), ptr(Lr), ptr(Lp), discard,
                             shift, min(max_nnz, len(Lv)), rel, int(keep))
    if nnz < 0:
        return nnz, None
    return nnz, (Lv[:nnz], Lr[:nnz], Lp)


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_factor_matches_oracle_on_random_spd():
    np.random.seed(0)  # tests/test_ichol.py
    for _ in range(10):
        n = 50
        A = np.random.rand(n, n)
        A[np.random.rand(n, n) < 0.8] = 0
        A += A.T + n * np.eye(n)
        for discard in (0.0, 1e-4, 0.05):
            nnz, L = factor(A, discard=discard, max_nnz=100000)
            nnz_o, L_o = oracle_factor(A, discard=discard, max_nnz=100000)
            assert nnz == nnz_o and _same(L, L_o)
        # No dropping: L L^T = A exactly up to rounding.
        _, (Lv, Lr, Lp) = factor(A, discard=0.0, max_nnz=100000)
        Lm = scipy.sparse.csc_matrix((Lv, Lr, Lp), (n, n)).toarray()
        assert np.abs(Lm @ Lm.T - A).max() < 1e-10 * np.abs(A).max()


def test_factor_matches_oracle_on_lemur_system():
    import oracle.cf as ocf
    d = load_golden("lemur_0p125")
    A, _ = ocf.make_linear_system(ocf.cf_laplacian(d["image"]), d["trimap"])
    for kw in (dict(), dict(rel=1e-3), dict(keep=False), dict(discard=1e-3, shift=2e-3)):
        nnz, L = factor(A, max_nnz=500000, **kw)
        nnz_o, L_o = oracle_factor(A, max_nnz=500000, **kw)
        assert nnz == nnz_o and nnz > A.shape[0] and _same(L, L_o), kw


def test_return_codes():
    A = np.array([[1.0, 2.0], [2.0, 1.0]])  # indefinite: the second pivot is -3
    assert factor(A, discard=0.0)[0] == -1
    assert oracle_factor(A, discard=0.0)[0] == -1
    nnz, L = factor(A, discard=0.0, shift=4.0)
    assert nnz == 3 and _same(L, oracle_factor(A, discard=0.0, shift=4.0)[1])
    B = np.random.default_rng(1).random((30, 30))
    B = B + B.T + 30 * np.eye(30)
    assert factor(B, discard=0.0, max_nnz=40)[0] == -2
