"""Preconditioners on the GPU.

Drop-ins for pymatting/preconditioner/jacobi.py:1-31 (jacobi),
pymatting/preconditioner/vcycle.py:103-153 (vcycle) and
pymatting/preconditioner/ichol.py:148-301 (ichol, CholeskyDecomposition).
All return callables (numpy or torch vectors in, the same type out) that
``cg`` recognises and runs inside its device loop. ichol is sequential by
construction: its factorisation and its triangular solves run in the
library on the host.
"""
import ctypes

import numpy as np
import scipy.sparse
import torch

from pymatting_amd import _ffi
from pymatting_amd._device import as_device, empty, output_like
from pymatting_amd.laplacian import CfOperator, DiaOperator, device_operator


def _diagonal_device(A):
    op = device_operator(A)
    if op is not None:
        return op.diagonal_device()
    if scipy.sparse.issparse(A):
        d = A.diagonal()
    else:
        d = np.diag(np.asarray(A))
    return as_device(np.asarray(d, dtype=np.float64))


class Jacobi:
    """x -> x / diag(A) (jacobi.py:24-31)."""

    def __init__(self, A):
        self.inv_diag = 1.0 / _diagonal_device(A)
        self.n = self.inv_diag.numel()

    @classmethod
    def from_inverse_diagonal(cls, inv_diag):
        """x -> x * inv_diag for a given device vector of inverse diagonal entries."""
        self = cls.__new__(cls)
        self.inv_diag = inv_diag.reshape(-1)
        self.n = self.inv_diag.numel()
        return self

    def __call__(self, x):
        if _ffi.numel(x) != self.n:
            raise ValueError("dimension mismatch: Jacobi preconditioner of size %d, vector of "
                             "size %d" % (self.n, _ffi.numel(x)))
        xd = as_device(x).reshape(-1)
        out = empty(self.n)
        _ffi.check(_ffi.lib().hm_diag_apply(_ffi.ptr(self.inv_diag), _ffi.ptr(xd), _ffi.ptr(out),
                                            self.n, _ffi.stream()))
        return output_like(out, x)


def jacobi(A):
    """jacobi.py:1-31: the Jacobi (diagonal) preconditioner of A."""
    return Jacobi(A)


class VCycle:
    """Geometric multigrid V-cycle (vcycle.py:46-153) built on the device.

    Levels: (h, w) -> (h//2, w//2) -> ... until h*w <= direct_solve_size,
    Galerkin coarse operators R A P, damped Jacobi smoothing, and a dense
    direct solve on the coarsest grid.

    Extensions beyond the reference: max_levels > 0 caps the number of grids;
    a capped coarsest grid above 2048 points is then solved approximately by
    coarse_sweeps damped Jacobi sweeps from zero.
    """

    def __init__(self, A, shape, num_pre_iter=1, num_post_iter=1, omega=0.8,
                 direct_solve_size=64, precision="f64", max_levels=0, coarse_sweeps=0):
        self.shape = (int(shape[0]), int(shape[1]))
        h, w = self.shape
        self.precision = precision
        prec = _ffi.HM_F64 if precision == "f64" else _ffi.HM_F32
        op = device_operator(A)
        handle = ctypes.c_void_p()
        L = _ffi.lib()
        self._keep = []
        if h * w <= direct_solve_size and not isinstance(op, CfOperator):
            # vcycle.py:59-60 solves the given matrix directly.
            if op is not None:
                dense = self._dense_from_dia(op)
            elif scipy.sparse.issparse(A):
                dense = A.toarray()
            else:
                dense = np.asarray(A, dtype=np.float64)
            dense = as_device(np.asarray(dense, dtype=np.float64))
            self.n = dense.shape[0]
            _ffi.check(L.hm_vcycle_create_dense(_ffi.ptr(dense), self.n, prec, _ffi.stream(),
                                                ctypes.byref(handle)))
        else:
            if op is None:
                op = DiaOperator.from_sparse(A, self.shape)
            self.n = op.n
            self._keep.append(op)
            cf = op.handle if isinstance(op, CfOperator) else None
            dia = op.handle if isinstance(op, DiaOperator) else None
            _ffi.check(L.hm_vcycle_create_ex(cf, dia, h, w, int(num_pre_iter),
                                             int(num_post_iter), float(omega),
                                             int(direct_solve_size), prec, int(max_levels),
                                             int(coarse_sweeps), _ffi.stream(),
                                             ctypes.byref(handle)))
        self.handle = handle
        self.levels = L.hm_vcycle_levels(handle)

    @staticmethod
    def _dense_from_dia(op):
        h, w = op.grid_shape
        n = op.n
        coef = op.coef.cpu().numpy() if isinstance(op, DiaOperator) else None
        if coef is None:
            raise ValueError("dense direct solve needs a stored operator")
        dense = np.zeros((n, n))
        for s in range(25):
            dy, dx = s // 5 - 2, s % 5 - 2
            for i in range(n):
                y, x = divmod(i, w)
                if 0 <= y + dy < h and 0 <= x + dx < w:
                    dense[i, (y + dy) * w + x + dx] = coef[s, i]
        return dense

    def __del__(self):
        _ffi.release(self, "hm_vcycle_destroy")

    def set_launches(self, mode):
        """Launch structure of the cycle, a bit mask: 1 runs consecutive
        coarse grids of the ascent in one launch, 2 pairs of grids of the
        descent, 4 applies the coarsest grids (those one workgroup runs in
        LDS) as their precomputed dense matrix; 0 runs every grid in a launch
        of its own. The default is 5. Bits 1 and 2 leave the result the same
        bit for bit; bit 4 changes it by rounding."""
        _ffi.check(_ffi.lib().hm_vcycle_set_launches(self.handle, int(mode)))
        return self

    def apply(self, r, out=None):
        dt = torch.float64 if self.precision == "f64" else torch.float32
        out = out if out is not None else empty(self.n, dt)
        _ffi.check(_ffi.lib().hm_vcycle_apply(self.handle, _ffi.ptr(r), _ffi.ptr(out),
                                              _ffi.stream()))
        return out

    def __call__(self, r):
        dt = torch.float64 if self.precision == "f64" else torch.float32
        rd = as_device(r, dt).reshape(-1)
        if rd.numel() != self.n:
            raise ValueError("dimension mismatch: V-cycle of size %d, vector of size %d"
                             % (self.n, rd.numel()))
        return output_like(self.apply(rd), r)


def vcycle(A, shape, num_pre_iter=1, num_post_iter=1, omega=0.8, direct_solve_size=64,
           cache=None):
    """vcycle.py:103-153: returns the V-cycle preconditioner of A on `shape`.

    `cache` (a dict) keeps the hierarchy per shape across calls, like the
    reference's per-shape cache of coarse operators.
    """
    key = (tuple(int(s) for s in shape), num_pre_iter, num_post_iter, omega, direct_solve_size)
    if cache is not None and key in cache:
        return cache[key]
    M = VCycle(A, shape, num_pre_iter, num_post_iter, omega, direct_solve_size)
    if cache is not None:
        cache[key] = M
    return M


class CholeskyDecomposition:
    """ichol.py:148-181: an incomplete Cholesky factor L (CSC) applied as
    x = L^-T L^-1 b by forward and back substitution.

    ``Ltuple`` is (Lv, Lr, Lp), the factor's values, row indices and column
    pointers (each column starts with its diagonal). Calling the object with
    a numpy vector returns a numpy vector; a torch tensor stays on the
    device. ``cg`` applies it inside its device loop: the substitutions, long
    sequential chains, run in the library on the host, between copies of
    the vector.
    """

    def __init__(self, Ltuple, _handle=None):
        Lv, Lr, Lp = Ltuple
        self.Ltuple = (np.ascontiguousarray(Lv, dtype=np.float64),
                       np.ascontiguousarray(Lr, dtype=np.int64),
                       np.ascontiguousarray(Lp, dtype=np.int64))
        self.n = len(self.Ltuple[2]) - 1
        if _handle is None:
            Lv, Lr, Lp = self.Ltuple
            _handle = ctypes.c_void_p()
            _ffi.check(_ffi.lib().hm_ichol_create(self.n, _ptr(Lv), _ptr(Lr), _ptr(Lp),
                                                  ctypes.byref(_handle)))
        self.handle = _handle

    def __del__(self):
        _ffi.release(self, "hm_ichol_destroy")

    @property
    def L(self):
        """The factor as a scipy.sparse.csc_matrix."""
        return scipy.sparse.csc_matrix(self.Ltuple, (self.n, self.n))

    def __call__(self, b):
        if _ffi.numel(b) != self.n:
            raise ValueError("dimension mismatch: factor of size %d, vector of size %d"
                             % (self.n, _ffi.numel(b)))
        bd = as_device(b).reshape(-1)
        out = empty(self.n)
        _ffi.check(_ffi.lib().hm_ichol_apply(self.handle, _ffi.ptr(bd), _ffi.ptr(out),
                                             _ffi.stream()))
        return output_like(out, b)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def ichol(A, discard_threshold=1e-4,
          shifts=[0.0, 1e-4, 1e-3, 1e-2, 0.1, 0.5, 1.0, 10.0, 100, 1e3, 1e4, 1e5],
          max_nnz=int(4e9 / 16), relative_discard_threshold=0.0, diag_keep_discarded=True):
    """ichol.py:184-301: thresholded incomplete Cholesky decomposition.

    A CSR matrix is transposed to CSC (A is symmetric), a non-canonical one
    has its duplicates summed in place, and the factorisation is retried with
    each diagonal shift in turn until every pivot is positive (printing the
    reference's PERFORMANCE WARNING for each failed shift). The factorisation
    runs in the library on the host, like the reference's, and so do the
    substitutions that apply it (CholeskyDecomposition); cg keeps every
    other step of the iteration on the GPU.
    """
    if isinstance(A, scipy.sparse.csr_matrix):
        A = A.T
    if not isinstance(A, scipy.sparse.csc_matrix):
        raise ValueError("Matrix A must be a scipy.sparse.csc_matrix")
    if not A.has_canonical_format:
        # On a copy: the transpose of a CSR shares its arrays, and summing
        # duplicates in place would leave the caller's CSR inconsistent.
        A = A.copy()
        A.sum_duplicates()
    m, n = A.shape
    assert m == n
    Av = np.ascontiguousarray(A.data, dtype=np.float64)
    Ar = np.ascontiguousarray(A.indices, dtype=np.int64)
    Ap = np.ascontiguousarray(A.indptr, dtype=np.int64)
    L = _ffi.lib()
    nnz = -1
    handle = None
    for shift in shifts:
        result = ctypes.c_int64()
        h = ctypes.c_void_p()
        _ffi.check(L.hm_ichol_factor(n, _ptr(Ap), _ptr(Ar), _ptr(Av), float(discard_threshold),
                                     float(shift), int(max_nnz),
                                     float(relative_discard_threshold),
                                     int(bool(diag_keep_discarded)), ctypes.byref(result),
                                     ctypes.byref(h)))
        nnz = result.value
        if nnz >= 0:
            handle = h
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
                             "than max_nnz non-zero elements were created. Try increasing "
                             "max_nnz or discard_threshold.")
    if nnz < 0:
        raise ValueError("Thresholded incomplete Cholesky decomposition failed due to "
                         "insufficient positive-definiteness of matrix A and diagonal shifts "
                         "did not help.")
    Lv = np.empty(nnz, dtype=np.float64)
    Lr = np.empty(nnz, dtype=np.int64)
    Lp = np.empty(n + 1, dtype=np.int64)
    _ffi.check(L.hm_ichol_copy_factor(handle, _ptr(Lv), _ptr(Lr), _ptr(Lp)))
    return CholeskyDecomposition((Lv, Lr, Lp), _handle=handle)
