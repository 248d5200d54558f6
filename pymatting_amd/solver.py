"""Preconditioned conjugate gradients on the GPU.

Drop-in for pymatting/solver/cg.py:4-117 and the callbacks of
pymatting/solver/callback.py:1-23. The iteration runs in the library with
every vector on the device. Operators and preconditioners from this package
run as device kernels; a scipy or numpy matrix is uploaded as CSR; Python
callables (A(x), M(r)) and the per-iteration callback are invoked from the
device loop through ctypes callbacks with host copies of the vectors.
"""
import ctypes

import numpy as np
import scipy.sparse
import torch

from pymatting_amd import _ffi
from pymatting_amd._device import as_device, empty, is_torch, output_like
from pymatting_amd.laplacian import CfOperator, DiaOperator, attached_csr, device_operator
from pymatting_amd.lkm import LkmOperator
from pymatting_amd.sparse import DeviceCSR
from pymatting_amd.preconditioner import CholeskyDecomposition, Jacobi, VCycle


class CounterCallback:
    """callback.py:1-8: counts iterations."""

    def __init__(self):
        self.n = 0

    def __call__(self, A, x, b, norm_b, r, norm_r):
        self.n += 1


class ProgressCallback:
    """callback.py:11-23: counts iterations and prints the residual norm."""

    def __init__(self):
        self.n = 0

    def __call__(self, A, x, b, norm_b, r, norm_r):
        self.n += 1
        print("iteration %7d - %e (%.20f)" % (self.n, norm_r, norm_r))


def _to_host(ptr, n):
    out = np.empty(n, dtype=np.float64)
    _ffi.check(_ffi.lib().hm_copy(out.ctypes.data_as(ctypes.c_void_p), ptr, 8 * n, 0,
                                  _ffi.stream()))
    return out


def _to_device(arr, ptr, n):
    arr = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).reshape(-1))
    if arr.size != n:
        raise ValueError("callable returned %d values, expected %d" % (arr.size, n))
    _ffi.check(_ffi.lib().hm_copy(ptr, arr.ctypes.data_as(ctypes.c_void_p), 8 * n, 1,
                                  _ffi.stream()))


class _HostApply:
    """Wraps a Python callable f(x) as a library callback on device vectors."""

    def __init__(self, fn, n):
        self.fn, self.n, self.error = fn, n, None
        self.cfunc = _ffi.APPLY_CB(self._call)

    def _call(self, ctx, x, y):
        try:
            _to_device(self.fn(_to_host(x, self.n)), y, self.n)
            return 0
        except BaseException as e:  # re-raised after the library returns
            self.error = e
            return 1


class _HostIter:
    def __init__(self, callback, A, b_host, n):
        self.callback, self.A, self.b, self.n, self.error = callback, A, b_host, n, None
        self.cfunc = _ffi.ITER_CB(self._call)

    def _call(self, ctx, x, r, norm_b, norm_r):
        try:
            self.callback(self.A, _to_host(x, self.n), self.b, norm_b, _to_host(r, self.n), norm_r)
            return 0
        except BaseException as e:
            self.error = e
            return 1


def _check_sizes(A, op, M, x0, n):
    """ValueError (dimension mismatch) unless A, M and x0 all match len(b)."""
    def mismatch(what):
        raise ValueError("dimension mismatch: %s, but b has size %d" % (what, n))
    if op is not None or isinstance(A, LkmOperator):
        m = (op if op is not None else A).n
        if m != n:
            mismatch("operator of size %d" % m)
    elif scipy.sparse.issparse(A) or isinstance(A, np.ndarray):
        if A.ndim != 2 or tuple(A.shape) != (n, n):
            mismatch("matrix of shape %s" % (tuple(A.shape),))
    if x0 is not None and _ffi.numel(x0) != n:
        mismatch("x0 of size %d" % _ffi.numel(x0))
    if isinstance(M, (Jacobi, VCycle, CholeskyDecomposition)):
        if M.n != n:
            mismatch("preconditioner of size %d" % M.n)
    elif scipy.sparse.issparse(M) or isinstance(M, np.ndarray):
        if M.ndim != 2 or tuple(M.shape) != (n, n):
            mismatch("preconditioner matrix of shape %s" % (tuple(M.shape),))


def cg(A, b, x0=None, atol=0.0, rtol=1e-7, maxiter=10000, callback=None, M=None,
       reorthogonalize=False, check_every=8, return_stats=False):
    """cg.py:4-117: solves A x = b by preconditioned conjugate gradients.

    A: a matrix from this package (matrix-free on the device), a scipy sparse
    matrix or numpy array (uploaded as CSR), or a callable x -> A x.
    M: None, ``jacobi(A)``, ``vcycle(A, shape)``, a matrix, or a callable.
    b/x0: numpy arrays or torch tensors; x is returned in b's type.
    Raises ValueError if the iteration does not converge within maxiter.
    """
    L = _ffi.lib()
    n = _ffi.numel(b)
    # Every size is checked before anything is uploaded or launched: the
    # library runs over the operator's n and trusts the vectors' lengths
    # (the reference raises ValueError from A.dot on a mismatch).
    op = device_operator(A)
    _check_sizes(A, op, M, x0, n)
    bd = as_device(b).reshape(-1)
    args = _ffi.CgArgs()
    keep = []
    if isinstance(op, CfOperator):
        args.cf = op.handle
    elif isinstance(op, DiaOperator):
        args.dia = op.handle
    elif isinstance(A, LkmOperator):
        args.lkm = A.handle
    elif attached_csr(A) is not None:
        args.csr = attached_csr(A).handle
    elif scipy.sparse.issparse(A) or isinstance(A, np.ndarray):
        csr = DeviceCSR(A)
        keep.append(csr)
        args.csr = csr.handle
    elif callable(A):
        host = _HostApply(A, n)
        keep.append(host)
        args.a_cb = host.cfunc
        args.n = n
    else:
        raise TypeError("A must be a matrix or a callable")

    if M is None:
        args.pc_kind = _ffi.HM_PC_NONE
    elif isinstance(M, Jacobi):
        args.pc_kind = _ffi.HM_PC_JACOBI
        args.inv_diag = _ffi.ptr(M.inv_diag)
    elif isinstance(M, VCycle):
        if M.precision != "f64":
            raise ValueError("cg needs a float64 V-cycle")
        args.pc_kind = _ffi.HM_PC_VCYCLE
        args.vc = M.handle
    elif isinstance(M, CholeskyDecomposition):
        args.pc_kind = _ffi.HM_PC_ICHOL
        args.ichol = M.handle
    elif attached_csr(M) is not None:
        args.pc_kind = _ffi.HM_PC_CSR
        args.pc_csr = attached_csr(M).handle
    elif scipy.sparse.issparse(M) or isinstance(M, np.ndarray):
        pc = DeviceCSR(M)
        keep.append(pc)
        args.pc_kind = _ffi.HM_PC_CSR
        args.pc_csr = pc.handle
    elif callable(M):
        host = _HostApply(M, n)
        keep.append(host)
        args.pc_kind = _ffi.HM_PC_HOST
        args.m_cb = host.cfunc
    else:
        raise TypeError("M must be None, a matrix or a callable")

    x = empty(n)
    x0d = as_device(x0).reshape(-1) if x0 is not None else None
    if callback is not None:
        it = _HostIter(callback, A, b if not is_torch(b) else b.detach().cpu().numpy(), n)
        keep.append(it)
        args.callback = it.cfunc
    args.b = _ffi.ptr(bd)
    args.x0 = _ffi.ptr(x0d)
    args.x = _ffi.ptr(x)
    args.atol = float(atol)
    args.rtol = float(rtol)
    args.maxiter = int(maxiter)
    args.reorthogonalize = int(bool(reorthogonalize))
    args.check_every = int(check_every)
    stats = _ffi.CgStats()
    rc = L.hm_cg_solve(ctypes.byref(args), ctypes.byref(stats), _ffi.stream())
    for k in keep:
        err = getattr(k, "error", None)
        if err is not None:
            raise err
    if rc == _ffi.HM_ERR_NOCONV:
        raise ValueError("Conjugate gradient descent did not converge within %d iterations"
                         % maxiter)
    _ffi.check(rc)
    out = output_like(x, b)
    if return_stats:
        return out, {"iterations": stats.iterations, "norm_b": stats.norm_b,
                     "norm_r": stats.norm_r}
    return out
