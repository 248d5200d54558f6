"""General sparse matrices on the device.

``DeviceCSR`` holds a float64 CSR matrix in device memory (cg on scipy or
numpy matrices, cg.py:68-71). ``csr_plus_diagonal`` forms the canonical CSR
of A + diag(c), the matrix make_linear_system builds from an arbitrary
Laplacian (pymatting/laplacian/laplacian.py:49-55), on the device.
"""
import ctypes

import numpy as np
import scipy.sparse
import torch

from pymatting_amd import _ffi
from pymatting_amd._device import as_device, empty


class DeviceCSR:
    """A general sparse (or dense) matrix as float64 CSR in device memory.

    Built from a host matrix (uploaded), or from device arrays
    (``from_device``). ``host()`` returns the scipy form.
    """

    def __init__(self, A):
        A = scipy.sparse.csr_matrix(A, dtype=np.float64)
        if A.shape[0] != A.shape[1]:
            raise ValueError("matrix must be square, got %s" % (A.shape,))
        self._make(A.shape[0], as_device(A.indptr.astype(np.int64), torch.int64),
                   as_device(A.indices.astype(np.int64), torch.int64), as_device(A.data))

    @classmethod
    def from_device(cls, n, indptr, indices, values):
        self = cls.__new__(cls)
        self._make(int(n), indptr, indices, values)
        return self

    def _make(self, n, indptr, indices, values):
        self.n = n
        self.shape = (n, n)
        self.nnz = int(values.numel())
        self._indptr, self._indices, self._values = indptr, indices, values
        handle = ctypes.c_void_p()
        _ffi.check(_ffi.lib().hm_csr_create(n, self.nnz, _ffi.ptr(indptr), _ffi.ptr(indices),
                                            _ffi.ptr(values), ctypes.byref(handle)))
        self.handle = handle

    def __del__(self):
        _ffi.release(self, "hm_csr_destroy")

    def host(self):
        """The matrix as scipy.sparse.csr_matrix (copied to the host)."""
        return scipy.sparse.csr_matrix(
            (self._values.cpu().numpy(), self._indices.cpu().numpy(),
             self._indptr.cpu().numpy()), shape=self.shape)


def _device_csr_arrays(L):
    """(n, indptr, indices, values) of L in device memory: an attached
    device CSR is used as it is, any other matrix is converted to CSR on the
    host (a format change, as scipy's own L + C makes) and uploaded."""
    from pymatting_amd.laplacian import attached_csr
    dev = attached_csr(L)
    if dev is not None:
        return dev.n, dev._indptr, dev._indices, dev._values
    A = L.tocsr() if scipy.sparse.issparse(L) else scipy.sparse.csr_matrix(L)
    A = scipy.sparse.csr_matrix(A, dtype=np.float64)
    if A.shape[0] != A.shape[1]:
        raise ValueError("matrix must be square, got %s" % (A.shape,))
    return (A.shape[0], as_device(A.indptr.astype(np.int64), torch.int64),
            as_device(A.indices.astype(np.int64), torch.int64), as_device(A.data))


def csr_plus_diagonal(L, c):
    """Canonical CSR of L + diag(c) computed on the device: equal columns of a
    row summed in stored order, c_i added to the diagonal, zero entries
    dropped, columns ascending (what scipy's (L + diags(c)).tocsr() and
    sum_duplicates() leave). Returns (DeviceCSR, scipy.sparse.csr_matrix)."""
    n, indptr, indices, values = _device_csr_arrays(L)
    c = as_device(np.asarray(c, dtype=np.float64).reshape(-1) if not torch.is_tensor(c) else c)
    if c.numel() != n:
        raise ValueError("dimension mismatch: matrix of size %d, diagonal of size %d"
                         % (n, c.numel()))
    L_ = _ffi.lib()
    out_ptr = empty(n + 1, torch.int64)
    nnz = ctypes.c_int64()
    _ffi.check(L_.hm_csr_add_diagonal_indptr(n, _ffi.ptr(indptr), _ffi.ptr(indices),
                                             _ffi.ptr(values), _ffi.ptr(c), _ffi.ptr(out_ptr),
                                             ctypes.byref(nnz), _ffi.stream()))
    out_idx = empty(max(nnz.value, 1), torch.int64)
    out_val = empty(max(nnz.value, 1))
    _ffi.check(L_.hm_csr_add_diagonal_fill(n, _ffi.ptr(indptr), _ffi.ptr(indices),
                                           _ffi.ptr(values), _ffi.ptr(c), _ffi.ptr(out_ptr),
                                           _ffi.ptr(out_idx), _ffi.ptr(out_val), _ffi.stream()))
    k = nnz.value
    dev = DeviceCSR.from_device(n, out_ptr, out_idx[:k], out_val[:k])
    host = scipy.sparse.csr_matrix((out_val[:k].cpu().numpy(), out_idx[:k].cpu().numpy(),
                                    out_ptr.cpu().numpy()), shape=(n, n))
    host.has_canonical_format = True
    return dev, host
