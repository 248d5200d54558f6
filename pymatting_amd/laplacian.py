"""Closed-form matting Laplacian on the GPU.

Drop-in for pymatting/laplacian/cf_laplacian.py:132-177 (cf_laplacian) and
pymatting/laplacian/laplacian.py:5-60 (make_linear_system).

``cf_laplacian`` returns a ``scipy.sparse.csr_matrix`` with the reference's
exact layout and entries (assembled by the library, bit-identical to the
reference's serial loop); the device-resident matrix-free operator rides along
as its ``.operator`` attribute, and ``cg``/``vcycle``/``jacobi`` use that
operator when they are handed the matrix. ``CfOperator`` is the operator
itself for callers that keep everything on the device.
"""
import ctypes

import numpy as np
import scipy.sparse
import torch
import xxhash

from pymatting_amd import _ffi
from pymatting_amd._device import as_device, empty, is_torch, output_like
from pymatting_amd.util import device_classes, trimap_split

MODES = {"laplacian": _ffi.HM_OP_LAPLACIAN, "lambda": _ffi.HM_OP_LAMBDA,
         "masked": _ffi.HM_OP_MASKED}


class CfOperator:
    """Device-resident closed-form operator on an (h, w) image.

    mode "laplacian": L (windows whose pixels are all known are skipped when a
    known mask is given, cf_laplacian.py:34-35); "lambda": L + lambda
    diag(known) (make_linear_system); "masked": [L_UU 0; 0 I], the hard
    constraints of estimate_alpha_cf.py:57-74 on the full grid.
    """

    def __init__(self, image, epsilon=1e-7, radius=1, mode="laplacian", classes=None,
                 trimap=None, lambda_value=0.0):
        if not is_torch(image):
            image = np.asarray(image)
            if image.dtype != np.float64:
                # cf_laplacian.py:5 only accepts float64 images.
                raise TypeError("image must be float64, got %s" % image.dtype)
        if image.ndim != 3 or image.shape[2] != 3:
            raise ValueError("image must have shape (h, w, 3), got %s" % (tuple(image.shape),))
        self.h, self.w = int(image.shape[0]), int(image.shape[1])
        self.n = self.h * self.w
        self.shape = (self.n, self.n)
        self.grid_shape = (self.h, self.w)
        self.epsilon = float(epsilon)
        self.radius = int(radius)
        self.mode = mode
        self.lambda_value = float(lambda_value)
        self.image = as_device(image, torch.float64)
        self.classes = classes
        self.trimap = trimap
        handle = ctypes.c_void_p()
        _ffi.check(_ffi.lib().hm_cf_create(
            _ffi.ptr(self.image), self.h, self.w, self.epsilon, self.radius,
            _ffi.ptr(classes) if classes is not None else None, self.lambda_value, MODES[mode],
            _ffi.stream(), ctypes.byref(handle)))
        self.handle = handle

    def __del__(self):
        _ffi.release(self, "hm_cf_destroy")

    # -- matrix-like interface -------------------------------------------
    def apply(self, x, out=None, precision="f64"):
        """A x for a device vector x (float64, or float32 with precision="f32")."""
        dt = torch.float64 if precision == "f64" else torch.float32
        out = out if out is not None else empty(self.n, dt)
        _ffi.check(_ffi.lib().hm_cf_apply(self.handle, _ffi.HM_F64 if precision == "f64"
                                          else _ffi.HM_F32, _ffi.ptr(x), _ffi.ptr(out),
                                          _ffi.stream()))
        return out

    def set_windows(self, recompute):
        """recompute=True: float32 passes recompute every window's mean and
        inverse covariance from the image tile they stage (float64
        arithmetic) instead of reading the stored float32 values."""
        _ffi.check(_ffi.lib().hm_cf_set_windows(self.handle, int(bool(recompute))))
        return self

    def dot(self, x):
        xd = as_device(x, torch.float64).reshape(-1)
        if xd.numel() != self.n:
            raise ValueError("dimension mismatch: operator of size %d, vector of size %d"
                             % (self.n, xd.numel()))
        return output_like(self.apply(xd), x)

    __call__ = dot

    def __matmul__(self, x):
        return self.dot(x)

    def diagonal_device(self):
        d = empty(self.n)
        _ffi.check(_ffi.lib().hm_cf_diagonal(self.handle, _ffi.ptr(d), _ffi.stream()))
        return d

    def diagonal(self):
        return self.diagonal_device().cpu().numpy()

    def rhs(self):
        """lambda * is_fg (lambda mode) or -(R m) on the unknown pixels (masked mode)."""
        b = empty(self.n)
        _ffi.check(_ffi.lib().hm_cf_rhs(self.handle, _ffi.ptr(b), _ffi.stream()))
        return b

    def reference_csr_device(self):
        """(indptr, indices, values) of the reference's 25-slot CSR layout
        (cf_laplacian.py:169-175) as device tensors."""
        n = self.n
        values = empty(25 * n)
        indices = empty(25 * n, torch.int64)
        indptr = empty(n + 1, torch.int64)
        _ffi.check(_ffi.lib().hm_cf_assemble_csr(self.handle, _ffi.ptr(values), _ffi.ptr(indices),
                                                 _ffi.ptr(indptr), _ffi.stream()))
        return indptr, indices, values

    def reference_csr(self):
        """The reference's 25-slot CSR layout (cf_laplacian.py:169-175)."""
        indptr, indices, values = self.reference_csr_device()
        return (values.cpu().numpy(), indices.cpu().numpy(), indptr.cpu().numpy())

    def canonical_csr(self):
        """Canonical CSR of the operator as scipy leaves (L + C) (laplacian.py:49-55)."""
        n = self.n
        indptr = empty(n + 1, torch.int64)
        nnz = ctypes.c_int64()
        L = _ffi.lib()
        _ffi.check(L.hm_cf_system_indptr(self.handle, _ffi.ptr(indptr), ctypes.byref(nnz),
                                         _ffi.stream()))
        indices = empty(max(nnz.value, 1), torch.int64)
        values = empty(max(nnz.value, 1))
        _ffi.check(L.hm_cf_system_fill(self.handle, _ffi.ptr(indptr), _ffi.ptr(indices),
                                       _ffi.ptr(values), _ffi.stream()))
        k = nnz.value
        return (values[:k].cpu().numpy(), indices[:k].cpu().numpy(), indptr.cpu().numpy())

    def derived(self, mode, classes=None, trimap=None, lambda_value=0.0):
        """Another operator on the same device image."""
        return CfOperator(self.image, self.epsilon, self.radius, mode, classes, trimap,
                          lambda_value)


class DiaOperator:
    """Stored 25-point stencil on an (h, w) grid (slot-major coefficients [25][n])."""

    def __init__(self, coef, h, w):
        self.h, self.w = int(h), int(w)
        self.n = self.h * self.w
        self.shape = (self.n, self.n)
        self.grid_shape = (self.h, self.w)
        self._coef = as_device(coef, torch.float64).reshape(25, self.n).contiguous()
        handle = ctypes.c_void_p()
        _ffi.check(_ffi.lib().hm_dia_create(_ffi.ptr(self._coef), self.h, self.w, _ffi.stream(),
                                            ctypes.byref(handle)))
        self.handle = handle

    @classmethod
    def from_cf(cls, op):
        """Stored form of a CfOperator, assembled exactly on the device."""
        self = cls.__new__(cls)
        self.h, self.w = op.h, op.w
        self.n = op.n
        self.shape = op.shape
        self.grid_shape = op.grid_shape
        handle = ctypes.c_void_p()
        _ffi.check(_ffi.lib().hm_dia_from_cf(op.handle, _ffi.stream(), ctypes.byref(handle)))
        self.handle = handle
        self._coef = None
        return self

    @property
    def coef(self):
        if getattr(self, "_coef", None) is None:
            c = empty(25 * self.n)
            _ffi.check(_ffi.lib().hm_dia_coefficients(self.handle, _ffi.ptr(c), _ffi.stream()))
            self._coef = c.reshape(25, self.n)
        return self._coef

    @classmethod
    def from_sparse(cls, A, shape):
        """Grid stencil form of a sparse matrix whose couplings reach at most 2 pixels."""
        h, w = int(shape[0]), int(shape[1])
        n = h * w
        A = scipy.sparse.coo_matrix(A)
        if A.shape != (n, n):
            raise ValueError("dimension mismatch: matrix %s for grid shape %s" % (A.shape, shape))
        dy = A.col // w - A.row // w
        dx = A.col % w - A.row % w
        if A.nnz and (np.abs(dy).max() > 2 or np.abs(dx).max() > 2):
            raise ValueError("matrix is not a 5x5 grid stencil on shape %s" % (shape,))
        coef = np.zeros((25, n))
        np.add.at(coef, ((dy + 2) * 5 + (dx + 2), A.row), A.data)
        S = scipy.sparse.csr_matrix(A)
        asym = abs(S - S.T).max() if S.nnz else 0.0
        if asym > 1e-10 * (abs(S).max() if S.nnz else 1.0):
            raise ValueError("the V-cycle needs a symmetric matrix (max |A - A^T| = %g)" % asym)
        return cls(coef, h, w)

    def __del__(self):
        _ffi.release(self, "hm_dia_destroy")

    def apply(self, x, out=None):
        out = out if out is not None else empty(self.n)
        _ffi.check(_ffi.lib().hm_dia_apply(self.handle, _ffi.HM_F64, _ffi.ptr(x), _ffi.ptr(out),
                                           _ffi.stream()))
        return out

    def dot(self, x):
        return output_like(self.apply(as_device(x).reshape(-1)), x)

    __call__ = dot

    def diagonal_device(self):
        return self.coef[12].clone()

    def diagonal(self):
        return self.coef[12].cpu().numpy()


class CfLaplacianMatrix(scipy.sparse.csr_matrix):
    """scipy CSR matrix carrying the device form it was assembled from (a
    stencil operator, or a device copy of the CSR itself).

    The operator stands for the matrix only while the matrix is unchanged:
    a fingerprint of its arrays (their addresses and sizes, and a hash of the
    values) is taken when the operator is attached, and ``device_operator``
    drops the operator once an in-place edit (``L *= 2``, ``L.data[k] = v``,
    ``setdiag``) has changed it. The edited matrix is then solved as the CSR
    matrix it now is.
    """

    operator = None
    _fingerprint = None

    def matches_operator(self):
        return self.operator is not None and self._fingerprint == _fingerprint(self)


def _fingerprint(csr):
    arrays = (csr.data, csr.indices, csr.indptr)
    return (tuple(a.__array_interface__["data"][0] for a in arrays),
            tuple(a.size for a in arrays), tuple(csr.shape),
            xxhash.xxh3_64_intdigest(np.ascontiguousarray(csr.data)))


def _attach(csr, operator):
    out = CfLaplacianMatrix(csr)
    out.operator = operator
    out._fingerprint = _fingerprint(out)
    return out


def cf_laplacian(image, epsilon=1e-7, radius=1, is_known=None):
    """cf_laplacian.py:132-177: the closed-form matting Laplacian as CSR.

    Entries and layout are the reference's: (4r+1)^2 slots per row (25 for
    the default radius 1), out-of-image slots with column 0 and value 0.
    Radius 1 carries the matrix-free device operator (``.operator``); other
    radii are assembled by the general-radius kernel and carry their device
    CSR, which ``cg`` and ``make_linear_system`` use without another upload.
    """
    if not is_torch(image) and np.asarray(image).ndim == 3:
        assert np.asarray(image).shape[2] == 3  # cf_laplacian.py:8
    h, w = int(image.shape[0]), int(image.shape[1])
    n = h * w
    known = None
    if is_known is not None:
        known = np.asarray(is_known.cpu() if is_torch(is_known) else is_known).reshape(n)
    if int(radius) != 1:
        return _cf_laplacian_radius(image, float(epsilon), int(radius), known)
    classes = None if known is None else as_device(known.astype(np.uint8) * 2, torch.uint8)
    op = CfOperator(image, epsilon, radius, "laplacian", classes)
    values, indices, indptr = op.reference_csr()
    return _attach(scipy.sparse.csr_matrix((values, indices, indptr), (n, n)), op)


def _cf_laplacian_radius(image, epsilon, radius, known):
    from pymatting_amd.sparse import DeviceCSR
    if not is_torch(image):
        image = np.asarray(image)
        if image.dtype != np.float64:
            raise TypeError("image must be float64, got %s" % image.dtype)
    if radius < 1:
        raise ValueError("radius must be at least 1, got %d" % radius)
    h, w = int(image.shape[0]), int(image.shape[1])
    n = h * w
    slots = (4 * radius + 1) ** 2
    img = as_device(image, torch.float64).reshape(h, w, 3).contiguous()
    kd = None if known is None else as_device(known.astype(np.uint8), torch.uint8)
    values = empty(n * slots)
    indices = empty(n * slots, torch.int64)
    indptr = empty(n + 1, torch.int64)
    _ffi.check(_ffi.lib().hm_cf_laplacian_csr(_ffi.ptr(img), h, w, epsilon, radius, _ffi.ptr(kd),
                                              _ffi.ptr(values), _ffi.ptr(indices),
                                              _ffi.ptr(indptr), _ffi.stream()))
    dev = DeviceCSR.from_device(n, indptr, indices, values)
    host = scipy.sparse.csr_matrix((values.cpu().numpy(), indices.cpu().numpy(),
                                    indptr.cpu().numpy()), (n, n))
    return _attach(host, dev)


def make_linear_system(L, trimap, lambda_value=100.0, return_c=False):
    """laplacian.py:5-60: A = L + lambda diag(known), b = lambda is_fg.

    L must come from ``cf_laplacian`` (or be a ``CfOperator``): A is assembled
    on the GPU in the canonical CSR scipy produces, and carries the
    matrix-free lambda operator as ``A.operator``.
    """
    is_fg, is_bg, is_known, is_unknown = trimap_split(trimap)
    c = lambda_value * is_known
    b = lambda_value * is_fg
    base = device_operator(L)
    if not isinstance(base, CfOperator) or base.classes is not None:
        # Any other Laplacian (another radius, an edited or user-built
        # matrix, one built with is_known): the canonical CSR of L + diag(c)
        # formed on the device.
        from pymatting_amd.sparse import csr_plus_diagonal
        if isinstance(L, (CfOperator, DiaOperator)):
            raise TypeError("make_linear_system needs a matrix or the operator returned by "
                            "cf_laplacian")
        dev, A = csr_plus_diagonal(L, c)
        A = _attach(A, dev)
        A.has_canonical_format = True
        if return_c:
            return A, b, c
        return A, b
    classes, tdev = device_classes(trimap)
    op = base.derived("lambda", classes, tdev, lambda_value)
    if isinstance(L, CfOperator):
        A = op
    else:
        values, indices, indptr = op.canonical_csr()
        A = _attach(scipy.sparse.csr_matrix((values, indices, indptr), (op.n, op.n)), op)
        A.has_canonical_format = True
    if return_c:
        return A, b, c
    return A, b


def device_operator(A):
    """The stencil operator behind A (CfOperator/DiaOperator), or None: a
    matrix from cf_laplacian/make_linear_system yields its operator unless it
    has been edited since."""
    if isinstance(A, (CfOperator, DiaOperator)):
        return A
    if isinstance(A, CfLaplacianMatrix) and isinstance(A.operator, (CfOperator, DiaOperator)):
        return A.operator if A.matches_operator() else None
    return None


def attached_csr(A):
    """The device CSR copy a matrix from this package carries (another radius
    of cf_laplacian, make_linear_system of a general Laplacian, the reduced
    system of estimate_alpha_cf), or None (also once the matrix is edited)."""
    from pymatting_amd.sparse import DeviceCSR
    if isinstance(A, CfLaplacianMatrix) and isinstance(A.operator, DeviceCSR):
        return A.operator if A.matches_operator() else None
    return None
