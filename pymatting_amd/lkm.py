"""Large-kernel matting on the GPU.

Drop-ins for pymatting/laplacian/lkm_laplacian.py:6-76 (lkm_laplacian) and
pymatting/alpha/estimate_alpha_lkm.py:8-64 (estimate_alpha_lkm). The
Laplacian is applied matrix-free with box filters in the library
(``hm_lkm_*``); ``cg`` runs it inside its device loop.
"""
import ctypes

import torch

from pymatting_amd import _ffi
from pymatting_amd._device import as_device, empty, is_torch, output_like
from pymatting_amd.util import check_trimap, device_classes, sanity_check_image


class LkmOperator:
    """The large-kernel Laplacian of an (h, w, 3) image as a device operator.

    Calling it with a vector of h*w values (numpy, or a torch tensor that
    stays on the device) returns L p flattened, like the reference's
    L_matvec. With ``classes`` (a device class map) and ``lambda_value`` it
    is L + lambda diag(known) instead.
    """

    def __init__(self, image, epsilon=1e-7, radius=10, classes=None, lambda_value=0.0):
        if image.ndim != 3 or image.shape[2] != 3:
            raise ValueError("image must have shape (h, w, 3), got %s" % (tuple(image.shape),))
        self.h, self.w = int(image.shape[0]), int(image.shape[1])
        self.n = self.h * self.w
        self.shape = (self.n, self.n)
        self.grid_shape = (self.h, self.w)
        self.radius = int(radius)
        # lkm_laplacian.py:33: the image is used as float64.
        self.image = as_device(image, torch.float64).reshape(self.h, self.w, 3).contiguous()
        handle = ctypes.c_void_p()
        _ffi.check(_ffi.lib().hm_lkm_create(_ffi.ptr(self.image), self.h, self.w, float(epsilon),
                                            self.radius, _ffi.ptr(classes), float(lambda_value),
                                            _ffi.stream(), ctypes.byref(handle)))
        self.handle = handle

    def __del__(self):
        _ffi.release(self, "hm_lkm_destroy")

    def apply(self, x, out=None):
        out = out if out is not None else empty(self.n)
        _ffi.check(_ffi.lib().hm_lkm_apply(self.handle, _ffi.ptr(x), _ffi.ptr(out),
                                           _ffi.stream()))
        return out

    def __call__(self, p):
        if _ffi.numel(p) != self.n:
            raise ValueError("dimension mismatch: operator of size %d, vector of size %d"
                             % (self.n, _ffi.numel(p)))
        pd = as_device(p, torch.float64).reshape(-1)
        return output_like(self.apply(pd), p)

    dot = __call__

    def diagonal_device(self, inverse=False):
        d = empty(self.n)
        _ffi.check(_ffi.lib().hm_lkm_diagonal(self.handle, int(bool(inverse)), _ffi.ptr(d),
                                              _ffi.stream()))
        return d

    def rhs(self):
        """lambda * is_fg (an operator made with a class map)."""
        b = empty(self.n)
        _ffi.check(_ffi.lib().hm_lkm_rhs(self.handle, _ffi.ptr(b), _ffi.stream()))
        return b

    def diagonal(self):
        return self.diagonal_device().cpu().numpy()


def lkm_laplacian(image, epsilon=1e-7, radius=10, return_diagonal=True):
    """lkm_laplacian.py:6-76: the large-kernel matting Laplacian (He et al. 2010).

    Returns the operator L (a callable p -> L p, applied on the GPU) and,
    with return_diagonal, its diagonal as a numpy array (a device tensor for
    a torch image).
    """
    op = LkmOperator(image, epsilon, radius)
    if return_diagonal:
        d = op.diagonal_device()
        return op, (d if is_torch(image) else d.cpu().numpy())
    return op


def estimate_alpha_lkm(image, trimap, laplacian_kwargs={}, cg_kwargs={}):
    """estimate_alpha_lkm.py:8-64: alpha from the large-kernel Laplacian.

    Solves (L + lambda diag(known)) x = lambda is_fg with lambda = 100 by
    Jacobi-preconditioned CG (inverse diagonal 1 / (diag(L) + lambda known)),
    every step on the device, and clips to [0, 1].
    """
    from pymatting_amd.preconditioner import Jacobi
    from pymatting_amd.solver import cg
    sanity_check_image(image)
    check_trimap(trimap, stacklevel=3)
    classes, tdev = device_classes(trimap)
    lambda_value = 100.0
    op = LkmOperator(image, classes=classes, lambda_value=lambda_value, **laplacian_kwargs)
    M = Jacobi.from_inverse_diagonal(op.diagonal_device(inverse=True))
    x = cg(op, op.rhs(), M=M, **cg_kwargs)
    alpha = empty(op.n)
    _ffi.check(_ffi.lib().hm_alpha_finalize(_ffi.ptr(x), None, None, _ffi.ptr(alpha), op.n,
                                            _ffi.stream()))
    alpha = alpha.reshape(tuple(trimap.shape))
    return alpha if is_torch(image) else alpha.cpu().numpy()
