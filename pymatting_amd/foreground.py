"""Foreground estimation on the GPU.

Drop-ins for pymatting/foreground/estimate_foreground_ml.py:186-253 (the
multilevel method: the pyramid sizes follow the reference, :101-105, and
every resize and Gauss-Seidel sweep runs in the library, hm_fb_ml) and
pymatting/foreground/estimate_foreground_cf.py:8-130 (the closed-form
method: its sparse system assembled on the device, ichol-preconditioned
device PCG per channel).
"""
import ctypes

import numpy as np
import torch

from pymatting_amd import _ffi
from pymatting_amd._device import as_device, is_torch


def level_sizes(h0, w0):
    """estimate_foreground_ml.py:101-105: n_levels and round(h0 ** (i / n_levels))."""
    n_levels = int(np.ceil(np.log2(max(w0, h0))))
    if n_levels == 0:
        raise ZeroDivisionError("division by zero")
    hs = [round(h0 ** (i / n_levels)) for i in range(n_levels + 1)]
    ws = [round(w0 ** (i / n_levels)) for i in range(n_levels + 1)]
    return n_levels, hs, ws


def estimate_foreground_ml(image, alpha, regularization=1e-5, n_small_iterations=10,
                           n_big_iterations=2, small_size=32, return_background=False,
                           gradient_weight=1.0, precision="f32"):
    """estimate_foreground_ml.py:186-253 (Germer et al. 2020).

    precision "f32" computes like the reference (float32 arrays, float64 for
    the 2x2 solve); "f64" runs in double precision. Returns F (and B) as
    float32 (float64 for "f64") arrays, or torch tensors if `image` is one.
    """
    h0, w0 = int(alpha.shape[0]), int(alpha.shape[1])
    n_levels, hs, ws = level_sizes(h0, w0)
    dt = torch.float32 if precision == "f32" else torch.float64
    img = as_device(image, dt).reshape(h0, w0, 3)
    a = as_device(alpha, dt).reshape(h0, w0)
    F = torch.empty((h0, w0, 3), dtype=dt, device=img.device)
    B = torch.empty((h0, w0, 3), dtype=dt, device=img.device)
    lh = (ctypes.c_int * (n_levels + 1))(*hs)
    lw = (ctypes.c_int * (n_levels + 1))(*ws)
    _ffi.check(_ffi.lib().hm_fb_ml(
        _ffi.ptr(img), _ffi.ptr(a), h0, w0, n_levels, lh, lw, float(regularization),
        int(n_small_iterations), int(n_big_iterations), int(small_size), float(gradient_weight),
        _ffi.HM_F32 if precision == "f32" else _ffi.HM_F64, _ffi.ptr(F), _ffi.ptr(B),
        _ffi.stream()))
    if not is_torch(image):
        F, B = F.cpu().numpy(), B.cpu().numpy()
    if return_background:
        return F, B
    return F


def foreground_cf_system(alpha, regularization=1e-5, neighbors=((-1, 0), (1, 0), (0, -1), (0, 1))):
    """The 2n x 2n system matrix of estimate_foreground_cf.py:69-90, assembled on
    the device: (DeviceCSR, scipy.sparse.csr_matrix carrying it)."""
    from pymatting_amd.laplacian import _attach
    from pymatting_amd.sparse import DeviceCSR
    import scipy.sparse
    h, w = int(alpha.shape[0]), int(alpha.shape[1])
    n2 = 2 * h * w
    a = as_device(alpha, torch.float64).reshape(-1)
    k = len(neighbors)
    dx = (ctypes.c_int * max(k, 1))(*[int(d[0]) for d in neighbors])
    dy = (ctypes.c_int * max(k, 1))(*[int(d[1]) for d in neighbors])
    L = _ffi.lib()
    indptr = torch.empty(n2 + 1, dtype=torch.int64, device=a.device)
    nnz = ctypes.c_int64()
    _ffi.check(L.hm_fgcf_system_indptr(_ffi.ptr(a), h, w, float(regularization), k, dx, dy,
                                       _ffi.ptr(indptr), ctypes.byref(nnz), _ffi.stream()))
    indices = torch.empty(max(nnz.value, 1), dtype=torch.int64, device=a.device)
    values = torch.empty(max(nnz.value, 1), dtype=torch.float64, device=a.device)
    _ffi.check(L.hm_fgcf_system_fill(_ffi.ptr(a), h, w, float(regularization), k, dx, dy,
                                     _ffi.ptr(indptr), _ffi.ptr(indices), _ffi.ptr(values),
                                     _ffi.stream()))
    m = nnz.value
    dev = DeviceCSR.from_device(n2, indptr, indices[:m], values[:m])
    host = scipy.sparse.csr_matrix((values[:m].cpu().numpy(), indices[:m].cpu().numpy(),
                                    indptr.cpu().numpy()), shape=(n2, n2))
    host = _attach(host, dev)
    host.has_canonical_format = True
    return dev, host


def estimate_foreground_cf(image, alpha, regularization=1e-5, rtol=1e-5,
                           neighbors=[(-1, 0), (1, 0), (0, -1), (0, 1)], return_background=False,
                           foreground_guess=None, background_guess=None, ichol_kwargs={},
                           cg_kwargs={}):
    """estimate_foreground_cf.py:8-130: closed-form foreground (Levin et al. 2007).

    The 2n x 2n system A = U^T U + blockdiag(S, S) is assembled on the device
    with the reference's entries; it is preconditioned by ichol(A) (the
    factorisation runs on the host like the reference's, its triangular
    solves on the GPU) and solved by one device PCG per colour channel from
    the guesses (the image by default), then clipped to [0, 1]. Float64
    throughout; numpy in gives numpy out, torch tensors stay on the device.
    """
    from pymatting_amd.preconditioner import ichol
    from pymatting_amd.solver import cg
    h, w, d = (int(s) for s in image.shape)
    assert tuple(alpha.shape) == (h, w)
    n = h * w
    img = as_device(image, torch.float64).reshape(h, w, d)
    a = as_device(alpha, torch.float64).reshape(h, w)
    _, A = foreground_cf_system(a, regularization, neighbors)
    M = ichol(A, **ichol_kwargs)
    F0 = img if foreground_guess is None else as_device(foreground_guess).reshape(h, w, d)
    B0 = img if background_guess is None else as_device(background_guess).reshape(h, w, d)
    F = torch.empty((h, w, d), dtype=torch.float64, device=img.device)
    B = torch.empty_like(F)
    b = torch.empty(2 * n, dtype=torch.float64, device=img.device)
    x0 = torch.empty_like(b)
    L = _ffi.lib()
    for c in range(d):
        _ffi.check(L.hm_fgcf_channel(_ffi.ptr(img), _ffi.ptr(a), _ffi.ptr(F0), _ffi.ptr(B0), n, d,
                                     c, _ffi.ptr(b), _ffi.ptr(x0), _ffi.stream()))
        x = cg(A, b, x0=x0, M=M, rtol=rtol, **cg_kwargs)
        _ffi.check(L.hm_fgcf_store(_ffi.ptr(x), n, d, c, _ffi.ptr(F), _ffi.ptr(B),
                                   _ffi.stream()))
    if not is_torch(image):
        F, B = F.cpu().numpy(), B.cpu().numpy()
    if return_background:
        return F, B
    return F
