"""Multilevel foreground estimation on the GPU.

Drop-in for pymatting/foreground/estimate_foreground_ml.py:186-253. The
pyramid sizes follow the reference (:101-105); every resize and
Gauss-Seidel sweep runs in the library (hm_fb_ml).
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
