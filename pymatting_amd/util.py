"""Trimap and image checks of the drop-in API.

Mirrors pymatting/util/util.py:439-578 (trimap_split, sanity_check_image):
same thresholds, warnings and ValueErrors. The device class map used by the
solvers is computed by the library (hm_trimap_classes).
"""
import ctypes
import warnings

import numpy as np
import torch

from pymatting_amd import _ffi
from pymatting_amd._device import as_device


def trimap_split(trimap, flatten=True, bg_threshold=0.1, fg_threshold=0.9):
    """util.py:439-525: boolean fg/bg/known/unknown masks of a trimap."""
    if isinstance(trimap, torch.Tensor):
        trimap = trimap.detach().cpu().numpy()
    if flatten:
        trimap = trimap.flatten()
    lo = trimap.min()
    hi = trimap.max()
    if lo < 0.0:
        warnings.warn("Trimap values should be in [0, 1], but trimap.min() is %s." % lo,
                      stacklevel=3)
    if hi > 1.0:
        # util.py:498 reports the minimum in this message; kept for parity.
        warnings.warn("Trimap values should be in [0, 1], but trimap.max() is %s." % lo,
                      stacklevel=3)
    if trimap.dtype not in [np.float32, np.float64]:
        warnings.warn("Unexpected trimap.dtype %s. Are you sure that you do not want to use "
                      "np.float32 or np.float64 instead?" % trimap.dtype, stacklevel=3)
    is_fg = trimap >= fg_threshold
    is_bg = trimap <= bg_threshold
    if is_bg.sum() == 0:
        raise ValueError("Trimap did not contain background values (values <= %f)" % bg_threshold)
    if is_fg.sum() == 0:
        raise ValueError("Trimap did not contain foreground values (values >= %f)" % fg_threshold)
    is_known = is_fg | is_bg
    return is_fg, is_bg, is_known, ~is_known


def sanity_check_image(image):
    """util.py:528-578: warnings for shape, range and dtype."""
    if isinstance(image, torch.Tensor):
        image = image.detach().cpu().numpy()
    if len(image.shape) != 3 or image.shape[2] != 3:
        warnings.warn("Expected RGB image of shape (?, ?, 3), but image.shape is %s."
                      % str(image.shape), stacklevel=3)
    if image.min() < 0.0:
        warnings.warn("Image values should be in [0, 1], but image.min() is %s." % image.min(),
                      stacklevel=3)
    if image.max() > 1.0:
        warnings.warn("Image values should be in [0, 1], but image.max() is %s." % image.max(),
                      stacklevel=3)
    if image.dtype not in [np.float32, np.float64]:
        warnings.warn("Unexpected image.dtype %s. Are you sure that you do not want to use "
                      "np.float32 or np.float64 instead?" % image.dtype, stacklevel=3)


def device_classes(trimap, fg_threshold=0.9, bg_threshold=0.1):
    """Device class map (0 unknown, 1 fg, 2 bg) and the device trimap.

    Raises ValueError like trimap_split when the trimap has no foreground or
    no background pixels.
    """
    t = as_device(trimap, torch.float64).reshape(-1)
    n = t.numel()
    cls = torch.empty(n, dtype=torch.uint8, device=t.device)
    counts = torch.zeros(2, dtype=torch.int64, device=t.device)
    L = _ffi.lib()
    _ffi.check(L.hm_trimap_classes(_ffi.ptr(t), n, fg_threshold, bg_threshold, _ffi.ptr(cls),
                                   _ffi.ptr(counts), _ffi.stream()))
    n_fg, n_bg = (int(v) for v in counts.cpu())
    if n_bg == 0:
        raise ValueError("Trimap did not contain background values (values <= %f)" % bg_threshold)
    if n_fg == 0:
        raise ValueError("Trimap did not contain foreground values (values >= %f)" % fg_threshold)
    return cls, t
