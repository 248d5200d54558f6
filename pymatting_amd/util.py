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


def _min_max(a):
    """(min, max) of a numpy array, or of a torch tensor reduced where it lives
    (a device tensor is not copied to the host: two scalars are read back)."""
    if isinstance(a, torch.Tensor):
        lo, hi = torch.aminmax(a.detach())
        return lo.item(), hi.item()
    return a.min(), a.max()


def _dtype_name(a):
    """(is float32/float64, dtype name as numpy spells it)."""
    if isinstance(a, torch.Tensor):
        return a.dtype in (torch.float32, torch.float64), str(a.dtype).replace("torch.", "")
    return a.dtype in [np.float32, np.float64], a.dtype


def check_trimap(trimap, stacklevel=4):
    """The warnings of util.py:490-507 (range and dtype of a trimap)."""
    lo, hi = _min_max(trimap)
    if lo < 0.0:
        warnings.warn("Trimap values should be in [0, 1], but trimap.min() is %s." % lo,
                      stacklevel=stacklevel)
    if hi > 1.0:
        # util.py:498 reports the minimum in this message; kept for parity.
        warnings.warn("Trimap values should be in [0, 1], but trimap.max() is %s." % lo,
                      stacklevel=stacklevel)
    ok, name = _dtype_name(trimap)
    if not ok:
        warnings.warn("Unexpected trimap.dtype %s. Are you sure that you do not want to use "
                      "np.float32 or np.float64 instead?" % name, stacklevel=stacklevel)


def trimap_split(trimap, flatten=True, bg_threshold=0.1, fg_threshold=0.9):
    """util.py:439-525: boolean fg/bg/known/unknown masks of a trimap."""
    if isinstance(trimap, torch.Tensor):
        trimap = trimap.detach().cpu().numpy()
    if flatten:
        trimap = trimap.flatten()
    check_trimap(trimap)
    is_fg = trimap >= fg_threshold
    is_bg = trimap <= bg_threshold
    if is_bg.sum() == 0:
        raise ValueError("Trimap did not contain background values (values <= %f)" % bg_threshold)
    if is_fg.sum() == 0:
        raise ValueError("Trimap did not contain foreground values (values >= %f)" % fg_threshold)
    is_known = is_fg | is_bg
    return is_fg, is_bg, is_known, ~is_known


def sanity_check_image(image):
    """util.py:528-578: warnings for shape, range and dtype (a device tensor is
    checked on the device)."""
    if len(image.shape) != 3 or image.shape[2] != 3:
        warnings.warn("Expected RGB image of shape (?, ?, 3), but image.shape is %s."
                      % str(tuple(image.shape)), stacklevel=3)
    lo, hi = _min_max(image)
    if lo < 0.0:
        warnings.warn("Image values should be in [0, 1], but image.min() is %s." % lo,
                      stacklevel=3)
    if hi > 1.0:
        warnings.warn("Image values should be in [0, 1], but image.max() is %s." % hi,
                      stacklevel=3)
    ok, name = _dtype_name(image)
    if not ok:
        warnings.warn("Unexpected image.dtype %s. Are you sure that you do not want to use "
                      "np.float32 or np.float64 instead?" % name, stacklevel=3)


def _check_counts(n_fg, n_bg, fg_threshold, bg_threshold):
    if n_bg == 0:
        raise ValueError("Trimap did not contain background values (values <= %f)" % bg_threshold)
    if n_fg == 0:
        raise ValueError("Trimap did not contain foreground values (values >= %f)" % fg_threshold)


def device_classes(trimap, fg_threshold=0.9, bg_threshold=0.1, deferred=False):
    """Device class map (0 unknown, 1 fg, 2 bg) and the device trimap.

    Raises ValueError like trimap_split when the trimap has no foreground or
    no background pixels. With deferred=True it returns (classes, trimap,
    check) instead: the counts travel to pinned host memory behind the
    stream's work and check() raises those errors once they have arrived, so
    that device work enqueued before check() overlaps the wait.
    """
    t = as_device(trimap, torch.float64).reshape(-1)
    n = t.numel()
    cls = torch.empty(n, dtype=torch.uint8, device=t.device)
    counts = torch.zeros(2, dtype=torch.int64, device=t.device)
    L = _ffi.lib()
    _ffi.check(L.hm_trimap_classes(_ffi.ptr(t), n, fg_threshold, bg_threshold, _ffi.ptr(cls),
                                   _ffi.ptr(counts), _ffi.stream()))
    if deferred:
        host = torch.empty(2, dtype=torch.int64, pin_memory=True)
        host.copy_(counts, non_blocking=True)
        arrived = torch.cuda.Event()
        arrived.record()

        def check():
            arrived.synchronize()
            _check_counts(int(host[0]), int(host[1]), fg_threshold, bg_threshold)
        return cls, t, check
    n_fg, n_bg = (int(v) for v in counts.cpu())
    _check_counts(n_fg, n_bg, fg_threshold, bg_threshold)
    return cls, t
