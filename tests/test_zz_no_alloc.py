"""The allocation-free foreground entry point (hm_fb_ml_noalloc) against
hm_fb_ml, through the C ABI: identical F and B on two calls in a row, the
second on a smaller image (the static workspace keeps the first call's
contents), in float32 and float64; and an image too large for the static
workspace is rejected before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(entry, image, alpha, prec):
    import torch
    from pymatting_amd import _ffi
    from pymatting_amd.foreground import level_sizes
    h0, w0 = alpha.shape
    n_levels, hs, ws = level_sizes(h0, w0)
    dt = torch.float32 if prec == _ffi.HM_F32 else torch.float64
    img = torch.from_numpy(image).to("cuda", dt).contiguous()
    a = torch.from_numpy(alpha).to("cuda", dt).contiguous()
    F = torch.full((h0, w0, 3), -1.0, dtype=dt, device="cuda")
    B = torch.full((h0, w0, 3), -1.0, dtype=dt, device="cuda")
    lh = (ctypes.c_int * (n_levels + 1))(*hs)
    lw = (ctypes.c_int * (n_levels + 1))(*ws)
    rc = getattr(_ffi.lib(), entry)(_ffi.ptr(img), _ffi.ptr(a), h0, w0, n_levels, lh, lw, 1e-5,
                                    10, 2, 32, 1.0, prec, _ffi.ptr(F), _ffi.ptr(B), _ffi.stream())
    _ffi.check(rc)
    torch.cuda.synchronize()
    return F.cpu().numpy(), B.cpu().numpy()


def _inputs(h, w, seed):
    rng = np.random.default_rng(seed)
    image = rng.random((h, w, 3))
    alpha = np.clip(rng.random((h, w)) * 1.4 - 0.2, 0, 1)
    return image, alpha


@pytest.mark.parametrize("prec", [1, 0])
def test_noalloc_matches_allocating_entry_point(prec):
    from pymatting_amd import _ffi
    # float32 at 700x900 needs ~48 MiB of the 64 MiB workspace; float64 half
    # the pixels.
    shapes = [(700, 900), (130, 201)] if prec == _ffi.HM_F32 else [(480, 640), (130, 201)]
    for k, (h, w) in enumerate(shapes):
        image, alpha = _inputs(h, w, 11 + k)
        print("hm_fb_ml_noalloc", "f32" if prec else "f64", (h, w), flush=True)
        F, B = _run("hm_fb_ml", image, alpha, prec)
        Fn, Bn = _run("hm_fb_ml_noalloc", image, alpha, prec)
        assert np.array_equal(F, Fn), (h, w, np.abs(F - Fn).max())
        assert np.array_equal(B, Bn), (h, w, np.abs(B - Bn).max())
        assert F.min() >= 0.0 and B.min() >= 0.0


def test_noalloc_rejects_images_beyond_the_workspace():
    from pymatting_amd import _ffi
    image, alpha = _inputs(1080, 1920, 3)
    with pytest.raises(ValueError, match="static workspace"):
        _run("hm_fb_ml_noalloc", image, alpha, _ffi.HM_F32)
