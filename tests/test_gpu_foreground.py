"""GPU parity of estimate_foreground_ml against the oracle.

The GPU sums the mean colours in float64 (the reference sums in float32,
serially); with the oracle in the same mean mode the float32 result must be
bit-identical, since the sweep kernel reproduces raster Gauss-Seidel order
and the reference's float32/float64 arithmetic exactly.
"""
import numpy as np
import pytest

from conftest import load_golden
from oracle.cf import trimap_split
from oracle.foreground import estimate_foreground_ml as ofg

pytestmark = pytest.mark.gpu

SHAPES = [(5, 7), (31, 33), (64, 40), (65, 40), (130, 201), (200, 129)]


@pytest.mark.parametrize("shape", SHAPES)
def test_bit_exact_float32(shape):
    import pymatting_amd as pm
    rng = np.random.default_rng(shape[0] * 1000 + shape[1])
    h, w = shape
    image = rng.random((h, w, 3))
    alpha = np.clip(rng.random((h, w)) * 1.4 - 0.2, 0, 1)
    print("shape", shape, flush=True)
    F, B = pm.estimate_foreground_ml(image, alpha, return_background=True)
    Fr, Br = ofg(image, alpha, return_background=True, mean_mode="f64")
    assert F.dtype == np.float32
    assert np.array_equal(F, Fr), np.abs(F - Fr).max()
    assert np.array_equal(B, Br), np.abs(B - Br).max()


@pytest.mark.parametrize("shape", [(1, 37), (37, 1), (2, 150), (150, 2), (3, 300), (1, 2),
                                   (70, 3)])
def test_bit_exact_float32_thin(shape):
    """Images one to three pixels thin: every level of the pyramid is a strip
    of one or two rows or columns (the sweep's strips and halos at their
    narrowest)."""
    import pymatting_amd as pm
    rng = np.random.default_rng(shape[0] * 1000 + shape[1])
    image = rng.random(shape + (3,))
    alpha = np.clip(rng.random(shape) * 1.4 - 0.2, 0, 1)
    print("shape", shape, flush=True)
    F, B = pm.estimate_foreground_ml(image, alpha, return_background=True)
    Fr, Br = ofg(image, alpha, return_background=True, mean_mode="f64")
    assert np.array_equal(F, Fr), np.abs(F - Fr).max()
    assert np.array_equal(B, Br), np.abs(B - Br).max()
    F64 = pm.estimate_foreground_ml(image, alpha, precision="f64")
    assert np.abs(F64 - ofg(image, alpha, precision="f64")).max() < 1e-12


@pytest.mark.parametrize("shape", [(9, 13), (130, 201)])
def test_float64(shape):
    import pymatting_amd as pm
    rng = np.random.default_rng(7)
    image = rng.random(shape + (3,))
    alpha = np.clip(rng.random(shape) * 1.4 - 0.2, 0, 1)
    F = pm.estimate_foreground_ml(image, alpha, precision="f64")
    Fr = ofg(image, alpha, precision="f64")
    assert np.abs(F - Fr).max() < 1e-12


def test_float32_within_1e4_of_float64_at_1080p():
    """The float32 form (the reference's typing) stays within 1e-4 of the
    float64 form on a 1920x1080 synthetic composite (F, B in [0, 1])."""
    import torch
    import pymatting_amd as pm
    from pymatting_amd import synthetic
    image, _, alpha = synthetic.make_image(1080, 1920, 2000, strands=40, erode=10)
    img, a = torch.from_numpy(image).cuda(), torch.from_numpy(alpha).cuda()
    F32, B32 = pm.estimate_foreground_ml(img.float(), a.float(), return_background=True)
    F64, B64 = pm.estimate_foreground_ml(img, a, return_background=True, precision="f64")
    assert float((F32.double() - F64).abs().max()) <= 1e-4
    assert float((B32.double() - B64).abs().max()) <= 1e-4


def test_lemur_known_answer_and_options():
    import pymatting_amd as pm
    d = load_golden("lemur_fg_0p1")
    F, B = pm.estimate_foreground_ml(d["image"], d["alpha"], return_background=True)
    unknown = trimap_split(d["trimap"], flatten=False)[3]
    diff = np.abs(F - d["expected_foreground"])
    assert np.mean((d["alpha"][:, :, None] * diff)[unknown]) <= 0.022
    kw = dict(regularization=1e-3, n_small_iterations=3, n_big_iterations=4, small_size=20,
              gradient_weight=0.5)
    F2 = pm.estimate_foreground_ml(d["image"], d["alpha"], **kw)
    F2r = ofg(d["image"], d["alpha"], mean_mode="f64", **kw)
    assert np.array_equal(F2, F2r)


def test_many_sweeps_both_level_kernels():
    """Levels that fit in LDS run all their sweeps in one workgroup, up to
    4096 sweeps per launch; more go to the strip kernel, one launch per
    sweep. Both are bit-exact against the oracle."""
    import pymatting_amd as pm
    rng = np.random.default_rng(12)
    image = rng.random((12, 9, 3))
    alpha = np.clip(rng.random((12, 9)) * 1.4 - 0.2, 0, 1)
    for n_small in (37, 4100):
        print("n_small_iterations", n_small, flush=True)
        F = pm.estimate_foreground_ml(image, alpha, n_small_iterations=n_small)
        Fr = ofg(image, alpha, n_small_iterations=n_small, mean_mode="f64")
        assert np.array_equal(F, Fr), (n_small, np.abs(F - Fr).max())


def test_single_pixel_raises():
    import pymatting_amd as pm
    with pytest.raises(ZeroDivisionError):
        pm.estimate_foreground_ml(np.zeros((1, 1, 3)), np.zeros((1, 1)))
