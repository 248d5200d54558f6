"""Image files and array helpers (pymatting/util/util.py:211-287, 581-643)
on the host: save/load round trips, the float-to-uint8 rule of save_image,
stack_images and blend; and, where the reference's data directory is
present, load_image reproduces the committed fixtures (which were decoded
with the reference tests' scale and filter choices)."""
import os

import numpy as np
import pytest

import pymatting_amd as pm
from conftest import load_golden

DATA = "/root/reference/data"


def test_round_trip_and_float_rule(tmp_path):
    rng = np.random.default_rng(0)
    rgb = rng.integers(0, 256, (7, 9, 3), dtype=np.uint8)
    path = str(tmp_path / "sub" / "a.png")
    pm.save_image(path, rgb)
    assert np.array_equal(pm.load_image(path), rgb / 255.0)
    assert np.array_equal(pm.load_image(path, "GRAY").shape, (7, 9))
    f = rng.random((5, 4)) * 1.2 - 0.1
    pm.save_image(str(tmp_path / "f.png"), f)
    expected = np.clip(f * 255, 0, 255).astype(np.uint8) / 255.0
    assert np.array_equal(pm.load_image(str(tmp_path / "f.png")), expected)
    half = pm.load_image(path, "RGB", 0.5, "nearest")
    assert half.shape == (3, 4, 3)
    assert pm.load_image(path, "RGB", (5, 2), "bilinear").shape == (2, 5, 3)
    with pytest.raises(AssertionError):
        pm.save_image(str(tmp_path / "i.png"), np.zeros((2, 2), np.int64))


def test_stack_and_blend():
    rng = np.random.default_rng(1)
    F, B, a = rng.random((4, 5, 3)), rng.random((4, 5, 3)), rng.random((4, 5))
    S = pm.stack_images(F, a)
    assert S.shape == (4, 5, 4) and np.array_equal(S[:, :, 3], a)
    assert np.array_equal(pm.blend(F, B, a), a[:, :, None] * F + (1 - a[:, :, None]) * B)


@pytest.mark.skipif(not os.path.isdir(DATA), reason="reference data not present")
def test_load_image_reproduces_fixtures():
    d = load_golden("lemur_0p1")
    assert np.array_equal(pm.load_image(os.path.join(DATA, "lemur/lemur.png"), "rgb", 0.1,
                                        "bilinear"), d["image"])
    assert np.array_equal(pm.load_image(os.path.join(DATA, "lemur/lemur_trimap.png"), "gray",
                                        0.1, "nearest"), d["trimap"])
    t = load_golden("test32")
    assert np.array_equal(pm.load_image(os.path.join(DATA, "test/test.png"), "RGB"), t["image"])
