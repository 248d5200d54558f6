"""Host-side checks of the drop-in API: trimap_split and sanity_check_image
(pymatting/util/util.py:439-578) on numpy inputs.

The thresholds are inclusive (fg: t >= 0.9, bg: t <= 0.1, util.py:509-510),
the warnings carry the reference's messages (the max-warning prints the
minimum, util.py:498), and a trimap without background or foreground raises
ValueError (util.py:512-520). Device tensors go through the same checks on
the device (tests/test_gpu_alpha.py).
"""
import warnings

import numpy as np
import pytest

import pymatting_amd as pm
import oracle.cf as ocf


def _caught(fn, *args):
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        out = fn(*args)
    return out, [str(x.message) for x in w]


def test_trimap_split_thresholds_inclusive():
    t = np.array([[0.0, 0.1, np.nextafter(0.1, 1.0), 0.5],
                  [np.nextafter(0.9, 0.0), 0.9, 1.0, 0.4]])
    fg, bg, known, unknown = pm.trimap_split(t)
    assert fg.tolist() == [False, False, False, False, False, True, True, False]
    assert bg.tolist() == [True, True, False, False, False, False, False, False]
    assert np.array_equal(known, fg | bg) and np.array_equal(unknown, ~known)
    for a, b in zip(pm.trimap_split(t), ocf.trimap_split(t)):
        assert np.array_equal(a, b)
    fg2 = pm.trimap_split(t, flatten=False)[0]
    assert fg2.shape == t.shape
    # Custom thresholds move the class boundaries.
    fg3, bg3, _, _ = pm.trimap_split(t, bg_threshold=0.4, fg_threshold=0.5)
    assert bool(fg3[3]) and bool(bg3[7])


def test_trimap_split_warnings():
    t = np.array([-0.5, 0.0, 0.5, 1.0])
    _, msgs = _caught(pm.trimap_split, t)
    assert msgs == ["Trimap values should be in [0, 1], but trimap.min() is -0.5."]
    t = np.array([0.0, 0.2, 0.5, 1.5])
    _, msgs = _caught(pm.trimap_split, t)
    # util.py:498 reports the minimum in the max-warning.
    assert msgs == ["Trimap values should be in [0, 1], but trimap.max() is 0.0."]
    t = np.array([0, 1, 0, 1], dtype=np.int64)
    _, msgs = _caught(pm.trimap_split, t)
    assert len(msgs) == 1 and msgs[0].startswith("Unexpected trimap.dtype int64.")
    _, msgs = _caught(pm.trimap_split, np.array([0.0, 0.5, 1.0]))
    assert msgs == []


def test_trimap_split_errors():
    with pytest.raises(ValueError, match="did not contain background"):
        pm.trimap_split(np.array([0.5, 1.0, 0.95]))
    with pytest.raises(ValueError, match="did not contain foreground"):
        pm.trimap_split(np.array([0.5, 0.0, 0.05]))
    # Both missing: the background check comes first (util.py:512-520).
    with pytest.raises(ValueError, match="did not contain background"):
        pm.trimap_split(np.array([0.5, 0.5]))


def test_sanity_check_image_warnings():
    ok = np.full((4, 5, 3), 0.5)
    _, msgs = _caught(pm.sanity_check_image, ok)
    assert msgs == []
    _, msgs = _caught(pm.sanity_check_image, np.full((4, 5), 0.5))
    assert msgs == ["Expected RGB image of shape (?, ?, 3), but image.shape is (4, 5)."]
    bad = ok.copy()
    bad[0, 0, 0], bad[1, 1, 1] = -0.25, 1.5
    _, msgs = _caught(pm.sanity_check_image, bad)
    assert msgs == ["Image values should be in [0, 1], but image.min() is -0.25.",
                    "Image values should be in [0, 1], but image.max() is 1.5."]
    _, msgs = _caught(pm.sanity_check_image, np.zeros((4, 5, 3), dtype=np.uint8))
    assert len(msgs) == 1 and msgs[0].startswith("Unexpected image.dtype uint8.")
