"""Writes the small input fixtures the parity tests use.

Reads the reference's own test images (pymatting data/lemur/*.png and
data/test/*.png) and resizes them with PIL exactly the way the reference's
tests call ``load_image(path, mode, scale, resample)``
(pymatting/util/util.py:211-260): convert the mode, resize to
(int(width*scale), int(height*scale)) with the named filter, divide by 255.
Only data is written; no reference code is imported or run.

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import os

import numpy as np
from PIL import Image

REF = "/root/reference/data"
OUT = os.path.dirname(os.path.abspath(__file__))
FILTERS = {"bilinear": Image.BILINEAR, "nearest": Image.NEAREST, "box": Image.BOX}


def load(path, mode, scale=None, resample="box"):
    im = Image.open(os.path.join(REF, path))
    im = im.convert("L" if mode.upper() == "GRAY" else mode.upper())
    if scale is not None:
        im = im.resize((int(im.width * scale), int(im.height * scale)), FILTERS[resample])
    return np.array(im) / 255.0


def main():
    # tests/test_laplacians.py:14-22 and tests/test_estimate_alpha.py:6-14 (scale 0.1).
    np.savez_compressed(
        os.path.join(OUT, "lemur_0p1.npz"),
        image=load("lemur/lemur.png", "rgb", 0.1, "bilinear"),
        trimap=load("lemur/lemur_trimap.png", "gray", 0.1, "nearest"),
        good_alpha=load("lemur/lemur_alpha_fba.png", "gray", 0.1, "bilinear"))
    # tests/test_preconditioners.py:15-22 and tests/test_lkm.py:14-22 (scale 0.125).
    np.savez_compressed(
        os.path.join(OUT, "lemur_0p125.npz"),
        image=load("lemur/lemur.png", "rgb", 0.125, "bilinear"),
        trimap=load("lemur/lemur_trimap.png", "gray", 0.125, "nearest"))
    # tests/test_foreground.py:35-41 (scale 0.1, box filter).
    np.savez_compressed(
        os.path.join(OUT, "lemur_fg_0p1.npz"),
        image=load("lemur/lemur.png", "RGB", 0.1, "box"),
        trimap=load("lemur/lemur_trimap.png", "GRAY", 0.1, "nearest"),
        alpha=load("lemur/lemur_alpha.png", "GRAY", 0.1, "box"),
        expected_foreground=load("lemur/lemur_foreground.png", "RGB", 0.1, "box"))
    # tests/test_simple_api.py:6-27 (32x32 test image, no resize).
    np.savez_compressed(
        os.path.join(OUT, "test32.npz"),
        image=load("test/test.png", "RGB"),
        trimap=load("test/test_trimap.png", "GRAY"),
        true_alpha=load("test/test_alpha.png", "GRAY"),
        true_foreground=load("test/test_foreground.png", "RGB"))


if __name__ == "__main__":
    main()
