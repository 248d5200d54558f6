"""Image files and array helpers of the drop-in API.

Mirrors pymatting/util/util.py:211-287 (load_image, save_image) and
:581-643 (blend, stack_images): PNG/JPEG decoding and encoding with PIL on
the host (files live there), float images in [0, 1].
"""
import os

import numpy as np
from PIL import Image

_FILTERS = {
    "bicubic": Image.BICUBIC,
    "bilinear": Image.BILINEAR,
    "box": Image.BOX,
    "hamming": Image.HAMMING,
    "lanczos": Image.LANCZOS,
    "nearest": Image.NEAREST,
    "none": Image.NEAREST,
}


def _resize(image, size, resample):
    """util.py:211-228: a scale factor or a (width, height) pair."""
    try:
        iter(size)
    except TypeError:
        size = (int(image.width * size), int(image.height * size))
    return image.resize(size, _FILTERS[resample.lower()])


def load_image(path, mode=None, size=None, resample="box"):
    """util.py:232-260: the image at `path` as a float64 array in [0, 1].

    mode: "GRAY", "RGB" or any PIL mode to convert to; size: a scale factor
    or (width, height), resized with the named filter.
    """
    image = Image.open(path)
    if mode is not None:
        mode = mode.upper()
        image = image.convert("L" if mode == "GRAY" else mode)
    if size is not None:
        image = _resize(image, size, resample)
    return np.array(image) / 255.0


def save_image(path, image, make_directory=True):
    """util.py:263-287: saves a uint8 image, or a float image in [0, 1] scaled
    to 0..255 (clipped, truncated), creating the directory if asked."""
    assert image.dtype in [np.uint8, np.float32, np.float64]
    if image.dtype in [np.float32, np.float64]:
        image = np.clip(image * 255, 0, 255).astype(np.uint8)
    if make_directory:
        directory = os.path.dirname(path)
        if directory:
            os.makedirs(directory, exist_ok=True)
    Image.fromarray(image).save(path)


def stack_images(*images):
    """util.py:617-643: concatenates images along the channel axis (2-D
    images count as one channel)."""
    return np.concatenate([im if im.ndim == 3 else im[:, :, None] for im in images], axis=2)


def blend(foreground, background, alpha):
    """util.py:581-614: alpha F + (1 - alpha) B."""
    if alpha.ndim == 2:
        alpha = alpha[:, :, None]
    return alpha * foreground + (1 - alpha) * background
