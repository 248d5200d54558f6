"""File-to-file matting (pymatting/cutout/cutout.py:6-36).

Loads the image and trimap, solves alpha (estimate_alpha_cf) and the
foreground (estimate_foreground_ml) on the GPU, and saves the RGBA cutout.
"""
from pymatting_amd.alpha import estimate_alpha_cf
from pymatting_amd.foreground import estimate_foreground_ml
from pymatting_amd.io import load_image, save_image, stack_images


def cutout(image_path, trimap_path, cutout_path):
    """cutout.py:6-36: writes foreground and alpha of the image as an RGBA file."""
    image = load_image(image_path, "RGB")
    trimap = load_image(trimap_path, "GRAY")
    if image.shape[:2] != trimap.shape[:2]:
        raise ValueError("Input image and trimap must have same size")
    alpha = estimate_alpha_cf(image, trimap)
    foreground = estimate_foreground_ml(image, alpha)
    save_image(cutout_path, stack_images(foreground, alpha))
