"""Multilevel foreground estimation on the CPU (test infrastructure).

Restates pymatting/foreground/estimate_foreground_ml.py:61-253 around the C
sweeps in ``fb_ml.c``.
"""
import numpy as np

from oracle._native import lib, ptr


def level_sizes(h0, w0):
    """estimate_foreground_ml.py:101-105: pyramid sizes round(w0 ** (i / n))."""
    n_levels = int(np.ceil(np.log2(max(w0, h0))))
    if n_levels == 0:
        raise ZeroDivisionError("division by zero")
    hs = [round(h0 ** (i / n_levels)) for i in range(n_levels + 1)]
    ws = [round(w0 ** (i / n_levels)) for i in range(n_levels + 1)]
    return n_levels, hs, ws


def estimate_foreground_ml(image, alpha, regularization=1e-5, n_small_iterations=10,
                           n_big_iterations=2, small_size=32, return_background=False,
                           gradient_weight=1.0, precision="f32", mean_mode="reference"):
    """estimate_foreground_ml.py:186-253.

    precision "f32" mirrors the reference's float32/float64 typing; "f64"
    runs everything in double. mean_mode "reference" sums colours in
    float32 like the reference, "f64" sums them in double.
    """
    h0, w0 = alpha.shape[:2]
    n_levels, hs, ws = level_sizes(h0, w0)
    lh = np.array(hs, dtype=np.int32)
    lw = np.array(ws, dtype=np.int32)
    if precision == "f32":
        img = np.ascontiguousarray(image, dtype=np.float32)
        a = np.ascontiguousarray(alpha, dtype=np.float32)
        F = np.empty((h0, w0, 3), dtype=np.float32)
        B = np.empty((h0, w0, 3), dtype=np.float32)
        lib().oracle_fb_ml_f32(ptr(img), ptr(a), h0, w0, n_levels, ptr(lh), ptr(lw),
                               float(np.float32(regularization)), n_small_iterations,
                               n_big_iterations, small_size, float(np.float32(gradient_weight)),
                               1 if mean_mode == "f64" else 0, ptr(F), ptr(B))
    else:
        img = np.ascontiguousarray(image, dtype=np.float64)
        a = np.ascontiguousarray(alpha, dtype=np.float64)
        F = np.empty((h0, w0, 3), dtype=np.float64)
        B = np.empty((h0, w0, 3), dtype=np.float64)
        lib().oracle_fb_ml_f64(ptr(img), ptr(a), h0, w0, n_levels, ptr(lh), ptr(lw),
                               float(regularization), n_small_iterations, n_big_iterations,
                               small_size, float(gradient_weight), ptr(F), ptr(B))
    if return_background:
        return F, B
    return F
