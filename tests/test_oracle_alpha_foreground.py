"""Pins the oracle's estimate_alpha_cf and estimate_foreground_ml.

Checks: the tolerances of tests/test_estimate_alpha.py, tests/test_foreground.py
and tests/test_simple_api.py on the reference's own images, the optimality
condition (L x)_U = 0 of the hard-constrained solve, and a pure-Python loop
restatement of the Gauss-Seidel pyramid on tiny images.
"""
import numpy as np

from conftest import load_golden
from oracle.cf import cf_laplacian, estimate_alpha_cf, trimap_split
from oracle.foreground import estimate_foreground_ml, level_sizes


def test_estimate_alpha_cf_lemur():
    d = load_golden("lemur_0p1")
    alpha = estimate_alpha_cf(d["image"], d["trimap"])
    assert np.linalg.norm(alpha - d["good_alpha"]) < 1.68


def test_estimate_alpha_cf_optimality():
    d = load_golden("lemur_0p1")
    image, trimap = d["image"], d["trimap"]
    alpha = estimate_alpha_cf(image, trimap, cg_kwargs={"rtol": 1e-12})
    is_fg, is_bg, is_known, is_unknown = trimap_split(trimap)
    L = cf_laplacian(image)
    # Unclipped solution: known pixels at 1/0, unknown pixels from the solve.
    x = np.where(is_known, is_fg.astype(float), 0.0)
    L_U = L[is_unknown][:, is_unknown]
    R = L[is_unknown][:, is_known]
    x_u = np.linalg.solve(L_U.toarray(), -R.dot(is_fg[is_known].astype(float)))
    x[is_unknown] = x_u
    assert np.abs(L.dot(x)[is_unknown]).max() < 1e-8
    assert np.abs(np.clip(x, 0, 1).reshape(trimap.shape) - alpha).max() < 1e-6


def test_simple_api_32():
    d = load_golden("test32")
    alpha = estimate_alpha_cf(d["image"], d["trimap"])
    F = estimate_foreground_ml(d["image"], alpha)
    assert np.linalg.norm(alpha - d["true_alpha"]) < 3.1
    assert np.linalg.norm(alpha[:, :, None] * np.abs(F - d["true_foreground"])) < 4


def test_foreground_lemur():
    d = load_golden("lemur_fg_0p1")
    F, B = estimate_foreground_ml(d["image"], d["alpha"], return_background=True)
    unknown = trimap_split(d["trimap"], flatten=False)[3]
    diff = np.abs(F - d["expected_foreground"])
    assert np.mean((d["alpha"][:, :, None] * diff)[unknown]) <= 0.022
    F64 = estimate_foreground_ml(d["image"], d["alpha"], precision="f64")
    assert np.abs(F - F64).max() < 1e-5


def _fb_ml_python(image, alpha, reg=1e-5, n_small=10, n_big=2, small_size=32, gw=1.0):
    """Loop restatement in NumPy scalars, float64 throughout."""
    h0, w0 = alpha.shape
    n_levels, hs, ws = level_sizes(h0, w0)
    fm = image[alpha > 0.9].sum(axis=0) / ((alpha > 0.9).sum() + 1e-5)
    bm = image[alpha < 0.1].sum(axis=0) / ((alpha < 0.1).sum() + 1e-5)
    Fp, Bp = fm.reshape(1, 1, 3), bm.reshape(1, 1, 3)

    def resize(src, h, w):
        hs_, ws_ = src.shape[:2]
        ys = np.minimum(np.arange(h) * hs_ // h, hs_ - 1)
        xs = np.minimum(np.arange(w) * ws_ // w, ws_ - 1)
        return src[ys][:, xs].copy()

    for l in range(n_levels + 1):
        h, w = hs[l], ws[l]
        img, a = resize(image, h, w), resize(alpha, h, w)
        F, B = resize(Fp, h, w), resize(Bp, h, w)
        for _ in range(n_small if (w <= small_size and h <= small_size) else n_big):
            for y in range(h):
                for x in range(w):
                    a0 = a[y, x]
                    a1 = 1.0 - a0
                    M = np.array([[a0 * a0, a0 * a1], [a0 * a1, a1 * a1]])
                    rhs = np.array([a0 * img[y, x], a1 * img[y, x]])
                    for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                        y2 = min(max(y + dy, 0), h - 1)
                        x2 = min(max(x + dx, 0), w - 1)
                        da = reg + gw * abs(a0 - a[y2, x2])
                        M += da * np.eye(2)
                        rhs += da * np.array([F[y2, x2], B[y2, x2]])
                    sol = np.linalg.solve(M, rhs)
                    F[y, x] = np.clip(sol[0], 0, 1)
                    B[y, x] = np.clip(sol[1], 0, 1)
        Fp, Bp = F, B
    return Fp, Bp


def test_foreground_matches_loop_restatement():
    rng = np.random.default_rng(4)
    for h, w in [(5, 7), (9, 4), (33, 20)]:
        image = rng.random((h, w, 3))
        alpha = np.clip(rng.random((h, w)) * 1.4 - 0.2, 0, 1)
        F, B = estimate_foreground_ml(image, alpha, return_background=True, precision="f64")
        Fr, Br = _fb_ml_python(image, alpha)
        assert np.abs(F - Fr).max() < 1e-9 and np.abs(B - Br).max() < 1e-9


def test_foreground_mean_modes_agree():
    d = load_golden("lemur_fg_0p1")
    F0 = estimate_foreground_ml(d["image"], d["alpha"], mean_mode="reference")
    F1 = estimate_foreground_ml(d["image"], d["alpha"], mean_mode="f64")
    assert np.abs(F0 - F1).max() < 1e-5
