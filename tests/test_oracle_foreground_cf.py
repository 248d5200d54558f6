"""Pins the oracle's estimate_foreground_cf
(pymatting/foreground/estimate_foreground_cf.py:8-130): its system matrix
equals half the Hessian of the energy the reference minimises (brute-force
loops, tiny images, several neighbour lists), its result equals a dense
solve of that Hessian's system on a tiny image, and the lemur fixture stays within the reference's
tests/test_foreground.py tolerance (alpha-weighted mean |F - F_expected|
<= 0.022 on the unknown pixels)."""
import numpy as np
import pytest

from conftest import load_golden
from oracle.cf import trimap_split
from oracle.foreground_cf import (estimate_foreground_cf, foreground_cf_hessian_bruteforce,
                                  foreground_cf_system)

NEIGHBOURS = [((-1, 0), (1, 0), (0, -1), (0, 1)), ((1, 0), (0, 1)),
              ((-1, -1), (2, 0), (0, 1), (1, 1))]


@pytest.mark.parametrize("shape", [(4, 5), (1, 6), (6, 1), (3, 3)])
@pytest.mark.parametrize("nb", NEIGHBOURS)
def test_system_is_the_energy_hessian(shape, nb):
    rng = np.random.default_rng(shape[0] * 10 + shape[1])
    alpha = np.clip(rng.random(shape) * 1.4 - 0.2, 0, 1)
    A, U = foreground_cf_system(alpha, 1e-5, nb)
    H = foreground_cf_hessian_bruteforce(alpha, 1e-5, nb)
    assert np.abs(A.toarray() - H).max() <= 1e-13
    assert A.has_canonical_format and not np.any(A.data == 0)


def test_matches_a_dense_solve_of_the_energy():
    """On a tiny image the whole function equals clip(H^-1 U^T I) per channel,
    with H the brute-force Hessian and a dense solve."""
    rng = np.random.default_rng(5)
    h, w = 6, 7
    image = rng.random((h, w, 3))
    alpha = np.clip(rng.random((h, w)) * 1.4 - 0.2, 0, 1)
    n = h * w
    H = foreground_cf_hessian_bruteforce(alpha)
    a = alpha.ravel()
    F, B = estimate_foreground_cf(image, alpha, rtol=1e-13, return_background=True)
    for c in range(3):
        I = image[:, :, c].ravel()
        x = np.linalg.solve(H, np.concatenate([a * I, (1 - a) * I]))
        assert np.abs(F[:, :, c].ravel() - np.clip(x[:n], 0, 1)).max() <= 1e-8
        assert np.abs(B[:, :, c].ravel() - np.clip(x[n:], 0, 1)).max() <= 1e-8


def test_lemur_tolerance():
    d = load_golden("lemur_fg_0p1")
    image, alpha, trimap = d["image"], d["alpha"], d["trimap"]
    F, B = estimate_foreground_cf(image, alpha, return_background=True)
    _, _, _, unknown = trimap_split(trimap, flatten=False)
    diff = np.abs(F - d["expected_foreground"])
    assert np.mean((alpha[:, :, None] * diff)[unknown]) <= 0.022
