"""Pins the oracle's lkm_laplacian (pymatting/laplacian/lkm_laplacian.py):
the large-kernel operator is the closed-form Laplacian of the same radius
computed with box filters, so its products and diagonal must equal those of
the oracle's cf_laplacian (itself bit-exact to the reference) to rounding;
its box filter matches a direct window sum, and the reference's
tests/test_lkm.py known answer holds (CG with Jacobi on the lemur at 0.125,
radius 2: ||x_lkm - x_cf|| < 1e-5, iteration counts within 2)."""
import numpy as np
import pytest
import scipy.signal

from conftest import load_golden
import oracle.cf as ocf
import oracle.lkm as olkm
import oracle.solver as osol


def test_boxfilter_matches_window_sums():
    rng = np.random.default_rng(0)
    a = rng.random((9, 13))
    for r in (1, 2, 4):
        k = np.ones((2 * r + 1, 2 * r + 1))
        assert np.abs(olkm.boxfilter(a, r, "valid")
                      - scipy.signal.correlate2d(a, k, "valid")).max() < 1e-12
        assert np.abs(olkm.boxfilter(a, r, "full")
                      - scipy.signal.correlate2d(a, k, "full")).max() < 1e-12


@pytest.mark.parametrize("r", [1, 2, 3])
def test_operator_is_the_cf_laplacian(r):
    rng = np.random.default_rng(r)
    h, w = 4 * r + 5, 4 * r + 7
    image = rng.random((h, w, 3))
    L_matvec, diag = olkm.lkm_laplacian(image, epsilon=1e-4, radius=r)
    L = ocf.cf_laplacian(image, epsilon=1e-4, radius=r)
    for _ in range(3):
        p = rng.standard_normal(h * w)
        ref = L @ p
        assert np.abs(L_matvec(p) - ref).max() <= 1e-8 * np.abs(ref).max()
    assert np.abs(diag - L.diagonal()).max() <= 1e-8 * np.abs(L.diagonal()).max()


def test_lemur_known_answer():
    d = load_golden("lemur_0p125")
    image, trimap = d["image"], d["trimap"]
    L_matvec, diag = olkm.lkm_laplacian(image, epsilon=1e-7, radius=2)
    A_cf, b, c = ocf.make_linear_system(ocf.cf_laplacian(image, epsilon=1e-7, radius=2), trimap,
                                        return_c=True)
    inv = 1.0 / (diag + c)
    n_lkm, n_cf = osol.CounterCallback(), osol.CounterCallback()
    x_lkm = osol.cg(lambda x: L_matvec(x) + c * x, b, M=lambda r: inv * r, callback=n_lkm)
    x_cf = osol.cg(A_cf, b, M=osol.jacobi(A_cf), callback=n_cf)
    assert np.linalg.norm(x_lkm - x_cf) < 1e-5
    assert abs(n_lkm.n - n_cf.n) <= 2
