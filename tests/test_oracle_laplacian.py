"""Pins the oracle's cf_laplacian and make_linear_system.

Checks: the reference's known answer (tests/test_laplacians.py:26), an
independent dense assembly of the Levin et al. Laplacian, closed-form
invariants (symmetry, L 1 = 0, positive semi-definiteness) and the storage
layout of pymatting/laplacian/cf_laplacian.py:12-25,169-175.
"""
import numpy as np
import pytest
import scipy.sparse.linalg

from conftest import load_golden
from oracle.cf import (cf_laplacian, cf_laplacian_dense_bruteforce, make_linear_system,
                       trimap_split)


def _tiny_images():
    rng = np.random.default_rng(7)
    yield rng.random((4, 5, 3)), 1e-7
    yield rng.random((7, 6, 3)), 1e-7
    yield np.full((5, 5, 3), 0.25), 1e-7                 # flat windows: Sigma = eps/9 I
    t = rng.random((6, 7, 1))
    # Collinear colours: Sigma has two eigenvalues eps/9, so a larger eps keeps
    # the comparison of two inverse formulas well-conditioned.
    yield np.concatenate([t, 0.5 * t, 0.2 + 0.3 * t], axis=2), 1e-3
    yield rng.random((3, 3, 3)), 1e-7                    # a single window


@pytest.mark.parametrize("k", range(5))
def test_matches_dense_bruteforce(k):
    image, eps = list(_tiny_images())[k]
    L = cf_laplacian(image, epsilon=eps).toarray()
    D = cf_laplacian_dense_bruteforce(image, epsilon=eps)
    scale = max(1.0, np.abs(D).max())
    assert np.abs(L - D).max() <= 1e-10 * scale


def test_known_mask_skips_all_known_windows():
    rng = np.random.default_rng(3)
    image = rng.random((6, 7, 3))
    known = rng.random((6, 7)) < 0.7
    L = cf_laplacian(image, is_known=known).toarray()
    D = cf_laplacian_dense_bruteforce(image, is_known=known)
    assert np.abs(L - D).max() <= 1e-9 * np.abs(D).max()


def test_invariants():
    rng = np.random.default_rng(11)
    image = rng.random((9, 8, 3))
    L = cf_laplacian(image, epsilon=1e-4)
    Ld = L.toarray()
    assert np.abs(Ld - Ld.T).max() < 1e-12 * np.abs(Ld).max()
    assert np.abs(L.dot(np.ones(72))).max() < 1e-9
    assert np.linalg.eigvalsh(0.5 * (Ld + Ld.T)).min() > -1e-9


def test_layout():
    rng = np.random.default_rng(5)
    h, w = 4, 6
    L = cf_laplacian(rng.random((h, w, 3)))
    assert L.indptr.tolist() == list(range(0, 25 * h * w + 1, 25))
    idx = L.indices.reshape(h, w, 5, 5)
    val = L.data.reshape(h, w, 5, 5)
    for y in range(h):
        for x in range(w):
            for oy in range(5):
                for ox in range(5):
                    yy, xx = y + oy - 2, x + ox - 2
                    if 0 <= yy < h and 0 <= xx < w:
                        assert idx[y, x, oy, ox] == xx + yy * w
                    else:
                        assert idx[y, x, oy, ox] == 0 and val[y, x, oy, ox] == 0.0


def test_known_answer_lemur():
    d = load_golden("lemur_0p1")
    image, trimap, good = d["image"], d["trimap"], d["good_alpha"]
    A, b = make_linear_system(cf_laplacian(image), trimap)
    x = scipy.sparse.linalg.spsolve(A, b)
    alpha = np.clip(x, 0, 1).reshape(trimap.shape)
    unknown = trimap_split(trimap, flatten=False)[3]
    err = np.linalg.norm(np.abs(alpha - good)[unknown])
    # tests/test_laplacians.py:26 holds 1.6383251091411692 (1 % regression allowed).
    assert abs(err - 1.6383251091411692) / 1.6383251091411692 < 1e-9


def test_make_linear_system():
    rng = np.random.default_rng(2)
    image = rng.random((5, 6, 3))
    trimap = np.full((5, 6), 0.5)
    trimap[0, :] = 1.0
    trimap[-1, :] = 0.0
    L = cf_laplacian(image)
    A, b, c = make_linear_system(L, trimap, lambda_value=7.0, return_c=True)
    known = trimap.ravel() != 0.5
    assert np.allclose(A.toarray(), L.toarray() + np.diag(7.0 * known), rtol=0, atol=1e-14)
    assert np.array_equal(b, 7.0 * (trimap.ravel() >= 0.9))
    assert np.array_equal(c, 7.0 * known)
    assert A.has_canonical_format


def test_trimap_split_errors():
    with pytest.raises(ValueError):
        trimap_split(np.full((3, 3), 0.5).clip(0, 0.05))
    with pytest.raises(ValueError):
        trimap_split(np.ones((3, 3)))
    fg, bg, known, unknown = trimap_split(np.array([[1, 0], [0.5, 0.2]]))
    assert fg.tolist() == [True, False, False, False]
    assert bg.tolist() == [False, True, False, False]
    assert unknown.tolist() == [False, False, True, True]
