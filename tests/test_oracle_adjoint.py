"""Pins the oracle's reference gradient (oracle/adjoint.py): its dense
Laplacian is the oracle's cf_laplacian, its solve reproduces the oracle's
alpha, and its autograd gradient matches central finite differences."""
import numpy as np

import oracle.cf as ocf
from oracle.adjoint import alpha_loss_and_gradient, dense_laplacian


def _case(h=6, w=7, seed=0):
    rng = np.random.default_rng(seed)
    image = rng.random((h, w, 3))
    trimap = np.full((h, w), 0.5)
    trimap[:, 0] = 0.0
    trimap[:, -1] = 1.0
    trimap[0, 2] = 1.0
    target = rng.random((h, w))
    return image, trimap, target


def test_dense_laplacian_is_cf_laplacian():
    import torch
    image, _, _ = _case(7, 8, 1)
    L = dense_laplacian(torch.tensor(image), 1e-3).numpy()
    Lr = ocf.cf_laplacian(image, 1e-3).toarray()
    assert np.abs(L - Lr).max() < 1e-10


def test_alpha_matches_oracle_solve():
    import scipy.sparse.linalg
    image, trimap, target = _case()
    _, alpha, _ = alpha_loss_and_gradient(image, trimap, target)
    A, b = ocf.make_linear_system(ocf.cf_laplacian(image), trimap)
    x = scipy.sparse.linalg.spsolve(A.tocsc(), b)
    assert np.abs(alpha.ravel() - np.clip(x, 0, 1)).max() < 1e-8


def test_gradient_matches_finite_differences():
    image, trimap, target = _case(6, 7, 2)
    eps = 1e-3  # a well-conditioned window regime for finite differences
    loss, _, grad = alpha_loss_and_gradient(image, trimap, target, epsilon=eps)
    h = 1e-6
    rng = np.random.default_rng(3)
    for _ in range(12):
        y, x, c = rng.integers(0, 6), rng.integers(0, 7), rng.integers(0, 3)
        ip, im = image.copy(), image.copy()
        ip[y, x, c] += h
        im[y, x, c] -= h
        fd = (alpha_loss_and_gradient(ip, trimap, target, epsilon=eps)[0]
              - alpha_loss_and_gradient(im, trimap, target, epsilon=eps)[0]) / (2 * h)
        assert abs(fd - grad[y, x, c]) <= 1e-6 * np.abs(grad).max() + 1e-4 * abs(fd), (y, x, c)
