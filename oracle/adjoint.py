"""Reference gradient of the lambda-penalised closed-form alpha solve.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py). pymatting has no
gradient; this module restates its forward computation densely in torch
float64 (the matting Laplacian of cf_laplacian.py:5-129 assembled window by
window, A = L + lambda diag(known) and b = lambda is_fg of
laplacian.py:46-47, an exact dense solve, the clip of
estimate_alpha_cf.py:76) and lets torch's autograd differentiate it. The
gradient so obtained owes nothing to the hand-written adjoint it checks.
Dense: for images of at most a few hundred pixels.
"""
import numpy as np
import torch


def dense_laplacian(image, epsilon=1e-7):
    """L(I) for a (h, w, 3) float64 torch tensor, differentiable in I."""
    h, w, _ = image.shape
    n = h * w
    flat = image.reshape(n, 3)
    cy, cx = np.mgrid[1:h - 1, 1:w - 1]
    cy, cx = cy.ravel(), cx.ravel()
    offs = [(dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    idx = torch.tensor(np.stack([(cy + dy) * w + (cx + dx) for dy, dx in offs], axis=1))  # (K, 9)
    colors = flat[idx]                                       # (K, 9, 3)
    mu = colors.mean(dim=1, keepdim=True)
    c = colors - mu
    cov = c.transpose(1, 2) @ c / 9.0 + epsilon / 9.0 * torch.eye(3, dtype=image.dtype)
    M = torch.linalg.inv(cov)                                # (K, 3, 3)
    Lk = torch.eye(9, dtype=image.dtype) - (1.0 + c @ M @ c.transpose(1, 2)) / 9.0  # (K, 9, 9)
    rows = idx[:, :, None].expand(-1, 9, 9).reshape(-1)
    cols = idx[:, None, :].expand(-1, 9, 9).reshape(-1)
    L = torch.zeros((n, n), dtype=image.dtype)
    return L.index_put((rows, cols), Lk.reshape(-1), accumulate=True)


def alpha_loss_and_gradient(image, trimap, target, epsilon=1e-7, lambda_value=100.0):
    """Loss mean((clip(x) - target)^2) with x = A(I)^-1 b, and d loss / d image.

    image (h, w, 3), trimap (h, w), target (h, w): numpy float64.
    Returns (loss, alpha, grad) as numpy arrays.
    """
    I = torch.tensor(np.asarray(image, dtype=np.float64), requires_grad=True)
    t = np.asarray(trimap, dtype=np.float64).ravel()
    is_fg = torch.tensor((t >= 0.9).astype(np.float64))
    known = torch.tensor(((t >= 0.9) | (t <= 0.1)).astype(np.float64))
    A = dense_laplacian(I, epsilon) + torch.diag(lambda_value * known)
    x = torch.linalg.solve(A, lambda_value * is_fg)
    alpha = x.clamp(0.0, 1.0)
    loss = ((alpha - torch.tensor(np.asarray(target, dtype=np.float64).ravel())) ** 2).mean()
    loss.backward()
    return (float(loss.detach()), alpha.detach().numpy().reshape(image.shape[:2]),
            I.grad.numpy().copy())
