"""Closed-form foreground estimation on the CPU (test infrastructure).

Restates pymatting/foreground/estimate_foreground_cf.py:8-130 with scipy:
directional difference matrices to the clamped neighbours
(sparse_conv_matrix_with_offsets, pymatting/util/util.py:757-802), the
smoothness term S = sum_k D_k^T diag(reg + |D_k a|) D_k, the system
A = U^T U + blockdiag(S, S) with U = [diag(a), diag(1 - a)], and one
ichol-preconditioned CG per colour channel started from the guesses.
"""
import numpy as np
import scipy.sparse


def difference_matrix(width, height, dx, dy):
    """(D x)_i = x_i - x_j with j the neighbour at (dx, dy), clamped into the
    image; a pixel whose clamped neighbour is itself gets an explicit zero
    (util.py:757-802 with kernel [1, -1] at offsets (0, 0) and (dx, dy))."""
    n = width * height
    x = np.tile(np.arange(width), height)
    y = np.repeat(np.arange(height), width)
    rows = np.concatenate([x + y * width, x + y * width])
    nx = np.clip(x + dx, 0, width - 1)
    ny = np.clip(y + dy, 0, height - 1)
    cols = np.concatenate([x + y * width, nx + ny * width])
    vals = np.concatenate([np.ones(n), -np.ones(n)])
    return scipy.sparse.csr_matrix((vals, (rows.astype(np.int32), cols.astype(np.int32))),
                                   shape=(n, n))


def foreground_cf_system(alpha, regularization=1e-5,
                         neighbors=((-1, 0), (1, 0), (0, -1), (0, 1))):
    """(A, U) of estimate_foreground_cf.py:69-90, in the reference's order of
    sparse products and sums."""
    h, w = alpha.shape
    a = alpha.flatten()
    S = None
    for dx, dy in neighbors:
        D = difference_matrix(w, h, dx, dy)
        term = D.T.dot(scipy.sparse.diags(regularization + np.abs(D.dot(a)))).dot(D)
        S = term if S is None else S + term
    V = scipy.sparse.bmat([[S, None], [None, S]])
    U = scipy.sparse.bmat([[scipy.sparse.diags(a), scipy.sparse.diags(1 - a)]])
    A = U.T.dot(U) + V
    A.sum_duplicates()
    return A, U


def estimate_foreground_cf(image, alpha, regularization=1e-5, rtol=1e-5,
                           neighbors=((-1, 0), (1, 0), (0, -1), (0, 1)), return_background=False,
                           foreground_guess=None, background_guess=None, ichol_kwargs={},
                           cg_kwargs={}):
    """estimate_foreground_cf.py:8-130."""
    from oracle.ichol import ichol
    from oracle.solver import cg
    h, w, d = image.shape
    assert alpha.shape == (h, w)
    n = w * h
    A, U = foreground_cf_system(alpha, regularization, neighbors)
    M = ichol(A, **ichol_kwargs)
    F = (image if foreground_guess is None else foreground_guess).copy()
    B = (image if background_guess is None else background_guess).copy()
    for c in range(d):
        b = U.T.dot(image[:, :, c].flatten())
        x0 = np.concatenate([F[:, :, c].flatten(), B[:, :, c].flatten()])
        x = cg(A, b, x0=x0, M=M, rtol=rtol, **cg_kwargs)
        F[:, :, c] = x[:n].reshape(h, w)
        B[:, :, c] = x[n:].reshape(h, w)
    F = np.clip(F, 0, 1)
    B = np.clip(B, 0, 1)
    if return_background:
        return F, B
    return F


def foreground_cf_hessian_bruteforce(alpha, regularization=1e-5,
                                     neighbors=((-1, 0), (1, 0), (0, -1), (0, 1))):
    """Half the Hessian of the energy the system minimises, by loops:
    sum_i (a_i F_i + (1 - a_i) B_i - I_i)^2
      + sum_k sum_i (reg + |a_i - a_j|) ((F_i - F_j)^2 + (B_i - B_j)^2),
    j the clamped neighbour of i at offset k (dense, tiny images only)."""
    h, w = alpha.shape
    n = h * w
    a = alpha.flatten()
    H = np.zeros((2 * n, 2 * n))
    for i in range(n):
        H[i, i] += a[i] * a[i]
        H[i, n + i] += a[i] * (1 - a[i])
        H[n + i, i] += a[i] * (1 - a[i])
        H[n + i, n + i] += (1 - a[i]) ** 2
    for dx, dy in neighbors:
        for yi in range(h):
            for xi in range(w):
                i = yi * w + xi
                j = min(max(yi + dy, 0), h - 1) * w + min(max(xi + dx, 0), w - 1)
                if i == j:
                    continue
                g = regularization + abs(a[i] - a[j])
                for o in (0, n):
                    H[o + i, o + i] += g
                    H[o + j, o + j] += g
                    H[o + i, o + j] -= g
                    H[o + j, o + i] -= g
    return H
