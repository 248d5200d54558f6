"""CPU restatement of the closed-form matting Laplacian and its linear systems.

Test infrastructure (see ``oracle/__init__.py``). The Laplacian is assembled
with exactly the reference's storage layout and floating-point operation
order so that its entries can be compared bit-for-bit.
"""
import warnings

import numpy as np
import scipy.sparse


def trimap_split(trimap, flatten=True, bg_threshold=0.1, fg_threshold=0.9):
    """pymatting/util/util.py:439-525: split a trimap into fg/bg/known/unknown."""
    if flatten:
        trimap = trimap.flatten()
    lo = trimap.min()
    hi = trimap.max()
    if lo < 0.0:
        warnings.warn("Trimap values should be in [0, 1], but trimap.min() is %s." % lo, stacklevel=3)
    if hi > 1.0:
        # util.py:498 prints the minimum in this message; kept as is.
        warnings.warn("Trimap values should be in [0, 1], but trimap.max() is %s." % lo, stacklevel=3)
    if trimap.dtype not in [np.float32, np.float64]:
        warnings.warn("Unexpected trimap.dtype %s. Are you sure that you do not want to use "
                      "np.float32 or np.float64 instead?" % trimap.dtype, stacklevel=3)
    is_fg = trimap >= fg_threshold
    is_bg = trimap <= bg_threshold
    if is_bg.sum() == 0:
        raise ValueError("Trimap did not contain background values (values <= %f)" % bg_threshold)
    if is_fg.sum() == 0:
        raise ValueError("Trimap did not contain foreground values (values >= %f)" % fg_threshold)
    is_known = is_fg | is_bg
    return is_fg, is_bg, is_known, ~is_known


def _window_sum_order(radius):
    size = 2 * radius + 1
    return [(dy, dx) for dy in range(size) for dx in range(size)]


def cf_laplacian(image, epsilon=1e-7, radius=1, is_known=None):
    """pymatting/laplacian/cf_laplacian.py:5-177.

    Returns the non-canonical CSR matrix of the reference: (4r+1)^2 slots per
    row in raster order of the neighbour offset, out-of-image slots holding
    column 0 and value 0.

    Every window's arithmetic follows cf_laplacian.py:38-129 operation by
    operation (sum in dy,dx order, centred colours, epsilon-regularised
    covariance divided by the window area, cofactor inverse with the negated
    determinant). Contributions are added to each slot in raster order of the
    window centre, which is the order the reference's serial loop produces,
    so the entries agree with the reference bit for bit.
    """
    image = np.asarray(image)
    h, w, d = image.shape
    assert d == 3
    assert image.dtype == np.float64
    r = int(radius)
    size = 2 * r + 1
    area = size * size
    K = 4 * r + 1
    n = h * w

    if is_known is None:
        is_known = np.zeros((h, w), dtype=np.bool_)
    is_known = np.asarray(is_known).reshape(h, w)

    # Index pass (cf_laplacian.py:12-25).
    yi, xi = np.mgrid[0:h, 0:w]
    indices = np.zeros((h, w, K, K), dtype=np.int64)
    for oy in range(K):
        for ox in range(K):
            yj = yi + oy - 2 * r
            xj = xi + ox - 2 * r
            ok = (xj >= 0) & (xj < w) & (yj >= 0) & (yj < h)
            indices[:, :, oy, ox][ok] = (xj + yj * w)[ok]
    indptr = np.arange(n + 1, dtype=np.int64) * (K * K)
    values = np.zeros((n, K, K), dtype=np.float64)

    if h > 2 * r and w > 2 * r:
        cy, cx = np.mgrid[r:h - r, r:w - r]
        cy = cy.ravel()
        cx = cx.ravel()
        # Skip windows whose pixels are all known (cf_laplacian.py:34-35).
        all_known = np.ones(cy.shape, dtype=np.bool_)
        for dy, dx in _window_sum_order(r):
            all_known &= is_known[cy + dy - r, cx + dx - r]
        cy = cy[~all_known]
        cx = cx[~all_known]

        # Centred window colours (cf_laplacian.py:38-50).
        c = np.zeros((size, size, 3, cy.size))
        for ch in range(3):
            s = np.zeros(cy.size)
            for dy, dx in _window_sum_order(r):
                s += image[cy + dy - r, cx + dx - r, ch]
            for dy, dx in _window_sum_order(r):
                c[dy, dx, ch] = image[cy + dy - r, cx + dx - r, ch] - s / area

        # Covariance with epsilon on the diagonal (cf_laplacian.py:53-74).
        a00 = np.full(cy.size, epsilon)
        a01 = np.zeros(cy.size)
        a02 = np.zeros(cy.size)
        a11 = np.full(cy.size, epsilon)
        a12 = np.zeros(cy.size)
        a22 = np.full(cy.size, epsilon)
        for dy, dx in _window_sum_order(r):
            c0, c1, c2 = c[dy, dx, 0], c[dy, dx, 1], c[dy, dx, 2]
            a00 += c0 * c0
            a01 += c0 * c1
            a02 += c0 * c2
            a11 += c1 * c1
            a12 += c1 * c2
            a22 += c2 * c2
        a00 /= area
        a01 /= area
        a02 /= area
        a11 /= area
        a12 /= area
        a22 /= area

        # Cofactor inverse with the negated determinant (cf_laplacian.py:76-92).
        det = (a00 * a12 * a12 + a01 * a01 * a22 + a02 * a02 * a11
               - a00 * a11 * a22 - 2 * a01 * a02 * a12)
        inv_det = 1.0 / det
        m00 = (a12 * a12 - a11 * a22) * inv_det
        m01 = (a01 * a22 - a02 * a12) * inv_det
        m02 = (a02 * a11 - a01 * a12) * inv_det
        m11 = (a02 * a02 - a00 * a22) * inv_det
        m12 = (a00 * a12 - a01 * a02) * inv_det
        m22 = (a01 * a01 - a00 * a11) * inv_det

        # Pair contributions (cf_laplacian.py:95-129). The outer loop runs the
        # pixel-in-window offset backwards so that, for a fixed slot, windows
        # arrive in raster order of their centre.
        for dyi in range(size - 1, -1, -1):
            for dxi in range(size - 1, -1, -1):
                s_, t_, u_ = c[dyi, dxi, 0], c[dyi, dxi, 1], c[dyi, dxi, 2]
                c0 = m00 * s_ + m01 * t_ + m02 * u_
                c1 = m01 * s_ + m11 * t_ + m12 * u_
                c2 = m02 * s_ + m12 * t_ + m22 * u_
                i = (cx + dxi - r) + (cy + dyi - r) * w
                for dyj in range(size):
                    for dxj in range(size):
                        temp = c0 * c[dyj, dxj, 0] + c1 * c[dyj, dxj, 1] + c2 * c[dyj, dxj, 2]
                        delta = 1.0 if (dyi == dyj and dxi == dxj) else 0.0
                        value = delta - (1 + temp) / area
                        values[i, dyj - dyi + 2 * r, dxj - dxi + 2 * r] += value

    return scipy.sparse.csr_matrix((values.ravel(), indices.ravel(), indptr), (n, n))


def cf_laplacian_dense_bruteforce(image, epsilon=1e-7, radius=1, is_known=None):
    """Independent dense assembly of the Levin et al. (2007) matting Laplacian.

    L = sum over windows k of (I_k - (1/|w|) (1 + (I_i - mu_k)^T
    (Sigma_k + eps/|w| I)^-1 (I_j - mu_k))) scattered into rows/cols of the
    window, with Sigma_k the biased window covariance and the inverse taken by
    numpy.linalg.inv. Used only to pin ``cf_laplacian`` on tiny images.
    """
    h, w, _ = image.shape
    r = radius
    size = 2 * r + 1
    area = size * size
    n = h * w
    L = np.zeros((n, n))
    if is_known is None:
        is_known = np.zeros((h, w), dtype=bool)
    is_known = np.asarray(is_known).reshape(h, w)
    for y in range(r, h - r):
        for x in range(r, w - r):
            if is_known[y - r:y + r + 1, x - r:x + r + 1].all():
                continue
            ys, xs = np.mgrid[y - r:y + r + 1, x - r:x + r + 1]
            idx = (xs + ys * w).ravel()
            I = image[ys, xs].reshape(area, 3)
            mu = I.mean(axis=0)
            C = I - mu
            Sigma = C.T @ C / area + epsilon / area * np.eye(3)
            G = np.eye(area) - (1.0 + C @ np.linalg.inv(Sigma) @ C.T) / area
            L[np.ix_(idx, idx)] += G
    return L


def make_linear_system(L, trimap, lambda_value=100.0, return_c=False):
    """pymatting/laplacian/laplacian.py:5-60: A = L + lambda*diag(known), b = lambda*fg."""
    is_fg, is_bg, is_known, is_unknown = trimap_split(trimap)
    c = lambda_value * is_known
    b = lambda_value * is_fg
    A = (L + scipy.sparse.diags(c)).tocsr()
    A.sum_duplicates()
    if return_c:
        return A, b, c
    return A, b


def sanity_check_image(image):
    """pymatting/util/util.py:528-578 (warnings only)."""
    if len(image.shape) != 3 or image.shape[2] != 3:
        warnings.warn("Expected RGB image of shape (?, ?, 3), but image.shape is %s." % str(image.shape), stacklevel=3)
    if image.min() < 0.0:
        warnings.warn("Image values should be in [0, 1], but image.min() is %s." % image.min(), stacklevel=3)
    if image.max() > 1.0:
        warnings.warn("Image values should be in [0, 1], but image.max() is %s." % image.max(), stacklevel=3)
    if image.dtype not in [np.float32, np.float64]:
        warnings.warn("Unexpected image.dtype %s. Are you sure that you do not want to use "
                      "np.float32 or np.float64 instead?" % image.dtype, stacklevel=3)


def estimate_alpha_cf(image, trimap, preconditioner=None, laplacian_kwargs={}, cg_kwargs={}):
    """pymatting/alpha/estimate_alpha_cf.py:8-78: hard-constrained closed-form alpha.

    Solves L_U x_U = -R m over the unknown pixels with the reference's
    default ichol preconditioner, keeps known pixels at their trimap value
    and clips to [0, 1].
    """
    from oracle.solver import cg
    if preconditioner is None:
        from oracle.ichol import ichol
        preconditioner = ichol
    sanity_check_image(image)
    is_fg, is_bg, is_known, is_unknown = trimap_split(trimap)
    L = cf_laplacian(image, **laplacian_kwargs, is_known=is_known)
    L_U = L[is_unknown, :][:, is_unknown]
    R = L[is_unknown, :][:, is_known]
    m = is_fg[is_known]
    x = trimap.copy().ravel()
    x[is_unknown] = cg(L_U, -R.dot(m), M=preconditioner(L_U), **cg_kwargs)
    return np.clip(x, 0, 1).reshape(trimap.shape)


def solve_alpha_lambda(image, trimap, epsilon=1e-7, lambda_value=100.0, rtol=1e-7, atol=0.0,
                       maxiter=10000, preconditioner="vcycle", callback=None):
    """The lambda-penalised solve used by the benchmark workloads.

    cf_laplacian -> make_linear_system -> cg(M=vcycle(A, shape)) -> clip, the
    recipe of pymatting's tests/test_preconditioners.py:24-45 and
    examples/expert_example.py:14-41.
    """
    from oracle.solver import cg, vcycle, jacobi
    L = cf_laplacian(image, epsilon=epsilon)
    A, b = make_linear_system(L, trimap, lambda_value)
    if preconditioner == "vcycle":
        M = vcycle(A, trimap.shape)
    elif preconditioner == "jacobi":
        M = jacobi(A)
    else:
        M = None
    x = cg(A, b, rtol=rtol, atol=atol, maxiter=maxiter, M=M, callback=callback)
    return np.clip(x, 0, 1).reshape(trimap.shape), x
