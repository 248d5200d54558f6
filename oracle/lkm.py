"""Large-kernel matting Laplacian on the CPU (test infrastructure).

Restates pymatting/laplacian/lkm_laplacian.py:6-76 and
pymatting/alpha/estimate_alpha_lkm.py:8-64 with numpy, and the running-sum
box filter of pymatting/util/boxfilter.py:6-155 (valid and full modes; the
rows' running sums are computed for all rows at once, each row in the
reference's order).
"""
import numpy as np


def _rows_valid(src, r):
    m, n = src.shape
    dst = np.zeros((m, n - 2 * r))
    s = np.zeros(m)
    for j in range(2 * r + 1):
        s = s + src[:, j]
    dst[:, 0] = s
    for j in range(1, dst.shape[1]):
        s = s - src[:, j - 1]
        s = s + src[:, j + 2 * r]
        dst[:, j] = s
    return dst


def _rows_full(src, r):
    m, n = src.shape
    dst = np.zeros((m, n + 2 * r))
    s = np.zeros(m)
    s = s + src[:, 0]
    dst[:, 0] = s
    for j in range(1, 2 * r + 1):
        s = s + src[:, j]
        dst[:, j] = s
    for j in range(2 * r + 1, dst.shape[1] - 2 * r):
        s = s - src[:, j - 2 * r - 1]
        s = s + src[:, j]
        dst[:, j] = s
    for j in range(dst.shape[1] - 2 * r, dst.shape[1]):
        s = s - src[:, j - 2 * r - 1]
        dst[:, j] = s
    return dst


def boxfilter(src, radius, mode):
    """boxfilter.py:89-155 for 'valid' and 'full' (columns first, then rows),
    channel by channel."""
    assert radius > 0 and mode in ("valid", "full")
    assert src.shape[0] >= 2 * radius + 1 and src.shape[1] >= 2 * radius + 1
    rows = _rows_valid if mode == "valid" else _rows_full
    if src.ndim == 2:
        return rows(rows(src.T, radius).T, radius)
    return np.stack([boxfilter(src[..., c], radius, mode) for c in range(src.shape[-1])],
                    axis=-1)


def lkm_laplacian(image, epsilon=1e-7, radius=10, return_diagonal=True):
    """lkm_laplacian.py:6-76."""
    image = image.astype(np.float64)
    area = (2 * radius + 1) ** 2
    h, w, depth = image.shape
    means = boxfilter(image, radius, "valid") / area
    outer = np.einsum("...i,...j->...ij", image, image).reshape(h, w, depth * depth)
    covs = boxfilter(outer, radius, "valid").reshape(h - 2 * radius, w - 2 * radius, depth,
                                                      depth) / area
    covs -= np.einsum("...i,...j->...ij", means, means)
    V = np.linalg.inv(covs + epsilon / area * np.eye(depth)) / area
    Vm = np.einsum("...ij,...j->...i", V, means)
    mVm = 1 / area + np.einsum("...i,...i->...", means, Vm)
    c = boxfilter(np.ones((h - 2 * radius, w - 2 * radius)), radius, "full")

    def L_matvec(p):
        p = p.reshape(h, w)
        p_sums = boxfilter(p, radius, "valid")
        pI_sums = boxfilter(p[:, :, None] * image, radius, "valid")
        p_L = c * p
        temp = p_sums[:, :, None] * Vm - np.einsum("...ij,...j->...i", V, pI_sums)
        p_L += np.einsum("...i,...i->...", image, boxfilter(temp, radius, "full"))
        temp = p_sums * mVm - np.einsum("...i,...i->...", pI_sums, Vm)
        p_L -= boxfilter(temp, radius, "full")
        return p_L.flatten()

    if return_diagonal:
        diag = boxfilter(1.0 - mVm, radius, "full")
        Vf = boxfilter(V.reshape(h - 2 * radius, w - 2 * radius, depth * depth), radius,
                       "full").reshape(h, w, depth, depth)
        temp = 2 * boxfilter(Vm, radius, "full") - np.einsum("...ij,...j->...i", Vf, image)
        diag += np.einsum("...i,...i->...", image, temp)
        return L_matvec, diag.flatten()
    return L_matvec


def estimate_alpha_lkm(image, trimap, laplacian_kwargs={}, cg_kwargs={}):
    """estimate_alpha_lkm.py:8-64."""
    from oracle.cf import trimap_split
    from oracle.solver import cg
    L_matvec, diag_L = lkm_laplacian(image, **laplacian_kwargs)
    is_fg, is_bg, is_known, is_unknown = trimap_split(trimap)
    lambda_value = 100.0
    c = lambda_value * is_known
    b = lambda_value * is_fg
    inv_diag_A = 1.0 / (diag_L + c)
    x = cg(lambda x: L_matvec(x) + c * x, b, M=lambda x: inv_diag_A * x, **cg_kwargs)
    return np.clip(x, 0, 1).reshape(trimap.shape)
