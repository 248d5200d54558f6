"""CPU restatement of PCG, the Jacobi preconditioner and the geometric V-cycle.

Test infrastructure (see ``oracle/__init__.py``).
"""
import numpy as np
import scipy.sparse
import scipy.sparse.linalg


class CounterCallback:
    """pymatting/solver/callback.py:1-8: counts iterations."""

    def __init__(self):
        self.n = 0

    def __call__(self, A, x, b, norm_b, r, norm_r):
        self.n += 1


def cg(A, b, x0=None, atol=0.0, rtol=1e-7, maxiter=10000, callback=None, M=None,
       reorthogonalize=False):
    """pymatting/solver/cg.py:4-117: preconditioned conjugate gradients.

    Recursive residual (never recomputed), callback before the convergence
    test, stop when ||r|| < atol or ||r|| < rtol*||b||, ValueError after
    ``maxiter`` iterations. ``reorthogonalize`` selects the Polak-Ribiere beta.
    """
    if M is None:
        apply_m = lambda v: v  # noqa: E731
    elif callable(M):
        apply_m = M
    else:
        apply_m = M.dot
    apply_a = A if callable(A) else A.dot

    x = np.zeros_like(b) if x0 is None else x0.copy()
    norm_b = np.linalg.norm(b)
    r = b - apply_a(x)
    norm_r = np.linalg.norm(r)
    if norm_r < atol or norm_r < rtol * norm_b:
        return x
    z = apply_m(r)
    p = z.copy()
    rz = np.inner(r, z)
    for _ in range(maxiter):
        r_prev = r.copy()
        q = apply_a(p)
        step = rz / np.inner(p, q)
        x += step * p
        r -= step * q
        norm_r = np.linalg.norm(r)
        if callback is not None:
            callback(A, x, b, norm_b, r, norm_r)
        if norm_r < atol or norm_r < rtol * norm_b:
            return x
        z = apply_m(r)
        if reorthogonalize:
            beta = np.inner(r - r_prev, z) / rz
            rz = np.inner(r, z)
        else:
            beta = 1.0 / rz
            rz = np.inner(r, z)
            beta *= rz
        p *= beta
        p += z
    raise ValueError("Conjugate gradient descent did not converge within %d iterations" % maxiter)


def jacobi(A):
    """pymatting/preconditioner/jacobi.py:1-31: x -> x / diag(A)."""
    inv_diag = 1.0 / A.diagonal()
    return lambda v: v * inv_diag


def make_P(shape):
    """pymatting/preconditioner/vcycle.py:6-29: full-weighting restriction R and P = R^T.

    Coarse point (X, Y) of an (h//2, w//2) grid gathers the 3x3 fine
    neighbourhood of (2X, 2Y) with weights [1 2 1; 2 4 2; 1 2 1]/16. Taps
    outside the image are dropped without renormalisation.
    """
    h, w = shape
    h2, w2 = h // 2, w // 2
    n2 = h2 * w2
    wts = np.array([1, 2, 1, 2, 4, 2, 1, 2, 1], dtype=np.float64) / 16
    ox = np.array([-1, 0, 1, -1, 0, 1, -1, 0, 1])
    oy = np.array([-1, -1, -1, 0, 0, 0, 1, 1, 1])
    X = np.repeat(np.tile(np.arange(w2), h2), 9)
    Y = np.repeat(np.repeat(np.arange(h2), w2), 9)
    x = 2 * X + np.tile(ox, n2)
    y = 2 * Y + np.tile(oy, n2)
    ok = (x >= 0) & (x < w) & (y >= 0) & (y < h)
    R = scipy.sparse.csr_matrix((np.tile(wts, n2)[ok], ((X + Y * w2)[ok], (x + y * w)[ok])),
                                (n2, h * w))
    return R.T, R


def make_P_bruteforce(shape):
    """Loop form of the restriction, used only to pin ``make_P``."""
    h, w = shape
    h2, w2 = h // 2, w // 2
    R = np.zeros((h2 * w2, h * w))
    k = np.array([[1, 2, 1], [2, 4, 2], [1, 2, 1]]) / 16.0
    for Y in range(h2):
        for X in range(w2):
            for ey in (-1, 0, 1):
                for ex in (-1, 0, 1):
                    y, x = 2 * Y + ey, 2 * X + ex
                    if 0 <= y < h and 0 <= x < w:
                        R[X + Y * w2, x + y * w] = k[ey + 1, ex + 1]
    return R


def _smooth(A, D, b, x, sweeps, omega):
    """Damped Jacobi (vcycle.py:32-43); ``x is None`` starts from x = omega*b/D."""
    if x is None:
        if sweeps <= 0:
            return np.zeros_like(b)
        x = omega * b / D
        sweeps -= 1
    for _ in range(sweeps):
        x = x + omega * (b - A.dot(x)) / D
    return x


def _vcycle(A, b, shape, cache, pre, post, omega, direct_size):
    """vcycle.py:46-100: one V-cycle applied to ``b``; coarse data cached per shape."""
    h, w = shape
    if h * w <= direct_size:
        return scipy.sparse.linalg.spsolve(A, b)
    if shape not in cache:
        P, R = make_P(shape)
        cache[shape] = (P, R, R.dot(A).dot(P), A.diagonal())
    P, R, Ac, D = cache[shape]
    x = _smooth(A, D, b, None, pre, omega)
    rc = R.dot(b - A.dot(x))
    xc = _vcycle(Ac, rc, (h // 2, w // 2), cache, pre, post, omega, direct_size)
    x += P.dot(xc)
    return _smooth(A, D, b, x, post, omega)


def vcycle(A, shape, num_pre_iter=1, num_post_iter=1, omega=0.8, direct_solve_size=64, cache=None):
    """pymatting/preconditioner/vcycle.py:103-153: returns r -> V-cycle(r)."""
    if cache is None:
        cache = {}
    return lambda r: _vcycle(A, r, tuple(shape), cache, num_pre_iter, num_post_iter, omega,
                             direct_solve_size)
