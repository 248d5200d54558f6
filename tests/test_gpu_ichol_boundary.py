"""GPU parity of ichol (pymatting/preconditioner/ichol.py) and of the
drop-in boundary of estimate_alpha_cf, make_linear_system and
cf_laplacian(radius) against the oracle.

- ichol: the docstring example (ichol.py:231-237); the exact solve of
  tests/test_ichol.py; the library's substitutions bit-identical to the
  oracle's; the iteration ceiling of
  tests/test_preconditioners.py (icholt <= 3 at lemur 0.125).
- estimate_alpha_cf with preconditioner factories: the factory receives the
  reference's reduced matrix L_U (bit-identical to the oracle's
  L[U, :][:, U], with b_U = -R m), and ichol, a factory calling tocsc(), a
  matrix and a callable M all reach the oracle's alpha.
- make_linear_system of a user-built or edited Laplacian: the canonical CSR
  of scipy's L + diags(c), exactly.
- cf_laplacian(radius=2): bit-identical to the oracle (tests/test_lkm.py
  uses radius 2), and solved by cg through the device CSR.
"""
import numpy as np
import pytest
import scipy.sparse

from conftest import load_golden
import oracle.cf as ocf
import oracle.ichol as oic
import oracle.solver as osol

pytestmark = pytest.mark.gpu


def _lemur(scale="0p125"):
    d = load_golden("lemur_" + scale)
    return d["image"], d["trimap"]


def test_ichol_docstring_and_exact_solve():
    import pymatting_amd as pm
    A = np.array([[2.0, 3.0], [3.0, 5.0]])
    assert np.allclose(pm.ichol(scipy.sparse.csc_matrix(A))(np.array([1.0, 2.0])), [-1.0, 1.0])
    np.random.seed(0)  # tests/test_ichol.py
    for _ in range(10):
        n = 50
        A = np.random.rand(n, n)
        A[np.random.rand(n, n) < 0.8] = 0
        A += A.T + n * np.eye(n)
        A = scipy.sparse.csc_matrix(A)
        dec = pm.ichol(A, discard_threshold=0.0, max_nnz=100000)
        x_true = np.random.rand(n)
        b = A.dot(x_true)
        x = dec(b)
        assert np.linalg.norm(x - x_true) < 1e-10
        ref = oic.ichol(A, discard_threshold=0.0, max_nnz=100000)
        assert all(np.array_equal(u, v) for u, v in zip(dec.Ltuple, ref.Ltuple))
        assert np.array_equal(x, ref(b))


def test_ichol_apply_bit_identical_on_lemur_system():
    import pymatting_amd as pm
    image, trimap = _lemur()
    A, b = ocf.make_linear_system(ocf.cf_laplacian(image), trimap)
    dec = pm.ichol(A, max_nnz=500000)
    ref = oic.ichol(A, max_nnz=500000)
    assert all(np.array_equal(u, v) for u, v in zip(dec.Ltuple, ref.Ltuple))
    r = np.random.default_rng(3).standard_normal(A.shape[0])
    assert np.array_equal(dec(r), ref(r))
    # A factor given as a tuple (CholeskyDecomposition(Ltuple)) applies alike.
    again = pm.CholeskyDecomposition(ref.Ltuple)
    assert np.array_equal(again(r), ref(r))
    assert (dec.L - ref.L).nnz == 0


def test_ichol_iteration_ceiling():
    """tests/test_preconditioners.py:24-61, icholt: at most 3 iterations."""
    import pymatting_amd as pm
    image, trimap = _lemur()
    A, b = pm.make_linear_system(pm.cf_laplacian(image), trimap)
    counter = pm.CounterCallback()
    x = pm.cg(A, b, M=pm.ichol(A, max_nnz=500000), atol=1e-6, rtol=0, maxiter=10000,
              callback=counter)
    assert np.linalg.norm(b - A.dot(x)) <= 1e-6
    assert counter.n <= 3
    Ar, br = ocf.make_linear_system(ocf.cf_laplacian(image), trimap)
    c_ref = osol.CounterCallback()
    osol.cg(Ar, br, M=oic.ichol(Ar, max_nnz=500000), atol=1e-6, rtol=0, callback=c_ref)
    assert counter.n == c_ref.n


def test_reduced_system_matches_fancy_indexing():
    import torch
    from pymatting_amd.alpha import reduced_system
    from pymatting_amd.util import device_classes
    for image, trimap in (_lemur("0p1"), _lemur()):
        classes, _ = device_classes(trimap)
        L_U, b_U = reduced_system(image, classes)
        fg, bg, known, unknown = ocf.trimap_split(trimap)
        L = ocf.cf_laplacian(image, is_known=known)
        ref_U = L[unknown, :][:, unknown]
        R = L[unknown, :][:, known]
        assert np.array_equal(L_U.indptr, ref_U.indptr)
        assert np.array_equal(L_U.indices, ref_U.indices)
        assert np.array_equal(L_U.data, ref_U.data)
        assert np.array_equal(b_U.cpu().numpy(), -R.dot(fg[known]))


@pytest.mark.parametrize("name", ["ichol", "tocsc_ichol", "matrix", "callable", "default"])
def test_estimate_alpha_cf_factories(name):
    import pymatting_amd as pm
    image, trimap = _lemur("0p1")
    tight = {"rtol": 1e-10}
    factories = {
        "ichol": (pm.ichol, oic.ichol),
        "tocsc_ichol": (lambda A: pm.ichol(A.tocsc()), lambda A: oic.ichol(A.tocsc())),
        "matrix": (lambda A: scipy.sparse.diags(1.0 / A.diagonal()),
                   lambda A: scipy.sparse.diags(1.0 / A.diagonal())),
        "callable": (lambda A: (lambda r, d=A.diagonal(): r / d),
                     lambda A: (lambda r, d=A.diagonal(): r / d)),
        "default": (None, None),
    }
    ours, ref = factories[name]
    alpha = pm.estimate_alpha_cf(image, trimap, preconditioner=ours, cg_kwargs=tight)
    alpha_ref = ocf.estimate_alpha_cf(image, trimap, preconditioner=ref, cg_kwargs=tight)
    assert np.abs(alpha - alpha_ref).max() <= 1e-6
    # tests/test_estimate_alpha.py:17-18 at the default tolerance.
    d = load_golden("lemur_0p1")
    alpha = pm.estimate_alpha_cf(image, trimap, preconditioner=ours)
    assert np.linalg.norm(alpha - d["good_alpha"]) < 1.68


def test_make_linear_system_of_any_laplacian():
    import pymatting_amd as pm
    image, trimap = _lemur()
    # A Laplacian the package did not build (the oracle's plain CSR, with
    # the reference's duplicates and explicit zeros), and a random one.
    L = ocf.cf_laplacian(image)
    rng = np.random.default_rng(2)
    n = L.shape[0]
    R = scipy.sparse.random(n, n, density=3e-4, random_state=4, format="csr")
    R = (R + R.T).tocsr()
    for Lx in (L, R, scipy.sparse.coo_matrix(L), scipy.sparse.csc_matrix(R)):
        A, b = pm.make_linear_system(Lx, trimap, lambda_value=37.0)
        Ar, br = ocf.make_linear_system(Lx, trimap, lambda_value=37.0)
        assert np.array_equal(A.indptr, Ar.indptr) and np.array_equal(A.indices, Ar.indices)
        assert np.array_equal(A.data, Ar.data)
        assert np.array_equal(b, br)
    # An edited cf_laplacian goes the same way.
    Lp = pm.cf_laplacian(image)
    Lp.data *= 1.5
    A, _ = pm.make_linear_system(Lp, trimap)
    Ar, _ = ocf.make_linear_system(ocf.cf_laplacian(image) * 1.5, trimap)
    assert np.array_equal(A.data, Ar.data) and np.array_equal(A.indices, Ar.indices)
    # The system solves on the device.
    x = pm.cg(A, np.asarray(trimap >= 0.9, float) * 100.0, rtol=1e-9, M=pm.jacobi(A))
    assert np.linalg.norm(A @ x - 100.0 * (trimap.ravel() >= 0.9)) <= 1e-8 * np.linalg.norm(
        100.0 * (trimap.ravel() >= 0.9))


def test_cf_laplacian_radius_two():
    import pymatting_amd as pm
    from pymatting_amd.laplacian import _cf_laplacian_radius, attached_csr
    image, trimap = _lemur()
    for r in (1, 2, 3):
        L = pm.cf_laplacian(image, radius=r) if r != 1 else _cf_laplacian_radius(image, 1e-7, 1,
                                                                                 None)
        Lr = ocf.cf_laplacian(image, radius=r)
        assert L.shape == Lr.shape
        assert np.array_equal(L.indptr, Lr.indptr) and np.array_equal(L.indices, Lr.indices)
        assert np.array_equal(L.data, Lr.data), r
    known = (trimap <= 0.1) | (trimap >= 0.9)
    L = pm.cf_laplacian(image, radius=2, is_known=known)
    assert np.array_equal(L.data, ocf.cf_laplacian(image, radius=2, is_known=known).data)
    # tests/test_lkm.py:24-36: the radius-2 system with Jacobi; the device
    # CSR carries the solve.
    A, b = pm.make_linear_system(pm.cf_laplacian(image, radius=2), trimap)
    assert attached_csr(A) is not None
    x = pm.cg(A, b, M=pm.jacobi(A), rtol=1e-10)
    Ar, br = ocf.make_linear_system(ocf.cf_laplacian(image, radius=2), trimap)
    xr = osol.cg(Ar, br, M=osol.jacobi(Ar), rtol=1e-10)
    assert np.abs(x - xr).max() <= 1e-7
