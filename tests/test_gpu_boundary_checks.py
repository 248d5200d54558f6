"""Argument checks of the drop-in API that run before any upload or launch,
and matrices edited after cf_laplacian returned them.

cg checks every size against len(b) first (the reference raises ValueError
from A.dot on a mismatch): the operator of a device matrix, a scipy or
numpy A, x0, and a Jacobi, V-cycle or matrix preconditioner. A matrix from
cf_laplacian that is edited in place (L *= 2, L.data[k] = v, setdiag) no
longer matches its device operator and is solved as the CSR matrix it now
is."""
import numpy as np
import pytest
import scipy.sparse

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _system():
    d = load_golden("lemur_0p125")
    return d["image"], d["trimap"]


def test_cg_rejects_every_size_mismatch():
    import pymatting_amd as pm
    A = scipy.sparse.random(20, 20, density=0.3, random_state=1) + 20 * scipy.sparse.eye(20)
    A = (A + A.T).tocsr()
    b = np.ones(20)
    with pytest.raises(ValueError, match="dimension mismatch"):
        pm.cg(A, np.ones(19))
    with pytest.raises(ValueError, match="x0 of size 21"):
        pm.cg(A, b, x0=np.zeros(21))
    with pytest.raises(ValueError, match="preconditioner of size 19"):
        pm.cg(A, b, M=pm.jacobi(A[:19, :19]))
    with pytest.raises(ValueError, match="preconditioner matrix"):
        pm.cg(A, b, M=scipy.sparse.eye(18).tocsr())
    with pytest.raises(ValueError, match="matrix of shape"):
        pm.cg(np.eye(20)[:, :19], b)
    image, trimap = _system()
    L = pm.cf_laplacian(image)
    with pytest.raises(ValueError, match="operator of size"):
        pm.cg(L, np.ones(L.shape[0] - 1))
    A2, b2 = pm.make_linear_system(L, trimap)
    with pytest.raises(ValueError, match="preconditioner of size 4"):
        pm.cg(A2, b2, M=pm.vcycle(2.0 * np.eye(4), (2, 2)))
    J = pm.jacobi(A)
    with pytest.raises(ValueError, match="Jacobi preconditioner of size 20"):
        J(np.ones(7))
    # Correct sizes still solve.
    x = pm.cg(A, b, M=J, rtol=1e-12)
    assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)


def test_in_place_edits_drop_the_device_operator():
    import pymatting_amd as pm
    from pymatting_amd.laplacian import device_operator
    image, trimap = _system()
    A, b = pm.make_linear_system(pm.cf_laplacian(image), trimap)
    assert device_operator(A) is not None
    x_dev = pm.cg(A, b, rtol=1e-10, M=pm.jacobi(A))
    plain = scipy.sparse.csr_matrix(A, copy=True)  # no operator: the CSR path
    x_csr = pm.cg(plain, b, rtol=1e-10, M=pm.jacobi(plain))
    assert np.abs(x_dev - x_csr).max() <= 1e-8
    A *= 2.0
    assert device_operator(A) is None
    x2 = pm.cg(A, b, rtol=1e-10, M=pm.jacobi(A))
    assert np.abs(2.0 * x2 - x_csr).max() <= 1e-8
    B, _ = pm.make_linear_system(pm.cf_laplacian(image), trimap)
    k = B.indptr[5] + int(np.flatnonzero(B.indices[B.indptr[5]:B.indptr[6]] == 5)[0])
    B.data[k] += 1e-3  # the diagonal of row 5: B stays symmetric
    assert device_operator(B) is None
    ref = scipy.sparse.csr_matrix(B, copy=True)
    assert np.abs(pm.cg(B, b, rtol=1e-10) - pm.cg(ref, b, rtol=1e-10)).max() <= 1e-8
