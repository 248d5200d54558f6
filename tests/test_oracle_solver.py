"""Pins the oracle's cg, jacobi, make_P, vcycle and ichol.

Checks: docstring examples of the reference (cg.py:44-50, jacobi.py:17-22,
vcycle.py:134-142, ichol.py:231-237), the random SPD test of
tests/test_cg.py, exactness of ichol without dropping (tests/test_ichol.py),
the iteration ceilings of tests/test_preconditioners.py:33-38, a loop form of
the restriction, and symmetry/linearity of the V(1,1) operator.
"""
import numpy as np
import pytest
import scipy.sparse

from conftest import load_golden
from oracle.cf import cf_laplacian, make_linear_system
from oracle.ichol import ichol
from oracle.solver import CounterCallback, cg, jacobi, make_P, make_P_bruteforce, vcycle


def test_docstring_examples():
    A = np.array([[3.0, 1.0], [1.0, 2.0]])
    assert np.allclose(cg(A, np.array([4.0, 3.0]), M=jacobi(A)), [1.0, 1.0])
    B = np.array([[2, 3], [3, 5]])
    assert np.allclose(jacobi(B)(np.array([1, 2])), [0.5, 0.4])
    assert np.allclose(vcycle(scipy.sparse.csr_matrix(B.astype(float)), (2, 2))(np.array([1.0, 2.0])), [-1.0, 1.0])
    assert np.allclose(ichol(scipy.sparse.csc_matrix(B.astype(float)))(np.array([1.0, 2.0])), [-1.0, 1.0])


def test_cg_random_spd():
    np.random.seed(0)
    for n in [1, 5, 10]:
        x_true = np.random.rand(n)
        A = np.random.rand(n, n)
        A += A.T + n * np.eye(n)
        b = A.dot(x_true)
        Mj = np.diag(1.0 / np.diag(A))
        for x0 in [None, np.random.rand(n)]:
            for M in [None, Mj, lambda v: Mj.dot(v)]:
                cb = CounterCallback()
                x = cg(A, b, x0=x0, M=M, rtol=0, atol=1e-10, callback=cb)
                assert np.linalg.norm(A.dot(x) - b) < 1e-10
                assert np.allclose(x, np.linalg.solve(A, b), atol=1e-9)
                assert 0 < cb.n < 20 or np.linalg.norm(b - A.dot(x0 if x0 is not None else 0 * b)) < 1e-10


def test_cg_maxiter():
    A = np.diag(np.arange(1.0, 50.0))
    with pytest.raises(ValueError, match="did not converge within 3 iterations"):
        cg(A, np.ones(49), maxiter=3)


def test_ichol_exact():
    np.random.seed(0)
    for _ in range(10):
        n = 50
        A = np.random.rand(n, n)
        A[np.random.rand(n, n) < 0.8] = 0
        A += A.T + n * np.eye(n)
        A = scipy.sparse.csc_matrix(A)
        dec = ichol(A, discard_threshold=0.0, max_nnz=100000)
        x_true = np.random.rand(n)
        assert np.linalg.norm(dec(A.dot(x_true)) - x_true) < 1e-10


@pytest.mark.parametrize("shape", [(4, 4), (5, 7), (8, 3), (9, 9)])
def test_make_P(shape):
    P, R = make_P(shape)
    assert np.array_equal(R.toarray(), make_P_bruteforce(shape))
    assert np.array_equal(P.toarray(), R.toarray().T)


def test_vcycle_operator_is_symmetric_and_linear():
    rng = np.random.default_rng(1)
    d = load_golden("lemur_0p125")
    image, trimap = d["image"][:24, :30], d["trimap"][:24, :30].copy()
    trimap[0, 0], trimap[-1, -1] = 1.0, 0.0
    A, b = make_linear_system(cf_laplacian(image), trimap)
    M = vcycle(A, trimap.shape)
    u, v = rng.random(A.shape[0]), rng.random(A.shape[0])
    assert abs(u @ M(v) - v @ M(u)) < 1e-9 * abs(u @ M(v))
    assert np.allclose(M(2.0 * u + v), 2.0 * M(u) + M(v), rtol=1e-12, atol=1e-12)
    # With a direct size at least n the V-cycle is an exact solve.
    Mx = vcycle(A, trimap.shape, direct_solve_size=10 ** 9)
    assert np.linalg.norm(A.dot(Mx(b)) - b) < 1e-8 * np.linalg.norm(b)


def test_preconditioner_iteration_ceilings():
    d = load_golden("lemur_0p125")
    image, trimap = d["image"], d["trimap"]
    A, b = make_linear_system(cf_laplacian(image), trimap)
    expected = {"none": 378, "jacobi": 197, "icholt": 3, "vcycle": 48}
    factories = {"none": lambda A: None, "jacobi": jacobi,
                 "icholt": lambda A: ichol(A, max_nnz=500000),
                 "vcycle": lambda A: vcycle(A, trimap.shape)}
    for name, make in factories.items():
        cb = CounterCallback()
        x = cg(A, b, M=make(A), atol=1e-6, rtol=0, maxiter=10000, callback=cb)
        assert np.linalg.norm(b - A.dot(x)) <= 1e-6
        assert cb.n <= expected[name], name
