"""GPU parity of cg, jacobi and vcycle against the oracle and the reference's
known answers (docstrings, tests/test_cg.py, tests/test_preconditioners.py)."""
import numpy as np
import pytest
import scipy.sparse

from conftest import load_golden
import oracle.cf as ocf
import oracle.solver as osol

pytestmark = pytest.mark.gpu


def test_docstring_examples():
    import pymatting_amd as pm
    A = np.array([[3.0, 1.0], [1.0, 2.0]])
    assert np.allclose(pm.cg(A, np.array([4.0, 3.0]), M=pm.jacobi(A)), [1.0, 1.0])
    B = np.array([[2.0, 3.0], [3.0, 5.0]])
    assert np.allclose(pm.jacobi(B)(np.array([1.0, 2.0])), [0.5, 0.4])
    assert np.allclose(pm.vcycle(B, (2, 2))(np.array([1.0, 2.0])), [-1.0, 1.0])


def test_cg_random_spd_all_modes():
    import pymatting_amd as pm
    np.random.seed(0)
    for n in [1, 5, 10]:
        x_true = np.random.rand(n)
        A = np.random.rand(n, n)
        A += A.T + n * np.eye(n)
        b = A.dot(x_true)
        Mj = np.diag(1.0 / np.diag(A))
        for x0 in [None, np.random.rand(n)]:
            for M in [None, Mj, lambda v: Mj.dot(v)]:
                for cb in [None, pm.CounterCallback(), pm.ProgressCallback()]:
                    x = pm.cg(A, b, x0=x0, M=M, rtol=0, atol=1e-10, callback=cb)
                    assert np.linalg.norm(A.dot(x) - b) < 1e-10
                    if cb is not None:
                        assert 0 < cb.n < 20
    # Callable operator.
    A = np.diag(np.arange(1.0, 6.0))
    x = pm.cg(lambda v: A.dot(v), np.ones(5), rtol=1e-12)
    assert np.allclose(x, 1.0 / np.arange(1.0, 6.0))


def test_cg_maxiter_raises():
    import pymatting_amd as pm
    A = np.diag(np.arange(1.0, 50.0))
    with pytest.raises(ValueError, match="did not converge within 3 iterations"):
        pm.cg(A, np.ones(49), maxiter=3)


def _lemur_system():
    d = load_golden("lemur_0p125")
    return d["image"], d["trimap"]


def test_vcycle_matches_oracle():
    import pymatting_amd as pm
    image, trimap = _lemur_system()
    A, b = pm.make_linear_system(pm.cf_laplacian(image), trimap)
    Ar, br = ocf.make_linear_system(ocf.cf_laplacian(image), trimap)
    for pre, post in [(1, 1), (2, 2), (0, 1), (1, 0), (3, 2)]:
        M = pm.vcycle(A, trimap.shape, num_pre_iter=pre, num_post_iter=post)
        Mr = osol.vcycle(Ar, trimap.shape, num_pre_iter=pre, num_post_iter=post)
        rng = np.random.default_rng(pre * 10 + post)
        r = rng.standard_normal(b.size)
        z, zr = M(r), Mr(r)
        # Level 0 is matrix-free on the GPU and CSR in the oracle; the
        # difference is rounding amplified by the operator's conditioning.
        assert np.abs(z - zr).max() <= 1e-8 * np.abs(zr).max(), (pre, post)
    # Stored-stencil input (plain scipy matrix) builds the same hierarchy.
    Md = pm.vcycle(scipy.sparse.csr_matrix(Ar), trimap.shape)
    r = np.random.default_rng(5).standard_normal(b.size)
    zr = osol.vcycle(Ar, trimap.shape)(r)
    assert np.abs(Md(r) - zr).max() <= 1e-8 * np.abs(zr).max()


def test_vcycle_level_cap_and_coarse_sweeps():
    """Extensions with no reference counterpart: max_levels and an approximate
    coarsest solve. A cap at or above the natural depth changes nothing; a cap
    that leaves the coarsest grid above the dense-solve limit needs
    coarse_sweeps, and then gives a symmetric, linear preconditioner with
    which PCG reaches the exact solution of the system."""
    import scipy.sparse.linalg
    import pymatting_amd as pm
    from pymatting_amd import synthetic
    from pymatting_amd.preconditioner import VCycle
    image, trimap, _ = synthetic.make_image(128, 160, 7, strands=6, erode=6)
    A, b = pm.make_linear_system(pm.cf_laplacian(image), trimap)
    op = A.operator
    rng = np.random.default_rng(3)
    r = rng.standard_normal(b.size)
    assert np.array_equal(VCycle(op, trimap.shape)(r), VCycle(op, trimap.shape, max_levels=32)(r))
    with pytest.raises(ValueError, match="coarse_sweeps"):  # coarsest 64 x 80 = 5120 points
        VCycle(op, trimap.shape, max_levels=2)
    Ms = VCycle(op, trimap.shape, num_pre_iter=2, num_post_iter=2, max_levels=2, coarse_sweeps=3)
    u, v = rng.standard_normal(b.size), rng.standard_normal(b.size)
    Mu, Mv = Ms(u), Ms(v)
    assert abs(u @ Mv - v @ Mu) <= 1e-9 * np.linalg.norm(u) * np.linalg.norm(Mv)
    assert np.abs(Ms(2 * u + 3 * v) - (2 * Mu + 3 * Mv)).max() <= 1e-10 * np.abs(Mu).max()
    x = pm.cg(A, b, M=Ms, rtol=1e-11)
    Ar, br = ocf.make_linear_system(ocf.cf_laplacian(image), trimap)
    xr = scipy.sparse.linalg.spsolve(Ar.tocsc(), br)
    assert np.abs(x - xr).max() < 1e-6


def test_iteration_ceilings_and_oracle_counts():
    import pymatting_amd as pm
    image, trimap = _lemur_system()
    A, b = pm.make_linear_system(pm.cf_laplacian(image), trimap)
    Ar, br = ocf.make_linear_system(ocf.cf_laplacian(image), trimap)
    expected = {"none": 378, "jacobi": 197, "vcycle": 48}
    for name, mk, mkr in [("none", lambda A: None, lambda A: None),
                          ("jacobi", pm.jacobi, osol.jacobi),
                          ("vcycle", lambda A: pm.vcycle(A, trimap.shape),
                           lambda A: osol.vcycle(A, trimap.shape))]:
        cb, cbr = pm.CounterCallback(), osol.CounterCallback()
        x = pm.cg(A, b, M=mk(A), atol=1e-6, rtol=0, maxiter=10000, callback=cb)
        xr = osol.cg(Ar, br, M=mkr(Ar), atol=1e-6, rtol=0, maxiter=10000, callback=cbr)
        assert np.linalg.norm(br - Ar.dot(x)) <= 1e-6
        assert cb.n <= expected[name]
        assert abs(cb.n - cbr.n) <= 1, (name, cb.n, cbr.n)
        assert np.abs(x - xr).max() < 1e-6


def test_dia_operator_and_csr_paths_agree():
    import pymatting_amd as pm
    image, trimap = _lemur_system()
    A, b = pm.make_linear_system(pm.cf_laplacian(image), trimap)
    plain = scipy.sparse.csr_matrix(A)  # no device operator: CSR upload path
    x1 = pm.cg(A, b, M=pm.vcycle(A, trimap.shape), rtol=1e-10)
    x2 = pm.cg(plain, b, M=pm.vcycle(plain, trimap.shape), rtol=1e-10)
    assert np.abs(x1 - x2).max() < 1e-8


@pytest.mark.parametrize("n", [1, 7, 28, 64, 65, 100])
def test_direct_coarse_inverse(n):
    """The coarsest-grid dense inverse (Gauss-Jordan with partial pivoting;
    LDS-resident up to 64 unknowns, global memory above): M(A x) = x for a
    random SPD matrix solved directly (vcycle.py:59-60), and a singular
    matrix is reported."""
    from pymatting_amd.preconditioner import VCycle
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, n))
    A = B @ B.T + n * np.eye(n)
    x = rng.standard_normal(n)
    M = VCycle(A, (1, n), direct_solve_size=n)
    got = M(A @ x)
    assert np.abs(got - x).max() <= 1e-10 * np.abs(x).max()
    S = A.copy()  # a zero last row and column stay exactly zero
    S[:, -1] = 0.0
    S[-1, :] = 0.0
    with pytest.raises(ValueError):
        VCycle(S, (1, n), direct_solve_size=n)


def _thin_case(shape):
    rng = np.random.default_rng(shape[0] * 1000 + shape[1])
    image = rng.random(shape + (3,))
    trimap = rng.choice([0.0, 0.5, 1.0], size=shape, p=[0.3, 0.4, 0.3])
    trimap.flat[0], trimap.flat[-1] = 0.0, 1.0
    return image, trimap, rng


@pytest.mark.parametrize("shape", [(3, 200), (200, 3), (12, 700)])
def test_thin_images_reach_an_empty_grid(shape):
    """Coarsening that reaches an empty grid above the direct-solve size
    (3x200 -> 1x100 -> 0x50; 12x700 -> ... -> 1x87 -> 0x43). The reference
    recurses into the empty grid (vcycle.py:59-60 solves the 0x0 system), so
    the last non-empty grid is smoothed with a zero coarse correction. The
    V-cycle, PCG with it, the device solvers and estimate_alpha_cf match the
    oracle."""
    import pymatting_amd as pm
    image, trimap, rng = _thin_case(shape)
    print("shape", shape, flush=True)
    A, b = pm.make_linear_system(pm.cf_laplacian(image), trimap)
    Ar, br = ocf.make_linear_system(ocf.cf_laplacian(image), trimap)
    for pre, post in [(1, 1), (2, 1), (0, 2)]:
        M = pm.vcycle(A, shape, num_pre_iter=pre, num_post_iter=post)
        Mr = osol.vcycle(Ar, shape, num_pre_iter=pre, num_post_iter=post)
        r = rng.standard_normal(b.size)
        z, zr = M(r), Mr(r)
        assert np.abs(z - zr).max() <= 1e-8 * np.abs(zr).max(), (pre, post)
    cb, cbr = pm.CounterCallback(), osol.CounterCallback()
    x = pm.cg(A, b, M=pm.vcycle(A, shape), rtol=1e-10, callback=cb)
    xr = osol.cg(Ar, br, M=osol.vcycle(Ar, shape), rtol=1e-10, callback=cbr)
    assert abs(cb.n - cbr.n) <= 1, (cb.n, cbr.n)
    assert np.abs(x - xr).max() < 1e-6
    ref = np.clip(xr, 0, 1).reshape(shape)
    for prec, level0, kw in [("f64", "stencil", {}), ("f64", "matrix_free", {}),
                             ("mixed", "stencil", {"mixed": "reliable"}),
                             ("mixed", "stencil", {"mixed": "refine", "inner_rtol": 1e-3})]:
        s = pm.LambdaSolver(image, trimap, precision=prec, level0=level0)
        alpha = s.solve(rtol=1e-10, **kw).cpu().numpy()
        assert np.abs(alpha - ref).max() < 1e-6, (prec, level0, kw)
    alpha = pm.estimate_alpha_cf(image, trimap, cg_kwargs={"rtol": 1e-12})
    ref_a1 = ocf.estimate_alpha_cf(image, trimap, cg_kwargs={"rtol": 1e-12})
    assert np.abs(alpha - ref_a1).max() < 1e-6


@pytest.mark.parametrize("shape", [(1, 200), (200, 1)])
def test_single_grid_that_cannot_be_coarsened(shape):
    """A 200-point grid of one row (or column) has an empty coarse grid: the
    V-cycle is the pre- and post-smoothing alone, pre + post damped Jacobi
    sweeps from zero, as in the oracle's recursion. A SPD 1-D operator of
    reach 2 along the grid."""
    import pymatting_amd as pm
    n = 200
    rng = np.random.default_rng(9)
    A = scipy.sparse.diags([6.0 + rng.random(n), -1.5 * np.ones(n - 1), -1.5 * np.ones(n - 1),
                            -0.5 * np.ones(n - 2), -0.5 * np.ones(n - 2)],
                           [0, 1, -1, 2, -2], format="csr")
    for pre, post in [(1, 1), (2, 2), (0, 1), (3, 0)]:
        M = pm.vcycle(A, shape, num_pre_iter=pre, num_post_iter=post)
        Mr = osol.vcycle(A, shape, num_pre_iter=pre, num_post_iter=post)
        r = rng.standard_normal(n)
        z, zr = M(r), Mr(r)
        assert np.abs(z - zr).max() <= 1e-13 * np.abs(zr).max(), (pre, post)
    b = rng.standard_normal(n)
    cb, cbr = pm.CounterCallback(), osol.CounterCallback()
    x = pm.cg(A, b, M=pm.vcycle(A, shape), rtol=1e-12, callback=cb)
    xr = osol.cg(A, b, M=osol.vcycle(A, shape), rtol=1e-12, callback=cbr)
    assert abs(cb.n - cbr.n) <= 1, (cb.n, cbr.n)
    assert np.abs(x - xr).max() < 1e-10 * np.abs(xr).max()
    with pytest.raises(ValueError, match="num_pre_iter"):
        pm.vcycle(A, shape, num_pre_iter=0, num_post_iter=0)
