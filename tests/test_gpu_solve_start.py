"""The start of a solve without a host round trip: ||b|| and the initial
residual test (cg.py:66-76) are finalised on the device and read with the
first batch's state. The reference's immediate return (x0 already within
tolerance, zero iterations), maxiter = 0, a zero right-hand side of the
reliable-update solve, and the trimap errors that LambdaSolver raises after
enqueueing its operator's set-up."""
import ctypes

import numpy as np
import pytest
import torch

import oracle.solver as osol

pytestmark = pytest.mark.gpu


def _diag_system(n=64, seed=3):
    rng = np.random.default_rng(seed)
    d = 1.0 + rng.random(n)
    x_true = rng.random(n)
    return np.diag(d), d * x_true, x_true


@pytest.mark.parametrize("precond", [None, "jacobi"])
def test_initial_guess_within_tolerance_returns_at_once(precond):
    import pymatting_amd as pm
    A, b, x_true = _diag_system()
    M = pm.jacobi(A) if precond else None
    # The reference returns x0 untouched when ||b - A x0|| < rtol ||b||.
    x0 = x_true + 1e-12
    ref = osol.cg(A, b, x0=x0, rtol=1e-7, M=None if M is None else (lambda v: v / np.diag(A)))
    x, st = pm.cg(A, b, x0=x0, rtol=1e-7, M=M, return_stats=True)
    assert st["iterations"] == 0
    assert np.array_equal(x, x0) and np.array_equal(ref, x0)
    assert st["norm_b"] == pytest.approx(np.linalg.norm(b), rel=1e-15)
    # Also with maxiter = 0; and a guess outside the tolerance then raises.
    x, st = pm.cg(A, b, x0=x0, rtol=1e-7, M=M, maxiter=0, return_stats=True)
    assert st["iterations"] == 0 and np.array_equal(x, x0)
    with pytest.raises(ValueError, match="did not converge within 0 iterations"):
        pm.cg(A, b, rtol=1e-7, M=M, maxiter=0)


def test_norm_b_and_iterations_from_the_first_batch():
    """||b|| reported by the solve is the device's; the iteration count of a
    solve that needs several batches equals the oracle's."""
    import pymatting_amd as pm
    rng = np.random.default_rng(7)
    n = 60
    Q = rng.random((n, n))
    A = Q @ Q.T + n * np.eye(n)
    b = rng.random(n)
    x, st = pm.cg(A, b, rtol=1e-12, return_stats=True)
    assert np.linalg.norm(A @ x - b) < 1e-10 * np.linalg.norm(b)
    assert st["norm_b"] == pytest.approx(np.linalg.norm(b), rel=1e-14)
    counter = pm.CounterCallback()
    pm.cg(A, b, rtol=1e-12, callback=counter)  # the per-iteration (host) path
    assert st["iterations"] == counter.n


@pytest.mark.parametrize("maxiter", [0, 100])
def test_reliable_solve_of_a_zero_rhs(maxiter):
    import pymatting_amd as pm
    from pymatting_amd import _ffi, synthetic
    image, trimap, _ = synthetic.make_image(40, 56, seed=2, erode=3)
    s = pm.LambdaSolver(torch.from_numpy(image).cuda(), torch.from_numpy(trimap).cuda(),
                        precision="mixed", level0="stencil")
    n = s.h * s.w
    b = torch.zeros(n, dtype=torch.float64, device="cuda")
    x = torch.full((n,), 7.0, dtype=torch.float64, device="cuda")
    stats = _ffi.CgStats()
    _ffi.check(_ffi.lib().hm_cg_solve_reliable(None, s.op.handle, s.M.handle, _ffi.ptr(b),
                                               _ffi.ptr(x), 1e-7, 0.01, maxiter, 8, 0,
                                               ctypes.byref(stats), _ffi.stream()))
    assert stats.converged == 1 and stats.iterations == 0 and stats.norm_b == 0.0
    assert float(x.abs().max()) == 0.0


def test_lambda_solver_trimap_errors_after_setup_is_enqueued():
    import pymatting_amd as pm
    from pymatting_amd import synthetic
    image, trimap, _ = synthetic.make_image(40, 56, seed=4, erode=3)
    im = torch.from_numpy(image).cuda()
    t = torch.from_numpy(trimap).cuda()
    with pytest.raises(ValueError, match="did not contain background"):
        pm.LambdaSolver(im, torch.where(t <= 0.1, 0.5, t), level0="stencil")
    with pytest.raises(ValueError, match="did not contain foreground"):
        pm.LambdaSolver(im, torch.where(t >= 0.9, 0.5, t), precision="mixed")
    # The same solver still works afterwards on a valid trimap.
    s = pm.LambdaSolver(im, t, precision="mixed", level0="stencil")
    a = s.solve(rtol=1e-7, mixed="reliable")
    assert s.stats["iterations"] > 0 and float(a.min()) >= 0.0 and float(a.max()) <= 1.0
