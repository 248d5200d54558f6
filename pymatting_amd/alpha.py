"""Closed-form alpha estimation on the GPU.

``estimate_alpha_cf`` is the drop-in for pymatting/alpha/estimate_alpha_cf.py:8-78
(hard constraints: L_U x_U = -R m over the unknown pixels). A preconditioner
factory gets the reference's reduced matrix L_U, built on the device; the
default instead solves the masked full-grid operator [L_UU 0; 0 I] with
right-hand side -(R m) on unknown pixels and 0 on known ones, whose solution
restricted to the unknown pixels is the reference's x_U, with the V-cycle on
that operator (the reference's default, ichol, is available as a factory:
its factorisation is sequential).

``solve_alpha_lambda`` is the lambda-penalised solve the benchmark workloads
use: A = L + lambda diag(known), b = lambda is_fg, PCG with the V-cycle, clip
(tests/test_preconditioners.py:24-45, examples/expert_example.py:14-41), in
float64 or in the mixed float32/float64 form.
"""
import ctypes

import numpy as np
import scipy.sparse
import torch

from pymatting_amd import _ffi
from pymatting_amd._device import as_device, empty, is_torch
from pymatting_amd.laplacian import CfOperator, DiaOperator, attached_csr, cf_laplacian
from pymatting_amd.preconditioner import VCycle, ichol, jacobi, vcycle
from pymatting_amd.solver import cg
from pymatting_amd.util import check_trimap, device_classes, sanity_check_image


def _finalize(x, trimap_dev, classes, shape):
    alpha = empty(x.numel())
    _ffi.check(_ffi.lib().hm_alpha_finalize(_ffi.ptr(x), _ffi.ptr(trimap_dev), _ffi.ptr(classes),
                                            _ffi.ptr(alpha), x.numel(), _ffi.stream()))
    return alpha.reshape(shape)


def estimate_alpha_cf(image, trimap, preconditioner=None, laplacian_kwargs={}, cg_kwargs={}):
    """estimate_alpha_cf.py:8-78: closed-form alpha matting (Levin et al. 2007).

    preconditioner: a factory A -> M, called as in the reference with the
    reduced matrix L_U = L[U, :][:, U] of the unknown pixels (a scipy CSR
    matrix assembled on the device, carrying its device copy); its M may be
    a matrix, a callable or one of this package's preconditioners, e.g.
    ``ichol`` (the reference's default), whose factor is applied on the GPU.
    None (or ``vcycle``) selects this package's default, the V-cycle on the
    full grid: PCG then runs on the stored stencil of the masked operator
    [L_UU 0; 0 I] (entries equal to cf_laplacian's bit for bit), whose
    solution on the unknown pixels is the reference's x_U. ``jacobi`` runs
    on the same masked operator. A radius other than 1 in laplacian_kwargs
    uses the reduced system, with ichol when no factory is given.
    Returns alpha in the type of `image` (numpy or torch), shape trimap.shape.
    """
    sanity_check_image(image)
    # util.py:490-520: the trimap's warnings here, its ValueErrors (no
    # foreground / no background) from device_classes.
    check_trimap(trimap, stacklevel=3)
    if "is_known" in laplacian_kwargs:
        raise TypeError("cf_laplacian() got multiple values for keyword argument 'is_known'")
    classes, tdev = device_classes(trimap)
    shape = tuple(trimap.shape)
    radius = int(laplacian_kwargs.get("radius", 1))
    full_grid = radius == 1 and (preconditioner is None or preconditioner is vcycle or
                                 preconditioner is jacobi)
    if full_grid:
        op = CfOperator(image, mode="masked", classes=classes, trimap=tdev, **laplacian_kwargs)
        b = op.rhs()
        if preconditioner is jacobi:
            A, M = op, jacobi(op)
        else:
            A = DiaOperator.from_cf(op)
            M = VCycle(A, op.grid_shape)
        x = cg(A, b, M=M, **cg_kwargs)
        alpha = _finalize(x, tdev, classes, shape)
    else:
        if preconditioner is None:
            preconditioner = ichol
        L_U, b_U = reduced_system(image, classes, laplacian_kwargs)
        if L_U.shape[0] == 0:
            x_u = empty(1)
        else:
            x_u = cg(L_U, b_U, M=preconditioner(L_U), **cg_kwargs)
            x_u = as_device(x_u).reshape(-1)
        alpha = empty(classes.numel())
        _ffi.check(_ffi.lib().hm_reduced_finalize(classes.numel(), _ffi.ptr(classes),
                                                  _ffi.ptr(x_u), _ffi.ptr(tdev), _ffi.ptr(alpha),
                                                  _ffi.stream()))
        alpha = alpha.reshape(shape)
    return alpha if is_torch(image) else alpha.cpu().numpy()


def reduced_system(image, classes, laplacian_kwargs={}):
    """(L_U, b_U) of estimate_alpha_cf.py:57-70, built on the device.

    L = cf_laplacian(image, is_known=known) in the reference's CSR layout,
    then L_U = L[U, :][:, U] as scipy's fancy indexing leaves it and
    b_U = -R m with R = L[U, :][:, K], m = is_fg[K]. L_U is returned as a
    scipy CSR matrix carrying its device copy, b_U as a device vector.
    """
    from pymatting_amd.laplacian import _attach
    from pymatting_amd.sparse import DeviceCSR
    n = classes.numel()
    if int(laplacian_kwargs.get("radius", 1)) == 1:
        # The reference layout assembled straight into device memory.
        op = CfOperator(image, mode="laplacian", classes=classes, **laplacian_kwargs)
        dev = DeviceCSR.from_device(n, *op.reference_csr_device())
    else:
        known = (classes != 0).cpu().numpy()
        dev = attached_csr(cf_laplacian(image, is_known=known, **laplacian_kwargs))
    lib = _ffi.lib()
    ptr_u = empty(n + 1, torch.int64)
    n_u, nnz_u = ctypes.c_int64(), ctypes.c_int64()
    _ffi.check(lib.hm_csr_reduced_indptr(n, _ffi.ptr(dev._indptr), _ffi.ptr(dev._indices),
                                         _ffi.ptr(classes), _ffi.ptr(ptr_u), ctypes.byref(n_u),
                                         ctypes.byref(nnz_u), _ffi.stream()))
    m, k = n_u.value, nnz_u.value
    idx_u = empty(max(k, 1), torch.int64)
    val_u = empty(max(k, 1))
    b_u = empty(max(m, 1))
    if m > 0:
        _ffi.check(lib.hm_csr_reduced_fill(n, _ffi.ptr(dev._indptr), _ffi.ptr(dev._indices),
                                           _ffi.ptr(dev._values), _ffi.ptr(classes),
                                           _ffi.ptr(ptr_u), _ffi.ptr(idx_u), _ffi.ptr(val_u),
                                           _ffi.ptr(b_u), _ffi.stream()))
    ptr_u = ptr_u[:m + 1]
    dev_u = DeviceCSR.from_device(m, ptr_u, idx_u[:k], val_u[:k])
    host = scipy.sparse.csr_matrix((val_u[:k].cpu().numpy(), idx_u[:k].cpu().numpy(),
                                    ptr_u.cpu().numpy()), shape=(m, m))
    return _attach(host, dev_u), b_u[:m]


estimate_alpha = estimate_alpha_cf


class LambdaSolver:
    """Device-resident lambda-penalised alpha solve for one image.

    Set-up (operator, right-hand side, V-cycle hierarchy) happens in the
    constructor; ``solve()`` runs PCG and returns alpha on the device.
    precision "f64": float64 PCG with a float64 V-cycle (the plain method).
    precision "mixed": float32 stencil, SpMV and V-cycle smoothing inside PCG
    with float64 dot products, wrapped in float64 iterative refinement.
    """

    def __init__(self, image, trimap, epsilon=1e-7, lambda_value=100.0, precision="f64",
                 num_pre_iter=1, num_post_iter=1, omega=0.8, direct_solve_size=64,
                 level0="matrix_free", max_levels=0, coarse_sweeps=0):
        self.h, self.w = int(trimap.shape[0]), int(trimap.shape[1])
        self.precision = precision
        # The trimap's errors are raised once the operator's set-up is
        # enqueued, which then overlaps the read-back of its class counts.
        self.classes, self.tdev, check_trimap_counts = device_classes(trimap, deferred=True)
        self.cf = CfOperator(image, epsilon, 1, "lambda", self.classes, self.tdev, lambda_value)
        self.b = self.cf.rhs()
        # level0 "matrix_free": the stencil formed from the image and the
        # stored window means and inverse covariances on every pass;
        # "matrix_free_recompute": float32 passes also recompute the windows
        # from the image tile (float64 arithmetic); "stencil": the 25-point
        # operator assembled once (exactly) and streamed (the stored ELL/DIA
        # form).
        check_trimap_counts()
        if level0 not in ("matrix_free", "matrix_free_recompute", "stencil"):
            raise ValueError("level0 must be 'matrix_free', 'matrix_free_recompute' or "
                             "'stencil'")
        if level0 == "matrix_free_recompute":
            self.cf.set_windows(True)
        self.op = DiaOperator.from_cf(self.cf) if level0 == "stencil" else self.cf
        self.M = VCycle(self.op, (self.h, self.w), num_pre_iter, num_post_iter, omega,
                        direct_solve_size, "f64" if precision == "f64" else "f32",
                        max_levels, coarse_sweeps)
        self.stats = {}

    def solve_system(self, b, rtol=1e-7, atol=0.0, maxiter=10000, inner_rtol=1e-4,
                     max_outer=20, check_every=8, mixed="refine", delta=0.01, graph=False):
        """x = A^-1 b for a device float64 right-hand side b (length h*w), with
        this solver's operator, preconditioner and precision.

        For precision "mixed", ``mixed`` picks the float64 outer loop:
        "refine" restarts a float32 PCG on each float64 residual (inner_rtol,
        max_outer); "reliable" runs one float32 PCG whose solution accumulates
        in float64 and whose residual is replaced by the float64 one each time
        it has dropped by ``delta`` (``graph``: its iterations without a
        replacement are replayed from HIP graphs, with identical results)."""
        import ctypes
        n = self.h * self.w
        b = b.reshape(-1)
        if b.numel() != n or b.dtype != torch.float64 or not b.is_cuda:
            raise ValueError("b must be a device float64 vector of length %d" % n)
        b = b.contiguous()
        if self.precision == "f64":
            x, st = cg(self.op, b, atol=atol, rtol=rtol, maxiter=maxiter, M=self.M,
                       check_every=check_every, return_stats=True)
            self.stats = st
        else:
            x = empty(n)
            stats = _ffi.CgStats()
            cf = self.op.handle if isinstance(self.op, CfOperator) else None
            dia = self.op.handle if isinstance(self.op, DiaOperator) else None
            if mixed == "reliable":
                rc = _ffi.lib().hm_cg_solve_reliable(
                    cf, dia, self.M.handle, _ffi.ptr(b), _ffi.ptr(x), float(rtol),
                    float(delta), int(maxiter), int(check_every), int(bool(graph)),
                    ctypes.byref(stats), _ffi.stream())
            elif mixed == "refine":
                rc = _ffi.lib().hm_cg_solve_mixed(
                    cf, dia, self.M.handle, _ffi.ptr(b), _ffi.ptr(x), float(rtol),
                    float(inner_rtol), int(max_outer), int(maxiter), ctypes.byref(stats),
                    _ffi.stream())
            else:
                raise ValueError("mixed must be 'refine' or 'reliable'")
            _ffi.check(rc)
            self.stats = {"iterations": stats.iterations, "norm_b": stats.norm_b,
                          "norm_r": stats.norm_r}
        return x

    def solve(self, rtol=1e-7, atol=0.0, maxiter=10000, inner_rtol=1e-4, max_outer=20,
              check_every=8, mixed="refine", delta=0.01, graph=False):
        self.x = self.solve_system(self.b, rtol, atol, maxiter, inner_rtol, max_outer,
                                   check_every, mixed, delta, graph)
        return _finalize(self.x, self.tdev, None, (self.h, self.w))


def solve_alpha_lambda(image, trimap, epsilon=1e-7, lambda_value=100.0, rtol=1e-7,
                       precision="f64", **kwargs):
    """cf_laplacian -> make_linear_system -> cg(M=vcycle) -> clip, on the device."""
    solver = LambdaSolver(image, trimap, epsilon, lambda_value, precision)
    alpha = solver.solve(rtol=rtol, **kwargs)
    return alpha, solver.stats
