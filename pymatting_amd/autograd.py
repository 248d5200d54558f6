"""Differentiable closed-form alpha solve (torch.autograd.Function).

pymatting itself is not differentiable. This module makes the lambda-penalised
closed-form solve (cf_laplacian.py:5-177, laplacian.py:5-60, cg.py with the
V-cycle of vcycle.py, the clip of estimate_alpha_cf.py:76) a torch operation
with a hand-written adjoint:

- forward: x = A(I)^-1 b with A = L(I) + lambda diag(known),
  b = lambda is_fg, by V-cycle PCG on the GPU (float64 or the mixed
  float32/float64 refinement); alpha = clip(x, 0, 1).
- backward: g = dl/dalpha on the pixels where the clip passes the gradient
  (0 <= x <= 1, as torch.clamp does); lam = A^-1 g by the same PCG (A is
  symmetric); dl/dI = -lam^T (dA/dI) x through the window means and inverse
  covariances of L (hm_cf_adjoint_image). The trimap is not differentiable.

A batch (B, h, w, 3) is solved as one batched PCG (BatchSolver) in the
forward pass and one batched adjoint PCG in the backward pass.
"""
import torch

from pymatting_amd import _ffi
from pymatting_amd.alpha import LambdaSolver


class SolveOptions:
    """Forward and adjoint solve settings."""

    def __init__(self, epsilon=1e-7, lambda_value=100.0, precision="mixed", rtol=1e-8,
                 inner_rtol=1e-3, adjoint_rtol=None, maxiter=20000, level0="stencil",
                 num_pre_iter=1, num_post_iter=1, direct_solve_size=64):
        self.epsilon = float(epsilon)
        self.lambda_value = float(lambda_value)
        self.precision = precision
        self.rtol = float(rtol)
        self.inner_rtol = float(inner_rtol)
        self.adjoint_rtol = float(rtol if adjoint_rtol is None else adjoint_rtol)
        self.maxiter = int(maxiter)
        self.level0 = level0
        self.num_pre_iter = int(num_pre_iter)
        self.num_post_iter = int(num_post_iter)
        self.direct_solve_size = int(direct_solve_size)


class AlphaSolve(torch.autograd.Function):
    """alpha = clip(A(image)^-1 b, 0, 1) for one (h, w, 3) image on the GPU."""

    @staticmethod
    def forward(ctx, image, trimap, opts):
        h, w = int(image.shape[0]), int(image.shape[1])
        solver = LambdaSolver(image.detach().to(torch.float64), trimap, opts.epsilon,
                              opts.lambda_value, opts.precision, opts.num_pre_iter,
                              opts.num_post_iter, direct_solve_size=opts.direct_solve_size,
                              level0=opts.level0)
        x = solver.solve_system(solver.b, rtol=opts.rtol, maxiter=opts.maxiter,
                                inner_rtol=opts.inner_rtol)
        ctx.solver, ctx.x, ctx.opts = solver, x, opts
        ctx.image_dtype = image.dtype
        ctx.forward_stats = dict(solver.stats)
        return x.clamp(0.0, 1.0).reshape(h, w).to(image.dtype)

    @staticmethod
    def backward(ctx, grad_alpha):
        solver, x, opts = ctx.solver, ctx.x, ctx.opts
        g = grad_alpha.reshape(-1).to(torch.float64) * ((x >= 0.0) & (x <= 1.0))
        grad = torch.zeros((solver.h, solver.w, 3), dtype=torch.float64, device=x.device)
        if bool(torch.any(g != 0)):
            lam = solver.solve_system(g.contiguous(), rtol=opts.adjoint_rtol,
                                      maxiter=opts.maxiter, inner_rtol=opts.inner_rtol)
            ctx.adjoint_stats = dict(solver.stats)
            _ffi.check(_ffi.lib().hm_cf_adjoint_image(solver.cf.handle, _ffi.ptr(lam),
                                                      _ffi.ptr(x), _ffi.ptr(grad), 0,
                                                      _ffi.stream()))
        return grad.to(ctx.image_dtype), None, None


class BatchAlphaSolve(torch.autograd.Function):
    """alpha_i = clip(A(image_i)^-1 b_i, 0, 1) for a (B, h, w, 3) batch: the
    forward solves and the adjoint solves each run as one batched PCG."""

    @staticmethod
    def forward(ctx, images, trimaps, opts):
        from pymatting_amd.batch import BatchSolver
        B, h, w = int(images.shape[0]), int(images.shape[1]), int(images.shape[2])
        bs = BatchSolver([images[i].detach().to(torch.float64) for i in range(B)],
                         [trimaps[i] for i in range(B)], opts.epsilon, opts.lambda_value,
                         opts.precision, opts.num_pre_iter, opts.num_post_iter,
                         direct_solve_size=opts.direct_solve_size)
        xs = bs.solve_x(rtol=opts.rtol, inner_rtol=opts.inner_rtol, maxiter=opts.maxiter)
        ctx.bs, ctx.xs, ctx.opts = bs, xs, opts
        ctx.image_dtype = images.dtype
        ctx.forward_stats = list(bs.stats)
        return torch.stack([x.clamp(0.0, 1.0).reshape(h, w) for x in xs]).to(images.dtype)

    @staticmethod
    def backward(ctx, grad_alpha):
        bs, xs, opts = ctx.bs, ctx.xs, ctx.opts
        gs = [grad_alpha[i].reshape(-1).to(torch.float64) * ((x >= 0.0) & (x <= 1.0))
              for i, x in enumerate(xs)]
        B = len(xs)
        h, w = bs.solvers[0].h, bs.solvers[0].w
        grad = torch.zeros((B, h, w, 3), dtype=torch.float64, device=xs[0].device)
        lams = bs.solve_x(rtol=opts.adjoint_rtol, inner_rtol=opts.inner_rtol,
                          maxiter=opts.maxiter, rhs=gs)
        ctx.adjoint_stats = list(bs.stats)
        for i, s in enumerate(bs.solvers):
            _ffi.check(_ffi.lib().hm_cf_adjoint_image(s.cf.handle, _ffi.ptr(lams[i]),
                                                      _ffi.ptr(xs[i]), _ffi.ptr(grad[i]), 0,
                                                      _ffi.stream()))
        return grad.to(ctx.image_dtype), None, None


def solve_alpha(image, trimap, batched=True, **options):
    """Differentiable alpha of one image (h, w, 3) or a batch (B, h, w, 3).

    image: torch tensor on the GPU (float32 or float64); gradients flow to it.
    trimap: (h, w) or (B, h, w) tensor or array (not differentiated).
    batched: a batch is solved as one batched PCG (forward and adjoint);
    False solves its images one at a time.
    options: see SolveOptions (precision "mixed" or "f64", rtol, ...).
    """
    opts = SolveOptions(**options)
    if image.dim() == 3:
        return AlphaSolve.apply(image, trimap, opts)
    if image.dim() != 4:
        raise ValueError("image must have shape (h, w, 3) or (B, h, w, 3)")
    if batched:
        return BatchAlphaSolve.apply(image, trimap, opts)
    return torch.stack([AlphaSolve.apply(image[i], trimap[i], opts)
                        for i in range(image.shape[0])])
