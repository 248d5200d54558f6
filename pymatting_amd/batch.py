"""Batched lambda-system alpha solves over many images (sizes may differ).

pymatting solves one image per call (estimate_alpha_cf.py, cg.py); a caller
with many images loops. ``BatchSolver`` sets up each image like
``LambdaSolver`` (stored 13-slot stencil, Galerkin V-cycle) and then runs all
PCG solves together in the library (hm_batch_solve): one launch per step over
the tiles of every image that still iterates, per-image CG scalars and
convergence masks, and iterations replayed from a HIP graph.

``autotune()`` picks the tile shapes (finest grid / Galerkin grids: 64x8 or
32x8 points per workgroup) and the iterations per graph by timing the batch
itself; results are cached per batch signature for the process.
"""
import ctypes
import time

import torch

from pymatting_amd import _ffi
from pymatting_amd._device import empty
from pymatting_amd.alpha import LambdaSolver, _finalize

_TUNED = {}  # (signature, mixed) -> (level0_shape, coarse_shape, graph_iters)

SHAPES = (0, 1)
GRAPH_ITERS = (0, 2, 4, 8, 16)


class BatchSolver:
    """Alpha of every (image, trimap) pair by batched V-cycle PCG.

    precision "f64": float64 PCG; "mixed": float32 PCG (float64 dot
    products) inside float64 iterative refinement.
    """

    def __init__(self, images, trimaps, epsilon=1e-7, lambda_value=100.0, precision="f64",
                 num_pre_iter=1, num_post_iter=1, omega=0.8, direct_solve_size=64,
                 level0_shape=0, coarse_shape=1, graph_iters=8):
        if len(images) != len(trimaps) or len(images) == 0:
            raise ValueError("need the same, non-zero number of images and trimaps")
        self.precision = precision
        self.solvers = [LambdaSolver(im, tr, epsilon, lambda_value, precision, num_pre_iter,
                                     num_post_iter, omega, direct_solve_size, level0="stencil")
                        for im, tr in zip(images, trimaps)]
        self.count = len(self.solvers)
        self.level0_shape = level0_shape
        self.coarse_shape = coarse_shape
        self.graph_iters = graph_iters
        self.stats = []
        self._handle = None
        self._handle_key = None

    def signature(self):
        return (self.precision, tuple((s.h, s.w) for s in self.solvers),
                self.solvers[0].M.handle is not None)

    def _batch(self):
        key = (self.level0_shape, self.coarse_shape)
        if self._handle is not None and self._handle_key == key:
            return self._handle
        self._destroy()
        ops = (ctypes.c_void_p * self.count)(*[s.op.handle.value for s in self.solvers])
        vcs = (ctypes.c_void_p * self.count)(*[s.M.handle.value for s in self.solvers])
        h = ctypes.c_void_p()
        prec = _ffi.HM_F64 if self.precision == "f64" else _ffi.HM_F32
        _ffi.check(_ffi.lib().hm_batch_create(self.count, ops, vcs, prec, int(self.level0_shape),
                                              int(self.coarse_shape), ctypes.byref(h)))
        self._handle, self._handle_key = h, key
        return h

    def _destroy(self):
        h = getattr(self, "_handle", None)
        self._handle = None
        if h is None:
            return
        try:  # at interpreter shutdown the handle is left to the process exit
            if h.value and _ffi._lib is not None:
                _ffi._lib.hm_batch_destroy(h)
        except Exception:
            pass

    def __del__(self):
        self._destroy()

    def solve_x(self, rtol=1e-7, inner_rtol=1e-3, max_outer=20, maxiter=10000, rhs=None,
                mixed="refine", delta=0.01):
        """The solutions x_i = A_i^-1 b_i (device float64 vectors), without the
        clip. b_i is each image's lambda right-hand side, or rhs[i] (device
        float64 vectors of h_i * w_i values, e.g. adjoint right-hand sides).
        For precision "mixed", ``mixed`` picks restarted refinement
        (inner_rtol, max_outer) or "reliable" updates (delta), as in
        LambdaSolver.solve_system."""
        h = self._batch()
        xs = [empty(s.h * s.w) for s in self.solvers]
        if rhs is None:
            rhs = [s.b for s in self.solvers]
        else:
            if len(rhs) != self.count:
                raise ValueError("need one right-hand side per image")
            rhs = [r.reshape(-1).to(torch.float64).contiguous() for r in rhs]
            for r, s in zip(rhs, self.solvers):
                if r.numel() != s.h * s.w or not r.is_cuda:
                    raise ValueError("right-hand sides must be device vectors of h*w values")
        bp = (ctypes.c_void_p * self.count)(*[r.data_ptr() for r in rhs])
        xp = (ctypes.c_void_p * self.count)(*[x.data_ptr() for x in xs])
        stats = (_ffi.CgStats * self.count)()
        if self.precision != "f64" and mixed == "reliable":
            rc = _ffi.lib().hm_batch_solve_reliable(h, bp, xp, float(rtol), float(delta),
                                                    int(maxiter), int(self.graph_iters > 0),
                                                    stats, _ffi.stream())
        elif self.precision == "f64" or mixed == "refine":
            rc = _ffi.lib().hm_batch_solve(h, bp, xp, float(rtol), float(inner_rtol),
                                           int(max_outer), int(maxiter), int(self.graph_iters),
                                           stats, _ffi.stream())
        else:
            raise ValueError("mixed must be 'refine' or 'reliable'")
        self.stats = [{"iterations": st.iterations, "converged": bool(st.converged),
                       "norm_b": st.norm_b, "norm_r": st.norm_r} for st in stats]
        _ffi.check(rc)
        return xs

    def solve(self, rtol=1e-7, inner_rtol=1e-3, max_outer=20, maxiter=10000, mixed="refine",
              delta=0.01):
        """Alpha of every image (device float64 arrays of shape (h, w))."""
        xs = self.solve_x(rtol, inner_rtol, max_outer, maxiter, mixed=mixed, delta=delta)
        return [_finalize(x, s.tdev, None, (s.h, s.w)) for x, s in zip(xs, self.solvers)]

    def autotune(self, rtol=1e-7, inner_rtol=1e-3, maxiter=10000, repeats=1, mixed="refine",
                 delta=0.01):
        """Chooses tile shapes, then iterations per graph, by timing full solves
        of this batch with the given solve options; cached per batch signature
        and form. For reliable updates the second choice is graph replay of
        each iteration's fixed parts (graph_iters 8) or direct launches (0).
        Returns the choice."""
        reliable = self.precision != "f64" and mixed == "reliable"
        key = (self.signature(), "reliable" if reliable else "refine")
        if key in _TUNED:
            self.level0_shape, self.coarse_shape, self.graph_iters = _TUNED[key]
            return _TUNED[key]

        def run():
            self.solve_x(rtol, inner_rtol, 20, maxiter, mixed=mixed, delta=delta)

        def timed():
            best = float("inf")
            for _ in range(repeats):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                run()
                torch.cuda.synchronize()
                best = min(best, time.perf_counter() - t0)
            return best

        self.graph_iters = 0 if reliable else 8
        run()  # warm the allocator and code
        results = {}
        for s0 in SHAPES:
            for sc in SHAPES:
                self.level0_shape, self.coarse_shape = s0, sc
                results[(s0, sc)] = timed()
        self.level0_shape, self.coarse_shape = min(results, key=results.get)
        gtimes = {}
        for g in ((0, 8) if reliable else GRAPH_ITERS):
            self.graph_iters = g
            gtimes[g] = timed()
        self.graph_iters = min(gtimes, key=gtimes.get)
        self.tuning = {"shapes_s": {"%d,%d" % k: v for k, v in results.items()},
                       "graph_iters_s": gtimes}
        _TUNED[key] = (self.level0_shape, self.coarse_shape, self.graph_iters)
        return _TUNED[key]
