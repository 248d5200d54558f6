"""pymatting_amd: closed-form alpha matting and multilevel foreground estimation
on AMD Instinct MI355X (gfx950), with pymatting's Python API as a drop-in.

Drop-ins (same names and arguments as pymatting 1.1.15):
  cf_laplacian (any radius), make_linear_system, cg, CounterCallback,
  ProgressCallback, jacobi, vcycle, ichol (CholeskyDecomposition),
  estimate_alpha_cf (estimate_alpha), lkm_laplacian, estimate_alpha_lkm,
  estimate_foreground_ml, estimate_foreground_cf, cutout, load_image,
  save_image, stack_images, blend, trimap_split, sanity_check_image.
Device-resident API: CfOperator, DiaOperator, LambdaSolver, solve_alpha_lambda.
Differentiable solve (torch.autograd, hand-written adjoint): solve_alpha.
Batched solves of many images: BatchSolver.

All numerical work runs in hand-written HIP kernels of libpymatting_amd.so
(C ABI: include/pymatting_amd.h), except ichol's factorisation, which is
sequential and runs in the library on the host as the reference's does, and
the image file decoding of load_image. There is no CPU fallback.
"""
from pymatting_amd.alpha import (LambdaSolver, estimate_alpha, estimate_alpha_cf,
                                 solve_alpha_lambda)
from pymatting_amd.autograd import AlphaSolve, BatchAlphaSolve, solve_alpha
from pymatting_amd.batch import BatchSolver
from pymatting_amd.cutout import cutout
from pymatting_amd.foreground import estimate_foreground_cf, estimate_foreground_ml
from pymatting_amd.io import blend, load_image, save_image, stack_images
from pymatting_amd.laplacian import (CfLaplacianMatrix, CfOperator, DiaOperator, cf_laplacian,
                                     make_linear_system)
from pymatting_amd.lkm import LkmOperator, estimate_alpha_lkm, lkm_laplacian
from pymatting_amd.preconditioner import CholeskyDecomposition, ichol, jacobi, vcycle
from pymatting_amd.solver import CounterCallback, ProgressCallback, cg
from pymatting_amd.util import sanity_check_image, trimap_split

__all__ = [
    "cf_laplacian", "make_linear_system", "cg", "CounterCallback", "ProgressCallback", "jacobi",
    "vcycle", "ichol", "estimate_alpha_cf", "estimate_alpha", "estimate_foreground_ml",
    "trimap_split", "sanity_check_image", "CfOperator", "DiaOperator", "CfLaplacianMatrix",
    "LambdaSolver", "solve_alpha_lambda", "AlphaSolve", "BatchAlphaSolve", "solve_alpha",
    "BatchSolver", "CholeskyDecomposition", "estimate_foreground_cf", "lkm_laplacian",
    "estimate_alpha_lkm", "LkmOperator", "cutout", "load_image", "save_image", "stack_images",
    "blend",
]
