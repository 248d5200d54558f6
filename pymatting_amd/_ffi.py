"""ctypes bindings of libpymatting_amd.so (the C ABI in include/pymatting_amd.h).

PyTorch is imported first: the library links the HIP runtime that ships with
torch, so one runtime serves both and torch's device pointers and streams are
valid in the library. There is no CPU fallback: if the library or a HIP
device is missing, every call raises.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpymatting_amd.so")

HM_OK = 0
HM_ERR_ARG = -1
HM_ERR_HIP = -2
HM_ERR_NOCONV = -3
HM_ERR_UNSUPPORTED = -4
HM_ERR_VALUE = -5
HM_ERR_CALLBACK = -6

HM_F64 = 0
HM_F32 = 1

HM_OP_LAPLACIAN = 0
HM_OP_LAMBDA = 1
HM_OP_MASKED = 2

HM_PC_NONE = 0
HM_PC_JACOBI = 1
HM_PC_VCYCLE = 2
HM_PC_CSR = 3
HM_PC_HOST = 4
HM_PC_ICHOL = 5

P = ctypes.c_void_p
i32 = ctypes.c_int
i64 = ctypes.c_int64
f64 = ctypes.c_double

APPLY_CB = ctypes.CFUNCTYPE(ctypes.c_int, P, P, P)
ITER_CB = ctypes.CFUNCTYPE(ctypes.c_int, P, P, P, f64, f64)


class CgStats(ctypes.Structure):
    _fields_ = [("iterations", i32), ("converged", i32), ("norm_b", f64), ("norm_r", f64)]


class CgArgs(ctypes.Structure):
    _fields_ = [
        ("cf", P), ("dia", P), ("csr", P), ("a_cb", APPLY_CB), ("a_ctx", P), ("n", i64),
        ("pc_kind", i32), ("inv_diag", P), ("vc", P), ("pc_csr", P), ("m_cb", APPLY_CB),
        ("m_ctx", P), ("b", P), ("x0", P), ("x", P), ("atol", f64), ("rtol", f64),
        ("maxiter", i32), ("reorthogonalize", i32), ("check_every", i32),
        ("callback", ITER_CB), ("cb_ctx", P), ("ichol", P), ("lkm", P),
    ]


class ConvergenceError(ValueError):
    """CG did not converge (the reference raises ValueError, cg.py:115-117)."""


_SIGS = {
    "hm_last_error": (ctypes.c_char_p, []),
    "hm_version": (i32, []),
    "hm_copy": (i32, [P, P, i64, i32, P]),
    "hm_trimap_classes": (i32, [P, i64, f64, f64, P, P, P]),
    "hm_cf_create": (i32, [P, i32, i32, f64, i32, P, f64, i32, P, ctypes.POINTER(P)]),
    "hm_cf_destroy": (i32, [P]),
    "hm_cf_assemble_csr": (i32, [P, P, P, P, P]),
    "hm_cf_system_indptr": (i32, [P, P, ctypes.POINTER(i64), P]),
    "hm_cf_system_fill": (i32, [P, P, P, P, P]),
    "hm_cf_apply": (i32, [P, i32, P, P, P]),
    "hm_cf_set_windows": (i32, [P, i32]),
    "hm_cf_diagonal": (i32, [P, P, P]),
    "hm_cf_rhs": (i32, [P, P, P]),
    "hm_cf_adjoint_image": (i32, [P, P, P, P, i32, P]),
    "hm_dia_create": (i32, [P, i32, i32, P, ctypes.POINTER(P)]),
    "hm_dia_from_cf": (i32, [P, P, ctypes.POINTER(P)]),
    "hm_dia_destroy": (i32, [P]),
    "hm_dia_apply": (i32, [P, i32, P, P, P]),
    "hm_dia_coefficients": (i32, [P, P, P]),
    "hm_csr_create": (i32, [i64, i64, P, P, P, ctypes.POINTER(P)]),
    "hm_csr_destroy": (i32, [P]),
    "hm_csr_apply": (i32, [P, P, P, P]),
    "hm_vcycle_create": (i32, [P, P, i32, i32, i32, i32, f64, i32, i32, P, ctypes.POINTER(P)]),
    "hm_vcycle_create_ex": (i32, [P, P, i32, i32, i32, i32, f64, i32, i32, i32, i32, P,
                                  ctypes.POINTER(P)]),
    "hm_vcycle_create_dense": (i32, [P, i32, i32, P, ctypes.POINTER(P)]),
    "hm_vcycle_destroy": (i32, [P]),
    "hm_vcycle_levels": (i32, [P]),
    "hm_vcycle_set_launches": (i32, [P, i32]),
    "hm_vcycle_apply": (i32, [P, P, P, P]),
    "hm_cg_solve": (i32, [ctypes.POINTER(CgArgs), ctypes.POINTER(CgStats), P]),
    "hm_cg_solve_mixed": (i32, [P, P, P, P, P, f64, f64, i32, i32, ctypes.POINTER(CgStats), P]),
    "hm_cg_solve_reliable": (i32, [P, P, P, P, P, f64, f64, i32, i32, i32,
                                   ctypes.POINTER(CgStats),
                                   P]),
    "hm_diag_apply": (i32, [P, P, P, i64, P]),
    "hm_alpha_finalize": (i32, [P, P, P, P, i64, P]),
    "hm_prof_set": (i32, [i32]),
    "hm_vcycle_time_fine": (i32, [P, P, i32, ctypes.POINTER(f64), ctypes.POINTER(f64), P]),
    "hm_prof_query": (i32, [i32, i32, ctypes.POINTER(f64), ctypes.POINTER(i64), ctypes.POINTER(i64)]),
    "hm_fb_ml": (i32, [P, P, i32, i32, i32, P, P, f64, i32, i32, i32, f64, i32, P, P, P]),
    "hm_fb_ml_noalloc": (i32, [P, P, i32, i32, i32, P, P, f64, i32, i32, i32, f64, i32, P, P,
                               P]),
    "hm_batch_create": (i32, [i32, P, P, i32, i32, i32, ctypes.POINTER(P)]),
    "hm_batch_solve": (i32, [P, P, P, f64, f64, i32, i32, i32, P, P]),
    "hm_batch_solve_reliable": (i32, [P, P, P, f64, f64, i32, i32, P, P]),
    "hm_batch_destroy": (i32, [P]),
    "hm_time_sync": (i32, [i32, i32, P, P, P, P, P]),
    "hm_ichol_factor": (i32, [i64, P, P, P, f64, f64, i64, f64, i32, ctypes.POINTER(i64),
                              ctypes.POINTER(P)]),
    "hm_ichol_create": (i32, [i64, P, P, P, ctypes.POINTER(P)]),
    "hm_ichol_nnz": (i64, [P]),
    "hm_ichol_copy_factor": (i32, [P, P, P, P]),
    "hm_ichol_apply": (i32, [P, P, P, P]),
    "hm_ichol_destroy": (i32, [P]),
    "hm_cf_laplacian_csr": (i32, [P, i32, i32, f64, i32, P, P, P, P, P]),
    "hm_csr_reduced_indptr": (i32, [i64, P, P, P, P, ctypes.POINTER(i64), ctypes.POINTER(i64),
                                    P]),
    "hm_csr_reduced_fill": (i32, [i64, P, P, P, P, P, P, P, P, P]),
    "hm_reduced_finalize": (i32, [i64, P, P, P, P, P]),
    "hm_csr_add_diagonal_indptr": (i32, [i64, P, P, P, P, P, ctypes.POINTER(i64), P]),
    "hm_csr_add_diagonal_fill": (i32, [i64, P, P, P, P, P, P, P, P]),
    "hm_fgcf_system_indptr": (i32, [P, i32, i32, f64, i32, P, P, P, ctypes.POINTER(i64), P]),
    "hm_fgcf_system_fill": (i32, [P, i32, i32, f64, i32, P, P, P, P, P, P]),
    "hm_fgcf_channel": (i32, [P, P, P, P, i64, i32, i32, P, P, P]),
    "hm_fgcf_store": (i32, [P, i64, i32, i32, P, P, P]),
    "hm_lkm_create": (i32, [P, i32, i32, f64, i32, P, f64, P, ctypes.POINTER(P)]),
    "hm_lkm_apply": (i32, [P, P, P, P]),
    "hm_lkm_diagonal": (i32, [P, i32, P, P]),
    "hm_lkm_rhs": (i32, [P, P, P]),
    "hm_lkm_destroy": (i32, [P]),
}

_lib = None


def lib():
    """The loaded library; raises if it was not built or no HIP device is visible."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("%s is missing: build it with `python -m pymatting_amd.build`"
                               % LIB_PATH)
        if not torch.cuda.is_available():
            raise RuntimeError("pymatting_amd needs a HIP device (MI355X); none is visible")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def load_for_symbols():
    """Loads the library without requiring a device (symbol checks only)."""
    return ctypes.CDLL(LIB_PATH)


def check(rc):
    if rc == HM_OK:
        return
    msg = lib().hm_last_error().decode(errors="replace")
    if rc == HM_ERR_NOCONV:
        raise ConvergenceError(msg)
    if rc in (HM_ERR_ARG, HM_ERR_VALUE):
        raise ValueError(msg)
    if rc == HM_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError("pymatting_amd: %s (status %d)" % (msg, rc))


def stream():
    """Handle of torch's current HIP stream; all library work is ordered on it."""
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def numel(a):
    """Element count of a numpy array, torch tensor or sequence, without copying it."""
    if hasattr(a, "shape"):
        n = 1
        for d in tuple(a.shape):
            n *= int(d)
        return n
    import numpy as np
    return int(np.asarray(a).size)


def release(obj, destroy):
    """Destroys obj.handle with the library function `destroy`, once. Safe at
    interpreter shutdown, when module globals and ctypes may already be gone:
    the handle is then left to the process exit."""
    h = getattr(obj, "handle", None)
    if h is None:
        return
    obj.handle = None
    try:
        if h.value and _lib is not None:
            getattr(_lib, destroy)(h)
    except Exception:
        pass
