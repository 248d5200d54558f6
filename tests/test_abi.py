"""CPU checks of the C ABI: the built library loads and exports every
function declared in include/pymatting_amd.h (no device calls)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "pymatting_amd.h")
LIB = os.path.join(ROOT, "pymatting_amd", "libpymatting_amd.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+(hm_[a-z0-9_]+)\s*\(",
                           text, flags=re.M))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared()
    for must in ["hm_cf_create", "hm_cf_assemble_csr", "hm_vcycle_create", "hm_cg_solve",
                 "hm_cg_solve_mixed", "hm_fb_ml", "hm_last_error"]:
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    import torch  # noqa: F401  (loads the HIP runtime the library links)
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def _lib():
    import torch  # noqa: F401  (loads the HIP runtime the library links)
    lib = ctypes.CDLL(LIB)
    lib.hm_last_error.restype = ctypes.c_char_p
    return lib


def _err(lib):
    return lib.hm_last_error().decode()


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_argument_checks_before_any_device_call():
    """Entry points reject bad arguments with HM_ERR_ARG (-1) and a message
    before they touch a device, so these run without a GPU."""
    lib = _lib()
    P, i32, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    ARG = -1
    assert lib.hm_version() == 1
    # Reliable-update and refinement solves need one operator and a V-cycle.
    lib.hm_cg_solve_reliable.argtypes = [P, P, P, P, P, f64, f64, i32, i32, i32, P, P]
    assert lib.hm_cg_solve_reliable(None, None, None, None, None, 1e-9, 0.01, 10, 8, 0, None,
                                    None) == ARG
    assert "0 < delta < 1" in _err(lib)
    lib.hm_cg_solve_mixed.argtypes = [P, P, P, P, P, f64, f64, i32, i32, P, P]
    assert lib.hm_cg_solve_mixed(None, None, None, None, None, 1e-8, 1e-3, 20, 10, None,
                                 None) == ARG
    assert "float32 V-cycle" in _err(lib)
    lib.hm_vcycle_create.argtypes = [P, P, i32, i32, i32, i32, f64, i32, i32, P, P]
    assert lib.hm_vcycle_create(None, None, 4, 4, 1, 1, 0.8, 64, 0, None, None) == ARG
    assert "null output" in _err(lib)
    # Batches: count, arrays and tile shapes.
    lib.hm_batch_create.argtypes = [i32, P, P, i32, i32, i32, P]
    out = P()
    assert lib.hm_batch_create(0, None, None, 0, 0, 0, ctypes.byref(out)) == ARG
    ops = (P * 1)()
    vcs = (P * 1)()
    assert lib.hm_batch_create(1, ops, vcs, 0, 2, 0, ctypes.byref(out)) == ARG
    assert "tile shapes" in _err(lib)
    lib.hm_batch_solve_reliable.argtypes = [P, P, P, f64, f64, i32, i32, P, P]
    assert lib.hm_batch_solve_reliable(None, None, None, 1e-9, 0.01, 10, 0, None, None) == ARG
    # Foreground: the pyramid must run from 1x1 to the image size.
    lib.hm_fb_ml.argtypes = [P, P, i32, i32, i32, P, P, f64, i32, i32, i32, f64, i32, P, P, P]
    buf = ctypes.create_string_buffer(64)
    lh = (i32 * 3)(1, 2, 4)
    lw = (i32 * 3)(1, 3, 5)
    assert lib.hm_fb_ml(None, buf, 4, 5, 2, lh, lw, 1e-5, 10, 2, 32, 1.0, 1, buf, buf,
                        None) == ARG
    bad_lh = (i32 * 3)(2, 2, 4)
    assert lib.hm_fb_ml(buf, buf, 4, 5, 2, bad_lh, lw, 1e-5, 10, 2, 32, 1.0, 1, buf, buf,
                        None) == ARG
    assert "from 1x1" in _err(lib)
    # The allocation-free foreground rejects images beyond its static
    # workspace before any device work (1080p float32 needs ~165 MB).
    lib.hm_fb_ml_noalloc.argtypes = lib.hm_fb_ml.argtypes
    n_levels = 11
    lh = (i32 * (n_levels + 1))(*[round(1080 ** (i / n_levels)) for i in range(n_levels + 1)])
    lw = (i32 * (n_levels + 1))(*[round(1920 ** (i / n_levels)) for i in range(n_levels + 1)])
    assert lib.hm_fb_ml_noalloc(buf, buf, 1080, 1920, n_levels, lh, lw, 1e-5, 10, 2, 32, 1.0, 1,
                                buf, buf, None) == ARG
    assert "static workspace" in _err(lib)
    # Vector helpers: negative sizes are rejected, empty ones do nothing.
    lib.hm_diag_apply.argtypes = [P, P, P, ctypes.c_int64, P]
    assert lib.hm_diag_apply(None, None, None, 4, None) == ARG
    assert lib.hm_diag_apply(buf, buf, buf, 0, None) == 0
    lib.hm_alpha_finalize.argtypes = [P, P, P, P, ctypes.c_int64, P]
    assert lib.hm_alpha_finalize(buf, None, None, buf, -1, None) == ARG
    lib.hm_prof_query.argtypes = [i32, i32, P, P, P]
    assert lib.hm_prof_query(0, 0, None, None, None) == ARG
    assert "null argument" in _err(lib)
