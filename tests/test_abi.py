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
