"""ctypes loader for the oracle's C restatements (test infrastructure)."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    """Load oracle/liboracle.so, building it with make if it is missing."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", _HERE])
        L = ctypes.CDLL(path)
        i64, f64, i32, f32 = ctypes.c_int64, ctypes.c_double, ctypes.c_int, ctypes.c_float
        P = ctypes.c_void_p
        L.oracle_ichol.restype = i64
        L.oracle_ichol.argtypes = [i64, P, P, P, P, P, P, f64, f64, i64, f64, i32]
        L.oracle_backsub_L.argtypes = [P, P, P, P, i64]
        L.oracle_backsub_LT.argtypes = [P, P, P, P, i64]
        L.oracle_fb_ml_f32.restype = i32
        L.oracle_fb_ml_f32.argtypes = [P, P, i32, i32, i32, P, P, f32, i32, i32, i32, f32, i32, P, P]
        L.oracle_fb_ml_f64.restype = i32
        L.oracle_fb_ml_f64.argtypes = [P, P, i32, i32, i32, P, P, f64, i32, i32, i32, f64, P, P]
        _LIB = L
    return _LIB


def ptr(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)
