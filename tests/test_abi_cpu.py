"""The C-ABI library loads and exports every function include/aperture/*.h
declares (no device calls), and the layouts match the reference sizes."""
import ctypes
import os
import re

import numpy as np

import aperture as ap

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(REPO, "include", "aperture")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            names |= set(re.findall(r"^\s*(?:int|const char\*|void)\s+(ap_[a-z0-9_]+)\s*\(", text, re.M))
    return names


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(ap.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert names <= set(ap.EXPORTED_SYMBOLS), names - set(ap.EXPORTED_SYMBOLS)


def test_layout_sizes():
    assert ap.SCENE_OBJECT_DTYPE.itemsize == 432
    assert ap.BVH_NODE_DTYPE.itemsize == 48
    assert ap.RAY_DTYPE.itemsize == 32 and ap.HIT_DTYPE.itemsize == 8
    assert ap.LIGHT_INFO_DTYPE.itemsize == 12
    assert ap.SCENE_OBJECT_DTYPE.fields["visible"][1] == 428
    assert ap.SCENE_OBJECT_DTYPE.fields["v0_position"][1] == 288


def test_errors_are_codes_not_aborts():
    lib = ap.lib()
    assert lib.ap_scene_create(None, 0, None, None) == -1
    assert b"out_scene" in lib.ap_last_error_string()
    assert lib.ap_trace_closest(None, None, 1, None, None) == -1
    sdf = np.zeros(1, dtype=ap.SCENE_OBJECT_DTYPE)
    sdf["type"] = ap.OBJ_SDF
    h = ctypes.c_void_p()
    assert lib.ap_scene_create(sdf.ctypes.data, 1, None, ctypes.byref(h)) == -4
    assert b"SDF" in lib.ap_last_error_string()
