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


def test_scene_object_offsets_match_the_c_layout():
    """Offsets of ap_layout.h (static_asserted there) seen through the dtype."""
    f = ap.SCENE_OBJECT_DTYPE.fields
    want = {"center": 0, "extent": 12, "shading_model": 32, "albedo": 64, "emissive": 76, "specular": 88,
            "metallic": 92, "roughness": 96, "ior": 100, "sigma_a": 104, "normal": 108, "abbe_number": 120,
            "tilt_number": 124, "textures": 128, "world_to_object": 160, "object_to_world": 196, "type": 232,
            "sdf_type": 240, "sdf_info0": 256, "cached_volume_index": 272, "v0_position": 288, "v0_normal": 300,
            "v0_uv": 312, "v1_position": 320, "v2_position": 352, "proxy_min": 384, "proxy_max": 396,
            "proxy_center": 408, "object_index": 420, "light_index": 424, "visible": 428}
    for k, off in want.items():
        assert f[k][1] == off, k
    # a scene written by the C generator reads back through the dtype
    box = ap.synth_box_scene(1, 0, 15.0)
    light = box[box["shading_model"] == ap.MAT_LIGHT]
    assert len(light) == 1 and np.allclose(light["emissive"], 15.0) and np.allclose(light["albedo"], 0.0)
    walls = box[box["type"] == ap.OBJ_TRIANGLE]
    assert ((walls["albedo"] >= 0.1) & (walls["albedo"] <= 0.9)).all()
    assert np.allclose(walls["ior"], 1.5) and np.allclose(walls["abbe_number"], 20.0)


def test_errors_are_codes_not_aborts():
    lib = ap.lib()
    assert lib.ap_scene_create(None, 0, None, None) == -1
    assert b"out_scene" in lib.ap_last_error_string()
    assert lib.ap_trace_closest(None, None, 1, None, None) == -1
    bad = np.zeros(1, dtype=ap.SCENE_OBJECT_DTYPE)
    bad["type"] = 7
    h = ctypes.c_void_p()
    assert lib.ap_scene_create(bad.ctypes.data, 1, None, ctypes.byref(h)) == -1
    assert b"unknown object type" in lib.ap_last_error_string()
