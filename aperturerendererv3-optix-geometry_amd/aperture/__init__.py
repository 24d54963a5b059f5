"""Python host binding of the Aperture MI355X core (libaperture_hip.so).

The core is a C ABI (include/aperture/ap_core.h); this module loads it with
ctypes and mirrors the reference's host classes over it:

* ``SceneRenderer`` -- the reference ``SceneRenderer`` (core/scene.hpp:19-154):
  constructed from a SceneObject array, it owns the device scene.
* ``Camera`` -- the reference ``Camera`` (core/camera.hpp:14-99).
* ``trace_closest`` / ``trace_any`` -- the reference ``TraverseBVH<false/true>``
  (core/render/intersection.cuh:235-397) as a batched operator over device
  buffers.

Device memory, streams and timing events come from the library's own HIP
runtime (include/aperture/ap_device.h, wrapped by ``DeviceBuffer``, ``Stream``
and ``Event``), so a process never mixes two HIP runtimes. There is no CPU
fallback: loading fails loudly when the shared library is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaperture_hip.so")

AP_OK = 0
ERR_INVALID_ARG, ERR_HIP, ERR_OUT_OF_MEMORY, ERR_UNSUPPORTED, ERR_NO_DEVICE = -1, -2, -3, -4, -5

# ---------------------------------------------------------------------------
# Layouts (include/aperture/ap_layout.h)
# ---------------------------------------------------------------------------
_F3 = ("<f4", (3,))
SCENE_OBJECT_DTYPE = np.dtype(
    {
        "names": [
            "center", "extent", "shading_model", "albedo", "emissive", "specular", "metallic",
            "roughness", "ior", "sigma_a", "normal", "abbe_number", "tilt_number", "textures",
            "world_to_object", "object_to_world", "type", "sdf_type", "sdf_info0",
            "cached_volume_index", "v0_position", "v0_normal", "v0_uv", "v1_position", "v1_normal",
            "v1_uv", "v2_position", "v2_normal", "v2_uv", "proxy_min", "proxy_max", "proxy_center",
            "object_index", "light_index", "visible",
        ],
        "formats": [
            _F3, _F3, "<i4", _F3, _F3, "<f4", "<f4", "<f4", "<f4", "<f4", _F3, "<f4", "<f4",
            ("<u8", (4,)), ("<f4", (3, 3)), ("<f4", (3, 3)), "<i4", "<i4", ("<f4", (4,)), "<i4",
            _F3, _F3, ("<f4", (2,)), _F3, _F3, ("<f4", (2,)), _F3, _F3, ("<f4", (2,)),
            _F3, _F3, _F3, "<i4", "<i4", "u1",
        ],
        "offsets": [
            0, 12, 32, 64, 76, 88, 92, 96, 100, 104, 108, 120, 124, 128, 160, 196, 232, 240, 256,
            272, 288, 300, 312, 320, 332, 344, 352, 364, 376, 384, 396, 408, 420, 424, 428,
        ],
        "itemsize": 432,
    }
)
RAY_DTYPE = np.dtype([("o", "<f4", (3,)), ("tmin", "<f4"), ("d", "<f4", (3,)), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("id", "<i4"), ("t", "<f4")])
STATS_DTYPE = np.dtype([("nodes", "<u4"), ("prims", "<u4")])
BVH_NODE_DTYPE = np.dtype(
    {
        "names": ["bound_min", "bound_max", "left", "right", "first", "count"],
        "formats": [_F3, _F3, "<i4", "<i4", "<i4", "<i4"],
        "offsets": [0, 12, 24, 28, 32, 36],
        "itemsize": 48,
    }
)
LIGHT_INFO_DTYPE = np.dtype([("object_index", "<i4"), ("importance_weight", "<f4"), ("select_prob", "<f4")])

OBJ_CUBE, OBJ_SPHERE, OBJ_SDF, OBJ_TRIANGLE = 0, 1, 2, 3
MAT_DIFFUSE, MAT_LIGHT, MAT_GLASS, MAT_GGX = 0, 1, 2, 3
TRAVERSE_REFERENCE, TRAVERSE_CULLED = 0, 1


class BuildParams(ctypes.Structure):
    _fields_ = [("leaf_size", ctypes.c_int32), ("bin_count", ctypes.c_int32)]


class BVHInfo(ctypes.Structure):
    _fields_ = [
        ("node_count", ctypes.c_int64),
        ("interior_count", ctypes.c_int64),
        ("prim_count", ctypes.c_int64),
        ("max_depth", ctypes.c_int32),
        ("stack_capacity", ctypes.c_int32),
        ("device_bytes", ctypes.c_int64),
        ("light_count", ctypes.c_int32),
        ("traversal_mode", ctypes.c_int32),
        ("line_margin", ctypes.c_float),
        ("pad", ctypes.c_int32),
    ]


class Camera(ctypes.Structure):
    """ap_camera (include/aperture/ap_render.h)."""
    _fields_ = [("origin", ctypes.c_float * 3), ("direction", ctypes.c_float * 3), ("fov_deg", ctypes.c_float),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32)]

    @classmethod
    def make(cls, origin, direction, fov_deg, width, height):
        return cls((ctypes.c_float * 3)(*origin), (ctypes.c_float * 3)(*direction), fov_deg, width, height)


class RenderParams(ctypes.Structure):
    """ap_render_params; defaults are the reference's constants (config.cuh)."""
    _fields_ = [("spp", ctypes.c_int32), ("frame_index", ctypes.c_int32), ("max_depth", ctypes.c_int32),
                ("rr_min_depth", ctypes.c_int32), ("nee_max_depth", ctypes.c_int32), ("use_nee", ctypes.c_int32),
                ("nee_always", ctypes.c_int32), ("sky", ctypes.c_int32), ("triangle_material", ctypes.c_int32),
                ("use_guiding", ctypes.c_int32), ("seed", ctypes.c_uint64), ("wavelengths", ctypes.c_int32),
                ("guide_frames", ctypes.c_int32)]

    @classmethod
    def defaults(cls, **kw):
        p = cls()
        lib().ap_render_params_default(ctypes.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        return p


class NrcRenderParams(ctypes.Structure):
    """ap_nrc_render_params (include/aperture/ap_nrc.h)."""
    _fields_ = [("spread", ctypes.c_float), ("train_paths", ctypes.c_uint32), ("train_steps", ctypes.c_uint32),
                ("batch", ctypes.c_uint32), ("learning_rate", ctypes.c_float), ("unbiased_every", ctypes.c_uint32),
                ("train_passes", ctypes.c_uint32)]

    @classmethod
    def defaults(cls, **kw):
        p = cls()
        lib().ap_nrc_render_params_default(ctypes.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        return p


class NrcFrameStats(ctypes.Structure):
    _fields_ = [("queries", ctypes.c_uint32), ("train_paths", ctypes.c_uint32), ("samples", ctypes.c_uint32),
                ("pad", ctypes.c_uint32)]


class Framebuffers(ctypes.Structure):
    _fields_ = [("hdr", ctypes.c_void_p), ("odd", ctypes.c_void_p), ("even", ctypes.c_void_p),
                ("sum_xyz", ctypes.c_void_p), ("primary_id", ctypes.c_void_p), ("ray_count", ctypes.c_void_p),
                ("gbuffer", ctypes.c_void_p), ("statistics", ctypes.c_void_p)]


# denoise::ScreenGBuffer / ScreenStatisticsBuffer (include/aperture/ap_layout.h).
SCREEN_GBUFFER_DTYPE = np.dtype({"names": ["albedo", "normal", "metallic", "roughness", "depth", "valid_frame_index",
                                           "frame_index", "shading_model"],
                                 "formats": [("<f2", 3), ("<f2", 3), "<f2", "<f2", "<f4", "<u8", "<u8", ("<f2", 5)],
                                 "offsets": [0, 6, 12, 14, 16, 24, 32, 48], "itemsize": 64})
SCREEN_STATISTICS_DTYPE = np.dtype({"names": ["num", "mean", "m2_div_num", "m3_div_num"],
                                    "formats": ["<u4", ("<f4", 3), ("<f4", 3), ("<f4", 3)],
                                    "offsets": [0, 4, 16, 28], "itemsize": 48})
TONE_LINEAR, TONE_GAMMA, TONE_ACES = 0, 1, 2
(DISPLAY_IMAGE, DISPLAY_ALBEDO, DISPLAY_NORMAL, DISPLAY_DEPTH, DISPLAY_METALLIC, DISPLAY_ROUGHNESS,
 DISPLAY_STATISTICS_NUM, DISPLAY_STATISTICS_MEAN, DISPLAY_STATISTICS_M2, DISPLAY_STATISTICS_M3) = range(10)


TRIANGLE_MATERIAL_OBJECT, TRIANGLE_MATERIAL_REFERENCE = 0, 1


class ApertureError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"aperture error {code}: {message}")
        self.code = code


_lib = None

# Every symbol include/aperture/ap_core.h and ap_synth.h declare.
EXPORTED_SYMBOLS = (
    "ap_scene_create", "ap_build_bvh", "ap_scene_destroy", "ap_scene_get_info", "ap_scene_set_traversal_mode",
    "ap_scene_set_launch", "ap_scene_get_launch", "ap_scene_autotune",
    "ap_scene_export_bvh",
    "ap_trace_closest", "ap_trace_any", "ap_trace_closest_stats", "ap_last_error_string",
    "ap_synth_triangle_soup", "ap_synth_camera_rays", "ap_synth_secondary_rays",
    "ap_device_count", "ap_device_set", "ap_device_name", "ap_device_malloc", "ap_device_free", "ap_memcpy",
    "ap_device_copy",
    "ap_memset_async", "ap_stream_create", "ap_stream_destroy", "ap_stream_synchronize", "ap_device_synchronize",
    "ap_event_create", "ap_event_destroy", "ap_event_record", "ap_event_elapsed_ms",
    "ap_spectrum_setup", "ap_spectrum_encode_rgb", "ap_spectrum_encode_lambda", "ap_spectrum_query_rgb",
    "ap_spectrum_query_lambda", "ap_spectrum_interact", "ap_spectrum_sample_lambda",
    "ap_render_params_default", "ap_render", "ap_render_offline", "ap_xyz_to_srgb", "ap_synth_box_scene",
    "ap_synth_heightfield_scene", "ap_guiding_reset", "ap_guiding_export",
    "ap_nrc_create", "ap_nrc_destroy", "ap_nrc_set_bounds", "ap_nrc_get_params", "ap_nrc_set_params",
    "ap_nrc_infer", "ap_nrc_train_step", "ap_nrc_gradient", "ap_nrc_last_gradient", "ap_nrc_get_bounds",
    "ap_nrc_set_ema", "ap_nrc_render_params_default", "ap_nrc_create_grouped", "ap_nrc_info",
    "ap_nrc_set_group_bounds", "ap_nrc_infer_grouped", "ap_nrc_gradient_grouped", "ap_nrc_train_step_grouped",
    "ap_batch_create", "ap_batch_destroy", "ap_batch_render", "ap_batch_set_launch", "ap_batch_get_launch",
    "ap_batch_set_graph", "ap_batch_autotune", "ap_render_nrc", "ap_nrc_last_frame", "ap_batch_render_nrc",
    "ap_batch_nrc_last_frame", "ap_tonemap", "ap_sdf_cache_create", "ap_sdf_cache_destroy",
    "ap_sdf_cache_add_functor", "ap_sdf_cache_add_volume", "ap_sdf_cache_get_volume", "ap_scene_set_sdf_cache",
    "ap_texture_cache_create", "ap_texture_cache_destroy", "ap_texture_cache_add_rgba8", "ap_texture_cache_count",
    "ap_texture_sample", "ap_scene_set_texture_cache", "ap_mesh_load_obj",
    "ap_mse_f64", "ap_mse_luminance_f64", "ap_relmse_f64", "ap_relmse_f64_noalloc",
    "ap_guiding_scan_u32", "ap_guiding_sort_pairs_u64",
)


def lib():
    """Loads libaperture_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        L.ap_scene_create.argtypes = [vp, sz, ctypes.POINTER(BuildParams), ctypes.POINTER(vp)]
        L.ap_scene_destroy.argtypes = [vp]
        L.ap_build_bvh.argtypes = [vp, sz, ctypes.POINTER(BuildParams), vp, vp, ctypes.POINTER(ctypes.c_size_t)]
        L.ap_scene_get_info.argtypes = [vp, ctypes.POINTER(BVHInfo)]
        L.ap_scene_set_traversal_mode.argtypes = [vp, i32]
        L.ap_scene_set_launch.argtypes = [vp, i32, i32]
        L.ap_scene_get_launch.argtypes = [vp, vp, vp]
        L.ap_scene_autotune.argtypes = [vp, vp, sz, vp, i32, vp, vp, vp, vp]
        L.ap_guiding_reset.argtypes = [vp]
        L.ap_nrc_create.argtypes = [ctypes.c_uint64, vp]
        L.ap_nrc_destroy.argtypes = [vp]
        L.ap_nrc_set_bounds.argtypes = [vp, vp, vp]
        L.ap_nrc_get_params.argtypes = [vp, vp]
        L.ap_nrc_set_params.argtypes = [vp, vp]
        L.ap_nrc_infer.argtypes = [vp, vp, sz, vp, vp]
        L.ap_nrc_train_step.argtypes = [vp, vp, sz, ctypes.c_float, vp]
        L.ap_nrc_gradient.argtypes = [vp, vp, sz, vp, vp, vp]
        L.ap_nrc_last_gradient.argtypes = [vp, vp]
        L.ap_nrc_get_bounds.argtypes = [vp, vp, vp]
        L.ap_nrc_set_ema.argtypes = [vp, ctypes.c_float]
        L.ap_nrc_create_grouped.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64, vp]
        L.ap_nrc_info.argtypes = [vp, vp, vp, vp, vp]
        L.ap_nrc_set_group_bounds.argtypes = [vp, ctypes.c_uint32, vp, vp]
        L.ap_nrc_infer_grouped.argtypes = [vp, vp, vp, vp, vp]
        L.ap_nrc_gradient_grouped.argtypes = [vp, vp, sz, vp, vp, vp]
        L.ap_nrc_train_step_grouped.argtypes = [vp, vp, sz, ctypes.c_float, vp]
        L.ap_batch_create.argtypes = [vp, vp, sz, vp]
        L.ap_batch_destroy.argtypes = [vp]
        L.ap_batch_render.argtypes = [vp, vp, vp, vp, vp, vp]
        L.ap_batch_set_launch.argtypes = [vp, i32, i32]
        L.ap_batch_get_launch.argtypes = [vp, vp, vp]
        L.ap_batch_set_graph.argtypes = [vp, i32]
        L.ap_batch_autotune.argtypes = [vp, vp, i32, vp, vp, vp, vp]
        L.ap_render_nrc.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.ap_nrc_last_frame.argtypes = [vp, vp, vp, sz]
        L.ap_batch_render_nrc.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
        L.ap_batch_nrc_last_frame.argtypes = [vp, vp, sz]
        L.ap_tonemap.argtypes = [vp, vp, vp, vp, i32, i32, i32, i32, ctypes.c_float, vp]
        L.ap_sdf_cache_create.argtypes = [vp]
        L.ap_sdf_cache_destroy.argtypes = [vp]
        L.ap_sdf_cache_add_functor.argtypes = [vp, i32, vp, vp, vp]
        L.ap_sdf_cache_add_volume.argtypes = [vp, vp, vp, vp]
        L.ap_sdf_cache_get_volume.argtypes = [vp, ctypes.c_uint32, vp, vp]
        L.ap_scene_set_sdf_cache.argtypes = [vp, vp]
        L.ap_texture_cache_create.argtypes = [vp]
        L.ap_texture_cache_destroy.argtypes = [vp]
        L.ap_texture_cache_add_rgba8.argtypes = [vp, vp, i32, i32, vp]
        L.ap_texture_cache_count.argtypes = [vp, vp]
        L.ap_texture_sample.argtypes = [vp, ctypes.c_uint64, vp, sz, vp]
        L.ap_scene_set_texture_cache.argtypes = [vp, vp]
        L.ap_mesh_load_obj.argtypes = [ctypes.c_char_p, vp, ctypes.c_float, vp, i32, vp, sz, vp]
        L.ap_guiding_export.argtypes = [vp, vp, sz, vp, sz, vp, vp]
        L.ap_guiding_scan_u32.argtypes = [vp, vp, sz, vp]
        L.ap_guiding_sort_pairs_u64.argtypes = [vp, vp, vp, vp, sz, i32, i32]
        L.ap_scene_export_bvh.argtypes = [vp, vp, vp, vp]
        for name in ("ap_trace_closest", "ap_trace_any"):
            getattr(L, name).argtypes = [vp, vp, sz, vp, vp]
        L.ap_trace_closest_stats.argtypes = [vp, vp, sz, vp, vp, vp]
        L.ap_last_error_string.restype = ctypes.c_char_p
        L.ap_last_error_string.argtypes = []
        L.ap_synth_triangle_soup.argtypes = [sz, ctypes.c_uint64, ctypes.c_float, ctypes.c_float, vp]
        L.ap_synth_camera_rays.argtypes = [i32, i32, vp, vp, ctypes.c_float, vp]
        L.ap_synth_secondary_rays.argtypes = [vp, vp, vp, sz, ctypes.c_uint64, ctypes.c_float, vp,
                                              ctypes.POINTER(ctypes.c_size_t)]
        pvp = ctypes.POINTER(vp)
        L.ap_device_count.argtypes = [ctypes.POINTER(i32)]
        L.ap_device_set.argtypes = [i32]
        L.ap_device_name.argtypes = [i32, ctypes.c_char_p, sz]
        L.ap_device_malloc.argtypes = [sz, pvp]
        L.ap_device_free.argtypes = [vp]
        L.ap_memcpy.argtypes = [vp, vp, sz, i32]
        L.ap_memset_async.argtypes = [vp, i32, sz, vp]
        L.ap_device_copy.argtypes = [vp, vp, sz, vp]
        L.ap_stream_create.argtypes = [pvp]
        L.ap_stream_destroy.argtypes = [vp]
        L.ap_stream_synchronize.argtypes = [vp]
        L.ap_device_synchronize.argtypes = []
        L.ap_event_create.argtypes = [pvp]
        L.ap_event_destroy.argtypes = [vp]
        L.ap_event_record.argtypes = [vp, vp]
        L.ap_event_elapsed_ms.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_float)]
        L.ap_spectrum_setup.argtypes = []
        for name in ("ap_spectrum_encode_rgb", "ap_spectrum_encode_lambda", "ap_spectrum_query_rgb",
                     "ap_spectrum_query_lambda"):
            getattr(L, name).argtypes = [vp, sz, vp, vp]
        L.ap_spectrum_interact.argtypes = [vp, vp, sz, vp, vp]
        L.ap_spectrum_sample_lambda.argtypes = [vp, sz, vp, vp, vp]
        L.ap_render_params_default.argtypes = [ctypes.POINTER(RenderParams)]
        L.ap_render.argtypes = [vp, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams),
                                ctypes.POINTER(Framebuffers), vp]
        L.ap_render_offline.argtypes = [vp, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams), vp]
        L.ap_xyz_to_srgb.argtypes = [vp, sz, vp, vp]
        L.ap_mse_f64.argtypes = [vp, vp, sz, vp, vp]
        L.ap_mse_luminance_f64.argtypes = [vp, vp, sz, vp, vp]
        L.ap_relmse_f64.argtypes = [ctypes.POINTER(ImageView), ctypes.POINTER(ImageView), sz, ctypes.c_double, vp, vp]
        L.ap_relmse_f64_noalloc.argtypes = L.ap_relmse_f64.argtypes
        L.ap_synth_box_scene.argtypes = [ctypes.c_uint64, i32, ctypes.c_float, vp, sz, ctypes.POINTER(sz)]
        L.ap_synth_heightfield_scene.argtypes = [i32, i32, ctypes.c_uint64, ctypes.c_float, vp, sz,
                                                 ctypes.POINTER(sz)]
        for name in EXPORTED_SYMBOLS:
            if name in ("ap_render_params_default",):
                getattr(L, name).restype = None
                continue
            if name != "ap_last_error_string":
                getattr(L, name).restype = i32
        _lib = L
    return _lib


def check(code):
    if code != AP_OK:
        raise ApertureError(code, lib().ap_last_error_string().decode())
    return code


def _ptr(x):
    """Address of a numpy array, DeviceBuffer or integer."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if isinstance(x, np.ndarray):
        return x.ctypes.data_as(ctypes.c_void_p)  # keeps a temporary array alive through the call
    if isinstance(x, DeviceBuffer):
        return x.ptr
    return x.data_ptr()


def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    if isinstance(stream, Stream):
        return stream.handle
    return stream.cuda_stream


# ---------------------------------------------------------------------------
# Device plumbing (include/aperture/ap_device.h)
# ---------------------------------------------------------------------------
COPY_H2D, COPY_D2H, COPY_D2D = 1, 2, 3


def device_count():
    n = ctypes.c_int(0)
    check(lib().ap_device_count(ctypes.byref(n)))
    return n.value


def device_name(device=0):
    buf = ctypes.create_string_buffer(128)
    check(lib().ap_device_name(device, buf, 128))
    return buf.value.decode()


def synchronize():
    check(lib().ap_device_synchronize())


class DeviceBuffer:
    """A device allocation of `nbytes` bytes (hipMalloc through the C ABI)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().ap_device_malloc(max(self.nbytes, 1), ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr):
        a = np.ascontiguousarray(arr)
        buf = cls(a.nbytes)
        if a.nbytes:
            check(lib().ap_memcpy(buf.ptr, a.ctypes.data, a.nbytes, COPY_H2D))
        return buf

    def to_numpy(self, dtype, count=None):
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else int(count)
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib().ap_memcpy(out.ctypes.data, self.ptr, out.nbytes, COPY_D2H))
        return out

    def zero(self, stream=None):
        check(lib().ap_memset_async(self.ptr, 0, self.nbytes, _stream(stream)))

    def copy_from(self, src, nbytes=None, stream=None):
        """Stream-ordered device-to-device copy by the library's streaming kernel."""
        n = self.nbytes if nbytes is None else nbytes
        check(lib().ap_device_copy(self.ptr, _ptr(src), n, _stream(stream)))

    def free(self):
        if self.ptr:
            lib().ap_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().ap_stream_create(ctypes.byref(h)))
        self.handle = h.value

    def synchronize(self):
        check(lib().ap_stream_synchronize(self.handle))

    def __del__(self):
        try:
            lib().ap_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().ap_event_create(ctypes.byref(h)))
        self.handle = h.value

    def record(self, stream=None):
        check(lib().ap_event_record(self.handle, _stream(stream)))

    def elapsed_ms(self, end):
        ms = ctypes.c_float(0.0)
        check(lib().ap_event_elapsed_ms(self.handle, end.handle, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            lib().ap_event_destroy(self.handle)
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Synthetic workloads (include/aperture/ap_synth.h)
# ---------------------------------------------------------------------------
def synth_triangle_soup(n, seed=2, edge_min=1e-3, edge_max=1e-2):
    objs = np.zeros(n, dtype=SCENE_OBJECT_DTYPE)
    check(lib().ap_synth_triangle_soup(n, seed, edge_min, edge_max, _ptr(objs)))
    return objs


def synth_camera_rays(width, height, origin, direction, fov_deg):
    rays = np.zeros(width * height, dtype=RAY_DTYPE)
    o = np.asarray(origin, dtype=np.float32)
    d = np.asarray(direction, dtype=np.float32)
    check(lib().ap_synth_camera_rays(width, height, _ptr(o), _ptr(d), fov_deg, _ptr(rays)))
    return rays


def synth_box_scene(seed=1, sphere_subdiv=3, light=15.0):
    """The closed-box workload (include/aperture/ap_synth.h)."""
    n = ctypes.c_size_t(0)
    check(lib().ap_synth_box_scene(seed, sphere_subdiv, light, None, 0, ctypes.byref(n)))
    objs = np.zeros(n.value, dtype=SCENE_OBJECT_DTYPE)
    check(lib().ap_synth_box_scene(seed, sphere_subdiv, light, _ptr(objs), n.value, ctypes.byref(n)))
    return objs


def synth_heightfield_scene(grid=1024, n_spheres=256, seed=4, light=20.0):
    """Heightfield + spheres + 4 lights (ap_synth_heightfield_scene)."""
    n = ctypes.c_size_t(0)
    check(lib().ap_synth_heightfield_scene(grid, n_spheres, seed, light, None, 0, ctypes.byref(n)))
    objs = np.zeros(n.value, dtype=SCENE_OBJECT_DTYPE)
    check(lib().ap_synth_heightfield_scene(grid, n_spheres, seed, light, _ptr(objs), n.value, ctypes.byref(n)))
    return objs


def synth_batch_scenes(n_scenes=48, seed=6):
    """The batched heterogeneous workload: n_scenes heightfield + icosphere
    scenes with triangle counts log-uniform in [1e4, 5e5] (4% of them in
    320-triangle spheres, the rest a square heightfield grid; scene k seeded
    100 + k, 4 thin-cube lights) and square resolutions uniform in
    {128, 256, 512, 1024}, all viewed by the same camera. Returns (objects
    per scene, Camera per scene)."""
    g = np.random.default_rng(seed)
    objs, cams = [], []
    for k in range(n_scenes):
        tris = float(np.exp(g.uniform(np.log(1e4), np.log(5e5))))
        spheres = max(1, int(tris * 0.04 / 320))
        grid = max(8, int(round(np.sqrt(max(tris - 320 * spheres, 128) / 2))))
        res = int(g.choice([128, 256, 512, 1024]))
        objs.append(synth_heightfield_scene(grid, spheres, 100 + k, 20.0))
        cams.append(Camera.make([0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0, res, res))
    return objs, cams


def spectrum_setup():
    check(lib().ap_spectrum_setup())


def _spec_call(name, host_in, width_in, width_out, extra_out=False):
    a = np.ascontiguousarray(host_in, dtype=np.float32).reshape(-1)
    n = a.size // width_in
    d_in = DeviceBuffer.from_numpy(a)
    d_out = DeviceBuffer(max(n * width_out * 4, 4))
    if extra_out:
        d_out2 = DeviceBuffer(max(n * 4, 4))
        check(getattr(lib(), name)(d_in.ptr, n, d_out.ptr, d_out2.ptr, None))
        synchronize()
        return d_out.to_numpy(np.float32, n), d_out2.to_numpy(np.float32, n)
    check(getattr(lib(), name)(d_in.ptr, n, d_out.ptr, None))
    synchronize()
    return d_out.to_numpy(np.float32, n * width_out).reshape(n, width_out)


def spectrum_encode_rgb(rgb):
    return _spec_call("ap_spectrum_encode_rgb", rgb, 3, 4)


def spectrum_encode_lambda(lam):
    return _spec_call("ap_spectrum_encode_lambda", lam, 1, 4)


def spectrum_query_rgb(rgb):
    return _spec_call("ap_spectrum_query_rgb", rgb, 3, 4)


def spectrum_query_lambda(lam):
    return _spec_call("ap_spectrum_query_lambda", lam, 1, 4)


def spectrum_sample_lambda(u3):
    return _spec_call("ap_spectrum_sample_lambda", u3, 3, 1, extra_out=True)


def spectrum_interact(lam, rgb):
    lam = np.ascontiguousarray(lam, dtype=np.float32).reshape(-1)
    rgb = np.ascontiguousarray(rgb, dtype=np.float32).reshape(-1, 3)
    d_l, d_c, d_o = DeviceBuffer.from_numpy(lam), DeviceBuffer.from_numpy(rgb), DeviceBuffer(max(lam.size * 4, 4))
    check(lib().ap_spectrum_interact(d_l.ptr, d_c.ptr, lam.size, d_o.ptr, None))
    synchronize()
    return d_o.to_numpy(np.float32, lam.size)


IMAGE_XYZ_F32, IMAGE_XYZ_SUMS_F64 = 0, 1
REDUCE_BLOCKS = 1024
REDUCE_MAX_ELEMENTS = 1 << 34


class ImageView(ctypes.Structure):
    """ap_image_view (include/aperture/ap_measure.h)."""
    _fields_ = [("data", ctypes.c_void_p), ("kind", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("scale", ctypes.c_double)]

    @classmethod
    def of(cls, buf, kind=IMAGE_XYZ_F32, scale=1.0):
        p = _ptr(buf)
        return cls(p.value if isinstance(p, ctypes.c_void_p) else p, kind, 0, scale)


def _reduce_out(out, stream):
    """A device double for a reduction's result: the caller's, or a fresh one
    read back (synchronously) when out is None."""
    return (DeviceBuffer(8), True) if out is None else (out, False)


def _reduce_result(d_out, fresh, stream):
    if not fresh:
        return None
    if stream is not None:
        stream.synchronize()
    synchronize()
    return float(d_out.to_numpy(np.float64, 1)[0])


def guiding_scan_u32(values):
    """ap_guiding_scan_u32 of a host uint32 array: (exclusive prefix sums,
    total), computed on the device."""
    a = np.ascontiguousarray(values, dtype=np.uint32)
    n = a.size
    d_in = DeviceBuffer.from_numpy(a) if n else DeviceBuffer(4)
    d_out = DeviceBuffer(max(4 * n, 4))
    d_tot = DeviceBuffer(4)
    check(lib().ap_guiding_scan_u32(d_in.ptr, d_out.ptr, n, d_tot.ptr))
    return d_out.to_numpy(np.uint32, n), int(d_tot.to_numpy(np.uint32, 1)[0])


def guiding_sort_pairs_u64(keys, vals, begin_bit=0, end_bit=64):
    """ap_guiding_sort_pairs_u64 of host uint64 keys and uint32 values:
    (sorted keys, values in the same order), stable, on the device."""
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    v = np.ascontiguousarray(vals, dtype=np.uint32)
    n = k.size
    dk = DeviceBuffer.from_numpy(k) if n else DeviceBuffer(8)
    dv = DeviceBuffer.from_numpy(v) if n else DeviceBuffer(4)
    ok, ov = DeviceBuffer(max(8 * n, 8)), DeviceBuffer(max(4 * n, 4))
    check(lib().ap_guiding_sort_pairs_u64(dk.ptr, ok.ptr, dv.ptr, ov.ptr, n, begin_bit, end_bit))
    return ok.to_numpy(np.uint64, n), ov.to_numpy(np.uint32, n)


def mse_f64(a, b, n, out=None, stream=None):
    """ap_mse_f64 of n floats of device buffers a, b (FP64 mean of squared
    differences); with out=None the value is returned."""
    d_out, fresh = _reduce_out(out, stream)
    check(lib().ap_mse_f64(_ptr(a), _ptr(b), n, _ptr(d_out), _stream(stream)))
    return _reduce_result(d_out, fresh, stream)


def mse_luminance_f64(a, b, n_pixels, out=None, stream=None):
    """ap_mse_luminance_f64 of two float3 device images (DeviceMSELuminance)."""
    d_out, fresh = _reduce_out(out, stream)
    check(lib().ap_mse_luminance_f64(_ptr(a), _ptr(b), n_pixels, _ptr(d_out), _stream(stream)))
    return _reduce_result(d_out, fresh, stream)


def relmse_f64(image, reference, n_pixels, eps=1e-2, out=None, stream=None):
    """ap_relmse_f64: image and reference are ImageView (or a float3 XYZ
    DeviceBuffer)."""
    if not isinstance(image, ImageView):
        image = ImageView.of(image)
    if not isinstance(reference, ImageView):
        reference = ImageView.of(reference)
    d_out, fresh = _reduce_out(out, stream)
    check(lib().ap_relmse_f64(ctypes.byref(image), ctypes.byref(reference), n_pixels, eps, _ptr(d_out),
                              _stream(stream)))
    return _reduce_result(d_out, fresh, stream)


def xyz_to_srgb(xyz):
    """XYZ2SRGBLinearD65 on the device for an (..., 3) host array."""
    a = np.ascontiguousarray(xyz, dtype=np.float32)
    n = a.size // 3
    d_in, d_out = DeviceBuffer.from_numpy(a), DeviceBuffer(max(a.nbytes, 4))
    check(lib().ap_xyz_to_srgb(d_in.ptr, n, d_out.ptr, None))
    synchronize()
    return d_out.to_numpy(np.float32, a.size).reshape(a.shape)


SDF_TORUS, SDF_VOLUME = 0, 1
SDF_FUNCTOR_TORUS, SDF_FUNCTOR_DIAMOND = 0, 1


class TextureCache:
    """RGBA8 images behind GBuffer texture handles (ap_texture_cache_*)."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        check(lib().ap_texture_cache_create(ctypes.byref(self._h)))

    @property
    def handle(self):
        return self._h.value

    def add_rgba8(self, image):
        """An (height, width, 4) uint8 image, rows from the top; returns its handle (>= 1)."""
        img = np.ascontiguousarray(image, np.uint8)
        if img.ndim != 3 or img.shape[2] != 4:
            raise ValueError("image must be (height, width, 4) uint8")
        h = ctypes.c_uint64()
        check(lib().ap_texture_cache_add_rgba8(self.handle, _ptr(img), img.shape[1], img.shape[0], ctypes.byref(h)))
        return h.value

    def count(self):
        n = ctypes.c_uint32()
        check(lib().ap_texture_cache_count(self.handle, ctypes.byref(n)))
        return n.value

    def sample(self, handle, uv):
        """tex2D of image `handle` at (n, 2) UV pairs, as the integrator samples it: (n, 4) RGBA."""
        u = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        out = np.zeros((len(u), 4), np.float32)
        check(lib().ap_texture_sample(self.handle, handle, _ptr(u), len(u), _ptr(out)))
        return out

    def close(self):
        if self._h is not None and self._h.value:
            lib().ap_texture_cache_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


MESH_REFERENCE_PROXY = 1


def load_obj(path, position=(0.0, 0.0, 0.0), scale=1.0, textures=(0, 0, 0, 0), reference_proxy=False):
    """LoadTexturedMesh: the triangles of an OBJ file as a SCENE_OBJECT_DTYPE array
    (ap_mesh_load_obj); textures = albedo, normal, metallic, roughness handles.
    Proxy centres are the triangles' centroids unless reference_proxy (the
    reference's mesh position for every triangle, one BVH leaf per mesh)."""
    pos = np.ascontiguousarray(position, np.float32).reshape(3)
    tex = np.ascontiguousarray(textures, np.uint64).reshape(4)
    flags = MESH_REFERENCE_PROXY if reference_proxy else 0
    n = ctypes.c_size_t()
    p = os.fsencode(path)
    check(lib().ap_mesh_load_obj(p, _ptr(pos), float(scale), _ptr(tex), flags, None, 0, ctypes.byref(n)))
    out = np.zeros(n.value, SCENE_OBJECT_DTYPE)
    if n.value:
        check(lib().ap_mesh_load_obj(p, _ptr(pos), float(scale), _ptr(tex), flags, _ptr(out), n.value,
                                     ctypes.byref(n)))
    return out


def load_image_rgba8(path):
    """stbi_load(path, ..., 4): an image file as (height, width, 4) uint8 (needs Pillow)."""
    from PIL import Image

    with Image.open(path) as im:
        return np.asarray(im.convert("RGBA"), np.uint8).copy()


class SdfCache:
    """Volumes of signed distances for SDF_VOLUME objects (ap_sdf_cache_*)."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        check(lib().ap_sdf_cache_create(ctypes.byref(self._h)))

    @property
    def handle(self):
        return self._h.value

    def add_functor(self, functor, info, dims):
        """A volume of the torus or diamond functor on a dims grid; returns its index."""
        inf = np.ascontiguousarray(info, np.float32).reshape(4)
        dm = np.ascontiguousarray(dims, np.int32).reshape(3)
        idx = ctypes.c_uint32()
        check(lib().ap_sdf_cache_add_functor(self.handle, functor, _ptr(inf), _ptr(dm), ctypes.byref(idx)))
        return idx.value

    def add_volume(self, values):
        """A volume from a host (nz, ny, nx) array; returns its index."""
        v = np.ascontiguousarray(values, np.float32)
        dm = np.array([v.shape[2], v.shape[1], v.shape[0]], np.int32)
        idx = ctypes.c_uint32()
        check(lib().ap_sdf_cache_add_volume(self.handle, _ptr(v), _ptr(dm), ctypes.byref(idx)))
        return idx.value

    def volume(self, index):
        """Volume `index` as a host (nz, ny, nx) array."""
        dm = np.zeros(3, np.int32)
        check(lib().ap_sdf_cache_get_volume(self.handle, index, _ptr(dm), None))
        out = np.zeros((dm[2], dm[1], dm[0]), np.float32)
        check(lib().ap_sdf_cache_get_volume(self.handle, index, _ptr(dm), _ptr(out)))
        return out

    def close(self):
        if self._h is not None and self._h.value:
            lib().ap_sdf_cache_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tonemap(hdr, width, height, tone=TONE_LINEAR, display=DISPLAY_IMAGE, exposure=1.0, gbuffer=None,
            statistics=None, ldr=None, stream=None):
    """ap_tonemap over device buffers; returns the device ldr buffer (uint32 ARGB)."""
    out = ldr if ldr is not None else DeviceBuffer(width * height * 4)
    check(lib().ap_tonemap(_ptr(hdr), _ptr(gbuffer), _ptr(statistics), _ptr(out), width, height, tone, display,
                           exposure, _stream(stream)))
    return out


NRC_PARAMS = 21504


class RadianceCache:
    """The learned radiance cache (include/aperture/ap_nrc.h): one network,
    or `groups` independent networks of hidden width 32 or 64 in "bf16" or
    "fp16" served by grouped launches."""

    PRECISIONS = {"bf16": 0, "fp16": 1}

    def __init__(self, seed=1, groups=1, width=64, precision="bf16"):
        self._h = ctypes.c_void_p()
        if groups == 1 and width == 64 and precision == "bf16":
            check(lib().ap_nrc_create(seed, ctypes.byref(self._h)))
        else:
            check(lib().ap_nrc_create_grouped(groups, width, self.PRECISIONS[precision], seed,
                                              ctypes.byref(self._h)))
        g, w, pr, n = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_int32(), ctypes.c_size_t()
        check(lib().ap_nrc_info(self.handle, ctypes.byref(g), ctypes.byref(w), ctypes.byref(pr), ctypes.byref(n)))
        self.groups, self.width, self.n_params = g.value, w.value, n.value
        self.precision = "fp16" if pr.value == 1 else "bf16"

    @property
    def handle(self):
        return self._h.value

    def loss_scale(self, batch):
        """Power of two an FP16 network's loss gradient is scaled by (1 for BF16)."""
        if self.precision != "fp16":
            return 1.0
        k = 0
        while (1 << k) < 3 * batch:
            k += 1
        return 2.0 ** (k - 4)

    def set_bounds(self, bmin, bmax, group=None):
        lo = np.ascontiguousarray(bmin, np.float32)
        hi = np.ascontiguousarray(bmax, np.float32)
        if group is None:
            check(lib().ap_nrc_set_bounds(self.handle, _ptr(lo), _ptr(hi)))
        else:
            check(lib().ap_nrc_set_group_bounds(self.handle, group, _ptr(lo), _ptr(hi)))

    def set_ema(self, decay):
        check(lib().ap_nrc_set_ema(self.handle, decay))

    def bounds(self):
        lo, hi = np.zeros(3, np.float32), np.zeros(3, np.float32)
        check(lib().ap_nrc_get_bounds(self.handle, _ptr(lo), _ptr(hi)))
        return lo, hi

    def _shape(self, a):
        return a if self.groups == 1 else a.reshape(self.groups, self.n_params)

    def params(self):
        p = np.zeros(self.groups * self.n_params, np.float32)
        check(lib().ap_nrc_get_params(self.handle, _ptr(p)))
        return self._shape(p)

    def set_params(self, p):
        host = np.ascontiguousarray(p, np.float32).reshape(-1)
        if host.size != self.groups * self.n_params:
            raise ValueError(f"expected {self.groups * self.n_params} parameters, got {host.size}")
        check(lib().ap_nrc_set_params(self.handle, _ptr(host)))

    def infer(self, queries):
        """Host (n, 16) query records -> host (n, 3) RGB."""
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 16)
        d_q, d_o = DeviceBuffer.from_numpy(q), DeviceBuffer(max(len(q) * 12, 4))
        check(lib().ap_nrc_infer(self.handle, d_q.ptr, len(q), d_o.ptr, None))
        synchronize()
        return d_o.to_numpy(np.float32, len(q) * 3).reshape(-1, 3)

    def infer_grouped(self, queries, seg):
        """Host (n, 16) records, rows [seg[g], seg[g+1]) for network g -> host (n, 3)."""
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 16)
        sg = np.ascontiguousarray(seg, np.uint32)
        d_q, d_o = DeviceBuffer.from_numpy(q if len(q) else np.zeros((1, 16), np.float32)), \
            DeviceBuffer(max(len(q) * 12, 4))
        check(lib().ap_nrc_infer_grouped(self.handle, d_q.ptr, _ptr(sg), d_o.ptr, None))
        synchronize()
        return d_o.to_numpy(np.float32, len(q) * 3).reshape(-1, 3)

    def gradient(self, samples):
        """Loss and gradient over host samples: (n, 20) for one network,
        (groups, batch, 20) for a grouped cache (losses per network)."""
        s = np.ascontiguousarray(samples, np.float32)
        d_s = DeviceBuffer.from_numpy(s.reshape(-1, 20))
        g = np.zeros(self.groups * self.n_params, np.float32)
        loss = np.zeros(self.groups, np.float32)
        if self.groups == 1:
            check(lib().ap_nrc_gradient(self.handle, d_s.ptr, s.size // 20, _ptr(g), _ptr(loss), None))
            return float(loss[0]), g
        batch = s.size // 20 // self.groups
        check(lib().ap_nrc_gradient_grouped(self.handle, d_s.ptr, batch, _ptr(g), _ptr(loss), None))
        return loss, self._shape(g)

    def train_step(self, samples_device, n, lr=1e-3, stream=None):
        """One Adam step; n is the batch per network."""
        if self.groups == 1:
            check(lib().ap_nrc_train_step(self.handle, _ptr(samples_device), n, lr, _stream(stream)))
        else:
            check(lib().ap_nrc_train_step_grouped(self.handle, _ptr(samples_device), n, lr, _stream(stream)))

    def last_gradient(self):
        g = np.zeros(self.groups * self.n_params, np.float32)
        check(lib().ap_nrc_last_gradient(self.handle, _ptr(g)))
        return self._shape(g)

    def close(self):
        if self._h is not None and self._h.value:
            lib().ap_nrc_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Batched multi-scene rendering (include/aperture/ap_batch.h) over
    SceneRenderer objects, which must outlive the batch."""

    def __init__(self, scenes, cameras):
        self._scenes = list(scenes)
        self._cams = (Camera * len(cameras))(*cameras)
        arr = (ctypes.c_void_p * len(self._scenes))(*[sc.handle for sc in self._scenes])
        self._h = ctypes.c_void_p()
        check(lib().ap_batch_create(arr, self._cams, len(self._scenes), ctypes.byref(self._h)))

    @property
    def handle(self):
        return self._h.value

    def render(self, params, hdrs, sums=None, stream=None, count_rays=True):
        """Adds params.spp frames of every scene; hdrs[k] / sums[k] are device
        buffers (DeviceBuffer or address, None to skip). Returns the rays
        traced; with count_rays=False the call does not wait for its frames
        and returns None."""
        n = len(self._scenes)
        h = (ctypes.c_void_p * n)(*[_ptr(x) if x is not None else None for x in hdrs])
        sm = None if sums is None else (ctypes.c_void_p * n)(*[_ptr(x) if x is not None else None for x in sums])
        rays = (ctypes.c_uint64 * 2)() if count_rays else None
        check(lib().ap_batch_render(self.handle, ctypes.byref(params), h, sm, rays, _stream(stream)))
        return (int(rays[0]), int(rays[1])) if count_rays else None

    def render_nrc(self, params, cache, nrc_params, hdrs, sums=None, stream=None):
        """render() with per-scene radiance caches: `cache` is a RadianceCache
        with one network per scene. Returns the rays traced."""
        n = len(self._scenes)
        h = (ctypes.c_void_p * n)(*[_ptr(x) if x is not None else None for x in hdrs])
        sm = None if sums is None else (ctypes.c_void_p * n)(*[_ptr(x) if x is not None else None for x in sums])
        rays = (ctypes.c_uint64 * 2)()
        check(lib().ap_batch_render_nrc(self.handle, ctypes.byref(params), cache.handle, ctypes.byref(nrc_params), h,
                                        sm, rays, _stream(stream)))
        return int(rays[0]), int(rays[1])

    def nrc_last_frame(self):
        """Per-scene NrcFrameStats of the last render_nrc frame."""
        n = len(self._scenes)
        st = (NrcFrameStats * n)()
        check(lib().ap_batch_nrc_last_frame(self.handle, st, n))
        return list(st)

    def set_launch(self, wave_rays, refill):
        check(lib().ap_batch_set_launch(self.handle, wave_rays, refill))

    def get_launch(self):
        w, r = ctypes.c_int(), ctypes.c_int()
        check(lib().ap_batch_get_launch(self.handle, ctypes.byref(w), ctypes.byref(r)))
        return w.value, r.value

    def set_graph(self, enable):
        check(lib().ap_batch_set_graph(self.handle, 1 if enable else 0))

    def autotune(self, params, frames=2, stream=None):
        w, r, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        check(lib().ap_batch_autotune(self.handle, ctypes.byref(params), frames, ctypes.byref(w), ctypes.byref(r),
                                      ctypes.byref(ms), _stream(stream)))
        return w.value, r.value, ms.value

    def close(self):
        if self._h is not None and self._h.value:
            lib().ap_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_bvh(scene_objects, leaf_size=8, bin_count=12):
    """Host BVH build (BuildBVH, core/bvh/bvh.cuh:142-422): (nodes, object_indices)."""
    objs = np.ascontiguousarray(scene_objects, dtype=SCENE_OBJECT_DTYPE)
    n = len(objs)
    nodes = np.zeros(max(2 * n, 1), dtype=BVH_NODE_DTYPE)
    idx = np.zeros(max(n, 1), dtype=np.int32)
    count = ctypes.c_size_t(0)
    check(lib().ap_build_bvh(_ptr(objs), n, ctypes.byref(BuildParams(leaf_size, bin_count)), _ptr(nodes),
                             _ptr(idx), ctypes.byref(count)))
    return nodes[: count.value], idx[:n]


class SceneRenderer:
    """Owns one device scene; mirrors the reference SceneRenderer
    (core/scene.hpp:113, core/scene.cu:1025-1118)."""

    def __init__(self, scene_objects, leaf_size=8, bin_count=12):
        objs = np.ascontiguousarray(scene_objects, dtype=SCENE_OBJECT_DTYPE)
        self._objects = objs
        self._handle = ctypes.c_void_p()
        params = BuildParams(leaf_size, bin_count)
        check(lib().ap_scene_create(_ptr(objs) if len(objs) else None, len(objs), ctypes.byref(params),
                                    ctypes.byref(self._handle)))

    @property
    def handle(self):
        return self._handle.value

    def info(self):
        out = BVHInfo()
        check(lib().ap_scene_get_info(self.handle, ctypes.byref(out)))
        return out

    def set_sdf_cache(self, cache):
        """The SdfCache the scene's SDF_VOLUME objects read (None: none)."""
        self._sdf_cache = cache
        check(lib().ap_scene_set_sdf_cache(self.handle, cache.handle if cache is not None else None))

    def set_texture_cache(self, cache):
        """The TextureCache the scene's texture handles name (None: none)."""
        self._tex_cache = cache
        check(lib().ap_scene_set_texture_cache(self.handle, cache.handle if cache is not None else None))

    def set_traversal_mode(self, mode):
        check(lib().ap_scene_set_traversal_mode(self.handle, mode))

    def set_launch(self, wave_rays=0, refill=0):
        """Traversal launch shape (rays per wave, idle lanes per refill); 0 keeps a value."""
        check(lib().ap_scene_set_launch(self.handle, wave_rays, refill))

    def get_launch(self):
        w, r = ctypes.c_int32(), ctypes.c_int32()
        check(lib().ap_scene_get_launch(self.handle, ctypes.byref(w), ctypes.byref(r)))
        return w.value, r.value

    def autotune(self, rays, hits, n=None, any_hit=False, stream=None):
        """Times the candidate launch shapes on a device ray batch, keeps the
        fastest and returns (wave_rays, refill, ms)."""
        n = n if n is not None else _numel_rays(rays)
        w, r, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_float()
        check(lib().ap_scene_autotune(self.handle, _ptr(rays), n, _ptr(hits), 1 if any_hit else 0, _stream(stream),
                                      ctypes.byref(w), ctypes.byref(r), ctypes.byref(ms)))
        return w.value, r.value, ms.value

    def guiding_reset(self):
        check(lib().ap_guiding_reset(self.handle))

    def guiding_export(self):
        """(probe texels (N, 86, 4) float32, counts (N,) uint64, index volume int32, bounds int32[9])."""
        bounds = np.zeros(9, np.int32)
        count = ctypes.c_int32()
        check(lib().ap_guiding_export(self.handle, None, 0, None, 0, _ptr(bounds), ctypes.byref(count)))
        n = count.value
        nvol = int(bounds[6]) * int(bounds[7]) * int(bounds[8]) * 20
        raw = np.zeros(max(n, 1) * 1392, np.uint8)
        vol = np.zeros(max(nvol, 1), np.int32)
        check(lib().ap_guiding_export(self.handle, _ptr(raw), n, _ptr(vol), nvol, _ptr(bounds), ctypes.byref(count)))
        rec = raw[: n * 1392].reshape(n, 1392)
        data = rec[:, :1376].copy().view(np.float32).reshape(n, 86, 4)
        counts = rec[:, 1376:1384].copy().view(np.uint64).reshape(n)
        return data, counts, vol[:nvol], bounds

    def export_bvh(self):
        inf = self.info()
        nodes = np.zeros(inf.node_count, dtype=BVH_NODE_DTYPE)
        idx = np.zeros(inf.prim_count, dtype=np.int32)
        lights = np.zeros(inf.light_count, dtype=LIGHT_INFO_DTYPE)
        check(lib().ap_scene_export_bvh(self.handle, _ptr(nodes), _ptr(idx), _ptr(lights)))
        return nodes, idx, lights

    def trace_closest(self, rays, hits, n=None, stream=None, stats=None):
        """Closest hit of n device rays (ap_ray, 32 B) into device hits (ap_hit, 8 B)."""
        n = n if n is not None else _numel_rays(rays)
        if stats is None:
            check(lib().ap_trace_closest(self.handle, _ptr(rays), n, _ptr(hits), _stream(stream)))
        else:
            check(lib().ap_trace_closest_stats(self.handle, _ptr(rays), n, _ptr(hits), _ptr(stats),
                                               _stream(stream)))

    def trace_any(self, rays, hits, n=None, stream=None):
        n = n if n is not None else _numel_rays(rays)
        check(lib().ap_trace_any(self.handle, _ptr(rays), n, _ptr(hits), _stream(stream)))

    def secondary_rays(self, primaries, hits, seed=3, offset=1e-4):
        """Host generator: one cosine-weighted ray per primary hit."""
        primaries = np.ascontiguousarray(primaries, dtype=RAY_DTYPE)
        hits = np.ascontiguousarray(hits, dtype=HIT_DTYPE)
        out = np.zeros(len(primaries), dtype=RAY_DTYPE)
        n_out = ctypes.c_size_t(0)
        check(lib().ap_synth_secondary_rays(self.handle, _ptr(primaries), _ptr(hits), len(primaries), seed,
                                            offset, _ptr(out), ctypes.byref(n_out)))
        return out[: n_out.value]

    def render(self, camera, params, hdr, odd=None, even=None, sum_xyz=None, primary_id=None, stream=None,
               ray_count=None, gbuffer=None, statistics=None):
        """SceneRenderer::Render into caller-owned device framebuffers (DeviceBuffer or addresses)."""
        fb = Framebuffers(_ptr(hdr), _ptr(odd), _ptr(even), _ptr(sum_xyz), _ptr(primary_id), _ptr(ray_count),
                          _ptr(gbuffer), _ptr(statistics))
        check(lib().ap_render(self.handle, ctypes.byref(camera), ctypes.byref(params), ctypes.byref(fb),
                              _stream(stream)))

    def render_nrc(self, camera, params, cache, nrc_params, hdr, odd=None, even=None, sum_xyz=None,
                   primary_id=None, stream=None, ray_count=None, gbuffer=None, statistics=None):
        """render() with the learned radiance cache `cache` (a RadianceCache)."""
        fb = Framebuffers(_ptr(hdr), _ptr(odd), _ptr(even), _ptr(sum_xyz), _ptr(primary_id), _ptr(ray_count),
                          _ptr(gbuffer), _ptr(statistics))
        check(lib().ap_render_nrc(self.handle, ctypes.byref(camera), ctypes.byref(params), cache.handle,
                                  ctypes.byref(nrc_params), ctypes.byref(fb), _stream(stream)))

    def nrc_last_frame(self, samples=True):
        """Counters of the last render_nrc frame and its training samples, (n, 20) float32."""
        st = NrcFrameStats()
        check(lib().ap_nrc_last_frame(self.handle, ctypes.byref(st), None, 0))
        out = np.zeros((st.samples, 20), np.float32)
        if samples and st.samples:
            check(lib().ap_nrc_last_frame(self.handle, ctypes.byref(st), _ptr(out), st.samples))
        return st, out

    def render_offline(self, camera, params):
        """SceneRenderer::Render(size, origin, dir, spp): host (H, W, 3) linear XYZ."""
        out = np.zeros(camera.width * camera.height * 3, np.float32)
        check(lib().ap_render_offline(self.handle, ctypes.byref(camera), ctypes.byref(params), _ptr(out)))
        return out.reshape(camera.height, camera.width, 3)

    def close(self):
        if self._handle is not None and self._handle.value:
            lib().ap_scene_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _numel_rays(rays):
    return int(rays.nbytes) // 32
