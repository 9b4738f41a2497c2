"""ctypes binding of liblrhip.so (include/luce_render_hip.h).

The library is built in-tree by luce-render_amd/build.py. Loading fails loudly
when it is missing: there is no CPU fallback for any entry point.
"""
import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "liblrhip.so"

OK, INVALID, LIMIT, STOPPED, DEVICE_ERROR, UNAVAILABLE = 0, 400, 401, 402, 500, 503


class LuceError(Exception):
    """A failed call: `code` as src/render/module.lucb:60-63 names them."""

    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


class GpuUnavailable(LuceError):
    pass


class BvhNode(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_float * 3), ("first", ctypes.c_uint32),
                ("hi", ctypes.c_float * 3), ("count", ctypes.c_uint32)]


class SceneArrays(ctypes.Structure):
    _fields_ = [("positions", ctypes.c_void_p), ("points", ctypes.c_size_t),
                ("indices", ctypes.c_void_p), ("faces", ctypes.c_void_p),
                ("shading", ctypes.c_void_p), ("uvs", ctypes.c_void_p),
                ("triangles", ctypes.c_size_t), ("nodes", ctypes.c_void_p),
                ("node_count", ctypes.c_size_t), ("splat_records", ctypes.c_void_p), ("splats", ctypes.c_size_t),
                ("splat_harmonics", ctypes.c_void_p), ("splat_words", ctypes.c_size_t),
                ("splat_nodes", ctypes.c_void_p), ("splat_node_count", ctypes.c_size_t)]


class TraceConfig(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("group", ctypes.c_uint32), ("stack", ctypes.c_uint32),
                ("refill", ctypes.c_uint32)]


class SceneInfo(ctypes.Structure):
    _fields_ = [("triangles", ctypes.c_size_t), ("binary_nodes", ctypes.c_size_t),
                ("wide_nodes", ctypes.c_size_t * 3), ("wide_bytes", ctypes.c_size_t * 3),
                ("triangle_bytes", ctypes.c_size_t), ("stack_bound", ctypes.c_uint32 * 3)]


class Loaded(ctypes.Structure):
    _fields_ = [("lights", ctypes.c_void_p), ("light_floats", ctypes.c_size_t), ("constants", ctypes.c_void_p),
                ("materials", ctypes.c_void_p), ("material_floats", ctypes.c_size_t), ("emitters", ctypes.c_void_p),
                ("emitter_words", ctypes.c_size_t), ("light_paths", ctypes.c_uint32), ("through", ctypes.c_uint32),
                ("programs", ctypes.c_void_p), ("program_words", ctypes.c_size_t), ("textures", ctypes.c_void_p),
                ("texture_floats", ctypes.c_size_t), ("texture_table", ctypes.c_void_p),
                ("texture_count", ctypes.c_size_t)]


class RenderSettingsC(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("width", "height", "samples", "max_bounces", "spectral", "pool", "seed",
                                                "guiding", "profile", "use_graph", "trace_width", "trace_group",
                                                "trace_stack", "train_samples", "train_steps", "pass_samples")]


class Image(ctypes.Structure):
    _fields_ = [("rgba", ctypes.c_void_p), ("display", ctypes.c_void_p), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32)]


class RenderStats(ctypes.Structure):
    _fields_ = [("samples_done", ctypes.c_uint32), ("light_samples", ctypes.c_uint32),
                ("closest_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64)]


class ServerSettings(ctypes.Structure):
    _fields_ = [("budget_bytes", ctypes.c_uint64), ("spectral", ctypes.c_uint32), ("trace_width", ctypes.c_uint32),
                ("trace_group", ctypes.c_uint32), ("trace_stack", ctypes.c_uint32), ("max_active", ctypes.c_uint32)]


class RequestStatus(ctypes.Structure):
    _fields_ = [("state", ctypes.c_uint32), ("samples_done", ctypes.c_uint32), ("samples", ctypes.c_uint32),
                ("bytes", ctypes.c_uint64)]


_P = ctypes.c_void_p
_SIGNATURES = {
    "lr_device_open": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
    "lr_device_close": (None, [_P]),
    "lr_last_error": (ctypes.c_size_t, [ctypes.c_char_p, ctypes.c_size_t]),
    "lr_bvh_build": (ctypes.c_int, [_P, ctypes.c_size_t, _P, ctypes.c_size_t, ctypes.c_int, _P,
                                    ctypes.POINTER(ctypes.c_size_t), _P]),
    "lr_bvh_build_boxes": (ctypes.c_int, [_P, ctypes.c_size_t, ctypes.c_int, _P, ctypes.POINTER(ctypes.c_size_t), _P]),
    "lr_bvh_split_triangles": (ctypes.c_int, [_P, ctypes.c_size_t, _P, ctypes.c_size_t, _P, ctypes.c_float,
                                              ctypes.c_uint32, ctypes.c_int, _P, _P, ctypes.c_size_t,
                                              ctypes.POINTER(ctypes.c_size_t)]),
    "lr_bvh_build_device": (ctypes.c_int, [_P, _P, ctypes.c_size_t, _P, ctypes.c_size_t, _P,
                                           ctypes.POINTER(ctypes.c_size_t), _P]),
    "lr_scene_upload": (ctypes.c_int, [_P, ctypes.POINTER(SceneArrays), ctypes.POINTER(_P)]),
    "lr_scene_destroy": (None, [_P]),
    "lr_scene_build": (ctypes.c_int, [_P, ctypes.POINTER(SceneArrays), ctypes.POINTER(_P), _P,
                                      ctypes.POINTER(ctypes.c_size_t), _P]),
    "lr_scene_refit": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "lr_scene_read_layout": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "lr_bvh_collapse": (ctypes.c_int, [_P, ctypes.c_size_t, ctypes.c_int, _P, ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_uint32)]),
    "lr_probe": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, ctypes.c_uint32, _P, _P, ctypes.c_size_t]),
    "lr_scene_set_trace_config": (ctypes.c_int, [_P, ctypes.POINTER(TraceConfig)]),
    "lr_scene_get_trace_config": (ctypes.c_int, [_P, ctypes.POINTER(TraceConfig)]),
    "lr_probe_device": (ctypes.c_int, [_P, _P, ctypes.c_uint32, ctypes.c_uint32, _P, _P]),
    "lr_trace_closest_device": (ctypes.c_int, [_P, _P, ctypes.c_size_t, _P, _P]),
    "lr_trace_frame_and_closest_device": (ctypes.c_int, [_P, _P, ctypes.c_uint32, ctypes.c_uint32, _P, _P,
                                                         ctypes.c_size_t, _P, _P]),
    "lr_trace_any_device": (ctypes.c_int, [_P, _P, ctypes.c_size_t, _P, _P]),
    "lr_trace_resident_grid": (ctypes.c_int, [_P, ctypes.c_size_t, ctypes.c_int, _P]),
    "lr_trace_closest_device_noalloc": (ctypes.c_int, [_P, _P, ctypes.c_size_t, _P, _P]),
    "lr_scene_info_get": (ctypes.c_int, [_P, ctypes.POINTER(SceneInfo)]),
    "lr_renderer_create": (ctypes.c_int, [_P, _P, ctypes.POINTER(Loaded), ctypes.POINTER(RenderSettingsC), _P,
                                          ctypes.POINTER(_P)]),
    "lr_renderer_advance": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int)]),
    "lr_renderer_finish": (ctypes.c_int, [_P]),
    "lr_renderer_samples_done": (ctypes.c_uint32, [_P]),
    "lr_renderer_image": (ctypes.c_int, [_P, ctypes.POINTER(Image)]),
    "lr_renderer_read": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "lr_renderer_read_display": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "lr_renderer_set_display": (ctypes.c_int, [_P, ctypes.c_float, ctypes.c_int]),
    "lr_renderer_read_film": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "lr_renderer_read_guide": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "lr_renderer_kernel_times": (ctypes.c_int, [_P, _P]),
    "lr_renderer_destroy": (None, [_P]),
    "lr_renderer_stats": (ctypes.c_int, [_P, ctypes.POINTER(RenderStats)]),
    "lr_server_create": (ctypes.c_int, [_P, ctypes.POINTER(ServerSettings), ctypes.POINTER(_P)]),
    "lr_server_submit": (ctypes.c_int, [_P, _P, ctypes.POINTER(Loaded), ctypes.POINTER(RenderSettingsC), _P,
                                        ctypes.POINTER(ctypes.c_uint64)]),
    "lr_server_step": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]),
    "lr_server_status": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.POINTER(RequestStatus)]),
    "lr_server_read": (ctypes.c_int, [_P, ctypes.c_uint64, _P, ctypes.c_size_t]),
    "lr_server_stats": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.POINTER(RenderStats)]),
    "lr_server_release": (ctypes.c_int, [_P, ctypes.c_uint64]),
    "lr_server_destroy": (None, [_P]),
    "lr_buffer_alloc": (ctypes.c_int, [_P, ctypes.c_size_t, ctypes.POINTER(_P)]),
    "lr_buffer_free": (None, [_P, _P]),
    "lr_buffer_write": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    "lr_buffer_read": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    "lr_buffer_copy": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    "lr_device_stream": (_P, [_P]),
    "lr_device_synchronize": (ctypes.c_int, [_P]),
    "lr_event_create": (ctypes.c_int, [ctypes.POINTER(_P)]),
    "lr_event_record": (ctypes.c_int, [_P, _P]),
    "lr_event_elapsed": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_float)]),
    "lr_event_destroy": (None, [_P]),
    "lr_renderer_reset": (ctypes.c_int, [_P, _P, ctypes.POINTER(Loaded), ctypes.POINTER(RenderSettingsC), _P]),
    "lr_renderer_converged_pixels": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_size_t)]),
    "lr_guide_net_create": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, ctypes.POINTER(_P)]),
    "lr_guide_net_destroy": (None, [_P]),
    "lr_guide_net_param_count": (ctypes.c_size_t, []),
    "lr_guide_net_set_params": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "lr_guide_net_get_params": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "lr_guide_net_gradient": (ctypes.c_int, [_P, _P, ctypes.c_size_t, _P, ctypes.c_size_t,
                                             ctypes.POINTER(ctypes.c_float)]),
    "lr_guide_net_step": (ctypes.c_int, [_P, _P, ctypes.c_size_t, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]),
    "lr_guide_net_mixtures": (ctypes.c_int, [_P, _P, ctypes.c_size_t, _P]),
}

_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built: run python luce-render_amd/build.py")
        handle = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | ctypes.RTLD_GLOBAL)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def exported_names():
    return list(_SIGNATURES)


def last_error():
    buf = ctypes.create_string_buffer(1024)
    lib().lr_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(code):
    if code == OK:
        return
    message = last_error()
    if code == UNAVAILABLE:
        raise GpuUnavailable(code, message)
    raise LuceError(code, message)


def ptr(array):
    """The data pointer of a numpy array (None for None)."""
    if array is None:
        return None
    return array.ctypes.data
