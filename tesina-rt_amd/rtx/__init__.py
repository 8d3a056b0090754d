"""Python binding of librtx.so (the gfx950 ray-tracing hot path) over its C ABI.

Thin ctypes layer: every call goes to the HIP library; there is no CPU
fallback.  Loading fails loudly if librtx.so has not been built
(python tesina-rt_amd/build.py).

Object model mirrors the reference's stage classes (include/rt/*.hpp):
Scene (geometry, materials, lights, LBVH), Rays (RayBundle), Hits
(HitBundle), FrameBuffer, Renderer.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent.parent
LIB_PATH = Path(os.environ.get("RTX_LIB", PKG / "librtx.so"))

DEVICE_COUNT = (1 << 64) - 1  # RTX_DEVICE_COUNT

LIGHT_SPOT, LIGHT_DIR, LIGHT_POINT = 0, 1, 2
MAX_LIGHTS = 8

# FrameStats stage ids (include/rt/frame-stats.hpp:5-17)
STAGES = ["BVH_BUILD", "PRIM_RAY_GEN", "SEC_RAY_GEN", "PRIM_TRACE", "PRIM_SHADOW_TRACE", "SEC_TRACE",
          "SEC_SHADOW_TRACE", "SHADE", "FB_CLEAR", "FB_COPY"]


class RtxError(RuntimeError):
    pass


class Material(C.Structure):
    _fields_ = [("diffuse", C.c_float * 4), ("shininess", C.c_float), ("reflectiveness", C.c_float),
                ("refractive_index", C.c_float), ("texture", C.c_int32)]


class Light(C.Structure):
    _fields_ = [("type", C.c_int32), ("pos", C.c_float * 3), ("dir", C.c_float * 3), ("color", C.c_float * 3),
                ("angle", C.c_float), ("radius", C.c_float)]


class Lights(C.Structure):
    _fields_ = [("ambient", C.c_float * 3), ("count", C.c_int32), ("l", Light * MAX_LIGHTS)]


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("dir", C.c_float * 3), ("right", C.c_float * 3), ("up", C.c_float * 3)]


SERVER_MAX_CLASSES = 4


class ServerParams(C.Structure):
    _fields_ = [("workers", C.c_int32), ("max_bounces", C.c_int32), ("classes", C.c_int32),
                ("class_w", C.c_int32 * SERVER_MAX_CLASSES), ("class_h", C.c_int32 * SERVER_MAX_CLASSES),
                ("hbm_budget", C.c_uint64), ("plain", C.c_int32), ("queue_cap", C.c_int32),
                ("arena_bytes", C.c_uint64), ("groups", C.c_int32), ("steps_ahead", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("scene", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("bounces", C.c_int32),
                ("camera", C.c_float * 12), ("arrival_ns", C.c_int64), ("tag", C.c_uint64),
                ("out_accum", C.c_void_p)]


class Completion(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("arrival_ns", C.c_int64), ("start_ns", C.c_int64), ("done_ns", C.c_int64),
                ("rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("status", C.c_int32), ("worker", C.c_int32)]


class WavefrontParams(C.Structure):
    _fields_ = [("arena_bytes", C.c_uint64), ("max_frames", C.c_int32), ("quad_size", C.c_int32)]


class LbvhParams(C.Structure):
    _fields_ = [("depth", C.c_int32), ("min_leaf", C.c_int32)]


class NodeRef(C.Structure):
    _fields_ = [("hi", C.c_float * 4), ("lo", C.c_float * 4), ("a", C.c_uint32), ("b", C.c_uint32),
                ("parent", C.c_uint32), ("split_axis", C.c_int8), ("leaf", C.c_int8), ("pad", C.c_uint8 * 2)]


NODE_DTYPE = np.dtype([("hi", "<f4", 4), ("lo", "<f4", 4), ("a", "<u4"), ("b", "<u4"), ("parent", "<u4"),
                       ("split_axis", "i1"), ("leaf", "i1"), ("pad", "u1", 2)])
assert NODE_DTYPE.itemsize == 48


class FrameStats(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("secondary_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("bounces_run", C.c_uint32), ("node_count", C.c_uint32), ("stage_ms", C.c_double * 10),
                ("frame_ms", C.c_double)]


class ObjGroup(C.Structure):
    _fields_ = [("first_triangle", C.c_uint32), ("triangle_count", C.c_uint32), ("material", C.c_int32),
                ("ambient", C.c_float * 4), ("diffuse", C.c_float * 4), ("specular", C.c_float * 4),
                ("shininess", C.c_float), ("alpha", C.c_float), ("name", C.c_char * 128),
                ("color_map", C.c_char * 256), ("bump_map", C.c_char * 256)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("quad_size", C.c_int32), ("max_bounces", C.c_int32),
                ("depth", C.c_int32), ("min_leaf", C.c_int32), ("ray_capacity", C.c_uint64)]


# Every symbol include/rtx.h declares (checked by tests/test_abi.py).
SYMBOLS = [
    "rtx_init", "rtx_shutdown", "rtx_last_error", "rtx_abi_version", "rtx_stream", "rtx_sync",
    "rtx_profile_enable", "rtx_profile_enabled", "rtx_profile_report", "rtx_profile_reset", "rtx_mem_in_use",
    "rtx_tune_set", "rtx_tune_get",
    "rtx_scene_create", "rtx_scene_destroy", "rtx_scene_set_vertices", "rtx_scene_set_vertices_ref",
    "rtx_scene_update_positions", "rtx_scene_set_triangles", "rtx_scene_set_materials", "rtx_scene_set_lights",
    "rtx_scene_set_textures", "rtx_scene_set_cubemap",
    "rtx_scene_triangle_count", "rtx_scene_displace_waves", "rtx_scene_set_waves", "rtx_scene_set_clock",
    "rtx_scene_get_clock", "rtx_scene_animate",
    "rtx_lbvh_build", "rtx_lbvh_build_noalloc", "rtx_lbvh_refit", "rtx_lbvh_info", "rtx_sah_build", "rtx_sah_build_host",
    "rtx_read_codes", "rtx_read_perm", "rtx_read_indices", "rtx_read_matmap", "rtx_read_orig_ids",
    "rtx_read_global_bbox", "rtx_read_vertices", "rtx_read_nodes_ref",
    "rtx_rays_create", "rtx_rays_destroy", "rtx_rays_capacity", "rtx_rays_count", "rtx_rays_read", "rtx_rays_write",
    "rtx_hits_create", "rtx_hits_destroy", "rtx_hits_read", "rtx_hits_write",
    "rtx_raygen_primary", "rtx_trace", "rtx_trace_shadow", "rtx_raygen_secondary",
    "rtx_bvh_create", "rtx_bvh_destroy", "rtx_bvh_copy", "rtx_trace_with", "rtx_trace_shadow_with",
    "rtx_fb_create", "rtx_fb_destroy", "rtx_fb_clear", "rtx_fb_resolve", "rtx_fb_read_accum", "rtx_fb_read_rgba8",
    "rtx_shade",
    "rtx_renderer_create", "rtx_renderer_destroy", "rtx_renderer_fb", "rtx_render_frame", "rtx_render_frame_async",
    "rtx_renderer_last_stats", "rtx_renderer_capture", "rtx_renderer_replay", "rtx_renderer_set_animation",
    "rtx_renderer_set_pipelined", "rtx_renderer_set_bounces",
    "rtx_wavefront_create", "rtx_wavefront_destroy", "rtx_wavefront_frame_bytes", "rtx_wavefront_add",
    "rtx_wavefront_step", "rtx_wavefront_run", "rtx_wavefront_sync", "rtx_wavefront_poll",
    "rtx_wavefront_frame_stats", "rtx_wavefront_read_accum", "rtx_wavefront_read_rgba8", "rtx_wavefront_release",
    "rtx_wavefront_memory",
    "rtx_now_ns", "rtx_server_create", "rtx_server_destroy", "rtx_server_add_scene", "rtx_server_memory",
    "rtx_server_submit", "rtx_server_poll", "rtx_server_drain",
    "rtx_batch_create", "rtx_batch_destroy", "rtx_batch_render", "rtx_batch_capture", "rtx_batch_replay",
    "rtx_batch_stats",
    "rtx_pipeline_create", "rtx_pipeline_destroy", "rtx_pipeline_set_animation", "rtx_pipeline_frame",
    "rtx_pipeline_wait", "rtx_pipeline_fb", "rtx_pipeline_stats",
    "rtx_obj_load", "rtx_obj_destroy", "rtx_obj_info", "rtx_obj_read", "rtx_obj_group_info", "rtx_obj_read_group",
    "rtx_instances_create", "rtx_instances_destroy", "rtx_instances_set_transforms", "rtx_instances_set_lights",
    "rtx_instances_tri_base", "rtx_trace_instances", "rtx_trace_shadow_instances", "rtx_instances_set_materials",
    "rtx_instances_shading_scene", "rtx_render_instances",
    "rtx_camera_set",
]

_lib = None


def lib() -> C.CDLL:
    """Load librtx.so (raises if it is missing: there is no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RtxError(f"{LIB_PATH} not built; run python tesina-rt_amd/build.py")
    L = C.CDLL(str(LIB_PATH))
    P, I32, U32, U64, F = C.c_void_p, C.c_int32, C.c_uint32, C.c_uint64, C.c_float
    sig = {
        "rtx_init": (I32, [I32, C.POINTER(P)]),
        "rtx_shutdown": (I32, [P]),
        "rtx_last_error": (C.c_char_p, []),
        "rtx_abi_version": (I32, []),
        "rtx_stream": (P, [P]),
        "rtx_sync": (I32, [P, P]),
        "rtx_profile_enable": (I32, [P, I32]),
        "rtx_profile_enabled": (I32, [P]),
        "rtx_profile_report": (C.c_int64, [P, C.c_char_p, U64]),
        "rtx_profile_reset": (I32, [P]),
        "rtx_mem_in_use": (U64, [P]),
        "rtx_tune_set": (I32, [P, C.c_char_p, C.c_int64]),
        "rtx_tune_get": (C.c_int64, [P, C.c_char_p]),
        "rtx_scene_create": (I32, [P, C.POINTER(P)]),
        "rtx_scene_destroy": (I32, [P]),
        "rtx_scene_set_vertices": (I32, [P, P, P, P, U32]),
        "rtx_scene_set_vertices_ref": (I32, [P, P, U32]),
        "rtx_scene_update_positions": (I32, [P, P, P]),
        "rtx_scene_set_triangles": (I32, [P, P, P, U32]),
        "rtx_scene_set_materials": (I32, [P, P, U32]),
        "rtx_scene_set_lights": (I32, [P, P]),
        "rtx_scene_triangle_count": (U32, [P]),
        "rtx_scene_displace_waves": (I32, [P, P, I32, F, P]),
        "rtx_scene_set_waves": (I32, [P, P, I32]),
        "rtx_scene_set_clock": (I32, [P, U32]),
        "rtx_scene_get_clock": (I32, [P, P]),
        "rtx_scene_animate": (I32, [P, F, F, P]),
        "rtx_read_vertices": (I32, [P, P, P, U32]),
        "rtx_renderer_set_animation": (I32, [P, I32, F, F]),
        "rtx_renderer_set_pipelined": (I32, [P, I32]),
        "rtx_renderer_set_bounces": (I32, [P, I32]),
        "rtx_now_ns": (C.c_int64, []),
        "rtx_server_create": (I32, [P, P, C.POINTER(P)]),
        "rtx_server_destroy": (I32, [P]),
        "rtx_server_add_scene": (I32, [P, P, P]),
        "rtx_server_memory": (I32, [P, P, P]),
        "rtx_server_submit": (I32, [P, P, U32]),
        "rtx_server_poll": (I32, [P, P, U32, P]),
        "rtx_server_drain": (I32, [P, C.c_int64]),
        "rtx_batch_create": (I32, [P, P, P, P, I32, I32, C.POINTER(P)]),
        "rtx_batch_destroy": (I32, [P]),
        "rtx_batch_render": (I32, [P, P]),
        "rtx_batch_capture": (I32, [P, P]),
        "rtx_batch_replay": (I32, [P, P]),
        "rtx_batch_stats": (I32, [P, P, P]),
        "rtx_pipeline_create": (I32, [P, P, P, I32, C.POINTER(P)]),
        "rtx_pipeline_destroy": (I32, [P]),
        "rtx_pipeline_set_animation": (I32, [P, I32, F, F]),
        "rtx_pipeline_frame": (I32, [P, P, I32, P]),
        "rtx_pipeline_wait": (I32, [P, U64]),
        "rtx_pipeline_fb": (P, [P, U64]),
        "rtx_pipeline_stats": (I32, [P, U64, P]),
        "rtx_lbvh_build": (I32, [P, P, P]),
        "rtx_lbvh_build_noalloc": (I32, [P, P, P]),
        "rtx_lbvh_refit": (I32, [P, P]),
        "rtx_lbvh_info": (I32, [P, P, P]),
        "rtx_read_codes": (I32, [P, P, U32]),
        "rtx_read_perm": (I32, [P, P, U32]),
        "rtx_read_indices": (I32, [P, P, U32]),
        "rtx_read_matmap": (I32, [P, P, U32]),
        "rtx_read_orig_ids": (I32, [P, P, U32]),
        "rtx_read_global_bbox": (I32, [P, P, P]),
        "rtx_read_nodes_ref": (I32, [P, P, U32, P]),
        "rtx_rays_create": (I32, [P, U64, C.POINTER(P)]),
        "rtx_rays_destroy": (I32, [P]),
        "rtx_rays_capacity": (U64, [P]),
        "rtx_rays_count": (I32, [P, C.POINTER(U64)]),
        "rtx_rays_read": (I32, [P, P, P, P, P, U64]),
        "rtx_rays_write": (I32, [P, P, P, P, P, U64]),
        "rtx_hits_create": (I32, [P, U64, C.POINTER(P)]),
        "rtx_hits_destroy": (I32, [P]),
        "rtx_hits_read": (I32, [P, P, P, P, P, P, P, P, U64]),
        "rtx_hits_write": (I32, [P, P, P, P, P, P, U64]),
        "rtx_raygen_primary": (I32, [P, P, I32, I32, I32, I32, P, U64, U64, P]),
        "rtx_trace": (I32, [P, P, P, U64, P]),
        "rtx_trace_shadow": (I32, [P, P, P, U64, P]),
        "rtx_bvh_create": (I32, [P, C.POINTER(P)]),
        "rtx_bvh_destroy": (I32, [P]),
        "rtx_bvh_copy": (I32, [P, P, P]),
        "rtx_trace_with": (I32, [P, P, P, P, U64, P]),
        "rtx_trace_shadow_with": (I32, [P, P, P, P, U64, P]),
        "rtx_raygen_secondary": (I32, [P, P, P, P, U64, C.POINTER(U64), P]),
        "rtx_fb_create": (I32, [P, I32, I32, C.POINTER(P)]),
        "rtx_fb_destroy": (I32, [P]),
        "rtx_fb_clear": (I32, [P, P]),
        "rtx_fb_resolve": (I32, [P, P]),
        "rtx_fb_read_accum": (I32, [P, P]),
        "rtx_fb_read_rgba8": (I32, [P, P]),
        "rtx_shade": (I32, [P, P, P, P, U64, I32, P]),
        "rtx_renderer_create": (I32, [P, P, C.POINTER(P)]),
        "rtx_renderer_destroy": (I32, [P]),
        "rtx_renderer_fb": (P, [P]),
        "rtx_render_frame": (I32, [P, P, P, I32, P, P]),
        "rtx_render_frame_async": (I32, [P, P, P, I32, P]),
        "rtx_renderer_last_stats": (I32, [P, P]),
        "rtx_renderer_capture": (I32, [P, P, P, I32, P]),
        "rtx_renderer_replay": (I32, [P, P]),
        "rtx_camera_set": (I32, [P, P, P, P, F, F, I32]),
        "rtx_scene_set_textures": (I32, [P, P, P, I32]),
        "rtx_scene_set_cubemap": (I32, [P, P, P]),
        "rtx_sah_build": (I32, [P, P]),
        "rtx_sah_build_host": (I32, [P, U32, P, U32, P, U32, P, P]),
        "rtx_obj_load": (I32, [C.c_char_p, I32, C.POINTER(P)]),
        "rtx_obj_destroy": (I32, [P]),
        "rtx_obj_info": (I32, [P, P, P, P]),
        "rtx_obj_read": (I32, [P, P, P, P, P]),
        "rtx_obj_group_info": (I32, [P, U32, P]),
        "rtx_obj_read_group": (I32, [P, U32, P, P, P, P, P]),
        "rtx_instances_create": (I32, [P, P, P, P, I32, C.POINTER(P)]),
        "rtx_instances_destroy": (I32, [P]),
        "rtx_instances_set_transforms": (I32, [P, P, P]),
        "rtx_instances_set_lights": (I32, [P, P]),
        "rtx_instances_tri_base": (I32, [P, P]),
        "rtx_trace_instances": (I32, [P, P, P, U64, P]),
        "rtx_trace_shadow_instances": (I32, [P, P, P, U64, P]),
        "rtx_instances_set_materials": (I32, [P, P]),
        "rtx_instances_shading_scene": (I32, [P, P]),
        "rtx_render_instances": (I32, [P, P, P, P, P]),
        "rtx_wavefront_create": (I32, [P, P, C.POINTER(P)]),
        "rtx_wavefront_destroy": (I32, [P]),
        "rtx_wavefront_frame_bytes": (U64, [I32, I32, I32, U64]),
        "rtx_wavefront_add": (I32, [P, P, P, I32, I32, I32, U64, U64, C.POINTER(I32)]),
        "rtx_wavefront_step": (I32, [P, P, C.POINTER(I32)]),
        "rtx_wavefront_run": (I32, [P, P]),
        "rtx_wavefront_sync": (I32, [P]),
        "rtx_wavefront_poll": (I32, [P, P, P, C.c_uint32, C.POINTER(C.c_uint32)]),
        "rtx_wavefront_frame_stats": (I32, [P, I32, P]),
        "rtx_wavefront_read_accum": (I32, [P, I32, P]),
        "rtx_wavefront_read_rgba8": (I32, [P, I32, P]),
        "rtx_wavefront_release": (I32, [P, I32]),
        "rtx_wavefront_memory": (I32, [P, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().rtx_last_error().decode(errors="replace")
        raise RtxError(f"{what} failed ({rc}): {msg}")


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to librtx must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def camera_set(pos, look_at=None, up=(0.0, 1.0, 0.0), fov=np.pi / 4, aspect=1.0, right_handed=False,
               direction=None) -> Camera:
    """Camera::set (src/rt/camera.cpp:31-59) evaluated by the library (host code)."""
    cam = Camera()
    p = np.asarray(pos, np.float32)
    d = np.asarray(direction if direction is not None else np.asarray(look_at, np.float32) - p, np.float32)
    u = np.asarray(up, np.float32)
    _check(lib().rtx_camera_set(C.byref(cam), _ptr(p), _ptr(d), _ptr(u), float(np.float32(fov)),
                                float(np.float32(aspect)), int(right_handed)), "rtx_camera_set")
    return cam


def make_lights(ambient, lights) -> Lights:
    """lights: iterable of dicts {type, pos, dir, color, angle, radius}."""
    L = Lights()
    L.ambient[:] = [float(x) for x in ambient]
    lights = list(lights)
    assert len(lights) <= MAX_LIGHTS
    L.count = len(lights)
    for i, d in enumerate(lights):
        L.l[i].type = int(d.get("type", LIGHT_POINT))
        L.l[i].pos[:] = [float(x) for x in d.get("pos", (0, 0, 0))]
        L.l[i].dir[:] = [float(x) for x in d.get("dir", (0, -1, 0))]
        L.l[i].color[:] = [float(x) for x in d.get("color", (1, 1, 1))]
        L.l[i].angle = float(d.get("angle", 0.0))
        L.l[i].radius = float(d.get("radius", 0.0))
    return L


def make_materials(rows) -> np.ndarray:
    """rows: iterable of (r, g, b, shininess, reflectiveness, ior). Returns the 32-B material array."""
    rows = list(rows)
    arr = (Material * len(rows))()
    for i, (r, g, b, sh, rf, ior) in enumerate(rows):
        arr[i].diffuse[:] = [r, g, b, 0.0]
        arr[i].shininess, arr[i].reflectiveness, arr[i].refractive_index = sh, rf, ior
        arr[i].texture = -1
    return np.frombuffer(bytes(arr), dtype=np.uint8).copy()


def _set_textures(h, images):
    imgs = [np.ascontiguousarray(im, np.uint8) for im in images]
    dims = np.asarray([d for im in imgs for d in (im.shape[1], im.shape[0])] or [0], np.int32)
    data = np.concatenate([im.reshape(-1) for im in imgs]) if imgs else np.zeros(4, np.uint8)
    _check(lib().rtx_scene_set_textures(h, _ptr(data), _ptr(dims), len(imgs)), "rtx_scene_set_textures")


def _set_cubemap(h, faces):
    if faces is None:
        _check(lib().rtx_scene_set_cubemap(h, None, None), "rtx_scene_set_cubemap")
        return
    imgs = [np.ascontiguousarray(f, np.uint8) for f in faces]
    assert len(imgs) == 6
    dims = np.asarray([d for im in imgs for d in (im.shape[1], im.shape[0])], np.int32)
    data = np.concatenate([im.reshape(-1) for im in imgs])
    _check(lib().rtx_scene_set_cubemap(h, _ptr(data), _ptr(dims)), "rtx_scene_set_cubemap")


class Context:
    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        _check(lib().rtx_init(device, C.byref(self.h)), "rtx_init")

    def close(self):
        if self.h:
            lib().rtx_shutdown(self.h)
            self.h = C.c_void_p()

    def sync(self):
        _check(lib().rtx_sync(self.h, None), "rtx_sync")

    @property
    def stream(self):
        return lib().rtx_stream(self.h)

    def device_sync(self):
        """Waits for every stream of the device (not only the context's)."""
        _hip_check(_hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


    def profile(self, on: bool = True):
        _check(lib().rtx_profile_enable(self.h, int(on)), "rtx_profile_enable")

    def profile_reset(self):
        _check(lib().rtx_profile_reset(self.h), "rtx_profile_reset")

    def profile_report(self) -> dict:
        n = lib().rtx_profile_report(self.h, None, 0)
        buf = C.create_string_buffer(int(n) + 16)
        lib().rtx_profile_report(self.h, buf, len(buf))
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms = line.split()
            out[name] = (int(cnt), float(ms))
        return out

    def mem_in_use(self) -> int:
        return int(lib().rtx_mem_in_use(self.h))

    def tune(self, key: str, value: int):
        _check(lib().rtx_tune_set(self.h, key.encode(), int(value)), "rtx_tune_set")

    def tuned(self, key: str) -> int:
        return int(lib().rtx_tune_get(self.h, key.encode()))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()



def _hip() -> C.CDLL:
    """The HIP runtime librtx is linked against (the instance it already
    loaded): streams given to the entry points must come from it."""
    lib()
    for name in ("libamdhip64.so.7", "libamdhip64.so"):
        try:
            return C.CDLL(name)
        except OSError:
            continue
    raise RtxError("HIP runtime library not found")


def _hip_check(rc: int, what: str) -> None:
    if rc != 0:
        raise RtxError(f"{what} failed (hipError_t {rc})")


class Stream:
    """A non-blocking HIP stream at the default priority, to pass as
    `stream=` to the entry points (e.g. a build beside a rendering frame)."""

    def __init__(self):
        self.h = C.c_void_p()
        _hip_check(_hip().hipStreamCreateWithFlags(C.byref(self.h), 1), "hipStreamCreateWithFlags")

    @property
    def handle(self) -> int:
        return self.h.value

    def synchronize(self):
        _hip_check(_hip().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def destroy(self):
        if self.h:
            _hip().hipStreamSynchronize(self.h)
            _hip().hipStreamDestroy(self.h)
            self.h = C.c_void_p()


class Scene:
    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.h = C.c_void_p()
        _check(lib().rtx_scene_create(ctx.h, C.byref(self.h)), "rtx_scene_create")
        self.T = 0
        self.V = 0

    def destroy(self):
        if self.h:
            lib().rtx_scene_destroy(self.h)
            self.h = C.c_void_p()

    def set_geometry(self, pos, idx, nrm=None, uv=None, matmap=None):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
        nrm = None if nrm is None else np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
        uv = None if uv is None else np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        _check(lib().rtx_scene_set_vertices(self.h, _ptr(pos), _ptr(nrm), _ptr(uv), pos.shape[0]),
               "rtx_scene_set_vertices")
        mm = None if matmap is None else np.ascontiguousarray(matmap, np.int32)
        _check(lib().rtx_scene_set_triangles(self.h, _ptr(idx), _ptr(mm), idx.shape[0]), "rtx_scene_set_triangles")
        self.V, self.T = pos.shape[0], idx.shape[0]

    def set_vertices_ref(self, vertex80: np.ndarray):
        v = np.ascontiguousarray(vertex80, np.float32).reshape(-1, 20)
        _check(lib().rtx_scene_set_vertices_ref(self.h, _ptr(v), v.shape[0]), "rtx_scene_set_vertices_ref")
        self.V = v.shape[0]

    def set_triangles(self, idx, matmap=None):
        idx = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
        mm = None if matmap is None else np.ascontiguousarray(matmap, np.int32)
        _check(lib().rtx_scene_set_triangles(self.h, _ptr(idx), _ptr(mm), idx.shape[0]), "rtx_scene_set_triangles")
        self.T = idx.shape[0]

    def update_positions(self, pos, stream=None):
        self._staged = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        _check(lib().rtx_scene_update_positions(self.h, _ptr(self._staged), stream), "rtx_scene_update_positions")

    def displace_waves(self, waves: np.ndarray, t: float, stream=None):
        w = np.ascontiguousarray(waves, np.float32).reshape(-1, 6)
        _check(lib().rtx_scene_displace_waves(self.h, _ptr(w), w.shape[0], float(t), stream),
               "rtx_scene_displace_waves")

    def set_waves(self, waves: np.ndarray):
        w = np.ascontiguousarray(waves, np.float32).reshape(-1, 6)
        _check(lib().rtx_scene_set_waves(self.h, _ptr(w), w.shape[0]), "rtx_scene_set_waves")

    def set_clock(self, frame: int):
        _check(lib().rtx_scene_set_clock(self.h, int(frame)), "rtx_scene_set_clock")

    def clock(self) -> int:
        f = C.c_uint32(0)
        _check(lib().rtx_scene_get_clock(self.h, C.byref(f)), "rtx_scene_get_clock")
        return int(f.value)

    def animate(self, t0: float, dt: float, stream=None):
        _check(lib().rtx_scene_animate(self.h, float(t0), float(dt), stream), "rtx_scene_animate")

    def vertices(self):
        """(pos V x 3, nrm V x 3) float32 read back from the device."""
        V = self.V
        pos = np.zeros((V, 3), np.float32)
        nrm = np.zeros((V, 3), np.float32)
        _check(lib().rtx_read_vertices(self.h, _ptr(pos), _ptr(nrm), V), "rtx_read_vertices")
        return pos, nrm

    def set_materials(self, mats: np.ndarray):
        mats = np.ascontiguousarray(mats, np.uint8)
        _check(lib().rtx_scene_set_materials(self.h, _ptr(mats), mats.size // 32), "rtx_scene_set_materials")

    def set_lights(self, lights: Lights):
        _check(lib().rtx_scene_set_lights(self.h, C.byref(lights)), "rtx_scene_set_lights")

    def set_textures(self, images):
        """Material textures: h x w x 4 uint8 arrays, texel (x, y) at [y, x]."""
        _set_textures(self.h, images)

    def set_cubemap(self, faces):
        """Six h x w x 4 uint8 faces (+x, -x, +y, -y, +z, -z), or None to disable."""
        _set_cubemap(self.h, faces)

    def build(self, depth: int = 30, min_leaf: int = 1, stream=None):
        p = LbvhParams(depth, min_leaf)
        _check(lib().rtx_lbvh_build(self.h, C.byref(p), stream), "rtx_lbvh_build")

    def build_noalloc(self, depth: int = 30, min_leaf: int = 1, stream=None):
        """The same build with no allocation: scratch in the library's fixed
        array, outputs in tree buffers reserved by an earlier build."""
        p = LbvhParams(depth, min_leaf)
        _check(lib().rtx_lbvh_build_noalloc(self.h, C.byref(p), stream), "rtx_lbvh_build_noalloc")

    def build_sah(self, stream=None):
        """The CPU SAH BVH (BVH::construct_and_map, src/rt/bvh.cpp:126-175)."""
        _check(lib().rtx_sah_build(self.h, stream), "rtx_sah_build")

    def refit(self, stream=None):
        _check(lib().rtx_lbvh_refit(self.h, stream), "rtx_lbvh_refit")

    def info(self):
        n = C.c_uint32()
        lc = np.zeros(64, np.uint32)
        _check(lib().rtx_lbvh_info(self.h, C.byref(n), _ptr(lc)), "rtx_lbvh_info")
        return int(n.value), lc

    def _read(self, fn, dtype, count):
        out = np.empty(count, dtype)
        _check(getattr(lib(), fn)(self.h, _ptr(out), self.T), fn)
        return out

    def codes(self):
        return self._read("rtx_read_codes", np.uint32, self.T)

    def perm(self):
        return self._read("rtx_read_perm", np.uint32, self.T)

    def indices(self):
        return self._read("rtx_read_indices", np.uint32, 3 * self.T).reshape(-1, 3)

    def matmap(self):
        return self._read("rtx_read_matmap", np.int32, self.T)

    def orig_ids(self):
        return self._read("rtx_read_orig_ids", np.uint32, self.T)

    def global_bbox(self):
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        _check(lib().rtx_read_global_bbox(self.h, _ptr(lo), _ptr(hi)), "rtx_read_global_bbox")
        return lo, hi

    def nodes(self) -> np.ndarray:
        n = C.c_uint32()
        _check(lib().rtx_read_nodes_ref(self.h, None, 0, C.byref(n)), "rtx_read_nodes_ref")
        out = np.zeros(n.value, NODE_DTYPE)
        _check(lib().rtx_read_nodes_ref(self.h, _ptr(out), n.value, C.byref(n)), "rtx_read_nodes_ref")
        return out


class Rays:
    def __init__(self, ctx: Context, capacity: int):
        self.ctx = ctx
        self.h = C.c_void_p()
        _check(lib().rtx_rays_create(ctx.h, int(capacity), C.byref(self.h)), "rtx_rays_create")
        self.cap = int(capacity)

    def destroy(self):
        if self.h:
            lib().rtx_rays_destroy(self.h)
            self.h = C.c_void_p()

    def count(self) -> int:
        c = C.c_uint64()
        _check(lib().rtx_rays_count(self.h, C.byref(c)), "rtx_rays_count")
        return int(c.value)

    def read(self, n: int):
        o = np.empty((n, 4), np.float32)
        d = np.empty((n, 4), np.float32)
        p = np.empty(n, np.int32)
        c = np.empty(n, np.float32)
        _check(lib().rtx_rays_read(self.h, _ptr(o), _ptr(d), _ptr(p), _ptr(c), n), "rtx_rays_read")
        return o, d, p, c

    def write(self, ori4, dir4, pixel=None, contrib=None):
        o = np.ascontiguousarray(ori4, np.float32).reshape(-1, 4)
        d = np.ascontiguousarray(dir4, np.float32).reshape(-1, 4)
        p = None if pixel is None else np.ascontiguousarray(pixel, np.int32)
        c = None if contrib is None else np.ascontiguousarray(contrib, np.float32)
        _check(lib().rtx_rays_write(self.h, _ptr(o), _ptr(d), _ptr(p), _ptr(c), o.shape[0]), "rtx_rays_write")

    def primary(self, cam: Camera, w: int, h: int, quad: int = 32, spp: int = 1, psamples=None, offset: int = 0,
                count: int | None = None, stream=None):
        ps = None if psamples is None else np.ascontiguousarray(psamples, np.float32)
        count = w * h * spp - offset if count is None else count
        _check(lib().rtx_raygen_primary(self.h, C.byref(cam), w, h, quad, spp, _ptr(ps), offset, count, stream),
               "rtx_raygen_primary")


class Hits:
    def __init__(self, ctx: Context, capacity: int):
        self.h = C.c_void_p()
        _check(lib().rtx_hits_create(ctx.h, int(capacity), C.byref(self.h)), "rtx_hits_create")
        self.cap = int(capacity)

    def destroy(self):
        if self.h:
            lib().rtx_hits_destroy(self.h)
            self.h = C.c_void_p()

    def read(self, n: int) -> dict:
        out = dict(flags=np.empty(n, np.uint32), t=np.empty(n, np.float32), id=np.empty(n, np.int32),
                   bary=np.empty((n, 2), np.float32), n4=np.empty((n, 4), np.float32),
                   p4=np.empty((n, 4), np.float32), uv2=np.empty((n, 2), np.float32))
        _check(lib().rtx_hits_read(self.h, *(_ptr(out[k]) for k in ("flags", "t", "id", "bary", "n4", "p4", "uv2")),
                                   n), "rtx_hits_read")
        return out

    def write(self, flags, t, id, n4, p4):
        arrs = [np.ascontiguousarray(flags, np.uint32), np.ascontiguousarray(t, np.float32),
                np.ascontiguousarray(id, np.int32), np.ascontiguousarray(n4, np.float32).reshape(-1, 4),
                np.ascontiguousarray(p4, np.float32).reshape(-1, 4)]
        _check(lib().rtx_hits_write(self.h, *(_ptr(a) for a in arrs), arrs[0].shape[0]), "rtx_hits_write")


class FrameBuffer:
    def __init__(self, ctx: Context, w: int, h: int, handle=None):
        self.w, self.h_ = w, h
        self.owned = handle is None
        self.h = C.c_void_p(handle) if handle is not None else C.c_void_p()
        if self.owned:
            _check(lib().rtx_fb_create(ctx.h, w, h, C.byref(self.h)), "rtx_fb_create")

    def destroy(self):
        if self.h and self.owned:
            lib().rtx_fb_destroy(self.h)
        self.h = C.c_void_p()

    def clear(self, stream=None):
        _check(lib().rtx_fb_clear(self.h, stream), "rtx_fb_clear")

    def resolve(self, stream=None):
        _check(lib().rtx_fb_resolve(self.h, stream), "rtx_fb_resolve")

    def accum(self) -> np.ndarray:
        out = np.empty((self.w * self.h_, 3), np.uint32)
        _check(lib().rtx_fb_read_accum(self.h, _ptr(out)), "rtx_fb_read_accum")
        return out

    def rgba8(self) -> np.ndarray:
        out = np.empty((self.w * self.h_, 4), np.uint8)
        _check(lib().rtx_fb_read_rgba8(self.h, _ptr(out)), "rtx_fb_read_rgba8")
        return out


class BVHCopy:
    """A copy of a built scene's accelerator (rtx_bvh): the reference's
    bvh_mem that Tracer::trace(Scene&, DeviceMemory& bvh_mem, ...) traces."""

    def __init__(self, ctx: "Context"):
        h = C.c_void_p()
        _check(lib().rtx_bvh_create(ctx.h, C.byref(h)), "rtx_bvh_create")
        self.h = h

    def copy(self, scene: "Scene", stream=None):
        _check(lib().rtx_bvh_copy(self.h, scene.h, stream), "rtx_bvh_copy")
        return self

    def destroy(self):
        if self.h:
            lib().rtx_bvh_destroy(self.h)
            self.h = None


def trace(scene: Scene, rays: Rays, hits: Hits, count=DEVICE_COUNT, stream=None, bvh: BVHCopy = None):
    """Tracer::trace; with bvh, the overload that traces an accelerator copy."""
    if bvh is not None:
        _check(lib().rtx_trace_with(scene.h, bvh.h, rays.h, hits.h, count, stream), "rtx_trace_with")
        return
    _check(lib().rtx_trace(scene.h, rays.h, hits.h, count, stream), "rtx_trace")


def trace_shadow(scene: Scene, rays: Rays, hits: Hits, count=DEVICE_COUNT, stream=None, bvh: BVHCopy = None):
    """Tracer::shadow_trace; with bvh, the overload that traces an accelerator copy."""
    if bvh is not None:
        _check(lib().rtx_trace_shadow_with(scene.h, bvh.h, rays.h, hits.h, count, stream), "rtx_trace_shadow_with")
        return
    _check(lib().rtx_trace_shadow(scene.h, rays.h, hits.h, count, stream), "rtx_trace_shadow")


def sah_build_host(pos, idx):
    """The library's CPU SAH builder on host arrays: (nodes as NodeRef records
    in the reference layout, triangle order)."""
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    idx = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
    T = idx.shape[0]
    nodes = (NodeRef * max(1, 2 * T))()
    count = C.c_uint32()
    order = np.zeros(T, np.uint32)
    _check(lib().rtx_sah_build_host(_ptr(pos), pos.shape[0], _ptr(idx), T, nodes, len(nodes), C.byref(count),
                                    _ptr(order)), "rtx_sah_build_host")
    raw = np.frombuffer(bytes(nodes), dtype=NODE_DTYPE, count=count.value).copy()
    return raw, order


class ObjModel:
    """A Wavefront OBJ model read by the library's loader (ModelOBJ::import,
    src/rt/obj-loader.cpp:163-227): `pos`, `nrm` (V x 3), `uv` (V x 2),
    `idx` (T x 3) as get_aggregate_mesh gives them, and `groups` (one dict per
    material run, the loader's order)."""

    def __init__(self, path, rebuild_normals: bool = False):
        h = C.c_void_p()
        _check(lib().rtx_obj_load(str(path).encode(), int(rebuild_normals), C.byref(h)), "rtx_obj_load")
        try:
            V, T, G = C.c_uint32(), C.c_uint32(), C.c_uint32()
            _check(lib().rtx_obj_info(h, C.byref(V), C.byref(T), C.byref(G)), "rtx_obj_info")
            self.pos = np.zeros((V.value, 3), np.float32)
            self.nrm = np.zeros((V.value, 3), np.float32)
            self.uv = np.zeros((V.value, 2), np.float32)
            self.idx = np.zeros((T.value, 3), np.uint32)
            _check(lib().rtx_obj_read(h, _ptr(self.pos), _ptr(self.nrm), _ptr(self.uv), _ptr(self.idx)),
                   "rtx_obj_read")
            self.groups = []
            for g in range(G.value):
                gi = ObjGroup()
                _check(lib().rtx_obj_group_info(h, g, C.byref(gi)), "rtx_obj_group_info")
                nv = C.c_uint32()
                _check(lib().rtx_obj_read_group(h, g, C.byref(nv), None, None, None, None), "rtx_obj_read_group")
                gp = np.zeros((nv.value, 3), np.float32)
                gn = np.zeros((nv.value, 3), np.float32)
                gu = np.zeros((nv.value, 2), np.float32)
                gx = np.zeros((gi.triangle_count, 3), np.uint32)
                _check(lib().rtx_obj_read_group(h, g, C.byref(nv), _ptr(gp), _ptr(gn), _ptr(gu), _ptr(gx)),
                       "rtx_obj_read_group")
                self.groups.append(dict(first=gi.first_triangle, count=gi.triangle_count,
                                        material=gi.material, ambient=list(gi.ambient), diffuse=list(gi.diffuse),
                                        specular=list(gi.specular), shininess=gi.shininess, alpha=gi.alpha,
                                        name=gi.name.decode(), color_map=gi.color_map.decode(),
                                        bump_map=gi.bump_map.decode(), pos=gp, nrm=gn, uv=gu, idx=gx))
        finally:
            lib().rtx_obj_destroy(h)


class Instances:
    """Objects of a scene over per-mesh BVHs (Scene::create_bvhs /
    update_bvh_roots, src/rt/scene.cpp:596-700): instance i is built scene
    scenes[i] under the 4x4 object-to-world transform tr[i] (row-major) with
    inverse tr_inv[i].  Triangle ids are shared: the distinct scenes in order
    of first use take consecutive blocks."""

    def __init__(self, ctx: Context, scenes, tr, tr_inv):
        self.ctx = ctx
        self.scenes = list(scenes)  # keeps the scenes alive with the set
        n = len(self.scenes)
        arr = (C.c_void_p * n)(*[sc.h.value for sc in self.scenes])
        tr = np.ascontiguousarray(tr, np.float32).reshape(n, 16)
        tri = np.ascontiguousarray(tr_inv, np.float32).reshape(n, 16)
        self.h = C.c_void_p()
        _check(lib().rtx_instances_create(ctx.h, arr, _ptr(tr), _ptr(tri), n, C.byref(self.h)),
               "rtx_instances_create")
        self.n = n

    def destroy(self):
        if self.h:
            _check(lib().rtx_instances_destroy(self.h), "rtx_instances_destroy")
            self.h = C.c_void_p()

    def set_transforms(self, tr, tr_inv):
        tr = np.ascontiguousarray(tr, np.float32).reshape(self.n, 16)
        tri = np.ascontiguousarray(tr_inv, np.float32).reshape(self.n, 16)
        _check(lib().rtx_instances_set_transforms(self.h, _ptr(tr), _ptr(tri)), "rtx_instances_set_transforms")

    def set_lights(self, lights: Lights):
        _check(lib().rtx_instances_set_lights(self.h, C.byref(lights)), "rtx_instances_set_lights")

    def set_materials(self, mats: np.ndarray):
        """One 32-B material per instance (make_materials rows)."""
        m = np.ascontiguousarray(mats, np.uint8)
        assert m.size == 32 * self.n
        _check(lib().rtx_instances_set_materials(self.h, _ptr(m)), "rtx_instances_set_materials")

    def _shading(self):
        sh = C.c_void_p()
        _check(lib().rtx_instances_shading_scene(self.h, C.byref(sh)), "rtx_instances_shading_scene")
        return sh

    def set_textures(self, images):
        """Material textures of frames over the instances (after set_materials)."""
        _set_textures(self._shading(), images)

    def set_cubemap(self, faces):
        """The background of frames over the instances, or None (after set_materials)."""
        _set_cubemap(self._shading(), faces)

    def tri_base(self) -> np.ndarray:
        out = np.zeros(self.n, np.uint32)
        _check(lib().rtx_instances_tri_base(self.h, _ptr(out)), "rtx_instances_tri_base")
        return out

    def trace(self, rays: Rays, hits: Hits, count=DEVICE_COUNT, stream=None):
        _check(lib().rtx_trace_instances(self.h, rays.h, hits.h, count, stream), "rtx_trace_instances")

    def trace_shadow(self, rays: Rays, hits: Hits, count=DEVICE_COUNT, stream=None):
        _check(lib().rtx_trace_shadow_instances(self.h, rays.h, hits.h, count, stream),
               "rtx_trace_shadow_instances")


def raygen_secondary(scene: Scene, rays_in: Rays, hits: Hits, rays_out: Rays, count_in=DEVICE_COUNT,
                     read_count: bool = True, stream=None):
    c = C.c_uint64()
    _check(lib().rtx_raygen_secondary(scene.h, rays_in.h, hits.h, rays_out.h, count_in,
                                      C.byref(c) if read_count else None, stream), "rtx_raygen_secondary")
    return int(c.value) if read_count else None


def shade(scene: Scene, rays: Rays, hits: Hits, fb: FrameBuffer, count=DEVICE_COUNT, primary=True, stream=None):
    _check(lib().rtx_shade(scene.h, rays.h, hits.h, fb.h, count, int(primary), stream), "rtx_shade")


class Renderer:
    def __init__(self, ctx: Context, width: int, height: int, max_bounces: int, depth: int = 30, min_leaf: int = 1,
                 quad: int = 32, ray_capacity: int = 0):
        self.ctx = ctx
        self.p = RenderParams(width, height, quad, max_bounces, depth, min_leaf, ray_capacity)
        self.h = C.c_void_p()
        _check(lib().rtx_renderer_create(ctx.h, C.byref(self.p), C.byref(self.h)), "rtx_renderer_create")
        self.fb = FrameBuffer(ctx, width, height, handle=lib().rtx_renderer_fb(self.h))

    def destroy(self):
        if self.h:
            lib().rtx_renderer_destroy(self.h)
            self.h = C.c_void_p()

    def frame(self, scene: Scene, cam: Camera, rebuild: bool = True, stream=None) -> FrameStats:
        st = FrameStats()
        _check(lib().rtx_render_frame(self.h, scene.h, C.byref(cam), int(rebuild), C.byref(st), stream),
               "rtx_render_frame")
        return st

    def frame_async(self, scene: Scene, cam: Camera, rebuild: bool = True, stream=None):
        _check(lib().rtx_render_frame_async(self.h, scene.h, C.byref(cam), int(rebuild), stream),
               "rtx_render_frame_async")

    def capture(self, scene: Scene, cam: Camera, rebuild: bool = True, stream=None):
        """Capture one read-back-free frame into a hipGraph (render once eagerly first)."""
        _check(lib().rtx_renderer_capture(self.h, scene.h, C.byref(cam), int(rebuild), stream),
               "rtx_renderer_capture")

    def set_bounces(self, bounces: int):
        _check(lib().rtx_renderer_set_bounces(self.h, int(bounces)), "rtx_renderer_set_bounces")

    def frame_instances(self, inst: "Instances", cam: Camera, stream=None) -> FrameStats:
        """A frame over a scene of objects (multi-root trace and shadow rays)."""
        st = FrameStats()
        _check(lib().rtx_render_instances(self.h, inst.h, C.byref(cam), C.byref(st), stream), "rtx_render_instances")
        return st

    def set_pipelined(self, enable: bool):
        _check(lib().rtx_renderer_set_pipelined(self.h, int(bool(enable))), "rtx_renderer_set_pipelined")

    def set_animation(self, enable: bool, t0: float = 0.0, dt: float = 0.0):
        _check(lib().rtx_renderer_set_animation(self.h, int(bool(enable)), float(t0), float(dt)),
               "rtx_renderer_set_animation")

    def replay(self, stream=None):
        _check(lib().rtx_renderer_replay(self.h, stream), "rtx_renderer_replay")

    def last_stats(self) -> FrameStats:
        st = FrameStats()
        _check(lib().rtx_renderer_last_stats(self.h, C.byref(st)), "rtx_renderer_last_stats")
        return st


def now_ns() -> int:
    """The server's monotonic clock (ns)."""
    return int(lib().rtx_now_ns())


class Wavefront:
    """Frames of many scenes rendered together (rtx_wavefront_*): each step
    advances every admitted frame by one bounce, one launch per stage over
    all of them.  Frame memory comes from a device arena of `arena_bytes`."""

    def __init__(self, ctx: Context, arena_bytes: int = 0, max_frames: int = 64, quad_size: int = 0):
        p = WavefrontParams(int(arena_bytes), int(max_frames), int(quad_size))
        self.h = C.c_void_p()
        _check(lib().rtx_wavefront_create(ctx.h, C.byref(p), C.byref(self.h)), "rtx_wavefront_create")
        self.sizes = {}

    @staticmethod
    def frame_bytes(width: int, height: int, bounces: int, ray_capacity: int = 0) -> int:
        return int(lib().rtx_wavefront_frame_bytes(int(width), int(height), int(bounces), int(ray_capacity)))

    def add(self, scene: Scene, cam: Camera, width: int, height: int, bounces: int, tag: int = 0,
            ray_capacity: int = 0) -> int:
        """Frame id; raises RtxError (-3, RTX_ENOMEM) when the arena has no room now."""
        fid = C.c_int32(-1)
        _check(lib().rtx_wavefront_add(self.h, scene.h, C.byref(cam), int(width), int(height), int(bounces),
                                       int(ray_capacity), int(tag), C.byref(fid)), "rtx_wavefront_add")
        self.sizes[int(fid.value)] = (int(width), int(height))
        return int(fid.value)

    def try_add(self, scene: Scene, cam: Camera, width: int, height: int, bounces: int, tag: int = 0,
                ray_capacity: int = 0):
        """Frame id, or None when the arena (or the frame table) is full now."""
        fid = C.c_int32(-1)
        rc = lib().rtx_wavefront_add(self.h, scene.h, C.byref(cam), int(width), int(height), int(bounces),
                                     int(ray_capacity), int(tag), C.byref(fid))
        if rc == -3:
            return None
        _check(rc, "rtx_wavefront_add")
        self.sizes[int(fid.value)] = (int(width), int(height))
        return int(fid.value)

    def step(self, stream=None) -> int:
        pend = C.c_int32(0)
        _check(lib().rtx_wavefront_step(self.h, stream, C.byref(pend)), "rtx_wavefront_step")
        return int(pend.value)

    def run(self, stream=None):
        _check(lib().rtx_wavefront_run(self.h, stream), "rtx_wavefront_run")

    def sync(self):
        _check(lib().rtx_wavefront_sync(self.h), "rtx_wavefront_sync")

    def poll(self, cap: int = 256):
        fr = (C.c_int32 * cap)()
        tg = (C.c_uint64 * cap)()
        n = C.c_uint32(0)
        _check(lib().rtx_wavefront_poll(self.h, fr, tg, cap, C.byref(n)), "rtx_wavefront_poll")
        return [(int(fr[i]), int(tg[i])) for i in range(n.value)]

    def stats(self, fid: int) -> FrameStats:
        st = FrameStats()
        _check(lib().rtx_wavefront_frame_stats(self.h, int(fid), C.byref(st)), "rtx_wavefront_frame_stats")
        return st

    def accum(self, fid: int) -> np.ndarray:
        w, h = self.sizes[int(fid)]
        out = np.zeros((w * h, 3), np.uint32)
        _check(lib().rtx_wavefront_read_accum(self.h, int(fid), _ptr(out)), "rtx_wavefront_read_accum")
        return out

    def rgba8(self, fid: int) -> np.ndarray:
        w, h = self.sizes[int(fid)]
        out = np.zeros((w * h, 4), np.uint8)
        _check(lib().rtx_wavefront_read_rgba8(self.h, int(fid), _ptr(out)), "rtx_wavefront_read_rgba8")
        return out

    def release(self, fid: int):
        _check(lib().rtx_wavefront_release(self.h, int(fid)), "rtx_wavefront_release")
        self.sizes.pop(int(fid), None)

    def memory(self):
        u, c = C.c_uint64(0), C.c_uint64(0)
        _check(lib().rtx_wavefront_memory(self.h, C.byref(u), C.byref(c)), "rtx_wavefront_memory")
        return int(u.value), int(c.value)

    def destroy(self):
        if self.h:
            _check(lib().rtx_wavefront_destroy(self.h), "rtx_wavefront_destroy")
            self.h = C.c_void_p()


class Server:
    """Request server over resident scenes (rtx_server_*): a native
    dispatcher thread admits requests at their arrival time (first come
    first served, by arena bytes, at most `workers` in flight) and renders
    them with the wavefront renderer, every step advancing all requests in
    flight by one bounce; it records each request's device completion time.
    plain=True: the per-request baseline, one at a time with the plain frame
    on a renderer per resolution class."""

    def __init__(self, ctx: Context, classes=(), workers: int = 64, max_bounces: int = 3, hbm_budget: int = 0,
                 plain: bool = False, queue_cap: int = 0, arena_bytes: int = 0, groups: int = 0,
                 steps_ahead: int = 0):
        assert len(classes) <= SERVER_MAX_CLASSES
        p = ServerParams()
        p.workers, p.max_bounces, p.classes = int(workers), int(max_bounces), len(classes)
        for i, (w, h) in enumerate(classes):
            p.class_w[i], p.class_h[i] = int(w), int(h)
        p.hbm_budget, p.plain, p.queue_cap = int(hbm_budget), int(bool(plain)), int(queue_cap)
        p.arena_bytes, p.groups, p.steps_ahead = int(arena_bytes), int(groups), int(steps_ahead)
        self.h = C.c_void_p()
        _check(lib().rtx_server_create(ctx.h, C.byref(p), C.byref(self.h)), "rtx_server_create")

    def add_scene(self, scene: Scene) -> int:
        sid = C.c_int32(-1)
        _check(lib().rtx_server_add_scene(self.h, scene.h, C.byref(sid)), "rtx_server_add_scene")
        return int(sid.value)

    def memory(self):
        w, sc = C.c_uint64(0), C.c_uint64(0)
        _check(lib().rtx_server_memory(self.h, C.byref(w), C.byref(sc)), "rtx_server_memory")
        return int(w.value), int(sc.value)

    def submit(self, reqs):
        """reqs: iterable of dicts {scene, width, height, bounces, camera (Camera), arrival_ns, tag[, out]}
        where out is an optional C-contiguous uint32 array of w*h*3 for the accumulators."""
        reqs = list(reqs)
        arr = (Request * len(reqs))()
        self._keep = getattr(self, "_keep", [])
        for i, r in enumerate(reqs):
            a = arr[i]
            a.scene, a.width, a.height, a.bounces = int(r["scene"]), int(r["width"]), int(r["height"]), \
                int(r["bounces"])
            C.memmove(C.addressof(a.camera), C.addressof(r["camera"]), 48)
            a.arrival_ns, a.tag = int(r["arrival_ns"]), int(r["tag"])
            out = r.get("out")
            if out is not None:
                assert out.flags["C_CONTIGUOUS"] and out.dtype == np.uint32 and out.size == a.width * a.height * 3
                self._keep.append(out)
                a.out_accum = out.ctypes.data
        _check(lib().rtx_server_submit(self.h, arr, len(reqs)), "rtx_server_submit")

    def poll(self, cap: int = 4096):
        arr = (Completion * cap)()
        n = C.c_uint32(0)
        _check(lib().rtx_server_poll(self.h, arr, cap, C.byref(n)), "rtx_server_poll")
        return [dict(tag=c.tag, arrival_ns=c.arrival_ns, start_ns=c.start_ns, done_ns=c.done_ns, rays=c.rays,
                     shadow_rays=c.shadow_rays, status=c.status, worker=c.worker) for c in arr[: n.value]]

    def drain(self, timeout_ms: int = 60000):
        _check(lib().rtx_server_drain(self.h, int(timeout_ms)), "rtx_server_drain")

    def destroy(self):
        if self.h:
            _check(lib().rtx_server_destroy(self.h), "rtx_server_destroy")
            self.h = C.c_void_p()
            self._keep = []


class Batch:
    """Frames of several resident scenes rendered together (rtx_batch_*)."""

    def __init__(self, ctx: Context, renderers, scenes, cams, streams: int = 4):
        n = len(renderers)
        assert n == len(scenes) == len(cams) and n > 0
        rs = (C.c_void_p * n)(*[r.h for r in renderers])
        ss = (C.c_void_p * n)(*[x.h for x in scenes])
        cs = (Camera * n)(*cams)
        self.h = C.c_void_p()
        _check(lib().rtx_batch_create(ctx.h, rs, ss, cs, n, int(streams), C.byref(self.h)), "rtx_batch_create")

    def render(self, stream=None):
        _check(lib().rtx_batch_render(self.h, stream), "rtx_batch_render")

    def capture(self, stream=None):
        _check(lib().rtx_batch_capture(self.h, stream), "rtx_batch_capture")

    def replay(self, stream=None):
        _check(lib().rtx_batch_replay(self.h, stream), "rtx_batch_replay")

    def stats(self):
        r, sh = C.c_uint64(0), C.c_uint64(0)
        _check(lib().rtx_batch_stats(self.h, C.byref(r), C.byref(sh)), "rtx_batch_stats")
        return int(r.value), int(sh.value)

    def destroy(self):
        if self.h:
            _check(lib().rtx_batch_destroy(self.h), "rtx_batch_destroy")
            self.h = C.c_void_p()


class Pipeline:
    """Frames of one scene with `depth` frames in flight (rtx_pipeline_*)."""

    def __init__(self, ctx: Context, scene: Scene, width: int, height: int, max_bounces: int, depth: int = 2,
                 bvh_depth: int = 30, min_leaf: int = 1, quad: int = 32, ray_capacity: int = 0):
        self.ctx = ctx
        self.w, self.hgt = width, height
        p = RenderParams(width, height, quad, max_bounces, bvh_depth, min_leaf, ray_capacity)
        self.h = C.c_void_p()
        _check(lib().rtx_pipeline_create(ctx.h, scene.h, C.byref(p), int(depth), C.byref(self.h)),
               "rtx_pipeline_create")

    def set_animation(self, enable: bool, t0: float = 0.0, dt: float = 0.0):
        _check(lib().rtx_pipeline_set_animation(self.h, int(bool(enable)), float(t0), float(dt)),
               "rtx_pipeline_set_animation")

    def frame(self, cam: Camera, rebuild: bool = True) -> int:
        f = C.c_uint64(0)
        _check(lib().rtx_pipeline_frame(self.h, C.byref(cam), int(bool(rebuild)), C.byref(f)),
               "rtx_pipeline_frame")
        return int(f.value)

    def wait(self, frame: int):
        _check(lib().rtx_pipeline_wait(self.h, int(frame)), "rtx_pipeline_wait")

    def fb(self, frame: int) -> "FrameBuffer":
        h = lib().rtx_pipeline_fb(self.h, int(frame))
        if not h:
            raise RtxError("rtx_pipeline_fb: frame not available")
        return FrameBuffer(self.ctx, self.w, self.hgt, handle=h)

    def stats(self, frame: int) -> FrameStats:
        st = FrameStats()
        _check(lib().rtx_pipeline_stats(self.h, int(frame), C.byref(st)), "rtx_pipeline_stats")
        return st

    def destroy(self):
        if self.h:
            _check(lib().rtx_pipeline_destroy(self.h), "rtx_pipeline_destroy")
            self.h = C.c_void_p()
