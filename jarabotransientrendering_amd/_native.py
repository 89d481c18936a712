"""ctypes binding of the C ABI in include/tr_api.h (libtransient_hip.so).

The library is loaded from the package's _lib/ directory. There is no fallback:
if the library is missing or cannot find a gfx950 device, calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .build import LIB

TR_OK = 0
TR_ERR_INVALID = -1
TR_ERR_DEVICE = -2
TR_ERR_NOMEM = -3
TR_ERR_STATE = -4

TR_MATERIAL_TWO_SIDED = 1
TR_NOALLOC_MAX_RAYS = 65536  # rays per tr_intersect_batch_noalloc call
TR_RENDER_PROFILE = 1
TR_RENDER_PROFILE_EXTEND = 256
TR_RENDER_BINARY_BVH = 2
TR_RENDER_ORTHONORMAL_FRAME = 4
TR_RENDER_GRAPH = 8
TR_RENDER_TIME_SAMPLING = 16
TR_RENDER_STACK32 = 32
TR_RENDER_STACK_KEYED = 64
TR_RENDER_PER_JOB = 128
TR_RENDER_POSTPONE_LEAVES = 512
TR_RENDER_NO_CAMERA_PACKETS = 1024
K_CAMERA, K_EXTEND, K_SHADE, K_SHADOW, K_FILM_MERGE, K_FILM, K_GUIDE = 0, 1, 2, 3, 4, 5, 6
K_NAMES = {K_CAMERA: "camera", K_EXTEND: "extend", K_SHADE: "shade", K_SHADOW: "shadow",
           K_FILM_MERGE: "film_merge", K_FILM: "film", K_GUIDE: "guide"}
# guiding network (tr_guide_*): parameter count, raw input / output / record widths
GUIDE_PARAMS = 18216
GUIDE_RAW = 12
GUIDE_OUT = 40
GUIDE_RECORD = 16
TR_K_COUNT = 8


class TrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Material(ctypes.Structure):
    _fields_ = [("kd", ctypes.c_float * 3)]


class SceneInfo(ctypes.Structure):
    _fields_ = [("n_tris", ctypes.c_uint32), ("n_bvh_nodes", ctypes.c_uint32), ("bvh_depth", ctypes.c_uint32),
                ("n_lights", ctypes.c_uint32), ("device_bytes", ctypes.c_uint64), ("build_ms", ctypes.c_double)]


class Ray(ctypes.Structure):
    _fields_ = [("org", ctypes.c_float * 3), ("tnear", ctypes.c_float), ("dir", ctypes.c_float * 3),
                ("tfar", ctypes.c_float)]


class Hit(ctypes.Structure):
    _fields_ = [("t", ctypes.c_float), ("geom_id", ctypes.c_uint32), ("prim_id", ctypes.c_uint32),
                ("tri", ctypes.c_uint32)]


class FilmDesc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("time_bins", ctypes.c_int32),
                ("exposure", ctypes.c_float), ("offset", ctypes.c_float), ("scanline", ctypes.c_int32),
                ("camera_unwarp", ctypes.c_int32), ("bounce_layers", ctypes.c_int32)]


class FilmSample(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int32), ("y", ctypes.c_int32), ("time", ctypes.c_float), ("rgb", ctypes.c_float * 3),
                ("weight", ctypes.c_float), ("pad0", ctypes.c_uint32)]


# numpy record layout of tr_film_sample (32 bytes)
FILM_SAMPLE_DTYPE = np.dtype([("x", "<i4"), ("y", "<i4"), ("time", "<f4"), ("rgb", "<f4", (3,)),
                              ("weight", "<f4"), ("pad0", "<u4")])


class CameraDesc(ctypes.Structure):
    _fields_ = [("position", ctypes.c_float * 3), ("focus", ctypes.c_float * 3), ("up", ctypes.c_float * 3),
                ("fov_deg", ctypes.c_float), ("view_plane_dist", ctypes.c_float)]


class RenderDesc(ctypes.Structure):
    _fields_ = [("sqrt_spp", ctypes.c_int32), ("n_incoming", ctypes.c_int32), ("max_bounces", ctypes.c_int32),
                ("pad0", ctypes.c_int32), ("seed", ctypes.c_uint64), ("wave_paths", ctypes.c_uint32),
                ("flags", ctypes.c_uint32)]


class RenderStats(ctypes.Structure):
    _fields_ = [("paths", ctypes.c_uint64), ("extension_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
                ("contributions", ctypes.c_uint64), ("waves", ctypes.c_uint32), ("bounces_max", ctypes.c_uint32),
                ("render_ms", ctypes.c_double), ("kernel_ms", ctypes.c_double * TR_K_COUNT),
                ("kernel_launches", ctypes.c_uint32 * TR_K_COUNT)]

    def as_dict(self):
        d = {k: int(getattr(self, k)) for k in ("paths", "extension_rays", "shadow_rays", "contributions", "waves",
                                                 "bounces_max")}
        d["render_ms"] = float(self.render_ms)
        d["kernel_ms"] = {K_NAMES[k]: float(self.kernel_ms[k]) for k in K_NAMES}
        d["kernel_launches"] = {K_NAMES[k]: int(self.kernel_launches[k]) for k in K_NAMES}
        return d


class LaunchDesc(ctypes.Structure):
    _fields_ = [("wave_paths", ctypes.c_uint32), ("flags", ctypes.c_uint32), ("trace_per_cu", ctypes.c_uint32),
                ("shade_per_cu", ctypes.c_uint32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class GuideDesc(ctypes.Structure):
    _fields_ = [("learning_rate", ctypes.c_float), ("bsdf_fraction", ctypes.c_float),
                ("train_paths", ctypes.c_uint32), ("batch_size", ctypes.c_uint32),
                ("steps_per_wave", ctypes.c_uint32), ("pad0", ctypes.c_uint32), ("init_seed", ctypes.c_uint64)]


class GuideInfo(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_uint64), ("records", ctypes.c_uint64), ("last_loss", ctypes.c_double),
                ("n_params", ctypes.c_uint32), ("flop_per_eval", ctypes.c_uint32)]


# exported symbol -> (restype, argtypes); the header declares exactly these
SIGNATURES = {
    "tr_scene_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    "tr_scene_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "tr_scene_add_mesh": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                         ctypes.c_uint32, ctypes.POINTER(Material), ctypes.POINTER(ctypes.c_uint32)]),
    "tr_scene_add_mesh_ex": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                            ctypes.c_uint32, ctypes.c_void_p, ctypes.POINTER(Material),
                                            ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
    "tr_scene_add_obj": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(Material), ctypes.c_uint32,
                                        ctypes.POINTER(ctypes.c_uint32)]),
    "tr_obj_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(ctypes.c_float)),
                                   ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.POINTER(ctypes.c_uint32)),
                                   ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.POINTER(ctypes.c_float))]),
    "tr_obj_free": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "tr_scene_add_point_light": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float * 3, ctypes.c_float * 3]),
    "tr_scene_add_rect_light": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float * 3, ctypes.c_float * 3,
                                               ctypes.c_float * 2, ctypes.c_float]),
    "tr_scene_set_medium": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "tr_scene_commit": (ctypes.c_int, [ctypes.c_void_p]),
    "tr_scene_info_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SceneInfo)]),
    "tr_intersect_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
    "tr_intersect_batch_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                 ctypes.c_void_p, ctypes.c_void_p]),
    "tr_intersect_batch_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                  ctypes.c_void_p]),
    "tr_occluded_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
    "tr_occluded_batch_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                ctypes.c_void_p, ctypes.c_void_p]),
    "tr_film_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(FilmDesc)]),
    "tr_film_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "tr_film_clear": (ctypes.c_int, [ctypes.c_void_p]),
    "tr_film_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "tr_film_read_rows": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                         ctypes.c_void_p]),
    "tr_film_read_sums": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "tr_film_device_buffers": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                              ctypes.POINTER(ctypes.c_void_p)]),
    "tr_film_add_samples": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]),
    "tr_film_write": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p]),
    "tr_image_save": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "tr_render_transient": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CameraDesc), ctypes.c_void_p,
                                           ctypes.POINTER(RenderDesc), ctypes.POINTER(RenderStats)]),
    "tr_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    "tr_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "tr_batch_add": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(CameraDesc), ctypes.c_void_p,
                                    ctypes.POINTER(RenderDesc), ctypes.POINTER(ctypes.c_uint32)]),
    "tr_batch_clear": (ctypes.c_int, [ctypes.c_void_p]),
    "tr_batch_render": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LaunchDesc), ctypes.POINTER(RenderStats)]),
    "tr_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LaunchDesc), ctypes.POINTER(ctypes.c_double)]),
    "tr_guide_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(GuideDesc)]),
    "tr_guide_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "tr_guide_get_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "tr_guide_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "tr_guide_set_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "tr_guide_eval": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
    "tr_guide_eval_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                            ctypes.POINTER(ctypes.c_float)]),
    "tr_guide_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                         ctypes.POINTER(ctypes.c_float)]),
    "tr_guide_train_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                            ctypes.POINTER(ctypes.c_float)]),
    "tr_guide_info_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(GuideInfo)]),
    "tr_render_transient_guided": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CameraDesc), ctypes.c_void_p,
                                                  ctypes.POINTER(RenderDesc), ctypes.c_void_p,
                                                  ctypes.POINTER(RenderStats)]),
    "tr_last_error": (ctypes.c_char_p, []),
    "tr_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_uint64)]),
}

_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"HIP library not built: {LIB} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != TR_OK:
        raise TrError(rc, lib().tr_last_error().decode(errors="replace"))


def f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


def device_info():
    name = ctypes.create_string_buffer(256)
    cu = ctypes.c_int()
    mem = ctypes.c_uint64()
    check(lib().tr_device_info(name, 256, ctypes.byref(cu), ctypes.byref(mem)))
    return {"name": name.value.decode(), "compute_units": cu.value, "global_mem": mem.value}


def as_ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def obj_load(path: str):
    """tr_obj_load: (vertices (nv,3), triangles (nt,3), corner normals (nt,3,3) or None)."""
    xyz = ctypes.POINTER(ctypes.c_float)()
    idx = ctypes.POINTER(ctypes.c_uint32)()
    nrm = ctypes.POINTER(ctypes.c_float)()
    nv, nt = ctypes.c_uint32(), ctypes.c_uint32()
    check(lib().tr_obj_load(path.encode(), ctypes.byref(xyz), ctypes.byref(nv), ctypes.byref(idx), ctypes.byref(nt),
                            ctypes.byref(nrm)))
    try:
        v = np.ctypeslib.as_array(xyz, shape=(max(nv.value, 1) * 3,))[:nv.value * 3].reshape(-1, 3).copy()
        t = np.ctypeslib.as_array(idx, shape=(max(nt.value, 1) * 3,))[:nt.value * 3].reshape(-1, 3).copy()
        n = None
        if nrm:
            n = np.ctypeslib.as_array(nrm, shape=(nt.value * 9,)).reshape(-1, 3, 3).copy()
    finally:
        lib().tr_obj_free(ctypes.cast(xyz, ctypes.c_void_p), ctypes.cast(idx, ctypes.c_void_p),
                          ctypes.cast(nrm, ctypes.c_void_p))
    return v, t, n
