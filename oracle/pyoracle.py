"""ctypes binding of the CPU oracle (oracle/oracle.c) — TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker / CPU baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

RNG_REFERENCE = 0
RNG_PER_PATH = 1
RNG_REFERENCE_PER_CS = 2
ISECT_MOLLER = 0
ISECT_WATERTIGHT = 1


class _Scene(ctypes.Structure):
    _fields_ = [
        ("n_tris", ctypes.c_int32), ("tris", ctypes.c_void_p), ("tri_mat", ctypes.c_void_p),
        ("n_mats", ctypes.c_int32), ("mats", ctypes.c_void_p),
        ("n_points", ctypes.c_int32), ("points", ctypes.c_void_p),
        ("n_rects", ctypes.c_int32), ("rects", ctypes.c_void_p),
        ("has_medium", ctypes.c_int32), ("medium", ctypes.c_float * 6),
        ("vn", ctypes.c_void_p), ("mat_flags", ctypes.c_void_p),
    ]


class _Render(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("time_bins", ctypes.c_int32),
        ("exposure", ctypes.c_float), ("offset", ctypes.c_float),
        ("scanline", ctypes.c_int32), ("camera_unwarp", ctypes.c_int32),
        ("camera", ctypes.c_float * 10),
        ("sqrt_spp", ctypes.c_int32), ("n_incoming", ctypes.c_int32), ("max_bounces", ctypes.c_int32),
        ("seed", ctypes.c_uint64), ("rng_mode", ctypes.c_int32),
        ("row_begin", ctypes.c_int32), ("row_end", ctypes.c_int32),
        ("tnear", ctypes.c_float), ("shadow_tnear", ctypes.c_float), ("shadow_tfar_eps", ctypes.c_float),
        ("bounce_layers", ctypes.c_int32), ("time_sampling", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [("camera_samples", ctypes.c_uint64), ("paths", ctypes.c_uint64),
                ("extension_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
                ("contributions", ctypes.c_uint64), ("render_seconds", ctypes.c_double)]

    def as_dict(self):
        d = {k: int(getattr(self, k)) for k, _ in self._fields_[:-1]}
        d["render_seconds"] = float(self.render_seconds)
        return d


class _PCG(ctypes.Structure):
    _fields_ = [("state", ctypes.c_uint64), ("inc", ctypes.c_uint64)]


_lib = None


def build() -> None:
    subprocess.check_call(["sh", os.path.join(HERE, "build_oracle.sh")])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_render.argtypes = [ctypes.POINTER(_Scene), ctypes.POINTER(_Render), ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.POINTER(Stats)]
        L.orc_render.restype = ctypes.c_int
        L.orc_render_ex.argtypes = [ctypes.POINTER(_Scene), ctypes.POINTER(_Render), ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Stats)]
        L.orc_render_ex.restype = ctypes.c_int
        L.orc_intersect.argtypes = [ctypes.POINTER(_Scene), ctypes.c_void_p, ctypes.c_int32, ctypes.c_float,
                                    ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_intersect.restype = ctypes.c_int
        L.orc_pcg32_seed.argtypes = [ctypes.POINTER(_PCG), ctypes.c_uint64, ctypes.c_uint64]
        L.orc_pcg32_default.argtypes = [ctypes.POINTER(_PCG)]
        L.orc_pcg32_next.argtypes = [ctypes.POINTER(_PCG)]
        L.orc_pcg32_next.restype = ctypes.c_uint32
        L.orc_pcg32_next_float.argtypes = [ctypes.POINTER(_PCG)]
        L.orc_pcg32_next_float.restype = ctypes.c_float
        _lib = L
    return _lib


class PCG32:
    """The reference's Random::RNG (RandomNumbers.h:59-212) as restated by the oracle."""

    def __init__(self, state=None, seq=0xda3e39cb94b95bdb):
        self._r = _PCG()
        if state is None:
            lib().orc_pcg32_default(ctypes.byref(self._r))
        else:
            lib().orc_pcg32_seed(ctypes.byref(self._r), state, seq)

    def next_int(self) -> int:
        return int(lib().orc_pcg32_next(ctypes.byref(self._r)))

    def next_float(self) -> float:
        return float(lib().orc_pcg32_next_float(ctypes.byref(self._r)))


def _scene_struct(scene):
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data if a.size else None

    s = _Scene()
    s.n_tris = scene.tris.shape[0]
    s.tris = arr(scene.tris.reshape(-1), np.float32)
    s.tri_mat = arr(scene.tri_mat, np.int32)
    s.n_mats = scene.materials.shape[0]
    s.mats = arr(scene.materials.reshape(-1), np.float32)
    s.n_points = scene.point_lights.shape[0]
    s.points = arr(scene.point_lights.reshape(-1), np.float32)
    s.n_rects = scene.rect_lights.shape[0]
    s.rects = arr(scene.rect_lights.reshape(-1), np.float32)
    if scene.medium is not None:
        s.has_medium = 1
        for i in range(6):
            s.medium[i] = float(scene.medium[i])
    vn = getattr(scene, "vertex_normals", None)
    if vn is not None:
        s.vn = arr(np.asarray(vn, dtype=np.float32).reshape(-1), np.float32)
    ts = getattr(scene, "two_sided", None)
    if ts is not None:
        s.mat_flags = arr(np.asarray(ts, dtype=np.int32).reshape(-1), np.int32)
    return s, keep


# ray epsilons of the reference's non-Embree World: Triangle's SMALLEST_DIST = 1e-4 (Ray.h:36),
# shadow rays shortened by 1e-5 * scene radius (World.h:687-691); the unit box has radius 0.5
NON_EMBREE_EPS = (1e-4, 1e-4, 0.5e-5)


def render(scene, rng_mode=RNG_PER_PATH, rows=None, seed=None, eps=None, bounce_layers=0, extra=None):
    """Returns (film_sum [rows][W][T][3] float64, weight [rows][W] float64, stats dict).
    With bounce_layers = B > 0 the film is [B][rows][W][T][3] (MultibounceStreakFilm layers) and,
    if `extra` is a dict, extra["layer_weight"] ([B][rows][W]) and extra["steady"] ([rows][W][3],
    every sample regardless of the time window) are filled in."""
    s, keep = _scene_struct(scene)
    f, c, r = scene.film, scene.camera, scene.render
    d = _Render()
    d.width, d.height, d.time_bins = f.width, f.height, f.time_bins
    d.exposure, d.offset = f.exposure, f.offset
    d.scanline, d.camera_unwarp = f.scanline, int(f.camera_unwarp)
    for i, v in enumerate((*c.position, *c.focus, *c.up, c.fov_deg)):
        d.camera[i] = float(v)
    d.sqrt_spp, d.n_incoming, d.max_bounces = r.sqrt_spp, r.n_incoming, r.max_bounces
    d.seed = r.seed if seed is None else seed
    d.rng_mode = rng_mode
    if eps is not None:
        d.tnear, d.shadow_tnear, d.shadow_tfar_eps = eps
    if f.scanline >= 0:
        nrows = 1
    else:
        r0, r1 = rows if rows is not None else (0, f.height)
        d.row_begin, d.row_end = r0, r1
        nrows = r1 - r0
    d.bounce_layers = int(bounce_layers)
    d.time_sampling = int(bool(getattr(r, "time_sampling", False)))
    shape = (nrows, f.width, f.time_bins, 3)
    film = np.zeros(((bounce_layers,) if bounce_layers else ()) + shape, dtype=np.float64)
    weight = np.zeros((nrows, f.width), dtype=np.float64)
    lw = np.zeros((max(bounce_layers, 1), nrows, f.width), dtype=np.float64)
    steady = np.zeros((nrows, f.width, 3), dtype=np.float64)
    st = Stats()
    rc = lib().orc_render_ex(ctypes.byref(s), ctypes.byref(d), film.ctypes.data, weight.ctypes.data,
                             lw.ctypes.data, steady.ctypes.data, ctypes.byref(st))
    if extra is not None:
        extra["layer_weight"] = lw[:bounce_layers] if bounce_layers else None
        extra["steady"] = steady
    if rc != 0:
        raise ValueError(f"orc_render failed with status {rc}")
    del keep
    return film, weight, st.as_dict()


def intersect(scene, rays, tnear=1e-5, method=ISECT_MOLLER, use_bvh=True):
    """Closest hit per ray; rays (n,6) float32 origin+direction. Returns (ids int32, t float32)."""
    s, keep = _scene_struct(scene)
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    n = rays.shape[0]
    ids = np.empty(n, dtype=np.int32)
    t = np.empty(n, dtype=np.float32)
    lib().orc_intersect(ctypes.byref(s), rays.ctypes.data, n, tnear, method, int(use_bvh),
                        ids.ctypes.data, t.ctypes.data)
    del keep
    return ids, t


def normalize(film_sum, weight):
    """StreakCameraFilm::average_and_save_slice (StreakCameraFilm.h:369-397) with each pixel
    divided by its own weight; 0/0 -> 0."""
    w = weight[:, :, None, None]
    with np.errstate(invalid="ignore", divide="ignore"):
        out = film_sum / w
    out[~np.isfinite(out)] = 0.0
    return out
