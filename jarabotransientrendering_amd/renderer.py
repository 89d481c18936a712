"""Host-side mirror of the reference renderer's interfaces for the transient path.

The reference wires (bunnykiller.cpp:1504-1827)::

    World w; w.add_triangle_mesh(...); w.add_light(...); w.set_medium(...); w.freeze();
    PinholeCamera camera(pos, focus, up, view_plane_dist);
    StratifiedSampler sampler(W, H, sqrt_spp, true);
    StreakCameraFilm film(W, H, T, exposure, camera_unwarp, BoxFilter(1), RADIANCE);
    BidirectionalPathTracing bpt(w, n_incoming, f_log, film);
    bpt.set_max_path_size(n); bpt.set_path_tracing_only(true); bpt.set_mode(SteadyState);
    TransientRenderer engine(f_log); engine.render(name, w, &bpt, camera, film, &sampler);

The classes below keep those names, arguments and error behaviour, and run the
whole render as one call into the HIP library (include/tr_api.h). Everything
executes on the GPU; there is no CPU path.
"""
from __future__ import annotations

import ctypes
import math
import os
from typing import List, Optional, Sequence

import numpy as np

from . import _native as nat


class Lambertian:
    """Lambertian BSDF inside a one-BSDF UberMaterial (bunnykiller.cpp:328-336)."""

    def __init__(self, kd: Sequence[float]):
        self.kd = tuple(float(x) for x in kd)


class PointLightSource:
    """PointLightSource (src/LightSource/PointLightSource.h:39-171)."""

    def __init__(self, position: Sequence[float], intensity: Sequence[float]):
        self.position = tuple(float(x) for x in position)
        self.intensity = tuple(float(x) for x in intensity)


class RectangularAreaLightSource:
    """RectangularAreaLightSource with the CLI parameterisation (bunnykiller.cpp:1242-1261):
    `target` is a point, the normal is normalize(target - corner)."""

    def __init__(self, corner: Sequence[float], target: Sequence[float], size: Sequence[float], intensity: float):
        self.corner = tuple(float(x) for x in corner)
        self.target = tuple(float(x) for x in target)
        self.size = (float(size[0]), float(size[1]))
        self.intensity = float(intensity)


class HomogeneousMedium:
    """HomogeneousMedium<Isotropic> without time delay (src/Media/HomogeneousMedium.h:45-127)."""

    def __init__(self, sigma_a: Sequence[float], sigma_s: Sequence[float]):
        self.sigma_a = tuple(float(x) for x in sigma_a)
        self.sigma_s = tuple(float(x) for x in sigma_s)


class World:
    """World (src/RayTracing/World.h:95-789): geometry, lights, medium; freeze() builds the BVH."""

    def __init__(self):
        h = ctypes.c_void_p()
        nat.check(nat.lib().tr_scene_create(ctypes.byref(h)))
        self._h = h
        self._lights: List[object] = []
        self._medium: Optional[HomogeneousMedium] = None
        self.frozen = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            nat.lib().tr_scene_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def add_triangle_mesh(self, vertices: np.ndarray, indices: np.ndarray, material: Lambertian,
                          normals: Optional[np.ndarray] = None, two_sided: bool = False) -> int:
        """World::add_triangle_mesh (World.h:209-541) from arrays: vertices (nv,3), indices (nt,3),
        optional per-corner shading normals (nt,3,3) and the -two-sided material flag."""
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        i = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
        n = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 9)
        m = nat.Material((ctypes.c_float * 3)(*material.kd))
        gid = ctypes.c_uint32()
        nat.check(nat.lib().tr_scene_add_mesh_ex(self._h, nat.as_ptr(v), v.shape[0], nat.as_ptr(i), i.shape[0],
                                                 nat.as_ptr(n) if n is not None else None, ctypes.byref(m),
                                                 nat.TR_MATERIAL_TWO_SIDED if two_sided else 0, ctypes.byref(gid)))
        self.frozen = False
        return int(gid.value)

    def add_obj(self, path: str, material: Lambertian, two_sided: bool = False) -> int:
        """-name-mesh file.obj (World::add_triangle_mesh, World.h:209-541)."""
        m = nat.Material((ctypes.c_float * 3)(*material.kd))
        gid = ctypes.c_uint32()
        nat.check(nat.lib().tr_scene_add_obj(self._h, path.encode(), ctypes.byref(m),
                                             nat.TR_MATERIAL_TWO_SIDED if two_sided else 0, ctypes.byref(gid)))
        self.frozen = False
        return int(gid.value)

    def add_triangles(self, tris: np.ndarray, material: Lambertian, normals: Optional[np.ndarray] = None,
                      two_sided: bool = False) -> int:
        """Triangle soup (n,3,3) as one mesh."""
        t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 3)
        return self.add_triangle_mesh(t, np.arange(t.shape[0], dtype=np.uint32).reshape(-1, 3), material,
                                      normals, two_sided)

    def add_light(self, light) -> None:
        """World::add_light (World.h:179-182)."""
        if isinstance(light, PointLightSource):
            nat.check(nat.lib().tr_scene_add_point_light(self._h, nat.f3(light.position), nat.f3(light.intensity)))
        elif isinstance(light, RectangularAreaLightSource):
            nat.check(nat.lib().tr_scene_add_rect_light(self._h, nat.f3(light.corner), nat.f3(light.target),
                                                        (ctypes.c_float * 2)(*light.size), light.intensity))
        else:
            raise TypeError(f"unsupported light source {type(light).__name__}")
        self._lights.append(light)
        self.frozen = False

    def nb_lights(self) -> int:
        return len(self._lights)

    def set_medium(self, medium: Optional[HomogeneousMedium]) -> None:
        """World::set_medium (World.h:711-715)."""
        if medium is None:
            nat.check(nat.lib().tr_scene_set_medium(self._h, None, None))
        else:
            a = np.asarray(medium.sigma_a, dtype=np.float32)
            s = np.asarray(medium.sigma_s, dtype=np.float32)
            nat.check(nat.lib().tr_scene_set_medium(self._h, nat.as_ptr(a), nat.as_ptr(s)))
        self._medium = medium
        self.frozen = False

    def get_medium(self):
        return self._medium

    def freeze(self) -> None:
        """World::freeze (World.h:543-550)."""
        nat.check(nat.lib().tr_scene_commit(self._h))
        self.frozen = True

    def info(self) -> dict:
        i = nat.SceneInfo()
        nat.check(nat.lib().tr_scene_info_get(self._h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in nat.SceneInfo._fields_}

    def first_intersection(self, origins: np.ndarray, directions: np.ndarray, tnear: float = 1e-5,
                           tfar: float = math.inf):
        """Batched World::first_intersection (World.h:566-615): returns (t, geom_id, prim_id, tri);
        misses have t = -1 and ids 0xFFFFFFFF."""
        o = np.asarray(origins, dtype=np.float32).reshape(-1, 3)
        d = np.asarray(directions, dtype=np.float32).reshape(-1, 3)
        n = o.shape[0]
        rays = np.zeros((n, 8), dtype=np.float32)
        rays[:, 0:3] = o
        rays[:, 3] = tnear
        rays[:, 4:7] = d
        rays[:, 7] = tfar
        hits = np.zeros((n, 4), dtype=np.uint32)
        nat.check(nat.lib().tr_intersect_batch(self._h, nat.as_ptr(rays), n, nat.as_ptr(hits)))
        t = hits[:, 0].view(np.float32).copy()
        return t, hits[:, 1].copy(), hits[:, 2].copy(), hits[:, 3].copy()

    def intersects(self, origins: np.ndarray, directions: np.ndarray, max_length=math.inf,
                   epsilon: float = 1e-5) -> np.ndarray:
        """Batched World::intersects (World.h:617-639): True where some triangle is hit in
        (epsilon, max_length - epsilon]; `max_length` may be a scalar or one value per ray, and
        epsilon defaults to _SIGMA_VISIBILITY_ (1e-5), as in the reference."""
        o = np.asarray(origins, dtype=np.float32).reshape(-1, 3)
        d = np.asarray(directions, dtype=np.float32).reshape(-1, 3)
        n = o.shape[0]
        rays = np.zeros((n, 8), dtype=np.float32)
        rays[:, 0:3] = o
        rays[:, 3] = epsilon
        rays[:, 4:7] = d
        rays[:, 7] = np.asarray(max_length, dtype=np.float64) - epsilon
        occ = np.zeros(n, dtype=np.uint8)
        nat.check(nat.lib().tr_occluded_batch(self._h, nat.as_ptr(rays), n, nat.as_ptr(occ)))
        return occ.astype(bool)


class PinholeCamera:
    """PinholeCamera (src/Camera/PinholeCamera.h:36-131); `fov_deg` as -camera-fov."""

    def __init__(self, position, focus, up, fov_deg: float = 60.0):
        self.position = tuple(float(x) for x in position)
        self.focus = tuple(float(x) for x in focus)
        self.up = tuple(float(x) for x in up)
        self.fov_deg = float(fov_deg)

    def desc(self) -> nat.CameraDesc:
        return nat.CameraDesc(nat.f3(self.position), nat.f3(self.focus), nat.f3(self.up), self.fov_deg)


class StratifiedSampler:
    """StratifiedSampler (src/Sampling/StratifiedSampler.h:32-125), jittered."""

    def __init__(self, size_x: int, size_y: int, sqrt_spp: int, jittered: bool = True):
        if not jittered:
            raise ValueError("only the jittered stratified sampler is used by the transient path")
        self.size_x, self.size_y, self.sqrt_spp = int(size_x), int(size_y), int(sqrt_spp)
        self.scanline = -1

    def set_scanline(self, y: int) -> None:
        self.scanline = int(y)

    def get_nb_samples(self) -> int:
        rows = 1 if self.scanline >= 0 else self.size_y
        return rows * self.size_x * self.sqrt_spp ** 2


class StreakCameraFilm:
    """StreakCameraFilm (src/Film/StreakCameraFilm.h:40-437) with a box filter of size 1."""

    bounce_layers = 0

    def __init__(self, width: int, height: int, time_bins: int, exposure: float, camera_unwarp: bool = False):
        self.width, self.height, self.time_bins = int(width), int(height), int(time_bins)
        self.exposure = float(exposure)
        self.camera_unwarp = bool(camera_unwarp)
        self.offset = 0.0
        self.streak = -1
        self._h = None

    # StreakCameraFilm::set_offset / set_streak / getters (StreakCameraFilm.h:197-230)
    def set_offset(self, offset: float) -> None:
        self._require_unallocated()
        self.offset = float(offset)

    def set_streak(self, y: int) -> None:
        self._require_unallocated()
        self.streak = int(y)

    def get_time_length(self) -> float:
        return self.time_bins * self.exposure + self.offset

    def get_exposure_time(self) -> float:
        return self.exposure

    def get_time_offset(self) -> float:
        return self.offset

    def get_time_resolution(self) -> int:
        return self.time_bins

    @property
    def rows(self) -> int:
        return 1 if self.streak >= 0 else self.height

    def _require_unallocated(self):
        if self._h is not None:
            raise RuntimeError("film geometry is fixed once rendering has started")

    @property
    def handle(self):
        if self._h is None:
            d = nat.FilmDesc(self.width, self.height, self.time_bins, self.exposure, self.offset, self.streak,
                             int(self.camera_unwarp), int(self.bounce_layers))
            h = ctypes.c_void_p()
            nat.check(nat.lib().tr_film_create(ctypes.byref(h), ctypes.byref(d)))
            self._h = h
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            nat.lib().tr_film_destroy(h)
            self._h = None

    def clean(self) -> None:
        nat.check(nat.lib().tr_film_clear(self.handle))

    def _lead(self):
        return (self.bounce_layers,) if self.bounce_layers else ()

    def read(self, steady: bool = False):
        """Normalised streak data [rows][W][T][3] (and the steady image [rows][W][3]); multibounce
        films return [layers][rows][W][T][3]."""
        out = np.empty(self._lead() + (self.rows, self.width, self.time_bins, 3), dtype=np.float32)
        st = np.empty((self.rows, self.width, 3), dtype=np.float32) if steady else None
        nat.check(nat.lib().tr_film_read(self.handle, nat.as_ptr(out), nat.as_ptr(st) if steady else None))
        return (out, st) if steady else out

    def read_rows(self, row_begin: int, row_count: int, steady: bool = False):
        """Normalised streak data of rows [row_begin, row_begin + row_count) only: [rows][W][T][3]
        (and their steady image [rows][W][3])."""
        out = np.empty((row_count, self.width, self.time_bins, 3), dtype=np.float32)
        st = np.empty((row_count, self.width, 3), dtype=np.float32) if steady else None
        nat.check(nat.lib().tr_film_read_rows(self.handle, int(row_begin), int(row_count), nat.as_ptr(out),
                                              nat.as_ptr(st) if steady else None))
        return (out, st) if steady else out

    def read_sums(self):
        """Compensated sums [rows][W][T][3] (float64) and weights [rows][W] (multibounce films:
        a leading layer axis on both)."""
        s = np.empty(self._lead() + (self.rows, self.width, self.time_bins, 3), dtype=np.float64)
        w = np.empty(self._lead() + (self.rows, self.width), dtype=np.float64)
        nat.check(nat.lib().tr_film_read_sums(self.handle, nat.as_ptr(s), nat.as_ptr(w)))
        return s, w

    def add_samples(self, samples: np.ndarray) -> None:
        """StreakCameraFilm::add_samples (StreakCameraFilm.h:139-192) for caller-made samples: a
        structured array of dtype FILM_SAMPLE_DTYPE (x, y, time, rgb, weight), deposited with
        compensated additions by the renders' film kernel."""
        a = np.ascontiguousarray(samples, dtype=nat.FILM_SAMPLE_DTYPE)
        nat.check(nat.lib().tr_film_add_samples(self.handle, nat.as_ptr(a), a.shape[0]))

    def write(self, name: str, extension: str = "hdr") -> List[str]:
        """StreakCameraFilm::write + Film::write (StreakCameraFilm.h:233-242, 357-437; Film.h:358-471):
        `<name>_%04d.<extension>` per scanline (an image T wide and W high, transposed when T < 8),
        `extension` "hdr" (Radiance RGBE) or "raw" (FP32), then the steady image `<name>.hdr`."""
        d = os.path.dirname(name)
        if d:
            os.makedirs(d, exist_ok=True)
        nat.check(nat.lib().tr_film_write(self.handle, name.encode(), extension.encode()))
        rows = [self.streak] if self.streak >= 0 else range(self.height)
        if self.bounce_layers:
            return [f"{name}_b{b + 1:02d}_{y:04d}.{extension}" for b in range(self.bounce_layers) for y in rows] + \
                [f"{name}.hdr"]
        return [f"{name}_{y:04d}.{extension}" for y in rows] + [f"{name}.hdr"]


class MultibounceStreakFilm(StreakCameraFilm):
    """MultibounceStreakFilm (src/Film/MultibounceStreakCameraFilm.h:36-200): one streak film per
    bounce (samples of bounce >= nb_bounces go to the last), each normalised by the camera samples
    that sent it samples, and a steady image of every sample."""

    def __init__(self, width: int, height: int, time_bins: int, exposure: float, nb_bounces: int,
                 camera_unwarp: bool = False):
        super().__init__(width, height, time_bins, exposure, camera_unwarp)
        if not 1 <= int(nb_bounces) <= 32:
            raise ValueError("nb_bounces must be in [1, 32]")
        self.bounce_layers = int(nb_bounces)


class BidirectionalPathTracing:
    """Integrator settings of BidirectionalPathTracing in its path-tracing-only configuration
    (BidirectionalPathTracing.h:122-217), with the SteadyState sampling techniques (the default) or
    the Transient ones (set_mode(Transient), -time-sampling; :190-198)."""

    SteadyState, Transient = 0, 1

    def __init__(self, world: World, incoming_samples: int = 1, f_log=None, film: Optional[StreakCameraFilm] = None):
        if incoming_samples <= 0:
            raise ValueError("incoming_samples must be positive")
        self.world = world
        self.incoming_samples = int(incoming_samples)
        self.film = film
        self.max_path_size = 30
        self.path_trace_only = True
        self.mode = self.SteadyState

    def set_max_path_size(self, n: int) -> None:
        self.max_path_size = int(n)

    def set_path_tracing_only(self, on: bool) -> None:
        if not on:
            raise NotImplementedError("light subpaths (full BDPT) are not part of this path")
        self.path_trace_only = True

    def set_mode(self, mode: int) -> None:
        if mode not in (self.SteadyState, self.Transient):
            raise ValueError(f"unknown sampling mode {mode}")
        self.mode = mode

    def _flags(self) -> int:
        return nat.TR_RENDER_TIME_SAMPLING if self.mode == self.Transient else 0

    def preprocess(self) -> None:
        pass


class PathGuide:
    """Online-trained vMF-mixture guiding network (tr_guide_*; no counterpart in the reference).
    Pass it to TransientRenderer.render(..., guide=...) to sample directions from the one-sample
    MIS mixture of the BSDF/phase function and the network; weights persist across renders."""

    def __init__(self, learning_rate: float = 1e-3, bsdf_fraction: float = 0.5, train_paths: int = 1 << 16,
                 batch_size: int = 1 << 14, steps_per_wave: int = 1, init_seed: int = 1):
        self.handle = ctypes.c_void_p()
        d = nat.GuideDesc(learning_rate, bsdf_fraction, train_paths, batch_size, steps_per_wave, 0, init_seed)
        nat.check(nat.lib().tr_guide_create(ctypes.byref(self.handle), ctypes.byref(d)))
        self.batch_size = batch_size

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                nat.lib().tr_guide_destroy(h)
            except Exception:
                pass

    @property
    def params(self) -> np.ndarray:
        p = np.empty(nat.GUIDE_PARAMS, np.float32)
        nat.check(nat.lib().tr_guide_get_params(self.handle, nat.as_ptr(p)))
        return p

    @params.setter
    def params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.shape != (nat.GUIDE_PARAMS,):
            raise ValueError(f"expected {nat.GUIDE_PARAMS} parameters")
        nat.check(nat.lib().tr_guide_set_params(self.handle, nat.as_ptr(p)))

    def eval(self, raw) -> np.ndarray:
        raw = np.ascontiguousarray(raw, dtype=np.float32).reshape(-1, nat.GUIDE_RAW)
        out = np.empty((len(raw), nat.GUIDE_OUT), np.float32)
        nat.check(nat.lib().tr_guide_eval(self.handle, nat.as_ptr(raw), len(raw), nat.as_ptr(out)))
        return out

    def gradient(self, records):
        rec = np.ascontiguousarray(records, dtype=np.float32).reshape(-1, nat.GUIDE_RECORD)
        g = np.empty(nat.GUIDE_PARAMS, np.float32)
        loss = ctypes.c_float()
        nat.check(nat.lib().tr_guide_gradient(self.handle, nat.as_ptr(rec), len(rec), nat.as_ptr(g),
                                              ctypes.byref(loss)))
        return g, float(loss.value)

    def train_batch(self, records) -> float:
        rec = np.ascontiguousarray(records, dtype=np.float32).reshape(-1, nat.GUIDE_RECORD)
        loss = ctypes.c_float()
        nat.check(nat.lib().tr_guide_train_batch(self.handle, nat.as_ptr(rec), len(rec), ctypes.byref(loss)))
        return float(loss.value)

    def set_steps(self, steps_per_wave: int) -> None:
        """Adam steps after each wavefront of later renders; 0 freezes the network."""
        nat.check(nat.lib().tr_guide_set_steps(self.handle, int(steps_per_wave)))

    def info(self) -> dict:
        i = nat.GuideInfo()
        nat.check(nat.lib().tr_guide_info_get(self.handle, ctypes.byref(i)))
        return {"steps": int(i.steps), "records": int(i.records), "last_loss": float(i.last_loss),
                "n_params": int(i.n_params), "flop_per_eval": int(i.flop_per_eval)}


class TransientRenderer:
    """TransientRenderer::render (src/Renderer/TransientRenderer.h:77-201)."""

    def __init__(self, f_log=None, time_sampling: bool = False):
        # time_sampling only changes the (unused) time arguments of World::Li
        # (TransientRenderer.h:160-164); the integrator's mode selects the sampling techniques
        self.time_sampling = bool(time_sampling)
        self.wave_paths = 0
        self.profile = False
        self.binary_bvh = False   # traverse the binary BVH instead of the 4-wide one
        self.orthonormal_frame = False  # Lambertian sampling frame without the reference's quirk
        self.graph = False              # replay each wavefront from a captured hipGraph
        self.extra_flags = 0            # launch tuning flags (TR_RENDER_STACK32 / TR_RENDER_STACK_KEYED)
        self.last_stats: Optional[dict] = None

    def render(self, name: Optional[str], world: World, integrator: Optional[BidirectionalPathTracing],
               camera: PinholeCamera, film: StreakCameraFilm, sampler: StratifiedSampler, seed: int = 0,
               guide: Optional[PathGuide] = None) -> dict:
        if integrator is None:
            raise RuntimeError("Error: Need to define an integrator for rendering\n")
        if not world.frozen:
            world.freeze()
        if sampler.scanline >= 0 and film.streak != sampler.scanline:
            film.set_streak(sampler.scanline)
        if (sampler.size_x, sampler.size_y) != (film.width, film.height):
            raise ValueError("sampler and film sizes differ")
        integrator.preprocess()
        flags = ((nat.TR_RENDER_PROFILE if self.profile else 0) | (nat.TR_RENDER_BINARY_BVH if self.binary_bvh else 0)
                 | (nat.TR_RENDER_ORTHONORMAL_FRAME if self.orthonormal_frame else 0)
                 | (nat.TR_RENDER_GRAPH if self.graph else 0) | integrator._flags() | int(self.extra_flags))
        d = nat.RenderDesc(sampler.sqrt_spp, integrator.incoming_samples, integrator.max_path_size, 0, int(seed),
                           int(self.wave_paths), flags)
        st = nat.RenderStats()
        cam = camera.desc()
        if guide is None:
            nat.check(nat.lib().tr_render_transient(world.handle, ctypes.byref(cam), film.handle, ctypes.byref(d),
                                                    ctypes.byref(st)))
        else:
            nat.check(nat.lib().tr_render_transient_guided(world.handle, ctypes.byref(cam), film.handle,
                                                           ctypes.byref(d), guide.handle, ctypes.byref(st)))
        self.last_stats = st.as_dict()
        if name:
            film.write(name)
        return self.last_stats


# ---------------------------------------------------------------------------------------------
class Batch:
    """Many render jobs in one launch sequence (tr_batch_*; no counterpart in the reference).
    Each job accumulates into its own film exactly as TransientRenderer.render would."""

    def __init__(self):
        self.handle = ctypes.c_void_p()
        nat.check(nat.lib().tr_batch_create(ctypes.byref(self.handle)))
        self._keep = []  # scenes and films must outlive the batch's renders

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                nat.lib().tr_batch_destroy(h)
            except Exception:
                pass

    def add(self, world: World, integrator: BidirectionalPathTracing, camera: PinholeCamera,
            film: StreakCameraFilm, sampler: StratifiedSampler, seed: int = 0) -> int:
        if not world.frozen:
            world.freeze()
        if sampler.scanline >= 0 and film.streak != sampler.scanline:
            film.set_streak(sampler.scanline)
        d = nat.RenderDesc(sampler.sqrt_spp, integrator.incoming_samples, integrator.max_path_size, 0, int(seed), 0,
                           integrator._flags())
        cam = camera.desc()
        job = ctypes.c_uint32()
        nat.check(nat.lib().tr_batch_add(self.handle, world.handle, ctypes.byref(cam), film.handle, ctypes.byref(d),
                                         ctypes.byref(job)))
        self._keep.append((world, film))
        return int(job.value)

    @staticmethod
    def _launch(wave_paths=0, flags=0, trace_per_cu=0, shade_per_cu=0):
        return nat.LaunchDesc(int(wave_paths), int(flags), int(trace_per_cu), int(shade_per_cu))

    def render(self, launch: Optional[dict] = None) -> dict:
        st = nat.RenderStats()
        L = self._launch(**launch) if launch is not None else None
        nat.check(nat.lib().tr_batch_render(self.handle, ctypes.byref(L) if L is not None else None,
                                            ctypes.byref(st)))
        return st.as_dict()

    def autotune(self):
        best = nat.LaunchDesc()
        ms = ctypes.c_double()
        nat.check(nat.lib().tr_autotune(self.handle, ctypes.byref(best), ctypes.byref(ms)))
        return best.as_dict(), float(ms.value)


def setup_scene(scene):
    """Build (world, camera, sampler, film, integrator) for a `scenes.Scene`, the way
    bunnykiller.cpp wires them from command-line flags."""
    w = World()
    # one mesh per run of equal material, so scene triangle indices keep the scene's order
    tm = np.asarray(scene.tri_mat)
    start = 0
    vn = getattr(scene, "vertex_normals", None)
    ts = getattr(scene, "two_sided", None)
    for i in range(1, len(tm) + 1):
        if i == len(tm) or tm[i] != tm[start]:
            nrm = None
            if vn is not None and np.any(vn[start:i] != 0):
                nrm = vn[start:i]
            w.add_triangles(scene.tris[start:i], Lambertian(scene.materials[tm[start]]), nrm,
                            bool(ts[tm[start]]) if ts is not None else False)
            start = i
    for p in scene.point_lights:
        w.add_light(PointLightSource(p[0:3], p[3:6]))
    for r in scene.rect_lights:
        w.add_light(RectangularAreaLightSource(r[0:3], r[3:6], r[6:8], r[8]))
    if scene.medium is not None:
        w.set_medium(HomogeneousMedium(scene.medium[0:3], scene.medium[3:6]))
    w.freeze()
    c = scene.camera
    cam = PinholeCamera(c.position, c.focus, c.up, c.fov_deg)
    f = scene.film
    sampler = StratifiedSampler(f.width, f.height, scene.render.sqrt_spp)
    if getattr(f, "bounce_layers", 0):
        film = MultibounceStreakFilm(f.width, f.height, f.time_bins, f.exposure, f.bounce_layers, f.camera_unwarp)
    else:
        film = StreakCameraFilm(f.width, f.height, f.time_bins, f.exposure, f.camera_unwarp)
    if f.offset > 0:
        film.set_offset(f.offset)
    if f.scanline >= 0:
        sampler.set_scanline(f.scanline)
        film.set_streak(f.scanline)
    integ = BidirectionalPathTracing(w, scene.render.n_incoming, None, film)
    integ.set_max_path_size(scene.render.max_bounces)
    if getattr(scene.render, "time_sampling", False):
        integ.set_mode(BidirectionalPathTracing.Transient)
    return w, cam, sampler, film, integ


def render_scene(scene, seed: Optional[int] = None, wave_paths: int = 0, profile: bool = False,
                 binary_bvh: bool = False, orthonormal_frame: bool = False, guide: Optional[PathGuide] = None,
                 graph: bool = False):
    """Render a `scenes.Scene`; returns (normalised film [rows][W][T][3], stats, film object)."""
    w, cam, sampler, film, integ = setup_scene(scene)
    eng = TransientRenderer()
    eng.wave_paths = wave_paths
    eng.profile = profile
    eng.binary_bvh = binary_bvh
    eng.orthonormal_frame = orthonormal_frame
    eng.graph = graph
    st = eng.render(None, w, integ, cam, film, sampler, seed=scene.render.seed if seed is None else seed,
                    guide=guide)
    return film.read(), st, film
