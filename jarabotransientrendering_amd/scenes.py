"""Procedural scenes for the transient path tracer.

The reference renders scenes loaded from OBJ files through ``-name-mesh``
(bunnykiller.cpp:995-1002, World.h:209-541).  The scenes used here are built
procedurally so that the reference harness, the CPU oracle and the GPU path can
all ingest exactly the same FP32 triangles:

* ``cornell_box``: unit box of 5 quads (front open) with inward normals under the
  reference's normal convention ``ng = normalize(cross(p2-p0, p1-p0))``
  (World.h:449), plus an inner box of 12 triangles with outward normals.
* ``icosphere``: subdivided icosahedron (20*4^k triangles), radially displaced by
  seeded value noise, outward normals under the same convention.
* point and rectangular area lights with the reference's CLI parameterisation
  (bunnykiller.cpp:1194-1207, 1242-1261): the rectangle is given by a corner, a
  target point (normal = normalize(target - corner)), its two side lengths and
  a scalar intensity.

A scene serialises to the small text format read by ``oracle/ref_harness.cpp``.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Tuple

import numpy as np

RED = (0.6, 0.0, 0.0)
GREEN = (0.0, 0.6, 0.0)
WHITE = (0.75, 0.75, 0.75)


@dataclasses.dataclass
class Film:
    """Streak film geometry (StreakCameraFilm.h:91-121)."""
    width: int = 64
    height: int = 64
    time_bins: int = 256
    exposure: float = 0.01
    offset: float = 0.0
    scanline: int = -1          # -film-scanline y (bunnykiller.cpp:1082-1086); -1 = whole film
    camera_unwarp: bool = False  # -camera-unwarp (bunnykiller.cpp:1442)
    bounce_layers: int = 0       # -multibounce-streak N (bunnykiller.cpp:1413-1417); 0 = one streak film


@dataclasses.dataclass
class Camera:
    """Pinhole camera (PinholeCamera.h:56-108); fov in degrees (bunnykiller.cpp:1156-1160)."""
    position: Tuple[float, float, float] = (0.5, 0.5, -1.2)
    focus: Tuple[float, float, float] = (0.5, 0.5, 0.5)
    up: Tuple[float, float, float] = (0.0, 1.0, 0.0)
    fov_deg: float = 45.0


@dataclasses.dataclass
class RenderParams:
    """Sampler/integrator settings (-camera-spp, -bidirectional-path-tracing,
    -max-nb-bounces, -new-seed, -time-sampling; bunnykiller.cpp:1151, 1389, 1367, 1356, 1337-1340)."""
    sqrt_spp: int = 4
    n_incoming: int = 1
    max_bounces: int = 4
    seed: int = 0
    time_sampling: bool = False  # the integrator's Transient sampling mode (BidirectionalPathTracing.h:190-198)


@dataclasses.dataclass
class Scene:
    tris: np.ndarray                 # (n, 3, 3) float32, vertex order as given to the reference
    tri_mat: np.ndarray              # (n,) int32 material index
    materials: np.ndarray            # (m, 3) float32 Lambertian Kd
    point_lights: np.ndarray         # (p, 6) float32: position, RGB intensity
    rect_lights: np.ndarray          # (r, 9) float32: corner, target, size_x, size_y, intensity
    medium: Optional[np.ndarray] = None  # (6,) float32 sigma_a RGB, sigma_s RGB (isotropic phase)
    vertex_normals: Optional[np.ndarray] = None  # (n, 3, 3) float32 corner normals, zeros = flat triangle
    two_sided: Optional[np.ndarray] = None       # (m,) bool: material is two-sided (-two-sided)
    camera: Camera = dataclasses.field(default_factory=Camera)
    film: Film = dataclasses.field(default_factory=Film)
    render: RenderParams = dataclasses.field(default_factory=RenderParams)
    name: str = "scene"

    @property
    def n_tris(self) -> int:
        return int(self.tris.shape[0])

    def n_paths(self) -> int:
        rows = 1 if self.film.scanline >= 0 else self.film.height
        return rows * self.film.width * self.render.sqrt_spp ** 2 * self.render.n_incoming

    def to_ref_text(self, out_prefix: str) -> str:
        f, c, r = self.film, self.camera, self.render
        lines = [
            f"film {f.width} {f.height} {f.time_bins} {f.exposure!r} {f.offset!r} {f.scanline} {int(f.camera_unwarp)}",
            "camera " + " ".join(repr(float(v)) for v in (*c.position, *c.focus, *c.up, c.fov_deg)),
            f"render {r.sqrt_spp} {r.n_incoming} {r.max_bounces} {r.seed}",
        ]
        for m in self.materials:
            lines.append("material " + " ".join(repr(float(v)) for v in m))
        vn = self.vertex_normals
        for i, (t, m) in enumerate(zip(self.tris, self.tri_mat)):
            if vn is not None and np.any(vn[i] != 0):
                lines.append("trin " + " ".join(repr(float(v)) for v in t.reshape(-1)) + " " +
                             " ".join(repr(float(v)) for v in vn[i].reshape(-1)) + f" {int(m)}")
            else:
                lines.append("tri " + " ".join(repr(float(v)) for v in t.reshape(-1)) + f" {int(m)}")
        if self.two_sided is not None:
            for m in np.flatnonzero(self.two_sided):
                lines.append(f"twosided {int(m)}")
        for p in self.point_lights:
            lines.append("point " + " ".join(repr(float(v)) for v in p))
        for q in self.rect_lights:
            lines.append("rect " + " ".join(repr(float(v)) for v in q))
        if self.medium is not None:
            lines.append("medium " + " ".join(repr(float(v)) for v in self.medium))
        if f.bounce_layers:
            lines.append(f"multibounce {int(f.bounce_layers)}")
        if r.time_sampling:
            lines.append("timesampling")
        lines.append(f"out {out_prefix}")
        return "\n".join(lines) + "\n"


def _quad(a, b, c, d) -> List[np.ndarray]:
    """Two triangles (a,b,c), (a,c,d)."""
    return [np.array([a, b, c]), np.array([a, c, d])]


def _orient(tri: np.ndarray, inward_to: Optional[np.ndarray], outward_from: Optional[np.ndarray]) -> np.ndarray:
    """Order the vertices so that ng = cross(p2-p0, p1-p0) points towards `inward_to`
    (or away from `outward_from`)."""
    p0, p1, p2 = tri
    ng = np.cross(p2 - p0, p1 - p0)
    ctr = tri.mean(axis=0)
    if inward_to is not None:
        want = inward_to - ctr
    else:
        want = ctr - outward_from
    if np.dot(ng, want) < 0:
        return np.array([p0, p2, p1])
    return tri


def _box_quads(lo, hi) -> List[Tuple[str, List[np.ndarray]]]:
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    P = lambda x, y, z: np.array([x, y, z], dtype=np.float64)
    return [
        ("floor", _quad(P(x0, y0, z0), P(x1, y0, z0), P(x1, y0, z1), P(x0, y0, z1))),
        ("ceiling", _quad(P(x0, y1, z0), P(x0, y1, z1), P(x1, y1, z1), P(x1, y1, z0))),
        ("back", _quad(P(x0, y0, z1), P(x1, y0, z1), P(x1, y1, z1), P(x0, y1, z1))),
        ("left", _quad(P(x0, y0, z0), P(x0, y0, z1), P(x0, y1, z1), P(x0, y1, z0))),
        ("right", _quad(P(x1, y0, z0), P(x1, y1, z0), P(x1, y1, z1), P(x1, y0, z1))),
        ("front", _quad(P(x0, y0, z0), P(x0, y1, z0), P(x1, y1, z0), P(x1, y0, z0))),
    ]


def box_shell(inner_box: bool = True):
    """Unit box (front open) with inward normals; optional inner box with outward normals.
    Returns (tris, tri_mat, materials)."""
    materials = [WHITE, RED, GREEN]
    centre = np.array([0.5, 0.5, 0.5])
    tris, mats = [], []
    for name, quad in _box_quads((0, 0, 0), (1, 1, 1)):
        if name == "front":
            continue
        m = 1 if name == "left" else 2 if name == "right" else 0
        for t in quad:
            tris.append(_orient(t, centre, None))
            mats.append(m)
    if inner_box:
        lo, hi = (0.15, 0.0, 0.45), (0.45, 0.35, 0.75)
        c = (np.array(lo) + np.array(hi)) / 2
        for _name, quad in _box_quads(lo, hi):
            for t in quad:
                tris.append(_orient(t, None, c))
                mats.append(0)
    return (np.asarray(tris, dtype=np.float32), np.asarray(mats, dtype=np.int32),
            np.asarray(materials, dtype=np.float32))


def cornell_box(width=64, height=64, time_bins=256, exposure=0.01, sqrt_spp=4, n_incoming=1,
                max_bounces=4, seed=0) -> Scene:
    """CPU-reference workload: 22-triangle box, point light at (0.5,0.95,0.5) with I=10."""
    tris, tri_mat, mats = box_shell(inner_box=True)
    return Scene(
        tris=tris, tri_mat=tri_mat, materials=mats,
        point_lights=np.array([[0.5, 0.95, 0.5, 10.0, 10.0, 10.0]], dtype=np.float32),
        rect_lights=np.zeros((0, 9), dtype=np.float32),
        film=Film(width, height, time_bins, exposure),
        render=RenderParams(sqrt_spp, n_incoming, max_bounces, seed),
        name="cornell_box")


# ---------------------------------------------------------------------------------------------
# icosphere + value noise

def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.uint64) * np.uint64(73856093)) ^ (iy.astype(np.uint64) * np.uint64(19349663)) \
        ^ (iz.astype(np.uint64) * np.uint64(83492791)) ^ (np.uint64(seed) * np.uint64(2654435761))
    h = h & np.uint64(0xFFFFFFFF)
    # integer avalanche (32-bit xorshift-multiply)
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x7FEB352D)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(15)
    h = (h * np.uint64(0x846CA68B)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(16)
    return h.astype(np.float64) / 4294967296.0


def value_noise(p: np.ndarray, seed: int, freq: float = 4.0) -> np.ndarray:
    """Trilinear value noise in [0,1) on a lattice of spacing 1/freq with smoothstep fade."""
    q = p * freq + 1000.0
    i = np.floor(q).astype(np.int64)
    f = q - i
    f = f * f * (3 - 2 * f)
    out = np.zeros(p.shape[0])
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * \
                    (f[:, 2] if dz else 1 - f[:, 2])
                out += w * _hash3(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
    return out


def icosphere_mesh(level: int):
    """Unit icosphere: (verts (v,3) float64, faces (f,3) int64 CCW seen from outside)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
         (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.asarray(verts), np.asarray(faces, dtype=np.int64)


def displaced_icosphere(level=6, centre=(0.5, 0.4, 0.5), radius=0.3, amplitude=0.05, seed=0):
    """Triangles (n,3,3) float32 with ng = cross(p2-p0, p1-p0) pointing outward."""
    verts, faces = icosphere_mesh(level)
    n = value_noise(verts, seed)
    r = radius + amplitude * (2.0 * n - 1.0)
    pos = np.asarray(centre) + verts * r[:, None]
    # CCW-outward faces have cross(p1-p0, p2-p0) outward; swap 1<->2 for the reference convention
    tris = pos[faces[:, [0, 2, 1]]]
    return tris.astype(np.float32)


def icosphere_box(level=6, width=256, height=256, time_bins=1024, exposure=0.01, sqrt_spp=32,
                  n_incoming=1, max_bounces=8, seed=0, albedo=0.7, light_intensity=20.0) -> Scene:
    """Plain-GPU workload: displaced icosphere (20*4^level triangles, albedo 0.7) in the open box,
    lit by a 0.3x0.3 area light just below the ceiling facing -y."""
    tris, tri_mat, mats = box_shell(inner_box=False)
    sphere = displaced_icosphere(level, seed=seed)
    mats = np.concatenate([mats, np.array([[albedo] * 3], dtype=np.float32)])
    tris = np.concatenate([tris, sphere])
    tri_mat = np.concatenate([tri_mat, np.full(sphere.shape[0], mats.shape[0] - 1, dtype=np.int32)])
    # corner (0.65, 0.99, 0.65); the rectangle spans -x and -z from it (RectangularAreaLightSource.h:66-80)
    rect = np.array([[0.65, 0.99, 0.65, 0.65, 0.0, 0.65, 0.3, 0.3, light_intensity]], dtype=np.float32)
    return Scene(
        tris=tris, tri_mat=tri_mat, materials=mats,
        point_lights=np.zeros((0, 6), dtype=np.float32), rect_lights=rect,
        film=Film(width, height, time_bins, exposure),
        render=RenderParams(sqrt_spp, n_incoming, max_bounces, seed),
        name=f"icosphere{level}_box")


def media_box(width=512, height=512, scanline=256, time_bins=1024, exposure=0.01, sqrt_spp=64,
              n_incoming=1, max_bounces=16, seed=0, sigma_a=0.05, sigma_s=0.5) -> Scene:
    """Media workload: the CPU-reference box filled with a homogeneous isotropic medium,
    lit by a pulsed point light (emission instant 0); one streak scanline."""
    s = cornell_box(width, height, time_bins, exposure, sqrt_spp, n_incoming, max_bounces, seed)
    s.medium = np.array([sigma_a] * 3 + [sigma_s] * 3, dtype=np.float32)
    s.film.scanline = scanline
    s.name = "media_box"
    return s


def batch_scene_params(seed: int):
    """Draws of one batched-workload job, log-uniform from the job's seed: a triangle target in
    [1e3, 5e5] (met by the nearest icosphere subdivision level, 20 * 4^level triangles), a square
    film of 32..512 px, 128..2048 time bins and 64..1024 samples per pixel."""
    rng = np.random.default_rng(seed)
    target = float(np.exp(rng.uniform(np.log(1e3), np.log(5e5))))
    level = min(range(3, 8), key=lambda l: abs(np.log(20 * 4 ** l) - np.log(target)))
    size = int(round(np.exp(rng.uniform(np.log(32), np.log(512)))))
    bins = int(round(np.exp(rng.uniform(np.log(128), np.log(2048)))))
    spp = float(np.exp(rng.uniform(np.log(64), np.log(1024))))
    return {"triangle_target": target, "level": level, "size": size, "time_bins": bins,
            "sqrt_spp": max(1, int(round(spp ** 0.5)))}


def batch_scenes(n: int = 64, first_seed: int = 0) -> List[Scene]:
    """The batched workload: n uneven icosphere-in-box scenes, seeds first_seed .. first_seed + n - 1."""
    out = []
    for k in range(first_seed, first_seed + n):
        p = batch_scene_params(k)
        s = icosphere_box(level=p["level"], width=p["size"], height=p["size"], time_bins=p["time_bins"],
                          exposure=0.01 * 1024 / p["time_bins"], sqrt_spp=p["sqrt_spp"], n_incoming=1,
                          max_bounces=8, seed=k)
        s.name = f"batch{k}_ico{p['level']}"
        out.append(s)
    return out


def smooth_scene(level=3, width=16, height=16, time_bins=256, exposure=0.01, sqrt_spp=2, n_incoming=1,
                 max_bounces=6, seed=0) -> Scene:
    """A smooth-shaded icosphere (corner normals = the unit sphere's normals, as an OBJ exporter
    writes "vn" records) and a two-sided vertical panel lit from one side, in the open box with
    the point light of the CPU-reference scene."""
    tris, tri_mat, mats = box_shell(inner_box=False)
    centre, radius = np.array([0.35, 0.3, 0.55]), 0.2
    verts, faces = icosphere_mesh(level)
    sph = (centre + radius * verts[faces]).astype(np.float32)
    sph = np.array([_orient(t, None, centre) for t in sph], dtype=np.float32)
    vn = np.zeros((tris.shape[0] + sph.shape[0] + 2, 3, 3), dtype=np.float32)
    nrm = sph - centre.astype(np.float32)
    vn[tris.shape[0]:tris.shape[0] + sph.shape[0]] = nrm / np.linalg.norm(nrm, axis=2, keepdims=True)
    # two-sided panel in the plane x = 0.7, light at x = 0.5 on one side of it
    panel = np.array([[[0.7, 0.05, 0.3], [0.7, 0.05, 0.8], [0.7, 0.6, 0.8]],
                      [[0.7, 0.05, 0.3], [0.7, 0.6, 0.8], [0.7, 0.6, 0.3]]], dtype=np.float32)
    mats = np.concatenate([mats, np.array([[0.7, 0.7, 0.7], [0.5, 0.6, 0.7]], dtype=np.float32)])
    tri_all = np.concatenate([tris, sph, panel])
    mat_all = np.concatenate([tri_mat, np.full(sph.shape[0], len(mats) - 2, np.int32),
                              np.full(2, len(mats) - 1, np.int32)])
    two = np.zeros(len(mats), dtype=bool)
    two[-1] = True
    return Scene(
        tris=tri_all, tri_mat=mat_all.astype(np.int32), materials=mats,
        point_lights=np.array([[0.5, 0.95, 0.5, 10.0, 10.0, 10.0]], dtype=np.float32),
        rect_lights=np.zeros((0, 9), dtype=np.float32), vertex_normals=vn, two_sided=two,
        film=Film(width, height, time_bins, exposure),
        render=RenderParams(sqrt_spp, n_incoming, max_bounces, seed),
        name=f"smooth{level}_box")
