"""Seeded scene generators for the bake workloads.

Each generator returns a SceneSpec: meshes (positions, normals, lightmap uvs, 32-bit indices,
albedo), lights, and the lightmap each mesh is rasterised into. Geometry is built in float64 and
stored as float32; the same arrays feed the GPU library and the oracle. Every planar face is one
UV chart; charts are shelf-packed into their lightmap with a one-texel gutter at a uniform texel
density (the reference's UvGenerator is asset preparation and is not on the bake path).

Workloads:
  box_room   closed 4x3x4 m room, walls 8x8 quads, two cubes (0.8 m, 1.2 m) with 4x4-quad faces,
             one 128x128 lightmap; point light (2, 2.6, 2), intensity 50, radius 10.
  city       200x200 m ground + 5,000 boxes (footprint U(2,12) m, height LogNormal(2.3, 0.6),
             centres uniform, seed 7), 32 tiles = 32 meshes/lightmaps of side round(256*8^U);
             directional sun + a 2x2 m area light of 16 triangles (16 light vertices).
"""
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from .relight import AmbientOcclusionSettings, IndirectLightSettings, LightDesc
from ._capi import RL_PHOTON_RUSSIAN_ROULETTE, RL_SAMPLE_COSINE, RL_SAMPLE_UNIFORM


@dataclass
class MeshSpec:
    positions: np.ndarray
    normals: np.ndarray
    uvs: np.ndarray
    indices: np.ndarray
    lightmap: int
    albedo: Optional[np.ndarray] = None
    photonmap: int = -1

    @property
    def num_tris(self):
        return self.indices.size // 3


@dataclass
class SceneSpec:
    name: str
    meshes: List[MeshSpec]
    lights: List[LightDesc]
    lightmaps: List[Tuple[int, int]]
    ao: Optional[AmbientOcclusionSettings] = None
    params: dict = field(default_factory=dict)
    photonmaps: List[Tuple[int, int]] = field(default_factory=list)
    indirect: Optional[IndirectLightSettings] = None

    @property
    def num_tris(self):
        return sum(m.num_tris for m in self.meshes)


@dataclass
class _Quad:
    """Planar rectangle origin + s*U + t*V, s,t in [0,1], split su x sv quads."""
    origin: np.ndarray
    U: np.ndarray
    V: np.ndarray
    normal: np.ndarray
    su: int
    sv: int
    charted: bool = True
    rect: Tuple[int, int, int, int] = (0, 0, 0, 0)  # x0, y0, w, h texels


def _pack(quads: List[_Quad], W: int, H: int, gutter: int = 1, fill: float = 0.85):
    """Shelf-packs charted quads into W x H at one texel density; sets quad.rect."""
    charted = [q for q in quads if q.charted]
    area = sum(np.linalg.norm(q.U) * np.linalg.norm(q.V) for q in charted)
    d = np.sqrt(W * H * fill / max(area, 1e-12))
    for _ in range(400):
        sizes = [(max(1, int(np.ceil(np.linalg.norm(q.U) * d))),
                  max(1, int(np.ceil(np.linalg.norm(q.V) * d)))) for q in charted]
        order = sorted(range(len(charted)), key=lambda i: (-sizes[i][1], -sizes[i][0], i))
        x, y, shelf_h = gutter, gutter, 0
        ok = True
        rects = {}
        for i in order:
            w, h = sizes[i]
            if x + w + gutter > W:
                x, y = gutter, y + shelf_h + gutter
                shelf_h = 0
            if x + w + gutter > W or y + h + gutter > H:
                ok = False
                break
            rects[i] = (x, y, w, h)
            x += w + gutter
            shelf_h = max(shelf_h, h)
        if ok:
            for i, q in enumerate(charted):
                q.rect = rects[i]
            return d
        d *= 0.97
    raise ValueError("charts do not fit the lightmap")


def _emit(quads: List[_Quad], W: int, H: int):
    pos, nrm, uv, idx = [], [], [], []
    base = 0
    for q in quads:
        x0, y0, cw, ch = q.rect
        for j in range(q.sv + 1):
            for i in range(q.su + 1):
                s, t = i / q.su, j / q.sv
                pos.append(q.origin + q.U * s + q.V * t)
                nrm.append(q.normal)
                if q.charted:
                    uv.append(((x0 + cw * s) / W, (y0 + ch * t) / H))
                else:
                    uv.append((0.0, 0.0))  # no texels: a zero-area chart
        for j in range(q.sv):
            for i in range(q.su):
                a = base + j * (q.su + 1) + i
                b, c, e = a + 1, a + q.su + 2, a + q.su + 1
                idx += [a, b, c, a, c, e]
        base += (q.su + 1) * (q.sv + 1)
    return (np.asarray(pos, np.float64).astype(np.float32),
            np.asarray(nrm, np.float64).astype(np.float32),
            np.asarray(uv, np.float64).astype(np.float32),
            np.asarray(idx, np.uint32))


def _box_quads(lo, hi, sub: int, inward: bool = False, bottom_charted: bool = False):
    """Six faces of an axis-aligned box; normals outward (or inward for a room)."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    ex = hi - lo
    X, Y, Z = np.eye(3)
    faces = [
        # origin, U, V, outward normal
        (lo, Z * ex[2], Y * ex[1], -X),                          # -x
        (lo + X * ex[0], Y * ex[1], Z * ex[2], X),               # +x
        (lo, X * ex[0], Z * ex[2], -Y),                          # -y (bottom)
        (lo + Y * ex[1], Z * ex[2], X * ex[0], Y),               # +y (top)
        (lo, Y * ex[1], X * ex[0], -Z),                          # -z
        (lo + Z * ex[2], X * ex[0], Y * ex[1], Z),               # +z
    ]
    quads = []
    for k, (o, U, V, n) in enumerate(faces):
        if inward:
            n = -n
        charted = True if inward else (k != 2 or bottom_charted)
        quads.append(_Quad(o, U, V, n, sub, sub, charted))
    return quads


def box_room(seed: int = 1) -> SceneSpec:
    """Box room: 4x3x4 m, inward walls of 8x8 quads, cubes of 0.8 m and 1.2 m, 128x128 map."""
    W = H = 128
    quads = _box_quads((0, 0, 0), (4, 3, 4), 8, inward=True)
    quads += _box_quads((0.6, 0.0, 0.7), (1.4, 0.8, 1.5), 4)
    quads += _box_quads((2.3, 0.0, 2.2), (3.5, 1.2, 3.4), 4)
    density = _pack(quads, W, H)
    p, n, uv, ix = _emit(quads, W, H)
    light = LightDesc.point((2.0, 2.6, 2.0), radius=10.0, intensity=50.0)
    ao = AmbientOcclusionSettings(samples=64, occluded_fraction=0.8, max_distance=1.5,
                                  exponent=1.0, seed=seed, sampling=RL_SAMPLE_COSINE)
    return SceneSpec("box_room", [MeshSpec(p, n, uv, ix, 0)], [light], [(W, H)], ao,
                     {"texels_per_m": density, "point_light_radius": 10.0, "seed": seed})


def _area_light_rect(size=2.0, nx=4, nz=2):
    """A size x size rectangle in the local xz plane as nx*nz quads (2*nx*nz triangles)."""
    pts = []
    for j in range(nz + 1):
        for i in range(nx + 1):
            pts.append((-size / 2 + size * i / nx, 0.0, -size / 2 + size * j / nz))
    idx = []
    for j in range(nz):
        for i in range(nx):
            a = j * (nx + 1) + i
            idx += [a, a + 1, a + nx + 2, a, a + nx + 2, a + nx + 1]
    return np.asarray(pts, np.float32), np.asarray(idx, np.uint32)


def city(seed: int = 7, num_boxes: int = 5000, tiles: Tuple[int, int] = (8, 4),
         side_range: Tuple[int, int] = (256, 2048), ao_samples: int = 256,
         lightmap_scale: float = 1.0) -> SceneSpec:
    """Procedural city: 200x200 m ground, `num_boxes` boxes, one mesh + lightmap per tile."""
    rng = np.random.default_rng(seed)
    wx = rng.uniform(2.0, 12.0, num_boxes)
    wz = rng.uniform(2.0, 12.0, num_boxes)
    hy = rng.lognormal(2.3, 0.6, num_boxes)
    cx = rng.uniform(-100.0, 100.0, num_boxes)
    cz = rng.uniform(-100.0, 100.0, num_boxes)
    tx, tz = tiles
    ntiles = tx * tz
    u = rng.uniform(0.0, 1.0, ntiles)
    lo_s, hi_s = side_range
    sides = [max(16, int(round(lo_s * (hi_s / lo_s) ** x * lightmap_scale))) for x in u]
    tile_of = (np.clip(((cx + 100.0) / (200.0 / tx)).astype(int), 0, tx - 1) +
               tx * np.clip(((cz + 100.0) / (200.0 / tz)).astype(int), 0, tz - 1))
    meshes = []
    for t in range(ntiles):
        i, k = t % tx, t // tx
        g0 = np.array([-100.0 + 200.0 / tx * i, 0.0, -100.0 + 200.0 / tz * k])
        quads = [_Quad(g0, np.array([0.0, 0.0, 200.0 / tz]), np.array([200.0 / tx, 0.0, 0.0]),
                       np.array([0.0, 1.0, 0.0]), 1, 1)]
        for b in np.nonzero(tile_of == t)[0]:
            lo = (cx[b] - wx[b] / 2, 0.0, cz[b] - wz[b] / 2)
            hi = (cx[b] + wx[b] / 2, hy[b], cz[b] + wz[b] / 2)
            quads += _box_quads(lo, hi, 1)
        S = sides[t]
        _pack(quads, S, S)
        p, n, uv, ix = _emit(quads, S, S)
        meshes.append(MeshSpec(p, n, uv, ix, t))
    sun_dir = np.array([-0.3, -0.85, -0.2])
    sun_dir = sun_dir / np.linalg.norm(sun_dir)
    sun = LightDesc.directional(tuple(sun_dir), color=(1.0, 0.96, 0.9), intensity=1.0)
    apos, aidx = _area_light_rect()
    area = LightDesc.area(apos, aidx, position=(0.0, 30.0, 0.0), radius=60.0, intensity=40.0)
    ao = AmbientOcclusionSettings(samples=ao_samples, occluded_fraction=0.8, max_distance=4.0,
                                  exponent=1.0, seed=1, sampling=RL_SAMPLE_COSINE)
    return SceneSpec("city", meshes, [sun, area], [(s, s) for s in sides], ao,
                     {"seed": seed, "boxes": num_boxes, "tiles": ntiles, "sides": sides,
                      "sun_direction": sun_dir.tolist(), "area_light_position": [0, 30, 0],
                      "area_light_radius": 60.0})


def box_room_gi(seed: int = 1, photonmap_size: int = 64, photons_per_light: int = 20000,
                fg_samples: int = 64) -> SceneSpec:
    """Box room with coloured albedo, a photonmap and final-gather settings (photons + FG)."""
    spec = box_room(seed)
    m = spec.meshes[0]
    m.albedo = np.array([0.8, 0.7, 0.6, 1.0], np.float32)
    m.photonmap = 0
    spec.photonmaps = [(photonmap_size, photonmap_size)]
    spec.indirect = IndirectLightSettings(
        photon_pass_count=2, photon_bounce_count=3, photon_energy_threshold=0.05,
        photon_max_distance=10.0, final_gather_samples=fg_samples, final_gather_distance=10.0,
        final_gather_radius=3, sky_color=(0.2, 0.3, 0.4), ambient_color=(0.01, 0.01, 0.01),
        seed=seed, sampling=RL_SAMPLE_UNIFORM, photons_per_light=photons_per_light)
    spec.name = "box_room_gi"
    return spec


def build_gpu(relight, spec: SceneSpec):
    """Builds the scene, its lightmaps and photonmaps in the library; returns
    (scene, lightmaps, mesh ids) and sets scene.photonmaps."""
    scene = relight.create_scene()
    scene.begin()
    ids = []
    for m in spec.meshes:
        ids.append(scene.add_mesh(m.positions, m.normals, m.uvs, m.indices, albedo=m.albedo))
    for L in spec.lights:
        scene.add_light(L)
    lms = [relight.create_lightmap(w, h) for (w, h) in spec.lightmaps]
    for i, lm in enumerate(lms):
        lm.set_seed(i)
    pms = [relight.create_photonmap(w, h) for (w, h) in spec.photonmaps]
    for mid, m in zip(ids, spec.meshes):
        if m.lightmap is not None and m.lightmap >= 0:
            lms[m.lightmap].add_mesh(scene, mid)
        if m.photonmap is not None and m.photonmap >= 0:
            pms[m.photonmap].add_mesh(scene, mid)
    scene.end()
    scene.photonmaps = pms
    return scene, lms, ids
