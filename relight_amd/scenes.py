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
  heightfield  512x512-cell fBm terrain (4 octaves of value noise, amplitude 3 m, 0.5 m cells,
             seed 3) in 64 tiles of 64x64 cells (524,288 triangles) + 2,000 one-metre boxes
             (24,000 triangles, no lightmap); one 2048x2048 lightmap over the terrain, a 128x128
             photonmap per tile; 4 rectangular 4x4 m area lights 30 m up; albedo U(0.3, 0.8) per
             mesh; 4M photons (1M per light), 3 bounces, Russian roulette; k = 64 nearest-photon
             gather within 1 m; 512 final-gather rays per texel.
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
    transform: Optional[np.ndarray] = None  # 16 floats, the reference's Matrix4 layout
    texture: Optional[int] = None           # index into SceneSpec.textures (TexturedMaterial)
    diffuse_uvs: Optional[np.ndarray] = None

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
    textures: List[np.ndarray] = field(default_factory=list)  # (h, w, 3 or 4) uint8 pixels

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


def _value_noise(rng, cells: int, period: int) -> np.ndarray:
    """Lattice noise in [-1, 1] with smoothstep-interpolated values, sampled at the
    (cells + 1)^2 grid vertices (rows z, columns x)."""
    g = cells // period + 2
    lat = rng.uniform(-1.0, 1.0, (g, g))
    x = np.arange(cells + 1) / period
    i = np.floor(x).astype(np.int64)
    f = x - i
    f = f * f * (3.0 - 2.0 * f)
    a, b = lat[np.ix_(i, i)], lat[np.ix_(i, i + 1)]
    c, d = lat[np.ix_(i + 1, i)], lat[np.ix_(i + 1, i + 1)]
    fx, fz = f[None, :], f[:, None]
    return (a * (1 - fx) + b * fx) * (1 - fz) + (c * (1 - fx) + d * fx) * fz


def _box_mesh(lo, hi):
    """Closed axis-aligned box, 12 triangles with outward per-face normals, no lightmap charts."""
    quads = _box_quads(lo, hi, 1)
    for q in quads:
        q.charted = False
    p, n, _, ix = _emit(quads, 1, 1)
    return p, n, ix


def heightfield(seed: int = 3, cells: int = 512, cell_size: float = 0.5, tile_cells: int = 64,
                octaves: int = 4, amplitude: float = 3.0, num_boxes: int = 2000, map_size: int = 2048,
                photonmap_size: int = 128, photons_per_light: int = 1_000_000, fg_samples: int = 512,
                knn_k: int = 64, knn_radius: float = 1.0, box_meshes: int = 8,
                tile_map_sides: Optional[Tuple[int, int]] = None) -> SceneSpec:
    """fBm terrain in tiles (one mesh each, one shared lightmap, a photonmap per tile), boxes
    flattened into a few meshes without a lightmap, four rectangular area lights. With
    tile_map_sides = (lo, hi), every tile gets its own lightmap of side round(lo * (hi/lo)^U)."""
    from ._capi import RL_GATHER_KNN
    rng = np.random.default_rng(seed)
    H = np.zeros((cells + 1, cells + 1))
    wsum = 0.0
    for o in range(octaves):
        w = 0.5 ** o
        H += w * _value_noise(rng, cells, max(1, 64 >> o))
        wsum += w
    H = amplitude * H / wsum
    # Per-vertex normals from central differences of the height grid.
    gx = np.gradient(H, cell_size, axis=1)
    gz = np.gradient(H, cell_size, axis=0)
    N = np.stack([-gx, np.ones_like(H), -gz], -1)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    # Lightmap uv: 4 texels per cell, scaled by map/(map + 0.5) and offset by a quarter texel so
    # that no texel centre lies on a cell diagonal (both triangles' strict inside tests would miss it).
    scale = 1.0 / (map_size + 0.5)
    tex_per_cell = map_size / cells
    meshes = []
    nt = cells // tile_cells
    tile_albedo = rng.uniform(0.3, 0.8, (nt * nt, 3))
    sides = None
    if tile_map_sides is not None:
        lo_s, hi_s = tile_map_sides
        sides = [max(16, int(round(lo_s * (hi_s / lo_s) ** x))) for x in rng.uniform(0.0, 1.0, nt * nt)]
    for tj in range(nt):
        for ti in range(nt):
            j0, i0 = tj * tile_cells, ti * tile_cells
            jj, ii = np.meshgrid(np.arange(j0, j0 + tile_cells + 1), np.arange(i0, i0 + tile_cells + 1), indexing="ij")
            pos = np.stack([ii * cell_size, H[jj, ii], jj * cell_size], -1).reshape(-1, 3)
            nrm = N[jj, ii].reshape(-1, 3)
            t = tj * nt + ti
            if sides is None:
                uv = np.stack([(ii * tex_per_cell + 0.25) * scale, (jj * tex_per_cell + 0.25) * scale], -1)
                lm_index = 0
            else:
                tpc, sc = sides[t] / tile_cells, 1.0 / (sides[t] + 0.5)
                uv = np.stack([((ii - i0) * tpc + 0.25) * sc, ((jj - j0) * tpc + 0.25) * sc], -1)
                lm_index = t
            uv = uv.reshape(-1, 2)
            r = tile_cells + 1
            a = (np.arange(tile_cells)[:, None] * r + np.arange(tile_cells)[None, :]).reshape(-1)
            idx = np.stack([a, a + r, a + r + 1, a, a + r + 1, a + 1], -1).reshape(-1)
            meshes.append(MeshSpec(pos.astype(np.float32), nrm.astype(np.float32), uv.astype(np.float32),
                                   idx.astype(np.uint32), lm_index,
                                   np.array([*tile_albedo[t], 1.0], np.float32), photonmap=t))
    # Boxes: 1 m cubes sunk 0.3 m into the terrain at uniform positions.
    ext = cells * cell_size
    bx = rng.uniform(1.0, ext - 1.0, num_boxes)
    bz = rng.uniform(1.0, ext - 1.0, num_boxes)
    by = H[np.clip((bz / cell_size).astype(int), 0, cells), np.clip((bx / cell_size).astype(int), 0, cells)] - 0.3
    box_albedo = rng.uniform(0.3, 0.8, (box_meshes, 3))
    per = (num_boxes + box_meshes - 1) // box_meshes
    for m in range(box_meshes):
        P, Nn, I = [], [], []
        for b in range(m * per, min(num_boxes, (m + 1) * per)):
            p, n, ix = _box_mesh(np.array([bx[b] - 0.5, by[b], bz[b] - 0.5]), np.array([bx[b] + 0.5, by[b] + 1.0, bz[b] + 0.5]))
            I.append(ix + sum(len(q) for q in P))
            P.append(p)
            Nn.append(n)
        if not P:
            continue
        P = np.concatenate(P)
        meshes.append(MeshSpec(P.astype(np.float32), np.concatenate(Nn).astype(np.float32),
                               np.zeros((P.shape[0], 2), np.float32), np.concatenate(I).astype(np.uint32), -1,
                               np.array([*box_albedo[m], 1.0], np.float32)))
    # Four 4x4 m area lights 30 m above the quadrant centres.
    apos, aidx = _area_light_rect(4.0, 4, 4)
    lights = []
    for k, (qx, qz) in enumerate([(0.25, 0.25), (0.75, 0.25), (0.25, 0.75), (0.75, 0.75)]):
        col = [(1.0, 0.95, 0.9), (0.9, 0.95, 1.0), (1.0, 1.0, 0.9), (0.95, 1.0, 0.95)][k]
        lights.append(LightDesc.area(apos, aidx, position=(qx * ext, 30.0, qz * ext), radius=60.0,
                                     color=col, intensity=40.0))
    ind = IndirectLightSettings(
        photon_pass_count=1, photon_bounce_count=3, photon_energy_threshold=0.0, photon_max_distance=100.0,
        final_gather_samples=fg_samples, final_gather_distance=50.0, final_gather_radius=7,
        sky_color=(0.2, 0.25, 0.3), ambient_color=(0.0, 0.0, 0.0), seed=seed, sampling=RL_SAMPLE_UNIFORM,
        photons_per_light=photons_per_light, termination=RL_PHOTON_RUSSIAN_ROULETTE,
        gather_mode=RL_GATHER_KNN, knn_k=knn_k, knn_radius=knn_radius)
    maps = [(map_size, map_size)] if sides is None else [(x, x) for x in sides]
    return SceneSpec("heightfield", meshes, lights, maps, None,
                     {"seed": seed, "cells": cells, "cell_size": cell_size, "octaves": octaves,
                      "amplitude": amplitude, "boxes": num_boxes, "tiles": nt * nt},
                     photonmaps=[(photonmap_size, photonmap_size)] * (nt * nt), indirect=ind)


def with_photonmaps(spec: SceneSpec, size: int = 64, photons_per_light: int = 200_000,
                    fg_samples: int = 64, knn_radius: float = 2.0, seed: int = 5) -> SceneSpec:
    """Gives every lightmapped mesh a photonmap of size x size and an albedo U(0.3, 0.8), and sets
    final-gather settings (Russian roulette, k = 64 nearest-photon gather)."""
    from ._capi import RL_GATHER_KNN
    rng = np.random.default_rng(seed)
    k = 0
    for m in spec.meshes:
        if m.lightmap is not None and m.lightmap >= 0:
            m.albedo = np.array([*rng.uniform(0.3, 0.8, 3), 1.0], np.float32)
            m.photonmap = k
            k += 1
    spec.photonmaps = [(size, size)] * k
    spec.indirect = IndirectLightSettings(
        photon_pass_count=1, photon_bounce_count=3, photon_energy_threshold=0.0, photon_max_distance=200.0,
        final_gather_samples=fg_samples, final_gather_distance=50.0, final_gather_radius=7,
        sky_color=(0.2, 0.25, 0.3), seed=seed, photons_per_light=photons_per_light,
        termination=RL_PHOTON_RUSSIAN_ROULETTE, gather_mode=RL_GATHER_KNN, knn_k=64, knn_radius=knn_radius)
    return spec


def build_gpu(relight, spec: SceneSpec):
    """Builds the scene, its lightmaps and photonmaps in the library; returns
    (scene, lightmaps, mesh ids) and sets scene.photonmaps."""
    scene = relight.create_scene()
    scene.begin()
    ids = []
    textures = [relight.create_texture(px) for px in spec.textures]
    for m in spec.meshes:
        tex = None if m.texture is None else textures[m.texture]
        ids.append(scene.add_mesh(m.positions, m.normals, m.uvs, m.indices, transform=m.transform,
                                  albedo=m.albedo, texture=tex, diffuse_uvs=m.diffuse_uvs))
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
