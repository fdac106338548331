"""Python mirror of the reference's Relight / Scene / Lightmap API over librelight_amd.so.

Names, argument meaning and error behaviour follow src/relight/Relight.h:107-189,
scene/Scene.h and Lightmap.h; every call goes through the C-ABI of include/relight.h, and every
bake runs on the GPU. A non-success status raises RelightError carrying the RelightStatus-style
code (Relight.h:82-86) and the library's message.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _capi as A


class RelightError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{A.STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


def _check(lib, status: int):
    if status != A.RL_SUCCESS:
        raise RelightError(status, lib.rl_last_error().decode())


def _fptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _uptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint32))


@dataclass
class AmbientOcclusionSettings:
    """AmbientOcclusionSettings (Relight.h:135-152) + sample-stream seed and sampling mode."""
    samples: int = 128
    occluded_fraction: float = 0.8
    max_distance: float = 0.6
    exponent: float = 1.0
    seed: int = 1
    sampling: int = A.RL_SAMPLE_UNIFORM

    # Presets of Relight.cpp:123-172.
    @classmethod
    def draft(cls, occluded_fraction=0.8, max_distance=0.6, exponent=1.0):
        return cls(128, occluded_fraction, max_distance, exponent)

    @classmethod
    def fast(cls, occluded_fraction=0.8, max_distance=0.6, exponent=1.0):
        return cls(256, occluded_fraction, max_distance, exponent)

    @classmethod
    def best(cls, occluded_fraction=0.8, max_distance=0.6, exponent=1.0):
        return cls(1024, occluded_fraction, max_distance, exponent)

    @classmethod
    def production(cls, occluded_fraction=0.8, max_distance=0.6, exponent=1.0):
        return cls(2048, occluded_fraction, max_distance, exponent)

    def to_c(self) -> A.rl_ao_settings:
        return A.rl_ao_settings(int(self.samples), float(self.occluded_fraction),
                                float(self.max_distance), float(self.exponent),
                                int(self.seed) & 0xFFFFFFFF, int(self.sampling))


@dataclass
class IndirectLightSettings:
    """IndirectLightSettings (Relight.h:108-132) + seed, sampling, photon count override,
    termination rule and gather rule (texel disc, or k nearest photons in 3D)."""
    photon_pass_count: int = 16
    photon_bounce_count: int = 3
    photon_energy_threshold: float = 0.05
    photon_max_distance: float = 10.0
    final_gather_samples: int = 64
    final_gather_distance: float = 50.0
    final_gather_radius: int = 7
    sky_color: Sequence[float] = (0.0, 0.0, 0.0)
    ambient_color: Sequence[float] = (0.0, 0.0, 0.0)
    seed: int = 1
    sampling: int = A.RL_SAMPLE_UNIFORM
    photons_per_light: int = 0
    termination: int = A.RL_PHOTON_THRESHOLD
    gather_mode: int = A.RL_GATHER_TEXEL_DISC
    knn_k: int = 64
    knn_radius: float = 1.0

    # Presets of Relight.cpp:42-120: (passes, bounces, final-gather samples).
    _PRESETS = {"draft": (8, 1, 32), "fast": (16, 3, 64), "best": (32, 3, 128),
                "production": (64, 4, 1024)}

    @classmethod
    def _preset(cls, name, sky, ambient, photon_max_distance, final_gather_distance):
        p, b, s = cls._PRESETS[name]
        return cls(p, b, 0.05, photon_max_distance, s, final_gather_distance, 7, tuple(sky),
                   tuple(ambient))

    @classmethod
    def draft(cls, sky=(0, 0, 0), ambient=(0, 0, 0), photon_max_distance=10.0, final_gather_distance=50.0):
        return cls._preset("draft", sky, ambient, photon_max_distance, final_gather_distance)

    @classmethod
    def fast(cls, sky=(0, 0, 0), ambient=(0, 0, 0), photon_max_distance=10.0, final_gather_distance=50.0):
        return cls._preset("fast", sky, ambient, photon_max_distance, final_gather_distance)

    @classmethod
    def best(cls, sky=(0, 0, 0), ambient=(0, 0, 0), photon_max_distance=10.0, final_gather_distance=50.0):
        return cls._preset("best", sky, ambient, photon_max_distance, final_gather_distance)

    @classmethod
    def production(cls, sky=(0, 0, 0), ambient=(0, 0, 0), photon_max_distance=10.0,
                   final_gather_distance=50.0):
        return cls._preset("production", sky, ambient, photon_max_distance, final_gather_distance)

    def to_c(self) -> A.rl_indirect_settings:
        s = A.rl_indirect_settings()
        s.photon_pass_count = int(self.photon_pass_count)
        s.photon_bounce_count = int(self.photon_bounce_count)
        s.photon_energy_threshold = float(self.photon_energy_threshold)
        s.photon_max_distance = float(self.photon_max_distance)
        s.final_gather_samples = int(self.final_gather_samples)
        s.final_gather_distance = float(self.final_gather_distance)
        s.final_gather_radius = int(self.final_gather_radius)
        s.sky_color[:] = [float(x) for x in self.sky_color]
        s.ambient_color[:] = [float(x) for x in self.ambient_color]
        s.seed = int(self.seed) & 0xFFFFFFFF
        s.sampling = int(self.sampling)
        s.photons_per_light = int(self.photons_per_light)
        s.termination = int(self.termination)
        s.gather_mode = int(self.gather_mode)
        s.knn_k = int(self.knn_k)
        s.knn_radius = float(self.knn_radius)
        return s


@dataclass
class LightDesc:
    """A light as Light::create* builds it (scene/Light.cpp:166-234)."""
    type: int
    position: Sequence[float] = (0.0, 0.0, 0.0)
    direction: Sequence[float] = (0.0, -1.0, 0.0)
    color: Sequence[float] = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    casts_shadow: bool = True
    has_attenuation: bool = True
    radius: float = 1.0
    att_constant: float = 0.0
    att_linear: float = 0.0
    att_quadratic: float = 25.0   # LinearLightAttenuation defaults, scene/Light.h:125
    spot_cutoff: float = 0.0
    spot_exponent: float = 1.0
    area_positions: Optional[np.ndarray] = None   # (V, 3) local to position
    area_indices: Optional[np.ndarray] = None     # (T*3,)

    @classmethod
    def point(cls, position, radius, color=(1, 1, 1), intensity=1.0, casts_shadow=True):
        return cls(A.RL_LIGHT_POINT, position=position, radius=radius, color=color,
                   intensity=intensity, casts_shadow=casts_shadow)

    @classmethod
    def spot(cls, position, direction, cutoff, radius, color=(1, 1, 1), intensity=1.0,
             casts_shadow=True):
        return cls(A.RL_LIGHT_SPOT, position=position, direction=direction, spot_cutoff=cutoff,
                   radius=radius, color=color, intensity=intensity, casts_shadow=casts_shadow)

    @classmethod
    def directional(cls, direction, color=(1, 1, 1), intensity=1.0, casts_shadow=True):
        return cls(A.RL_LIGHT_DIRECTIONAL, direction=direction, color=color, intensity=intensity,
                   casts_shadow=casts_shadow, has_attenuation=False)

    @classmethod
    def area(cls, positions, indices, position, radius, color=(1, 1, 1), intensity=1.0,
             casts_shadow=True):
        return cls(A.RL_LIGHT_AREA, position=position, radius=radius, color=color,
                   intensity=intensity, casts_shadow=casts_shadow,
                   area_positions=np.ascontiguousarray(positions, np.float32),
                   area_indices=np.ascontiguousarray(indices, np.uint32))

    def to_c(self) -> A.rl_light_desc:
        d = A.rl_light_desc()
        d.type = int(self.type)
        d.casts_shadow = 1 if self.casts_shadow else 0
        d.position[:] = [float(x) for x in self.position]
        d.direction[:] = [float(x) for x in self.direction]
        d.color[:] = [float(x) for x in self.color]
        d.intensity = float(self.intensity)
        d.has_attenuation = 1 if self.has_attenuation else 0
        d.radius = float(self.radius)
        d.att_constant = float(self.att_constant)
        d.att_linear = float(self.att_linear)
        d.att_quadratic = float(self.att_quadratic)
        d.spot_cutoff = float(self.spot_cutoff)
        d.spot_exponent = float(self.spot_exponent)
        if self.area_positions is not None:
            d.area_positions = _fptr(self.area_positions)
            d.area_indices = _uptr(self.area_indices)
            d.area_vertex_count = int(self.area_positions.shape[0])
            d.area_index_count = int(self.area_indices.size)
        return d


def generate_uv(positions, normals, indices, angle=88.0, margin=0.0, padding=0.0, diffuse_uvs=None):
    """Mesh::generateUv (scene/Mesh.cpp:229-238) through rl_mesh_generate_uv: returns the mesh
    (positions, normals, diffuse uvs, lightmap uvs, indices) with a generated lightmap uv set."""
    lib = A.load()
    p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    n = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    ix = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    d = None if diffuse_uvs is None else np.ascontiguousarray(diffuse_uvs, np.float32).reshape(-1, 2)
    nv, ni = C.c_uint32(), C.c_uint32()
    args = (_fptr(p), _fptr(n), _fptr(d), p.shape[0], _uptr(ix), ix.size, float(angle), float(margin), float(padding))
    _check(lib, lib.rl_mesh_generate_uv(*args, C.byref(nv), C.byref(ni), None, None, None, None, None))
    op = np.zeros((nv.value, 3), np.float32)
    on = np.zeros((nv.value, 3), np.float32)
    od = np.zeros((nv.value, 2), np.float32)
    ol = np.zeros((nv.value, 2), np.float32)
    oi = np.zeros(ni.value, np.uint32)
    _check(lib, lib.rl_mesh_generate_uv(*args, C.byref(nv), C.byref(ni), _fptr(op), _fptr(on), _fptr(od), _fptr(ol),
                                        _uptr(oi)))
    return op, on, od, ol, oi


class Texture:
    """Texture::create (scene/Material.h:95): (h, w, channels) uint8 pixels, channels 3 or 4."""

    def __init__(self, relight: "Relight", pixels):
        px = np.ascontiguousarray(pixels, np.uint8)
        if px.ndim != 3 or px.shape[2] not in (3, 4):
            raise ValueError("pixels must be (height, width, 3 or 4) uint8")
        self._lib = relight._lib
        h = C.c_void_p()
        _check(self._lib, self._lib.rl_texture_create(relight._h, px.shape[1], px.shape[0], px.shape[2],
                                                      px.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            self._lib.rl_texture_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """Scene (scene/Scene.h): begin -> add_mesh/add_light -> end (BVH build + upload)."""

    def __init__(self, relight: "Relight"):
        self._relight = relight
        self._lib = relight._lib
        h = C.c_void_p()
        _check(self._lib, self._lib.rl_scene_create(relight._h, C.byref(h)))
        self._h = h
        self._keep = []

    def close(self):
        """Raises RelightError (and keeps the handle) while an unfinished server job uses it."""
        if self._h:
            _check(self._lib, self._lib.rl_scene_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def begin(self):
        _check(self._lib, self._lib.rl_scene_begin(self._h))

    def add_mesh(self, positions, normals, lightmap_uvs, indices, transform=None,
                 albedo=None, texture: Optional["Texture"] = None, diffuse_uvs=None) -> int:
        """Scene::addMesh with a Material (albedo rgba) or, when `texture` is given, a
        TexturedMaterial(texture, albedo[:3]) sampled at `diffuse_uvs`."""
        p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        n = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        uv = np.ascontiguousarray(lightmap_uvs, np.float32).reshape(-1, 2)
        ix = np.ascontiguousarray(indices, np.uint32).reshape(-1)
        tr = None if transform is None else np.ascontiguousarray(transform, np.float32).reshape(16)
        out = C.c_int32()
        if texture is not None:
            duv = np.ascontiguousarray(diffuse_uvs, np.float32).reshape(-1, 2)
            if duv.shape[0] != p.shape[0]:
                raise ValueError("diffuse_uvs needs one uv per vertex")
            col = None if albedo is None else np.ascontiguousarray(np.asarray(albedo)[:3], np.float32)
            _check(self._lib, self._lib.rl_scene_add_mesh_textured(
                self._h, _fptr(p), _fptr(n), _fptr(uv), _fptr(duv), p.shape[0], _uptr(ix), ix.size,
                _fptr(tr), _fptr(col), texture._h, C.byref(out)))
            return out.value
        al = None if albedo is None else np.ascontiguousarray(albedo, np.float32).reshape(4)
        _check(self._lib, self._lib.rl_scene_add_mesh(
            self._h, _fptr(p), _fptr(n), _fptr(uv), p.shape[0], _uptr(ix), ix.size, _fptr(tr),
            _fptr(al), C.byref(out)))
        return out.value

    def add_light(self, light: LightDesc) -> int:
        d = light.to_c()
        self._keep.append(light)
        out = C.c_int32()
        _check(self._lib, self._lib.rl_scene_add_light(self._h, C.byref(d), C.byref(out)))
        return out.value

    def end(self):
        _check(self._lib, self._lib.rl_scene_end(self._h))

    def info(self):
        t, n, s = C.c_uint32(), C.c_uint32(), C.c_int32()
        _check(self._lib, self._lib.rl_scene_info(self._h, C.byref(t), C.byref(n), C.byref(s)))
        return {"triangles": t.value, "nodes": n.value, "stack_bound": s.value}

    def trace_segments(self, starts, ends, mode=A.RL_TRACE_ANY):
        """rt::ITracer::traceSegment over arrays of segments (N, 3) -> flags or hit records."""
        s = np.ascontiguousarray(starts, np.float32).reshape(-1, 3)
        e = np.ascontiguousarray(ends, np.float32).reshape(-1, 3)
        n = s.shape[0]
        if mode == A.RL_TRACE_ANY:
            flags = np.zeros(n, np.uint8)
            _check(self._lib, self._lib.rl_trace_segments(
                self._relight._h, self._h, _fptr(s), _fptr(e), n, mode,
                flags.ctypes.data_as(C.POINTER(C.c_uint8)), None))
            return flags
        hits = (A.rl_hit * max(n, 1))()
        _check(self._lib, self._lib.rl_trace_segments(
            self._relight._h, self._h, _fptr(s), _fptr(e), n, mode, None, hits))
        arr = np.ctypeslib.as_array(hits)[:n]
        return {k: np.array(arr[k]) for k in ("t", "triangle", "mesh", "face", "u", "v")}


class Lightmap:
    """Lightmap (Lightmap.h:58-150): texel raster at add_mesh, HDR readback."""

    def __init__(self, relight: "Relight", width: int, height: int):
        self._lib = relight._lib
        self.width, self.height = int(width), int(height)
        h = C.c_void_p()
        _check(self._lib, self._lib.rl_lightmap_create(relight._h, self.width, self.height,
                                                       C.byref(h)))
        self._h = h

    def close(self):
        """Raises RelightError (and keeps the handle) while an unfinished server job uses it."""
        if self._h:
            _check(self._lib, self._lib.rl_lightmap_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_mesh(self, scene: Scene, mesh: int):
        _check(self._lib, self._lib.rl_lightmap_add_mesh(self._h, scene._h, int(mesh)))

    def set_seed(self, seed: int):
        _check(self._lib, self._lib.rl_lightmap_set_seed(self._h, int(seed) & 0xFFFFFFFF))

    def valid_count(self) -> int:
        c = C.c_uint32()
        _check(self._lib, self._lib.rl_lightmap_valid_count(self._h, C.byref(c)))
        return c.value

    def lumels(self):
        n = self.valid_count()
        texel = np.zeros(n, np.uint32)
        pos = np.zeros((n, 3), np.float32)
        nrm = np.zeros((n, 3), np.float32)
        _check(self._lib, self._lib.rl_lightmap_read_lumels(self._h, _uptr(texel), _fptr(pos),
                                                            _fptr(nrm)))
        return texel, pos, nrm

    def expand(self):
        """Lightmap::expand: invalid texels take the colour of a valid diagonal neighbour."""
        _check(self._lib, self._lib.rl_lightmap_expand(self._h))

    def blur(self):
        """Lightmap::blur: in-place 3x3 mean over valid texels, in raster order."""
        _check(self._lib, self._lib.rl_lightmap_blur(self._h))

    def encode(self, fmt: int) -> np.ndarray:
        """toDoubleLdr (fmt RL_FORMAT_TGA_DOUBLE_LDR, h x w x 3) or toRgbmLdr (RL_FORMAT_TGA_RGBM,
        h x w x 4) bytes."""
        ch = 3 if fmt == A.RL_FORMAT_TGA_DOUBLE_LDR else 4
        out = np.zeros((self.height, self.width, ch), np.uint8)
        _check(self._lib, self._lib.rl_lightmap_encode(self._h, int(fmt), out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    def save(self, path: str, fmt: int = A.RL_FORMAT_RAW_HDR):
        """Lightmap::save as RawHdr or TGA (double LDR / RGBM)."""
        _check(self._lib, self._lib.rl_lightmap_save(self._h, str(path).encode(), int(fmt)))

    def to_hdr(self) -> np.ndarray:
        """Lightmap::toHdr: (height, width, 3) float32, invalid texels 0."""
        out = np.zeros((self.height, self.width, 3), np.float32)
        _check(self._lib, self._lib.rl_lightmap_read_hdr(self._h, _fptr(out)))
        return out


class Photonmap(Lightmap):
    """Photonmap (Lightmap.h:153-170): receives photon stores; gathered colour per texel."""

    def __init__(self, relight: "Relight", width: int, height: int):
        self._lib = relight._lib
        self.width, self.height = int(width), int(height)
        h = C.c_void_p()
        _check(self._lib, self._lib.rl_photonmap_create(relight._h, self.width, self.height,
                                                        C.byref(h)))
        self._h = h

    def add_mesh(self, scene: Scene, mesh: int):
        _check(self._lib, self._lib.rl_photonmap_add_mesh(self._h, scene._h, int(mesh)))

    def read(self):
        """(color (h, w, 3), photons (h, w), gathered (h, w, 3))."""
        c = np.zeros((self.height, self.width, 3), np.float32)
        p = np.zeros((self.height, self.width), np.uint32)
        g = np.zeros((self.height, self.width, 3), np.float32)
        _check(self._lib, self._lib.rl_photonmap_read(self._h, _fptr(c), _uptr(p), _fptr(g)))
        return c, p, g


class Radiosity:
    """Radiosity patches, GPU form factors and the BF16-MFMA gather solve
    (radiosity/RadiosityBuilder.cpp:44-176, radiosity/RadiositySolver.cpp:34-36)."""

    def __init__(self, relight: "Relight", scene: Scene, patches, tri_patch):
        self._relight = relight
        self._lib = relight._lib
        self.scene = scene
        p = np.ascontiguousarray(patches, np.float32).reshape(-1, 12)
        t = np.ascontiguousarray(tri_patch, np.uint32).reshape(-1)
        self.num_patches = p.shape[0]
        h = C.c_void_p()
        _check(self._lib, self._lib.rl_radiosity_create(relight._h, scene._h, _fptr(p), p.shape[0],
                                                        _uptr(t), C.byref(h)))
        self._h = h

    def close(self):
        """Raises RelightError (and keeps the handle) while an unfinished server job uses it."""
        if self._h:
            _check(self._lib, self._lib.rl_radiosity_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_form_factors(self, rays_per_patch=1024, seed=1, max_distance=1000.0):
        _check(self._lib, self._lib.rl_radiosity_build_form_factors(
            self._relight._h, self._h, int(rays_per_patch), int(seed) & 0xFFFFFFFF, float(max_distance)))

    def form_factor_counts(self, first=0, rows=None) -> np.ndarray:
        rows = self.num_patches - first if rows is None else rows
        out = np.zeros((rows, self.num_patches), np.uint16)
        _check(self._lib, self._lib.rl_radiosity_read_form_factors(
            self._h, int(first), int(rows), out.ctypes.data_as(C.POINTER(C.c_uint16))))
        return out

    def solve(self, emission, reflectance, tolerance=1e-5, max_outer=20, inner_sweeps=16,
              check=True):
        """B = E + rho o (F B); emission/reflectance (num_patches, num_rhs). Returns (B, info)."""
        E = np.ascontiguousarray(emission, np.float32)
        R = np.ascontiguousarray(reflectance, np.float32)
        assert E.shape == R.shape and E.shape[0] == self.num_patches
        out = np.zeros_like(E)
        info = A.rl_solve_info()
        st = self._lib.rl_radiosity_solve(self._relight._h, self._h, _fptr(E), _fptr(R), E.shape[1],
                                          float(tolerance), int(max_outer), int(inner_sweeps),
                                          _fptr(out), C.byref(info))
        res = {"outer_iterations": info.outer_iterations, "sweeps": info.sweeps,
               "residual": info.residual, "gpu_ms": info.gpu_ms, "status": st}
        if check and st != A.RL_SUCCESS:
            _check(self._lib, st)
        return out, res


class Relight:
    """Relight (Relight.h:155-189): factories and bake entry points, all on the GPU."""

    def __init__(self):
        self._lib = A.load()
        h = C.c_void_p()
        _check(self._lib, self._lib.rl_create(C.byref(h)))
        self._h = h

    @classmethod
    def create(cls) -> "Relight":
        return cls()

    def close(self):
        if self._h:
            self._lib.rl_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def create_scene(self) -> Scene:
        return Scene(self)

    def create_lightmap(self, width: int, height: int) -> Lightmap:
        return Lightmap(self, width, height)

    def create_texture(self, pixels) -> Texture:
        return Texture(self, pixels)

    def create_photonmap(self, width: int, height: int) -> Photonmap:
        return Photonmap(self, width, height)

    def create_radiosity(self, scene: Scene, patches, tri_patch) -> Radiosity:
        return Radiosity(self, scene, patches, tri_patch)

    def tuning(self) -> dict:
        t = A.rl_tuning()
        _check(self._lib, self._lib.rl_get_tuning(self._h, C.byref(t)))
        return {name: getattr(t, name) for name, _ in A.rl_tuning._fields_}

    def set_tuning(self, **kw):
        """Launch tuning (ao_kernel, refill_min, ao_chunk_lumels, blocks_per_cu, fg_kernel,
        direct_kernel, local_trees, occluder_hints); results do not depend on it. The refill AO
        kernel uses chunks of at most 16 lumels, halved down to 4 for passes with fewer chunks
        than it has resident wavefronts."""
        cur = self.tuning()
        cur.update(kw)
        t = A.rl_tuning(*[int(cur[name]) for name, _ in A.rl_tuning._fields_])
        _check(self._lib, self._lib.rl_set_tuning(self._h, C.byref(t)))

    def emit_photons(self, scene: Scene, settings: IndirectLightSettings):
        s = settings.to_c()
        _check(self._lib, self._lib.rl_emit_photons(self._h, scene._h, C.byref(s)))

    @staticmethod
    def _iter_args(iterator, progress, updates):
        """(rl_bake_iterator or None, rl_progress or None, keep-alive) for the *_iter entry points.
        iterator: ("lumels" | "faces", first, step); progress: callable(mesh, step, step_count)."""
        it = None
        if iterator is not None:
            kind, first, step = iterator
            it = A.rl_bake_iterator({"lumels": A.RL_ITERATE_LUMELS, "faces": A.RL_ITERATE_FACES}[kind],
                                    int(first), int(step))
        pr, fn = None, None
        if progress is not None:
            fn = A.rl_progress_fn(lambda user, mesh, step, count: progress(mesh, step, count))
            pr = A.rl_progress(fn, None, int(updates))
        return (C.byref(it) if it is not None else None, C.byref(pr) if pr is not None else None, (it, pr, fn))

    def bake_indirect_light(self, scene: Scene, mesh: int, settings: IndirectLightSettings, iterator=None,
                            progress=None, updates: int = 8):
        """Relight::bakeIndirectLight (Relight.cpp:230-241); iterator ("lumels"|"faces", first,
        step) and progress callback(mesh, step, step_count) as rl_bake_indirect_iter."""
        s = settings.to_c()
        if iterator is None and progress is None:
            _check(self._lib, self._lib.rl_bake_indirect(self._h, scene._h, int(mesh), C.byref(s)))
            return
        it, pr, keep = self._iter_args(iterator, progress, updates)
        _check(self._lib, self._lib.rl_bake_indirect_iter(self._h, scene._h, int(mesh), C.byref(s), it, pr))

    def bake_direct_light(self, scene: Scene, mesh: int, iterator=None, progress=None, updates: int = 8):
        if iterator is None and progress is None:
            _check(self._lib, self._lib.rl_bake_direct(self._h, scene._h, int(mesh)))
            return
        it, pr, keep = self._iter_args(iterator, progress, updates)
        _check(self._lib, self._lib.rl_bake_direct_iter(self._h, scene._h, int(mesh), it, pr))

    def bake_ambient_occlusion(self, scene: Scene, mesh: int, settings: AmbientOcclusionSettings, iterator=None,
                               progress=None, updates: int = 8):
        s = settings.to_c()
        if iterator is None and progress is None:
            _check(self._lib, self._lib.rl_bake_ambient_occlusion(self._h, scene._h, int(mesh), C.byref(s)))
            return
        it, pr, keep = self._iter_args(iterator, progress, updates)
        _check(self._lib, self._lib.rl_bake_ambient_occlusion_iter(self._h, scene._h, int(mesh), C.byref(s), it,
                                                                   pr))

    def progress(self):
        """(done, total) lumels of the bake call running on this context (pollable from another
        thread)."""
        d, t = C.c_uint64(), C.c_uint64()
        _check(self._lib, self._lib.rl_get_progress(self._h, C.byref(d), C.byref(t)))
        return d.value, t.value

    def bake_batch(self, scene: Scene, lightmaps: List[Lightmap], direct: bool = True,
                   ao: Optional[AmbientOcclusionSettings] = None):
        arr = (C.c_void_p * len(lightmaps))(*[lm._h.value for lm in lightmaps])
        s = ao.to_c() if ao is not None else None
        _check(self._lib, self._lib.rl_bake_batch(self._h, scene._h, arr, len(lightmaps),
                                                  1 if direct else 0,
                                                  C.byref(s) if s is not None else None))

    def record_pass(self, scene: Scene, lightmaps: List[Lightmap], direct: bool = True,
                    ao: Optional[AmbientOcclusionSettings] = None) -> "Pass":
        """The bake_batch pass recorded once as a HIP graph; Pass.launch() replays it."""
        return Pass(self, scene, lightmaps, direct, ao)

    def bake_items(self, items):
        """One multi-scene launch per pass over (scene, lightmap, first, count, direct, ao) items
        (count None: to the end of the lightmap); equals the per-lightmap bakes."""
        keep = []
        arr = (A.rl_bake_item * len(items))()
        for k, (scene, lm, first, count, direct, ao) in enumerate(items):
            s = ao.to_c() if ao is not None else None
            keep.append(s)
            arr[k] = A.rl_bake_item(scene._h.value, lm._h.value, int(first),
                                    0xFFFFFFFF if count is None else int(count), 1 if direct else 0,
                                    C.pointer(s) if s is not None else None)
        _check(self._lib, self._lib.rl_bake_items(self._h, arr, len(items)))

    def bake(self, scene: Scene, direct: bool = True, ao: Optional[AmbientOcclusionSettings] = None,
             indirect: Optional[IndirectLightSettings] = None):
        """Relight::bake with a FullBakeJob that bakes direct light, indirect light and AO over every
        lightmap of the scene: photons, direct, final gather, then AO, one launch per pass."""
        a = ao.to_c() if ao is not None else None
        i = indirect.to_c() if indirect is not None else None
        _check(self._lib, self._lib.rl_bake_scene(self._h, scene._h, 1 if direct else 0,
                                                  C.byref(a) if a is not None else None,
                                                  C.byref(i) if i is not None else None))

    def create_server(self, memory_budget=64 << 30, max_active=16, quantum_rays=1 << 28) -> "Server":
        return Server(self, memory_budget, max_active, quantum_rays)

    def stats(self) -> dict:
        st = A.rl_stats()
        _check(self._lib, self._lib.rl_get_stats(self._h, C.byref(st)))
        kernels = {name: {"ms": st.kernel_ms[i], "launches": st.kernel_launches[i]}
                   for i, name in A.KERNEL_NAMES.items() if st.kernel_launches[i]}
        return {"rays": st.rays, "texels": st.texels, "gpu_ms": st.gpu_ms,
                "wall_ms": st.wall_ms, "launches": st.launches, "kernels": kernels}


class Pass:
    """A recorded bake_batch pass (rl_pass_create): launch() replays it with one graph launch and
    equals the same bake_batch call bit for bit. Close it before its scene and lightmaps."""

    def __init__(self, relight: "Relight", scene: Scene, lightmaps: List[Lightmap], direct: bool,
                 ao: Optional[AmbientOcclusionSettings]):
        self._relight = relight
        self._lib = relight._lib
        self.scene, self.lightmaps = scene, list(lightmaps)
        arr = (C.c_void_p * len(lightmaps))(*[lm._h.value for lm in lightmaps])
        s = ao.to_c() if ao is not None else None
        h = C.c_void_p()
        _check(self._lib, self._lib.rl_pass_create(relight._h, scene._h, arr, len(lightmaps), 1 if direct else 0,
                                                   C.byref(s) if s is not None else None, C.byref(h)))
        self._h = h

    def launch(self):
        _check(self._lib, self._lib.rl_pass_launch(self._relight._h, self._h))

    def close(self):
        if self._h:
            self._lib.rl_pass_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BakeJob:
    """A job submitted to a Server. It holds what the native job borrows until it finishes: the
    scene, the lightmaps, the radiosity object, the settings and the radiosity output array (the
    server also keeps the job until then), so dropping the caller's references cannot free them
    under the server's worker threads."""

    _FINAL = ("done", "failed", "cancelled")

    def __init__(self, server: "Server", job_id: int, kind: int, out=None, keep=()):
        self.server = server
        self.id = job_id
        self.kind = kind
        self.result = out
        self._keep = keep

    def _settle(self, status: dict) -> dict:
        if status["state"] in self._FINAL:
            self.server._jobs.pop(self.id, None)
        return status

    @staticmethod
    def _status(st: A.rl_job_status) -> dict:
        return {"state": A.JOB_STATE_NAMES[st.state], "status": st.status, "submit_ms": st.submit_ms,
                "admit_ms": st.admit_ms, "done_ms": st.done_ms, "rays": st.rays,
                "footprint": st.footprint}

    def poll(self) -> dict:
        st = A.rl_job_status()
        _check(self.server._lib, self.server._lib.rl_server_poll(self.server._h, self.id, C.byref(st)))
        return self._settle(self._status(st))

    def wait(self, timeout: Optional[float] = None) -> dict:
        """Blocks until the job finishes (timeout in seconds); raises RelightError if it failed
        and TimeoutError if it is still running at the timeout."""
        lib = self.server._lib
        st = A.rl_job_status()
        rc = lib.rl_server_wait(self.server._h, self.id, -1.0 if timeout is None else 1e3 * timeout,
                                C.byref(st))
        if rc == A.RL_REJECTED and st.state in (A.RL_JOB_QUEUED, A.RL_JOB_RUNNING):
            raise TimeoutError(f"job {self.id} still {A.JOB_STATE_NAMES[st.state]}")
        _check(lib, rc)
        status = self._settle(self._status(st))
        if st.state == A.RL_JOB_FAILED:
            raise RelightError(st.status, lib.rl_last_error().decode())
        return status

    def cancel(self):
        _check(self.server._lib, self.server._lib.rl_server_cancel(self.server._h, self.id))


class Server:
    """Continuous-batching bake server (no counterpart in the reference; it replaces the
    Worker/ThreadWorker job loop of src/relight/Worker.cpp for serving). A native thread admits
    queued jobs under a device-memory budget and bakes the direct/AO work of every admitted job
    in shared launches of quantum_rays rays; results equal the same job baked alone. While the
    server runs, make no other call on its Relight context."""

    def __init__(self, relight: Relight, memory_budget=64 << 30, max_active=16, quantum_rays=1 << 28):
        self._relight = relight
        self._lib = relight._lib
        cfg = A.rl_server_config(int(memory_budget), int(max_active), int(quantum_rays))
        h = C.c_void_p()
        _check(self._lib, self._lib.rl_server_create(relight._h, C.byref(cfg), C.byref(h)))
        self._h = h
        self._jobs = {}  # id -> BakeJob, until the job is seen finished

    def close(self):
        """Cancels queued jobs, finishes running ones and stops the loop."""
        if self._h:
            self._lib.rl_server_destroy(self._h)
            self._h = None
            self._jobs.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _submit(self, d: A.rl_job_desc, out=None, keep=()) -> BakeJob:
        if len(self._jobs) >= 256:
            for job in list(self._jobs.values()):
                job.poll()  # drops finished jobs
        jid = C.c_uint64()
        _check(self._lib, self._lib.rl_server_submit(self._h, C.byref(d), C.byref(jid)))
        job = BakeJob(self, jid.value, d.kind, out, keep)
        self._jobs[job.id] = job
        return job

    @staticmethod
    def _maps(lightmaps):
        return (C.c_void_p * len(lightmaps))(*[lm._h.value for lm in lightmaps])

    def submit_direct_ao(self, scene: Scene, lightmaps: Sequence[Lightmap], direct: bool = True,
                         ao: Optional[AmbientOcclusionSettings] = None) -> BakeJob:
        maps = self._maps(lightmaps)
        s = ao.to_c() if ao is not None else None
        d = A.rl_job_desc(kind=A.RL_JOB_DIRECT_AO, scene=scene._h.value, lightmaps=maps,
                          num_lightmaps=len(lightmaps), direct=1 if direct else 0,
                          ao=C.pointer(s) if s is not None else None)
        return self._submit(d, keep=(scene, list(lightmaps)))

    def submit_indirect(self, scene: Scene, lightmaps: Sequence[Lightmap],
                        settings: IndirectLightSettings) -> BakeJob:
        maps = self._maps(lightmaps)
        s = settings.to_c()
        d = A.rl_job_desc(kind=A.RL_JOB_INDIRECT, scene=scene._h.value, lightmaps=maps,
                          num_lightmaps=len(lightmaps), indirect=C.pointer(s))
        return self._submit(d, keep=(scene, list(lightmaps)))

    def submit_final_gather(self, scene: Scene, lightmaps: Sequence[Lightmap],
                            settings: IndirectLightSettings) -> BakeJob:
        """Final gather from the scene's photonmaps as they are (photons emitted earlier)."""
        maps = self._maps(lightmaps)
        s = settings.to_c()
        d = A.rl_job_desc(kind=A.RL_JOB_FINAL_GATHER, scene=scene._h.value, lightmaps=maps,
                          num_lightmaps=len(lightmaps), indirect=C.pointer(s))
        return self._submit(d, keep=(scene, list(lightmaps)))

    def submit_full_bake(self, scene: Scene, lightmaps: Sequence[Lightmap], direct: bool = True,
                         ao: Optional[AmbientOcclusionSettings] = None,
                         indirect: Optional[IndirectLightSettings] = None) -> BakeJob:
        """Photons (indirect), direct, final gather (indirect), then AO over the lightmaps."""
        maps = self._maps(lightmaps)
        a = ao.to_c() if ao is not None else None
        i = indirect.to_c() if indirect is not None else None
        d = A.rl_job_desc(kind=A.RL_JOB_FULL_BAKE, scene=scene._h.value, lightmaps=maps,
                          num_lightmaps=len(lightmaps), direct=1 if direct else 0,
                          ao=C.pointer(a) if a is not None else None,
                          indirect=C.pointer(i) if i is not None else None)
        return self._submit(d, keep=(scene, list(lightmaps)))

    def submit_radiosity(self, radiosity: Radiosity, emission, reflectance, rays_per_patch=1024,
                         seed=1, max_distance=1000.0, tolerance=1e-5) -> BakeJob:
        E = np.ascontiguousarray(emission, np.float32)
        R = np.ascontiguousarray(reflectance, np.float32)
        assert E.shape == R.shape and E.shape[0] == radiosity.num_patches
        out = np.zeros_like(E)
        d = A.rl_job_desc(kind=A.RL_JOB_RADIOSITY, scene=radiosity.scene._h.value,
                          radiosity=radiosity._h.value, rays_per_patch=int(rays_per_patch),
                          ff_seed=int(seed) & 0xFFFFFFFF, ff_max_distance=float(max_distance),
                          emission=_fptr(E), reflectance=_fptr(R), num_rhs=E.shape[1],
                          tolerance=float(tolerance), out_radiosity=_fptr(out))
        return self._submit(d, out, keep=(radiosity, radiosity.scene))

    def stats(self) -> dict:
        st = A.rl_server_stats()
        _check(self._lib, self._lib.rl_server_stats_get(self._h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in A.rl_server_stats._fields_}
