"""ORACLE — test infrastructure only (never shipped, never measured as the product).

ctypes wrapper of liboracle.so, the CPU restatement of the reference's bake math (oracle.cpp
holds the file:line citations). Used by tests/, by the smoke check and by bench.py's CPU-baseline
leg, only as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class or_light(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("casts_shadow", C.c_int32),
        ("position", C.c_float * 3), ("direction", C.c_float * 3), ("color", C.c_float * 3),
        ("intensity", C.c_float), ("has_attenuation", C.c_int32),
        ("radius", C.c_float), ("att_constant", C.c_float), ("att_linear", C.c_float),
        ("att_quadratic", C.c_float), ("spot_cutoff", C.c_float), ("spot_exponent", C.c_float),
        ("area_positions", C.POINTER(C.c_float)), ("area_indices", C.POINTER(C.c_uint32)),
        ("area_vertex_count", C.c_uint32), ("area_index_count", C.c_uint32),
    ]


class or_ao(C.Structure):
    _fields_ = [("samples", C.c_int32), ("occluded_fraction", C.c_float),
                ("max_distance", C.c_float), ("exponent", C.c_float), ("seed", C.c_uint32),
                ("sampling", C.c_int32)]


class or_indirect(C.Structure):
    _fields_ = [("photon_pass_count", C.c_int32), ("photon_bounce_count", C.c_int32),
                ("photon_energy_threshold", C.c_float), ("photon_max_distance", C.c_float),
                ("final_gather_samples", C.c_int32), ("final_gather_distance", C.c_float),
                ("final_gather_radius", C.c_int32), ("sky_color", C.c_float * 3),
                ("ambient_color", C.c_float * 3), ("seed", C.c_uint32), ("sampling", C.c_int32),
                ("photons_per_light", C.c_int32), ("termination", C.c_int32),
                ("gather_mode", C.c_int32), ("knn_k", C.c_int32), ("knn_radius", C.c_float)]


def _indirect(s):
    o = or_indirect()
    for name, _ in or_indirect._fields_:
        v = getattr(s, name, 0)
        if name in ("sky_color", "ambient_color"):
            getattr(o, name)[:] = [float(x) for x in v]
        else:
            setattr(o, name, v)
    return o


def photon_layout(spec):
    """Photonmaps in first-appearance mesh order (the library's order) and mesh -> map index."""
    order, mesh_pm = [], []
    for m in spec.meshes:
        k = getattr(m, "photonmap", -1)
        if k is not None and k >= 0:
            if k not in order:
                order.append(k)
            mesh_pm.append(order.index(k))
        else:
            mesh_pm.append(-1)
    return order, np.asarray(mesh_pm, np.int32)


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)


_lib = None
_F = C.POINTER(C.c_float)
_U32 = C.POINTER(C.c_uint32)
_U8 = C.POINTER(C.c_uint8)
_I32 = C.POINTER(C.c_int32)
_U64 = C.POINTER(C.c_uint64)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "or_scene_create": (C.c_void_p, []),
            "or_scene_destroy": (None, [C.c_void_p]),
            "or_scene_add_mesh": (C.c_int32, [C.c_void_p, _F, _F, _F, C.c_uint32, _U32, C.c_uint32, _F, _F]),
            "or_scene_add_light": (C.c_int32, [C.c_void_p, C.POINTER(or_light)]),
            "or_scene_add_texture": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _U8]),
            "or_scene_add_mesh_textured": (C.c_int32, [C.c_void_p, _F, _F, _F, _F, C.c_uint32, _U32, C.c_uint32,
                                                       _F, _F, C.c_int32]),
            "or_texture_color": (None, [C.c_void_p, C.c_int32, C.c_float, C.c_float, _F]),
            "or_generate_uv": (C.c_int32, [_F, _F, _F, C.c_uint32, _U32, C.c_uint32, C.c_float, C.c_float, _U32,
                                           _U32]),
            "or_generate_uv_get": (None, [_F, _F, _F, _F, _U32]),
            "or_face_color": (None, [C.c_void_p, C.c_uint32, C.c_float, C.c_float, _F]),
            "or_scene_build": (C.c_int32, [C.c_void_p, C.c_int32]),
            "or_raster": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _U8, _U32, _F, _F]),
            "or_bake_direct": (C.c_int32, [C.c_void_p, C.c_uint32, _F, _F, _F, C.c_int32, _U64]),
            "or_bake_ao": (C.c_int32, [C.c_void_p, C.c_uint32, _F, _F, _U32, C.c_uint32, C.POINTER(or_ao),
                                       _F, _U32, C.c_int32, _U64]),
            "or_trace": (C.c_int32, [C.c_void_p, C.c_uint32, _F, _F, C.c_int32, _U8, _F, _I32, _F, _F,
                                     _U64, _U64, C.c_int32]),
            "or_tri_info": (C.c_int32, [C.c_void_p, C.c_int32, _I32, _I32]),
            "or_pcg32": (None, [C.c_uint64, C.c_uint64, C.c_uint32, _U32]),
            "or_stream_u32": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _U32]),
            "or_hemisphere": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, _F, _F]),
            "or_splitmix64": (C.c_uint64, [C.c_uint64]),
            "or_emit_photons": (C.c_int32, [C.c_void_p, C.POINTER(or_indirect), _I32, C.c_int32, _I32, _I32, _U8,
                                            _F, _F, _U32, _F, C.c_int32, _U64]),
            "or_photon_records": (C.c_int32, [C.c_void_p, _U64, _U32, _F, _F]),
            "or_expand": (C.c_int32, [C.c_int32, C.c_int32, _U8, _F]),
            "or_blur": (C.c_int32, [C.c_int32, C.c_int32, _U8, _F]),
            "or_encode": (C.c_int32, [C.c_int32, C.c_int32, _F, C.c_int32, _U8]),
            "or_final_gather": (C.c_int32, [C.c_void_p, C.c_uint32, _F, _F, _U32, C.c_uint32, C.POINTER(or_indirect),
                                            _I32, C.c_int32, _I32, _I32, _F, _F, C.c_int32, _U64]),
            "or_form_factors": (C.c_int32, [C.c_void_p, _F, _U32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float,
                                            C.c_uint32, C.c_uint32, C.POINTER(C.c_uint16), C.c_int32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _f(a):
    return None if a is None else a.ctypes.data_as(_F)


def _u(a):
    return None if a is None else a.ctypes.data_as(_U32)


def pcg32(initstate, initseq, n):
    out = np.zeros(n, np.uint32)
    lib().or_pcg32(initstate, initseq, n, _u(out))
    return out


def stream_u32(seed, lm_seed, texel, tech, sample, n):
    out = np.zeros(n, np.uint32)
    lib().or_stream_u32(seed, lm_seed, texel, tech, sample, n, _u(out))
    return out


def hemisphere(seed, lm_seed, texel, tech, count, mode, normal):
    out = np.zeros((count, 3), np.float32)
    nrm = np.ascontiguousarray(normal, np.float32)
    lib().or_hemisphere(seed, lm_seed, texel, tech, count, mode, _f(nrm), _f(out))
    return out


class OracleScene:
    """The oracle's scene: same meshes and lights as the library, its own binary BVH."""

    def __init__(self, use_bvh=True):
        self._L = lib()
        self._h = self._L.or_scene_create()
        self._keep = []
        self.use_bvh = use_bvh
        self.meshes = 0

    def __del__(self):
        try:
            if self._h:
                self._L.or_scene_destroy(self._h)
        except Exception:
            pass

    def add_mesh(self, positions, normals, uvs, indices, transform=None, albedo=None):
        p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        n = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        uv = np.ascontiguousarray(uvs, np.float32).reshape(-1, 2)
        ix = np.ascontiguousarray(indices, np.uint32).reshape(-1)
        tr = None if transform is None else np.ascontiguousarray(transform, np.float32).reshape(16)
        al = None if albedo is None else np.ascontiguousarray(albedo, np.float32).reshape(4)
        self._keep += [p, n, uv, ix, tr, al]
        self.meshes += 1
        return self._L.or_scene_add_mesh(self._h, _f(p), _f(n), _f(uv), p.shape[0], _u(ix), ix.size,
                                         _f(tr), _f(al))

    def add_texture(self, pixels):
        """Texture::create from (h, w, channels) uint8 pixels; returns the texture index."""
        px = np.ascontiguousarray(pixels, np.uint8)
        self._keep.append(px)
        return self._L.or_scene_add_texture(self._h, px.shape[1], px.shape[0], px.shape[2], px.ctypes.data_as(_U8))

    def add_mesh_textured(self, positions, normals, uvs, diffuse_uvs, indices, texture, color=None, transform=None):
        p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        n = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        uv = np.ascontiguousarray(uvs, np.float32).reshape(-1, 2)
        duv = np.ascontiguousarray(diffuse_uvs, np.float32).reshape(-1, 2)
        ix = np.ascontiguousarray(indices, np.uint32).reshape(-1)
        tr = None if transform is None else np.ascontiguousarray(transform, np.float32).reshape(16)
        col = np.ascontiguousarray([1.0, 1.0, 1.0] if color is None else np.asarray(color)[:3], np.float32)
        self._keep += [p, n, uv, duv, ix, tr, col]
        self.meshes += 1
        return self._L.or_scene_add_mesh_textured(self._h, _f(p), _f(n), _f(uv), _f(duv), p.shape[0], _u(ix), ix.size,
                                                  _f(tr), _f(col), int(texture))

    def texture_color(self, texture, u, v):
        out = np.zeros(4, np.float32)
        self._L.or_texture_color(self._h, int(texture), C.c_float(u), C.c_float(v), _f(out))
        return out

    def face_color(self, tri, u, v):
        out = np.zeros(4, np.float32)
        self._L.or_face_color(self._h, int(tri), C.c_float(u), C.c_float(v), _f(out))
        return out

    def add_light(self, L):
        d = or_light()
        d.type = L.type
        d.casts_shadow = 1 if L.casts_shadow else 0
        d.position[:] = [float(x) for x in L.position]
        d.direction[:] = [float(x) for x in L.direction]
        d.color[:] = [float(x) for x in L.color]
        d.intensity = L.intensity
        d.has_attenuation = 1 if L.has_attenuation else 0
        d.radius, d.att_constant, d.att_linear, d.att_quadratic = (
            L.radius, L.att_constant, L.att_linear, L.att_quadratic)
        d.spot_cutoff, d.spot_exponent = L.spot_cutoff, L.spot_exponent
        if L.area_positions is not None:
            ap = np.ascontiguousarray(L.area_positions, np.float32)
            ai = np.ascontiguousarray(L.area_indices, np.uint32)
            self._keep += [ap, ai]
            d.area_positions = _f(ap)
            d.area_indices = _u(ai)
            d.area_vertex_count = ap.shape[0]
            d.area_index_count = ai.size
        return self._L.or_scene_add_light(self._h, C.byref(d))

    def build(self):
        self._L.or_scene_build(self._h, 1 if self.use_bvh else 0)

    def raster(self, mesh, w, h, grids=None):
        """Lightmap::initializeLumels of one mesh into (valid, tri, pos, nrm) grids."""
        if grids is None:
            grids = (np.zeros(w * h, np.uint8), np.full(w * h, 0xFFFFFFFF, np.uint32),
                     np.zeros((w * h, 3), np.float32), np.zeros((w * h, 3), np.float32))
        valid, tri, pos, nrm = grids
        self._L.or_raster(self._h, mesh, w, h, valid.ctypes.data_as(_U8), _u(tri), _f(pos), _f(nrm))
        return grids

    def bake_direct(self, pos, nrm, color, threads=0):
        pos = np.ascontiguousarray(pos, np.float32)
        nrm = np.ascontiguousarray(nrm, np.float32)
        assert color.dtype == np.float32 and color.flags.c_contiguous
        rays = C.c_uint64()
        self._L.or_bake_direct(self._h, pos.shape[0], _f(pos), _f(nrm), _f(color), threads, C.byref(rays))
        return rays.value

    def bake_ao(self, pos, nrm, texel, lm_seed, ao, color, threads=0):
        pos = np.ascontiguousarray(pos, np.float32)
        nrm = np.ascontiguousarray(nrm, np.float32)
        texel = np.ascontiguousarray(texel, np.uint32)
        s = or_ao(int(ao.samples), float(ao.occluded_fraction), float(ao.max_distance), float(ao.exponent),
                  int(ao.seed), int(ao.sampling))
        occ = np.zeros(pos.shape[0], np.uint32)
        rays = C.c_uint64()
        self._L.or_bake_ao(self._h, pos.shape[0], _f(pos), _f(nrm), _u(texel), lm_seed, C.byref(s), _f(color),
                           _u(occ), threads, C.byref(rays))
        return occ, rays.value

    def trace(self, starts, ends, closest=False, threads=0):
        s = np.ascontiguousarray(starts, np.float32).reshape(-1, 3)
        e = np.ascontiguousarray(ends, np.float32).reshape(-1, 3)
        n = s.shape[0]
        flags = np.zeros(n, np.uint8)
        t = np.zeros(n, np.float32)
        tri = np.zeros(n, np.int32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        nv, nt = C.c_uint64(), C.c_uint64()
        self._L.or_trace(self._h, n, _f(s), _f(e), 1 if closest else 0, flags.ctypes.data_as(_U8), _f(t),
                         tri.ctypes.data_as(_I32), _f(u), _f(v), C.byref(nv), C.byref(nt), threads)
        return {"flags": flags, "t": t, "triangle": tri, "u": u, "v": v,
                "node_visits": nv.value, "tri_tests": nt.value}

    @classmethod
    def from_spec(cls, spec, use_bvh=True):
        S = cls(use_bvh)
        tex = [S.add_texture(px) for px in getattr(spec, "textures", [])]
        for m in spec.meshes:
            if getattr(m, "texture", None) is not None:
                S.add_mesh_textured(m.positions, m.normals, m.uvs, m.diffuse_uvs, m.indices, tex[m.texture],
                                    color=None if m.albedo is None else m.albedo[:3],
                                    transform=getattr(m, "transform", None))
            else:
                S.add_mesh(m.positions, m.normals, m.uvs, m.indices, transform=getattr(m, "transform", None),
                           albedo=m.albedo)
        for L in spec.lights:
            S.add_light(L)
        S.build()
        return S

    def emit_photons(self, spec, settings, threads=0):
        """Relight::emitPhotons: returns {spec photonmap index: (color, photons, gathered)} and
        the number of photon rays."""
        order, mesh_pm = photon_layout(spec)
        ws = np.asarray([spec.photonmaps[k][0] for k in order], np.int32)
        hs = np.asarray([spec.photonmaps[k][1] for k in order], np.int32)
        valid, ppos = [], []
        for k in order:
            w, h = spec.photonmaps[k]
            grids = None
            for mi, m in enumerate(spec.meshes):
                if getattr(m, "photonmap", -1) == k:
                    grids = self.raster(mi, w, h, grids)
            valid.append(grids[0])
            ppos.append(grids[2])
        n = int(sum(int(w) * int(h) for w, h in zip(ws, hs)))
        valid = np.concatenate(valid) if valid else np.zeros(1, np.uint8)
        ppos = np.ascontiguousarray(np.concatenate(ppos) if ppos else np.zeros((1, 3)), np.float32)
        color = np.zeros(max(n, 1) * 3, np.float32)
        photons = np.zeros(max(n, 1), np.uint32)
        gathered = np.zeros(max(n, 1) * 3, np.float32)
        s = _indirect(settings)
        rays = C.c_uint64()
        self._L.or_emit_photons(self._h, C.byref(s), mesh_pm.ctypes.data_as(_I32), len(order),
                                ws.ctypes.data_as(_I32), hs.ctypes.data_as(_I32), valid.ctypes.data_as(_U8),
                                _f(ppos), _f(color), _u(photons), _f(gathered), threads, C.byref(rays))
        out, base = {}, 0
        for k, w, h in zip(order, ws, hs):
            sz = int(w) * int(h)
            out[k] = (color[3 * base:3 * (base + sz)].reshape(h, w, 3),
                      photons[base:base + sz].reshape(h, w),
                      gathered[3 * base:3 * (base + sz)].reshape(h, w, 3))
            base += sz
        return out, rays.value

    def photon_records(self):
        """Records of the last emit_photons call in slot order: (key, power (n,3), hit point (n,3));
        key 0xFFFFFFFF marks an empty slot."""
        n = C.c_uint64()
        self._L.or_photon_records(self._h, C.byref(n), None, None, None)
        key = np.zeros(n.value, np.uint32)
        power = np.zeros((n.value, 3), np.float32)
        pos = np.zeros((n.value, 3), np.float32)
        self._L.or_photon_records(self._h, C.byref(n), _u(key), _f(power), _f(pos))
        return key, power, pos

    def final_gather(self, spec, pos, nrm, texel, lm_seed, settings, maps, color, threads=0):
        """IndirectLight::bakeLumel over lumels (color in/out); maps from emit_photons."""
        order, mesh_pm = photon_layout(spec)
        ws = np.asarray([spec.photonmaps[k][0] for k in order], np.int32)
        hs = np.asarray([spec.photonmaps[k][1] for k in order], np.int32)
        g = (np.concatenate([maps[k][2].reshape(-1) for k in order]).astype(np.float32)
             if order else np.zeros(3, np.float32))
        pos = np.ascontiguousarray(pos, np.float32)
        nrm = np.ascontiguousarray(nrm, np.float32)
        texel = np.ascontiguousarray(texel, np.uint32)
        assert color.dtype == np.float32 and color.flags.c_contiguous
        s = _indirect(settings)
        rays = C.c_uint64()
        self._L.or_final_gather(self._h, pos.shape[0], _f(pos), _f(nrm), _u(texel), lm_seed, C.byref(s),
                                mesh_pm.ctypes.data_as(_I32), len(order), ws.ctypes.data_as(_I32),
                                hs.ctypes.data_as(_I32), _f(g), _f(color), threads, C.byref(rays))
        return rays.value

    def form_factors(self, patches, tri_patch, rays, seed, max_distance, rows=None, row0=0, threads=0):
        """Hit counts of form-factor rows [row0, row0+rows) (u16, rows x num_patches)."""
        p = np.ascontiguousarray(patches, np.float32).reshape(-1, 12)
        t = np.ascontiguousarray(tri_patch, np.uint32)
        n = p.shape[0]
        rows = n - row0 if rows is None else rows
        out = np.zeros((rows, n), np.uint16)
        self._L.or_form_factors(self._h, _f(p), _u(t), n, rays, seed, max_distance, row0, rows,
                                out.ctypes.data_as(C.POINTER(C.c_uint16)), threads)
        return out

    def lumels(self, spec, lightmap):
        """Valid lumels of one lightmap in texel order: texel, pos, nrm (rasterising its meshes
        in mesh order, as Lightmap::addMesh calls would)."""
        w, h = spec.lightmaps[lightmap]
        grids = None
        for mi, m in enumerate(spec.meshes):
            if m.lightmap == lightmap:
                grids = self.raster(mi, w, h, grids)
        valid, tri, pos, nrm = grids
        texel = np.nonzero(valid)[0].astype(np.uint32)
        return texel, pos[texel].copy(), nrm[texel].copy()


def generate_uv(positions, normals, indices, angle=88.0, padding=0.0, diffuse_uvs=None):
    """Mesh::generateUv(angle, margin, padding): (positions, normals, diffuse uvs, lightmap uvs,
    indices) of the regenerated mesh."""
    p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    n = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    ix = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    d = None if diffuse_uvs is None else np.ascontiguousarray(diffuse_uvs, np.float32).reshape(-1, 2)
    nv, ni = C.c_uint32(), C.c_uint32()
    lib().or_generate_uv(_f(p), _f(n), _f(d), p.shape[0], _u(ix), ix.size, angle, padding, C.byref(nv),
                         C.byref(ni))
    op = np.zeros((nv.value, 3), np.float32)
    on = np.zeros((nv.value, 3), np.float32)
    od = np.zeros((nv.value, 2), np.float32)
    ol = np.zeros((nv.value, 2), np.float32)
    oi = np.zeros(ni.value, np.uint32)
    lib().or_generate_uv_get(_f(op), _f(on), _f(od), _f(ol), _u(oi))
    return op, on, od, ol, oi


def expand(valid, rgb):
    """Lightmap::expand on an (h, w) validity grid and an (h, w, 3) float32 colour grid (copy)."""
    h, w = valid.shape
    out = np.ascontiguousarray(rgb, np.float32).copy()
    v = np.ascontiguousarray(valid, np.uint8)
    lib().or_expand(w, h, v.ctypes.data_as(_U8), _f(out))
    return out


def blur(valid, rgb):
    """Lightmap::blur (in place, raster order) on a copy."""
    h, w = valid.shape
    out = np.ascontiguousarray(rgb, np.float32).copy()
    v = np.ascontiguousarray(valid, np.uint8)
    lib().or_blur(w, h, v.ctypes.data_as(_U8), _f(out))
    return out


def encode(rgb, fmt):
    """toDoubleLdr (fmt 1, h x w x 3) / toRgbmLdr (fmt 2, h x w x 4) bytes."""
    h, w = rgb.shape[:2]
    c = np.ascontiguousarray(rgb, np.float32)
    out = np.zeros((h, w, 3 if fmt == 1 else 4), np.uint8)
    lib().or_encode(w, h, _f(c), int(fmt), out.ctypes.data_as(_U8))
    return out


def tga_bytes(pixels):
    """The file Lightmap::save writes for TGA formats: 12-byte header, 6-byte info, b g r (a)."""
    h, w, ch = pixels.shape
    head = bytes([0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0]) + bytes([w % 256, w // 256, h % 256, h // 256, ch * 8, 0])
    px = pixels.reshape(-1, ch)
    order = [2, 1, 0] + ([3] if ch == 4 else [])
    return head + np.ascontiguousarray(px[:, order]).tobytes()
