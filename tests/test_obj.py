"""OBJ ingestion with MeshLoader::load semantics (scene/MeshLoader.cpp:44-226): the library's
loader (rl_obj_load, C++) against the oracle's restatement on the reference's own OBJ files and on
hand-written cases whose expected arrays are computed by hand."""
import ctypes as C
import os

import numpy as np
import pytest

from relight_amd import _capi as A
from oracle import objload

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def lib_load(path):
    lib = A.load()
    nv, ni = C.c_uint32(), C.c_uint32()
    st = lib.rl_obj_load(str(path).encode(), C.byref(nv), C.byref(ni), None, None, None, None)
    if st != 0:
        raise ValueError(lib.rl_last_error().decode())
    p = np.zeros((nv.value, 3), np.float32)
    n = np.zeros((nv.value, 3), np.float32)
    uv = np.zeros((nv.value, 2), np.float32)
    ix = np.zeros(ni.value, np.uint32)
    F = C.POINTER(C.c_float)
    assert lib.rl_obj_load(str(path).encode(), C.byref(nv), C.byref(ni), p.ctypes.data_as(F), n.ctypes.data_as(F),
                           uv.ctypes.data_as(F), ix.ctypes.data_as(C.POINTER(C.c_uint32))) == 0
    return p, n, uv, ix


needs_lib = pytest.mark.skipif(not os.path.exists(A.LIB_PATH), reason="librelight_amd.so not built")


@needs_lib
@pytest.mark.parametrize("name", ["box_uv.obj", "boxes_uv.obj"])
def test_reference_obj_files_match_oracle(name):
    path = os.path.join(GOLD, name)
    got = lib_load(path)
    exp = objload.load(path)
    for g, e in zip(got, exp):
        assert g.shape == e.shape and np.array_equal(g, e)
    p, n, uv, ix = got
    assert ix.size % 3 == 0 and ix.max() < p.shape[0]
    assert (uv >= 0).all() and (uv <= 1).all()


def _write(tmp_path, text):
    f = tmp_path / "m.obj"
    f.write_text(text)
    return f


@needs_lib
def test_hand_cases(tmp_path):
    # No uv/normals: "v v v" faces, normal (0, 1, 0), uv (0, 0); a quad keeps its first triangle.
    f = _write(tmp_path, "v 0 0 0\nv 1 0 0\nv 0 0 1\nv 1 0 1\nf 1 2 3 4\n")
    p, n, uv, ix = lib_load(f)
    assert p.tolist() == [[0, 0, 0], [1, 0, 0], [0, 0, 1]] and ix.tolist() == [0, 1, 2]
    assert n.tolist() == [[0, 1, 0]] * 3 and uv.tolist() == [[0, 0]] * 3
    # uv clamped; shared corners deduplicated; a v line after faces commits a submesh whose flags
    # are then cleared (the next faces use the plain format).
    f = _write(tmp_path, "v 0 0 0\nv 1 0 0\nv 0 0 1\nvt -0.5 1.5\nvt 0.25 0.75\nvn 0 1 0\n"
                         "f 1/1/1 2/2/1 3/2/1\nf 1/1/1 3/2/1 2/2/1\nv 2 0 0\nf 2 3 4\n")
    p, n, uv, ix = lib_load(f)
    assert uv[:3].tolist() == [[0.0, 1.0], [0.25, 0.75], [0.25, 0.75]]
    assert ix[:6].tolist() == [0, 1, 2, 0, 2, 1]
    assert ix[6:].tolist() == [3, 4, 5] and p[3:].tolist() == [[1, 0, 0], [0, 0, 1], [2, 0, 0]]
    assert n[3:].tolist() == [[0, 1, 0]] * 3
    for g, e in zip((p, n, uv, ix), objload.load(f)):
        assert np.array_equal(g, e)


@needs_lib
def test_malformed_files_are_rejected(tmp_path):
    with pytest.raises(ValueError, match="out of range"):
        lib_load(_write(tmp_path, "v 0 0 0\nv 1 0 0\nf 1 2 3\n"))
    with pytest.raises(ValueError, match="cannot open"):
        lib_load(tmp_path / "missing.obj")


@pytest.mark.gpu
def test_gpu_bake_of_obj_box_matches_oracle():
    """The reference's box_uv.obj through rl_scene_add_obj, baked (direct + AO) on the GPU, against
    the oracle scene built from the oracle's parse of the same file."""
    import oracle as O
    from relight_amd import AmbientOcclusionSettings, LightDesc, Relight, RL_SAMPLE_COSINE
    print("[gpu] OBJ box bake", flush=True)
    path = os.path.join(GOLD, "box_uv.obj")
    r = Relight()
    scene = r.create_scene()
    scene.begin()
    mid = C.c_int32()
    assert r._lib.rl_scene_add_obj(scene._h, path.encode(), None, None, C.byref(mid)) == 0
    light = LightDesc.point((0.0, 2.5, 0.0), 10.0, intensity=5.0)
    scene.add_light(light)
    lm = r.create_lightmap(64, 64)
    lm.set_seed(0)
    lm.add_mesh(scene, mid.value)
    scene.end()
    ao = AmbientOcclusionSettings(32, 0.8, 2.0, 1.0, 1, RL_SAMPLE_COSINE)
    r.bake_direct_light(scene, mid.value)
    r.bake_ambient_occlusion(scene, mid.value, ao)
    gpu = lm.to_hdr().reshape(-1, 3)
    p, n, uv, ix = objload.load(path)
    S = O.OracleScene()
    S.add_mesh(p, n, uv, ix)
    S.add_light(light)
    S.build()
    valid, _, pos, nrm = S.raster(0, 64, 64)
    texel = np.nonzero(valid)[0].astype(np.uint32)
    col = np.zeros((texel.size, 3), np.float32)
    S.bake_direct(pos[texel], nrm[texel], col)
    S.bake_ao(pos[texel], nrm[texel], texel, 0, ao, col)
    assert texel.size > 1000
    assert np.array_equal(gpu[texel], col)
    r.close()
