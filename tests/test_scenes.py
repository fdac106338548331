"""Scene generators: the seeded workloads have the stated sizes and are deterministic."""
import numpy as np

from relight_amd import scenes


def test_box_room_shape():
    s = scenes.box_room()
    assert s.num_tris == 768 + 384
    assert s.lightmaps == [(128, 128)]
    m = s.meshes[0]
    assert m.positions.dtype == np.float32 and m.indices.dtype == np.uint32
    assert m.uvs.min() >= 0 and m.uvs.max() <= 1
    assert s.ao.samples == 64 and s.ao.max_distance == 1.5


def test_city_shape_and_determinism():
    a = scenes.city()
    assert len(a.meshes) == 32 and len(a.lightmaps) == 32
    assert a.num_tris == 5000 * 12 + 32 * 2
    sides = [w for w, _ in a.lightmaps]
    assert min(sides) >= 256 and max(sides) <= 2048
    b = scenes.city()
    for ma, mb in zip(a.meshes, b.meshes):
        assert np.array_equal(ma.positions, mb.positions) and np.array_equal(ma.uvs, mb.uvs)
    assert a.params["sides"] == b.params["sides"]
