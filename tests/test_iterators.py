"""BakeIterator and Progress semantics (baker/Baker.cpp:52-77,148-234; Relight.h:96-105) through
the *_iter entry points: worker-split iterators (i, N) together equal the whole bake bit for bit; a
single iterator bakes exactly the texels (LumelBakeIterator) or the faces' lumels
(FaceBakeIterator) it names, equal to the oracle there and untouched elsewhere; progress reports
increasing steps and ends with (itemCount, itemCount)."""
import numpy as np
import pytest

import oracle as O
from relight_amd import Relight, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def relight():
    r = Relight()
    yield r
    r.close()


def _oracle_full(spec, direct=True, ao=None, fg=False):
    S = O.OracleScene.from_spec(spec)
    texel, pos, nrm = S.lumels(spec, 0)
    col = np.zeros((texel.size, 3), np.float32)
    if fg:
        maps, _ = S.emit_photons(spec, spec.indirect)
        S.final_gather(spec, pos, nrm, texel, 0, spec.indirect, maps, col)
    if direct:
        S.bake_direct(pos, nrm, col)
    if ao is not None:
        S.bake_ao(pos, nrm, texel, 0, ao, col)
    valid, tri, _, _ = S.raster(0, *spec.lightmaps[0])
    return texel, col, tri


@pytest.mark.parametrize("kind", ["lumels", "faces"])
def test_worker_split_equals_whole_bake(relight, kind):
    print(f"[gpu] iterators: worker split ({kind})", flush=True)
    spec = scenes.box_room()
    whole, lw, iw = scenes.build_gpu(relight, spec)
    relight.bake_direct_light(whole, iw[0])
    relight.bake_ambient_occlusion(whole, iw[0], spec.ao)
    split, ls, is_ = scenes.build_gpu(relight, spec)
    N = 3
    for i in range(N):
        relight.bake_direct_light(split, is_[0], iterator=(kind, i, N))
    for i in range(N):
        relight.bake_ambient_occlusion(split, is_[0], spec.ao, iterator=(kind, i, N))
    assert np.array_equal(lw[0].to_hdr(), ls[0].to_hdr())
    texel, col, _ = _oracle_full(spec, True, spec.ao)
    assert np.array_equal(ls[0].to_hdr().reshape(-1, 3)[texel], col)


@pytest.mark.parametrize("kind,first,step", [("lumels", 1, 3), ("lumels", 5000, 7), ("faces", 2, 5),
                                             ("faces", 0, 1000)])
def test_single_iterator_bakes_exactly_its_lumels(relight, kind, first, step):
    print(f"[gpu] iterators: single {kind} ({first}, {step})", flush=True)
    spec = scenes.box_room()
    scene, lms, ids = scenes.build_gpu(relight, spec)
    relight.bake_direct_light(scene, ids[0], iterator=(kind, first, step))
    texel, col, tri = _oracle_full(spec, True, None)
    if kind == "lumels":
        want = (texel >= first) & ((texel.astype(np.int64) - first) % step == 0)
    else:
        face = tri[texel].astype(np.int64)  # one mesh: scene-order triangle = face index
        want = (face >= first) & ((face - first) % step == 0)
    got = lms[0].to_hdr().reshape(-1, 3)[texel]
    assert want.any() and not want.all()
    assert np.array_equal(got[want], col[want])
    assert (got[~want] == 0).all()
    done, total = relight.progress()
    assert done == total == int(want.sum())


def test_iterated_final_gather_and_progress(relight):
    print("[gpu] iterators: final gather by faces with progress", flush=True)
    spec = scenes.box_room_gi(photons_per_light=10000, fg_samples=32)
    whole, lw, iw = scenes.build_gpu(relight, spec)
    relight.emit_photons(whole, spec.indirect)
    relight.bake_indirect_light(whole, iw[0], spec.indirect)
    split, ls, is_ = scenes.build_gpu(relight, spec)
    relight.emit_photons(split, spec.indirect)
    calls = []
    for i in range(2):
        calls.append([])
        relight.bake_indirect_light(split, is_[0], spec.indirect, iterator=("faces", i, 2),
                                    progress=lambda m, s, c, k=i: calls[k].append((m, s, c)), updates=4)
    assert np.array_equal(lw[0].to_hdr(), ls[0].to_hdr())
    faces = spec.meshes[0].num_tris
    for seq in calls:
        # Three intermediate notifications (4 parts), steps rising, then (faces, faces).
        assert len(seq) == 4 and seq[-1] == (is_[0], faces, faces)
        steps = [s for _, s, _ in seq[:-1]]
        assert steps == sorted(steps) and all(0 < s <= (faces + 1) // 2 for s in steps)
        assert all(c == faces for _, _, c in seq)
    texel, col, _ = _oracle_full(spec, False, None, fg=True)
    assert np.array_equal(ls[0].to_hdr().reshape(-1, 3)[texel], col)
