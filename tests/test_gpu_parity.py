"""GPU parity: librelight_amd.so (HIP, gfx950) against the CPU oracle on identical inputs.

Bars: lumel setup, direct light and AO bit-exact (per-texel relative RMSE reported, <= 1e-3
required); segment queries hit/miss and closest-triangle agreement >= 99.9%; batched bake
bit-identical to per-lightmap bakes.
"""
import numpy as np
import pytest

import oracle as O
from relight_amd import RL_TRACE_CLOSEST, AmbientOcclusionSettings, Relight, scenes
from relight_amd.relight import LightDesc

pytestmark = pytest.mark.gpu


def rel_rmse(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.sqrt(np.mean(b * b))
    return float(np.sqrt(np.mean((a - b) ** 2)) / den) if den > 0 else float(np.abs(a - b).max())


@pytest.fixture(scope="module")
def relight():
    print("[gpu] creating Relight context", flush=True)
    r = Relight()
    yield r
    r.close()


def oracle_bake(spec, lm_index, direct=True, ao=None, lm_seed=None):
    S = O.OracleScene.from_spec(spec)
    texel, pos, nrm = S.lumels(spec, lm_index)
    col = np.zeros((texel.size, 3), np.float32)
    occ = None
    if direct:
        S.bake_direct(pos, nrm, col)
    if ao is not None:
        occ, _ = S.bake_ao(pos, nrm, texel, lm_index if lm_seed is None else lm_seed, ao, col)
    return texel, pos, nrm, col, occ


def test_box_room_lumels_match_oracle(relight):
    print("[gpu] box_room lumels", flush=True)
    spec = scenes.box_room()
    scene, lms, ids = scenes.build_gpu(relight, spec)
    texel, pos, nrm = lms[0].lumels()
    S = O.OracleScene.from_spec(spec)
    ot, op, on = S.lumels(spec, 0)
    assert np.array_equal(texel, ot)
    assert np.array_equal(pos, op) and np.array_equal(nrm, on)


def test_box_room_direct_and_ao_bit_exact(relight):
    print("[gpu] box_room direct + AO", flush=True)
    spec = scenes.box_room()
    scene, lms, ids = scenes.build_gpu(relight, spec)
    relight.bake_direct_light(scene, ids[0])
    st_direct = relight.stats()
    relight.bake_ambient_occlusion(scene, ids[0], spec.ao)
    st_ao = relight.stats()
    hdr = lms[0].to_hdr().reshape(-1, 3)
    texel, pos, nrm, col, occ = oracle_bake(spec, 0, True, spec.ao)
    gpu = hdr[texel]
    assert rel_rmse(gpu, col) <= 1e-3
    assert np.array_equal(gpu, col), f"max abs diff {np.abs(gpu - col).max()}"
    assert st_ao["rays"] == texel.size * spec.ao.samples
    assert 0 < st_direct["rays"] <= texel.size
    # Invalid texels stay 0 (Lightmap::toHdr).
    mask = np.ones(hdr.shape[0], bool)
    mask[texel] = False
    assert (hdr[mask] == 0).all()


@pytest.mark.parametrize("sampling", [0, 1])
def test_ao_sampling_modes_bit_exact(relight, sampling):
    print(f"[gpu] box_room AO sampling={sampling}", flush=True)
    spec = scenes.box_room()
    # Exponent 2 (not 1): exercises powf; a fractional exponent of the reference's unclamped
    # negative values would give NaN on both sides.
    spec.ao = AmbientOcclusionSettings(48, 0.7, 1.1, 2.0, 77, sampling)
    scene, lms, ids = scenes.build_gpu(relight, spec)
    relight.bake_direct_light(scene, ids[0])
    relight.bake_ambient_occlusion(scene, ids[0], spec.ao)
    gpu = lms[0].to_hdr().reshape(-1, 3)
    texel, _, _, col, occ = oracle_bake(spec, 0, True, spec.ao)
    assert rel_rmse(gpu[texel], col) <= 1e-3
    # powf is not correctly rounded in glibc or on the GPU: allow an ulp-level difference.
    assert np.allclose(gpu[texel], col, rtol=1e-6, atol=0)


def _random_segments(spec, n, seed):
    rng = np.random.default_rng(seed)
    allp = np.concatenate([m.positions for m in spec.meshes])
    lo, hi = allp.min(0), allp.max(0)
    return (rng.uniform(lo, hi, (n, 3)).astype(np.float32),
            rng.uniform(lo, hi, (n, 3)).astype(np.float32))


@pytest.mark.parametrize("which", ["box_room", "city_small"])
def test_trace_segments_agreement(relight, which):
    print(f"[gpu] trace_segments {which}", flush=True)
    spec = scenes.box_room() if which == "box_room" else scenes.city(num_boxes=600, tiles=(2, 2),
                                                                    side_range=(64, 128))
    scene, lms, ids = scenes.build_gpu(relight, spec)
    a, b = _random_segments(spec, 200000, 3)
    S = O.OracleScene.from_spec(spec)
    flags = scene.trace_segments(a, b)
    ref = S.trace(a, b)
    agree_any = np.mean(flags == ref["flags"])
    hits = scene.trace_segments(a, b, RL_TRACE_CLOSEST)
    refc = S.trace(a, b, closest=True)
    agree_tri = np.mean(hits["triangle"] == refc["triangle"])
    print(f"[gpu] any-hit agreement {agree_any:.6f}, closest-triangle agreement {agree_tri:.6f}")
    assert agree_any >= 0.999 and agree_tri >= 0.999
    assert agree_any == 1.0 and agree_tri == 1.0
    hit = refc["triangle"] >= 0
    assert np.array_equal(hits["t"][hit], refc["t"][hit])
    assert np.array_equal(hits["u"][hit], refc["u"][hit]) and np.array_equal(hits["v"][hit], refc["v"][hit])
    # Mesh / face attribution of the closest hit.
    for i in np.nonzero(hit)[0][:50]:
        m, f = np.zeros(1, np.int32), np.zeros(1, np.int32)
        O.lib().or_tri_info(S._h, int(refc["triangle"][i]), m.ctypes.data_as(O._I32), f.ctypes.data_as(O._I32))
        assert hits["mesh"][i] == m[0] and hits["face"][i] == f[0]


def test_far_origin_segments_agree(relight):
    """Segments from origins far outside the scene (2 to 18 times the largest coordinate
    magnitude away, inside the range the slab test's rounding analysis and box padding cover)
    into random points of the city, any and closest hit, against the oracle."""
    print("[gpu] far-origin segments", flush=True)
    spec = scenes.city(num_boxes=600, tiles=(2, 2), side_range=(64, 128))
    scene, lms, ids = scenes.build_gpu(relight, spec)
    allp = np.concatenate([m.positions for m in spec.meshes])
    maxc = float(np.abs(allp).max())
    rng = np.random.default_rng(23)
    n = 60000
    d = rng.normal(size=(n, 3))
    d[:, 2] = np.abs(d[:, 2])  # from above the ground
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    a = (d * rng.uniform(2.0, 18.0, (n, 1)) * maxc).astype(np.float32)
    b = rng.uniform(allp.min(0), allp.max(0), (n, 3)).astype(np.float32)
    S = O.OracleScene.from_spec(spec)
    assert np.array_equal(scene.trace_segments(a, b), S.trace(a, b)["flags"])
    h = scene.trace_segments(a, b, RL_TRACE_CLOSEST)
    ref = S.trace(a, b, closest=True)
    assert np.array_equal(h["triangle"], ref["triangle"])
    hit = ref["triangle"] >= 0
    assert hit.mean() > 0.2
    assert np.array_equal(h["t"][hit], ref["t"][hit])


def test_city_small_batch_equals_plain_and_oracle(relight):
    print("[gpu] city_small batch vs plain vs oracle", flush=True)
    spec = scenes.city(num_boxes=400, tiles=(2, 2), side_range=(64, 160), ao_samples=32)
    # Plain: one bake call per mesh, direct then AO (the reference's per-mesh call shape).
    scene, lms, ids = scenes.build_gpu(relight, spec)
    for mid in ids:
        relight.bake_direct_light(scene, mid)
    for mid in ids:
        relight.bake_ambient_occlusion(scene, mid, spec.ao)
    plain = [lm.to_hdr() for lm in lms]
    # Batched: every lightmap in one launch per pass.
    scene2, lms2, ids2 = scenes.build_gpu(relight, spec)
    relight.bake_batch(scene2, lms2, direct=True, ao=spec.ao)
    st = relight.stats()
    batch = [lm.to_hdr() for lm in lms2]
    for p, b in zip(plain, batch):
        assert np.array_equal(p, b)
    lumels = 0
    for li in range(len(spec.lightmaps)):
        texel, _, _, col, _ = oracle_bake(spec, li, True, spec.ao)
        lumels += texel.size
        gpu = batch[li].reshape(-1, 3)[texel]
        assert rel_rmse(gpu, col) <= 1e-3
        assert np.array_equal(gpu, col)
    assert st["texels"] == 2 * lumels


@pytest.mark.parametrize("direct_kernel", ["lumel", "pairs"])
def test_direct_kernels_bit_exact(relight, direct_kernel):
    """Both direct-light kernels forced (a lane per lumel; (lumel, shadow ray) pairs whose
    influences are summed per lumel afterwards, in the reference's light and vertex order) against
    the oracle bit for bit: the box room (point light) and a small city (sun and a 16-vertex area
    light), the city's lightmaps in one launch."""
    from relight_amd._capi import RL_DIRECT_KERNEL_LUMEL, RL_DIRECT_KERNEL_PAIRS
    print(f"[gpu] direct kernel {direct_kernel}: box room + small city", flush=True)
    saved = relight.tuning()
    relight.set_tuning(direct_kernel=RL_DIRECT_KERNEL_LUMEL if direct_kernel == "lumel" else RL_DIRECT_KERNEL_PAIRS)
    try:
        room = scenes.box_room()
        scene, lms, ids = scenes.build_gpu(relight, room)
        relight.bake_direct_light(scene, ids[0])
        texel, _, _, col, _ = oracle_bake(room, 0, True, None)
        assert np.array_equal(lms[0].to_hdr().reshape(-1, 3)[texel], col)
        city = scenes.city(num_boxes=400, tiles=(2, 2), side_range=(64, 160), ao_samples=32)
        scene, lms, ids = scenes.build_gpu(relight, city)
        relight.bake_batch(scene, lms, direct=True, ao=None)
        st = relight.stats()
        lumels = 0
        for li in range(len(city.lightmaps)):
            texel, _, _, col, _ = oracle_bake(city, li, True, None)
            lumels += texel.size
            assert np.array_equal(lms[li].to_hdr().reshape(-1, 3)[texel], col)
        assert st["texels"] == lumels and 0 < st["rays"] <= 17 * lumels
    finally:
        relight.set_tuning(**saved)


def test_city_full_size_properties(relight):
    """The full procedural city (60k triangles, 32 lightmaps 256^2-2048^2) at reduced AO count:
    size-independent properties (ray counts exact, values finite and in range, lightmaps
    independent of each other)."""
    print("[gpu] city full-size properties", flush=True)
    spec = scenes.city(ao_samples=8)
    scene, lms, ids = scenes.build_gpu(relight, spec)
    relight.bake_batch(scene, lms, direct=False, ao=AmbientOcclusionSettings(8, 0.8, 4.0, 1.0, 1, 1))
    st = relight.stats()
    nvalid = sum(lm.valid_count() for lm in lms)
    assert st["rays"] == 8 * nvalid
    assert st["texels"] == 2 * nvalid or st["texels"] == nvalid
    # AO alone multiplies zero colour: every texel stays 0.
    assert all((lm.to_hdr() == 0).all() for lm in lms[:4])
    # Direct light on one lightmap alone equals the same lightmap inside the full batch.
    relight.bake_batch(scene, lms, direct=True, ao=None)
    full = lms[5].to_hdr()
    scene2, lms2, ids2 = scenes.build_gpu(relight, spec)
    relight.bake_direct_light(scene2, ids2[5])
    assert np.array_equal(full, lms2[5].to_hdr())
    assert np.isfinite(full).all() and (full >= 0).all()


def _gi_compare(relight, spec):
    scene, lms, ids = scenes.build_gpu(relight, spec)
    relight.emit_photons(scene, spec.indirect)
    st = relight.stats()
    S = O.OracleScene.from_spec(spec)
    maps, rays = S.emit_photons(spec, spec.indirect)
    assert st["rays"] == rays
    for k, pm in enumerate(scene.photonmaps):
        c, p, g = pm.read()
        oc, op, og = maps[k]
        assert np.array_equal(p, op), f"photon counts differ in map {k}: {np.abs(p.astype(int) - op).sum()}"
        assert np.array_equal(c, oc) and np.array_equal(g, og)
    for li, mid in enumerate(ids[:len(lms)]):
        relight.bake_indirect_light(scene, mid, spec.indirect)
        fst = relight.stats()
        texel, pos, nrm = S.lumels(spec, li)
        col = np.zeros((texel.size, 3), np.float32)
        S.final_gather(spec, pos, nrm, texel, li, spec.indirect, maps, col)
        gpu = lms[li].to_hdr().reshape(-1, 3)[texel]
        assert fst["rays"] == texel.size * spec.indirect.final_gather_samples
        assert rel_rmse(gpu, col) <= 1e-3
        assert np.array_equal(gpu, col)


def test_box_room_photons_and_final_gather_bit_exact(relight):
    print("[gpu] box_room photons + final gather", flush=True)
    _gi_compare(relight, scenes.box_room_gi())


def test_photons_russian_roulette_directional_and_cosine_fg(relight):
    print("[gpu] city_small photons (RR, sun + area) + cosine final gather", flush=True)
    from relight_amd import RL_PHOTON_RUSSIAN_ROULETTE, RL_SAMPLE_COSINE, IndirectLightSettings
    spec = scenes.city(num_boxes=150, tiles=(2, 1), side_range=(48, 64))
    rng = np.random.default_rng(3)
    for i, m in enumerate(spec.meshes):
        m.albedo = np.array([*rng.uniform(0.3, 0.8, 3), 1.0], np.float32)
        m.photonmap = i
    spec.photonmaps = [(32, 32) for _ in spec.meshes]
    spec.indirect = IndirectLightSettings(
        photon_pass_count=1, photon_bounce_count=3, photon_energy_threshold=0.0,
        photon_max_distance=300.0, final_gather_samples=40, final_gather_distance=60.0,
        final_gather_radius=2, sky_color=(0.3, 0.4, 0.5), ambient_color=(0.0, 0.0, 0.0), seed=9,
        sampling=RL_SAMPLE_COSINE, photons_per_light=15000, termination=RL_PHOTON_RUSSIAN_ROULETTE)
    _gi_compare(relight, spec)


@pytest.mark.parametrize("k,radius", [(64, 0.6), (16, 0.25), (1, 0.3)])
def test_knn_photon_gather_bit_exact(relight, k, radius):
    """Hash-grid k-nearest gather (one wavefront per photonmap texel) against the oracle's grid
    search, then the final gather that reads it."""
    print(f"[gpu] box_room k-NN gather k={k} r={radius}", flush=True)
    from relight_amd import RL_GATHER_KNN
    spec = scenes.box_room_gi()
    spec.indirect.gather_mode, spec.indirect.knn_k, spec.indirect.knn_radius = RL_GATHER_KNN, k, radius
    _gi_compare(relight, spec)


def test_knn_gather_city_multi_map(relight):
    print("[gpu] city k-NN gather, 8 photonmaps, RR", flush=True)
    from relight_amd import RL_GATHER_KNN, RL_PHOTON_RUSSIAN_ROULETTE, IndirectLightSettings
    spec = scenes.city(num_boxes=200, tiles=(4, 2), side_range=(32, 64))
    rng = np.random.default_rng(5)
    for i, m in enumerate(spec.meshes):
        m.albedo = np.array([*rng.uniform(0.3, 0.8, 3), 1.0], np.float32)
        m.photonmap = i
    spec.photonmaps = [(48, 48) for _ in spec.meshes]
    spec.indirect = IndirectLightSettings(
        photon_pass_count=1, photon_bounce_count=3, photon_energy_threshold=0.0,
        photon_max_distance=300.0, final_gather_samples=24, final_gather_distance=60.0,
        final_gather_radius=2, sky_color=(0.3, 0.4, 0.5), seed=4, photons_per_light=40000,
        termination=RL_PHOTON_RUSSIAN_ROULETTE, gather_mode=RL_GATHER_KNN, knn_k=64, knn_radius=3.0)
    _gi_compare(relight, spec)


def test_heightfield_small_photons_knn_final_gather(relight):
    """The heightfield workload at reduced size (64x64 cells in 4 tiles, 40 boxes without
    photonmaps, 4 area lights, Russian roulette, k = 64 gather) against the oracle, bit for bit."""
    print("[gpu] small heightfield photons + k-NN + final gather", flush=True)
    spec = scenes.heightfield(cells=64, tile_cells=32, num_boxes=40, map_size=256, photonmap_size=32,
                              photons_per_light=20000, fg_samples=64, box_meshes=2)
    _gi_compare(relight, spec)


@pytest.mark.parametrize("fg_kernel", ["wave", "lane"])
def test_final_gather_kernels_bit_exact(relight, fg_kernel):
    """Both final-gather kernels forced on small scenes (a wavefront per lumel summing through an
    in-order LDS ring; a lane per lumel summing in registers), against the oracle bit for bit:
    the city with 40 cosine samples per texel (lumels end inside a wavefront's sample batch) and
    the small heightfield with the k-NN gather."""
    from relight_amd._capi import RL_FG_KERNEL_LANE, RL_FG_KERNEL_WAVE
    from relight_amd import RL_PHOTON_RUSSIAN_ROULETTE, RL_SAMPLE_COSINE, IndirectLightSettings
    print(f"[gpu] final gather kernel {fg_kernel}: city (40 cosine samples) + small heightfield", flush=True)
    saved = relight.tuning()
    relight.set_tuning(fg_kernel=RL_FG_KERNEL_WAVE if fg_kernel == "wave" else RL_FG_KERNEL_LANE, refill_min=8)
    try:
        spec = scenes.city(num_boxes=120, tiles=(2, 1), side_range=(40, 56))
        rng = np.random.default_rng(11)
        for i, m in enumerate(spec.meshes):
            m.albedo = np.array([*rng.uniform(0.3, 0.8, 3), 1.0], np.float32)
            m.photonmap = i
        spec.photonmaps = [(32, 32) for _ in spec.meshes]
        spec.indirect = IndirectLightSettings(
            photon_pass_count=1, photon_bounce_count=3, photon_energy_threshold=0.0,
            photon_max_distance=300.0, final_gather_samples=40, final_gather_distance=60.0,
            final_gather_radius=2, sky_color=(0.3, 0.4, 0.5), ambient_color=(0.01, 0.0, 0.02), seed=13,
            sampling=RL_SAMPLE_COSINE, photons_per_light=12000, termination=RL_PHOTON_RUSSIAN_ROULETTE)
        _gi_compare(relight, spec)
        _gi_compare(relight, scenes.heightfield(cells=64, tile_cells=32, num_boxes=40, map_size=256,
                                                photonmap_size=32, photons_per_light=20000, fg_samples=64,
                                                box_meshes=2))
    finally:
        relight.set_tuning(**saved)


def test_axis_parallel_segments_agree(relight):
    """Segments with one or two direction components exactly zero (the slab test's 1/0 case),
    from origins on box faces and inside the city, against the oracle."""
    print("[gpu] axis-parallel segments", flush=True)
    spec = scenes.city(num_boxes=600, tiles=(2, 2), side_range=(64, 128))
    scene, lms, ids = scenes.build_gpu(relight, spec)
    rng = np.random.default_rng(17)
    allp = np.concatenate([m.positions for m in spec.meshes])
    n = 60000
    a = allp[rng.integers(0, allp.shape[0], n)].astype(np.float32)  # vertices: on box planes
    a[n // 2:] = rng.uniform(allp.min(0), allp.max(0), (n - n // 2, 3)).astype(np.float32)
    b = a.copy()
    axis = rng.integers(0, 3, n)
    step = rng.uniform(-30, 30, n).astype(np.float32)
    b[np.arange(n), axis] += step
    two = rng.random(n) < 0.3   # some segments move along two axes
    b[np.arange(n)[two], (axis[two] + 1) % 3] += step[two]
    S = O.OracleScene.from_spec(spec)
    assert np.array_equal(scene.trace_segments(a, b), S.trace(a, b)["flags"])
    h = scene.trace_segments(a, b, RL_TRACE_CLOSEST)
    assert np.array_equal(h["triangle"], S.trace(a, b, closest=True)["triangle"])


@pytest.mark.parametrize("samples,sampling", [(4, "cosine"), (5, "cosine"), (48, "cosine"), (100, "cosine"),
                                              (256, "cosine"), (5, "uniform"), (100, "uniform")])
def test_ao_refill_equals_chunked_and_oracle(relight, samples, sampling):
    """The lane-refill AO kernel against the chunked kernel and the oracle, for sample counts
    below, between and above multiples of 64, under several refill thresholds and chunk sizes
    (64 is above the refill kernel's 16-lumel chunk limit, which the runtime applies), with cosine
    and uniform hemisphere sampling."""
    print(f"[gpu] AO refill vs chunked, samples={samples}, {sampling}", flush=True)
    from relight_amd import RL_SAMPLE_COSINE, RL_SAMPLE_UNIFORM
    spec = scenes.city(num_boxes=300, tiles=(2, 2), side_range=(48, 96))
    ao = AmbientOcclusionSettings(samples, 0.8, 4.0, 1.0, 3,
                                  RL_SAMPLE_COSINE if sampling == "cosine" else RL_SAMPLE_UNIFORM)
    outs = []
    for kernel, rmin, chunk in [(0, 16, 8), (1, 1, 1), (1, 16, 8), (1, 64, 3), (1, 8, 64)]:
        relight.set_tuning(ao_kernel=kernel, refill_min=rmin, ao_chunk_lumels=chunk)
        scene, lms, ids = scenes.build_gpu(relight, spec)
        relight.bake_batch(scene, lms, direct=True, ao=ao)
        assert relight.stats()["kernels"]["ao_count"]["launches"] == 1
        outs.append([lm.to_hdr() for lm in lms])
    relight.set_tuning(ao_kernel=1, refill_min=16, ao_chunk_lumels=8)
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)
    for li in (0, 3):
        texel, _, _, col, _ = oracle_bake(spec, li, True, ao)
        assert np.array_equal(outs[2][li].reshape(-1, 3)[texel], col)


@pytest.mark.parametrize("sampling,direct_kernel", [("cosine", "lumel"), ("uniform", "lumel"), ("cosine", "pairs")])
def test_any_hit_shortcuts_change_no_result(relight, sampling, direct_kernel):
    """Local trees (AO rays start at a tree built in LDS over the subtrees their chunk can reach)
    and occluder hints (AO and shadow rays first test recently blocking triangles) switched off and
    on through rl_tuning: direct light and AO equal bit for bit in all four combinations, and equal
    the oracle. 64 AO rays per lumel, so every 16-lumel chunk builds a tree."""
    print(f"[gpu] any-hit shortcuts on/off, {sampling} AO, {direct_kernel} direct kernel", flush=True)
    from relight_amd import RL_SAMPLE_COSINE, RL_SAMPLE_UNIFORM
    from relight_amd._capi import RL_DIRECT_KERNEL_LUMEL, RL_DIRECT_KERNEL_PAIRS
    spec = scenes.city(num_boxes=300, tiles=(2, 2), side_range=(48, 96))
    ao = AmbientOcclusionSettings(64, 0.8, 4.0, 1.0, 3,
                                  RL_SAMPLE_COSINE if sampling == "cosine" else RL_SAMPLE_UNIFORM)
    dk = RL_DIRECT_KERNEL_LUMEL if direct_kernel == "lumel" else RL_DIRECT_KERNEL_PAIRS
    saved = relight.tuning()
    outs = []
    try:
        for trees, hints in [(0, 0), (1, 0), (0, 1), (1, 1)]:
            relight.set_tuning(ao_kernel=1, refill_min=16, ao_chunk_lumels=16, direct_kernel=dk,
                               local_trees=trees, occluder_hints=hints)
            scene, lms, ids = scenes.build_gpu(relight, spec)
            relight.bake_batch(scene, lms, direct=True, ao=ao)
            outs.append([lm.to_hdr() for lm in lms])
    finally:
        relight.set_tuning(**saved)
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)
    texel, _, _, col, _ = oracle_bake(spec, 1, True, ao)
    assert np.array_equal(outs[3][1].reshape(-1, 3)[texel], col)


def _spot_spec():
    """Box room (photons + final gather) lit by two spot lights: exponent 1 (createSpotLight) and
    exponent 2 (the powf branch of LightSpotCutoff, scene/Light.cpp:452-477)."""
    spec = scenes.box_room_gi(photons_per_light=20000, fg_samples=32)
    d1 = np.array([0.2, -1.0, 0.1])
    d2 = np.array([-0.4, -1.0, -0.3])
    s1 = LightDesc.spot((2.0, 2.9, 2.0), tuple(d1 / np.linalg.norm(d1)), 0.6, radius=10.0, intensity=40.0)
    s2 = LightDesc.spot((1.0, 2.5, 3.0), tuple(d2 / np.linalg.norm(d2)), 0.8, radius=8.0, color=(1.0, 0.7, 0.4),
                        intensity=30.0)
    s2.spot_exponent = 2.0
    spec.lights = [s1, s2]
    spec.name = "box_room_spots"
    return spec


def test_spot_lights_direct_photons_and_final_gather_bit_exact(relight):
    print("[gpu] spot lights: direct, photons, final gather", flush=True)
    spec = _spot_spec()
    scene, lms, ids = scenes.build_gpu(relight, spec)
    relight.bake_direct_light(scene, ids[0])
    gpu = lms[0].to_hdr().reshape(-1, 3)
    texel, pos, nrm, col, _ = oracle_bake(spec, 0, True, None)
    assert np.array_equal(gpu[texel], col)
    lit = (col.max(1) > 0).mean()
    assert 0.05 < lit < 0.95  # the cones light part of the room
    _gi_compare(relight, spec)


def _rot(axis, ang):
    a = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


def _matrix4(R, t):
    """Matrix4 (Types.h:1188-1196) for p' = R p + t: m[4*i + j] = R[j][i], m[12 + j] = t[j]."""
    M = np.eye(4)
    M[:3, :3] = np.asarray(R).T
    M[3, :3] = t
    return M.reshape(16).astype(np.float32)


def _transformed_spec():
    """Box room walls plus two copies of one cube mesh added with Scene::addMesh transforms
    (rotation + translation, input normals of length 1.7 that the transform renormalises), each
    cube in its own lightmap and photonmap."""
    spec = scenes.box_room_gi(photons_per_light=10000, fg_samples=24)
    walls = scenes._box_quads((0, 0, 0), (4, 3, 4), 8, inward=True)
    scenes._pack(walls, 128, 128)
    p, n, uv, ix = scenes._emit(walls, 128, 128)
    meshes = [scenes.MeshSpec(p, n, uv, ix, 0, albedo=np.array([0.8, 0.7, 0.6, 1.0], np.float32), photonmap=0)]
    cube = scenes._box_quads((-0.4, 0.0, -0.4), (0.4, 0.8, 0.4), 4)
    scenes._pack(cube, 64, 64)
    cp, cn, cuv, cix = scenes._emit(cube, 64, 64)
    cn = (cn * 1.7).astype(np.float32)
    for k, (R, t) in enumerate([(_rot((0, 1, 0), 0.5), (1.2, 0.0, 1.3)),
                                (_rot((1, 2, 0.5), 0.35), (2.8, 0.45, 2.7))]):
        meshes.append(scenes.MeshSpec(cp, cn, cuv, cix, 1 + k, albedo=np.array([0.5, 0.6, 0.7, 1.0], np.float32),
                                      photonmap=1 + k, transform=_matrix4(R, t)))
    spec.meshes = meshes
    spec.lightmaps = [(128, 128), (64, 64), (64, 64)]
    spec.photonmaps = [(32, 32), (16, 16), (16, 16)]
    spec.name = "box_room_transformed"
    return spec


def test_transformed_meshes_bit_exact(relight):
    print("[gpu] transformed meshes: lumels, direct + AO, segments, photons, final gather", flush=True)
    spec = _transformed_spec()
    scene, lms, ids = scenes.build_gpu(relight, spec)
    for li in range(3):
        texel, pos, nrm = lms[li].lumels()
        S = O.OracleScene.from_spec(spec)
        ot, op, on = S.lumels(spec, li)
        assert np.array_equal(texel, ot) and np.array_equal(pos, op) and np.array_equal(nrm, on)
        relight.bake_direct_light(scene, ids[li])
        relight.bake_ambient_occlusion(scene, ids[li], spec.ao)
        gpu = lms[li].to_hdr().reshape(-1, 3)[texel]
        _, _, _, col, _ = oracle_bake(spec, li, True, spec.ao)
        assert np.array_equal(gpu, col), f"lightmap {li}"
    a, b = _random_segments(spec, 50000, 21)
    S = O.OracleScene.from_spec(spec)
    hits = scene.trace_segments(a, b, RL_TRACE_CLOSEST)
    ref = S.trace(a, b, closest=True)
    hit = ref["triangle"] >= 0
    assert np.array_equal(hits["triangle"], ref["triangle"]) and np.array_equal(hits["t"][hit], ref["t"][hit])
    _gi_compare(relight, spec)


def test_bake_scene_equals_oracle_sequence(relight):
    """Relight::bake with a direct + indirect + AO job (rl_bake_scene) against the oracle's
    Photons::emit -> DirectLight -> IndirectLight -> AmbientOcclusion sequence (AO multiplies the
    accumulated colour, baker/AmbientOcclusion.cpp:59-65)."""
    print("[gpu] rl_bake_scene vs oracle sequence", flush=True)
    spec = scenes.box_room_gi(photons_per_light=10000, fg_samples=32)
    scene, lms, ids = scenes.build_gpu(relight, spec)
    relight.bake(scene, True, spec.ao, spec.indirect)
    S = O.OracleScene.from_spec(spec)
    maps, _ = S.emit_photons(spec, spec.indirect)
    texel, pos, nrm = S.lumels(spec, 0)
    col = np.zeros((texel.size, 3), np.float32)
    S.bake_direct(pos, nrm, col)
    S.final_gather(spec, pos, nrm, texel, 0, spec.indirect, maps, col)
    S.bake_ao(pos, nrm, texel, 0, spec.ao, col)
    gpu = lms[0].to_hdr().reshape(-1, 3)[texel]
    assert rel_rmse(gpu, col) <= 1e-3
    assert np.array_equal(gpu, col)
