"""Multi-scene batching and the continuous-batching server on the GPU.

Bars: a lumel-range item bake across scenes equals the per-lightmap bakes bit for bit; a job's
results are bit-identical whether it runs alone or in a mix with other scenes and job kinds;
admission, rejection and cancellation follow the server's rules.
"""
import numpy as np
import pytest

from relight_amd import AmbientOcclusionSettings, RelightError, Relight, scenes
from relight_amd import patches as PT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def relight():
    print("[gpu] creating Relight context", flush=True)
    r = Relight()
    yield r
    r.close()


def _small_city(seed=7, ao_samples=32):
    return scenes.city(seed=seed, num_boxes=300, tiles=(2, 2), side_range=(48, 96), ao_samples=ao_samples)


def _plain(relight, spec, direct=True, ao=True):
    scene, lms, ids = scenes.build_gpu(relight, spec)
    if direct:
        for mid in ids:
            relight.bake_direct_light(scene, mid)
    if ao:
        for mid in ids:
            relight.bake_ambient_occlusion(scene, mid, spec.ao)
    return [lm.to_hdr() for lm in lms]


def test_bake_items_across_scenes_equal_plain(relight):
    print("[gpu] bake_items: box room + city ranges in one launch per pass", flush=True)
    a, b = scenes.box_room(), _small_city()
    plain_a, plain_b = _plain(relight, a), _plain(relight, b)
    sa, la, _ = scenes.build_gpu(relight, a)
    sb, lb, _ = scenes.build_gpu(relight, b)
    rng = np.random.default_rng(5)
    items = []
    for scene, lms, spec in ((sa, la, a), (sb, lb, b)):
        for lm in lms:
            n = lm.valid_count()
            cuts = np.unique(np.concatenate([[0, n], rng.integers(0, n, 3)]))
            for lo, hi in zip(cuts[:-1], cuts[1:]):
                items.append((scene, lm, int(lo), int(hi - lo), True, spec.ao))
    rng.shuffle(items)
    relight.bake_items(items)
    st = relight.stats()
    assert st["kernels"]["ao_count"]["launches"] == 1
    for p, lm in zip(plain_a + plain_b, la + lb):
        assert np.array_equal(p, lm.to_hdr())


def test_bake_items_rejects_bad_ranges(relight):
    spec = scenes.box_room()
    scene, lms, _ = scenes.build_gpu(relight, spec)
    n = lms[0].valid_count()
    with pytest.raises(RelightError):
        relight.bake_items([(scene, lms[0], n - 4, 8, True, None)])
    other, olms, _ = scenes.build_gpu(relight, spec)
    with pytest.raises(RelightError):
        relight.bake_items([(scene, olms[0], 0, None, True, None)])


def _radiosity_case(relight):
    ps = PT.block_city(target=1024, ground=16, lots=2, sets=4, emitters=8, seed=2)
    scene, rad = PT.build_gpu(relight, ps)
    return ps, scene, rad


def test_server_job_alone_equals_job_in_mix(relight):
    """A DIRECT_AO job baked alone (many small steps) and again inside a mix of three other
    jobs (another scene's direct/AO, photons + final gather, radiosity) gives identical lightmaps;
    every other job of the mix equals its own plain bake."""
    print("[gpu] server: alone vs mix", flush=True)
    city = _small_city()
    room = scenes.box_room()
    gi = scenes.box_room_gi()
    # Alone.
    sc, lc, _ = scenes.build_gpu(relight, city)
    with relight.create_server(quantum_rays=1 << 16, max_active=4) as srv:
        job = srv.submit_direct_ao(sc, lc, True, city.ao)
        st = job.wait(timeout=120)
        assert st["state"] == "done" and st["rays"] > 0
        assert srv.stats()["steps"] > 2
    alone = [lm.to_hdr() for lm in lc]
    # Plain references of the other jobs.
    plain_room = _plain(relight, room)
    sg, lg, ig = scenes.build_gpu(relight, gi)
    relight.emit_photons(sg, gi.indirect)
    for mid in ig:
        relight.bake_indirect_light(sg, mid, gi.indirect)
    plain_gi = [lm.to_hdr() for lm in lg]
    ps, _, rad = _radiosity_case(relight)
    rad.build_form_factors(256, seed=9, max_distance=200.0)
    plain_B, _ = rad.solve(ps.emission, ps.reflectance, tolerance=1e-5, max_outer=40, inner_sweeps=16)
    # Mix.
    sc2, lc2, _ = scenes.build_gpu(relight, city)
    sr2, lr2, _ = scenes.build_gpu(relight, room)
    sg2, lg2, _ = scenes.build_gpu(relight, gi)
    ps2, _, rad2 = _radiosity_case(relight)
    # A final-gather-only job on photonmaps emitted before the server starts.
    sg3, lg3, _ = scenes.build_gpu(relight, gi)
    relight.emit_photons(sg3, gi.indirect)
    with relight.create_server(quantum_rays=1 << 21, max_active=8) as srv:
        jobs = [srv.submit_direct_ao(sr2, lr2, True, room.ao),
                srv.submit_indirect(sg2, lg2, gi.indirect),
                srv.submit_direct_ao(sc2, lc2, True, city.ao),
                srv.submit_radiosity(rad2, ps2.emission, ps2.reflectance, 256, 9, 200.0, 1e-5),
                srv.submit_final_gather(sg3, lg3, gi.indirect)]
        sts = [j.wait(timeout=300) for j in jobs]
        assert all(s["state"] == "done" for s in sts)
        stats = srv.stats()
        assert stats["completed"] == 5 and stats["failed"] == 0
        assert stats["peak_memory"] >= max(s["footprint"] for s in sts)
    for a, lm in zip(alone, lc2):
        assert np.array_equal(a, lm.to_hdr())
    for p, lm in zip(plain_room, lr2):
        assert np.array_equal(p, lm.to_hdr())
    for p, lm in zip(plain_gi, lg2):
        assert np.array_equal(p, lm.to_hdr())
    for p, lm in zip(plain_gi, lg3):
        assert np.array_equal(p, lm.to_hdr())
    assert np.array_equal(plain_B, jobs[3].result)


def test_server_rejects_and_cancels(relight):
    print("[gpu] server: rejection, admission, cancel", flush=True)
    spec = _small_city(ao_samples=16)
    scene, lms, _ = scenes.build_gpu(relight, spec)
    with relight.create_server(memory_budget=1 << 10, max_active=2) as srv:
        with pytest.raises(RelightError, match="budget"):
            srv.submit_direct_ao(scene, lms, True, spec.ao)
        assert srv.stats()["rejected"] == 1
    other, olms, _ = scenes.build_gpu(relight, scenes.box_room())
    with relight.create_server(max_active=1, quantum_rays=1 << 18) as srv:
        with pytest.raises(RelightError):
            srv.submit_direct_ao(scene, olms, True, spec.ao)  # lightmap of another scene
        first = srv.submit_direct_ao(scene, lms, True, spec.ao)
        queued = srv.submit_direct_ao(scene, lms, True, spec.ao)
        queued.cancel()
        assert first.wait(timeout=120)["state"] == "done"
        st = queued.wait(timeout=120)
        assert st["state"] in ("cancelled", "done")
        stats = srv.stats()
        assert stats["completed"] + stats["cancelled"] == 2
    bad = AmbientOcclusionSettings(0, 0.8, 4.0, 1.0, 1, 1)
    with relight.create_server() as srv:
        with pytest.raises(RelightError):
            srv.submit_direct_ao(scene, lms, True, bad)


def _full_plain(relight, spec):
    scene, lms, ids = scenes.build_gpu(relight, spec)
    relight.emit_photons(scene, spec.indirect)
    for mid in ids:
        relight.bake_direct_light(scene, mid)
    for mid in ids:
        relight.bake_indirect_light(scene, mid, spec.indirect)
    for mid in ids:
        relight.bake_ambient_occlusion(scene, mid, spec.ao)
    return [lm.to_hdr() for lm in lms]


def test_full_bake_scene_and_server_job_equal_plain_sequence(relight):
    """Relight::bake with a direct + indirect + AO job: rl_bake_scene (one launch per pass over every
    lightmap) and a FULL_BAKE server job equal the per-mesh call sequence bit for bit."""
    print("[gpu] full bake: scene call and server job vs plain calls", flush=True)
    spec = scenes.box_room_gi(photons_per_light=10000, fg_samples=32)
    plain = _full_plain(relight, spec)
    scene, lms, _ = scenes.build_gpu(relight, spec)
    relight.bake(scene, True, spec.ao, spec.indirect)
    st = relight.stats()
    for p, lm in zip(plain, lms):
        assert np.array_equal(p, lm.to_hdr())
    assert st["rays"] > 0
    s2, l2, _ = scenes.build_gpu(relight, spec)
    s3, l3, _ = scenes.build_gpu(relight, _small_city(ao_samples=16))
    with relight.create_server(quantum_rays=1 << 18, max_active=4) as srv:
        jobs = [srv.submit_full_bake(s2, l2, True, spec.ao, spec.indirect),
                srv.submit_direct_ao(s3, l3, True, AmbientOcclusionSettings(16, 0.8, 4.0, 1.0, 1, 1))]
        assert all(j.wait(timeout=120)["state"] == "done" for j in jobs)
    for p, lm in zip(plain, l2):
        assert np.array_equal(p, lm.to_hdr())


@pytest.mark.gpu
def test_recorded_pass_replays_equal_eager_passes():
    """A bake_batch pass recorded as a HIP graph, replayed four times with other eager bakes and
    synchronisations in between, equals the same eager passes on an identical second set of
    lightmaps after every replay (direct light accumulates, so each pass changes the colours)."""
    print("[gpu] recorded pass (HIP graph) vs eager bake_batch, 4 replays", flush=True)
    r = Relight()
    spec = scenes.city(num_boxes=200, tiles=(3, 2), side_range=(48, 96), ao_samples=16)
    scene_a, lms_a, _ = scenes.build_gpu(r, spec)
    scene_b, lms_b, _ = scenes.build_gpu(r, spec)
    scene_c, lms_c, ids_c = scenes.build_gpu(r, spec)
    ao = AmbientOcclusionSettings(16, 0.8, 4.0, 1.0, 3, spec.ao.sampling)
    p = r.record_pass(scene_b, lms_b, direct=True, ao=ao)
    for k in range(4):
        r.bake_batch(scene_a, lms_a, direct=True, ao=ao)
        p.launch()
        st = r.stats()
        assert st["rays"] > 0 and st["texels"] == sum(lm.valid_count() for lm in lms_b)
        for a, b in zip(lms_a, lms_b):
            assert np.array_equal(a.to_hdr(), b.to_hdr()), f"replay {k}"
        r.bake_direct_light(scene_c, ids_c[k % len(ids_c)])  # unrelated eager work between replays
        r.bake_ambient_occlusion(scene_c, ids_c[0], ao)
    p.close()
    r.close()


def test_server_jobs_keep_their_objects_alive(relight):
    """A job borrows its scene, lightmaps and radiosity object until it finishes: dropping the
    caller's references after submit neither frees them nor changes the result, and destroying a
    borrowed object through the C-ABI is refused (RL_INVALID_CALL) until the job is done."""
    import gc

    from relight_amd import _capi as A
    print("[gpu] server: job lifetimes", flush=True)
    spec = _small_city(ao_samples=16)
    ref_scene, ref_lms, _ = scenes.build_gpu(relight, spec)
    relight.bake_batch(ref_scene, ref_lms, direct=True, ao=spec.ao)
    want = [lm.to_hdr() for lm in ref_lms]
    big = scenes.city(num_boxes=800, tiles=(2, 2), side_range=(512, 768), ao_samples=256)
    with relight.create_server(max_active=1, quantum_rays=1 << 24) as srv:
        s1, l1, _ = scenes.build_gpu(relight, big)
        first = srv.submit_direct_ao(s1, l1, True, big.ao)  # keeps the only slot busy for a while
        s2, l2, _ = scenes.build_gpu(relight, spec)
        job = srv.submit_direct_ao(s2, l2, True, spec.ao)   # queued behind it
        handle = s2._h
        assert relight._lib.rl_scene_destroy(handle) == A.RL_INVALID_CALL
        assert b"server job" in relight._lib.rl_last_error()
        with pytest.raises(RelightError):
            l2[0].close()
        kept = job._keep[1]
        del s1, l1, s2, l2
        gc.collect()
        assert first.wait(timeout=120)["state"] == "done"
        assert job.wait(timeout=120)["state"] == "done"
        assert job.id not in srv._jobs
    for lm, w in zip(kept, want):
        assert np.array_equal(lm.to_hdr(), w)
    # Finished: the objects can be destroyed again.
    del job
    for lm in kept:
        lm.close()


def test_server_interleaved_scenes_and_running_cancel(relight):
    """Jobs of four scenes submitted alternately with a small quantum, so the server's steps (one
    scene each, the next queued while one runs) switch scenes many times: every job equals its
    plain bake bit for bit. A large job cancelled while its steps run ends cancelled, the other
    jobs still finish, and its scene can be destroyed once the server is done."""
    print("[gpu] server: interleaved scenes, pipelined steps, running cancel", flush=True)
    specs = [scenes.box_room(seed=1), _small_city(seed=3, ao_samples=16), scenes.box_room(seed=2),
             _small_city(seed=4, ao_samples=16)]
    plain = [_plain(relight, sp) for sp in specs]
    built = [scenes.build_gpu(relight, sp) for sp in specs * 2]
    big_spec = _small_city(seed=5, ao_samples=256)
    big_scene, big_lms, _ = scenes.build_gpu(relight, big_spec)
    with relight.create_server(quantum_rays=1 << 17, max_active=16) as srv:
        jobs = []
        for k, (sc, lms, _) in enumerate(built):
            jobs.append(srv.submit_direct_ao(sc, lms, True, specs[k % 4].ao))
            if k == 1:
                big = srv.submit_direct_ao(big_scene, big_lms, True, big_spec.ao)
        for _ in range(2000):
            if big.poll()["state"] != "queued":
                break
        big.cancel()
        sts = [j.wait(timeout=300) for j in jobs]
        assert all(s["state"] == "done" for s in sts)
        assert big.wait(timeout=300)["state"] in ("cancelled", "done")
        assert srv.stats()["steps"] > 8
    for k, (sc, lms, _) in enumerate(built):
        for want, lm in zip(plain[k % 4], lms):
            assert np.array_equal(want, lm.to_hdr())
    big_scene.close()
