"""Secondary benchmark legs reported beside the headline city bake (imported by bench.py).

  gi         the heightfield workload: 4M photons (Russian roulette, 3 bounces), k = 64 hash-grid
             gather over 64 photonmaps, then 512 final-gather rays per texel on one 2048^2 map.
  radiosity  16,384 patches of the block city: form factors by 1,024 cosine rays per patch into a
             dense 16,384^2 count matrix, then the BF16-MFMA gather solve of B = E + rho o (F B) for
             96 right-hand sides to residual 1e-5, with the residual recomputed on the host in FP64.
  serving    a stream of bake requests over resident scenes (cities and heightfields, 10k-485k
             triangles; lightmaps 128^2-2048^2; 50% direct+AO, 30% final gather, 20% radiosity):
             saturated throughput and Poisson-arrival latency, served (continuous batching) vs
             plain (one request at a time through the per-call entry points).
  serving_small  the same comparison for a stream of small requests: 256 direct+AO requests
             (64 AO rays) on 128^2-256^2 lightmaps of one resident city, where one request does not
             fill the GPU and the per-call overhead of the plain path dominates.
"""
import sys
import threading
import time

import numpy as np


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def percentile(x, q):
    return float(np.percentile(np.asarray(x, np.float64), q)) if len(x) else None


HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0


# ---------------------------------------------------------------------------------------------- gi

def bench_gi(relight, steps=2):
    from relight_amd import scenes
    t0 = time.time()
    spec = scenes.heightfield()
    scene, lms, ids = scenes.build_gpu(relight, spec)
    info = scene.info()
    texels = lms[0].valid_count()
    ind = spec.indirect
    log(f"gi: heightfield built: {info}, {texels} texels, {len(spec.photonmaps)} photonmaps ({time.time()-t0:.1f} s)")
    # Check pass (also the warm-up): one photon emission and one final gather from zero, kept for
    # the comparison with the oracle in bench.py's CPU leg (later passes accumulate onto them).
    relight.emit_photons(scene, ind)
    check_maps = [pm.read() for pm in scene.photonmaps]
    relight.bake_indirect_light(scene, ids[0], ind)
    check = {"spec": spec, "maps": check_maps, "fg": lms[0].to_hdr().reshape(-1, 3).copy()}
    rec = []
    for s in range(steps):
        t = time.perf_counter()
        relight.emit_photons(scene, ind)
        pst = relight.stats()
        relight.bake_indirect_light(scene, ids[0], ind)
        fst = relight.stats()
        ms = (time.perf_counter() - t) * 1e3
        rec.append((ms, pst, fst))
        log(f"gi step {s}: {ms:.1f} ms, photon rays {pst['rays']}, fg rays {fst['rays']}, "
            f"kernels {pst['kernels']} {fst['kernels']}")
    ms = float(np.mean([r[0] for r in rec]))
    prays = float(np.mean([r[1]["rays"] for r in rec]))
    frays = float(np.mean([r[2]["rays"] for r in rec]))
    k = lambda r, name: r.get("kernels", {}).get(name, {"ms": 0.0})["ms"]
    fg_ms = float(np.mean([k(r[2], "final_gather") for r in rec]))
    knn_ms = float(np.mean([k(r[1], "photon_gather") for r in rec]))
    grid_ms = float(np.mean([k(r[1], "photon_grid") for r in rec]))
    trace_ms = float(np.mean([k(r[1], "photon_trace") for r in rec]))
    store_ms = float(np.mean([k(r[1], "photon_store") for r in rec]))
    nq = sum(w * h for w, h in spec.photonmaps)
    roofline = None
    import bench
    pmc = bench.load_pmc("k_fg_lanes", "heightfield")
    if pmc and fg_ms and relight.tuning()["fg_kernel"] in (0, 2):
        roofline = bench.valu_roofline(pmc, fg_ms / 1e3, bench.MI355X_CUS, texels)
        roofline["kernel"] = "k_fg_lanes"
    return check, {
        "workload": f"heightfield: 512^2-cell fBm terrain + 2000 boxes ({spec.num_tris} tris), 4 area lights, "
                    f"{ind.photons_per_light * len(spec.lights)} photons (RR, 3 bounces), k={ind.knn_k} "
                    f"hash-grid gather r={ind.knn_radius} m on 64 photonmaps 128^2, FG "
                    f"{ind.final_gather_samples} rays/texel on 2048^2 ({texels} texels)",
        "ms_per_step": ms, "steps": steps,
        "mrays_per_s": (prays + frays) / (ms / 1e3) / 1e6,
        "final_gather": {"rays": frays, "kernel_ms": fg_ms, "mrays_per_s": frays / (fg_ms / 1e3) / 1e6 if fg_ms else None,
                         "texels_per_s": texels / (fg_ms / 1e3) if fg_ms else None},
        "photons": {"rays": prays, "trace_ms": trace_ms, "store_ms": store_ms, "grid_ms": grid_ms,
                    "knn_gather_ms": knn_ms, "knn_queries": nq,
                    "knn_queries_per_s": nq / (knn_ms / 1e3) if knn_ms else None},
        "triangles": spec.num_tris, "bvh_nodes": info["nodes"],
        "roofline": roofline,
    }


# --------------------------------------------------------------------------------------- radiosity

def bench_radiosity(relight, rays=1024, max_distance=1000.0):
    from relight_amd import patches as PT
    t0 = time.time()
    ps = PT.block_city(target=16384)
    scene, rad = PT.build_gpu(relight, ps)
    n, m = ps.num_patches, ps.emission.shape[1]
    log(f"radiosity: {n} patches, {m} right-hand sides ({time.time()-t0:.1f} s)")
    rad.build_form_factors(rays, seed=1, max_distance=max_distance)  # warm-up (allocates F)
    t = time.perf_counter()
    rad.build_form_factors(rays, seed=1, max_distance=max_distance)
    ff_ms = (time.perf_counter() - t) * 1e3
    ff = relight.stats()
    log(f"radiosity: form factors {ff_ms:.1f} ms, {ff['rays']} rays")
    rad.solve(ps.emission, ps.reflectance, tolerance=1e-5, max_outer=40, inner_sweeps=16)  # warm-up
    t = time.perf_counter()
    B, info = rad.solve(ps.emission, ps.reflectance, tolerance=1e-5, max_outer=40, inner_sweeps=16)
    solve_ms = (time.perf_counter() - t) * 1e3
    st = relight.stats()
    log(f"radiosity: solve {solve_ms:.1f} ms, {info}, kernels {st['kernels']}")
    # FP64 residual on the host: ||E + rho o (F B) - B|| / ||E|| with F = counts / rays.
    C = rad.form_factor_counts()
    B64 = B.astype(np.float64)
    FB = np.empty_like(B64)
    for r0 in range(0, n, 2048):
        FB[r0:r0 + 2048] = (C[r0:r0 + 2048].astype(np.float64) / rays) @ B64
    res64 = float(np.linalg.norm(ps.emission + ps.reflectance * FB - B64) / np.linalg.norm(ps.emission))
    del FB
    err64, lu_s = fp64_solve_error(C, rays, ps, B64)
    log(f"radiosity: error vs FP64 LU solve {err64:.3e} ({lu_s:.1f} s)")
    npad = (n + 63) // 64 * 64
    gemm = {}
    for name in ("ff_gemm_fast", "ff_gemm_accurate"):
        k = st["kernels"].get(name)
        if not k or not k["launches"]:
            continue
        avg_s = k["ms"] / k["launches"] / 1e3
        flops = 2.0 * n * n * m          # one F * X block product (algorithmic)
        fbytes = 2.0 * npad * npad       # the u16 count matrix streamed once per launch
        gemm[name] = {"launches": k["launches"], "avg_launch_ms": avg_s * 1e3,
                      "tflops": flops / avg_s / 1e12, "mfma_frac": flops / avg_s / 1e12 / BF16_PEAK_TFLOPS,
                      "hbm_gbs": fbytes / avg_s / 1e9, "hbm_frac": fbytes / avg_s / 1e9 / HBM_PEAK_GBS}
    return {
        "workload": f"block city: {n} patches of 1 m, {rays} cosine rays/patch, dense {n}^2 counts (u16), "
                    f"{m} right-hand sides, BF16 MFMA gather + FP32 refinement to 1e-5",
        "form_factors_ms": ff_ms, "form_factor_rays": ff["rays"],
        "form_factor_mrays_per_s": ff["rays"] / (ff_ms / 1e3) / 1e6,
        "solve_ms": solve_ms, "outer_iterations": info["outer_iterations"], "sweeps": info["sweeps"],
        "residual_fp32": info["residual"], "residual_fp64_host": res64,
        "error_vs_fp64": err64,
        "fp64_check": f"host FP64 LU (scipy getrf) of (I - diag(rho_c) F) for the 3 RGB channels of {n}^2, "
                      f"{m // 3} right-hand sides each, {lu_s:.1f} s; error = ||B - B64|| / ||B64||",
        "gemm": gemm,
    }, {"ps": ps, "counts": C, "rays": rays, "max_distance": max_distance}


def fp64_solve_error(C, rays, ps, B):
    """Relative error of B against an FP64 LU solve of (I - diag(rho_c) F) B_c = E_c per RGB
    channel c (rho is the same for every emission set, so three factorizations cover all
    right-hand sides)."""
    import scipy.linalg as sl
    t = time.perf_counter()
    n, m = ps.emission.shape
    R = ps.reflectance.astype(np.float64)
    E = ps.emission.astype(np.float64)
    ref = np.empty((n, m))
    for ch in range(3):
        cols = list(range(ch, m, 3))
        rho = R[:, ch]
        assert all(np.array_equal(R[:, c], rho) for c in cols)
        A = np.empty((n, n))
        for r0 in range(0, n, 2048):
            A[r0:r0 + 2048] = C[r0:r0 + 2048].astype(np.float64) * (-rho[r0:r0 + 2048, None] / rays)
        A[np.arange(n), np.arange(n)] += 1.0
        lu = sl.lu_factor(A, overwrite_a=True, check_finite=False)
        ref[:, cols] = sl.lu_solve(lu, E[:, cols], check_finite=False)
        del A, lu
    err = float(np.linalg.norm(B - ref) / np.linalg.norm(ref))
    return err, time.perf_counter() - t


def gi_oracle_check(check, stride=256, threads=0):
    """The gi leg's first pass against the oracle (bench.py's CPU leg): photon counts and k-NN
    gathered colour of every photonmap, and the final gather on every `stride`-th valid lumel of
    the 2048^2 map. Also times the oracle on that sample."""
    import oracle as O
    spec = check["spec"]
    t = time.perf_counter()
    S = O.OracleScene.from_spec(spec)
    maps, prays = S.emit_photons(spec, spec.indirect, threads)
    emit_s = time.perf_counter() - t
    counts_eq = all(np.array_equal(check["maps"][k][1], maps[k][1]) for k in range(len(maps)))
    color_eq = all(np.array_equal(check["maps"][k][0], maps[k][0]) for k in range(len(maps)))
    gathered_eq = all(np.array_equal(check["maps"][k][2], maps[k][2]) for k in range(len(maps)))
    texel, pos, nrm = S.lumels(spec, 0)
    sel = slice(0, None, stride)
    col = np.zeros((texel[sel].size, 3), np.float32)
    t = time.perf_counter()
    frays = S.final_gather(spec, pos[sel], nrm[sel], texel[sel], 0, spec.indirect, maps, col, threads)
    fg_s = time.perf_counter() - t
    g = check["fg"][texel[sel]].astype(np.float64)
    c = col.astype(np.float64)
    rmse = float(np.sqrt(np.mean((g - c) ** 2)) / np.sqrt(np.mean(c * c)))
    return {"photonmaps": len(maps), "photon_counts_equal": counts_eq, "photon_power_equal": color_eq,
            "knn_gathered_equal": gathered_eq, "fg_sample_lumels": int(texel[sel].size),
            "rmse_vs_oracle": rmse, "fg_bit_exact_frac": float(np.mean(np.all(g == c, axis=1))),
            "oracle_mrays_per_s": {"photons": prays / emit_s / 1e6, "final_gather": frays / fg_s / 1e6},
            "sample": f"every {stride}th valid lumel of the 2048^2 map ({texel[sel].size} lumels, {frays} rays)"}


def radiosity_oracle_check(check, rows=32, threads=0):
    """Form-factor rows of the radiosity leg against the oracle: `rows` evenly spaced rows of the
    16,384^2 hit-count matrix, compared exactly (bench.py's CPU leg)."""
    import oracle as O
    ps, C = check["ps"], check["counts"]
    S = O.OracleScene()
    S.add_mesh(ps.positions, ps.normals, np.zeros((ps.positions.shape[0], 2), np.float32), ps.indices)
    S.build()
    n = ps.num_patches
    idx = np.linspace(0, n - 1, rows).astype(int)
    t = time.perf_counter()
    same = 0
    for r in idx:
        Co = S.form_factors(ps.patches, ps.tri_patch, check["rays"], 1, check["max_distance"], rows=1, row0=int(r),
                            threads=threads)
        same += int(np.array_equal(Co[0], C[r]))
    return {"rows_checked": int(rows), "rows_equal": same, "form_factor_rows_equal": same == rows,
            "oracle_s": time.perf_counter() - t}


# ----------------------------------------------------------------------------------------- serving

def _serving_setup(relight, n_requests, seed, ao_samples, fg_samples):
    """Resident scenes, photon maps, lightmaps and radiosity objects, and the request list."""
    from relight_amd import AmbientOcclusionSettings, RL_SAMPLE_COSINE, scenes
    from relight_amd import patches as PT
    specs = [
        scenes.city(seed=21, num_boxes=1000, tiles=(4, 2), side_range=(128, 2048), ao_samples=ao_samples),
        scenes.city(seed=22, num_boxes=5000, tiles=(4, 4), side_range=(128, 2048), ao_samples=ao_samples),
        scenes.city(seed=23, num_boxes=12000, tiles=(4, 4), side_range=(128, 2048), ao_samples=ao_samples),
        scenes.heightfield(seed=24, cells=256, tile_cells=64, num_boxes=500, tile_map_sides=(128, 2048)),
        scenes.heightfield(seed=25, cells=480, tile_cells=120, num_boxes=2000, tile_map_sides=(128, 2048)),
    ]
    ao = AmbientOcclusionSettings(ao_samples, 0.8, 4.0, 1.0, 1, RL_SAMPLE_COSINE)
    built = []
    for k, sp in enumerate(specs):
        scenes.with_photonmaps(sp, size=64, photons_per_light=200_000, fg_samples=fg_samples,
                               knn_radius=2.0, seed=30 + k)
        t = time.time()
        scene, lms, ids = scenes.build_gpu(relight, sp)
        relight.emit_photons(scene, sp.indirect)    # resident photon maps
        relight.bake_items([(scene, lm, 0, 0, False, None) for lm in lms])  # resident lumels
        mesh_of = {}
        for mid, m in zip(ids, sp.meshes):
            if m.lightmap is not None and m.lightmap >= 0:
                mesh_of.setdefault(m.lightmap, mid)
        built.append((sp, scene, lms, mesh_of))
        log(f"serving: scene {k} {sp.name}: {sp.num_tris} tris, {len(lms)} lightmaps "
            f"({sum(lm.valid_count() for lm in lms)} texels) in {time.time()-t:.1f} s")
    patch_scenes = [PT.block_city(target=t, ground=32, lots=4, sets=4, emitters=16, seed=40 + i)
                    for i, t in enumerate((2048, 3072, 4096))]
    patch_built = [PT.build_gpu(relight, ps) for ps in patch_scenes]
    rng = np.random.default_rng(seed)
    pool = [(si, li) for si, (_, _, lms, _) in enumerate(built) for li in range(len(lms))]
    order = rng.permutation(len(pool))
    kinds = rng.choice(3, n_requests, p=[0.5, 0.3, 0.2])
    reqs, used = [], 0
    for kind in kinds:
        if kind == 2:
            pi = int(rng.integers(0, len(patch_scenes)))
            ps = patch_scenes[pi]
            rad = relight.create_radiosity(patch_built[pi][0], ps.patches, ps.tri_patch)
            rad.build_form_factors(256, seed=int(rng.integers(1 << 30)), max_distance=100.0)  # allocates F
            reqs.append({"kind": "radiosity", "rad": rad, "ps": ps, "seed": int(rng.integers(1 << 30))})
        else:
            if used >= len(pool):
                continue
            si, li = pool[order[used]]
            used += 1
            sp, scene, lms, mesh_of = built[si]
            reqs.append({"kind": "direct_ao" if kind == 0 else "final_gather", "scene": scene, "lm": lms[li],
                         "mesh": mesh_of[li], "ind": sp.indirect, "texels": lms[li].valid_count()})
    return built, reqs, ao


def _run_plain(relight, req, ao):
    """One request through the per-call entry points; returns rays cast."""
    if req["kind"] == "direct_ao":
        relight.bake_direct_light(req["scene"], req["mesh"])
        r = relight.stats()["rays"]
        relight.bake_ambient_occlusion(req["scene"], req["mesh"], ao)
        return r + relight.stats()["rays"]
    if req["kind"] == "final_gather":
        relight.bake_indirect_light(req["scene"], req["mesh"], req["ind"])
        return relight.stats()["rays"]
    ps = req["ps"]
    req["rad"].build_form_factors(256, seed=req["seed"], max_distance=100.0)
    r = relight.stats()["rays"]
    req["rad"].solve(ps.emission, ps.reflectance, tolerance=1e-5, max_outer=40, inner_sweeps=16)
    return r


def _submit(srv, req, ao):
    if req["kind"] == "direct_ao":
        return srv.submit_direct_ao(req["scene"], [req["lm"]], True, ao)
    if req["kind"] == "final_gather":
        return srv.submit_final_gather(req["scene"], [req["lm"]], req["ind"])
    ps = req["ps"]
    return srv.submit_radiosity(req["rad"], ps.emission, ps.reflectance, 256, req["seed"], 100.0, 1e-5)


def bench_serving(relight, n_requests=64, seed=5, ao_samples=64, fg_samples=64, load=0.8,
                  quantum_rays=1 << 27, max_active=32):
    t0 = time.time()
    built, reqs, ao = _serving_setup(relight, n_requests, seed, ao_samples, fg_samples)
    n = len(reqs)
    mix = {k: sum(1 for r in reqs if r["kind"] == k) for k in ("direct_ao", "final_gather", "radiosity")}
    texels = [r["texels"] for r in reqs if "texels" in r]
    log(f"serving: {n} requests {mix}, setup {time.time()-t0:.1f} s")

    # Saturated: every request available at t = 0.
    t = time.perf_counter()
    plain_done, plain_rays = [], 0
    for i, r in enumerate(reqs):
        plain_rays += _run_plain(relight, r, ao)
        plain_done.append((time.perf_counter() - t) * 1e3)
        if i % 16 == 15:
            log(f"serving: plain saturated {i+1}/{n} at {plain_done[-1]:.0f} ms")
    plain_ms = (time.perf_counter() - t) * 1e3
    with relight.create_server(quantum_rays=quantum_rays, max_active=max_active) as srv:
        t = time.perf_counter()
        jobs = [_submit(srv, r, ao) for r in reqs]
        sts = [j.wait(timeout=600) for j in jobs]
        served_ms = (time.perf_counter() - t) * 1e3
        sstats = srv.stats()
    assert all(s["state"] == "done" for s in sts), [s for s in sts if s["state"] != "done"][:3]
    served_lat = [s["done_ms"] - s["submit_ms"] for s in sts]
    log(f"serving: saturated plain {plain_ms:.0f} ms, served {served_ms:.0f} ms, steps {sstats['steps']}")

    # Poisson arrivals at `load` x the served saturated rate.
    rate = load * n / (served_ms / 1e3)
    arrivals = np.cumsum(np.random.default_rng(seed + 1).exponential(1.0 / rate, n))
    with relight.create_server(quantum_rays=quantum_rays, max_active=max_active) as srv:
        jobs = [None] * n
        t = time.perf_counter()

        def feeder():
            for i, r in enumerate(reqs):
                d = arrivals[i] - (time.perf_counter() - t)
                if d > 0:
                    time.sleep(d)
                jobs[i] = _submit(srv, r, ao)

        th = threading.Thread(target=feeder)
        th.start()
        th.join()
        sts_p = [j.wait(timeout=600) for j in jobs]
        poisson_served_ms = (time.perf_counter() - t) * 1e3
    poisson_served_lat = [s["done_ms"] - s["submit_ms"] for s in sts_p]
    log(f"serving: poisson served p50 {percentile(poisson_served_lat, 50):.1f} ms "
        f"p99 {percentile(poisson_served_lat, 99):.1f} ms")
    t = time.perf_counter()
    poisson_plain_lat = []
    for i, r in enumerate(reqs):
        d = arrivals[i] - (time.perf_counter() - t)
        if d > 0:
            time.sleep(d)
        _run_plain(relight, r, ao)
        poisson_plain_lat.append((time.perf_counter() - t) * 1e3 - arrivals[i] * 1e3)
    poisson_plain_ms = (time.perf_counter() - t) * 1e3
    log(f"serving: poisson plain p50 {percentile(poisson_plain_lat, 50):.1f} ms "
        f"p99 {percentile(poisson_plain_lat, 99):.1f} ms")
    return {
        "workload": f"{n} requests over 5 resident scenes (cities 10k-144k tris, heightfields 137k-485k tris), "
                    f"lightmaps 128^2-2048^2 (mean {np.mean(texels):.0f} texels), mix {mix}; "
                    f"AO {ao_samples} cosine rays + direct, FG {fg_samples} rays (k=64 gather), "
                    "radiosity 2048-4096 patches x 256 rays, 12 RHS",
        "saturated": {
            "plain_ms": plain_ms, "served_ms": served_ms,
            "plain_requests_per_s": n / (plain_ms / 1e3), "served_requests_per_s": n / (served_ms / 1e3),
            "plain_mrays_per_s": plain_rays / (plain_ms / 1e3) / 1e6,
            "served_mrays_per_s": plain_rays / (served_ms / 1e3) / 1e6,
            "served_vs_plain": plain_ms / served_ms,
            "served_latency_ms": {"p50": percentile(served_lat, 50), "p99": percentile(served_lat, 99)},
            "plain_latency_ms": {"p50": percentile(plain_done, 50), "p99": percentile(plain_done, 99)},
            "server_steps": sstats["steps"], "peak_memory_bytes": sstats["peak_memory"],
        },
        "poisson": {
            "rate_per_s": rate, "load_of_served_capacity": load,
            "served_latency_ms": {"p50": percentile(poisson_served_lat, 50),
                                  "p99": percentile(poisson_served_lat, 99)},
            "plain_latency_ms": {"p50": percentile(poisson_plain_lat, 50),
                                 "p99": percentile(poisson_plain_lat, 99)},
            "served_makespan_ms": poisson_served_ms, "plain_makespan_ms": poisson_plain_ms,
        },
        "rays": plain_rays,
    }


# ---------------------------------------------------------------------------------- serving_small

def bench_serving_small(relight, n_requests=256, ao_samples=64, quantum_rays=1 << 27, max_active=256):
    from relight_amd import AmbientOcclusionSettings, RL_SAMPLE_COSINE, scenes
    t0 = time.time()
    spec = scenes.city(seed=31, num_boxes=2500, tiles=(16, 16), side_range=(128, 256), ao_samples=ao_samples)
    scene, lms, ids = scenes.build_gpu(relight, spec)
    relight.bake_items([(scene, lm, 0, 0, False, None) for lm in lms])  # resident lumels
    ao = AmbientOcclusionSettings(ao_samples, 0.8, 4.0, 1.0, 1, RL_SAMPLE_COSINE)
    n = min(n_requests, len(lms))
    texels = [lms[i].valid_count() for i in range(n)]
    log(f"serving_small: {n} direct+AO requests, {sum(texels)} texels, setup {time.time()-t0:.1f} s")
    for i in range(2):  # warm both paths
        relight.bake_direct_light(scene, ids[i])
        relight.bake_ambient_occlusion(scene, ids[i], ao)
    t = time.perf_counter()
    plain_done, plain_rays = [], 0
    for i in range(n):
        relight.bake_direct_light(scene, ids[i])
        plain_rays += relight.stats()["rays"]
        relight.bake_ambient_occlusion(scene, ids[i], ao)
        plain_rays += relight.stats()["rays"]
        plain_done.append((time.perf_counter() - t) * 1e3)
    plain_ms = (time.perf_counter() - t) * 1e3
    with relight.create_server(quantum_rays=quantum_rays, max_active=max_active) as srv:
        t = time.perf_counter()
        jobs = [srv.submit_direct_ao(scene, [lms[i]], True, ao) for i in range(n)]
        sts = [j.wait(timeout=600) for j in jobs]
        served_ms = (time.perf_counter() - t) * 1e3
        sstats = srv.stats()
    assert all(s_["state"] == "done" for s_ in sts)
    served_lat = [s_["done_ms"] - s_["submit_ms"] for s_ in sts]
    log(f"serving_small: saturated plain {plain_ms:.1f} ms, served {served_ms:.1f} ms "
        f"({plain_ms / served_ms:.2f}x), steps {sstats['steps']}")
    # One small lightmap baked repeatedly: eager bake_batch calls vs replays of the recorded pass.
    reps = 50
    lm0 = [lms[0]]
    relight.bake_batch(scene, lm0, direct=True, ao=ao)
    t = time.perf_counter()
    for _ in range(reps):
        relight.bake_batch(scene, lm0, direct=True, ao=ao)
    eager_ms = (time.perf_counter() - t) * 1e3 / reps
    pas = relight.record_pass(scene, lm0, direct=True, ao=ao)
    pas.launch()
    t = time.perf_counter()
    for _ in range(reps):
        pas.launch()
    graph_ms = (time.perf_counter() - t) * 1e3 / reps
    pas.close()
    log(f"serving_small: one {lms[0].valid_count()}-texel lightmap, bake_batch {eager_ms:.3f} ms/pass, "
        f"recorded pass {graph_ms:.3f} ms/pass")
    return {
        "workload": f"{n} direct+AO requests ({ao_samples} cosine AO rays + sun and 16-vertex area light) on "
                    f"128^2-256^2 lightmaps of one resident city (2,500 boxes, mean {np.mean(texels):.0f} texels)",
        "plain_ms": plain_ms, "served_ms": served_ms, "served_vs_plain": plain_ms / served_ms,
        "plain_mrays_per_s": plain_rays / (plain_ms / 1e3) / 1e6,
        "served_mrays_per_s": plain_rays / (served_ms / 1e3) / 1e6,
        "served_latency_ms": {"p50": percentile(served_lat, 50), "p99": percentile(served_lat, 99)},
        "plain_latency_ms": {"p50": percentile(plain_done, 50), "p99": percentile(plain_done, 99)},
        "server_steps": sstats["steps"], "rays": plain_rays,
        "repeated_small_pass": {"texels": lms[0].valid_count(), "bake_batch_ms": eager_ms,
                                "recorded_pass_ms": graph_ms},
    }
