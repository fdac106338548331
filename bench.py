"""Benchmark: batched direct + AO bake of the procedural city on one MI355X.

One step = one full bake pass (direct light from the sun and the 16-vertex area light, then
256 cosine AO rays per texel) over all 32 lightmaps of the city, through rl_bake_batch (one launch
per pass). The plain path (one bake call per mesh per pass, the reference's per-mesh call shape)
is timed beside it. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "Mrays/s + bake p50/p99 latency, batched vs plain GPU at ≤1e-3 texel RMSE"
HBM_PEAK_GBS = 8000.0
# MI355X peak engine clock; a compute unit's four SIMDs execute 16 lanes of a VALU instruction
# per clock (a wave64 instruction issues over 4 clocks), so the VALU ceiling is 64 lane-ops per
# CU per clock.
ENGINE_CLOCK_GHZ = 2.4
MI355X_CUS = 256
ORACLE_BUILD = "g++ -O3 -march=x86-64-v3 -ffp-contract=off (oracle/Makefile)"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def load_traffic():
    path = os.path.join(HERE, "tests", "golden", "city_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def load_pmc(kernel, workload="city"):
    name = f"{kernel}_{workload}_pmc.json"
    path = os.path.join(HERE, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    d["file"] = name
    return d


def cpu_share():
    """Host threads this process may run on (the box's CPU share for one GPU)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def valu_roofline(pmc, avg_s, cus, texels):
    """VALU-issue roofline of a traversal kernel from its committed counters (same workload and
    kernel): useful lane-operations per launch (SQ_THREAD_CYCLES_VALU: active lanes summed over
    the VALU instructions) over the live launch time, against the VALU ceiling."""
    sq = pmc["sq"]
    lane_ops = sq["SQ_THREAD_CYCLES_VALU"]
    peak = cus * 64 * ENGINE_CLOCK_GHZ  # G lane-op/s
    achieved = lane_ops / avg_s / 1e9
    return {
        "bound": "valu_issue", "achieved": achieved, "peak": peak, "unit": "G lane-op/s",
        "frac": achieved / peak, "traffic": pmc["hbm_bytes_per_launch"],
        "counters": os.path.join("profiles", pmc["file"]),
        "lane_ops_per_launch": lane_ops,
        # Wave-instruction issue (1 = each of a CU's four SIMDs issues a wave64 VALU instruction
        # every 4 clocks) and the mean fraction of a VALU instruction's 64 lanes that are active;
        # frac = their product.
        "valu_issue_util": sq["SQ_INSTS_VALU"] / (cus * ENGINE_CLOCK_GHZ * 1e9 * avg_s),
        "lane_efficiency": sq["SQ_THREAD_CYCLES_VALU"] / (64.0 * sq["SQ_INSTS_VALU"]),
        "hbm_gbs_measured": pmc["hbm_bytes_per_launch"] / avg_s / 1e9,
        # Compulsory traffic: per texel the 24-B lumel (position, normal) read and the 12-B RGB write.
        "hbm_vs_compulsory": pmc["hbm_bytes_per_launch"] / (36.0 * texels),
        "l2_hit_rate": pmc["l2_hit_rate"],
    }


def percentile(x, q):
    return float(np.percentile(np.asarray(x, np.float64), q))


def cpu_baseline(spec, gpu_maps, stride, threads):
    """Times the oracle on every `stride`-th valid lumel of every lightmap (direct + AO) and
    returns (baseline dict, relative RMSE of the GPU's first pass on the same texels)."""
    import oracle as O

    S = O.OracleScene.from_spec(spec)
    rays = 0
    elapsed = 0.0
    num = den = 0.0
    lumels = 0
    for li in range(len(spec.lightmaps)):
        texel, pos, nrm = S.lumels(spec, li)
        sel = slice(li % stride, None, stride)
        t, p, n = texel[sel], pos[sel], nrm[sel]
        col = np.zeros((t.size, 3), np.float32)
        t0 = time.perf_counter()
        rays += S.bake_direct(p, n, col, threads)
        _, r = S.bake_ao(p, n, t, li, spec.ao, col, threads)
        elapsed += time.perf_counter() - t0
        rays += r
        lumels += t.size
        g = gpu_maps[li].reshape(-1, 3)[t].astype(np.float64)
        num += float(((g - col) ** 2).sum())
        den += float((col.astype(np.float64) ** 2).sum())
    base = {"value": rays / elapsed / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"oracle direct+AO on every {stride}th valid lumel of all 32 lightmaps "
                      f"({lumels} lumels, {rays} rays, {elapsed:.1f} s)",
            "cpu_model": cpu_model(), "build": ORACLE_BUILD,
            "threads_note": "all host threads this process may use (sched_getaffinity)"}
    rmse = (num / den) ** 0.5 if den > 0 else 0.0
    return base, rmse, lumels


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-stride", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ao-samples", type=int, default=256)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--workloads", default="gi,radiosity,serving,serving_small",
                    help="secondary legs reported under 'workloads' (comma list of gi, radiosity, "
                         "serving, serving_small; empty for none)")
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device with a plain copy kernel (torch) and exit; "
                         "used to scale rocprof FETCH_SIZE/WRITE_SIZE readings")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 4 * n}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the bake path "
              "would shard lightmaps across GPUs with no exchange; this build runs on one).",
              file=sys.stderr)
        sys.exit(2)

    from relight_amd import Relight, scenes

    t0 = time.time()
    spec = scenes.city(ao_samples=args.ao_samples)
    log(f"city generated: {spec.num_tris} tris, {len(spec.lightmaps)} lightmaps ({time.time()-t0:.1f} s)")
    relight = Relight()
    t0 = time.time()
    scene, lms, ids = scenes.build_gpu(relight, spec)
    info = scene.info()
    texels = sum(lm.valid_count() for lm in lms)
    log(f"scene built: {info}, {texels} valid texels ({time.time()-t0:.1f} s)")

    tuning, tune_results = None, None
    if not args.no_autotune:
        from relight_amd.autotune import autotune
        t0 = time.time()
        # Probe passes at the timed pass's own sample count (the autotuner's default probes at 64).
        tuning, tune_results = autotune(relight, scene, lms, spec.ao, probe_samples=int(spec.ao.samples))
        log(f"autotune ({time.time()-t0:.1f} s): {tuning}; "
            + ", ".join(f"{r[1]:.1f}" for r in tune_results))
    # Check pass (also the first warmup): one batched pass from zero, kept for the parity check.
    relight.bake_batch(scene, lms, direct=True, ao=spec.ao)
    first = [lm.to_hdr() for lm in lms]
    log(f"check pass done: {relight.stats()['rays']} rays")
    for _ in range(max(0, args.warmup - 1)):
        relight.bake_batch(scene, lms, direct=True, ao=spec.ao)

    lat, rays, ao_ms, ao_launch, ao_rays, gpu_ms = [], [], 0.0, 0, 0, []
    for s in range(args.steps):
        t = time.perf_counter()
        relight.bake_batch(scene, lms, direct=True, ao=spec.ao)
        lat.append((time.perf_counter() - t) * 1e3)
        st = relight.stats()
        rays.append(st["rays"])
        gpu_ms.append(st["gpu_ms"])
        k = st["kernels"].get("ao_count", {"ms": 0.0, "launches": 0})
        ao_ms += k["ms"]
        ao_launch += k["launches"]
        ao_rays += texels * args.ao_samples
        log(f"step {s}: {lat[-1]:.1f} ms, {st['rays']} rays, kernels {st['kernels']}")
    total_ms = float(np.sum(lat))
    value = float(np.sum(rays)) / (total_ms / 1e3) / 1e6

    plain_lat, plain_rays = [], []
    for s in range(args.plain_steps):
        t = time.perf_counter()
        r = 0
        for mid in ids:
            relight.bake_direct_light(scene, mid)
            r += relight.stats()["rays"]
        for mid in ids:
            relight.bake_ambient_occlusion(scene, mid, spec.ao)
            r += relight.stats()["rays"]
        plain_lat.append((time.perf_counter() - t) * 1e3)
        plain_rays.append(r)
        log(f"plain step {s}: {plain_lat[-1]:.1f} ms, {r} rays")
    plain_value = (float(np.sum(plain_rays)) / (float(np.sum(plain_lat)) / 1e3) / 1e6
                   if plain_lat else None)

    # Roofline of the dominant kernel (the AO traversal): VALU issue, from its committed counters.
    traffic = load_traffic()
    roofline = None
    kernel = "k_ao_refill" if relight.tuning()["ao_kernel"] == 1 else "k_ao_count"
    pmc = load_pmc(kernel)
    if ao_launch and pmc:
        cus = MI355X_CUS
        avg_s = ao_ms / ao_launch / 1e3
        rays_per_launch = ao_rays / ao_launch
        roofline = valu_roofline(pmc, avg_s, cus, texels)
        roofline.update({"kernel": kernel, "avg_launch_ms": avg_s * 1e3, "rays_per_launch": rays_per_launch})
        if traffic:
            # Work metric: bytes a binary BVH of the oracle's kind would fetch per ray (node and
            # triangle visits counted on the oracle's traversal of this workload's rays), not HBM
            # traffic: the BVH stays in the caches.
            bpr = traffic["ao"]["bytes_per_ray"]
            roofline["work_metric"] = {"bytes_per_ray": bpr, "gbs": rays_per_launch * bpr / avg_s / 1e9,
                                       "source": "tests/golden/city_traffic.json"}

    out = {
        "metric": METRIC, "value": value, "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": total_ms / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {
            "workload": "city: 200x200 m ground + 5000 boxes (seed 7), "
                        f"{spec.num_tris} tris, 32 lightmaps 256^2-2048^2, {texels} valid texels; "
                        "sun + 2x2 m area light (16 vertices); AO "
                        f"{args.ao_samples} cosine rays/texel, max 4 m",
            "lightmaps": len(lms), "texels": texels, "triangles": spec.num_tris,
            "bvh_nodes": info["nodes"], "path": "rl_bake_batch (direct + AO)"},
        "latency_ms": {"p50": percentile(lat, 50), "p99": percentile(lat, 99)},
        "texels_per_s": texels / (total_ms / args.steps / 1e3),
        "rays_per_step": float(np.mean(rays)),
        "gpu_ms_per_step": float(np.mean(gpu_ms)),
        "plain": {"value": plain_value, "unit": "Mrays/s",
                  "latency_ms": {"p50": percentile(plain_lat, 50), "p99": percentile(plain_lat, 99)}
                  if plain_lat else None},
        "batched_vs_plain": (value / plain_value) if plain_value else None,
        "rmse_vs_oracle": None,
        "tuning": relight.tuning(),
        "roofline": roofline,
        "cpu_baseline": None,
    }
    # Secondary legs (each on its own resident scenes), after the headline measurement.
    import bench_workloads as W
    legs, checks = {}, {}
    for name in [w for w in args.workloads.split(",") if w]:
        t0 = time.time()
        if name == "gi":
            checks["gi"], legs["gi"] = W.bench_gi(relight)
        elif name == "radiosity":
            legs["radiosity"], checks["radiosity"] = W.bench_radiosity(relight)
        elif name == "serving":
            legs["serving"] = W.bench_serving(relight, n_requests=args.serve_requests)
        elif name == "serving_small":
            legs["serving_small"] = W.bench_serving_small(relight)
        else:
            raise SystemExit(f"unknown workload {name}")
        log(f"{name} leg done in {time.time()-t0:.1f} s")
    if legs:
        out["workloads"] = legs
    # CPU leg: the oracle timed on a sample of the headline workload, and every GPU result of this
    # run compared with it on the same inputs.
    if not args.no_cpu:
        threads = cpu_share()
        cpu, rmse, n = cpu_baseline(spec, first, args.cpu_stride, threads)
        out["cpu_baseline"], out["rmse_vs_oracle"] = cpu, rmse
        log(f"cpu baseline: {cpu}, rel RMSE on {n} texels: {rmse:.3e}")
        if "gi" in checks:
            t0 = time.time()
            legs["gi"]["parity"] = W.gi_oracle_check(checks["gi"], threads=threads)
            legs["gi"]["rmse_vs_oracle"] = legs["gi"]["parity"]["rmse_vs_oracle"]
            log(f"gi oracle check ({time.time()-t0:.1f} s): {legs['gi']['parity']}")
        if "radiosity" in checks:
            t0 = time.time()
            legs["radiosity"]["parity"] = W.radiosity_oracle_check(checks["radiosity"], threads=threads)
            log(f"radiosity oracle check ({time.time()-t0:.1f} s): {legs['radiosity']['parity']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
