"""Launch autotuner: picks the AO kernel variant, the refill threshold (rays per wavefront that must
be idle before they are refilled), the AO chunk size (lumels per queue entry) and the persistent
grid size (blocks per compute unit) for a scene by timing short bake passes on the GPU.

Results are bit-identical for every tuning (work assignment changes, per-texel arithmetic does
not), so tuning never changes a lightmap. Timing uses the library's HIP-event kernel times.
"""
import itertools

from ._capi import RL_AO_KERNEL_CHUNKED, RL_AO_KERNEL_REFILL
from .relight import AmbientOcclusionSettings

DEFAULT_GRID = {
    "ao_kernel": [RL_AO_KERNEL_REFILL],
    "refill_min": [8, 16, 20, 24, 32],
    "ao_chunk_lumels": [4, 16],
    "blocks_per_cu": [8, 16],
}


def candidates(grid=None):
    grid = grid or DEFAULT_GRID
    keys = list(grid)
    out = [dict(zip(keys, vals)) for vals in itertools.product(*(grid[k] for k in keys))]
    out.append({"ao_kernel": RL_AO_KERNEL_CHUNKED, "refill_min": 16, "ao_chunk_lumels": 8,
                "blocks_per_cu": 16})
    return out


def autotune(relight, scene, lightmaps, ao: AmbientOcclusionSettings, probe_samples=None,
             grid=None, repeats=1):
    """Times an AO pass with `probe_samples` rays per texel for each candidate; keeps the fastest.
    Returns (best tuning, [(tuning, ms)]). By default the probe uses min(ao.samples, 64) rays per
    texel: the refill kernel pairs rays with lumels one way for 64 or more samples per lumel and
    another way below, so a probe of at least 64 times the production path of a large pass.

    The probe passes bake AO into `lightmaps`: AO multiplies the texel colour, so call this
    before baking them (zero colours stay zero)."""
    if probe_samples is None:
        probe_samples = min(int(ao.samples), 64)
    probe = AmbientOcclusionSettings(probe_samples, ao.occluded_fraction, ao.max_distance,
                                     ao.exponent, ao.seed, ao.sampling)
    results = []
    for cand in candidates(grid):
        relight.set_tuning(**cand)
        best = None
        for _ in range(repeats + 1):  # first pass warms caches
            relight.bake_batch(scene, lightmaps, direct=False, ao=probe)
            ms = relight.stats()["kernels"].get("ao_count", {"ms": 0.0})["ms"]
            best = ms if best is None else min(best, ms)
        results.append((cand, best))
    best_cand = min(results, key=lambda r: r[1])[0]
    relight.set_tuning(**best_cand)
    return best_cand, results
