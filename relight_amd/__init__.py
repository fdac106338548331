"""relight_amd — MI355X lightmap baker for the ray-cast hot path of dmsovetov/Relight.

The product is librelight_amd.so (HIP kernels for gfx950 + a C-ABI, include/relight.h); this
package mirrors the reference's Relight / Scene / Lightmap API over that ABI and provides the
seeded scene generators the tests and the benchmark use.
"""
from ._capi import (RL_FORMAT_RAW_HDR, RL_FORMAT_TGA_DOUBLE_LDR, RL_FORMAT_TGA_RGBM,
                    RL_GATHER_KNN, RL_GATHER_TEXEL_DISC, RL_LIGHT_AREA, RL_LIGHT_DIRECTIONAL, RL_LIGHT_POINT, RL_LIGHT_SPOT,
                    RL_PHOTON_RUSSIAN_ROULETTE, RL_PHOTON_THRESHOLD, RL_SAMPLE_COSINE,
                    RL_SAMPLE_UNIFORM, RL_TRACE_ANY, RL_TRACE_CLOSEST)
from .relight import (AmbientOcclusionSettings, BakeJob, IndirectLightSettings, LightDesc,
                      Lightmap, Pass, Photonmap, Radiosity, Relight, RelightError, Scene, Server, Texture,
                      generate_uv)

__all__ = [
    "Relight", "Scene", "Lightmap", "Photonmap", "LightDesc", "AmbientOcclusionSettings",
    "IndirectLightSettings", "RelightError", "Radiosity", "Server", "BakeJob", "Pass", "Texture", "generate_uv",
    "RL_LIGHT_POINT", "RL_LIGHT_SPOT", "RL_LIGHT_DIRECTIONAL", "RL_LIGHT_AREA",
    "RL_SAMPLE_UNIFORM", "RL_SAMPLE_COSINE", "RL_TRACE_ANY", "RL_TRACE_CLOSEST",
    "RL_PHOTON_THRESHOLD", "RL_PHOTON_RUSSIAN_ROULETTE", "RL_GATHER_TEXEL_DISC", "RL_GATHER_KNN",
    "RL_FORMAT_RAW_HDR", "RL_FORMAT_TGA_DOUBLE_LDR", "RL_FORMAT_TGA_RGBM",
]
