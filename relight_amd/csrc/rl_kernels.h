// Host-callable launchers of the kernels in rl_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "rl_device.h"

namespace rl {

// LDS traversal-stack entries per lane for a BVH whose worst-case stack bound is `bound`;
// -1 when no compiled capacity is large enough.
int stack_capacity_for(int bound);

hipError_t launch_fill_u32(hipStream_t st, uint32_t* p, uint64_t n, uint32_t v);

hipError_t launch_bake_direct(hipStream_t st, const SceneDev& S, int cap, const BakeSeg* segs,
                              int nsegs, uint64_t total, const LightDev* lights, int nlights,
                              const float4* lverts, unsigned long long* rays);

hipError_t launch_ao_count(hipStream_t st, const SceneDev& S, int cap, const BakeSeg* segs,
                           int nsegs, uint64_t total_pairs, const AoParams& P,
                           unsigned long long* rays);

hipError_t launch_ao_finalize(hipStream_t st, const BakeSeg* segs, int nsegs, uint64_t total,
                              const AoParams& P);

hipError_t launch_trace(hipStream_t st, const SceneDev& S, int cap, bool any, const float4* starts,
                        const float4* ends, uint32_t n, uint8_t* flags, float4* hits,
                        unsigned long long* rays);

// Photons (rl_photon.hip).
hipError_t launch_photon_trace(hipStream_t st, const SceneDev& S, int cap, const LightDev* lights,
                               const uint64_t* light_first, const PhotonParams& P,
                               const int32_t* mesh_pm, const PhotonMapDev* pms, PhotonRec* out,
                               unsigned long long* rays);
hipError_t photon_sort_temp_bytes(uint32_t n, size_t* bytes);
hipError_t launch_photon_store(hipStream_t st, const PhotonRec* rec, uint32_t n, uint32_t* keys_a,
                               uint32_t* keys_b, uint32_t* idx_a, uint32_t* idx_b, void* temp,
                               size_t temp_bytes, const PhotonMapDev* pms, int npm);
hipError_t launch_photon_gather(hipStream_t st, const PhotonMapDev* pms, int npm,
                                uint32_t total_texels, int radius);
hipError_t launch_final_gather(hipStream_t st, const SceneDev& S, int cap, const BakeSeg* segs,
                               int nsegs, uint64_t total_lumels, const FinalGatherParams& P,
                               const int32_t* mesh_pm, const PhotonMapDev* pms,
                               unsigned long long* rays);

}  // namespace rl
