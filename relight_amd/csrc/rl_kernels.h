// Host-callable launchers of the HIP kernels (rl_kernels.hip, rl_photon.hip, rl_radiosity.hip).
//
// Ray kernels are persistent: `max_blocks` caps the grid (the runtime sizes it to the device's
// compute units) and `queue` is a device u64 work counter the launcher clears first.
#pragma once

#include <hip/hip_runtime.h>

#include "rl_device.h"

namespace rl {

// LDS traversal-stack entries per lane for a BVH whose worst-case stack bound is `bound`;
// -1 when no compiled capacity is large enough.
int stack_capacity_for(int bound);

hipError_t launch_fill_u32(hipStream_t st, uint32_t* p, uint64_t n, uint32_t v);

// Copies lumels sel[0..n) of a lightmap's compact arrays to a compact segment, and the segment's
// colours back (bakes of an iterator's subset of a lightmap).
hipError_t launch_gather_lumels(hipStream_t st, const uint32_t* sel, uint32_t n, const float4* pos, const float4* nrm,
                                const float4* col, float4* opos, float4* onrm, float4* ocol);
hipError_t launch_scatter_colors(hipStream_t st, const uint32_t* sel, uint32_t n, const float4* col, float4* out);

// Segments carry their scene, lights and AO settings; `cap` is the largest traversal-stack
// capacity any of their scenes needs. work_begin units: lumels (direct, finalize), (lumel,
// sample) pairs (AO chunked) or chunks of chunk_lumels lumels (AO refill).
// alpha: some segment's scene has a material whose alpha can be <= 0.1, so shadow rays take the
// closest-hit alpha re-trace (HitUseAlpha) instead of any-hit.
// shortcuts: kShortcut* bits, exact shortcuts of the any-hit kernels (local trees for AO rays,
// occluder hints for AO and shadow rays); results do not depend on them.
constexpr uint32_t kShortcutTrees = 1u, kShortcutHints = 2u;
hipError_t launch_bake_direct(hipStream_t st, int cap, uint32_t max_blocks, const BakeSeg* segs,
                              int nsegs, uint64_t total, bool alpha, uint32_t shortcuts, unsigned long long* queue,
                              unsigned long long* rays);
// Direct light of small passes over (lumel, shadow ray) pairs: influences into `inf` (one float per
// pair, pairs counted from pair_segs' work_begin), then summed per lumel in the reference's light
// order (lumel_segs: the same segments in lumel units). Results equal launch_bake_direct's.
hipError_t launch_bake_direct_pairs(hipStream_t st, int cap, uint32_t max_blocks, const BakeSeg* pair_segs,
                                   const BakeSeg* lumel_segs, int nsegs, uint64_t total_pairs, uint64_t total_lumels,
                                   bool alpha, uint32_t shortcuts, float* inf, unsigned long long* queue,
                                   unsigned long long* rays);

hipError_t launch_ao_count(hipStream_t st, int cap, uint32_t max_blocks, const BakeSeg* segs,
                           int nsegs, uint64_t total_pairs, unsigned long long* queue,
                           unsigned long long* rays);

// AO with lane refill (counts written, not accumulated); chunks of at most kAoChunkMax lumels.
// The kernel runs one workgroup of kAoRefillWavesPerCu wavefronts per compute unit; passes with
// fewer chunks than that many wavefronts on the device use smaller chunks, down to kAoSmallChunk.
constexpr uint32_t kAoChunkMax = 16;
constexpr uint32_t kAoRefillWavesPerCu = 16;
constexpr uint32_t kAoSmallChunk = 4;
hipError_t launch_ao_refill(hipStream_t st, int cap, uint32_t max_blocks, const BakeSeg* segs,
                            int nsegs, uint64_t total_chunks, uint32_t chunk_lumels,
                            uint32_t refill_min, uint32_t shortcuts, unsigned long long* queue,
                            unsigned long long* rays);

hipError_t launch_ao_finalize(hipStream_t st, const BakeSeg* segs, int nsegs, uint64_t total);

hipError_t launch_trace(hipStream_t st, const SceneDev& S, int cap, uint32_t max_blocks, bool any,
                        const float4* starts, const float4* ends, uint32_t n, uint8_t* flags,
                        float4* hits, unsigned long long* queue, unsigned long long* rays);

// Radiosity (rl_radiosity.hip).
hipError_t launch_form_factor_rows(hipStream_t st, const SceneDev& S, int cap, const PatchDev* patches,
                                   const uint32_t* tri_patch, const FormFactorParams& P, uint16_t* F,
                                   unsigned long long* rays);
hipError_t launch_ff_gemm(hipStream_t st, bool accurate, const uint16_t* F, uint32_t stride,
                          const __bf16* X1, const __bf16* X2, const __bf16* X3, uint32_t npad,
                          uint32_t mpad, uint32_t nsplit, float* Yp);
hipError_t launch_ff_untile(hipStream_t st, const uint16_t* F, uint32_t npad, uint32_t n, uint32_t first,
                           uint32_t rows, uint16_t* out);
hipError_t launch_ff_to_cols(hipStream_t st, const float* in, uint32_t n, uint32_t m, uint32_t npad, uint32_t mpad,
                            float* out);
hipError_t launch_ff_to_rows(hipStream_t st, const float* cols, uint32_t n, uint32_t m, uint32_t npad, float* out);
hipError_t launch_ff_residual(hipStream_t st, const float* Yp, uint32_t nsplit, uint32_t npad,
                              uint32_t mpad, float inv_rays, const float* E, const float* rho,
                              const float* B, float* R, float* D, __bf16* X1, float* part);
hipError_t launch_ff_sweep(hipStream_t st, const float* Yp, uint32_t nsplit, uint32_t npad,
                           uint32_t mpad, float inv_rays, const float* rho, const float* R, float* D,
                           __bf16* X1);
hipError_t launch_ff_apply(hipStream_t st, uint32_t npad, uint32_t mpad, float* B, const float* D,
                           int init, __bf16* X1, __bf16* X2, __bf16* X3);

// Photons (rl_photon.hip).
hipError_t launch_photon_trace(hipStream_t st, const SceneDev& S, int cap, uint32_t max_blocks,
                               const LightDev* lights, const uint64_t* light_first,
                               const PhotonParams& P, const int32_t* mesh_pm,
                               const PhotonMapDev* pms, PhotonRec* out, float4* out_pos,
                               unsigned long long* queue, unsigned long long* rays);
// Photons::store of n records (sorted by texel key with the radix sort of rl_sort.h, scratch
// sort_pairs_temp_bytes(n)); total_keys = texels of all photonmaps (the keys' upper bound).
hipError_t launch_photon_store(hipStream_t st, const PhotonRec* rec, uint32_t n, uint32_t* keys_a,
                               uint32_t* keys_b, uint32_t* idx_a, uint32_t* idx_b, void* temp,
                               size_t temp_bytes, const PhotonMapDev* pms, int npm, uint32_t total_keys);
hipError_t launch_photon_gather(hipStream_t st, const PhotonMapDev* pms, int npm,
                                uint32_t total_texels, int radius);
// k-nearest gather: hash grid over the stored records (key_bits = log2(mask + 1) + 1), then one
// wavefront per query (float4: surface point, w = global texel key bits).
hipError_t launch_knn_build(hipStream_t st, const PhotonRec* rec, const float4* ppos, uint32_t n,
                            const KnnGrid& G, int key_bits, uint32_t* keys_a, uint32_t* keys_b, uint32_t* idx_a,
                            uint32_t* idx_b, void* temp, size_t temp_bytes, float4* spos, uint32_t* cstart,
                            uint32_t* cend);
hipError_t launch_knn_gather(hipStream_t st, uint32_t max_blocks, const float4* queries, uint32_t nq,
                             const float4* spos, const PhotonRec* rec, const uint32_t* cstart,
                             const uint32_t* cend, const KnnGrid& G, const PhotonMapDev* pms, int npm,
                             unsigned long long* queue);
// Final gather over lumel segments; each segment carries its scene, settings and photon tables.
// Final-gather kernel choice, the values of RL_FG_KERNEL_AUTO / _WAVE / _LANE.
constexpr int kFgKernelAuto = 0, kFgKernelWave = 1, kFgKernelLane = 2;
// kernel: kFgKernel*.
hipError_t launch_final_gather(hipStream_t st, int cap, uint32_t max_blocks, const BakeSeg* segs, int nsegs,
                               uint64_t total_lumels, int kernel, unsigned long long* queue, unsigned long long* rays);

// Texel raster of one mesh (rl_raster.hip): records of the texels it claims (last face wins).
hipError_t launch_raster(hipStream_t st, const RasterMesh& M, uint32_t max_blocks, uint32_t* owner, RasterRec* out,
                         unsigned int* count, unsigned long long* queue);

// Lightmap post-process and encodings (rl_post.hip) on a full w x h grid.
hipError_t launch_expand(hipStream_t st, float4* grid, const uint8_t* valid, int w, int h);
hipError_t launch_blur(hipStream_t st, float4* grid, const uint8_t* valid, int w, int h);
hipError_t launch_encode(hipStream_t st, const float4* grid, uint64_t n, int format, uint8_t* out);

}  // namespace rl
