// Texel raster of a mesh into a lightmap on the GPU: Lightmap::initializeLumels (Lightmap.cpp:118-164)
// with Face::uvRect / isUvInside / positionAt / normalAt (scene/Mesh.cpp:397-476).
//
// The reference visits faces in order and every inside texel takes the face's values, so a texel
// ends with the last face that contains it. Pass 1 (one wavefront per face) claims texels with
// atomicMax(face + 1); pass 2 recomputes the winning face's barycentrics at each claimed texel (the
// same arithmetic, rounded as written) and appends one record per texel.
#include <hip/hip_runtime.h>

#include "rl_kernels.h"
#include "rl_traverse.h"

namespace rl {

namespace {

using namespace trav;

struct FaceUv {
    float ax, ay, bx, by, cx, cy;
    int x1, x2, y1, y2;  // texel rectangle, clamped to the map
};

__device__ __forceinline__ FaceUv face_uv(const RasterMesh& M, uint32_t f) {
    const uint32_t ia = M.idx[3 * f], ib = M.idx[3 * f + 1], ic = M.idx[3 * f + 2];
    FaceUv F;
    F.ax = M.uv[2 * ia]; F.ay = M.uv[2 * ia + 1];
    F.bx = M.uv[2 * ib]; F.by = M.uv[2 * ib + 1];
    F.cx = M.uv[2 * ic]; F.cy = M.uv[2 * ic + 1];
    const float minx = fminf(F.ax, fminf(F.bx, F.cx)), maxx = fmaxf(F.ax, fmaxf(F.bx, F.cx));
    const float miny = fminf(F.ay, fminf(F.by, F.cy)), maxy = fmaxf(F.ay, fmaxf(F.by, F.cy));
    F.x1 = max((int)(minx * M.w), 0);
    F.x2 = min((int)(maxx * M.w), M.w - 1);
    F.y1 = max((int)(miny * M.h), 0);
    F.y2 = min((int)(maxy * M.h), M.h - 1);
    return F;
}

// Face::isUvInside at the texel centre: barycentric weights (bu on c, bv on b) and the test.
__device__ __forceinline__ bool inside(const FaceUv& F, int u, int v, int w, int h, float& bu, float& bv) {
    const float uvx = ((float)u + 0.5f) / (float)w, uvy = ((float)v + 0.5f) / (float)h;
    const float v0x = F.cx - F.ax, v0y = F.cy - F.ay;
    const float v1x = F.bx - F.ax, v1y = F.by - F.ay;
    const float v2x = uvx - F.ax, v2y = uvy - F.ay;
    const float dot00 = v0x * v0x + v0y * v0y;
    const float dot01 = v0x * v1x + v0y * v1y;
    const float dot02 = v0x * v2x + v0y * v2y;
    const float dot11 = v1x * v1x + v1y * v1y;
    const float dot12 = v1x * v2x + v1y * v2y;
    const float inv = 1.0f / (dot00 * dot11 - dot01 * dot01);
    bu = (dot11 * dot02 - dot01 * dot12) * inv;
    bv = (dot00 * dot12 - dot01 * dot02) * inv;
    return (bu >= 0.0f) && (bv >= 0.0f) && (bu + bv < 1.0f);
}

__global__ __launch_bounds__(256) void k_raster_claim(RasterMesh M, uint32_t* __restrict__ owner,
                                                      unsigned long long* queue) {
    const uint32_t lane = lane_id();
    for (;;) {
        const uint64_t f = wave_next_chunk(queue);
        if (f >= M.nfaces) break;
        const FaceUv F = face_uv(M, (uint32_t)f);
        if (F.x2 < F.x1 || F.y2 < F.y1) continue;
        const uint32_t nx = (uint32_t)(F.x2 - F.x1 + 1), ny = (uint32_t)(F.y2 - F.y1 + 1);
        const uint32_t n = nx * ny;  // maps are at most 16384 texels a side
        for (uint32_t k = lane; k < n; k += 64) {
            const int u = F.x1 + (int)(k % nx), v = F.y1 + (int)(k / nx);
            float bu, bv;
            if (inside(F, u, v, M.w, M.h, bu, bv)) atomicMax(&owner[(size_t)v * M.w + u], (uint32_t)f + 1u);
        }
    }
}

__global__ void k_raster_fill(RasterMesh M, const uint32_t* __restrict__ owner, RasterRec* __restrict__ out,
                              unsigned int* __restrict__ count) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)M.w * M.h) return;
    const uint32_t o = owner[t];
    if (o == 0) return;
    const uint32_t f = o - 1u;
    const FaceUv F = face_uv(M, f);
    const int u = (int)(t % (uint64_t)M.w), v = (int)(t / (uint64_t)M.w);
    float bu, bv;
    inside(F, u, v, M.w, M.h, bu, bv);
    const uint32_t ia = M.idx[3 * f], ib = M.idx[3 * f + 1], ic = M.idx[3 * f + 2];
    const float* pa = M.pos + 3 * ia; const float* pb = M.pos + 3 * ib; const float* pc = M.pos + 3 * ic;
    const float* na = M.nrm + 3 * ia; const float* nb = M.nrm + 3 * ib; const float* nc = M.nrm + 3 * ic;
    RasterRec r;
    r.texel = (uint32_t)t;
    r.tri = M.tri_base + f;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        r.pos[k] = pa[k] + (pb[k] - pa[k]) * bv + (pc[k] - pa[k]) * bu;
        r.nrm[k] = na[k] + (nb[k] - na[k]) * bv + (nc[k] - na[k]) * bu;
    }
    out[atomicAdd(count, 1u)] = r;
}

}  // namespace

hipError_t launch_raster(hipStream_t st, const RasterMesh& M, uint32_t max_blocks, uint32_t* owner, RasterRec* out,
                         unsigned int* count, unsigned long long* queue) {
    const uint64_t texels = (uint64_t)M.w * M.h;
    if (M.nfaces == 0 || texels == 0) return hipSuccess;
    hipError_t e = launch_fill_u32(st, owner, texels, 0u);
    if (e == hipSuccess) e = launch_fill_u32(st, count, 1, 0u);
    if (e == hipSuccess) e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(queue), 2, 0u);
    if (e != hipSuccess) return e;
    const uint64_t need = ((uint64_t)M.nfaces + 3) / 4;
    const uint32_t grid = (uint32_t)(need < max_blocks ? (need ? need : 1) : max_blocks);
    hipLaunchKernelGGL(k_raster_claim, dim3(grid), dim3(256), 0, st, M, owner, queue);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_raster_fill, dim3((uint32_t)((texels + 255) / 256)), dim3(256), 0, st, M, owner, out, count);
    return hipGetLastError();
}

}  // namespace rl
