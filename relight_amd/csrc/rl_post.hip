// Lightmap post-process and output encodings on the GPU: Lightmap::expand / blur
// (Lightmap.cpp:167-220) and the LDR encodings of Lightmap::save (Lightmap.cpp:223-331 with
// Rgb::doubleLdr / Rgb::rgbm, Types.h:216-234), on a full w x h colour grid (float4, w unused)
// and the lumel validity grid.
#include <hip/hip_runtime.h>

#include "rl_kernels.h"

namespace rl {

namespace {

__device__ __forceinline__ bool interior(int x, int y, int w, int h) { return x >= 1 && y >= 1 && x <= w - 2 && y <= h - 2; }

// Lightmap::expand: every valid interior texel writes its colour to its four invalid diagonal
// neighbours, in raster order, so an invalid texel ends with the colour of its last writer: the
// first valid interior texel of (x+1,y+1), (x-1,y+1), (x+1,y-1), (x-1,y-1). Writers keep their
// colours (only invalid texels change), so every texel is computed independently.
__global__ void k_expand(float4* __restrict__ grid, const uint8_t* __restrict__ valid, int w, int h) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (uint64_t)w * h) return;
    const int x = (int)(g % (uint64_t)w), y = (int)(g / (uint64_t)w);
    if (valid[g]) return;
    const int dx[4] = {1, -1, 1, -1}, dy[4] = {1, 1, -1, -1};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int sx = x + dx[k], sy = y + dy[k];
        if (interior(sx, sy, w, h) && valid[(size_t)sy * w + sx]) {
            grid[g] = grid[(size_t)sy * w + sx];
            return;
        }
    }
}

// Lightmap::blur runs in place in raster order: texel (x, y) reads the already-blurred (x-1, y),
// (x-1, y-1), (x, y-1) and (x+1, y-1). Those have x' + 2y' < x + 2y and every neighbour it reads
// unblurred has x' + 2y' > x + 2y, so the texels of one anti-diagonal t = x + 2y are independent:
// one workgroup walks t with a barrier per step; thread k owns rows k, k + 1024, ...
__global__ __launch_bounds__(1024) void k_blur(float4* __restrict__ grid, const uint8_t* __restrict__ valid, int w,
                                               int h) {
    const int tmax = (w - 2) + 2 * (h - 2);
    for (int t = 3; t <= tmax; t++) {
        for (int y = 1 + (int)threadIdx.x; y <= h - 2; y += (int)blockDim.x) {
            const int x = t - 2 * y;
            if (x < 1 || x > w - 2) continue;
            float r = 0.0f, gg = 0.0f, b = 0.0f;
            int count = 0;
            for (int j = y - 1; j <= y + 1; j++) {
                for (int i = x - 1; i <= x + 1; i++) {
                    const size_t k = (size_t)j * w + i;
                    if (valid[k]) {
                        const float4 c = grid[k];
                        r += c.x;
                        gg += c.y;
                        b += c.z;
                        count++;
                    }
                }
            }
            if (count) {
                const float s = 1.0f / (float)count;
                grid[(size_t)y * w + x] = make_float4(r * s, gg * s, b * s, 0.0f);
            }
        }
        __syncthreads();
    }
}

// static_cast<unsigned char> of a value in [0, 255] (truncation); values the reference leaves
// undefined (negative, NaN) give 0.
__device__ __forceinline__ uint8_t to_byte(float v) { return v >= 1.0f ? (uint8_t)(int)v : (uint8_t)0; }

// min2(255, v) = 255 < v ? 255 : v (Types.h:87-89).
__device__ __forceinline__ float min255(float v) { return 255.0f < v ? 255.0f : v; }

// Rgb::doubleLdr (3 bytes) or Rgb::rgbm (4 bytes) per texel, in r, g, b(, m) order.
__global__ void k_encode(const float4* __restrict__ grid, uint64_t n, int format, uint8_t* __restrict__ out) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const float4 c = grid[g];
    if (format == 1) {
        uint8_t* p = out + 3 * g;
        p[0] = to_byte(min255(c.x * 0.5f * 255.0f));
        p[1] = to_byte(min255(c.y * 0.5f * 255.0f));
        p[2] = to_byte(min255(c.z * 0.5f * 255.0f));
    } else {
        // max3(r, g, b) = max2(r, max2(g, b)), max2(a, b) = a > b ? a : b.
        const float gb = c.y > c.z ? c.y : c.z;
        const float mx = c.x > gb ? c.x : gb;
        const float m = 255.0f / mx;
        uint8_t* p = out + 4 * g;
        p[0] = to_byte(min255(c.x / m * 255.0f));
        p[1] = to_byte(min255(c.y / m * 255.0f));
        p[2] = to_byte(min255(c.z / m * 255.0f));
        p[3] = to_byte(min255(m * 255.0f));
    }
}

uint32_t blocks_of(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

hipError_t launch_expand(hipStream_t st, float4* grid, const uint8_t* valid, int w, int h) {
    const uint64_t n = (uint64_t)w * h;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_expand, dim3(blocks_of(n, 256)), dim3(256), 0, st, grid, valid, w, h);
    return hipGetLastError();
}

hipError_t launch_blur(hipStream_t st, float4* grid, const uint8_t* valid, int w, int h) {
    if (w < 3 || h < 3) return hipSuccess;
    hipLaunchKernelGGL(k_blur, dim3(1), dim3(1024), 0, st, grid, valid, w, h);
    return hipGetLastError();
}

hipError_t launch_encode(hipStream_t st, const float4* grid, uint64_t n, int format, uint8_t* out) {
    if (n == 0) return hipSuccess;
    if (format != 1 && format != 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_encode, dim3(blocks_of(n, 256)), dim3(256), 0, st, grid, n, format, out);
    return hipGetLastError();
}

}  // namespace rl
