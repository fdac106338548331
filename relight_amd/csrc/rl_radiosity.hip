// Radiosity on the GPU: form-factor rows by hemisphere rays, and the gather B = E + rho o (F B)
// solved by Jacobi sweeps on BF16 MFMA (fp32 accumulation) inside fp32 iterative refinement.
//
// Reference: RadiosityBuilder (radiosity/RadiosityBuilder.cpp:44-176) builds pairwise
// point-to-point form factors with one visibility ray per pair; RadiositySolver::solve is empty
// (radiosity/RadiositySolver.cpp:34-36) and the only gather iteration is the legacy
// indirect_i = rho_i * sum_j F_ij * injected_j (Radiosity.cpp:130-146). Here F_ij is estimated as
// the fraction of `rays` cosine-weighted rays from patch i whose closest hit lies on patch j
// (rows sum to <= 1), stored exactly as 16-bit hit counts.
//
// F layout: u16 counts, npad x npad (npad a multiple of 64), in 16 x 32 tiles (below). Vectors (E, rho, B,
// residual, correction) are column-major fp32 [rhs][patch]; MFMA operands X are column-major
// bf16 splits of a vector (x = x1 + x2 + x3 to ~24 bits).
#include <hip/hip_runtime.h>

#include "rl_device.h"
#include "rl_kernels.h"
#include "rl_traverse.h"

namespace rl {

namespace {

using namespace trav;

constexpr int kFfBlock = 128;

// F (u16 hit counts, npad x npad) is stored in 16 x 32 tiles, tile-row-major: the 16 rows x 32
// columns one MFMA A operand covers are 1 KB contiguous, so a wavefront's load of them is one
// contiguous 1 KB.
__device__ __forceinline__ size_t f_index(uint32_t r, uint32_t k, uint32_t npad) {
    return ((size_t)(r >> 4) * (npad >> 5) + (k >> 5)) * 512u + (r & 15u) * 32u + (k & 31u);
}

// The BF16 parts of the iterate (x1, x2, x3; npad x mpad) are stored in 32-row blocks, each block
// column-major: the 32 rows x 16 columns one MFMA B operand covers are 1 KB contiguous.
__device__ __forceinline__ size_t x_index(uint32_t i, uint32_t c, uint32_t mpad) {
    return (size_t)(i >> 5) * (mpad * 32u) + c * 32u + (i & 31u);
}

// One workgroup per patch row: `rays` closest-hit rays from uniform points on the patch with
// cosine-weighted directions; hits are histogrammed per patch in LDS (two u16 counters per word,
// counts <= 65535) and the row is written out as counts.
template <int CAP>
__global__ __launch_bounds__(kFfBlock) void k_form_factor_rows(SceneDev S, const PatchDev* __restrict__ patches,
                                                               const uint32_t* __restrict__ tri_patch,
                                                               FormFactorParams P, uint16_t* __restrict__ F,
                                                               unsigned long long* rays_out) {
    extern __shared__ uint32_t smem[];
    uint32_t* stack = smem;
    uint32_t* hist = smem + kLdsStack * kFfBlock;
    const uint32_t row = P.first_row + blockIdx.x;
    const uint32_t words = P.row_stride / 2;
    for (uint32_t w = threadIdx.x; w < words; w += kFfBlock) hist[w] = 0u;
    __syncthreads();
    const PatchDev pd = patches[row];
    const V3 o = v3(pd.o[0], pd.o[1], pd.o[2]), u = v3(pd.u[0], pd.u[1], pd.u[2]);
    const V3 v = v3(pd.v[0], pd.v[1], pd.v[2]), n = v3(pd.n[0], pd.n[1], pd.n[2]);
    uint32_t rays = 0;
    for (uint32_t k = threadIdx.x; k < P.rays; k += kFfBlock) {
        Pcg32 rng = sample_stream(P.seed, 0xFFFFFFFEu, row, TECH_FORM_FACTOR, k);
        const float s = pcg32_float(rng), t = pcg32_float(rng);
        const V3 p = o + u * s + v * t;
        const V3 dir = sample_cosine_hemisphere(rng, n);
        const Ray ray = ray_from_segment(p, p + dir * P.max_distance);
        TraceResult tr;
        rays++;
        if (traverse<false, CAP>(S, ray, stack + threadIdx.x, kFfBlock, tr)) {
            const uint32_t j = tri_patch[tr.tri];
            if (j < P.num_patches) atomicAdd(&hist[j >> 1], 1u << ((j & 1u) * 16u));
        }
    }
    __syncthreads();
    uint32_t* out = reinterpret_cast<uint32_t*>(F);
    for (uint32_t w = threadIdx.x; w < words; w += kFfBlock) out[f_index(row, 2u * w, P.row_stride) >> 1] = hist[w];
    count_rays(rays, rays_out);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Y_part[split][c][row] = sum over this split's k of F[row][k] * X[k][c] with
// v_mfma_f32_16x16x32_bf16: each wavefront owns 16 rows and all NT*16 columns. A operand lane l
// holds F[row0 + l%16][k + 8*(l/16) .. +7]; B operand lane l holds X[k + 8*(l/16) .. +7][16t + l%16];
// the accumulator lane l holds rows row0 + 4*(l/16) + i, column 16t + l%16.
// ACC = false: F_hi * x1 (one BF16 product). ACC = true: (F_hi + F_lo) * (x1 + x2 + x3) without the
// F_lo * x3 term, where F_hi = bf16(count) and F_lo = count - F_hi are both exact.
template <int NT, bool ACC>
__global__ __launch_bounds__(256) void k_ff_gemm(const uint16_t* __restrict__ F, uint32_t stride,
                                                 const __bf16* __restrict__ X1, const __bf16* __restrict__ X2,
                                                 const __bf16* __restrict__ X3, uint32_t npad,
                                                 uint32_t klen, float* __restrict__ Yp) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t row0 = (blockIdx.x * 4u + wave) * 16u;
    const uint32_t k0 = blockIdx.y * klen, kend = k0 + klen;
    const uint32_t sub = lane >> 4, col = lane & 15u;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    constexpr uint32_t mp = NT * 16;
    // Lane offsets inside a 16 x 32 F tile and inside a 32 x 16 block of x.
    const uint16_t* ftile = F + (size_t)(row0 >> 4) * (stride >> 5) * 512u + col * 32u + 8u * sub;
    const uint32_t xl = col * 32u + 8u * sub;
    // F is streamed once per launch (non-temporal, so x stays in L2); the next step's F and x
    // fragments are loaded before this step's conversions and MFMAs.
    u32x4 raw_next = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(ftile + (size_t)(k0 >> 5) * 512u));
    bf16x8 b1n[NT], b2n[NT], b3n[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const size_t xo = (size_t)(k0 >> 5) * (mp * 32u) + 512u * t + xl;
        b1n[t] = *reinterpret_cast<const bf16x8*>(X1 + xo);
        if (ACC) {
            b2n[t] = *reinterpret_cast<const bf16x8*>(X2 + xo);
            b3n[t] = *reinterpret_cast<const bf16x8*>(X3 + xo);
        }
    }
    for (uint32_t k = k0; k < kend; k += 32) {
        const u32x4 raw = raw_next;
        bf16x8 b1[NT], b2[NT], b3[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            b1[t] = b1n[t];
            if (ACC) {
                b2[t] = b2n[t];
                b3[t] = b3n[t];
            }
        }
        if (k + 32 < kend) {
            raw_next = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(ftile + (size_t)((k + 32) >> 5) * 512u));
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const size_t xo = (size_t)((k + 32) >> 5) * (mp * 32u) + 512u * t + xl;
                b1n[t] = *reinterpret_cast<const bf16x8*>(X1 + xo);
                if (ACC) {
                    b2n[t] = *reinterpret_cast<const bf16x8*>(X2 + xo);
                    b3n[t] = *reinterpret_cast<const bf16x8*>(X3 + xo);
                }
            }
        }
        const uint32_t w[4] = {raw.x, raw.y, raw.z, raw.w};
        bf16x8 ahi, alo;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const float fc = (float)((w[j >> 1] >> ((j & 1) * 16)) & 0xFFFFu);
            const __bf16 h = (__bf16)fc;
            ahi[j] = h;
            if (ACC) alo[j] = (__bf16)(fc - (float)h);
        }
#pragma unroll
        for (int t = 0; t < NT; t++) {
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, b1[t], acc[t], 0, 0, 0);
            if (ACC) {
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, b2[t], acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, b3[t], acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, b1[t], acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, b2[t], acc[t], 0, 0, 0);
            }
        }
    }
    const uint32_t mpad = NT * 16;
    float* out = Yp + (size_t)blockIdx.y * mpad * npad;
#pragma unroll
    for (int t = 0; t < NT; t++) {
#pragma unroll
        for (int i = 0; i < 4; i++) out[(size_t)(16u * t + col) * npad + row0 + 4u * sub + i] = acc[t][i];
    }
}

// Staged gather products. A workgroup (four wavefronts, 128 rows) stages x through LDS AHEAD K
// steps at a time, double-buffered, so x is read from L2 once per workgroup instead of once per
// wavefront; each wavefront owns two 16-row tiles (every B fragment feeds two row tiles' MFMAs)
// and keeps F loads AHEAD K steps ahead in a register ring. ACC = false: F_hi * x1, the Jacobi
// sweeps' product. ACC = true: the refinement residual's (F_hi + F_lo)(x1 + x2 + x3) without
// F_lo * x3, with x1, x2 and x3 all staged. Same lane layout, products and summation order per
// row tile as k_ff_gemm<NT, ACC>.
constexpr int kFastRows = 32;  // rows per wavefront
// K steps of F in flight = K steps per LDS stage of x. For the sweeps eight (96 KB of x stages for
// 96 columns, one workgroup per compute unit) measured 121 us per 16,384-patch product against
// 129 us with four; the residual product stages three x vectors, two steps at a time (72 KB).
constexpr int kFastAhead = 8;
constexpr int kAccAhead = 2;

template <int NT, bool ACC>
__global__ __launch_bounds__(256) void k_ff_gemm_staged(const uint16_t* __restrict__ F, uint32_t stride,
                                                        const __bf16* __restrict__ X1,
                                                        const __bf16* __restrict__ X2,
                                                        const __bf16* __restrict__ X3, uint32_t npad,
                                                        uint32_t klen, float* __restrict__ Yp) {
    constexpr int AHEAD = ACC ? kAccAhead : kFastAhead;
    constexpr int NX = ACC ? 3 : 1;                           // x vectors staged
    constexpr uint32_t mp = NT * 16;
    constexpr uint32_t kStepVecs = mp * 32u / 8u;             // bf16x8 vectors of one x per K step
    constexpr uint32_t kStageVecs = kStepVecs * AHEAD;        // per x per LDS stage
    constexpr uint32_t kPerThread = (kStageVecs + 255u) / 256u;
    __shared__ bf16x8 xs[2][NX][kStageVecs];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t row0 = (blockIdx.x * 4u + wave) * (uint32_t)kFastRows;
    const uint32_t k0 = blockIdx.y * klen;
    const uint32_t nstage = klen / (32u * AHEAD);
    const uint32_t sub = lane >> 4, col = lane & 15u;
    f32x4 acc[2][NT];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[r][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const uint16_t* ftile[2] = {F + (size_t)(row0 >> 4) * (stride >> 5) * 512u + col * 32u + 8u * sub,
                                F + (size_t)((row0 >> 4) + 1u) * (stride >> 5) * 512u + col * 32u + 8u * sub};
    const bf16x8* xg[3] = {reinterpret_cast<const bf16x8*>(X1) + (size_t)(k0 >> 5) * kStepVecs,
                           reinterpret_cast<const bf16x8*>(ACC ? X2 : X1) + (size_t)(k0 >> 5) * kStepVecs,
                           reinterpret_cast<const bf16x8*>(ACC ? X3 : X1) + (size_t)(k0 >> 5) * kStepVecs};
    const uint32_t xl = col * 4u + sub;  // vector offset of this lane's fragment in a step
    u32x4 ring[2][AHEAD];
#pragma unroll
    for (int j = 0; j < AHEAD; j++)
#pragma unroll
        for (int r = 0; r < 2; r++)
            ring[r][j] = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(ftile[r] + (size_t)((k0 >> 5) + j) * 512u));
    bf16x8 stg[NX][kPerThread];
#pragma unroll
    for (int q = 0; q < NX; q++)
#pragma unroll
        for (uint32_t p = 0; p < kPerThread; p++) {
            const uint32_t v = threadIdx.x + 256u * p;
            if (v < kStageVecs) stg[q][p] = xg[q][v];
        }
#pragma unroll
    for (int q = 0; q < NX; q++)
#pragma unroll
        for (uint32_t p = 0; p < kPerThread; p++) {
            const uint32_t v = threadIdx.x + 256u * p;
            if (v < kStageVecs) xs[0][q][v] = stg[q][p];
        }
    __syncthreads();
    for (uint32_t s = 0; s < nstage; s++) {
        const uint32_t buf = s & 1u;
        const bool more = s + 1u < nstage;
        if (more) {
#pragma unroll
            for (int q = 0; q < NX; q++)
#pragma unroll
                for (uint32_t p = 0; p < kPerThread; p++) {
                    const uint32_t v = threadIdx.x + 256u * p;
                    if (v < kStageVecs) stg[q][p] = xg[q][(size_t)(s + 1u) * kStageVecs + v];
                }
        }
#pragma unroll
        for (int j = 0; j < AHEAD; j++) {
            const uint32_t step = (k0 >> 5) + s * AHEAD + j;  // global K step (32 columns)
            u32x4 cur[2] = {ring[0][j], ring[1][j]};
            if (more) {
#pragma unroll
                for (int r = 0; r < 2; r++)
                    ring[r][j] = __builtin_nontemporal_load(
                        reinterpret_cast<const u32x4*>(ftile[r] + (size_t)(step + AHEAD) * 512u));
            }
            bf16x8 b[NX][NT];
#pragma unroll
            for (int q = 0; q < NX; q++)
#pragma unroll
                for (int t = 0; t < NT; t++) b[q][t] = xs[buf][q][(uint32_t)j * kStepVecs + 64u * t + xl];
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const uint32_t w[4] = {cur[r].x, cur[r].y, cur[r].z, cur[r].w};
                bf16x8 ahi, alo;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const float fc = (float)((w[e >> 1] >> ((e & 1) * 16)) & 0xFFFFu);
                    const __bf16 h = (__bf16)fc;
                    ahi[e] = h;
                    if (ACC) alo[e] = (__bf16)(fc - (float)h);
                }
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, b[0][t], acc[r][t], 0, 0, 0);
                    if (ACC) {
                        acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, b[NX > 1 ? 1 : 0][t], acc[r][t], 0, 0, 0);
                        acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, b[NX > 2 ? 2 : 0][t], acc[r][t], 0, 0, 0);
                        acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, b[0][t], acc[r][t], 0, 0, 0);
                        acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, b[NX > 1 ? 1 : 0][t], acc[r][t], 0, 0, 0);
                    }
                }
            }
        }
        if (more) {
#pragma unroll
            for (int q = 0; q < NX; q++)
#pragma unroll
                for (uint32_t p = 0; p < kPerThread; p++) {
                    const uint32_t v = threadIdx.x + 256u * p;
                    if (v < kStageVecs) xs[buf ^ 1u][q][v] = stg[q][p];
                }
        }
        __syncthreads();
    }
    float* out = Yp + (size_t)blockIdx.y * mp * npad;
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) out[(size_t)(16u * t + col) * npad + row0 + 16u * r + 4u * sub + i] = acc[r][t][i];
}

// Row-major (n x m) -> column-major padded (npad x mpad, zeros outside), and back.
__global__ void k_ff_to_cols(const float* __restrict__ in, uint32_t n, uint32_t m, uint32_t npad, uint32_t mpad,
                             float* __restrict__ out) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (uint64_t)npad * mpad) return;
    const uint32_t c = (uint32_t)(g / npad), i = (uint32_t)(g % npad);
    out[g] = (i < n && c < m) ? in[(size_t)i * m + c] : 0.0f;
}

__global__ void k_ff_to_rows(const float* __restrict__ cols, uint32_t n, uint32_t m, uint32_t npad,
                             float* __restrict__ out) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (uint64_t)n * m) return;
    const uint32_t i = (uint32_t)(g / m), c = (uint32_t)(g % m);
    out[g] = cols[(size_t)c * npad + i];
}

__device__ __forceinline__ void split3(float x, __bf16& a, __bf16& b, __bf16& c) {
    a = (__bf16)x;
    const float r1 = x - (float)a;
    b = (__bf16)r1;
    c = (__bf16)(r1 - (float)b);
}

__device__ __forceinline__ float sum_parts(const float* Yp, uint32_t nsplit, size_t plane, size_t idx) {
    float y = 0.0f;
    for (uint32_t s = 0; s < nsplit; s++) y += Yp[s * plane + idx];
    return y;
}

// Block-level sum of squares in a fixed order (deterministic), one partial per (column, block).
__device__ __forceinline__ void block_sq_sum(float v, float* part) {
    __shared__ float red[256];
    red[threadIdx.x] = v * v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *part = red[0];
}

// r = E + rho o (F B) / rays - B; d = r; x1 = bf16(r). grid (npad/256, mpad).
__global__ __launch_bounds__(256) void k_ff_residual(const float* __restrict__ Yp, uint32_t nsplit, uint32_t npad,
                                                     uint32_t mpad, float inv_rays, const float* __restrict__ E,
                                                     const float* __restrict__ rho, const float* __restrict__ B,
                                                     float* __restrict__ R, float* __restrict__ D,
                                                     __bf16* __restrict__ X1, float* __restrict__ part) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x, c = blockIdx.y;
    const size_t idx = (size_t)c * npad + i;
    const float y = sum_parts(Yp, nsplit, (size_t)mpad * npad, idx);
    const float r = E[idx] + rho[idx] * (y * inv_rays) - B[idx];
    R[idx] = r;
    D[idx] = r;
    X1[x_index(i, c, mpad)] = (__bf16)r;
    block_sq_sum(r, part + (size_t)c * gridDim.x + blockIdx.x);
}

// One Jacobi sweep on the correction: d = r + rho o (F_hi d) / rays; x1 = bf16(d).
__global__ __launch_bounds__(256) void k_ff_sweep(const float* __restrict__ Yp, uint32_t nsplit, uint32_t npad,
                                                  uint32_t mpad, float inv_rays, const float* __restrict__ rho,
                                                  const float* __restrict__ R, float* __restrict__ D,
                                                  __bf16* __restrict__ X1) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x, c = blockIdx.y;
    const size_t idx = (size_t)c * npad + i;
    const float y = sum_parts(Yp, nsplit, (size_t)mpad * npad, idx);
    const float d = R[idx] + rho[idx] * (y * inv_rays);
    D[idx] = d;
    X1[x_index(i, c, mpad)] = (__bf16)d;
}

// B += d (or B = d when init), then the three-way bf16 split of B.
__global__ __launch_bounds__(256) void k_ff_apply(uint32_t npad, float* __restrict__ B, const float* __restrict__ D,
                                                  int init, __bf16* __restrict__ X1, __bf16* __restrict__ X2,
                                                  __bf16* __restrict__ X3) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x, c = blockIdx.y;
    const size_t idx = (size_t)c * npad + i;
    const float b = init ? D[idx] : B[idx] + D[idx];
    B[idx] = b;
    __bf16 a1, a2, a3;
    split3(b, a1, a2, a3);
    const size_t xi = x_index(i, c, gridDim.y);
    X1[xi] = a1;
    X2[xi] = a2;
    X3[xi] = a3;
}

// Rows [first, first + rows) of the tiled count matrix as row-major rows of n counts.
__global__ void k_ff_untile(const uint16_t* __restrict__ F, uint32_t npad, uint32_t n, uint32_t first, uint32_t rows,
                            uint16_t* __restrict__ out) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (uint64_t)rows * n) return;
    const uint32_t r = (uint32_t)(g / n), k = (uint32_t)(g % n);
    out[g] = F[f_index(first + r, k, npad)];
}

uint32_t blocks_for(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

hipError_t launch_form_factor_rows(hipStream_t st, const SceneDev& S, int cap, const PatchDev* patches,
                                   const uint32_t* tri_patch, const FormFactorParams& P, uint16_t* F,
                                   unsigned long long* rays) {
    if (P.rows == 0) return hipSuccess;
    const size_t lds = (size_t)kLdsStack * kFfBlock * 4 + (size_t)P.row_stride * 2;
    if (lds > 65536) return hipErrorInvalidValue;
    if (cap == 32) {
        hipLaunchKernelGGL(k_form_factor_rows<32>, dim3(P.rows), dim3(kFfBlock), lds, st, S, patches, tri_patch, P, F,
                           rays);
    } else {
        hipLaunchKernelGGL(k_form_factor_rows<64>, dim3(P.rows), dim3(kFfBlock), lds, st, S, patches, tri_patch, P, F,
                           rays);
    }
    return hipGetLastError();
}

hipError_t launch_ff_gemm(hipStream_t st, bool accurate, const uint16_t* F, uint32_t stride, const __bf16* X1,
                          const __bf16* X2, const __bf16* X3, uint32_t npad, uint32_t mpad, uint32_t nsplit,
                          float* Yp) {
    const uint32_t klen = npad / nsplit;
    if (klen % 32u != 0u || npad % 64u != 0u) return hipErrorInvalidValue;
    const dim3 grid(npad / 64, nsplit), block(256);
    // The staged products take 128 rows per block and whole LDS stages of K when the sizes allow it.
    const uint32_t ahead = accurate ? (uint32_t)kAccAhead : (uint32_t)kFastAhead;
    const bool wide = npad % (4u * kFastRows) == 0u && klen % (32u * ahead) == 0u;
    const dim3 grid_fast(npad / (4u * kFastRows), nsplit);
#define RL_GEMM_CASE(NTV)                                                                                     \
    case NTV:                                                                                                 \
        if (wide && accurate)                                                                                 \
            hipLaunchKernelGGL((k_ff_gemm_staged<NTV, true>), grid_fast, block, 0, st, F, stride, X1, X2, X3, npad, \
                               klen, Yp);                                                                     \
        else if (wide)                                                                                        \
            hipLaunchKernelGGL((k_ff_gemm_staged<NTV, false>), grid_fast, block, 0, st, F, stride, X1, X1, X1, npad, \
                               klen, Yp);                                                                     \
        else if (accurate)                                                                                    \
            hipLaunchKernelGGL((k_ff_gemm<NTV, true>), grid, block, 0, st, F, stride, X1, X2, X3, npad, klen, Yp); \
        else                                                                                                  \
            hipLaunchKernelGGL((k_ff_gemm<NTV, false>), grid, block, 0, st, F, stride, X1, X2, X3, npad, klen, Yp); \
        break;
    switch (mpad / 16) {
        RL_GEMM_CASE(1)
        RL_GEMM_CASE(2)
        RL_GEMM_CASE(3)
        RL_GEMM_CASE(4)
        RL_GEMM_CASE(6)
        RL_GEMM_CASE(8)
        default:
            return hipErrorInvalidValue;
    }
#undef RL_GEMM_CASE
    return hipGetLastError();
}

hipError_t launch_ff_untile(hipStream_t st, const uint16_t* F, uint32_t npad, uint32_t n, uint32_t first,
                           uint32_t rows, uint16_t* out) {
    if ((uint64_t)rows * n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ff_untile, dim3(blocks_for((uint64_t)rows * n, 256)), dim3(256), 0, st, F, npad, n, first,
                       rows, out);
    return hipGetLastError();
}

hipError_t launch_ff_to_cols(hipStream_t st, const float* in, uint32_t n, uint32_t m, uint32_t npad, uint32_t mpad,
                            float* out) {
    hipLaunchKernelGGL(k_ff_to_cols, dim3(blocks_for((uint64_t)npad * mpad, 256)), dim3(256), 0, st, in, n, m, npad,
                       mpad, out);
    return hipGetLastError();
}

hipError_t launch_ff_to_rows(hipStream_t st, const float* cols, uint32_t n, uint32_t m, uint32_t npad, float* out) {
    if ((uint64_t)n * m == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ff_to_rows, dim3(blocks_for((uint64_t)n * m, 256)), dim3(256), 0, st, cols, n, m, npad, out);
    return hipGetLastError();
}

hipError_t launch_ff_residual(hipStream_t st, const float* Yp, uint32_t nsplit, uint32_t npad, uint32_t mpad,
                              float inv_rays, const float* E, const float* rho, const float* B, float* R, float* D,
                              __bf16* X1, float* part) {
    hipLaunchKernelGGL(k_ff_residual, dim3(npad / 256, mpad), dim3(256), 0, st, Yp, nsplit, npad, mpad, inv_rays, E,
                       rho, B, R, D, X1, part);
    return hipGetLastError();
}

hipError_t launch_ff_sweep(hipStream_t st, const float* Yp, uint32_t nsplit, uint32_t npad, uint32_t mpad,
                           float inv_rays, const float* rho, const float* R, float* D, __bf16* X1) {
    hipLaunchKernelGGL(k_ff_sweep, dim3(npad / 256, mpad), dim3(256), 0, st, Yp, nsplit, npad, mpad, inv_rays, rho, R,
                       D, X1);
    return hipGetLastError();
}

hipError_t launch_ff_apply(hipStream_t st, uint32_t npad, uint32_t mpad, float* B, const float* D, int init,
                           __bf16* X1, __bf16* X2, __bf16* X3) {
    hipLaunchKernelGGL(k_ff_apply, dim3(npad / 256, mpad), dim3(256), 0, st, npad, B, D, init, X1, X2, X3);
    return hipGetLastError();
}

}  // namespace rl
