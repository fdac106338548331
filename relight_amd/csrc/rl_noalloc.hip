// Segment tracing that allocates no device memory: the staging of rl_trace_segments (packed
// segments, raw hit records, output records), its work counter and its ray counter live in one
// file-scope device array instead of per-call hipMalloc buffers.
//
// The array is one per GPU. Its device address is looked up once per device with
// hipGetSymbolAddress (rl::noalloc_setup, called by rl_create with that device current) and kept
// in g_scratch[device]; host code passes only that address to kernels and copies, never the
// array's own name (on the host that name is a stand-in, not a device address).
//
// Every call on one GPU shares the array and leaves its contents behind; a call clears the
// counters it needs before use. Calls from one process are serialised by a per-device mutex;
// overlapping calls from different processes on one GPU would race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/relight.h"
#include "rl_device.h"
#include "rl_kernels.h"
#include "rl_noalloc.h"

namespace rl {

namespace {

// Scratch layout (byte offsets; every region 256-B aligned):
//   [0, 256)          header: u64 work queue[2] (cleared by launch_trace), u64 ray counter
//   kOffIn            starts then ends, 3 floats per segment as the caller passes them
//   kOffPacked        starts then ends as float4 (the trace kernel's operand layout)
//   kOffRaw           float4 per segment: t, triangle bits, u, v (closest hit)
//   kOffHits          rl_hit per segment (closest hit)
//   kOffFlags         u8 per segment (any hit)
constexpr size_t kMaxSeg = kNoAllocMaxSegments;
constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t kOffQueue = 0;
constexpr size_t kOffRays = 32;
constexpr size_t kOffIn = 256;
constexpr size_t kOffPacked = kOffIn + align256(2 * kMaxSeg * 3 * sizeof(float));
constexpr size_t kOffRaw = kOffPacked + align256(2 * kMaxSeg * sizeof(float4));
constexpr size_t kOffHits = kOffRaw + align256(kMaxSeg * sizeof(float4));
constexpr size_t kOffFlags = kOffHits + align256(kMaxSeg * sizeof(rl_hit));
constexpr size_t kScratchBytes = kOffFlags + align256(kMaxSeg);
static_assert(kScratchBytes == kNoAllocScratchBytes, "scratch size in rl_noalloc.h is stale");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
unsigned char* g_scratch[kMaxDevices] = {};
std::mutex g_setup_mutex;
std::mutex g_call_mutex[kMaxDevices];

// Segment start/end triples -> float4 (w = 0), as rl_trace_segments packs them on the host.
__global__ void k_pack_segments(const float* __restrict__ in, uint32_t n, float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * n) return;
    out[i] = make_float4(in[3 * i], in[3 * i + 1], in[3 * i + 2], 0.0f);
}

// Raw closest-hit records -> rl_hit, with the triangle's mesh and face from the scene-order
// attributes (a miss has triangle, mesh and face -1).
__global__ void k_unpack_hits(const float4* __restrict__ raw, uint32_t n, const TriAttr* __restrict__ attr,
                              rl_hit* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 h = raw[i];
    const uint32_t tri = __float_as_uint(h.y);
    rl_hit o;
    o.t = h.x;
    o.u = h.z;
    o.v = h.w;
    if (tri == 0xFFFFFFFFu) {
        o.triangle = -1;
        o.mesh = -1;
        o.face = -1;
    } else {
        o.triangle = (int32_t)tri;
        o.mesh = (int32_t)attr[tri].mesh;
        o.face = (int32_t)attr[tri].face;
    }
    out[i] = o;
}

}  // namespace

hipError_t noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_setup_mutex);
    if (g_scratch[device]) return hipSuccess;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<unsigned char*>(ptr);
    return hipSuccess;
}

hipError_t noalloc_trace(int device, hipStream_t st, const SceneDev& S, int cap, uint32_t max_blocks, bool any,
                         const float* starts, const float* ends, uint32_t n, uint8_t* out_flags, rl_hit* out_hits,
                         unsigned long long* out_rays) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    if (n > kMaxSeg) return hipErrorInvalidValue;
    unsigned char* base = g_scratch[device];
    if (!base) return hipErrorNotInitialized;
    std::lock_guard<std::mutex> lock(g_call_mutex[device]);
    unsigned long long* queue = reinterpret_cast<unsigned long long*>(base + kOffQueue);
    unsigned long long* rays = reinterpret_cast<unsigned long long*>(base + kOffRays);
    float* in = reinterpret_cast<float*>(base + kOffIn);
    float4* packed = reinterpret_cast<float4*>(base + kOffPacked);
    float4* raw = reinterpret_cast<float4*>(base + kOffRaw);
    rl_hit* hits = reinterpret_cast<rl_hit*>(base + kOffHits);
    uint8_t* flags = base + kOffFlags;
    *out_rays = 0;
    // The ray counter keeps the last call's value: clear it (a fill kernel, no memset node).
    hipError_t e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(rays), 2, 0u);
    if (e != hipSuccess || n == 0) return e;
    const size_t bytes3 = (size_t)n * 3 * sizeof(float);
    e = hipMemcpyAsync(in, starts, bytes3, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(in + (size_t)n * 3, ends, bytes3, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pack_segments, dim3((2 * n + 255) / 256), dim3(256), 0, st, in, n, packed);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    e = launch_trace(st, S, cap, max_blocks, any, packed, packed + n, n, flags, raw, queue, rays);
    if (e != hipSuccess) return e;
    if (any) {
        e = hipMemcpyAsync(out_flags, flags, n, hipMemcpyDeviceToHost, st);
    } else {
        hipLaunchKernelGGL(k_unpack_hits, dim3((n + 255) / 256), dim3(256), 0, st, raw, n, S.attr, hits);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        e = hipMemcpyAsync(out_hits, hits, (size_t)n * sizeof(rl_hit), hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out_rays, rays, sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

}  // namespace rl
