// Photon emission and tracing, the deterministic photon store, the photonmap gather and the
// final gather: Photons::emit/trace/store (baker/Photons.cpp:50-144), Photonmap::gather
// (Lightmap.cpp:376-413) and IndirectLight::bakeLumel (baker/IndirectLight.cpp:47-75).
//
// Photon stores land in fixed record slots (photon-major, bounce-minor), are grouped by texel with
// a stable radix sort and summed in record order, so the photonmap equals a sequential store in
// the reference's emission order, bit for bit, on every run.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "rl_device.h"
#include "rl_kernels.h"
#include "rl_traverse.h"

namespace rl {

namespace {

using namespace trav;

constexpr int kBlock = 128;

__device__ __forceinline__ int find_map(const PhotonMapDev* pms, int npm, uint32_t key) {
    int lo = 0, hi = npm - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (pms[mid].key_base <= key) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Hit attributes through Face::normalAt / uvAt with Embree's (u, v) (rt/Embree.cpp:160-171).
struct HitAttr {
    V3 n;
    float uvx, uvy;
    uint32_t mesh;
    float albedo[4];
};

__device__ __forceinline__ HitAttr hit_attributes(const SceneDev& S, const TraceResult& tr) {
    const TriAttr& A = S.attr[tr.tri];
    const float u = tr.V / tr.det, v = tr.W / tr.det;
    HitAttr h;
    h.n = v3(face_lerp(A.n[0], A.n[3], A.n[6], u, v), face_lerp(A.n[1], A.n[4], A.n[7], u, v),
             face_lerp(A.n[2], A.n[5], A.n[8], u, v));
    h.uvx = face_lerp(A.uv[0], A.uv[2], A.uv[4], u, v);
    h.uvy = face_lerp(A.uv[1], A.uv[3], A.uv[5], u, v);
    h.mesh = A.mesh;
    h.albedo[0] = A.albedo[0]; h.albedo[1] = A.albedo[1]; h.albedo[2] = A.albedo[2]; h.albedo[3] = A.albedo[3];
    return h;
}

// One thread per photon: emission (PhotonEmitter::emit / DirectionalPhotonEmitter::emit,
// scene/Light.cpp:339-361), spot cutoff, then the bounce loop of Photons::trace. All draws of a
// photon come from one PCG32 stream keyed by its global index.
template <int CAP>
__global__ __launch_bounds__(kBlock) void k_photon_trace(SceneDev S, const LightDev* __restrict__ lights,
                                                         const uint64_t* __restrict__ light_first,
                                                         PhotonParams P, const int32_t* __restrict__ mesh_pm,
                                                         const PhotonMapDev* __restrict__ pms,
                                                         PhotonRec* __restrict__ out,
                                                         float4* __restrict__ out_pos,
                                                         unsigned long long* queue,
                                                         unsigned long long* rays_out) {
    __shared__ uint32_t stack[kLdsStack * kBlock];
    uint32_t* stk = stack + threadIdx.x;
    const uint64_t nchunks = (P.total + 63) / 64;
    uint32_t rays = 0;
    for (;;) {
        const uint64_t chunk = wave_next_chunk(queue);
        if (chunk >= nchunks) break;
        const uint64_t p = chunk * 64 + lane_id();
        if (p >= P.total) continue;
        const uint64_t r = p % P.per_pass;
        int l = 0;
        while (l + 1 < P.nlights && light_first[l + 1] <= r) l++;
        const LightDev L = lights[l];
        Pcg32 rng = sample_stream(P.seed, 0xFFFFFFFFu, (uint32_t)p, TECH_PHOTON, (uint32_t)(p >> 32));
        V3 pos, dir;
        if (L.type == LIGHT_DIRECTIONAL) {
            // Plane::calculate(dir, light position) applied to Bounds::randomPointInside, minus
            // dir * 5 (Types.h:1445-1450,1655-1658,1715-1718; lerp = a*(1-t) + b*t, Types.h:75-78).
            const V3 n = v3(L.dir[0], L.dir[1], L.dir[2]);
            const float tx = pcg32_float(rng), ty = pcg32_float(rng), tz = pcg32_float(rng);
            const V3 q = v3(P.scene_lo[0] * (1.0f - tx) + P.scene_hi[0] * tx,
                            P.scene_lo[1] * (1.0f - ty) + P.scene_hi[1] * ty,
                            P.scene_lo[2] * (1.0f - tz) + P.scene_hi[2] * tz);
            const float d = -dot(n, v3(L.pos[0], L.pos[1], L.pos[2]));
            const float dist = dot(n, q) + d;
            pos = (q - n * dist) - n * 5.0f;
            dir = n;
        } else {
            pos = v3(L.pos[0], L.pos[1], L.pos[2]);
            dir = sample_sphere(rng);
        }
        float cut = 1.0f;
        if (L.has_spot) {
            // LightSpotCutoff::cutoffForDirection (scene/Light.cpp:462-476).
            float value = dot(dir, v3(L.dir[0], L.dir[1], L.dir[2]));
            if (value <= L.cutoff) {
                cut = 0.0f;
            } else {
                value = (1.0f - (1.0f - value) * 1.0f / (1.0f - L.cutoff));
                if (fabsf(1.0f - L.spot_exponent) > 0.01f) value = powf(value, L.spot_exponent);
                cut = value;
            }
        }
        if (cut > 0.0f) {
            float cr = L.color[0] * L.intensity * cut;
            float cg = L.color[1] * L.intensity * cut;
            float cb = L.color[2] * L.intensity * cut;
            PhotonRec* slot = out + p * (uint64_t)(P.max_depth + 1);
            for (int depth = 0; depth <= P.max_depth; depth++) {
                if (P.termination == PHOTON_THRESHOLD && luminance(cr, cg, cb) < P.energy_threshold) break;
                const Ray ray = ray_from_segment(pos, pos + dir * P.max_distance);
                TraceResult tr;
                rays++;
                if (!traverse<false, CAP>(S, ray, stk, kBlock, tr)) break;
                const V3 hp = ray.o + ray.d * tr.t;
                const HitAttr h = hit_attributes(S, tr);
                const float influence = lambert(-dir, h.n) * 1.0f;
                const float hr = cr * h.albedo[0] * influence;
                const float hg = cg * h.albedo[1] * influence;
                const float hb = cb * h.albedo[2] * influence;
                const int32_t pm = mesh_pm[h.mesh];
                if (pm >= 0) {
                    const PhotonMapDev M = pms[pm];
                    const int x = lumel_coord(h.uvx, M.w), y = lumel_coord(h.uvy, M.h);
                    const uint32_t t = (uint32_t)y * (uint32_t)M.w + (uint32_t)x;
                    if (M.valid[t]) {
                        PhotonRec rec;
                        rec.key = M.key_base + t;
                        rec.r = hr; rec.g = hg; rec.b = hb;
                        slot[depth] = rec;
                        if (out_pos) out_pos[p * (uint64_t)(P.max_depth + 1) + depth] = make_float4(hp.x, hp.y, hp.z, 0.0f);
                    }
                }
                if (P.termination == PHOTON_RUSSIAN_ROULETTE) {
                    const float q = (h.albedo[0] + h.albedo[1] + h.albedo[2]) / 3.0f;
                    if (!(pcg32_float(rng) < q)) break;
                    cr = hr / q; cg = hg / q; cb = hb / q;
                } else {
                    cr = hr; cg = hg; cb = hb;
                }
                pos = hp;
                dir = sample_uniform_hemisphere(rng, h.n);
            }
        }
    }
    count_rays(rays, rays_out);
}

__global__ void k_photon_clear(PhotonRec* out, uint64_t n) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) {
        PhotonRec r;
        r.key = 0xFFFFFFFFu;
        r.r = r.g = r.b = 0.0f;
        out[g] = r;
    }
}

__global__ void k_keys_iota(const PhotonRec* __restrict__ rec, uint32_t* __restrict__ keys,
                            uint32_t* __restrict__ idx, uint32_t n) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) {
        keys[g] = rec[g].key;
        idx[g] = g;
    }
}

// Photons::store (baker/Photons.cpp:129-144) over records grouped by texel: one thread per run
// head, summing in record order on top of the texel's current value.
__global__ void k_photon_accumulate(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ idx,
                                    const PhotonRec* __restrict__ rec, uint32_t n,
                                    const PhotonMapDev* __restrict__ pms, int npm) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const uint32_t key = keys[g];
    if (key == 0xFFFFFFFFu) return;
    if (g > 0 && keys[g - 1] == key) return;
    const PhotonMapDev M = pms[find_map(pms, npm, key)];
    const uint32_t t = key - M.key_base;
    float4 c = M.color[t];
    uint32_t cnt = M.photons[t];
    for (uint32_t j = g; j < n && keys[j] == key; j++) {
        const PhotonRec r = rec[idx[j]];
        c.x += r.r; c.y += r.g; c.z += r.b;
        cnt++;
    }
    M.color[t] = c;
    M.photons[t] = cnt;
}

// Photonmap::gather(x, y, radius) (Lightmap.cpp:386-413) for every texel of every map.
__global__ void k_photon_gather(const PhotonMapDev* __restrict__ pms, int npm, uint32_t total, int radius) {
    const uint32_t key = blockIdx.x * blockDim.x + threadIdx.x;
    if (key >= total) return;
    const PhotonMapDev M = pms[find_map(pms, npm, key)];
    const uint32_t t = key - M.key_base;
    const int x = (int)(t % (uint32_t)M.w), y = (int)(t / (uint32_t)M.w);
    float cr = 0.0f, cg = 0.0f, cb = 0.0f;
    int photons = 0;
    for (int j = y - radius; j <= y + radius; j++) {
        for (int i = x - radius; i <= x + radius; i++) {
            if (i < 0 || j < 0 || i >= M.w || j >= M.h) continue;
            const float distance = sqrtf((float)((x - i) * (x - i) + (y - j) * (y - j)));
            if (distance > (float)radius) continue;
            const uint32_t k = (uint32_t)j * (uint32_t)M.w + (uint32_t)i;
            const float4 c = M.color[k];
            cr += c.x; cg += c.y; cb += c.z;
            photons += (int)M.photons[k];
        }
    }
    float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (photons != 0) {
        const float fp = (float)photons;
        out = make_float4(cr / fp, cg / fp, cb / fp, 0.0f);
    }
    M.gathered[t] = out;
}

// IndirectLight::bakeLumel: one wavefront per lumel (taken from the persistent queue), 64
// samples per step; the per-sample terms are added in sample order, as the reference's loop does,
// by broadcasting them lane by lane.
// Workgroups of four wavefronts, each keeping the first 192 nodes of the first segment's scene in
// LDS (21 KB; with the stacks a workgroup takes 37 KB, so four fit a compute unit's 160 KB).
constexpr int kFgBlock = 256;
constexpr uint32_t kFgTopNodes = 192;

template <int CAP>
__global__ __launch_bounds__(kFgBlock) void k_final_gather(const BakeSeg* __restrict__ segs, int nsegs,
                                                           uint64_t total, unsigned long long* queue,
                                                           unsigned long long* rays_out) {
    __shared__ uint32_t stack[kLdsStack * kFgBlock];
    __shared__ u32x4 top_rows[kFgTopNodes * kNodeRows];
    const SceneDev top_scene = segs[0].scene;
    const uint32_t top_count = load_top_nodes(top_scene, top_rows, kFgTopNodes);
    __syncthreads();
    uint32_t* stk = stack + threadIdx.x;
    const uint32_t lane = lane_id();
    uint32_t rays = 0;
    for (;;) {
    const uint64_t g = wave_next_chunk(queue);  // lumel of this wavefront
    if (g >= total) break;
    const BakeSeg& sg = segs[find_seg(segs, nsegs, g)];
    const uint32_t i = (uint32_t)(g - sg.work_begin);
    const SceneDev S = sg.scene;
    const TopCache top{top_rows, S.nodes == top_scene.nodes ? top_count : 0u};
    const FinalGatherParams P = sg.fg;
    const int32_t* mesh_pm = sg.mesh_pm;
    const PhotonMapDev* pms = sg.pms;
    const float4 p4 = sg.pos[i], n4 = sg.nrm[i];
    const V3 p = v3(p4.x, p4.y, p4.z), n = v3(n4.x, n4.y, n4.z);
    const uint32_t texel = f_bits(p4.w);
    float ar = 0.0f, ag = 0.0f, ab = 0.0f;
    for (int base = 0; base < P.samples; base += 64) {
        const int s = base + (int)lane;
        float tr_ = 0.0f, tg_ = 0.0f, tb_ = 0.0f;
        int use = 0;
        if (s < P.samples) {
            Pcg32 rng = sample_stream(P.seed, sg.lm_seed, texel, TECH_FINAL_GATHER, (uint32_t)s);
            const V3 dir = sample_hemisphere(rng, n, P.sampling);
            const float d0 = dot(n, dir);
            const float influence = d0 > 0.0f ? d0 : 0.0f;
            const Ray ray = ray_from_segment(p, p + dir * P.distance);
            TraceResult tr;
            rays++;
            if (!traverse<false, CAP>(S, top, ray, stk, kFgBlock, tr)) {
                tr_ = P.sky[0] * influence + P.ambient[0];
                tg_ = P.sky[1] * influence + P.ambient[1];
                tb_ = P.sky[2] * influence + P.ambient[2];
                use = 1;
            } else {
                const HitAttr h = hit_attributes(S, tr);
                if (dot(dir, h.n) < 0.0f) {
                    const int32_t pm = mesh_pm[h.mesh];
                    if (pm >= 0) {
                        const PhotonMapDev M = pms[pm];
                        const int x = lumel_coord(h.uvx, M.w), y = lumel_coord(h.uvy, M.h);
                        const float4 gc = M.gathered[(uint32_t)y * (uint32_t)M.w + (uint32_t)x];
                        tr_ = gc.x * influence + P.ambient[0];
                        tg_ = gc.y * influence + P.ambient[1];
                        tb_ = gc.z * influence + P.ambient[2];
                        use = 1;
                    }
                }
            }
        }
        const int nb = min(64, P.samples - base);
        for (int k = 0; k < nb; k++) {
            const float xr = __shfl(tr_, k, 64), xg = __shfl(tg_, k, 64), xb = __shfl(tb_, k, 64);
            if (__shfl(use, k, 64)) { ar += xr; ag += xg; ab += xb; }
        }
    }
    if (lane == 0) {
        float4 c = sg.color[i];
        const float fs = (float)P.samples;
        c.x += ar / fs; c.y += ag / fs; c.z += ab / fs;
        sg.color[i] = c;
    }
    }
    count_rays(rays, rays_out);
}

// ---------------------------------------------------------------- k-nearest photon gather

__device__ __forceinline__ int knn_axis(float x, float lo, float inv_h) { return (int)floorf((x - lo) * inv_h); }

__device__ __forceinline__ uint32_t knn_bucket(const KnnGrid& G, int cx, int cy, int cz) {
    return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u)) & G.mask;
}

// Bucket key of every record slot (mask + 1 = empty slot, sorted last) and the slot index.
__global__ void k_knn_keys(const PhotonRec* __restrict__ rec, const float4* __restrict__ ppos, uint32_t n,
                           KnnGrid G, uint32_t* __restrict__ keys, uint32_t* __restrict__ idx) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    uint32_t key = G.mask + 1u;
    if (rec[g].key != 0xFFFFFFFFu) {
        const float4 p = ppos[g];
        key = knn_bucket(G, knn_axis(p.x, G.lo[0], G.inv_h), knn_axis(p.y, G.lo[1], G.inv_h),
                         knn_axis(p.z, G.lo[2], G.inv_h));
    }
    keys[g] = key;
    idx[g] = g;
}

// Photons in bucket order (position, w = slot bits) and each bucket's [start, end) range.
__global__ void k_knn_scatter(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ idx, uint32_t n,
                              const float4* __restrict__ ppos, KnnGrid G, float4* __restrict__ spos,
                              uint32_t* __restrict__ cstart, uint32_t* __restrict__ cend) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = keys[i];
    if (b > G.mask) return;
    const uint32_t slot = idx[i];
    const float4 p = ppos[slot];
    spos[i] = make_float4(p.x, p.y, p.z, __uint_as_float(slot));
    if (i == 0 || keys[i - 1] != b) cstart[b] = i;
    if (i + 1 == n || keys[i + 1] != b) cend[b] = i + 1;
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
    const uint32_t lo = __shfl((uint32_t)v, src, 64), hi = __shfl((uint32_t)(v >> 32), src, 64);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
    const uint32_t lo = __shfl_xor((uint32_t)v, m, 64), hi = __shfl_xor((uint32_t)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}

// Bitonic sort of one key per lane, ascending by lane.
__device__ __forceinline__ uint64_t wave_sort(uint64_t v, int lane) {
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint64_t o = shfl_xor_u64(v, j);
            const bool up = (lane & k) == 0, low = (lane & j) == 0;
            v = (low == up) ? (v < o ? v : o) : (v < o ? o : v);
        }
    }
    return v;
}

// The 64 smallest of two ascending lane-sorted sequences, ascending: min(a, reverse(b)) is
// bitonic and holds them; a half-cleaner cascade sorts it.
__device__ __forceinline__ uint64_t wave_merge_low(uint64_t a, uint64_t b, int lane) {
    const uint64_t br = shfl_u64(b, 63 - lane);
    uint64_t v = a < br ? a : br;
#pragma unroll
    for (int j = 32; j > 0; j >>= 1) {
        const uint64_t o = shfl_xor_u64(v, j);
        v = (lane & j) == 0 ? (v < o ? v : o) : (v < o ? o : v);
    }
    return v;
}

// One wavefront per photonmap texel (persistent queue): the k stored photons nearest to the
// texel's surface point within the radius, ordered by (squared distance, emission slot), their
// power summed in that order and divided by their number (Photonmap::gather's normalisation,
// Lightmap.cpp:407-412). Keys are (float bits of d^2) << 32 | slot: non-negative floats order
// as their bit patterns.
__global__ __launch_bounds__(256) void k_knn_gather(const float4* __restrict__ queries, uint32_t nq,
                                                    const float4* __restrict__ spos,
                                                    const PhotonRec* __restrict__ rec,
                                                    const uint32_t* __restrict__ cstart,
                                                    const uint32_t* __restrict__ cend, KnnGrid G,
                                                    const PhotonMapDev* __restrict__ pms, int npm,
                                                    unsigned long long* queue) {
    const int lane = (int)lane_id();
    for (;;) {
        const uint64_t qi = wave_next_chunk(queue);
        if (qi >= nq) break;
        const float4 Q = queries[qi];
        const int qx = knn_axis(Q.x, G.lo[0], G.inv_h), qy = knn_axis(Q.y, G.lo[1], G.inv_h),
                  qz = knn_axis(Q.z, G.lo[2], G.inv_h);
        uint64_t best = ~0ull;
        for (int s = 0; s <= G.max_shell; s++) {
            for (int dz = -s; dz <= s; dz++) {
                for (int dy = -s; dy <= s; dy++) {
                    for (int dx = -s; dx <= s; dx++) {
                        const int ax = dx < 0 ? -dx : dx, ay = dy < 0 ? -dy : dy, az = dz < 0 ? -dz : dz;
                        if ((ax > ay ? (ax > az ? ax : az) : (ay > az ? ay : az)) != s) continue;
                        const int cx = qx + dx, cy = qy + dy, cz = qz + dz;
                        const uint32_t b = knn_bucket(G, cx, cy, cz);
                        const uint32_t st = cstart[b], en = cend[b];
                        for (uint32_t base = st; base < en; base += 64) {
                            const uint32_t i = base + (uint32_t)lane;
                            uint64_t key = ~0ull;
                            if (i < en) {
                                const float4 P = spos[i];
                                if (knn_axis(P.x, G.lo[0], G.inv_h) == cx && knn_axis(P.y, G.lo[1], G.inv_h) == cy &&
                                    knn_axis(P.z, G.lo[2], G.inv_h) == cz) {
                                    const float ex = P.x - Q.x, ey = P.y - Q.y, ez = P.z - Q.z;
                                    const float d2 = ex * ex + ey * ey + ez * ez;
                                    if (d2 <= G.r2) key = ((uint64_t)__float_as_uint(d2) << 32) | __float_as_uint(P.w);
                                }
                            }
                            const uint64_t kth = wave_uniform(shfl_u64(best, G.k - 1));
                            if (!__any(key < kth)) continue;
                            best = wave_merge_low(best, wave_sort(key, lane), lane);
                        }
                    }
                }
            }
            const uint64_t kth = wave_uniform(shfl_u64(best, G.k - 1));
            if ((uint32_t)(kth >> 32) < __float_as_uint(G.shell_d2[s])) break;
        }
        const bool mine = lane < G.k && best != ~0ull;
        const int found = __popcll(__ballot(mine));
        float pr = 0.0f, pg = 0.0f, pb = 0.0f;
        if (mine) {
            const PhotonRec r = rec[(uint32_t)best];
            pr = r.r; pg = r.g; pb = r.b;
        }
        float sr = 0.0f, sg = 0.0f, sb = 0.0f;
        for (int j = 0; j < found; j++) {
            sr += __shfl(pr, j, 64);
            sg += __shfl(pg, j, 64);
            sb += __shfl(pb, j, 64);
        }
        if (lane == 0) {
            const uint32_t key = __float_as_uint(Q.w);
            const PhotonMapDev M = pms[find_map(pms, npm, key)];
            float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (found) {
                const float fn = (float)found;
                out = make_float4(sr / fn, sg / fn, sb / fn, 0.0f);
            }
            M.gathered[key - M.key_base] = out;
        }
    }
}

uint32_t blocks_for(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

#define RL_DISPATCH_CAP(cap, ...)   \
    do {                            \
        if ((cap) == 32) {          \
            constexpr int CAP = 32; \
            __VA_ARGS__;            \
        } else {                    \
            constexpr int CAP = 64; \
            __VA_ARGS__;            \
        }                           \
    } while (0)

hipError_t launch_photon_trace(hipStream_t st, const SceneDev& S, int cap, uint32_t max_blocks,
                               const LightDev* lights, const uint64_t* light_first, const PhotonParams& P,
                               const int32_t* mesh_pm, const PhotonMapDev* pms, PhotonRec* out,
                               float4* out_pos, unsigned long long* queue, unsigned long long* rays) {
    const uint64_t nrec = P.total * (uint64_t)(P.max_depth + 1);
    if (nrec == 0) return hipSuccess;
    hipLaunchKernelGGL(k_photon_clear, dim3(blocks_for(nrec, 256)), dim3(256), 0, st, out, nrec);
    hipError_t e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(queue), 2, 0u);
    if (e != hipSuccess) return e;
    const uint64_t need = blocks_for(P.total, kBlock);
    const uint32_t grid = (uint32_t)(need < max_blocks ? need : max_blocks);
    RL_DISPATCH_CAP(cap, hipLaunchKernelGGL(k_photon_trace<CAP>, dim3(grid), dim3(kBlock), 0, st, S, lights,
                                            light_first, P, mesh_pm, pms, out, out_pos, queue, rays));
    return hipGetLastError();
}

hipError_t photon_sort_temp_bytes(uint32_t n, size_t* bytes) {
    *bytes = 0;
    return hipcub::DeviceRadixSort::SortPairs(nullptr, *bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                              (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n);
}

hipError_t launch_photon_store(hipStream_t st, const PhotonRec* rec, uint32_t n, uint32_t* keys_a,
                               uint32_t* keys_b, uint32_t* idx_a, uint32_t* idx_b, void* temp,
                               size_t temp_bytes, const PhotonMapDev* pms, int npm) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_keys_iota, dim3(blocks_for(n, 256)), dim3(256), 0, st, rec, keys_a, idx_a, n);
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys_a, keys_b, idx_a, idx_b, (int)n, 0,
                                                      32, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_photon_accumulate, dim3(blocks_for(n, 256)), dim3(256), 0, st, keys_b, idx_b, rec, n,
                       pms, npm);
    return hipGetLastError();
}

hipError_t launch_photon_gather(hipStream_t st, const PhotonMapDev* pms, int npm, uint32_t total_texels,
                                int radius) {
    if (total_texels == 0) return hipSuccess;
    if (total_texels > 0xFFFFFF00u) return hipErrorInvalidValue;  // one dispatch stays below 2^32 items
    hipLaunchKernelGGL(k_photon_gather, dim3(blocks_for(total_texels, 256)), dim3(256), 0, st, pms, npm,
                       total_texels, radius);
    return hipGetLastError();
}

hipError_t knn_sort_temp_bytes(uint32_t n, int key_bits, size_t* bytes) {
    *bytes = 0;
    return hipcub::DeviceRadixSort::SortPairs(nullptr, *bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                              (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, key_bits);
}

hipError_t launch_knn_build(hipStream_t st, const PhotonRec* rec, const float4* ppos, uint32_t n,
                            const KnnGrid& G, int key_bits, uint32_t* keys_a, uint32_t* keys_b, uint32_t* idx_a,
                            uint32_t* idx_b, void* temp, size_t temp_bytes, float4* spos, uint32_t* cstart,
                            uint32_t* cend) {
    if (n == 0) return hipSuccess;
    hipError_t e = launch_fill_u32(st, cstart, (uint64_t)G.mask + 1u, 0u);
    if (e == hipSuccess) e = launch_fill_u32(st, cend, (uint64_t)G.mask + 1u, 0u);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_knn_keys, dim3(blocks_for(n, 256)), dim3(256), 0, st, rec, ppos, n, G, keys_a, idx_a);
    e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys_a, keys_b, idx_a, idx_b, (int)n, 0, key_bits, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_knn_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, st, keys_b, idx_b, n, ppos, G, spos,
                       cstart, cend);
    return hipGetLastError();
}

hipError_t launch_knn_gather(hipStream_t st, uint32_t max_blocks, const float4* queries, uint32_t nq,
                             const float4* spos, const PhotonRec* rec, const uint32_t* cstart,
                             const uint32_t* cend, const KnnGrid& G, const PhotonMapDev* pms, int npm,
                             unsigned long long* queue) {
    if (nq == 0) return hipSuccess;
    if (G.k < 1 || G.k > 64 || G.max_shell < 0 || G.max_shell > 3) return hipErrorInvalidValue;
    hipError_t e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(queue), 2, 0u);
    if (e != hipSuccess) return e;
    const uint64_t need = blocks_for(nq, 256 / 64);
    const uint32_t grid = (uint32_t)(need < max_blocks ? need : max_blocks);
    hipLaunchKernelGGL(k_knn_gather, dim3(grid), dim3(256), 0, st, queries, nq, spos, rec, cstart, cend, G, pms,
                       npm, queue);
    return hipGetLastError();
}

hipError_t launch_final_gather(hipStream_t st, int cap, uint32_t max_blocks, const BakeSeg* segs, int nsegs,
                               uint64_t total_lumels, unsigned long long* queue, unsigned long long* rays) {
    if (total_lumels == 0) return hipSuccess;
    hipError_t e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(queue), 2, 0u);
    if (e != hipSuccess) return e;
    // max_blocks counts kBlock-thread workgroups; this kernel's are kFgBlock threads.
    const uint64_t need = blocks_for(total_lumels, kFgBlock / 64);
    const uint64_t cap_blocks = std::max<uint64_t>(1, (uint64_t)max_blocks * kBlock / kFgBlock);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(need, cap_blocks);
    RL_DISPATCH_CAP(cap, hipLaunchKernelGGL(k_final_gather<CAP>, dim3(grid), dim3(kFgBlock), 0, st, segs, nsegs,
                                            total_lumels, queue, rays));
    return hipGetLastError();
}

}  // namespace rl
