// CDNA4 (gfx950) kernels of the direct-light and AO bake passes and of segment tracing:
// DirectLight::bakeLumel (baker/DirectLight.cpp:46-125), AmbientOcclusion::bakeLumel
// (baker/AmbientOcclusion.cpp:46-66) and rt::ITracer::traceSegment (rt/Tracer.h:87).
//
// All ray kernels are persistent: a grid sized to the machine takes chunks from an atomic queue,
// so one launch covers any number of lightmaps of any size (a single dispatch may not exceed
// 2^32 work-items). Every segment carries its own scene, lights and AO settings, so one launch
// also covers lightmaps of different scenes and jobs. Per-texel reductions are integer (ballot +
// popcount or LDS counters), so results are exact and independent of scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rl_device.h"
#include "rl_kernels.h"
#include "rl_traverse.h"

namespace rl {

namespace {

constexpr int kBlock = 128;  // two wavefronts
constexpr int kDirectBlock = 768;  // direct light: twelve wavefronts, one workgroup per compute unit
using namespace trav;
// An empty occluder hint (above every triangle index).
constexpr uint32_t kNoHint = 0xFFFFFFFFu;

// LightInfluence::calculate / DirectionalLightInfluence::calculate with LightCutoff and
// LinearLightAttenuation, i.e. DirectLight::influenceFromPoint (baker/DirectLight.cpp:101-125,
// scene/Light.cpp:372-421,452-500).
// Whether the shadow segment r is blocked (any hit). A scalar-base variant for wavefronts of one
// scene (as the AO kernel has) raised this kernel to 171 VGPRs, two wavefronts per SIMD, and
// measured 25.7 ms against 21.2 ms on the city.
// `hint` is the leaf-order index of the triangle that blocked the lane's previous shadow segment
// (or above every index): it is tested first, and a hit there is a hit the traversal would also
// find, so the result does not depend on it; a blocking triangle found by traversal replaces it.
template <int CAP>
__device__ __forceinline__ bool occluded(const SceneDev& S, const TopCache& top, const Ray& r, uint32_t* stk,
                                         uint32_t& hint) {
    const WtRay w = wt_setup(r);
    if (hint < S.num_tris && hits_triangle(S, w, hint)) return true;
    TraceResult tr;
    const bool h = traverse<true, CAP>(S, top, r, w, stk, kDirectBlock, tr);
    if (h) hint = tr.leaf;
    return h;
}

// Embree::traceSegment with HitUseAlpha (rt/Embree.cpp:145-177), as the shadow tests of
// LightInfluence::calculate and DirectionalLightInfluence::calculate call it (scene/Light.cpp:385,
// 417): the closest hit of start -> end blocks the light unless its colour's alpha is <= 0.1, in
// which case the segment is traced again from the hit point plus one unit along the direction
// (the reference's `point + direction`), at most 25 times; the 26th closest hit blocks whatever
// its alpha. Every trace counts as a ray.
template <int CAP>
__device__ bool occluded_alpha(const SceneDev& S, const TopCache& top, V3 start, V3 end, uint32_t* stk,
                               uint32_t& rays) {
    for (int step = 0;; step++) {
        const Ray r = ray_from_segment(start, end);
        TraceResult tr;
        rays++;
        if (!traverse<false, CAP>(S, top, r, stk, kDirectBlock, tr)) return false;
        const float4 c = face_color(S, tr.tri, tr.V / tr.det, tr.W / tr.det);
        if (!(c.w <= 0.1f) || step >= 25) return true;
        const V3 point = start + r.d * tr.t;
        start = point + r.d;
    }
}

// A shadow segment p -> e: any hit (ALPHA = false: no material of the scene can be transparent,
// and then the closest hit exists exactly when any hit does) or the alpha re-trace.
template <int CAP, bool ALPHA>
__device__ __forceinline__ bool shadowed(const SceneDev& S, const TopCache& top, V3 p, V3 e, uint32_t* stk,
                                         uint32_t& rays, uint32_t& hint) {
    if (ALPHA) return occluded_alpha<CAP>(S, top, p, e, stk, rays);
    Ray r = ray_from_segment(p, e);
    rays++;
    return occluded<CAP>(S, top, r, stk, hint);
}

template <int CAP, bool ALPHA>
__device__ float influence_from_point(const SceneDev& S, const TopCache& top, const LightDev& L, V3 lp, V3 p, V3 n,
                                      uint32_t* stk, uint32_t& rays, uint32_t& hint) {
    float inf = 0.0f, att = 1.0f, cut = 1.0f, distance = 0.0f;
    if (L.type == LIGHT_DIRECTIONAL) {
        V3 ld = v3(L.dir[0], L.dir[1], L.dir[2]);
        inf = lambert(-ld, n);
        if (inf <= 0.001f) {
            inf = 0.0f;
        } else if (L.casts_shadow) {
            inf *= shadowed<CAP, ALPHA>(S, top, p, p - ld * 1000.0f, stk, rays, hint) ? 0.0f : 1.0f;
        }
    } else {
        V3 dir = lp - p;
        distance = normalize(dir);
        inf = lambert(dir, n);
        if (inf <= 0.001f) {
            inf = 0.0f;
        } else if (L.casts_shadow) {
            inf *= shadowed<CAP, ALPHA>(S, top, p, lp, stk, rays, hint) ? 0.0f : 1.0f;
        }
    }
    if (L.has_spot) {
        V3 sd = p - v3(L.pos[0], L.pos[1], L.pos[2]);
        normalize(sd);
        float value = dot(sd, v3(L.dir[0], L.dir[1], L.dir[2]));
        if (value <= L.cutoff) {
            cut = 0.0f;
        } else {
            value = (1.0f - (1.0f - value) * 1.0f / (1.0f - L.cutoff));
            if (fabsf(1.0f - L.spot_exponent) > 0.01f) value = ref_powf(value, L.spot_exponent);
            cut = value;
        }
    }
    if (L.has_attenuation) att = attenuation(distance, L.radius, L.att_c, L.att_l, L.att_q);
    return inf * att * cut;
}

// One lane per lumel; lights and light vertices in the reference's order. One workgroup of twelve
// wavefronts per compute unit (the occupancy the kernel's registers allow) keeps the first 880
// nodes of the first segment's scene in LDS (110 KB; with the stacks 158 KB): city direct light
// 21.2 ms without the cache, 20.0 ms with 120 nodes per two-wavefront workgroup, 19.6 ms so.
constexpr uint32_t kDirectTopNodes = 880;

template <int CAP, bool ALPHA>
__global__ __launch_bounds__(kDirectBlock) void k_bake_direct(const BakeSeg* __restrict__ segs, int nsegs,
                                                              uint64_t total, uint32_t shortcuts,
                                                              unsigned long long* queue,
                                                              unsigned long long* rays_out) {
    const bool use_hints = (shortcuts & kShortcutHints) != 0;
    __shared__ uint32_t stack[kLdsStack * kDirectBlock];
    __shared__ u32x4 top_rows[kDirectTopNodes * kNodeRows];
    const SceneDev top_scene = segs[0].scene;
    const uint32_t top_count = load_top_nodes(top_scene, top_rows, kDirectTopNodes);
    __syncthreads();
    uint32_t* stk = stack + threadIdx.x;
    const uint64_t nchunks = (total + 63) / 64;
    uint32_t rays = 0, hint = kNoHint;  // hint: the lane's last blocking triangle
    for (;;) {
        const uint64_t chunk = wave_next_chunk(queue);
        if (chunk >= nchunks) break;
        // Segments start at multiples of 64 lumels, so the chunk's lanes share one segment: its
        // scene, lights and light vertices are read into scalar registers.
        const BakeSeg& sg = segs[find_seg(segs, nsegs, chunk * 64)];
        const uint64_t g = chunk * 64 + lane_id();
        const uint32_t i = (uint32_t)(g - sg.work_begin);
        if (g >= total || i >= sg.count) continue;  // past the end, or the gap before the next segment
        const SceneDev S = sg.scene;
        const TopCache top{top_rows, S.nodes == top_scene.nodes ? top_count : 0u};
        const float4 p4 = sg.pos[i], n4 = sg.nrm[i];
        const V3 p = v3(p4.x, p4.y, p4.z), n = v3(n4.x, n4.y, n4.z);
        float4 col = sg.color[i];
        for (int l = 0; l < sg.nlights; l++) {
            const LightDev L = sg.lights[l];
            float cr, cg, cb;
            if (L.num_vertices > 0) {
                // DirectLight::lightFromPointSet (baker/DirectLight.cpp:75-98).
                cr = cg = cb = 0.0f;
                for (int v = 0; v < L.num_vertices; v++) {
                    const float4 lv = sg.lverts[L.vertex_offset + v];
                    V3 lp = v3(lv.x + L.pos[0], lv.y + L.pos[1], lv.z + L.pos[2]);
                    if (!use_hints) hint = kNoHint;
                    float inf = influence_from_point<CAP, ALPHA>(S, top, L, lp, p, n, stk, rays, hint);
                    if (inf > 0.0f) {
                        cr += L.color[0] * L.intensity * inf;
                        cg += L.color[1] * L.intensity * inf;
                        cb += L.color[2] * L.intensity * inf;
                    }
                }
                float nv = (float)L.num_vertices;
                cr = cr / nv; cg = cg / nv; cb = cb / nv;
            } else {
                // DirectLight::lightFromPoint (baker/DirectLight.cpp:63-72).
                if (!use_hints) hint = kNoHint;
                float inf = influence_from_point<CAP, ALPHA>(S, top, L, v3(L.pos[0], L.pos[1], L.pos[2]), p, n, stk, rays,
                                                             hint);
                if (inf > 0.0f) {
                    cr = L.color[0] * L.intensity * inf;
                    cg = L.color[1] * L.intensity * inf;
                    cb = L.color[2] * L.intensity * inf;
                } else {
                    cr = cg = cb = 0.0f;
                }
            }
            col.x += cr; col.y += cg; col.z += cb;
        }
        sg.color[i] = col;
    }
    count_rays(rays, rays_out);
}

// DirectLight::bakeLumel over (lumel, shadow ray) pairs, for passes with too few lumels to fill
// the machine with a lane per lumel (a lane then traces its lumel's rays one after another, so a
// pass takes the time of the slowest lumel's rays in series). Pass 1 (k_direct_pairs) computes
// each pair's influence, the same influence_from_point call k_bake_direct makes for that light
// and vertex, into `inf`, pairs lumel-major with a lumel's rays in the reference's light and
// vertex order; pass 2 (k_direct_sum) adds them per lumel in that order with k_bake_direct's
// arithmetic, so the colour is the same bit for bit. Segments' work_begin counts pairs in pass 1
// and lumels in pass 2 (two tables of the same segments in the same order).

// Shadow rays per lumel of a segment: one per point, spot or directional light, one per vertex of
// a vertex-generated (area) light; at least 1 (a pair without a light has influence 0).
__device__ __forceinline__ uint32_t direct_rays(const BakeSeg& sg) {
    uint32_t r = 0;
    for (int l = 0; l < sg.nlights; l++)
        r += sg.lights[l].num_vertices > 0 ? (uint32_t)sg.lights[l].num_vertices : 1u;
    return r > 0u ? r : 1u;
}

template <int CAP, bool ALPHA>
__global__ __launch_bounds__(kDirectBlock) void k_direct_pairs(const BakeSeg* __restrict__ segs, int nsegs,
                                                               uint64_t total, uint32_t shortcuts,
                                                               float* __restrict__ inf_out, unsigned long long* queue,
                                                               unsigned long long* rays_out) {
    const bool use_hints = (shortcuts & kShortcutHints) != 0;
    __shared__ uint32_t stack[kLdsStack * kDirectBlock];
    __shared__ u32x4 top_rows[kDirectTopNodes * kNodeRows];
    const SceneDev top_scene = segs[0].scene;
    const uint32_t top_count = load_top_nodes(top_scene, top_rows, kDirectTopNodes);
    __syncthreads();
    uint32_t* stk = stack + threadIdx.x;
    const uint64_t nchunks = (total + 63) / 64;
    uint32_t rays = 0, hint = kNoHint;  // hint: the lane's last blocking triangle
    for (;;) {
        const uint64_t chunk = wave_next_chunk(queue);
        if (chunk >= nchunks) break;
        const uint64_t g = chunk * 64 + lane_id();
        if (g >= total) continue;
        const BakeSeg& sg = segs[find_seg(segs, nsegs, g)];
        const uint64_t local = g - sg.work_begin;
        const uint32_t R = direct_rays(sg);
        const uint32_t i = (uint32_t)(local / R), r = (uint32_t)(local % R);
        // The light of ray r and, for an area light, its vertex.
        int l = 0, v = -1;
        for (uint32_t first = 0; l < sg.nlights; l++) {
            const int nv = sg.lights[l].num_vertices;
            const uint32_t k = nv > 0 ? (uint32_t)nv : 1u;
            if (r < first + k) {
                v = nv > 0 ? (int)(r - first) : -1;
                break;
            }
            first += k;
        }
        float inf = 0.0f;
        if (l < sg.nlights) {
            const SceneDev S = sg.scene;
            const TopCache top{top_rows, S.nodes == top_scene.nodes ? top_count : 0u};
            const LightDev L = sg.lights[l];
            const float4 p4 = sg.pos[i], n4 = sg.nrm[i];
            const V3 p = v3(p4.x, p4.y, p4.z), n = v3(n4.x, n4.y, n4.z);
            V3 lp = v3(L.pos[0], L.pos[1], L.pos[2]);
            if (v >= 0) {
                const float4 lv = sg.lverts[L.vertex_offset + v];
                lp = v3(lv.x + L.pos[0], lv.y + L.pos[1], lv.z + L.pos[2]);
            }
            if (!use_hints) hint = kNoHint;
            inf = influence_from_point<CAP, ALPHA>(S, top, L, lp, p, n, stk, rays, hint);
        }
        inf_out[g] = inf;
    }
    count_rays(rays, rays_out);
}

__global__ __launch_bounds__(kBlock) void k_direct_sum(const BakeSeg* __restrict__ lumel_segs,
                                                       const BakeSeg* __restrict__ pair_segs, int nsegs,
                                                       uint64_t total, const float* __restrict__ inf) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x; g < total; g += stride) {
        const int s = find_seg(lumel_segs, nsegs, g);
        const BakeSeg& sg = lumel_segs[s];
        const uint32_t i = (uint32_t)(g - sg.work_begin);
        if (i >= sg.count) continue;  // the gap before the next segment (lumel tables aligned to 64)
        const float* f = inf + pair_segs[s].work_begin + (uint64_t)i * direct_rays(sg);
        float4 col = sg.color[i];
        uint32_t r = 0;
        for (int l = 0; l < sg.nlights; l++) {
            const LightDev L = sg.lights[l];
            float cr, cg, cb;
            if (L.num_vertices > 0) {
                cr = cg = cb = 0.0f;
                for (int v = 0; v < L.num_vertices; v++) {
                    const float x = f[r++];
                    if (x > 0.0f) {
                        cr += L.color[0] * L.intensity * x;
                        cg += L.color[1] * L.intensity * x;
                        cb += L.color[2] * L.intensity * x;
                    }
                }
                float nv = (float)L.num_vertices;
                cr = cr / nv; cg = cg / nv; cb = cb / nv;
            } else {
                const float x = f[r++];
                if (x > 0.0f) {
                    cr = L.color[0] * L.intensity * x;
                    cg = L.color[1] * L.intensity * x;
                    cb = L.color[2] * L.intensity * x;
                } else {
                    cr = cg = cb = 0.0f;
                }
            }
            col.x += cr; col.y += cg; col.z += cb;
        }
        sg.color[i] = col;
    }
}

// AO pass 1 (chunked): one lane per (lumel, sample) pair, pairs lumel-major so a lumel's rays
// share a wavefront; segments' work_begin counts pairs. Occluded rays are counted per lumel with
// a ballot, one integer atomic per run.
template <int CAP>
__global__ __launch_bounds__(kBlock) void k_ao_count(const BakeSeg* __restrict__ segs, int nsegs, uint64_t total,
                                                     unsigned long long* queue, unsigned long long* rays_out) {
    __shared__ uint32_t stack[kLdsStack * kBlock];
    uint32_t* stk = stack + threadIdx.x;
    const uint32_t lane = lane_id();
    const uint64_t nchunks = (total + 63) / 64;
    uint64_t rays = 0;
    for (;;) {
        const uint64_t chunk = wave_next_chunk(queue);
        if (chunk >= nchunks) break;
        const uint64_t g = chunk * 64 + lane;
        const bool active = g < total;
        bool occ = false;
        uint32_t* counter = nullptr;
        uint64_t key = ~0ull;
        if (active) {
            const int s = find_seg(segs, nsegs, g);
            const BakeSeg& sg = segs[s];
            const AoParams P = sg.ao;
            const uint64_t local = g - sg.work_begin;
            const uint32_t li = (uint32_t)(local / (uint64_t)P.samples);
            const uint32_t si = (uint32_t)(local % (uint64_t)P.samples);
            const float4 p4 = sg.pos[li], n4 = sg.nrm[li];
            const V3 p = v3(p4.x, p4.y, p4.z), n = v3(n4.x, n4.y, n4.z);
            Pcg32 rng = sample_stream(P.seed, sg.lm_seed, f_bits(p4.w), TECH_AO, si);
            V3 dir = sample_hemisphere(rng, n, P.sampling);
            Ray r = ray_from_segment(p, p + dir * P.max_distance);
            TraceResult tr;
            occ = traverse<true, CAP>(sg.scene, r, stk, kBlock, tr);
            counter = sg.counts + li;
            key = ((uint64_t)s << 32) | li;
        }
        const uint64_t occ_mask = __ballot(occ);
        const uint64_t act_mask = __ballot(active);
        const uint64_t prev_key = __shfl_up(key, 1, 64);
        const bool head = active && (lane == 0 || prev_key != key);
        const uint64_t heads = __ballot(head);
        if (head) {
            const uint64_t above = (lane == 63) ? 0ull : (heads & (~0ull << (lane + 1)));
            const uint64_t end_mask = above ? (above & (~above + 1ull)) : 0ull;  // lowest head above
            const uint64_t run = (end_mask ? (end_mask - 1ull) : ~0ull) & (~0ull << lane);
            const uint32_t c = (uint32_t)__popcll(occ_mask & run & act_mask);
            if (c) atomicAdd(counter, c);
        }
        rays += (uint64_t)__popcll(act_mask);
    }
    if (lane == 0 && rays) atomicAdd(rays_out, (unsigned long long)rays);
}

// AO pass 1 with lane refill: each wavefront takes chunks of up to `chunk_lumels` lumels (one
// segment each; segments' work_begin counts chunks) and keeps a pool of their (lumel, sample)
// rays; whenever at least `refill_min` lanes have finished their ray, those lanes take the next
// rays of the pool (ballot + prefix count), so a wavefront no longer waits for its slowest ray.
// Occlusion counts per lumel accumulate in an LDS ring of kRing slots per wavefront; a ray may
// only start a lumel fewer than kRing - 2 places after the oldest lumel still in flight, and the
// lane that completes a lumel's last ray writes its count.
constexpr int kRing = 32;
// A wavefront's node phase stops once no more than this many lanes are still searching, so that
// finished lanes take new rays instead of idling until the slowest ray of the wavefront reaches a
// leaf (city AO: 399 ms with no early stop, 397/393/387/384/387/398 ms at 1/2/4/8/16/24 lanes).
constexpr uint32_t kAoNodeExit = 8;
// One workgroup of sixteen wavefronts per compute unit (the occupancy the kernel's registers
// allow), so one LDS copy of the scene's top nodes serves all of them: before local trees, city AO
// 360 ms with four-wavefront workgroups and 152 nodes, 350 ms with eight and 304, 352/347 ms with
// sixteen and 400/600. With local trees, 768-thread workgroups (no register spills, 560 nodes)
// measured 256.6 ms against 237.8 ms for sixteen wavefronts.
constexpr int kAoBlock = 1024;
static_assert(kAoBlock / 64 == (int)kAoRefillWavesPerCu, "the launcher's chunk choice counts these wavefronts");
constexpr uint32_t kAoTopNodes = 296;
// Each wavefront builds a local tree (rl_traverse.h) for a chunk of at least kVtMinRays rays, over
// the union of the boxes its lumels' AO segments stay in, and starts the chunk's rays at its
// virtual root. Rays of several chunks can be in flight at once, so a wavefront keeps three trees
// and builds into one only when no ray of the chunk that used it last is still in flight; otherwise
// (and for smaller chunks) the chunk's rays start at the scene's root. LDS: 30 KB of virtual
// nodes, paid for with the top-node cache (600 -> 296 nodes with the hint rows and ring records
// below), which local trees mostly skip.
constexpr uint32_t kVtBuffers = 3;
constexpr uint32_t kVtMinRays = 64;

// The box that the AO segments of a lumel stay in: from p, length D, directions of the sampler.
// Cosine: d = t1 x + t2 y + n z with x^2 + y^2 <= 1 and 0 <= z <= 1, so each component of d lies
// within sqrt(t1_a^2 + t2_a^2) of [min(0, n_a), max(0, n_a)]; uniform: any unit vector. The margin
// covers the rounding of the direction, its normalisation and the segment end point.
__device__ __forceinline__ QueryBox ao_segment_box(V3 p, const CosineFrame& f, float D, bool cosine) {
    QueryBox q;
    const float pa[3] = {p.x, p.y, p.z};
    const float t1[3] = {f.t1.x, f.t1.y, f.t1.z}, t2[3] = {f.t2.x, f.t2.y, f.t2.z}, n[3] = {f.n.x, f.n.y, f.n.z};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float dlo = -1.0f, dhi = 1.0f;
        if (cosine) {
            const float s = sqrtf(t1[a] * t1[a] + t2[a] * t2[a]);
            dlo = fminf(n[a], 0.0f) - s;
            dhi = fmaxf(n[a], 0.0f) + s;
        }
        const float pad = 1e-3f * D + 1e-5f * (fabsf(pa[a]) + D);
        q.lo[a] = pa[a] + D * dlo - pad;
        q.hi[a] = pa[a] + D * dhi + pad;
    }
    return q;
}

template <int CAP>
__global__ __launch_bounds__(kAoBlock) void k_ao_refill(const BakeSeg* __restrict__ segs, int nsegs,
                                                      uint64_t total_chunks, uint32_t chunk_lumels,
                                                      uint32_t refill_min, uint32_t shortcuts,
                                                      unsigned long long* queue, unsigned long long* rays_out) {
    const bool use_trees = (shortcuts & kShortcutTrees) != 0, use_hints = (shortcuts & kShortcutHints) != 0;
    __shared__ uint32_t stack[kLdsStack * kAoBlock];
    // Per wavefront and ring slot (one lumel in flight): occluded rays so far, rays still to
    // finish, and where the lumel's count goes. A lumel's first ray sets them up.
    __shared__ uint32_t ring_occ[kAoBlock / 64][kRing];
    __shared__ uint32_t ring_left[kAoBlock / 64][kRing];
    __shared__ uint32_t* ring_out[kAoBlock / 64][kRing];
    // The scene's first kAoTopNodes nodes, then the wavefronts' local trees (virtual nodes).
    __shared__ u32x4 top_rows[(kAoTopNodes + kAoBlock / 64 * kVtBuffers * kVtNodes) * kNodeRows];
    // Per wavefront, rows of the lumels of the chunks taken, lumel stream number s in row
    // s % kAoChunkMax: position and texel key, cosine-sampling frame (t1, t2, n), computed once
    // per lumel when its chunk is taken instead of once per ray, and row 4: the lumel's three
    // occluder hints and its stream number (which tells whether a later chunk has reused the row).
    __shared__ u32x4 lumel_rows[kAoBlock / 64][kAoChunkMax][5];
    const SceneDev top_scene = segs[0].scene;
    const uint32_t top_count = load_top_nodes(top_scene, top_rows, kAoTopNodes);
    uint32_t* stk = stack + threadIdx.x;
    uint32_t spill[CAP - kLdsStack];
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;

    // Pool state, identical in every lane of the wavefront.
    int pseg = 0;
    uint32_t chunk_first = 0, chunk_count = 0, chunk_ns = 1;
    uint32_t pool_next = 0, pool_end = 0, stream_base = 0;
    uint32_t pool_lc = 0, pool_si = 0;  // pool_next = pool_lc * chunk_ns + pool_si
    bool pool_open = true;
    // Local trees: trees built, each buffer's end stream (one past its chunk's last lumel), the
    // current chunk's start node.
    uint32_t vt_seq = 0, vt_end0 = 0, vt_end1 = 0, vt_end2 = 0, chunk_root = 0;
    uint32_t rays = 0;  // rays finished by the wavefront
    // Lane state.
    bool active = false, fresh = false;
    TravRay tr;
    SceneDev scene;
    uint32_t node = 0, my_stream = 0;
    uint32_t hint0 = kNoHint, hint1 = kNoHint, hint2 = kNoHint;
    int sp = 0;
    TopCache top{top_rows, 0u};
    // A lane's ray is done: add it to its lumel's ring slot and publish the lumel's occlusion count
    // when this was its last ray.
    auto finish = [&](bool occluded) {
        const uint32_t slot = my_stream % (uint32_t)kRing;
        if (occluded) atomicAdd(&ring_occ[wave][slot], 1u);
        if (atomicSub(&ring_left[wave][slot], 1u) == 1u) {
            // Last ray of the lumel: every earlier increment of this slot precedes this one in
            // the wavefront's LDS order.
            *ring_out[wave][slot] = atomicAdd(&ring_occ[wave][slot], 0u);
        }
        active = false;
    };
    __syncthreads();
    for (;;) {
        uint64_t idle = __ballot(!active);
        if (pool_open && ((uint32_t)__popcll(idle) >= refill_min || idle == ~0ull)) {
            uint32_t oldest = active ? my_stream : 0xFFFFFFFFu;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) oldest = min(oldest, (uint32_t)__shfl_xor((int)oldest, off, 64));
            // Every lane holds the minimum now. Said so, the pool arithmetic below (the ring window,
            // room, take, the pool counters) runs in scalar registers and branches on the sample
            // count are scalar branches instead of selects over both paths.
            oldest = wave_uniform32(oldest);
            const uint64_t allow = oldest == 0xFFFFFFFFu ? ~0ull : (uint64_t)oldest + (uint64_t)(kRing - 2);
            // The lowest stream in flight: the oldest active ray's, or the first one this refill
            // hands out (streams are handed out in increasing order).
            uint32_t lowest = oldest;
            uint64_t want = idle;
            while (want && pool_open) {
                if (pool_next == pool_end) {
                    const uint64_t c = wave_next_chunk(queue);
                    if (c >= total_chunks) {
                        pool_open = false;
                        break;
                    }
                    stream_base += chunk_count;
                    pseg = find_seg(segs, nsegs, c);
                    const BakeSeg& sg = segs[pseg];
                    chunk_first = (uint32_t)(c - sg.work_begin) * chunk_lumels;
                    chunk_count = min(chunk_lumels, sg.count - chunk_first);
                    chunk_ns = (uint32_t)sg.ao.samples;
                    pool_next = 0;
                    pool_lc = pool_si = 0;
                    pool_end = chunk_count * chunk_ns;
                    QueryBox q = {{__builtin_inff(), __builtin_inff(), __builtin_inff()},
                                  {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()}};
                    if (lane < chunk_count) {  // chunk_count <= kAoChunkMax (launcher)
                        const uint32_t li = chunk_first + lane;
                        const uint32_t stream = stream_base + lane;
                        const float4 p4 = sg.pos[li], n4 = sg.nrm[li];
                        const uint64_t key = texel_key(sg.ao.seed, sg.lm_seed, f_bits(p4.w));
                        const CosineFrame f = cosine_frame(v3(n4.x, n4.y, n4.z));
                        u32x4* row = lumel_rows[wave][stream % kAoChunkMax];
                        row[0] = u32x4{f_bits(p4.x), f_bits(p4.y), f_bits(p4.z), (uint32_t)key};
                        row[1] = u32x4{f_bits(f.t1.x), f_bits(f.t1.y), f_bits(f.t1.z), (uint32_t)(key >> 32)};
                        row[2] = u32x4{f_bits(f.t2.x), f_bits(f.t2.y), f_bits(f.t2.z), 0u};
                        row[3] = u32x4{f_bits(n4.x), f_bits(n4.y), f_bits(n4.z), 0u};
                        row[4] = u32x4{kNoHint, kNoHint, kNoHint, stream};
                        q = ao_segment_box(v3(p4.x, p4.y, p4.z), f, sg.ao.max_distance,
                                           sg.ao.sampling == SAMPLE_COSINE);
                    }
                    // Other lanes of this wavefront read the rows below; the wavefront's LDS
                    // operations complete in order.
                    __builtin_amdgcn_wave_barrier();
                    // The chunk's local tree over the union of its lumels' segment boxes.
#pragma unroll
                    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
                        for (int a = 0; a < 3; a++) {
                            q.lo[a] = fminf(q.lo[a], __shfl_xor(q.lo[a], off, 64));
                            q.hi[a] = fmaxf(q.hi[a], __shfl_xor(q.hi[a], off, 64));
                        }
                    }
                    const uint32_t buf = vt_seq % kVtBuffers;
                    const uint32_t buf_end = buf == 0 ? vt_end0 : (buf == 1 ? vt_end1 : vt_end2);
                    chunk_root = 0u;
                    if (use_trees && chunk_count * chunk_ns >= kVtMinRays && lowest >= buf_end) {
                        const uint32_t first = kAoTopNodes + (wave * kVtBuffers + buf) * kVtNodes;
                        const TopCache btop{top_rows, sg.scene.nodes == top_scene.nodes ? top_count : 0u};
                        chunk_root = build_local_tree(sg.scene, btop, q, top_rows + first * kNodeRows, first);
                        const uint32_t end = stream_base + chunk_count;
                        if (buf == 0) vt_end0 = end;
                        else if (buf == 1) vt_end1 = end;
                        else vt_end2 = end;
                        vt_seq++;
                    }
                }
                uint64_t room = pool_end - pool_next;
                if (allow != ~0ull) {
                    const uint64_t lim = allow > stream_base ? (allow - stream_base) * (uint64_t)chunk_ns : 0ull;
                    room = lim > pool_next ? min(room, lim - pool_next) : 0ull;
                }
                const uint32_t take = wave_uniform32((uint32_t)min(room, (uint64_t)__popcll(want)));
                if (take == 0) break;
                lowest = min(lowest, stream_base + pool_lc);
                const uint32_t rank = rank_below(want);
                const bool mine = ((want >> lane) & 1ull) && rank < take;
                // Pair pool_next + rank as (lumel, sample): with at least 64 samples per lumel it
                // is at most one lumel past (pool_lc, pool_si). The branch is on a wave-uniform
                // count and outside the lanes' own branch, so only one of the two ways runs.
                const uint32_t t = pool_si + rank;
                uint32_t lc, si;
                if (chunk_ns >= 64u) {
                    const bool wrap = t >= chunk_ns;
                    lc = pool_lc + (wrap ? 1u : 0u);
                    si = wrap ? t - chunk_ns : t;
                } else {
                    lc = pool_lc + t / chunk_ns;
                    si = t % chunk_ns;
                }
                if (mine) {
                    // The segment index is the same in every lane: read as such, the segment's
                    // fields load through the scalar path.
                    const BakeSeg& sg = segs[__builtin_amdgcn_readfirstlane(pseg)];
                    const uint32_t li = chunk_first + lc;
                    my_stream = stream_base + lc;
                    const u32x4* row = lumel_rows[wave][my_stream % kAoChunkMax];
                    const u32x4 r0 = row[0], r1 = row[1], r2 = row[2], r3 = row[3];
                    const V3 p = v3(bits_f(r0.x), bits_f(r0.y), bits_f(r0.z));
                    CosineFrame f;
                    f.t1 = v3(bits_f(r1.x), bits_f(r1.y), bits_f(r1.z));
                    f.t2 = v3(bits_f(r2.x), bits_f(r2.y), bits_f(r2.z));
                    f.n = v3(bits_f(r3.x), bits_f(r3.y), bits_f(r3.z));
                    Pcg32 rng = texel_stream(((uint64_t)r1.w << 32) | r0.w, TECH_AO, si);
                    const V3 dir = sg.ao.sampling == SAMPLE_COSINE ? sample_cosine_in_frame(rng, f)
                                                                   : sample_uniform_hemisphere(rng, f.n);
                    tr = trav_ray(ray_from_segment(p, p + dir * sg.ao.max_distance));
                    scene = sg.scene;
                    top.count = sg.scene.nodes == top_scene.nodes ? top_count : 0u;
                    node = chunk_root;
                    sp = 0;
                    if (si == 0) {
                        // The lumel's first ray: its ring slot is free (a ray starts only lumels
                        // fewer than kRing - 2 places after the oldest one in flight).
                        const uint32_t slot = my_stream % (uint32_t)kRing;
                        ring_occ[wave][slot] = 0u;
                        ring_left[wave][slot] = chunk_ns;
                        ring_out[wave][slot] = sg.counts + li;
                    }
                    // A lumel none of whose rays has been occluded yet starts from the hints of
                    // the lumel before it (the previous texel of its lightmap row, mostly), if its
                    // row still holds them.
                    u32x4 h = row[4];
                    if (h.x == kNoHint) {
                        const u32x4 hp = lumel_rows[wave][(my_stream - 1u) % kAoChunkMax][4];
                        if (hp.w == my_stream - 1u) h = hp;
                    }
                    hint0 = h.x;
                    hint1 = h.y;
                    hint2 = h.z;
                    active = true;
                    fresh = true;
                }
                pool_next += take;
                pool_si += take;
                if (chunk_ns >= 64u) {
                    if (pool_si >= chunk_ns) { pool_si -= chunk_ns; pool_lc++; }
                } else {
                    pool_lc += pool_si / chunk_ns;
                    pool_si %= chunk_ns;
                }
                want &= ~__ballot(mine);
            }
            // Occluder hints: a new ray first tests the last three triangles that occluded rays of
            // its lumel. A hit there is a hit the traversal would also find (the same watertight
            // test in the ray's interval), so the ray is occluded and completes at once; the lanes
            // it frees take new rays before the wavefront traverses. Hints only choose which
            // triangles are tested first: the occlusion result does not depend on them.
            bool by_hint = false;
            if (fresh && use_hints) {
                if (hint0 < scene.num_tris) by_hint = hits_triangle(scene, tr, hint0);
                if (!by_hint && hint1 < scene.num_tris) by_hint = hits_triangle(scene, tr, hint1);
                if (!by_hint && hint2 < scene.num_tris) by_hint = hits_triangle(scene, tr, hint2);
                if (by_hint) finish(true);
            }
            fresh = false;
            const uint64_t hinted = __ballot(by_hint);
            rays += (uint32_t)__popcll(hinted);
            if (hinted != 0ull && pool_open && (uint32_t)__popcll(__ballot(!active)) >= refill_min) continue;
            if (!pool_open && __ballot(active) == 0ull) break;
        } else if (!pool_open && idle == ~0ull) {
            break;
        }
        bool ended = false;
        if (active) {
            const uint32_t min_lanes = pool_open ? kAoNodeExit : 0u;
            uint32_t hit = kNoHint;
            const int st = one_scene(scene)
                               ? any_hit_run<CAP, true, true>(scene, top, tr, node, sp, stk, kAoBlock, spill, min_lanes,
                                                              hit)
                               : any_hit_run<CAP, false, true>(scene, top, tr, node, sp, stk, kAoBlock, spill,
                                                               min_lanes, hit);
            if (st > 0 && use_hints) {
                // Move the occluder to the front of the lumel's hints (most recent first) while the
                // row still holds this ray's lumel.
                uint32_t* w = reinterpret_cast<uint32_t*>(lumel_rows[wave][my_stream % kAoChunkMax]);
                if (w[4 * 4 + 3] == my_stream) {
                    const uint32_t h0 = w[4 * 4 + 0], h1 = w[4 * 4 + 1];
                    if (h0 != hit) {
                        if (h1 != hit) w[4 * 4 + 2] = h1;
                        w[4 * 4 + 1] = h0;
                        w[4 * 4 + 0] = hit;
                    }
                }
            }
            if (st != 0) finish(st > 0);
            ended = st != 0;
        }
        rays += (uint32_t)__popcll(__ballot(ended));
    }
    if (lane == 0 && rays) atomicAdd(rays_out, (unsigned long long)rays);
}

// AO pass 2: AmbientOcclusion::bakeLumel's value and multiply (baker/AmbientOcclusion.cpp:59-65).
__global__ __launch_bounds__(kBlock) void k_ao_finalize(const BakeSeg* __restrict__ segs, int nsegs, uint64_t total) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x; g < total; g += stride) {
        const BakeSeg& sg = segs[find_seg(segs, nsegs, g)];
        const AoParams P = sg.ao;
        const uint32_t i = (uint32_t)(g - sg.work_begin);
        const uint32_t occluded = sg.counts[i];
        float value = 1.0f - (float)occluded / ((float)P.samples * P.occluded_fraction);
        if (fabsf(1.0f - P.exponent) > 0.01f) value = ref_powf(value, P.exponent);
        float4 c = sg.color[i];
        c.x *= value; c.y *= value; c.z *= value;
        sg.color[i] = c;
        sg.counts[i] = 0u;
    }
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += stride) p[g] = v;
}

template <bool ANY, int CAP>
__global__ __launch_bounds__(kBlock) void k_trace(SceneDev S, const float4* __restrict__ starts,
                                                  const float4* __restrict__ ends, uint32_t n,
                                                  uint8_t* __restrict__ flags, float4* __restrict__ hits,
                                                  unsigned long long* queue, unsigned long long* rays_out) {
    __shared__ uint32_t stack[kLdsStack * kBlock];
    const uint64_t nchunks = ((uint64_t)n + 63) / 64;
    uint32_t rays = 0;
    for (;;) {
        const uint64_t chunk = wave_next_chunk(queue);
        if (chunk >= nchunks) break;
        const uint64_t g = chunk * 64 + lane_id();
        if (g >= n) continue;
        const float4 a = starts[g], b = ends[g];
        Ray r = ray_from_segment(v3(a.x, a.y, a.z), v3(b.x, b.y, b.z));
        TraceResult tr;
        rays++;
        bool h = traverse<ANY, CAP>(S, r, stack + threadIdx.x, kBlock, tr);
        if (ANY) {
            flags[g] = h ? 1 : 0;
        } else {
            // Embree barycentrics: u weights v1 (V/det), v weights v2 (W/det).
            const float u = h ? tr.V / tr.det : 0.0f;
            const float v = h ? tr.W / tr.det : 0.0f;
            hits[g] = make_float4(tr.t, bits_f(tr.tri), u, v);
        }
    }
    count_rays(rays, rays_out);
}

// Lumels `sel` of a lightmap copied to a compact segment (position, normal, colour) and the
// baked colours copied back: the bake kernels run on the selection as on a whole lightmap (the
// texel id in position.w keys the sample streams, so results equal a whole-lightmap bake).
__global__ void k_gather_lumels(const uint32_t* __restrict__ sel, uint32_t n, const float4* __restrict__ pos,
                                const float4* __restrict__ nrm, const float4* __restrict__ col,
                                float4* __restrict__ opos, float4* __restrict__ onrm, float4* __restrict__ ocol) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = sel[i];
    opos[i] = pos[j];
    onrm[i] = nrm[j];
    ocol[i] = col[j];
}

__global__ void k_scatter_colors(const uint32_t* __restrict__ sel, uint32_t n, const float4* __restrict__ col,
                                 float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[sel[i]] = col[i];
}

uint32_t blocks_for(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

// Persistent grid: enough blocks for `waves` wavefronts of work, at most max_blocks.
uint32_t grid_for_waves(uint64_t waves, uint32_t max_blocks) {
    const uint64_t need = (waves + (kBlock / 64) - 1) / (kBlock / 64);
    return (uint32_t)(need < max_blocks ? (need ? need : 1) : max_blocks);
}

}  // namespace

int stack_capacity_for(int bound) {
    if (bound <= 32) return 32;
    if (bound <= 64) return 64;
    return -1;
}

#define RL_DISPATCH_CAP(cap, ...)          \
    do {                                   \
        if ((cap) == 32) {                 \
            constexpr int CAP = 32;        \
            __VA_ARGS__;                   \
        } else {                           \
            constexpr int CAP = 64;        \
            __VA_ARGS__;                   \
        }                                  \
    } while (0)

hipError_t launch_gather_lumels(hipStream_t st, const uint32_t* sel, uint32_t n, const float4* pos, const float4* nrm,
                                const float4* col, float4* opos, float4* onrm, float4* ocol) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_lumels, dim3(blocks_for(n, 256)), dim3(256), 0, st, sel, n, pos, nrm, col, opos, onrm,
                       ocol);
    return hipGetLastError();
}

hipError_t launch_scatter_colors(hipStream_t st, const uint32_t* sel, uint32_t n, const float4* col, float4* out) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_colors, dim3(blocks_for(n, 256)), dim3(256), 0, st, sel, n, col, out);
    return hipGetLastError();
}

hipError_t launch_fill_u32(hipStream_t st, uint32_t* p, uint64_t n, uint32_t v) {
    if (n == 0) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(blocks_for(n, 256), 65536);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_bake_direct(hipStream_t st, int cap, uint32_t max_blocks, const BakeSeg* segs, int nsegs,
                              uint64_t total, bool alpha, uint32_t shortcuts, unsigned long long* queue,
                              unsigned long long* rays) {
    if (total == 0) return hipSuccess;
    hipError_t e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(queue), 2, 0u);
    if (e != hipSuccess) return e;
    // max_blocks counts kBlock-thread workgroups; this kernel's are kDirectBlock threads.
    const uint64_t need = ((total + 63) / 64 + (kDirectBlock / 64) - 1) / (kDirectBlock / 64);
    const uint32_t cap_blocks = std::max<uint32_t>(1u, max_blocks * (uint32_t)kBlock / (uint32_t)kDirectBlock);
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(need, cap_blocks));
    if (alpha) {
        RL_DISPATCH_CAP(cap, hipLaunchKernelGGL((k_bake_direct<CAP, true>), dim3(grid), dim3(kDirectBlock), 0, st, segs,
                                                nsegs, total, shortcuts, queue, rays));
    } else {
        RL_DISPATCH_CAP(cap, hipLaunchKernelGGL((k_bake_direct<CAP, false>), dim3(grid), dim3(kDirectBlock), 0, st, segs,
                                                nsegs, total, shortcuts, queue, rays));
    }
    return hipGetLastError();
}

hipError_t launch_bake_direct_pairs(hipStream_t st, int cap, uint32_t max_blocks, const BakeSeg* pair_segs,
                                   const BakeSeg* lumel_segs, int nsegs, uint64_t total_pairs, uint64_t total_lumels,
                                   bool alpha, uint32_t shortcuts, float* inf, unsigned long long* queue,
                                   unsigned long long* rays) {
    if (total_pairs == 0) return hipSuccess;
    hipError_t e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(queue), 2, 0u);
    if (e != hipSuccess) return e;
    const uint64_t need = ((total_pairs + 63) / 64 + (kDirectBlock / 64) - 1) / (kDirectBlock / 64);
    const uint32_t cap_blocks = std::max<uint32_t>(1u, max_blocks * (uint32_t)kBlock / (uint32_t)kDirectBlock);
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(need, cap_blocks));
    if (alpha) {
        RL_DISPATCH_CAP(cap, hipLaunchKernelGGL((k_direct_pairs<CAP, true>), dim3(grid), dim3(kDirectBlock), 0, st,
                                                pair_segs, nsegs, total_pairs, shortcuts, inf, queue, rays));
    } else {
        RL_DISPATCH_CAP(cap, hipLaunchKernelGGL((k_direct_pairs<CAP, false>), dim3(grid), dim3(kDirectBlock), 0, st,
                                                pair_segs, nsegs, total_pairs, shortcuts, inf, queue, rays));
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(blocks_for(total_lumels, kBlock), 65536);
    hipLaunchKernelGGL(k_direct_sum, dim3(blocks), dim3(kBlock), 0, st, lumel_segs, pair_segs, nsegs, total_lumels, inf);
    return hipGetLastError();
}

hipError_t launch_ao_count(hipStream_t st, int cap, uint32_t max_blocks, const BakeSeg* segs, int nsegs,
                           uint64_t total_pairs, unsigned long long* queue, unsigned long long* rays) {
    if (total_pairs == 0) return hipSuccess;
    hipError_t e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(queue), 2, 0u);
    if (e != hipSuccess) return e;
    const uint32_t grid = grid_for_waves((total_pairs + 63) / 64, max_blocks);
    RL_DISPATCH_CAP(cap, hipLaunchKernelGGL(k_ao_count<CAP>, dim3(grid), dim3(kBlock), 0, st, segs, nsegs,
                                            total_pairs, queue, rays));
    return hipGetLastError();
}

hipError_t launch_ao_refill(hipStream_t st, int cap, uint32_t max_blocks, const BakeSeg* segs, int nsegs,
                            uint64_t total_chunks, uint32_t chunk_lumels, uint32_t refill_min, uint32_t shortcuts,
                            unsigned long long* queue, unsigned long long* rays) {
    if (total_chunks == 0) return hipSuccess;
    if (chunk_lumels == 0 || chunk_lumels > kAoChunkMax || refill_min == 0 || refill_min > 64)
        return hipErrorInvalidValue;
    hipError_t e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(queue), 2, 0u);
    if (e != hipSuccess) return e;
    // max_blocks counts kBlock-thread workgroups; this kernel's are kAoBlock threads.
    const uint64_t need = (total_chunks + (kAoBlock / 64) - 1) / (kAoBlock / 64);
    const uint32_t cap_blocks = std::max<uint32_t>(1u, max_blocks * (uint32_t)kBlock / (uint32_t)kAoBlock);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(need, cap_blocks);
    RL_DISPATCH_CAP(cap, hipLaunchKernelGGL(k_ao_refill<CAP>, dim3(grid), dim3(kAoBlock), 0, st, segs, nsegs,
                                            total_chunks, chunk_lumels, refill_min, shortcuts, queue, rays));
    return hipGetLastError();
}

hipError_t launch_ao_finalize(hipStream_t st, const BakeSeg* segs, int nsegs, uint64_t total) {
    if (total == 0) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(blocks_for(total, kBlock), 65536);
    hipLaunchKernelGGL(k_ao_finalize, dim3(blocks), dim3(kBlock), 0, st, segs, nsegs, total);
    return hipGetLastError();
}

hipError_t launch_trace(hipStream_t st, const SceneDev& S, int cap, uint32_t max_blocks, bool any,
                        const float4* starts, const float4* ends, uint32_t n, uint8_t* flags,
                        float4* hits, unsigned long long* queue, unsigned long long* rays) {
    if (n == 0) return hipSuccess;
    hipError_t e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(queue), 2, 0u);
    if (e != hipSuccess) return e;
    const dim3 grid(grid_for_waves(((uint64_t)n + 63) / 64, max_blocks)), block(kBlock);
    if (any) {
        RL_DISPATCH_CAP(cap, hipLaunchKernelGGL((k_trace<true, CAP>), grid, block, 0, st, S, starts, ends, n, flags,
                                                hits, queue, rays));
    } else {
        RL_DISPATCH_CAP(cap, hipLaunchKernelGGL((k_trace<false, CAP>), grid, block, 0, st, S, starts, ends, n, flags,
                                                hits, queue, rays));
    }
    return hipGetLastError();
}

}  // namespace rl
