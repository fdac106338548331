// Photon emission and tracing, the deterministic photon store, the photonmap gather and the
// final gather: Photons::emit/trace/store (baker/Photons.cpp:50-144), Photonmap::gather
// (Lightmap.cpp:376-413) and IndirectLight::bakeLumel (baker/IndirectLight.cpp:47-75).
//
// Photon stores land in fixed record slots (photon-major, bounce-minor), are grouped by texel with
// a stable radix sort and summed in record order, so the photonmap equals a sequential store in
// the reference's emission order, bit for bit, on every run.
#include <hip/hip_runtime.h>

#include "rl_device.h"
#include "rl_kernels.h"
#include "rl_sort.h"
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
    if (S.mat) {
        // Rgb(hit.m_color) of Photons::trace (baker/Photons.cpp:119) for textured materials.
        const float4 c = face_color(S, tr.tri, u, v);
        h.albedo[0] = c.x; h.albedo[1] = c.y; h.albedo[2] = c.z; h.albedo[3] = c.w;
    }
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
                if (fabsf(1.0f - L.spot_exponent) > 0.01f) value = ref_powf(value, L.spot_exponent);
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

// Sort keys of the photon records: the texel key, or `empty` (the number of texel keys, above
// every texel key) for a slot with nothing stored, so the keys need only bit_length(empty) bits.
__global__ void k_keys_iota(const PhotonRec* __restrict__ rec, uint32_t* __restrict__ keys,
                            uint32_t* __restrict__ idx, uint32_t n, uint32_t empty) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) {
        const uint32_t k = rec[g].key;
        keys[g] = k == 0xFFFFFFFFu ? empty : k;
        idx[g] = g;
    }
}

// Photons::store (baker/Photons.cpp:129-144) over records grouped by texel: one thread per run
// head, summing in record order on top of the texel's current value.
__global__ void k_photon_accumulate(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ idx,
                                    const PhotonRec* __restrict__ rec, uint32_t n,
                                    const PhotonMapDev* __restrict__ pms, int npm, uint32_t empty) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const uint32_t key = keys[g];
    if (key >= empty) return;
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

// Final-gather workgroups: sixteen wavefronts, one workgroup per compute unit (the occupancy the
// registers allow), keeping the first nodes of the first segment's scene in LDS (kFgTopNodes for
// the lane-per-lumel kernel: 88 KB, with the stacks 152 KB). Heightfield final gather: 289 ms with
// four-wavefront workgroups and 184 nodes, 282 ms with sixteen and 700; later, against 254 ms for
// sixteen wavefronts and 700 nodes, twelve (768 threads, 860 nodes) measured 294.7 ms and eight (512
// threads, 990 nodes) 383.2 ms: more cached nodes do not make up for fewer wavefronts.
constexpr int kFgBlock = 1024;
constexpr uint32_t kFgTopNodes = 700;
// A closest-hit round's node phase stops once no more than this many lanes are still searching
// without a leaf (heightfield final gather: 338 ms with no early stop, 327/328/328/330/345 ms at
// 2/3/4/6/16 lanes, 8 to 32 idle lanes before a refill within 1 ms of each other at 4).
constexpr uint32_t kFgNodeExit = 4;
// Likewise its leaf phase, once no more than this many lanes still have triangles to test (the
// rest of their leaf carries over). Heightfield final gather: 299 ms with a per-lane leaf loop;
// this loop 307/290/289/288/292/293 ms at 0/8/12/16/20/24 lanes; at 16, a node-phase stop at 2
// or 8 lanes instead of 4: 288/293 ms.
constexpr uint32_t kFgLeafExit = 16;

// IndirectLight::bakeLumel (baker/IndirectLight.cpp:47-75) with lane refill. A wavefront takes
// lumels from the persistent queue and hands their (lumel, sample) pairs, in order, to idle lanes
// whenever at least `refill_min` lanes are idle; pair number j of the wavefront's stream owns slot
// j mod kFgRing of the wavefront's LDS ring. A lane traces its sample in rounds (closest_run), so
// finished lanes take new samples while others still traverse; when its ray is done it shades the
// sample and writes the sample's term into its slot. The wavefront adds finished terms to the
// current lumel's sum strictly in stream order, that is in sample order as the reference's loop
// does, and writes the lumel when its last sample has been added. A lane may only start pair j <
// consumed + kFgRing.
constexpr uint32_t kFgRing = 128;
// Workgroups of four wavefronts, four to a compute unit (registers allow four wavefronts per
// SIMD): passes of a few thousand lumels, a lumel per wavefront, still spread over every compute
// unit. LDS per workgroup: 16 KB stacks + 10 KB rings + 13 KB of the scene's first 104 nodes.
constexpr int kFgRefillBlock = 256;
constexpr uint32_t kFgRefillTopNodes = 104;

template <int CAP>
__global__ __launch_bounds__(kFgRefillBlock) void k_fg_refill(const BakeSeg* __restrict__ segs, int nsegs,
                                                              uint64_t total, uint32_t refill_min,
                                                              unsigned long long* queue, unsigned long long* rays_out) {
    __shared__ uint32_t stack[kLdsStack * kFgRefillBlock];
    __shared__ u32x4 top_rows[kFgRefillTopNodes * kNodeRows];
    __shared__ float4 ring[kFgRefillBlock / 64][kFgRing];          // x, y, z: term; w: flags (0 = pending)
    __shared__ uint32_t ring_lumel[kFgRefillBlock / 64][kFgRing];  // queue index of the slot's lumel
    const SceneDev top_scene = segs[0].scene;
    const uint32_t top_count = load_top_nodes(top_scene, top_rows, kFgRefillTopNodes);
    uint32_t* stk = stack + threadIdx.x;
    uint32_t spill[CAP - kLdsStack];
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    for (uint32_t k = lane; k < kFgRing; k += 64) ring[wave][k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    __syncthreads();
    constexpr uint32_t kUsed = 1u, kUnused = 2u, kLast = 4u;

    // Wavefront state, identical in every lane.
    uint64_t take_g = 0;                 // lumel being handed out (queue index)
    int take_seg = 0;
    uint32_t take_s = 0, take_n = 0;     // its next sample and its sample count
    uint32_t j_take = 0, j_cons = 0;     // stream pairs handed out / added
    bool pool_open = true;
    float ar = 0.0f, ag = 0.0f, ab = 0.0f;  // sum of the lumel being added
    // Lane state.
    bool active = false;
    TravRay tr;
    SceneDev S;
    TopCache top{top_rows, 0u};
    TraceResult res;
    uint32_t node = 0, pend = 0, my_j = 0, my_seg = 0, my_last = 0;
    int sp = 0;
    V3 dir = v3(0.0f, 0.0f, 0.0f);
    float influence = 0.0f;
    uint32_t rays = 0;
    for (;;) {
        const uint64_t idle = __ballot(!active);
        if (pool_open && ((uint32_t)__popcll(idle) >= refill_min || idle == ~0ull)) {
            uint64_t want = idle;
            while (want && pool_open) {
                if (take_s == take_n) {
                    const uint64_t c = wave_next_chunk(queue);
                    if (c >= total) {
                        pool_open = false;
                        break;
                    }
                    take_g = c;
                    take_seg = find_seg(segs, nsegs, c);
                    take_s = 0;
                    take_n = (uint32_t)segs[take_seg].fg.samples;
                }
                const uint32_t room = j_cons + kFgRing - j_take;
                const uint32_t take = min(min(room, take_n - take_s), (uint32_t)__popcll(want));
                if (take == 0) break;
                const uint32_t rank = rank_below(want);
                const bool mine = ((want >> lane) & 1ull) && rank < take;
                if (mine) {
                    const BakeSeg& sg = segs[take_seg];
                    const FinalGatherParams& P = sg.fg;
                    const uint32_t i = (uint32_t)(take_g - sg.work_begin);
                    const uint32_t s = take_s + rank;
                    my_j = j_take + rank;
                    my_seg = (uint32_t)take_seg;
                    my_last = s == take_n - 1u ? kLast : 0u;
                    ring_lumel[wave][my_j % kFgRing] = (uint32_t)take_g;
                    const float4 p4 = sg.pos[i], n4 = sg.nrm[i];
                    const V3 p = v3(p4.x, p4.y, p4.z), n = v3(n4.x, n4.y, n4.z);
                    Pcg32 rng = sample_stream(P.seed, sg.lm_seed, f_bits(p4.w), TECH_FINAL_GATHER, s);
                    dir = sample_hemisphere(rng, n, P.sampling);
                    const float d0 = dot(n, dir);
                    influence = d0 > 0.0f ? d0 : 0.0f;
                    const Ray ray = ray_from_segment(p, p + dir * P.distance);
                    tr = trav_ray(ray);
                    S = sg.scene;
                    top.count = S.nodes == top_scene.nodes ? top_count : 0u;
                    node = 0;
                    pend = 0;
                    sp = 0;
                    res.t = ray.tfar;
                    res.tri = 0xFFFFFFFFu;
                    res.U = res.V = res.W = res.det = 0.0f;
                    active = true;
                }
                take_s += take;
                j_take += take;
                want &= ~__ballot(mine);
            }
        }
        if (active && (one_scene(S) ? closest_run<CAP, true>(S, top, tr, node, sp, stk, kFgRefillBlock, spill, res,
                                                             pend, kFgNodeExit, kFgLeafExit)
                                    : closest_run<CAP, false>(S, top, tr, node, sp, stk, kFgRefillBlock, spill,
                                                              res, pend, kFgNodeExit, kFgLeafExit)) != 0) {
            rays++;
            const BakeSeg& sg = segs[my_seg];
            const FinalGatherParams& P = sg.fg;
            float xr = 0.0f, xg = 0.0f, xb = 0.0f;
            uint32_t f = kUnused;
            if (res.tri == 0xFFFFFFFFu) {
                xr = P.sky[0] * influence + P.ambient[0];
                xg = P.sky[1] * influence + P.ambient[1];
                xb = P.sky[2] * influence + P.ambient[2];
                f = kUsed;
            } else {
                const HitAttr h = hit_attributes(S, res);
                if (dot(dir, h.n) < 0.0f) {
                    const int32_t pm = sg.mesh_pm[h.mesh];
                    if (pm >= 0) {
                        const PhotonMapDev M = sg.pms[pm];
                        const int x = lumel_coord(h.uvx, M.w), y = lumel_coord(h.uvy, M.h);
                        const float4 gc = M.gathered[(uint32_t)y * (uint32_t)M.w + (uint32_t)x];
                        xr = gc.x * influence + P.ambient[0];
                        xg = gc.y * influence + P.ambient[1];
                        xb = gc.z * influence + P.ambient[2];
                        f = kUsed;
                    }
                }
            }
            ring[wave][my_j % kFgRing] = make_float4(xr, xg, xb, __uint_as_float(f | my_last));
            active = false;
        }
        // Add finished terms in stream order; a slot read here is the same LDS word in every lane.
        while (j_cons != j_take) {
            const uint32_t slot = j_cons % kFgRing;
            const float4 e = ring[wave][slot];
            const uint32_t f = (uint32_t)wave_uniform(__float_as_uint(e.w));
            if (f == 0u) break;
            if (f & kUsed) {
                ar += e.x;
                ag += e.y;
                ab += e.z;
            }
            if (f & kLast) {
                const uint32_t g = ring_lumel[wave][slot];
                if (lane == 0) {
                    const BakeSeg& sg = segs[find_seg(segs, nsegs, g)];
                    const uint32_t i = (uint32_t)(g - sg.work_begin);
                    const float fs = (float)sg.fg.samples;
                    float4 c = sg.color[i];
                    c.x += ar / fs; c.y += ag / fs; c.z += ab / fs;
                    sg.color[i] = c;
                }
                ar = ag = ab = 0.0f;
            }
            if (lane == 0) ring[wave][slot].w = 0.0f;
            j_cons++;
        }
        if (!pool_open && __ballot(active) == 0ull && j_cons == j_take) break;
    }
    count_rays(rays, rays_out);
}

// IndirectLight::bakeLumel (baker/IndirectLight.cpp:47-75) with one lumel per lane. Each lane takes
// a lumel from the persistent queue and traces its samples one after another, adding each sample's
// term to the lumel's sum as it finishes, so the sum runs in sample order as the reference's loop
// does, with no exchange between lanes. A lane traces in rounds (closest_run); lanes whose ray is
// done wait until at least `refill_min` lanes of the wavefront are without a ray, then start their
// next samples together (lanes whose lumel is done first take new lumels), so the ray set-up runs
// for many lanes at once. LDS per workgroup: 64 KB stacks + 88 KB of the scene's first 700 nodes.
// Lumels per resident lane from which the lane-per-lumel kernel runs. Heightfield maps at 64 samples
// (wave / lane kernel): 65 K lumels 1.14 / 3.13 ms, 147 K 2.29 / 3.22 ms, 262 K 3.71 / 3.25 ms, 590 K
// 7.91 / 6.72 ms, 1 M 13.87 / 10.99 ms; at 256 samples, 262 K: 12.30 / 12.57 ms.
constexpr uint64_t kFgLanesMinPerLane = 1;

template <int CAP>
__global__ __launch_bounds__(kFgBlock) void k_fg_lanes(const BakeSeg* __restrict__ segs, int nsegs, uint64_t total,
                                                       uint32_t refill_min, unsigned long long* queue,
                                                       unsigned long long* rays_out) {
    __shared__ uint32_t stack[kLdsStack * kFgBlock];
    __shared__ u32x4 top_rows[kFgTopNodes * kNodeRows];
    const SceneDev top_scene = segs[0].scene;
    const uint32_t top_count = load_top_nodes(top_scene, top_rows, kFgTopNodes);
    __syncthreads();
    uint32_t* stk = stack + threadIdx.x;
    uint32_t spill[CAP - kLdsStack];
    const uint32_t lane = lane_id();
    bool pool_open = true;  // identical in every lane
    // Lane state: the lumel (queue index g, its segment, next sample, sample count, sums) and the
    // sample's ray in flight.
    bool has_lumel = false, ray_on = false;
    uint32_t g = 0, seg = 0, s = 0, nsamples = 0;
    float ar = 0.0f, ag = 0.0f, ab = 0.0f;
    // The lumel's position, sampling frame, texel key and settings, read and computed once for
    // all its samples.
    V3 lp = v3(0.0f, 0.0f, 0.0f);
    CosineFrame frame{};
    uint64_t key = 0;
    float fg_distance = 0.0f;
    int32_t fg_sampling = 0;
    TravRay tr;
    SceneDev S;
    TopCache top{top_rows, 0u};
    TraceResult res;
    uint32_t node = 0, pend = 0;
    int sp = 0;
    V3 dir = v3(0.0f, 0.0f, 0.0f);
    float influence = 0.0f;
    uint32_t rays = 0;
    for (;;) {
        const uint64_t idle = __ballot(!ray_on);
        if ((uint32_t)__popcll(idle) >= refill_min || idle == ~0ull) {
            const uint64_t fresh = __ballot(!ray_on && !has_lumel);
            if (fresh && pool_open) {
                const uint32_t n = (uint32_t)__popcll(fresh);
                const uint64_t base = wave_take(queue, n);
                if (base + n >= total) pool_open = false;
                if (!ray_on && !has_lumel) {
                    const uint64_t mine = base + rank_below(fresh);
                    if (mine < total) {
                        has_lumel = true;
                        g = (uint32_t)mine;
                        seg = (uint32_t)find_seg(segs, nsegs, mine);
                        const BakeSeg& sg = segs[seg];
                        s = 0;
                        nsamples = (uint32_t)sg.fg.samples;
                        ar = ag = ab = 0.0f;
                        const uint32_t i = g - (uint32_t)sg.work_begin;
                        const float4 p4 = sg.pos[i], n4 = sg.nrm[i];
                        lp = v3(p4.x, p4.y, p4.z);
                        frame = cosine_frame(v3(n4.x, n4.y, n4.z));
                        key = texel_key(sg.fg.seed, sg.lm_seed, f_bits(p4.w));
                        fg_distance = sg.fg.distance;
                        fg_sampling = sg.fg.sampling;
                        S = sg.scene;
                        top.count = S.nodes == top_scene.nodes ? top_count : 0u;
                    }
                }
            }
            if (!ray_on && has_lumel) {
                Pcg32 rng = texel_stream(key, TECH_FINAL_GATHER, s);
                dir = fg_sampling == SAMPLE_COSINE ? sample_cosine_in_frame(rng, frame)
                                                   : sample_uniform_hemisphere(rng, frame.n);
                const float d0 = dot(frame.n, dir);
                influence = d0 > 0.0f ? d0 : 0.0f;
                const Ray ray = ray_from_segment(lp, lp + dir * fg_distance);
                tr = trav_ray(ray);
                node = 0;
                pend = 0;
                sp = 0;
                res.t = ray.tfar;
                res.tri = 0xFFFFFFFFu;
                res.U = res.V = res.W = res.det = 0.0f;
                ray_on = true;
            }
        }
        if (ray_on && (one_scene(S) ? closest_run<CAP, true>(S, top, tr, node, sp, stk, kFgBlock, spill, res, pend,
                                                             kFgNodeExit, kFgLeafExit)
                                    : closest_run<CAP, false>(S, top, tr, node, sp, stk, kFgBlock, spill, res, pend,
                                                              kFgNodeExit, kFgLeafExit)) != 0) {
            rays++;
            ray_on = false;
            const BakeSeg& sg = segs[seg];
            const FinalGatherParams& P = sg.fg;
            if (res.tri == 0xFFFFFFFFu) {
                ar += P.sky[0] * influence + P.ambient[0];
                ag += P.sky[1] * influence + P.ambient[1];
                ab += P.sky[2] * influence + P.ambient[2];
            } else {
                const HitAttr h = hit_attributes(S, res);
                if (dot(dir, h.n) < 0.0f) {
                    const int32_t pm = sg.mesh_pm[h.mesh];
                    if (pm >= 0) {
                        const PhotonMapDev M = sg.pms[pm];
                        const int x = lumel_coord(h.uvx, M.w), y = lumel_coord(h.uvy, M.h);
                        const float4 gc = M.gathered[(uint32_t)y * (uint32_t)M.w + (uint32_t)x];
                        ar += gc.x * influence + P.ambient[0];
                        ag += gc.y * influence + P.ambient[1];
                        ab += gc.z * influence + P.ambient[2];
                    }
                }
            }
            if (++s == nsamples) {
                const uint32_t i = g - (uint32_t)sg.work_begin;
                const float fs = (float)nsamples;
                float4 c = sg.color[i];
                c.x += ar / fs; c.y += ag / fs; c.z += ab / fs;
                sg.color[i] = c;
                has_lumel = false;
            }
        }
        if (!pool_open && __ballot(ray_on || has_lumel) == 0ull) break;
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

hipError_t launch_photon_store(hipStream_t st, const PhotonRec* rec, uint32_t n, uint32_t* keys_a,
                               uint32_t* keys_b, uint32_t* idx_a, uint32_t* idx_b, void* temp,
                               size_t temp_bytes, const PhotonMapDev* pms, int npm, uint32_t total_keys) {
    if (n == 0) return hipSuccess;
    if (total_keys == 0xFFFFFFFFu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_keys_iota, dim3(blocks_for(n, 256)), dim3(256), 0, st, rec, keys_a, idx_a, n, total_keys);
    int bits = 1;
    while (bits < 32 && (total_keys >> bits) != 0u) bits++;
    uint32_t *keys, *idx;
    hipError_t e = launch_sort_pairs(st, keys_a, keys_b, idx_a, idx_b, n, bits, temp, temp_bytes, &keys, &idx);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_photon_accumulate, dim3(blocks_for(n, 256)), dim3(256), 0, st, keys, idx, rec, n,
                       pms, npm, total_keys);
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

hipError_t launch_knn_build(hipStream_t st, const PhotonRec* rec, const float4* ppos, uint32_t n,
                            const KnnGrid& G, int key_bits, uint32_t* keys_a, uint32_t* keys_b, uint32_t* idx_a,
                            uint32_t* idx_b, void* temp, size_t temp_bytes, float4* spos, uint32_t* cstart,
                            uint32_t* cend) {
    if (n == 0) return hipSuccess;
    hipError_t e = launch_fill_u32(st, cstart, (uint64_t)G.mask + 1u, 0u);
    if (e == hipSuccess) e = launch_fill_u32(st, cend, (uint64_t)G.mask + 1u, 0u);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_knn_keys, dim3(blocks_for(n, 256)), dim3(256), 0, st, rec, ppos, n, G, keys_a, idx_a);
    uint32_t *keys, *idx;
    e = launch_sort_pairs(st, keys_a, keys_b, idx_a, idx_b, n, key_bits, temp, temp_bytes, &keys, &idx);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_knn_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, st, keys, idx, n, ppos, G, spos,
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
                               uint64_t total_lumels, int kernel, unsigned long long* queue, unsigned long long* rays) {
    if (total_lumels == 0) return hipSuccess;
    // Idle lanes a wavefront waits for before starting more samples: the heightfield pass measured
    // 370 ms at 8 and 16, 388 ms at 32 (lane kernel); 407-411 ms from 4 to 48 (ring kernel); after
    // the early node-phase stop 329/328/329/330 ms at 8/16/24/32 and after the per-lumel set-up
    // 298/298/299 ms at 8/16/32.
    const uint32_t refill_min = 16;
    hipError_t e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(queue), 2, 0u);
    if (e != hipSuccess) return e;
    // max_blocks counts kBlock-thread workgroups; these kernels' are kFgBlock and kFgRefillBlock threads.
    const uint64_t cap_blocks = std::max<uint64_t>(1, (uint64_t)max_blocks * kBlock / kFgBlock);
    // One lumel per lane needs many lumels per resident lane to keep lanes busy to the end; smaller
    // passes give each lumel a wavefront.
    const bool lanes = kernel == kFgKernelLane ||
                       (kernel == kFgKernelAuto && total_lumels >= kFgLanesMinPerLane * cap_blocks * kFgBlock);
    if (lanes) {
        const uint32_t grid = (uint32_t)std::min<uint64_t>(blocks_for(total_lumels, kFgBlock), cap_blocks);
        RL_DISPATCH_CAP(cap, hipLaunchKernelGGL(k_fg_lanes<CAP>, dim3(grid), dim3(kFgBlock), 0, st, segs, nsegs,
                                                total_lumels, refill_min, queue, rays));
    } else {
        const uint64_t refill_cap = std::max<uint64_t>(1, (uint64_t)max_blocks * kBlock / kFgRefillBlock);
        const uint32_t grid = (uint32_t)std::min<uint64_t>(blocks_for(total_lumels, kFgRefillBlock / 64), refill_cap);
        RL_DISPATCH_CAP(cap, hipLaunchKernelGGL(k_fg_refill<CAP>, dim3(grid), dim3(kFgRefillBlock), 0, st, segs, nsegs,
                                                total_lumels, refill_min, queue, rays));
    }
    return hipGetLastError();
}

}  // namespace rl
