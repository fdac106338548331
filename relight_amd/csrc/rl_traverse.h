// BVH4 traversal shared by every kernel that casts rays (device code only).
//
// Near-first traversal with a per-lane LDS stack (CAP entries, one column per lane), a
// conservative fma slab test and the watertight triangle test of rl_math.h. Each axis's entry and
// exit planes are picked by the sign of the ray direction when the node is loaded (a per-lane
// load offset), so the slab test needs no per-box min/max of the two plane distances.
// Closest-hit ties at equal t go to the smaller scene-order triangle id, so results do not depend
// on traversal order.
#pragma once

#include <hip/hip_runtime.h>

#include "rl_device.h"

namespace rl {
namespace trav {

// Slab-test safety factor: scales the exit distances so that the rounded slab distances never
// reject a box the exact ray passes through (1 + 5 * 2^-23 once rounded; the bound is derived
// at SlabRay).
constexpr float kBoxScale = 1.0000006f;

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// A 64-bit value equal in every active lane, moved to scalar registers so the compiler treats
// branches on it as wave-uniform.
__device__ __forceinline__ uint64_t wave_uniform(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// The same for a 32-bit value (a cross-lane reduction's result, for instance, which the compiler
// cannot tell is uniform: values computed from it would otherwise live in vector registers).
__device__ __forceinline__ uint32_t wave_uniform32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }


// Global-memory views of the scene arrays. A pointer read from a struct is generic to the
// compiler, and a generic (flat) load also counts against the LDS wait counter, so every wait for
// a node or triangle fetch would also wait for the stack's LDS traffic.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef const f32x4 __attribute__((address_space(1))) gfloat4;
typedef const u32x4 __attribute__((address_space(1))) guint4;
typedef const float __attribute__((address_space(1))) gfloat;
__device__ __forceinline__ const gfloat4* gmem(const float4* p) { return (const gfloat4*)p; }

// Triangles of the any-hit traversals (SceneDev::tris_rot), in leaf order, 64 B each: vertex k as
// (x, y, z, x, y) at floats 5k .. 5k + 4, so its components in a ray's watertight order (wt_rotate:
// a rotation by WtRay::rot) are the three floats from 5k + rot, one 12-B load with no selects.
// The closest-hit traversals read the 48-B triangles (SceneDev::tris: v0, v1, v2 as float4, the
// scene-order id in v0.w) and rotate the components with selects: the larger triangles measured
// 4 % slower in the final gather, which waits on triangle and node fetches, while the AO kernel,
// which issues VALU at the ceiling, gained 3 % from the 24 VALU fewer per triangle test.
constexpr uint32_t kRotTriFloats = 16;

__device__ __forceinline__ const gfloat* rtri_ptr(const float4* tris_rot, uint32_t tri) {
    return reinterpret_cast<const gfloat*>(gmem(tris_rot)) + kRotTriFloats * tri;
}

__device__ __forceinline__ V3 rtri_vertex(const gfloat* t, uint32_t k, uint32_t rot) {
    const gfloat* p = t + 5u * k + rot;
    return v3(p[0], p[1], p[2]);
}

__device__ __forceinline__ uint32_t rtri_id(const gfloat* t) { return __float_as_uint(t[15]); }

__device__ __forceinline__ const gfloat* tri_ptr(const float4* tris, uint32_t tri) {
    return reinterpret_cast<const gfloat*>(gmem(tris) + 3u * tri);
}

// The scene's array addresses taken from the first active lane into scalar registers, for callers
// whose active lanes all trace the same scene (one_scene): loads then use a scalar base and a
// 32-bit per-lane offset instead of per-lane 64-bit address arithmetic.
__device__ __forceinline__ SceneDev uniform_scene(const SceneDev& S) {
    SceneDev U = S;
    U.nodes = reinterpret_cast<const Bvh4Node*>(wave_uniform(reinterpret_cast<uint64_t>(S.nodes)));
    U.tris = reinterpret_cast<const float4*>(wave_uniform(reinterpret_cast<uint64_t>(S.tris)));
    U.tris_rot = reinterpret_cast<const float4*>(wave_uniform(reinterpret_cast<uint64_t>(S.tris_rot)));
    U.attr = reinterpret_cast<const TriAttr*>(wave_uniform(reinterpret_cast<uint64_t>(S.attr)));
    return U;
}

// Whether every active lane holds the scene of the first active lane.
__device__ __forceinline__ bool one_scene(const SceneDev& S) {
    const uint64_t lead = wave_uniform(reinterpret_cast<uint64_t>(S.nodes));
    return __ballot(reinterpret_cast<uint64_t>(S.nodes) != lead) == 0ull;
}

__device__ __forceinline__ float bits_f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t f_bits(float f) { return __float_as_uint(f); }

struct TraceResult {
    float t;
    uint32_t tri;  // scene-order id, 0xFFFFFFFF on a miss
    float U, V, W, det;
    uint32_t leaf;  // any-hit traversal: the hit triangle's leaf-order index
};

__device__ __forceinline__ void cswap(float& ka, uint32_t& va, float& kb, uint32_t& vb) {
    if (kb < ka) {
        float tk = ka; ka = kb; kb = tk;
        uint32_t tv = va; va = vb; vb = tv;
    }
}

// Entries of a lane's traversal stack kept in LDS; deeper entries (rare) spill to a private
// per-lane array. Blocks allocate kLdsStack words of LDS per thread. 12 entries (the final gather
// caching 828 top nodes instead of 700) measured city AO 206.5 ms and heightfield final gather
// 252.1 ms against 205.0 and 254.2 ms at 16; 8 entries (956 nodes) 210.9 and 254.7 ms.
constexpr int kLdsStack = 16;

// Number of set bits of `mask` below this lane (a wavefront-local prefix count).
__device__ __forceinline__ uint32_t rank_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Per-lane constants of the slab test. Entry planes: t = fma(b, i, oi) with i = 1/d and
// oi = -o/d per axis; exit planes: t = fma(b, si, soi) with the same two constants multiplied by
// kBoxScale. re*/rf*: byte offsets in a node of each axis's entry-plane row (lo for a positive
// direction, hi for a negative one) and exit-plane row.
//
// Rounding: i is the hardware reciprocal (within 1 ulp, a relative 2^-23, of 1/d). An entry
// distance is the exact distance to its plane shifted by at most |o| * 2^-24, times a factor
// within 3 * 2^-24 of 1; an exit distance is the exact distance to its plane shifted by at most
// 3 |o| * 2^-24, times kBoxScale * (1 +- 4 * 2^-24). kBoxScale (1 + 10 * 2^-24) covers the seven
// relative roundings, and the builder's box padding (2^-18 of the largest coordinate magnitude)
// covers the shifts for any origin within 20 times that magnitude, so no box the exact ray
// crosses within (tnear, lim) is rejected.
struct SlabRay {
    float ix, iy, iz, oix, oiy, oiz;
    float six, siy, siz, soix, soiy, soiz;
    uint32_t rex, rfx, rey, rfy, rez, rfz;
};

__device__ __forceinline__ SlabRay slab_ray(const Ray& ray) {
    SlabRay r;
    // A zero direction component would give fma(b, inf, -inf) = NaN; the slab test uses
    // +-1e-20 there instead (a displacement below 1e-17 m over any ray, inside the box padding).
    const float dx = fabsf(ray.d.x) > 1e-20f ? ray.d.x : copysignf(1e-20f, ray.d.x);
    const float dy = fabsf(ray.d.y) > 1e-20f ? ray.d.y : copysignf(1e-20f, ray.d.y);
    const float dz = fabsf(ray.d.z) > 1e-20f ? ray.d.z : copysignf(1e-20f, ray.d.z);
    r.ix = __builtin_amdgcn_rcpf(dx); r.iy = __builtin_amdgcn_rcpf(dy); r.iz = __builtin_amdgcn_rcpf(dz);
    r.oix = -(ray.o.x * r.ix); r.oiy = -(ray.o.y * r.iy); r.oiz = -(ray.o.z * r.iz);
    r.six = r.ix * kBoxScale; r.siy = r.iy * kBoxScale; r.siz = r.iz * kBoxScale;
    r.soix = r.oix * kBoxScale; r.soiy = r.oiy * kBoxScale; r.soiz = r.oiz * kBoxScale;
    // Node rows: lox, hix, loy, hiy, loz, hiz (16 B each).
    r.rex = r.ix < 0.0f ? 16u : 0u;
    r.rfx = 16u - r.rex;
    r.rey = r.iy < 0.0f ? 48u : 32u;
    r.rfy = 80u - r.rey;
    r.rez = r.iz < 0.0f ? 80u : 64u;
    r.rfz = 144u - r.rez;
    return r;
}

// Per-lane constants of a ray for the step-wise traversal below.
struct TravRay {
    WtRay w;
    SlabRay s;
};

__device__ __forceinline__ TravRay trav_ray(const Ray& ray) {
    TravRay r;
    r.w = wt_setup(ray);
    r.s = slab_ray(ray);
    return r;
}

// Four child boxes of a node against a ray, given the node's entry-plane (e*) and exit-plane (f*)
// rows of each axis: the entry distance of each box (its sort key), +inf for a box the ray misses
// within (tnear, lim). The entry plane of an axis is lo for a positive direction component and hi
// for a negative one, so the entry distance is the min of the two plane distances (fma rounding is
// monotone in the plane) and the exit distance the max.
__device__ __forceinline__ void box4(f32x4 ex, f32x4 fx, f32x4 ey, f32x4 fy, f32x4 ez, f32x4 fz, const SlabRay& r,
                                     float tnear, float lim, float* key) {
    const float exa[4] = {ex.x, ex.y, ex.z, ex.w}, fxa[4] = {fx.x, fx.y, fx.z, fx.w};
    const float eya[4] = {ey.x, ey.y, ey.z, ey.w}, fya[4] = {fy.x, fy.y, fy.z, fy.w};
    const float eza[4] = {ez.x, ez.y, ez.z, ez.w}, fza[4] = {fz.x, fz.y, fz.z, fz.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float t0x = fmaf(exa[i], r.ix, r.oix), t1x = fmaf(fxa[i], r.six, r.soix);
        const float t0y = fmaf(eya[i], r.iy, r.oiy), t1y = fmaf(fya[i], r.siy, r.soiy);
        const float t0z = fmaf(eza[i], r.iz, r.oiz), t1z = fmaf(fza[i], r.siz, r.soiz);
        const float tmin = fmaxf(fmaxf(t0x, t0y), fmaxf(t0z, tnear));
        const float tmax = fminf(fminf(t1x, t1y), t1z);
        const bool hit = tmin <= fminf(tmax, lim);  // empty slots hold inverted boxes
        key[i] = hit ? tmin : __builtin_inff();
    }
}

__device__ __forceinline__ void stack_push(uint32_t v, int& sp, uint32_t* stk, uint32_t stride, uint32_t* spill) {
    if (sp < kLdsStack) stk[sp * stride] = v;
    else spill[sp - kLdsStack] = v;
    sp++;
}

__device__ __forceinline__ uint32_t stack_pop(int& sp, const uint32_t* stk, uint32_t stride, const uint32_t* spill) {
    --sp;
    if (sp < kLdsStack) return stk[sp * stride];
    // volatile keeps the compiler from merging the two loads into one generic load of a selected
    // address (which would make every pop a flat load)
    return *(volatile const uint32_t*)&spill[sp - kLdsStack];
}

// The first nodes of a scene in breadth-first order (the levels nearest the root, which most rays
// visit) copied into LDS by a workgroup, each node's 128 bytes as they are in memory, so a node's
// rows lie at the same byte offsets in LDS as in the scene's node array and one offset computation
// serves both. A node visit reads a cached node from LDS instead of through the vector-memory path.
constexpr uint32_t kNodeRows = 8;

struct TopCache {
    const u32x4* rows;  // LDS, kNodeRows rows per node
    uint32_t count;     // nodes held for this lane's scene (0: none)
};

// Virtual nodes (nodes built in LDS for a box of rays, see build_local_tree) sit in the same LDS
// array as the top nodes, after its top-node capacity C: the node in slot C + k has the node index
// kVirtualBase + C + k. Real node indices stay below kVirtualBase = 2^25, so the node's byte offset
// (index << 7, in 32 bits) is (C + k) * 128, its offset in that array, and one test
// (index < count or index >= 2^25) tells LDS nodes from nodes in memory.
constexpr uint32_t kVirtualBase = 1u << 25;

// Copies the first min(max_nodes, S.num_nodes) nodes of S into `rows` (LDS) with the whole
// workgroup and returns how many it holds; the caller synchronises the workgroup before use.
__device__ __forceinline__ uint32_t load_top_nodes(const SceneDev& S, u32x4* rows, uint32_t max_nodes) {
    const uint32_t n = S.num_nodes < max_nodes ? S.num_nodes : max_nodes;
    const guint4* src = reinterpret_cast<const guint4*>(gmem(reinterpret_cast<const float4*>(S.nodes)));
    for (uint32_t i = threadIdx.x; i < n * kNodeRows; i += blockDim.x) rows[i] = src[i];
    return n;
}

// ---------------------------------------------------------------- local trees
// Rays known to stay inside a box q (the AO rays of a chunk of lumels: segments of length D from
// each lumel into its hemisphere) can only hit triangles in BVH subtrees whose boxes overlap q: a
// hit point lies in q and in every box on the path from the root to the hit triangle's leaf. A
// local tree replaces the top of the BVH for such rays: a virtual root whose entries are a
// frontier of subtrees overlapping q (copies of their child boxes and words), and up to four
// second-level virtual nodes that refine the root's internal entries the same way. Traversal from
// the virtual root tests the same child boxes below the frontier as traversal from the real root,
// and skips only subtrees that cannot hold a hit, so any-hit results are unchanged; it skips the
// several node visits the real root's levels cost every ray (city AO: 14.1 node visits per ray
// from the root, 7.7 plus 2.3 virtual ones from a 16-lumel chunk's local tree).

struct QueryBox {
    float lo[3], hi[3];
};

__device__ __forceinline__ bool box_overlaps(const QueryBox& q, float lx, float hx, float ly, float hy, float lz,
                                             float hz) {
    return !(hx < q.lo[0] || lx > q.hi[0] || hy < q.lo[1] || ly > q.hi[1] || hz < q.lo[2] || lz > q.hi[2]);
}

// A virtual node in LDS: the node layout (rows lox, hix, loy, hiy, loz, hiz, child, pad) written
// entry by entry.
__device__ __forceinline__ void vt_set(u32x4* v, uint32_t j, const float* lo, const float* hi, uint32_t child) {
    float* f = reinterpret_cast<float*>(v);
    f[0 * 4 + j] = lo[0]; f[1 * 4 + j] = hi[0];
    f[2 * 4 + j] = lo[1]; f[3 * 4 + j] = hi[1];
    f[4 * 4 + j] = lo[2]; f[5 * 4 + j] = hi[2];
    reinterpret_cast<uint32_t*>(v)[6 * 4 + j] = child;
}

__device__ __forceinline__ void vt_get(const u32x4* v, uint32_t j, float* lo, float* hi, uint32_t& child) {
    const float* f = reinterpret_cast<const float*>(v);
    lo[0] = f[0 * 4 + j]; hi[0] = f[1 * 4 + j];
    lo[1] = f[2 * 4 + j]; hi[1] = f[3 * 4 + j];
    lo[2] = f[4 * 4 + j]; hi[2] = f[5 * 4 + j];
    child = reinterpret_cast<const uint32_t*>(v)[6 * 4 + j];
}

__device__ __forceinline__ void vt_clear(u32x4* v, uint32_t j) {
    const float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    vt_set(v, j, lo, hi, RL_EMPTY_CHILD);  // an inverted box: no ray enters it
}

// A real node's six box rows (lox, hix, loy, hiy, loz, hiz) and child words.
__device__ __forceinline__ void read_node(const SceneDev& S, const TopCache& top, uint32_t c, f32x4* rows, u32x4& ch) {
    if (c < top.count) {
        const u32x4* b = top.rows + c * kNodeRows;
#pragma unroll
        for (int k = 0; k < 6; k++) rows[k] = __builtin_bit_cast(f32x4, b[k]);
        ch = b[6];
    } else {
        const guint4* b = reinterpret_cast<const guint4*>(gmem(reinterpret_cast<const float4*>(S.nodes))) +
                          (size_t)c * kNodeRows;
#pragma unroll
        for (int k = 0; k < 6; k++) rows[k] = __builtin_bit_cast(f32x4, b[k]);
        ch = b[6];
    }
}

// Builds virtual node v (cleared beforehand) for box q from the children of real node `start`
// that overlap q, then repeatedly replaces an internal entry by its children overlapping q while
// the entries still fit in four (in entry order, re-examining an entry after it is replaced; an
// entry none of whose children overlaps q is dropped). At most max_loads nodes are read; every
// step keeps the entries a cover of the subtrees under `start` that overlap q, so stopping early
// is still correct. Returns the number of entries. Run by one lane per node.
__device__ __forceinline__ uint32_t build_local_node(const SceneDev& S, const TopCache& top, uint32_t start,
                                                     const QueryBox& q, u32x4* v, uint32_t max_loads) {
    uint32_t n = 0, loads = 0;
    for (uint32_t j = 0; loads < max_loads;) {
        uint32_t c;
        if (loads == 0) {
            c = start;
        } else {
            if (j >= n) break;
            float lo[3], hi[3];
            vt_get(v, j, lo, hi, c);
            if (c & RL_LEAF_BIT) {
                j++;
                continue;
            }
        }
        f32x4 rows[6];
        u32x4 ch;
        read_node(S, top, c, rows, ch);
        loads++;
        const float lox[4] = {rows[0].x, rows[0].y, rows[0].z, rows[0].w}, hix[4] = {rows[1].x, rows[1].y, rows[1].z, rows[1].w};
        const float loy[4] = {rows[2].x, rows[2].y, rows[2].z, rows[2].w}, hiy[4] = {rows[3].x, rows[3].y, rows[3].z, rows[3].w};
        const float loz[4] = {rows[4].x, rows[4].y, rows[4].z, rows[4].w}, hiz[4] = {rows[5].x, rows[5].y, rows[5].z, rows[5].w};
        const uint32_t cw[4] = {ch.x, ch.y, ch.z, ch.w};
        bool ov[4];
        uint32_t k = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            ov[i] = cw[i] != RL_EMPTY_CHILD && box_overlaps(q, lox[i], hix[i], loy[i], hiy[i], loz[i], hiz[i]);
            k += ov[i] ? 1u : 0u;
        }
        if (loads == 1) {
            // The start node's overlapping children are the initial entries (at most four).
#pragma unroll
            for (int i = 0; i < 4; i++) {
                if (ov[i]) {
                    const float lo[3] = {lox[i], loy[i], loz[i]}, hi[3] = {hix[i], hiy[i], hiz[i]};
                    vt_set(v, n++, lo, hi, cw[i]);
                }
            }
            continue;
        }
        if (n - 1u + k > 4u) {
            j++;
            continue;
        }
        // Entry j becomes the first overlapping child, the others are appended; with none, the last
        // entry moves into slot j.
        uint32_t slot = j;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (ov[i]) {
                const float lo[3] = {lox[i], loy[i], loz[i]}, hi[3] = {hix[i], hiy[i], hiz[i]};
                vt_set(v, slot, lo, hi, cw[i]);
                slot = slot == j ? n : slot + 1u;
            }
        }
        if (k == 0) {
            n--;
            if (j != n) {
                float lo[3], hi[3];
                uint32_t c2;
                vt_get(v, n, lo, hi, c2);
                vt_set(v, j, lo, hi, c2);
            }
            vt_clear(v, n);
        } else {
            n += k - 1u;
        }
    }
    return n;
}

// Builds a wavefront's local tree for box q in the kVtNodes virtual nodes at `v` (node 0 the root,
// node 1 + j refining root entry j), which are slots first .. first + 4 of the LDS node array, and
// returns the root's node index. Every lane of the wavefront calls it with the same arguments.
constexpr uint32_t kVtNodes = 5;
constexpr uint32_t kVtMaxLoads = 48;

__device__ __forceinline__ uint32_t build_local_tree(const SceneDev& S, const TopCache& top, const QueryBox& q,
                                                     u32x4* v, uint32_t first) {
    const uint32_t lane = lane_id();
    if (lane < kVtNodes * 4u) vt_clear(v + (lane >> 2) * kNodeRows, lane & 3u);
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) build_local_node(S, top, 0u, q, v, kVtMaxLoads);
    __builtin_amdgcn_wave_barrier();
    uint32_t c = RL_EMPTY_CHILD, k = 0;
    float lo[3], hi[3];
    if (lane < 4u) {
        vt_get(v, lane, lo, hi, c);
        if (c != RL_EMPTY_CHILD && !(c & RL_LEAF_BIT))
            k = build_local_node(S, top, c, q, v + (1u + lane) * kNodeRows, kVtMaxLoads);
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < 4u && c != RL_EMPTY_CHILD && !(c & RL_LEAF_BIT)) {
        if (k == 0) {
            vt_clear(v, lane);  // no subtree under this entry overlaps q
        } else if (k == 1) {
            uint32_t c1;
            vt_get(v + (1u + lane) * kNodeRows, 0, lo, hi, c1);
            vt_set(v, lane, lo, hi, c1);
        } else {
            reinterpret_cast<uint32_t*>(v)[6 * 4 + lane] = kVirtualBase + first + 1u + lane;
        }
    }
    __builtin_amdgcn_wave_barrier();
    return kVirtualBase + first;
}

// One internal node: its four child boxes tested against (tnear, lim), hit children pushed
// far-first, the nearest taken next; with no child hit the next node is popped. Returns false when
// the stack was empty (traversal ended).
// VT: the node may be a virtual node (index >= kVirtualBase, held in top.rows).
template <int CAP, bool VT = false>
__device__ __forceinline__ bool visit_node(const SceneDev& S, const TopCache& top, const SlabRay& r, float tnear,
                                           float lim, uint32_t& node, int& sp, uint32_t* stk, uint32_t stride,
                                           uint32_t* spill) {
    f32x4 ex, fx, ey, fy, ez, fz;
    u32x4 ch;
    const uint32_t off = node << 7;  // byte offset of the node (in 32 bits, also for virtual nodes)
    const uint32_t aex = off + r.rex, afx = off + r.rfx, aey = off + r.rey, afy = off + r.rfy, aez = off + r.rez,
                   afz = off + r.rfz;
    if (node < top.count || (VT && node >= kVirtualBase)) {
        const char* b = reinterpret_cast<const char*>(top.rows);
        ex = __builtin_bit_cast(f32x4, *reinterpret_cast<const u32x4*>(b + aex));
        fx = __builtin_bit_cast(f32x4, *reinterpret_cast<const u32x4*>(b + afx));
        ey = __builtin_bit_cast(f32x4, *reinterpret_cast<const u32x4*>(b + aey));
        fy = __builtin_bit_cast(f32x4, *reinterpret_cast<const u32x4*>(b + afy));
        ez = __builtin_bit_cast(f32x4, *reinterpret_cast<const u32x4*>(b + aez));
        fz = __builtin_bit_cast(f32x4, *reinterpret_cast<const u32x4*>(b + afz));
        ch = *reinterpret_cast<const u32x4*>(b + off + 96u);
    } else {
        typedef const char __attribute__((address_space(1))) gchar;
        gchar* b = reinterpret_cast<gchar*>(gmem(reinterpret_cast<const float4*>(S.nodes)));
        ex = *reinterpret_cast<const gfloat4*>(b + aex);
        fx = *reinterpret_cast<const gfloat4*>(b + afx);
        ey = *reinterpret_cast<const gfloat4*>(b + aey);
        fy = *reinterpret_cast<const gfloat4*>(b + afy);
        ez = *reinterpret_cast<const gfloat4*>(b + aez);
        fz = *reinterpret_cast<const gfloat4*>(b + afz);
        ch = *reinterpret_cast<const guint4*>(b + off + 96u);
    }
    float key[4];
    uint32_t cid[4] = {ch.x, ch.y, ch.z, ch.w};
    box4(ex, fx, ey, fy, ez, fz, r, tnear, lim, key);
    // The child ids are only used when a box was hit; this keeps their load with the box rows
    // instead of letting the compiler issue it (and wait for it) behind the test.
    asm volatile("" ::"v"(ch.x), "v"(ch.y), "v"(ch.z), "v"(ch.w));
    // Sort the four (distance, child) pairs: near child next, the rest pushed far-first.
    cswap(key[0], cid[0], key[1], cid[1]);
    cswap(key[2], cid[2], key[3], cid[3]);
    cswap(key[0], cid[0], key[2], cid[2]);
    cswap(key[1], cid[1], key[3], cid[3]);
    cswap(key[1], cid[1], key[2], cid[2]);
    if (key[0] == __builtin_inff()) {
        if (sp == 0) return false;
        node = stack_pop(sp, stk, stride, spill);
        return true;
    }
    const bool h1 = key[1] != __builtin_inff(), h2 = key[2] != __builtin_inff(), h3 = key[3] != __builtin_inff();
    if (sp <= kLdsStack - 3) {
        // The k = h1 + h2 + h3 other hits (sorted to the front) go to slots sp .. sp+k-1, farthest
        // first. All three slots are written: the ones from sp+k up are above the new top.
        uint32_t* p = stk + (uint32_t)sp * stride;
        p[0] = h3 ? cid[3] : (h2 ? cid[2] : cid[1]);
        p[stride] = h3 ? cid[2] : cid[1];
        p[2u * stride] = cid[1];
        sp += (int)h1 + (int)h2 + (int)h3;
    } else {
        if (h3 && sp < CAP) stack_push(cid[3], sp, stk, stride, spill);
        if (h2 && sp < CAP) stack_push(cid[2], sp, stk, stride, spill);
        if (h1 && sp < CAP) stack_push(cid[1], sp, stk, stride, spill);
    }
    node = cid[0];
    return true;
}

template <int CAP, bool VT = false>
__device__ __forceinline__ bool any_hit_node(const SceneDev& S, const TopCache& top, const TravRay& r, uint32_t& node,
                                             int& sp, uint32_t* stk, uint32_t stride, uint32_t* spill) {
    return visit_node<CAP, VT>(S, top, r.s, r.w.tnear, r.w.tfar * kBoxScale, node, sp, stk, stride, spill);
}

// One step of any-hit traversal: either one internal node or one leaf (its triangles tested).
// Returns 1 when the ray is occluded, -1 when traversal ended without a hit, 0 to continue.
// Same box and triangle tests as traverse(), so results equal traverse<true>().
template <int CAP>
__device__ __forceinline__ int any_hit_step(const SceneDev& S, const TopCache& top, const TravRay& r, uint32_t& node,
                                            int& sp, uint32_t* stk, uint32_t stride, uint32_t* spill) {
    if (!(node & RL_LEAF_BIT)) return any_hit_node<CAP>(S, top, r, node, sp, stk, stride, spill) ? 0 : -1;
    const uint32_t first = node & RL_LEAF_FIRST_MASK;
    const uint32_t count = ((node >> 27) & 15u) + 1u;
    for (uint32_t j = 0; j < count; j++) {
        const gfloat* tf = rtri_ptr(S.tris_rot, first + j);
        float t, U, V, W, det;
        if (wt_intersect_rotated(r.w, rtri_vertex(tf, 0, r.w.rot), rtri_vertex(tf, 1, r.w.rot),
                                 rtri_vertex(tf, 2, r.w.rot), t, U, V, W, det))
            return 1;
    }
    if (sp == 0) return -1;
    node = stack_pop(sp, stk, stride, spill);
    return 0;
}

// A longer step of any-hit traversal in two phases: internal nodes until this lane holds a leaf
// (a lane that reaches one first waits for the rest of its wavefront, so the triangle tests that
// follow run with the lanes together), then two triangles per iteration through that leaf and
// through any leaves popped right after it. After its first node, the node phase ends early for
// the whole wavefront once no more than `min_lanes` of the calling lanes are still searching;
// those lanes return 0 and continue on the next call (every call visits at least one node of
// each searching lane, so repeated calls always finish). Returns like any_hit_step; the set of
// nodes and triangles a ray visits is the same (only the grouping into steps differs), so the
// occlusion result is the same.
// UNI: the caller has checked one_scene(S0). VT: nodes may be virtual (local trees). On a hit,
// hit_tri receives the occluding triangle's leaf-order index.
template <int CAP, bool UNI, bool VT = false>
__device__ __forceinline__ int any_hit_run(const SceneDev& S0, const TopCache& top, const TravRay& r, uint32_t& node,
                                           int& sp, uint32_t* stk, uint32_t stride, uint32_t* spill,
                                           uint32_t min_lanes, uint32_t& hit_tri) {
    const SceneDev S = UNI ? uniform_scene(S0) : S0;
    // min_lanes is the same in every lane; said so, the loop's exit test is a scalar compare.
    const uint32_t exit_lanes = wave_uniform32(min_lanes);
    bool ended = false;
    for (uint32_t stop_at = 0;; stop_at = exit_lanes) {
        const bool searching = !ended && !(node & RL_LEAF_BIT);
        if ((uint32_t)__popcll(__ballot(searching)) <= stop_at) break;
        if (searching && !any_hit_node<CAP, VT>(S, top, r, node, sp, stk, stride, spill)) ended = true;
    }
    if (ended) return -1;
    if (!(node & RL_LEAF_BIT)) return 0;
    uint32_t cur = node & RL_LEAF_FIRST_MASK;
    uint32_t end = cur + ((node >> 27) & 15u) + 1u;
    for (;;) {
        // Two triangles per iteration, each as three 12-B loads of its vertices already in the
        // ray's component order (rtri_vertex), both triangles' loads issued before either test (the
        // last triangle of a leaf is tested twice when one is left): this loop is bound by
        // vector-memory data return (city AO 346.5 -> 341.6 ms against one triangle per iteration;
        // nine per-component loads of one triangle measured slower still).
        const uint32_t c2 = cur + 1u < end ? cur + 1u : cur;
        const gfloat* ta = rtri_ptr(S.tris_rot, cur);
        const gfloat* tb = rtri_ptr(S.tris_rot, c2);
        const uint32_t rot = r.w.rot;
        const V3 a0 = rtri_vertex(ta, 0, rot), a1 = rtri_vertex(ta, 1, rot), a2 = rtri_vertex(ta, 2, rot);
        const V3 b0 = rtri_vertex(tb, 0, rot), b1 = rtri_vertex(tb, 1, rot), b2 = rtri_vertex(tb, 2, rot);
        float t, U, V, W, det;
        if (wt_intersect_rotated(r.w, a0, a1, a2, t, U, V, W, det)) {
            hit_tri = cur;
            return 1;
        }
        if (wt_intersect_rotated(r.w, b0, b1, b2, t, U, V, W, det)) {
            hit_tri = c2;
            return 1;
        }
        cur = c2;
        if (++cur == end) {
            if (sp == 0) return -1;
            node = stack_pop(sp, stk, stride, spill);
            if (!(node & RL_LEAF_BIT)) return 0;
            cur = node & RL_LEAF_FIRST_MASK;
            end = cur + ((node >> 27) & 15u) + 1u;
        }
    }
}

// Whether the ray hits triangle `tri` (leaf order) in its interval: the watertight test of the
// traversal's leaves, so a hit here is a hit the traversal would also find.
__device__ __forceinline__ bool hits_triangle(const SceneDev& S, const WtRay& w, uint32_t tri) {
    const gfloat* ta = rtri_ptr(S.tris_rot, tri);
    float t, U, V, W, det;
    return wt_intersect_rotated(w, rtri_vertex(ta, 0, w.rot), rtri_vertex(ta, 1, w.rot), rtri_vertex(ta, 2, w.rot), t, U,
                                V, W, det);
}

__device__ __forceinline__ bool hits_triangle(const SceneDev& S, const TravRay& r, uint32_t tri) {
    return hits_triangle(S, r.w, tri);
}

// One round of closest-hit traversal: a node phase, then the triangles of the leaves it found.
// In the node phase a lane that reaches a leaf holds it and goes on visiting nodes (speculatively:
// its interval has not yet been shortened by that leaf's triangles) until it reaches a second leaf
// or runs out of nodes; the phase ends when every lane holds a leaf or has finished (Aila and
// Laine 2009, speculative while-while), so lanes do not idle while others search, or, after its
// first node, once no more than `min_lanes` lanes are still searching without a leaf (those return
// 0 and continue on the next call). The leaf phase tests the held leaf's triangles, two per
// iteration, then any leaf the lane stopped at and the leaves popped right after it. Returns 0
// when the round ends at an internal node (call again), -1 when traversal is finished: res then
// holds the closest hit (tri = 0xFFFFFFFF for a miss). Start with node = 0, sp = 0, pend = 0 and
// res = {r.w.tfar, 0xFFFFFFFF, 0, 0, 0, 0}. Every node a lane skips is culled against an interval
// no shorter than the final one, so the lane tests every triangle that traverse<false>() could
// return, and the result (nearest t, ties to the smaller id) equals traverse<false>()'s; rounds
// only let a caller give finished lanes new rays.
// UNI: the caller has checked one_scene(S0).
template <int CAP, bool UNI>
__device__ __forceinline__ int closest_run(const SceneDev& S0, const TopCache& top, const TravRay& r,
                                           uint32_t& node, int& sp, uint32_t* stk, uint32_t stride, uint32_t* spill,
                                           TraceResult& res, uint32_t& pend, uint32_t min_lanes, uint32_t leaf_lanes) {
    const SceneDev S = UNI ? uniform_scene(S0) : S0;
    uint32_t held = pend;  // the rest of a leaf an earlier round stopped in
    bool has = pend != 0u, done = false;
    pend = 0u;
    for (uint32_t stop_at = 0;; stop_at = min_lanes) {
        if (!done && !has && (node & RL_LEAF_BIT)) {
            held = node;
            has = true;
            if (sp == 0) done = true;
            else node = stack_pop(sp, stk, stride, spill);
        }
        if ((uint32_t)__popcll(__ballot(!has && !done)) <= stop_at) break;
        if (!done && !(node & RL_LEAF_BIT) &&
            !visit_node<CAP>(S, top, r.s, r.w.tnear, res.t * kBoxScale, node, sp, stk, stride, spill))
            done = true;
    }
    if (!has) return done ? -1 : 0;  // finished without another leaf, or still searching
    uint32_t cur = held & RL_LEAF_FIRST_MASK;
    uint32_t end = cur + ((held >> 27) & 15u) + 1u;
    // The leaf phase, two triangles per iteration (the last one of a leaf tested twice when one is
    // left: the tie rule makes the repeat a no-op); after its first, it ends for the wavefront once
    // no more than `leaf_lanes` lanes still have triangles, and those keep the rest of their leaf
    // for the next round (in `pend`, or as the next node when nothing follows it).
    bool leafing = true;
    for (uint32_t stop_at = 0;; stop_at = leaf_lanes) {
        if ((uint32_t)__popcll(__ballot(leafing)) <= stop_at) break;
        if (leafing) {
            // Each triangle as three 16-B row loads with per-lane component selects, both
            // triangles' loads issued before either test: this loop waits on the loads' return
            // (heightfield final gather: nine per-component loads of one triangle 283 ms, three
            // row loads 261 ms, two triangles per iteration 252 ms, three 265 ms, four 287 ms).
            const uint32_t c2 = cur + 1u < end ? cur + 1u : cur;
            const gfloat* ta = tri_ptr(S.tris, cur);
            const gfloat* tb = tri_ptr(S.tris, c2);
            const V3 a0 = v3(ta[0], ta[1], ta[2]), a1 = v3(ta[4], ta[5], ta[6]), a2 = v3(ta[8], ta[9], ta[10]);
            const V3 b0 = v3(tb[0], tb[1], tb[2]), b1 = v3(tb[4], tb[5], tb[6]), b2 = v3(tb[8], tb[9], tb[10]);
            const uint32_t ida = f_bits(ta[3]), idb = f_bits(tb[3]);
            float t, U, V, W, det;
            if (wt_intersect(r.w, a0, a1, a2, t, U, V, W, det) && (t < res.t || (t == res.t && ida < res.tri))) {
                res.t = t;
                res.tri = ida;
                res.U = U; res.V = V; res.W = W; res.det = det;
            }
            if (wt_intersect(r.w, b0, b1, b2, t, U, V, W, det) && (t < res.t || (t == res.t && idb < res.tri))) {
                res.t = t;
                res.tri = idb;
                res.U = U; res.V = V; res.W = W; res.det = det;
            }
            cur = c2;
            if (++cur == end) {
                if (done || !(node & RL_LEAF_BIT)) {
                    leafing = false;
                } else {
                    cur = node & RL_LEAF_FIRST_MASK;
                    end = cur + ((node >> 27) & 15u) + 1u;
                    if (sp == 0) done = true;
                    else node = stack_pop(sp, stk, stride, spill);
                }
            }
        }
    }
    if (leafing) {
        const uint32_t rest = RL_LEAF_BIT | ((end - cur - 1u) << 27) | cur;
        if (done) node = rest;  // the stack is empty: the next round takes it as its leaf and ends
        else pend = rest;
        return 0;
    }
    return done ? -1 : 0;
}

// Traverses the BVH4 for one ray. ANY: stop at the first triangle with t in (tnear, tfar).
// Closest: the hit with the smallest t, ties broken by the smaller scene-order triangle id, so
// the result does not depend on traversal order. stk points at this lane's first LDS stack slot;
// consecutive entries are `stride` words apart (one column per lane, conflict-free). CAP is the
// total stack capacity (LDS + spill), at least the BVH's worst-case bound.
//
// Slab test: as described at SlabRay.
// w: wt_setup(ray), for callers that have it already.
template <bool ANY, int CAP>
__device__ bool traverse(const SceneDev& S, const TopCache& top, const Ray& ray, const WtRay& w, uint32_t* stk,
                         uint32_t stride, TraceResult& res) {
    const SlabRay sr = slab_ray(ray);
    const float tnear = ray.tnear;
    float tlimit = ray.tfar;  // closest: shrinks to the best hit so far
    res.t = ray.tfar;
    res.tri = 0xFFFFFFFFu;
    res.U = res.V = res.W = res.det = 0.0f;
    uint32_t spill[CAP - kLdsStack];

    // Two phases per round (Aila and Laine 2009, "while-while"): internal nodes until this lane
    // holds a leaf (a lane that reaches one first waits for the rest of its wavefront, so the
    // triangle tests run with the lanes together), then one triangle per iteration through that
    // leaf and any leaves popped right after it. Each lane visits the same nodes in the same order
    // as a one-node-or-one-leaf-per-iteration loop would.
    uint32_t node = 0;
    int sp = 0;
    for (;;) {
        while (!(node & RL_LEAF_BIT))
            if (!visit_node<CAP>(S, top, sr, tnear, tlimit * kBoxScale, node, sp, stk, stride, spill))
                goto done;
        uint32_t cur = node & RL_LEAF_FIRST_MASK;
        uint32_t end = cur + ((node >> 27) & 15u) + 1u;
        for (;;) {
            float t, U, V, W, det;
            bool hit;
            uint32_t id;
            if (ANY) {
                const gfloat* tf = rtri_ptr(S.tris_rot, cur);
                hit = wt_intersect_rotated(w, rtri_vertex(tf, 0, w.rot), rtri_vertex(tf, 1, w.rot),
                                           rtri_vertex(tf, 2, w.rot), t, U, V, W, det);
                id = rtri_id(tf);
            } else {
                const gfloat* tf = tri_ptr(S.tris, cur);
                hit = wt_intersect(w, v3(tf[0], tf[1], tf[2]), v3(tf[4], tf[5], tf[6]), v3(tf[8], tf[9], tf[10]), t, U,
                                   V, W, det);
                id = f_bits(tf[3]);
            }
            if (hit) {
                if (ANY) {
                    res.t = t;
                    res.tri = id;
                    res.leaf = cur;
                    return true;
                }
                if (t < res.t || (t == res.t && id < res.tri)) {
                    res.t = t;
                    res.tri = id;
                    res.U = U; res.V = V; res.W = W; res.det = det;
                    tlimit = t;
                }
            }
            if (++cur == end) {
                if (sp == 0) goto done;
                node = stack_pop(sp, stk, stride, spill);
                if (!(node & RL_LEAF_BIT)) break;
                cur = node & RL_LEAF_FIRST_MASK;
                end = cur + ((node >> 27) & 15u) + 1u;
            }
        }
    }
done:
    return res.tri != 0xFFFFFFFFu;
}

// traverse() without an LDS copy of the top nodes.
template <bool ANY, int CAP>
__device__ __forceinline__ bool traverse(const SceneDev& S, const TopCache& top, const Ray& ray, uint32_t* stk,
                                         uint32_t stride, TraceResult& res) {
    return traverse<ANY, CAP>(S, top, ray, wt_setup(ray), stk, stride, res);
}

template <bool ANY, int CAP>
__device__ __forceinline__ bool traverse(const SceneDev& S, const Ray& ray, uint32_t* stk, uint32_t stride,
                                         TraceResult& res) {
    return traverse<ANY, CAP>(S, TopCache{nullptr, 0u}, ray, wt_setup(ray), stk, stride, res);
}

// Face::colorAt (scene/Mesh.cpp:478-485) at Embree's (u, v): the mesh's material colour (rgba;
// (1, 1, 1, 1) without a material), times Texture::colorAt at the diffuse uv for a textured
// material (TexturedMaterial::colorAt, scene/Material.cpp:67-70: Rgba(colour, 1) * texel).
__device__ __forceinline__ float4 face_color(const SceneDev& S, uint32_t tri, float u, float v) {
    const TriAttr& A = S.attr[tri];
    float4 c = make_float4(A.albedo[0], A.albedo[1], A.albedo[2], A.albedo[3]);
    if (S.mat) {
        const TriMat m = S.mat[tri];
        if (m.tex >= 0) {
            const TexDev T = S.texs[m.tex];
            float tc[4];
            texture_color(T.px, T.w, T.h, T.channels, face_lerp(m.duv[0], m.duv[2], m.duv[4], u, v),
                          face_lerp(m.duv[1], m.duv[3], m.duv[5], u, v), tc);
            c = make_float4(c.x * tc[0], c.y * tc[1], c.z * tc[2], c.w * tc[3]);
        }
    }
    return c;
}

__device__ __forceinline__ int find_seg(const BakeSeg* segs, int nsegs, uint64_t g) {
    int lo = 0, hi = nsegs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (segs[mid].work_begin <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}


// Persistent work queue: a wavefront takes the next n items (wave_take returns the first index,
// wave_next_chunk takes one) and every lane gets the index. A wavefront keeps taking until the
// index passes the item count, so every wavefront reaches its exit and the grid drains. Every
// lane executes the atomic (lane 0 adds n, the others 0) and
// the index is returned wave-uniform: with a lane-0-only atomic at the loop head and a lane-0-only
// store at the loop end, the compiler may rotate the loop so that the two regions merge and the
// other lanes re-read the index from lane 0 after it has left their loop, never reaching the exit.
__device__ __forceinline__ uint64_t wave_take(unsigned long long* counter, uint32_t n) {
    const unsigned long long c = atomicAdd(counter, lane_id() == 0 ? (unsigned long long)n : 0ull);
    return wave_uniform(__shfl(c, 0, 64));
}

// wave_take of one item (the common case: one chunk per wavefront).
__device__ __forceinline__ uint64_t wave_next_chunk(unsigned long long* counter) { return wave_take(counter, 1u); }

__device__ __forceinline__ void count_rays(uint32_t n, unsigned long long* counter) {
    // Wave sum of per-lane ray counts, one atomic per wavefront.
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) n += __shfl_xor(n, off, 64);
    if (lane_id() == 0 && n) atomicAdd(counter, (unsigned long long)n);
}

}  // namespace trav
}  // namespace rl
