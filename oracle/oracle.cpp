// ORACLE — test infrastructure only. A CPU restatement of dmsovetov/Relight's bake math used to
// check the GPU library; nothing in the product links or calls it. Only tests/, the smoke check
// and bench.py's CPU-baseline leg load liboracle.so.
//
// What it restates (file:line in the reference):
//   segment -> ray            rt/Embree.cpp:107-130 (dir = normalize(end-start), (0.01, len))
//   hit semantics             rt/Embree.cpp:145-177 (closest hit; Embree u,v)
//   lumel raster              Lightmap.cpp:118-164, scene/Mesh.cpp:397-476 (Face::isUvInside etc.)
//   direct light              baker/DirectLight.cpp:46-125, scene/Light.cpp:372-500
//   area light vertices       scene/Light.cpp:293-323, scene/Mesh.cpp:322-327 (centroids)
//   ambient occlusion         baker/AmbientOcclusion.cpp:46-66
//   hemisphere sampling       Types.h:924-949 (uniform, with PCG32 instead of rand())
//   photons                   baker/Photons.cpp:50-144, scene/Light.cpp:334-361, Relight.cpp:244-257
//   photonmap gather          Lightmap.cpp:376-413 (texel disc); k-nearest photons in 3D with the
//                             same normalisation (the configuration's hash-grid gather)
//   final gather              baker/IndirectLight.cpp:47-75
//   form factors              hemisphere-ray estimator of radiosity/RadiosityBuilder.cpp:142-176
//   post-process, encodings   Lightmap.cpp:167-331 (expand, blur, toDoubleLdr, toRgbmLdr), Types.h:216-234
//   lightmap uv generation    scene/Mesh.cpp:229-238, Types.h:2293-3326 (charts, flatten, rotate, pack)
// Project-defined parts shared by contract with the GPU path (not in the reference): PCG32 streams
// keyed by (seed, lightmap seed, texel, technique, sample); cosine sampling (Malley + Duff et al.
// 2017 basis); the watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013); closest-hit ties
// broken by the smaller scene-order triangle id.
//
// Pinning (tests/test_oracle.py): PCG32 against the published pcg32-demo output and an
// independent Python integer model; sampler moments; ray/triangle known answers and shared-edge
// watertightness; BVH == brute force; direct light closed forms; AO closed-box and open-plane
// cases; first-bounce photon energy; texel-disc gather by hand; k-nearest gather against a
// numpy brute force; final gather under a constant sky in closed form; form factors against
// Howell's parallel and perpendicular plate formulas and closed-enclosure row sums.
// tests/test_oracle_pins.py: the photon bounce loop in a closed sphere (per-depth power
// I * albedo^(d+1) / 2^d; Russian roulette survivors and per-depth sums), the directional
// emitter's start plane, the final gather's hit branch (front face, back face, the
// floor(uv * (size - 1)) lookup at the barycentric-swapped uv) against ray/plane predictions,
// and the mesh transform against R p + t with renormalised normals.
// tests/test_materials.py: Texture::colorAt known answers (truncating wrap, channel order, magenta
// outside), the alpha re-trace (one-unit restart, 25-step limit, checker alpha texture) and
// textured photon reflectance. tests/test_radiosity.py: also reciprocity A_i F_ij = A_j F_ji on
// unequal plates and the solve against an FP64 LU solve. tests/test_uv.py: uv generation against
// hand-derived layouts and invariants (uvs in [0, 1], constant uv/3-D area ratio per chart, no
// texel centre in two charts). tests/test_postprocess.py and test_obj.py: expand, blur and the
// encodings by hand; the OBJ parse on the reference's own data files.
// Compiled with -ffp-contract=off so every expression rounds as written.
#include <stdint.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

// ------------------------------------------------------------------ vectors
struct Vec {
    float x, y, z;
};
inline Vec mk(float x, float y, float z) { Vec v = {x, y, z}; return v; }
inline Vec add(Vec a, Vec b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline Vec sub(Vec a, Vec b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline Vec scale(Vec a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline Vec divs(Vec a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
inline Vec neg(Vec a) { return mk(-a.x, -a.y, -a.z); }
inline float dot3(Vec a, Vec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float len3(Vec a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
// Vec3::normalize, Types.h:852-859.
inline float normalize3(Vec& a) {
    float l = len3(a);
    if (l) { a.x /= l; a.y /= l; a.z /= l; }
    return l;
}
// powf of LightSpotCutoff (scene/Light.cpp:464-477) and AmbientOcclusion::bakeLumel
// (baker/AmbientOcclusion.cpp:61-63), evaluated as the double pow rounded to float: the
// correctly rounded powf except within about 2^-52 of a float rounding midpoint, and the same
// bits as the GPU library (single-precision powf implementations differ in the last bit).
inline float ref_powf(float x, float y) { return (float)pow((double)x, (double)y); }
inline float axis(Vec a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// ------------------------------------------------------------------ PCG32 (O'Neill 2014)
struct Rng {
    uint64_t s, inc;
};
uint32_t rng_u32(Rng& r) {
    const uint64_t prev = r.s;
    r.s = prev * 6364136223846793005ULL + r.inc;
    const uint32_t x = (uint32_t)(((prev >> 18u) ^ prev) >> 27u);
    const uint32_t k = (uint32_t)(prev >> 59u);
    return (x >> k) | (x << ((0u - k) & 31u));
}
Rng rng_seed(uint64_t initstate, uint64_t initseq) {
    Rng r;
    r.s = 0;
    r.inc = (initseq << 1u) | 1u;
    rng_u32(r);
    r.s += initstate;
    rng_u32(r);
    return r;
}
float rng_f(Rng& r) { return (float)(rng_u32(r) >> 8) * (1.0f / 16777216.0f); }
uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
Rng stream(uint32_t seed, uint32_t lm_seed, uint32_t texel, uint32_t tech, uint32_t sample) {
    uint64_t k = mix64(((uint64_t)seed << 32) | lm_seed);
    k = mix64(k ^ (uint64_t)texel);
    return rng_seed(k, ((uint64_t)tech << 32) | sample);
}

// Vec3::randomDirection, Types.h:924-937, with PCG draws (zero-length draws rejected).
Vec hemi_uniform_sphere(Rng& r) {
    Vec d;
    float l;
    for (;;) {
        d.x = rng_f(r) * 2.0f - 1.0f;
        d.y = rng_f(r) * 2.0f - 1.0f;
        d.z = rng_f(r) * 2.0f - 1.0f;
        l = len3(d);
        if (!(l > 1.0f || l == 0.0f)) break;
    }
    return divs(d, l);
}

// Vec3::randomHemisphereDirection, Types.h:940-949.
Vec hemi_uniform(Rng& r, Vec n) {
    Vec d = hemi_uniform_sphere(r);
    if (dot3(d, n) < 0.0f) d = neg(d);
    return d;
}

// Cosine-weighted: Malley's method on a rejection-sampled disk, Duff et al. 2017 basis.
Vec hemi_cosine(Rng& r, Vec n) {
    float x, y, q;
    for (;;) {
        x = rng_f(r) * 2.0f - 1.0f;
        y = rng_f(r) * 2.0f - 1.0f;
        q = x * x + y * y;
        if (q < 1.0f) break;
    }
    const float z = sqrtf(1.0f - q);
    const float sg = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sg + n.z);
    const float b = n.x * n.y * a;
    const Vec e1 = mk(1.0f + sg * n.x * n.x * a, sg * b, -sg * n.x);
    const Vec e2 = mk(b, sg + n.y * n.y * a, -n.y);
    return mk(e1.x * x + e2.x * y + n.x * z, e1.y * x + e2.y * y + n.y * z, e1.z * x + e2.z * y + n.z * z);
}

Vec hemi(Rng& r, Vec n, int mode) { return mode == 1 ? hemi_cosine(r, n) : hemi_uniform(r, n); }

// ------------------------------------------------------------------ rays and triangles
struct Seg {
    Vec o, d;
    float t0, t1;
    int kx, ky, kz;
    float sx, sy, sz;
    Vec inv;
};

Seg make_seg(Vec start, Vec end) {
    Seg s;
    s.o = start;
    s.d = sub(end, start);
    s.t1 = normalize3(s.d);
    s.t0 = 0.01f;
    const float ax = fabsf(s.d.x), ay = fabsf(s.d.y), az = fabsf(s.d.z);
    int kz;
    if (ax > ay) kz = ax > az ? 0 : 2;
    else kz = ay > az ? 1 : 2;
    int kx = (kz + 1) % 3, ky = (kz + 2) % 3;
    if (axis(s.d, kz) < 0.0f) std::swap(kx, ky);
    s.kx = kx; s.ky = ky; s.kz = kz;
    s.sx = axis(s.d, kx) / axis(s.d, kz);
    s.sy = axis(s.d, ky) / axis(s.d, kz);
    s.sz = 1.0f / axis(s.d, kz);
    s.inv = mk(1.0f / s.d.x, 1.0f / s.d.y, 1.0f / s.d.z);
    return s;
}

// Woop-Benthin-Wald watertight test; on success t in (t0, t1) and the Embree barycentrics.
bool tri_hit(const Seg& s, const float* v, float& t, float& bu, float& bv) {
    const Vec A = sub(mk(v[0], v[1], v[2]), s.o);
    const Vec B = sub(mk(v[3], v[4], v[5]), s.o);
    const Vec C = sub(mk(v[6], v[7], v[8]), s.o);
    const float az_ = axis(A, s.kz), bz_ = axis(B, s.kz), cz_ = axis(C, s.kz);
    const float Ax = axis(A, s.kx) - s.sx * az_, Ay = axis(A, s.ky) - s.sy * az_;
    const float Bx = axis(B, s.kx) - s.sx * bz_, By = axis(B, s.ky) - s.sy * bz_;
    const float Cx = axis(C, s.kx) - s.sx * cz_, Cy = axis(C, s.ky) - s.sy * cz_;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    const bool anyneg = U < 0.0f || V < 0.0f || W < 0.0f;
    const bool anypos = U > 0.0f || V > 0.0f || W > 0.0f;
    if (anyneg && anypos) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float T = U * (s.sz * az_) + V * (s.sz * bz_) + W * (s.sz * cz_);
    t = T / det;
    if (!(t > s.t0 && t < s.t1)) return false;
    bu = V / det;
    bv = W / det;
    return true;
}

// ------------------------------------------------------------------ scene + binary BVH
struct Light {
    int32_t type, shadow;
    Vec pos, dir, color;
    float intensity;
    int32_t has_att;
    float radius, c, l, q, cutoff, sexp;
    std::vector<Vec> verts;
};

struct BNode {  // binned-SAH binary node: box, then children or a leaf range
    float lo[3], hi[3];
    int32_t left, right;
    uint32_t first, count;
};

struct Scene {
    std::vector<float> tri;       // 9 floats per scene-order triangle
    std::vector<uint32_t> tmesh, tface;
    std::vector<float> tuv;       // 6 per triangle (lightmap uv of a, b, c)
    std::vector<float> tnrm;      // 9 per triangle
    std::vector<float> talb;      // 4 per triangle (material colour rgba)
    std::vector<int32_t> ttex;    // per triangle: texture index or -1 (TexturedMaterial)
    std::vector<float> tduv;      // 6 per triangle (diffuse uv of a, b, c)
    struct Tex { int32_t w, h, ch; std::vector<uint8_t> px; };
    std::vector<Tex> textures;
    bool alpha = false;           // some material can have alpha <= 0.1 (set by or_scene_build)
    // per mesh
    std::vector<std::vector<Vec>> mpos, mnrm;
    std::vector<std::vector<float>> muv;
    std::vector<std::vector<uint32_t>> midx;
    std::vector<uint32_t> mfirst;
    std::vector<Light> lights;
    std::vector<BNode> nodes;
    std::vector<uint32_t> order;
    bool use_bvh = true;
    // Records of the last photon emission (slot order): texel key (0xFFFFFFFF: none), power, hit point.
    std::vector<uint32_t> last_key;
    std::vector<float> last_power, last_pos;
};

struct Counters {
    uint64_t nodes = 0, tris = 0;
};

void build_bvh(Scene& S) {
    const uint32_t n = (uint32_t)(S.tri.size() / 9);
    S.order.resize(n);
    std::vector<float> bl(3 * n), bh(3 * n), ce(3 * n);
    for (uint32_t i = 0; i < n; i++) {
        S.order[i] = i;
        for (int k = 0; k < 3; k++) {
            const float a = S.tri[9 * i + k], b = S.tri[9 * i + 3 + k], c = S.tri[9 * i + 6 + k];
            bl[3 * i + k] = std::min(a, std::min(b, c));
            bh[3 * i + k] = std::max(a, std::max(b, c));
            ce[3 * i + k] = 0.5f * (bl[3 * i + k] + bh[3 * i + k]);
        }
    }
    S.nodes.clear();
    S.nodes.push_back(BNode());
    struct W { int32_t node; uint32_t first, count; };
    std::vector<W> work;
    work.push_back({0, 0, n});
    auto area = [](const float* lo, const float* hi) {
        double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0) return 0.0;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    };
    while (!work.empty()) {
        W w = work.back();
        work.pop_back();
        BNode nd;
        float clo[3], chi[3];
        for (int k = 0; k < 3; k++) { nd.lo[k] = clo[k] = FLT_MAX; nd.hi[k] = chi[k] = -FLT_MAX; }
        for (uint32_t i = w.first; i < w.first + w.count; i++) {
            const uint32_t t = S.order[i];
            for (int k = 0; k < 3; k++) {
                nd.lo[k] = std::min(nd.lo[k], bl[3 * t + k]);
                nd.hi[k] = std::max(nd.hi[k], bh[3 * t + k]);
                clo[k] = std::min(clo[k], ce[3 * t + k]);
                chi[k] = std::max(chi[k], ce[3 * t + k]);
            }
        }
        nd.left = nd.right = -1;
        nd.first = w.first;
        nd.count = w.count;
        uint32_t mid = 0;
        bool do_split = false;
        if (w.count > 4) {
            int ax = 0;
            for (int k = 1; k < 3; k++) if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
            const float ext = chi[ax] - clo[ax];
            uint32_t* ord = &S.order[w.first];
            if (ext > 0.0f) {
                const int NB = 16;
                float blo[NB][3], bhi[NB][3];
                uint32_t cnt[NB] = {0};
                for (int b = 0; b < NB; b++) for (int k = 0; k < 3; k++) { blo[b][k] = FLT_MAX; bhi[b][k] = -FLT_MAX; }
                auto binof = [&](uint32_t t) { int b = (int)((ce[3 * t + ax] - clo[ax]) * (NB / ext)); return b < 0 ? 0 : (b >= NB ? NB - 1 : b); };
                for (uint32_t i = 0; i < w.count; i++) {
                    const uint32_t t = ord[i];
                    const int b = binof(t);
                    cnt[b]++;
                    for (int k = 0; k < 3; k++) { blo[b][k] = std::min(blo[b][k], bl[3 * t + k]); bhi[b][k] = std::max(bhi[b][k], bh[3 * t + k]); }
                }
                double best = DBL_MAX;
                int bestb = -1;
                for (int s = 1; s < NB; s++) {
                    float l0[3], l1[3], r0[3], r1[3];
                    uint32_t nl = 0, nr = 0;
                    for (int k = 0; k < 3; k++) { l0[k] = r0[k] = FLT_MAX; l1[k] = r1[k] = -FLT_MAX; }
                    for (int b = 0; b < NB; b++) {
                        if (!cnt[b]) continue;
                        float* L0 = b < s ? l0 : r0; float* L1 = b < s ? l1 : r1;
                        for (int k = 0; k < 3; k++) { L0[k] = std::min(L0[k], blo[b][k]); L1[k] = std::max(L1[k], bhi[b][k]); }
                        (b < s ? nl : nr) += cnt[b];
                    }
                    if (!nl || !nr) continue;
                    const double c = area(l0, l1) * nl + area(r0, r1) * nr;
                    if (c < best) { best = c; bestb = s; }
                }
                if (bestb > 0) {
                    uint32_t* p = std::partition(ord, ord + w.count, [&](uint32_t t) { return binof(t) < bestb; });
                    mid = w.first + (uint32_t)(p - ord);
                    do_split = mid > w.first && mid < w.first + w.count;
                }
            }
            if (!do_split) {
                std::nth_element(ord, ord + w.count / 2, ord + w.count, [&](uint32_t a, uint32_t b) { return ce[3 * a + ax] < ce[3 * b + ax]; });
                mid = w.first + w.count / 2;
                do_split = true;
            }
        }
        if (do_split) {
            nd.left = (int32_t)S.nodes.size();
            nd.right = nd.left + 1;
            S.nodes.push_back(BNode());
            S.nodes.push_back(BNode());
            work.push_back({nd.left, w.first, mid - w.first});
            work.push_back({nd.right, mid, w.first + w.count - mid});
        }
        S.nodes[w.node] = nd;
    }
}

// Robust slab test (Ize 2013): exit distances scaled by 1 + 2*gamma(3). An axis the ray does not
// move along constrains nothing when the origin lies in the slab (ends included) and rejects the
// box otherwise; (b - o) * (1/0) would give NaN or a one-sided infinity there.
inline bool box_hit(const Seg& s, const float* lo, const float* hi, float tlim, float& tenter) {
    const float K = 1.0000004f;
    const float o[3] = {s.o.x, s.o.y, s.o.z}, d[3] = {s.d.x, s.d.y, s.d.z}, inv[3] = {s.inv.x, s.inv.y, s.inv.z};
    float tmin = s.t0, tmax = INFINITY;
    for (int k = 0; k < 3; k++) {
        if (fabsf(d[k]) < 1e-30f) {
            if (o[k] < lo[k] || o[k] > hi[k]) return false;
            continue;
        }
        const float t0 = (lo[k] - o[k]) * inv[k], t1 = (hi[k] - o[k]) * inv[k];
        tmin = fmaxf(tmin, fminf(t0, t1));
        tmax = fminf(tmax, fmaxf(t0, t1));
    }
    tenter = tmin;
    return tmin <= fminf(tmax * K, tlim * K);
}

struct HitRec {
    float t = 0.0f, u = 0.0f, v = 0.0f;
    int32_t tri = -1;
};

// Closest hit (any=false) or any hit. Ties at equal t go to the smaller scene-order id.
bool trace(const Scene& S, const Seg& s, bool any, HitRec& h, Counters* ctr) {
    h = HitRec();
    float best = s.t1;
    auto test_tri = [&](uint32_t id) -> bool {
        float t, bu, bv;
        if (ctr) ctr->tris++;
        if (!tri_hit(s, &S.tri[9 * (size_t)id], t, bu, bv)) return false;
        if (h.tri < 0 || t < best || (t == best && (int32_t)id < h.tri)) {
            best = t; h.t = t; h.u = bu; h.v = bv; h.tri = (int32_t)id;
        }
        return true;
    };
    if (!S.use_bvh) {
        const uint32_t n = (uint32_t)(S.tri.size() / 9);
        for (uint32_t i = 0; i < n; i++)
            if (test_tri(i) && any) return true;
        return h.tri >= 0;
    }
    // Binary traversal, each step fetches one node and tests both children's boxes.
    int32_t stack[128];
    int sp = 0;
    const BNode& root = S.nodes[0];
    float te;
    if (ctr) ctr->nodes++;
    if (!box_hit(s, root.lo, root.hi, best, te)) return false;
    int32_t cur = 0;
    for (;;) {
        const BNode& nd = S.nodes[cur];
        if (nd.left < 0) {
            for (uint32_t i = nd.first; i < nd.first + nd.count; i++)
                if (test_tri(S.order[i]) && any) return true;
            if (!sp) break;
            cur = stack[--sp];
            continue;
        }
        if (ctr) ctr->nodes++;
        float ta, tb;
        const bool ha = box_hit(s, S.nodes[nd.left].lo, S.nodes[nd.left].hi, best, ta);
        const bool hb = box_hit(s, S.nodes[nd.right].lo, S.nodes[nd.right].hi, best, tb);
        if (ha && hb) {
            int32_t nearc = ta <= tb ? nd.left : nd.right;
            int32_t farc = ta <= tb ? nd.right : nd.left;
            stack[sp++] = farc;
            cur = nearc;
        } else if (ha) {
            cur = nd.left;
        } else if (hb) {
            cur = nd.right;
        } else {
            if (!sp) break;
            cur = stack[--sp];
        }
    }
    return h.tri >= 0;
}

// ------------------------------------------------------------------ materials
// Texture::colorAt (scene/Material.cpp:87-98): texel (int(u * w) % w, int(v * h) % h) with C's
// truncating conversion and remainder; magenta (1, 0, 1, 1) outside; channels / 255 (alpha 1
// with 3 channels). A product outside the int range (undefined in the reference) is magenta.
void texture_color(const Scene::Tex& T, float u, float v, float out[4]) {
    const float fx = u * (float)T.w, fy = v * (float)T.h;
    int x = -1, y = -1;
    if (fx > -2147483648.0f && fx < 2147483648.0f && fy > -2147483648.0f && fy < 2147483648.0f) {
        x = (int)fx % T.w;
        y = (int)fy % T.h;
    }
    if (x >= 0 && x < T.w && y >= 0 && y < T.h) {
        const uint8_t* p = &T.px[(size_t)y * T.w * T.ch + (size_t)x * T.ch];
        out[0] = p[0] / 255.0f;
        out[1] = p[1] / 255.0f;
        out[2] = p[2] / 255.0f;
        out[3] = T.ch == 4 ? p[3] / 255.0f : 1.0f;
    } else {
        out[0] = 1.0f; out[1] = 0.0f; out[2] = 1.0f; out[3] = 1.0f;
    }
}

// Face::colorAt (scene/Mesh.cpp:478-485) at Embree's (u, v): the material colour, times the
// texel at the diffuse uv a + (b - a) v + (c - a) u for a TexturedMaterial (Rgba(colour, 1) *
// Texture::colorAt, scene/Material.cpp:67-70).
void face_color(const Scene& S, uint32_t tri, float u, float v, float out[4]) {
    const float* A = &S.talb[4 * (size_t)tri];
    out[0] = A[0]; out[1] = A[1]; out[2] = A[2]; out[3] = A[3];
    const int32_t t = S.ttex[tri];
    if (t < 0) return;
    const float* D = &S.tduv[6 * (size_t)tri];
    const float du = D[0] + (D[2] - D[0]) * v + (D[4] - D[0]) * u;
    const float dv = D[1] + (D[3] - D[1]) * v + (D[5] - D[1]) * u;
    float c[4];
    texture_color(S.textures[t], du, dv, c);
    out[0] = out[0] * c[0]; out[1] = out[1] * c[1]; out[2] = out[2] * c[2]; out[3] = out[3] * c[3];
}

// The shadow test of LightInfluence::calculate / DirectionalLightInfluence::calculate
// (scene/Light.cpp:385,417): Embree::traceSegment(start, end, HitUseAlpha) (rt/Embree.cpp:145-177)
// as written: the closest hit blocks unless its colour's alpha is <= 0.1 and fewer than 25
// re-traces were made, in which case it recurses from point + direction (one unit past the
// hit). Each trace is a ray.
bool shadow_hit(const Scene& S, Vec start, Vec end, uint64_t& rays, int step = 0) {
    const Seg sg = make_seg(start, end);
    HitRec h;
    rays++;
    // With no transparent material the closest hit exists exactly when any hit does.
    if (!S.alpha) return trace(S, sg, true, h, nullptr);
    if (!trace(S, sg, false, h, nullptr)) return false;
    const Vec point = add(start, scale(sg.d, h.t));
    float c[4];
    face_color(S, (uint32_t)h.tri, h.u, h.v, c);
    if (c[3] <= 0.1f && step < 25) return shadow_hit(S, add(point, sg.d), end, rays, step + 1);
    return true;
}

// ------------------------------------------------------------------ bakers
// DirectLight::influenceFromPoint (baker/DirectLight.cpp:101-125) and the light models of
// scene/Light.cpp:372-500.
float influence(const Scene& S, const Light& L, Vec lp, Vec p, Vec n, uint64_t& rays) {
    float inf = 0.0f, att = 1.0f, cut = 1.0f, dist = 0.0f;
    HitRec h;
    if (L.type == 2) {  // directional: lambert(-dir, n), shadow towards p - dir*1000
        float dp = dot3(neg(L.dir), n);
        inf = dp < 0.0f ? 0.0f : dp;
        if (inf <= 0.001f) {
            inf = 0.0f;
        } else if (L.shadow) {
            if (shadow_hit(S, p, sub(p, scale(L.dir, 1000.0f)), rays)) inf *= 0.0f;
            else inf *= 1.0f;
        }
    } else {
        Vec d = sub(lp, p);
        dist = normalize3(d);
        float dp = dot3(d, n);
        inf = dp < 0.0f ? 0.0f : dp;
        if (inf <= 0.001f) {
            inf = 0.0f;
        } else if (L.shadow) {
            if (shadow_hit(S, p, lp, rays)) inf *= 0.0f;
            else inf *= 1.0f;
        }
    }
    if (L.type == 1) {  // LightSpotCutoff::calculate at the light position
        Vec sd = sub(p, L.pos);
        normalize3(sd);
        float v = dot3(sd, L.dir);
        if (v <= L.cutoff) cut = 0.0f;
        else {
            v = (1.0f - (1.0f - v) * 1.0f / (1.0f - L.cutoff));
            if (fabsf(1.0f - L.sexp) > 0.01f) v = ref_powf(v, L.sexp);
            cut = v;
        }
    }
    if (L.has_att && L.type != 2) {
        const float r = dist / L.radius;
        const float a = 1.0f / (1.0f + L.c + L.l * r + L.q * r * r);
        att = a > 0.0f ? a : 0.0f;
    }
    return inf * att * cut;
}

void direct_lumel(const Scene& S, Vec p, Vec n, float* col, uint64_t& rays) {
    for (const Light& L : S.lights) {
        float r = 0.0f, g = 0.0f, b = 0.0f;
        if (!L.verts.empty()) {  // DirectLight::lightFromPointSet
            for (const Vec& v : L.verts) {
                const float f = influence(S, L, mk(v.x + L.pos.x, v.y + L.pos.y, v.z + L.pos.z), p, n, rays);
                if (f > 0.0f) {
                    r += L.color.x * L.intensity * f;
                    g += L.color.y * L.intensity * f;
                    b += L.color.z * L.intensity * f;
                }
            }
            const float cnt = (float)L.verts.size();
            r = r / cnt; g = g / cnt; b = b / cnt;
        } else {  // DirectLight::lightFromPoint
            const float f = influence(S, L, L.pos, p, n, rays);
            if (f > 0.0f) {
                r = L.color.x * L.intensity * f;
                g = L.color.y * L.intensity * f;
                b = L.color.z * L.intensity * f;
            }
        }
        col[0] += r; col[1] += g; col[2] += b;
    }
}

template <class F>
void parallel_for(uint32_t n, int threads, F f) {
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    if (threads == 1 || n < 64) {
        for (uint32_t i = 0; i < n; i++) f(i, 0);
        return;
    }
    std::atomic<uint32_t> next(0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) {
        pool.emplace_back([&, t]() {
            for (;;) {
                const uint32_t b = next.fetch_add(64);
                if (b >= n) break;
                const uint32_t e = std::min(n, b + 64);
                for (uint32_t i = b; i < e; i++) f(i, t);
            }
        });
    }
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

struct or_light {
    int32_t type;
    int32_t casts_shadow;
    float position[3];
    float direction[3];
    float color[3];
    float intensity;
    int32_t has_attenuation;
    float radius, att_constant, att_linear, att_quadratic;
    float spot_cutoff, spot_exponent;
    const float* area_positions;
    const uint32_t* area_indices;
    uint32_t area_vertex_count;
    uint32_t area_index_count;
};

struct or_ao {
    int32_t samples;
    float occluded_fraction, max_distance, exponent;
    uint32_t seed;
    int32_t sampling;
};

void* or_scene_create(void) { return new Scene(); }
void or_scene_destroy(void* s) { delete (Scene*)s; }

// Scene::addMesh + Mesh::transformed (scene/Scene.cpp:98-107, scene/Mesh.cpp:277-297).
int32_t or_scene_add_mesh(void* sp, const float* pos, const float* nrm, const float* uv, uint32_t nv,
                          const uint32_t* idx, uint32_t ni, const float* M, const float* albedo) {
    Scene& S = *(Scene*)sp;
    std::vector<Vec> P(nv), N(nv);
    for (uint32_t i = 0; i < nv; i++) {
        Vec p = mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        Vec n = mk(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
        if (M) {
            Vec tp = mk(p.x * M[0] + p.y * M[4] + p.z * M[8] + M[12], p.x * M[1] + p.y * M[5] + p.z * M[9] + M[13],
                        p.x * M[2] + p.y * M[6] + p.z * M[10] + M[14]);
            Vec tn = mk(n.x * M[0] + n.y * M[4] + n.z * M[8], n.x * M[1] + n.y * M[5] + n.z * M[9],
                        n.x * M[2] + n.y * M[6] + n.z * M[10]);
            normalize3(tn);
            p = tp;
            n = tn;
        }
        P[i] = p;
        N[i] = n;
    }
    const uint32_t mesh = (uint32_t)S.mpos.size();
    const uint32_t first = (uint32_t)(S.tri.size() / 9);
    for (uint32_t f = 0; f < ni / 3; f++) {
        for (int k = 0; k < 3; k++) {
            const uint32_t v = idx[3 * f + k];
            S.tri.push_back(P[v].x); S.tri.push_back(P[v].y); S.tri.push_back(P[v].z);
            S.tuv.push_back(uv[2 * v]); S.tuv.push_back(uv[2 * v + 1]);
            S.tnrm.push_back(N[v].x); S.tnrm.push_back(N[v].y); S.tnrm.push_back(N[v].z);
        }
        for (int k = 0; k < 4; k++) S.talb.push_back(albedo ? albedo[k] : 1.0f);
        S.ttex.push_back(-1);
        for (int k = 0; k < 6; k++) S.tduv.push_back(0.0f);
        S.tmesh.push_back(mesh);
        S.tface.push_back(f);
    }
    S.mpos.push_back(P);
    S.mnrm.push_back(N);
    S.muv.push_back(std::vector<float>(uv, uv + 2 * (size_t)nv));
    S.midx.push_back(std::vector<uint32_t>(idx, idx + ni));
    S.mfirst.push_back(first);
    return (int32_t)mesh;
}

// Texture::create (scene/Material.cpp:78-82): a copy of w*h*ch bytes; returns its index.
int32_t or_scene_add_texture(void* sp, int32_t w, int32_t h, int32_t ch, const uint8_t* px) {
    Scene& S = *(Scene*)sp;
    Scene::Tex t{w, h, ch, std::vector<uint8_t>(px, px + (size_t)w * h * ch)};
    S.textures.push_back(std::move(t));
    return (int32_t)S.textures.size() - 1;
}

// Scene::addMesh with TexturedMaterial(texture, colour): colour rgb (alpha 1), diffuse uvs
// (2 floats per vertex).
int32_t or_scene_add_mesh_textured(void* sp, const float* pos, const float* nrm, const float* uv, const float* duv,
                                   uint32_t nv, const uint32_t* idx, uint32_t ni, const float* M, const float* color,
                                   int32_t tex) {
    Scene& S = *(Scene*)sp;
    const float c4[4] = {color[0], color[1], color[2], 1.0f};
    const uint32_t first = (uint32_t)S.ttex.size();
    const int32_t mesh = or_scene_add_mesh(sp, pos, nrm, uv, nv, idx, ni, M, c4);
    for (uint32_t f = 0; f < ni / 3; f++) {
        S.ttex[first + f] = tex;
        for (int k = 0; k < 3; k++) {
            const uint32_t v = idx[3 * f + k];
            S.tduv[6 * (size_t)(first + f) + 2 * k] = duv[2 * v];
            S.tduv[6 * (size_t)(first + f) + 2 * k + 1] = duv[2 * v + 1];
        }
    }
    return mesh;
}

// Face::colorAt of triangle `tri` at Embree barycentrics (u, v), and Texture::colorAt of texture
// `tex` at (u, v) (for the pins).
void or_face_color(void* sp, uint32_t tri, float u, float v, float* out) { face_color(*(Scene*)sp, tri, u, v, out); }
void or_texture_color(void* sp, int32_t tex, float u, float v, float* out) {
    texture_color(((Scene*)sp)->textures[tex], u, v, out);
}

int32_t or_scene_add_light(void* sp, const or_light* d) {
    Scene& S = *(Scene*)sp;
    Light L;
    L.type = d->type;
    L.shadow = d->casts_shadow;
    L.pos = mk(d->position[0], d->position[1], d->position[2]);
    L.dir = mk(d->direction[0], d->direction[1], d->direction[2]);
    L.color = mk(d->color[0], d->color[1], d->color[2]);
    L.intensity = d->intensity;
    L.has_att = d->has_attenuation;
    L.radius = d->radius; L.c = d->att_constant; L.l = d->att_linear; L.q = d->att_quadratic;
    L.cutoff = d->spot_cutoff; L.sexp = d->spot_exponent;
    if (d->type == 3) {
        // Triangle(face) = (face.vertex(0), vertex(1), vertex(2)) = (a, c, b); centroid = sum / 3.
        for (uint32_t f = 0; f < d->area_index_count / 3; f++) {
            const float* P = d->area_positions;
            const uint32_t ia = d->area_indices[3 * f], ib = d->area_indices[3 * f + 1], ic = d->area_indices[3 * f + 2];
            Vec a = mk(P[3 * ia], P[3 * ia + 1], P[3 * ia + 2]);
            Vec b = mk(P[3 * ib], P[3 * ib + 1], P[3 * ib + 2]);
            Vec c = mk(P[3 * ic], P[3 * ic + 1], P[3 * ic + 2]);
            L.verts.push_back(divs(add(add(a, c), b), 3.0f));
        }
    }
    S.lights.push_back(L);
    return (int32_t)S.lights.size() - 1;
}

int32_t or_scene_build(void* sp, int32_t use_bvh) {
    Scene& S = *(Scene*)sp;
    S.use_bvh = use_bvh != 0;
    if (S.use_bvh) build_bvh(S);
    std::vector<bool> cut(S.textures.size(), false);
    for (size_t t = 0; t < S.textures.size(); t++)
        if (S.textures[t].ch == 4)
            for (size_t i = 3; i < S.textures[t].px.size(); i += 4)
                if (S.textures[t].px[i] / 255.0f <= 0.1f) cut[t] = true;
    S.alpha = false;
    for (size_t i = 0; i < S.ttex.size(); i++)
        if (S.ttex[i] >= 0 ? cut[S.ttex[i]] : S.talb[4 * i + 3] <= 0.1f) S.alpha = true;
    return 0;
}

// Lightmap::initializeLumels for every face of `mesh` into a w*h grid (in/out arrays).
int32_t or_raster(void* sp, int32_t mesh, int32_t w, int32_t h, uint8_t* valid, uint32_t* tri,
                  float* pos, float* nrm) {
    const Scene& S = *(Scene*)sp;
    const std::vector<Vec>& P = S.mpos[mesh];
    const std::vector<Vec>& N = S.mnrm[mesh];
    const std::vector<float>& UV = S.muv[mesh];
    const std::vector<uint32_t>& I = S.midx[mesh];
    for (uint32_t f = 0; f < I.size() / 3; f++) {
        const uint32_t a = I[3 * f], b = I[3 * f + 1], c = I[3 * f + 2];
        const float ua = UV[2 * a], va = UV[2 * a + 1], ub = UV[2 * b], vb = UV[2 * b + 1], uc = UV[2 * c], vc = UV[2 * c + 1];
        // Lightmap::rect of Face::uvRect: int(min * size), int(max * size).
        const int x1 = (int)(std::min(ua, std::min(ub, uc)) * w), x2 = (int)(std::max(ua, std::max(ub, uc)) * w);
        const int y1 = (int)(std::min(va, std::min(vb, vc)) * h), y2 = (int)(std::max(va, std::max(vb, vc)) * h);
        for (int y = std::max(0, y1); y <= std::min(y2, h - 1); y++) {
            for (int x = std::max(0, x1); x <= std::min(x2, w - 1); x++) {
                const float px = (x + 0.5f) / float(w), py = (y + 0.5f) / float(h);
                // Face::isUvInside: v0 = c-a, v1 = b-a, v2 = uv-a.
                const float e0x = uc - ua, e0y = vc - va, e1x = ub - ua, e1y = vb - va, e2x = px - ua, e2y = py - va;
                const float d00 = e0x * e0x + e0y * e0y, d01 = e0x * e1x + e0y * e1y, d02 = e0x * e2x + e0y * e2y;
                const float d11 = e1x * e1x + e1y * e1y, d12 = e1x * e2x + e1y * e2y;
                const float inv = 1.0f / (d00 * d11 - d01 * d01);
                const float bu = (d11 * d02 - d01 * d12) * inv;
                const float bv = (d00 * d12 - d01 * d02) * inv;
                if (!(bu >= 0.0f && bv >= 0.0f && bu + bv < 1.0f)) continue;
                const size_t t = (size_t)y * w + x;
                valid[t] = 1;
                tri[t] = S.mfirst[mesh] + f;
                // Face::positionAt / normalAt: a + (b-a)*y + (c-a)*x with (x, y) = (bu, bv).
                const Vec A = P[a], B = P[b], C = P[c];
                pos[3 * t] = A.x + (B.x - A.x) * bv + (C.x - A.x) * bu;
                pos[3 * t + 1] = A.y + (B.y - A.y) * bv + (C.y - A.y) * bu;
                pos[3 * t + 2] = A.z + (B.z - A.z) * bv + (C.z - A.z) * bu;
                const Vec NA = N[a], NB = N[b], NC = N[c];
                nrm[3 * t] = NA.x + (NB.x - NA.x) * bv + (NC.x - NA.x) * bu;
                nrm[3 * t + 1] = NA.y + (NB.y - NA.y) * bv + (NC.y - NA.y) * bu;
                nrm[3 * t + 2] = NA.z + (NB.z - NA.z) * bv + (NC.z - NA.z) * bu;
            }
        }
    }
    return 0;
}

// DirectLight::bakeLumel over n lumels; color is in/out (3 per lumel).
int32_t or_bake_direct(void* sp, uint32_t n, const float* pos, const float* nrm, float* color,
                       int32_t threads, uint64_t* rays_out) {
    const Scene& S = *(Scene*)sp;
    std::vector<uint64_t> rays(256, 0);
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = std::min(threads, 255);
    parallel_for(n, threads, [&](uint32_t i, int t) {
        direct_lumel(S, mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]), mk(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]),
                     &color[3 * i], rays[t]);
    });
    uint64_t tot = 0;
    for (uint64_t r : rays) tot += r;
    if (rays_out) *rays_out = tot;
    return 0;
}

// AmbientOcclusion::bakeLumel over n lumels (texel ids key the sample streams).
int32_t or_bake_ao(void* sp, uint32_t n, const float* pos, const float* nrm, const uint32_t* texel,
                   uint32_t lm_seed, const or_ao* a, float* color, uint32_t* occluded_out,
                   int32_t threads, uint64_t* rays_out) {
    const Scene& S = *(Scene*)sp;
    parallel_for(n, threads, [&](uint32_t i, int) {
        const Vec p = mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        const Vec nn = mk(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
        int occ = 0;
        HitRec h;
        for (int k = 0; k < a->samples; k++) {
            Rng r = stream(a->seed, lm_seed, texel[i], 1u, (uint32_t)k);
            const Vec d = hemi(r, nn, a->sampling);
            if (trace(S, make_seg(p, add(p, scale(d, a->max_distance))), true, h, nullptr)) occ++;
        }
        float value = 1.0f - occ / (a->samples * a->occluded_fraction);
        if (fabsf(1.0f - a->exponent) > 0.01f) value = ref_powf(value, a->exponent);
        color[3 * i] *= value;
        color[3 * i + 1] *= value;
        color[3 * i + 2] *= value;
        if (occluded_out) occluded_out[i] = (uint32_t)occ;
    });
    if (rays_out) *rays_out = (uint64_t)n * (uint64_t)a->samples;
    return 0;
}

// rt::ITracer::traceSegment over n segments. mode 0 = any hit (flags), 1 = closest hit.
// Counters (node fetches, triangle tests) feed the per-ray traffic model.
int32_t or_trace(void* sp, uint32_t n, const float* st, const float* en, int32_t mode, uint8_t* flags,
                 float* t, int32_t* tri, float* u, float* v, uint64_t* node_visits, uint64_t* tri_tests,
                 int32_t threads) {
    const Scene& S = *(Scene*)sp;
    std::vector<Counters> ctr(256);
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = std::min(threads, 255);
    parallel_for(n, threads, [&](uint32_t i, int th) {
        HitRec h;
        const bool hit = trace(S, make_seg(mk(st[3 * i], st[3 * i + 1], st[3 * i + 2]), mk(en[3 * i], en[3 * i + 1], en[3 * i + 2])),
                               mode == 0, h, &ctr[th]);
        if (flags) flags[i] = hit ? 1 : 0;
        if (t) t[i] = hit ? h.t : 0.0f;
        if (tri) tri[i] = hit ? h.tri : -1;
        if (u) u[i] = hit ? h.u : 0.0f;
        if (v) v[i] = hit ? h.v : 0.0f;
    });
    uint64_t nv = 0, nt = 0;
    for (auto& c : ctr) { nv += c.nodes; nt += c.tris; }
    if (node_visits) *node_visits = nv;
    if (tri_tests) *tri_tests = nt;
    return 0;
}

// Scene-order triangle -> mesh, face.
int32_t or_tri_info(void* sp, int32_t tri, int32_t* mesh, int32_t* face) {
    const Scene& S = *(Scene*)sp;
    if (tri < 0 || (size_t)tri >= S.tmesh.size()) return -1;
    *mesh = (int32_t)S.tmesh[tri];
    *face = (int32_t)S.tface[tri];
    return 0;
}

// Raw PCG32 outputs for pinning against published vectors.
void or_pcg32(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* out) {
    Rng r = rng_seed(initstate, initseq);
    for (uint32_t i = 0; i < n; i++) out[i] = rng_u32(r);
}

// Sample-stream outputs and hemisphere directions for the sampler checks.
void or_stream_u32(uint32_t seed, uint32_t lm_seed, uint32_t texel, uint32_t tech, uint32_t sample, uint32_t n,
                   uint32_t* out) {
    Rng r = stream(seed, lm_seed, texel, tech, sample);
    for (uint32_t i = 0; i < n; i++) out[i] = rng_u32(r);
}

void or_hemisphere(uint32_t seed, uint32_t lm_seed, uint32_t texel, uint32_t tech, uint32_t count, int32_t mode,
                   const float* normal, float* out) {
    const Vec n = mk(normal[0], normal[1], normal[2]);
    for (uint32_t k = 0; k < count; k++) {
        Rng r = stream(seed, lm_seed, texel, tech, k);
        const Vec d = hemi(r, n, mode);
        out[3 * k] = d.x; out[3 * k + 1] = d.y; out[3 * k + 2] = d.z;
    }
}

uint64_t or_splitmix64(uint64_t z) { return mix64(z); }

// ------------------------------------------------------------------ photons + final gather
struct or_indirect {
    int32_t photon_pass_count, photon_bounce_count;
    float photon_energy_threshold, photon_max_distance;
    int32_t final_gather_samples;
    float final_gather_distance;
    int32_t final_gather_radius;
    float sky_color[3], ambient_color[3];
    uint32_t seed;
    int32_t sampling, photons_per_light, termination;
    int32_t gather_mode, knn_k;
    float knn_radius;
};

// Photonmaps are passed concatenated: map k covers keys [base_k, base_k + w_k*h_k) of the
// valid/color/photons/gathered arrays (color and gathered hold 3 floats per texel).
struct PMaps {
    int32_t n;
    const int32_t *w, *h;
    std::vector<uint32_t> base;
    const uint8_t* valid;
    float* color;
    uint32_t* photons;
    float* gathered;
};

static PMaps make_pmaps(int32_t n, const int32_t* w, const int32_t* h, const uint8_t* valid, float* color,
                        uint32_t* photons, float* gathered) {
    PMaps M{n, w, h, {}, valid, color, photons, gathered};
    uint32_t acc = 0;
    for (int k = 0; k < n; k++) {
        M.base.push_back(acc);
        acc += (uint32_t)(w[k] * h[k]);
    }
    return M;
}

// Face::normalAt / uvAt at Embree's (u, v): a + (b - a) * v + (c - a) * u.
static inline float flerp(float a, float b, float c, float u, float v) { return a + (b - a) * v + (c - a) * u; }

static inline int lumel_xy(float uv, int size) {
    int x = (int)(uv * (float)(size - 1));
    return x < 0 ? 0 : (x >= size ? size - 1 : x);
}

struct PRec {
    uint32_t key;
    float r, g, b;
    float x, y, z;
};

// k-nearest gather (gather_mode 1): for every valid photonmap texel, the k stored photons nearest
// to its surface point (squared distance ex*ex + ey*ey + ez*ez, e = photon - point) within the
// radius, ordered by (squared distance, emission slot); their power summed in that order and
// divided by their number, as Photonmap::gather normalises (Lightmap.cpp:407-412). A uniform grid
// of cell 1.01 * radius holds every photon within the radius in the 27 cells around a point.
static void knn_gather(const std::vector<PRec>& rec, PMaps& M, const float* pm_pos, int k, float radius,
                       int threads) {
    std::vector<uint32_t> ids;
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    for (uint32_t i = 0; i < (uint32_t)rec.size(); i++) {
        if (rec[i].key == 0xFFFFFFFFu) continue;
        ids.push_back(i);
        lo[0] = std::min(lo[0], rec[i].x); lo[1] = std::min(lo[1], rec[i].y); lo[2] = std::min(lo[2], rec[i].z);
    }
    const float cell = radius * 1.01f, r2 = radius * radius;
    auto cell_of = [&](float x, float y, float z, int64_t* c) {
        c[0] = (int64_t)floor((double)(x - lo[0]) / cell);
        c[1] = (int64_t)floor((double)(y - lo[1]) / cell);
        c[2] = (int64_t)floor((double)(z - lo[2]) / cell);
    };
    auto pack = [](int64_t a, int64_t b, int64_t c) {
        return ((uint64_t)(a + (1 << 20)) << 42) | ((uint64_t)(b + (1 << 20)) << 21) | (uint64_t)(c + (1 << 20));
    };
    std::unordered_map<uint64_t, std::vector<uint32_t>> grid;
    for (uint32_t i : ids) {
        int64_t c[3];
        cell_of(rec[i].x, rec[i].y, rec[i].z, c);
        grid[pack(c[0], c[1], c[2])].push_back(i);
    }
    uint32_t total = 0;
    for (int m = 0; m < M.n; m++) total += (uint32_t)(M.w[m] * M.h[m]);
    parallel_for(total, threads, [&](uint32_t t, int) {
        float* out = &M.gathered[3 * (size_t)t];
        out[0] = out[1] = out[2] = 0.0f;
        if (!M.valid[t] || ids.empty()) return;
        const float qx = pm_pos[3 * (size_t)t], qy = pm_pos[3 * (size_t)t + 1], qz = pm_pos[3 * (size_t)t + 2];
        int64_t c[3];
        cell_of(qx, qy, qz, c);
        std::vector<std::pair<float, uint32_t>> cand;
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    auto it = grid.find(pack(c[0] + dx, c[1] + dy, c[2] + dz));
                    if (it == grid.end()) continue;
                    for (uint32_t i : it->second) {
                        const float ex = rec[i].x - qx, ey = rec[i].y - qy, ez = rec[i].z - qz;
                        const float d2 = ex * ex + ey * ey + ez * ez;
                        if (d2 <= r2) cand.push_back({d2, i});
                    }
                }
        std::sort(cand.begin(), cand.end());
        const size_t n = std::min<size_t>(cand.size(), (size_t)k);
        if (n == 0) return;
        float sr = 0.0f, sg = 0.0f, sb = 0.0f;
        for (size_t j = 0; j < n; j++) {
            sr += rec[cand[j].second].r;
            sg += rec[cand[j].second].g;
            sb += rec[cand[j].second].b;
        }
        out[0] = sr / (float)n;
        out[1] = sg / (float)n;
        out[2] = sb / (float)n;
    });
}

// Photons::emit / trace / store (baker/Photons.cpp:50-144) with PhotonEmitter::emit,
// DirectionalPhotonEmitter::emit (scene/Light.cpp:339-361) and LightSpotCutoff, then
// Photonmap::gather (Lightmap.cpp:376-413) on every map.
int32_t or_emit_photons(void* sp, const or_indirect* s, const int32_t* mesh_pm, int32_t npm, const int32_t* pw,
                        const int32_t* ph, const uint8_t* valid, const float* pm_pos, float* color,
                        uint32_t* photons, float* gathered, int32_t threads, uint64_t* rays_out) {
    Scene& S = *(Scene*)sp;
    PMaps M = make_pmaps(npm, pw, ph, valid, color, photons, gathered);
    const int nl = (int)S.lights.size();
    std::vector<uint64_t> first(nl);
    uint64_t per_pass = 0;
    for (int l = 0; l < nl; l++) {
        first[l] = per_pass;
        int64_t c = s->photons_per_light > 0 ? s->photons_per_light : (int64_t)(int32_t)(S.lights[l].intensity * 25000.0f);
        per_pass += (uint64_t)std::max<int64_t>(c, 0);
    }
    const uint64_t total = per_pass * (uint64_t)s->photon_pass_count;
    const int D = s->photon_bounce_count;
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (size_t i = 0; i < S.tri.size(); i += 3)
        for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], S.tri[i + k]); hi[k] = std::max(hi[k], S.tri[i + k]); }
    std::vector<PRec> rec((size_t)total * (D + 1), PRec{0xFFFFFFFFu, 0, 0, 0, 0, 0, 0});
    std::vector<uint64_t> rays(256, 0);
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = std::min(threads, 255);
    parallel_for((uint32_t)total, threads, [&](uint32_t p, int th) {
        const uint64_t r = p % per_pass;
        int l = 0;
        while (l + 1 < nl && first[l + 1] <= r) l++;
        const Light& L = S.lights[l];
        Rng g = stream(s->seed, 0xFFFFFFFFu, p, 3u, 0u);
        Vec pos, dir;
        if (L.type == 2) {
            const float tx = rng_f(g), ty = rng_f(g), tz = rng_f(g);
            const Vec q = mk(lo[0] * (1.0f - tx) + hi[0] * tx, lo[1] * (1.0f - ty) + hi[1] * ty, lo[2] * (1.0f - tz) + hi[2] * tz);
            const float d = -dot3(L.dir, L.pos);
            const float dist = dot3(L.dir, q) + d;
            pos = sub(sub(q, scale(L.dir, dist)), scale(L.dir, 5.0f));
            dir = L.dir;
        } else {
            pos = L.pos;
            dir = hemi_uniform_sphere(g);
        }
        float cut = 1.0f;
        if (L.type == 1) {
            float v = dot3(dir, L.dir);
            if (v <= L.cutoff) cut = 0.0f;
            else {
                v = (1.0f - (1.0f - v) * 1.0f / (1.0f - L.cutoff));
                if (fabsf(1.0f - L.sexp) > 0.01f) v = ref_powf(v, L.sexp);
                cut = v;
            }
        }
        if (!(cut > 0.0f)) return;
        float cr = L.color.x * L.intensity * cut, cg = L.color.y * L.intensity * cut, cb = L.color.z * L.intensity * cut;
        for (int depth = 0; depth <= D; depth++) {
            if (s->termination == 0 && 0.2126f * cr + 0.7152f * cg + 0.0722f * cb < s->photon_energy_threshold) break;
            const Seg sg = make_seg(pos, add(pos, scale(dir, s->photon_max_distance)));
            HitRec h;
            rays[th]++;
            if (!trace(S, sg, false, h, nullptr)) break;
            const Vec hp = add(sg.o, scale(sg.d, h.t));
            const float* N = &S.tnrm[9 * (size_t)h.tri];
            const float* UV = &S.tuv[6 * (size_t)h.tri];
            float A[4];
            face_color(S, (uint32_t)h.tri, h.u, h.v, A);
            const Vec nh = mk(flerp(N[0], N[3], N[6], h.u, h.v), flerp(N[1], N[4], N[7], h.u, h.v), flerp(N[2], N[5], N[8], h.u, h.v));
            const float ux = flerp(UV[0], UV[2], UV[4], h.u, h.v), uy = flerp(UV[1], UV[3], UV[5], h.u, h.v);
            float lam = dot3(neg(dir), nh);
            lam = lam < 0.0f ? 0.0f : lam;
            const float inf = lam * 1.0f;
            const float hr = cr * A[0] * inf, hg = cg * A[1] * inf, hb = cb * A[2] * inf;
            const int32_t pm = mesh_pm[S.tmesh[h.tri]];
            if (pm >= 0) {
                const int x = lumel_xy(ux, M.w[pm]), y = lumel_xy(uy, M.h[pm]);
                const uint32_t t = (uint32_t)(y * M.w[pm] + x);
                if (M.valid[M.base[pm] + t])
                    rec[(size_t)p * (D + 1) + depth] = PRec{M.base[pm] + t, hr, hg, hb, hp.x, hp.y, hp.z};
            }
            if (s->termination == 1) {
                const float q = (A[0] + A[1] + A[2]) / 3.0f;
                if (!(rng_f(g) < q)) break;
                cr = hr / q; cg = hg / q; cb = hb / q;
            } else {
                cr = hr; cg = hg; cb = hb;
            }
            pos = hp;
            dir = hemi_uniform(g, nh);
        }
    });
    // Sequential store in emission order (Photons::store: m_color += color; m_photons++).
    for (const PRec& q : rec) {
        if (q.key == 0xFFFFFFFFu) continue;
        M.color[3 * (size_t)q.key] += q.r;
        M.color[3 * (size_t)q.key + 1] += q.g;
        M.color[3 * (size_t)q.key + 2] += q.b;
        M.photons[q.key]++;
    }
    S.last_key.resize(rec.size());
    S.last_power.resize(3 * rec.size());
    S.last_pos.resize(3 * rec.size());
    for (size_t i = 0; i < rec.size(); i++) {
        S.last_key[i] = rec[i].key;
        S.last_power[3 * i] = rec[i].r; S.last_power[3 * i + 1] = rec[i].g; S.last_power[3 * i + 2] = rec[i].b;
        S.last_pos[3 * i] = rec[i].x; S.last_pos[3 * i + 1] = rec[i].y; S.last_pos[3 * i + 2] = rec[i].z;
    }
    const int R = s->final_gather_radius;
    if (s->gather_mode == 1) {
        knn_gather(rec, M, pm_pos, s->knn_k, s->knn_radius, threads);
        uint64_t tot = 0;
        for (uint64_t x : rays) tot += x;
        if (rays_out) *rays_out = tot;
        return 0;
    }
    for (int k = 0; k < npm; k++) {
        const int W = M.w[k], H = M.h[k];
        for (int y = 0; y < H; y++) {
            for (int x = 0; x < W; x++) {
                float sr = 0.0f, sg_ = 0.0f, sb = 0.0f;
                int cnt = 0;
                for (int j = y - R; j <= y + R; j++) {
                    for (int i = x - R; i <= x + R; i++) {
                        if (i < 0 || j < 0 || i >= W || j >= H) continue;
                        const float dd = sqrtf((float)((x - i) * (x - i) + (y - j) * (y - j)));
                        if (dd > (float)R) continue;
                        const size_t t = M.base[k] + (size_t)j * W + i;
                        sr += M.color[3 * t]; sg_ += M.color[3 * t + 1]; sb += M.color[3 * t + 2];
                        cnt += (int)M.photons[t];
                    }
                }
                const size_t o = M.base[k] + (size_t)y * W + x;
                if (cnt == 0) {
                    M.gathered[3 * o] = M.gathered[3 * o + 1] = M.gathered[3 * o + 2] = 0.0f;
                } else {
                    M.gathered[3 * o] = sr / (float)cnt;
                    M.gathered[3 * o + 1] = sg_ / (float)cnt;
                    M.gathered[3 * o + 2] = sb / (float)cnt;
                }
            }
        }
    }
    uint64_t tot = 0;
    for (uint64_t x : rays) tot += x;
    if (rays_out) *rays_out = tot;
    return 0;
}

// Lightmap::expand (Lightmap.cpp:167-193) on a full grid (3 floats per texel), as written: raster
// order over interior texels, each valid one writing its colour to invalid diagonal neighbours.
int32_t or_expand(int32_t w, int32_t h, const uint8_t* valid, float* rgb) {
    for (int y = 1; y < h - 1; y++) {
        for (int x = 1; x < w - 1; x++) {
            const size_t c = (size_t)y * w + x;
            if (!valid[c]) continue;
            const int nx[4] = {x - 1, x - 1, x + 1, x + 1}, ny[4] = {y - 1, y + 1, y - 1, y + 1};
            for (int k = 0; k < 4; k++) {
                const size_t t = (size_t)ny[k] * w + nx[k];
                if (valid[t]) continue;
                rgb[3 * t] = rgb[3 * c];
                rgb[3 * t + 1] = rgb[3 * c + 1];
                rgb[3 * t + 2] = rgb[3 * c + 2];
            }
        }
    }
    return 0;
}

// Lightmap::blur (Lightmap.cpp:196-220): in place, raster order.
int32_t or_blur(int32_t w, int32_t h, const uint8_t* valid, float* rgb) {
    for (int y = 1; y < h - 1; y++) {
        for (int x = 1; x < w - 1; x++) {
            float r = 0.0f, g = 0.0f, b = 0.0f;
            int count = 0;
            for (int j = y - 1; j <= y + 1; j++) {
                for (int i = x - 1; i <= x + 1; i++) {
                    const size_t t = (size_t)j * w + i;
                    if (!valid[t]) continue;
                    r += rgb[3 * t];
                    g += rgb[3 * t + 1];
                    b += rgb[3 * t + 2];
                    count++;
                }
            }
            if (count) {
                const float s = 1.0f / (float)count;
                const size_t c = (size_t)y * w + x;
                rgb[3 * c] = r * s;
                rgb[3 * c + 1] = g * s;
                rgb[3 * c + 2] = b * s;
            }
        }
    }
    return 0;
}

// Rgb::doubleLdr (format 1) / Rgb::rgbm (format 2) per texel (Types.h:216-234) with
// min2(a, b) = a < b ? a : b and max2(a, b) = a > b ? a : b; the byte cast truncates values in
// [0, 255] and gives 0 where the reference's cast is undefined (negative, NaN).
static inline uint8_t byte_of(float v) {
    const float m = 255.0f < v ? 255.0f : v;
    return m >= 1.0f ? (uint8_t)(int)m : (uint8_t)0;
}

int32_t or_encode(int32_t w, int32_t h, const float* rgb, int32_t format, uint8_t* out) {
    const size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; i++) {
        const float r = rgb[3 * i], g = rgb[3 * i + 1], b = rgb[3 * i + 2];
        if (format == 1) {
            out[3 * i] = byte_of(r * 0.5f * 255.0f);
            out[3 * i + 1] = byte_of(g * 0.5f * 255.0f);
            out[3 * i + 2] = byte_of(b * 0.5f * 255.0f);
        } else {
            const float gb = g > b ? g : b;
            const float mx = r > gb ? r : gb;
            const float m = 255.0f / mx;
            out[4 * i] = byte_of(r / m * 255.0f);
            out[4 * i + 1] = byte_of(g / m * 255.0f);
            out[4 * i + 2] = byte_of(b / m * 255.0f);
            out[4 * i + 3] = byte_of(m * 255.0f);
        }
    }
    return 0;
}

// Records of the last or_emit_photons call in slot order (photon-major, bounce-minor): count, then
// key / power (3) / hit point (3) when the arrays are given.
int32_t or_photon_records(void* sp, uint64_t* n, uint32_t* key, float* power, float* pos) {
    const Scene& S = *(Scene*)sp;
    *n = S.last_key.size();
    if (key) memcpy(key, S.last_key.data(), S.last_key.size() * 4);
    if (power) memcpy(power, S.last_power.data(), S.last_power.size() * 4);
    if (pos) memcpy(pos, S.last_pos.data(), S.last_pos.size() * 4);
    return 0;
}

// IndirectLight::bakeLumel (baker/IndirectLight.cpp:47-75) over n lumels; color is in/out.
int32_t or_final_gather(void* sp, uint32_t n, const float* pos, const float* nrm, const uint32_t* texel,
                        uint32_t lm_seed, const or_indirect* s, const int32_t* mesh_pm, int32_t npm, const int32_t* pw,
                        const int32_t* ph, const float* gathered, float* color, int32_t threads, uint64_t* rays_out) {
    const Scene& S = *(Scene*)sp;
    PMaps M = make_pmaps(npm, pw, ph, nullptr, nullptr, nullptr, const_cast<float*>(gathered));
    parallel_for(n, threads, [&](uint32_t i, int) {
        const Vec p = mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        const Vec nn = mk(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
        float gr = 0.0f, gg = 0.0f, gb = 0.0f;
        for (int k = 0; k < s->final_gather_samples; k++) {
            Rng g = stream(s->seed, lm_seed, texel[i], 2u, (uint32_t)k);
            const Vec dir = hemi(g, nn, s->sampling);
            float inf = dot3(nn, dir);
            inf = inf > 0.0f ? inf : 0.0f;
            HitRec h;
            if (!trace(S, make_seg(p, add(p, scale(dir, s->final_gather_distance))), false, h, nullptr)) {
                gr += s->sky_color[0] * inf + s->ambient_color[0];
                gg += s->sky_color[1] * inf + s->ambient_color[1];
                gb += s->sky_color[2] * inf + s->ambient_color[2];
                continue;
            }
            const float* N = &S.tnrm[9 * (size_t)h.tri];
            const Vec nh = mk(flerp(N[0], N[3], N[6], h.u, h.v), flerp(N[1], N[4], N[7], h.u, h.v), flerp(N[2], N[5], N[8], h.u, h.v));
            if (dot3(dir, nh) >= 0.0f) continue;
            const int32_t pm = mesh_pm[S.tmesh[h.tri]];
            if (pm < 0) continue;
            const float* UV = &S.tuv[6 * (size_t)h.tri];
            const float ux = flerp(UV[0], UV[2], UV[4], h.u, h.v), uy = flerp(UV[1], UV[3], UV[5], h.u, h.v);
            const int x = lumel_xy(ux, M.w[pm]), y = lumel_xy(uy, M.h[pm]);
            const size_t t = M.base[pm] + (size_t)y * M.w[pm] + x;
            gr += M.gathered[3 * t] * inf + s->ambient_color[0];
            gg += M.gathered[3 * t + 1] * inf + s->ambient_color[1];
            gb += M.gathered[3 * t + 2] * inf + s->ambient_color[2];
        }
        const float fs = (float)s->final_gather_samples;
        color[3 * i] += gr / fs;
        color[3 * i + 1] += gg / fs;
        color[3 * i + 2] += gb / fs;
    });
    if (rays_out) *rays_out = (uint64_t)n * (uint64_t)s->final_gather_samples;
    return 0;
}

// Form-factor rows by hemisphere rays (the estimator the GPU uses in place of the pairwise
// RadiosityBuilder::distanceFormFactor, radiosity/RadiosityBuilder.cpp:142-176): for patch i,
// `rays` segments from uniform points o + s*U + t*V with cosine-weighted directions about n,
// counting closest hits per patch. patches: 12 floats each (o, U, V, n); out: rows x n counts.
int32_t or_form_factors(void* sp, const float* patches, const uint32_t* tri_patch, uint32_t n, uint32_t rays,
                        uint32_t seed, float max_distance, uint32_t row0, uint32_t rows, uint16_t* out,
                        int32_t threads) {
    const Scene& S = *(Scene*)sp;
    parallel_for(rows, threads, [&](uint32_t rr, int) {
        const uint32_t i = row0 + rr;
        const float* P = patches + 12 * (size_t)i;
        const Vec o = mk(P[0], P[1], P[2]), U = mk(P[3], P[4], P[5]), V = mk(P[6], P[7], P[8]), N = mk(P[9], P[10], P[11]);
        uint16_t* row = out + (size_t)rr * n;
        for (uint32_t j = 0; j < n; j++) row[j] = 0;
        HitRec h;
        for (uint32_t k = 0; k < rays; k++) {
            Rng g = stream(seed, 0xFFFFFFFEu, i, 4u, k);
            const float s = rng_f(g), t = rng_f(g);
            const Vec p = add(add(o, scale(U, s)), scale(V, t));
            const Vec d = hemi_cosine(g, N);
            if (!trace(S, make_seg(p, add(p, scale(d, max_distance))), false, h, nullptr)) continue;
            const uint32_t j = tri_patch[h.tri];
            if (j < n) row[j]++;
        }
    });
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ lightmap uv generation
// Mesh::generateUv (scene/Mesh.cpp:229-238): UvGenerator::generate (Types.h:3063-3160) with
// AngularChartifier (Types.h:2427-2595), DCEL::create (Types.h:2380-2423), Face::flatten / normal
// (Types.h:2825-2886), UvGenerator::rotate (Types.h:3031-3060; covarianceMatrix, CovMatrix::converge,
// Types.h:2200-2290) and RectanglePacker::place (Types.h:3163-3326), as written (recursive flood,
// the plain grow-by-0.1 loop).
namespace uvgen {

struct V {
    float p[3], n[3], t0[2], t1[2];
    bool operator<(const V& o) const {  // Vertex::Compare (scene/Mesh.h:54-70)
        for (int i = 0; i < 3; i++) if (p[i] != o.p[i]) return p[i] < o.p[i];
        for (int i = 0; i < 3; i++) if (n[i] != o.n[i]) return n[i] < o.n[i];
        for (int i = 0; i < 2; i++) if (t0[i] != o.t0[i]) return t0[i] < o.t0[i];
        for (int i = 0; i < 2; i++) if (t1[i] != o.t1[i]) return t1[i] < o.t1[i];
        return false;
    }
};

struct Indexer {
    std::map<V, uint32_t> cache;
    std::vector<V> vb;
    std::vector<uint32_t> ib;
    uint32_t operator+=(const V& v) {
        auto it = cache.find(v);
        uint32_t idx;
        if (it != cache.end()) idx = it->second;
        else { idx = (uint32_t)vb.size(); cache[v] = idx; vb.push_back(v); }
        ib.push_back(idx);
        return idx;
    }
};

struct Half { uint32_t face, vertex; Half* next; Half* twin; };

struct Gen {
    const std::vector<V>& vs;
    const std::vector<uint32_t>& is;
    float angle;
    std::map<int, int> chart_by_face;
    std::vector<std::vector<uint32_t>> charts;
    Gen(const std::vector<V>& v, const std::vector<uint32_t>& i, float a) : vs(v), is(i), angle(a) {}

    Vec normal(uint32_t f) const {
        const float* a = vs[is[3 * f]].p; const float* b = vs[is[3 * f + 1]].p; const float* c = vs[is[3 * f + 2]].p;
        const Vec u = mk(b[0] - a[0], b[1] - a[1], b[2] - a[2]), w = mk(c[0] - a[0], c[1] - a[1], c[2] - a[2]);
        Vec n = mk(u.y * w.z - u.z * w.y, u.z * w.x - u.x * w.z, u.x * w.y - u.y * w.x);
        const float len = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
        if (len) { n.x /= len; n.y /= len; n.z /= len; }
        return n;
    }

    void add_to_chart(const Half* e, Vec axis, int index) {
        if (chart_by_face.count((int)e->face)) return;
        const float ang = acosf(dot3(axis, normal(e->face))) * 180.0f / 3.1415926535897932f;
        if (ang > angle) return;
        chart_by_face[(int)e->face] = index;
        if ((int)charts.size() <= index) charts.resize(index + 1);
        charts[index].push_back(e->face);
        const Half* i = e;
        do {
            if (i->twin) add_to_chart(i->twin, axis, index);
            i = i->next;
        } while (i != e);
    }
};

struct R { float x, y, w, h; };

struct Node {
    bool free = true;
    std::unique_ptr<Node> l, r;
    R rc;
    explicit Node(R q) : rc(q) {}
    bool assign(R& pl) {
        if (l || r) return l->assign(pl) ? true : r->assign(pl);
        if ((pl.w > rc.w || pl.h > rc.h) || !free) return false;
        if (pl.w == rc.w && pl.h == rc.h) { pl.x = rc.x; pl.y = rc.y; free = false; return true; }
        const float dw = rc.w - pl.w, dh = rc.h - pl.h;
        if (dw > dh) {
            l.reset(new Node(R{rc.x, rc.y, pl.w, rc.h}));
            r.reset(new Node(R{rc.x + pl.w, rc.y, rc.w - pl.w, rc.h}));
        } else {
            l.reset(new Node(R{rc.x, rc.y, rc.w, pl.h}));
            r.reset(new Node(R{rc.x, rc.y + pl.h, rc.w, rc.h - pl.h}));
        }
        return assign(pl);
    }
};

bool place(std::vector<R>& rects, float w, float h) {
    std::vector<R*> order;
    for (R& q : rects) order.push_back(&q);
    std::sort(order.begin(), order.end(), [](const R* a, const R* b) { return a->w * a->h > b->w * b->h; });
    Node root(R{0.0f, 0.0f, w - 0.0f, h - 0.0f});
    for (R* q : order) if (!root.assign(*q)) return false;
    return true;
}

void bounds(const std::vector<std::pair<float, float>>& uv, float mn[2], float mx[2]) {
    mn[0] = mn[1] = FLT_MAX; mx[0] = mx[1] = -FLT_MAX;
    for (const auto& t : uv) {
        mn[0] = std::min(mn[0], t.first); mx[0] = std::max(mx[0], t.first);
        mn[1] = std::min(mn[1], t.second); mx[1] = std::max(mx[1], t.second);
    }
}

void rotate(std::vector<std::pair<float, float>>& uv) {
    const uint32_t n = (uint32_t)uv.size();
    float sx = 0.0f, sy = 0.0f;
    for (const auto& t : uv) { sx += t.first; sy += t.second; }
    const float mx0 = sx / (float)n, my0 = sy / (float)n;
    std::vector<std::pair<float, float>> c;
    for (const auto& t : uv) c.push_back({t.first - mx0, t.second - my0});
    float cx = 0.0f, cy = 0.0f;
    for (const auto& t : c) { cx += t.first; cy += t.second; }
    const float m[2] = {cx / (float)n, cy / (float)n};
    float M[4];  // M[row * 2 + col]
    for (int row = 0; row < 2; row++)
        for (int col = 0; col < 2; col++) {
            float cov = 0.0f;
            for (uint32_t i = 0; i < n; i++) {
                const float a = (row == 0 ? c[i].first : c[i].second) - m[row];
                const float b = (col == 0 ? c[i].first : c[i].second) - m[col];
                cov += a * b;
            }
            M[row * 2 + col] = cov / (float)n;
        }
    float v[2] = {1.0f, 1.0f};
    for (int it = 0; it < 10; it++) {
        const float r0 = (0.0f + M[0] * v[0]) + M[1] * v[1];
        const float r1 = (0.0f + M[2] * v[0]) + M[3] * v[1];
        v[0] = r0; v[1] = r1;
    }
    float len = 0.0f;
    len += v[0] * v[0];
    len += v[1] * v[1];
    len = sqrtf(len);
    if (len != 0.0f) { v[0] /= len; v[1] /= len; }
    const float e2[2] = {v[1], -v[0]};
    std::vector<std::pair<float, float>> rot;
    for (const auto& t : uv) rot.push_back({v[0] * t.first + v[1] * t.second, e2[0] * t.first + e2[1] * t.second});
    float bmn[2], bmx[2], amn[2], amx[2];
    bounds(uv, bmn, bmx);
    bounds(rot, amn, amx);
    const float bw = bmx[0] - bmn[0], bh = bmx[1] - bmn[1], aw = amx[0] - amn[0], ah = amx[1] - amn[1];
    if ((aw * ah) < (bw * bh)) uv = rot;
}

}  // namespace uvgen

// The generated mesh is kept until the next call; or_generate_uv_get copies it out.
static std::vector<uvgen::V> g_uv_vb;
static std::vector<uint32_t> g_uv_ib;

extern "C" {

int32_t or_generate_uv(const float* pos, const float* nrm, const float* duv, uint32_t nv, const uint32_t* idx,
                       uint32_t ni, float angle, float padding, uint32_t* out_nv, uint32_t* out_ni) {
    using namespace uvgen;
    std::vector<V> vs(nv);
    for (uint32_t i = 0; i < nv; i++) {
        for (int k = 0; k < 3; k++) { vs[i].p[k] = pos[3 * i + k]; vs[i].n[k] = nrm[3 * i + k]; }
        vs[i].t0[0] = duv ? duv[2 * i] : 0.0f;
        vs[i].t0[1] = duv ? duv[2 * i + 1] : 0.0f;
        vs[i].t1[0] = vs[i].t1[1] = 0.0f;
    }
    std::vector<uint32_t> is(idx, idx + ni);
    const uint32_t nf = ni / 3;
    // DCEL::create: a multimap of (vertex, next vertex) -> edge, twins by lookup.
    std::vector<Half> edges(3 * (size_t)nf);
    std::multimap<std::pair<uint32_t, uint32_t>, Half*> cache;
    for (uint32_t f = 0; f < nf; f++) {
        Half* fe = &edges[3 * (size_t)f];
        for (int i = 0; i < 3; i++) { fe[i].face = f; fe[i].vertex = is[3 * f + i]; fe[i].twin = nullptr; }
        for (int i = 0; i < 3; i++) {
            fe[i].next = &fe[(i + 1) % 3];
            cache.insert({{fe[i].vertex, fe[i].next->vertex}, &fe[i]});
        }
    }
    for (Half& e : edges) {
        auto z = cache.find({e.next->vertex, e.vertex});
        if (z != cache.end() && &e != z->second) { e.twin = z->second; z->second->twin = &e; }
    }
    Gen G(vs, is, angle);
    for (Half& e : edges) G.add_to_chart(&e, G.normal(e.face), (int)G.charts.size());
    Indexer ix;
    std::vector<std::vector<uint32_t>> cv(G.charts.size());
    for (size_t c = 0; c < G.charts.size(); c++) {
        Vec n = mk(0, 0, 0);
        for (uint32_t f : G.charts[c]) n = add(n, G.normal(f));
        n = divs(n, (float)G.charts[c].size());
        normalize3(n);
        const float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
        int axis = 2;
        if (ax > ay && ax > az) axis = 0;
        else if (ay > ax && ay > az) axis = 1;
        std::vector<std::pair<float, float>> uv;
        for (uint32_t f : G.charts[c])
            for (int k = 0; k < 3; k++) {
                const float* p = vs[is[3 * f + k]].p;
                uv.push_back(axis == 0 ? std::make_pair(p[1], p[2]) : axis == 1 ? std::make_pair(p[0], p[2])
                                                                                 : std::make_pair(p[0], p[1]));
            }
        rotate(uv);
        float mn[2], mx[2];
        bounds(uv, mn, mx);
        const float sw = mx[0] - mn[0], sh = mx[1] - mn[1];
        size_t q = 0;
        for (uint32_t f : G.charts[c])
            for (int k = 0; k < 3; k++) {
                V v = vs[is[3 * f + k]];
                const auto t = uv[q++];
                if (sw > sh) { v.t1[0] = t.first - mn[0]; v.t1[1] = t.second - mn[1]; }
                else { v.t1[0] = t.second - mn[1]; v.t1[1] = t.first - mn[0]; }
                cv[c].push_back(ix += v);
            }
    }
    std::vector<R> rects;
    for (size_t c = 0; c < G.charts.size(); c++) {
        float mn[2] = {FLT_MAX, FLT_MAX}, mx[2] = {-FLT_MAX, -FLT_MAX};
        for (uint32_t i : cv[c]) {
            const float* t = ix.vb[i].t1;
            mn[0] = std::min(t[0], mn[0]); mn[1] = std::min(t[1], mn[1]);
            mx[0] = std::max(t[0], mx[0]); mx[1] = std::max(t[1], mx[1]);
        }
        const float sw = mx[0] - mn[0], sh = mx[1] - mn[1];
        rects.push_back(R{0.0f, 0.0f, std::max(sw, sh) + padding, std::min(sw, sh) + padding});
    }
    float w = 1.0f, h = 1.0f;
    bool ew = true;
    while (!place(rects, w, h)) {
        if (ew) { w += 0.1f; ew = false; } else { h += 0.1f; ew = true; }
    }
    Indexer out;
    for (size_t c = 0; c < G.charts.size(); c++)
        for (uint32_t i : cv[c]) {
            V v = ix.vb[i];
            v.t1[0] = rects[c].x + v.t1[0];
            v.t1[1] = rects[c].y + v.t1[1];
            out += v;
        }
    float mn[2] = {FLT_MAX, FLT_MAX}, mx[2] = {-FLT_MAX, -FLT_MAX};
    for (uint32_t i : out.ib) {
        const float* t = out.vb[i].t1;
        mn[0] = std::min(t[0], mn[0]); mn[1] = std::min(t[1], mn[1]);
        mx[0] = std::max(t[0], mx[0]); mx[1] = std::max(t[1], mx[1]);
    }
    for (V& v : out.vb) {
        v.t1[0] = (v.t1[0] - mn[0]) / (mx[0] - mn[0]);
        v.t1[1] = (v.t1[1] - mn[1]) / (mx[1] - mn[1]);
    }
    g_uv_vb = out.vb;
    g_uv_ib = out.ib;
    *out_nv = (uint32_t)out.vb.size();
    *out_ni = (uint32_t)out.ib.size();
    return 0;
}

void or_generate_uv_get(float* pos, float* nrm, float* duv, float* luv, uint32_t* idx) {
    for (size_t i = 0; i < g_uv_vb.size(); i++) {
        for (int k = 0; k < 3; k++) { pos[3 * i + k] = g_uv_vb[i].p[k]; nrm[3 * i + k] = g_uv_vb[i].n[k]; }
        for (int k = 0; k < 2; k++) { duv[2 * i + k] = g_uv_vb[i].t0[k]; luv[2 * i + k] = g_uv_vb[i].t1[k]; }
    }
    std::copy(g_uv_ib.begin(), g_uv_ib.end(), idx);
}

}  // extern "C"
