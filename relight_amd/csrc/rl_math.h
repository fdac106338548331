// Host/device arithmetic shared by the bake kernels and the host-side scene code.
//
// Every expression here is evaluated in the order the reference writes it, and the library is
// compiled with -ffp-contract=off, so the host (lumel setup, light vertices) and the device
// (ray setup, sampling, shading) produce the same bits for the same inputs. Division and sqrt are
// IEEE correctly rounded on both sides (HIP's default for fp32 div/sqrt on gfx950).
#pragma once

#include <stddef.h>
#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RL_HD __host__ __device__ __forceinline__
#else
#define RL_HD inline
#endif

namespace rl {

struct V3 {
    float x, y, z;
};

RL_HD V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
RL_HD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
RL_HD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
RL_HD V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
RL_HD V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
RL_HD V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
// Vec3::operator*(Vec3) is the dot product, summed left to right (Types.h:792-794).
RL_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// The reference's powf (LightSpotCutoff::cutoffForDirection, scene/Light.cpp:464-477, and
// AmbientOcclusion::bakeLumel, baker/AmbientOcclusion.cpp:61-63) as the double-precision pow
// rounded to float. Single-precision powf implementations differ in the last bit (the GPU's and
// glibc's disagree on about 1 % of spot-light cutoffs); a double pow within an ulp of the exact
// value rounds to the correctly rounded powf except when x^y lies within about 2^-52 of a float
// rounding midpoint, so device and host agree bit for bit.
RL_HD float ref_powf(float x, float y) { return (float)pow((double)x, (double)y); }
RL_HD float length(V3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
// Vec3::normalize (Types.h:852-859): divide by the length unless it is zero, return the length.
RL_HD float normalize(V3& a) {
    float len = length(a);
    if (len != 0.0f) { a.x /= len; a.y /= len; a.z /= len; }
    return len;
}
RL_HD float comp(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// ---------------------------------------------------------------- random streams
// PCG32 (M. E. O'Neill, "PCG: A Family of Simple Fast Space-Efficient Statistically Good
// Algorithms for Random Number Generation", 2014): 64-bit LCG state, XSH-RR 32-bit output.
// It replaces the reference's unseeded global rand() (Types.h:110-113), so every sample below is
// a pure function of (seed, lightmap seed, texel, technique, sample).
struct Pcg32 {
    uint64_t state, inc;
};

#define RL_PCG_MULT 6364136223846793005ull

RL_HD uint32_t pcg32_next(Pcg32& r) {
    uint64_t old = r.state;
    r.state = old * RL_PCG_MULT + r.inc;
    uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
    uint32_t rot = (uint32_t)(old >> 59);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

RL_HD Pcg32 pcg32_seed(uint64_t initstate, uint64_t initseq) {
    Pcg32 r;
    r.state = 0;
    r.inc = (initseq << 1) | 1ull;
    pcg32_next(r);
    r.state += initstate;
    pcg32_next(r);
    return r;
}

// Uniform float in [0,1) from the top 24 bits (exact in fp32).
RL_HD float pcg32_float(Pcg32& r) { return (float)(pcg32_next(r) >> 8) * (1.0f / 16777216.0f); }

// SplitMix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom number generators",
// OOPSLA 2014) to spread (seed, lightmap, texel) keys over PCG initial states.
RL_HD uint64_t splitmix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

enum Technique : uint32_t {
    TECH_AO = 1,
    TECH_FINAL_GATHER = 2,
    TECH_PHOTON = 3,
    TECH_FORM_FACTOR = 4,
};

// The initial state shared by a texel's streams: (bake seed, lightmap seed, texel) spread by
// SplitMix64.
RL_HD uint64_t texel_key(uint32_t seed, uint32_t lm_seed, uint32_t texel) {
    const uint64_t key = splitmix64(((uint64_t)seed << 32) | (uint64_t)lm_seed);
    return splitmix64(key ^ (uint64_t)texel);
}

// The stream of one (technique, sample) of a texel, from its texel_key.
RL_HD Pcg32 texel_stream(uint64_t key, uint32_t tech, uint32_t sample) {
    return pcg32_seed(key, ((uint64_t)tech << 32) | (uint64_t)sample);
}

// One independent stream per (bake seed, lightmap seed, texel, technique, sample).
RL_HD Pcg32 sample_stream(uint32_t seed, uint32_t lm_seed, uint32_t texel, uint32_t tech,
                          uint32_t sample) {
    return texel_stream(texel_key(seed, lm_seed, texel), tech, sample);
}

// ---------------------------------------------------------------- hemisphere sampling
enum Sampling : int32_t {
    SAMPLE_UNIFORM = 0,  // the reference's rejection-in-cube + flip (Types.h:924-949)
    SAMPLE_COSINE = 1,   // Malley's method over a rejection-sampled unit disk
};

// Vec3::randomDirection (Types.h:924-937) with PCG draws: rejection in [-1,1]^3, normalised.
// A zero-length draw is rejected as well (the reference would divide by zero there).
RL_HD V3 sample_sphere(Pcg32& r) {
    V3 d;
    float len;
    do {
        d.x = pcg32_float(r) * 2.0f - 1.0f;
        d.y = pcg32_float(r) * 2.0f - 1.0f;
        d.z = pcg32_float(r) * 2.0f - 1.0f;
        len = length(d);
    } while (len > 1.0f || len == 0.0f);
    return d / len;
}

// Vec3::randomHemisphereDirection (Types.h:940-949): a sphere direction flipped towards n.
RL_HD V3 sample_uniform_hemisphere(Pcg32& r, V3 n) {
    V3 d = sample_sphere(r);
    if (dot(d, n) < 0.0f) d = -d;
    return d;
}

// Cosine-weighted direction about n: a point (x,y) uniform in the unit disk lifted to
// z = sqrt(1 - x^2 - y^2) (Malley), rotated into the frame of n with the branchless orthonormal
// basis of Duff et al., "Building an Orthonormal Basis, Revisited", JCGT 6(1), 2017.
// The orthonormal frame (t1, t2, n) the cosine sampler rotates into.
struct CosineFrame {
    V3 t1, t2, n;
};

RL_HD CosineFrame cosine_frame(V3 n) {
    CosineFrame f;
    float s = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (s + n.z);
    float b = n.x * n.y * a;
    f.t1 = v3(1.0f + s * n.x * n.x * a, s * b, -s * n.x);
    f.t2 = v3(b, s + n.y * n.y * a, -n.y);
    f.n = n;
    return f;
}

RL_HD V3 sample_cosine_in_frame(Pcg32& r, const CosineFrame& f) {
    float x, y, r2;
    do {
        x = pcg32_float(r) * 2.0f - 1.0f;
        y = pcg32_float(r) * 2.0f - 1.0f;
        r2 = x * x + y * y;
    } while (r2 >= 1.0f);
    float z = sqrtf(1.0f - r2);
    return v3(f.t1.x * x + f.t2.x * y + f.n.x * z,
              f.t1.y * x + f.t2.y * y + f.n.y * z,
              f.t1.z * x + f.t2.z * y + f.n.z * z);
}

RL_HD V3 sample_cosine_hemisphere(Pcg32& r, V3 n) { return sample_cosine_in_frame(r, cosine_frame(n)); }

RL_HD V3 sample_hemisphere(Pcg32& r, V3 n, int mode) {
    return mode == SAMPLE_COSINE ? sample_cosine_hemisphere(r, n) : sample_uniform_hemisphere(r, n);
}

// ---------------------------------------------------------------- rays
// Segment query semantics of Embree::initializeRay (rt/Embree.cpp:107-130): direction is the
// normalised end-start, the valid interval is (0.01, |end-start|).
#define RL_TNEAR 0.01f

struct Ray {
    V3 o, d;
    float tnear, tfar;
};

RL_HD Ray ray_from_segment(V3 start, V3 end) {
    Ray r;
    r.o = start;
    r.d = end - start;
    r.tfar = normalize(r.d);
    r.tnear = RL_TNEAR;
    return r;
}

// Per-ray constants of the watertight ray/triangle test (Woop, Benthin, Wald, "Watertight
// Ray/Triangle Intersection", JCGT 2(1), 2013): the dominant axis kz, the two others
// kx = kz + 1 and ky = kz + 2 (mod 3), and the shear constants. The paper swaps kx and ky when
// d[kz] < 0 to keep the winding; the swap exchanges the two sheared coordinates, which negates U,
// V, W, det and T exactly (each is an antisymmetric expression of them, and IEEE subtraction is
// antisymmetric), so the hit, t and the ratios V/det, W/det are the same without it. Without the
// swap the component order (kx, ky, kz) is a rotation of (x, y, z) by kx, and the origin is kept
// in that order: rot 0 = (x, y, z), 1 = (y, z, x), 2 = (z, x, y).
struct WtRay {
    V3 o;          // origin components in the order (kx, ky, kz)
    uint32_t rot;  // kx
    float Sx, Sy, Sz;
    float tnear, tfar;
};

RL_HD WtRay wt_setup(const Ray& r) {
    WtRay w;
    float ax = fabsf(r.d.x), ay = fabsf(r.d.y), az = fabsf(r.d.z);
    int kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dz = comp(r.d, kz);
    w.o = v3(comp(r.o, kx), comp(r.o, ky), comp(r.o, kz));
    w.rot = (uint32_t)kx;
    w.Sx = comp(r.d, kx) / dz;
    w.Sy = comp(r.d, ky) / dz;
    w.Sz = 1.0f / dz;
    w.tnear = r.tnear;
    w.tfar = r.tfar;
    return w;
}

// A point's components in the ray's order (kx, ky, kz), selected component by component.
RL_HD V3 wt_rotate(const WtRay& r, V3 p) {
    const bool r0 = r.rot == 0, r1 = r.rot == 1;
    return v3(r0 ? p.x : (r1 ? p.y : p.z), r0 ? p.y : (r1 ? p.z : p.x), r0 ? p.z : (r1 ? p.x : p.y));
}

// The watertight test from the vertices relative to the ray origin, each given by its kx, ky and
// kz components: shear, edge functions (an exactly-zero edge function is recomputed in double
// precision, as the paper prescribes, so shared edges are never missed), inside test and distance.
// U, V, W are the unnormalised barycentrics of v0, v1, v2.
RL_HD bool wt_finish(const WtRay& r, float Akx, float Aky, float Akz, float Bkx, float Bky, float Bkz, float Ckx,
                     float Cky, float Ckz, float& t, float& U, float& V, float& W, float& det) {
    float Ax = Akx - r.Sx * Akz;
    float Ay = Aky - r.Sy * Akz;
    float Bx = Bkx - r.Sx * Bkz;
    float By = Bky - r.Sy * Bkz;
    float Cx = Ckx - r.Sx * Ckz;
    float Cy = Cky - r.Sy * Ckz;
    U = Cx * By - Cy * Bx;
    V = Ax * Cy - Ay * Cx;
    W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    t = T / det;
    return t > r.tnear && t < r.tfar;
}

// Returns true and the hit distance t when the triangle is hit with tnear < t < tfar; the
// vertices are given with their components in the ray's order (wt_rotate).
RL_HD bool wt_intersect_rotated(const WtRay& r, V3 v0, V3 v1, V3 v2, float& t, float& U, float& V,
                                float& W, float& det) {
    V3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
    return wt_finish(r, A.x, A.y, A.z, B.x, B.y, B.z, C.x, C.y, C.z, t, U, V, W, det);
}

// The same for vertices in (x, y, z) order.
RL_HD bool wt_intersect(const WtRay& r, V3 v0, V3 v1, V3 v2, float& t, float& U, float& V,
                        float& W, float& det) {
    return wt_intersect_rotated(r, wt_rotate(r, v0), wt_rotate(r, v1), wt_rotate(r, v2), t, U, V, W, det);
}

// ---------------------------------------------------------------- light models
// Light kinds of Light::create* (scene/Light.cpp:166-234).
enum LightType : int32_t {
    LIGHT_POINT = 0,
    LIGHT_SPOT = 1,
    LIGHT_DIRECTIONAL = 2,
    LIGHT_AREA = 3,
};

// LinearLightAttenuation::calculate (scene/Light.cpp:496-500).
RL_HD float attenuation(float distance, float radius, float c, float l, float q) {
    float r = distance / radius;
    float v = 1.0f / (1.0f + c + l * r + q * r * r);
    return v > 0.0f ? v : 0.0f;
}

// Rgb::luminance (Types.h:211-213).
RL_HD float luminance(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }

// Face::*At interpolation with Embree's hit barycentrics (rt/Embree.cpp:160,169-171 pass (u, v)
// straight into Face::normalAt/uvAt, scene/Mesh.cpp:454-495): a + (b - a) * v + (c - a) * u.
RL_HD float face_lerp(float a, float b, float c, float u, float v) { return a + (b - a) * v + (c - a) * u; }

// Texture::colorAt (scene/Material.cpp:87-98): texel (int(u * w) % w, int(v * h) % h) with C's
// truncating conversion and remainder, so a negative coordinate below -1 texel gives the
// reference's magenta (1, 0, 1, 1); channels / 255 (alpha 1 for 3-channel textures). A product
// u * w outside the int range (undefined in the reference) is also magenta.
RL_HD void texture_color(const uint8_t* px, int w, int h, int channels, float u, float v, float out[4]) {
    const float fx = u * (float)w, fy = v * (float)h;
    int x = -1, y = -1;
    if (fx > -2147483648.0f && fx < 2147483648.0f && fy > -2147483648.0f && fy < 2147483648.0f) {
        x = (int)fx % w;
        y = (int)fy % h;
    }
    if (x >= 0 && x < w && y >= 0 && y < h) {
        const uint8_t* p = px + ((size_t)y * (size_t)w * (size_t)channels + (size_t)x * (size_t)channels);
        out[0] = p[0] / 255.0f;
        out[1] = p[1] / 255.0f;
        out[2] = p[2] / 255.0f;
        out[3] = channels == 4 ? p[3] / 255.0f : 1.0f;
    } else {
        out[0] = 1.0f; out[1] = 0.0f; out[2] = 1.0f; out[3] = 1.0f;
    }
}

// Lightmap::lumel(const Uv&) (Lightmap.cpp:61-67): int(uv * (size - 1)); clamped to the map
// (the reference asserts instead).
RL_HD int lumel_coord(float uv, int size) {
    int x = (int)(uv * (float)(size - 1));
    return x < 0 ? 0 : (x >= size ? size - 1 : x);
}

// LightInfluence::lambert (scene/Light.cpp:392-396).
RL_HD float lambert(V3 dir, V3 n) {
    float dp = dot(dir, n);
    return dp < 0.0f ? 0.0f : dp;
}

}  // namespace rl
