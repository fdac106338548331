// Device-side data layout shared by the host runtime and the HIP kernels.
//
// HBM layout of a built scene (one allocation each, 128-B aligned by hipMalloc):
//   nodes : Bvh4Node[num_nodes]       128 B per node, child boxes stored SoA inside the node so
//                                     one lane reads a node as eight 16-B loads.
//   tris  : float4[3 * num_tris]      48 B per triangle in leaf order: v0, v1, v2; v0.w carries
//                                     the scene-order triangle id (bit pattern of a uint32). Read
//                                     by the closest-hit traversals.
//   tris_rot : float4[4 * num_tris]   the same triangles, 64 B each: vertex k as (x, y, z, x, y)
//                                     at floats 5k .. 5k + 4 (a ray's watertight component order,
//                                     a rotation of (x, y, z), is three consecutive floats), float
//                                     15 the scene-order id. Read by the any-hit traversals.
//   tri_attr : TriAttr[num_tris]      scene-order attributes used only after a closest hit.
// Lumels of a lightmap are stored compacted (valid texels only) as three float4 streams: position
// (w = texel index y*width+x), normal (w = scene-order triangle id), colour (w unused).
#pragma once

#include <stdint.h>
#include <hip/hip_runtime.h>
#include "rl_math.h"

namespace rl {

#define RL_LEAF_BIT 0x80000000u
#define RL_EMPTY_CHILD 0xFFFFFFFFu
#define RL_LEAF_MAX 8
// Leaf child word: bit 31 set, bits 27..30 = count-1, bits 0..26 = first triangle (leaf order).
#define RL_LEAF_FIRST_MASK 0x07FFFFFFu

struct alignas(128) Bvh4Node {
    float lox[4], hix[4];
    float loy[4], hiy[4];
    float loz[4], hiz[4];
    uint32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(Bvh4Node) == 128, "node size");

// Per scene-order triangle: owning mesh, face index in the mesh, and lightmap uvs of its three
// vertices in the reference's vertex order a, b, c (indices 0, 1, 2 of the face).
struct TriAttr {
    uint32_t mesh;
    uint32_t face;
    float uv[6];
    float n[9];
    float albedo[4];
};

// A texture (Texture, scene/Material.h:66-123): width x height pixels of `channels` (3 or 4)
// bytes, rows in the caller's order.
struct TexDev {
    const uint8_t* px;
    int32_t w, h, channels;
    int32_t pad;
};

// Per scene-order triangle of a scene with textured meshes: the diffuse uvs of its vertices a, b,
// c (Vertex::Diffuse, scene/Mesh.h:37-72) and its texture (-1: the mesh's constant colour only).
struct TriMat {
    float duv[6];
    int32_t tex;
    uint32_t pad;
};

struct SceneDev {
    const Bvh4Node* nodes;
    const float4* tris;
    const float4* tris_rot;
    const TriAttr* attr;
    uint32_t num_nodes;
    uint32_t num_tris;
    const TriMat* mat;    // NULL when no mesh of the scene is textured
    const TexDev* texs;
};

// One light as the bake kernels see it (scene/Light.h models, flattened).
struct LightDev {
    int32_t type;
    int32_t casts_shadow;
    int32_t has_attenuation;
    int32_t num_vertices;   // > 0: vertex-generated (area) light
    int32_t vertex_offset;  // into the light-vertex array (xyz float4, local to position)
    int32_t has_spot;
    float pos[3];
    float dir[3];
    float color[3];
    float intensity;
    float radius, att_c, att_l, att_q;
    float cutoff, spot_exponent;
};

struct AoParams {
    int32_t samples;
    float occluded_fraction;
    float max_distance;
    float exponent;
    uint32_t seed;
    int32_t sampling;
};

// A photonmap (Lightmap.h:153-170) as full w*h grids: accumulated photon power (rgb), photon
// count, gathered colour, and the lumel validity the photon store checks.
struct PhotonMapDev {
    float4* color;
    uint32_t* photons;
    float4* gathered;
    const uint8_t* valid;
    int32_t w, h;
    uint32_t key_base;  // first global texel key of this map (maps are concatenated)
    uint32_t pad;
};

enum PhotonTermination : int32_t {
    PHOTON_THRESHOLD = 0,        // depth > maxDepth or luminance < threshold (baker/Photons.cpp:98)
    PHOTON_RUSSIAN_ROULETTE = 1, // survive with probability mean(albedo), power divided by it
};

struct PhotonParams {
    uint64_t per_pass;  // photons per pass over all lights
    uint64_t total;     // passes * per_pass
    uint32_t seed;
    int32_t max_depth;
    float energy_threshold;
    float max_distance;
    int32_t termination;
    int32_t nlights;
    float scene_lo[3], scene_hi[3];
};

// One photon store: global texel key (0xFFFFFFFF = nothing stored) and its power.
struct PhotonRec {
    uint32_t key;
    float r, g, b;
};

// Hash grid of stored photons for the k-nearest gather: cell = floor((p - lo) * inv_h) per axis,
// bucket = spatial hash of the cell & mask. Cell size h = radius / 1.98, so every photon within
// the radius lies in the 5x5x5 cells around a query's cell (max_shell = 2), with a margin for
// rounding; shell_d2[s] = ((s - 0.01) h)^2 lets a query stop after shell s once its k-th
// distance is below it (unvisited photons are farther).
struct KnnGrid {
    float lo[3];
    float inv_h;
    uint32_t mask;
    int32_t max_shell;
    int32_t k;
    float r2;
    float shell_d2[4];
};

// A radiosity patch: the quad origin + s*U + t*V (s, t in [0,1]) with normal n.
struct PatchDev {
    float o[3], u[3], v[3], n[3];
};

struct FormFactorParams {
    uint32_t num_patches;
    uint32_t row_stride;  // u16 entries per F row (num_patches padded to a multiple of 64)
    uint32_t rays;
    uint32_t seed;
    float max_distance;
    uint32_t first_row;
    uint32_t rows;
    uint32_t pad;
};

struct FinalGatherParams {
    int32_t samples;
    float distance;
    float sky[3];
    float ambient[3];
    uint32_t seed;
    int32_t sampling;
};

// A mesh being rasterised into a lightmap's texel grid (positions, normals: 3 floats per vertex;
// lightmap uvs: 2; indices: 3 per face; tri_base: scene-order id of its first face).
struct RasterMesh {
    const float* pos;
    const float* nrm;
    const float* uv;
    const uint32_t* idx;
    uint32_t nfaces;
    int32_t w, h;
    uint32_t tri_base;
};

// One rasterised texel: index, scene-order triangle, position and (unnormalised) normal.
struct RasterRec {
    uint32_t texel, tri;
    float pos[3], nrm[3];
};

// One lightmap (or a lumel range of one) inside a bake launch, with the scene, lights, AO and
// final-gather settings and photon tables it is baked with, so one launch can cover lightmaps of
// different scenes and jobs.
struct BakeSeg {
    const float4* pos;
    const float4* nrm;
    float4* color;
    uint32_t* counts;     // per-lumel AO occluded counts
    uint32_t count;       // lumels in this segment
    uint32_t lm_seed;
    uint64_t work_begin;  // first global work item (or chunk, for chunked launches) of this segment
    SceneDev scene;
    const LightDev* lights;
    const float4* lverts;
    int32_t nlights;
    AoParams ao;
    FinalGatherParams fg;
    const int32_t* mesh_pm;     // scene mesh -> photonmap index (final gather)
    const PhotonMapDev* pms;    // the scene's photonmaps (final gather)
};

}  // namespace rl
