/*
 * relight.h — C-ABI of the MI355X lightmap baker (librelight_amd.so).
 *
 * Drop-in boundary for the ray-cast hot path of dmsovetov/Relight: the public Relight / Scene /
 * Lightmap API (src/relight/Relight.h:155-189, scene/Scene.h, Lightmap.h:58-150) and the tracer
 * plugin seam rt::ITracer (rt/Tracer.h:75-108) that Scene::end hard-wires to Embree
 * (scene/Scene.cpp:129). Every entry point is extern "C" with plain pointers and sizes.
 *
 * Ownership: every *_create returns a handle the caller owns and releases with *_destroy (the
 * reference leaks its factories' objects; this ABI adds the destroy calls).
 * Errors: int32 status codes mirroring RelightStatus (Relight.h:82-86) plus codes the GPU path
 * needs; rl_last_error() returns a message for the calling thread's last failure.
 * Threading: calls on one handle must not overlap; GPU work is ordered on the context's stream
 * and every bake call returns after its work has finished.
 */
#ifndef RELIGHT_AMD_RELIGHT_H
#define RELIGHT_AMD_RELIGHT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The first three are RelightStatus (Relight.h:82-86). */
#define RL_SUCCESS 0
#define RL_INVALID_CALL 1
#define RL_NOT_IMPLEMENTED 2
#define RL_OUT_OF_MEMORY 3
#define RL_REJECTED 4
#define RL_DEVICE_ERROR 5
#define RL_INVALID_ARGUMENT 6
#define RL_NOT_CONVERGED 7

/* Light kinds: Light::createPointLight / createSpotLight / createDirectionalLight /
 * createAreaLight (scene/Light.cpp:166-234). */
#define RL_LIGHT_POINT 0
#define RL_LIGHT_SPOT 1
#define RL_LIGHT_DIRECTIONAL 2
#define RL_LIGHT_AREA 3

/* Hemisphere sampling for AO and final gather. RL_SAMPLE_UNIFORM is the reference's
 * rejection-in-cube + flip (Types.h:924-949); RL_SAMPLE_COSINE is cosine-weighted (Malley). Both
 * draw from per-(seed, lightmap seed, texel, sample) PCG32 streams instead of rand(). */
#define RL_SAMPLE_UNIFORM 0
#define RL_SAMPLE_COSINE 1

/* Trace modes of rl_trace_segments. */
#define RL_TRACE_ANY 0     /* occlusion: hit flag only */
#define RL_TRACE_CLOSEST 1 /* closest hit: t, triangle, barycentrics */

typedef struct rl_relight rl_relight;
typedef struct rl_scene rl_scene;
typedef struct rl_lightmap rl_lightmap;

/* A light (scene/Light.h). Attenuation is LinearLightAttenuation (scene/Light.cpp:496-500) with
 * radius and constant/linear/quadratic terms (defaults 0, 0, 25 at Light.h:125). has_attenuation=0
 * disables it (the directional light has none). Area lights take their light mesh in local
 * coordinates; one light vertex per face centroid is generated (FaceLightVertexGenerator with
 * excludeVertices=true, maxSubdivisions=0, scene/Light.cpp:224,293-323) and offset by position. */
typedef struct rl_light_desc {
    int32_t type;
    int32_t casts_shadow;
    float position[3];
    float direction[3];
    float color[3];
    float intensity;
    int32_t has_attenuation;
    float radius;
    float att_constant;
    float att_linear;
    float att_quadratic;
    float spot_cutoff;    /* RL_LIGHT_SPOT: cosine cutoff (LightSpotCutoff) */
    float spot_exponent;  /* RL_LIGHT_SPOT: 1 in createSpotLight */
    const float* area_positions; /* RL_LIGHT_AREA: xyz per vertex, local to position */
    const uint32_t* area_indices;
    uint32_t area_vertex_count;
    uint32_t area_index_count;
} rl_light_desc;

/* AmbientOcclusionSettings (Relight.h:135-152) plus the sample-stream seed and sampling mode. */
typedef struct rl_ao_settings {
    int32_t samples;
    float occluded_fraction;
    float max_distance;
    float exponent;
    uint32_t seed;
    int32_t sampling;
} rl_ao_settings;

/* Closest-hit record of rl_trace_segments: Embree's ray record fields the reference reads
 * (rt/Embree.cpp:151-160). triangle = scene-order id (-1 on a miss); mesh/face locate it;
 * (u, v) are Embree barycentrics: hit = (1-u-v)*v0 + u*v1 + v*v2. */
typedef struct rl_hit {
    float t;
    int32_t triangle;
    int32_t mesh;
    int32_t face;
    float u;
    float v;
} rl_hit;

/* Photon termination: the reference's depth/energy test (baker/Photons.cpp:98) or Russian
 * roulette on the hit albedo (survive with probability mean(albedo), power divided by it). */
#define RL_PHOTON_THRESHOLD 0
#define RL_PHOTON_RUSSIAN_ROULETTE 1

/* Photonmap gather: the reference's texel disc (Photonmap::gather, Lightmap.cpp:376-413), or the
 * k nearest stored photons in 3D (hash grid) around each photonmap texel's surface point, within
 * knn_radius, normalised as the reference does (sum of power / number of photons). */
#define RL_GATHER_TEXEL_DISC 0
#define RL_GATHER_KNN 1

/* IndirectLightSettings (Relight.h:108-132) plus the sample-stream seed, the final-gather
 * hemisphere sampling, an optional photon count per light per pass (0: intensity * 25000 as
 * PhotonEmitter::photonCount, scene/Light.cpp:334-337), the photon termination rule and the
 * gather rule (k-NN: knn_k in 1..64 photons within knn_radius; the photons of one
 * rl_emit_photons call). */
typedef struct rl_indirect_settings {
    int32_t photon_pass_count;
    int32_t photon_bounce_count;
    float photon_energy_threshold;
    float photon_max_distance;
    int32_t final_gather_samples;
    float final_gather_distance;
    int32_t final_gather_radius;
    float sky_color[3];
    float ambient_color[3];
    uint32_t seed;
    int32_t sampling;
    int32_t photons_per_light;
    int32_t termination;
    int32_t gather_mode;
    int32_t knn_k;
    float knn_radius;
} rl_indirect_settings;

/* Kernel slots of rl_stats.kernel_ms. */
#define RL_KERNEL_DIRECT 0
#define RL_KERNEL_AO_COUNT 1
#define RL_KERNEL_AO_FINALIZE 2
#define RL_KERNEL_TRACE 3
#define RL_KERNEL_PHOTON_TRACE 4
#define RL_KERNEL_PHOTON_STORE 5
#define RL_KERNEL_PHOTON_GATHER 6
#define RL_KERNEL_FINAL_GATHER 7
#define RL_KERNEL_FORM_FACTORS 8
#define RL_KERNEL_FF_GEMM_ACCURATE 9
#define RL_KERNEL_FF_GEMM_FAST 10
#define RL_KERNEL_PHOTON_GRID 11
#define RL_KERNEL_SLOTS 16

/* Counters of the last bake or trace call on a context. */
typedef struct rl_stats {
    uint64_t rays;          /* rays actually cast by the GPU */
    uint64_t texels;        /* valid lumels processed */
    double gpu_ms;          /* device time of the call, first to last kernel (HIP events) */
    double wall_ms;         /* host wall time of the call */
    int32_t launches;       /* kernel launches issued */
    int32_t pad;
    double kernel_ms[RL_KERNEL_SLOTS];        /* per-kernel device time (HIP events on the
                                                 context's stream around each launch) */
    int32_t kernel_launches[RL_KERNEL_SLOTS];
} rl_stats;

/* Last error message of the calling thread ("" if none). */
const char* rl_last_error(void);

/* Relight::create (Relight.h:183). Creates the context: device, stream, events. */
int32_t rl_create(rl_relight** out);
int32_t rl_destroy(rl_relight* ctx);

/* Relight::createScene (Relight.h:165) and the Scene state machine begin -> add* -> end
 * (scene/Scene.cpp:98-142). add_mesh copies and transforms the mesh like Scene::addMesh +
 * Mesh::transformed (scene/Mesh.cpp:277-297): transform is 16 floats in the reference's
 * Matrix4 layout (Types.h:1188-1196) or NULL; albedo is the material colour (rgba) or NULL for
 * Face::colorAt's (1,1,1,1) default (scene/Mesh.cpp:478-485). The colour is the photon
 * reflectance (baker/Photons.cpp:119), and its alpha is the one shadow rays test (HitUseAlpha,
 * rt/Embree.cpp:163-165: a hit with alpha <= 0.1 does not block the light; the reference's
 * Material(Rgb) always has alpha 1, so a constant alpha here acts as a 1x1 RGBA texture would).
 * Indices are 32-bit; meshes are not limited to the reference's 16-bit Index (Relight.h:59).
 * end() builds the BVH and uploads the scene to HBM (the rtcCommit of rt/Embree.cpp:55-58). */
int32_t rl_scene_create(rl_relight* ctx, rl_scene** out);
int32_t rl_scene_destroy(rl_scene* scene);
int32_t rl_scene_begin(rl_scene* scene);
int32_t rl_scene_add_mesh(rl_scene* scene, const float* positions, const float* normals,
                          const float* lightmap_uvs, uint32_t vertex_count,
                          const uint32_t* indices, uint32_t index_count, const float* transform,
                          const float* albedo, int32_t* out_mesh);
/* Texture::create (scene/Material.h:95, scene/Material.cpp:78-82): width x height pixels of
 * `channels` (3 or 4) bytes, copied. A scene keeps its own copy of the textures its meshes use,
 * so a texture may be destroyed after rl_scene_add_mesh_textured. */
typedef struct rl_texture rl_texture;
int32_t rl_texture_create(rl_relight* ctx, int32_t width, int32_t height, int32_t channels,
                          const uint8_t* pixels, rl_texture** out);
int32_t rl_texture_destroy(rl_texture* texture);
/* Scene::addMesh with a TexturedMaterial(texture, color) (scene/Material.cpp:61-70): the mesh's
 * colour at a hit is Rgba(color, 1) times Texture::colorAt (scene/Material.cpp:87-98: texel
 * (int(u*w) % w, int(v*h) % h), magenta outside) at the hit's diffuse uv (Face::colorAt, the
 * Vertex::Diffuse layer, scene/Mesh.cpp:478-485), with 2 floats per vertex in diffuse_uvs. color
 * NULL is white. Other arguments as rl_scene_add_mesh. */
int32_t rl_scene_add_mesh_textured(rl_scene* scene, const float* positions, const float* normals,
                                   const float* lightmap_uvs, const float* diffuse_uvs,
                                   uint32_t vertex_count, const uint32_t* indices,
                                   uint32_t index_count, const float* transform, const float* color,
                                   const rl_texture* texture, int32_t* out_mesh);
int32_t rl_scene_add_light(rl_scene* scene, const rl_light_desc* light, int32_t* out_light);
int32_t rl_scene_end(rl_scene* scene);
/* Scene statistics after end(): triangle count, BVH node count, traversal stack bound. */
int32_t rl_scene_info(const rl_scene* scene, uint32_t* triangles, uint32_t* nodes,
                      int32_t* stack_bound);

/* Wavefront OBJ with the semantics of MeshLoader::load (scene/MeshLoader.cpp:44-226): submeshes
 * split at v/vt/vn lines after faces, corners deduplicated exactly in first-appearance order,
 * uvs clamped to [0, 1], missing normals (0, 1, 0), triangles only. rl_obj_load parses a file:
 * with NULL arrays it returns the counts, else it fills arrays of at least those sizes (3 floats
 * per position and normal, 2 per uv). rl_scene_add_obj adds the file as one mesh (as
 * rl_scene_add_mesh). Malformed files (indices out of range, unparsable numbers) give
 * RL_INVALID_ARGUMENT where the reference reads undefined values. */
int32_t rl_obj_load(const char* path, uint32_t* vertex_count, uint32_t* index_count, float* positions,
                    float* normals, float* uvs, uint32_t* indices);
int32_t rl_scene_add_obj(rl_scene* scene, const char* path, const float* transform, const float* albedo,
                         int32_t* out_mesh);

/* Mesh::generateUv(angle, margin, padding) (scene/Mesh.cpp:229-238; UvGenerator, AngularChartifier,
 * RectanglePacker, Types.h:2293-3326) on a mesh given as arrays: charts of faces within `angle`
 * degrees of a seed face (connected through shared vertex indices), each projected onto its
 * dominant axis plane and turned onto its principal axis when that shrinks it, packed with
 * `padding` (in the mesh's units) and scaled to [0, 1]; the returned mesh's lightmap uvs are the
 * generated set. Vertices are split and merged as the reference's indexer does (exact equality of
 * position, normal and uvs), so the counts change. margin is unused, as in the reference. Faces
 * with a zero normal join no chart at angles below 90 and are dropped, as in the reference. With
 * NULL output arrays only the counts are written. diffuse_uvs may be NULL (zeros). */
int32_t rl_mesh_generate_uv(const float* positions, const float* normals, const float* diffuse_uvs,
                            uint32_t vertex_count, const uint32_t* indices, uint32_t index_count,
                            float angle, float margin, float padding, uint32_t* out_vertex_count,
                            uint32_t* out_index_count, float* out_positions, float* out_normals,
                            float* out_diffuse_uvs, float* out_lightmap_uvs, uint32_t* out_indices);

/* Relight::createLightmap (Relight.h:159) and Lightmap::addMesh (Lightmap.cpp:93-103), which
 * rasterises the mesh's texels (Lightmap.cpp:133-164). A mesh may belong to one lightmap. The
 * lightmap seed keys its texels' sample streams (default: creation order in the context). */
int32_t rl_lightmap_create(rl_relight* ctx, int32_t width, int32_t height, rl_lightmap** out);
int32_t rl_lightmap_destroy(rl_lightmap* lm);
int32_t rl_lightmap_add_mesh(rl_lightmap* lm, rl_scene* scene, int32_t mesh);
int32_t rl_lightmap_set_seed(rl_lightmap* lm, uint32_t seed);
/* Number of valid lumels (texels covered by a face). */
int32_t rl_lightmap_valid_count(const rl_lightmap* lm, uint32_t* count);
/* Lumel setup of valid texels in texel order: texel index, position xyz, normal xyz. Any output
 * may be NULL. */
int32_t rl_lightmap_read_lumels(rl_lightmap* lm, uint32_t* texel, float* position, float* normal);
/* Lightmap::toHdr (Lightmap.cpp:332-349): width*height*3 floats, invalid texels 0. */
int32_t rl_lightmap_read_hdr(rl_lightmap* lm, float* out_rgb);

/* Lightmap::save storage formats (StorageFormat, Relight.h:89-93). */
#define RL_FORMAT_RAW_HDR 0
#define RL_FORMAT_TGA_DOUBLE_LDR 1
#define RL_FORMAT_TGA_RGBM 2
/* Lightmap::expand (Lightmap.cpp:167-193): valid interior texels copy their colour to invalid
 * diagonal neighbours, in the reference's raster order (on the GPU). */
int32_t rl_lightmap_expand(rl_lightmap* lm);
/* Lightmap::blur (Lightmap.cpp:196-220): in-place 3x3 mean over valid texels in raster order
 * (on the GPU, bit-identical to the sequential order). */
int32_t rl_lightmap_blur(rl_lightmap* lm);
/* Lightmap::toDoubleLdr / toRgbmLdr (Lightmap.cpp:289-331; Rgb::doubleLdr / rgbm, Types.h:216-234):
 * width*height*3 (RL_FORMAT_TGA_DOUBLE_LDR) or *4 (RL_FORMAT_TGA_RGBM) bytes, r g b (m) per texel.
 * Channels the reference leaves undefined (negative or NaN before the byte cast) are 0; RGBM uses
 * 4 bytes per texel (the reference's toRgbmLdr steps 3 bytes within a 4-byte row stride). */
int32_t rl_lightmap_encode(rl_lightmap* lm, int32_t format, uint8_t* out);
/* Lightmap::save (Lightmap.cpp:223-285): RawHdr (int width, int height, float rgb) or an
 * uncompressed 24/32-bit TGA. */
int32_t rl_lightmap_save(rl_lightmap* lm, const char* path, int32_t format);

/* Relight::bakeDirectLight (Relight.cpp:216-227) -> DirectLight::bakeLumel over every valid lumel
 * of the mesh's lightmap (LumelBakeIterator, baker/Baker.cpp:52-77,176-185). One launch. */
int32_t rl_bake_direct(rl_relight* ctx, rl_scene* scene, int32_t mesh);
/* Relight::bakeAmbientOcclusion (Relight.cpp:260-271) -> AmbientOcclusion::bakeLumel. */
int32_t rl_bake_ambient_occlusion(rl_relight* ctx, rl_scene* scene, int32_t mesh,
                                  const rl_ao_settings* settings);
/* Batched bake pass: direct light (if do_direct) then AO (if ao != NULL) over every valid lumel of
 * num_maps lightmaps in one launch per pass. Results equal per-lightmap rl_bake_direct followed by
 * rl_bake_ambient_occlusion. */
int32_t rl_bake_batch(rl_relight* ctx, rl_scene* scene, rl_lightmap* const* maps,
                      int32_t num_maps, int32_t do_direct, const rl_ao_settings* ao);

/* The rl_bake_batch pass recorded once as a HIP graph (segment tables in buffers of its own,
 * kernel choice and launch shape fixed at recording) and replayed by rl_pass_launch with one
 * graph launch; each launch equals the same rl_bake_batch call bit for bit. Replays check that
 * the scene and lightmaps still hold the device arrays the pass was recorded with (rebuilding the
 * scene or re-rasterising a lightmap makes the pass stale: RL_INVALID_CALL). Destroy a pass before
 * its scene, its lightmaps and its context. */
typedef struct rl_pass rl_pass;
int32_t rl_pass_create(rl_relight* ctx, rl_scene* scene, rl_lightmap* const* maps, int32_t num_maps,
                       int32_t do_direct, const rl_ao_settings* ao, rl_pass** out);
int32_t rl_pass_launch(rl_relight* ctx, rl_pass* pass);
int32_t rl_pass_destroy(rl_pass* pass);

/* rt::ITracer::traceSegment (rt/Tracer.h:87) for n segments start[i] -> end[i] (xyz each):
 * direction normalize(end-start), interval (0.01, |end-start|) as Embree::initializeRay
 * (rt/Embree.cpp:107-130). RL_TRACE_ANY writes hit flags (0/1) to out_flags; RL_TRACE_CLOSEST
 * writes out_hits. Host buffers in, host buffers out. */
int32_t rl_trace_segments(rl_relight* ctx, rl_scene* scene, const float* starts,
                          const float* ends, uint32_t n, int32_t mode, uint8_t* out_flags,
                          rl_hit* out_hits);

/* rl_trace_segments for callers that cannot allocate device memory (code that must not call
 * hipMalloc/hipFree: serving loops with a fixed footprint, harnesses that forbid allocation). No
 * counterpart in the reference. Same arguments, results and host buffers as rl_trace_segments,
 * but every device buffer of the call (segment staging, hit records, the work and ray counters)
 * is a static device array of the library, about 6.1 MiB per GPU, held until the process ends;
 * the call allocates nothing. At most RL_NOALLOC_MAX_SEGMENTS segments per call (more:
 * RL_INVALID_ARGUMENT). Calls share the array of their GPU: calls from one process are
 * serialised; calls that overlap from several processes on one GPU race. */
#define RL_NOALLOC_MAX_SEGMENTS 65536
int32_t rl_trace_segments_noalloc(rl_relight* ctx, rl_scene* scene, const float* starts,
                                  const float* ends, uint32_t n, int32_t mode, uint8_t* out_flags,
                                  rl_hit* out_hits);

/* Relight::createPhotonmap (Relight.h:162) and Photonmap::addMesh (Lightmap.cpp:359-373): a
 * photonmap is a lightmap whose texels receive photon stores. read() returns its full grids:
 * accumulated photon power (w*h*3), photon counts (w*h) and gathered colour (w*h*3); any output
 * may be NULL. */
int32_t rl_photonmap_create(rl_relight* ctx, int32_t width, int32_t height, rl_lightmap** out);
int32_t rl_photonmap_add_mesh(rl_lightmap* pm, rl_scene* scene, int32_t mesh);
int32_t rl_photonmap_read(rl_lightmap* pm, float* color, uint32_t* photons, float* gathered);

/* Relight::emitPhotons (Relight.cpp:244-257): Photons::emit over all passes and lights
 * (baker/Photons.cpp:50-144) storing into the hit meshes' photonmaps, then Photonmap::gather
 * with final_gather_radius (Lightmap.cpp:376-413) on every photonmap of the scene. */
int32_t rl_emit_photons(rl_relight* ctx, rl_scene* scene, const rl_indirect_settings* settings);
/* Relight::bakeIndirectLight (Relight.cpp:230-241) -> IndirectLight::bakeLumel
 * (baker/IndirectLight.cpp:47-75) over every valid lumel of the mesh's lightmap. */
int32_t rl_bake_indirect(rl_relight* ctx, rl_scene* scene, int32_t mesh,
                         const rl_indirect_settings* settings);

/* BakeIterator (baker/Baker.h, baker/Baker.cpp:148-234): which lumels a bake visits.
 * RL_ITERATE_LUMELS is LumelBakeIterator(first, step): texel indices first, first + step, ... of the
 * mesh's whole lightmap (the texels of every mesh in it). RL_ITERATE_FACES is
 * FaceBakeIterator(first, step): the lumels owned by faces first, first + step, ... of the mesh (a
 * lumel belongs to the last face of the lightmap that covers its texel centre; the reference would
 * also bake another mesh's lumel of the same face index where two meshes' charts overlap). Workers
 * split a bake with iterators (i, N), i = 0..N-1; the union equals the whole bake bit for bit.
 * first >= 0 and step >= 1 (the reference reads out of range for a first past the end; here such
 * an iterator visits nothing). */
#define RL_ITERATE_LUMELS 0
#define RL_ITERATE_FACES 1
typedef struct rl_bake_iterator {
    int32_t kind;
    int32_t first;
    int32_t step;
} rl_bake_iterator;

/* Progress::notify (Relight.h:96-105, called by Baker::bakeMesh, baker/Baker.cpp:71-74): with a
 * progress record the selected lumels are baked in `updates` parts on the calling thread, and
 * notify(user, mesh, step, step_count) follows each part but the last (step = iterator steps whose
 * lumels are all baked, step_count = the iterator's itemCount: texels or faces), then
 * notify(user, mesh, step_count, step_count) when the mesh is done, as the reference ends. */
typedef void (*rl_progress_fn)(void* user, int32_t mesh, int32_t step, int32_t step_count);
typedef struct rl_progress {
    rl_progress_fn notify;
    void* user;
    int32_t updates;
} rl_progress;

/* Relight::bakeDirectLight / bakeAmbientOcclusion / bakeIndirectLight with their progress and
 * iterator arguments (Relight.cpp:216-271). iterator NULL = LumelBakeIterator(0, 1), the whole
 * lightmap (Relight.cpp:218-220); progress may be NULL. */
int32_t rl_bake_direct_iter(rl_relight* ctx, rl_scene* scene, int32_t mesh,
                            const rl_bake_iterator* iterator, const rl_progress* progress);
int32_t rl_bake_ambient_occlusion_iter(rl_relight* ctx, rl_scene* scene, int32_t mesh,
                                       const rl_ao_settings* settings,
                                       const rl_bake_iterator* iterator,
                                       const rl_progress* progress);
int32_t rl_bake_indirect_iter(rl_relight* ctx, rl_scene* scene, int32_t mesh,
                              const rl_indirect_settings* settings,
                              const rl_bake_iterator* iterator, const rl_progress* progress);

/* Progress as a polled counter: lumels queued (`total`) and finished (`done`) by the bake call
 * running on ctx, readable from any thread while it runs (between calls: the last call's). A
 * single-launch phase counts as finished when its launch has completed. */
int32_t rl_get_progress(const rl_relight* ctx, uint64_t* done, uint64_t* total);

/* Radiosity: patches are quads origin + s*U + t*V with normal n (12 floats each: o, U, V, n);
 * tri_patch maps every scene-order triangle to its patch or 0xFFFFFFFF. Replaces
 * RadiosityBuilder::build / computeFormFactors (radiosity/RadiosityBuilder.cpp:44-176) and the empty
 * RadiositySolver::solve (radiosity/RadiositySolver.cpp:34-36).
 * build_form_factors: F_ij = (closest hits on patch j) / rays_per_patch over rays_per_patch
 * cosine-weighted rays from uniform points on patch i (segments of max_distance), kept as exact
 * 16-bit counts (at most 16,384 patches, 65,535 rays per patch).
 * solve: B = E + rho o (F B) for num_rhs right-hand sides (emission and reflectance are row-major
 * num_patches x num_rhs): Jacobi sweeps with F and the correction in BF16 on MFMA (fp32
 * accumulation) inside fp32 iterative refinement; the residual ||E + rho o FB - B|| / ||E|| uses the
 * exact F and a three-way BF16 split of B. Returns RL_NOT_CONVERGED if `tolerance` is not reached
 * within max_outer refinement steps (the best solution is still written). */
typedef struct rl_radiosity rl_radiosity;
typedef struct rl_solve_info {
    int32_t outer_iterations;  /* residual evaluations */
    int32_t sweeps;            /* BF16 Jacobi sweeps */
    double residual;           /* final relative residual */
    double gpu_ms;
} rl_solve_info;
int32_t rl_radiosity_create(rl_relight* ctx, rl_scene* scene, const float* patches,
                            uint32_t num_patches, const uint32_t* tri_patch, rl_radiosity** out);
int32_t rl_radiosity_destroy(rl_radiosity* r);
int32_t rl_radiosity_build_form_factors(rl_relight* ctx, rl_radiosity* r, uint32_t rays_per_patch,
                                        uint32_t seed, float max_distance);
/* Rows [first_row, first_row+rows) of the hit counts, row-major rows x num_patches. */
int32_t rl_radiosity_read_form_factors(rl_radiosity* r, uint32_t first_row, uint32_t rows,
                                       uint16_t* counts);
int32_t rl_radiosity_solve(rl_relight* ctx, rl_radiosity* r, const float* emission,
                           const float* reflectance, uint32_t num_rhs, float tolerance,
                           int32_t max_outer, int32_t inner_sweeps, float* out_radiosity,
                           rl_solve_info* info);

/* Multi-scene batched bake (the FullBakeJob fan-out of Worker.cpp:43-83 across scenes): one
 * direct-light launch over every item with direct != 0, then one AO launch over every item with
 * ao != NULL. An item is lumels [first_lumel, first_lumel + lumel_count) of a lightmap in valid-
 * texel order (lumel_count = 0xFFFFFFFF: to the end). Results equal per-lightmap rl_bake_direct
 * and rl_bake_ambient_occlusion calls. Items must not overlap. */
typedef struct rl_bake_item {
    rl_scene* scene;
    rl_lightmap* lightmap;
    uint32_t first_lumel;
    uint32_t lumel_count;
    int32_t direct;
    const rl_ao_settings* ao;
} rl_bake_item;
int32_t rl_bake_items(rl_relight* ctx, const rl_bake_item* items, int32_t num_items);

/* Relight::bake with FullBakeJob (Relight.cpp:205-213, Worker.cpp:43-83) for a job that bakes
 * direct light, indirect light and AO: every lightmap of the scene, one launch per pass, in the
 * order photons (indirect != NULL), direct (direct != 0), final gather (indirect), AO (ao != NULL;
 * it multiplies the accumulated colour). The reference's per-mesh order by bounds volume does not
 * change results here (sample streams are keyed by lightmap seed and texel). */
int32_t rl_bake_scene(rl_relight* ctx, rl_scene* scene, int32_t direct, const rl_ao_settings* ao,
                      const rl_indirect_settings* indirect);

/* Bake server (no counterpart in the reference; replaces the Worker/ThreadWorker job loop of
 * Worker.cpp and src/demo/Lightmapping.cpp:619-649 for serving). A background thread owns the
 * context's stream; submitted jobs wait in a FIFO queue and are admitted while the sum of admitted
 * jobs' device footprints stays within memory_budget and at most max_active run. Every step the
 * loop takes up to quantum_rays rays of direct, AO and final-gather work, one launch per phase, from
 * the jobs of one scene: that of the oldest job with ray work. The step takes that scene's jobs
 * older than any other scene's oldest job and its small jobs (whole remaining work within a
 * quarter of an even share; none while that oldest job has waited over 60 ms since submission),
 * each first a quarter of an even share, the rest in admission order
 * (continuous batching: a job submitted while others run joins at the next step). The next step is
 * planned and queued while one runs. Photon emission runs as a whole step; radiosity jobs run
 * beside the steps on a second stream. Scenes, lightmaps
 * (and the scene's photonmaps) and radiosity objects are borrowed until the job finishes, fails or
 * is cancelled: their destroy calls return RL_INVALID_CALL meanwhile. While a server runs, make no
 * other call on its context. Results are identical to the same job run alone. */
#define RL_JOB_DIRECT_AO 0
#define RL_JOB_INDIRECT 1      /* photons, then final gather per lightmap */
#define RL_JOB_RADIOSITY 2
#define RL_JOB_FINAL_GATHER 3  /* final gather from the scene's resident photonmaps */
#define RL_JOB_FULL_BAKE 4     /* photons (indirect != NULL), direct (direct != 0), final gather, AO (ao) */
#define RL_JOB_QUEUED 0
#define RL_JOB_RUNNING 1
#define RL_JOB_DONE 2
#define RL_JOB_FAILED 3
#define RL_JOB_CANCELLED 4
typedef struct rl_server rl_server;
typedef struct rl_server_config {
    uint64_t memory_budget;  /* bytes of per-job device memory admitted at once */
    int32_t max_active;      /* jobs admitted at once */
    uint64_t quantum_rays;   /* rays of direct/AO/final-gather work per step */
} rl_server_config;
typedef struct rl_job_desc {
    int32_t kind;
    rl_scene* scene;
    rl_lightmap* const* lightmaps;   /* DIRECT_AO and INDIRECT: lightmaps to bake */
    int32_t num_lightmaps;
    int32_t direct;                  /* DIRECT_AO: bake direct light first */
    const rl_ao_settings* ao;        /* DIRECT_AO: AO after direct (NULL: none) */
    const rl_indirect_settings* indirect; /* INDIRECT, FINAL_GATHER */
    rl_radiosity* radiosity;         /* RADIOSITY: form factors then solve */
    uint32_t rays_per_patch;
    uint32_t ff_seed;
    float ff_max_distance;
    const float* emission;           /* RADIOSITY: num_patches x num_rhs (copied at submit) */
    const float* reflectance;
    uint32_t num_rhs;
    float tolerance;
    float* out_radiosity;            /* RADIOSITY: written when the job is done */
} rl_job_desc;
typedef struct rl_job_status {
    int32_t state;       /* RL_JOB_* */
    int32_t status;      /* RL_SUCCESS or the failing call's code */
    double submit_ms;    /* times since the server started */
    double admit_ms;
    double done_ms;
    uint64_t rays;
    uint64_t footprint;  /* admitted device bytes */
} rl_job_status;
typedef struct rl_server_stats {
    uint64_t submitted, completed, failed, cancelled, rejected;
    uint64_t steps;
    uint64_t rays;
    double busy_ms;      /* wall time with a step on the GPU (the next step is planned and queued
                          * while one runs) */
    uint64_t peak_memory;
} rl_server_stats;
int32_t rl_server_create(rl_relight* ctx, const rl_server_config* config, rl_server** out);
/* RL_REJECTED if the job's footprint exceeds the budget or the description is invalid. */
int32_t rl_server_submit(rl_server* server, const rl_job_desc* job, uint64_t* out_id);
int32_t rl_server_poll(rl_server* server, uint64_t id, rl_job_status* out);
/* Blocks until the job has finished (timeout_ms < 0: no limit); RL_REJECTED on timeout. */
int32_t rl_server_wait(rl_server* server, uint64_t id, double timeout_ms, rl_job_status* out);
/* Cancels a queued job (running jobs finish their current step and stop). */
int32_t rl_server_cancel(rl_server* server, uint64_t id);
int32_t rl_server_stats_get(rl_server* server, rl_server_stats* out);
/* Stops the loop after cancelling queued jobs and finishing running ones. */
int32_t rl_server_destroy(rl_server* server);

/* Launch tuning of a context (no counterpart in the reference; chosen by the autotuner in
 * relight_amd/autotune.py). ao_kernel: RL_AO_KERNEL_REFILL (lanes take new rays from a per-
 * wavefront pool when at least refill_min of them are idle; chunks of ao_chunk_lumels lumels, at
 * most 16, halved down to 4 for passes with fewer chunks than the kernel has resident wavefronts) or
 * RL_AO_KERNEL_CHUNKED (64 rays per wavefront per chunk). blocks_per_cu sizes the persistent
 * grids (128-thread blocks). Results do not depend on the tuning. */
#define RL_AO_KERNEL_CHUNKED 0
#define RL_AO_KERNEL_REFILL 1
/* Final-gather kernel: AUTO gives each lane its own lumel when a pass has at least one lumel per
   resident lane, else each lumel a wavefront; WAVE and LANE force one of the two. */
#define RL_FG_KERNEL_AUTO 0
#define RL_FG_KERNEL_WAVE 1
#define RL_FG_KERNEL_LANE 2
/* Direct-light kernel: AUTO traces small passes as (lumel, shadow ray) pairs summed per lumel
   afterwards and larger ones with a lane per lumel; LUMEL and PAIRS force one of the two. */
#define RL_DIRECT_KERNEL_AUTO 0
#define RL_DIRECT_KERNEL_LUMEL 1
#define RL_DIRECT_KERNEL_PAIRS 2
/* local_trees (1, the default; or 0): a wavefront's AO rays start at a small tree built in LDS over
   the BVH subtrees they can reach instead of at the scene's root. occluder_hints (1 or 0): AO and
   shadow rays first test the triangles that blocked recent rays of their lumel (or lane). Both
   only change which exact tests run and in which order, never a result. */
typedef struct rl_tuning {
    int32_t ao_kernel;
    uint32_t refill_min;
    uint32_t ao_chunk_lumels;
    uint32_t blocks_per_cu;
    int32_t fg_kernel;
    int32_t direct_kernel;
    int32_t local_trees;
    int32_t occluder_hints;
} rl_tuning;
int32_t rl_set_tuning(rl_relight* ctx, const rl_tuning* tuning);
int32_t rl_get_tuning(const rl_relight* ctx, rl_tuning* out);

/* Counters of the last bake/trace call on ctx. */
int32_t rl_get_stats(const rl_relight* ctx, rl_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* RELIGHT_AMD_RELIGHT_H */
