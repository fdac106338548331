// Host runtime behind include/relight.h: context, scene state machine, lumel rasterisation,
// device residency and the bake entry points.
//
// Mirrors Relight (Relight.cpp:174-271), Scene (scene/Scene.cpp:90-142), Lightmap
// (Lightmap.cpp:55-164,332-349) and the tracer seam (rt/Tracer.h:75-108). Host arithmetic that
// feeds the kernels (lumel positions and normals, light vertices) follows the reference's
// expression order and is compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/relight.h"
#include "rl_bvh.h"
#include "rl_device.h"
#include "rl_kernels.h"
#include "rl_math.h"
#include "rl_noalloc.h"
#include "rl_sort.h"

using namespace rl;

namespace {

thread_local std::string g_error;

int32_t fail(int32_t code, const std::string& msg) {
    g_error = msg;
    return code;
}

}  // namespace

// Error reporting for the library's other host translation units (rl_last_error reads it).
namespace rl {
int32_t report_error(int32_t code, const std::string& msg) { return fail(code, msg); }
}  // namespace rl

namespace {

#define RL_HIP(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            return fail(e_ == hipErrorOutOfMemory ? RL_OUT_OF_MEMORY : RL_DEVICE_ERROR,  \
                        std::string(#call) + ": " + hipGetErrorString(e_));              \
        }                                                                                \
    } while (0)

template <class T>
hipError_t dev_alloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc((void**)p, count * sizeof(T));
}

template <class T>
void dev_free(T*& p) {
    if (p) hipFree((void*)p);
    p = nullptr;
}

// A texture's pixels (Texture::create copies them, scene/Material.cpp:78-82). cutout: some texel's
// alpha is <= 0.1 (4 channels), so shadow rays through its meshes need the alpha re-trace.
struct TexHost {
    int32_t w = 0, h = 0, channels = 0;
    std::vector<uint8_t> px;
    bool cutout = false;
};

struct MeshHost {
    std::vector<V3> pos, nrm;
    std::vector<float> uv;  // lightmap uv, 2 per vertex
    std::vector<float> duv; // diffuse uv, 2 per vertex (textured meshes)
    std::vector<uint32_t> idx;
    float albedo[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    int32_t tex = -1;       // index into the scene's textures (TexturedMaterial), -1: constant colour
    uint32_t first_tri = 0;
    uint32_t num_tris = 0;
    rl_lightmap* lightmap = nullptr;
    rl_lightmap* photonmap = nullptr;
};

struct LightHost {
    LightDev dev;
    std::vector<V3> verts;
};

enum SceneState { kInitial = 0, kAdding = 1, kReady = 2 };

}  // namespace

struct rl_relight {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    unsigned long long* d_rays = nullptr;
    unsigned long long* d_queue = nullptr;  // persistent-kernel work counter
    uint32_t max_blocks = 4096;             // persistent grid: blocks_per_cu blocks per compute unit
    int cus = 256;
    rl_tuning tuning{RL_AO_KERNEL_REFILL, 20u, 8u, 16u, RL_FG_KERNEL_AUTO, RL_DIRECT_KERNEL_AUTO, 1, 1};
    rl_stats stats{};
    uint32_t next_lm_seed = 0;
    // Event pairs around each kernel launch of the current call: (slot, start, end).
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    struct Mark { int slot; hipEvent_t a, b; };
    std::vector<Mark> marks;
    // Segment tables of the launches in flight (at most two at once), grown as needed and reused:
    // every bake helper synchronises its stream before returning.
    // Segment tables, one slot per phase of a pass (kSegDirect..kSegDirectPairs), uploaded on the
    // stream from pinned staging; seg_ev[k] marks when the staging of slot k may be rewritten.
    static constexpr int kSegSlots = 5;
    BakeSeg* seg_buf[kSegSlots] = {};
    size_t seg_cap[kSegSlots] = {};
    BakeSeg* seg_host[kSegSlots] = {};
    size_t seg_host_cap[kSegSlots] = {};
    hipEvent_t seg_ev[kSegSlots] = {};
    bool seg_ev_live[kSegSlots] = {};
    // Lumels queued and finished by the current bake call (rl_get_progress reads them from other
    // threads while the call runs).
    std::atomic<uint64_t> progress_done{0}, progress_total{0};
    // Texel-raster scratch: per-texel owner and claimed-texel records (grown to the largest map).
    uint32_t* raster_owner = nullptr;
    RasterRec* raster_rec = nullptr;
    unsigned int* raster_count = nullptr;
    size_t raster_cap = 0;
    // Influence of each (lumel, shadow ray) pair of a small direct-light pass (grown as needed).
    float* direct_inf = nullptr;
    size_t direct_inf_cap = 0;
};

struct rl_scene {
    rl_relight* ctx = nullptr;
    int state = kInitial;
    std::vector<MeshHost> meshes;
    std::vector<LightHost> lights;
    uint32_t num_tris = 0;
    uint32_t num_nodes = 0;
    int stack_bound = 0;
    int cap = 0;
    float bounds_lo[3] = {0, 0, 0}, bounds_hi[3] = {0, 0, 0};
    Bvh4Node* d_nodes = nullptr;
    float4* d_tris = nullptr;
    float4* d_tris_rot = nullptr;
    TriAttr* d_attr = nullptr;
    LightDev* d_lights = nullptr;
    float4* d_lverts = nullptr;
    std::vector<uint32_t> tri_mesh, tri_face;
    // Textures of the scene's textured meshes (shared with the rl_texture handles they came from)
    // and their device copies; alpha_shadows: some material can have alpha <= 0.1.
    std::vector<std::shared_ptr<const TexHost>> textures;
    TriMat* d_mat = nullptr;
    TexDev* d_texs = nullptr;
    uint8_t* d_texpx = nullptr;
    bool alpha_shadows = false;
    std::atomic<int32_t> server_jobs{0};  // unfinished server jobs that use the scene

    SceneDev dev() const {
        SceneDev s;
        s.nodes = d_nodes;
        s.tris = d_tris;
        s.tris_rot = d_tris_rot;
        s.attr = d_attr;
        s.num_nodes = num_nodes;
        s.num_tris = num_tris;
        s.mat = d_mat;
        s.texs = d_texs;
        return s;
    }
    ~rl_scene() {
        dev_free(d_nodes);
        dev_free(d_tris);
        dev_free(d_tris_rot);
        dev_free(d_attr);
        dev_free(d_lights);
        dev_free(d_lverts);
        dev_free(d_mat);
        dev_free(d_texs);
        dev_free(d_texpx);
    }
};

struct rl_texture {
    std::shared_ptr<const TexHost> t;
};

struct rl_lightmap {
    rl_relight* ctx = nullptr;
    rl_scene* scene = nullptr;
    std::atomic<int32_t> server_jobs{0};  // unfinished server jobs that bake into the lightmap
    int w = 0, h = 0;
    uint32_t seed = 0;
    // Host lumel grid (Lightmap::m_lumels).
    std::vector<uint8_t> valid;
    std::vector<uint32_t> tri;
    std::vector<V3> pos, nrm;
    std::vector<float> color;  // 3 per texel
    // Device copy of the valid lumels, compacted in texel order.
    bool dirty = true;
    bool device_newer = false;
    std::vector<uint32_t> texels;
    float4* d_pos = nullptr;
    float4* d_nrm = nullptr;
    float4* d_color = nullptr;
    uint32_t* d_counts = nullptr;
    uint32_t d_capacity = 0;
    // Photonmap role (Photonmap, Lightmap.h:153-170): full-grid photon power, count, gathered.
    bool is_photonmap = false;
    float4* d_pm_color = nullptr;
    uint32_t* d_pm_photons = nullptr;
    float4* d_pm_gathered = nullptr;
    uint8_t* d_pm_valid = nullptr;
    bool pm_valid_dirty = true;

    ~rl_lightmap() {
        release();
        dev_free(d_pm_color);
        dev_free(d_pm_photons);
        dev_free(d_pm_gathered);
        dev_free(d_pm_valid);
    }
    void release() {
        dev_free(d_pos);
        dev_free(d_nrm);
        dev_free(d_color);
        dev_free(d_counts);
        d_capacity = 0;
    }
};

namespace {

// Copies device colours back into the host grid when a bake has run since the last upload.
int32_t pull_colors(rl_lightmap* lm) {
    if (!lm->device_newer) return RL_SUCCESS;
    const size_t n = lm->texels.size();
    std::vector<float4> c(n);
    if (n) RL_HIP(hipMemcpy(c.data(), lm->d_color, n * sizeof(float4), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) {
        float* dst = &lm->color[3 * (size_t)lm->texels[i]];
        dst[0] = c[i].x; dst[1] = c[i].y; dst[2] = c[i].z;
    }
    lm->device_newer = false;
    return RL_SUCCESS;
}

// Uploads the valid lumels (position, normal, colour) when the host grid changed.
int32_t push_lumels(rl_lightmap* lm) {
    if (!lm->dirty) return RL_SUCCESS;
    lm->texels.clear();
    const uint32_t nt = (uint32_t)lm->w * (uint32_t)lm->h;
    for (uint32_t t = 0; t < nt; t++)
        if (lm->valid[t]) lm->texels.push_back(t);
    const uint32_t n = (uint32_t)lm->texels.size();
    if (n > lm->d_capacity) {
        lm->release();
        RL_HIP(dev_alloc(&lm->d_pos, n));
        RL_HIP(dev_alloc(&lm->d_nrm, n));
        RL_HIP(dev_alloc(&lm->d_color, n));
        RL_HIP(dev_alloc(&lm->d_counts, n));
        lm->d_capacity = n;
        RL_HIP(launch_fill_u32(lm->ctx->stream, lm->d_counts, n, 0u));
    }
    std::vector<float4> p(n), q(n), c(n);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t t = lm->texels[i];
        uint32_t tb = t, fb = lm->tri[t];
        float tw, fw;
        std::memcpy(&tw, &tb, 4);
        std::memcpy(&fw, &fb, 4);
        p[i] = make_float4(lm->pos[t].x, lm->pos[t].y, lm->pos[t].z, tw);
        q[i] = make_float4(lm->nrm[t].x, lm->nrm[t].y, lm->nrm[t].z, fw);
        c[i] = make_float4(lm->color[3 * t], lm->color[3 * t + 1], lm->color[3 * t + 2], 0.0f);
    }
    if (n) {
        RL_HIP(hipMemcpy(lm->d_pos, p.data(), n * sizeof(float4), hipMemcpyHostToDevice));
        RL_HIP(hipMemcpy(lm->d_nrm, q.data(), n * sizeof(float4), hipMemcpyHostToDevice));
        RL_HIP(hipMemcpy(lm->d_color, c.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    }
    lm->dirty = false;
    return RL_SUCCESS;
}

// Lightmap::initializeLumels for every face of mesh m (Lightmap.cpp:118-164; Face::isUvInside /
// positionAt / normalAt, scene/Mesh.cpp:397-476) on the GPU; the claimed texels are applied to the
// host grid (a texel keeps the last face of the mesh that contains it, over earlier meshes).
int32_t raster_mesh(rl_lightmap* lm, const MeshHost& m, uint32_t first_tri) {
    rl_relight* ctx = lm->ctx;
    const size_t texels = (size_t)lm->w * lm->h;
    if (m.num_tris == 0 || texels == 0) return RL_SUCCESS;
    if (lm->w > 16384 || lm->h > 16384) return fail(RL_INVALID_ARGUMENT, "lightmap side above 16384");
    if (ctx->raster_cap < texels) {
        dev_free(ctx->raster_owner);
        dev_free(ctx->raster_rec);
        ctx->raster_cap = 0;
        RL_HIP(dev_alloc(&ctx->raster_owner, texels));
        RL_HIP(dev_alloc(&ctx->raster_rec, texels));
        ctx->raster_cap = texels;
    }
    if (!ctx->raster_count) RL_HIP(dev_alloc(&ctx->raster_count, 1));
    float *dp = nullptr, *dn = nullptr, *du = nullptr;
    uint32_t* di = nullptr;
    auto cleanup = [&]() { dev_free(dp); dev_free(dn); dev_free(du); dev_free(di); };
    hipError_t e = dev_alloc(&dp, m.pos.size() * 3);
    if (e == hipSuccess) e = dev_alloc(&dn, m.nrm.size() * 3);
    if (e == hipSuccess) e = dev_alloc(&du, m.uv.size());
    if (e == hipSuccess) e = dev_alloc(&di, m.idx.size());
    if (e == hipSuccess) e = hipMemcpy(dp, m.pos.data(), m.pos.size() * sizeof(V3), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dn, m.nrm.data(), m.nrm.size() * sizeof(V3), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(du, m.uv.data(), m.uv.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(di, m.idx.data(), m.idx.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    RasterMesh M{dp, dn, du, di, m.num_tris, lm->w, lm->h, first_tri};
    if (e == hipSuccess)
        e = launch_raster(ctx->stream, M, ctx->max_blocks, ctx->raster_owner, ctx->raster_rec, ctx->raster_count,
                          ctx->d_queue);
    unsigned int n = 0;
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(&n, ctx->raster_count, sizeof(n), hipMemcpyDeviceToHost);
    std::vector<RasterRec> rec(n);
    if (e == hipSuccess && n) e = hipMemcpy(rec.data(), ctx->raster_rec, n * sizeof(RasterRec), hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return fail(RL_DEVICE_ERROR, std::string("texel raster: ") + hipGetErrorString(e));
    for (const RasterRec& r : rec) {
        const size_t t = r.texel;
        lm->valid[t] = 1;
        lm->tri[t] = r.tri;
        lm->pos[t] = v3(r.pos[0], r.pos[1], r.pos[2]);
        lm->nrm[t] = v3(r.nrm[0], r.nrm[1], r.nrm[2]);
        lm->color[3 * t] = lm->color[3 * t + 1] = lm->color[3 * t + 2] = 0.0f;
    }
    return RL_SUCCESS;
}

struct CallTimer {
    rl_relight* ctx;
    std::chrono::steady_clock::time_point t0;
    explicit CallTimer(rl_relight* c) : ctx(c), t0(std::chrono::steady_clock::now()) {}
};

hipEvent_t pool_event(rl_relight* ctx) {
    if (ctx->pool_used == ctx->pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->pool.push_back(e);
    }
    return ctx->pool[ctx->pool_used++];
}

// Records the start event of a kernel launch in `slot`; kernel_end records its end.
hipError_t kernel_begin(rl_relight* ctx, int slot) {
    rl_relight::Mark m{slot, pool_event(ctx), pool_event(ctx)};
    if (!m.a || !m.b) return hipErrorOutOfMemory;
    ctx->marks.push_back(m);
    return hipEventRecord(m.a, ctx->stream);
}

hipError_t kernel_end(rl_relight* ctx) { return hipEventRecord(ctx->marks.back().b, ctx->stream); }

int32_t begin_call(rl_relight* ctx) {
    ctx->stats = rl_stats{};
    ctx->progress_done = 0;
    ctx->progress_total = 0;
    ctx->pool_used = 0;
    ctx->marks.clear();
    RL_HIP(launch_fill_u32(ctx->stream, reinterpret_cast<uint32_t*>(ctx->d_rays), 2, 0u));
    RL_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    return RL_SUCCESS;
}

int32_t end_call(rl_relight* ctx, const CallTimer& tm) {
    RL_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    RL_HIP(hipEventSynchronize(ctx->ev1));
    ctx->progress_done = ctx->progress_total.load();
    float ms = 0.0f;
    RL_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    unsigned long long rays = 0;
    RL_HIP(hipMemcpy(&rays, ctx->d_rays, sizeof(rays), hipMemcpyDeviceToHost));
    ctx->stats.rays = rays;
    ctx->stats.gpu_ms = ms;
    for (const rl_relight::Mark& m : ctx->marks) {
        float km = 0.0f;
        RL_HIP(hipEventElapsedTime(&km, m.a, m.b));
        ctx->stats.kernel_ms[m.slot] += km;
        ctx->stats.kernel_launches[m.slot] += 1;
    }
    ctx->stats.wall_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tm.t0).count();
    return RL_SUCCESS;
}

int32_t check_scene_ready(rl_scene* scene) {
    if (!scene) return fail(RL_INVALID_ARGUMENT, "scene is NULL");
    if (scene->state != kReady) return fail(RL_INVALID_CALL, "scene has not been ended");
    return RL_SUCCESS;
}

// Builds the device segment table for a set of lightmaps; `per_lumel` work items per lumel.
// One unit of bake work: lumels [first, first + count) of a lightmap of a scene, with its AO
// settings. A whole lightmap is first = 0, count = its valid lumels.
struct BakeItem {
    rl_lightmap* lm;
    rl_scene* scene;
    AoParams ao;
    uint32_t first;
    uint32_t count;
    FinalGatherParams fg{};
    const int32_t* mesh_pm = nullptr;
    const PhotonMapDev* pms = nullptr;
};

// kUnitLumelAligned: lumels, each segment's range starting at a multiple of 64, so that a wavefront's
// 64 consecutive work items lie in one segment (the direct-light lumel kernel reads the segment, its
// scene and lights into scalar registers); the items in the gaps are skipped.
enum SegUnit { kUnitLumel = 0, kUnitPair = 1, kUnitChunk = 2, kUnitDirectPair = 3, kUnitLumelAligned = 4 };
// Segment-table slots: one per phase, so the phases of a pass can be queued back to back.
enum SegSlot { kSegDirect = 0, kSegAoWork = 1, kSegAoLumel = 2, kSegGather = 3, kSegDirectPairs = 4 };

// A device segment table for one launch (in one of the context's reusable buffers).
struct DeviceSegs {
    BakeSeg* d = nullptr;
    int n = 0;
    uint64_t total = 0;   // work items in `unit`
    uint64_t texels = 0;  // lumels covered
    int cap = 32;         // traversal-stack capacity for every scene in the table
    bool alpha = false;   // some scene in the table needs the alpha re-trace for shadow rays
};

AoParams ao_params(const rl_ao_settings& s) {
    AoParams P;
    P.samples = s.samples;
    P.occluded_fraction = s.occluded_fraction;
    P.max_distance = s.max_distance;
    P.exponent = s.exponent;
    P.seed = s.seed;
    P.sampling = s.sampling;
    return P;
}

int32_t check_ao(const rl_ao_settings* s) {
    if (!s) return fail(RL_INVALID_ARGUMENT, "AO settings are NULL");
    if (s->samples <= 0) return fail(RL_INVALID_ARGUMENT, "AO sample count must be positive");
    if (s->sampling != RL_SAMPLE_UNIFORM && s->sampling != RL_SAMPLE_COSINE)
        return fail(RL_INVALID_ARGUMENT, "unknown sampling mode");
    return RL_SUCCESS;
}

// Whole-lightmap items of `maps` in `scene`, uploading lumels where needed.
int32_t whole_items(rl_scene* scene, rl_lightmap* const* maps, int n, const AoParams& ao,
                    std::vector<BakeItem>& items) {
    for (int i = 0; i < n; i++) {
        rl_lightmap* lm = maps[i];
        if (!lm) return fail(RL_INVALID_ARGUMENT, "lightmap is NULL");
        if (lm->scene != scene) return fail(RL_INVALID_CALL, "lightmap holds no mesh of this scene");
        int32_t st = push_lumels(lm);
        if (st) return st;
        items.push_back(BakeItem{lm, scene, ao, 0u, (uint32_t)lm->texels.size()});
    }
    return RL_SUCCESS;
}

// Shadow rays per lumel: one per point, spot or directional light, one per vertex of an area light
// (the count the direct-light kernels derive from the scene's device lights), at least 1.
uint32_t direct_rays_per_lumel(const rl_scene* s) {
    uint32_t r = 0;
    for (const LightHost& L : s->lights) r += L.verts.empty() ? 1u : (uint32_t)L.verts.size();
    return std::max(r, 1u);
}

// The segment table of `items` in work units `unit` (out gets the counts; out.d stays unset).
std::vector<BakeSeg> build_segments(const std::vector<BakeItem>& items, int unit, uint32_t chunk_lumels,
                                    DeviceSegs& out) {
    std::vector<BakeSeg> segs;
    uint64_t acc = 0;
    for (const BakeItem& it : items) {
        if (it.count == 0) continue;
        rl_lightmap* lm = it.lm;
        BakeSeg s;
        s.pos = lm->d_pos + it.first;
        s.nrm = lm->d_nrm + it.first;
        s.color = lm->d_color + it.first;
        s.counts = lm->d_counts + it.first;
        s.count = it.count;
        s.lm_seed = lm->seed;
        if (unit == kUnitLumelAligned) acc = (acc + 63u) & ~uint64_t(63);
        s.work_begin = acc;
        s.scene = it.scene->dev();
        s.lights = it.scene->d_lights;
        s.lverts = it.scene->d_lverts;
        s.nlights = (int32_t)it.scene->lights.size();
        s.ao = it.ao;
        s.fg = it.fg;
        s.mesh_pm = it.mesh_pm;
        s.pms = it.pms;
        segs.push_back(s);
        if (unit == kUnitLumel || unit == kUnitLumelAligned) acc += it.count;
        else if (unit == kUnitPair) acc += (uint64_t)it.count * (uint64_t)it.ao.samples;
        else if (unit == kUnitDirectPair) acc += (uint64_t)it.count * direct_rays_per_lumel(it.scene);
        else acc += (it.count + chunk_lumels - 1) / chunk_lumels;
        out.texels += it.count;
        out.cap = std::max(out.cap, it.scene->cap);
        out.alpha = out.alpha || it.scene->alpha_shadows;
    }
    out.total = acc;
    out.n = (int)segs.size();
    return segs;
}

int32_t upload_segments(rl_relight* ctx, int slot, const std::vector<BakeItem>& items, int unit,
                        uint32_t chunk_lumels, DeviceSegs& out) {
    const std::vector<BakeSeg> segs = build_segments(items, unit, chunk_lumels, out);
    if (segs.empty()) return RL_SUCCESS;
    if (ctx->seg_cap[slot] < segs.size()) {
        RL_HIP(hipStreamSynchronize(ctx->stream));  // queued kernels may still read the old table
        dev_free(ctx->seg_buf[slot]);
        ctx->seg_cap[slot] = 0;
        const size_t cap = std::max<size_t>(segs.size(), 256);
        RL_HIP(dev_alloc(&ctx->seg_buf[slot], cap));
        ctx->seg_cap[slot] = cap;
    }
    out.d = ctx->seg_buf[slot];
    // The copy runs on the stream, so it lands after the kernels queued before it (which may still
    // read the slot's previous table) and before the kernels queued after it.
    if (ctx->seg_ev_live[slot]) RL_HIP(hipEventSynchronize(ctx->seg_ev[slot]));
    if (ctx->seg_host_cap[slot] < segs.size()) {
        if (ctx->seg_host[slot]) RL_HIP(hipHostFree(ctx->seg_host[slot]));
        ctx->seg_host[slot] = nullptr;
        ctx->seg_host_cap[slot] = 0;
        const size_t cap = std::max<size_t>(segs.size(), 256);
        RL_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->seg_host[slot]), cap * sizeof(BakeSeg), hipHostMallocDefault));
        ctx->seg_host_cap[slot] = cap;
    }
    if (!ctx->seg_ev[slot]) RL_HIP(hipEventCreateWithFlags(&ctx->seg_ev[slot], hipEventDisableTiming));
    std::memcpy(ctx->seg_host[slot], segs.data(), segs.size() * sizeof(BakeSeg));
    RL_HIP(hipMemcpyAsync(out.d, ctx->seg_host[slot], segs.size() * sizeof(BakeSeg), hipMemcpyHostToDevice,
                          ctx->stream));
    RL_HIP(hipEventRecord(ctx->seg_ev[slot], ctx->stream));
    ctx->seg_ev_live[slot] = true;
    return RL_SUCCESS;
}

// Lumels below which a direct-light pass runs as (lumel, shadow ray) pairs. City lumel ranges
// with a sun and a 16-vertex area light (lane per lumel / pairs): 64 lumels 33 / 30 us, 1,024
// 45 / 40 us, 8,192 278 / 84 us, 32,768 347 / 220 us, 65,536 454 / 325 us, 131,072 446 / 536 us,
// 524,288 466 / 1,816 us: the lumel kernel is bound by one lane's serial rays until the machine
// is full, the pair kernel grows with the rays.
constexpr uint64_t kDirectPairsBelow = 98304;

// The any-hit kernels' exact shortcuts the context's tuning enables (kShortcut* bits).
uint32_t shortcuts(const rl_relight* ctx) {
    return (ctx->tuning.local_trees ? kShortcutTrees : 0u) | (ctx->tuning.occluder_hints ? kShortcutHints : 0u);
}

// Whether a direct-light pass over `items` (`lumels` lumels) runs as pairs.
bool direct_uses_pairs(const rl_relight* ctx, const std::vector<BakeItem>& items, uint64_t lumels) {
    uint32_t rays_per_lumel = 1;
    for (const BakeItem& it : items) rays_per_lumel = std::max(rays_per_lumel, direct_rays_per_lumel(it.scene));
    if (lumels == 0 || rays_per_lumel < 2) return false;
    return ctx->tuning.direct_kernel == RL_DIRECT_KERNEL_PAIRS ||
           (ctx->tuning.direct_kernel == RL_DIRECT_KERNEL_AUTO && lumels < kDirectPairsBelow);
}

// Direct light over items (one launch, or two for a small pass). sync = false leaves the launch
// queued on the stream (a later synchronisation, e.g. end_call's, reports its errors).
int32_t run_direct_items(rl_relight* ctx, const std::vector<BakeItem>& items, bool sync = true) {
    DeviceSegs ds;
    int32_t st = upload_segments(ctx, kSegDirect, items, kUnitLumelAligned, 1, ds);
    if (st) return st;
    // A lane per lumel traces the lumel's shadow rays one after another; a pass with fewer
    // lumels than kDirectPairsBelow takes the pair kernels, which spread them over lanes.
    const bool pairs = direct_uses_pairs(ctx, items, ds.texels);
    DeviceSegs ps;
    if (pairs) {
        if ((st = upload_segments(ctx, kSegDirectPairs, items, kUnitDirectPair, 1, ps))) return st;
        if (ctx->direct_inf_cap < ps.total) {
            RL_HIP(hipStreamSynchronize(ctx->stream));  // queued kernels may still write the old buffer
            dev_free(ctx->direct_inf);
            ctx->direct_inf_cap = 0;
            RL_HIP(dev_alloc(&ctx->direct_inf, (size_t)ps.total));
            ctx->direct_inf_cap = ps.total;
        }
    }
    hipError_t e = hipSuccess;
    if (ds.total) {
        e = kernel_begin(ctx, RL_KERNEL_DIRECT);
        if (e == hipSuccess && pairs)
            e = launch_bake_direct_pairs(ctx->stream, ds.cap, ctx->max_blocks, ps.d, ds.d, ds.n, ps.total, ds.total,
                                         ds.alpha, shortcuts(ctx), ctx->direct_inf, ctx->d_queue, ctx->d_rays);
        if (e == hipSuccess && !pairs)
            e = launch_bake_direct(ctx->stream, ds.cap, ctx->max_blocks, ds.d, ds.n, ds.total, ds.alpha, shortcuts(ctx),
                                   ctx->d_queue, ctx->d_rays);
        if (e == hipSuccess) e = kernel_end(ctx);
    }
    ctx->stats.texels += ds.texels;
    ctx->stats.launches += ds.total ? (pairs ? 2 : 1) : 0;
    ctx->progress_total += ds.texels;
    if (e == hipSuccess && sync) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(RL_DEVICE_ERROR, std::string("direct bake: ") + hipGetErrorString(e));
    if (sync) ctx->progress_done += ds.texels;
    for (const BakeItem& it : items) it.lm->device_newer = true;
    return RL_SUCCESS;
}

// Lumels per AO work chunk: the refill kernel's chunks hold at most kAoChunkMax lumels. A pass with
// fewer chunks than the refill kernel has resident wavefronts (kAoRefillWavesPerCu per compute
// unit) halves the chunk, down to kAoSmallChunk lumels, so more wavefronts share it: a 13,120-lumel
// city pass at 64 samples takes 0.184 ms with 8-lumel chunks and 0.154 ms with 4 (2: 0.197 ms, the
// local-tree builds per chunk then cost more than the wavefronts gain); an 80,583-lumel pass,
// already more chunks than wavefronts at 8, takes 0.410 ms with 8 and 0.436 ms with 4.
uint32_t ao_chunk(const rl_relight* ctx, bool refill, const std::vector<BakeItem>& items) {
    if (!refill) return ctx->tuning.ao_chunk_lumels;
    uint32_t c = std::min(ctx->tuning.ao_chunk_lumels, kAoChunkMax);
    uint64_t chunks = 0;
    const uint64_t waves = (uint64_t)ctx->cus * kAoRefillWavesPerCu;
    for (;;) {
        chunks = 0;
        for (const BakeItem& it : items) chunks += ((uint64_t)it.count + c - 1) / c;
        if (c / 2 < kAoSmallChunk || chunks >= waves) return c;
        c /= 2;
    }
}

// AO over items: count pass (lane refill when every item has >= 4 samples, else chunked) and
// the finalize pass.
int32_t run_ao_items(rl_relight* ctx, const std::vector<BakeItem>& items, bool sync = true) {
    bool refill = ctx->tuning.ao_kernel == RL_AO_KERNEL_REFILL;
    for (const BakeItem& it : items) refill = refill && it.ao.samples >= 4;
    const uint32_t chunk = ao_chunk(ctx, refill, items);
    DeviceSegs work, lum;
    int32_t st = upload_segments(ctx, kSegAoWork, items, refill ? kUnitChunk : kUnitPair, chunk, work);
    if (st) return st;
    if ((st = upload_segments(ctx, kSegAoLumel, items, kUnitLumel, 1, lum))) return st;
    hipError_t e = hipSuccess;
    if (work.total) {
        e = kernel_begin(ctx, RL_KERNEL_AO_COUNT);
        if (e == hipSuccess && refill)
            e = launch_ao_refill(ctx->stream, work.cap, ctx->max_blocks, work.d, work.n, work.total, chunk,
                                 ctx->tuning.refill_min, shortcuts(ctx), ctx->d_queue, ctx->d_rays);
        if (e == hipSuccess && !refill) {
            for (const BakeItem& it : items)
                if (e == hipSuccess && it.count)
                    e = launch_fill_u32(ctx->stream, it.lm->d_counts + it.first, it.count, 0u);
            if (e == hipSuccess)
                e = launch_ao_count(ctx->stream, work.cap, ctx->max_blocks, work.d, work.n, work.total, ctx->d_queue,
                                    ctx->d_rays);
        }
        if (e == hipSuccess) e = kernel_end(ctx);
        if (e == hipSuccess) e = kernel_begin(ctx, RL_KERNEL_AO_FINALIZE);
        if (e == hipSuccess) e = launch_ao_finalize(ctx->stream, lum.d, lum.n, lum.total);
        if (e == hipSuccess) e = kernel_end(ctx);
    }
    ctx->stats.texels += lum.texels;
    ctx->stats.launches += work.total ? 2 : 0;
    ctx->progress_total += lum.texels;
    if (e == hipSuccess && sync) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(RL_DEVICE_ERROR, std::string("AO bake: ") + hipGetErrorString(e));
    if (sync) ctx->progress_done += lum.texels;
    for (const BakeItem& it : items) it.lm->device_newer = true;
    return RL_SUCCESS;
}

int32_t run_direct(rl_relight* ctx, rl_scene* scene, rl_lightmap* const* maps, int n, bool sync = true) {
    std::vector<BakeItem> items;
    AoParams none{};
    int32_t st = whole_items(scene, maps, n, none, items);
    if (st) return st;
    return run_direct_items(ctx, items, sync);
}

int32_t run_ao(rl_relight* ctx, rl_scene* scene, rl_lightmap* const* maps, int n,
               const rl_ao_settings* s, bool sync = true) {
    int32_t st = check_ao(s);
    if (st) return st;
    std::vector<BakeItem> items;
    if ((st = whole_items(scene, maps, n, ao_params(*s), items))) return st;
    return run_ao_items(ctx, items, sync);
}

// Final gather over items (each with its FG settings and its scene's photon tables).
int32_t run_fg_items(rl_relight* ctx, const std::vector<BakeItem>& items, bool sync = true) {
    DeviceSegs ds;
    int32_t st = upload_segments(ctx, kSegGather, items, kUnitLumel, 1, ds);
    if (st) return st;
    hipError_t e = hipSuccess;
    if (ds.total) {
        e = kernel_begin(ctx, RL_KERNEL_FINAL_GATHER);
        if (e == hipSuccess)
            e = launch_final_gather(ctx->stream, ds.cap, ctx->max_blocks, ds.d, ds.n, ds.total,
                                    ctx->tuning.fg_kernel, ctx->d_queue, ctx->d_rays);
        if (e == hipSuccess) e = kernel_end(ctx);
    }
    ctx->stats.texels += ds.texels;
    ctx->stats.launches += ds.total ? 1 : 0;
    ctx->progress_total += ds.texels;
    if (e == hipSuccess && sync) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(RL_DEVICE_ERROR, std::string("final gather: ") + hipGetErrorString(e));
    if (sync) ctx->progress_done += ds.texels;
    for (const BakeItem& it : items) it.lm->device_newer = true;
    return RL_SUCCESS;
}

FinalGatherParams fg_params(const rl_indirect_settings* s) {
    FinalGatherParams P;
    P.samples = s->final_gather_samples;
    P.distance = s->final_gather_distance;
    for (int k = 0; k < 3; k++) {
        P.sky[k] = s->sky_color[k];
        P.ambient[k] = s->ambient_color[k];
    }
    P.seed = s->seed;
    P.sampling = s->sampling;
    return P;
}

// Device tables of a scene's photonmaps (distinct maps in mesh order) and the mesh -> map index.
struct PhotonTables {
    std::vector<rl_lightmap*> maps;
    std::vector<int32_t> mesh_pm;
    std::vector<PhotonMapDev> dev;
    uint64_t total_texels = 0;
    int32_t* d_mesh_pm = nullptr;
    PhotonMapDev* d_pms = nullptr;
    ~PhotonTables() {
        dev_free(d_mesh_pm);
        dev_free(d_pms);
    }
};

int32_t ensure_photonmap_device(rl_lightmap* pm) {
    const size_t n = (size_t)pm->w * (size_t)pm->h;
    if (!pm->d_pm_color) {
        RL_HIP(dev_alloc(&pm->d_pm_color, n));
        RL_HIP(dev_alloc(&pm->d_pm_photons, n));
        RL_HIP(dev_alloc(&pm->d_pm_gathered, n));
        RL_HIP(dev_alloc(&pm->d_pm_valid, n));
        RL_HIP(hipMemset(pm->d_pm_color, 0, n * sizeof(float4)));
        RL_HIP(hipMemset(pm->d_pm_photons, 0, n * sizeof(uint32_t)));
        RL_HIP(hipMemset(pm->d_pm_gathered, 0, n * sizeof(float4)));
        pm->pm_valid_dirty = true;
    }
    if (pm->pm_valid_dirty) {
        RL_HIP(hipMemcpy(pm->d_pm_valid, pm->valid.data(), n, hipMemcpyHostToDevice));
        pm->pm_valid_dirty = false;
    }
    return RL_SUCCESS;
}

int32_t make_photon_tables(rl_scene* scene, PhotonTables& T) {
    T.mesh_pm.assign(scene->meshes.size(), -1);
    for (size_t i = 0; i < scene->meshes.size(); i++) {
        rl_lightmap* pm = scene->meshes[i].photonmap;
        if (!pm) continue;
        auto it = std::find(T.maps.begin(), T.maps.end(), pm);
        if (it == T.maps.end()) {
            T.maps.push_back(pm);
            it = T.maps.end() - 1;
        }
        T.mesh_pm[i] = (int32_t)(it - T.maps.begin());
    }
    for (rl_lightmap* pm : T.maps) {
        int32_t st = ensure_photonmap_device(pm);
        if (st) return st;
        PhotonMapDev d;
        d.color = pm->d_pm_color;
        d.photons = pm->d_pm_photons;
        d.gathered = pm->d_pm_gathered;
        d.valid = pm->d_pm_valid;
        d.w = pm->w;
        d.h = pm->h;
        d.key_base = (uint32_t)T.total_texels;
        d.pad = 0;
        T.total_texels += (uint64_t)pm->w * (uint64_t)pm->h;
        if (T.total_texels >= 0xFFFFFFFFull) return fail(RL_REJECTED, "photonmaps exceed 2^32 texels");
        T.dev.push_back(d);
    }
    RL_HIP(dev_alloc(&T.d_mesh_pm, T.mesh_pm.size()));
    RL_HIP(hipMemcpy(T.d_mesh_pm, T.mesh_pm.data(), T.mesh_pm.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!T.dev.empty()) {
        RL_HIP(dev_alloc(&T.d_pms, T.dev.size()));
        RL_HIP(hipMemcpy(T.d_pms, T.dev.data(), T.dev.size() * sizeof(PhotonMapDev), hipMemcpyHostToDevice));
    }
    return RL_SUCCESS;
}

int32_t check_indirect_settings(const rl_indirect_settings* s) {
    if (!s) return fail(RL_INVALID_ARGUMENT, "indirect settings are NULL");
    if (s->photon_pass_count < 0 || s->photon_bounce_count < 0 || s->final_gather_samples <= 0 ||
        s->final_gather_radius < 0)
        return fail(RL_INVALID_ARGUMENT, "bad indirect settings");
    if (s->sampling != RL_SAMPLE_UNIFORM && s->sampling != RL_SAMPLE_COSINE)
        return fail(RL_INVALID_ARGUMENT, "unknown sampling mode");
    if (s->termination != RL_PHOTON_THRESHOLD && s->termination != RL_PHOTON_RUSSIAN_ROULETTE)
        return fail(RL_INVALID_ARGUMENT, "unknown photon termination");
    if (s->gather_mode != RL_GATHER_TEXEL_DISC && s->gather_mode != RL_GATHER_KNN)
        return fail(RL_INVALID_ARGUMENT, "unknown gather mode");
    if (s->gather_mode == RL_GATHER_KNN &&
        (s->knn_k < 1 || s->knn_k > 64 || !(s->knn_radius > 0.0f) || !std::isfinite(s->knn_radius)))
        return fail(RL_INVALID_ARGUMENT, "k-NN gather needs knn_k in 1..64 and a positive finite knn_radius");
    return RL_SUCCESS;
}

}  // namespace

extern "C" {

const char* rl_last_error(void) { return g_error.c_str(); }

int32_t rl_create(rl_relight** out) {
    if (!out) return fail(RL_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(RL_DEVICE_ERROR, "no HIP device is available");
    rl_relight* ctx = new (std::nothrow) rl_relight();
    if (!ctx) return fail(RL_OUT_OF_MEMORY, "context allocation failed");
    hipGetDevice(&ctx->device);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        hipMalloc((void**)&ctx->d_rays, sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc((void**)&ctx->d_queue, 4 * sizeof(unsigned long long)) != hipSuccess) {
        delete ctx;
        return fail(RL_DEVICE_ERROR, "stream/event/counter creation failed");
    }
    // The static scratch of rl_trace_segments_noalloc: its address on this device, looked up once.
    if (hipError_t se = noalloc_setup(ctx->device); se != hipSuccess) {
        rl_destroy(ctx);
        return fail(RL_DEVICE_ERROR, std::string("scratch lookup: ") + hipGetErrorString(se));
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) == hipSuccess && cus > 0) {
        ctx->cus = cus;
        ctx->max_blocks = (uint32_t)cus * ctx->tuning.blocks_per_cu;
    }
    *out = ctx;
    return RL_SUCCESS;
}

int32_t rl_destroy(rl_relight* ctx) {
    if (!ctx) return RL_SUCCESS;
    if (ctx->d_rays) hipFree(ctx->d_rays);
    if (ctx->d_queue) hipFree(ctx->d_queue);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (BakeSeg* b : ctx->seg_buf)
        if (b) hipFree(b);
    for (BakeSeg* b : ctx->seg_host)
        if (b) hipHostFree(b);
    for (hipEvent_t e : ctx->seg_ev)
        if (e) hipEventDestroy(e);
    if (ctx->raster_owner) hipFree(ctx->raster_owner);
    if (ctx->raster_rec) hipFree(ctx->raster_rec);
    if (ctx->raster_count) hipFree(ctx->raster_count);
    if (ctx->direct_inf) hipFree(ctx->direct_inf);
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    for (hipEvent_t e : ctx->pool) hipEventDestroy(e);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return RL_SUCCESS;
}

int32_t rl_scene_create(rl_relight* ctx, rl_scene** out) {
    if (!ctx || !out) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    rl_scene* s = new (std::nothrow) rl_scene();
    if (!s) return fail(RL_OUT_OF_MEMORY, "scene allocation failed");
    s->ctx = ctx;
    *out = s;
    return RL_SUCCESS;
}

int32_t rl_scene_destroy(rl_scene* scene) {
    if (scene && scene->server_jobs.load() > 0)
        return fail(RL_INVALID_CALL, "scene is in use by an unfinished server job");
    delete scene;
    return RL_SUCCESS;
}

int32_t rl_scene_begin(rl_scene* scene) {
    if (!scene) return fail(RL_INVALID_ARGUMENT, "scene is NULL");
    if (scene->state != kInitial) return fail(RL_INVALID_CALL, "Scene::begin called twice");
    scene->state = kAdding;
    return RL_SUCCESS;
}

}  // extern "C"

namespace {

// Scene::addMesh (scene/Scene.cpp:98-107) with a Material (colour rgba, or (1,1,1,1) when albedo
// is NULL) or a TexturedMaterial (texture != NULL: colour rgb with alpha 1, diffuse uvs).
int32_t add_mesh_impl(rl_scene* scene, const float* positions, const float* normals, const float* lightmap_uvs,
                      const float* diffuse_uvs, uint32_t vertex_count, const uint32_t* indices, uint32_t index_count,
                      const float* transform, const float* albedo, int albedo_channels, const rl_texture* texture,
                      int32_t* out_mesh) {
    if (!scene) return fail(RL_INVALID_ARGUMENT, "scene is NULL");
    if (scene->state != kAdding) return fail(RL_INVALID_CALL, "add_mesh outside begin/end");
    if (!positions || !normals || !lightmap_uvs || !indices)
        return fail(RL_INVALID_ARGUMENT, "mesh arrays must not be NULL");
    if (index_count % 3 != 0) return fail(RL_INVALID_ARGUMENT, "index count is not a multiple of 3");
    for (uint32_t i = 0; i < index_count; i++)
        if (indices[i] >= vertex_count) return fail(RL_INVALID_ARGUMENT, "index out of range");
    MeshHost m;
    m.pos.resize(vertex_count);
    m.nrm.resize(vertex_count);
    m.uv.assign(lightmap_uvs, lightmap_uvs + 2 * (size_t)vertex_count);
    m.idx.assign(indices, indices + index_count);
    for (uint32_t i = 0; i < vertex_count; i++) {
        V3 p = v3(positions[3 * i], positions[3 * i + 1], positions[3 * i + 2]);
        V3 n = v3(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]);
        if (transform) {
            // Mesh::transform (scene/Mesh.cpp:277-284): Matrix4 * Vec3 and Matrix4::rotate
            // (Types.h:1188-1196,1303-1312), then the normal is renormalised.
            const float* M = transform;
            V3 tp = v3(p.x * M[0] + p.y * M[4] + p.z * M[8] + M[12],
                       p.x * M[1] + p.y * M[5] + p.z * M[9] + M[13],
                       p.x * M[2] + p.y * M[6] + p.z * M[10] + M[14]);
            V3 tn = v3(n.x * M[0] + n.y * M[4] + n.z * M[8],
                       n.x * M[1] + n.y * M[5] + n.z * M[9],
                       n.x * M[2] + n.y * M[6] + n.z * M[10]);
            normalize(tn);
            p = tp;
            n = tn;
        }
        m.pos[i] = p;
        m.nrm[i] = n;
    }
    if (albedo)
        for (int k = 0; k < albedo_channels; k++) m.albedo[k] = albedo[k];
    if (texture) {
        m.duv.assign(diffuse_uvs, diffuse_uvs + 2 * (size_t)vertex_count);
        int32_t ti = -1;
        for (size_t k = 0; k < scene->textures.size(); k++)
            if (scene->textures[k] == texture->t) ti = (int32_t)k;
        if (ti < 0) {
            ti = (int32_t)scene->textures.size();
            scene->textures.push_back(texture->t);
        }
        m.tex = ti;
    }
    m.num_tris = index_count / 3;
    scene->meshes.push_back(std::move(m));
    if (out_mesh) *out_mesh = (int32_t)scene->meshes.size() - 1;
    return RL_SUCCESS;
}

}  // namespace

extern "C" {

int32_t rl_scene_add_mesh(rl_scene* scene, const float* positions, const float* normals,
                          const float* lightmap_uvs, uint32_t vertex_count,
                          const uint32_t* indices, uint32_t index_count, const float* transform,
                          const float* albedo, int32_t* out_mesh) {
    return add_mesh_impl(scene, positions, normals, lightmap_uvs, nullptr, vertex_count, indices, index_count,
                         transform, albedo, 4, nullptr, out_mesh);
}

int32_t rl_texture_create(rl_relight* ctx, int32_t width, int32_t height, int32_t channels, const uint8_t* pixels,
                          rl_texture** out) {
    if (!ctx || !out || !pixels) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (width <= 0 || height <= 0) return fail(RL_INVALID_ARGUMENT, "texture size must be positive");
    if (channels != 3 && channels != 4) return fail(RL_INVALID_ARGUMENT, "textures have 3 or 4 channels");
    auto t = std::make_shared<TexHost>();
    t->w = width;
    t->h = height;
    t->channels = channels;
    t->px.assign(pixels, pixels + (size_t)width * (size_t)height * (size_t)channels);
    if (channels == 4)
        for (size_t i = 3; i < t->px.size(); i += 4)
            if (t->px[i] / 255.0f <= 0.1f) t->cutout = true;
    rl_texture* h = new (std::nothrow) rl_texture();
    if (!h) return fail(RL_OUT_OF_MEMORY, "texture allocation failed");
    h->t = std::move(t);
    *out = h;
    return RL_SUCCESS;
}

int32_t rl_texture_destroy(rl_texture* tex) {
    delete tex;
    return RL_SUCCESS;
}

int32_t rl_scene_add_mesh_textured(rl_scene* scene, const float* positions, const float* normals,
                                   const float* lightmap_uvs, const float* diffuse_uvs, uint32_t vertex_count,
                                   const uint32_t* indices, uint32_t index_count, const float* transform,
                                   const float* color, const rl_texture* texture, int32_t* out_mesh) {
    if (!texture) return fail(RL_INVALID_ARGUMENT, "texture is NULL");
    if (!diffuse_uvs) return fail(RL_INVALID_ARGUMENT, "diffuse uvs are NULL");
    // TexturedMaterial(texture, color): Material::colorAt is Rgba(color) with alpha 1.
    const float white[3] = {1.0f, 1.0f, 1.0f};
    return add_mesh_impl(scene, positions, normals, lightmap_uvs, diffuse_uvs, vertex_count, indices, index_count,
                         transform, color ? color : white, 3, texture, out_mesh);
}

int32_t rl_scene_add_light(rl_scene* scene, const rl_light_desc* d, int32_t* out_light) {
    if (!scene || !d) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    if (scene->state == kReady) return fail(RL_INVALID_CALL, "add_light after end");
    if (d->type < RL_LIGHT_POINT || d->type > RL_LIGHT_AREA) return fail(RL_INVALID_ARGUMENT, "unknown light type");
    LightHost L;
    std::memset(&L.dev, 0, sizeof(L.dev));
    L.dev.type = d->type;
    L.dev.casts_shadow = d->casts_shadow ? 1 : 0;
    for (int k = 0; k < 3; k++) {
        L.dev.pos[k] = d->position[k];
        L.dev.dir[k] = d->direction[k];
        L.dev.color[k] = d->color[k];
    }
    L.dev.intensity = d->intensity;
    L.dev.has_attenuation = (d->type != RL_LIGHT_DIRECTIONAL && d->has_attenuation) ? 1 : 0;
    L.dev.radius = d->radius;
    L.dev.att_c = d->att_constant;
    L.dev.att_l = d->att_linear;
    L.dev.att_q = d->att_quadratic;
    L.dev.has_spot = d->type == RL_LIGHT_SPOT ? 1 : 0;
    L.dev.cutoff = d->spot_cutoff;
    L.dev.spot_exponent = d->spot_exponent;
    if (d->type == RL_LIGHT_AREA) {
        if (!d->area_positions || !d->area_indices || d->area_index_count % 3 != 0 || d->area_index_count == 0)
            return fail(RL_INVALID_ARGUMENT, "area light needs a triangle mesh");
        for (uint32_t i = 0; i < d->area_index_count; i++)
            if (d->area_indices[i] >= d->area_vertex_count) return fail(RL_INVALID_ARGUMENT, "area index out of range");
        // FaceLightVertexGenerator::generateFromTriangle at subdivision 0 pushes
        // Triangle(face).centroid(); Triangle(face) takes face.vertex(0..2) = a, c, b
        // (scene/Mesh.cpp:322-327,427-437), so the sum is (a + c) + b.
        const float* P = d->area_positions;
        for (uint32_t f = 0; f < d->area_index_count / 3; f++) {
            const uint32_t ia = d->area_indices[3 * f], ib = d->area_indices[3 * f + 1], ic = d->area_indices[3 * f + 2];
            V3 a = v3(P[3 * ia], P[3 * ia + 1], P[3 * ia + 2]);
            V3 b = v3(P[3 * ib], P[3 * ib + 1], P[3 * ib + 2]);
            V3 c = v3(P[3 * ic], P[3 * ic + 1], P[3 * ic + 2]);
            L.verts.push_back((a + c + b) / 3.0f);
        }
    }
    scene->lights.push_back(std::move(L));
    if (out_light) *out_light = (int32_t)scene->lights.size() - 1;
    return RL_SUCCESS;
}

int32_t rl_scene_end(rl_scene* scene) {
    if (!scene) return fail(RL_INVALID_ARGUMENT, "scene is NULL");
    if (scene->state != kAdding) return fail(RL_INVALID_CALL, "Scene::end without begin");
    uint64_t ntri = 0;
    for (MeshHost& m : scene->meshes) {
        m.first_tri = (uint32_t)ntri;
        ntri += m.num_tris;
    }
    if (ntri == 0) return fail(RL_INVALID_CALL, "scene has no triangles");
    if (ntri > RL_LEAF_FIRST_MASK) return fail(RL_REJECTED, "too many triangles");
    std::vector<float> tri9((size_t)ntri * 9);
    std::vector<TriAttr> attr((size_t)ntri);
    scene->tri_mesh.resize(ntri);
    scene->tri_face.resize(ntri);
    for (uint32_t mi = 0; mi < scene->meshes.size(); mi++) {
        const MeshHost& m = scene->meshes[mi];
        for (uint32_t f = 0; f < m.num_tris; f++) {
            const uint32_t t = m.first_tri + f;
            for (int k = 0; k < 3; k++) {
                const uint32_t vi = m.idx[3 * f + k];
                tri9[9 * (size_t)t + 3 * k + 0] = m.pos[vi].x;
                tri9[9 * (size_t)t + 3 * k + 1] = m.pos[vi].y;
                tri9[9 * (size_t)t + 3 * k + 2] = m.pos[vi].z;
                attr[t].uv[2 * k] = m.uv[2 * vi];
                attr[t].uv[2 * k + 1] = m.uv[2 * vi + 1];
                attr[t].n[3 * k] = m.nrm[vi].x;
                attr[t].n[3 * k + 1] = m.nrm[vi].y;
                attr[t].n[3 * k + 2] = m.nrm[vi].z;
            }
            attr[t].mesh = mi;
            attr[t].face = f;
            std::memcpy(attr[t].albedo, m.albedo, sizeof(m.albedo));
            scene->tri_mesh[t] = mi;
            scene->tri_face[t] = f;
        }
    }
    // Scene::bounds (scene/Scene.cpp:81-88): union of the meshes' vertex bounds.
    for (int k = 0; k < 3; k++) {
        scene->bounds_lo[k] = tri9[k];
        scene->bounds_hi[k] = tri9[k];
    }
    for (size_t i = 0; i < tri9.size(); i += 3) {
        for (int k = 0; k < 3; k++) {
            scene->bounds_lo[k] = std::min(scene->bounds_lo[k], tri9[i + k]);
            scene->bounds_hi[k] = std::max(scene->bounds_hi[k], tri9[i + k]);
        }
    }
    Bvh4Build bvh;
    int rc = build_bvh4(tri9.data(), (uint32_t)ntri, bvh);
    if (rc) return fail(RL_REJECTED, "BVH build failed");
    int cap = stack_capacity_for(bvh.stack_bound);
    if (cap < 0) return fail(RL_REJECTED, "BVH traversal stack bound exceeds the LDS stack");
    // Triangles in leaf order twice: 48 B each (v0, v1, v2 as float4, the scene-order id in v0.w)
    // for the closest-hit traversals, and 64 B each for the any-hit ones, vertex k as
    // (x, y, z, x, y) at floats 5k .. 5k + 4 (a ray's watertight component order is a rotation of
    // (x, y, z), read as three consecutive floats) and the id in float 15.
    std::vector<float4> tris((size_t)ntri * 3), tris_rot((size_t)ntri * 4);
    for (uint32_t i = 0; i < ntri; i++) {
        const uint32_t t = bvh.order[i];
        const float* v = &tri9[9 * (size_t)t];
        float idw;
        std::memcpy(&idw, &t, 4);
        tris[3 * (size_t)i + 0] = make_float4(v[0], v[1], v[2], idw);
        tris[3 * (size_t)i + 1] = make_float4(v[3], v[4], v[5], 0.0f);
        tris[3 * (size_t)i + 2] = make_float4(v[6], v[7], v[8], 0.0f);
        float* f = reinterpret_cast<float*>(&tris_rot[4 * (size_t)i]);
        for (int k = 0; k < 3; k++) {
            const float* p = v + 3 * k;
            f[5 * k + 0] = p[0];
            f[5 * k + 1] = p[1];
            f[5 * k + 2] = p[2];
            f[5 * k + 3] = p[0];
            f[5 * k + 4] = p[1];
        }
        f[15] = idw;
    }
    // Textured meshes: per-triangle diffuse uvs and texture, and the textures' pixels.
    bool textured = false, alpha = false;
    for (const MeshHost& m : scene->meshes) {
        if (m.tex >= 0) {
            textured = true;
            if (scene->textures[m.tex]->cutout) alpha = true;
        } else if (m.albedo[3] <= 0.1f && m.num_tris > 0) {
            alpha = true;
        }
    }
    std::vector<TriMat> mats;
    std::vector<TexDev> texs;
    std::vector<uint8_t> texpx;
    if (textured) {
        mats.resize(ntri);
        for (const MeshHost& m : scene->meshes) {
            for (uint32_t f = 0; f < m.num_tris; f++) {
                TriMat& tm = mats[m.first_tri + f];
                tm.tex = m.tex;
                tm.pad = 0;
                for (int k = 0; k < 3; k++) {
                    const uint32_t vi = m.idx[3 * f + k];
                    tm.duv[2 * k] = m.tex >= 0 ? m.duv[2 * vi] : 0.0f;
                    tm.duv[2 * k + 1] = m.tex >= 0 ? m.duv[2 * vi + 1] : 0.0f;
                }
            }
        }
        std::vector<size_t> offs;
        for (const auto& t : scene->textures) {
            offs.push_back(texpx.size());
            texpx.insert(texpx.end(), t->px.begin(), t->px.end());
        }
        for (size_t k = 0; k < scene->textures.size(); k++) {
            TexDev d{};
            d.w = scene->textures[k]->w;
            d.h = scene->textures[k]->h;
            d.channels = scene->textures[k]->channels;
            texs.push_back(d);
        }
    }
    std::vector<LightDev> lights;
    std::vector<float4> lverts;
    for (LightHost& L : scene->lights) {
        L.dev.num_vertices = (int32_t)L.verts.size();
        L.dev.vertex_offset = (int32_t)lverts.size();
        for (const V3& v : L.verts) lverts.push_back(make_float4(v.x, v.y, v.z, 0.0f));
        lights.push_back(L.dev);
    }
    RL_HIP(dev_alloc(&scene->d_nodes, bvh.nodes.size()));
    RL_HIP(dev_alloc(&scene->d_tris, tris.size()));
    RL_HIP(dev_alloc(&scene->d_tris_rot, tris_rot.size()));
    RL_HIP(dev_alloc(&scene->d_attr, attr.size()));
    RL_HIP(hipMemcpy(scene->d_nodes, bvh.nodes.data(), bvh.nodes.size() * sizeof(Bvh4Node), hipMemcpyHostToDevice));
    RL_HIP(hipMemcpy(scene->d_tris, tris.data(), tris.size() * sizeof(float4), hipMemcpyHostToDevice));
    RL_HIP(hipMemcpy(scene->d_tris_rot, tris_rot.data(), tris_rot.size() * sizeof(float4), hipMemcpyHostToDevice));
    RL_HIP(hipMemcpy(scene->d_attr, attr.data(), attr.size() * sizeof(TriAttr), hipMemcpyHostToDevice));
    if (!lights.empty()) {
        RL_HIP(dev_alloc(&scene->d_lights, lights.size()));
        RL_HIP(hipMemcpy(scene->d_lights, lights.data(), lights.size() * sizeof(LightDev), hipMemcpyHostToDevice));
    }
    if (!lverts.empty()) {
        RL_HIP(dev_alloc(&scene->d_lverts, lverts.size()));
        RL_HIP(hipMemcpy(scene->d_lverts, lverts.data(), lverts.size() * sizeof(float4), hipMemcpyHostToDevice));
    }
    if (textured) {
        RL_HIP(dev_alloc(&scene->d_texpx, texpx.size()));
        RL_HIP(hipMemcpy(scene->d_texpx, texpx.data(), texpx.size(), hipMemcpyHostToDevice));
        size_t off = 0;
        for (size_t k = 0; k < texs.size(); k++) {
            texs[k].px = scene->d_texpx + off;
            off += scene->textures[k]->px.size();
        }
        RL_HIP(dev_alloc(&scene->d_texs, texs.size()));
        RL_HIP(hipMemcpy(scene->d_texs, texs.data(), texs.size() * sizeof(TexDev), hipMemcpyHostToDevice));
        RL_HIP(dev_alloc(&scene->d_mat, mats.size()));
        RL_HIP(hipMemcpy(scene->d_mat, mats.data(), mats.size() * sizeof(TriMat), hipMemcpyHostToDevice));
    }
    scene->alpha_shadows = alpha;
    scene->num_tris = (uint32_t)ntri;
    scene->num_nodes = (uint32_t)bvh.nodes.size();
    scene->stack_bound = bvh.stack_bound;
    scene->cap = cap;
    scene->state = kReady;
    return RL_SUCCESS;
}

int32_t rl_scene_info(const rl_scene* scene, uint32_t* triangles, uint32_t* nodes, int32_t* stack_bound) {
    if (!scene) return fail(RL_INVALID_ARGUMENT, "scene is NULL");
    if (triangles) *triangles = scene->num_tris;
    if (nodes) *nodes = scene->num_nodes;
    if (stack_bound) *stack_bound = scene->stack_bound;
    return RL_SUCCESS;
}

int32_t rl_lightmap_create(rl_relight* ctx, int32_t width, int32_t height, rl_lightmap** out) {
    if (!ctx || !out) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    if (width <= 0 || height <= 0 || (int64_t)width * height > (1ll << 31))
        return fail(RL_INVALID_ARGUMENT, "bad lightmap size");
    rl_lightmap* lm = new (std::nothrow) rl_lightmap();
    if (!lm) return fail(RL_OUT_OF_MEMORY, "lightmap allocation failed");
    lm->ctx = ctx;
    lm->w = width;
    lm->h = height;
    lm->seed = ctx->next_lm_seed++;
    const size_t n = (size_t)width * height;
    lm->valid.assign(n, 0);
    lm->tri.assign(n, 0xFFFFFFFFu);
    lm->pos.assign(n, v3(0, 0, 0));
    lm->nrm.assign(n, v3(0, 0, 0));
    lm->color.assign(3 * n, 0.0f);
    *out = lm;
    return RL_SUCCESS;
}

int32_t rl_lightmap_destroy(rl_lightmap* lm) {
    if (lm && lm->server_jobs.load() > 0)
        return fail(RL_INVALID_CALL, "lightmap is in use by an unfinished server job");
    delete lm;
    return RL_SUCCESS;
}

int32_t rl_lightmap_set_seed(rl_lightmap* lm, uint32_t seed) {
    if (!lm) return fail(RL_INVALID_ARGUMENT, "lightmap is NULL");
    lm->seed = seed;
    return RL_SUCCESS;
}

int32_t rl_lightmap_add_mesh(rl_lightmap* lm, rl_scene* scene, int32_t mesh) {
    if (!lm || !scene) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    if (mesh < 0 || mesh >= (int32_t)scene->meshes.size()) return fail(RL_INVALID_ARGUMENT, "bad mesh index");
    if (lm->scene && lm->scene != scene) return fail(RL_INVALID_CALL, "lightmap already holds another scene's meshes");
    MeshHost& m = scene->meshes[mesh];
    // Lightmap::addMesh refuses a mesh that already has a lightmap (Lightmap.cpp:95-97).
    if (m.lightmap) return fail(RL_INVALID_CALL, "mesh already has a lightmap");
    int32_t st = pull_colors(lm);
    if (st) return st;
    m.lightmap = lm;
    lm->scene = scene;
    // The scene-order triangle id is assigned at end(); until then faces are numbered per mesh.
    uint32_t first = 0;
    for (int32_t i = 0; i < mesh; i++) first += scene->meshes[i].num_tris;
    if ((st = raster_mesh(lm, m, first))) return st;
    lm->dirty = true;
    return RL_SUCCESS;
}

int32_t rl_lightmap_valid_count(const rl_lightmap* lm, uint32_t* count) {
    if (!lm || !count) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    uint32_t c = 0;
    for (uint8_t v : lm->valid) c += v;
    *count = c;
    return RL_SUCCESS;
}

int32_t rl_lightmap_read_lumels(rl_lightmap* lm, uint32_t* texel, float* position, float* normal) {
    if (!lm) return fail(RL_INVALID_ARGUMENT, "lightmap is NULL");
    size_t k = 0;
    for (size_t t = 0; t < lm->valid.size(); t++) {
        if (!lm->valid[t]) continue;
        if (texel) texel[k] = (uint32_t)t;
        if (position) { position[3 * k] = lm->pos[t].x; position[3 * k + 1] = lm->pos[t].y; position[3 * k + 2] = lm->pos[t].z; }
        if (normal) { normal[3 * k] = lm->nrm[t].x; normal[3 * k + 1] = lm->nrm[t].y; normal[3 * k + 2] = lm->nrm[t].z; }
        k++;
    }
    return RL_SUCCESS;
}

int32_t rl_lightmap_read_hdr(rl_lightmap* lm, float* out_rgb) {
    if (!lm || !out_rgb) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    int32_t st = pull_colors(lm);
    if (st) return st;
    std::memcpy(out_rgb, lm->color.data(), lm->color.size() * sizeof(float));
    return RL_SUCCESS;
}

}  // extern "C"

namespace {

// The full colour grid and validity of a lightmap on the device, for the post-process kernels.
struct GridDev {
    float4* grid = nullptr;
    uint8_t* valid = nullptr;
    ~GridDev() {
        dev_free(grid);
        dev_free(valid);
    }
};

int32_t grid_upload(rl_lightmap* lm, GridDev& G) {
    int32_t st = pull_colors(lm);
    if (st) return st;
    const size_t n = (size_t)lm->w * lm->h;
    std::vector<float4> h(n);
    for (size_t i = 0; i < n; i++) h[i] = make_float4(lm->color[3 * i], lm->color[3 * i + 1], lm->color[3 * i + 2], 0.0f);
    RL_HIP(dev_alloc(&G.grid, n));
    RL_HIP(dev_alloc(&G.valid, n));
    RL_HIP(hipMemcpy(G.grid, h.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    RL_HIP(hipMemcpy(G.valid, lm->valid.data(), n, hipMemcpyHostToDevice));
    return RL_SUCCESS;
}

// Copies a post-processed grid back into the lightmap (host grid; lumels re-uploaded on the next bake).
int32_t grid_download(rl_lightmap* lm, const GridDev& G) {
    const size_t n = (size_t)lm->w * lm->h;
    std::vector<float4> h(n);
    RL_HIP(hipStreamSynchronize(lm->ctx->stream));
    RL_HIP(hipMemcpy(h.data(), G.grid, n * sizeof(float4), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) {
        lm->color[3 * i] = h[i].x;
        lm->color[3 * i + 1] = h[i].y;
        lm->color[3 * i + 2] = h[i].z;
    }
    lm->dirty = true;
    lm->device_newer = false;
    return RL_SUCCESS;
}

int32_t post_process(rl_lightmap* lm, bool blur) {
    if (!lm) return fail(RL_INVALID_ARGUMENT, "lightmap is NULL");
    if (lm->is_photonmap) return fail(RL_INVALID_CALL, "not a lightmap");
    GridDev G;
    int32_t st = grid_upload(lm, G);
    if (st) return st;
    const hipError_t e = blur ? launch_blur(lm->ctx->stream, G.grid, G.valid, lm->w, lm->h)
                              : launch_expand(lm->ctx->stream, G.grid, G.valid, lm->w, lm->h);
    if (e != hipSuccess) return fail(RL_DEVICE_ERROR, std::string("post-process: ") + hipGetErrorString(e));
    return grid_download(lm, G);
}

int32_t encode_ldr(rl_lightmap* lm, int32_t format, std::vector<uint8_t>& out) {
    const int ch = format == RL_FORMAT_TGA_DOUBLE_LDR ? 3 : 4;
    const size_t n = (size_t)lm->w * lm->h;
    GridDev G;
    int32_t st = grid_upload(lm, G);
    if (st) return st;
    uint8_t* d = nullptr;
    hipError_t e = dev_alloc(&d, n * ch);
    if (e == hipSuccess) e = launch_encode(lm->ctx->stream, G.grid, n, format, d);
    if (e == hipSuccess) e = hipStreamSynchronize(lm->ctx->stream);
    out.resize(n * ch);
    if (e == hipSuccess) e = hipMemcpy(out.data(), d, n * ch, hipMemcpyDeviceToHost);
    dev_free(d);
    if (e != hipSuccess) return fail(RL_DEVICE_ERROR, std::string("encode: ") + hipGetErrorString(e));
    return RL_SUCCESS;
}

}  // namespace

extern "C" {

int32_t rl_lightmap_expand(rl_lightmap* lm) { return post_process(lm, false); }

int32_t rl_lightmap_blur(rl_lightmap* lm) { return post_process(lm, true); }

int32_t rl_lightmap_encode(rl_lightmap* lm, int32_t format, uint8_t* out) {
    if (!lm || !out) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    if (format != RL_FORMAT_TGA_DOUBLE_LDR && format != RL_FORMAT_TGA_RGBM)
        return fail(RL_INVALID_ARGUMENT, "encode takes RL_FORMAT_TGA_DOUBLE_LDR or RL_FORMAT_TGA_RGBM");
    std::vector<uint8_t> px;
    int32_t st = encode_ldr(lm, format, px);
    if (st) return st;
    std::memcpy(out, px.data(), px.size());
    return RL_SUCCESS;
}

int32_t rl_lightmap_save(rl_lightmap* lm, const char* path, int32_t format) {
    if (!lm || !path) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    if (format != RL_FORMAT_RAW_HDR && format != RL_FORMAT_TGA_DOUBLE_LDR && format != RL_FORMAT_TGA_RGBM)
        return fail(RL_INVALID_ARGUMENT, "unknown storage format");
    std::vector<uint8_t> bytes;
    const int32_t w = lm->w, h = lm->h;
    if (format == RL_FORMAT_RAW_HDR) {
        int32_t st = pull_colors(lm);
        if (st) return st;
        bytes.resize(8 + lm->color.size() * sizeof(float));
        std::memcpy(bytes.data(), &w, 4);
        std::memcpy(bytes.data() + 4, &h, 4);
        std::memcpy(bytes.data() + 8, lm->color.data(), lm->color.size() * sizeof(float));
    } else {
        std::vector<uint8_t> px;
        int32_t st = encode_ldr(lm, format, px);
        if (st) return st;
        const int ch = format == RL_FORMAT_TGA_DOUBLE_LDR ? 3 : 4;
        // Uncompressed true-colour TGA: 12-byte header, width, height, bits per pixel, descriptor 0,
        // then pixels bottom-up as stored, b g r (a).
        const uint8_t head[12] = {0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        const uint8_t info[6] = {(uint8_t)(w % 256), (uint8_t)(w / 256), (uint8_t)(h % 256), (uint8_t)(h / 256),
                                 (uint8_t)(ch * 8), 0};
        bytes.insert(bytes.end(), head, head + 12);
        bytes.insert(bytes.end(), info, info + 6);
        const size_t n = (size_t)w * h;
        for (size_t i = 0; i < n; i++) {
            const uint8_t* p = &px[i * ch];
            bytes.push_back(p[2]);
            bytes.push_back(p[1]);
            bytes.push_back(p[0]);
            if (ch == 4) bytes.push_back(p[3]);
        }
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(RL_INVALID_ARGUMENT, std::string("cannot open ") + path);
    const size_t wrote = std::fwrite(bytes.data(), 1, bytes.size(), f);
    std::fclose(f);
    if (wrote != bytes.size()) return fail(RL_DEVICE_ERROR, std::string("short write to ") + path);
    return RL_SUCCESS;
}

}  // extern "C"

namespace {

// The lumels of `lm` a BakeIterator visits for `mesh` (baker/Baker.cpp:148-234), as indices into
// the lightmap's valid lumels, in visiting order: LumelBakeIterator(first, step) takes texel
// indices first, first + step, ... of the whole lightmap (every mesh's texels, as the reference
// iterates lightmap->lumels()); FaceBakeIterator(first, step) takes the lumels owned by faces
// first, first + step, ... of `mesh`, face by face. `steps` gets the number of iterator steps
// (indices visited), `items` the iterator's itemCount.
int32_t iterator_selection(const rl_scene* scene, int32_t mesh, const rl_lightmap* lm, const rl_bake_iterator& it,
                           std::vector<uint32_t>& sel, std::vector<uint32_t>& step_of, int64_t& steps,
                           int64_t& items) {
    if (it.first < 0 || it.step < 1) return fail(RL_INVALID_ARGUMENT, "iterator needs first >= 0 and step >= 1");
    const MeshHost& m = scene->meshes[mesh];
    sel.clear();
    step_of.clear();
    if (it.kind == RL_ITERATE_LUMELS) {
        items = (int64_t)lm->w * lm->h;
        steps = it.first < items ? (items - it.first + it.step - 1) / it.step : 0;
        for (uint32_t j = 0; j < (uint32_t)lm->texels.size(); j++) {
            const int64_t t = lm->texels[j];
            if (t >= it.first && (t - it.first) % it.step == 0) {
                sel.push_back(j);
                step_of.push_back((uint32_t)((t - it.first) / it.step));
            }
        }
        return RL_SUCCESS;
    }
    if (it.kind != RL_ITERATE_FACES) return fail(RL_INVALID_ARGUMENT, "unknown iterator kind");
    items = m.num_tris;
    steps = it.first < items ? (items - it.first + it.step - 1) / it.step : 0;
    std::vector<std::pair<uint32_t, uint32_t>> by_face;  // (iterator step, lumel)
    for (uint32_t j = 0; j < (uint32_t)lm->texels.size(); j++) {
        const uint32_t tri = lm->tri[lm->texels[j]];
        if (tri < m.first_tri || tri >= m.first_tri + m.num_tris) continue;
        const int64_t f = tri - m.first_tri;
        if (f >= it.first && (f - it.first) % it.step == 0) by_face.push_back({(uint32_t)((f - it.first) / it.step), j});
    }
    std::stable_sort(by_face.begin(), by_face.end(),
                     [](const std::pair<uint32_t, uint32_t>& a, const std::pair<uint32_t, uint32_t>& b) {
                         return a.first < b.first;
                     });
    for (const auto& p : by_face) {
        step_of.push_back(p.first);
        sel.push_back(p.second);
    }
    return RL_SUCCESS;
}

// Bakes the selected lumels of lm (indices into its valid lumels) with `bake`, which runs on the
// selection copied to a compact segment (same texel ids and lightmap seed, so the same sample
// streams), then writes the colours back. With a progress callback the selection is baked in
// progress->updates parts, and Progress::notify's (mesh, step, stepCount) follows each part
// (steps of the iterator whose lumels are all baked), then the final (items, items) as
// Baker::bakeMesh ends (baker/Baker.cpp:71-74).
template <class Bake>
int32_t bake_selection(rl_relight* ctx, rl_scene* scene, int32_t mesh, rl_lightmap* lm,
                       const std::vector<uint32_t>& sel, const std::vector<uint32_t>& step_of, int64_t steps,
                       int64_t items, const rl_progress* progress, const AoParams& ao, Bake bake) {
    int32_t st = RL_SUCCESS;
    const uint32_t n = (uint32_t)sel.size();
    const uint32_t parts = progress && progress->updates > 1 ? std::min<uint32_t>((uint32_t)progress->updates,
                                                                                    std::max<uint32_t>(n, 1u))
                                                              : 1u;
    if (n) {
        rl_lightmap tmp;
        tmp.ctx = ctx;
        tmp.scene = scene;
        tmp.seed = lm->seed;
        tmp.w = lm->w;
        tmp.h = lm->h;
        tmp.dirty = false;
        uint32_t* d_sel = nullptr;
        hipError_t e = dev_alloc(&d_sel, n);
        if (e == hipSuccess) e = dev_alloc(&tmp.d_pos, n);
        if (e == hipSuccess) e = dev_alloc(&tmp.d_nrm, n);
        if (e == hipSuccess) e = dev_alloc(&tmp.d_color, n);
        if (e == hipSuccess) e = dev_alloc(&tmp.d_counts, n);
        tmp.d_capacity = n;
        if (e == hipSuccess) e = hipMemcpy(d_sel, sel.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = launch_gather_lumels(ctx->stream, d_sel, n, lm->d_pos, lm->d_nrm, lm->d_color, tmp.d_pos, tmp.d_nrm,
                                     tmp.d_color);
        if (e != hipSuccess) {
            dev_free(d_sel);
            return fail(RL_DEVICE_ERROR, std::string("iterator selection: ") + hipGetErrorString(e));
        }
        for (uint32_t p = 0; p < parts && st == RL_SUCCESS; p++) {
            const uint32_t a = (uint32_t)((uint64_t)n * p / parts), b = (uint32_t)((uint64_t)n * (p + 1) / parts);
            std::vector<BakeItem> part{BakeItem{&tmp, scene, ao, a, b - a}};
            st = bake(part);
            if (st == RL_SUCCESS && progress && progress->notify && p + 1 < parts) {
                // Steps below the next part's first step are finished.
                const int64_t done = b < n ? (int64_t)step_of[b] : steps;
                progress->notify(progress->user, mesh, (int32_t)std::min<int64_t>(done, INT32_MAX),
                                 (int32_t)std::min<int64_t>(items, INT32_MAX));
            }
        }
        if (st == RL_SUCCESS) {
            e = launch_scatter_colors(ctx->stream, d_sel, n, tmp.d_color, lm->d_color);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) st = fail(RL_DEVICE_ERROR, std::string("iterator write-back: ") + hipGetErrorString(e));
            lm->device_newer = true;
        }
        dev_free(d_sel);
    }
    if (st == RL_SUCCESS && progress && progress->notify)
        progress->notify(progress->user, mesh, (int32_t)std::min<int64_t>(items, INT32_MAX),
                         (int32_t)std::min<int64_t>(items, INT32_MAX));
    return st;
}

// Baker::bakeMesh preconditions (baker/Baker.cpp:54-65) and the mesh's lightmap.
int32_t mesh_lightmap(rl_relight* ctx, rl_scene* scene, int32_t mesh, rl_lightmap** out) {
    if (!ctx) return fail(RL_INVALID_ARGUMENT, "context is NULL");
    int32_t st = check_scene_ready(scene);
    if (st) return st;
    if (mesh < 0 || mesh >= (int32_t)scene->meshes.size()) return fail(RL_INVALID_ARGUMENT, "bad mesh index");
    if (scene->meshes[mesh].num_tris == 0) return fail(RL_INVALID_CALL, "mesh has no faces");
    *out = scene->meshes[mesh].lightmap;
    if (!*out) return fail(RL_INVALID_CALL, "mesh has no lightmap");
    return RL_SUCCESS;
}

// A bake call through an iterator: selection, then bake_selection.
template <class Bake>
int32_t bake_iterated(rl_relight* ctx, rl_scene* scene, int32_t mesh, const rl_bake_iterator* iterator,
                      const rl_progress* progress, const AoParams& ao, Bake bake) {
    rl_lightmap* lm = nullptr;
    int32_t st = mesh_lightmap(ctx, scene, mesh, &lm);
    if (st) return st;
    // Relight::bake* with a NULL iterator uses LumelBakeIterator(0, 1) (Relight.cpp:218-220).
    const rl_bake_iterator all{RL_ITERATE_LUMELS, 0, 1};
    const rl_bake_iterator& it = iterator ? *iterator : all;
    if ((st = push_lumels(lm))) return st;
    std::vector<uint32_t> sel, step_of;
    int64_t steps = 0, items = 0;
    if ((st = iterator_selection(scene, mesh, lm, it, sel, step_of, steps, items))) return st;
    CallTimer tm(ctx);
    if ((st = begin_call(ctx))) return st;
    if ((st = bake_selection(ctx, scene, mesh, lm, sel, step_of, steps, items, progress, ao, bake))) return st;
    return end_call(ctx, tm);
}

}  // namespace

extern "C" {

int32_t rl_bake_direct_iter(rl_relight* ctx, rl_scene* scene, int32_t mesh, const rl_bake_iterator* iterator,
                            const rl_progress* progress) {
    return bake_iterated(ctx, scene, mesh, iterator, progress, AoParams{},
                         [ctx](const std::vector<BakeItem>& items) { return run_direct_items(ctx, items); });
}

int32_t rl_bake_ambient_occlusion_iter(rl_relight* ctx, rl_scene* scene, int32_t mesh, const rl_ao_settings* settings,
                                       const rl_bake_iterator* iterator, const rl_progress* progress) {
    int32_t st = check_ao(settings);
    if (st) return st;
    return bake_iterated(ctx, scene, mesh, iterator, progress, ao_params(*settings),
                         [ctx](const std::vector<BakeItem>& items) { return run_ao_items(ctx, items); });
}

int32_t rl_get_progress(const rl_relight* ctx, uint64_t* done, uint64_t* total) {
    if (!ctx) return fail(RL_INVALID_ARGUMENT, "context is NULL");
    if (done) *done = ctx->progress_done.load();
    if (total) *total = ctx->progress_total.load();
    return RL_SUCCESS;
}

int32_t rl_bake_direct(rl_relight* ctx, rl_scene* scene, int32_t mesh) {
    if (!ctx) return fail(RL_INVALID_ARGUMENT, "context is NULL");
    int32_t st = check_scene_ready(scene);
    if (st) return st;
    if (mesh < 0 || mesh >= (int32_t)scene->meshes.size()) return fail(RL_INVALID_ARGUMENT, "bad mesh index");
    // Baker::bakeMesh preconditions (baker/Baker.cpp:54-65).
    if (scene->meshes[mesh].num_tris == 0) return fail(RL_INVALID_CALL, "mesh has no faces");
    rl_lightmap* lm = scene->meshes[mesh].lightmap;
    if (!lm) return fail(RL_INVALID_CALL, "mesh has no lightmap");
    CallTimer tm(ctx);
    if ((st = begin_call(ctx))) return st;
    if ((st = run_direct(ctx, scene, &lm, 1))) return st;
    return end_call(ctx, tm);
}

int32_t rl_bake_ambient_occlusion(rl_relight* ctx, rl_scene* scene, int32_t mesh,
                                  const rl_ao_settings* settings) {
    if (!ctx) return fail(RL_INVALID_ARGUMENT, "context is NULL");
    int32_t st = check_scene_ready(scene);
    if (st) return st;
    if (mesh < 0 || mesh >= (int32_t)scene->meshes.size()) return fail(RL_INVALID_ARGUMENT, "bad mesh index");
    if (scene->meshes[mesh].num_tris == 0) return fail(RL_INVALID_CALL, "mesh has no faces");
    rl_lightmap* lm = scene->meshes[mesh].lightmap;
    if (!lm) return fail(RL_INVALID_CALL, "mesh has no lightmap");
    CallTimer tm(ctx);
    if ((st = begin_call(ctx))) return st;
    if ((st = run_ao(ctx, scene, &lm, 1, settings))) return st;
    return end_call(ctx, tm);
}

int32_t rl_bake_batch(rl_relight* ctx, rl_scene* scene, rl_lightmap* const* maps,
                      int32_t num_maps, int32_t do_direct, const rl_ao_settings* ao) {
    if (!ctx || !maps || num_maps <= 0) return fail(RL_INVALID_ARGUMENT, "bad batch arguments");
    int32_t st = check_scene_ready(scene);
    if (st) return st;
    CallTimer tm(ctx);
    if ((st = begin_call(ctx))) return st;
    // Both phases queued back to back; end_call synchronises.
    if (do_direct && (st = run_direct(ctx, scene, maps, num_maps, false))) return st;
    if (ao && (st = run_ao(ctx, scene, maps, num_maps, ao, false))) return st;
    return end_call(ctx, tm);
}

}  // extern "C"

// ---------------------------------------------------------------- recorded passes (HIP graphs)

// A batched direct/AO pass recorded once as a HIP graph and replayed: its segment tables live in
// buffers of its own, and each replay checks that the scene and lightmaps still have the device
// arrays the graph was recorded with.
struct rl_pass {
    rl_relight* ctx = nullptr;
    rl_scene* scene = nullptr;
    const Bvh4Node* nodes = nullptr;
    struct MapRef {
        rl_lightmap* lm;
        const float4* pos;
        float4* color;
        uint32_t* counts;
        size_t texels;
    };
    std::vector<MapRef> maps;
    BakeSeg* d_segs[4] = {};  // direct (lumels), AO work units, AO lumels, direct (pairs)
    float* d_inf = nullptr;   // pair influences of a small pass's direct light
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t texels = 0;
    int launches = 0;

    ~rl_pass() {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        for (BakeSeg* d : d_segs) dev_free(d);
        dev_free(d_inf);
    }
};

namespace {

int32_t pass_upload(const std::vector<BakeSeg>& segs, BakeSeg** d) {
    if (segs.empty()) return RL_SUCCESS;
    RL_HIP(dev_alloc(d, segs.size()));
    RL_HIP(hipMemcpy(*d, segs.data(), segs.size() * sizeof(BakeSeg), hipMemcpyHostToDevice));
    return RL_SUCCESS;
}

// The launches of one pass on ctx->stream (recorded into the graph while it is captured). The
// fills are kernels, not memset nodes.
hipError_t pass_launches(rl_relight* ctx, const rl_pass& p, const DeviceSegs& dir, const DeviceSegs& pairs,
                         const DeviceSegs& work, const DeviceSegs& lum, bool refill, uint32_t chunk,
                         const std::vector<BakeItem>& ao_items) {
    hipError_t e = hipSuccess;
    if (dir.total && pairs.total)
        e = launch_bake_direct_pairs(ctx->stream, dir.cap, ctx->max_blocks, p.d_segs[3], p.d_segs[0], dir.n,
                                     pairs.total, dir.total, dir.alpha, shortcuts(ctx), p.d_inf, ctx->d_queue,
                                     ctx->d_rays);
    else if (dir.total)
        e = launch_bake_direct(ctx->stream, dir.cap, ctx->max_blocks, p.d_segs[0], dir.n, dir.total, dir.alpha,
                               shortcuts(ctx), ctx->d_queue, ctx->d_rays);
    if (e == hipSuccess && work.total) {
        if (refill) {
            e = launch_ao_refill(ctx->stream, work.cap, ctx->max_blocks, p.d_segs[1], work.n, work.total, chunk,
                                 ctx->tuning.refill_min, shortcuts(ctx), ctx->d_queue, ctx->d_rays);
        } else {
            for (const BakeItem& it : ao_items)
                if (e == hipSuccess && it.count) e = launch_fill_u32(ctx->stream, it.lm->d_counts + it.first, it.count, 0u);
            if (e == hipSuccess)
                e = launch_ao_count(ctx->stream, work.cap, ctx->max_blocks, p.d_segs[1], work.n, work.total,
                                    ctx->d_queue, ctx->d_rays);
        }
        if (e == hipSuccess) e = launch_ao_finalize(ctx->stream, p.d_segs[2], lum.n, lum.total);
    }
    return e;
}

}  // namespace

extern "C" {

int32_t rl_pass_create(rl_relight* ctx, rl_scene* scene, rl_lightmap* const* maps, int32_t num_maps,
                       int32_t do_direct, const rl_ao_settings* ao, rl_pass** out) {
    if (!ctx || !maps || num_maps <= 0 || !out) return fail(RL_INVALID_ARGUMENT, "bad pass arguments");
    *out = nullptr;
    int32_t st = check_scene_ready(scene);
    if (st) return st;
    if (ao && (st = check_ao(ao))) return st;
    if (!do_direct && !ao) return fail(RL_INVALID_ARGUMENT, "a pass needs direct light or AO");
    std::vector<BakeItem> items;
    if ((st = whole_items(scene, maps, num_maps, ao ? ao_params(*ao) : AoParams{}, items))) return st;
    std::unique_ptr<rl_pass> p(new rl_pass());
    p->ctx = ctx;
    p->scene = scene;
    p->nodes = scene->dev().nodes;
    for (const BakeItem& it : items) {
        p->maps.push_back({it.lm, it.lm->d_pos, it.lm->d_color, it.lm->d_counts, it.lm->texels.size()});
        p->texels += it.count;
    }
    // The same kernel choice as a direct bake_batch call, made once at recording.
    bool refill = ctx->tuning.ao_kernel == RL_AO_KERNEL_REFILL;
    for (const BakeItem& it : items) refill = refill && it.ao.samples >= 4;
    const uint32_t chunk = ao_chunk(ctx, refill, items);
    DeviceSegs dir, pairs, work, lum;
    if (do_direct && (st = pass_upload(build_segments(items, kUnitLumelAligned, 1, dir), &p->d_segs[0])))
        return st;
    if (do_direct && direct_uses_pairs(ctx, items, dir.texels)) {
        if ((st = pass_upload(build_segments(items, kUnitDirectPair, 1, pairs), &p->d_segs[3]))) return st;
        RL_HIP(dev_alloc(&p->d_inf, (size_t)pairs.total));
    }
    if (ao) {
        if ((st = pass_upload(build_segments(items, refill ? kUnitChunk : kUnitPair, chunk, work), &p->d_segs[1])))
            return st;
        if ((st = pass_upload(build_segments(items, kUnitLumel, 1, lum), &p->d_segs[2]))) return st;
    }
    p->launches = (dir.total ? (pairs.total ? 2 : 1) : 0) + (work.total ? 2 : 0);
    RL_HIP(hipStreamSynchronize(ctx->stream));
    RL_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = pass_launches(ctx, *p, dir, pairs, work, lum, refill, chunk, items);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
    p->graph = g;
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return fail(RL_DEVICE_ERROR, std::string("recording the pass: ") + hipGetErrorString(e));
    *out = p.release();
    return RL_SUCCESS;
}

int32_t rl_pass_launch(rl_relight* ctx, rl_pass* pass) {
    if (!ctx || !pass) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    if (pass->ctx != ctx) return fail(RL_INVALID_CALL, "the pass was recorded on another context");
    if (pass->scene->dev().nodes != pass->nodes) return fail(RL_INVALID_CALL, "the scene changed since recording");
    for (const rl_pass::MapRef& m : pass->maps) {
        int32_t st = push_lumels(m.lm);
        if (st) return st;
        if (m.lm->d_pos != m.pos || m.lm->d_color != m.color || m.lm->d_counts != m.counts ||
            m.lm->texels.size() != m.texels)
            return fail(RL_INVALID_CALL, "a lightmap changed since recording");
    }
    CallTimer tm(ctx);
    int32_t st = begin_call(ctx);
    if (st) return st;
    RL_HIP(hipGraphLaunch(pass->exec, ctx->stream));
    ctx->stats.texels += pass->texels;
    ctx->stats.launches += pass->launches;
    for (const rl_pass::MapRef& m : pass->maps) m.lm->device_newer = true;
    return end_call(ctx, tm);
}

int32_t rl_pass_destroy(rl_pass* pass) {
    if (!pass) return RL_SUCCESS;
    hipStreamSynchronize(pass->ctx->stream);
    delete pass;
    return RL_SUCCESS;
}

}  // extern "C"

extern "C" {

int32_t rl_trace_segments(rl_relight* ctx, rl_scene* scene, const float* starts, const float* ends,
                          uint32_t n, int32_t mode, uint8_t* out_flags, rl_hit* out_hits) {
    if (!ctx) return fail(RL_INVALID_ARGUMENT, "context is NULL");
    int32_t st = check_scene_ready(scene);
    if (st) return st;
    if (mode != RL_TRACE_ANY && mode != RL_TRACE_CLOSEST) return fail(RL_INVALID_ARGUMENT, "bad trace mode");
    if (n == 0) return RL_SUCCESS;
    if (!starts || !ends) return fail(RL_INVALID_ARGUMENT, "segment arrays are NULL");
    if (mode == RL_TRACE_ANY && !out_flags) return fail(RL_INVALID_ARGUMENT, "out_flags is NULL");
    if (mode == RL_TRACE_CLOSEST && !out_hits) return fail(RL_INVALID_ARGUMENT, "out_hits is NULL");
    std::vector<float4> a(n), b(n);
    for (uint32_t i = 0; i < n; i++) {
        a[i] = make_float4(starts[3 * i], starts[3 * i + 1], starts[3 * i + 2], 0.0f);
        b[i] = make_float4(ends[3 * i], ends[3 * i + 1], ends[3 * i + 2], 0.0f);
    }
    float4 *da = nullptr, *db = nullptr, *dh = nullptr;
    uint8_t* df = nullptr;
    auto cleanup = [&]() { dev_free(da); dev_free(db); dev_free(dh); dev_free(df); };
    hipError_t e = dev_alloc(&da, n);
    if (e == hipSuccess) e = dev_alloc(&db, n);
    if (e == hipSuccess) e = mode == RL_TRACE_ANY ? dev_alloc(&df, n) : dev_alloc(&dh, n);
    if (e != hipSuccess) {
        cleanup();
        return fail(RL_OUT_OF_MEMORY, "trace buffers");
    }
    CallTimer tm(ctx);
    if (e == hipSuccess) e = hipMemcpy(da, a.data(), n * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, b.data(), n * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess && (st = begin_call(ctx))) { cleanup(); return st; }
    if (e == hipSuccess) e = kernel_begin(ctx, RL_KERNEL_TRACE);
    if (e == hipSuccess)
        e = launch_trace(ctx->stream, scene->dev(), scene->cap, ctx->max_blocks, mode == RL_TRACE_ANY, da, db, n, df,
                         dh, ctx->d_queue, ctx->d_rays);
    if (e == hipSuccess) e = kernel_end(ctx);
    ctx->stats.launches = 1;
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    std::vector<float4> h;
    if (e == hipSuccess) {
        if (mode == RL_TRACE_ANY) {
            e = hipMemcpy(out_flags, df, n, hipMemcpyDeviceToHost);
        } else {
            h.resize(n);
            e = hipMemcpy(h.data(), dh, n * sizeof(float4), hipMemcpyDeviceToHost);
        }
    }
    if (e != hipSuccess) {
        cleanup();
        return fail(RL_DEVICE_ERROR, std::string("trace: ") + hipGetErrorString(e));
    }
    if (mode == RL_TRACE_CLOSEST) {
        for (uint32_t i = 0; i < n; i++) {
            uint32_t tri;
            std::memcpy(&tri, &h[i].y, 4);
            rl_hit& o = out_hits[i];
            o.t = h[i].x;
            o.u = h[i].z;
            o.v = h[i].w;
            if (tri == 0xFFFFFFFFu) {
                o.triangle = -1;
                o.mesh = -1;
                o.face = -1;
            } else {
                o.triangle = (int32_t)tri;
                o.mesh = (int32_t)scene->tri_mesh[tri];
                o.face = (int32_t)scene->tri_face[tri];
            }
        }
    }
    cleanup();
    return end_call(ctx, tm);
}

int32_t rl_trace_segments_noalloc(rl_relight* ctx, rl_scene* scene, const float* starts, const float* ends,
                                  uint32_t n, int32_t mode, uint8_t* out_flags, rl_hit* out_hits) {
    if (!ctx) return fail(RL_INVALID_ARGUMENT, "context is NULL");
    int32_t st = check_scene_ready(scene);
    if (st) return st;
    if (mode != RL_TRACE_ANY && mode != RL_TRACE_CLOSEST) return fail(RL_INVALID_ARGUMENT, "bad trace mode");
    if (n > kNoAllocMaxSegments)
        return fail(RL_INVALID_ARGUMENT, "rl_trace_segments_noalloc takes at most " +
                                             std::to_string(kNoAllocMaxSegments) + " segments per call, got " +
                                             std::to_string(n));
    if (n == 0) return RL_SUCCESS;
    if (!starts || !ends) return fail(RL_INVALID_ARGUMENT, "segment arrays are NULL");
    if (mode == RL_TRACE_ANY && !out_flags) return fail(RL_INVALID_ARGUMENT, "out_flags is NULL");
    if (mode == RL_TRACE_CLOSEST && !out_hits) return fail(RL_INVALID_ARGUMENT, "out_hits is NULL");
    const auto t0 = std::chrono::steady_clock::now();
    unsigned long long rays = 0;
    hipError_t e = noalloc_trace(ctx->device, ctx->stream, scene->dev(), scene->cap, ctx->max_blocks,
                                 mode == RL_TRACE_ANY, starts, ends, n, out_flags, out_hits, &rays);
    if (e != hipSuccess) return fail(RL_DEVICE_ERROR, std::string("trace (no-alloc): ") + hipGetErrorString(e));
    ctx->stats = rl_stats{};
    ctx->stats.rays = rays;
    ctx->stats.launches = 1;
    ctx->stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return RL_SUCCESS;
}

int32_t rl_photonmap_create(rl_relight* ctx, int32_t width, int32_t height, rl_lightmap** out) {
    int32_t st = rl_lightmap_create(ctx, width, height, out);
    if (st) return st;
    (*out)->is_photonmap = true;
    return RL_SUCCESS;
}

int32_t rl_photonmap_add_mesh(rl_lightmap* pm, rl_scene* scene, int32_t mesh) {
    if (!pm || !scene) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    if (!pm->is_photonmap) return fail(RL_INVALID_CALL, "not a photonmap");
    if (mesh < 0 || mesh >= (int32_t)scene->meshes.size()) return fail(RL_INVALID_ARGUMENT, "bad mesh index");
    if (pm->scene && pm->scene != scene) return fail(RL_INVALID_CALL, "photonmap already holds another scene's meshes");
    MeshHost& m = scene->meshes[mesh];
    // Photonmap::addMesh refuses a mesh that already has a photonmap (Lightmap.cpp:362-367).
    if (m.photonmap) return fail(RL_INVALID_CALL, "mesh already has a photonmap");
    m.photonmap = pm;
    pm->scene = scene;
    uint32_t first = 0;
    for (int32_t i = 0; i < mesh; i++) first += scene->meshes[i].num_tris;
    int32_t st = raster_mesh(pm, m, first);
    if (st) return st;
    pm->pm_valid_dirty = true;
    return RL_SUCCESS;
}

int32_t rl_photonmap_read(rl_lightmap* pm, float* color, uint32_t* photons, float* gathered) {
    if (!pm) return fail(RL_INVALID_ARGUMENT, "photonmap is NULL");
    if (!pm->is_photonmap) return fail(RL_INVALID_CALL, "not a photonmap");
    const size_t n = (size_t)pm->w * (size_t)pm->h;
    if (!pm->d_pm_color) {
        if (color) std::memset(color, 0, 3 * n * sizeof(float));
        if (photons) std::memset(photons, 0, n * sizeof(uint32_t));
        if (gathered) std::memset(gathered, 0, 3 * n * sizeof(float));
        return RL_SUCCESS;
    }
    std::vector<float4> tmp(n);
    for (int which = 0; which < 2; which++) {
        float* dst = which == 0 ? color : gathered;
        if (!dst) continue;
        RL_HIP(hipMemcpy(tmp.data(), which == 0 ? pm->d_pm_color : pm->d_pm_gathered, n * sizeof(float4),
                         hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) {
            dst[3 * i] = tmp[i].x;
            dst[3 * i + 1] = tmp[i].y;
            dst[3 * i + 2] = tmp[i].z;
        }
    }
    if (photons) RL_HIP(hipMemcpy(photons, pm->d_pm_photons, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return RL_SUCCESS;
}

int32_t rl_emit_photons(rl_relight* ctx, rl_scene* scene, const rl_indirect_settings* s) {
    if (!ctx) return fail(RL_INVALID_ARGUMENT, "context is NULL");
    int32_t st = check_scene_ready(scene);
    if (st) return st;
    if ((st = check_indirect_settings(s))) return st;
    PhotonTables T;
    if ((st = make_photon_tables(scene, T))) return st;
    CallTimer tm(ctx);
    if ((st = begin_call(ctx))) return st;
    if (T.maps.empty() || scene->lights.empty()) return end_call(ctx, tm);
    // Photons::emit: passes x lights x PhotonEmitter::photonCount() = intensity * 25000
    // (baker/Photons.cpp:50-92, scene/Light.cpp:334-337) unless photons_per_light overrides it.
    const int nl = (int)scene->lights.size();
    std::vector<uint64_t> first(nl);
    uint64_t per_pass = 0;
    for (int l = 0; l < nl; l++) {
        first[l] = per_pass;
        int64_t c = s->photons_per_light > 0 ? s->photons_per_light
                                             : (int64_t)(int32_t)(scene->lights[l].dev.intensity * 25000.0f);
        per_pass += (uint64_t)std::max<int64_t>(c, 0);
    }
    PhotonParams P;
    P.per_pass = per_pass;
    P.total = per_pass * (uint64_t)s->photon_pass_count;
    P.seed = s->seed;
    P.max_depth = s->photon_bounce_count;
    P.energy_threshold = s->photon_energy_threshold;
    P.max_distance = s->photon_max_distance;
    P.termination = s->termination;
    P.nlights = nl;
    for (int k = 0; k < 3; k++) {
        P.scene_lo[k] = scene->bounds_lo[k];
        P.scene_hi[k] = scene->bounds_hi[k];
    }
    const uint64_t nrec = P.total * (uint64_t)(P.max_depth + 1);
    if (nrec == 0) return end_call(ctx, tm);
    if (nrec > 0x7FFFFFFFull) return fail(RL_REJECTED, "too many photon records for one emit call");
    const bool knn = s->gather_mode == RL_GATHER_KNN;
    // k-NN grid: cell = radius / 1.98 (all photons within the radius lie in the 5x5x5 cells
    // around a query's cell), a power-of-two bucket table of at least the record count.
    KnnGrid G{};
    int key_bits = 0;
    std::vector<float4> queries;
    if (knn) {
        const float h = s->knn_radius / 1.98f;
        for (int k = 0; k < 3; k++) G.lo[k] = scene->bounds_lo[k];
        G.inv_h = 1.0f / h;
        uint64_t buckets = 1024;
        while (buckets < nrec && buckets < (1ull << 28)) buckets <<= 1;
        G.mask = (uint32_t)(buckets - 1);
        for (key_bits = 0; (1ull << key_bits) <= buckets; key_bits++) {}
        G.max_shell = 2;
        G.k = s->knn_k;
        G.r2 = s->knn_radius * s->knn_radius;
        G.shell_d2[0] = 0.0f;
        for (int k = 1; k < 4; k++) {
            const float e = ((float)k - 0.01f) * h;
            G.shell_d2[k] = e * e;
        }
        for (size_t m = 0; m < T.maps.size(); m++) {
            const rl_lightmap* pm = T.maps[m];
            const uint32_t n = (uint32_t)pm->w * (uint32_t)pm->h;
            for (uint32_t t = 0; t < n; t++) {
                if (!pm->valid[t]) continue;
                const uint32_t key = T.dev[m].key_base + t;
                float kw;
                std::memcpy(&kw, &key, 4);
                queries.push_back(make_float4(pm->pos[t].x, pm->pos[t].y, pm->pos[t].z, kw));
            }
        }
    }
    uint64_t* d_first = nullptr;
    PhotonRec* d_rec = nullptr;
    uint32_t *ka = nullptr, *kb = nullptr, *ia = nullptr, *ib = nullptr;
    uint8_t* temp = nullptr;
    size_t temp_bytes = 0, knn_temp = 0;
    float4 *d_ppos = nullptr, *d_spos = nullptr, *d_queries = nullptr;
    uint32_t *d_cstart = nullptr, *d_cend = nullptr;
    auto cleanup = [&]() {
        dev_free(d_first); dev_free(d_rec); dev_free(ka); dev_free(kb); dev_free(ia); dev_free(ib); dev_free(temp);
        dev_free(d_ppos); dev_free(d_spos); dev_free(d_queries); dev_free(d_cstart); dev_free(d_cend);
    };
    hipError_t e = dev_alloc(&d_first, (size_t)nl);
    if (e == hipSuccess) e = dev_alloc(&d_rec, (size_t)nrec);
    if (e == hipSuccess) e = dev_alloc(&ka, (size_t)nrec);
    if (e == hipSuccess) e = dev_alloc(&kb, (size_t)nrec);
    if (e == hipSuccess) e = dev_alloc(&ia, (size_t)nrec);
    if (e == hipSuccess) e = dev_alloc(&ib, (size_t)nrec);
    temp_bytes = knn_temp = sort_pairs_temp_bytes((uint32_t)nrec);
    if (e == hipSuccess) e = dev_alloc(&temp, std::max<size_t>(std::max(temp_bytes, knn_temp), 1));
    if (e == hipSuccess && knn) e = dev_alloc(&d_ppos, (size_t)nrec);
    if (e == hipSuccess && knn) e = dev_alloc(&d_spos, (size_t)nrec);
    if (e == hipSuccess && knn) e = dev_alloc(&d_cstart, (size_t)G.mask + 1);
    if (e == hipSuccess && knn) e = dev_alloc(&d_cend, (size_t)G.mask + 1);
    if (e == hipSuccess && knn && !queries.empty()) e = dev_alloc(&d_queries, queries.size());
    if (e != hipSuccess) {
        cleanup();
        return fail(RL_OUT_OF_MEMORY, std::string("photon buffers: ") + hipGetErrorString(e));
    }
    e = hipMemcpy(d_first, first.data(), nl * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && d_queries)
        e = hipMemcpy(d_queries, queries.data(), queries.size() * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = kernel_begin(ctx, RL_KERNEL_PHOTON_TRACE);
    if (e == hipSuccess)
        e = launch_photon_trace(ctx->stream, scene->dev(), scene->cap, ctx->max_blocks, scene->d_lights, d_first, P,
                                T.d_mesh_pm, T.d_pms, d_rec, d_ppos, ctx->d_queue, ctx->d_rays);
    if (e == hipSuccess) e = kernel_end(ctx);
    if (e == hipSuccess) e = kernel_begin(ctx, RL_KERNEL_PHOTON_STORE);
    if (e == hipSuccess)
        e = launch_photon_store(ctx->stream, d_rec, (uint32_t)nrec, ka, kb, ia, ib, temp, temp_bytes, T.d_pms,
                                (int)T.dev.size(), (uint32_t)T.total_texels);
    if (e == hipSuccess) e = kernel_end(ctx);
    // Relight::emitPhotons gathers every photonmap after the passes (Relight.cpp:250-254).
    if (knn) {
        // Texels without a surface point get 0; the rest the k-nearest estimate of this call's photons.
        for (size_t m = 0; m < T.maps.size() && e == hipSuccess; m++)
            e = launch_fill_u32(ctx->stream, reinterpret_cast<uint32_t*>(T.dev[m].gathered),
                                4ull * (uint64_t)T.dev[m].w * (uint64_t)T.dev[m].h, 0u);
        if (e == hipSuccess) e = kernel_begin(ctx, RL_KERNEL_PHOTON_GRID);
        if (e == hipSuccess)
            e = launch_knn_build(ctx->stream, d_rec, d_ppos, (uint32_t)nrec, G, key_bits, ka, kb, ia, ib, temp,
                                 std::max(temp_bytes, knn_temp), d_spos, d_cstart, d_cend);
        if (e == hipSuccess) e = kernel_end(ctx);
        if (e == hipSuccess) e = kernel_begin(ctx, RL_KERNEL_PHOTON_GATHER);
        if (e == hipSuccess)
            e = launch_knn_gather(ctx->stream, ctx->max_blocks, d_queries, (uint32_t)queries.size(), d_spos, d_rec,
                                  d_cstart, d_cend, G, T.d_pms, (int)T.dev.size(), ctx->d_queue);
        if (e == hipSuccess) e = kernel_end(ctx);
        ctx->stats.launches += 3;
    } else {
        if (e == hipSuccess) e = kernel_begin(ctx, RL_KERNEL_PHOTON_GATHER);
        if (e == hipSuccess)
            e = launch_photon_gather(ctx->stream, T.d_pms, (int)T.dev.size(), (uint32_t)T.total_texels,
                                     s->final_gather_radius);
        if (e == hipSuccess) e = kernel_end(ctx);
    }
    ctx->stats.launches += 4;
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    cleanup();
    if (e != hipSuccess) return fail(RL_DEVICE_ERROR, std::string("photon pass: ") + hipGetErrorString(e));
    return end_call(ctx, tm);
}

int32_t rl_bake_indirect_iter(rl_relight* ctx, rl_scene* scene, int32_t mesh, const rl_indirect_settings* s,
                              const rl_bake_iterator* iterator, const rl_progress* progress) {
    if (!ctx) return fail(RL_INVALID_ARGUMENT, "context is NULL");
    int32_t st = check_scene_ready(scene);
    if (st) return st;
    if ((st = check_indirect_settings(s))) return st;
    PhotonTables T;
    if ((st = make_photon_tables(scene, T))) return st;
    const FinalGatherParams fg = fg_params(s);
    return bake_iterated(ctx, scene, mesh, iterator, progress, AoParams{},
                         [ctx, &T, &fg](std::vector<BakeItem>& items) {
                             for (BakeItem& it : items) {
                                 it.fg = fg;
                                 it.mesh_pm = T.d_mesh_pm;
                                 it.pms = T.d_pms;
                             }
                             return run_fg_items(ctx, items);
                         });
}

int32_t rl_bake_indirect(rl_relight* ctx, rl_scene* scene, int32_t mesh, const rl_indirect_settings* s) {
    if (!ctx) return fail(RL_INVALID_ARGUMENT, "context is NULL");
    int32_t st = check_scene_ready(scene);
    if (st) return st;
    if ((st = check_indirect_settings(s))) return st;
    if (mesh < 0 || mesh >= (int32_t)scene->meshes.size()) return fail(RL_INVALID_ARGUMENT, "bad mesh index");
    if (scene->meshes[mesh].num_tris == 0) return fail(RL_INVALID_CALL, "mesh has no faces");
    rl_lightmap* lm = scene->meshes[mesh].lightmap;
    if (!lm) return fail(RL_INVALID_CALL, "mesh has no lightmap");
    PhotonTables T;
    if ((st = make_photon_tables(scene, T))) return st;
    CallTimer tm(ctx);
    if ((st = begin_call(ctx))) return st;
    std::vector<BakeItem> items;
    AoParams none{};
    if ((st = whole_items(scene, &lm, 1, none, items))) return st;
    for (BakeItem& it : items) {
        it.fg = fg_params(s);
        it.mesh_pm = T.d_mesh_pm;
        it.pms = T.d_pms;
    }
    if ((st = run_fg_items(ctx, items))) return st;
    return end_call(ctx, tm);
}

}  // extern "C"

struct rl_radiosity {
    rl_relight* ctx = nullptr;
    rl_scene* scene = nullptr;
    std::atomic<int32_t> server_jobs{0};  // unfinished server jobs that solve it
    uint32_t n = 0, npad = 0;
    PatchDev* d_patches = nullptr;
    uint32_t* d_tri_patch = nullptr;
    uint16_t* d_F = nullptr;
    bool built = false;
    uint32_t rays = 0;
    // Solve buffers, kept between solves (sized for sol_mpad columns and sol_nsplit partials).
    uint32_t sol_mpad = 0, sol_nsplit = 0;
    float *E = nullptr, *RHO = nullptr, *B = nullptr, *R = nullptr, *D = nullptr, *Yp = nullptr, *part = nullptr;
    float* stage = nullptr;  // row-major emission / reflectance / result
    __bf16 *X1 = nullptr, *X2 = nullptr, *X3 = nullptr;
    void free_solve() {
        dev_free(E); dev_free(RHO); dev_free(B); dev_free(R); dev_free(D); dev_free(Yp); dev_free(part);
        dev_free(stage); dev_free(X1); dev_free(X2); dev_free(X3);
        sol_mpad = sol_nsplit = 0;
    }
    ~rl_radiosity() {
        free_solve();
        dev_free(d_patches);
        dev_free(d_tri_patch);
        dev_free(d_F);
    }
};

extern "C" {

int32_t rl_radiosity_create(rl_relight* ctx, rl_scene* scene, const float* patches, uint32_t num_patches,
                            const uint32_t* tri_patch, rl_radiosity** out) {
    if (!ctx || !out || !patches || !tri_patch) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    int32_t st = check_scene_ready(scene);
    if (st) return st;
    if (num_patches == 0 || num_patches > 16384)
        return fail(RL_REJECTED, "patch count must be in [1, 16384] (one LDS histogram per row)");
    for (uint32_t t = 0; t < scene->num_tris; t++)
        if (tri_patch[t] != 0xFFFFFFFFu && tri_patch[t] >= num_patches)
            return fail(RL_INVALID_ARGUMENT, "triangle patch index out of range");
    rl_radiosity* r = new (std::nothrow) rl_radiosity();
    if (!r) return fail(RL_OUT_OF_MEMORY, "radiosity allocation failed");
    r->ctx = ctx;
    r->scene = scene;
    r->n = num_patches;
    r->npad = (num_patches + 255u) / 256u * 256u;
    std::vector<PatchDev> pd(num_patches);
    for (uint32_t i = 0; i < num_patches; i++) std::memcpy(&pd[i], patches + 12 * (size_t)i, sizeof(PatchDev));
    hipError_t e = dev_alloc(&r->d_patches, num_patches);
    if (e == hipSuccess) e = dev_alloc(&r->d_tri_patch, scene->num_tris);
    if (e == hipSuccess) e = dev_alloc(&r->d_F, (size_t)r->npad * r->npad);
    if (e == hipSuccess) e = hipMemcpy(r->d_patches, pd.data(), pd.size() * sizeof(PatchDev), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(r->d_tri_patch, tri_patch, scene->num_tris * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(r->d_F, 0, (size_t)r->npad * r->npad * sizeof(uint16_t));
    if (e != hipSuccess) {
        delete r;
        return fail(e == hipErrorOutOfMemory ? RL_OUT_OF_MEMORY : RL_DEVICE_ERROR,
                    std::string("radiosity buffers: ") + hipGetErrorString(e));
    }
    *out = r;
    return RL_SUCCESS;
}

int32_t rl_radiosity_destroy(rl_radiosity* r) {
    if (r && r->server_jobs.load() > 0)
        return fail(RL_INVALID_CALL, "radiosity object is in use by an unfinished server job");
    delete r;
    return RL_SUCCESS;
}

int32_t rl_radiosity_build_form_factors(rl_relight* ctx, rl_radiosity* r, uint32_t rays_per_patch, uint32_t seed,
                                        float max_distance) {
    if (!ctx || !r) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    if (rays_per_patch == 0 || rays_per_patch > 65535) return fail(RL_INVALID_ARGUMENT, "rays per patch must be in [1, 65535]");
    if (!(max_distance > 0.0f)) return fail(RL_INVALID_ARGUMENT, "max distance must be positive");
    CallTimer tm(ctx);
    int32_t st = begin_call(ctx);
    if (st) return st;
    FormFactorParams P;
    P.num_patches = r->n;
    P.row_stride = r->npad;
    P.rays = rays_per_patch;
    P.seed = seed;
    P.max_distance = max_distance;
    P.first_row = 0;
    P.rows = r->n;
    P.pad = 0;
    hipError_t e = kernel_begin(ctx, RL_KERNEL_FORM_FACTORS);
    if (e == hipSuccess)
        e = launch_form_factor_rows(ctx->stream, r->scene->dev(), r->scene->cap, r->d_patches, r->d_tri_patch, P,
                                    r->d_F, ctx->d_rays);
    if (e == hipSuccess) e = kernel_end(ctx);
    ctx->stats.launches = 1;
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(RL_DEVICE_ERROR, std::string("form factors: ") + hipGetErrorString(e));
    r->built = true;
    r->rays = rays_per_patch;
    return end_call(ctx, tm);
}

int32_t rl_radiosity_read_form_factors(rl_radiosity* r, uint32_t first_row, uint32_t rows, uint16_t* counts) {
    if (!r || !counts) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    if ((uint64_t)first_row + rows > r->n) return fail(RL_INVALID_ARGUMENT, "row range out of bounds");
    // F is stored in 16 x 32 tiles on the device: untile 1024 rows at a time into row-major rows.
    const uint32_t chunk = 1024;
    uint16_t* tmp = nullptr;
    hipError_t e = dev_alloc(&tmp, (size_t)std::min(rows, chunk) * r->n);
    hipStream_t st = r->ctx->stream;
    for (uint32_t r0 = 0; e == hipSuccess && r0 < rows; r0 += chunk) {
        const uint32_t nr = std::min(chunk, rows - r0);
        e = launch_ff_untile(st, r->d_F, r->npad, r->n, first_row + r0, nr, tmp);
        if (e == hipSuccess)
            e = hipMemcpyAsync(counts + (size_t)r0 * r->n, tmp, (size_t)nr * r->n * sizeof(uint16_t),
                               hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    dev_free(tmp);
    if (e != hipSuccess) return fail(RL_DEVICE_ERROR, std::string("read form factors: ") + hipGetErrorString(e));
    return RL_SUCCESS;
}

int32_t rl_radiosity_solve(rl_relight* ctx, rl_radiosity* r, const float* emission, const float* reflectance,
                           uint32_t num_rhs, float tolerance, int32_t max_outer, int32_t inner_sweeps,
                           float* out_radiosity, rl_solve_info* info) {
    if (!ctx || !r || !emission || !reflectance || !out_radiosity) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    if (!r->built) return fail(RL_INVALID_CALL, "form factors have not been built");
    if (num_rhs == 0 || num_rhs > 128) return fail(RL_INVALID_ARGUMENT, "right-hand sides must be in [1, 128]");
    if (max_outer <= 0 || inner_sweeps < 0) return fail(RL_INVALID_ARGUMENT, "bad iteration limits");
    const uint32_t tiles[] = {1, 2, 3, 4, 6, 8};
    uint32_t nt = 8;
    for (uint32_t t : tiles) {
        if (t * 16 >= num_rhs) { nt = t; break; }
    }
    const uint32_t n = r->n, npad = r->npad, mpad = nt * 16;
    // Split-K so that a launch has about 4 wavefronts per SIMD (npad/16 row tiles x nsplit), with
    // at least 64 columns of K per split (16,384 patches: split 4 measured 129 us per sweep product
    // against 137 us at 8 and 144 us at 16; fewer partial sums to write and add up).
    uint32_t nsplit = 4;
    while (nsplit < 32 && (npad / 16) * nsplit < 4096 && npad / (nsplit * 2) >= 64 && npad % (nsplit * 2 * 32) == 0)
        nsplit *= 2;
    const size_t vec = (size_t)mpad * npad;
    double e2 = 0.0;
    for (size_t i = 0; i < (size_t)n * num_rhs; i++) e2 += (double)emission[i] * emission[i];
    const uint32_t nblk = npad / 256;
    hipError_t e = hipSuccess;
    if (r->sol_mpad != mpad || r->sol_nsplit != nsplit) {
        r->free_solve();
        e = dev_alloc(&r->E, vec);
        if (e == hipSuccess) e = dev_alloc(&r->RHO, vec);
        if (e == hipSuccess) e = dev_alloc(&r->B, vec);
        if (e == hipSuccess) e = dev_alloc(&r->R, vec);
        if (e == hipSuccess) e = dev_alloc(&r->D, vec);
        if (e == hipSuccess) e = dev_alloc(&r->Yp, vec * nsplit);
        if (e == hipSuccess) e = dev_alloc(&r->part, (size_t)mpad * nblk);
        if (e == hipSuccess) e = dev_alloc(&r->stage, 2 * (size_t)npad * mpad);
        if (e == hipSuccess) e = dev_alloc(&r->X1, vec);
        if (e == hipSuccess) e = dev_alloc(&r->X2, vec);
        if (e == hipSuccess) e = dev_alloc(&r->X3, vec);
        if (e != hipSuccess) {
            r->free_solve();
            return fail(RL_OUT_OF_MEMORY, "radiosity solve buffers");
        }
        r->sol_mpad = mpad;
        r->sol_nsplit = nsplit;
    }
    float *E = r->E, *RHO = r->RHO, *B = r->B, *R = r->R, *D = r->D, *Yp = r->Yp, *part = r->part;
    __bf16 *X1 = r->X1, *X2 = r->X2, *X3 = r->X3;
    float* stE = r->stage;
    float* stR = r->stage + (size_t)npad * mpad;
    CallTimer tm(ctx);
    int32_t st = begin_call(ctx);
    if (st) return st;
    const size_t rowbytes = (size_t)n * num_rhs * sizeof(float);
    e = hipMemcpyAsync(stE, emission, rowbytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(stR, reflectance, rowbytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = launch_ff_to_cols(ctx->stream, stE, n, num_rhs, npad, mpad, E);
    if (e == hipSuccess) e = launch_ff_to_cols(ctx->stream, stR, n, num_rhs, npad, mpad, RHO);
    if (e == hipSuccess) e = hipMemcpyAsync(D, E, vec * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream);
    // B0 = E.
    if (e == hipSuccess) e = launch_ff_apply(ctx->stream, npad, mpad, B, D, 1, X1, X2, X3);
    const float inv_rays = 1.0f / (float)r->rays;
    const double enorm = std::sqrt(e2);
    std::vector<float> hp((size_t)mpad * nblk);
    double rel = 0.0;
    int outer = 0, sweeps = 0;
    for (; e == hipSuccess && outer < max_outer; outer++) {
        e = kernel_begin(ctx, RL_KERNEL_FF_GEMM_ACCURATE);
        if (e == hipSuccess) e = launch_ff_gemm(ctx->stream, true, r->d_F, npad, X1, X2, X3, npad, mpad, nsplit, Yp);
        if (e == hipSuccess) e = kernel_end(ctx);
        if (e == hipSuccess) e = launch_ff_residual(ctx->stream, Yp, nsplit, npad, mpad, inv_rays, E, RHO, B, R, D, X1, part);
        if (e == hipSuccess) e = hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) break;
        double r2 = 0.0;
        for (float v : hp) r2 += (double)v;
        rel = enorm > 0.0 ? std::sqrt(r2) / enorm : std::sqrt(r2);
        if (rel <= (double)tolerance) break;
        for (int s = 0; e == hipSuccess && s < inner_sweeps; s++) {
            e = kernel_begin(ctx, RL_KERNEL_FF_GEMM_FAST);
            if (e == hipSuccess) e = launch_ff_gemm(ctx->stream, false, r->d_F, npad, X1, X2, X3, npad, mpad, nsplit, Yp);
            if (e == hipSuccess) e = kernel_end(ctx);
            if (e == hipSuccess) e = launch_ff_sweep(ctx->stream, Yp, nsplit, npad, mpad, inv_rays, RHO, R, D, X1);
            sweeps++;
        }
        if (e == hipSuccess) e = launch_ff_apply(ctx->stream, npad, mpad, B, D, 0, X1, X2, X3);
    }
    if (e == hipSuccess) e = launch_ff_to_rows(ctx->stream, B, n, num_rhs, npad, stE);
    if (e == hipSuccess) e = hipMemcpyAsync(out_radiosity, stE, rowbytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(RL_DEVICE_ERROR, std::string("radiosity solve: ") + hipGetErrorString(e));
    st = end_call(ctx, tm);
    if (info) {
        info->outer_iterations = outer + (outer < max_outer ? 1 : 0);
        info->sweeps = sweeps;
        info->residual = rel;
        info->gpu_ms = ctx->stats.gpu_ms;
    }
    if (st) return st;
    if (rel > (double)tolerance) return fail(RL_NOT_CONVERGED, "radiosity solve did not reach the tolerance");
    return RL_SUCCESS;
}

static_assert(RL_FG_KERNEL_AUTO == rl::kFgKernelAuto && RL_FG_KERNEL_WAVE == rl::kFgKernelWave &&
                  RL_FG_KERNEL_LANE == rl::kFgKernelLane,
              "final-gather kernel ids");

int32_t rl_set_tuning(rl_relight* ctx, const rl_tuning* t) {
    if (!ctx || !t) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    if (t->ao_kernel != RL_AO_KERNEL_CHUNKED && t->ao_kernel != RL_AO_KERNEL_REFILL)
        return fail(RL_INVALID_ARGUMENT, "unknown AO kernel");
    if (t->refill_min < 1 || t->refill_min > 64) return fail(RL_INVALID_ARGUMENT, "refill_min must be in [1, 64]");
    if (t->ao_chunk_lumels < 1 || t->ao_chunk_lumels > 4096) return fail(RL_INVALID_ARGUMENT, "bad chunk size");
    if (t->blocks_per_cu < 1 || t->blocks_per_cu > 64) return fail(RL_INVALID_ARGUMENT, "blocks_per_cu must be in [1, 64]");
    if (t->fg_kernel != RL_FG_KERNEL_AUTO && t->fg_kernel != RL_FG_KERNEL_WAVE && t->fg_kernel != RL_FG_KERNEL_LANE)
        return fail(RL_INVALID_ARGUMENT, "unknown final-gather kernel");
    if (t->direct_kernel != RL_DIRECT_KERNEL_AUTO && t->direct_kernel != RL_DIRECT_KERNEL_LUMEL &&
        t->direct_kernel != RL_DIRECT_KERNEL_PAIRS)
        return fail(RL_INVALID_ARGUMENT, "unknown direct-light kernel");
    if ((t->local_trees != 0 && t->local_trees != 1) || (t->occluder_hints != 0 && t->occluder_hints != 1))
        return fail(RL_INVALID_ARGUMENT, "local_trees and occluder_hints must be 0 or 1");
    ctx->tuning = *t;
    ctx->max_blocks = (uint32_t)ctx->cus * t->blocks_per_cu;
    return RL_SUCCESS;
}

int32_t rl_get_tuning(const rl_relight* ctx, rl_tuning* out) {
    if (!ctx || !out) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    *out = ctx->tuning;
    return RL_SUCCESS;
}

int32_t rl_get_stats(const rl_relight* ctx, rl_stats* out) {
    if (!ctx || !out) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    *out = ctx->stats;
    return RL_SUCCESS;
}

}  // extern "C"

// ---------------------------------------------------------------- multi-scene batch and server

namespace {

int32_t add_item(const rl_bake_item& d, std::vector<BakeItem>& direct, std::vector<BakeItem>& ao) {
    if (!d.scene || !d.lightmap) return fail(RL_INVALID_ARGUMENT, "item scene or lightmap is NULL");
    int32_t st = check_scene_ready(d.scene);
    if (st) return st;
    if (d.lightmap->scene != d.scene) return fail(RL_INVALID_CALL, "lightmap holds no mesh of this scene");
    if ((st = push_lumels(d.lightmap))) return st;
    const uint32_t n = (uint32_t)d.lightmap->texels.size();
    if (d.first_lumel > n) return fail(RL_INVALID_ARGUMENT, "lumel range out of bounds");
    const uint32_t cnt = d.lumel_count == 0xFFFFFFFFu ? n - d.first_lumel : d.lumel_count;
    if ((uint64_t)d.first_lumel + cnt > n) return fail(RL_INVALID_ARGUMENT, "lumel range out of bounds");
    if (d.direct) direct.push_back(BakeItem{d.lightmap, d.scene, AoParams{}, d.first_lumel, cnt});
    if (d.ao) {
        if ((st = check_ao(d.ao))) return st;
        ao.push_back(BakeItem{d.lightmap, d.scene, ao_params(*d.ao), d.first_lumel, cnt});
    }
    return RL_SUCCESS;
}

uint32_t valid_lumels(const rl_lightmap* lm) {
    uint32_t c = 0;
    for (uint8_t v : lm->valid) c += v;
    return c;
}

// Ray phases of a job, baked in shared multi-job launches.
enum RayPhase { kPhaseNone = 0, kPhaseDirect = 1, kPhaseAo = 2, kPhaseGather = 3 };
// Steps of a job's plan: photon emission (a whole step) or one ray phase over all its lightmaps.
enum JobStep { kStepEmit = 0, kStepDirect = 1, kStepAo = 2, kStepGather = 3 };

// Wait of the oldest job (ms since submission) beyond which younger small jobs stop joining its
// steps. Serving mix of 64 requests at 80 % of the served capacity (Poisson): without the rule
// p50/p99 29.4/128 ms (plain 43/126 ms); at 20, 40 and 60 ms 30-32/122-124 ms; at 100 ms
// 29.4/128 ms. Holding radiosity jobs back the same way raised p99 to 134-143 ms.
constexpr double kBacklogMs = 60.0;

struct Job {
    uint64_t id = 0;
    rl_job_desc desc{};
    std::vector<rl_lightmap*> maps;
    std::vector<uint32_t> valid;  // valid lumels per map
    rl_ao_settings ao{};
    bool has_ao = false;
    rl_indirect_settings ind{};
    std::vector<float> E, R;
    // Written by the worker threads while callers poll: atomics (the error string is written
    // before the state turns FAILED).
    std::atomic<int32_t> state{RL_JOB_QUEUED};
    std::atomic<int32_t> status{RL_SUCCESS};
    std::atomic<uint64_t> rays{0};
    double submit_ms = 0, admit_ms = 0, done_ms = 0;
    uint64_t footprint = 0;
    // Plan (JobStep values) and the current step; radiosity jobs run on the auxiliary worker.
    std::vector<int> plan;
    size_t pi = 0;
    size_t map_i = 0;     // next lightmap of the phase
    uint32_t cursor = 0;  // next lumel of that lightmap
    std::atomic<bool> cancel{false};
    bool taken = false;                    // radiosity: picked up by the auxiliary worker
    bool prepared = false;                 // lumels uploaded
    std::unique_ptr<PhotonTables> tables;  // final gather: the scene's photon tables
    std::string error;    // message of the failing call, handed to the caller's rl_last_error
    bool holds = false;   // counted in the server_jobs of its scene, lightmaps and radiosity
    // Steps queued on the stream that bake part of this job and have not finished yet, and
    // whether every step of its plan has been queued (it is done once those have finished).
    int inflight = 0;
    bool planned = false;

    // Marks the job's scene, lightmaps and radiosity object as used (at submit) or no longer used
    // (once the job has finished, failed or been cancelled): their destroy calls refuse meanwhile.
    void hold(bool on) {
        if (holds == on) return;
        holds = on;
        const int32_t d = on ? 1 : -1;
        if (desc.scene) {
            desc.scene->server_jobs += d;
            // Photon steps store into, and final gathers read, the scene's photonmaps.
            for (const MeshHost& m : desc.scene->meshes)
                if (m.photonmap) m.photonmap->server_jobs += d;
        }
        for (rl_lightmap* lm : maps) lm->server_jobs += d;
        if (desc.kind == RL_JOB_RADIOSITY && desc.radiosity) desc.radiosity->server_jobs += d;
    }
};

}  // namespace

struct rl_server {
    rl_relight* ctx = nullptr;
    rl_relight* aux = nullptr;  // second context (own stream) for radiosity jobs
    rl_server_config cfg{};
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::thread worker, aux_worker;
    bool stop = false;
    std::map<uint64_t, std::unique_ptr<Job>> jobs;
    std::deque<Job*> queue;
    std::vector<Job*> active;
    uint64_t next_id = 1, mem_used = 0;
    rl_server_stats stats{};
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    // A step queued on the context's stream: the worker plans and queues step k + 1 while step k
    // runs, then waits for step k, so the GPU does not idle while the host prepares a step.
    struct StepSlot {
        hipEvent_t done = nullptr;             // recorded after the step's launches
        std::vector<Job*> jobs;                // jobs with lumels in the step
        unsigned long long* d_rays = nullptr;  // the step's ray counter
        unsigned long long* h_rays = nullptr;  // pinned copy of it, valid once `done` has completed
        uint64_t extra_rays = 0;               // photon rays of the emissions run while planning it
        double t_queued = 0;
        bool live = false;
    };
    StepSlot slots[2];
    double last_finish_ms = 0;
    // Photon tables of finished final gathers, freed once no step is in flight (hipFree waits for
    // the device, which would drain the queued step).
    std::vector<std::unique_ptr<PhotonTables>> retired_tables;

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    void loop();
    void aux_loop();
    void admit_locked();
    void retire_locked();
    bool queue_step(std::vector<Job*>& js, StepSlot& sl);
    void finish_step(StepSlot& sl);
};

namespace {

bool job_finished(const Job* j) {
    return j->state == RL_JOB_DONE || j->state == RL_JOB_FAILED || j->state == RL_JOB_CANCELLED;
}

void fail_job(Job* j, int32_t st) {
    j->error = g_error;
    j->status = st;
    j->state = RL_JOB_FAILED;
}

int current_step(const Job* j) { return j->pi < j->plan.size() ? j->plan[j->pi] : -1; }

RayPhase ray_phase(const Job* j) {
    if (job_finished(j) || j->desc.kind == RL_JOB_RADIOSITY) return kPhaseNone;
    switch (current_step(j)) {
        case kStepDirect: return kPhaseDirect;
        case kStepAo: return kPhaseAo;
        case kStepGather: return kPhaseGather;
        default: return kPhaseNone;
    }
}

// Moves a job to the next step of its plan; after the last, the job is planned (it becomes done
// once the steps that bake it have finished).
void advance(Job* j) {
    j->map_i = 0;
    j->cursor = 0;
    j->pi++;
    if (j->pi >= j->plan.size()) j->planned = true;
}

uint64_t phase_rpl(const Job* j, int step) {
    if (step == kStepDirect) return direct_rays_per_lumel(j->desc.scene);
    if (step == kStepAo) return (uint64_t)j->ao.samples;
    if (step == kStepGather) return (uint64_t)j->ind.final_gather_samples;
    return 0;
}

// Rays of the current ray phase of a job that are not planned yet.
uint64_t phase_rays_left(const Job* j) {
    uint64_t left = 0;
    for (size_t k = j->map_i; k < j->valid.size(); k++) left += j->valid[k];
    left -= std::min<uint64_t>(left, j->cursor);
    return left * phase_rpl(j, current_step(j));
}

// Rays of the ray phases a job has still to plan (photon emission not counted).
uint64_t remaining_rays(const Job* j) {
    uint64_t lumels = 0, total = 0;
    for (uint32_t v : j->valid) lumels += v;
    uint64_t done_now = 0;  // lumels of the current phase already planned
    for (size_t k = 0; k < j->map_i && k < j->valid.size(); k++) done_now += j->valid[k];
    done_now += j->cursor;
    for (size_t p = j->pi; p < j->plan.size(); p++)
        total += (p == j->pi ? lumels - std::min(lumels, done_now) : lumels) * phase_rpl(j, j->plan[p]);
    return total;
}

// Takes up to `budget` lumels of the job's current phase into `items`; returns lumels taken.
uint64_t take_lumels(Job* j, uint64_t budget, const BakeItem& proto, std::vector<BakeItem>& items) {
    uint64_t taken = 0;
    while (budget > 0 && j->map_i < j->maps.size()) {
        const uint32_t n = j->valid[j->map_i];
        const uint32_t take = (uint32_t)std::min<uint64_t>(budget, n - j->cursor);
        if (take) {
            BakeItem it = proto;
            it.lm = j->maps[j->map_i];
            it.scene = j->desc.scene;
            it.first = j->cursor;
            it.count = take;
            items.push_back(it);
        }
        j->cursor += take;
        budget -= take;
        taken += take;
        if (j->cursor >= n) {
            j->map_i++;
            j->cursor = 0;
        }
    }
    return taken;
}

}  // namespace

// Plans the next step of the jobs `js` and queues its launches on the stream without waiting for
// them: photon emissions first (synchronous calls), then the ray phases. Job cursors and plans
// advance here, at planning time; the stream runs a job's phases in plan order. Returns false when
// nothing was queued.
bool rl_server::queue_step(std::vector<Job*>& js, StepSlot& sl) {
    sl.jobs.clear();
    sl.extra_rays = 0;
    for (Job* j : js) {
        if (job_finished(j) || j->planned) continue;
        if (current_step(j) < 0) {
            j->planned = true;
            continue;
        }
        if (current_step(j) != kStepEmit) continue;
        const int32_t st = rl_emit_photons(ctx, j->desc.scene, &j->ind);
        j->rays += ctx->stats.rays;
        sl.extra_rays += ctx->stats.rays;
        if (st) fail_job(j, st);
        else advance(j);
    }
    // Ray phases. A step bakes one scene, that of the oldest job with ray work, in one launch per
    // phase: the launch holds the scene's top nodes in LDS and reads scene data wave-uniformly (a
    // launch over several scenes loses both; launches per scene within a step measured slower
    // still, being smaller). The step's jobs are the scene's jobs older than the oldest job of any
    // other scene, so large jobs keep admission order across scenes, plus the scene's small jobs
    // (whole remaining plan within a floor share), which are batched alongside. Each gets the floor
    // (a quarter of an even share of the quantum) and the rest goes to them in admission order.
    // A step queues its direct, AO and gather launches in that order, whatever the jobs' plans,
    // hence one phase per job per step. Lumels (and photon tables) are uploaded on entry to a phase.
    std::vector<Job*> cand;
    for (Job* j : js)
        if (!j->planned && !job_finished(j) && ray_phase(j) != kPhaseNone) cand.push_back(j);
    if (cand.empty()) return false;
    const rl_scene* scene = cand[0]->desc.scene;
    const uint64_t floor_rays = std::max<uint64_t>(cfg.quantum_rays / (4 * cand.size()), 4096);
    // Under a backlog (the oldest job waiting longer than kBacklogMs) younger small jobs no longer
    // join its steps, so the oldest jobs, which set the latency tail, are not delayed further.
    const bool backlog = now_ms() - cand[0]->submit_ms > kBacklogMs;
    std::vector<Job*> mine;
    bool prefix = true;
    for (Job* j : cand) {
        if (j->desc.scene != scene) {
            prefix = false;
            continue;
        }
        if (prefix || (!backlog && remaining_rays(j) <= floor_rays)) mine.push_back(j);
    }
    for (size_t k = 0; k < mine.size();) {
        Job* j = mine[k];
        int32_t st = RL_SUCCESS;
        if (!j->prepared) {
            for (size_t m = 0; m < j->maps.size() && !st; m++) {
                st = push_lumels(j->maps[m]);
                if (!st && j->maps[m]->texels.size() != j->valid[m])
                    st = fail(RL_INVALID_CALL, "lightmap changed after submit");
            }
            j->prepared = true;
        }
        if (!st && ray_phase(j) == kPhaseGather && !j->tables) {
            j->tables.reset(new PhotonTables());
            st = make_photon_tables(j->desc.scene, *j->tables);
        }
        if (st) {
            fail_job(j, st);
            mine.erase(mine.begin() + k);
        } else {
            k++;
        }
    }
    if (mine.empty()) return false;
    std::vector<BakeItem> items[4];
    std::vector<std::pair<Job*, RayPhase>> phase_of;
    for (Job* j : mine) phase_of.emplace_back(j, ray_phase(j));
    uint64_t left = cfg.quantum_rays;
    auto take = [&](Job* j, RayPhase ph, uint64_t rays_wanted) {
        if (j->planned || ray_phase(j) != ph) return;  // its phase was completed earlier in this step
        BakeItem proto{};
        uint64_t rpl = 1;
        if (ph == kPhaseDirect) {
            rpl = direct_rays_per_lumel(j->desc.scene);
        } else if (ph == kPhaseAo) {
            proto.ao = ao_params(j->ao);
            rpl = (uint64_t)j->ao.samples;
        } else {
            proto.fg = fg_params(&j->ind);
            proto.mesh_pm = j->tables->d_mesh_pm;
            proto.pms = j->tables->d_pms;
            rpl = (uint64_t)j->ind.final_gather_samples;
        }
        const uint64_t got = take_lumels(j, std::max<uint64_t>(rays_wanted / rpl, 64), proto, items[ph]) * rpl;
        j->rays += got;
        left -= std::min(left, got);
        if (got && std::find(sl.jobs.begin(), sl.jobs.end(), j) == sl.jobs.end()) sl.jobs.push_back(j);
        if (j->map_i >= j->maps.size()) advance(j);
    };
    for (const auto& jp : phase_of) take(jp.first, jp.second, floor_rays);
    for (const auto& jp : phase_of) {
        if (left == 0) break;
        // A phase that would leave less than half a quantum for a later step is taken whole: a
        // launch's tail (lanes that finish early while the last lumels run) costs every launch.
        const uint64_t rest = phase_rays_left(jp.first);
        take(jp.first, jp.second, rest <= left + left / 2 ? rest : left);
    }
    if (sl.jobs.empty()) return false;
    // The phases are queued back to back, counting rays into the step's own counter.
    ctx->pool_used = 0;
    ctx->marks.clear();
    unsigned long long* const own_rays = ctx->d_rays;
    ctx->d_rays = sl.d_rays;
    int32_t st = RL_SUCCESS;
    hipError_t e = launch_fill_u32(ctx->stream, reinterpret_cast<uint32_t*>(sl.d_rays), 2, 0u);
    if (e != hipSuccess) st = fail(RL_DEVICE_ERROR, std::string("server step: ") + hipGetErrorString(e));
    if (!st && !items[kPhaseDirect].empty()) st = run_direct_items(ctx, items[kPhaseDirect], false);
    if (!st && !items[kPhaseAo].empty()) st = run_ao_items(ctx, items[kPhaseAo], false);
    if (!st && !items[kPhaseGather].empty()) st = run_fg_items(ctx, items[kPhaseGather], false);
    ctx->d_rays = own_rays;
    if (!st) {
        e = hipMemcpyAsync(sl.h_rays, sl.d_rays, sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(sl.done, ctx->stream);
        if (e != hipSuccess) st = fail(RL_DEVICE_ERROR, std::string("server step: ") + hipGetErrorString(e));
    }
    if (st) {
        // Launches queued before the failure may still run: wait for them before the jobs retire.
        hipStreamSynchronize(ctx->stream);
        for (Job* j : sl.jobs) fail_job(j, st);
        sl.jobs.clear();
        return false;
    }
    for (Job* j : sl.jobs) j->inflight++;
    sl.t_queued = now_ms();
    sl.live = true;
    return true;
}

// Waits for a queued step; its jobs that no other queued step bakes retire at once.
void rl_server::finish_step(StepSlot& sl) {
    const hipError_t e = hipEventSynchronize(sl.done);
    int32_t st = RL_SUCCESS;
    if (e != hipSuccess) st = fail(RL_DEVICE_ERROR, std::string("server step: ") + hipGetErrorString(e));
    const double t = now_ms();
    std::lock_guard<std::mutex> lk(m);
    for (Job* j : sl.jobs) {
        j->inflight--;
        if (st && !job_finished(j)) fail_job(j, st);
    }
    retire_locked();
    stats.rays += (st ? 0 : *sl.h_rays) + sl.extra_rays;
    stats.busy_ms += t - std::max(sl.t_queued, last_finish_ms);
    stats.steps++;
    last_finish_ms = t;
    sl.jobs.clear();
    sl.live = false;
}

void rl_server::admit_locked() {
    while (!queue.empty() && (int)active.size() < cfg.max_active) {
        Job* j = queue.front();
        if (j->cancel) {
            queue.pop_front();
            j->state = RL_JOB_CANCELLED;
            j->hold(false);
            j->done_ms = now_ms();
            stats.cancelled++;
            cv_done.notify_all();
            continue;
        }
        if (mem_used + j->footprint > cfg.memory_budget) break;
        queue.pop_front();
        j->state = RL_JOB_RUNNING;
        j->admit_ms = now_ms();
        mem_used += j->footprint;
        stats.peak_memory = std::max(stats.peak_memory, mem_used);
        active.push_back(j);
        cv_work.notify_all();
    }
}

// Ends jobs that no queued step bakes any more: planned jobs become done, cancelled ones
// cancelled; finished jobs release their memory and their hold on the caller's objects.
void rl_server::retire_locked() {
    bool any = false;
    for (size_t k = 0; k < active.size();) {
        Job* j = active[k];
        const bool mine = j->desc.kind != RL_JOB_RADIOSITY;  // the auxiliary worker settles its own
        if (mine && j->inflight == 0 && !job_finished(j)) {
            if (j->cancel) j->state = RL_JOB_CANCELLED;
            else if (j->planned) j->state = RL_JOB_DONE;
        }
        if (job_finished(j) && j->inflight == 0) {
            j->done_ms = now_ms();
            if (j->tables) retired_tables.push_back(std::move(j->tables));
            j->hold(false);
            mem_used -= j->footprint;
            if (j->state == RL_JOB_DONE) stats.completed++;
            else if (j->state == RL_JOB_FAILED) stats.failed++;
            else stats.cancelled++;
            active.erase(active.begin() + k);
            any = true;
        } else {
            k++;
        }
    }
    if (any) {
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

// Ray work of every admitted job (direct, AO, final gather, photon emission) in shared steps, one
// step queued ahead of the one running.
void rl_server::loop() {
    std::unique_lock<std::mutex> lk(m);
    int cur = 0;
    for (;;) {
        retire_locked();
        admit_locked();
        std::vector<Job*> js;
        for (Job* j : active)
            if (j->desc.kind != RL_JOB_RADIOSITY && !job_finished(j) && !j->planned && !j->cancel) js.push_back(j);
        StepSlot& prev = slots[1 - cur];
        if (js.empty() && !prev.live) {
            if (!retired_tables.empty() && !slots[cur].live) {
                lk.unlock();
                retired_tables.clear();  // nothing is queued: the frees do not stall a step
                lk.lock();
                continue;
            }
            if (stop && active.empty() && queue.empty()) break;
            cv_work.wait(lk);
            continue;
        }
        if (retired_tables.size() > 64) {  // bounded: free them even with a step in flight
            lk.unlock();
            retired_tables.clear();
            lk.lock();
        }
        lk.unlock();
        const bool queued = !js.empty() && queue_step(js, slots[cur]);
        if (prev.live) finish_step(prev);
        lk.lock();
        if (queued) cur = 1 - cur;
    }
    cv_work.notify_all();
}

// Radiosity jobs (form factors, then the solve) on the auxiliary context's stream, so that their
// GEMM sweeps overlap the ray steps.
void rl_server::aux_loop() {
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
        Job* j = nullptr;
        for (Job* a : active)
            if (a->desc.kind == RL_JOB_RADIOSITY && !a->taken && !job_finished(a)) {
                j = a;
                break;
            }
        if (!j) {
            if (stop && active.empty() && queue.empty()) break;
            cv_work.wait(lk);
            continue;
        }
        j->taken = true;
        lk.unlock();
        int32_t st = RL_SUCCESS;
        uint64_t rays = 0;
        if (!j->cancel) {
            st = rl_radiosity_build_form_factors(aux, j->desc.radiosity, j->desc.rays_per_patch, j->desc.ff_seed,
                                                 j->desc.ff_max_distance);
            rays = aux->stats.rays;
        }
        if (!st && !j->cancel) {
            rl_solve_info info{};
            st = rl_radiosity_solve(aux, j->desc.radiosity, j->E.data(), j->R.data(), j->desc.num_rhs,
                                    j->desc.tolerance, 40, 16, j->desc.out_radiosity, &info);
        }
        lk.lock();
        j->rays += rays;
        stats.rays += rays;
        if (st) fail_job(j, st);
        else j->state = j->cancel ? RL_JOB_CANCELLED : RL_JOB_DONE;
        cv_work.notify_all();
    }
}

namespace {

// Device bytes a job needs while admitted (resident scene data excluded).
uint64_t job_footprint(const Job& j) {
    uint64_t lumels = 0;
    for (uint32_t v : j.valid) lumels += v;
    uint64_t bytes = lumels * (3 * sizeof(float4) + sizeof(uint32_t));
    const bool emits = j.desc.kind == RL_JOB_INDIRECT || (j.desc.kind == RL_JOB_FULL_BAKE && j.desc.indirect);
    if (emits) {
        uint64_t per_pass = 0;
        for (const LightHost& L : j.desc.scene->lights)
            per_pass += j.ind.photons_per_light > 0 ? (uint64_t)j.ind.photons_per_light
                                                    : (uint64_t)std::max(0, (int32_t)(L.dev.intensity * 25000.0f));
        const uint64_t recs = per_pass * (uint64_t)std::max(0, j.ind.photon_pass_count) * (uint64_t)(j.ind.photon_bounce_count + 1);
        bytes += recs * (sizeof(PhotonRec) + 4 * sizeof(uint32_t)) + recs * 2;  // records, keys/indices, sort scratch
        if (j.ind.gather_mode == RL_GATHER_KNN) bytes += recs * (2 * sizeof(float4) + 2 * sizeof(uint32_t));
        for (const MeshHost& mh : j.desc.scene->meshes)
            if (mh.photonmap) bytes += (uint64_t)mh.photonmap->w * mh.photonmap->h * (2 * sizeof(float4) + sizeof(uint32_t) + 1);
    } else if (j.desc.kind == RL_JOB_RADIOSITY) {
        const uint64_t npad = j.desc.radiosity->npad, mpad = (j.desc.num_rhs + 15) / 16 * 16;
        bytes += npad * npad * sizeof(uint16_t) + npad * mpad * (9 * sizeof(float) + 3 * 2);
    }
    return bytes;
}

}  // namespace

extern "C" {

int32_t rl_bake_items(rl_relight* ctx, const rl_bake_item* items, int32_t num_items) {
    if (!ctx || !items || num_items <= 0) return fail(RL_INVALID_ARGUMENT, "bad item list");
    std::vector<BakeItem> direct, ao;
    for (int32_t i = 0; i < num_items; i++) {
        int32_t st = add_item(items[i], direct, ao);
        if (st) return st;
    }
    CallTimer tm(ctx);
    int32_t st = begin_call(ctx);
    if (st) return st;
    if (!direct.empty() && (st = run_direct_items(ctx, direct, false))) return st;
    if (!ao.empty() && (st = run_ao_items(ctx, ao, false))) return st;
    return end_call(ctx, tm);
}

int32_t rl_bake_scene(rl_relight* ctx, rl_scene* scene, int32_t direct, const rl_ao_settings* ao,
                      const rl_indirect_settings* indirect) {
    if (!ctx) return fail(RL_INVALID_ARGUMENT, "context is NULL");
    int32_t st = check_scene_ready(scene);
    if (st) return st;
    if (ao && (st = check_ao(ao))) return st;
    if (indirect && (st = check_indirect_settings(indirect))) return st;
    std::vector<rl_lightmap*> maps;
    for (const MeshHost& m : scene->meshes)
        if (m.lightmap && std::find(maps.begin(), maps.end(), m.lightmap) == maps.end()) maps.push_back(m.lightmap);
    uint64_t photon_rays = 0;
    if (indirect) {
        if ((st = rl_emit_photons(ctx, scene, indirect))) return st;
        photon_rays = ctx->stats.rays;
    }
    PhotonTables T;
    if (indirect && (st = make_photon_tables(scene, T))) return st;
    CallTimer tm(ctx);
    if ((st = begin_call(ctx))) return st;
    std::vector<BakeItem> items;
    if ((st = whole_items(scene, maps.data(), (int)maps.size(), AoParams{}, items))) return st;
    if (direct && (st = run_direct_items(ctx, items, false))) return st;
    if (indirect) {
        std::vector<BakeItem> fg(items);
        for (BakeItem& it : fg) {
            it.fg = fg_params(indirect);
            it.mesh_pm = T.d_mesh_pm;
            it.pms = T.d_pms;
        }
        if ((st = run_fg_items(ctx, fg, false))) return st;
    }
    if (ao) {
        std::vector<BakeItem> a(items);
        for (BakeItem& it : a) it.ao = ao_params(*ao);
        if ((st = run_ao_items(ctx, a, false))) return st;
    }
    st = end_call(ctx, tm);
    ctx->stats.rays += photon_rays;
    return st;
}

static void free_slots(rl_server* s) {
    for (rl_server::StepSlot& sl : s->slots) {
        if (sl.done) hipEventDestroy(sl.done);
        sl.done = nullptr;
        dev_free(sl.d_rays);
        if (sl.h_rays) hipHostFree(sl.h_rays);
        sl.h_rays = nullptr;
    }
}

int32_t rl_server_create(rl_relight* ctx, const rl_server_config* cfg, rl_server** out) {
    if (!ctx || !cfg || !out) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    if (cfg->max_active <= 0 || cfg->quantum_rays == 0 || cfg->memory_budget == 0)
        return fail(RL_INVALID_ARGUMENT, "bad server configuration");
    rl_server* s = new (std::nothrow) rl_server();
    if (!s) return fail(RL_OUT_OF_MEMORY, "server allocation failed");
    s->ctx = ctx;
    s->cfg = *cfg;
    int32_t st = rl_create(&s->aux);
    if (st) {
        delete s;
        return st;
    }
    s->aux->tuning = ctx->tuning;
    s->aux->max_blocks = ctx->max_blocks;
    hipError_t e = hipSuccess;
    for (rl_server::StepSlot& sl : s->slots) {
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e == hipSuccess) e = dev_alloc(&sl.d_rays, 1);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_rays), sizeof(unsigned long long), hipHostMallocDefault);
    }
    if (e != hipSuccess) {
        free_slots(s);
        rl_destroy(s->aux);
        delete s;
        return fail(RL_DEVICE_ERROR, std::string("server set-up: ") + hipGetErrorString(e));
    }
    s->worker = std::thread([s]() { s->loop(); });
    s->aux_worker = std::thread([s]() { s->aux_loop(); });
    *out = s;
    return RL_SUCCESS;
}

int32_t rl_server_submit(rl_server* s, const rl_job_desc* d, uint64_t* out_id) {
    if (!s || !d) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    std::unique_ptr<Job> j(new (std::nothrow) Job());
    if (!j) return fail(RL_OUT_OF_MEMORY, "job allocation failed");
    j->desc = *d;
    int32_t st = check_scene_ready(d->scene);
    if (st) return fail(RL_REJECTED, std::string("job rejected: ") + g_error);
    const bool lightmap_job = d->kind == RL_JOB_DIRECT_AO || d->kind == RL_JOB_INDIRECT ||
                              d->kind == RL_JOB_FINAL_GATHER || d->kind == RL_JOB_FULL_BAKE;
    if (lightmap_job) {
        if (!d->lightmaps || d->num_lightmaps <= 0) return fail(RL_REJECTED, "job rejected: no lightmaps");
        for (int32_t i = 0; i < d->num_lightmaps; i++) {
            rl_lightmap* lm = d->lightmaps[i];
            if (!lm || lm->scene != d->scene) return fail(RL_REJECTED, "job rejected: lightmap not in the job's scene");
            j->maps.push_back(lm);
            j->valid.push_back(valid_lumels(lm));
        }
    }
    if (lightmap_job) {
        const bool needs_ind = d->kind != RL_JOB_DIRECT_AO && (d->kind != RL_JOB_FULL_BAKE || d->indirect);
        const bool takes_ao = d->kind == RL_JOB_DIRECT_AO || d->kind == RL_JOB_FULL_BAKE;
        if (takes_ao && d->ao) {
            if (check_ao(d->ao)) return fail(RL_REJECTED, std::string("job rejected: ") + g_error);
            j->ao = *d->ao;
            j->has_ao = true;
        }
        if (needs_ind) {
            if (check_indirect_settings(d->indirect)) return fail(RL_REJECTED, std::string("job rejected: ") + g_error);
            j->ind = *d->indirect;
        }
        // FULL_BAKE: photons, direct, final gather, AO (AO multiplies the accumulated colour).
        const bool direct = (d->kind == RL_JOB_DIRECT_AO || d->kind == RL_JOB_FULL_BAKE) && d->direct;
        if (needs_ind && d->kind != RL_JOB_FINAL_GATHER) j->plan.push_back(kStepEmit);
        if (direct) j->plan.push_back(kStepDirect);
        if (needs_ind) j->plan.push_back(kStepGather);
        if (j->has_ao) j->plan.push_back(kStepAo);
    } else if (d->kind == RL_JOB_RADIOSITY) {
        if (!d->radiosity || !d->emission || !d->reflectance || !d->out_radiosity || d->num_rhs == 0 ||
            d->num_rhs > 128 || d->radiosity->scene != d->scene)
            return fail(RL_REJECTED, "job rejected: bad radiosity job");
        const size_t n = (size_t)d->radiosity->n * d->num_rhs;
        j->E.assign(d->emission, d->emission + n);
        j->R.assign(d->reflectance, d->reflectance + n);
    } else {
        return fail(RL_REJECTED, "job rejected: unknown kind");
    }
    j->footprint = job_footprint(*j);
    std::lock_guard<std::mutex> lk(s->m);
    s->stats.submitted++;
    if (j->footprint > s->cfg.memory_budget) {
        s->stats.rejected++;
        return fail(RL_REJECTED, "job rejected: footprint exceeds the memory budget");
    }
    if (s->stop) return fail(RL_INVALID_CALL, "server is stopping");
    j->id = s->next_id++;
    j->submit_ms = s->now_ms();
    Job* raw = j.get();
    raw->hold(true);
    s->jobs[raw->id] = std::move(j);
    s->queue.push_back(raw);
    if (out_id) *out_id = raw->id;
    s->cv_work.notify_all();
    return RL_SUCCESS;
}

static void fill_status(const Job* j, rl_job_status* out) {
    if (j->state == RL_JOB_FAILED) g_error = j->error;
    out->state = j->state;
    out->status = j->status;
    out->submit_ms = j->submit_ms;
    out->admit_ms = j->admit_ms;
    out->done_ms = j->done_ms;
    out->rays = j->rays;
    out->footprint = j->footprint;
}

int32_t rl_server_poll(rl_server* s, uint64_t id, rl_job_status* out) {
    if (!s || !out) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(RL_INVALID_ARGUMENT, "unknown job id");
    fill_status(it->second.get(), out);
    return RL_SUCCESS;
}

int32_t rl_server_wait(rl_server* s, uint64_t id, double timeout_ms, rl_job_status* out) {
    if (!s) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(RL_INVALID_ARGUMENT, "unknown job id");
    Job* j = it->second.get();
    auto done = [j]() { return job_finished(j); };
    if (timeout_ms < 0) {
        s->cv_done.wait(lk, done);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), done)) {
        if (out) fill_status(j, out);
        return fail(RL_REJECTED, "timed out waiting for the job");
    }
    if (out) fill_status(j, out);
    return RL_SUCCESS;
}

int32_t rl_server_cancel(rl_server* s, uint64_t id) {
    if (!s) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(RL_INVALID_ARGUMENT, "unknown job id");
    it->second->cancel = true;
    s->cv_work.notify_all();
    return RL_SUCCESS;
}

int32_t rl_server_stats_get(rl_server* s, rl_server_stats* out) {
    if (!s || !out) return fail(RL_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(s->m);
    *out = s->stats;
    return RL_SUCCESS;
}

int32_t rl_server_destroy(rl_server* s) {
    if (!s) return RL_SUCCESS;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
        for (Job* j : s->queue) j->cancel = true;
        s->cv_work.notify_all();
    }
    s->worker.join();
    s->aux_worker.join();
    // Jobs still queued when the loop stopped were never admitted.
    for (Job* j : s->queue) {
        j->state = RL_JOB_CANCELLED;
        j->hold(false);
        s->stats.cancelled++;
    }
    for (Job* j : s->active) j->hold(false);
    s->retired_tables.clear();
    free_slots(s);
    rl_destroy(s->aux);
    delete s;
    return RL_SUCCESS;
}

}  // extern "C"
