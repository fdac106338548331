// Lightmap UV generation: Mesh::generateUv (scene/Mesh.cpp:229-238) with UvGenerator,
// AngularChartifier, the half-edge mesh and RectanglePacker (Types.h:2293-2600,2825-3326). Host
// code: it prepares meshes before the bake (the GPU then rasterises their texels).
//
// The steps, in the reference's order and float arithmetic:
//  1. Half edges per face (vertex i -> i+1); an edge's twin is the earliest-added edge running the
//     other way between the same two vertex indices (faces that only share positions are not
//     connected).
//  2. Charts: for every edge in order, a depth-first flood from its face over twin edges, taking
//     faces whose normal is within `angle` degrees of the seed face's normal.
//  3. Per chart: positions projected onto the plane of the dominant axis of the chart's mean
//     normal; rotated onto the principal axis of the points (2x2 covariance, ten power steps from
//     (1, 1)) when that shrinks the bounding box; moved to the origin and transposed when taller
//     than wide.
//  4. Vertices de-duplicated on exact equality of position, normal and both uv layers.
//  5. Charts packed as rectangles (longer side first, plus padding) by a guillotine packer into
//     a w x h area grown from 1 x 1 by 0.1 alternately in width and height until all fit,
//     rectangles placed in decreasing area (std::sort); uvs offset by their rectangle, vertices
//     de-duplicated again and the uv set scaled to [0, 1] per axis.
#include <stdint.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/relight.h"

namespace rl {
int32_t report_error(int32_t code, const std::string& msg);
}

namespace {

struct UvVertex {
    float p[3], n[3], uv0[2], uv1[2];
};

// Vertex::Compare (scene/Mesh.h:54-70): position, normal, then the uv layers, by operator<.
struct VertexLess {
    bool operator()(const UvVertex& a, const UvVertex& b) const {
        for (int i = 0; i < 3; i++)
            if (a.p[i] != b.p[i]) return a.p[i] < b.p[i];
        for (int i = 0; i < 3; i++)
            if (a.n[i] != b.n[i]) return a.n[i] < b.n[i];
        for (int i = 0; i < 2; i++)
            if (a.uv0[i] != b.uv0[i]) return a.uv0[i] < b.uv0[i];
        for (int i = 0; i < 2; i++)
            if (a.uv1[i] != b.uv1[i]) return a.uv1[i] < b.uv1[i];
        return false;
    }
};

// MeshIndexer (Types.h:2458-2556).
struct Indexer {
    std::map<UvVertex, uint32_t, VertexLess> cache;
    std::vector<UvVertex> vertices;
    std::vector<uint32_t> indices;
    uint32_t add(const UvVertex& v) {
        auto it = cache.find(v);
        uint32_t idx;
        if (it != cache.end()) {
            idx = it->second;
        } else {
            idx = (uint32_t)vertices.size();
            cache[v] = idx;
            vertices.push_back(v);
        }
        indices.push_back(idx);
        return idx;
    }
};

struct V3f {
    float x, y, z;
};

// TriMesh::Face::normal (Types.h:2873-2886): (v1 - v0) x (v2 - v0), Vec3::normalize (Types.h:852-859).
V3f face_normal(const std::vector<UvVertex>& V, const std::vector<uint32_t>& I, uint32_t f) {
    const float* a = V[I[3 * f]].p;
    const float* b = V[I[3 * f + 1]].p;
    const float* c = V[I[3 * f + 2]].p;
    const float ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
    const float vx = c[0] - a[0], vy = c[1] - a[1], vz = c[2] - a[2];
    V3f n{uy * vz - uz * vy, uz * vx - ux * vz, ux * vy - uy * vx};
    const float len = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
    if (len) {
        n.x /= len;
        n.y /= len;
        n.z /= len;
    }
    return n;
}

struct Chartifier {
    struct Edge {
        uint32_t face, vertex;
        int32_t next, twin;
    };
    std::vector<Edge> edges;
    std::vector<V3f> normals;
    std::vector<int32_t> chart_of;                // face -> chart, -1: none yet
    std::vector<std::vector<uint32_t>> charts;   // faces in the order they were added
    float angle = 88.0f;

    // DCEL::create (Types.h:2380-2423).
    void build_dcel(const std::vector<uint32_t>& I) {
        const uint32_t nf = (uint32_t)I.size() / 3;
        edges.assign(3 * (size_t)nf, Edge{0, 0, -1, -1});
        std::multimap<std::pair<uint32_t, uint32_t>, int32_t> cache;
        for (uint32_t f = 0; f < nf; f++) {
            for (int i = 0; i < 3; i++) {
                edges[3 * f + i].face = f;
                edges[3 * f + i].vertex = I[3 * f + i];
            }
            for (int i = 0; i < 3; i++) {
                Edge& e = edges[3 * f + i];
                e.next = (int32_t)(3 * f + (i + 1) % 3);
                cache.insert({{e.vertex, edges[e.next].vertex}, (int32_t)(3 * f + i)});
            }
        }
        for (size_t i = 0; i < edges.size(); i++) {
            Edge& e = edges[i];
            auto z = cache.find({edges[e.next].vertex, e.vertex});
            if (z != cache.end() && (int32_t)i != z->second) {
                e.twin = z->second;
                edges[z->second].twin = (int32_t)i;
            }
        }
    }

    // AngularChartifier::addToChart (Types.h:2566-2595), the recursion unrolled onto a stack that
    // keeps its visiting order.
    void add_to_chart(int32_t edge0, V3f axis, int32_t index) {
        struct Frame {
            int32_t edge;
            int k;  // 0: not entered; 1..3: next edge of the face to follow; 4: done
        };
        std::vector<Frame> stack{{edge0, 0}};
        while (!stack.empty()) {
            Frame& fr = stack.back();
            const Edge& e = edges[fr.edge];
            if (fr.k == 0) {
                if (chart_of[e.face] >= 0) {
                    stack.pop_back();
                    continue;
                }
                const V3f n = normals[e.face];
                const float d = axis.x * n.x + axis.y * n.y + axis.z * n.z;
                const float deg = acosf(d) * 180.0f / 3.14159265358979323846f;
                if (deg > angle) {
                    stack.pop_back();
                    continue;
                }
                chart_of[e.face] = index;
                if ((int32_t)charts.size() <= index) charts.resize(index + 1);
                charts[index].push_back(e.face);
                fr.k = 1;
            }
            if (fr.k > 3) {
                stack.pop_back();
                continue;
            }
            // The face's edges from the entering one: k = 1, 2, 3.
            int32_t cur = fr.edge;
            for (int s = 1; s < fr.k; s++) cur = edges[cur].next;
            fr.k++;
            const int32_t tw = edges[cur].twin;
            if (tw >= 0) stack.push_back(Frame{tw, 0});
        }
    }
};

struct Vec2f {
    float x, y;
};

// UvGenerator::calculateBoundingRect (Types.h:2977-2993).
Vec2f bounding_rect(const std::vector<Vec2f>& uv, Vec2f& mn, Vec2f& mx) {
    mn = {FLT_MAX, FLT_MAX};
    mx = {-FLT_MAX, -FLT_MAX};
    for (const Vec2f& v : uv) {
        mn.x = std::min(mn.x, v.x);
        mx.x = std::max(mx.x, v.x);
        mn.y = std::min(mn.y, v.y);
        mx.y = std::max(mx.y, v.y);
    }
    return {mx.x - mn.x, mx.y - mn.y};
}

// UvGenerator::rotate (Types.h:3031-3060): Samples::centered, covarianceMatrix (population),
// CovMatrix::converge (ten products, no rescaling) and Vector::normalize.
void rotate(std::vector<Vec2f>& uv) {
    const uint32_t n = (uint32_t)uv.size();
    Vec2f sum{0.0f, 0.0f};
    for (const Vec2f& v : uv) {
        sum.x += v.x;
        sum.y += v.y;
    }
    const float nf = (float)n;
    const Vec2f mean{sum.x / nf, sum.y / nf};
    std::vector<Vec2f> c(n);
    for (uint32_t i = 0; i < n; i++) c[i] = {uv[i].x - mean.x, uv[i].y - mean.y};
    Vec2f csum{0.0f, 0.0f};
    for (const Vec2f& v : c) {
        csum.x += v.x;
        csum.y += v.y;
    }
    const Vec2f cm{csum.x / nf, csum.y / nf};
    float cov[2][2];
    for (int row = 0; row < 2; row++) {
        for (int col = 0; col < 2; col++) {
            float s = 0.0f;
            for (uint32_t i = 0; i < n; i++) {
                const float a = (row ? c[i].y : c[i].x) - (row ? cm.y : cm.x);
                const float b = (col ? c[i].y : c[i].x) - (col ? cm.y : cm.x);
                s += a * b;
            }
            cov[row][col] = s / (float)n;  // stored at result(col, row): m[row * 2 + col]
        }
    }
    // Matrix * Vector: C[i] = sum_k m[i * 2 + k] * v[k] (Types.h:1962-1978).
    float e[2] = {1.0f, 1.0f};
    for (int it = 0; it < 10; it++) {
        float r[2];
        for (int i = 0; i < 2; i++) {
            float s = 0.0f;
            for (int k = 0; k < 2; k++) s = s + cov[i][k] * e[k];
            r[i] = s;
        }
        e[0] = r[0];
        e[1] = r[1];
    }
    const float len = std::sqrt(0.0f + e[0] * e[0] + e[1] * e[1]);
    if (len != 0.0f) {
        e[0] /= len;
        e[1] /= len;
    }
    const float f[2] = {e[1], -e[0]};
    std::vector<Vec2f> rot(n);
    for (uint32_t i = 0; i < n; i++) rot[i] = {e[0] * uv[i].x + e[1] * uv[i].y, f[0] * uv[i].x + f[1] * uv[i].y};
    Vec2f bmn, bmx, amn, amx;
    const Vec2f before = bounding_rect(uv, bmn, bmx);
    const Vec2f after = bounding_rect(rot, amn, amx);
    if ((after.x * after.y) < (before.x * before.y)) uv = rot;
}

// RectanglePacker (Types.h:3163-3326). margin is accepted and unused, as in the reference
// (its constructor sets m_margin to 0).
struct Packer {
    struct Rect {
        float x, y, w, h;
    };
    std::vector<Rect> rects;
    float padding = 0.0f;

    struct Node {
        int32_t left = -1, right = -1;
        bool free = true;
        Rect r;
    };

    void add(float w, float h) { rects.push_back(Rect{0.0f, 0.0f, w + padding, h + padding}); }

    // Node::assign.
    static bool assign(std::vector<Node>& nodes, int32_t ni, Rect& placed) {
        for (;;) {
            Node& nd = nodes[ni];
            if (nd.left >= 0 || nd.right >= 0) {
                const int32_t l = nd.left, r = nd.right;
                if (assign(nodes, l, placed)) return true;
                ni = r;
                continue;
            }
            if (placed.w > nd.r.w || placed.h > nd.r.h || !nd.free) return false;
            if (placed.w == nd.r.w && placed.h == nd.r.h) {
                placed.x = nd.r.x;
                placed.y = nd.r.y;
                nd.free = false;
                return true;
            }
            const float dw = nd.r.w - placed.w, dh = nd.r.h - placed.h;
            const Rect rr = nd.r;
            Node a, b;
            if (dw > dh) {
                a.r = Rect{rr.x, rr.y, placed.w, rr.h};
                b.r = Rect{rr.x + placed.w, rr.y, rr.w - placed.w, rr.h};
            } else {
                a.r = Rect{rr.x, rr.y, rr.w, placed.h};
                b.r = Rect{rr.x, rr.y + placed.h, rr.w, rr.h - placed.h};
            }
            nodes.push_back(a);
            nodes.push_back(b);
            nodes[ni].left = (int32_t)nodes.size() - 2;
            nodes[ni].right = (int32_t)nodes.size() - 1;
            // The node is no longer a leaf: assign again from it (the reference recurses).
        }
    }

    // RectanglePacker::place with Rect::compareByArea.
    bool place(float width, float height) {
        std::vector<Rect*> order;
        for (Rect& r : rects) order.push_back(&r);
        std::sort(order.begin(), order.end(), [](const Rect* a, const Rect* b) { return a->w * a->h > b->w * b->h; });
        std::vector<Node> nodes(1);
        nodes[0].r = Rect{0.0f, 0.0f, width - 0.0f, height - 0.0f};
        for (Rect* r : order)
            if (!assign(nodes, 0, *r)) return false;
        return true;
    }
};

}  // namespace

extern "C" {

int32_t rl_mesh_generate_uv(const float* positions, const float* normals, const float* diffuse_uvs,
                            uint32_t vertex_count, const uint32_t* indices, uint32_t index_count, float angle,
                            float margin, float padding, uint32_t* out_vertex_count, uint32_t* out_index_count,
                            float* out_positions, float* out_normals, float* out_diffuse_uvs,
                            float* out_lightmap_uvs, uint32_t* out_indices) {
    (void)margin;
    if (!positions || !normals || !indices || !out_vertex_count || !out_index_count)
        return rl::report_error(RL_INVALID_ARGUMENT, "NULL argument");
    if (index_count % 3 != 0) return rl::report_error(RL_INVALID_ARGUMENT, "index count is not a multiple of 3");
    for (uint32_t i = 0; i < index_count; i++)
        if (indices[i] >= vertex_count) return rl::report_error(RL_INVALID_ARGUMENT, "index out of range");
    std::vector<UvVertex> V(vertex_count);
    for (uint32_t i = 0; i < vertex_count; i++) {
        UvVertex& v = V[i];
        std::memcpy(v.p, positions + 3 * (size_t)i, 3 * sizeof(float));
        std::memcpy(v.n, normals + 3 * (size_t)i, 3 * sizeof(float));
        v.uv0[0] = diffuse_uvs ? diffuse_uvs[2 * (size_t)i] : 0.0f;
        v.uv0[1] = diffuse_uvs ? diffuse_uvs[2 * (size_t)i + 1] : 0.0f;
        v.uv1[0] = v.uv1[1] = 0.0f;
    }
    const std::vector<uint32_t> I(indices, indices + index_count);
    const uint32_t nf = index_count / 3;
    // Charts (AngularChartifier::build, Types.h:2552-2563).
    Chartifier C;
    C.angle = angle;
    C.build_dcel(I);
    C.normals.resize(nf);
    for (uint32_t f = 0; f < nf; f++) C.normals[f] = face_normal(V, I, f);
    C.chart_of.assign(nf, -1);
    for (size_t e = 0; e < C.edges.size(); e++)
        C.add_to_chart((int32_t)e, C.normals[C.edges[e].face], (int32_t)C.charts.size());
    // Flatten, rotate, place each chart at the origin (UvGenerator::generate, Types.h:3063-3110).
    Indexer ix;
    std::vector<std::vector<uint32_t>> chart_vertices(C.charts.size());
    for (size_t ci = 0; ci < C.charts.size(); ci++) {
        const std::vector<uint32_t>& faces = C.charts[ci];
        // Chart::normal (Types.h:2790-2803): the mean face normal, normalised; its ordinal axis.
        V3f cn{0.0f, 0.0f, 0.0f};
        for (uint32_t f : faces) {
            cn.x += C.normals[f].x;
            cn.y += C.normals[f].y;
            cn.z += C.normals[f].z;
        }
        const float cnt = (float)faces.size();
        cn.x /= cnt;
        cn.y /= cnt;
        cn.z /= cnt;
        const float cl = sqrtf(cn.x * cn.x + cn.y * cn.y + cn.z * cn.z);
        if (cl) {
            cn.x /= cl;
            cn.y /= cl;
            cn.z /= cl;
        }
        const float ax = fabsf(cn.x), ay = fabsf(cn.y), az = fabsf(cn.z);
        const int drop = (ax > ay && ax > az) ? 0 : ((ay > ax && ay > az) ? 1 : 2);
        // Face::flatten (Types.h:2825-2871): the position with the ordinal axis' component removed.
        std::vector<Vec2f> uv;
        for (uint32_t f : faces) {
            for (int k = 0; k < 3; k++) {
                const float* p = V[I[3 * f + k]].p;
                if (drop == 0) uv.push_back({p[1], p[2]});
                else if (drop == 1) uv.push_back({p[0], p[2]});
                else uv.push_back({p[0], p[1]});
            }
        }
        rotate(uv);
        Vec2f mn, mx;
        const Vec2f size = bounding_rect(uv, mn, mx);
        size_t q = 0;
        for (uint32_t f : faces) {
            for (int k = 0; k < 3; k++) {
                UvVertex vtx = V[I[3 * f + k]];
                const Vec2f v = uv[q++];
                if (size.x > size.y) {
                    vtx.uv1[0] = v.x - mn.x;
                    vtx.uv1[1] = v.y - mn.y;
                } else {
                    vtx.uv1[0] = v.y - mn.y;
                    vtx.uv1[1] = v.x - mn.x;
                }
                chart_vertices[ci].push_back(ix.add(vtx));
            }
        }
    }
    const std::vector<UvVertex> verts = ix.vertices;
    Packer P;
    P.padding = padding;
    float max_w = 0.0f, max_h = 0.0f;
    double area = 0.0;  // exact products of the float sides, summed in double
    for (size_t ci = 0; ci < C.charts.size(); ci++) {
        Vec2f mn{FLT_MAX, FLT_MAX}, mx{-FLT_MAX, -FLT_MAX};
        for (uint32_t vi : chart_vertices[ci]) {
            const float* t = verts[vi].uv1;
            mn.x = std::min(t[0], mn.x);
            mn.y = std::min(t[1], mn.y);
            mx.x = std::max(t[0], mx.x);
            mx.y = std::max(t[1], mx.y);
        }
        const Vec2f size{mx.x - mn.x, mx.y - mn.y};
        P.add(std::max(size.x, size.y), std::min(size.x, size.y));
        max_w = std::max(max_w, P.rects.back().w);
        max_h = std::max(max_h, P.rects.back().h);
        area += (double)P.rects.back().w * (double)P.rects.back().h;
    }
    // UvGenerator::pack (Types.h:3010-3028). An area that cannot hold every rectangle (one is wider
    // or taller, or their areas add up to more: placed rectangles are disjoint) would fail to
    // place; it is skipped without building the tree.
    float w = 1.0f, h = 1.0f;
    bool expand_width = true;
    for (;;) {
        const bool hopeless = max_w > w || max_h > h || area > (double)w * (double)h * (1.0 + 1e-9);
        if (!hopeless && P.place(w, h)) break;
        if (expand_width) {
            w += 0.1f;
            expand_width = false;
        } else {
            h += 0.1f;
            expand_width = true;
        }
    }
    Indexer out;
    for (size_t ci = 0; ci < C.charts.size(); ci++) {
        const Packer::Rect& r = P.rects[ci];
        for (uint32_t vi : chart_vertices[ci]) {
            UvVertex vtx = verts[vi];
            vtx.uv1[0] = r.x + vtx.uv1[0];
            vtx.uv1[1] = r.y + vtx.uv1[1];
            out.add(vtx);
        }
    }
    // Normalise the uv set to [0, 1] per axis.
    Vec2f mn{FLT_MAX, FLT_MAX}, mx{-FLT_MAX, -FLT_MAX};
    for (uint32_t vi : out.indices) {
        const float* t = out.vertices[vi].uv1;
        mn.x = std::min(t[0], mn.x);
        mn.y = std::min(t[1], mn.y);
        mx.x = std::max(t[0], mx.x);
        mx.y = std::max(t[1], mx.y);
    }
    for (UvVertex& v : out.vertices) {
        const float u = v.uv1[0], t = v.uv1[1];
        v.uv1[0] = (u - mn.x) / (mx.x - mn.x);
        v.uv1[1] = (t - mn.y) / (mx.y - mn.y);
    }
    *out_vertex_count = (uint32_t)out.vertices.size();
    *out_index_count = (uint32_t)out.indices.size();
    for (size_t i = 0; i < out.vertices.size(); i++) {
        const UvVertex& v = out.vertices[i];
        if (out_positions) std::memcpy(out_positions + 3 * i, v.p, 3 * sizeof(float));
        if (out_normals) std::memcpy(out_normals + 3 * i, v.n, 3 * sizeof(float));
        if (out_diffuse_uvs) std::memcpy(out_diffuse_uvs + 2 * i, v.uv0, 2 * sizeof(float));
        if (out_lightmap_uvs) std::memcpy(out_lightmap_uvs + 2 * i, v.uv1, 2 * sizeof(float));
    }
    if (out_indices) std::memcpy(out_indices, out.indices.data(), out.indices.size() * sizeof(uint32_t));
    return RL_SUCCESS;
}

}  // extern "C"
