// Wavefront OBJ ingestion with the reference loader's semantics (scene/MeshLoader.cpp:44-226):
//   - `v`, `vt`, `vn` lines first commit the faces read so far as a submesh and then clear the
//     uv/normal flags; `f` lines are read with the format the flags select: any of uv/normal seen
//     -> `v/vt/vn` triples, else `v v v`; only the first three corners count; indices are 1-based;
//   - a submesh's corners are deduplicated on (position, normal, colour 1, lightmap uv) with exact
//     float comparison, indices in first-appearance order (std::map<Vertex, int>);
//   - uvs are clamped to [0, 1]; a missing normal is (0, 1, 0); submeshes append to one mesh
//     (Mesh::addFaces, scene/Mesh.cpp:146-165).
// Where the reference reads undefined values (an index out of range, an unparsable number) this
// loader reports RL_INVALID_ARGUMENT instead.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "../../include/relight.h"

namespace {

struct ObjMesh {
    std::vector<float> pos, nrm, uv;
    std::vector<uint32_t> idx;
};

struct Corner {
    float v[11];  // position, normal, colour, uv
    bool operator<(const Corner& o) const {
        for (int i = 0; i < 11; i++)
            if (v[i] != o.v[i]) return v[i] < o.v[i];
        return false;
    }
};

class ObjReader {
public:
    std::string error;

    bool load(const char* path, ObjMesh& out) {
        std::ifstream file(path);
        if (!file.is_open()) {
            error = std::string("cannot open ") + path;
            return false;
        }
        out_ = &out;
        std::string line;
        while (std::getline(file, line)) {
            const char c1 = line.size() > 0 ? line[0] : '\0', c2 = line.size() > 1 ? line[1] : '\0';
            if (c1 == 'v' && c2 == 't') {
                if (!commit()) return false;
                float x = 0.0f, y = 0.0f;
                if (std::sscanf(line.c_str() + 2, "%f %f", &x, &y) != 2) return bad("unparsable vt line");
                uv_.push_back(x);
                uv_.push_back(y);
                flags_ |= kUv;
            } else if (c1 == 'v' && c2 == 'n') {
                if (!commit()) return false;
                float x = 0.0f, y = 0.0f, z = 0.0f;
                if (std::sscanf(line.c_str() + 2, "%f %f %f", &x, &y, &z) != 3) return bad("unparsable vn line");
                nrm_.push_back(x);
                nrm_.push_back(y);
                nrm_.push_back(z);
                flags_ |= kNormals;
            } else if (c1 == 'v') {
                if (!commit()) return false;
                float x = 0.0f, y = 0.0f, z = 0.0f;
                if (std::sscanf(line.c_str() + 1, "%f %f %f", &x, &y, &z) != 3) return bad("unparsable v line");
                xyz_.push_back(x);
                xyz_.push_back(y);
                xyz_.push_back(z);
            }
            if (c1 == 'f' && !face(line)) return false;
        }
        return commit();
    }

private:
    enum { kUv = 2, kNormals = 4 };
    ObjMesh* out_ = nullptr;
    std::vector<float> xyz_, nrm_, uv_;
    std::vector<int> fxyz_, fuv_, fnrm_;  // pending submesh corners (0-based)
    int flags_ = 0;

    bool bad(const char* what) {
        error = what;
        return false;
    }

    bool face(const std::string& line) {
        int xyz[3] = {0, 0, 0}, uv[3] = {0, 0, 0}, n[3] = {0, 0, 0};
        const char* s = line.size() > 2 ? line.c_str() + 2 : "";  // the reference skips "f" and one character
        int got;
        if (flags_ & (kNormals | kUv)) {
            got = std::sscanf(s, "%i/%i/%i %i/%i/%i %i/%i/%i", &xyz[0], &uv[0], &n[0], &xyz[1], &uv[1], &n[1], &xyz[2],
                              &uv[2], &n[2]);
            if (got != 9) return bad("face without v/vt/vn triples");
        } else {
            got = std::sscanf(s, "%i %i %i", &xyz[0], &xyz[1], &xyz[2]);
            if (got != 3) return bad("face without three vertex indices");
        }
        for (int i = 0; i < 3; i++) {
            fxyz_.push_back(xyz[i] - 1);
            fuv_.push_back(uv[i] - 1);
            fnrm_.push_back(n[i] - 1);
        }
        return true;
    }

    // MeshLoader::commit: the pending faces as one submesh with deduplicated corners.
    bool commit() {
        if (fxyz_.empty()) return true;
        std::map<Corner, uint32_t> index;
        std::vector<Corner> verts;
        std::vector<uint32_t> idx;
        const bool has_n = flags_ & kNormals, has_uv = flags_ & kUv;
        for (size_t c = 0; c < fxyz_.size(); c++) {
            const int p = fxyz_[c], q = fnrm_[c], t = fuv_[c];
            if (p < 0 || (size_t)p * 3 + 2 >= xyz_.size()) return bad("vertex index out of range");
            if (has_n && (q < 0 || (size_t)q * 3 + 2 >= nrm_.size())) return bad("normal index out of range");
            if (has_uv && (t < 0 || (size_t)t * 2 + 1 >= uv_.size())) return bad("uv index out of range");
            Corner k;
            for (int i = 0; i < 3; i++) k.v[i] = xyz_[3 * (size_t)p + i];
            k.v[3] = has_n ? nrm_[3 * (size_t)q] : 0.0f;
            k.v[4] = has_n ? nrm_[3 * (size_t)q + 1] : 1.0f;
            k.v[5] = has_n ? nrm_[3 * (size_t)q + 2] : 0.0f;
            k.v[6] = k.v[7] = k.v[8] = 1.0f;
            k.v[9] = k.v[10] = 0.0f;
            if (has_uv) {
                float u = uv_[2 * (size_t)t], v = uv_[2 * (size_t)t + 1];
                if (u > 1.0f) u = 1.0f;
                if (v > 1.0f) v = 1.0f;
                if (u < 0.0f) u = 0.0f;
                if (v < 0.0f) v = 0.0f;
                k.v[9] = u;
                k.v[10] = v;
            }
            auto it = index.find(k);
            if (it == index.end()) {
                const uint32_t id = (uint32_t)verts.size();
                index.emplace(k, id);
                verts.push_back(k);
                idx.push_back(id);
            } else {
                idx.push_back(it->second);
            }
        }
        const uint32_t base = (uint32_t)(out_->pos.size() / 3);
        for (const Corner& k : verts) {
            out_->pos.insert(out_->pos.end(), k.v, k.v + 3);
            out_->nrm.insert(out_->nrm.end(), k.v + 3, k.v + 6);
            out_->uv.insert(out_->uv.end(), k.v + 9, k.v + 11);
        }
        for (uint32_t i : idx) out_->idx.push_back(base + i);
        fxyz_.clear();
        fuv_.clear();
        fnrm_.clear();
        flags_ = 0;
        return true;
    }
};

}  // namespace

namespace rl {
int32_t report_error(int32_t code, const std::string& msg);  // rl_runtime.cpp (rl_last_error)
}

namespace {
int32_t obj_fail(int32_t code, const std::string& msg) { return rl::report_error(code, msg); }
}  // namespace

extern "C" {

int32_t rl_obj_load(const char* path, uint32_t* vertex_count, uint32_t* index_count, float* positions, float* normals,
                    float* uvs, uint32_t* indices) {
    if (!path || !vertex_count || !index_count) return obj_fail(RL_INVALID_ARGUMENT, "NULL argument");
    ObjMesh m;
    ObjReader r;
    if (!r.load(path, m)) return obj_fail(RL_INVALID_ARGUMENT, r.error);
    const uint32_t nv = (uint32_t)(m.pos.size() / 3), ni = (uint32_t)m.idx.size();
    if (positions || normals || uvs || indices) {
        if (*vertex_count < nv || *index_count < ni) return obj_fail(RL_INVALID_ARGUMENT, "output arrays too small");
        if (positions) std::memcpy(positions, m.pos.data(), m.pos.size() * sizeof(float));
        if (normals) std::memcpy(normals, m.nrm.data(), m.nrm.size() * sizeof(float));
        if (uvs) std::memcpy(uvs, m.uv.data(), m.uv.size() * sizeof(float));
        if (indices) std::memcpy(indices, m.idx.data(), m.idx.size() * sizeof(uint32_t));
    }
    *vertex_count = nv;
    *index_count = ni;
    return RL_SUCCESS;
}

int32_t rl_scene_add_obj(rl_scene* scene, const char* path, const float* transform, const float* albedo,
                         int32_t* out_mesh) {
    if (!scene || !path) return obj_fail(RL_INVALID_ARGUMENT, "NULL argument");
    ObjMesh m;
    ObjReader r;
    if (!r.load(path, m)) return obj_fail(RL_INVALID_ARGUMENT, r.error);
    return rl_scene_add_mesh(scene, m.pos.data(), m.nrm.data(), m.uv.data(), (uint32_t)(m.pos.size() / 3),
                             m.idx.data(), (uint32_t)m.idx.size(), transform, albedo, out_mesh);
}

}  // extern "C"
