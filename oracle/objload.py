"""ORACLE — test infrastructure only. A Python restatement of MeshLoader::load
(scene/MeshLoader.cpp:44-226) used to check the library's OBJ loader: submeshes committed at
v/vt/vn lines that follow faces (flags cleared), face format by flags (any of uv/normal ->
v/vt/vn triples), corners deduplicated exactly in first-appearance order, uvs clamped to [0, 1],
missing normals (0, 1, 0), triangles only, submeshes appended into one mesh (Mesh::addFaces)."""
import numpy as np


def load(path):
    xyz, nrm, uv = [], [], []
    faces = []  # pending (xyz, uv, n) 0-based corner indices
    flags = {"uv": False, "n": False}
    out_pos, out_nrm, out_uv, out_idx = [], [], [], []

    def f32(s):
        return float(np.float32(s))

    def commit():
        if not faces:
            return
        index, verts, idx = {}, [], []
        for p, t, q in faces:
            key = tuple(xyz[p]) + ((tuple(nrm[q]) if flags["n"] else (0.0, 1.0, 0.0)) + (1.0, 1.0, 1.0))
            if flags["uv"]:
                u, v = uv[t]
                key += (min(max(u, 0.0), 1.0) if u == u else u, min(max(v, 0.0), 1.0) if v == v else v)
            else:
                key += (0.0, 0.0)
            if key not in index:
                index[key] = len(verts)
                verts.append(key)
            idx.append(index[key])
        base = len(out_pos)
        for k in verts:
            out_pos.append(k[0:3])
            out_nrm.append(k[3:6])
            out_uv.append(k[9:11])
        out_idx.extend(base + i for i in idx)
        faces.clear()
        flags["uv"] = flags["n"] = False

    with open(path) as fh:
        for line in fh.read().split("\n"):
            c1, c2 = line[:1], line[1:2]
            if c1 == "v" and c2 == "t":
                commit()
                a = line[2:].split()
                uv.append((f32(a[0]), f32(a[1])))
                flags["uv"] = True
            elif c1 == "v" and c2 == "n":
                commit()
                a = line[2:].split()
                nrm.append((f32(a[0]), f32(a[1]), f32(a[2])))
                flags["n"] = True
            elif c1 == "v":
                commit()
                a = line[1:].split()
                xyz.append((f32(a[0]), f32(a[1]), f32(a[2])))
            if c1 == "f":
                corners = line[2:].split()[:3]
                for c in corners:
                    if flags["uv"] or flags["n"]:
                        p, t, q = (int(x) for x in c.split("/")[:3])
                    else:
                        p, t, q = int(c), 0, 0
                    faces.append((p - 1, t - 1, q - 1))
    commit()
    return (np.asarray(out_pos, np.float32).reshape(-1, 3), np.asarray(out_nrm, np.float32).reshape(-1, 3),
            np.asarray(out_uv, np.float32).reshape(-1, 2), np.asarray(out_idx, np.uint32))
