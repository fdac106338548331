"""Radiosity patch workloads: a block city built from 1 m quad patches.

The ground is a G x G m grid of 1 m patches on which integer-sized boxes stand, one per lot of an
L x L lot grid; every visible face is split into 1 m quads (box bottoms are omitted and the ground
patches they cover are removed). Each box adds 2 (w + d) h patches net of the ground it covers, so
heights are chosen until the total is exactly `target`. Each patch is two triangles; tri_patch
maps a triangle to its patch.

Emission: `sets` random emission sets, each lighting `emitters` random patches with an RGB
emission U(1, 10); reflectance rho per patch per channel U(0.2, 0.8). Right-hand side c = 3*set + ch.
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class PatchScene:
    positions: np.ndarray   # (V, 3) float32
    normals: np.ndarray     # (V, 3)
    indices: np.ndarray     # (T*3,) uint32
    patches: np.ndarray     # (N, 12) float32: origin, U, V, normal
    tri_patch: np.ndarray   # (T,) uint32
    emission: np.ndarray    # (N, 3*sets) float32
    reflectance: np.ndarray  # (N, 3*sets) float32
    params: dict

    @property
    def num_patches(self):
        return self.patches.shape[0]


def _boxes(rng, target, ground, lots, lot):
    need = (target - ground * ground) // 2  # sum over boxes of (w + d) * h
    if (target - ground * ground) % 2 or need <= 0:
        raise ValueError("target - ground^2 must be a positive even number")
    nlots = lots * lots
    for _ in range(1000):
        boxes, acc = [], 0
        order = rng.permutation(nlots)
        ok = True
        for k, li in enumerate(order):
            left = need - acc
            if k == nlots - 1 or left <= 40:
                # Close with one box whose (w + d) * h equals what is left.
                sol = [(w, d, left // (w + d)) for w in range(1, lot - 1) for d in range(1, lot - 1)
                       if left % (w + d) == 0 and left // (w + d) >= 1]
                if not sol:
                    ok = False
                    break
                w, d, h = sol[int(rng.integers(len(sol)))]
                boxes.append((li, w, d, h))
                acc += (w + d) * h
                break
            w, d = int(rng.integers(2, lot - 1)), int(rng.integers(2, lot - 1))
            mean_h = max(1.0, (need - acc) / max(1, nlots - k) / (w + d))
            h = int(np.clip(round(rng.uniform(0.5, 1.5) * mean_h), 1, 60))
            if (w + d) * h > left:
                h = max(1, left // (w + d) - 1)
            boxes.append((li, w, d, h))
            acc += (w + d) * h
        if ok and acc == need:
            return boxes
    raise RuntimeError("could not hit the patch target")


def block_city(target=16384, ground=64, lots=8, sets=32, emitters=64, seed=11) -> PatchScene:
    rng = np.random.default_rng(seed)
    lot = ground // lots
    boxes = _boxes(rng, target, ground, lots, lot)
    covered = np.zeros((ground, ground), bool)
    quads = []  # (origin, U, V, n)
    placed = []
    for li, w, d, h in boxes:
        lx, lz = (li % lots) * lot, (li // lots) * lot
        ox = lx + int(rng.integers(1, lot - w)) if lot - w > 1 else lx
        oz = lz + int(rng.integers(1, lot - d)) if lot - d > 1 else lz
        covered[ox:ox + w, oz:oz + d] = True
        placed.append((ox, oz, w, d, h))
    X, Y, Z = np.eye(3)
    for x in range(ground):
        for z in range(ground):
            if not covered[x, z]:
                quads.append((np.array([x, 0.0, z]), Z, X, Y))
    for ox, oz, w, d, h in placed:
        for a in range(w):
            for b in range(d):
                quads.append((np.array([ox + a, h, oz + b], float), Z, X, Y))   # top
        for a in range(w):
            for y in range(h):
                quads.append((np.array([ox + a, y, oz], float), Y, X, -Z))     # -z side
                quads.append((np.array([ox + a, y, oz + d], float), X, Y, Z))  # +z side
        for b in range(d):
            for y in range(h):
                quads.append((np.array([ox, y, oz + b], float), Z, Y, -X))     # -x side
                quads.append((np.array([ox + w, y, oz + b], float), Y, Z, X))  # +x side
    n = len(quads)
    assert n == target, (n, target)
    patches = np.zeros((n, 12), np.float32)
    pos = np.zeros((4 * n, 3), np.float32)
    nrm = np.zeros((4 * n, 3), np.float32)
    idx = np.zeros(6 * n, np.uint32)
    for i, (o, U, V, nn) in enumerate(quads):
        patches[i] = np.concatenate([o, U, V, nn])
        pos[4 * i:4 * i + 4] = [o, o + U, o + U + V, o + V]
        nrm[4 * i:4 * i + 4] = nn
        idx[6 * i:6 * i + 6] = [4 * i, 4 * i + 1, 4 * i + 2, 4 * i, 4 * i + 2, 4 * i + 3]
    tri_patch = np.repeat(np.arange(n, dtype=np.uint32), 2)
    rho = rng.uniform(0.2, 0.8, (n, 3)).astype(np.float32)
    E = np.zeros((n, 3 * sets), np.float32)
    for s in range(sets):
        em = rng.choice(n, emitters, replace=False)
        E[em, 3 * s:3 * s + 3] = rng.uniform(1.0, 10.0, (emitters, 3)).astype(np.float32)
    R = np.tile(rho, (1, sets))
    return PatchScene(pos, nrm, idx, patches, tri_patch, E, R,
                      {"target": target, "ground": ground, "lots": lots, "boxes": len(boxes),
                       "sets": sets, "emitters": emitters, "seed": seed})


def build_gpu(relight, ps: PatchScene):
    """The patch geometry as one mesh (no lightmap) and its Radiosity object."""
    scene = relight.create_scene()
    scene.begin()
    scene.add_mesh(ps.positions, ps.normals, np.zeros((ps.positions.shape[0], 2), np.float32),
                   ps.indices)
    scene.end()
    return scene, relight.create_radiosity(scene, ps.patches, ps.tri_patch)
