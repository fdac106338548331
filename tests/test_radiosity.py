"""Radiosity: form factors by hemisphere rays and the BF16-MFMA gather solve.

CPU: the oracle's form-factor estimator against closed forms (closed enclosure row sums; Howell
catalogue factors for parallel and perpendicular unit plates) and reciprocity (A_i F_ij = A_j F_ji
for unequal plates).
GPU: form-factor counts bit-exact against the oracle; the solve against an FP64 solve of the same
system (numpy), with the FP64 residual reported.
"""
import math

import numpy as np
import pytest

import oracle as O
from relight_amd import patches as PT


def _plates(quads):
    """Scene of quads (origin, U, V, n), one patch each."""
    ps = []
    pos, nrm, idx = [], [], []
    for i, (o, U, V, n) in enumerate(quads):
        o, U, V, n = (np.asarray(x, np.float32) for x in (o, U, V, n))
        ps.append(np.concatenate([o, U, V, n]))
        pos += [o, o + U, o + U + V, o + V]
        nrm += [n] * 4
        idx += [4 * i, 4 * i + 1, 4 * i + 2, 4 * i, 4 * i + 2, 4 * i + 3]
    patches = np.asarray(ps, np.float32)
    S = O.OracleScene()
    S.add_mesh(np.asarray(pos, np.float32), np.asarray(nrm, np.float32), np.zeros((len(pos), 2), np.float32),
               np.asarray(idx, np.uint32))
    S.build()
    return S, patches, np.repeat(np.arange(len(quads), dtype=np.uint32), 2)


def howell_parallel(a, b, c):
    X, Y = a / c, b / c
    return (2 / (math.pi * X * Y)) * (
        math.log(math.sqrt((1 + X * X) * (1 + Y * Y) / (1 + X * X + Y * Y)))
        + X * math.sqrt(1 + Y * Y) * math.atan(X / math.sqrt(1 + Y * Y))
        + Y * math.sqrt(1 + X * X) * math.atan(Y / math.sqrt(1 + X * X))
        - X * math.atan(X) - Y * math.atan(Y))


def test_form_factor_parallel_plates_closed_form():
    # Unit squares at distance 1, facing each other: F = 0.19982 (Howell C-11).
    S, P, tp = _plates([((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)),
                        ((0, 0, 1), (0, 1, 0), (1, 0, 0), (0, 0, -1))])
    rays = 60000
    F = S.form_factors(P, tp, rays, 3, 100.0).astype(np.float64) / rays
    expect = howell_parallel(1.0, 1.0, 1.0)
    assert abs(expect - 0.19982) < 1e-4
    sd = math.sqrt(expect * (1 - expect) / rays)
    assert abs(F[0, 1] - expect) < 5 * sd and abs(F[1, 0] - expect) < 5 * sd
    assert F[0, 0] == 0 and F[1, 1] == 0


def test_form_factor_perpendicular_plates_closed_form():
    # Unit squares sharing an edge at 90 degrees: F = 0.20004 (Howell C-14).
    S, P, tp = _plates([((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)),
                        ((0, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 0))])
    rays = 60000
    F = S.form_factors(P, tp, rays, 5, 100.0).astype(np.float64) / rays
    sd = math.sqrt(0.2 * 0.8 / rays)
    assert abs(F[0, 1] - 0.20004) < 5 * sd and abs(F[1, 0] - 0.20004) < 5 * sd


def test_form_factor_reciprocity_unequal_plates():
    # A_i F_ij = A_j F_ji: a unit square under a parallel 1 x 2 rectangle 0.7 m above (offset
    # towards -x) and a 1 x 0.7 wall standing on the square's +x edge; every patch sees only the
    # others' front faces. Each estimate is a binomial fraction of the rays from one patch, so the
    # check is within five standard deviations of the difference.
    S, P, tp = _plates([((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)),
                        ((-1.5, 0.3, 0.7), (0, 1, 0), (2, 0, 0), (0, 0, -1)),
                        ((1, 0, 0), (0, 1, 0), (0, 0, 0.7), (-1, 0, 0))])
    areas = np.array([1.0, 2.0, 0.7])
    rays = 80000
    F = S.form_factors(P, tp, rays, 11, 100.0).astype(np.float64) / rays
    for i, j in [(0, 1), (0, 2), (1, 2)]:
        assert F[i, j] > 0.02 and F[j, i] > 0.02
        lhs, rhs = areas[i] * F[i, j], areas[j] * F[j, i]
        sd = math.sqrt(areas[i] ** 2 * F[i, j] * (1 - F[i, j]) / rays + areas[j] ** 2 * F[j, i] * (1 - F[j, i]) / rays)
        assert abs(lhs - rhs) < 5 * sd, (i, j, lhs, rhs, sd)


def test_form_factor_closed_enclosure_rows_sum_to_one():
    # Inside a closed cube of 6 patches every ray hits a wall, except rays that start within the
    # 0.01 m tnear band of an edge and meet the adjacent wall before tnear; with a 50 m cube that
    # band is 0.08% of each wall. By symmetry the factors are the plate values.
    L = 50.0
    q = [((0, 0, 0), (0, 0, L), (0, L, 0), (1, 0, 0)), ((L, 0, 0), (0, L, 0), (0, 0, L), (-1, 0, 0)),
         ((0, 0, 0), (L, 0, 0), (0, 0, L), (0, 1, 0)), ((0, L, 0), (0, 0, L), (L, 0, 0), (0, -1, 0)),
         ((0, 0, 0), (0, L, 0), (L, 0, 0), (0, 0, 1)), ((0, 0, L), (L, 0, 0), (0, L, 0), (0, 0, -1))]
    S, P, tp = _plates(q)
    rays = 20000
    C = S.form_factors(P, tp, rays, 7, 1000.0)
    assert (C.sum(1) >= rays - 10).all() and (C.sum(1) <= rays).all()
    F = C / rays
    assert abs(F[0, 1] - 0.19982) < 0.015 and abs(F[0, 2] - 0.20004) < 0.015


def test_block_city_patch_count_and_layout():
    ps = PT.block_city(target=16384)
    assert ps.num_patches == 16384
    assert ps.tri_patch.size == 2 * 16384 and ps.indices.size == 6 * 16384
    assert ps.emission.shape == (16384, 96) and ps.reflectance.shape == (16384, 96)
    r = ps.reflectance
    assert r.min() >= 0.2 and r.max() <= 0.8
    small = PT.block_city(target=1024, ground=16, lots=2, sets=4, emitters=8, seed=2)
    assert small.num_patches == 1024


def _fp64_solve(C, rays, E, R):
    F = C.astype(np.float64) / rays
    n, m = E.shape
    B = np.zeros((n, m))
    for ch in range(m):
        A = np.eye(n) - R[:, ch:ch + 1].astype(np.float64) * F
        B[:, ch] = np.linalg.solve(A, E[:, ch].astype(np.float64))
    return B


@pytest.mark.gpu
def test_gpu_form_factors_and_solve():
    from relight_amd import Relight
    print("[gpu] radiosity small block city", flush=True)
    ps = PT.block_city(target=1024, ground=16, lots=2, sets=4, emitters=8, seed=2)
    r = Relight()
    scene, rad = PT.build_gpu(r, ps)
    rad.build_form_factors(256, seed=9, max_distance=200.0)
    st = r.stats()
    C = rad.form_factor_counts()
    S = O.OracleScene()
    S.add_mesh(ps.positions, ps.normals, np.zeros((ps.positions.shape[0], 2), np.float32), ps.indices)
    S.build()
    Co = S.form_factors(ps.patches, ps.tri_patch, 256, 9, 200.0)
    assert st["rays"] == 1024 * 256
    assert np.array_equal(C, Co), f"{(C != Co).sum()} counts differ"
    B, info = rad.solve(ps.emission, ps.reflectance, tolerance=1e-5, max_outer=30, inner_sweeps=12)
    print(f"[gpu] solve: {info}", flush=True)
    ref = _fp64_solve(Co, 256, ps.emission, ps.reflectance)
    err = np.linalg.norm(B - ref) / np.linalg.norm(ref)
    F = Co.astype(np.float64) / 256
    res = np.linalg.norm(ps.emission + ps.reflectance * (F @ B) - B) / np.linalg.norm(ps.emission)
    print(f"[gpu] relative error vs FP64 solve {err:.3e}, FP64 residual {res:.3e}")
    assert info["residual"] <= 1e-5
    assert res <= 2e-5
    assert err <= 1e-4


@pytest.mark.gpu
def test_gpu_staged_gemm_solve_4096_patches():
    """4,096 patches: the split-K length (npad / 16 = 256) takes the LDS-staged MFMA kernels for
    both the Jacobi sweep and the refinement residual products (the 1,024-patch case above runs the
    unstaged ones). Form factors bit-exact vs the oracle, solve within 1e-4 of an FP64 LU solve."""
    from relight_amd import Relight
    print("[gpu] radiosity 4096-patch block city (staged GEMM)", flush=True)
    ps = PT.block_city(target=4096, ground=32, lots=4, sets=2, emitters=16, seed=4)
    r = Relight()
    scene, rad = PT.build_gpu(r, ps)
    rad.build_form_factors(128, seed=3, max_distance=200.0)
    C = rad.form_factor_counts()
    S = O.OracleScene()
    S.add_mesh(ps.positions, ps.normals, np.zeros((ps.positions.shape[0], 2), np.float32), ps.indices)
    S.build()
    Co = S.form_factors(ps.patches, ps.tri_patch, 128, 3, 200.0)
    assert np.array_equal(C, Co), f"{(C != Co).sum()} counts differ"
    B, info = rad.solve(ps.emission, ps.reflectance, tolerance=1e-5, max_outer=30, inner_sweeps=12)
    st = r.stats()
    print(f"[gpu] solve: {info}, kernels {st['kernels']}", flush=True)
    ref = _fp64_solve(Co, 128, ps.emission, ps.reflectance)
    err = np.linalg.norm(B - ref) / np.linalg.norm(ref)
    print(f"[gpu] relative error vs FP64 solve {err:.3e}")
    assert info["residual"] <= 1e-5
    assert err <= 1e-4
