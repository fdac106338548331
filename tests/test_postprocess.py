"""Lightmap post-process and output formats: Lightmap::expand / blur / toDoubleLdr / toRgbmLdr /
save (Lightmap.cpp:167-331). The oracle is pinned by hand-computed cases and an independent
numpy restatement; the GPU path is compared with it bit for bit."""
import os

import numpy as np
import pytest

import oracle as O


def _numpy_blur(valid, rgb):
    """Independent restatement: the same raster-order in-place recurrence, in numpy float32."""
    out = rgb.astype(np.float32).copy()
    h, w = valid.shape
    for y in range(1, h - 1):
        for x in range(1, w - 1):
            acc = np.zeros(3, np.float32)
            n = 0
            for j in range(y - 1, y + 2):
                for i in range(x - 1, x + 2):
                    if valid[j, i]:
                        acc = acc + out[j, i]
                        n += 1
            if n:
                out[y, x] = acc * (np.float32(1.0) / np.float32(n))
    return out


def test_expand_hand_case():
    valid = np.zeros((4, 5), np.uint8)
    valid[1, 1] = valid[2, 3] = 1
    rgb = np.zeros((4, 5, 3), np.float32)
    rgb[1, 1] = (1, 2, 3)
    rgb[2, 3] = (4, 5, 6)
    out = O.expand(valid, rgb)
    # (1,1) writes (0,0),(2,0),(0,2),(2,2); (3,2) writes (2,1),(2,3),(4,1),(4,3): (x, y) pairs.
    for x, y in [(0, 0), (2, 0), (0, 2), (2, 2)]:
        assert tuple(out[y, x]) == (1, 2, 3)
    for x, y in [(2, 1), (2, 3), (4, 1), (4, 3)]:
        assert tuple(out[y, x]) == (4, 5, 6)
    assert tuple(out[1, 1]) == (1, 2, 3) and tuple(out[3, 0]) == (0, 0, 0)
    # A texel written by two valid texels keeps the later one in raster order.
    valid = np.zeros((4, 4), np.uint8)
    valid[1, 1] = valid[1, 2] = 1
    rgb = np.zeros((4, 4, 3), np.float32)
    rgb[1, 1], rgb[1, 2] = (1, 1, 1), (2, 2, 2)
    out = O.expand(valid, rgb)
    assert out[2, 1, 0] == 2.0 and out[0, 3, 0] == 2.0 and out[0, 0, 0] == 1.0


def test_blur_constant_and_recurrence():
    rng = np.random.default_rng(0)
    valid = (rng.random((9, 11)) < 0.7).astype(np.uint8)
    const = np.full((9, 11, 3), 0.375, np.float32)
    out = O.blur(valid, const)
    # A constant field stays constant up to the rounding of sum * (1 / count) (one ulp per step).
    assert np.allclose(out, 0.375, rtol=4e-7, atol=0)
    rgb = rng.random((9, 11, 3)).astype(np.float32)
    assert np.array_equal(O.blur(valid, rgb), _numpy_blur(valid, rgb))
    # The recurrence is not a plain 3x3 mean of the input (earlier outputs feed later texels).
    plain = rgb.copy()
    y, x = 4, 5
    win = valid[y - 1:y + 2, x - 1:x + 2].astype(bool)
    assert not np.allclose(O.blur(valid, rgb)[y, x], rgb[y - 1:y + 2, x - 1:x + 2][win].mean(0), atol=0)


def test_encoders_known_values():
    rgb = np.array([[[1.0, 0.5, 0.0], [4.0, -1.0, np.nan], [0.25, 0.25, 0.25]]], np.float32)
    d = O.encode(rgb, 1)
    assert d[0, 0].tolist() == [127, 63, 0]      # 1*0.5*255 = 127.5 -> 127
    assert d[0, 1].tolist() == [255, 0, 0]       # saturates; undefined casts give 0
    assert d[0, 2].tolist() == [31, 31, 31]      # 31.875
    m = O.encode(rgb, 2)
    # m = 255 / max; channel = c / m * 255; m byte = min(255, m * 255).
    assert m[0, 0].tolist() == [1, 0, 0, 255]
    assert m[0, 2].tolist() == [0, 0, 0, 255]


def test_tga_layout():
    px = np.arange(2 * 3 * 4, dtype=np.uint8).reshape(2, 3, 4)
    b = O.tga_bytes(px)
    assert b[:18] == bytes([0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 3, 0, 2, 0, 32, 0])
    assert b[18:22] == bytes([2, 1, 0, 3])
    assert len(b) == 18 + 2 * 3 * 4


@pytest.mark.gpu
def test_gpu_postprocess_and_formats_match_oracle(tmp_path):
    from relight_amd import (RL_FORMAT_RAW_HDR, RL_FORMAT_TGA_DOUBLE_LDR, RL_FORMAT_TGA_RGBM, Relight,
                             scenes)
    print("[gpu] expand / blur / encode / save on the box room", flush=True)
    r = Relight()
    spec = scenes.box_room()
    scene, lms, ids = scenes.build_gpu(r, spec)
    r.bake_direct_light(scene, ids[0])
    r.bake_ambient_occlusion(scene, ids[0], spec.ao)
    lm = lms[0]
    W, H = spec.lightmaps[0]
    valid = np.zeros(W * H, np.uint8)
    texel, _, _ = lm.lumels()
    valid[texel] = 1
    valid = valid.reshape(H, W)
    base = lm.to_hdr().reshape(H, W, 3)
    lm.expand()
    e = O.expand(valid, base)
    assert np.array_equal(lm.to_hdr().reshape(H, W, 3), e)
    lm.blur()
    b = O.blur(valid, e)
    got = lm.to_hdr().reshape(H, W, 3)
    assert np.array_equal(got, b), f"{(got != b).sum()} values differ"
    for fmt in (RL_FORMAT_TGA_DOUBLE_LDR, RL_FORMAT_TGA_RGBM):
        assert np.array_equal(lm.encode(fmt), O.encode(b, fmt))
        p = tmp_path / f"lm{fmt}.tga"
        lm.save(p, fmt)
        assert p.read_bytes() == O.tga_bytes(O.encode(b, fmt))
    p = tmp_path / "lm.raw"
    lm.save(p, RL_FORMAT_RAW_HDR)
    raw = p.read_bytes()
    assert np.frombuffer(raw[:8], np.int32).tolist() == [W, H]
    assert np.array_equal(np.frombuffer(raw[8:], np.float32).reshape(H, W, 3), b)
    # A bake after post-processing starts from the post-processed colours (valid texels re-uploaded).
    r.bake_direct_light(scene, ids[0])
    assert np.isfinite(lm.to_hdr()).all()
    r.close()
