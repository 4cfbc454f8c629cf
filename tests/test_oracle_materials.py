"""Pins for the oracle's per-triangle material (triangleMaterial / triangleColor with sampleColorNearest and
srgbToLinear: src/convert/sdf_baker.cpp:42-74, src/convert/voxelize.cpp:12-38).  Runs without a GPU.

Checked against closed forms (untextured factors, the fallbacks of materialForTriangle, texel choice under
wrapping and clamping, bytes 0 and 255, the two branches of the sRGB curve) and an FP64 numpy statement of
the sRGB curve.
"""
import numpy as np
import pytest


def _uv_tri(u, v):
    return np.array([[u, v]] * 3, np.float32)


def test_untextured_and_fallbacks(oracle):
    mats = [((0.2, 0.4, 0.6), 0.1, 0.9, -1), ((1.0, 0.5, 0.25), 0.7, 0.3, -1)]
    uvs = np.zeros((12, 2), np.float32)
    # triangle 0: material 1; triangle 1: material index 9 (out of range) -> material 0; triangles 2, 3: past
    # the per-triangle table -> material 0
    out = oracle.triangle_materials(4, uvs, [1, 9], mats, [])
    assert np.array_equal(out[0], np.float32([1.0, 0.5, 0.25, 0.7, 0.3]))
    for t in (1, 2, 3):
        assert np.array_equal(out[t], np.float32([0.2, 0.4, 0.6, 0.1, 0.9]))
    # no materials at all: MaterialAsset's defaults (colour 1, metallic 0, roughness 1)
    out = oracle.triangle_materials(2, uvs, [], [], [])
    assert np.array_equal(out, np.tile(np.float32([1, 1, 1, 0, 1]), (2, 1)))


def test_texel_choice_wraps_and_clamps(oracle):
    # a 4 x 1 linear texture whose texel i has red = 10 * (i + 1)
    tex = np.zeros((1, 4, 4), np.uint8)
    tex[0, :, 0] = [10, 20, 30, 40]
    mats = [((1, 1, 1), 0, 1, 0)]
    cases = {0.1: 10, 0.3: 20, 0.6: 30, 0.99: 40, 1.25: 20, -0.25: 40, 1.0: 10, 0.0: 10, -1.0: 10}
    for u, red in cases.items():
        out = oracle.triangle_materials(1, _uv_tri(u, 0.5), [0], mats, [(tex, False)])
        assert out[0, 0] == np.float32(red) / np.float32(255), (u, out[0, 0])


def test_centroid_uv(oracle):
    tex = np.zeros((1, 4, 4), np.uint8)
    tex[0, :, 1] = [0, 85, 170, 255]
    uvs = np.array([[0.0, 0.5], [0.5, 0.5], [0.4, 0.5]], np.float32)  # centroid u = 0.3 -> texel 1
    out = oracle.triangle_materials(1, uvs, [0], [((1, 1, 1), 0, 1, 0)], [(tex, False)])
    assert out[0, 1] == np.float32(85) / np.float32(255)
    # a triangle without three UVs keeps the factor
    out = oracle.triangle_materials(1, uvs[:2], [0], [((0.5, 0.5, 0.5), 0, 1, 0)], [(tex, False)])
    assert np.array_equal(out[0, :3], np.float32([0.5, 0.5, 0.5]))


def _srgb64(b):
    c = np.float64(np.float32(b) / np.float32(255))
    return c / 12.92 if c <= 0.04045 else ((c + 0.055) / 1.055) ** 2.4


def test_srgb_curve(oracle):
    tex = np.zeros((1, 256, 4), np.uint8)
    tex[0, :, 0] = np.arange(256)
    uvs = np.array([[(i + 0.5) / 256, 0.5] for i in range(256) for _ in range(3)], np.float32)
    out = oracle.triangle_materials(256, uvs, np.zeros(256, np.uint32), [((1, 1, 1), 0, 1, 0)], [(tex, True)])
    red = out[:, 0]
    assert red[0] == 0.0 and red[255] == 1.0
    # byte 10 is on the linear branch (10/255 <= 0.04045), byte 11 on the power branch
    assert red[10] == np.float32(np.float32(10) / np.float32(255)) / np.float32(12.92)
    ref = np.array([_srgb64(b) for b in range(256)])
    assert np.max(np.abs(red - ref) / np.maximum(ref, 1e-30)) < 4e-7
    assert np.all(np.diff(red) > 0)  # increasing


@pytest.mark.parametrize("seed", [1, 2])
def test_random_tables_against_numpy(oracle, seed):
    rng = np.random.default_rng(seed)
    ntri = 500
    texs = [(rng.integers(0, 256, (h, w, 4), dtype=np.uint8), bool(s)) for h, w, s in ((3, 5, 1), (7, 2, 0), (1, 1, 1))]
    mats = [(tuple(rng.uniform(0, 2, 3)), float(rng.uniform()), float(rng.uniform()), int(t)) for t in (-1, 0, 1, 2, 7)]
    uvs = rng.uniform(-3, 3, (3 * ntri - 2, 2)).astype(np.float32)  # the last triangle has no full set of UVs
    tm = rng.integers(0, 6, ntri - 10).astype(np.uint32)
    out = oracle.triangle_materials(ntri, uvs, tm, mats, texs)
    for t in range(ntri):
        mi = int(tm[t]) if t < len(tm) else 0
        base, met, rough, ti = mats[mi if mi < len(mats) else 0]
        col = np.float32(base)
        if 0 <= ti < len(texs) and 3 * t + 2 < len(uvs):
            px, srgb = texs[ti]
            c = ((uvs[3 * t] + uvs[3 * t + 1]) + uvs[3 * t + 2]) / np.float32(3)
            w = c - np.floor(c)
            ix = min(max(int(w[0] * np.float32(px.shape[1])), 0), px.shape[1] - 1)
            iy = min(max(int(w[1] * np.float32(px.shape[0])), 0), px.shape[0] - 1)
            texel = px[iy, ix, :3]
            lin = np.array([_srgb64(b) if srgb else np.float32(b) / np.float32(255) for b in texel])
            expect = lin * np.float64(col)
            assert np.allclose(out[t, :3], expect, rtol=1e-6, atol=0), t
        else:
            assert np.array_equal(out[t, :3], col), t
        assert out[t, 3] == np.float32(met) and out[t, 4] == np.float32(rough)
