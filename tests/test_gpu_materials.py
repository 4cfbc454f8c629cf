"""Per-triangle materials on the device (ib_triangle_materials: triangleMaterial / triangleColor with
sampleColorNearest and srgbToLinear, src/convert/sdf_baker.cpp:42-74, src/convert/voxelize.cpp:12-38) equal
the oracle's bit for bit: random tables, sRGB and linear textures, wrapped UVs and every fallback."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tables(rng, ntri, nmat, ntex):
    texs = []
    for i in range(ntex):
        h, w = int(rng.integers(1, 9)), int(rng.integers(1, 9))
        texs.append((rng.integers(0, 256, (h, w, 4), dtype=np.uint8), bool(i % 2 == 0)))
    mats = []
    for i in range(nmat):
        tex = int(rng.integers(-2, ntex + 2))  # includes indices past the texture table: the factor
        mats.append((tuple(float(x) for x in rng.uniform(0, 2, 3)), float(rng.uniform()), float(rng.uniform()), tex))
    uvs = (rng.uniform(-4, 4, (3 * ntri, 2)) * rng.choice([1e-3, 1.0, 37.0], (3 * ntri, 1))).astype(np.float32)
    tm = rng.integers(0, nmat + 3, max(ntri - 7, 0)).astype(np.uint32)  # some past the table, the tail unnamed
    return uvs, tm, mats, texs


@pytest.mark.parametrize("seed,ntri,nmat,ntex", [(1, 1, 1, 1), (2, 500, 5, 3), (3, 40000, 17, 6), (4, 9, 3, 0)])
def test_random_tables_equal_oracle(ctx, oracle, seed, ntri, nmat, ntex):
    rng = np.random.default_rng(seed)
    uvs, tm, mats, texs = _tables(rng, ntri, nmat, ntex)
    got = ctx.triangle_materials(ntri, uvs, tm, mats, texs)
    want = oracle.triangle_materials(ntri, uvs, tm, mats, texs)
    assert got.shape == (ntri, 5)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_every_byte_through_both_curves(ctx, oracle):
    """Each byte value through the sRGB table and the linear path, one triangle per texel."""
    px = np.zeros((1, 256, 4), np.uint8)
    px[0, :, 0] = np.arange(256)
    px[0, :, 1] = 255 - np.arange(256)
    px[0, :, 2] = (np.arange(256) * 7) % 256
    uvs = np.repeat(np.array([[(i + 0.5) / 256, 0.5] for i in range(256)], np.float32), 3, axis=0)
    for srgb in (True, False):
        mats = [((0.5, 1.0, 1.75), 0.25, 0.5, 0)]
        got = ctx.triangle_materials(256, uvs, np.zeros(256, np.uint32), mats, [(px, srgb)])
        want = oracle.triangle_materials(256, uvs, np.zeros(256, np.uint32), mats, [(px, srgb)])
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), srgb


def test_fallbacks(ctx, oracle):
    uvs = np.zeros((4, 2), np.float32)  # only the first triangle has three UVs
    tex = [(np.full((2, 2, 4), 200, np.uint8), True)]
    for mats, tm in (([], []), ([((0.1, 0.2, 0.3), 0.4, 0.6, 0)], [5, 0]), ([((1, 1, 1), 0, 1, 0)], [0, 0, 0])):
        got = ctx.triangle_materials(3, uvs, tm, mats, tex)
        want = oracle.triangle_materials(3, uvs, tm, mats, tex)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (mats, tm)
    assert ctx.triangle_materials(0, uvs, [], [], []).shape == (0, 5)
