"""LBVH refit at the sizes where its 256-leaf blocks begin and end: closed soups of 4 .. 1,028 triangles
(seeded random tetrahedra, 4 triangles each) baked through the C ABI equal the oracle: distance magnitudes
bit for bit, signs away from the surface, occupancy bit for bit; with and without collapsed leaf blocks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [4, 8, 252, 256, 260, 508, 512, 516, 1024, 1028]


def _tetrahedra(n, seed):
    """n/4 tetrahedra of edge 0.04 .. 0.12 at random places and orientations, each wound consistently."""
    rng = np.random.default_rng(seed)
    base = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], np.float64)
    faces = [(0, 1, 2), (0, 3, 1), (0, 2, 3), (1, 3, 2)]
    tris = []
    for _ in range(n // 4):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        v = base @ q.T * rng.uniform(0.015, 0.045) + rng.uniform(-0.5, 0.5, 3)
        tris += [[v[a], v[b], v[c]] for a, b, c in faces]
    return np.array(tris, np.float32)


@pytest.mark.parametrize("block", [1, 8])
def test_block_boundary_sizes(ctx, oracle, meshgen, block):
    from italy_bake import bake

    g = bake.dense_grid(-0.6, 0.6, 20)
    ctx.set_option(bake.OPT_LEAF_BLOCK, block)
    try:
        for n in SIZES:
            xyz = _tetrahedra(n, 100 + n)
            dist, occ = ctx.bake_dense(xyz, g, g)
            ref, _ = oracle.bake_dense(xyz, list(g.origin), g.h, g.nx, g.ny, g.nz)
            assert np.array_equal(np.abs(dist).view(np.uint32), np.abs(ref).view(np.uint32)), n
            far = np.abs(ref) > 1.2e-4
            assert int(((dist < 0) != (ref < 0))[far].sum()) == 0, n
            assert np.array_equal(occ, oracle.occupancy(xyz, list(g.origin), g.h, g.nx, g.ny, g.nz)), n
    finally:
        ctx.set_option(bake.OPT_LEAF_BLOCK, bake.DEFAULT_OPTIONS[bake.OPT_LEAF_BLOCK])
