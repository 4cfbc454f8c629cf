"""The rare sign paths of the bake, forced and checked voxel by voxel against the oracle:

- the parity tie-break (sdf_bake.cu:53-69): voxel centres on the plane of a cube's missing face, where the
  winding number of the open cube is exactly 1/2;
- voxel centres lying on a hole cap, where the closure identity does not hold and the FP64 recheck uses the
  mesh's own winding number;
- the column-overflow recount: more than 32 crossings on a +z column (40 nested cubes).
"""
import numpy as np
import pytest

import shapes

pytestmark = pytest.mark.gpu


def _centres(g):
    from italy_bake import bake  # noqa: F401

    x = np.float32(g.origin[0]) + (np.arange(g.nx, dtype=np.float32) + np.float32(0.5)) * np.float32(g.h)
    y = np.float32(g.origin[1]) + (np.arange(g.ny, dtype=np.float32) + np.float32(0.5)) * np.float32(g.h)
    z = np.float32(g.origin[2]) + (np.arange(g.nz, dtype=np.float32) + np.float32(0.5)) * np.float32(g.h)
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    return np.stack([xx, yy, zz], -1).reshape(-1, 3).astype(np.float32)


def _check_against_gwn_oracle(ctx, oracle, xyz, g, extent):
    dist, _ = ctx.bake_dense(xyz, g, None)
    st = ctx.stats()
    ref, tie = oracle.signed_points_gwn(xyz, _centres(g), threads=0)
    dist = dist.reshape(-1)
    assert np.array_equal(np.abs(dist).view(np.uint32), np.abs(ref).view(np.uint32)), "distance bits"
    far = np.abs(ref) > 1e-4 * extent
    bad = np.nonzero((dist < 0)[far] != (ref < 0)[far])[0]
    assert bad.size == 0, f"{bad.size} sign mismatches, first voxel {np.nonzero(far)[0][bad[:5]]}"
    return st, tie, far


def test_parity_tie_break_on_an_open_face(ctx, oracle):
    """Cube [-0.5, 0.5]^3 without its +z face on a grid whose z = 0.5 layer of centres lies in the hole's
    plane: those voxels are ties, decided by the 5-ray vote, on the device and in the oracle alike."""
    from italy_bake import Grid

    box = shapes.box_faces(0.5, 4, skip=("+z",))
    g = Grid.make((-0.78125, -0.78125, -0.78125), 0.0625, 24, 24, 24)  # centres -0.75 + i / 16: z = 0.5 at i = 20
    st, tie, far = _check_against_gwn_oracle(ctx, oracle, box, g, 1.5)
    hole = np.zeros((24, 24, 24), bool)
    hole[20, 5:20, 5:20] = True  # the 15 x 15 centres strictly inside the hole
    assert np.array_equal(tie.reshape(24, 24, 24)[20][hole[20]], np.ones(225, bool))
    assert st["ties"] >= 225, st
    assert st["rechecked"] >= st["ties"]
    assert st["boundary_edges"] == 16


def test_voxels_on_hole_caps(ctx, oracle):
    """Open boxes whose holes are crossed by planes of voxel centres in several places: two faces missing,
    a box off centre, and a grid whose centres hit the cap fans' spokes and apices; every sign equals the
    FP64 winding number's (or its parity tie-break)."""
    from italy_bake import Grid

    a = shapes.box_faces(0.25, 2, skip=("+z", "-x"), centre=(-0.25, 0.0, 0.0))
    b = shapes.box_faces(0.1875, 3, skip=("-y",), centre=(0.3125, 0.125, -0.125), inward=True)
    mesh = np.concatenate([a, b])
    # centres -0.75 + k / 16: every face plane of both boxes, the holes' fan apices and spokes among them
    g = Grid.make((-0.78125, -0.78125, -0.78125), 0.0625, 24, 24, 24)
    st, tie, far = _check_against_gwn_oracle(ctx, oracle, mesh, g, 1.5)
    assert st["rechecked"] > 0 and st["boundary_edges"] > 0


def test_column_overflow_recount(ctx, oracle):
    """40 nested cubes, wound alternately outward and inward: the columns through the middle cross 80
    faces, more than the 32 a column stores, so their voxels are recounted one by one; every voxel equals
    the oracle's dense bake (exact crossings of the closed mesh)."""
    from italy_bake import bake

    mesh = shapes.nested_boxes(40, 0.1, 0.015)
    g = bake.dense_grid(-0.8, 0.8, 64)
    half = 0.1 + 0.015 * np.arange(40)
    c = -0.8 + (np.arange(64) + 0.5) * 0.025
    assert np.min(np.abs(c[:, None] - half[None, :])) > 1e-3 and np.min(np.abs(c[:, None] + half[None, :])) > 1e-3
    dist, _ = ctx.bake_dense(mesh, g, None)
    st = ctx.stats()
    ref, _ = oracle.bake_dense(mesh, list(g.origin), g.h, g.nx, g.ny, g.nz)
    assert np.array_equal(dist.view(np.uint32), ref.view(np.uint32))
    assert st["column_overflows"] > 0, st
    assert st["stack_overflows"] == 0
    inside = (ref < 0).reshape(64, 64, 64)
    assert inside[32, 32, 32 + 2] == False and inside.any()  # noqa: E712  the innermost layer is outside


def test_tree_depth_stays_within_the_key_bound(ctx, meshgen):
    """Duplicate centroids (every triangle of a soup repeated at one point) make the Karras tree split on the
    sorted positions alone: its depth is at most 30 + ceil(log2 n), and the bake runs with no stack overflow."""
    from italy_bake import bake

    n = 1 << 12
    tri = np.array([[[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.1, 0.0]]], np.float32)
    soup = np.concatenate([np.repeat(tri, n, 0), meshgen.icosphere(3, 0.5).xyz])
    g = bake.dense_grid(-0.8, 0.8, 32)
    dist, _ = ctx.bake_dense(soup, g, None)
    st = ctx.stats()
    assert st["bvh_depth"] <= 30 + int(np.ceil(np.log2(len(soup)))), st
    assert st["stack_overflows"] == 0
    alone, _ = ctx.bake_dense(soup[n:], g, None)  # the repeated triangle only adds its own distances
    single, _ = ctx.bake_dense(tri, g, None)
    assert np.array_equal(np.abs(dist), np.minimum(np.abs(alone), np.abs(single)))
