"""Narrow-band bakes (8^3 FP32 bricks, +-band voxels) against the oracle and the dense bake."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pool_voxels(index, pool):
    """(z, y, x) voxel coordinates and values of every pool entry."""
    slots = np.argwhere(index >= 0)  # (bz, by, bx)
    order = index[index >= 0]
    coords = np.zeros((len(order), 8, 8, 8, 3), np.int64)
    lz, ly, lx = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    for (bz, by, bx), slot in zip(slots, order):
        coords[slot, ..., 0] = bz * 8 + lz
        coords[slot, ..., 1] = by * 8 + ly
        coords[slot, ..., 2] = bx * 8 + lx
    return coords.reshape(-1, 3), pool.reshape(-1)


def test_band_closed_mesh_equals_dense(ctx, oracle, meshgen):
    from italy_bake import bake

    m = meshgen.displaced(5, 31)
    g = bake.dense_grid(-0.8, 0.8, 96)
    index, pool = ctx.bake_band(m.xyz, g, 4)
    keep = oracle.band_bricks(m.xyz, list(g.origin), g.h, g.nx, g.ny, g.nz, 4)
    assert np.array_equal(index.reshape(-1) >= 0, keep)
    assert sorted(index[index >= 0].tolist()) == list(range(int(keep.sum())))
    dense, _ = ctx.bake_dense(m.xyz, g, None)
    zyx, vals = _pool_voxels(index, pool)
    assert np.array_equal(vals.view(np.uint32), dense[zyx[:, 0], zyx[:, 1], zyx[:, 2]].view(np.uint32))
    # every voxel with |d| <= 4h lies in a kept brick
    band = np.abs(dense) <= 4 * g.h
    covered = np.zeros_like(band)
    covered[zyx[:, 0], zyx[:, 1], zyx[:, 2]] = True
    assert not (band & ~covered).any()


def test_band_open_mesh_against_oracle(ctx, oracle, meshgen):
    """Four seeded tori/boxes with 2% of triangles removed: brick set, FP32 distances bit-exact,
    sign = |GWN_64| > 0.5 on sampled band voxels away from ties."""
    from italy_bake import bake

    m = meshgen.stress(seed=11, ncomp=4, remove_frac=0.02)
    g = bake.dense_grid(-1.0, 1.0, 96)
    index, pool = ctx.bake_band(m.xyz, g, 4)
    st = ctx.stats()
    assert st["boundary_edges"] > 0
    keep = oracle.band_bricks(m.xyz, list(g.origin), g.h, g.nx, g.ny, g.nz, 4)
    assert np.array_equal(index.reshape(-1) >= 0, keep)
    zyx, vals = _pool_voxels(index, pool)
    pts = oracle.cell_centres(list(g.origin), g.h, g.nx, g.ny, g.nz)[zyx[:, 0], zyx[:, 1], zyx[:, 2]]
    ud = oracle.udf_points(m.xyz, pts)
    assert np.array_equal(np.abs(vals), ud)
    rng = np.random.default_rng(0)
    pick = rng.choice(len(pts), size=min(2500, len(pts)), replace=False)
    d64, w = oracle.sdf_f64_points(m.xyz, pts[pick].astype(np.float64))
    extent = 2.0
    decided = (np.abs(d64) > 1e-4 * extent) & (np.abs(np.abs(w) - 0.5) > 1e-6)
    assert np.array_equal((vals[pick] < 0)[decided], (np.abs(w) > 0.5)[decided])
    assert np.abs(vals[pick] - d64).max() <= 1e-5 * extent


def test_fp64_recheck_through_the_cap_tree_equals_the_full_sum(ctx, meshgen):
    """The FP64 recheck first sums the hole caps through their tree with a rigorous far-field bound and
    decides when |w| stays clear of 1/2 and of the tie band by more than that bound; otherwise, or with
    IB_OPT_RESOLVE_FULL set, it sums every cap fan.  Both give the same bits on every band voxel of the
    513,803-triangle precision-stress union and on a dense open-mesh bake, and the tree decides most
    rechecks on its own.  With IB_OPT_SHARED_CHIP set a single bake's rechecks take one voxel per wave
    instead of one per block: the same bits again."""
    import shapes
    from italy_bake import Grid, bake

    m, c = meshgen.workload_stress()
    g = bake.dense_grid(c["lo"], c["hi"], c["res"])
    box = shapes.box_faces(0.5, 6, skip=("+z", "-x"))
    gd = Grid.make((-0.7, -0.66, -0.69), 0.021, 64, 64, 64)
    runs = []
    try:
        for full, shared in ((0, 0), (1, 0), (0, 1)):
            ctx.set_option(bake.OPT_RESOLVE_FULL, full)
            ctx.set_option(bake.OPT_SHARED_CHIP, shared)
            index, pool = ctx.bake_band(m.xyz, g, c["band"])
            st_band = ctx.stats()
            dense, _ = ctx.bake_dense(box, gd, None)
            st_dense = ctx.stats()
            runs.append((index, pool, st_band, dense, st_dense))
    finally:
        ctx.set_option(bake.OPT_RESOLVE_FULL, 0)
        ctx.set_option(bake.OPT_SHARED_CHIP, 0)
    (i0, p0, b0, d0, s0), (i1, p1, b1, d1, s1), (i2, p2, b2, d2, s2) = runs
    for i, p, d in ((i1, p1, d1), (i2, p2, d2)):
        assert np.array_equal(i0, i)
        assert np.array_equal(p0.view(np.uint32), p.view(np.uint32))
        assert np.array_equal(d0.view(np.uint32), d.view(np.uint32))
    for tree, wave in ((b0, b2), (s0, s2)):
        assert tree["rechecked"] == wave["rechecked"] and tree["rechecked_full"] == wave["rechecked_full"]
    for tree, full in ((b0, b1), (s0, s1)):
        assert tree["rechecked"] == full["rechecked"] > 0
        assert full["rechecked_full"] == full["rechecked"]
        assert tree["ties"] == full["ties"] and tree["on_cap"] == full["on_cap"]
    assert b0["rechecked"] > 1000 and b0["rechecked_full"] < b0["rechecked"] // 4, b0
