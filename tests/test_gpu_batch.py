"""Batched bake (one persistent query launch over many meshes) == each mesh baked alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _meshes(meshgen):
    return [
        (meshgen.icosphere(3, 0.5, 1e-3, 3), 24),
        (meshgen.displaced(4, 9), 40),
        (meshgen.uv_sphere(0.6, 32, 16), 33),           # slivers, degenerate poles, hairline holes
        (meshgen.stress(seed=8, ncomp=2, remove_frac=0.02), 28),  # open: hole caps
        (meshgen.decimated(900, 12), 17),
        (meshgen.Mesh(np.zeros((0, 3, 3), np.float32), [0, 0, 0], [0, 0, 0]), 8),  # empty
        (meshgen.Mesh(np.array([[[0.1, 0.0, 0.0], [0.0, 0.3, 0.0], [0.0, 0.0, 0.2]]], np.float32),
                      [0, 0, 0], [0.3, 0.3, 0.3]), 9),  # one triangle
        (meshgen.displaced(5, 4), 64),
    ]


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4])
def test_batch_equals_alone(ctx, meshgen, shape):
    from italy_bake import Grid, Job, bake

    items = _meshes(meshgen)
    ctx.set_option(bake.OPT_BRICK_SHAPE, shape)
    try:
        jobs, bufs, grids = [], [], []
        for m, res in items:
            g = bake.dense_grid(-0.8, 0.8, res)
            og = bake.dense_grid(-0.8, 0.8, res + 3)
            d_xyz = ctx.alloc(max(m.xyz.nbytes, 4))
            if m.ntri:
                ctx.upload(d_xyz, m.xyz)
            d_dist = ctx.alloc(4 * res ** 3)
            d_occ = ctx.alloc(4 * ((og.nx + 31) // 32) * og.ny * og.nz)
            jobs.append(Job(d_xyz, m.ntri, g, d_dist, og, d_occ))
            bufs.append((d_xyz, d_dist, d_occ))
            grids.append((g, og))
        ctx.bake_batch_device(jobs)
        st = ctx.stats()
        assert st["voxels"] == sum(res ** 3 for m, res in items if m.ntri)
        for (m, res), (d_xyz, d_dist, d_occ), (g, og) in zip(items, bufs, grids):
            got = ctx.download(d_dist, np.empty(g.shape(), np.float32))
            gocc = ctx.download(d_occ, np.empty(((og.nx + 31) // 32) * og.ny * og.nz, np.uint32))
            alone, aocc = ctx.bake_dense(m.xyz, g, og)
            assert np.array_equal(got.view(np.uint32), alone.view(np.uint32)), (m, res)
            assert np.array_equal(gocc, aocc)
        for b in bufs:
            for p in b:
                ctx.free(p)
    finally:
        ctx.set_option(bake.OPT_BRICK_SHAPE, 0)


def test_brick_shapes_agree(ctx, meshgen):
    """The brick shape changes speed only: the plain bake gives the same bits for every shape."""
    from italy_bake import bake

    m = meshgen.displaced(5, 21)
    g = bake.Grid.make((-0.61, -0.66, -0.59), 0.0137, 83, 91, 77)
    ref, _ = ctx.bake_dense(m.xyz, g, None)
    try:
        for shape in (1, 2, 3, 4):
            ctx.set_option(bake.OPT_BRICK_SHAPE, shape)
            d, _ = ctx.bake_dense(m.xyz, g, None)
            assert np.array_equal(d.view(np.uint32), ref.view(np.uint32)), shape
        ctx.set_option(bake.OPT_BRICK_SHAPE, 0)
        for order in (0, 1):  # tile launch order: grid order, costliest first (2, the default, is ref)
            ctx.set_option(bake.OPT_TILE_ORDER, order)
            d, _ = ctx.bake_dense(m.xyz, g, None)
            assert np.array_equal(d.view(np.uint32), ref.view(np.uint32)), order
        ctx.set_option(bake.OPT_TILE_ORDER, 2)
        for block in (1, 2, 4, 8, 16):  # LBVH leaf blocks (ref: the default)
            ctx.set_option(bake.OPT_LEAF_BLOCK, block)
            d, _ = ctx.bake_dense(m.xyz, g, None)
            assert np.array_equal(d.view(np.uint32), ref.view(np.uint32)), block
        ctx.set_option(bake.OPT_LEAF_BLOCK, bake.DEFAULT_OPTIONS[bake.OPT_LEAF_BLOCK])
        ctx.set_option(bake.OPT_QUERY_VARIANT, 0)
        d, _ = ctx.bake_dense(m.xyz, g, None)
        assert np.array_equal(d.view(np.uint32), ref.view(np.uint32))
    finally:
        for opt in (bake.OPT_BRICK_SHAPE, bake.OPT_TILE_ORDER, bake.OPT_LEAF_BLOCK, bake.OPT_QUERY_VARIANT):
            ctx.set_option(opt, bake.DEFAULT_OPTIONS[opt])
