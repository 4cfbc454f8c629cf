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


def test_resident_meshes_equal_soups(ctx, meshgen):
    """Jobs that name resident meshes (ib_mesh_create) give the same bits as jobs that bring soups, for
    closed and open meshes, slivers, a narrow band and occupancy, twice over the same meshes."""
    from italy_bake import Job, bake

    items = [it for it in _meshes(meshgen) if it[0].ntri]
    band_mesh = meshgen.stress(seed=31, ncomp=3, remove_frac=0.02)
    bufs, meshes = [], []
    try:
        soup_jobs, mesh_jobs, outs = [], [], []
        for m, res in items:
            g = bake.dense_grid(-0.8, 0.8, res)
            og = bake.dense_grid(-0.8, 0.8, res + 5)
            d_xyz = ctx.alloc(m.xyz.nbytes)
            ctx.upload(d_xyz, m.xyz)
            h = ctx.mesh_create(d_xyz, m.ntri)
            meshes.append(h)
            assert ctx.mesh_bytes(h) > 0
            pair = []
            for _ in range(2):
                d_dist, d_occ = ctx.alloc(4 * res ** 3), ctx.alloc(4 * ((og.nx + 31) // 32) * og.ny * og.nz)
                bufs += [d_dist, d_occ]
                pair.append((d_dist, d_occ))
            soup_jobs.append(Job(d_xyz, m.ntri, g, pair[0][0], og, pair[0][1]))
            mesh_jobs.append(Job(None, 0, g, pair[1][0], og, pair[1][1], mesh=h))
            outs.append((g, og, pair))
            bufs.append(d_xyz)
        # a narrow band over an open mesh
        g = bake.dense_grid(-1.0, 1.0, 64)
        nb = (g.nx // 8) ** 3
        d_xyz = ctx.alloc(band_mesh.xyz.nbytes)
        ctx.upload(d_xyz, band_mesh.xyz)
        hb = ctx.mesh_create(d_xyz, band_mesh.ntri)
        meshes.append(hb)
        band = []
        for _ in range(2):
            d_idx, d_pool = ctx.alloc(4 * nb), ctx.alloc(4 * 512 * nb)
            bufs += [d_idx, d_pool]
            band.append((d_idx, d_pool))
        bufs.append(d_xyz)
        soup_jobs.append(Job(d_xyz, band_mesh.ntri, g, band[0][1], bake.Grid(), None, 4, 0, band[0][0], nb))
        mesh_jobs.append(Job(None, 0, g, band[1][1], bake.Grid(), None, 4, 0, band[1][0], nb, mesh=hb))
        ctx.bake_batch_device(soup_jobs)
        for rep in range(2):
            ctx.bake_batch_device(mesh_jobs)
            assert mesh_jobs[0].ntri == items[0][0].ntri  # written back from the mesh
            for g, og, pair in outs:
                a = ctx.download(pair[0][0], np.empty(g.shape(), np.float32))
                b = ctx.download(pair[1][0], np.empty(g.shape(), np.float32))
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), rep
                nw = ((og.nx + 31) // 32) * og.ny * og.nz
                assert np.array_equal(ctx.download(pair[0][1], np.empty(nw, np.uint32)),
                                      ctx.download(pair[1][1], np.empty(nw, np.uint32)))
            assert mesh_jobs[-1].nbricks == soup_jobs[-1].nbricks > 0
            n = soup_jobs[-1].nbricks
            assert np.array_equal(ctx.download(band[0][0], np.empty(nb, np.int32)),
                                  ctx.download(band[1][0], np.empty(nb, np.int32)))
            assert np.array_equal(ctx.download(band[0][1], np.empty(512 * n, np.uint32)),
                                  ctx.download(band[1][1], np.empty(512 * n, np.uint32)))
    finally:
        for h in meshes:
            ctx.mesh_destroy(h)
        for p in bufs:
            ctx.free(p)
