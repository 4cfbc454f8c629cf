"""The batched and served configurations at their own sizes, checked against the oracle.

- Batched: the 512-mesh workload (log-uniform 500..200k triangles, SDF + occupancy at 32/64/128/256^3)
  through the autotuned persistent launch.  Every mesh equals its plain per-mesh bake bit for bit; every
  32nd mesh equals the oracle on 20,000 sampled voxels (half within 3 voxels of the surface) and in its
  whole occupancy grid.
- Served: one 2M-triangle 512^3 narrow band (+-4 voxels) in a mix with precision-stress unions (384^3
  bands of open meshes), under the 200 GB budget.  Every result equals the bake alone; the large job equals
  the oracle on sampled band voxels, holds every surface voxel's brick, and is within 1e-5 * extent of the
  FP64 distance; the unions' distances equal the oracle's and their signs the FP64 winding number's.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sample(rng, dist, h, n):
    """n voxel indices of a dense grid, half anywhere and half within 3 voxels of the surface."""
    flat = np.abs(dist.reshape(-1))
    near = np.nonzero(flat < 3 * h)[0]
    a = rng.integers(0, flat.size, n // 2)
    b = rng.choice(near, min(n // 2, near.size), replace=False) if near.size else np.zeros(0, np.int64)
    return np.unique(np.concatenate([a, b]))


def test_batched_512_meshes_against_plain_and_oracle(ctx, meshgen, oracle):
    from italy_bake import Job, bake

    t0 = time.perf_counter()
    meshes, res = meshgen.workload_batched(512)
    print(f"generated 512 meshes in {time.perf_counter() - t0:.1f} s", flush=True)
    jobs, bufs = [], []
    try:
        for m, r in zip(meshes, res):
            g = bake.dense_grid(-0.8, 0.8, r)
            d_xyz, d_dist, d_occ = ctx.alloc(m.xyz.nbytes), ctx.alloc(4 * r ** 3), ctx.alloc(4 * ((r + 31) // 32) * r * r)
            ctx.upload(d_xyz, m.xyz)
            jobs.append(Job(d_xyz, m.ntri, g, d_dist, g, d_occ))
            bufs.append((d_xyz, d_dist, d_occ))
        shape, blocks, _ = ctx.autotune_batch(jobs, sample_stride=32)
        print(f"autotuned: shape {shape}, {blocks} blocks per CU", flush=True)
        ctx.bake_batch_device(jobs)
        st = ctx.stats()
        assert st["voxels"] == sum(r ** 3 for r in res) and st["stack_overflows"] == 0
        for opt in (bake.OPT_BRICK_SHAPE, bake.OPT_BLOCKS_PER_CU):
            ctx.set_option(opt, bake.DEFAULT_OPTIONS[opt])
        rng = np.random.default_rng(512)
        checked = 0
        for i, ((m, r), (d_xyz, d_dist, d_occ), j) in enumerate(zip(zip(meshes, res), bufs, jobs)):
            got = ctx.download(d_dist, np.empty((r, r, r), np.float32))
            gocc = ctx.download(d_occ, np.empty(((r + 31) // 32) * r * r, np.uint32))
            # the plain bake of the same mesh into the same buffers
            ctx.bake_dense_device(d_xyz, m.ntri, j.sdf, d_dist, j.occ, d_occ)
            alone = ctx.download(d_dist, np.empty((r, r, r), np.float32))
            aocc = ctx.download(d_occ, np.empty_like(gocc))
            assert np.array_equal(got.view(np.uint32), alone.view(np.uint32)), (i, m.ntri, r)
            assert np.array_equal(gocc, aocc), (i, m.ntri, r)
            if i % 32 == 0:
                g = j.sdf
                idx = _sample(rng, got, g.h, 20000)
                pts = oracle.cell_centres_at(list(g.origin), g.h, idx, r, r)
                ref = oracle.signed_points_closed(m.xyz, pts)
                assert ref is not None, "the batched workload's meshes are closed"
                assert np.array_equal(got.reshape(-1)[idx].view(np.uint32), ref.view(np.uint32)), (i, m.ntri, r)
                assert np.array_equal(gocc, oracle.occupancy(m.xyz, list(g.origin), g.h, r, r, r)), (i, m.ntri, r)
                checked += 1
        assert checked == 16
    finally:
        for b in bufs:
            for p in b:
                ctx.free(p)


def _band_voxels(idx, pool, slots):
    """(linear brick index, offset in the brick, distance) of the given pool slots' 512 voxels."""
    kept = np.nonzero(idx.reshape(-1) >= 0)[0]
    slot_of = idx.reshape(-1)[kept]
    brick = np.empty(len(slot_of), np.int64)
    brick[slot_of] = kept
    b = np.repeat(brick[slots], 512)
    off = np.tile(np.arange(512), len(slots))
    return b, off, pool.reshape(-1, 512)[slots].reshape(-1)


def _voxel_of(b, off, nb):
    """grid voxel (x, y, z) of brick b (x fastest over nb[0] x nb[1] bricks), offset (z%8*8 + y%8)*8 + x%8"""
    bx, by, bz = b % nb[0], (b // nb[0]) % nb[1], b // (nb[0] * nb[1])
    return bx * 8 + off % 8, by * 8 + (off // 8) % 8, bz * 8 + off // 64


def test_served_2m_band_in_a_mix_against_alone_and_oracle(ctx, meshgen, oracle):
    from italy_bake import bake
    from italy_bake.serve import Server

    t0 = time.perf_counter()
    big = meshgen.decimated(2_000_000, meshgen.BASE_SEED + 200)
    unions = [meshgen.stress(meshgen.BASE_SEED + 100 + i, 64, 0.02) for i in range(3)]
    print(f"generated the meshes in {time.perf_counter() - t0:.1f} s", flush=True)
    g512 = bake.dense_grid(-0.8, 0.8, 512)
    g384 = bake.dense_grid(-1.0, 1.0, 384)
    reqs = [(unions[0].xyz, g384), (big.xyz, g512), (unions[1].xyz, g384), (unions[2].xyz, g384)]
    with Server(hbm_budget_bytes=200 << 30, workers=3, max_batch=64) as srv:
        tickets = [srv.submit(x, g, 4, None, mesh_key=k + 1) for k, (x, g) in enumerate(reqs)]
        srv.drain()
        served = [srv.fetch(t) for t in tickets]
        for t in tickets:
            srv.release(t)
        c = srv.counters()
    assert c["failed"] == 0 and c["alloc_failures"] == 0, c
    rng = np.random.default_rng(2048)
    for k, ((xyz, g), (pool, idx, _)) in enumerate(zip(reqs, served)):
        ai, ap = ctx.bake_band(xyz, g, 4)
        assert np.array_equal(idx, ai), k
        assert np.array_equal(pool.view(np.uint32), ap.view(np.uint32)), k
        nb = (g.nx // 8, g.ny // 8, g.nz // 8)
        slots = rng.choice(len(pool), min(64 if k != 1 else 160, len(pool)), replace=False)
        b, off, d = _band_voxels(idx, pool, slots)
        x, y, z = _voxel_of(b, off, nb)
        lin = (z * g.ny + y) * g.nx + x
        pts = oracle.cell_centres_at(list(g.origin), g.h, lin, g.nx, g.ny)
        extent = np.float32(g.h) * g.nx
        if k == 1:  # the 2M-triangle displaced icosphere is closed: the dense CPU bake at the sampled voxels
            ref = oracle.signed_points_closed(xyz, pts)
            assert ref is not None
            assert np.array_equal(d.view(np.uint32), ref.view(np.uint32))
            d64 = oracle.udf_points_f64(xyz, pts.astype(np.float64))
            assert np.max(np.abs(np.abs(d).astype(np.float64) - d64)) <= 1e-5 * extent
            # every voxel that holds a surface point lies in a kept brick
            tri = xyz.reshape(-1, 3, 3)[rng.integers(0, xyz.size // 9, 20000)]
            w = rng.dirichlet([1, 1, 1], len(tri)).astype(np.float32)
            p = np.einsum("nk,nkc->nc", w, tri.astype(np.float32))
            v = np.floor((p - np.float32(g.origin[0])) / np.float32(g.h)).astype(np.int64)
            assert (idx[v[:, 2] // 8, v[:, 1] // 8, v[:, 0] // 8] >= 0).all()
        else:  # open unions: FP32 distances equal the oracle's, signs the FP64 winding number's
            ref = oracle.udf_points(xyz, pts)
            assert np.array_equal(np.abs(d).view(np.uint32), ref.view(np.uint32)), k
            near = np.nonzero(np.abs(d) > 1e-4 * extent)[0]
            pick = rng.choice(near, min(256, near.size), replace=False)
            sref, _ = oracle.signed_points_gwn(xyz, pts[pick])
            assert np.array_equal(d[pick] < 0, sref < 0), k
