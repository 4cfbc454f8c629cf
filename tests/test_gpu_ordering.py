"""Calls on one context that use different streams, a batch that fails part way, and a server whose budget
forces evictions: every later result is still the bake alone (and the oracle's)."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    for p in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
        if os.path.exists(p) or "/" not in p:
            return ctypes.CDLL(p)
    raise RuntimeError("libamdhip64.so not found")


def test_async_bake_on_another_stream_then_host_bake(ctx, meshgen):
    """A device bake left running on a caller's stream, then at once a host-buffer bake on the context's
    own stream: the second call orders itself after the first, which still owns the shared scratch."""
    from italy_bake import bake

    hip = _hip()
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreateWithFlags(ctypes.byref(stream), ctypes.c_uint(1)) == 0
    a = meshgen.displaced(6, 31)     # 81,920 triangles: a query long enough to overlap the next call
    b = meshgen.displaced(5, 32)
    g = bake.dense_grid(-0.8, 0.8, 128)
    og = bake.dense_grid(-0.8, 0.8, 96)
    nw = ((og.nx + 31) // 32) * og.ny * og.nz
    d_xyz = ctx.alloc(a.xyz.nbytes)
    d_dist = ctx.alloc(4 * 128 ** 3)
    d_occ = ctx.alloc(4 * nw)
    try:
        ctx.upload(d_xyz, a.xyz)
        for _ in range(3):
            ctx.bake_dense_device(d_xyz, a.ntri, g, d_dist, og, d_occ, stream=stream.value)
            got_b, got_bocc = ctx.bake_dense(b.xyz, g, og)       # no wait for the first bake
            assert hip.hipStreamSynchronize(stream) == 0
            got_a = ctx.download(d_dist, np.empty(g.shape(), np.float32))
            got_aocc = ctx.download(d_occ, np.empty(nw, np.uint32))
            ref_a, ref_aocc = ctx.bake_dense(a.xyz, g, og)
            ref_b, ref_bocc = ctx.bake_dense(b.xyz, g, og)
            assert np.array_equal(got_a.view(np.uint32), ref_a.view(np.uint32))
            assert np.array_equal(got_aocc, ref_aocc)
            assert np.array_equal(got_b.view(np.uint32), ref_b.view(np.uint32))
            assert np.array_equal(got_bocc, ref_bocc)
    finally:
        for p in (d_xyz, d_dist, d_occ):
            ctx.free(p)
        hip.hipStreamDestroy(stream)


def test_batch_failing_mid_build_leaves_the_context_usable(ctx, meshgen, oracle):
    """A batch that runs out of its context's budget while its meshes are being built returns
    IB_ERR_OUT_OF_BUDGET after its queued kernels end; the next bake on that context equals the oracle."""
    from italy_bake import BakeError, Job, bake

    small = bake.Context(0, hbm_budget_bytes=48 << 20)
    meshes = [meshgen.displaced(5, 40 + i) for i in range(8)]    # ~12 MB of tree and tables each
    bufs = []
    try:
        jobs = []
        g = bake.dense_grid(-0.8, 0.8, 32)
        for m in meshes:
            d_xyz, d_dist = ctx.alloc(m.xyz.nbytes), ctx.alloc(4 * 32 ** 3)   # outside the small budget
            ctx.upload(d_xyz, m.xyz)
            bufs += [d_xyz, d_dist]
            jobs.append(Job(d_xyz, m.ntri, g, d_dist, bake.Grid(), None))
        with pytest.raises(BakeError) as e:
            small.bake_batch_device(jobs)
        assert e.value.status == bake.IB_ERR_OUT_OF_BUDGET
        mesh, cfg = meshgen.workload_plain_small()
        g1 = bake.dense_grid(cfg["lo"], cfg["hi"], cfg["sdf_res"])
        dist, occ = small.bake_dense(mesh.xyz, g1, g1)
        ref, _ = oracle.bake_dense(mesh.xyz, list(g1.origin), g1.h, g1.nx, g1.ny, g1.nz)
        assert np.array_equal(dist.view(np.uint32), ref.view(np.uint32))
        assert np.array_equal(occ, oracle.occupancy(mesh.xyz, list(g1.origin), g1.h, g1.nx, g1.ny, g1.nz))
    finally:
        for p in bufs:
            ctx.free(p)
        small.close()


def test_server_evictions_under_a_small_budget(ctx, meshgen):
    """Keyed requests whose resident meshes do not all fit the budget: the least recently used are evicted
    (and destroyed by their owning worker between batches), every result equals the bake alone, and no
    staging allocation fails."""
    from italy_bake import bake
    from italy_bake.serve import Server

    meshes = [meshgen.decimated(30000 + 3000 * i, 700 + i) for i in range(6)]   # ~12-16 MiB resident each
    g = bake.dense_grid(-0.8, 0.8, 40)
    # a request reserves ~28 MiB while it runs: at most about three meshes stay resident beside it
    with Server(hbm_budget_bytes=64 << 20, workers=2, max_batch=2) as srv:
        tickets = []
        for rnd in range(3):
            for i, m in enumerate(meshes):
                t = srv.submit(m.xyz, g, 0, g, mesh_key=500 + i)
                tickets.append((t, m))
                srv.wait(t)
        for t, m in tickets:
            d, _, occ = srv.fetch(t)
            ad, aocc = ctx.bake_dense(m.xyz, g, g)
            assert np.array_equal(d.view(np.uint32), ad.view(np.uint32))
            assert np.array_equal(occ, aocc)
            srv.release(t)
        c = srv.counters()
        assert c["evictions"] > 0, c
        assert c["failed"] == 0 and c["alloc_failures"] == 0, c
        assert c["in_use"] == c["mesh_bytes"]
