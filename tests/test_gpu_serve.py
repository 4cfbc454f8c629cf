"""Served bakes: every result is the same alone and in the mix; budget admission; latencies."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(meshgen, bake):
    reqs = []
    for i in range(6):
        m = meshgen.decimated(800 + 3000 * i, 500 + i)
        res = [24, 40, 64, 33, 48, 72][i]
        reqs.append((m.xyz, bake.dense_grid(-0.8, 0.8, res), 0, bake.dense_grid(-0.8, 0.8, res)))
    s = meshgen.stress(seed=77, ncomp=4, remove_frac=0.02)
    reqs.append((s.xyz, bake.dense_grid(-1.0, 1.0, 96), 4, None))     # narrow band, open
    reqs.append((meshgen.uv_sphere(0.7).xyz, bake.dense_grid(-0.9, 0.9, 40), 4, None))
    return reqs


def test_served_results_equal_alone(ctx, meshgen):
    from italy_bake import bake
    from italy_bake.serve import Server

    reqs = _requests(meshgen, bake)
    with Server(workers=3, max_batch=4) as srv:
        # submit everything at once, twice over, so batches mix and overlap on the device
        tickets = [(srv.submit(*r), k) for k in range(2) for r in [reqs[i] for i in range(len(reqs))]]
        srv.drain()
        for (t, _), r in zip(tickets, reqs + reqs):
            xyz, g, band, og = r
            d, idx, occ = srv.fetch(t)
            if band:
                ai, ap = ctx.bake_band(xyz, g, band)
                assert np.array_equal(idx, ai)
                assert np.array_equal(d.view(np.uint32), ap.view(np.uint32))
            else:
                ad, aocc = ctx.bake_dense(xyz, g, og)
                assert np.array_equal(d.view(np.uint32), ad.view(np.uint32))
                assert np.array_equal(occ, aocc)
            srv.release(t)
        c = srv.counters()
        assert c["in_use"] == 0 and c["completed"] == 2 * len(reqs)


def test_budget_admission_and_refusal(meshgen):
    from italy_bake import BakeError, bake
    from italy_bake.serve import Server

    m = meshgen.displaced(5, 3)
    g = bake.dense_grid(-0.8, 0.8, 64)
    with Server(hbm_budget_bytes=64 << 20, workers=2, max_batch=8) as srv:
        ts = [srv.submit(m.xyz, g) for _ in range(12)]  # ~1 MiB each + 20k triangles: a few at a time
        srv.drain()
        infos = [srv.info(t) for t in ts]
        assert all(i["state"] == 2 for i in infos)
        for t in ts:
            srv.release(t)
        big = srv.submit(m.xyz, bake.dense_grid(-0.8, 0.8, 400))  # 256 MiB of output: can never fit
        with pytest.raises(BakeError) as e:
            srv.wait(big, 60)
        assert e.value.status == 3
        srv.release(big)


def test_open_loop_latency(meshgen):
    from italy_bake import bake
    from italy_bake.serve import Server, open_loop

    reqs = [(meshgen.decimated(2000 + 1000 * i, 900 + i).xyz, bake.dense_grid(-0.8, 0.8, 48), 0, None)
            for i in range(4)]
    with Server(workers=3, max_batch=32) as srv:
        r = open_loop(srv, reqs, rate_hz=100, duration_s=2.0, seed=1)
    assert r["failed"] == 0 and r["requests"] > 100
    assert r["latency_s"]["p99"] >= r["latency_s"]["p50"] > 0


def test_resident_meshes_in_the_server(ctx, meshgen):
    """Requests with mesh keys reuse the resident LBVH: every result equals the bake alone, the cache
    counts one build per key, and a key whose soup has another triangle count is baked without it."""
    from italy_bake import bake
    from italy_bake.serve import Server

    reqs = _requests(meshgen, bake)
    with Server(workers=3, max_batch=4) as srv:
        tickets = []
        for rnd in range(3):
            for i, r in enumerate(reqs):
                tickets.append((srv.submit(*r, mesh_key=100 + i), r))
            srv.drain()
        other = meshgen.decimated(1200, 5)  # reuses key 100 with a different triangle count
        og = bake.dense_grid(-0.8, 0.8, 30)
        tickets.append((srv.submit(other.xyz, og, 0, og, mesh_key=100), (other.xyz, og, 0, og)))
        srv.drain()
        for t, (xyz, g, band, occ_g) in tickets:
            d, idx, occ = srv.fetch(t)
            if band:
                ai, ap = ctx.bake_band(xyz, g, band)
                assert np.array_equal(idx, ai)
                assert np.array_equal(d.view(np.uint32), ap.view(np.uint32))
            else:
                ad, aocc = ctx.bake_dense(xyz, g, occ_g)
                assert np.array_equal(d.view(np.uint32), ad.view(np.uint32))
                if occ_g is not None:
                    assert np.array_equal(occ, aocc)
            srv.release(t)
        c = srv.counters()
        assert c["mesh_builds"] == len(reqs)
        assert c["mesh_hits"] >= 2 * len(reqs)
        assert c["meshes_resident"] == len(reqs) and c["mesh_bytes"] > 0
        assert c["in_use"] == c["mesh_bytes"]  # every result released: only the resident meshes remain
