"""Bake benchmark: SDF voxel-queries/s of the plain single-mesh bake on one MI355X.

Workload (one step): a 327,680-triangle displaced icosphere (subdivision 7, seeded 8-mode radial
displacement) already resident in HBM is baked to a dense 256^3 FP32 signed-distance grid over
[-0.8, 0.8]^3 plus a 512^3 bit-packed occupancy grid: weld + closedness check, LBVH build,
occupancy, +z column crossings, per-voxel closest triangle + sign.

Prints ONE JSON line.  `roofline` describes the dominant kernel (the per-voxel query); its FLOPs
are algorithmic: traversal counters from one counted run times per-test FLOP constants.
`cpu_baseline` times the oracle's all-core CPU bake (same exact math) on a bounded z-slab sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "italy-rs_amd"))

METRIC = "SDF voxel-queries/s (327k tris, 256³) batched vs plain; p99 bake latency"
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak (published, with packed FP32)
PEAK_HBM_GBS = 8000.0
# FLOPs per test, counted from the kernel source (ib_device.hpp; add/sub/mul = 1, fma = 2, min/max,
# clamps and compares not counted): point/box distance 6 sub + 1 mul + 2 fma = 11; branch-free point/
# triangle distance in the triangle's plane frame: offset 3 + three frame dots 15 + edge 0 (scale 1,
# fma 2, square 3) 6 + edge 1 (offset 1, dot 3, scale 1, two fma 4, square 3) 12 + edge 2 (offsets 2,
# dot 3, scale 1, two fma 4, square 3) 13 + two edge functions 6 + w^2 added to the in-plane term (one fma) 2 = 57
FLOP_BOX = 11
FLOP_TRI = 57
# point/oriented-box distance of a leaf block: offset 3 + three frame dots 15 + three |u| - e 3 + sum of
# squares 5 = 26
FLOP_OBB = 26
# FLOPs per cell of the glass cup's distance lambda (sdf_procedural.cpp:123-130; sqrt and divide = 1):
# cell centre 3 x 3 = 9, three spheres 3 x (3 sub + 3 mul + 2 add + sqrt + sub) = 30, two plane
# offsets p.y - h = 2, smooth intersect (b - a, * 0.5, / k, + 0.5, two 1 - h, three mul, add, mul, sub) = 12
FLOP_CUP_CELL = 53


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-zstep", type=int, default=4, help="the CPU bakes every zstep-th z-slab of the grid")
    ap.add_argument("--cpu-zstep1", type=int, default=64, help="z-slab stride of the single-thread CPU sample")
    ap.add_argument("--no-voxelize", action="store_true", help="skip the voxelizeMesh drop-in comparison")
    ap.add_argument("--variant", type=int, default=None, help="query kernel: 0 per-voxel, 1 per-brick wave")
    ap.add_argument("--batch-meshes", type=int, default=512, help="meshes in the batched workload (0: skip)")
    ap.add_argument("--no-band", action="store_true", help="skip the narrow-band precision-stress bake")
    ap.add_argument("--no-consumer", action="store_true", help="skip the grid consumer (sampling / tracing)")
    ap.add_argument("--no-analytic", action="store_true", help="skip the analytic CSG bake (glass cup at 320)")
    ap.add_argument("--no-ingest", action="store_true", help="skip the weld / watertight gate")
    ap.add_argument("--no-material", action="store_true", help="skip the per-cell material splat")
    ap.add_argument("--serve-seconds", type=float, default=8.0,
                    help="arrival window of the served workload on the batched mesh distribution (0: skip)")
    ap.add_argument("--serve-stress-seconds", type=float, default=2.0,
                    help="arrival window of the served precision-stress + large-job mix (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--no-serve-capacity", dest="serve_capacity", action="store_false",
                    help="skip the served capacity measurements (closed-loop saturation and rate sweep)")
    ap.add_argument("--batch-profile", type=str, default=None, metavar="SHAPE,BLOCKS",
                    help="only run the batched workload twice at this brick shape and blocks per CU (no autotuner, no "
                         "plain comparison) and print its timings: the command a kernel trace of the batch runs")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one plain device-to-device copy of this many MiB (a known-size kernel that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters of a profiling run)")
    return ap.parse_args()


def host_cores():
    """CPUs this process may run on: its affinity set, limited by a cgroup CPU quota when one is set."""
    n = len(os.sched_getaffinity(0))
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return n


def plain_band_rate(ctx, meshes, grids, weights, reps=2):
    """Plain per-mesh narrow-band bakes (one device call per mesh, soup resident): meshes/s of a request
    mix drawn with `weights`, from each mesh's best device-call time."""
    from italy_bake import Job

    times = []
    for m, g in zip(meshes, grids):
        nb = (g.nx // 8) * (g.ny // 8) * (g.nz // 8)
        d_xyz, d_idx, d_pool = ctx.alloc(m.xyz.nbytes), ctx.alloc(4 * nb), ctx.alloc(2048 * nb)
        ctx.upload(d_xyz, m.xyz)
        best = np.inf
        for _ in range(reps + 1):
            job = Job(d_xyz, m.ntri, g, d_pool, type(g)(), None)
            job.band_voxels, job.d_brick_index, job.pool_bricks = 4, d_idx, nb
            ctx.synchronize()
            t0 = time.perf_counter()
            ctx.bake_batch_device([job])
            best = min(best, time.perf_counter() - t0)
        times.append(best)
        for p in (d_xyz, d_idx, d_pool):
            ctx.free(p)
    w = np.asarray(weights, np.float64) / np.sum(weights)
    return 1.0 / float(np.dot(w, times)), times


def served_section(args, batched, ctx):
    """The bake server (3 worker streams, HBM budget 200 GB) under two readings of the served configuration.

    batched_distribution: requests drawn from the batched workload's distribution (dense SDF + occupancy,
      500..200k triangles, 32..256^3), keyed by mesh so that each LBVH stays resident after its first
      request; batched_distribution_rebuilt: the same requests without keys (every request uploaded and
      built).  The plain per-mesh bake of the batched workload is the comparison point.
    stress_mix: 95% precision-stress unions (384^3 narrow band) + 5% large jobs (2M-triangle displaced
      icosphere, 512^3 narrow band), compared with the plain per-mesh band bake of the same mix.
    Each reading reports the open-loop Poisson run at the configured rate and the server's capacity: the
    closed-loop saturation throughput and an open-loop sweep below it (serve.capacity)."""
    from italy_bake import bake, meshgen
    from italy_bake.serve import Server, capacity, open_loop

    out = {}
    budget = 200 << 30
    if args.serve_seconds > 0:
        meshes, res = meshgen.workload_batched(128, seed=meshgen.BASE_SEED + 9000)
        grids = [bake.dense_grid(-0.8, 0.8, r) for r in res]
        # keyed: each mesh's LBVH stays resident after its first request; unkeyed: every request is built
        for name, keyed in (("batched_distribution", True), ("batched_distribution_rebuilt", False)):
            reqs = [(m.xyz, g, 0, g, (i + 1) if keyed else 0) for i, (m, g) in enumerate(zip(meshes, grids))]
            with Server(hbm_budget_bytes=budget, workers=3, max_batch=256) as srv:
                open_loop(srv, [q[:4] for q in reqs[:8]], rate_hz=50, duration_s=0.5, seed=7)  # warm the contexts
                r = open_loop(srv, reqs, rate_hz=args.serve_rate, duration_s=args.serve_seconds, seed=1)
                if args.serve_capacity:
                    r["capacity"] = capacity(srv, reqs, n=1536, concurrency=256, rates=(0.5, 0.8, 0.95),
                                             duration_s=2.0, seed=3)
                c = srv.counters()
            r.pop("_latencies")
            r["workload"] = ("requests drawn from 128 meshes of the batched distribution (dense SDF + occupancy at "
                             "32/64/128/256^3)" + (", each mesh's LBVH kept resident under its key" if keyed else
                                                   ", every request uploaded and built"))
            r["mesh_builds"], r["mesh_hits"] = c["mesh_builds"], c["mesh_hits"]
            if batched:
                r["plain_per_mesh_meshes_per_s"] = batched["plain_meshes_per_s"]
                if "capacity" in r:
                    r["capacity_vs_plain"] = r["capacity"]["capacity_meshes_per_s"] / batched["plain_meshes_per_s"]
            out[name] = r
    if args.serve_stress_seconds > 0:
        stress = [meshgen.stress(meshgen.BASE_SEED + 100 + i, 64, 0.02) for i in range(6)]
        big = meshgen.decimated(2_000_000, meshgen.BASE_SEED + 200)
        g384 = bake.dense_grid(-1.0, 1.0, 384)
        g512 = bake.dense_grid(-0.8, 0.8, 512)
        # 95% / 5%: 19 stress entries per large entry in the uniform draw
        # keyed by mesh: a mesh's LBVH stays resident after its first request
        reqs = [(m.xyz, g384, 4, None, 1 + i) for i, m in enumerate(stress)] * 19 + [(big.xyz, g512, 4, None, 99)] * 6
        with Server(hbm_budget_bytes=budget, workers=3, max_batch=64) as srv:
            r = open_loop(srv, reqs, rate_hz=args.serve_rate, duration_s=args.serve_stress_seconds, seed=2)
            if args.serve_capacity:
                r["capacity"] = capacity(srv, reqs, n=160, concurrency=24, rates=(0.5, 0.8, 0.95), duration_s=3.0,
                                         seed=4)
        lat = r.pop("_latencies")
        big_lat = [l for k, l in lat if reqs[k][1] is g512]
        r["large_jobs"] = len(big_lat)
        r["large_job_latency_s_max"] = max(big_lat) if big_lat else None
        r["workload"] = ("95% precision-stress unions (64 tori/boxes, 2% removed, 384^3 narrow band +-4) + 5% "
                         "2M-triangle displaced icospheres (512^3 narrow band +-4)")
        plain, times = plain_band_rate(ctx, stress + [big], [g384] * 6 + [g512], [0.95 / 6] * 6 + [0.05])
        r["plain_per_mesh_meshes_per_s"] = plain
        r["plain_ms"] = {"precision_stress": 1e3 * float(np.mean(times[:6])), "large_2m_512": 1e3 * times[6]}
        if "capacity" in r:
            r["capacity_vs_plain"] = r["capacity"]["capacity_meshes_per_s"] / plain
            r["offered_vs_capacity"] = args.serve_rate / r["capacity"]["capacity_meshes_per_s"]
        out["stress_mix"] = r
    return out or None


def consumer_section(ctx, sdf, d_dist):
    """The renderer's grid consumer on the device-resident 256^3 bake: 2^24 trilinear samples with
    B-spline normals and nearest cells, and 2^20 sphere-traced rays from a shell of radius 0.75
    towards random points near the centre (no host round trip of the grid)."""
    rng = np.random.default_rng(3)
    npts, nray = 1 << 24, 1 << 20
    pts = rng.uniform(-0.8, 0.8, (npts, 3)).astype(np.float32)
    u = rng.normal(size=(nray, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    org = (0.75 * u).astype(np.float32)
    dirs = rng.uniform(-0.2, 0.2, (nray, 3)) - org
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    bufs = {k: ctx.alloc(n) for k, n in (("pts", pts.nbytes), ("val", 4 * npts), ("nrm", 12 * npts),
                                          ("near", 4 * npts), ("org", org.nbytes), ("dir", dirs.nbytes),
                                          ("t", 4 * nray))}
    try:
        ctx.upload(bufs["pts"], pts)
        ctx.upload(bufs["org"], org)
        ctx.upload(bufs["dir"], dirs)

        def sample():
            ctx.sample_grid_device(sdf, d_dist, bufs["pts"], npts, bufs["val"], bufs["nrm"], bufs["near"])

        def trace():
            ctx.trace_grid_device(sdf, d_dist, bufs["org"], bufs["dir"], nray, bufs["t"])

        res = {}
        for name, fn, n in (("sample", sample, npts), ("trace", trace, nray)):
            fn()
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(5):
                fn()
            ctx.synchronize()
            res[name] = n * 5 / (time.perf_counter() - t0)
        t = ctx.download(bufs["t"], np.empty(nray, np.float32))
    finally:
        for b in bufs.values():
            ctx.free(b)
    return {"workload": "256^3 grid of the 327,680-triangle bake, device-resident",
            "samples_per_s": res["sample"], "sample_outputs": "value + B-spline normal + nearest cell",
            "rays_per_s": res["trace"], "rays_hit": int((t > 0).sum()), "rays": nray}


def analytic_section(ctx, cpu=True, steps=20):
    """The viewer's glass cup (sdf_procedural.cpp:108-140 at kGlassCupResolution 320, main.cpp:116):
    324 x 264 x 324 cells of an analytic CSG program (two spheres, subtract, rim plane, smooth
    intersect, two-branch select) into device-resident distances + branch bytes.  The reference
    evaluates it on one host core; its notes put the bake at about 2 s of wall time
    (humans.md:420-423).  The GPU grid is compared with the oracle's per-cell restatement."""
    from italy_bake import procedural as pr

    prog, bmin, bmax = pr.cup_program(-1.0, (0.0, 0.0), glass=True)
    g = pr.analytic_framing(bmin, bmax, 320)
    n = g.nx * g.ny * g.nz
    d_dist, d_branch = ctx.alloc(4 * n), ctx.alloc(n)
    try:
        pr.bake_analytic_device(ctx, prog, g, d_dist, d_branch)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            pr.bake_analytic_device(ctx, prog, g, d_dist, d_branch)
        ctx.synchronize()
        wall_s = (time.perf_counter() - t0) / steps
        ev = []
        for _ in range(steps):  # the launch alone, HIP events on its stream
            pr.bake_analytic_device(ctx, prog, g, d_dist, d_branch)
            ev.append(ctx.stats()["ms_total"])
        dev_s = float(np.mean(ev)) * 1e-3
        dist = ctx.download(d_dist, np.empty(n, np.float32))
        branch = ctx.download(d_branch, np.empty(n, np.uint8))
    finally:
        ctx.free(d_dist)
        ctx.free(d_branch)
    t0 = time.perf_counter()
    hd, hb = pr.bake_analytic(ctx, prog, g)
    host_s = time.perf_counter() - t0
    out = {"workload": f"glass cup, analytic CSG at res 320: {g.nx}x{g.ny}x{g.nz} = {n} cells, f32 distance + "
                       "branch byte", "cells": n, "ms_per_bake": dev_s * 1e3, "cells_per_s": n / dev_s,
           "ms_per_bake_wall": wall_s * 1e3, "ms_per_bake_with_host_copy": host_s * 1e3,
           "gflop_per_bake": FLOP_CUP_CELL * n / 1e9,
           "roofline": {"bound": "valu", "achieved": FLOP_CUP_CELL * n / dev_s / 1e12, "peak": PEAK_FP32_TFLOPS,
                        "unit": "TFLOP/s", "frac": FLOP_CUP_CELL * n / dev_s / 1e12 / PEAK_FP32_TFLOPS},
           "algorithmic_bytes_per_bake": 5 * n, "hbm_GBs_algorithmic": 5 * n / dev_s / 1e9}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import ib_oracle

        t0 = time.perf_counter()
        _, _, cd, cb = ib_oracle.cup_cells(True, 320, -1.0, (0.0, 0.0))
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n / cpu_s, "unit": "cells/s", "cores": 1, "kind": "port",
                               "seconds": cpu_s, "gpu_multiple": cpu_s / dev_s,
                               "sample": "the whole grid, the reference's lambda per cell on one core"}
        out["matches_oracle"] = bool(np.array_equal(dist, cd.reshape(-1)) and np.array_equal(branch, cb.reshape(-1))
                                     and np.array_equal(hd, dist) and np.array_equal(hb, branch))
    return out


def ingest_section(ctx, mesh, d_xyz, cpu=True, steps=10):
    """The reference's ingestion gate before a bake (main.cpp:227-241): isWatertight (mesh_validate.cpp:21-59,
    which welds with weldPositions, mesh_weld.cpp:26-48) on the device-resident 327,680-triangle soup.  The
    verdict and edge counts are compared with the oracle's restatement, timed on one host core as the
    reference runs it."""
    bmin, bmax = mesh.xyz.reshape(-1, 3).min(0), mesh.xyz.reshape(-1, 3).max(0)
    got = ctx.watertight_device(d_xyz, mesh.ntri, bmin, bmax)
    ctx.synchronize()
    ev = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.watertight_device(d_xyz, mesh.ntri, bmin, bmax)
        ev.append(ctx.stats()["ms_total"])
    wall_s = (time.perf_counter() - t0) / steps
    dev_s = float(np.mean(ev)) * 1e-3
    nbytes = mesh.xyz.nbytes  # the soup, read once: the algorithmic traffic of the gate
    out = {"workload": f"isWatertight on the {mesh.ntri}-triangle displaced icosphere ({3 * mesh.ntri} corners "
                       "welded on the 1e-5 * diagonal lattice)",
           "watertight": got[0], "boundary_edges": got[1], "nonmanifold_edges": got[2],
           "ms_device": dev_s * 1e3, "ms_wall": wall_s * 1e3, "triangles_per_s": mesh.ntri / dev_s,
           "hbm_GBs_algorithmic": nbytes / dev_s / 1e9}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import ib_oracle

        t0 = time.perf_counter()
        ref = ib_oracle.watertight(mesh.xyz, bmin, bmax)
        cpu_s = time.perf_counter() - t0
        out["matches_oracle"] = tuple(ref) == tuple(got)
        out["cpu_baseline"] = {"value": mesh.ntri / cpu_s, "unit": "triangles/s", "cores": 1, "kind": "port",
                               "seconds": cpu_s, "gpu_multiple": cpu_s / dev_s,
                               "sample": "the whole mesh on one core (hash-map weld + sorted edge count)"}
    return out


def material_section(ctx, mesh, res=256, cpu=True, steps=3):
    """bakePerCellMaterial (sdf_baker.cpp:76-128), the per-cell material half of bakeSdf(mesh, 256): every
    cell a triangle's box test touches (the bakeSdf framing of the 327,680-triangle mesh) averages the
    triangle's (r, g, b, metallic, roughness) in triangle order.  Host buffers in and out, as bakeSdf
    returns them; beside it the oracle's single-threaded restatement (the reference's loop is
    single-threaded)."""
    from italy_bake import bake

    g = bake.sdf_framing(mesh.bmin, mesh.bmax, res)
    mat = np.random.default_rng(5).uniform(0, 1, (mesh.ntri, 5)).astype(np.float32)
    ctx.splat_materials(mesh.xyz, mat, g)
    t0 = time.perf_counter()
    for _ in range(steps):
        rgb, met, rough = ctx.splat_materials(mesh.xyz, mat, g)
    gpu_s = (time.perf_counter() - t0) / steps
    ncell = g.nx * g.ny * g.nz
    out = {"workload": f"bakePerCellMaterial of the {mesh.ntri}-triangle mesh on bakeSdf's {g.nx}x{g.ny}x{g.nz} grid "
                       "(res 256), host buffers in and out (PCIe included)",
           "cells": ncell, "ms_per_splat": gpu_s * 1e3, "triangles_per_s": mesh.ntri / gpu_s}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import ib_oracle

        t0 = time.perf_counter()
        ref = ib_oracle.splat_materials(mesh.xyz, mat, list(g.origin), g.h, g.nx, g.ny, g.nz)
        cpu_s = time.perf_counter() - t0
        ref = ref.reshape(5, -1)
        out["matches_oracle"] = bool(np.array_equal(rgb.T, ref[:3]) and np.array_equal(met, ref[3])
                                     and np.array_equal(rough, ref[4]))
        out["cpu_baseline"] = {"value": mesh.ntri / cpu_s, "unit": "triangles/s", "cores": 1, "kind": "port",
                               "seconds": cpu_s, "gpu_multiple": cpu_s / gpu_s,
                               "sample": "the whole mesh on one core, in triangle order"}
    out["triangle_materials"] = triangle_materials_timing(ctx, mesh.ntri, cpu, steps)
    return out


def triangle_materials_timing(ctx, ntri, cpu=True, steps=3):
    """triangleMaterial (sdf_baker.cpp:42-74), the per-triangle input of the splat: 8 materials, 6 of them
    textured by four 1024x1024 RGBA8 textures (two sRGB), random UVs over the mesh's triangles.  Host
    buffers in and out (the textures are uploaded by every call)."""
    rng = np.random.default_rng(9)
    texs = [(rng.integers(0, 256, (1024, 1024, 4), dtype=np.uint8), i % 2 == 0) for i in range(4)]
    mats = [(tuple(rng.uniform(0, 1, 3)), float(rng.uniform()), float(rng.uniform()), int(t))
            for t in (-1, 0, 1, 2, 3, 0, 1, -1)]
    uvs = rng.uniform(-2, 3, (3 * ntri, 2)).astype(np.float32)
    tm = rng.integers(0, len(mats), ntri).astype(np.uint32)
    got = ctx.triangle_materials(ntri, uvs, tm, mats, texs)
    t0 = time.perf_counter()
    for _ in range(steps):
        got = ctx.triangle_materials(ntri, uvs, tm, mats, texs)
    gpu_s = (time.perf_counter() - t0) / steps
    out = {"workload": f"{ntri} triangles, 8 materials, four 1024x1024 RGBA8 textures (host buffers, PCIe "
                       "included)", "ms": gpu_s * 1e3, "triangles_per_s": ntri / gpu_s}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import ib_oracle

        t0 = time.perf_counter()
        ref = ib_oracle.triangle_materials(ntri, uvs, tm, mats, texs)
        cpu_s = time.perf_counter() - t0
        out["matches_oracle"] = bool(np.array_equal(got.view(np.uint32), ref.view(np.uint32)))
        out["cpu_baseline"] = {"value": ntri / cpu_s, "unit": "triangles/s", "cores": 1, "kind": "port",
                               "seconds": cpu_s, "gpu_multiple": cpu_s / gpu_s,
                               "sample": "every triangle on one core"}
    return out


def voxelize_section(ctx, mesh, res=256, cpu=True, steps=3):
    """italy::voxelizeMesh (voxelize.cpp:52-107) as the reference calls it: mesh in host memory, cells in
    first-touch order back in host memory, on the 327,680-triangle mesh at res 256.  Beside it the
    oracle's single-threaded restatement (the reference runs it on one thread)."""
    from italy_bake import bake

    grid = bake.voxelizeMesh(mesh, res, ctx)
    t0 = time.perf_counter()
    for _ in range(steps):
        grid = bake.voxelizeMesh(mesh, res, ctx)
    gpu_s = (time.perf_counter() - t0) / steps
    out = {"workload": f"voxelizeMesh(327,680-triangle mesh, {res}): {len(grid.cells)} cells in first-touch order, "
                       "host mesh in, host cells out", "cells": int(len(grid.cells)), "gpu_ms": gpu_s * 1e3}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import ib_oracle

        t0 = time.perf_counter()
        cells, _ = ib_oracle.voxelize_first_touch(mesh.xyz, mesh.bmin, mesh.bmax, res)
        cpu_s = time.perf_counter() - t0
        out.update(cpu_ms=cpu_s * 1e3, cpu_cores=1, gpu_multiple=cpu_s / gpu_s,
                   matches_oracle=bool(np.array_equal(cells, grid.cells)))
    return out


def band_section(ctx, steps=3, cpu=True):
    """The precision-stress union (64 tori/boxes, 2% of triangles removed) as a 384^3 narrow band, at the
    configuration's explicit parameters (513,803 triangles) and at its stated ~1M triangles (64x64-segment tori,
    46x46-quad box faces).  cpu_baseline: the oracle's band bake (same distance, FP64 winding number by the
    closure identity) on every 8th band brick, on all usable host cores."""
    import ctypes
    from italy_bake import bake, meshgen

    def one(mesh, cfg, label):
        g = bake.dense_grid(cfg["lo"], cfg["hi"], cfg["res"])
        nbrick = (g.nx // 8) * (g.ny // 8) * (g.nz // 8)
        d_xyz = ctx.alloc(mesh.xyz.nbytes)
        ctx.upload(d_xyz, mesh.xyz)
        d_idx = ctx.alloc(4 * nbrick)
        d_pool = ctx.alloc(4 * 512 * nbrick)
        n = ctypes.c_uint64()

        def run():
            bake.check(ctx.lib.ib_bake_band_device(ctx.handle, ctypes.c_void_p(d_xyz), mesh.ntri, ctypes.byref(g),
                                                   cfg["band"], ctypes.c_void_p(d_idx), ctypes.c_void_p(d_pool),
                                                   nbrick, ctypes.byref(n), None))

        run()
        ctx.synchronize()
        times = []
        for _ in range(steps):
            t0 = time.perf_counter()
            run()
            times.append(time.perf_counter() - t0)
        st = ctx.stats()
        ctx.set_option(bake.OPT_COUNT_WORK, 1)
        run()
        cst = ctx.stats()
        ctx.set_option(bake.OPT_COUNT_WORK, 0)
        for p in (d_xyz, d_idx, d_pool):
            ctx.free(p)
        ms = 1e3 * float(np.median(times))
        exec_flops = cst["box_tests"] * FLOP_BOX + cst["tri_tests"] * FLOP_TRI + cst["obb_tests"] * FLOP_OBB
        rate = int(n.value) * 512 / (ms * 1e-3)
        out = {
            "workload": f"{label} ({mesh.ntri} triangles after 2% removed), {cfg['res']}^3 narrow band "
                        f"+-{cfg['band']} voxels in 8^3 FP32 bricks over [-1,1]^3",
            "triangles": mesh.ntri, "bricks": int(n.value), "band_voxels_computed": int(n.value) * 512,
            "ms_per_bake": ms, "ms_spread": [1e3 * min(times), 1e3 * max(times)], "voxel_queries_per_s": rate,
            "boundary_edges": st["boundary_edges"], "fp64_rechecked": st["rechecked"],
            "fp64_summed_over_every_fan": st["rechecked_full"], "fp64_on_cap": st["on_cap"], "parity_ties": st["ties"],
            "ms_hole_caps": st["ms_caps"], "ms_query_launch": st["ms_sdf"],
            "query_executed_work": {"flops": exec_flops, "TFLOP/s": exec_flops / (st["ms_sdf"] * 1e-3) / 1e12,
                                    "note": "executed FP32 work of the band's persistent query launch (counting "
                                            "build, per lane), over its event-timed duration, coarse pass excluded"},
        }
        if cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import ib_oracle

            _, _, ost = ib_oracle.bake_band(mesh.xyz, list(g.origin), g.h, g.nx, g.ny, g.nz, cfg["band"], stride=8,
                                            threads=host_cores())
            crate = ost["voxels"] / ost["seconds"]
            out["cpu_baseline"] = {
                "value": crate, "unit": "band voxel-queries/s", "cores": ost["threads"], "kind": "port",
                "gpu_multiple": rate / crate,
                "sample": f"every 8th of the {int(n.value)} band bricks ({ost['voxels']} voxels: FP32 closest triangle, "
                          f"FP64 winding number by the closure identity over {ost['cap_triangles']} cap triangles; "
                          f"brick selection and LBVH build included), {ost['seconds']:.2f} s"}
        return out

    mesh, cfg = meshgen.workload_stress()
    res = one(mesh, cfg, "64-component stress union")
    mesh1m, cfg1m = meshgen.workload_stress_1m()
    res["stress_1m"] = one(mesh1m, cfg1m, "64-component stress union at ~1M triangles")
    return res


def batched_section(ctx, n, cpu=True):
    """n meshes (log-uniform 500..200k triangles, SDF + occupancy at 32/64/128/256^3): the plain
    per-mesh bake loop against one batched bake with an autotuned persistent query launch."""
    from italy_bake import Job, bake, meshgen

    t0 = time.perf_counter()
    meshes, res = meshgen.workload_batched(n)
    gen_s = time.perf_counter() - t0
    jobs, ptrs = [], []
    for m, r in zip(meshes, res):
        g = bake.dense_grid(-0.8, 0.8, r)
        d_xyz = ctx.alloc(m.xyz.nbytes)
        ctx.upload(d_xyz, m.xyz)
        d_dist = ctx.alloc(4 * r ** 3)
        d_occ = ctx.alloc(4 * ((r + 31) // 32) * r * r)
        jobs.append(Job(d_xyz, m.ntri, g, d_dist, g, d_occ))
        ptrs += [d_xyz, d_dist, d_occ]
    voxels = sum(r ** 3 for r in res)
    tris = sum(m.ntri for m in meshes)
    # plain: one ib_bake_dense_device per mesh, as a caller without the batch API would (3 runs)
    plain_runs = []
    for _ in range(3):
        ctx.synchronize()
        t0 = time.perf_counter()
        for j in jobs:
            ctx.bake_dense_device(j.d_xyz, j.ntri, j.sdf, j.d_distances, j.occ, j.d_occ_words)
        ctx.synchronize()
        plain_runs.append(time.perf_counter() - t0)
    plain_s = float(np.median(plain_runs))
    # batched: autotune on a sample, then one batched bake of all meshes (3 runs)
    t0 = time.perf_counter()
    shape, blocks, _ = ctx.autotune_batch(jobs, sample_stride=32)
    tune_s = time.perf_counter() - t0
    batch_runs, query_ms = [], []
    for _ in range(3):
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.bake_batch_device(jobs)
        ctx.synchronize()
        batch_runs.append(time.perf_counter() - t0)
        query_ms.append(ctx.stats()["ms_sdf"])
    batch_s = float(np.median(batch_runs))
    st = ctx.stats()
    # work the persistent query executes (wave level: every lane's tests over its brick's candidates)
    ctx.set_option(bake.OPT_COUNT_WORK, 1)
    ctx.bake_batch_device(jobs)
    cst = ctx.stats()
    ctx.set_option(bake.OPT_COUNT_WORK, 0)
    exec_flops = cst["box_tests"] * FLOP_BOX + cst["tri_tests"] * FLOP_TRI + cst["obb_tests"] * FLOP_OBB
    qms = float(np.median(query_ms))
    executed = {"flops": exec_flops, "TFLOP/s": exec_flops / (qms * 1e-3) / 1e12,
                "frac_of_157_3": exec_flops / (qms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS,
                "note": "executed FP32 work of the persistent query (counting build, per lane: the brick's union of "
                        "candidates), over its median event-timed duration, columns and resolve included"}
    for p in ptrs:
        ctx.free(p)
    for opt in (bake.OPT_BRICK_SHAPE, bake.OPT_BLOCKS_PER_CU):  # the autotuner's choice applies to batches only
        ctx.set_option(opt, bake.DEFAULT_OPTIONS[opt])
    cpu_out = None
    if cpu:
        # the oracle's CPU bake (same exact math, all cores) on every 32nd mesh over about 8 z-slabs of its
        # grid; each mesh's full bake time is its measured voxel rate applied to its whole grid
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import ib_oracle

        cores = host_cores()
        t_cpu = time.perf_counter()
        est_s, sampled = 0.0, 0
        for i in range(0, n, 32):
            m, r = meshes[i], res[i]
            g = bake.dense_grid(-0.8, 0.8, r)
            _, cst = ib_oracle.bake_dense(m.xyz, list(g.origin), g.h, r, r, r, 0, r, threads=cores,
                                          zstep=max(1, r // 8))
            if cst is None:  # open mesh: the oracle's dense bake signs closed meshes only
                continue
            est_s += cst["seconds"] * r ** 3 / cst["voxels"]
            sampled += 1
        if sampled:
            cpu_s = est_s * n / sampled
            cpu_out = {"value": n / cpu_s, "unit": "meshes/s", "cores": cores, "kind": "port",
                       "gpu_multiple": cpu_s / batch_s,
                       "sample": f"every 32nd mesh ({sampled} meshes), about 8 z-slabs of each grid, each mesh's "
                                 f"rate applied to its whole grid and the mean to all {n} meshes (LBVH build "
                                 f"included per sample); {time.perf_counter() - t_cpu:.1f} s"}
    return {
        "cpu_baseline": cpu_out,
        "workload": f"{n} displaced icospheres, log-uniform 500..200k triangles ({tris} total), SDF + occupancy "
                    f"at 32/64/128/256^3 over [-0.8,0.8]^3 ({voxels} voxels)",
        "meshes": n, "triangles": tris, "voxels": voxels,
        "plain_s": plain_s, "plain_meshes_per_s": n / plain_s, "plain_voxel_queries_per_s": voxels / plain_s,
        "batched_s": batch_s, "batched_meshes_per_s": n / batch_s, "batched_voxel_queries_per_s": voxels / batch_s,
        "speedup_vs_plain": plain_s / batch_s, "autotuned": {"brick_shape": shape, "blocks_per_cu": blocks,
                                                              "tune_s": tune_s},
        "plain_runs_s": plain_runs, "batched_runs_s": batch_runs,
        "query_ms": qms, "query_executed_work": executed, "host_generation_s": gen_s,
    }


def batch_profile(shape, blocks, n=512):
    """The batched workload alone, baked twice at a fixed brick shape and block count (for profilers)."""
    from italy_bake import Job, bake, meshgen

    ctx = bake.Context(0)
    meshes, res = meshgen.workload_batched(n)
    jobs = []
    for m, r in zip(meshes, res):
        g = bake.dense_grid(-0.8, 0.8, r)
        d_xyz = ctx.alloc(m.xyz.nbytes)
        ctx.upload(d_xyz, m.xyz)
        jobs.append(Job(d_xyz, m.ntri, g, ctx.alloc(4 * r ** 3), g, ctx.alloc(4 * ((r + 31) // 32) * r * r)))
    ctx.set_option(bake.OPT_BRICK_SHAPE, shape)
    ctx.set_option(bake.OPT_BLOCKS_PER_CU, blocks)
    out = []
    for _ in range(2):
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.bake_batch_device(jobs)
        ctx.synchronize()
        st = ctx.stats()
        out.append({"wall_ms": (time.perf_counter() - t0) * 1e3, "query_ms": st["ms_sdf"], "total_ms": st["ms_total"]})
    print(json.dumps({"batch_profile": {"meshes": n, "brick_shape": shape, "blocks_per_cu": blocks, "runs": out}}))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (batched and served bakes "
              "shard by whole meshes and one large bake by brick ranges, with no communication between GPUs; "
              "only the one-GPU share is measured here)", file=sys.stderr)
        sys.exit(2)
    if args.batch_profile:
        shape, blocks = (int(v) for v in args.batch_profile.split(","))
        batch_profile(shape, blocks, args.batch_meshes or 512)
        return
    from italy_bake import bake, meshgen

    mesh, cfg = meshgen.workload_plain_large()
    sdf = bake.dense_grid(cfg["lo"], cfg["hi"], cfg["sdf_res"])
    occ = bake.dense_grid(cfg["lo"], cfg["hi"], cfg["occ_res"])
    nvox = sdf.nx * sdf.ny * sdf.nz
    ctx = bake.Context(0)
    if args.variant is not None:
        ctx.set_option(2, args.variant)
    d_xyz = ctx.alloc(mesh.xyz.nbytes)
    ctx.upload(d_xyz, mesh.xyz)
    d_dist = ctx.alloc(4 * nvox)
    d_occ = ctx.alloc(4 * ((occ.nx + 31) // 32) * occ.ny * occ.nz)

    def step():
        ctx.bake_dense_device(d_xyz, mesh.ntri, sdf, d_dist, occ, d_occ)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    sdf_ms, total_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        st = ctx.stats()  # event timings of this bake (waits for its last event)
        sdf_ms.append(st["ms_query"])  # HIP events around the query kernel's launch, on its stream
        total_ms.append(st["ms_total"])
    ctx.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = nvox / (ms_per_step * 1e-3)

    # algorithmic work of the query kernel from one counted run
    counters = ctx.query_counters(d_xyz, mesh.ntri, sdf) if hasattr(ctx, "query_counters") else None
    sdf_avg_ms = float(np.mean(sdf_ms))
    roofline = {"bound": "valu", "achieved": None, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": None,
                "traffic": None, "kernel": "k_sdf_wave", "kernel_ms": sdf_avg_ms}
    if counters:
        flops = counters["box_tests"] * FLOP_BOX + counters["tri_tests"] * FLOP_TRI
        achieved = flops / (sdf_avg_ms * 1e-3) / 1e12
        roofline.update(achieved=achieved, frac=achieved / PEAK_FP32_TFLOPS, flops_per_launch=flops,
                        box_tests=counters["box_tests"], tri_tests=counters["tri_tests"],
                        work_note="algorithmic FLOPs = box and triangle tests of a per-voxel closest-triangle "
                                  "search over a binary LBVH with one triangle per leaf (counted by a counting run), "
                                  "x FLOP_BOX and x FLOP_TRI FLOPs; the wave kernel shares candidates across a brick and culls "
                                  "by oriented boxes, so it executes fewer tests than this count")
    alg_bytes = 4 * nvox + mesh.xyz.nbytes
    roofline["hbm_algorithmic_GBs"] = alg_bytes / (sdf_avg_ms * 1e-3) / 1e9
    # HBM-side bytes per launch, from the committed FETCH_SIZE / WRITE_SIZE profile of this kernel
    # and workload (counters scaled by a known-size copy kernel); null when the profile is absent
    prof = os.path.join(ROOT, "profiles", "k_sdf_wave_plain_327k_256_traffic.json")
    if os.path.exists(prof) and args.variant in (None, 1):
        with open(prof) as f:
            tr = json.load(f)
        roofline["traffic"] = tr["traffic_bytes"]
        roofline["traffic_source"] = "profiles/" + tr["source"]
        if "valu_busy" in tr:  # executed work from the same profile's SQ counters
            roofline["executed"] = {k: tr[k] for k in ("valu_insts", "valu_insts_per_brick", "valu_busy")}

    if args.copy_calibration_mib > 0:
        nbytes = args.copy_calibration_mib << 20
        src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
        ctx.copy_device(dst, src, nbytes)
        ctx.free(src)
        ctx.free(dst)

    consumer = consumer_section(ctx, sdf, d_dist) if not args.no_consumer else None

    batched = batched_section(ctx, args.batch_meshes, cpu=not args.no_cpu_baseline) if args.batch_meshes > 0 else None
    band = band_section(ctx, cpu=not args.no_cpu_baseline) if not args.no_band else None
    analytic = analytic_section(ctx, cpu=not args.no_cpu_baseline) if not args.no_analytic else None
    voxelize = voxelize_section(ctx, mesh, cpu=not args.no_cpu_baseline) if not args.no_voxelize else None
    ingest = ingest_section(ctx, mesh, d_xyz, cpu=not args.no_cpu_baseline) if not args.no_ingest else None
    material = material_section(ctx, mesh, cpu=not args.no_cpu_baseline) if not args.no_material else None
    served = served_section(args, batched, ctx)

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import ib_oracle

        cores = host_cores()
        _, cst = ib_oracle.bake_dense(mesh.xyz, list(sdf.origin), sdf.h, sdf.nx, sdf.ny, sdf.nz, 0, sdf.nz,
                                      threads=cores, zstep=args.cpu_zstep)
        cpu = {"value": cst["voxels"] / cst["seconds"], "unit": "voxel-queries/s", "cores": cst["threads"],
               "kind": "port",
               "gpu_multiple": value / (cst["voxels"] / cst["seconds"]),
               "sample": f"every {args.cpu_zstep}th z-slab of the same 256^3 bake ({cst['voxels']} voxels, exact "
                         f"closest triangle + crossing sign, LBVH build included), {cst['seconds']:.2f} s"}
        # one thread, as the reference's CPU code runs
        _, c1 = ib_oracle.bake_dense(mesh.xyz, list(sdf.origin), sdf.h, sdf.nx, sdf.ny, sdf.nz, 0, sdf.nz,
                                     threads=1, zstep=args.cpu_zstep1)
        cpu["single_thread"] = {"value": c1["voxels"] / c1["seconds"], "cores": 1,
                                "gpu_multiple": value / (c1["voxels"] / c1["seconds"]),
                                "sample": f"every {args.cpu_zstep1}th z-slab ({c1['voxels']} voxels), "
                                          f"{c1['seconds']:.2f} s"}

    out = {
        "metric": METRIC, "value": value, "unit": "voxel-queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain single-mesh bake: 327,680-triangle displaced icosphere (level 7, seed "
                               "0x17A1E5), 256^3 FP32 SDF + 512^3 bit-packed occupancy over [-0.8,0.8]^3",
                   "triangles": mesh.ntri, "sdf_res": sdf.nx, "occ_res": occ.nx, "parallelism": "single GPU"},
        "device_ms_per_bake": float(np.mean(total_ms)),
        "roofline": roofline, "cpu_baseline": cpu, "batched": batched, "narrow_band": band, "served": served,
        "grid_consumer": consumer, "analytic_bake": analytic, "voxelize_mesh": voxelize,
        "ingest_gate": ingest, "material_splat": material,
        "p99_bake_latency_s": served["batched_distribution"]["latency_s"]["p99"]
        if served and served.get("batched_distribution") else None,
        "p99_bake_latency_s_stress_mix": served["stress_mix"]["latency_s"]["p99"]
        if served and served.get("stress_mix") else None,
        "host": {"cpus_usable": host_cores(), "cpus_affinity": len(os.sched_getaffinity(0)),
                 "cpu_count": os.cpu_count()},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
