"""Error of the GPU bake against the FP64 oracle, and every voxel of the large bakes against the CPU oracle.

- 5,120-triangle 64^3 bake: every voxel against the brute-force FP64 distance and winding number.
- 327,680-triangle 256^3 bake: all 16,777,216 voxels bit for bit against the oracle's dense bake (whose sign is
  the exact winding number of the closed mesh, its signed +z crossing count); 1,048,576 seeded voxels against
  the FP64 distance through the oracle's LBVH (equal to the brute force bit for bit); 131,072 of them (102,400
  within 3 voxels of the surface) against the FP64 winding number.
- Precision-stress narrow bands (513,803 and 1,053,195 triangles, 384^3, +-4 voxels): every band voxel (every
  4th brick for the larger) bit for bit against the oracle's band bake (FP64 winding number by the closure
  identity), 1,048,576 voxels against the FP64 distance, 131,072 (65,536) against the FP64 winding number,
  102,400 (49,152) of them within 3 voxels of the surface.
The FP64 winding numbers of the large samples are the oracle's orc_gwn_f64_points: a first-order tree walk with
a rigorous bound where it decides |w| > 1/2 and the tie band as the sum over every triangle does, that sum
elsewhere (pinned against the brute force in test_oracle.py).
Max and RMS |d - d64| (bound 1e-5 * extent) and sign agreement on every checked voxel with |d64| > 1e-4 * extent
whose |w64| is not within 1e-6 of 0.5.  The numbers are written as JSON to the file the environment variable
IB_ACCURACY_JSON names, when it is set."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
REPORT = {}


def _write():
    path = os.environ.get("IB_ACCURACY_JSON")
    if path:
        with open(path, "w") as f:
            json.dump(REPORT, f, indent=1)


def _record(name, dist, pts, xyz, extent, oracle, sign_sample=None, h=None):
    """dist: GPU signed distances at pts.  sign_sample: indices whose sign is checked against the FP64
    winding number (None: every point, with brute-force FP64 distances too); h: the voxel size (the count of
    checked signs within 3 voxels of the surface is reported)."""
    pts = pts.astype(np.float64)
    if sign_sample is None:
        d64s, w = oracle.sdf_f64_points(xyz, pts)
        d64 = np.abs(d64s)
        sign_sample = np.arange(len(pts))
    else:
        d64 = oracle.udf_points_f64(xyz, pts)
        w, exact = oracle.gwn_f64_points(xyz, pts[sign_sample])
    err = np.abs(dist.astype(np.float64)) - d64
    ds = d64[sign_sample]
    decided = (ds > 1e-4 * extent) & (np.abs(np.abs(w) - 0.5) > 1e-6)
    sign_bad = int(((dist[sign_sample] < 0) != (np.abs(w) > 0.5))[decided].sum())
    REPORT[name] = {"samples": int(len(pts)), "max_abs_err": float(np.abs(err).max()),
                    "rms_err": float(np.sqrt(np.mean(err ** 2))), "max_err_over_extent": float(np.abs(err).max() / extent),
                    "sign_mismatches": sign_bad, "sign_checked": int(decided.sum()),
                    "sign_checked_near_surface": int((decided & (ds < 3 * (h or 0.0))).sum())}
    _write()
    assert np.abs(err).max() <= 1e-5 * extent
    assert sign_bad == 0


def _sample(g, n, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, [g.nz, g.ny, g.nx], size=(n, 3))


def test_accuracy_plain_small(ctx, oracle, meshgen):
    from italy_bake import bake

    m, cfg = meshgen.workload_plain_small()
    g = bake.dense_grid(cfg["lo"], cfg["hi"], cfg["sdf_res"])
    dist, _ = ctx.bake_dense(m.xyz, g, None)
    pts = oracle.cell_centres(list(g.origin), g.h, g.nx, g.ny, g.nz).reshape(-1, 3)
    _record("plain_small_64_all_voxels", dist.reshape(-1), pts, m.xyz, cfg["hi"] - cfg["lo"], oracle)


def _near_half(zyx, near, seed):
    """replace the first half of a sample by voxels drawn from `near`"""
    half = len(zyx) // 2
    zyx[:half] = near[np.random.default_rng(seed).choice(len(near), half, replace=len(near) < half)]
    return zyx


def test_accuracy_plain_large(ctx, oracle, meshgen):
    from italy_bake import bake

    m, cfg = meshgen.workload_plain_large()
    g = bake.dense_grid(cfg["lo"], cfg["hi"], cfg["sdf_res"])
    dist, _ = ctx.bake_dense(m.xyz, g, None)
    zyx = _sample(g, 1 << 20, 2)
    zyx = _near_half(zyx, np.argwhere(np.abs(dist) < 3 * g.h), 3)
    pts = oracle.cell_centres(list(g.origin), g.h, g.nx, g.ny, g.nz)[zyx[:, 0], zyx[:, 1], zyx[:, 2]]
    signs = np.concatenate([np.arange(102400), (1 << 19) + np.arange(28672)])  # near the surface, and anywhere
    _record("plain_large_256", dist[zyx[:, 0], zyx[:, 1], zyx[:, 2]], pts, m.xyz, cfg["hi"] - cfg["lo"], oracle,
            sign_sample=signs, h=g.h)


def test_plain_large_every_voxel_equals_oracle(ctx, oracle, meshgen):
    """All 16.8M voxels of the 327,680-triangle 256^3 bake, bit for bit, against the oracle's dense bake."""
    from italy_bake import bake

    m, cfg = meshgen.workload_plain_large()
    g = bake.dense_grid(cfg["lo"], cfg["hi"], cfg["sdf_res"])
    dist, _ = ctx.bake_dense(m.xyz, g, None)
    ref, st = oracle.bake_dense(m.xyz, list(g.origin), g.h, g.nx, g.ny, g.nz)
    bad = int(np.count_nonzero(dist.view(np.uint32) != ref.view(np.uint32)))
    near = np.abs(ref) < 3 * g.h
    REPORT["plain_large_256_every_voxel"] = {"voxels": int(ref.size), "bit_mismatches": bad,
                                             "near_surface_voxels": int(near.sum()), "inside": int((ref < 0).sum()),
                                             "oracle_seconds": st["seconds"], "oracle_threads": st["threads"]}
    _write()
    assert bad == 0


def _band_check(ctx, oracle, name, m, cfg, sign_n, stride):
    from italy_bake import bake

    g = bake.dense_grid(cfg["lo"], cfg["hi"], cfg["res"])
    index, pool = ctx.bake_band(m.xyz, g, cfg["band"])
    st = ctx.stats()
    slots = np.argwhere(index >= 0)
    rng = np.random.default_rng(4)
    n = 1 << 20
    pick = rng.choice(len(slots), n)
    loc = rng.integers(0, 8, size=(n, 3))
    vals = pool[index[tuple(slots[pick].T)], loc[:, 0], loc[:, 1], loc[:, 2]]
    # half the sample within 3 voxels of the surface
    near = np.argwhere(np.abs(pool) < 3 * g.h)
    sel = near[rng.choice(len(near), n // 2)]
    pick[: n // 2] = -1
    zyx = slots[np.maximum(pick, 0)] * 8 + loc
    brick_of_slot = np.full(len(pool), -1, np.int64)
    brick_of_slot[index[tuple(slots.T)]] = np.arange(len(slots))
    zyx[: n // 2] = slots[brick_of_slot[sel[:, 0]]] * 8 + sel[:, 1:]
    vals[: n // 2] = pool[sel[:, 0], sel[:, 1], sel[:, 2], sel[:, 3]]
    pts = oracle.cell_centres(list(g.origin), g.h, g.nx, g.ny, g.nz)[zyx[:, 0], zyx[:, 1], zyx[:, 2]]
    near_n = sign_n * 25 // 32
    signs = np.concatenate([np.arange(near_n), (n // 2) + np.arange(sign_n - near_n)])
    _record(name, vals, pts, m.xyz, cfg["hi"] - cfg["lo"], oracle, sign_sample=signs, h=g.h)
    # every (stride-th) band brick against the oracle's band bake
    oi, op, ost = oracle.bake_band(m.xyz, list(g.origin), g.h, g.nx, g.ny, g.nz, cfg["band"], stride=stride)
    assert np.array_equal(oi, index)
    baked = np.arange(0, len(pool), stride)
    bad = int(np.count_nonzero(pool[baked].view(np.uint32) != op[baked].view(np.uint32)))
    REPORT[name].update(bricks=int((index >= 0).sum()), fp64_rechecked=int(st["rechecked"]),
                        parity_ties=int(st["ties"]), boundary_edges=int(st["boundary_edges"]),
                        triangles=int(m.ntri), oracle_band_voxels=int(ost["voxels"]), oracle_band_mismatches=bad,
                        oracle_band_seconds=ost["seconds"], oracle_band_threads=ost["threads"])
    _write()
    assert bad == 0


def test_accuracy_stress_band(ctx, oracle, meshgen):
    m, cfg = meshgen.workload_stress()
    _band_check(ctx, oracle, "stress_band_384", m, cfg, 131072, 1)


def test_accuracy_stress_band_1m(ctx, oracle, meshgen):
    """The precision-stress union at its stated ~1M triangles (64x64-segment tori, 46x46-quad box faces)."""
    m, cfg = meshgen.workload_stress_1m()
    _band_check(ctx, oracle, "stress_band_384_1m", m, cfg, 65536, 4)
