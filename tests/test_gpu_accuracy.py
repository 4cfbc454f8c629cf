"""Error of the GPU bake against the FP64 oracle (brute-force distance + winding number): every voxel
of the 5,120-triangle 64^3 bake, seeded samples of the larger workloads.  Max and RMS |d - d64|
(bound 1e-5 * extent) and sign agreement on every checked voxel with |d64| > 1e-4 * extent whose
|w64| is not within 1e-6 of 0.5.  The numbers are written to gpurun_out/accuracy.json when that
folder exists."""
import json
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
REPORT = {}


def _record(name, dist, pts, xyz, extent, oracle):
    d64, w = oracle.sdf_f64_points(xyz, pts.astype(np.float64))
    err = dist.astype(np.float64) - d64
    decided = (np.abs(d64) > 1e-4 * extent) & (np.abs(np.abs(w) - 0.5) > 1e-6)
    sign_bad = int(((dist < 0) != (np.abs(w) > 0.5))[decided].sum())
    REPORT[name] = {"samples": int(len(pts)), "max_abs_err": float(np.abs(err).max()),
                    "rms_err": float(np.sqrt(np.mean(err ** 2))), "max_err_over_extent": float(np.abs(err).max() / extent),
                    "sign_mismatches": sign_bad, "sign_checked": int(decided.sum())}
    out = os.path.join(ROOT, "gpurun_out")
    if os.path.isdir(out):
        with open(os.path.join(out, "accuracy.json"), "w") as f:
            json.dump(REPORT, f, indent=1)
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


def test_accuracy_plain_large(ctx, oracle, meshgen):
    from italy_bake import bake

    m, cfg = meshgen.workload_plain_large()
    g = bake.dense_grid(cfg["lo"], cfg["hi"], cfg["sdf_res"])
    dist, _ = ctx.bake_dense(m.xyz, g, None)
    zyx = _sample(g, 1024, 2)
    # half the sample within 3 voxels of the surface, where sign and distance are hardest
    near = np.argwhere(np.abs(dist) < 3 * g.h)
    zyx[:512] = near[np.random.default_rng(3).choice(len(near), 512, replace=False)]
    pts = oracle.cell_centres(list(g.origin), g.h, g.nx, g.ny, g.nz)[zyx[:, 0], zyx[:, 1], zyx[:, 2]]
    _record("plain_large_256", dist[zyx[:, 0], zyx[:, 1], zyx[:, 2]], pts, m.xyz, cfg["hi"] - cfg["lo"], oracle)


def test_accuracy_stress_band(ctx, oracle, meshgen):
    from italy_bake import bake

    m, cfg = meshgen.workload_stress()
    g = bake.dense_grid(cfg["lo"], cfg["hi"], cfg["res"])
    index, pool = ctx.bake_band(m.xyz, g, cfg["band"])
    st = ctx.stats()
    slots = np.argwhere(index >= 0)
    rng = np.random.default_rng(4)
    pick = rng.choice(len(slots), 1024)
    loc = rng.integers(0, 8, size=(1024, 3))
    zyx = slots[pick] * 8 + loc
    vals = pool[index[tuple(slots[pick].T)], loc[:, 0], loc[:, 1], loc[:, 2]]
    pts = oracle.cell_centres(list(g.origin), g.h, g.nx, g.ny, g.nz)[zyx[:, 0], zyx[:, 1], zyx[:, 2]]
    _record("stress_band_384", vals, pts, m.xyz, cfg["hi"] - cfg["lo"], oracle)
    REPORT["stress_band_384"].update(bricks=int((index >= 0).sum()), fp64_rechecked=int(st["rechecked"]),
                                     parity_ties=int(st["ties"]), boundary_edges=int(st["boundary_edges"]))
    out = os.path.join(ROOT, "gpurun_out")
    if os.path.isdir(out):
        with open(os.path.join(out, "accuracy.json"), "w") as f:
            json.dump(REPORT, f, indent=1)
