"""The committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from FP64
brute force, exact clipping and the sdBox closed form) against the oracle (CPU) and the device bake
(GPU, through the C ABI).

Bar: first-touch cells equal; |distance - FP64| <= 1e-5 * extent; sign equal wherever the FP64
distance exceeds 1e-4 * extent.  The fixtures keep their own grids; the library's framing is checked
against them.
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name)) as z:
        return {k: z[k] for k in z.files}


def check_signed(dist, fx_d, fx_w, extent, what):
    ref = np.where(np.abs(fx_w) > 0.5, -fx_d, fx_d)
    err = np.abs(dist.astype(np.float64) - ref)
    assert err.max() <= 1e-5 * extent, f"{what}: max |d - FP64| = {err.max():.3g}"
    far = fx_d > 1e-4 * extent
    bad = int(((dist < 0) != (ref < 0))[far].sum())
    assert bad == 0, f"{what}: {bad} sign mismatches"


def grid_of(fx):
    return list(map(float, fx["sdf_origin"])), float(fx["sdf_h"]), [int(v) for v in fx["sdf_n"]]


# ---- oracle (CPU) --------------------------------------------------------------------------------
def test_golden_cube_shell_oracle(oracle):
    fx = load("cube_shell.npz")
    cells, _ = oracle.voxelize_first_touch(fx["xyz"], fx["bmin"], fx["bmax"], int(fx["vox_res"]))
    assert np.array_equal(cells, fx["vox_cells"])
    origin, h, n = grid_of(fx)
    o2, h2, n2 = oracle.sdf_framing(fx["bmin"], fx["bmax"], int(fx["sdf_res"]))
    assert np.array_equal(np.float32(o2), np.float32(origin)) and np.float32(h2) == np.float32(h)
    assert list(n2) == n
    dist, _ = oracle.bake_dense(fx["xyz"], origin, h, *n)
    ref = fx["sdf_f64"]
    check_signed(dist.reshape(-1), np.abs(ref), np.where(ref < 0, 1.0, 0.0), 1.0, "cube shell")


def test_golden_uv_sphere_oracle(oracle):
    """The UV sphere is open (its FP32 seam and pole leave hairline gaps), so the oracle's legs for it
    are the FP32 LBVH distance and the FP64 winding number."""
    fx = load("uv_sphere.npz")
    origin, h, n = grid_of(fx)
    pts = oracle.cell_centres(origin, h, *n).reshape(-1, 3)
    d32 = oracle.udf_points(fx["xyz"], pts)
    _, w = oracle.sdf_f64_points(fx["xyz"], pts.astype(np.float64))
    check_signed(np.where(np.abs(w) > 0.5, -d32, d32), fx["dist_f64"], fx["winding_f64"], 2.0, "uv sphere")
    assert np.abs(w - fx["winding_f64"]).max() < 1e-9


def test_golden_icosphere_oracle(oracle):
    fx = load("icosphere_5120.npz")
    origin, h, n = grid_of(fx)
    dist, _ = oracle.bake_dense(fx["xyz"], origin, h, *n)
    check_signed(dist.reshape(-1)[fx["voxel_index"]], fx["dist_f64"], fx["winding_f64"], 1.2, "icosphere")


def test_golden_fixture_inputs_match_generators(meshgen):
    """the generators still produce the fixtures' inputs (so the GPU tests of other files bake the same meshes)"""
    assert np.array_equal(load("cube_shell.npz")["xyz"], meshgen.cube_shell().xyz.reshape(-1, 9))
    assert np.array_equal(load("uv_sphere.npz")["xyz"], meshgen.uv_sphere().xyz.reshape(-1, 9))
    assert np.array_equal(load("icosphere_5120.npz")["xyz"],
                          meshgen.icosphere(4, 0.5, 1e-3, 0x17A1E5).xyz.reshape(-1, 9))


# ---- device (GPU) --------------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("name,extent", [("cube_shell.npz", 1.0), ("uv_sphere.npz", 2.0),
                                         ("icosphere_5120.npz", 1.2)])
def test_golden_device_bake(ctx, name, extent):
    from italy_bake import bake

    fx = load(name)
    origin, h, n = grid_of(fx)
    print(f"golden {name}: grid {n}", flush=True)
    g = bake.Grid.make(origin, h, *n)
    dist, _ = ctx.bake_dense(fx["xyz"], g, None)
    dist = dist.reshape(-1)
    if name == "cube_shell.npz":
        ref = fx["sdf_f64"]
        check_signed(dist, np.abs(ref), np.where(ref < 0, 1.0, 0.0), extent, name)
    elif name == "uv_sphere.npz":
        check_signed(dist, fx["dist_f64"], fx["winding_f64"], extent, name)
    else:
        check_signed(dist[fx["voxel_index"]], fx["dist_f64"], fx["winding_f64"], extent, name)


@pytest.mark.gpu
def test_golden_device_framing_and_voxelize(ctx):
    from italy_bake import bake

    fx = load("cube_shell.npz")

    class M:
        xyz = fx["xyz"].reshape(-1, 3, 3)
        bmin, bmax = fx["bmin"], fx["bmax"]
        ntri = len(fx["xyz"])

    vg = bake.voxelizeMesh(M, int(fx["vox_res"]), ctx=ctx)
    assert np.array_equal(vg.cells, fx["vox_cells"])
    g = bake.sdf_framing(fx["bmin"], fx["bmax"], int(fx["sdf_res"]))
    origin, h, n = grid_of(fx)
    assert [g.nx, g.ny, g.nz] == n and np.float32(g.h) == np.float32(h)
    assert np.array_equal(np.float32(list(g.origin)), np.float32(origin))
