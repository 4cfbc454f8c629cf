"""GPU parity: the gfx950 bake against the CPU oracle, through the C ABI.

Bar (per workload): occupancy bit-exact; sign bit-exact on every voxel with |d| > 1e-4 * extent;
|distance - oracle| <= 1e-5 * extent.  Full-size grids are checked by size-independent properties
(1-Lipschitz distances, sign changes only across the surface) plus oracle z-slab samples.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check_sdf(dist, ref, extent, what):
    tol_d = 1e-5 * extent
    far = np.abs(ref) > 1e-4 * extent
    err = np.abs(dist - ref)
    assert np.isfinite(dist).all(), what
    assert err.max() <= tol_d, f"{what}: max |d - oracle| = {err.max():.3g} > {tol_d:.3g}"
    bad = int(((dist < 0) != (ref < 0))[far].sum())
    assert bad == 0, f"{what}: {bad} sign mismatches away from the surface"
    return float(err.max())


def test_plain_small_parity(ctx, oracle, meshgen):
    """5,120-triangle jittered icosphere, 64^3 SDF + occupancy over [-0.6, 0.6]^3."""
    from italy_bake import bake

    mesh, cfg = meshgen.workload_plain_small()
    g = bake.dense_grid(cfg["lo"], cfg["hi"], cfg["sdf_res"])
    dist, occ = ctx.bake_dense(mesh.xyz, g, g)
    ref, _ = oracle.bake_dense(mesh.xyz, list(g.origin), g.h, g.nx, g.ny, g.nz)
    _check_sdf(dist, ref, cfg["hi"] - cfg["lo"], "plain small")
    occ_ref = oracle.occupancy(mesh.xyz, list(g.origin), g.h, g.nx, g.ny, g.nz)
    assert np.array_equal(occ, occ_ref)
    st = ctx.stats()
    assert st["boundary_edges"] == 0 and st["column_overflows"] == 0


def test_material_splat_matches_oracle(ctx, oracle, meshgen):
    """bakePerCellMaterial (sdf_baker.cpp:76-128) on the device: every plane bit-exact against the
    oracle, with random per-triangle materials on the bakeSdf grid of the 5,120-triangle icosphere;
    the drop-in bakeSdf fills the same planes (default material without a callback)."""
    from italy_bake import bake

    mesh, _ = meshgen.workload_plain_small()
    rng = np.random.default_rng(11)
    mat = rng.uniform(0, 1, (mesh.ntri, 5)).astype(np.float32)
    g = bake.sdf_framing(mesh.bmin, mesh.bmax, 48)
    rgb, met, rough = ctx.splat_materials(mesh.xyz, mat, g)
    ref = oracle.splat_materials(mesh.xyz, mat, list(g.origin), g.h, g.nx, g.ny, g.nz).reshape(5, -1)
    got = np.stack([rgb[:, 0], rgb[:, 1], rgb[:, 2], met, rough])
    assert (ref[3] != 0).sum() > 1000  # touched cells
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    grid = bake.bakeSdf(mesh, 48, ctx, materials=mat)
    assert np.array_equal(grid.metallic.view(np.uint32), ref[3].view(np.uint32))
    plain = bake.bakeSdf(mesh, 48, ctx)
    assert np.array_equal(plain.baseColor, np.ones_like(plain.baseColor))
    assert (plain.metallic == 0).all() and (plain.roughness == 1).all()


def test_reference_sdf_sign_test(ctx, meshgen):
    """tests/sdf_sign_test.cpp:46-82 through the bakeSdf drop-in: UV sphere r=1, 32x16, res 24;
    no sign errors beyond 1.5h, max |error| < 3h."""
    from italy_bake import bake

    sphere = meshgen.uv_sphere(1.0, 32, 16)
    grid = bake.bakeSdf(sphere, 24, ctx)
    assert grid.distances.size == grid.nx * grid.ny * grid.nz > 0
    tol = 1.5 * float(grid.voxelSize)
    import ib_oracle

    c = ib_oracle.cell_centres(grid.origin, grid.voxelSize, grid.nx, grid.ny, grid.nz).reshape(-1, 3)
    analytic = np.linalg.norm(c.astype(np.float64), axis=1) - 1.0
    chk = np.abs(analytic) > tol
    assert chk.sum() > 0
    assert int(((analytic < 0) != (grid.distances < 0))[chk].sum()) == 0
    assert np.abs(grid.distances - analytic)[chk].max() < 2 * tol
    st = ctx.stats()
    assert st["boundary_edges"] == 64  # the sphere's float seam and south pole leave hairline holes


def test_reference_sdf_sign_test_exact(ctx, oracle, meshgen):
    """Same mesh against the FP64 oracle: distances to 1e-5 * extent, sign = |GWN| > 0.5."""
    from italy_bake import bake

    sphere = meshgen.uv_sphere(1.0, 32, 16)
    grid = bake.bakeSdf(sphere, 24, ctx)
    c = oracle.cell_centres(grid.origin, grid.voxelSize, grid.nx, grid.ny, grid.nz).reshape(-1, 3)
    d64, w = oracle.sdf_f64_points(sphere.xyz, c.astype(np.float64))
    extent = 2.0
    assert np.abs(grid.distances - d64).max() <= 1e-5 * extent
    far = np.abs(d64) > 1e-4 * extent
    assert np.array_equal((grid.distances < 0)[far], (d64 < 0)[far])


def test_reference_voxelize_test(ctx, oracle, meshgen):
    """tests/voxelize_test.cpp:37-58 through the voxelizeMesh drop-in, plus first-touch order."""
    from italy_bake import bake

    cube = meshgen.cube_shell()
    vg = bake.voxelizeMesh(cube, 8, ctx)
    cells = vg.cells
    assert 0 < len(cells) <= 6 * 8 * 8
    s = {tuple(x) for x in cells.tolist()}
    assert (0, 0, 0) in s and (4, 4, 4) not in s
    ref, _ = oracle.voxelize_first_touch(cube.xyz, cube.bmin, cube.bmax, 8)
    assert np.array_equal(cells, ref)


@pytest.mark.parametrize("res", [1, 7, 32, 64])
def test_voxelize_first_touch_order(ctx, oracle, meshgen, res):
    from italy_bake import bake

    m = meshgen.icosphere(3, 0.5, 1e-3, 77)
    vg = bake.voxelizeMesh(m, res, ctx)
    ref, _ = oracle.voxelize_first_touch(m.xyz, m.bmin, m.bmax, res)
    assert np.array_equal(vg.cells, ref)


@pytest.mark.parametrize("case", ["sphere64", "coarse256", "displaced128"])
def test_voxelize_colours_against_oracle(ctx, oracle, meshgen, case):
    """Cells in first-touch order and each cell's colour, the mean of the colours of the triangles that
    overlap it summed in triangle order (voxelizer.cpp:70-104), bit for bit.  The coarse case (80 triangles
    on a 256 grid, ~20 cells a triangle) outgrows the first capacity guess and takes the exact-count retry."""
    from italy_bake import bake

    m = {"sphere64": lambda: meshgen.icosphere(3, 0.5, 1e-3, 77),
         "coarse256": lambda: meshgen.icosphere(1, 0.6, 0.0, 5),
         "displaced128": lambda: meshgen.displaced(5, 9)}[case]()
    res = {"sphere64": 64, "coarse256": 256, "displaced128": 128}[case]
    rng = np.random.default_rng(len(case))
    colours = rng.random((m.ntri, 3), dtype=np.float32)
    vg = bake.voxelizeMesh(m, res, ctx, tri_colors=colours)
    ref, rcol = oracle.voxelize_first_touch(m.xyz, m.bmin, m.bmax, res, colours)
    assert len(ref) > 0
    assert np.array_equal(vg.cells, ref)
    assert np.array_equal(vg.colors.view(np.uint32), rcol.view(np.uint32))
    if case == "coarse256":
        assert len(ref) > 4 * m.ntri  # beyond the first guess of the capacity
    plain = bake.voxelizeMesh(m, res, ctx)  # no colours: the same cells, white
    assert np.array_equal(plain.cells, ref)


def test_plain_large_occupancy_bitexact(ctx, oracle, meshgen):
    """327,680-triangle displaced icosphere, 512^3 occupancy over [-0.8, 0.8]^3: bit-exact."""
    from italy_bake import bake

    mesh, cfg = meshgen.workload_plain_large()
    og = bake.dense_grid(cfg["lo"], cfg["hi"], cfg["occ_res"])
    _, occ = ctx.bake_dense(mesh.xyz, None, og)
    ref = oracle.occupancy(mesh.xyz, list(og.origin), og.h, og.nx, og.ny, og.nz)
    assert np.array_equal(occ, ref)
    assert int(np.unpackbits(occ.view(np.uint8)).sum()) > 300_000  # ~ surface area / h^2


def test_plain_large_sdf(ctx, oracle, meshgen):
    """256^3 SDF of the large workload: oracle z-slabs + full-grid properties."""
    from italy_bake import bake

    mesh, cfg = meshgen.workload_plain_large()
    g = bake.dense_grid(cfg["lo"], cfg["hi"], cfg["sdf_res"])
    dist, _ = ctx.bake_dense(mesh.xyz, g, None)
    extent = cfg["hi"] - cfg["lo"]
    for z0 in (3, 128, 200):
        ref, _ = oracle.bake_dense(mesh.xyz, list(g.origin), g.h, g.nx, g.ny, g.nz, z0, z0 + 1)
        _check_sdf(dist[z0], ref[z0], extent, f"slab {z0}")
    h = float(g.h)
    for ax in range(3):  # 1-Lipschitz between neighbouring cell centres
        dd = np.abs(np.diff(dist, axis=ax))
        assert dd.max() <= h * (1 + 1e-4) + 1e-6
    for ax in range(3):  # a sign change needs the surface between the two centres
        a = np.moveaxis(dist, ax, 0)
        flip = (a[1:] < 0) != (a[:-1] < 0)
        assert np.all(np.abs(a[1:])[flip] + np.abs(a[:-1])[flip] <= h * (1 + 1e-4) + 1e-6)
    inside = (dist < 0).mean()
    assert 0.02 < inside < 0.2


def test_deterministic(ctx, meshgen):
    from italy_bake import bake

    mesh, cfg = meshgen.workload_plain_small()
    g = bake.dense_grid(cfg["lo"], cfg["hi"], 48)
    a, oa = ctx.bake_dense(mesh.xyz, g, g)
    b, ob = ctx.bake_dense(mesh.xyz, g, g)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)) and np.array_equal(oa, ob)


def test_ragged_grid_and_degenerates(ctx, oracle, meshgen):
    """Grid dims that are not multiples of the 8x8x4 block, with zero-area and sliver triangles."""
    from italy_bake import Grid

    m = meshgen.icosphere(2, 0.5, 0.0)
    extra = np.array([[[0.1, 0.1, 0.1], [0.2, 0.2, 0.2], [0.3, 0.3, 0.3]],      # collinear
                      [[0.0, -0.3, 0.2]] * 3,                                   # a point
                      [[0.7, 0.0, 0.0], [0.9, 0.0, 0.0], [0.8, 1e-7, 0.0]]],   # sliver
                     np.float32)
    xyz = np.concatenate([m.xyz, extra])
    g = Grid.make((-0.71, -0.63, -0.57), 0.0913, 13, 7, 5)
    dist, occ = ctx.bake_dense(xyz, g, g)
    pts = oracle.cell_centres(list(g.origin), g.h, g.nx, g.ny, g.nz).reshape(-1, 3)
    ud = np.sqrt(oracle.udf2_f32_brute(xyz, pts)).reshape(dist.shape)
    assert np.abs(np.abs(dist) - ud).max() <= 2e-5
    occ_ref = oracle.occupancy(xyz, list(g.origin), g.h, g.nx, g.ny, g.nz)
    assert np.array_equal(occ, occ_ref)


def test_open_mesh_sign_matches_fp64_winding(ctx, oracle, meshgen):
    """Seeded torus + box with 2% of triangles removed: sign = |GWN| > 0.5 (FP64 brute force)."""
    from italy_bake import bake

    m = meshgen.stress(seed=5, ncomp=2, remove_frac=0.02)
    lo = float(m.bmin.min()) - 0.05
    hi = float(m.bmax.max()) + 0.05
    g = bake.dense_grid(lo, hi, 20)
    dist, _ = ctx.bake_dense(m.xyz, g, None)
    st = ctx.stats()
    assert st["boundary_edges"] > 0
    pts = oracle.cell_centres(list(g.origin), g.h, g.nx, g.ny, g.nz).reshape(-1, 3)
    d64, w = oracle.sdf_f64_points(m.xyz, pts.astype(np.float64))
    extent = hi - lo
    assert np.abs(dist.reshape(-1) - d64).max() <= 1e-5 * extent
    decided = (np.abs(d64) > 1e-4 * extent) & (np.abs(np.abs(w) - 0.5) > 1e-6)
    assert np.array_equal((dist.reshape(-1) < 0)[decided], (np.abs(w) > 0.5)[decided])


def test_empty_mesh(ctx):
    from italy_bake import bake

    g = bake.dense_grid(-1, 1, 8)
    dist, occ = ctx.bake_dense(np.zeros((0, 3, 3), np.float32), g, g)
    assert np.isinf(dist).all() and (dist > 0).all()
    assert not occ.any()


def test_non_finite_mesh_rejected(ctx, meshgen):
    from italy_bake import BakeError, bake

    m = meshgen.icosphere(1, 0.5)
    xyz = m.xyz.copy()
    xyz[3, 1, 2] = np.nan
    with pytest.raises(BakeError) as e:
        ctx.bake_dense(xyz, bake.dense_grid(-1, 1, 8), None)
    assert e.value.status == 2


def test_budget_refusal(meshgen):
    from italy_bake import BakeError, bake

    with bake.Context(0, hbm_budget_bytes=1 << 20) as small:
        m = meshgen.icosphere(5, 0.5)
        with pytest.raises(BakeError) as e:
            small.bake_dense(m.xyz, bake.dense_grid(-1, 1, 64), None)
        assert e.value.status == 3
