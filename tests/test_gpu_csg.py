"""Analytic SDF bake on the device (src/convert/sdf_procedural.cpp) against the oracle: distances and
branch bytes bit-exact on every cell, for the reference's three scenes and for a program that uses
every primitive and operator; malformed programs are rejected before any launch."""
import numpy as np
import pytest

from italy_bake import procedural as pr
from italy_bake._lib import BakeError

pytestmark = pytest.mark.gpu


def _oracle_bake(oracle, prog, g):
    return oracle.bake_csg(prog.array() if isinstance(prog, pr.Csg) else prog, np.array(g.origin[:], np.float32),
                           g.h, g.nx, g.ny, g.nz, threads=8)


@pytest.mark.parametrize("res", [96, 320])
def test_glass_cup_matches_oracle(ctx, oracle, res):
    grid = pr.makeGlassCupSdf(res, -1.0, (0.0, 0.0), ctx=ctx)
    o, h, dist, branch = oracle.cup_cells(True, res, -1.0, (0.0, 0.0))
    assert (grid.nx, grid.ny, grid.nz) == dist.shape[::-1] and np.array_equal(grid.origin, o) and grid.voxelSize == h
    assert np.array_equal(grid.distances, dist.reshape(-1))
    assert np.array_equal(grid.branch, branch.reshape(-1))
    assert [p.ior for p in grid.palette] == [1.47, 1.333]


def test_cup_matches_oracle(ctx, oracle):
    grid = pr.makeCupSdf(96, -1.0, (-1.6, 1.6), ctx=ctx)
    _, _, dist, _ = oracle.cup_cells(False, 96, -1.0, (-1.6, 1.6))
    assert np.array_equal(grid.distances, dist.reshape(-1))
    assert np.allclose(grid.baseColor, [0.85, 0.83, 0.78]) and np.all(grid.roughness == np.float32(0.35))


def test_material_probes_match_oracle(ctx, oracle):
    grids = pr.makeMaterialProbeSdfs(48, ctx=ctx)
    assert len(grids) == 9
    centres = [(x, 0.4, -0.6) for x in (-2.2, -1.1, 0.0, 1.1, 2.2)] + [(x, 0.4, 0.6) for x in (-1.65, -0.55, 0.55, 1.65)]
    for i, grid in enumerate(grids):
        prog = pr.Csg().sphere(centres[i], 0.4)
        ref, _ = oracle.bake_csg(prog.array(), grid.origin, grid.voxelSize, grid.nx, grid.ny, grid.nz)
        assert np.array_equal(grid.distances, ref.reshape(-1)), i
    assert [g.palette[0].roughness for g in grids[:5]] == [0.05, 0.3, 0.55, 0.8, 1.0]


def _every_op_program():
    return (pr.Csg()
            .box((0.1, 0.0, -0.1), (0.5, 0.3, 0.4), branch=1)
            .cylinder((0.0, 0.1, 0.2), 0.35, 0.6, branch=2).smooth_union(0.15)
            .torus((-0.2, -0.1, 0.0), 0.5, 0.12, branch=3).union()
            .sphere((0.3, 0.3, 0.3), 0.3, branch=4).subtract()
            .plane((0.6, 0.0, 0.8), 0.4, branch=5).intersect()
            .sphere((-0.4, 0.2, -0.3), 0.25, branch=6)
            .box((-0.4, 0.2, -0.3), (0.2, 0.2, 0.2), branch=7).smooth_intersect(0.05).select())


def test_every_operator_matches_oracle(ctx, oracle):
    prog = _every_op_program()
    g = pr.analytic_framing((-1, -1, -1), (1, 1, 1), 80)
    dist, branch = pr.bake_analytic(ctx, prog, g)
    ref, refb = _oracle_bake(oracle, prog, g)
    assert np.array_equal(dist, ref.reshape(-1)) and np.array_equal(branch, refb.reshape(-1))
    assert set(np.unique(branch)) == {1, 3, 5, 6}  # smooth operators and subtract keep the first branch


def test_device_entry_point(ctx, oracle):
    prog = _every_op_program()
    g = pr.analytic_framing((-1, -0.5, -1), (1, 0.5, 1), 50)
    n = g.nx * g.ny * g.nz
    d_dist, d_branch = ctx.alloc(4 * n), ctx.alloc(n)
    try:
        pr.bake_analytic_device(ctx, prog, g, d_dist, d_branch)
        pr.bake_analytic_device(ctx, pr.Csg().sphere((0, 0, 0), 0.5), g, d_dist)  # branch untouched
        ctx.synchronize()
        dist = ctx.download(d_dist, np.empty(n, np.float32))
        branch = ctx.download(d_branch, np.empty(n, np.uint8))
    finally:
        ctx.free(d_dist)
        ctx.free(d_branch)
    ref, _ = _oracle_bake(oracle, pr.Csg().sphere((0, 0, 0), 0.5), g)
    _, refb = _oracle_bake(oracle, prog, g)
    assert np.array_equal(dist, ref.reshape(-1)) and np.array_equal(branch, refb.reshape(-1))


@pytest.mark.parametrize("prog", [
    pr.Csg().sphere((0, 0, 0), 1).union(),
    pr.Csg().sphere((0, 0, 0), 1).sphere((0, 0, 0), 1),
    pr.Csg()._push(99, 0, 1.0),
    pr.Csg().sphere((0, 0, 0), 1, branch=300),
    pr.Csg().sphere((0, 0, 0), float("nan")),
])
def test_malformed_programs_rejected(ctx, prog):
    g = pr.analytic_framing((-1, -1, -1), (1, 1, 1), 8)
    with pytest.raises(BakeError):
        pr.bake_analytic(ctx, prog, g)


def test_long_program_from_device_memory(ctx, oracle):
    # past 64 instructions the program is read from a device copy instead of the kernel arguments
    rng = np.random.default_rng(11)
    prog = pr.Csg().sphere(rng.uniform(-0.6, 0.6, 3), 0.2, branch=0)
    for i in range(1, 60):
        prog.sphere(rng.uniform(-0.6, 0.6, 3), float(rng.uniform(0.05, 0.2)), branch=i)
        (prog.union() if i % 3 else prog.smooth_union(0.05))
    assert len(prog.ops) == 119
    g = pr.analytic_framing((-1, -1, -1), (1, 1, 1), 48)
    dist, branch = pr.bake_analytic(ctx, prog, g)
    ref, refb = _oracle_bake(oracle, prog, g)
    assert np.array_equal(dist, ref.reshape(-1)) and np.array_equal(branch, refb.reshape(-1))


def test_stack_depth_limit(ctx, oracle):
    deep = pr.Csg()
    for i in range(16):
        deep.sphere((0.05 * i, 0, 0), 0.3, branch=i)
    for _ in range(15):
        deep.union()
    g = pr.analytic_framing((-1, -1, -1), (2, 1, 1), 24)
    dist, branch = pr.bake_analytic(ctx, deep, g)
    ref, refb = _oracle_bake(oracle, deep, g)
    assert np.array_equal(dist, ref.reshape(-1)) and np.array_equal(branch, refb.reshape(-1))
    too_deep = pr.Csg()
    for i in range(17):
        too_deep.sphere((0, 0, 0), 0.3)
    for _ in range(16):
        too_deep.union()
    with pytest.raises(BakeError):
        pr.bake_analytic(ctx, too_deep, g)
