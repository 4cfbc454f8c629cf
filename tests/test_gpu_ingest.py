"""GPU parity of the mesh ingestion gate: ib_weld_positions / ib_watertight (weldPositions,
src/io/mesh_weld.cpp:26-48; isWatertight, src/io/mesh_validate.cpp:21-59) against the oracle.

Bar: bit-exact (integer indices, edge counts, unique positions, and the reference's message text).
"""
import numpy as np
import pytest

from test_oracle_ingest import unit_cube_soup

pytestmark = pytest.mark.gpu

MSG_TAIL = " — fill holes and merge duplicate vertices in your DCC tool and re-export"


def _bounds(xyz):
    p = np.asarray(xyz, np.float32).reshape(-1, 3)
    return p.min(0), p.max(0)


def _cases(meshgen):
    rng = np.random.default_rng(23)
    cube = unit_cube_soup()
    yield "cube", cube
    yield "holey_cube", cube[:-2]
    yield "duplicated_face", np.concatenate([cube, cube[:1]])
    jit = (cube.reshape(-1, 3).astype(np.float64) + rng.uniform(-5e-6, 5e-6, (cube.size // 3, 3)))
    yield "jittered_cube", jit.astype(np.float32).reshape(-1, 9)
    yield "icosphere_5120", meshgen.icosphere(4, 0.5, 1e-3).xyz
    yield "stress_8", meshgen.stress(ncomp=8).xyz
    yield "random_soup", rng.uniform(-1, 1, (5000, 9)).astype(np.float32)
    lattice = rng.integers(-4, 5, (20000, 9)) * 0.25 + rng.uniform(-2e-6, 2e-6, (20000, 9))
    yield "quantised_lattice", lattice.astype(np.float32)
    yield "displaced_327680", meshgen.displaced(7).xyz


@pytest.mark.parametrize("idx", range(9))
def test_weld_and_watertight_match_oracle(ctx, oracle, meshgen, idx):
    name, xyz = list(_cases(meshgen))[idx]
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 9)
    bmin, bmax = _bounds(xyz)
    print(f"case {name}: {len(xyz)} triangles", flush=True)
    w, uniq = ctx.weld_positions(xyz.reshape(-1, 3), bmin, bmax)
    w_ref, uniq_ref = oracle.weld_positions(xyz.reshape(-1, 3), bmin, bmax)
    assert np.array_equal(w, w_ref), name
    assert np.array_equal(uniq, uniq_ref), name
    ok, b, m, msg = ctx.watertight(xyz, bmin, bmax)
    assert (ok, b, m) == oracle.watertight(xyz, bmin, bmax), name
    if ok:
        assert msg == ""
    else:
        head = f"mesh is not watertight ({b} boundary edge{'' if b == 1 else 's'}"
        if m:
            head += f", {m} non-manifold edge{'' if m == 1 else 's'}"
        assert msg == head + ")" + MSG_TAIL, msg


def test_reference_messages(ctx):
    cube = unit_cube_soup()
    assert ctx.watertight(cube, (0, 0, 0), (1, 1, 1)) == (True, 0, 0, "")
    assert ctx.watertight(cube[:-2], (0, 0, 0), (1, 1, 1)) == (
        False, 4, 0, "mesh is not watertight (4 boundary edges)" + MSG_TAIL)
    assert ctx.watertight(np.zeros((0, 9), np.float32), (0, 0, 0), (0, 0, 0)) == (False, 0, 0, "mesh has no triangles")


def test_device_entry_and_repeat(ctx, oracle, meshgen):
    """ib_watertight_device on a device-resident soup, twice (the second smaller), equal to the oracle."""
    for level in (6, 3):
        xyz = meshgen.displaced(level).xyz
        bmin, bmax = _bounds(xyz)
        d = ctx.alloc(xyz.nbytes)
        try:
            ctx.upload(d, xyz)
            got = ctx.watertight_device(d, len(xyz), bmin, bmax)
        finally:
            ctx.free(d)
        assert got == oracle.watertight(xyz, bmin, bmax)[:3]
        assert got[0]


def test_non_representable_position_rejected(ctx):
    from italy_bake.bake import BakeError

    with pytest.raises(BakeError, match="IB_ERR_INVALID_MESH"):
        ctx.weld_positions(np.array([[np.nan, 0, 0]], np.float32), (0, 0, 0), (1, 1, 1))
