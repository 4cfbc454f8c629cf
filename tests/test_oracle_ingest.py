"""Pins for the oracle's mesh ingestion gate: weldPositions (src/io/mesh_weld.cpp:26-48) and
isWatertight (src/io/mesh_validate.cpp:21-59).  Runs without a GPU.

The oracle numbers welded vertices with a hash map in insertion order and counts edges by sorting;
the checks below use the reference test's known answers (tests/mesh_validate_test.cpp:37-62),
lattice points where the pitch and the rounding rule decide the answer, and a numpy restatement
that numbers vertices through np.unique (a sort) instead of a map.
"""
import numpy as np
import pytest


def _quad(a, b, c, d):
    return [a, b, c, a, c, d]


def unit_cube_soup(lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """makeUnitCube of tests/mesh_validate_test.cpp:20-34: six quads, two triangles each."""
    (x0, y0, z0), (x1, y1, z1) = lo, hi
    p = []
    p += _quad((x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0))
    p += _quad((x0, y0, z1), (x0, y1, z1), (x1, y1, z1), (x1, y0, z1))
    p += _quad((x0, y0, z0), (x0, y1, z0), (x0, y1, z1), (x0, y0, z1))
    p += _quad((x1, y0, z0), (x1, y0, z1), (x1, y1, z1), (x1, y1, z0))
    p += _quad((x0, y0, z0), (x0, y0, z1), (x1, y0, z1), (x1, y0, z0))
    p += _quad((x0, y1, z0), (x1, y1, z0), (x1, y1, z1), (x0, y1, z1))
    return np.asarray(p, np.float32).reshape(-1, 9)


def np_weld(pos, bmin, bmax):
    """numpy restatement: keys by the reference's formula, numbering through a sort."""
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    ext = np.asarray(bmax, np.float32) - np.asarray(bmin, np.float32)
    diag = np.sqrt((ext[0] * ext[0] + ext[1] * ext[1]) + ext[2] * ext[2], dtype=np.float32)
    cell = max(np.float32(diag * np.float32(1e-5)), np.float32(1e-6))
    q = (pos / np.float32(cell)).astype(np.float64)  # float32 division, exact in float64
    keys = (np.sign(q) * np.floor(np.abs(q) + 0.5)).astype(np.int64)  # round half away from zero
    if len(pos) == 0:
        return np.zeros(0, np.uint32), np.zeros((0, 3), np.float32)
    _, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    rank = np.empty(len(first), np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(len(first))
    welded = rank[inv.reshape(-1)].astype(np.uint32)
    return welded, pos[np.sort(first)]


def np_watertight(xyz, bmin, bmax):
    xyz = np.asarray(xyz, np.float32).reshape(-1, 9)
    if len(xyz) == 0:
        return False, 0, 0
    w, _ = np_weld(xyz.reshape(-1, 3), bmin, bmax)
    t = w.reshape(-1, 3).astype(np.int64)
    keep = (t[:, 0] != t[:, 1]) & (t[:, 1] != t[:, 2]) & (t[:, 2] != t[:, 0])
    t = t[keep]
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    e.sort(axis=1)
    if len(e) == 0:
        return True, 0, 0
    _, counts = np.unique(e, axis=0, return_counts=True)
    b, m = int((counts == 1).sum()), int((counts > 2).sum())
    return b == 0 and m == 0, b, m


def bounds(xyz):
    p = np.asarray(xyz, np.float32).reshape(-1, 3)
    return p.min(0), p.max(0)


# ---- known answers of tests/mesh_validate_test.cpp ------------------------------------------------
def test_reference_kat_cube(oracle):
    cube = unit_cube_soup()
    assert oracle.watertight(cube, (0, 0, 0), (1, 1, 1)) == (True, 0, 0)


def test_reference_kat_holey_cube(oracle):
    holey = unit_cube_soup()[:-2]  # positions.resize(size - 6): the last face's two triangles
    ok, b, m = oracle.watertight(holey, (0, 0, 0), (1, 1, 1))
    assert not ok and b == 4 and m == 0


def test_reference_kat_empty(oracle):
    ok, b, m = oracle.watertight(np.zeros((0, 9), np.float32), (0, 0, 0), (0, 0, 0))
    assert not ok and b == 0 and m == 0


# ---- closed forms -------------------------------------------------------------------------------
def test_duplicated_face_is_nonmanifold(oracle):
    cube = unit_cube_soup()
    dup = np.concatenate([cube, cube[:1]])
    assert oracle.watertight(dup, (0, 0, 0), (1, 1, 1)) == (False, 0, 3)


def test_degenerate_triangles_are_skipped(oracle):
    cube = unit_cube_soup()
    deg = np.array([[0, 0, 0, 0, 0, 0, 1, 1, 1]], np.float32)  # two equal corners
    assert oracle.watertight(np.concatenate([cube, deg]), (0, 0, 0), (1, 1, 1)) == (True, 0, 0)


def test_jitter_within_the_pitch_welds(oracle):
    # pitch = |(1,1,1)| * 1e-5 = 1.73e-5: corners moved by less than 0.3 pitch still weld, so the
    # soup is watertight although no two copies of a corner are bit-equal
    cube = unit_cube_soup().reshape(-1, 3).astype(np.float64)
    rng = np.random.default_rng(5)
    cube += rng.uniform(-5e-6, 5e-6, cube.shape)
    cube = cube.astype(np.float32)
    ok, b, m = oracle.watertight(cube, (0, 0, 0), (1, 1, 1))
    assert ok, (b, m)
    assert oracle.boundary_edge_count(cube) > 0  # exact-position welding sees the seams


def _pow2_pitch_bounds(k):
    """bounds (0,0,0)..(d,0,0) whose FP32 pitch d * 1e-5 is exactly 2^-k"""
    target = np.float32(2.0 ** -k)
    d = np.float32(target / np.float32(1e-5))
    for _ in range(64):
        p = np.float32(d * np.float32(1e-5))
        if p == target:
            return (0.0, 0.0, 0.0), (float(d), 0.0, 0.0), float(target)
        d = np.nextafter(d, np.float32(np.inf) if p < target else np.float32(-np.inf), dtype=np.float32)
    raise AssertionError("no float extent gives a power-of-two pitch")


def test_half_lattice_rounds_away_from_zero(oracle):
    bmin, bmax, cell = _pow2_pitch_bounds(10)
    # x / cell = 2.5 rounds to 3 (lround: half away from zero), 3.4 also to 3, 2.4 to 2;
    # -2.5 rounds to -3 and -2.6 to -3, -2.4 to -2
    xs = [2.5, 3.4, 2.4, -2.5, -2.6, -2.4]
    pos = np.array([[x * cell, 0, 0] for x in xs], np.float32)
    w, uniq = oracle.weld_positions(pos, bmin, bmax)
    assert list(w) == [0, 0, 1, 2, 2, 3]
    assert np.array_equal(uniq, pos[[0, 2, 3, 5]])


def test_minimum_pitch(oracle):
    # tiny bounds: the pitch is the 1e-6 floor; 0.4e-6 rounds to 0, 0.6e-6 to 1
    pos = np.array([[0, 0, 0], [4e-7, 0, 0], [6e-7, 0, 0], [0, -4e-7, 0], [0, 0, -6e-7]], np.float32)
    w, _ = oracle.weld_positions(pos, (0, 0, 0), (1e-3, 0, 0))
    assert list(w) == [0, 0, 1, 0, 2]


def test_first_occurrence_numbering(oracle):
    a, b, c = [1, 2, 3], [4, 5, 6], [7, 8, 9]
    w, uniq = oracle.weld_positions(np.array([b, a, b, c, a], np.float32), (0, 0, 0), (10, 10, 10))
    assert list(w) == [0, 1, 0, 2, 1]
    assert np.array_equal(uniq, np.array([b, a, c], np.float32))


def test_non_representable_key_rejected(oracle):
    with pytest.raises(ValueError):
        oracle.weld_positions(np.array([[np.nan, 0, 0]], np.float32), (0, 0, 0), (1, 1, 1))
    with pytest.raises(ValueError):
        oracle.weld_positions(np.array([[3e38, 0, 0]], np.float32), (0, 0, 0), (1e-6, 0, 0))


# ---- against the numpy restatement on workload meshes --------------------------------------------
@pytest.mark.parametrize("name", ["icosphere", "stress", "random_soup", "quantised"])
def test_weld_and_watertight_match_numpy(oracle, meshgen, name):
    rng = np.random.default_rng(11)
    if name == "icosphere":
        xyz = meshgen.icosphere(4, 0.5, 1e-3).xyz
    elif name == "stress":
        xyz = meshgen.stress(ncomp=8).xyz
    elif name == "random_soup":
        xyz = rng.uniform(-1, 1, (2000, 9)).astype(np.float32)
    else:  # coordinates on a coarse lattice: many coincident and near-coincident corners
        xyz = (rng.integers(-4, 5, (3000, 9)) * 0.25 + rng.uniform(-2e-6, 2e-6, (3000, 9))).astype(np.float32)
    bmin, bmax = bounds(xyz)
    w, uniq = oracle.weld_positions(xyz.reshape(-1, 3), bmin, bmax)
    w2, uniq2 = np_weld(xyz.reshape(-1, 3), bmin, bmax)
    assert np.array_equal(w, w2)
    assert np.array_equal(uniq, uniq2)
    assert oracle.watertight(xyz, bmin, bmax) == np_watertight(xyz, bmin, bmax)
    if name == "icosphere":
        assert oracle.watertight(xyz, bmin, bmax) == (True, 0, 0)
    if name == "stress":
        assert not oracle.watertight(xyz, bmin, bmax)[0]
