"""Pins for the CPU oracle (runs without a GPU).

Each oracle function is checked against something that does not restate its own formula: the
reference tests' known answers, closed forms, an independent algorithm (polygon clipping, bitwise
Morton interleave, FP64 brute force) or an invariant, chosen so that a plausible mistake (wrong
axis, sign, operand, dropped term) fails.
"""
import math

import numpy as np
import pytest


# ------------------------------------------------------------------------------------------
# triBoxOverlap (src/convert/tri_box_overlap.h:72-114)
# ------------------------------------------------------------------------------------------
def _clip_nonempty(tri, lo, hi):
    """Exact-ish FP64 triangle/box intersection by Sutherland-Hodgman clipping (closed box)."""
    poly = [np.asarray(p, np.float64) for p in tri]
    for axis in range(3):
        for bound, sgn in ((lo[axis], 1.0), (hi[axis], -1.0)):
            out = []
            n = len(poly)
            for i in range(n):
                a, b = poly[i], poly[(i + 1) % n]
                da, db = sgn * (a[axis] - bound), sgn * (b[axis] - bound)
                if da >= 0:
                    out.append(a)
                if (da >= 0) != (db >= 0):
                    t = da / (da - db)
                    out.append(a + t * (b - a))
            poly = out
            if not poly:
                return False
    return True


def test_tri_box_against_polygon_clipping(oracle):
    rng = np.random.default_rng(1)
    agree = 0
    for _ in range(4000):
        c = rng.uniform(-1, 1, 3).astype(np.float32)
        h = np.float32(rng.uniform(0.05, 0.5))
        tri = (c + rng.uniform(-1.2, 1.2, (3, 3)) * rng.uniform(0.1, 1.0)).astype(np.float32)
        # skip near-ties: the answer must be the same for a slightly grown and shrunk box
        eps = 1e-4 * float(h)
        grown = _clip_nonempty(tri, c - h - eps, c + h + eps)
        shrunk = _clip_nonempty(tri, c - h + eps, c + h - eps)
        if grown != shrunk:
            continue
        got = oracle.tri_box_overlap(c, [h, h, h], tri[0], tri[1], tri[2])
        assert got == grown, (c, h, tri)
        agree += 1
    assert agree > 3500


def test_tri_box_touching_counts_as_overlap(oracle):
    # a face lying exactly on the plane x = 0.5 touches both cells that share it (`>` / `<` tests)
    tri = np.array([[0.5, 0.1, 0.1], [0.5, 0.4, 0.1], [0.5, 0.1, 0.4]], np.float32)
    for cx in (0.25, 0.75):
        assert oracle.tri_box_overlap([cx, 0.25, 0.25], [0.25] * 3, *tri)
    assert not oracle.tri_box_overlap([1.25, 0.25, 0.25], [0.25] * 3, *tri)


def test_voxelize_reference_known_answers(oracle, meshgen):
    """tests/voxelize_test.cpp:37-58: cube shell at resolution 8."""
    cube = meshgen.cube_shell()
    cells, _ = oracle.voxelize_first_touch(cube.xyz, cube.bmin, cube.bmax, 8)
    assert 0 < len(cells) <= 6 * 8 * 8
    s = {tuple(c) for c in cells.tolist()}
    assert (0, 0, 0) in s
    assert (4, 4, 4) not in s
    assert len(s) == len(cells)  # each cell once


def test_voxelize_first_touch_order_brute_force(oracle, meshgen):
    """First-touch order (voxelize.cpp:88-96) == a brute-force replay over every cell of the grid."""
    m = meshgen.icosphere(1, 0.5, 0.0)
    res = 6
    # bounds offset so that no vertex sits exactly on a cell face: there the reference's AABB cell
    # range and an all-cells scan legitimately differ (a face-touching neighbour outside the range)
    bmin, bmax = m.bmin - np.float32(0.0131), m.bmax + np.float32(0.0217)
    origin, h = oracle.vox_framing(bmin, bmax, res)
    half = [h * np.float32(0.5)] * 3
    order, seen = [], set()
    for t in range(m.ntri):
        tri = m.xyz[t]
        for z in range(res):
            for y in range(res):
                for x in range(res):
                    c = origin + (np.array([x, y, z], np.float32) + np.float32(0.5)) * h
                    if (x, y, z) not in seen and oracle.tri_box_overlap(c, half, *tri):
                        seen.add((x, y, z))
                        order.append((x, y, z))
    cells, _ = oracle.voxelize_first_touch(m.xyz, bmin, bmax, res)
    assert [tuple(c) for c in cells.tolist()] == order


def test_occupancy_contains_every_vertex_cell(oracle, meshgen):
    m = meshgen.icosphere(3, 0.5, 1e-3)
    n, lo, hi = 32, -0.6, 0.6
    h = np.float32((np.float32(hi) - np.float32(lo)) / np.float32(n))
    words = oracle.occupancy(m.xyz, [lo] * 3, h, n, n, n)
    bits = np.unpackbits(words.view(np.uint8), bitorder="little").reshape(n, n, n)
    v = m.xyz.reshape(-1, 3)
    idx = np.floor((v - np.float32(lo)) / h).astype(int)
    # vertices strictly inside a cell (not on a face) must mark it
    frac = (v - np.float32(lo)) / h - idx
    ok = np.all((frac > 1e-3) & (frac < 1 - 1e-3), axis=1)
    assert ok.sum() > 100
    assert bits[idx[ok, 2], idx[ok, 1], idx[ok, 0]].all()


def test_splat_materials_brute_force(oracle, meshgen):
    """bakePerCellMaterial (sdf_baker.cpp:76-128) == a replay over every (triangle, cell) pair in
    triangle order with float32 sums; untouched cells keep (1, 1, 1, 0, 1)."""
    m = meshgen.icosphere(1, 0.5, 0.0)
    rng = np.random.default_rng(7)
    mat = rng.uniform(0, 1, (m.ntri, 5)).astype(np.float32)
    n = 7
    origin = (m.bmin - np.float32(0.0131)).astype(np.float32)
    h = np.float32((m.bmax.max() - m.bmin.min() + np.float32(0.04)) / n)
    half = [h * np.float32(0.5)] * 3
    sums = np.zeros((5, n, n, n), np.float32)
    cnt = np.zeros((n, n, n), np.int64)
    for t in range(m.ntri):
        for z in range(n):
            for y in range(n):
                for x in range(n):
                    c = origin + (np.array([x, y, z], np.float32) + np.float32(0.5)) * h
                    if oracle.tri_box_overlap(c, half, *m.xyz[t]):
                        sums[:, z, y, x] = (sums[:, z, y, x] + mat[t]).astype(np.float32)
                        cnt[z, y, x] += 1
    want = np.where(cnt > 0, sums / np.maximum(cnt, 1).astype(np.float32),
                    np.array([1, 1, 1, 0, 1], np.float32)[:, None, None, None]).astype(np.float32)
    got = oracle.splat_materials(m.xyz, mat, origin, h, n, n, n)
    assert (cnt > 0).sum() > 20 and (cnt > 1).sum() > 5
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


# ------------------------------------------------------------------------------------------
# framing (sdf_baker.cpp:137-148, voxelize.cpp:54-57)
# ------------------------------------------------------------------------------------------
def test_sdf_framing_known_values(oracle):
    o, h, n = oracle.sdf_framing([-1, -1, -1], [1, 1, 1], 24)
    assert h == np.float32(2.0) / np.float32(24.0)
    assert n == (29, 29, 29)  # ceil((2 + 4h) / h) + 1
    assert np.allclose(o, -1 - 2 * h)
    o, h, n = oracle.sdf_framing([0, 0, 0], [2, 1, 0.5], 8)
    assert h == np.float32(0.25) and n == (13, 9, 7)
    o, h, n = oracle.sdf_framing([0, 0, 0], [0, 0, 0], 0)  # degenerate: longest clamps to 1e-6, res to 1
    assert h == np.float32(1e-6) and n == (5, 5, 5)


# ------------------------------------------------------------------------------------------
# distance
# ------------------------------------------------------------------------------------------
TRI = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)


@pytest.mark.parametrize("q,d2", [
    ((0.2, 0.2, 0.5), 0.25),       # face region: height above the plane
    ((-1, -1, 0), 2.0),            # vertex a
    ((2, -1, 0), 2.0),             # vertex b
    ((-1, 2, 0.5), 1.0 + 1.0 + 0.25),  # vertex c
    ((0.5, -1, 0.3), 1.09),        # edge ab: distance to the line y = 0
    ((1, 1, 0), 0.5),              # edge bc: line x + y = 1
    ((-0.5, 0.3, -0.2), 0.25 + 0.04),  # edge ca: line x = 0
])
def test_point_triangle_regions(oracle, q, d2):
    assert oracle.tri_dist2_f64(q, TRI) == pytest.approx(d2, rel=1e-12)
    assert oracle.tri_dist2_f32(q, TRI) == pytest.approx(d2, rel=1e-6)


def test_degenerate_triangles(oracle):
    seg = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0]], np.float32)  # collinear
    assert oracle.tri_is_sliver(seg)
    assert oracle.tri_dist2_f32([1, 1, 0], seg) == pytest.approx(1.0)
    assert oracle.tri_dist2_f32([3, 0, 0], seg) == pytest.approx(1.0)
    pt = np.array([[1, 1, 1]] * 3, np.float32)  # all three vertices equal
    assert oracle.tri_dist2_f32([1, 1, 3], pt) == pytest.approx(4.0)
    assert not oracle.tri_is_sliver(TRI)


def test_f32_distance_matches_f64(oracle):
    rng = np.random.default_rng(2)
    worst = 0.0
    for _ in range(3000):
        tri = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
        q = rng.uniform(-1.5, 1.5, 3).astype(np.float32)
        d64 = math.sqrt(oracle.tri_dist2_f64(q.astype(np.float64), tri))
        d32 = math.sqrt(oracle.tri_dist2_f32(q, tri))
        worst = max(worst, abs(d32 - d64))
    assert worst < 2e-6


def test_f32_distance_error_within_pruning_slack(oracle):
    """The traversals enter a box when it could hold a triangle within sqrt(best) + 2^-17 (Mq + 3 Mp);
    that slack must exceed the FP32 distance's error by a wide margin (128 units here)."""
    below, above = oracle.f32_error_ratio(400000, seed=7)
    assert 0.0 < below < 16.0 and above < 16.0


def test_lbvh_distance_equals_brute_force_on_boxes(oracle, meshgen):
    """Axis-aligned box faces lie on their LBVH boxes' faces, where a box bound and the FP32 triangle
    distance can disagree in the last bits: the slack keeps the traversal equal to brute force."""
    m = meshgen.stress(seed=11, ncomp=4, remove_frac=0.0)
    rng = np.random.default_rng(5)
    g = np.linspace(-1.0, 1.0, 11, dtype=np.float32)
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pts = np.concatenate([pts, rng.uniform(-1, 1, (500, 3)).astype(np.float32)])
    ud = oracle.udf_points(m.xyz, pts)
    brute = np.sqrt(oracle.udf2_f32_brute(m.xyz, pts))
    assert np.array_equal(ud, brute)


def test_box_mesh_distance_is_sdbox(oracle, meshgen):
    """The cube shell's exact SDF is sdBox (src/convert/sdf_primitives.h:12-15)."""
    cube = meshgen.cube_shell()
    rng = np.random.default_rng(3)
    pts = rng.uniform(-0.5, 1.5, (400, 3))
    d, w = oracle.sdf_f64_points(cube.xyz, pts)
    q = np.abs(pts - 0.5) - 0.5
    sd = np.linalg.norm(np.maximum(q, 0), axis=1) + np.minimum(q.max(axis=1), 0)
    assert np.max(np.abs(d - sd)) < 1e-6
    d32 = np.sqrt(oracle.udf2_f32_brute(cube.xyz, pts.astype(np.float32)))
    assert np.max(np.abs(d32 - np.abs(sd))) < 1e-6


# ------------------------------------------------------------------------------------------
# winding number, crossings, parity
# ------------------------------------------------------------------------------------------
def test_gwn_closed_forms(oracle, meshgen):
    octant = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    assert oracle.gwn_f64([0, 0, 0], octant) == pytest.approx(1 / 8, abs=1e-12)
    assert oracle.gwn_f64([0, 0, 0], octant[::-1].copy()) == pytest.approx(-1 / 8, abs=1e-12)
    cube = meshgen.cube_shell()  # the reference's cube shell is wound inward
    assert oracle.gwn_f64([0.5, 0.5, 0.5], cube.xyz) == pytest.approx(-1.0, abs=1e-12)
    # one face of the cube seen from the centre subtends 4 pi / 6
    assert oracle.gwn_f64([0.5, 0.5, 0.5], cube.xyz[:2]) == pytest.approx(-1 / 6, abs=1e-12)
    assert abs(oracle.gwn_f64([50, 40, 30], cube.xyz)) < 1e-9


def test_crossings_and_parity_match_gwn_on_closed_meshes(oracle, meshgen):
    rng = np.random.default_rng(4)
    for m in (meshgen.icosphere(3, 0.5, 1e-3), meshgen.displaced(3, 11), meshgen.cube_shell(),
              meshgen.decimated(600, 5)):
        assert oracle.boundary_edge_count(m.xyz) == 0
        lo, hi = m.bmin - 0.2, m.bmax + 0.2
        pts = rng.uniform(lo, hi, (150, 3)).astype(np.float32)
        d, w = oracle.sdf_f64_points(m.xyz, pts)
        for p, dd, ww in zip(pts, d, w):
            if abs(dd) < 1e-4:
                continue
            assert abs(ww - round(ww)) < 1e-9
            assert oracle.crossings_up(p, m.xyz) == round(ww)
            assert oracle.parity_inside(p, m.xyz) == (abs(ww) > 0.5)


def test_parity_dirs_are_unit_and_start_with_plus_z(oracle):
    d = oracle.parity_dirs()
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert d[0].tolist() == [0.0, 0.0, 1.0]  # sdf_bake.cu:56


def test_crossings_watertight_through_vertices(oracle, meshgen):
    """Columns through exact mesh vertices and edges still count each crossing once."""
    m = meshgen.icosphere(2, 0.5, 0.0)
    v = np.unique(m.xyz.reshape(-1, 3), axis=0)
    for p in v[:60]:
        for zoff in (-2.0, 2.0):
            q = np.array([p[0], p[1], zoff], np.float32)
            c = oracle.crossings_up(q, m.xyz)
            assert c == 0  # a full column through a closed surface enters and leaves
    # midpoints of edges, from inside
    q = np.array([0.0, 0.0, 0.0], np.float32)
    assert oracle.crossings_up(q, m.xyz) == 1


# ------------------------------------------------------------------------------------------
# LBVH
# ------------------------------------------------------------------------------------------
def _morton_bitwise(q):
    code = 0
    for bit in range(10):
        for k, shift in ((0, 2), (1, 1), (2, 0)):
            code |= ((int(q[k]) >> bit) & 1) << (3 * bit + shift)
    return code


def test_morton_codes_independent(oracle, meshgen):
    m = meshgen.displaced(2, 3)
    codes = oracle.morton_codes(m.xyz)
    c = ((m.xyz[:, 0] + m.xyz[:, 1]) + m.xyz[:, 2]) / np.float32(3.0)
    lo, hi = c.min(0), c.max(0)
    u = (c - lo) / (hi - lo)
    q = np.clip(u * np.float32(1024.0), 0, 1023).astype(np.uint32)
    assert [int(x) for x in codes] == [_morton_bitwise(r) for r in q]


def test_lbvh_invariants(oracle, meshgen):
    m = meshgen.displaced(3, 9)
    b = oracle.bvh(m.xyz)
    n = m.ntri
    assert np.all(np.diff(b["codes"].astype(np.int64)) >= 0)
    assert sorted(b["order"].tolist()) == list(range(n))
    # every node except the root has exactly one parent; leaves reached once
    ch = np.concatenate([b["left"], b["right"]])
    ids = np.where(ch >= 0, ch, (n - 1) + ~ch)
    assert sorted(ids.tolist()) == list(range(1, 2 * n - 1))
    for i in range(n - 1):
        for c in (b["left"][i], b["right"][i]):
            j = c if c >= 0 else (n - 1) + ~c
            assert np.all(b["lo"][i] <= b["lo"][j]) and np.all(b["hi"][i] >= b["hi"][j])
    assert b["depth"] < 64


# ------------------------------------------------------------------------------------------
# dense CPU bake (the host-core baseline)
# ------------------------------------------------------------------------------------------
def test_dense_bake_equals_brute_force(oracle, meshgen):
    m = meshgen.displaced(2, 21)
    n, lo, hi = 12, -0.6, 0.6
    h = np.float32((np.float32(hi) - np.float32(lo)) / np.float32(n))
    d, st = oracle.bake_dense(m.xyz, [lo] * 3, h, n, n, n, threads=2)
    pts = oracle.cell_centres([lo] * 3, h, n, n, n).reshape(-1, 3)
    brute = np.sqrt(oracle.udf2_f32_brute(m.xyz, pts)).reshape(n, n, n)
    assert np.array_equal(np.abs(d), brute)  # bit-exact: same per-triangle FP32 values, min is exact
    d64, w = oracle.sdf_f64_points(m.xyz, pts.astype(np.float64))
    assert np.array_equal(d.reshape(-1) < 0, np.abs(w) > 0.5)
    assert np.max(np.abs(d.reshape(-1) - d64)) < 1e-6


def test_open_mesh_rejected_by_closed_bake(oracle, meshgen):
    u = meshgen.uv_sphere()
    assert oracle.boundary_edge_count(u.xyz) == 64  # seam crack + south-pole ring (float trig)
    d, st = oracle.bake_dense(u.xyz, [-1.2] * 3, 0.1, 4, 4, 4)
    assert d is None


# ------------------------------------------------------------------------------------------
# the renderer's grid consumer (pathtracer.cu:587-613, :684-709, :740-846)
# ------------------------------------------------------------------------------------------
def _affine_grid(n=12, h=0.1):
    origin = np.array([-0.6, -0.55, -0.5], np.float32)
    a = np.array([0.3, -0.7, 0.64], np.float64)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    c = origin.astype(np.float64) + (np.stack([x, y, z], -1) + 0.5) * np.float64(np.float32(h))
    return origin, np.float32(h), (c @ a + 0.25).astype(np.float32), a


def test_sample_grid_affine_field(oracle):
    """Trilinear interpolation reproduces an affine field, and the B-spline gradient of an affine
    field is its constant gradient (the weights sum to 1 and their derivatives to 0)."""
    origin, h, dist, a = _affine_grid()
    n = dist.shape[0]
    rng = np.random.default_rng(5)
    # inside the region where no stencil index is clamped: local coordinates in [1, n - 3]
    loc = rng.uniform(1.0, n - 3.0, (500, 3))
    pts = (origin + (loc + 0.5) * h).astype(np.float32)
    vals, nrm, near = oracle.sample_grid(origin, h, dist, pts)
    want = pts.astype(np.float64) @ a + 0.25
    assert np.abs(vals - want).max() < 2e-6
    assert np.abs(nrm - a / np.linalg.norm(a)).max() < 1e-5
    l = (pts - origin) / h - np.float32(0.5)
    r = np.clip(np.sign(l) * np.floor(np.abs(l) + 0.5), 0, n - 1).astype(int)  # lroundf: half away from zero
    assert np.array_equal(near, (r[:, 2] * n + r[:, 1]) * n + r[:, 0])


def test_sample_grid_clamps_outside(oracle):
    """Outside the grid sampleSdf clamps to the border cell pair and the nearest cell to the border."""
    origin, h, dist, _ = _affine_grid(6)
    vals, _, near = oracle.sample_grid(origin, h, dist, [[-5.0, -5.0, -5.0], [5.0, 5.0, 5.0]])
    assert vals[0] == dist[0, 0, 0] and vals[1] == dist[-1, -1, -1]
    assert near.tolist() == [0, dist.size - 1]


def test_trace_grid_sphere(oracle):
    """Sphere tracing the sampled SDF of a sphere of radius 0.5 hits at |o| - 0.5 (within a tenth of
    a voxel) for rays aimed at the centre, and misses for rays that leave the grid first."""
    n, h = 48, np.float32(1.6 / 48)
    origin = np.full(3, -0.8, np.float32)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    c = origin + (np.stack([x, y, z], -1).astype(np.float32) + np.float32(0.5)) * h
    dist = (np.linalg.norm(c.astype(np.float64), axis=-1) - 0.5).astype(np.float32)
    rng = np.random.default_rng(9)
    u = rng.normal(size=(200, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    org = (0.75 * u).astype(np.float32)
    t = oracle.trace_grid(origin, h, dist, org, -u)
    assert np.abs(t - 0.25).max() < 0.1 * h
    away = oracle.trace_grid(origin, h, dist, org, u)
    assert (away == -1).all()


@pytest.mark.parametrize("which", ["icosphere", "displaced", "stress", "cube"])
def test_udf_points_f64_equals_brute_force(oracle, meshgen, which):
    """The FP64 LBVH query returns the brute-force FP64 minimum bit for bit (seeded points in and around
    the mesh, including points on its vertices)."""
    m = {"icosphere": lambda: meshgen.icosphere(4, 0.5, 1e-3), "displaced": lambda: meshgen.displaced(5),
         "stress": lambda: meshgen.stress(ncomp=8), "cube": meshgen.cube_shell}[which]()
    rng = np.random.default_rng(17)
    lo, hi = m.bmin - 0.3, m.bmax + 0.3
    pts = rng.uniform(lo, hi, (3000, 3))
    pts[:200] = m.xyz.reshape(-1, 3)[rng.choice(m.ntri * 3, 200)]
    d_bvh = oracle.udf_points_f64(m.xyz, pts)
    d_brute, _ = oracle.sdf_f64_points(m.xyz, pts)
    assert np.array_equal(d_bvh, np.abs(d_brute))
    assert (d_bvh[:200] == 0).all()


def _open_box_hits(q, dirs):
    """Parity-ray hits of the cube [-0.5, 0.5]^3 without its +z face, from a point q inside it or on the open
    face's plane, by slab arithmetic: a ray leaves a convex box through exactly one face, and crosses the
    surface once unless that face is the missing one."""
    hits = []
    for d in dirs.reshape(5, 3).astype(np.float64):
        t_exit, face = np.inf, None
        for k in range(3):
            if d[k] != 0.0:
                t = ((0.5 if d[k] > 0 else -0.5) - q[k]) / d[k]
                if t < t_exit:
                    t_exit, face = t, (k, d[k] > 0)
        hits.append(0 if face == (2, True) else 1)
    return np.array(hits)


def test_parity_hits_and_tie_on_an_open_face(oracle):
    """On the plane of a missing cube face, inside the hole, the winding number of the open box is exactly 1/2
    (the five faces subtend half the sphere from either side), so the sign is a tie and the 5-ray parity vote
    decides (sdf_bake.cu:53-69): rays leaving through the hole see no surface, the others one face."""
    import shapes

    box = shapes.box_faces(0.5, 4, skip=("+z",))
    dirs = oracle.parity_dirs()
    assert oracle.boundary_edge_count(box) == 16
    for q in ([0.1, 0.2, 0.5], [-0.3, 0.05, 0.5], [0.45, -0.45, 0.5]):
        hits = oracle.parity_hits(q, box)
        assert np.array_equal(hits, _open_box_hits(np.array(q, np.float64), dirs)), (q, hits)
        assert np.array_equal(hits, [0, 1, 0, 1, 0])  # +z and the two upward rays leave through the hole
        assert abs(oracle.gwn_f64(q, box) - 0.5) < 1e-12
        assert oracle.gwn_sign(q, box) == (False, True)  # 2 of 5 odd: outside, by parity
    rng = np.random.default_rng(3)
    for q in rng.uniform(-0.45, 0.45, (20, 3)):  # strictly inside: the vote agrees with |w| > 1/2
        hits = oracle.parity_hits(q, box)
        assert np.array_equal(hits, _open_box_hits(q.astype(np.float32).astype(np.float64), dirs))
        inside, tie = oracle.gwn_sign(q, box)
        assert not tie and inside == (abs(oracle.gwn_f64(q, box)) > 0.5)
    # a planted mistake: counting the +z ray as a hit turns the vote at the tie
    q = [0.1, 0.2, 0.5]
    h = oracle.parity_hits(q, box)
    h[0] += 1
    assert (np.count_nonzero(h & 1) * 2 > 5) != oracle.parity_inside(q, box)


def test_signed_points_gwn_matches_brute_force(oracle):
    """signed_points_gwn: FP32 brute-force distance, sign from the FP64 winding number, on an open mesh."""
    import shapes

    box = shapes.box_faces(0.5, 3, skip=("+z", "-x"))
    rng = np.random.default_rng(11)
    pts = rng.uniform(-0.8, 0.8, (400, 3)).astype(np.float32)
    d, tie = oracle.signed_points_gwn(box, pts, threads=4)
    ref = np.sqrt(oracle.udf2_f32_brute(box, pts))
    assert np.array_equal(np.abs(d), ref)
    w = np.array([oracle.gwn_f64(p, box) for p in pts.astype(np.float64)])
    assert not tie.any()
    assert np.array_equal(d < 0, np.abs(w) > 0.5)


def test_signed_points_closed_equals_the_dense_bake(oracle, meshgen):
    """signed_points_closed is the dense CPU bake evaluated per point: equal bits at every cell centre."""
    m = meshgen.displaced(4, 12)
    h = np.float32(1.6 / 24)
    origin = [-0.8, -0.8, -0.8]
    dense, _ = oracle.bake_dense(m.xyz, origin, h, 24, 24, 24)
    pts = oracle.cell_centres_at(origin, h, np.arange(24 ** 3), 24, 24)
    got = oracle.signed_points_closed(m.xyz, pts, threads=4)
    assert np.array_equal(got.view(np.uint32), dense.reshape(-1).view(np.uint32))
    import shapes

    assert oracle.signed_points_closed(shapes.box_faces(0.5, 2, skip=("+z",)), pts[:10]) is None


def test_bake_band_against_brute_force(oracle, meshgen):
    """orc_bake_band: the bake's brick rule (equal to band_bricks), FP32 distances equal to the LBVH point
    query, signs equal to the FP64 brute-force winding number (or its parity tie) on an open union; on a
    closed mesh, every baked voxel equal to the dense CPU bake."""
    m = meshgen.stress(seed=8, ncomp=4, remove_frac=0.02)
    h = np.float32(2.0 / 48)
    origin = [-1.0, -1.0, -1.0]
    idx, pool, st = oracle.bake_band(m.xyz, origin, h, 48, 48, 48, 4, stride=1, threads=4)
    assert np.array_equal(idx.reshape(-1) >= 0, oracle.band_bricks(m.xyz, origin, h, 48, 48, 48, 4))
    assert st["cap_triangles"] == oracle.boundary_edge_count(m.xyz) > 0
    kept = np.nonzero(idx.reshape(-1) >= 0)[0]
    slots = idx.reshape(-1)[kept]
    rng = np.random.default_rng(5)
    pick = rng.choice(len(kept), 24, replace=False)
    lin, vals = [], []
    for i in pick:
        b = kept[i]
        bx, by, bz = b % 6, (b // 6) % 6, b // 36
        off = np.arange(0, 512, 5)
        lin.append(((bz * 8 + off // 64) * 48 + by * 8 + (off // 8) % 8) * 48 + bx * 8 + off % 8)
        vals.append(pool[slots[i]].reshape(-1)[off])
    lin, vals = np.concatenate(lin), np.concatenate(vals)
    pts = oracle.cell_centres_at(origin, h, lin, 48, 48)
    ref, _ = oracle.signed_points_gwn(m.xyz, pts, threads=4)
    assert np.array_equal(np.abs(vals).view(np.uint32), np.abs(ref).view(np.uint32))
    assert np.array_equal(vals < 0, ref < 0)
    closed = meshgen.displaced(4, 2)
    h2 = np.float32(1.6 / 32)
    idx, pool, st = oracle.bake_band(closed.xyz, [-0.8] * 3, h2, 32, 32, 32, 4, stride=1, threads=4)
    assert st["cap_triangles"] == 0
    dense, _ = oracle.bake_dense(closed.xyz, [-0.8] * 3, h2, 32, 32, 32)
    for b in np.nonzero(idx.reshape(-1) >= 0)[0]:
        bx, by, bz = b % 4, (b // 4) % 4, b // 16
        want = dense[bz * 8:bz * 8 + 8, by * 8:by * 8 + 8, bx * 8:bx * 8 + 8]
        assert np.array_equal(pool[idx.reshape(-1)[b]].view(np.uint32), want.view(np.uint32))


def test_bake_band_ties_on_an_open_face(oracle):
    """The band of the open cube on the tie grid: the hole-plane voxels of the kept bricks are parity ties,
    signed as orc_gwn_sign signs them."""
    import shapes

    box = shapes.box_faces(0.5, 4, skip=("+z",))
    origin, h = [-0.78125] * 3, np.float32(0.0625)
    idx, pool, st = oracle.bake_band(box, origin, h, 24, 24, 24, 4, stride=1, threads=4)
    assert st["ties"] > 0 and st["on_cap"] > 0
    kept = np.nonzero(idx.reshape(-1) >= 0)[0]
    lin, vals = [], []
    for b in kept:
        bx, by, bz = b % 3, (b // 3) % 3, b // 9
        off = np.arange(512)
        lin.append(((bz * 8 + off // 64) * 24 + by * 8 + (off // 8) % 8) * 24 + bx * 8 + off % 8)
        vals.append(pool[idx.reshape(-1)[b]].reshape(-1))
    lin, vals = np.concatenate(lin), np.concatenate(vals)
    ref, tie = oracle.signed_points_gwn(box, oracle.cell_centres_at(origin, h, lin, 24, 24), threads=4)
    far = np.abs(ref) > 1.5e-4
    assert tie.sum() > 0
    assert np.array_equal(np.abs(vals), np.abs(ref))
    assert np.array_equal((vals < 0)[far], (ref < 0)[far])


def test_gwn_f64_points_decides_as_the_brute_force_sum(oracle, meshgen):
    """orc_gwn_f64_points (a bounded tree walk, the brute-force sum where the bound leaves the decision open)
    against orc_gwn_f64 over every triangle: the same |w| > 1/2 and the same 1e-6 tie band at points far from
    and near a closed displaced icosphere, near an open stress union, and on and just off the plane of an open
    cube's missing face (w = 1/2 there); points on a triangle take the brute force."""
    import shapes

    rng = np.random.default_rng(7)
    closed = meshgen.displaced(4, 3).xyz
    open_union = meshgen.stress(seed=9, ncomp=4, remove_frac=0.05).xyz
    cube = shapes.box_faces(0.5, 3, skip=("+z",))
    cases = []
    for xyz in (closed, open_union):
        t = xyz.reshape(-1, 3, 3).astype(np.float64)
        c = t.mean(1)
        n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
        n /= np.linalg.norm(n, axis=1, keepdims=True) + 1e-300
        pick = rng.choice(len(t), 600, replace=False)
        off = rng.choice([-3e-2, -1e-3, -1e-5, 1e-5, 1e-3, 3e-2], size=600)[:, None]
        near = c[pick] + off * n[pick]
        far = rng.uniform(-1.2, 1.2, (300, 3))
        on = c[pick[:20]]  # on a triangle
        cases.append((xyz, np.concatenate([near, far, on])))
    u = rng.uniform(-0.45, 0.45, (200, 2))
    dz = rng.choice([0.0, 1e-9, -1e-9, 1e-7, -1e-7, 1e-3, -1e-3, 0.1, -0.1], size=200)
    cases.append((cube, np.stack([u[:, 0], u[:, 1], 0.5 + dz], 1)))
    fast = 0
    total = 0
    for xyz, pts in cases:
        w, exact = oracle.gwn_f64_points(xyz, pts, threads=4)
        _, wb = oracle.sdf_f64_points(xyz, pts, threads=4)
        assert np.array_equal(np.abs(w) > 0.5, np.abs(wb) > 0.5)
        assert np.array_equal(np.abs(np.abs(w) - 0.5) < 1e-6, np.abs(np.abs(wb) - 0.5) < 1e-6)
        assert np.array_equal(w[exact], wb[exact])
        fast += int((~exact).sum())
        total += len(pts)
    assert fast > total // 2  # the tree decides most points on its own
    w, exact = oracle.gwn_f64_points(cube, np.array([[0.0, 0.0, 0.5]]), threads=1)  # w = 1/2 exactly
    assert exact[0] and abs(w[0] - 0.5) < 1e-12
