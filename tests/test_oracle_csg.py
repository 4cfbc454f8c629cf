"""Pins of the oracle's analytic SDF restatement (sdf_primitives.h:10-36, sdf_procedural.cpp:12-140).

The reference has no tests or stored grids for this path, so each primitive is checked against a
different computation of the same distance (nearest point on a densely sampled surface, exact 2D
point-to-segment distances, the float64 closed form), each operator against the properties that fix
it (smooth union is min outside the blend band and min - k/4 on the diagonal), the program
encoding of the reference's cups against the lambdas written out per cell, and the framing
against worked sizes.  A plausible slip (a swapped operand, a missing abs, max for min, a wrong
offset) fails one of these."""
import numpy as np
import pytest

from italy_bake import procedural as pr

f32 = np.float32


def _pts(n, seed, lo=-1.5, hi=1.5):
    return np.random.default_rng(seed).uniform(lo, hi, size=(n, 3)).astype(np.float32)


def _eval(oracle, prog, pts):
    out = [oracle.csg_eval(prog, p) for p in pts]
    return np.array([d for d, _ in out], np.float32), np.array([b for _, b in out])


def test_sphere_closed_form(oracle):
    c, r = (0.2, -0.1, 0.3), 0.7
    pts = _pts(400, 1)
    d, b = _eval(oracle, pr.Csg().sphere(c, r, branch=7).array(), pts)
    exact = np.linalg.norm(pts.astype(np.float64) - np.float32(c).astype(np.float64), axis=1) - np.float32(r)
    assert np.all(b == 7)
    assert np.abs(d - exact).max() < 4e-7 * 3


def test_box_against_nearest_point(oracle):
    c, e = np.array([0.1, 0.2, -0.3], np.float32), np.array([0.5, 0.3, 0.8], np.float32)
    pts = _pts(500, 2)
    d, _ = _eval(oracle, pr.Csg().box(c, e).array(), pts)
    p = pts.astype(np.float64) - c
    lo, hi = -e.astype(np.float64), e.astype(np.float64)
    outside = np.any((p < lo) | (p > hi), axis=1)
    ref = np.where(outside, np.linalg.norm(p - np.clip(p, lo, hi), axis=1),
                   -np.min(np.concatenate([p - lo, hi - p], axis=1), axis=1))
    assert outside.any() and (~outside).any()
    assert np.abs(d - ref).max() < 2e-6


def _seg_dist(px, py, ax, ay, bx, by):
    vx, vy = bx - ax, by - ay
    t = np.clip(((px - ax) * vx + (py - ay) * vy) / (vx * vx + vy * vy), 0, 1)
    return np.hypot(px - (ax + t * vx), py - (ay + t * vy))


def test_cylinder_against_profile_segments(oracle):
    # a capped cylinder about +y is the revolution of the (radius, y) rectangle: the distance to its
    # surface is the 2D distance to the side and the two caps, negative inside
    c, r, h = np.array([0.0, 0.1, 0.2], np.float32), f32(0.6), f32(0.4)
    pts = _pts(500, 3)
    d, _ = _eval(oracle, pr.Csg().cylinder(c, r, h).array(), pts)
    p = pts.astype(np.float64) - c
    rho, y = np.hypot(p[:, 0], p[:, 2]), p[:, 1]
    r64, h64 = float(r), float(h)
    surf = np.minimum.reduce([_seg_dist(rho, y, r64, -h64, r64, h64), _seg_dist(rho, y, 0, h64, r64, h64),
                              _seg_dist(rho, y, 0, -h64, r64, -h64)])
    inside = (rho < r64) & (np.abs(y) < h64)
    ref = np.where(inside, -surf, surf)
    assert inside.any() and (~inside).any()
    assert np.abs(d - ref).max() < 2e-6


def test_torus_against_sampled_ring(oracle):
    c, R, r = np.array([0.1, -0.2, 0.0], np.float32), f32(0.7), f32(0.2)
    pts = _pts(300, 4)
    d, _ = _eval(oracle, pr.Csg().torus(c, R, r).array(), pts)
    ang = np.linspace(0, 2 * np.pi, 200001)
    ring = np.stack([float(R) * np.cos(ang), np.zeros_like(ang), float(R) * np.sin(ang)], 1)
    p = pts.astype(np.float64) - c
    ref = np.array([np.sqrt(((ring - q) ** 2).sum(1).min()) for q in p]) - float(r)
    assert np.abs(d - ref).max() < 2e-5  # ring sampling step 2.2e-5


def test_plane_is_signed_distance(oracle):
    n = np.array([0.6, 0.0, 0.8], np.float32)
    pts = _pts(200, 5)
    d, _ = _eval(oracle, pr.Csg().plane(n, 0.25).array(), pts)
    assert np.abs(d - (pts.astype(np.float64) @ n.astype(np.float64) - 0.25)).max() < 1e-6
    # the axis plane used by the cups gives p.y - h exactly
    d, _ = _eval(oracle, pr.Csg().plane((0, 1, 0), 0.3).array(), pts)
    assert np.array_equal(d, pts[:, 1] - f32(0.3))


def _two_spheres(op, *k):
    g = pr.Csg().sphere((-0.3, 0, 0), 0.5, branch=1).sphere((0.3, 0, 0), 0.5, branch=2)
    return getattr(g, op)(*k).array()


def test_boolean_operators_and_branches(oracle):
    pts = _pts(400, 6, -1, 1)
    da, _ = _eval(oracle, pr.Csg().sphere((-0.3, 0, 0), 0.5).array(), pts)
    db, _ = _eval(oracle, pr.Csg().sphere((0.3, 0, 0), 0.5).array(), pts)
    u, bu = _eval(oracle, _two_spheres("union"), pts)
    i, bi = _eval(oracle, _two_spheres("intersect"), pts)
    s, bs = _eval(oracle, _two_spheres("subtract"), pts)
    sel, bsel = _eval(oracle, _two_spheres("select"), pts)
    ina, inb = da < 0, db < 0
    assert np.array_equal(u < 0, ina | inb) and np.array_equal(i < 0, ina & inb)
    assert np.array_equal(s < 0, ina & ~(db <= 0))
    assert np.array_equal(u, np.minimum(da, db)) and np.array_equal(i, np.maximum(da, db))
    assert np.array_equal(bu, np.where(db < da, 2, 1)) and np.array_equal(bi, np.where(da < db, 2, 1))
    assert np.all(bs == 1)
    assert np.array_equal(sel, u) and np.array_equal(bsel, np.where(da <= db, 1, 2))
    # ties: select and union keep the first operand
    d, b = oracle.csg_eval(_two_spheres("select"), (0.0, 0.3, 0.1))
    assert b == 1 and oracle.csg_eval(_two_spheres("union"), (0.0, 0.3, 0.1))[1] == 1


def test_smooth_union_properties(oracle):
    k = 0.2
    pts = _pts(600, 7, -1, 1)
    da, _ = _eval(oracle, pr.Csg().sphere((-0.3, 0, 0), 0.5).array(), pts)
    db, _ = _eval(oracle, pr.Csg().sphere((0.3, 0, 0), 0.5).array(), pts)
    su, _ = _eval(oracle, _two_spheres("smooth_union", k), pts)
    si, _ = _eval(oracle, _two_spheres("smooth_intersect", k), pts)
    far = np.abs(da - db) >= f32(k) * 1.0001
    assert far.any() and (~far).any()
    # outside the blend band the blend weight is 0 or 1: exactly the hard min / max
    assert np.array_equal(su[far], np.minimum(da, db)[far])
    assert np.array_equal(si[far], np.maximum(da, db)[far])
    # inside it the blend lies between min - k/4 and min (max and max + k/4)
    lo, hi = np.minimum(da, db), np.maximum(da, db)
    assert np.all(su <= lo + 1e-6) and np.all(su >= lo - k / 4 - 1e-6)
    assert np.all(si >= hi - 1e-6) and np.all(si <= hi + k / 4 + 1e-6)
    # on the bisecting plane x = 0 (a == b) it is min - k/4
    on = np.array([[0.0, 0.37, -0.21], [0.0, -0.1, 0.05]], np.float32)
    for q in on:
        a, _ = oracle.csg_eval(pr.Csg().sphere((-0.3, 0, 0), 0.5).array(), q)
        assert oracle.csg_eval(_two_spheres("smooth_union", k), q)[0] == f32(a - f32(f32(k) * 0.25))


def test_malformed_programs_give_nan(oracle):
    assert np.isnan(oracle.csg_eval(pr.Csg().sphere((0, 0, 0), 1).union().array(), (0, 0, 0))[0])
    assert np.isnan(oracle.csg_eval(pr.Csg().sphere((0, 0, 0), 1).sphere((0, 0, 0), 1).array(), (0, 0, 0))[0])


def test_framing_worked_sizes(oracle):
    o, h, n = oracle.analytic_framing((0, 0, 0), (1, 1, 1), 4)
    assert h == f32(0.25) and n == (8, 8, 8) and np.array_equal(o, np.full(3, -0.5, np.float32))
    o, h, n = oracle.analytic_framing((0, 0, 0), (2, 1, 0.5), 2)  # resolution clamps to 4
    assert h == f32(0.5) and n == (8, 6, 5)
    # the glass cup of the viewer at 320 (main.cpp:116, :261): 1.1 x 0.891 x 1.1 -> 324 x 264 x 324
    _, bmin, bmax = pr.cup_program(-1.0, (0.0, 0.0), glass=True)
    o, h, n = oracle.analytic_framing(bmin, bmax, 320)
    assert n == (324, 264, 324) and h == f32(f32(1.1) / f32(320))


def test_library_framing_matches_oracle(oracle):
    rng = np.random.default_rng(8)
    for _ in range(50):
        lo = rng.uniform(-3, 3, 3).astype(np.float32)
        hi = (lo + rng.uniform(0.01, 4, 3)).astype(np.float32)
        res = int(rng.integers(1, 400))
        g = pr.analytic_framing(lo, hi, res)
        o, h, n = oracle.analytic_framing(lo, hi, res)
        assert (g.nx, g.ny, g.nz) == n and f32(g.h) == h and np.array_equal(np.array(g.origin[:], np.float32), o)


@pytest.mark.parametrize("glass", [False, True])
def test_cup_programs_match_the_written_out_lambdas(oracle, glass):
    prog, bmin, bmax = pr.cup_program(-1.0, (-1.6, 1.6) if not glass else (0.0, 0.0), glass=glass)
    o, h, n = oracle.analytic_framing(bmin, bmax, 64)
    dist, branch = oracle.bake_csg(prog.array(), o, h, *n, threads=4)
    o2, h2, dist2, branch2 = oracle.cup_cells(glass, 64, -1.0, (-1.6, 1.6) if not glass else (0.0, 0.0))
    assert np.array_equal(o, o2) and h == h2
    assert np.array_equal(dist, dist2) and np.array_equal(branch, branch2)
    assert (dist < 0).any() and (dist > 0).any()
    if glass:
        assert set(np.unique(branch)) == {0, 1}


def test_glass_cup_geometry(oracle):
    """Points whose side of the surface follows from the cup's dimensions (sdf_procedural.cpp:108-117)."""
    prog = pr.cup_program(-1.0, (0.0, 0.0), glass=True)[0].array()
    # outer sphere centre y = -0.45, radius 0.55; inner sphere centre y = -0.29, radius 0.46
    wall = oracle.csg_eval(prog, (0.0, -0.45 - 0.5, 0.0))        # between the two spheres, below the rim
    assert wall[0] < 0 and wall[1] == 0
    liquid = oracle.csg_eval(prog, (0.0, -0.6, 0.0))              # inside the inner sphere, below the fill
    assert liquid[0] < 0 and liquid[1] == 1
    air = oracle.csg_eval(prog, (0.0, -0.2, 0.0))                 # inside the cup, above the fill line
    assert air[0] > 0
    out = oracle.csg_eval(prog, (0.0, -0.45 + 0.62 * 0.55 + 0.05, 0.0))  # above the rim
    assert out[0] > 0
