"""The renderer's grid consumer on the device (ib_sample_grid_device, ib_trace_grid_device) against
the oracle's restatement of pathtracer.cu:587-613, :684-709, :740-846, on a grid baked on the
device: values, normals, nearest cells and traced hit distances bit-exact; traced hits agree with
exact ray/mesh intersections to a quarter voxel (the baked grid as a verifier of itself)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _baked_small(ctx, meshgen):
    from italy_bake import bake

    mesh, cfg = meshgen.workload_plain_small()
    g = bake.dense_grid(cfg["lo"], cfg["hi"], cfg["sdf_res"])
    dist, _ = ctx.bake_dense(mesh.xyz, g, None)
    return mesh, g, dist.reshape(g.nz, g.ny, g.nx)


def _ray_mesh(org, dirs, xyz):
    """Nearest positive hit of each ray with the triangle soup (Moller-Trumbore in float64)."""
    tri = xyz.reshape(-1, 3, 3).astype(np.float64)
    e1, e2 = tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]
    out = np.full(len(org), np.inf)
    for i, (o, d) in enumerate(zip(org.astype(np.float64), dirs.astype(np.float64))):
        p = np.cross(d, e2)
        det = np.einsum("ij,ij->i", e1, p)
        ok = np.abs(det) > 1e-15
        inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
        s = o - tri[:, 0]
        u = np.einsum("ij,ij->i", s, p) * inv
        q = np.cross(s, e1)
        v = (q @ d) * inv
        t = np.einsum("ij,ij->i", e2, q) * inv
        hit = ok & (u >= 0) & (v >= 0) & (u + v <= 1) & (t > 0)
        if hit.any():
            out[i] = t[hit].min()
    return out


def test_sample_grid_matches_oracle(ctx, oracle, meshgen):
    mesh, g, dist = _baked_small(ctx, meshgen)
    rng = np.random.default_rng(21)
    pts = rng.uniform(-0.7, 0.7, (20000, 3)).astype(np.float32)  # includes points outside the grid
    vals, nrm, near = ctx.sample_grid(g, dist, pts)
    rv, rn, rnear = oracle.sample_grid(list(g.origin), g.h, dist, pts)
    assert np.array_equal(vals.view(np.uint32), rv.view(np.uint32))
    assert np.array_equal(nrm.view(np.uint32), rn.view(np.uint32))
    assert np.array_equal(near, rnear)
    # near the surface the B-spline normal points away from the centre of the icosphere
    c = np.abs(vals) < 0.02
    radial = pts[c] / np.linalg.norm(pts[c], axis=1, keepdims=True)
    assert c.sum() > 100 and (np.einsum("ij,ij->i", nrm[c], radial) > 0.9).all()


def test_trace_grid_matches_oracle_and_mesh(ctx, oracle, meshgen):
    mesh, g, dist = _baked_small(ctx, meshgen)
    rng = np.random.default_rng(22)
    u = rng.normal(size=(400, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    org = (0.58 * u).astype(np.float32)
    aim = rng.uniform(-0.2, 0.2, (400, 3))
    dirs = aim - org
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    t = ctx.trace_grid(g, dist, org, dirs)
    rt = oracle.trace_grid(list(g.origin), g.h, dist, org, dirs)
    assert np.array_equal(t.view(np.uint32), rt.view(np.uint32))
    exact = _ray_mesh(org, dirs, mesh.xyz)
    hit = np.isfinite(exact)
    assert hit.sum() > 350 and (t[hit] > 0).all()
    assert np.abs(t[hit] - exact[hit]).max() < 0.25 * g.h
