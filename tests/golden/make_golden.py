"""Writes the golden fixtures of this directory (inputs and expected outputs, .npz).

The expected values come from computations that share no code with the oracle or the device bake:
  cube_shell.npz      the unit cube shell of tests/voxelize_test.cpp:11-33 (inward wound):
                      - voxelizeMesh(mesh, 8) cells in first-touch order, decided by FP64 polygon
                        clipping of each triangle against each closed cell box (every coordinate is a
                        multiple of 1/8, so the clipping is exact);
                      - bakeSdf(mesh, 16) distances from the sdBox closed form (planar faces: the mesh
                        distance is the box distance, negative inside).
  uv_sphere.npz       the UV sphere of tests/sdf_sign_test.cpp:18-42 (r 1, 32 x 16, inward wound) at
                      bakeSdf resolution 24: FP64 brute-force point/triangle distance (vectorised
                      Voronoi-region form) and FP64 generalized winding number (Van Oosterom-Strackee),
                      inside where |w| > 0.5.
  icosphere_5120.npz  the 5,120-triangle jittered icosphere (level 4, r 0.5, seed 0x17A1E5) on the
                      64^3 grid over [-0.6, 0.6]^3: 4,096 seeded voxels with FP64 brute-force
                      distance and winding number.
Grids are stored as (origin, h, n) next to the expected values, so the tests need no framing code; the
query points are the FP32 cell centres the bake samples, and everything after them is FP64.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "italy-rs_amd"))


# ---- FP64 geometry, written for these fixtures ----------------------------------------------------
def seg_dist2(p, a, b):
    ab = b - a
    den = np.einsum("...k,...k->...", ab, ab)
    t = np.where(den > 0, np.einsum("...k,...k->...", p - a, ab) / np.where(den > 0, den, 1.0), 0.0)
    t = np.clip(t, 0.0, 1.0)
    d = p - (a + t[..., None] * ab)
    return np.einsum("...k,...k->...", d, d)


def tri_dist2(p, t0, t1, t2):
    """squared distance from points p (..., 3) to triangles (..., 3): the plane distance when the
    projection falls inside the triangle (three same-sign edge tests), else the nearest edge"""
    n = np.cross(t1 - t0, t2 - t0)
    nn = np.einsum("...k,...k->...", n, n)
    e = np.minimum(np.minimum(seg_dist2(p, t0, t1), seg_dist2(p, t1, t2)), seg_dist2(p, t2, t0))
    s0 = np.einsum("...k,...k->...", np.cross(t1 - t0, p - t0), n)
    s1 = np.einsum("...k,...k->...", np.cross(t2 - t1, p - t1), n)
    s2 = np.einsum("...k,...k->...", np.cross(t0 - t2, p - t2), n)
    inside = (nn > 0) & (s0 >= 0) & (s1 >= 0) & (s2 >= 0)
    h = np.einsum("...k,...k->...", p - t0, n)
    plane = np.where(nn > 0, h * h / np.where(nn > 0, nn, 1.0), np.inf)
    return np.where(inside, plane, e)


def solid_angle(p, t0, t1, t2):
    a, b, c = t0 - p, t1 - p, t2 - p
    la, lb, lc = (np.sqrt(np.einsum("...k,...k->...", v, v)) for v in (a, b, c))
    det = np.einsum("...k,...k->...", a, np.cross(b, c))
    den = (la * lb * lc + np.einsum("...k,...k->...", a, b) * lc + np.einsum("...k,...k->...", a, c) * lb
           + np.einsum("...k,...k->...", b, c) * la)
    return 2.0 * np.arctan2(det, den)


def brute_sdf(xyz, pts, chunk=1 << 21):
    """FP64 (distance, winding number) of every point against every triangle"""
    tri = np.asarray(xyz, np.float64).reshape(-1, 3, 3)
    pts = np.asarray(pts, np.float64).reshape(-1, 3)
    per = max(1, chunk // len(tri))
    d = np.empty(len(pts))
    w = np.empty(len(pts))
    for i in range(0, len(pts), per):
        p = pts[i:i + per, None, :]
        d[i:i + per] = np.sqrt(tri_dist2(p, tri[None, :, 0], tri[None, :, 1], tri[None, :, 2]).min(axis=1))
        w[i:i + per] = solid_angle(p, tri[None, :, 0], tri[None, :, 1], tri[None, :, 2]).sum(axis=1) / (4 * np.pi)
    return d, w


def clip_nonempty(tri, lo, hi):
    """does the triangle meet the closed box [lo, hi]? (Sutherland-Hodgman against the six planes)"""
    poly = [np.asarray(v, np.float64) for v in tri]
    for axis in range(3):
        for bound, sgn in ((lo[axis], 1.0), (hi[axis], -1.0)):
            out = []
            for i in range(len(poly)):
                a, b = poly[i], poly[(i + 1) % len(poly)]
                da, db = sgn * (a[axis] - bound), sgn * (b[axis] - bound)
                if da >= 0:
                    out.append(a)
                if (da >= 0) != (db >= 0):
                    out.append(a + da / (da - db) * (b - a))
            poly = out
            if not poly:
                return False
    return True


def centres(origin, h, n):
    """the points the bake samples: origin + (float(i) + 0.5f) * h in FP32 (src/render/kernels/sdf_bake.cu:34-37),
    x fastest; the FP64 evaluation then starts from these exact floats"""
    nx, ny, nz = n
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    c = np.stack([x, y, z], -1).reshape(-1, 3).astype(np.float32)
    return (np.asarray(origin, np.float32) + (c + np.float32(0.5)) * np.float32(h)).astype(np.float64)


def bake_framing(bmin, bmax, res):
    """bakeSdf framing (src/convert/sdf_baker.cpp:139-148) in FP32, as numpy evaluates it"""
    bmin, bmax = np.float32(bmin), np.float32(bmax)
    ext = bmax - bmin
    h = np.float32(max(float(ext.max()), 1e-6)) / np.float32(max(res, 1))
    origin = bmin - np.float32(2) * h
    n = [int(np.ceil(np.float32(ext[k] + np.float32(4) * h) / h)) + 1 for k in range(3)]
    return origin.astype(np.float32), np.float32(h), n


# ---- fixtures -------------------------------------------------------------------------------------
def cube_shell():
    from italy_bake import meshgen

    m = meshgen.cube_shell()
    xyz = m.xyz.reshape(-1, 9).astype(np.float32)
    res = 8
    h = 1.0 / res
    cells = []
    seen = set()
    for t in range(len(xyz)):  # first-touch order: triangle order, then z, y, x (voxelize.cpp:80-96)
        tri = xyz[t].reshape(3, 3).astype(np.float64)
        lo = np.clip(np.floor(tri.min(0) / h), 0, res - 1).astype(int)
        hi = np.clip(np.floor(tri.max(0) / h), 0, res - 1).astype(int)
        for z in range(lo[2], hi[2] + 1):
            for y in range(lo[1], hi[1] + 1):
                for x in range(lo[0], hi[0] + 1):
                    c = np.array([x, y, z], np.float64) * h
                    if (x, y, z) not in seen and clip_nonempty(tri, c, c + h):
                        seen.add((x, y, z))
                        cells.append((x, y, z))
    origin, sh, n = bake_framing(m.bmin, m.bmax, 16)
    q = centres(origin, sh, n) - 0.5
    qa = np.abs(q) - 0.5  # sdBox(p - centre, half = 0.5), sdf_primitives.h:12-15, in FP64
    sd = np.linalg.norm(np.maximum(qa, 0.0), axis=1) + np.minimum(qa.max(axis=1), 0.0)
    np.savez_compressed(os.path.join(HERE, "cube_shell.npz"), xyz=xyz, bmin=np.float32(m.bmin),
                        bmax=np.float32(m.bmax), vox_res=res, vox_cells=np.asarray(cells, np.int32),
                        sdf_res=16, sdf_origin=origin, sdf_h=sh, sdf_n=np.int32(n), sdf_f64=sd)
    return len(cells)


def uv_sphere():
    from italy_bake import meshgen

    m = meshgen.uv_sphere()
    xyz = m.xyz.reshape(-1, 9).astype(np.float32)
    origin, h, n = bake_framing(m.bmin, m.bmax, 24)
    d, w = brute_sdf(xyz, centres(origin, h, n))
    np.savez_compressed(os.path.join(HERE, "uv_sphere.npz"), xyz=xyz, bmin=np.float32(m.bmin),
                        bmax=np.float32(m.bmax), sdf_res=24, sdf_origin=origin, sdf_h=h, sdf_n=np.int32(n),
                        dist_f64=d, winding_f64=w)
    return int((np.abs(w) > 0.5).sum())


def icosphere():
    from italy_bake import meshgen

    m = meshgen.icosphere(4, 0.5, 1e-3, 0x17A1E5)
    xyz = m.xyz.reshape(-1, 9).astype(np.float32)
    lo, hi, res = np.float32(-0.6), np.float32(0.6), 64
    h = np.float32((hi - lo) / np.float32(res))
    origin = np.array([lo, lo, lo], np.float32)
    rng = np.random.default_rng(0x17A1E5)
    idx = np.sort(rng.choice(res ** 3, 4096, replace=False)).astype(np.int64)
    d, w = brute_sdf(xyz, centres(origin, h, (res, res, res))[idx])
    np.savez_compressed(os.path.join(HERE, "icosphere_5120.npz"), xyz=xyz, sdf_origin=origin, sdf_h=h,
                        sdf_n=np.int32([res, res, res]), voxel_index=idx, dist_f64=d, winding_f64=w)
    return len(xyz)


if __name__ == "__main__":
    print("cube_shell: first-touch cells", cube_shell())
    print("uv_sphere: inside voxels", uv_sphere())
    print("icosphere_5120: triangles", icosphere())
