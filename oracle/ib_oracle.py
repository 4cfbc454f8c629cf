"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker.  The product package never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)


class BakeStats(ctypes.Structure):
    _fields_ = [
        ("voxels", ctypes.c_uint64),
        ("node_visits", ctypes.c_uint64),
        ("tri_tests", ctypes.c_uint64),
        ("column_tri_tests", ctypes.c_uint64),
        ("inside", ctypes.c_uint64),
        ("seconds", ctypes.c_double),
        ("threads", ctypes.c_int),
    ]


class Bvh(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("order", _i32p),
        ("codes", _u32p),
        ("left", _i32p),
        ("right", _i32p),
        ("parent", _i32p),
        ("lo", _f32p),
        ("hi", _f32p),
        ("depth", ctypes.c_int32),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = ctypes.CDLL(_LIB)
    lib.orc_tri_box_overlap.argtypes = [_f32p] * 5
    lib.orc_tri_box_overlap.restype = ctypes.c_int
    lib.orc_sdf_framing.argtypes = [_f32p, _f32p, ctypes.c_int, _f32p, _f32p, _i32p]
    lib.orc_vox_framing.argtypes = [_f32p, _f32p, ctypes.c_int, _f32p, _f32p]
    lib.orc_voxelize_first_touch.argtypes = [_f32p, ctypes.c_size_t, _f32p, _f32p, _f32p, ctypes.c_int, _i32p,
                                             _f32p, ctypes.c_size_t]
    lib.orc_voxelize_first_touch.restype = ctypes.c_size_t
    lib.orc_sample_grid.argtypes = [_f32p, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, _f32p,
                                    ctypes.c_size_t, _f32p, _f32p, _i32p]
    lib.orc_sample_grid.restype = None
    lib.orc_trace_grid.argtypes = [_f32p, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, _f32p,
                                   _f32p, ctypes.c_size_t, ctypes.c_float, ctypes.c_float, _f32p]
    lib.orc_trace_grid.restype = None
    lib.orc_splat_materials.argtypes = [_f32p, ctypes.c_size_t, _f32p, _f32p, ctypes.c_float, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, _f32p]
    lib.orc_splat_materials.restype = None
    lib.orc_occupancy.argtypes = [_f32p, ctypes.c_size_t, _f32p, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, _u32p]
    lib.orc_tri_is_sliver.argtypes = [_f32p]
    lib.orc_tri_dist2_f32.argtypes = [_f32p, _f32p]
    lib.orc_tri_dist2_f32.restype = ctypes.c_float
    lib.orc_tri_dist2_f64.argtypes = [_f64p, _f32p]
    lib.orc_tri_dist2_f64.restype = ctypes.c_double
    lib.orc_gwn_f64.argtypes = [_f64p, _f32p, ctypes.c_size_t]
    lib.orc_gwn_f64.restype = ctypes.c_double
    lib.orc_crossings_up.argtypes = [_f32p, _f32p, ctypes.c_size_t]
    lib.orc_parity_inside.argtypes = [_f32p, _f32p, ctypes.c_size_t]
    lib.orc_parity_dirs.argtypes = [_f32p]
    lib.orc_parity_hits.argtypes = [_f32p, _f32p, ctypes.c_size_t, _i32p]
    lib.orc_gwn_sign.argtypes = [_f32p, _f32p, ctypes.c_size_t, _i32p]
    lib.orc_signed_points_gwn.argtypes = [_f32p, ctypes.c_size_t, _f32p, ctypes.c_size_t, ctypes.c_int, _f32p,
                                          ctypes.c_void_p]
    lib.orc_signed_points_closed.argtypes = [_f32p, ctypes.c_size_t, _f32p, ctypes.c_size_t, ctypes.c_int, _f32p]
    lib.orc_bake_band.argtypes = [_f32p, ctypes.c_size_t, _f32p, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _f32p,
                                  ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(BandStats)]
    lib.orc_bvh_build.argtypes = [_f32p, ctypes.c_size_t, ctypes.POINTER(Bvh)]
    lib.orc_bvh_free.argtypes = [ctypes.POINTER(Bvh)]
    lib.orc_morton_codes.argtypes = [_f32p, ctypes.c_size_t, _u32p]
    lib.orc_bake_dense.argtypes = [_f32p, ctypes.c_size_t, _f32p, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p,
                                   ctypes.POINTER(BakeStats)]
    lib.orc_bake_dense_strided.argtypes = [_f32p, ctypes.c_size_t, _f32p, ctypes.c_float, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, _f32p, ctypes.POINTER(BakeStats)]
    lib.orc_sdf_f64_points.argtypes = [_f32p, ctypes.c_size_t, _f64p, ctypes.c_size_t, ctypes.c_int, _f64p, _f64p]
    lib.orc_gwn_f64_points.argtypes = [_f32p, ctypes.c_size_t, _f64p, ctypes.c_size_t, ctypes.c_int, _f64p,
                                       ctypes.POINTER(ctypes.c_uint8)]
    lib.orc_udf2_f32_brute.argtypes = [_f32p, ctypes.c_size_t, _f32p, ctypes.c_size_t, _f32p]
    lib.orc_udf_points_f32.argtypes = [_f32p, ctypes.c_size_t, _f32p, ctypes.c_size_t, ctypes.c_int, _f32p]
    lib.orc_boundary_edge_count.argtypes = [_f32p, ctypes.c_size_t]
    lib.orc_f32_error_ratio.argtypes = [ctypes.c_long, ctypes.c_uint64, _f64p]
    lib.orc_f32_error_ratio.restype = None
    lib.orc_csg_eval.argtypes = [ctypes.c_void_p, ctypes.c_int, _f32p, _i32p]
    lib.orc_csg_eval.restype = ctypes.c_float
    lib.orc_analytic_framing.argtypes = [_f32p, _f32p, ctypes.c_int, _f32p, _f32p, _i32p]
    lib.orc_analytic_framing.restype = None
    lib.orc_bake_csg.argtypes = [ctypes.c_void_p, ctypes.c_int, _f32p, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_void_p]
    lib.orc_bake_csg.restype = None
    lib.orc_cup_cells.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, _f32p,
                                  _f32p, _i32p, _f32p, ctypes.c_void_p]
    lib.orc_cup_cells.restype = None
    lib.orc_boundary_edge_count.restype = ctypes.c_int64
    lib.orc_udf_points_f64.argtypes = [_f32p, ctypes.c_size_t, _f64p, ctypes.c_size_t, ctypes.c_int, _f64p]
    lib.orc_udf_points_f64.restype = ctypes.c_int
    lib.orc_triangle_materials.argtypes = [ctypes.c_size_t, _f32p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                           ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, _f32p]
    lib.orc_triangle_materials.restype = None
    lib.orc_weld_positions.argtypes = [_f32p, ctypes.c_size_t, _f32p, _f32p, ctypes.c_void_p, ctypes.c_void_p]
    lib.orc_weld_positions.restype = ctypes.c_int64
    lib.orc_watertight.argtypes = [_f32p, ctypes.c_size_t, _f32p, _f32p, ctypes.POINTER(ctypes.c_int64),
                                   ctypes.POINTER(ctypes.c_int64)]
    lib.orc_watertight.restype = ctypes.c_int
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def tri_box_overlap(center, half, t0, t1, t2):
    args = [_f32(x) for x in (center, half, t0, t1, t2)]
    return bool(lib().orc_tri_box_overlap(*[_p(a, _f32p) for a in args]))


def sdf_framing(bmin, bmax, res):
    o = np.zeros(3, np.float32)
    h = np.zeros(1, np.float32)
    n = np.zeros(3, np.int32)
    lib().orc_sdf_framing(_p(_f32(bmin), _f32p), _p(_f32(bmax), _f32p), int(res), _p(o, _f32p), _p(h, _f32p),
                          _p(n, _i32p))
    return o, np.float32(h[0]), tuple(int(v) for v in n)


def vox_framing(bmin, bmax, res):
    o = np.zeros(3, np.float32)
    h = np.zeros(1, np.float32)
    lib().orc_vox_framing(_p(_f32(bmin), _f32p), _p(_f32(bmax), _f32p), int(res), _p(o, _f32p), _p(h, _f32p))
    return o, np.float32(h[0])


def voxelize_first_touch(xyz, bmin, bmax, res, colors=None):
    xyz = _f32(xyz)
    ntri = xyz.size // 9
    cp = _p(_f32(colors), _f32p) if colors is not None else None
    n = lib().orc_voxelize_first_touch(_p(xyz, _f32p), ntri, cp, _p(_f32(bmin), _f32p), _p(_f32(bmax), _f32p),
                                       int(res), None, None, 0)
    cells = np.zeros((max(n, 1), 3), np.int32)
    cols = np.zeros((max(n, 1), 3), np.float32)
    lib().orc_voxelize_first_touch(_p(xyz, _f32p), ntri, cp, _p(_f32(bmin), _f32p), _p(_f32(bmax), _f32p),
                                   int(res), _p(cells, _i32p), _p(cols, _f32p), n)
    return cells[:n], cols[:n]


def occupancy(xyz, origin, h, nx, ny, nz):
    xyz = _f32(xyz)
    words = np.zeros(((nx + 31) // 32) * ny * nz, np.uint32)
    lib().orc_occupancy(_p(xyz, _f32p), xyz.size // 9, _p(_f32(origin), _f32p), float(h), nx, ny, nz,
                        _p(words, _u32p))
    return words


def splat_materials(xyz, mat, origin, h, nx, ny, nz):
    """bakePerCellMaterial (sdf_baker.cpp:76-128): (5, nz, ny, nx) planes r, g, b, metallic, roughness."""
    xyz = _f32(xyz)
    mat = _f32(mat)
    out = np.zeros((5, nz, ny, nx), np.float32)
    lib().orc_splat_materials(_p(xyz, _f32p), xyz.size // 9, _p(mat, _f32p), _p(_f32(origin), _f32p), float(h),
                              nx, ny, nz, _p(out, _f32p))
    return out


def sample_grid(origin, h, dist, pts):
    """The renderer's grid sampling (pathtracer.cu:684-709, :794-846): (values, unit normals, nearest
    cell) at pts (n, 3); dist is (nz, ny, nx)."""
    dist = _f32(dist)
    nz, ny, nx = dist.shape
    pts = _f32(pts).reshape(-1, 3)
    n = len(pts)
    vals = np.zeros(n, np.float32)
    nrm = np.zeros((n, 3), np.float32)
    near = np.zeros(n, np.int32)
    lib().orc_sample_grid(_p(_f32(origin), _f32p), float(h), nx, ny, nz, _p(dist, _f32p), _p(pts, _f32p), n,
                          _p(vals, _f32p), _p(nrm, _f32p), _p(near, _i32p))
    return vals, nrm, near


def trace_grid(origin, h, dist, org, dirs, tmin=0.0, tmax=1e16):
    """The SDF intersection program's march (pathtracer.cu:587-613, :740-792): hit t per ray, -1 = miss."""
    dist = _f32(dist)
    nz, ny, nx = dist.shape
    org = _f32(org).reshape(-1, 3)
    dirs = _f32(dirs).reshape(-1, 3)
    t = np.zeros(len(org), np.float32)
    lib().orc_trace_grid(_p(_f32(origin), _f32p), float(h), nx, ny, nz, _p(dist, _f32p), _p(org, _f32p),
                         _p(dirs, _f32p), len(org), float(tmin), float(tmax), _p(t, _f32p))
    return t


def f32_error_ratio(n=200000, seed=1):
    """Worst (exact - FP32) and (FP32 - exact) triangle distance in units of 2^-24 (Mq + 3 Mp)."""
    out = np.zeros(2, np.float64)
    lib().orc_f32_error_ratio(int(n), int(seed), _p(out, _f64p))
    return float(out[0]), float(out[1])


CSG_DTYPE = np.dtype([("op", "<i4"), ("branch", "<i4"), ("a", "<f4", (8,))])


def _csg(prog):
    return np.ascontiguousarray(prog, CSG_DTYPE)


def csg_eval(prog, p):
    """One point of a CSG program (sdf_primitives.h:10-36): (distance, branch); NaN if malformed."""
    prog = _csg(prog)
    b = ctypes.c_int32(-1)
    d = lib().orc_csg_eval(prog.ctypes.data_as(ctypes.c_void_p), len(prog), _p(_f32(p), _f32p), ctypes.byref(b))
    return np.float32(d), b.value


def analytic_framing(bmin, bmax, res):
    """bakeSdfGridShape (sdf_procedural.cpp:12-26): (origin, h, (nx, ny, nz))."""
    o = np.zeros(3, np.float32)
    h = np.zeros(1, np.float32)
    n = np.zeros(3, np.int32)
    lib().orc_analytic_framing(_p(_f32(bmin), _f32p), _p(_f32(bmax), _f32p), int(res), _p(o, _f32p), _p(h, _f32p),
                               _p(n, _i32p))
    return o, np.float32(h[0]), tuple(int(v) for v in n)


def bake_csg(prog, origin, h, nx, ny, nz, threads=1):
    """Every cell centre of the grid (sdf_procedural.cpp:28-44): (distances (nz, ny, nx), branch)."""
    prog = _csg(prog)
    dist = np.empty((nz, ny, nx), np.float32)
    branch = np.empty((nz, ny, nx), np.uint8)
    lib().orc_bake_csg(prog.ctypes.data_as(ctypes.c_void_p), len(prog), _p(_f32(origin), _f32p), float(h), nx, ny,
                       nz, int(threads), _p(dist, _f32p), branch.ctypes.data_as(ctypes.c_void_p))
    return dist, branch


def cup_cells(glass, res, origin_y=-1.0, center_xz=(0.0, 0.0)):
    """makeCupSdf / makeGlassCupSdf (sdf_procedural.cpp:81-140) written out per cell, without a program:
    (origin, h, distances (nz, ny, nx), branch)."""
    o = np.zeros(3, np.float32)
    h = np.zeros(1, np.float32)
    n = np.zeros(3, np.int32)
    lib().orc_cup_cells(int(glass), int(res), float(origin_y), float(center_xz[0]), float(center_xz[1]),
                        _p(o, _f32p), _p(h, _f32p), _p(n, _i32p), None, None)
    dist = np.empty((n[2], n[1], n[0]), np.float32)
    branch = np.empty((n[2], n[1], n[0]), np.uint8)
    lib().orc_cup_cells(int(glass), int(res), float(origin_y), float(center_xz[0]), float(center_xz[1]),
                        _p(o, _f32p), _p(h, _f32p), _p(n, _i32p), _p(dist, _f32p),
                        branch.ctypes.data_as(ctypes.c_void_p))
    return o, np.float32(h[0]), dist, branch


def tri_dist2_f32(q, tri):
    return float(lib().orc_tri_dist2_f32(_p(_f32(q), _f32p), _p(_f32(tri), _f32p)))


def tri_dist2_f64(q, tri):
    return float(lib().orc_tri_dist2_f64(_p(np.ascontiguousarray(q, np.float64), _f64p), _p(_f32(tri), _f32p)))


def tri_is_sliver(tri):
    return bool(lib().orc_tri_is_sliver(_p(_f32(tri), _f32p)))


def gwn_f64(q, xyz):
    xyz = _f32(xyz)
    return float(lib().orc_gwn_f64(_p(np.ascontiguousarray(q, np.float64), _f64p), _p(xyz, _f32p), xyz.size // 9))


def crossings_up(q, xyz):
    xyz = _f32(xyz)
    return int(lib().orc_crossings_up(_p(_f32(q), _f32p), _p(xyz, _f32p), xyz.size // 9))


def parity_inside(q, xyz):
    xyz = _f32(xyz)
    return bool(lib().orc_parity_inside(_p(_f32(q), _f32p), _p(xyz, _f32p), xyz.size // 9))


def parity_hits(q, xyz):
    """Hits of each of the 5 parity rays from q (t > 0)."""
    xyz = _f32(xyz)
    h = np.zeros(5, np.int32)
    lib().orc_parity_hits(_p(_f32(q), _f32p), _p(xyz, _f32p), xyz.size // 9, _p(h, _i32p))
    return h


def gwn_sign(q, xyz):
    """(inside, tie): the bake's sign rule with the FP64 brute-force winding number, parity on ties."""
    xyz = _f32(xyz)
    t = np.zeros(1, np.int32)
    r = lib().orc_gwn_sign(_p(_f32(q), _f32p), _p(xyz, _f32p), xyz.size // 9, _p(t, _i32p))
    return bool(r), bool(t[0])


def signed_points_gwn(xyz, pts, threads=0):
    """(signed FP32 distances, tie flags) at FP32 points of any mesh (FP64 brute-force winding sign)."""
    xyz = _f32(xyz)
    pts = _f32(pts).reshape(-1, 3)
    out = np.empty(len(pts), np.float32)
    tie = np.zeros(len(pts), np.uint8)
    r = lib().orc_signed_points_gwn(_p(xyz, _f32p), xyz.size // 9, _p(pts, _f32p), len(pts), int(threads),
                                    _p(out, _f32p), tie.ctypes.data)
    if r != 0:
        raise ValueError("empty mesh")
    return out, tie.astype(bool)


class BandStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("voxels", "bricks_baked", "cap_triangles", "rechecked",
                                               "full_cap_sums", "on_cap", "ties")] + \
               [("seconds", ctypes.c_double), ("threads", ctypes.c_int)]


def bake_band(xyz, origin, h, nx, ny, nz, band, stride=1, threads=0):
    """Narrow band of any mesh: (brick index (nz/8, ny/8, nx/8) int32, pool (nbricks, 8, 8, 8) float32 with
    NaN in the slots not baked (every stride-th slot is), stats dict).  See orc_bake_band."""
    xyz = _f32(xyz)
    nb = (nx // 8) * (ny // 8) * (nz // 8)
    idx = np.empty(nb, np.int32)
    n = ctypes.c_size_t()
    o = _f32(origin)
    st = BandStats()
    args = (_p(xyz, _f32p), xyz.size // 9, _p(o, _f32p), float(h), nx, ny, nz, int(band), int(stride), int(threads),
            _p(idx, _i32p))
    dummy = np.zeros(1, np.float32)
    rc = lib().orc_bake_band(*args, _p(dummy, _f32p), 0, ctypes.byref(n), ctypes.byref(st))
    if rc == -2:
        raise ValueError("empty mesh or grid dims not divisible by 8")
    pool = np.full((max(n.value, 1), 8, 8, 8), np.nan, np.float32)
    rc = lib().orc_bake_band(*args, _p(pool, _f32p), n.value, ctypes.byref(n), ctypes.byref(st))
    assert rc == 0
    stats = {k: getattr(st, k) for k, _ in BandStats._fields_}
    return idx.reshape(nz // 8, ny // 8, nx // 8), pool[: n.value], stats


def signed_points_closed(xyz, pts, threads=0):
    """The dense CPU bake's signed distances at FP32 points of a closed mesh (None if the mesh is open)."""
    xyz = _f32(xyz)
    pts = _f32(pts).reshape(-1, 3)
    out = np.empty(len(pts), np.float32)
    r = lib().orc_signed_points_closed(_p(xyz, _f32p), xyz.size // 9, _p(pts, _f32p), len(pts), int(threads),
                                       _p(out, _f32p))
    if r == -1:
        return None
    if r != 0:
        raise ValueError("empty mesh")
    return out


def cell_centres_at(origin, h, idx, nx, ny):
    """FP32 centres, as the bake computes them, of the cells with linear indices idx ((z*ny + y)*nx + x)."""
    idx = np.asarray(idx, np.int64)
    x, y, z = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    h = np.float32(h)
    o = [np.float32(v) for v in origin]
    c = [o[k] + (v.astype(np.float32) + np.float32(0.5)) * h for k, v in enumerate((x, y, z))]
    return np.stack(c, -1).astype(np.float32)


def parity_dirs():
    d = np.zeros(15, np.float32)
    lib().orc_parity_dirs(_p(d, _f32p))
    return d.reshape(5, 3)


def morton_codes(xyz):
    xyz = _f32(xyz)
    out = np.zeros(xyz.size // 9, np.uint32)
    lib().orc_morton_codes(_p(xyz, _f32p), xyz.size // 9, _p(out, _u32p))
    return out


def bvh(xyz):
    """Karras LBVH: returns dict of numpy arrays (order, codes, left, right, parent, lo, hi, depth)."""
    xyz = _f32(xyz)
    b = Bvh()
    n = xyz.size // 9
    if lib().orc_bvh_build(_p(xyz, _f32p), n, ctypes.byref(b)) != 0:
        raise ValueError("empty mesh")
    ni = max(n - 1, 0)
    out = {
        "order": np.ctypeslib.as_array(b.order, (n,)).copy(),
        "codes": np.ctypeslib.as_array(b.codes, (n,)).copy(),
        "left": np.ctypeslib.as_array(b.left, (max(ni, 1),)).copy()[:ni],
        "right": np.ctypeslib.as_array(b.right, (max(ni, 1),)).copy()[:ni],
        "parent": np.ctypeslib.as_array(b.parent, (2 * n - 1,)).copy(),
        "lo": np.ctypeslib.as_array(b.lo, (3 * (2 * n - 1),)).copy().reshape(-1, 3),
        "hi": np.ctypeslib.as_array(b.hi, (3 * (2 * n - 1),)).copy().reshape(-1, 3),
        "depth": int(b.depth),
    }
    lib().orc_bvh_free(ctypes.byref(b))
    return out


def bake_dense(xyz, origin, h, nx, ny, nz, z0=0, z1=-1, threads=0, zstep=1):
    """Dense signed distance grid (z, y, x) with the oracle's BVH bake; None if the mesh is open.
    Only every zstep-th slab of [z0, z1) is computed (the others stay NaN)."""
    xyz = _f32(xyz)
    dist = np.full((nz, ny, nx), np.nan, np.float32)
    st = BakeStats()
    rc = lib().orc_bake_dense_strided(_p(xyz, _f32p), xyz.size // 9, _p(_f32(origin), _f32p), float(h), nx, ny, nz,
                                      z0, z1, zstep, threads, _p(dist, _f32p), ctypes.byref(st))
    if rc != 0:
        return None, None
    return dist, {k: getattr(st, k) for k, _ in BakeStats._fields_}


def sdf_f64_points(xyz, pts, threads=0):
    """FP64 brute-force signed distance (sign from |w| > 0.5) and winding number at points."""
    xyz = _f32(xyz)
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    d = np.zeros(len(pts), np.float64)
    w = np.zeros(len(pts), np.float64)
    lib().orc_sdf_f64_points(_p(xyz, _f32p), xyz.size // 9, _p(pts, _f64p), len(pts), threads, _p(d, _f64p),
                             _p(w, _f64p))
    return d, w


def gwn_f64_points(xyz, pts, threads=0):
    """FP64 winding numbers at points (orc_gwn_f64_points): a bounded hierarchical estimate where it decides
    |w| > 0.5 and the 1e-6 tie band as the exact sum does, else the brute-force sum.  Returns (w, exact)."""
    xyz = _f32(xyz)
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    w = np.zeros(len(pts), np.float64)
    exact = np.zeros(len(pts), np.uint8)
    rc = lib().orc_gwn_f64_points(_p(xyz, _f32p), xyz.size // 9, _p(pts, _f64p), len(pts), threads, _p(w, _f64p),
                                  exact.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    if rc != 0:
        raise RuntimeError("orc_gwn_f64_points: the tree could not be built")
    return w, exact.astype(bool)


def boundary_edge_count(xyz):
    xyz = _f32(xyz)
    return int(lib().orc_boundary_edge_count(_p(xyz, _f32p), xyz.size // 9))


def udf_points_f64(xyz, pts, threads=0):
    """FP64 unsigned distance at points through the LBVH (equal to the brute-force FP64 minimum)."""
    xyz = _f32(xyz)
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    out = np.zeros(len(pts), np.float64)
    rc = lib().orc_udf_points_f64(_p(xyz, _f32p), xyz.size // 9, _p(pts, _f64p), len(pts), threads, _p(out, _f64p))
    if rc != 0:
        raise ValueError(f"orc_udf_points_f64 failed ({rc})")
    return out


class Material(ctypes.Structure):
    _fields_ = [("base", ctypes.c_float * 3), ("metallic", ctypes.c_float), ("roughness", ctypes.c_float),
                ("tex", ctypes.c_int32)]


class Texture(ctypes.Structure):
    _fields_ = [("rgba", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("srgb", ctypes.c_int32)]


def triangle_materials(ntri, uvs, tri_material, materials, textures):
    """triangleMaterial (sdf_baker.cpp:42-74) per triangle: (ntri, 5) = r, g, b, metallic, roughness.
    uvs (ncorner, 2); tri_material (m,) uint32; materials: list of (base rgb, metallic, roughness, texture);
    textures: list of (rgba uint8 (h, w, 4), srgb)."""
    uvs = np.ascontiguousarray(uvs, np.float32).reshape(-1, 2)
    tm = np.ascontiguousarray(tri_material, np.uint32)
    mats = (Material * max(len(materials), 1))()
    for i, (base, met, rough, tex) in enumerate(materials):
        mats[i] = Material((ctypes.c_float * 3)(*base), met, rough, tex)
    keep = [np.ascontiguousarray(px, np.uint8) for px, _ in textures]
    texs = (Texture * max(len(textures), 1))()
    for i, (px, srgb) in enumerate(textures):
        texs[i] = Texture(keep[i].ctypes.data, keep[i].shape[1], keep[i].shape[0], int(srgb))
    out = np.zeros((ntri, 5), np.float32)
    lib().orc_triangle_materials(ntri, _p(uvs, _f32p), len(uvs), tm.ctypes.data, len(tm), ctypes.addressof(mats),
                                 len(materials), ctypes.addressof(texs), len(textures), _p(out, _f32p))
    return out


def weld_positions(positions, bmin, bmax):
    """weldPositions (src/io/mesh_weld.cpp:26-48): (welded index per position, first position of each index)."""
    pos = _f32(positions).reshape(-1, 3)
    welded = np.zeros(len(pos), np.uint32)
    unique = np.zeros((max(len(pos), 1), 3), np.float32)
    n = lib().orc_weld_positions(_p(pos, _f32p), len(pos), _p(_f32(bmin), _f32p), _p(_f32(bmax), _f32p),
                                 welded.ctypes.data, unique.ctypes.data)
    if n < 0:
        raise ValueError("position not representable on the weld lattice")
    return welded, unique[:n].copy()


def watertight(xyz, bmin, bmax):
    """isWatertight (src/io/mesh_validate.cpp:21-59) as counts: (watertight, boundary edges, non-manifold edges)."""
    xyz = _f32(xyz)
    b, m = ctypes.c_int64(), ctypes.c_int64()
    r = lib().orc_watertight(_p(xyz, _f32p), xyz.size // 9, _p(_f32(bmin), _f32p), _p(_f32(bmax), _f32p),
                             ctypes.byref(b), ctypes.byref(m))
    if r < 0:
        raise ValueError("position not representable on the weld lattice")
    return bool(r), int(b.value), int(m.value)


def udf2_f32_brute(xyz, pts):
    xyz = _f32(xyz)
    pts = _f32(pts).reshape(-1, 3)
    out = np.zeros(len(pts), np.float32)
    lib().orc_udf2_f32_brute(_p(xyz, _f32p), xyz.size // 9, _p(pts, _f32p), len(pts), _p(out, _f32p))
    return out


def cell_centres(origin, h, nx, ny, nz):
    """(nz, ny, nx, 3) float32 centres origin + (i + 0.5) * h, evaluated in float32 as the bake does."""
    o = np.asarray(origin, np.float32)
    h = np.float32(h)
    zs = o[2] + (np.arange(nz, dtype=np.float32) + np.float32(0.5)) * h
    ys = o[1] + (np.arange(ny, dtype=np.float32) + np.float32(0.5)) * h
    xs = o[0] + (np.arange(nx, dtype=np.float32) + np.float32(0.5)) * h
    zz, yy, xx = np.meshgrid(zs, ys, xs, indexing="ij")
    return np.stack([xx, yy, zz], -1).astype(np.float32)


def udf_points(xyz, pts, threads=0):
    """FP32 unsigned distance at points (LBVH; equal to udf2_f32_brute's square root)."""
    xyz = _f32(xyz)
    pts = _f32(pts).reshape(-1, 3)
    out = np.zeros(len(pts), np.float32)
    lib().orc_udf_points_f32(_p(xyz, _f32p), xyz.size // 9, _p(pts, _f32p), len(pts), threads, _p(out, _f32p))
    return out


def band_bricks(xyz, origin, h, nx, ny, nz, band):
    """Bricks (8^3) a narrow band of +-band voxels keeps: unsigned distance at the brick centre
    (cell centres of the grid of spacing 8h) <= h * (band + 3.5 sqrt 3), FP32 as the bake does it."""
    h8 = np.float32(h) * np.float32(8.0)
    c = cell_centres(origin, h8, nx // 8, ny // 8, nz // 8).reshape(-1, 3)
    d = udf_points(xyz, c)
    thresh = np.float32(h) * (np.float32(band) + np.float32(6.0621778))
    return d <= thresh
