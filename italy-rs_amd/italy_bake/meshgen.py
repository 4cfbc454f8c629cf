"""Seeded synthetic meshes (triangle soups) for the bake workloads, via libitaly_meshgen.so."""
import ctypes
import os

import numpy as np

from ._paths import PKG_ROOT

_f32p = ctypes.POINTER(ctypes.c_float)


class _Soup(ctypes.Structure):
    _fields_ = [("xyz", _f32p), ("ntri", ctypes.c_int64), ("bmin", ctypes.c_float * 3),
                ("bmax", ctypes.c_float * 3)]


class Mesh:
    """Triangle soup: positions (ntri, 3, 3) float32 and caller bounds, as MeshAsset
    (src/io/mesh_asset.h:33-54) holds them."""

    def __init__(self, xyz, bmin, bmax):
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3, 3)
        self.bmin = np.asarray(bmin, np.float32).copy()
        self.bmax = np.asarray(bmax, np.float32).copy()

    @property
    def ntri(self):
        return self.xyz.shape[0]

    def __repr__(self):
        return f"Mesh(ntri={self.ntri}, bmin={self.bmin.tolist()}, bmax={self.bmax.tolist()})"


_lib = None


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(PKG_ROOT, "lib", "libitaly_meshgen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        S = ctypes.POINTER(_Soup)
        lib.ib_gen_uv_sphere.argtypes = [ctypes.c_float, ctypes.c_int, ctypes.c_int, S]
        lib.ib_gen_cube_shell.argtypes = [S]
        lib.ib_gen_icosphere.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, S]
        lib.ib_gen_displaced.argtypes = [ctypes.c_int, ctypes.c_uint64, S]
        lib.ib_gen_decimated.argtypes = [ctypes.c_int64, ctypes.c_uint64, S]
        lib.ib_gen_stress.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_double, S]
        lib.ib_gen_stress_res.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, S]
        lib.ib_soup_free.argtypes = [S]
        lib.ib_gen_decimated_batch.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                               ctypes.POINTER(ctypes.c_uint64), ctypes.c_int, S]
        _lib = lib
    return _lib


def _take(fn, *args):
    s = _Soup()
    rc = fn(*args, ctypes.byref(s))
    if rc != 0:
        raise ValueError(f"mesh generator failed ({rc})")
    n = int(s.ntri)
    xyz = np.ctypeslib.as_array(s.xyz, (max(n, 1) * 9,)).copy()[: n * 9]
    m = Mesh(xyz, list(s.bmin), list(s.bmax))
    _load().ib_soup_free(ctypes.byref(s))
    return m


def uv_sphere(r=1.0, lon=32, lat=16):
    """The UV sphere of tests/sdf_sign_test.cpp:18-42 (bounds set to +-r as the test does)."""
    return _take(_load().ib_gen_uv_sphere, float(r), int(lon), int(lat))


def cube_shell():
    """The unit cube shell of tests/voxelize_test.cpp:11-33."""
    return _take(_load().ib_gen_cube_shell)


def icosphere(level, radius=0.5, jitter=0.0, seed=0x17A1E5):
    return _take(_load().ib_gen_icosphere, int(level), float(radius), float(jitter), int(seed))


def displaced(level, seed=0x17A1E5):
    return _take(_load().ib_gen_displaced, int(level), int(seed))


def decimated(target, seed):
    return _take(_load().ib_gen_decimated, int(target), int(seed))


def stress(seed=0x17A1E5, ncomp=64, remove_frac=0.02, torus_segments=(64, 32), box_quads=32):
    return _take(_load().ib_gen_stress_res, int(seed), int(ncomp), float(remove_frac), int(torus_segments[0]),
                 int(torus_segments[1]), int(box_quads))


# ---- the workloads -----------------------------------------------------------------------
BASE_SEED = 0x17A1E5


def workload_plain_small():
    """5,120-triangle icosphere (level 4, r 0.5) with radial jitter U(+-1e-3) of the welded
    vertices; baked to 64^3 over [-0.6, 0.6]^3 (SDF and occupancy)."""
    m = icosphere(4, 0.5, 1e-3, BASE_SEED)
    return m, dict(lo=-0.6, hi=0.6, sdf_res=64, occ_res=64)


def workload_plain_large():
    """327,680-triangle displaced icosphere (level 7); 256^3 SDF and 512^3 occupancy over
    [-0.8, 0.8]^3."""
    m = displaced(7, BASE_SEED)
    return m, dict(lo=-0.8, hi=0.8, sdf_res=256, occ_res=512)


def decimated_batch(targets, seeds, threads=0):
    """Decimated displaced icospheres generated in parallel on the host."""
    n = len(targets)
    t = (ctypes.c_int64 * n)(*[int(x) for x in targets])
    sd = (ctypes.c_uint64 * n)(*[int(x) for x in seeds])
    outs = (_Soup * n)()
    lib = _load()
    if lib.ib_gen_decimated_batch(n, t, sd, int(threads), outs) != 0:
        raise ValueError("mesh generator failed")
    meshes = []
    for i in range(n):
        k = int(outs[i].ntri)
        xyz = np.ctypeslib.as_array(outs[i].xyz, (max(k, 1) * 9,)).copy()[: k * 9]
        meshes.append(Mesh(xyz, list(outs[i].bmin), list(outs[i].bmax)))
        lib.ib_soup_free(ctypes.byref(outs[i]))
    return meshes


def workload_batched(n=512, seed=BASE_SEED + 4096, threads=0):
    """n displaced icospheres, triangle counts log-uniform in [500, 200k] (subdivision level, then
    random edge collapses down to the count), each with a dense SDF + occupancy grid over
    [-0.8, 0.8]^3 at a resolution drawn uniformly from {32, 64, 128, 256}.
    Returns (meshes, resolutions)."""
    rng = np.random.default_rng(seed)
    targets = np.exp(rng.uniform(np.log(500.0), np.log(200000.0), n)).astype(np.int64)
    res = rng.choice([32, 64, 128, 256], n)
    seeds = [seed + 1 + i for i in range(n)]
    return decimated_batch(targets, seeds, threads), [int(r) for r in res]


def workload_stress(seed=BASE_SEED):
    """64 seeded closed components (32 tori of 64x32 segments, 32 boxes of 32x32 quads per face)
    with random rotations and positions in [-0.7, 0.7]^3, overlaps allowed, 2% of triangles removed
    at random; baked as a 384^3 narrow band of +-4 voxels in 8^3 bricks over [-1, 1]^3."""
    m = stress(seed, 64, 0.02)
    return m, dict(lo=-1.0, hi=1.0, res=384, band=4)


def workload_stress_1m(seed=BASE_SEED):
    """The precision-stress union at the configuration's stated size (~1M triangles): 32 tori of 64x64
    segments and 32 boxes of 46x46 quads per face (1,074,688 triangles, 1,053,195 after 2% are removed),
    otherwise as workload_stress."""
    m = stress(seed, 64, 0.02, torus_segments=(64, 64), box_quads=46)
    return m, dict(lo=-1.0, hi=1.0, res=384, band=4)
