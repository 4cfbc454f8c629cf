"""Host-side mirror of the reference bake entry points over the HIP C ABI.

    bakeSdf(mesh, resolution)      -> SdfGrid      src/convert/sdf_baker.h:8
    voxelizeMesh(mesh, resolution) -> VoxelGrid    src/convert/voxelize.h:8

plus the explicit-grid bake the workloads use (Context.bake_dense / bake_dense_device).
Every call runs the gfx950 kernels; a missing library raises (there is no CPU fallback).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (IB_ERR_CAPACITY, IB_ERR_INTERNAL, IB_ERR_INVALID_ARGUMENT, IB_ERR_OUT_OF_BUDGET, IB_OK,  # noqa: F401
                   NOALLOC_MAX_COLUMNS, NOALLOC_MAX_TRIANGLES, NOALLOC_MAX_VOXELS, BakeError, Grid, Job, Stats, check)

OPT_COUNT_WORK = 1
OPT_QUERY_VARIANT = 2
OPT_BRICK_SHAPE = 3
OPT_BLOCKS_PER_CU = 4
OPT_TILE_ORDER = 5
OPT_LEAF_BLOCK = 6
OPT_RESOLVE_FULL = 7
OPT_SHARED_CHIP = 8
# the library's defaults (ib_create), per option
DEFAULT_OPTIONS = {OPT_COUNT_WORK: 0, OPT_QUERY_VARIANT: 1, OPT_BRICK_SHAPE: 0, OPT_BLOCKS_PER_CU: 8,
                   OPT_TILE_ORDER: 2, OPT_LEAF_BLOCK: 8, OPT_RESOLVE_FULL: 0, OPT_SHARED_CHIP: 0}


def _f3(v):
    a = (ctypes.c_float * 3)()
    for k in range(3):
        a[k] = float(v[k])
    return a


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class Context:
    """One device context: stream, growable device scratch, HBM budget (ib_create)."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        lib = _lib.load()
        h = ctypes.c_void_p()
        check(lib.ib_create(int(device), int(hbm_budget_bytes), ctypes.byref(h)))
        self._h = h
        self.lib = lib
        self._options = dict(DEFAULT_OPTIONS)  # the values set on this context

    def close(self):
        if self._h:
            self.lib.ib_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    # ---- explicit grids ---------------------------------------------------------------------
    def bake_dense(self, xyz, sdf_grid=None, occ_grid=None):
        """Bake one soup (ntri,3,3) float32: returns (distances (nz,ny,nx) or None, occ words or None)."""
        xyz = np.ascontiguousarray(xyz, np.float32)
        ntri = xyz.size // 9
        dist = np.empty(sdf_grid.shape(), np.float32) if sdf_grid is not None else None
        occ = None
        if occ_grid is not None:
            occ = np.empty(((occ_grid.nx + 31) // 32) * occ_grid.ny * occ_grid.nz, np.uint32)
        check(self.lib.ib_bake_dense(self._h, _ptr(xyz) if ntri else None, ntri,
                                     ctypes.byref(sdf_grid) if sdf_grid is not None else None, _ptr(dist),
                                     ctypes.byref(occ_grid) if occ_grid is not None else None, _ptr(occ)))
        return dist, occ

    def bake_dense_device(self, d_xyz, ntri, sdf_grid, d_dist, occ_grid=None, d_occ=None, stream=None):
        check(self.lib.ib_bake_dense_device(self._h, ctypes.c_void_p(d_xyz), int(ntri),
                                            ctypes.byref(sdf_grid) if sdf_grid is not None else None,
                                            ctypes.c_void_p(d_dist) if d_dist else None,
                                            ctypes.byref(occ_grid) if occ_grid is not None else None,
                                            ctypes.c_void_p(d_occ) if d_occ else None,
                                            ctypes.c_void_p(stream) if stream else None))

    def bake_dense_device_noalloc(self, d_xyz, ntri, sdf_grid, d_dist, occ_grid=None, d_occ=None, stream=None):
        """ib_bake_dense_device_noalloc: bake_dense_device without any device allocation (fixed scratch)."""
        check(self.lib.ib_bake_dense_device_noalloc(self._h, ctypes.c_void_p(d_xyz), int(ntri),
                                                    ctypes.byref(sdf_grid) if sdf_grid is not None else None,
                                                    ctypes.c_void_p(d_dist) if d_dist else None,
                                                    ctypes.byref(occ_grid) if occ_grid is not None else None,
                                                    ctypes.c_void_p(d_occ) if d_occ else None,
                                                    ctypes.c_void_p(stream) if stream else None))

    def bake_batch_device(self, jobs, stream=None):
        """One batched bake (ib_bake_batch_device) of a list of Job structures; synchronous.
        Narrow-band jobs get their nbricks written back into the list, jobs naming a resident mesh its
        soup pointer and triangle count."""
        arr = (Job * len(jobs))(*jobs)
        st = self.lib.ib_bake_batch_device(self._h, arr, len(jobs), ctypes.c_void_p(stream) if stream else None)
        for i in range(len(jobs)):  # written back by the library: band sizes, a resident mesh's soup
            jobs[i].nbricks = arr[i].nbricks
            jobs[i].d_xyz = arr[i].d_xyz
            jobs[i].ntri = arr[i].ntri
        check(st)

    def bake_band(self, xyz, grid, band_voxels):
        """Narrow-band bake of one host soup: returns (brick index (nz/8, ny/8, nx/8) int32, pool
        (nbricks, 8, 8, 8) float32)."""
        xyz = np.ascontiguousarray(xyz, np.float32)
        ntri = xyz.size // 9
        nbrick = (grid.nx // 8) * (grid.ny // 8) * (grid.nz // 8)
        d_xyz = self.alloc(max(xyz.nbytes, 4))
        d_idx = self.alloc(4 * max(nbrick, 1))
        try:
            if ntri:
                self.upload(d_xyz, xyz)
            n = ctypes.c_uint64()
            cap = max(nbrick, 1)  # worst case: every brick
            d_pool = self.alloc(4 * 512 * cap)
            try:
                check(self.lib.ib_bake_band_device(self._h, ctypes.c_void_p(d_xyz), ntri, ctypes.byref(grid),
                                                   int(band_voxels), ctypes.c_void_p(d_idx),
                                                   ctypes.c_void_p(d_pool), cap, ctypes.byref(n), None))
                pool = self.download(d_pool, np.empty((n.value, 8, 8, 8), np.float32))
            finally:
                self.free(d_pool)
            index = self.download(d_idx, np.empty((grid.nz // 8, grid.ny // 8, grid.nx // 8), np.int32))
        finally:
            self.free(d_xyz)
            self.free(d_idx)
        return index, pool

    def autotune_batch(self, jobs, sample_stride=16):
        """ib_autotune_batch: returns (brick shape, blocks per CU, query ms of the sample)."""
        arr = (Job * len(jobs))(*jobs)
        shape, blocks, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        check(self.lib.ib_autotune_batch(self._h, arr, len(jobs), int(sample_stride), ctypes.byref(shape),
                                         ctypes.byref(blocks), ctypes.byref(ms)))
        return shape.value, blocks.value, ms.value

    # ---- resident meshes -----------------------------------------------------------------------
    def mesh_create(self, d_xyz, ntri, stream=None):
        """ib_mesh_create: build a device soup's LBVH once and keep it resident; returns the handle
        (a Job may name it in place of its soup).  Free it with mesh_destroy before the context closes."""
        h = ctypes.c_void_p()
        check(self.lib.ib_mesh_create(self._h, ctypes.c_void_p(d_xyz), int(ntri),
                                      ctypes.c_void_p(stream) if stream else None, ctypes.byref(h)))
        return h.value

    def mesh_destroy(self, mesh):
        self.lib.ib_mesh_destroy(ctypes.c_void_p(mesh))

    def mesh_bytes(self, mesh):
        return int(self.lib.ib_mesh_bytes(ctypes.c_void_p(mesh)))

    # ---- mesh ingestion gate ------------------------------------------------------------------
    def weld_positions(self, positions, bmin, bmax):
        """ib_weld_positions (weldPositions, src/io/mesh_weld.cpp:26-48): (welded index per position
        (uint32), first position of each index (nunique, 3) float32)."""
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        welded = np.zeros(max(len(pos), 1), np.uint32)
        unique = np.zeros((max(len(pos), 1), 3), np.float32)
        n = ctypes.c_uint64()
        check(self.lib.ib_weld_positions(self._h, _ptr(pos) if len(pos) else None, len(pos), _f3(bmin), _f3(bmax),
                                         _ptr(welded), _ptr(unique), ctypes.byref(n)))
        return welded[:len(pos)], unique[:n.value].copy()

    def watertight(self, xyz, bmin, bmax):
        """ib_watertight (isWatertight, src/io/mesh_validate.cpp:21-59): (watertight, boundary edges,
        non-manifold edges, err text)."""
        xyz = np.ascontiguousarray(xyz, np.float32)
        ntri = xyz.size // 9
        ok, b, m = ctypes.c_int32(), ctypes.c_uint64(), ctypes.c_uint64()
        msg = ctypes.create_string_buffer(512)
        check(self.lib.ib_watertight(self._h, _ptr(xyz) if ntri else None, ntri, _f3(bmin), _f3(bmax),
                                     ctypes.byref(ok), ctypes.byref(b), ctypes.byref(m), msg, len(msg)))
        return bool(ok.value), b.value, m.value, msg.value.decode("utf-8")

    def watertight_device(self, d_xyz, ntri, bmin, bmax, stream=None):
        """ib_watertight_device on a device soup: (watertight, boundary edges, non-manifold edges)."""
        ok, b, m = ctypes.c_int32(), ctypes.c_uint64(), ctypes.c_uint64()
        check(self.lib.ib_watertight_device(self._h, ctypes.c_void_p(d_xyz), int(ntri), _f3(bmin), _f3(bmax),
                                            ctypes.byref(ok), ctypes.byref(b), ctypes.byref(m), None, 0,
                                            ctypes.c_void_p(stream) if stream else None))
        return bool(ok.value), b.value, m.value

    def triangle_materials(self, ntri, uvs, tri_material, materials, textures):
        """ib_triangle_materials (triangleMaterial, sdf_baker.cpp:42-74): (ntri, 5) r, g, b, metallic,
        roughness.  uvs (ncorner, 2); tri_material (m,) uint32; materials: list of (base rgb, metallic,
        roughness, texture index); textures: list of (rgba uint8 (h, w, 4), srgb)."""
        from ._lib import Material, Texture

        uvs = np.ascontiguousarray(uvs, np.float32).reshape(-1, 2)
        tm = np.ascontiguousarray(tri_material, np.uint32)
        mats = (Material * max(len(materials), 1))()
        for i, (base, met, rough, tex) in enumerate(materials):
            mats[i] = Material((ctypes.c_float * 3)(*base), met, rough, tex)
        keep = [np.ascontiguousarray(px, np.uint8) for px, _ in textures]
        texs = (Texture * max(len(textures), 1))()
        for i, (px, srgb) in enumerate(textures):
            texs[i] = Texture(keep[i].ctypes.data if keep[i].size else None, keep[i].shape[1], keep[i].shape[0],
                              int(srgb), 0)
        out = np.zeros((max(ntri, 1), 5), np.float32)
        check(self.lib.ib_triangle_materials(self._h, int(ntri), _ptr(uvs) if len(uvs) else None, len(uvs),
                                             _ptr(tm) if len(tm) else None, len(tm), ctypes.addressof(mats),
                                             len(materials), ctypes.addressof(texs), len(textures), _ptr(out)))
        return out[:ntri]

    def splat_materials(self, xyz, materials, grid):
        """ib_splat_materials (bakePerCellMaterial, sdf_baker.cpp:76-128): per-triangle materials
        (ntri, 5) = r, g, b, metallic, roughness, averaged per cell in triangle order.  Returns
        (baseColor (ncell, 3), metallic (ncell,), roughness (ncell,))."""
        xyz = np.ascontiguousarray(xyz, np.float32)
        ntri = xyz.size // 9
        mat = np.ascontiguousarray(materials, np.float32).reshape(ntri, 5)
        n = grid.nx * grid.ny * grid.nz
        rgb = np.empty((n, 3), np.float32)
        met = np.empty(n, np.float32)
        rough = np.empty(n, np.float32)
        check(self.lib.ib_splat_materials(self._h, _ptr(xyz) if ntri else None, ntri, _ptr(mat) if ntri else None,
                                          ctypes.byref(grid), _ptr(rgb), _ptr(met), _ptr(rough)))
        return rgb, met, rough

    def sample_grid_device(self, grid, d_dist, d_points, n, d_values=None, d_normals=None, d_nearest=None):
        """ib_sample_grid_device: the renderer's trilinear sample, B-spline normal and nearest cell of a
        device-resident grid at n device points (enqueued; call synchronize() before reading)."""
        vp = lambda p: ctypes.c_void_p(p) if p else None
        check(self.lib.ib_sample_grid_device(self._h, ctypes.byref(grid), vp(d_dist), vp(d_points), int(n),
                                             vp(d_values), vp(d_normals), vp(d_nearest), None))

    def trace_grid_device(self, grid, d_dist, d_org, d_dir, n, d_t, tmin=0.0, tmax=1e16):
        """ib_trace_grid_device: the SDF intersection program's march for n device rays."""
        check(self.lib.ib_trace_grid_device(self._h, ctypes.byref(grid), ctypes.c_void_p(d_dist),
                                            ctypes.c_void_p(d_org), ctypes.c_void_p(d_dir), int(n), float(tmin),
                                            float(tmax), ctypes.c_void_p(d_t), None))

    def sample_grid(self, grid, dist, points):
        """Host convenience: upload a grid (nz, ny, nx) and points (n, 3), sample, download
        (values, normals, nearest)."""
        dist = np.ascontiguousarray(dist, np.float32)
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        n = len(pts)
        bufs = [self.alloc(max(dist.nbytes, 4)), self.alloc(max(pts.nbytes, 4)), self.alloc(4 * max(n, 1)),
                self.alloc(12 * max(n, 1)), self.alloc(4 * max(n, 1))]
        try:
            self.upload(bufs[0], dist)
            if n:
                self.upload(bufs[1], pts)
            self.sample_grid_device(grid, bufs[0], bufs[1], n, bufs[2], bufs[3], bufs[4])
            self.synchronize()
            vals = self.download(bufs[2], np.empty(n, np.float32))
            nrm = self.download(bufs[3], np.empty((n, 3), np.float32))
            near = self.download(bufs[4], np.empty(n, np.int32))
        finally:
            for b in bufs:
                self.free(b)
        return vals, nrm, near

    def trace_grid(self, grid, dist, org, dirs, tmin=0.0, tmax=1e16):
        """Host convenience for trace_grid_device: hit t per ray (-1 = miss)."""
        dist = np.ascontiguousarray(dist, np.float32)
        org = np.ascontiguousarray(org, np.float32).reshape(-1, 3)
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = len(org)
        bufs = [self.alloc(max(dist.nbytes, 4)), self.alloc(12 * max(n, 1)), self.alloc(12 * max(n, 1)),
                self.alloc(4 * max(n, 1))]
        try:
            self.upload(bufs[0], dist)
            if n:
                self.upload(bufs[1], org)
                self.upload(bufs[2], dirs)
            self.trace_grid_device(grid, bufs[0], bufs[1], bufs[2], n, bufs[3], tmin, tmax)
            self.synchronize()
            t = self.download(bufs[3], np.empty(n, np.float32))
        finally:
            for b in bufs:
                self.free(b)
        return t

    def set_option(self, option, value):
        check(self.lib.ib_set_option(self._h, int(option), int(value)))
        self._options[int(option)] = int(value)

    def option(self, option):
        """the value last set for `option` on this context (the library default if never set)"""
        return self._options[int(option)]

    def query_counters(self, d_xyz, ntri, sdf_grid):
        """Run one counting bake and return the traversal counters of the query kernel."""
        d_out = self.alloc(4 * sdf_grid.nx * sdf_grid.ny * sdf_grid.nz)
        saved = {k: self.option(k) for k in (OPT_COUNT_WORK, OPT_QUERY_VARIANT, OPT_LEAF_BLOCK)}
        try:
            # the algorithm's own work: one traversal per voxel over the binary LBVH with single-
            # triangle leaves (no wave union, no leaf blocks)
            self.set_option(OPT_COUNT_WORK, 1)
            self.set_option(OPT_QUERY_VARIANT, 0)
            self.set_option(OPT_LEAF_BLOCK, 1)
            self.bake_dense_device(d_xyz, ntri, sdf_grid, d_out)
            st = self.stats()
        finally:
            for k, v in saved.items():
                self.set_option(k, v)
            self.free(d_out)
        return {"box_tests": st["box_tests"], "tri_tests": st["tri_tests"]}

    def stats(self):
        st = Stats()
        check(self.lib.ib_last_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    # ---- device memory ----------------------------------------------------------------------
    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        check(self.lib.ib_device_alloc(self._h, int(nbytes), ctypes.byref(p)))
        return p.value

    def free(self, p):
        check(self.lib.ib_device_free(self._h, ctypes.c_void_p(p)))

    def upload(self, d_ptr, arr):
        arr = np.ascontiguousarray(arr)
        check(self.lib.ib_memcpy_h2d(self._h, ctypes.c_void_p(d_ptr), _ptr(arr), arr.nbytes))

    def download(self, d_ptr, arr):
        check(self.lib.ib_memcpy_d2h(self._h, _ptr(arr), ctypes.c_void_p(d_ptr), arr.nbytes))
        return arr

    def copy_device(self, d_dst, d_src, nbytes):
        check(self.lib.ib_memcpy_d2d(self._h, ctypes.c_void_p(d_dst), ctypes.c_void_p(d_src), int(nbytes)))

    def synchronize(self):
        check(self.lib.ib_synchronize(self._h))


# ---- the reference's data types and free functions --------------------------------------------
class SdfGrid:
    """src/convert/sdf_grid.h:21-36: origin, voxelSize, nx, ny, nz, distances (x fastest)."""

    def __init__(self, grid, distances):
        self.origin = np.array(list(grid.origin), np.float32)
        self.voxelSize = np.float32(grid.h)
        self.nx, self.ny, self.nz = grid.nx, grid.ny, grid.nz
        self.distances = distances.reshape(-1)
        self.baseColor = np.empty((0, 3), np.float32)
        self.metallic = np.empty(0, np.float32)
        self.roughness = np.empty(0, np.float32)
        self.branch = np.empty(0, np.uint8)
        self.palette = []

    def index(self, x, y, z):
        return (z * self.ny + y) * self.nx + x

    def boundsMax(self):
        return self.origin + np.array([self.nx, self.ny, self.nz], np.float32) * self.voxelSize


class VoxelGrid:
    """src/convert/voxel_grid.h:9-17: origin, voxelSize, cells in first-touch order."""

    def __init__(self, origin, voxel_size, cells, first_tri):
        self.origin = np.asarray(origin, np.float32)
        self.voxelSize = np.float32(voxel_size)
        self.cells = cells
        self.first_triangle = first_tri

    def cellMin(self, c):
        return self.origin + np.asarray(c, np.float32) * self.voxelSize

    def cellMax(self, c):
        return self.cellMin(c) + self.voxelSize


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context()
    return _default_ctx


def sdf_framing(bmin, bmax, resolution):
    g = Grid()
    check(_lib.load().ib_sdf_framing(_f3(bmin), _f3(bmax), int(resolution), ctypes.byref(g)))
    return g


def bakeSdf(mesh, resolution, ctx=None, materials=None):
    """italy::bakeSdf (src/convert/sdf_baker.h:8): reference framing, exact signed distances, and the
    per-cell materials of bakePerCellMaterial (sdf_baker.cpp:76-128).  materials: (ntri, 5) per-triangle
    r, g, b, metallic, roughness (triangleMaterial, :61-74); None gives every triangle the default
    material (1, 1, 1, 0, 1) of MaterialAsset (mesh_asset.h:18-21)."""
    ctx = ctx or default_context()
    g = sdf_framing(mesh.bmin, mesh.bmax, resolution)
    dist = np.empty(g.nx * g.ny * g.nz, np.float32)
    xyz = np.ascontiguousarray(mesh.xyz, np.float32)
    check(ctx.lib.ib_bake_sdf(ctx.handle, _ptr(xyz) if mesh.ntri else None, mesh.ntri, _f3(mesh.bmin),
                              _f3(mesh.bmax), int(resolution), ctypes.byref(g), _ptr(dist), dist.size))
    grid = SdfGrid(g, dist)
    if materials is None:
        materials = np.tile(np.array([1, 1, 1, 0, 1], np.float32), (mesh.ntri, 1))
    grid.baseColor, grid.metallic, grid.roughness = ctx.splat_materials(xyz, materials, g)
    return grid


def voxelizeMesh(mesh, resolution, ctx=None, tri_colors=None):
    """italy::voxelizeMesh (src/convert/voxelize.h:8): occupied cells in first-touch order, each cell's first
    triangle, and VoxelGrid.colors (the mean of the touching triangles' colours; tri_colors: (ntri, 3) float32,
    None: white).  Room for a few cells per triangle first, the exact count on a second call if needed."""
    ctx = ctx or default_context()
    lib = ctx.lib
    xyz = np.ascontiguousarray(mesh.xyz, np.float32)
    origin = (ctypes.c_float * 3)()
    vs = ctypes.c_float()
    check(lib.ib_vox_framing(_f3(mesh.bmin), _f3(mesh.bmax), int(resolution), origin, ctypes.byref(vs)))
    rgb = None if tri_colors is None else np.ascontiguousarray(tri_colors, np.float32).reshape(-1, 3)
    if rgb is not None and len(rgb) != mesh.ntri:
        raise ValueError(f"tri_colors holds {len(rgb)} colours for {mesh.ntri} triangles")
    count = ctypes.c_uint64()
    cap = max(1024, 4 * mesh.ntri)
    for attempt in range(2):
        cells = np.zeros((cap, 3), np.int32)
        tri = np.zeros(cap, np.int32)
        st = lib.ib_voxelize(ctx.handle, _ptr(xyz) if mesh.ntri else None, mesh.ntri, _f3(mesh.bmin),
                             _f3(mesh.bmax), int(resolution), _ptr(cells), _ptr(tri), cap, ctypes.byref(count))
        if st == IB_ERR_CAPACITY and attempt == 0:
            cap = count.value
            continue
        check(st)
        break
    n = count.value
    # white triangles average to exactly 1.0 (n / n): only given colours need the second pass
    colors = np.ones((max(n, 1), 3), np.float32)
    if n and rgb is not None:
        check(lib.ib_voxelize_colors(ctx.handle, _ptr(xyz), mesh.ntri, _f3(mesh.bmin), _f3(mesh.bmax),
                                     int(resolution), _ptr(rgb), None, _ptr(colors), n,
                                     ctypes.byref(count)))
    g = VoxelGrid(list(origin), vs.value, cells[:n], tri[:n])
    g.colors = colors[:n]
    return g


def weldPositions(positions, boundsMin, boundsMax, ctx=None):
    """italy::weldPositions (src/io/mesh_weld.h:10-12): (welded indices, unique positions)."""
    return (ctx or default_context()).weld_positions(positions, boundsMin, boundsMax)


def weldedIndexCount(welded):
    """italy::weldedIndexCount (src/io/mesh_weld.cpp:50-54): largest index + 1, 0 when empty."""
    welded = np.asarray(welded)
    return int(welded.max()) + 1 if welded.size else 0


def isWatertight(mesh, ctx=None):
    """italy::isWatertight (src/io/mesh_validate.h:9): (bool, err).  err is the reference's text
    ("" when the mesh is watertight)."""
    ok, _, _, err = (ctx or default_context()).watertight(mesh.xyz, mesh.bmin, mesh.bmax)
    return ok, err


def dense_grid(lo, hi, res):
    """Cell-centred res^3 grid over [lo, hi]^3."""
    h = np.float32((np.float32(hi) - np.float32(lo)) / np.float32(res))
    return Grid.make((lo, lo, lo), h, res, res, res)


__all__ = ["Context", "SdfGrid", "VoxelGrid", "bakeSdf", "voxelizeMesh", "sdf_framing", "dense_grid", "BakeError",
           "Grid", "Job", "weldPositions", "weldedIndexCount", "isWatertight"]
