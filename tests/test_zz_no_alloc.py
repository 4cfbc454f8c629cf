"""The scratch-free dense bake (ib_bake_dense_device_noalloc) equals ib_bake_dense_device bit for bit.

Both entry points are called through the C ABI.  Two scratch-free calls run back to back, the second on
a smaller input than the first, so the fixed region holds the first call's contents when the second
starts; then the largest input the documented limits admit, and the refusals beyond them.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from italy_bake import _lib

    return _lib.load()


def _bake(ctx, lib, fn, d_xyz, ntri, g, og):
    """Bakes with entry point `fn`; returns (status, distances, occupancy words) on the host."""
    nw = ((og.nx + 31) // 32) * og.ny * og.nz
    d_dist = ctx.alloc(4 * g.nx * g.ny * g.nz)
    d_occ = ctx.alloc(4 * nw)
    try:
        # poison the outputs so that a voxel the bake does not write cannot match
        ctx.upload(d_dist, np.full(g.nx * g.ny * g.nz, np.nan, np.float32))
        ctx.upload(d_occ, np.full(nw, 0xDEADBEEF, np.uint32))
        st = fn(ctx.handle, ctypes.c_void_p(d_xyz), ntri, ctypes.byref(g), ctypes.c_void_p(d_dist),
                ctypes.byref(og), ctypes.c_void_p(d_occ), None)
        ctx.synchronize()
        dist = ctx.download(d_dist, np.empty(g.shape(), np.float32))
        occ = ctx.download(d_occ, np.empty(nw, np.uint32))
        return st, dist, occ
    finally:
        ctx.free(d_dist)
        ctx.free(d_occ)


def _compare(ctx, lib, xyz, g, og):
    from italy_bake import bake

    ntri = xyz.size // 9
    d_xyz = ctx.alloc(max(xyz.nbytes, 4))
    try:
        if ntri:
            ctx.upload(d_xyz, np.ascontiguousarray(xyz, np.float32))
        st0, ref, rocc = _bake(ctx, lib, lib.ib_bake_dense_device, d_xyz, ntri, g, og)
        assert st0 == bake.IB_OK, lib.ib_last_error()
        ref_stats = ctx.stats()
        st1, got, gocc = _bake(ctx, lib, lib.ib_bake_dense_device_noalloc, d_xyz, ntri, g, og)
        assert st1 == bake.IB_OK, lib.ib_last_error()
        st = ctx.stats()
    finally:
        ctx.free(d_xyz)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(gocc, rocc)
    for k in ("boundary_edges", "column_overflows", "rechecked", "ties", "stack_overflows", "bvh_depth"):
        assert st[k] == ref_stats[k], k
    return st


def test_noalloc_equals_allocating_twice(ctx, meshgen):
    from italy_bake import bake

    lib = _lib()
    big = meshgen.displaced(5, 21)                        # 20,480 triangles, closed
    print(f"noalloc: {big.ntri} triangles at 128^3", flush=True)
    _compare(ctx, lib, big.xyz, bake.dense_grid(-0.8, 0.8, 128), bake.dense_grid(-0.8, 0.8, 160))
    small = meshgen.stress(seed=5, ncomp=2, remove_frac=0.05)  # smaller and open: hole caps
    assert small.ntri < big.ntri
    print(f"noalloc: {small.ntri} triangles (open) at 40^3", flush=True)
    st = _compare(ctx, lib, small.xyz, bake.dense_grid(-0.8, 0.8, 40), bake.dense_grid(-0.8, 0.8, 48))
    assert st["boundary_edges"] > 0


def test_noalloc_at_its_limits(ctx):
    """The largest input the limits admit: 65,536 disjoint triangles (every edge open, so the hole caps
    are as large as they can be) on a 128^3 grid (16,384 columns, 2^21 voxels).  The triangles are about
    1e-4 across, so that their caps bound the winding number tightly and few voxels need the FP64
    recheck."""
    from italy_bake import bake

    lib = _lib()
    rng = np.random.default_rng(0x5C8A7C)
    n = int(bake.NOALLOC_MAX_TRIANGLES)
    centres = rng.uniform(-0.6, 0.6, (n, 1, 3))
    xyz = (centres + rng.uniform(-1e-4, 1e-4, (n, 3, 3))).astype(np.float32)
    print(f"noalloc: {n} disjoint triangles at 128^3", flush=True)
    st = _compare(ctx, lib, xyz, bake.dense_grid(-0.8, 0.8, 128), bake.dense_grid(-0.8, 0.8, 64))
    assert st["boundary_edges"] == 3 * n


def test_noalloc_refuses_beyond_its_limits(ctx):
    from italy_bake import Grid, bake

    lib = _lib()
    d = ctx.alloc(4 * 9 * 65537)
    try:
        g = bake.dense_grid(-0.8, 0.8, 16)
        st = lib.ib_bake_dense_device_noalloc(ctx.handle, ctypes.c_void_p(d), 65537, ctypes.byref(g),
                                              ctypes.c_void_p(d), None, None, None)
        assert st == bake.IB_ERR_CAPACITY and b"at most 65536 triangles" in lib.ib_last_error()
        wide = Grid.make((0, 0, 0), 0.01, 129, 128, 4)     # 16,512 columns
        st = lib.ib_bake_dense_device_noalloc(ctx.handle, ctypes.c_void_p(d), 1, ctypes.byref(wide),
                                              ctypes.c_void_p(d), None, None, None)
        assert st == bake.IB_ERR_CAPACITY
        deep = Grid.make((0, 0, 0), 0.01, 128, 128, 129)   # 2,113,536 voxels
        st = lib.ib_bake_dense_device_noalloc(ctx.handle, ctypes.c_void_p(d), 1, ctypes.byref(deep),
                                              ctypes.c_void_p(d), None, None, None)
        assert st == bake.IB_ERR_CAPACITY
    finally:
        ctx.free(d)
