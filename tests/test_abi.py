"""C-ABI checks that run without a GPU: the library loads, exports every declared symbol, the
pure-host framing entry points agree with the oracle, and device calls fail cleanly."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ib_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from italy_bake import _lib

    lib = _lib.load()
    names = _declared("italy_bake.h")
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTS) == names


def test_sdf_framing_matches_oracle(oracle):
    from italy_bake import bake

    rng = np.random.default_rng(0)
    for _ in range(200):
        bmin = rng.uniform(-2, 1, 3).astype(np.float32)
        bmax = (bmin + rng.uniform(0, 3, 3)).astype(np.float32)
        res = int(rng.integers(-2, 300))
        g = bake.sdf_framing(bmin, bmax, res)
        o, h, n = oracle.sdf_framing(bmin, bmax, res)
        assert np.float32(g.h) == h
        assert (g.nx, g.ny, g.nz) == n
        assert np.array_equal(np.array(list(g.origin), np.float32), o)


def test_vox_framing_matches_oracle(oracle):
    from italy_bake import _lib

    lib = _lib.load()
    rng = np.random.default_rng(1)
    for _ in range(100):
        bmin = rng.uniform(-2, 1, 3).astype(np.float32)
        bmax = (bmin + rng.uniform(0, 3, 3)).astype(np.float32)
        res = int(rng.integers(1, 300))
        o = (ctypes.c_float * 3)()
        h = ctypes.c_float()
        f3 = lambda v: (ctypes.c_float * 3)(*[float(x) for x in v])
        assert lib.ib_vox_framing(f3(bmin), f3(bmax), res, o, ctypes.byref(h)) == 0
        oo, hh = oracle.vox_framing(bmin, bmax, res)
        assert np.float32(h.value) == hh and np.array_equal(np.array(list(o), np.float32), oo)


def test_create_without_device_reports_status():
    from italy_bake import _lib, bake

    if os.environ.get("HIP_VISIBLE_DEVICES", None) == "" or not os.path.exists("/dev/kfd"):
        with pytest.raises(_lib.BakeError) as e:
            bake.Context()
        assert e.value.status == 5  # IB_ERR_NO_DEVICE
    else:
        pytest.skip("a GPU is visible")


def test_null_arguments_are_rejected():
    from italy_bake import _lib

    lib = _lib.load()
    assert lib.ib_sdf_framing(None, None, 4, None) == 1
    assert lib.ib_bake_dense(None, None, 0, None, None, None, None) == 1
    assert b"null" in lib.ib_last_error()


def test_new_entry_points_reject_null_arguments():
    """The ingestion gate and resident-mesh entry points validate their arguments before any device work."""
    import ctypes

    from italy_bake import _lib

    lib = _lib.load()
    ok = ctypes.c_int32()
    n = ctypes.c_uint64()
    h = ctypes.c_void_p()
    b3 = (ctypes.c_float * 3)(0, 0, 0)
    assert lib.ib_watertight(None, None, 0, b3, b3, ctypes.byref(ok), None, None, None, 0) == 1
    assert lib.ib_weld_positions(None, None, 0, b3, b3, None, None, ctypes.byref(n)) == 1
    assert lib.ib_mesh_create(None, None, 0, None, ctypes.byref(h)) == 1
    assert lib.ib_mesh_bytes(None) == 0 and lib.ib_mesh_triangles(None) == 0
    lib.ib_mesh_destroy(None)  # no-op
    out = (ctypes.c_float * 5)()
    assert lib.ib_triangle_materials(None, 1, None, 0, None, 0, None, 0, None, 0, out) == 1
