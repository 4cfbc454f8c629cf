"""ctypes binding of lib/libitaly_bake.so (C ABI declared in include/italy_bake.h)."""
import ctypes
import os

from ._paths import LIB_DIR

LIB_PATH = os.path.join(LIB_DIR, "libitaly_bake.so")

IB_OK = 0
IB_ERR_INVALID_ARGUMENT = 1
IB_ERR_OUT_OF_BUDGET = 3
IB_ERR_CAPACITY = 7
IB_ERR_INTERNAL = 9
# limits of ib_bake_dense_device_noalloc (italy_bake.h)
NOALLOC_MAX_TRIANGLES = 65536
NOALLOC_MAX_COLUMNS = 16384
NOALLOC_MAX_VOXELS = 2097152
STATUS_NAMES = {
    0: "IB_OK", 1: "IB_ERR_INVALID_ARGUMENT", 2: "IB_ERR_INVALID_MESH", 3: "IB_ERR_OUT_OF_BUDGET",
    4: "IB_ERR_DEVICE", 5: "IB_ERR_NO_DEVICE", 6: "IB_ERR_UNSUPPORTED", 7: "IB_ERR_CAPACITY",
    8: "IB_ERR_NOT_READY", 9: "IB_ERR_INTERNAL",
}

# every symbol include/italy_bake.h declares
EXPORTS = [
    "ib_create", "ib_destroy", "ib_last_error", "ib_version", "ib_sdf_framing", "ib_vox_framing", "ib_bake_sdf",
    "ib_voxelize", "ib_splat_materials", "ib_bake_dense", "ib_bake_dense_device", "ib_device_alloc",
    "ib_device_free", "ib_memcpy_h2d", "ib_memcpy_d2h", "ib_memcpy_d2d", "ib_synchronize", "ib_sample_grid_device",
    "ib_trace_grid_device", "ib_last_stats", "ib_set_option",
    "ib_bake_batch_device", "ib_autotune_batch", "ib_bake_band_device", "ib_server_create", "ib_server_destroy",
    "ib_server_last_error", "ib_server_submit", "ib_server_info", "ib_server_wait", "ib_server_drain",
    "ib_server_fetch", "ib_server_device_result", "ib_server_release", "ib_server_stats",
    "ib_analytic_framing", "ib_bake_analytic_device", "ib_bake_analytic",
    "ib_weld_positions", "ib_weld_positions_device", "ib_watertight", "ib_watertight_device",
    "ib_mesh_create", "ib_mesh_destroy", "ib_mesh_bytes", "ib_mesh_triangles", "ib_triangle_materials",
    "ib_bake_dense_device_noalloc", "ib_voxelize_colors",
]


class BakeError(RuntimeError):
    """A non-zero ib_status, raised like the reference's CUDA_CHECK/OPTIX_CHECK runtime_error."""

    def __init__(self, status, message):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")


class Grid(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_float * 3), ("h", ctypes.c_float), ("nx", ctypes.c_int32),
                ("ny", ctypes.c_int32), ("nz", ctypes.c_int32), ("reserved", ctypes.c_int32)]

    @classmethod
    def make(cls, origin, h, nx, ny, nz):
        g = cls()
        for k in range(3):
            g.origin[k] = float(origin[k])
        g.h = float(h)
        g.nx, g.ny, g.nz = int(nx), int(ny), int(nz)
        return g

    def shape(self):
        return (self.nz, self.ny, self.nx)


class Stats(ctypes.Structure):
    _fields_ = [
        ("ms_total", ctypes.c_double), ("ms_ingest", ctypes.c_double), ("ms_bvh", ctypes.c_double),
        ("ms_occupancy", ctypes.c_double), ("ms_columns", ctypes.c_double), ("ms_sdf", ctypes.c_double),
        ("triangles", ctypes.c_uint64), ("voxels", ctypes.c_uint64), ("occ_cells", ctypes.c_uint64),
        ("bvh_depth", ctypes.c_int32), ("pad0", ctypes.c_int32), ("boundary_edges", ctypes.c_int64),
        ("column_overflows", ctypes.c_uint64), ("rechecked", ctypes.c_uint64), ("ties", ctypes.c_uint64),
        ("box_tests", ctypes.c_uint64), ("tri_tests", ctypes.c_uint64), ("stack_overflows", ctypes.c_uint64),
        ("ms_caps", ctypes.c_double), ("ms_query", ctypes.c_double), ("obb_tests", ctypes.c_uint64),
        ("on_cap", ctypes.c_uint64), ("rechecked_full", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k not in ("reserved", "pad0")}


class Material(ctypes.Structure):
    """ib_material: baseColorFactor, metallic, roughness, baseColorTexture (-1: none)."""
    _fields_ = [("base_color", ctypes.c_float * 3), ("metallic", ctypes.c_float), ("roughness", ctypes.c_float),
                ("base_color_texture", ctypes.c_int32)]


class Texture(ctypes.Structure):
    """ib_texture: RGBA8 texels, width, height, sRGB flag."""
    _fields_ = [("rgba", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("srgb", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class Job(ctypes.Structure):
    """ib_job: one mesh of a batched bake (device pointers)."""
    _fields_ = [("d_xyz", ctypes.c_void_p), ("ntri", ctypes.c_uint64), ("sdf", Grid), ("d_distances", ctypes.c_void_p),
                ("occ", Grid), ("d_occ_words", ctypes.c_void_p), ("band_voxels", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("d_brick_index", ctypes.c_void_p), ("pool_bricks", ctypes.c_uint64),
                ("nbricks", ctypes.c_uint64), ("mesh", ctypes.c_void_p)]


_lib = None
_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_i32 = ctypes.c_int32
_f3 = ctypes.POINTER(ctypes.c_float)


def load():
    """Load the HIP library; raises if it is not built (there is no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() (there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    G = ctypes.POINTER(Grid)
    lib.ib_create.argtypes = [ctypes.c_int, _u64, ctypes.POINTER(_vp)]
    lib.ib_destroy.argtypes = [_vp]
    lib.ib_destroy.restype = None
    lib.ib_last_error.restype = ctypes.c_char_p
    lib.ib_version.restype = ctypes.c_char_p
    lib.ib_sdf_framing.argtypes = [_f3, _f3, _i32, G]
    lib.ib_vox_framing.argtypes = [_f3, _f3, _i32, _f3, _f3]
    lib.ib_bake_sdf.argtypes = [_vp, _vp, _u64, _f3, _f3, _i32, G, _vp, _u64]
    lib.ib_voxelize.argtypes = [_vp, _vp, _u64, _f3, _f3, _i32, _vp, _vp, _u64, ctypes.POINTER(_u64)]
    lib.ib_voxelize_colors.argtypes = [_vp, _vp, _u64, _f3, _f3, _i32, _vp, _vp, _vp, _u64, ctypes.POINTER(_u64)]
    lib.ib_splat_materials.argtypes = [_vp, _vp, _u64, _vp, G, _vp, _vp, _vp]
    lib.ib_bake_dense.argtypes = [_vp, _vp, _u64, G, _vp, G, _vp]
    lib.ib_bake_dense_device.argtypes = [_vp, _vp, _u64, G, _vp, G, _vp, _vp]
    lib.ib_bake_dense_device_noalloc.argtypes = [_vp, _vp, _u64, G, _vp, G, _vp, _vp]
    lib.ib_device_alloc.argtypes = [_vp, _u64, ctypes.POINTER(_vp)]
    lib.ib_device_free.argtypes = [_vp, _vp]
    lib.ib_memcpy_h2d.argtypes = [_vp, _vp, _vp, _u64]
    lib.ib_memcpy_d2h.argtypes = [_vp, _vp, _vp, _u64]
    lib.ib_memcpy_d2d.argtypes = [_vp, _vp, _vp, _u64]
    lib.ib_sample_grid_device.argtypes = [_vp, G, _vp, _vp, _u64, _vp, _vp, _vp, _vp]
    lib.ib_trace_grid_device.argtypes = [_vp, G, _vp, _vp, _vp, _u64, ctypes.c_float, ctypes.c_float, _vp, _vp]
    lib.ib_analytic_framing.argtypes = [_f3, _f3, _i32, G]
    lib.ib_bake_analytic_device.argtypes = [_vp, _vp, ctypes.c_uint32, G, _vp, _vp, _vp]
    lib.ib_bake_analytic.argtypes = [_vp, _vp, ctypes.c_uint32, G, _vp, _vp]
    lib.ib_synchronize.argtypes = [_vp]
    lib.ib_last_stats.argtypes = [_vp, ctypes.POINTER(Stats)]
    lib.ib_set_option.argtypes = [_vp, ctypes.c_int32, ctypes.c_int64]
    lib.ib_bake_batch_device.argtypes = [_vp, ctypes.POINTER(Job), ctypes.c_uint32, _vp]
    lib.ib_bake_band_device.argtypes = [_vp, _vp, _u64, G, ctypes.c_int32, _vp, _vp, _u64, ctypes.POINTER(_u64), _vp]
    lib.ib_autotune_batch.argtypes = [_vp, ctypes.POINTER(Job), ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_double)]
    lib.ib_mesh_create.argtypes = [_vp, _vp, _u64, _vp, ctypes.POINTER(_vp)]
    lib.ib_mesh_destroy.argtypes = [_vp]
    lib.ib_mesh_destroy.restype = None
    lib.ib_mesh_bytes.argtypes = [_vp]
    lib.ib_mesh_bytes.restype = _u64
    lib.ib_mesh_triangles.argtypes = [_vp]
    lib.ib_mesh_triangles.restype = _u64
    lib.ib_triangle_materials.argtypes = [_vp, _u64, _vp, _u64, _vp, _u64, _vp, ctypes.c_uint32, _vp, ctypes.c_uint32,
                                          _vp]
    lib.ib_weld_positions.argtypes = [_vp, _vp, _u64, _f3, _f3, _vp, _vp, ctypes.POINTER(_u64)]
    lib.ib_weld_positions_device.argtypes = [_vp, _vp, _u64, _f3, _f3, _vp, _vp, ctypes.POINTER(_u64), _vp]
    lib.ib_watertight.argtypes = [_vp, _vp, _u64, _f3, _f3, ctypes.POINTER(_i32), ctypes.POINTER(_u64),
                                  ctypes.POINTER(_u64), ctypes.c_char_p, _u64]
    lib.ib_watertight_device.argtypes = [_vp, _vp, _u64, _f3, _f3, ctypes.POINTER(_i32), ctypes.POINTER(_u64),
                                         ctypes.POINTER(_u64), ctypes.c_char_p, _u64, _vp]
    _lib = lib
    return lib


def check(status):
    if status != IB_OK:
        raise BakeError(status, load().ib_last_error().decode(errors="replace"))
