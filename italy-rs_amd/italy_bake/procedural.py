"""Analytic SDF bake: src/convert/sdf_procedural.{h,cpp} over the C-ABI library.

The reference takes a std::function per cell (SdfDistanceFn, sdf_procedural.h:25); the device
evaluates a postfix CSG program instead (ib_csg_op in include/italy_bake.h), built here with
`Csg`.  Scene constants are computed in float32 one operation at a time, as the reference's host
code does, so the programs reproduce makeCupSdf / makeGlassCupSdf / makeMaterialProbeSdfs
(sdf_procedural.cpp:81-195) cell for cell.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import Grid, check
from .bake import SdfGrid, _f3, default_context

CSG_DTYPE = np.dtype([("op", "<i4"), ("branch", "<i4"), ("a", "<f4", (8,))])
SPHERE, BOX, CYLINDER, TORUS, PLANE = 1, 2, 3, 4, 5
UNION, SUBTRACT, INTERSECT, SMOOTH_UNION, SMOOTH_INTERSECT, SELECT = 10, 11, 12, 13, 14, 15

f32 = np.float32


class Csg:
    """Postfix program builder: primitives push (distance, branch), operators pop two and push one.

    Primitive and operator semantics are those of src/convert/sdf_primitives.h:10-36; `select` is the
    two-branch pick `a <= b ? a : b` of makeGlassCupSdf (sdf_procedural.cpp:129)."""

    def __init__(self):
        self.ops = []

    def _push(self, op, branch=0, *a):
        rec = np.zeros((), CSG_DTYPE)
        rec["op"] = op
        rec["branch"] = branch
        vals = np.zeros(8, np.float32)
        vals[:len(a)] = np.asarray(a, np.float32)
        rec["a"] = vals
        self.ops.append(rec)
        return self

    def sphere(self, center, radius, branch=0):
        return self._push(SPHERE, branch, *center, radius)

    def box(self, center, half_extent, branch=0):
        return self._push(BOX, branch, *center, *half_extent)

    def cylinder(self, center, radius, half_height, branch=0):
        return self._push(CYLINDER, branch, *center, radius, half_height)

    def torus(self, center, major, minor, branch=0):
        return self._push(TORUS, branch, *center, major, minor)

    def plane(self, normal, offset, branch=0):
        """dot(normal, p) - offset; the axis plane (0, 1, 0), h gives p.y - h exactly."""
        return self._push(PLANE, branch, 0.0, 0.0, 0.0, *normal, offset)

    def union(self):
        return self._push(UNION)

    def subtract(self):
        return self._push(SUBTRACT)

    def intersect(self):
        return self._push(INTERSECT)

    def smooth_union(self, k):
        return self._push(SMOOTH_UNION, 0, k)

    def smooth_intersect(self, k):
        return self._push(SMOOTH_INTERSECT, 0, k)

    def select(self):
        return self._push(SELECT)

    def array(self):
        return np.array(self.ops, CSG_DTYPE) if self.ops else np.zeros(0, CSG_DTYPE)


def _program(prog):
    arr = prog.array() if isinstance(prog, Csg) else np.ascontiguousarray(prog, CSG_DTYPE)
    return arr, arr.ctypes.data_as(ctypes.c_void_p)


def analytic_framing(bmin, bmax, resolution):
    """bakeSdfGridShape (sdf_procedural.cpp:12-26) through ib_analytic_framing."""
    g = Grid()
    check(_lib.load().ib_analytic_framing(_f3(bmin), _f3(bmax), int(resolution), ctypes.byref(g)))
    return g


def bake_analytic_device(ctx, prog, grid, d_dist, d_branch=None, stream=None):
    """ib_bake_analytic_device: distances (and branch bytes) of every cell of `grid` into device memory."""
    arr, ptr = _program(prog)
    check(ctx.lib.ib_bake_analytic_device(ctx.handle, ptr, len(arr), ctypes.byref(grid), ctypes.c_void_p(d_dist),
                                          ctypes.c_void_p(d_branch) if d_branch else None,
                                          ctypes.c_void_p(stream) if stream else None))


def bake_analytic(ctx, prog, grid, with_branch=True):
    """ib_bake_analytic: host arrays (distances (n,), branch (n,) uint8 or None), x fastest."""
    arr, ptr = _program(prog)
    n = grid.nx * grid.ny * grid.nz
    dist = np.empty(n, np.float32)
    branch = np.empty(n, np.uint8) if with_branch else None
    check(ctx.lib.ib_bake_analytic(ctx.handle, ptr, len(arr), ctypes.byref(grid),
                                   dist.ctypes.data_as(ctypes.c_void_p),
                                   branch.ctypes.data_as(ctypes.c_void_p) if with_branch else None))
    return dist, branch


@dataclass
class SdfSurfaceMaterial:
    """sdf_procedural.h:17-21"""
    baseColor: tuple = (1.0, 1.0, 1.0)
    metallic: float = 0.0
    roughness: float = 1.0


@dataclass
class SdfPaletteMaterial:
    """sdf_grid.h:12-19 (kind 0 = Opaque, 1 = Dielectric)"""
    kind: int = 0
    baseColor: tuple = (1.0, 1.0, 1.0)
    metallic: float = 0.0
    roughness: float = 1.0
    ior: float = 1.0
    extinction: tuple = field(default=(0.0, 0.0, 0.0))


OPAQUE, DIELECTRIC = 0, 1


def bakeAnalyticSdf(boundsMin, boundsMax, resolution, distance, material, ctx=None, max_branches=256):
    """italy::bakeAnalyticSdf (sdf_procedural.cpp:46-64).  `distance` is a CSG program; `material`
    maps a branch id to an SdfSurfaceMaterial (a per-branch function: the cells' positions are not
    passed, and no scene of the reference's uses them)."""
    ctx = ctx or default_context()
    g = analytic_framing(boundsMin, boundsMax, resolution)
    dist, branch = bake_analytic(ctx, distance, g)
    grid = SdfGrid(g, dist)
    used = np.unique(branch)
    table = np.zeros((max_branches, 5), np.float32)
    for b in used:
        m = material(int(b))
        table[b] = [*m.baseColor, m.metallic, m.roughness]
    per_cell = table[branch]
    grid.baseColor = np.ascontiguousarray(per_cell[:, :3])
    grid.metallic = np.ascontiguousarray(per_cell[:, 3])
    grid.roughness = np.ascontiguousarray(per_cell[:, 4])
    return grid


def bakeAnalyticSdfPalette(boundsMin, boundsMax, resolution, distance, palette, maxBranches, ctx=None):
    """italy::bakeAnalyticSdfPalette (sdf_procedural.cpp:66-79): distances, a branch byte per cell and
    the palette palette(0 .. maxBranches-1)."""
    ctx = ctx or default_context()
    g = analytic_framing(boundsMin, boundsMax, resolution)
    dist, branch = bake_analytic(ctx, distance, g)
    grid = SdfGrid(g, dist)
    grid.branch = branch
    grid.palette = [palette(b) for b in range(maxBranches)]
    return grid


# ---- the reference's procedural scenes (sdf_procedural.cpp:81-195) -----------------------------------
def _cup_constants(originY, centerXZ):
    R, wall, lift, rim_frac = f32(0.55), f32(0.09), f32(0.16), f32(0.62)
    ri = f32(R - wall)
    oc = (f32(centerXZ[0]), f32(f32(originY) + R), f32(centerXZ[1]))
    ic = (f32(oc[0] + f32(0)), f32(oc[1] + lift), f32(oc[2] + f32(0)))
    rim_y = f32(oc[1] + f32(R * rim_frac))
    bmin = (f32(oc[0] - R), f32(oc[1] - R), f32(oc[2] - R))
    bmax = (f32(oc[0] + R), rim_y, f32(oc[2] + R))
    return R, wall, ri, oc, ic, rim_y, bmin, bmax


def cup_program(originY=-1.0, centerXZ=(0.0, 0.0), glass=False):
    """The distance lambdas of makeCupSdf (:95-100) and makeGlassCupSdf (:123-130) as CSG programs,
    with the bounds they are baked over."""
    R, wall, ri, oc, ic, rim_y, bmin, bmax = _cup_constants(originY, centerXZ)
    p = Csg().sphere(oc, R).sphere(ic, ri).subtract().plane((0, 1, 0), rim_y).intersect()
    if glass:
        # fillY = innerCenter.y - innerRadius + wallThickness * 0.5 + 0.42 * (2 * innerRadius) (:117)
        fill_y = f32(f32(f32(ic[1] - ri) + f32(wall * f32(0.5))) + f32(f32(0.42) * f32(f32(2) * ri)))
        p.sphere(ic, ri, branch=1).plane((0, 1, 0), fill_y, branch=1).smooth_intersect(0.012).select()
    return p, bmin, bmax


def makeCupSdf(resolution, originY, centerXZ=(0.0, 0.0), ctx=None):
    """italy::makeCupSdf (sdf_procedural.cpp:81-106)."""
    prog, bmin, bmax = cup_program(originY, centerXZ)
    return bakeAnalyticSdf(bmin, bmax, resolution, prog,
                           lambda b: SdfSurfaceMaterial((0.85, 0.83, 0.78), 0.0, 0.35), ctx=ctx)


def makeGlassCupSdf(resolution, originY, centerXZ=(0.0, 0.0), ctx=None):
    """italy::makeGlassCupSdf (sdf_procedural.cpp:108-140)."""
    prog, bmin, bmax = cup_program(originY, centerXZ, glass=True)

    def palette(b):
        if b == 0:
            return SdfPaletteMaterial(DIELECTRIC, (1.0, 1.0, 1.0), 0.0, 0.05, 1.47, (0.02, 0.02, 0.018))
        return SdfPaletteMaterial(DIELECTRIC, (1.0, 1.0, 1.0), 0.0, 0.02, 1.333, (0.9, 0.35, 0.12))

    return bakeAnalyticSdfPalette(bmin, bmax, resolution, prog, palette, 2, ctx=ctx)


def makeMaterialProbeSdfs(resolution, ctx=None):
    """italy::makeMaterialProbeSdfs (sdf_procedural.cpp:142-193): five opaque spheres of rising
    roughness and four dielectric spheres of rising IOR, radius 0.4, one grid each."""
    radius = f32(0.4)
    grids = []

    def sphere_grid(center, mat):
        c = tuple(f32(v) for v in center)
        bmin = tuple(f32(v - radius) for v in c)
        bmax = tuple(f32(v + radius) for v in c)
        return bakeAnalyticSdfPalette(bmin, bmax, resolution, Csg().sphere(c, radius), lambda b: mat, 1, ctx=ctx)

    for x, rough in zip((-2.2, -1.1, 0.0, 1.1, 2.2), (0.05, 0.3, 0.55, 0.8, 1.0)):
        grids.append(sphere_grid((x, radius, -0.6), SdfPaletteMaterial(OPAQUE, (0.7, 0.7, 0.72), 0.0, rough)))
    for x, ior in zip((-1.65, -0.55, 0.55, 1.65), (1.1, 1.35, 1.6, 2.0)):
        grids.append(sphere_grid((x, radius, 0.6),
                                 SdfPaletteMaterial(DIELECTRIC, (1.0, 1.0, 1.0), 0.0, 0.02, ior, (0.05, 0.05, 0.05))))
    return grids


__all__ = ["Csg", "CSG_DTYPE", "analytic_framing", "bake_analytic", "bake_analytic_device", "bakeAnalyticSdf",
           "bakeAnalyticSdfPalette", "makeCupSdf", "makeGlassCupSdf", "makeMaterialProbeSdfs", "cup_program",
           "SdfSurfaceMaterial", "SdfPaletteMaterial"]
