// italy_bake.hpp -- C++ drop-ins for the reference's bake entry points, over the C ABI.
//
//   italy_bake::bakeSdf(mesh, resolution)       replaces italy::bakeSdf      (src/convert/sdf_baker.h:8)
//   italy_bake::voxelizeMesh(mesh, resolution)  replaces italy::voxelizeMesh (src/convert/voxelize.h:8)
//   italy_bake::makeCupSdf / makeGlassCupSdf / makeMaterialProbeSdfs, bakeAnalyticSdf(Palette)
//                                               replace src/convert/sdf_procedural.h:28-40
//   italy_bake::weldPositions / weldedIndexCount replace src/io/mesh_weld.h:10-14
//   italy_bake::isWatertight(mesh, err)         replaces italy::isWatertight (src/io/mesh_validate.h:9)
//
// `Mesh` is any type shaped like italy::MeshAsset (src/io/mesh_asset.h:33-54): a contiguous
// `positions` container of 3-float vectors (triangle soup, 3 per triangle) and `boundsMin` /
// `boundsMax` with .x/.y/.z.  glm::vec3 qualifies (three packed floats), so the reference's own
// MeshAsset can be passed unchanged.  Non-zero statuses are rethrown as std::runtime_error, the
// reference's error convention (src/render/optix_check.h:13-48).
//
// Per-cell materials (sdf_baker.cpp:76-128) and per-cell voxel colours (voxelize.cpp:94-104) are
// averaged on the device in triangle order from a per-triangle material the caller supplies
// (`triangleMaterial` / colour callbacks), because texture lookups stay in the caller's domain.
#pragma once

#include <array>
#include <cstdint>
#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "italy_bake.h"

namespace italy_bake {

inline void check(ib_status s, const char *call) {
  if (s != IB_OK) throw std::runtime_error(std::string(call) + " failed: " + ib_last_error());
}

// one device context per process (the reference creates a context per call: sdf_baker.cpp:150-156)
inline ib_context *context() {
  struct Holder {
    ib_context *c = nullptr;
    Holder() { check(ib_create(0, 0, &c), "ib_create"); }
    ~Holder() { ib_destroy(c); }
  };
  static Holder h;
  return h.c;
}

enum class SdfMaterialKind : uint8_t { Opaque = 0, Dielectric = 1 };

struct SdfPaletteMaterial {  // src/convert/sdf_grid.h:12-19
  SdfMaterialKind kind = SdfMaterialKind::Opaque;
  std::array<float, 3> baseColor{1.0f, 1.0f, 1.0f};
  float metallic = 0.0f;
  float roughness = 1.0f;
  float ior = 1.0f;
  std::array<float, 3> extinction{0.0f, 0.0f, 0.0f};
};

struct SdfSurfaceMaterial {  // src/convert/sdf_procedural.h:17-21
  std::array<float, 3> baseColor{1.0f, 1.0f, 1.0f};
  float metallic = 0.0f;
  float roughness = 1.0f;
};

struct SdfGrid {  // src/convert/sdf_grid.h:21-36
  std::array<float, 3> origin{0.0f, 0.0f, 0.0f};
  float voxelSize = 1.0f;
  int nx = 0, ny = 0, nz = 0;
  std::vector<float> distances;
  std::vector<std::array<float, 3>> baseColor;
  std::vector<float> metallic;
  std::vector<float> roughness;
  std::vector<uint8_t> branch;
  std::vector<SdfPaletteMaterial> palette;
  size_t index(int x, int y, int z) const { return (static_cast<size_t>(z) * ny + y) * nx + x; }
  std::array<float, 3> boundsMax() const {
    return {origin[0] + nx * voxelSize, origin[1] + ny * voxelSize, origin[2] + nz * voxelSize};
  }
};

struct VoxelGrid {  // src/convert/voxel_grid.h:9-17
  std::array<float, 3> origin{0.0f, 0.0f, 0.0f};
  float voxelSize = 1.0f;
  std::vector<std::array<int, 3>> cells;   // first-touch order
  std::vector<std::array<float, 3>> colors;
};

template <class V>
inline void bounds(const V &lo, const V &hi, float bmin[3], float bmax[3]) {
  bmin[0] = lo.x;
  bmin[1] = lo.y;
  bmin[2] = lo.z;
  bmax[0] = hi.x;
  bmax[1] = hi.y;
  bmax[2] = hi.z;
}

template <class Mesh>
inline const float *soup(const Mesh &mesh) {
  static_assert(sizeof(mesh.positions[0]) == 3 * sizeof(float), "positions must be packed 3-float vectors");
  return mesh.positions.empty() ? nullptr : reinterpret_cast<const float *>(mesh.positions.data());
}

namespace detail {
// a MeshAsset-shaped type: materials, textures, per-corner uvs and per-triangle material indices
template <class M, class = void>
struct has_materials : std::false_type {};
template <class M>
struct has_materials<M, std::void_t<decltype(std::declval<const M &>().materials[0].baseColorFactor.x),
                                    decltype(std::declval<const M &>().materials[0].baseColorTexture),
                                    decltype(std::declval<const M &>().textures[0].pixelsRGBA.data()),
                                    decltype(std::declval<const M &>().uvs.data()),
                                    decltype(std::declval<const M &>().triangleMaterial.data())>> : std::true_type {};

// (r, g, b, metallic, roughness) per triangle: triangleMaterial (src/convert/sdf_baker.cpp:61-74) on the device
// for a MeshAsset, the default material for a type without materials
template <class Mesh>
std::vector<float> triangle_materials(const Mesh &mesh) {
  const uint64_t ntri = mesh.positions.size() / 3;
  std::vector<float> out(5 * ntri);
  if constexpr (has_materials<Mesh>::value) {
    std::vector<ib_material> mats(mesh.materials.size());
    for (size_t i = 0; i < mats.size(); ++i) {
      const auto &m = mesh.materials[i];
      mats[i] = ib_material{{m.baseColorFactor.x, m.baseColorFactor.y, m.baseColorFactor.z}, m.metallic, m.roughness,
                            m.baseColorTexture};
    }
    std::vector<ib_texture> texs(mesh.textures.size());
    for (size_t i = 0; i < texs.size(); ++i) {
      const auto &t = mesh.textures[i];
      if (t.width > 0 && t.height > 0 && t.pixelsRGBA.size() < (size_t)t.width * (size_t)t.height * 4)
        throw std::runtime_error("texture " + std::to_string(i) + " holds fewer texels than width * height");
      texs[i] = ib_texture{t.pixelsRGBA.data(), t.width, t.height, t.srgb ? 1 : 0, 0};
    }
    static_assert(sizeof(mesh.uvs[0]) == 2 * sizeof(float), "uvs must be packed 2-float vectors");
    check(ib_triangle_materials(context(), ntri, mesh.uvs.empty() ? nullptr : &mesh.uvs[0].x, mesh.uvs.size(),
                                mesh.triangleMaterial.data(), mesh.triangleMaterial.size(), mats.data(),
                                (uint32_t)mats.size(), texs.data(), (uint32_t)texs.size(), out.data()),
          "ib_triangle_materials");
  } else {
    for (uint64_t t = 0; t < ntri; ++t) {
      out[5 * t] = out[5 * t + 1] = out[5 * t + 2] = 1.0f;
      out[5 * t + 3] = 0.0f;
      out[5 * t + 4] = 1.0f;
    }
  }
  return out;
}
}  // namespace detail

// Per-triangle material: 5 floats (r, g, b, metallic, roughness) per triangle; MaterialFn(t, out5).
template <class Mesh, class MaterialFn>
SdfGrid bakeSdf(const Mesh &mesh, int resolution, MaterialFn &&material) {
  SdfGrid grid;
  float bmin[3], bmax[3];
  bounds(mesh.boundsMin, mesh.boundsMax, bmin, bmax);
  ib_grid g{};
  check(ib_sdf_framing(bmin, bmax, resolution, &g), "ib_sdf_framing");
  grid.origin = {g.origin[0], g.origin[1], g.origin[2]};
  grid.voxelSize = g.h;
  grid.nx = g.nx;
  grid.ny = g.ny;
  grid.nz = g.nz;
  const uint64_t ntri = mesh.positions.size() / 3;
  const uint64_t ncell = static_cast<uint64_t>(g.nx) * g.ny * g.nz;
  grid.distances.resize(ncell);
  check(ib_bake_sdf(context(), soup(mesh), ntri, bmin, bmax, resolution, &g, grid.distances.data(), ncell),
        "ib_bake_sdf");
  std::vector<float> mat(5 * ntri);
  for (uint64_t t = 0; t < ntri; ++t) material(t, &mat[5 * t]);
  grid.baseColor.resize(ncell);
  grid.metallic.resize(ncell);
  grid.roughness.resize(ncell);
  check(ib_splat_materials(context(), soup(mesh), ntri, mat.data(), &g, grid.baseColor[0].data(),
                           grid.metallic.data(), grid.roughness.data()),
        "ib_splat_materials");
  return grid;
}

// Without a material callback: a MeshAsset's own materials and textures (triangleMaterial, computed on the
// device), or the default material (colour 1, metallic 0, roughness 1) for a type without materials.
template <class Mesh>
SdfGrid bakeSdf(const Mesh &mesh, int resolution) {
  const std::vector<float> mat = detail::triangle_materials(mesh);
  return bakeSdf(mesh, resolution, [&](uint64_t t, float *m) {
    for (int k = 0; k < 5; ++k) m[k] = mat[5 * t + k];
  });
}

// ColorFn(t, out3): the triangle's colour (voxelize.cpp:30-38 computes it from the material).
template <class Mesh, class ColorFn>
VoxelGrid voxelizeMesh(const Mesh &mesh, int resolution, ColorFn &&color) {
  VoxelGrid grid;
  float bmin[3], bmax[3];
  bounds(mesh.boundsMin, mesh.boundsMax, bmin, bmax);
  check(ib_vox_framing(bmin, bmax, resolution, grid.origin.data(), &grid.voxelSize), "ib_vox_framing");
  const uint64_t ntri = mesh.positions.size() / 3;
  std::vector<float> rgb(3 * ntri);
  for (uint64_t t = 0; t < ntri; ++t) {
    float m[5] = {1.0f, 1.0f, 1.0f, 0.0f, 0.0f};
    color(t, m);
    for (int k = 0; k < 3; ++k) rgb[3 * t + k] = m[k];
  }
  // one call for the common case: room for a few cells per triangle, a second call with the exact count
  // when the triangles are large against the voxels
  uint64_t count = 0, cap = std::max<uint64_t>(1024, 4 * ntri);
  for (int attempt = 0; attempt < 2; ++attempt) {
    grid.cells.resize(cap);
    grid.colors.resize(cap);
    const ib_status st = ib_voxelize_colors(context(), soup(mesh), ntri, bmin, bmax, resolution, rgb.data(),
                                            grid.cells[0].data(), grid.colors[0].data(), cap, &count);
    if (st == IB_ERR_CAPACITY && attempt == 0) {
      cap = count;
      continue;
    }
    check(st, "ib_voxelize_colors");
    break;
  }
  grid.cells.resize(count);
  grid.colors.resize(count);
  return grid;
}

// Without a colour callback: triangleColor (src/convert/voxelize.cpp:26-38) of a MeshAsset, computed on the
// device, or white for a type without materials.
template <class Mesh>
VoxelGrid voxelizeMesh(const Mesh &mesh, int resolution) {
  const std::vector<float> mat = detail::triangle_materials(mesh);
  return voxelizeMesh(mesh, resolution, [&](uint64_t t, float *c) {
    for (int k = 0; k < 3; ++k) c[k] = mat[5 * t + k];
  });
}

// ---- mesh ingestion gate (src/io/mesh_weld.h, src/io/mesh_validate.h) ------------------------------
// weldPositions(positions, boundsMin, boundsMax, &unique): the reference's signature with any container
// of packed 3-float vectors (std::vector<glm::vec3> qualifies).
template <class Positions, class V, class Unique = Positions>
std::vector<uint32_t> weldPositions(const Positions &positions, const V &boundsMin, const V &boundsMax,
                                    Unique *uniquePositions = nullptr) {
  static_assert(sizeof(positions[0]) == 3 * sizeof(float), "positions must be packed 3-float vectors");
  float bmin[3], bmax[3];
  bounds(boundsMin, boundsMax, bmin, bmax);
  const uint64_t n = positions.size();
  std::vector<uint32_t> welded(n);
  std::vector<float> unique(uniquePositions ? 3 * n : 0);
  uint64_t count = 0;
  check(ib_weld_positions(context(), n ? reinterpret_cast<const float *>(positions.data()) : nullptr, n, bmin, bmax,
                          welded.data(), uniquePositions ? unique.data() : nullptr, &count),
        "ib_weld_positions");
  if (uniquePositions) {
    uniquePositions->clear();
    uniquePositions->resize(count);
    for (uint64_t i = 0; i < count; ++i) {
      auto &u = (*uniquePositions)[i];
      u.x = unique[3 * i];
      u.y = unique[3 * i + 1];
      u.z = unique[3 * i + 2];
    }
  }
  return welded;
}

// src/io/mesh_weld.cpp:50-54
inline size_t weldedIndexCount(const std::vector<uint32_t> &welded) {
  size_t m = 0;
  for (uint32_t w : welded) m = w + 1 > m ? w + 1 : m;
  return welded.empty() ? 0 : m;
}

// isWatertight(mesh, err): false with the reference's message in err, true with err untouched
template <class Mesh>
bool isWatertight(const Mesh &mesh, std::string &err) {
  float bmin[3], bmax[3];
  bounds(mesh.boundsMin, mesh.boundsMax, bmin, bmax);
  int32_t ok = 0;
  char msg[512];
  check(ib_watertight(context(), soup(mesh), mesh.positions.size() / 3, bmin, bmax, &ok, nullptr, nullptr, msg,
                      sizeof(msg)),
        "ib_watertight");
  if (!ok) err = msg;
  return ok != 0;
}

// ---- analytic SDF bake (src/convert/sdf_procedural.cpp) ------------------------------------------
// The reference's SdfDistanceFn is a host std::function; here the distance is a CSG program
// (ib_csg_op, include/italy_bake.h) evaluated on the device.  Csg builds one in postfix order.
struct Csg {
  std::vector<ib_csg_op> ops;
  Csg &push(int op, int branch, std::initializer_list<float> a) {
    ib_csg_op o{};
    o.op = op;
    o.branch = branch;
    int i = 0;
    for (float v : a) o.a[i++] = v;
    ops.push_back(o);
    return *this;
  }
  Csg &sphere(float cx, float cy, float cz, float r, int branch = 0) { return push(IB_CSG_SPHERE, branch, {cx, cy, cz, r}); }
  Csg &box(float cx, float cy, float cz, float bx, float by, float bz, int branch = 0) {
    return push(IB_CSG_BOX, branch, {cx, cy, cz, bx, by, bz});
  }
  Csg &cylinder(float cx, float cy, float cz, float r, float h, int branch = 0) {
    return push(IB_CSG_CYLINDER, branch, {cx, cy, cz, r, h});
  }
  Csg &torus(float cx, float cy, float cz, float major, float minor, int branch = 0) {
    return push(IB_CSG_TORUS, branch, {cx, cy, cz, major, minor});
  }
  Csg &plane(float nx, float ny, float nz, float d, int branch = 0) {
    return push(IB_CSG_PLANE, branch, {0.0f, 0.0f, 0.0f, nx, ny, nz, d});
  }
  Csg &opUnion() { return push(IB_CSG_UNION, 0, {}); }
  Csg &opSubtract() { return push(IB_CSG_SUBTRACT, 0, {}); }
  Csg &opIntersect() { return push(IB_CSG_INTERSECT, 0, {}); }
  Csg &opSmoothUnion(float k) { return push(IB_CSG_SMOOTH_UNION, 0, {k}); }
  Csg &opSmoothIntersect(float k) { return push(IB_CSG_SMOOTH_INTERSECT, 0, {k}); }
  Csg &select() { return push(IB_CSG_SELECT, 0, {}); }
};

inline SdfGrid analyticGrid(const float bmin[3], const float bmax[3], int resolution, const Csg &distance,
                            bool withBranch) {
  ib_grid g{};
  check(ib_analytic_framing(bmin, bmax, resolution, &g), "ib_analytic_framing");
  SdfGrid grid;
  grid.origin = {g.origin[0], g.origin[1], g.origin[2]};
  grid.voxelSize = g.h;
  grid.nx = g.nx;
  grid.ny = g.ny;
  grid.nz = g.nz;
  const size_t n = static_cast<size_t>(g.nx) * g.ny * g.nz;
  grid.distances.resize(n);
  grid.branch.resize(n);
  check(ib_bake_analytic(context(), distance.ops.data(), static_cast<uint32_t>(distance.ops.size()), &g,
                         grid.distances.data(), grid.branch.data()),
        "ib_bake_analytic");
  if (!withBranch) grid.branch.clear();
  return grid;
}

// MaterialFn(branch) -> SdfSurfaceMaterial: per branch (sdf_procedural.cpp:46-64 passes the cell
// position too; none of the reference's scenes reads it).
template <class V, class MaterialFn>
SdfGrid bakeAnalyticSdf(const V &boundsMin, const V &boundsMax, int resolution, const Csg &distance,
                        MaterialFn &&material) {
  float bmin[3], bmax[3];
  bounds(boundsMin, boundsMax, bmin, bmax);
  SdfGrid grid = analyticGrid(bmin, bmax, resolution, distance, true);
  const size_t n = grid.distances.size();
  grid.baseColor.resize(n);
  grid.metallic.resize(n);
  grid.roughness.resize(n);
  SdfSurfaceMaterial table[256];
  bool have[256] = {};
  for (size_t i = 0; i < n; ++i) {
    const uint8_t b = grid.branch[i];
    if (!have[b]) {
      table[b] = material(static_cast<int>(b));
      have[b] = true;
    }
    grid.baseColor[i] = table[b].baseColor;
    grid.metallic[i] = table[b].metallic;
    grid.roughness[i] = table[b].roughness;
  }
  grid.branch.clear();
  return grid;
}

template <class V, class PaletteFn>
SdfGrid bakeAnalyticSdfPalette(const V &boundsMin, const V &boundsMax, int resolution, const Csg &distance,
                               PaletteFn &&palette, int maxBranches) {  // sdf_procedural.cpp:66-79
  float bmin[3], bmax[3];
  bounds(boundsMin, boundsMax, bmin, bmax);
  SdfGrid grid = analyticGrid(bmin, bmax, resolution, distance, true);
  grid.palette.resize(static_cast<size_t>(maxBranches));
  for (int b = 0; b < maxBranches; ++b) grid.palette[b] = palette(b);
  return grid;
}

namespace detail {
struct V3 {
  float x, y, z;
};
struct CupShape {  // the constants of sdf_procedural.cpp:82-93 / :109-121, in the same float order
  float R = 0.55f, wall = 0.09f, ri = R - wall, lift = 0.16f, rimFraction = 0.62f;
  V3 oc, ic, bmin, bmax;
  float rimY, fillY;
  CupShape(float originY, float cx, float cz) {
    oc = {cx, originY + R, cz};
    ic = {oc.x + 0.0f, oc.y + lift, oc.z + 0.0f};
    rimY = oc.y + R * rimFraction;
    fillY = ic.y - ri + wall * 0.5f + 0.42f * (2.0f * ri);
    bmin = {oc.x - R, oc.y - R, oc.z - R};
    bmax = {oc.x + R, rimY, oc.z + R};
  }
  Csg shell() const {  // opIntersect(opSubtract(outer, inner), p.y - rimY)
    return Csg().sphere(oc.x, oc.y, oc.z, R).sphere(ic.x, ic.y, ic.z, ri).opSubtract().plane(0, 1, 0, rimY).opIntersect();
  }
};
}  // namespace detail

inline SdfGrid makeCupSdf(int resolution, float originY, float centerX = 0.0f, float centerZ = 0.0f) {
  const detail::CupShape c(originY, centerX, centerZ);  // sdf_procedural.cpp:81-106
  return bakeAnalyticSdf(c.bmin, c.bmax, resolution, c.shell(),
                         [](int) { return SdfSurfaceMaterial{{0.85f, 0.83f, 0.78f}, 0.0f, 0.35f}; });
}
template <class V2>
SdfGrid makeCupSdf(int resolution, float originY, const V2 &centerXZ) {
  return makeCupSdf(resolution, originY, centerXZ.x, centerXZ.y);
}

inline SdfGrid makeGlassCupSdf(int resolution, float originY, float centerX = 0.0f, float centerZ = 0.0f) {
  const detail::CupShape c(originY, centerX, centerZ);  // sdf_procedural.cpp:108-140
  Csg d = c.shell();
  d.sphere(c.ic.x, c.ic.y, c.ic.z, c.ri, 1).plane(0, 1, 0, c.fillY, 1).opSmoothIntersect(0.012f).select();
  return bakeAnalyticSdfPalette(c.bmin, c.bmax, resolution, d, [](int branch) {
    if (branch == 0)
      return SdfPaletteMaterial{SdfMaterialKind::Dielectric, {1.0f, 1.0f, 1.0f}, 0.0f, 0.05f, 1.47f, {0.02f, 0.02f, 0.018f}};
    return SdfPaletteMaterial{SdfMaterialKind::Dielectric, {1.0f, 1.0f, 1.0f}, 0.0f, 0.02f, 1.333f, {0.9f, 0.35f, 0.12f}};
  }, 2);
}
template <class V2>
SdfGrid makeGlassCupSdf(int resolution, float originY, const V2 &centerXZ) {
  return makeGlassCupSdf(resolution, originY, centerXZ.x, centerXZ.y);
}

inline std::vector<SdfGrid> makeMaterialProbeSdfs(int resolution) {  // sdf_procedural.cpp:142-193
  const float radius = 0.4f;
  auto sphereGrid = [&](detail::V3 c, const SdfPaletteMaterial &mat) {
    const detail::V3 lo{c.x - radius, c.y - radius, c.z - radius}, hi{c.x + radius, c.y + radius, c.z + radius};
    return bakeAnalyticSdfPalette(lo, hi, resolution, Csg().sphere(c.x, c.y, c.z, radius),
                                  [&](int) { return mat; }, 1);
  };
  std::vector<SdfGrid> grids;
  grids.reserve(9);
  const float row1X[5] = {-2.2f, -1.1f, 0.0f, 1.1f, 2.2f};
  const float row1Roughness[5] = {0.05f, 0.3f, 0.55f, 0.8f, 1.0f};
  for (int i = 0; i < 5; ++i) {
    SdfPaletteMaterial m;
    m.baseColor = {0.7f, 0.7f, 0.72f};
    m.roughness = row1Roughness[i];
    grids.push_back(sphereGrid({row1X[i], radius, -0.6f}, m));
  }
  const float row2X[4] = {-1.65f, -0.55f, 0.55f, 1.65f};
  const float row2Ior[4] = {1.1f, 1.35f, 1.6f, 2.0f};
  for (int i = 0; i < 4; ++i) {
    SdfPaletteMaterial m;
    m.kind = SdfMaterialKind::Dielectric;
    m.roughness = 0.02f;
    m.ior = row2Ior[i];
    m.extinction = {0.05f, 0.05f, 0.05f};
    grids.push_back(sphereGrid({row2X[i], radius, 0.6f}, m));
  }
  return grids;
}

}  // namespace italy_bake
