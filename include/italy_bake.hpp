// italy_bake.hpp -- C++ drop-ins for the reference's bake entry points, over the C ABI.
//
//   italy_bake::bakeSdf(mesh, resolution)       replaces italy::bakeSdf      (src/convert/sdf_baker.h:8)
//   italy_bake::voxelizeMesh(mesh, resolution)  replaces italy::voxelizeMesh (src/convert/voxelize.h:8)
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
#include <memory>
#include <stdexcept>
#include <string>
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

struct SdfGrid {  // src/convert/sdf_grid.h:21-36
  std::array<float, 3> origin{0.0f, 0.0f, 0.0f};
  float voxelSize = 1.0f;
  int nx = 0, ny = 0, nz = 0;
  std::vector<float> distances;
  std::vector<std::array<float, 3>> baseColor;
  std::vector<float> metallic;
  std::vector<float> roughness;
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

// Without a material callback every triangle carries the default material (colour 1, metallic 0,
// roughness 1), as a MeshAsset with one default MaterialAsset does.
template <class Mesh>
SdfGrid bakeSdf(const Mesh &mesh, int resolution) {
  return bakeSdf(mesh, resolution, [](uint64_t, float *m) {
    m[0] = m[1] = m[2] = 1.0f;
    m[3] = 0.0f;
    m[4] = 1.0f;
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
  uint64_t count = 0;
  check(ib_voxelize(context(), soup(mesh), ntri, bmin, bmax, resolution, nullptr, nullptr, 0, &count),
        "ib_voxelize");
  grid.cells.resize(count);
  if (count)
    check(ib_voxelize(context(), soup(mesh), ntri, bmin, bmax, resolution, grid.cells[0].data(), nullptr, count,
                      &count),
          "ib_voxelize");
  // colours: average over every touching triangle in triangle order, on the voxel grid
  const int r = resolution < 1 ? 1 : resolution;
  ib_grid vg{{grid.origin[0], grid.origin[1], grid.origin[2]}, grid.voxelSize, r, r, r, 0};
  std::vector<float> mat(5 * ntri);
  for (uint64_t t = 0; t < ntri; ++t) {
    color(t, &mat[5 * t]);
    mat[5 * t + 3] = 0.0f;
    mat[5 * t + 4] = 0.0f;
  }
  const uint64_t ncell = static_cast<uint64_t>(r) * r * r;
  std::vector<float> rgb(3 * ncell), m(ncell), rough(ncell);
  if (count) check(ib_splat_materials(context(), soup(mesh), ntri, mat.data(), &vg, rgb.data(), m.data(), rough.data()),
                   "ib_splat_materials");
  grid.colors.resize(count);
  for (uint64_t i = 0; i < count; ++i) {
    const auto &c = grid.cells[i];
    const uint64_t idx = (static_cast<uint64_t>(c[2]) * r + c[1]) * r + c[0];
    grid.colors[i] = {rgb[3 * idx], rgb[3 * idx + 1], rgb[3 * idx + 2]};
  }
  return grid;
}

template <class Mesh>
VoxelGrid voxelizeMesh(const Mesh &mesh, int resolution) {
  return voxelizeMesh(mesh, resolution, [](uint64_t, float *c) { c[0] = c[1] = c[2] = 1.0f; });
}

}  // namespace italy_bake
