// The reference's two hot-path tests, run through the C++ drop-ins of include/italy_bake.hpp:
//   tests/sdf_sign_test.cpp:46-82  (UV sphere r=1, 32x16, bakeSdf res 24, sign + magnitude)
//   tests/voxelize_test.cpp:37-58  (unit cube shell, voxelizeMesh res 8, membership)
// the reference's watertight gate test (tests/mesh_validate_test.cpp:37-62) with its messages,
// a textured MeshAsset's per-triangle materials (src/convert/sdf_baker.cpp:42-74) against the oracle,
// and the procedural scenes of src/convert/sdf_procedural.cpp:81-193 through the analytic drop-ins,
// checked bit for bit against the oracle's per-cell restatement of the same lambdas.
// Checks print a line and set a non-zero exit status instead of relying on assert.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ib_oracle.h"
#include "italy_bake.hpp"

namespace {

struct vec3 {  // three packed floats, like glm::vec3
  float x, y, z;
};

struct Mesh {  // the fields of italy::MeshAsset the bake reads
  std::vector<vec3> positions;
  vec3 boundsMin{0, 0, 0}, boundsMax{0, 0, 0};
};

struct vec2 {
  float x, y;
};
struct TextureAsset {  // italy::TextureAsset
  std::vector<uint8_t> pixelsRGBA;
  int width = 0, height = 0;
  bool srgb = false;
};
struct MaterialAsset {  // the fields of italy::MaterialAsset the bake reads
  vec3 baseColorFactor{1, 1, 1};
  float metallic = 0.0f, roughness = 1.0f;
  int baseColorTexture = -1;
};
struct TexturedMesh : Mesh {  // italy::MeshAsset with materials
  std::vector<vec2> uvs;
  std::vector<uint32_t> triangleMaterial;
  std::vector<MaterialAsset> materials;
  std::vector<TextureAsset> textures;
};

int failures = 0;
void expect(bool ok, const char *what) {
  if (!ok) {
    std::printf("FAILED: %s\n", what);
    ++failures;
  }
}

Mesh makeUvSphere(float r, int lonSegs, int latSegs) {  // sdf_sign_test.cpp:18-42
  Mesh mesh;
  auto vertex = [&](int lat, int lon) {
    const float theta = static_cast<float>(lat) / latSegs * static_cast<float>(M_PI);
    const float phi = static_cast<float>(lon) / lonSegs * 2.0f * static_cast<float>(M_PI);
    return vec3{r * std::sin(theta) * std::cos(phi), r * std::cos(theta), r * std::sin(theta) * std::sin(phi)};
  };
  for (int lat = 0; lat < latSegs; ++lat)
    for (int lon = 0; lon < lonSegs; ++lon) {
      const vec3 a = vertex(lat, lon), b = vertex(lat + 1, lon), c = vertex(lat + 1, lon + 1), d = vertex(lat, lon + 1);
      for (vec3 v : {a, b, c, a, c, d}) mesh.positions.push_back(v);
    }
  mesh.boundsMin = {-r, -r, -r};
  mesh.boundsMax = {r, r, r};
  return mesh;
}

void addQuad(Mesh &mesh, vec3 a, vec3 b, vec3 c, vec3 d) {  // voxelize_test.cpp:11-17
  for (vec3 v : {a, b, c, a, c, d}) mesh.positions.push_back(v);
}

Mesh makeUnitCubeShell() {  // voxelize_test.cpp:19-33
  Mesh mesh;
  const vec3 lo{0, 0, 0}, hi{1, 1, 1};
  addQuad(mesh, {lo.x, lo.y, lo.z}, {hi.x, lo.y, lo.z}, {hi.x, hi.y, lo.z}, {lo.x, hi.y, lo.z});
  addQuad(mesh, {lo.x, lo.y, hi.z}, {lo.x, hi.y, hi.z}, {hi.x, hi.y, hi.z}, {hi.x, lo.y, hi.z});
  addQuad(mesh, {lo.x, lo.y, lo.z}, {lo.x, hi.y, lo.z}, {lo.x, hi.y, hi.z}, {lo.x, lo.y, hi.z});
  addQuad(mesh, {hi.x, lo.y, lo.z}, {hi.x, lo.y, hi.z}, {hi.x, hi.y, hi.z}, {hi.x, hi.y, lo.z});
  addQuad(mesh, {lo.x, lo.y, lo.z}, {lo.x, lo.y, hi.z}, {hi.x, lo.y, hi.z}, {hi.x, lo.y, lo.z});
  addQuad(mesh, {lo.x, hi.y, lo.z}, {hi.x, hi.y, lo.z}, {hi.x, hi.y, hi.z}, {lo.x, hi.y, hi.z});
  mesh.boundsMin = lo;
  mesh.boundsMax = hi;
  return mesh;
}

void sdfSignTest() {
  const float radius = 1.0f;
  const Mesh sphere = makeUvSphere(radius, 32, 16);
  const italy_bake::SdfGrid grid = italy_bake::bakeSdf(sphere, 24);
  expect(!grid.distances.empty(), "distances non-empty");
  const float tolerance = 1.5f * grid.voxelSize;
  size_t checked = 0, wrongSign = 0;
  double maxAbsError = 0.0;
  for (int z = 0; z < grid.nz; ++z)
    for (int y = 0; y < grid.ny; ++y)
      for (int x = 0; x < grid.nx; ++x) {
        const float cx = grid.origin[0] + (static_cast<float>(x) + 0.5f) * grid.voxelSize;
        const float cy = grid.origin[1] + (static_cast<float>(y) + 0.5f) * grid.voxelSize;
        const float cz = grid.origin[2] + (static_cast<float>(z) + 0.5f) * grid.voxelSize;
        const float analytic = std::sqrt(cx * cx + cy * cy + cz * cz) - radius;
        const float baked = grid.distances[grid.index(x, y, z)];
        if (std::fabs(analytic) > tolerance) {
          ++checked;
          if ((analytic < 0.0f) != (baked < 0.0f)) ++wrongSign;
          maxAbsError = std::max(maxAbsError, static_cast<double>(std::fabs(baked - analytic)));
        }
      }
  std::printf("sdf_sign_test: %zu/%zu cells checked away from the surface, max |error| = %.4f (tolerance %.4f)\n",
              checked - wrongSign, checked, maxAbsError, static_cast<double>(tolerance));
  expect(checked > 0, "cells outside the tolerance band");
  expect(wrongSign == 0, "no sign disagreement with the analytic sphere");
  expect(maxAbsError < 2.0 * tolerance, "magnitude within 3h of the analytic sphere");
  // materials: the default material everywhere it was splatted, defaults elsewhere
  expect(grid.baseColor.size() == grid.distances.size(), "per-cell materials present");
  bool matOk = true;
  for (size_t i = 0; i < grid.baseColor.size(); ++i)
    matOk &= grid.baseColor[i][0] == 1.0f && grid.metallic[i] == 0.0f && grid.roughness[i] == 1.0f;
  expect(matOk, "default material");
}

void voxelizeTest() {
  const int resolution = 8;
  const Mesh cube = makeUnitCubeShell();
  const auto grid = italy_bake::voxelizeMesh(cube, resolution);
  const size_t maxShellCells = static_cast<size_t>(6 * resolution * resolution);
  expect(!grid.cells.empty(), "cells non-empty");
  expect(grid.cells.size() <= maxShellCells, "at most the shell's cells");
  bool foundCorner = false, foundCenter = false;
  for (const auto &c : grid.cells) {
    if (c[0] == 0 && c[1] == 0 && c[2] == 0) foundCorner = true;
    if (c[0] == 4 && c[1] == 4 && c[2] == 4) foundCenter = true;
  }
  expect(foundCorner, "corner cell should be on the shell");
  expect(!foundCenter, "center cell is interior");
  expect(grid.colors.size() == grid.cells.size(), "one colour per cell");
  std::printf("voxelize_test: %zu cells\n", grid.cells.size());
}

// tests/mesh_validate_test.cpp:37-62 through the drop-ins (its makeUnitCube has the quads of the cube
// shell above), with the reference's message text; weldPositions against the oracle's restatement
void meshValidateTest() {
  {
    const Mesh cube = makeUnitCubeShell();
    std::string err;
    expect(italy_bake::isWatertight(cube, err), "unit cube is watertight");
    expect(err.empty(), "no message for a watertight cube");
    std::vector<vec3> unique;
    const std::vector<uint32_t> w = italy_bake::weldPositions(cube.positions, cube.boundsMin, cube.boundsMax, &unique);
    expect(italy_bake::weldedIndexCount(w) == 8 && unique.size() == 8, "unit cube welds to 8 corners");
    std::vector<uint32_t> ow(cube.positions.size());
    std::vector<float> ou(3 * cube.positions.size());
    const float bmin[3] = {0, 0, 0}, bmax[3] = {1, 1, 1};
    const int64_t on = orc_weld_positions(&cube.positions[0].x, cube.positions.size(), bmin, bmax, ow.data(), ou.data());
    expect(on == 8 && ow == w, "weld indices equal the oracle's");
    expect(std::memcmp(ou.data(), unique.data(), 8 * sizeof(vec3)) == 0, "unique positions equal the oracle's");
  }
  {
    Mesh holey = makeUnitCubeShell();
    holey.positions.resize(holey.positions.size() - 6);
    std::string err;
    expect(!italy_bake::isWatertight(holey, err), "holey cube is not watertight");
    expect(err == "mesh is not watertight (4 boundary edges) \u2014 fill holes and merge duplicate vertices in "
                  "your DCC tool and re-export",
           "holey cube message");
  }
  {
    Mesh dup = makeUnitCubeShell();
    for (int k = 0; k < 3; ++k) dup.positions.push_back(dup.positions[k]);
    std::string err;
    expect(!italy_bake::isWatertight(dup, err), "duplicated face is not watertight");
    expect(err == "mesh is not watertight (0 boundary edges, 3 non-manifold edges) \u2014 fill holes and merge "
                  "duplicate vertices in your DCC tool and re-export",
           "non-manifold message");
  }
  {
    Mesh empty;
    std::string err;
    expect(!italy_bake::isWatertight(empty, err), "empty mesh is not watertight");
    expect(err == "mesh has no triangles", "empty mesh message");
  }
}

}  // namespace

void cupTest(bool glass, float cx, float cz) {
  const int res = 96;
  const italy_bake::SdfGrid g =
      glass ? italy_bake::makeGlassCupSdf(res, -1.0f, cx, cz) : italy_bake::makeCupSdf(res, -1.0f, cx, cz);
  float o[3], h;
  int32_t n[3];
  orc_cup_cells(glass, res, -1.0f, cx, cz, o, &h, n, nullptr, nullptr);
  const size_t cells = static_cast<size_t>(n[0]) * n[1] * n[2];
  std::vector<float> d(cells);
  std::vector<uint8_t> b(cells);
  orc_cup_cells(glass, res, -1.0f, cx, cz, o, &h, n, d.data(), b.data());
  expect(g.nx == n[0] && g.ny == n[1] && g.nz == n[2] && g.voxelSize == h, "cup framing");
  expect(g.origin[0] == o[0] && g.origin[1] == o[1] && g.origin[2] == o[2], "cup origin");
  expect(g.distances.size() == cells && std::memcmp(g.distances.data(), d.data(), cells * 4) == 0,
         "cup distances bit-exact");
  if (glass) {
    expect(g.branch.size() == cells && std::memcmp(g.branch.data(), b.data(), cells) == 0, "glass cup branches");
    expect(g.palette.size() == 2 && g.palette[0].ior == 1.47f && g.palette[1].ior == 1.333f, "glass cup palette");
  } else {
    expect(g.roughness.size() == cells && g.roughness[0] == 0.35f && g.baseColor[cells / 2][0] == 0.85f,
           "cup material");
  }
}

// every triangle of the cube shell gets a material; the device table equals the oracle's bit for bit, and
// voxelizeMesh / bakeSdf carry a uniform textured colour into every cell
void texturedTest() {
  TexturedMesh mesh;
  static_cast<Mesh &>(mesh) = makeUnitCubeShell();
  const size_t ntri = mesh.positions.size() / 3;
  TextureAsset srgbTex, linTex;
  srgbTex.width = 3, srgbTex.height = 2, srgbTex.srgb = true;
  linTex.width = 2, linTex.height = 2;
  for (int i = 0; i < 3 * 2 * 4; ++i) srgbTex.pixelsRGBA.push_back(static_cast<uint8_t>(37 * i + 11));
  for (int i = 0; i < 2 * 2 * 4; ++i) linTex.pixelsRGBA.push_back(static_cast<uint8_t>(53 * i + 5));
  mesh.textures = {srgbTex, linTex};
  mesh.materials.resize(4);
  mesh.materials[0].baseColorFactor = {0.25f, 0.5f, 0.75f};
  mesh.materials[1].baseColorTexture = 0, mesh.materials[1].metallic = 0.5f;
  mesh.materials[2].baseColorTexture = 1, mesh.materials[2].roughness = 0.25f;
  mesh.materials[2].baseColorFactor = {0.9f, 0.8f, 0.7f};
  mesh.materials[3].baseColorTexture = 7;  // no such texture: the factor
  for (size_t t = 0; t + 2 < ntri; ++t) mesh.triangleMaterial.push_back(static_cast<uint32_t>(t % 5));
  for (size_t c = 0; c < 3 * ntri; ++c)
    mesh.uvs.push_back({0.37f * static_cast<float>(c) - 2.1f, 1.3f - 0.29f * static_cast<float>(c)});
  const std::vector<float> got = italy_bake::detail::triangle_materials(mesh);

  std::vector<orc_material> om;
  for (const auto &m : mesh.materials)
    om.push_back(orc_material{{m.baseColorFactor.x, m.baseColorFactor.y, m.baseColorFactor.z}, m.metallic,
                              m.roughness, m.baseColorTexture});
  std::vector<orc_texture> ot;
  for (const auto &t : mesh.textures) ot.push_back(orc_texture{t.pixelsRGBA.data(), t.width, t.height, t.srgb});
  std::vector<float> want(5 * ntri);
  orc_triangle_materials(ntri, &mesh.uvs[0].x, mesh.uvs.size(), mesh.triangleMaterial.data(),
                         mesh.triangleMaterial.size(), om.data(), om.size(), ot.data(), ot.size(), want.data());
  expect(std::memcmp(got.data(), want.data(), want.size() * sizeof(float)) == 0,
         "device triangle materials equal the oracle's");

  // five different materials over the faces: voxelizeMesh's cells and colour means equal the oracle's
  {
    std::vector<float> rgb(3 * ntri);
    for (size_t t = 0; t < ntri; ++t)
      for (int k = 0; k < 3; ++k) rgb[3 * t + k] = want[5 * t + k];
    const float bmin[3] = {mesh.boundsMin.x, mesh.boundsMin.y, mesh.boundsMin.z};
    const float bmax[3] = {mesh.boundsMax.x, mesh.boundsMax.y, mesh.boundsMax.z};
    const size_t n = orc_voxelize_first_touch(&mesh.positions[0].x, ntri, rgb.data(), bmin, bmax, 13, nullptr,
                                              nullptr, 0);
    std::vector<int32_t> rc(3 * n);
    std::vector<float> rcol(3 * n);
    orc_voxelize_first_touch(&mesh.positions[0].x, ntri, rgb.data(), bmin, bmax, 13, rc.data(), rcol.data(), n);
    const auto vg = italy_bake::voxelizeMesh(mesh, 13);
    expect(n > 0 && vg.cells.size() == n && std::memcmp(vg.cells.data(), rc.data(), rc.size() * 4) == 0 &&
               std::memcmp(vg.colors.data(), rcol.data(), rcol.size() * 4) == 0,
           "voxelizeMesh cells and per-cell colours equal the oracle's");
  }

  // one textured material on every triangle, all UVs on one texel: every cell carries its colour
  for (auto &m : mesh.triangleMaterial) m = 2;
  mesh.triangleMaterial.resize(ntri, 2);
  for (auto &uv : mesh.uvs) uv = {0.75f, 0.25f};  // texel (1, 0)
  const float r = 0.9f * (static_cast<float>(linTex.pixelsRGBA[4]) / 255.0f);
  const auto vg = italy_bake::voxelizeMesh(mesh, 8);
  // cells average the colours of the triangles they touch: equal colours average to r within rounding
  bool colOk = !vg.colors.empty();
  for (const auto &c : vg.colors) colOk &= std::fabs(c[0] - r) <= 1e-6f;
  expect(colOk, "voxelizeMesh carries the texel colour");
  // bakeSdf's per-cell planes equal the oracle's splat of the oracle's per-triangle materials.  At 13 the
  // cube's faces cut through cells; at 12 (and 24) they fall on cell faces, where triBoxOverlap's rounding
  // leaves both neighbours untouched and every cell keeps the default material, on the device as in the oracle
  const auto sg = italy_bake::bakeSdf(mesh, 13);
  std::vector<float> tm(5 * ntri);
  orc_triangle_materials(ntri, &mesh.uvs[0].x, mesh.uvs.size(), mesh.triangleMaterial.data(),
                         mesh.triangleMaterial.size(), om.data(), om.size(), ot.data(), ot.size(), tm.data());
  const size_t ncell = sg.distances.size();
  std::vector<float> planes(5 * ncell);
  orc_splat_materials(&mesh.positions[0].x, ntri, tm.data(), sg.origin.data(), sg.voxelSize, sg.nx, sg.ny, sg.nz,
                      planes.data());
  size_t splatted = 0, bad = 0;
  for (size_t i = 0; i < ncell; ++i) {
    const float want5[5] = {planes[i], planes[ncell + i], planes[2 * ncell + i], planes[3 * ncell + i],
                            planes[4 * ncell + i]};
    const float got5[5] = {sg.baseColor[i][0], sg.baseColor[i][1], sg.baseColor[i][2], sg.metallic[i],
                           sg.roughness[i]};
    splatted += want5[4] != 1.0f;
    if (std::memcmp(want5, got5, sizeof(want5)) != 0 && bad++ == 0)
      std::printf("cell %zu: got %.9g %.9g %.9g %.9g %.9g, oracle %.9g %.9g %.9g %.9g %.9g\n", i, got5[0], got5[1],
                  got5[2], got5[3], got5[4], want5[0], want5[1], want5[2], want5[3], want5[4]);
  }
  expect(bad == 0 && splatted > 0, "bakeSdf carries the textured material (oracle splat)");
  std::printf("textured_test: %zu triangles, %zu voxel cells\n", ntri, vg.colors.size());
}

void probeTest() {
  const std::vector<italy_bake::SdfGrid> grids = italy_bake::makeMaterialProbeSdfs(48);
  expect(grids.size() == 9, "nine probes");
  for (const auto &g : grids) {
    // a sphere of radius 0.4 framed at 48: the cell nearest the centre is inside by about the radius
    float lo = 1e30f;
    for (float v : g.distances) lo = std::min(lo, v);
    expect(g.nx == 52 && g.ny == 52 && g.nz == 52, "probe framing 48 + 4");
    expect(lo < -0.38f && lo > -0.4f, "probe centre depth");
  }
  expect(grids[4].palette[0].roughness == 1.0f && grids[8].palette[0].ior == 2.0f, "probe palettes");
}

int main() {
  try {
    sdfSignTest();
    voxelizeTest();
    meshValidateTest();
    texturedTest();
    cupTest(false, -1.6f, 1.6f);
    cupTest(true, 0.0f, 0.0f);
    probeTest();
  } catch (const std::exception &e) {
    std::printf("FAILED: exception %s\n", e.what());
    return 2;
  }
  std::printf(failures ? "dropin_test: FAILED\n" : "dropin_test: OK\n");
  return failures ? 1 : 0;
}
