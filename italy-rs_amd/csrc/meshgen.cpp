// Seeded synthetic mesh generators for the bake workloads (host only).
#include "meshgen.h"

#include <algorithm>
#include <atomic>
#include <thread>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

struct V3 {
  double x, y, z;
};
static V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
static double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static V3 normalize(V3 a) { return a * (1.0 / std::sqrt(dot(a, a))); }

struct IndexedMesh {
  std::vector<V3> v;
  std::vector<int> f; // 3 per face
};

void emit(const IndexedMesh &m, ib_soup *out) {
  const size_t nf = m.f.size() / 3;
  out->ntri = (int64_t)nf;
  out->xyz = (float *)std::malloc(sizeof(float) * 9 * (nf ? nf : 1));
  for (size_t t = 0; t < nf; ++t)
    for (int k = 0; k < 3; ++k) {
      const V3 p = m.v[m.f[3 * t + k]];
      out->xyz[9 * t + 3 * k + 0] = (float)p.x;
      out->xyz[9 * t + 3 * k + 1] = (float)p.y;
      out->xyz[9 * t + 3 * k + 2] = (float)p.z;
    }
}

void set_bounds(ib_soup *s) {
  for (int k = 0; k < 3; ++k) {
    s->bmin[k] = INFINITY;
    s->bmax[k] = -INFINITY;
  }
  for (int64_t i = 0; i < 3 * s->ntri; ++i)
    for (int k = 0; k < 3; ++k) {
      s->bmin[k] = std::min(s->bmin[k], s->xyz[3 * i + k]);
      s->bmax[k] = std::max(s->bmax[k], s->xyz[3 * i + k]);
    }
  if (s->ntri == 0)
    for (int k = 0; k < 3; ++k) s->bmin[k] = s->bmax[k] = 0.0f;
}

// Unit icosphere: icosahedron refined `level` times by edge midpoints projected to the sphere.
IndexedMesh unit_icosphere(int level) {
  const double p = (1.0 + std::sqrt(5.0)) / 2.0;
  IndexedMesh m;
  const double raw[12][3] = {{-1, p, 0}, {1, p, 0}, {-1, -p, 0}, {1, -p, 0}, {0, -1, p}, {0, 1, p},
                             {0, -1, -p}, {0, 1, -p}, {p, 0, -1}, {p, 0, 1}, {-p, 0, -1}, {-p, 0, 1}};
  for (auto &r : raw) m.v.push_back(normalize({r[0], r[1], r[2]}));
  const int faces[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                            {11, 10, 2}, {10, 7, 6}, {7, 1, 8},   {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
                            {3, 8, 9},  {4, 9, 5},  {2, 4, 11},  {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
  for (auto &f : faces) {
    int a = f[0], b = f[1], c = f[2];
    const V3 n = cross(m.v[b] - m.v[a], m.v[c] - m.v[a]);
    if (dot(n, m.v[a] + m.v[b] + m.v[c]) < 0) std::swap(b, c); // outward (counter-clockwise seen from outside)
    m.f.insert(m.f.end(), {a, b, c});
  }
  for (int l = 0; l < level; ++l) {
    std::unordered_map<uint64_t, int> mid;
    auto midpoint = [&](int a, int b) {
      const uint64_t key = a < b ? ((uint64_t)a << 32 | (uint32_t)b) : ((uint64_t)b << 32 | (uint32_t)a);
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      m.v.push_back(normalize((m.v[a] + m.v[b]) * 0.5));
      const int id = (int)m.v.size() - 1;
      mid.emplace(key, id);
      return id;
    };
    std::vector<int> nf;
    nf.reserve(m.f.size() * 4);
    for (size_t t = 0; t < m.f.size() / 3; ++t) {
      const int a = m.f[3 * t], b = m.f[3 * t + 1], c = m.f[3 * t + 2];
      const int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      nf.insert(nf.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
    }
    m.f.swap(nf);
  }
  return m;
}

struct Modes {
  double a[8], f[8], ph[8];
  V3 u[8];
};

V3 random_unit(SplitMix64 &r) {
  const double z = r.uniform(-1.0, 1.0);
  const double phi = r.uniform(0.0, 2.0 * M_PI);
  const double s = std::sqrt(std::max(0.0, 1.0 - z * z));
  return {s * std::cos(phi), s * std::sin(phi), z};
}

// a_k ~ U(0,1) normalised to sum 1, f_k ~ U(2,12), u_k random unit vectors, phi_k ~ U(0, 2pi)
Modes draw_modes(SplitMix64 &r) {
  Modes m;
  double sum = 0.0;
  for (int k = 0; k < 8; ++k) {
    m.a[k] = r.uniform();
    sum += m.a[k];
  }
  for (int k = 0; k < 8; ++k) m.a[k] /= sum;
  for (int k = 0; k < 8; ++k) m.f[k] = r.uniform(2.0, 12.0);
  for (int k = 0; k < 8; ++k) m.u[k] = random_unit(r);
  for (int k = 0; k < 8; ++k) m.ph[k] = r.uniform(0.0, 2.0 * M_PI);
  return m;
}

// r = 0.5 * (1 + 0.15 * sum_k a_k sin(f_k (n . u_k) + phi_k))
double displaced_radius(const Modes &m, V3 n) {
  double s = 0.0;
  for (int k = 0; k < 8; ++k) s += m.a[k] * std::sin(m.f[k] * dot(n, m.u[k]) + m.ph[k]);
  return 0.5 * (1.0 + 0.15 * s);
}

IndexedMesh displaced_icosphere(int level, uint64_t seed) {
  SplitMix64 r(seed);
  const Modes modes = draw_modes(r);
  IndexedMesh m = unit_icosphere(level);
  for (auto &p : m.v) p = p * displaced_radius(modes, p);
  return m;
}

// Random half-edge collapses (u <- v, u keeps its position) on a closed manifold mesh, each kept
// only if the link condition holds and no remaining face flips, until `target` faces remain.
void decimate(IndexedMesh &m, int64_t target, SplitMix64 &r) {
  const int nv = (int)m.v.size();
  const int nf = (int)(m.f.size() / 3);
  std::vector<std::vector<int>> vf(nv);
  for (int t = 0; t < nf; ++t)
    for (int k = 0; k < 3; ++k) vf[m.f[3 * t + k]].push_back(t);
  std::vector<char> alive(nf, 1);
  std::vector<int> live(nf), pos(nf);
  for (int t = 0; t < nf; ++t) live[t] = pos[t] = t;
  auto kill = [&](int t) {
    alive[t] = 0;
    const int p = pos[t], last = live.back();
    live[p] = last;
    pos[last] = p;
    live.pop_back();
  };
  auto normal = [&](int a, int b, int c) { return cross(m.v[b] - m.v[a], m.v[c] - m.v[a]); };
  std::vector<int> nu, nvv;
  int64_t count = nf;
  int64_t attempts = 0;
  while (count > target && attempts < 50 * (int64_t)nf) {
    ++attempts;
    const int t = live[r.next() % live.size()];
    const int k = (int)(r.next() % 3);
    const int u = m.f[3 * t + k], v = m.f[3 * t + (k + 1) % 3];
    // link condition: common neighbours of u and v are exactly the two opposite vertices
    nu.clear();
    nvv.clear();
    for (int f : vf[u])
      for (int q = 0; q < 3; ++q)
        if (m.f[3 * f + q] != u) nu.push_back(m.f[3 * f + q]);
    for (int f : vf[v])
      for (int q = 0; q < 3; ++q)
        if (m.f[3 * f + q] != v) nvv.push_back(m.f[3 * f + q]);
    std::sort(nu.begin(), nu.end());
    nu.erase(std::unique(nu.begin(), nu.end()), nu.end());
    std::sort(nvv.begin(), nvv.end());
    nvv.erase(std::unique(nvv.begin(), nvv.end()), nvv.end());
    int common = 0;
    bool thin = false;
    for (size_t i = 0, j = 0; i < nu.size() && j < nvv.size();) {
      if (nu[i] == nvv[j]) {
        ++common;
        thin |= vf[nu[i]].size() <= 3; // would drop to valence 2
        ++i;
        ++j;
      } else if (nu[i] < nvv[j]) {
        ++i;
      } else {
        ++j;
      }
    }
    if (common != 2 || thin) continue;
    if (vf[u].size() + vf[v].size() < 7) continue; // u keeps valence >= 3 after the merge
    bool ok = true;
    for (int f : vf[v]) {
      int a = m.f[3 * f], b = m.f[3 * f + 1], c = m.f[3 * f + 2];
      if (a == u || b == u || c == u) continue;
      const V3 n0 = normal(a, b, c);
      if (a == v) a = u;
      if (b == v) b = u;
      if (c == v) c = u;
      const V3 n1 = normal(a, b, c);
      if (dot(n0, n1) <= 0.25 * std::sqrt(dot(n0, n0) * dot(n1, n1))) {
        ok = false;
        break;
      }
    }
    if (!ok) continue;
    std::vector<int> moved;
    for (int f : vf[v]) {
      bool has_u = false;
      for (int q = 0; q < 3; ++q) has_u |= m.f[3 * f + q] == u;
      if (has_u) {
        kill(f);
        --count;
        for (int q = 0; q < 3; ++q) {
          auto &lst = vf[m.f[3 * f + q]];
          if (m.f[3 * f + q] == v) continue;
          lst.erase(std::remove(lst.begin(), lst.end(), f), lst.end());
        }
      } else {
        for (int q = 0; q < 3; ++q)
          if (m.f[3 * f + q] == v) m.f[3 * f + q] = u;
        moved.push_back(f);
      }
    }
    vf[v].clear();
    for (int f : moved) vf[u].push_back(f);
  }
  // compact faces and vertices
  std::vector<int> remap(nv, -1);
  IndexedMesh out;
  for (int t = 0; t < nf; ++t) {
    if (!alive[t]) continue;
    for (int q = 0; q < 3; ++q) {
      int &id = remap[m.f[3 * t + q]];
      if (id < 0) {
        id = (int)out.v.size();
        out.v.push_back(m.v[m.f[3 * t + q]]);
      }
      out.f.push_back(id);
    }
  }
  m.v.swap(out.v);
  m.f.swap(out.f);
}

// uniform random rotation from a unit quaternion (Shoemake 1992)
void random_rotation(SplitMix64 &r, double R[9]) {
  const double u1 = r.uniform(), u2 = r.uniform(), u3 = r.uniform();
  const double a = std::sqrt(1 - u1), b = std::sqrt(u1);
  const double qx = a * std::sin(2 * M_PI * u2), qy = a * std::cos(2 * M_PI * u2);
  const double qz = b * std::sin(2 * M_PI * u3), qw = b * std::cos(2 * M_PI * u3);
  R[0] = 1 - 2 * (qy * qy + qz * qz);
  R[1] = 2 * (qx * qy - qz * qw);
  R[2] = 2 * (qx * qz + qy * qw);
  R[3] = 2 * (qx * qy + qz * qw);
  R[4] = 1 - 2 * (qx * qx + qz * qz);
  R[5] = 2 * (qy * qz - qx * qw);
  R[6] = 2 * (qx * qz - qy * qw);
  R[7] = 2 * (qy * qz + qx * qw);
  R[8] = 1 - 2 * (qx * qx + qy * qy);
}

V3 xform(const double R[9], V3 t, V3 p) {
  return {R[0] * p.x + R[1] * p.y + R[2] * p.z + t.x, R[3] * p.x + R[4] * p.y + R[5] * p.z + t.y,
          R[6] * p.x + R[7] * p.y + R[8] * p.z + t.z};
}

// torus in the xy plane, nu x nv segments, outward oriented
IndexedMesh torus(double R, double rr, int nu, int nvs) {
  IndexedMesh m;
  for (int i = 0; i < nu; ++i)
    for (int j = 0; j < nvs; ++j) {
      const double u = 2 * M_PI * i / nu, v = 2 * M_PI * j / nvs;
      m.v.push_back({(R + rr * std::cos(v)) * std::cos(u), (R + rr * std::cos(v)) * std::sin(u), rr * std::sin(v)});
    }
  auto id = [&](int i, int j) { return (i % nu) * nvs + (j % nvs); };
  for (int i = 0; i < nu; ++i)
    for (int j = 0; j < nvs; ++j) {
      const int a = id(i, j), b = id(i + 1, j), c = id(i + 1, j + 1), d = id(i, j + 1);
      m.f.insert(m.f.end(), {a, b, c, a, c, d});
    }
  // orientation: normal of the first face should point away from the tube centre line
  const V3 n = cross(m.v[m.f[1]] - m.v[m.f[0]], m.v[m.f[2]] - m.v[m.f[0]]);
  const V3 cen = (m.v[m.f[0]] + m.v[m.f[1]] + m.v[m.f[2]]) * (1.0 / 3.0);
  const V3 ring = normalize(V3{cen.x, cen.y, 0}) * R;
  if (dot(n, cen - ring) < 0)
    for (size_t t = 0; t < m.f.size() / 3; ++t) std::swap(m.f[3 * t + 1], m.f[3 * t + 2]);
  return m;
}

// box with sides s, each face an n x n grid of quads; vertices come from an integer lattice so
// face seams are shared exactly; outward oriented
IndexedMesh subdivided_box(const double s[3], int n) {
  IndexedMesh m;
  std::unordered_map<uint64_t, int> ids;
  auto vid = [&](int i, int j, int k) {
    const uint64_t key = ((uint64_t)i << 40) | ((uint64_t)j << 20) | (uint64_t)k;
    auto it = ids.find(key);
    if (it != ids.end()) return it->second;
    m.v.push_back({((double)i / n - 0.5) * s[0], ((double)j / n - 0.5) * s[1], ((double)k / n - 0.5) * s[2]});
    ids.emplace(key, (int)m.v.size() - 1);
    return (int)m.v.size() - 1;
  };
  for (int axis = 0; axis < 3; ++axis)
    for (int side = 0; side < 2; ++side) {
      const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
      for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) {
          int c[4][3];
          const int pp[4] = {p, p + 1, p + 1, p}, qq[4] = {q, q, q + 1, q + 1};
          for (int k = 0; k < 4; ++k) {
            c[k][axis] = side * n;
            c[k][a1] = pp[k];
            c[k][a2] = qq[k];
          }
          int v[4];
          for (int k = 0; k < 4; ++k) v[k] = vid(c[k][0], c[k][1], c[k][2]);
          // (a1 x a2) = +axis; flip on the negative side
          if (side == 1) m.f.insert(m.f.end(), {v[0], v[1], v[2], v[0], v[2], v[3]});
          else m.f.insert(m.f.end(), {v[0], v[2], v[1], v[0], v[3], v[2]});
        }
    }
  return m;
}

} // namespace

extern "C" {

// tests/sdf_sign_test.cpp:18-42: the float expressions are kept as written there
int ib_gen_uv_sphere(float r, int lonSegs, int latSegs, ib_soup *out) {
  std::memset(out, 0, sizeof(*out));
  if (lonSegs <= 0 || latSegs <= 0) return -1;
  auto vertex = [&](int lat, int lon, float *p) {
    const float theta = static_cast<float>(lat) / latSegs * static_cast<float>(M_PI);
    const float phi = static_cast<float>(lon) / lonSegs * 2.0f * static_cast<float>(M_PI);
    p[0] = r * std::sin(theta) * std::cos(phi);
    p[1] = r * std::cos(theta);
    p[2] = r * std::sin(theta) * std::sin(phi);
  };
  out->ntri = (int64_t)2 * lonSegs * latSegs;
  out->xyz = (float *)std::malloc(sizeof(float) * 9 * out->ntri);
  float *w = out->xyz;
  for (int lat = 0; lat < latSegs; ++lat)
    for (int lon = 0; lon < lonSegs; ++lon) {
      float a[3], b[3], c[3], d[3];
      vertex(lat, lon, a);
      vertex(lat + 1, lon, b);
      vertex(lat + 1, lon + 1, c);
      vertex(lat, lon + 1, d);
      const float *seq[6] = {a, b, c, a, c, d};
      for (auto *v : seq) {
        std::memcpy(w, v, 3 * sizeof(float));
        w += 3;
      }
    }
  // the test sets the bounds by hand (sdf_sign_test.cpp:38-39)
  for (int k = 0; k < 3; ++k) {
    out->bmin[k] = -r;
    out->bmax[k] = r;
  }
  return 0;
}

// tests/voxelize_test.cpp:11-33
int ib_gen_cube_shell(ib_soup *out) {
  std::memset(out, 0, sizeof(*out));
  const float lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  const float quads[6][4][3] = {
      {{lo[0], lo[1], lo[2]}, {hi[0], lo[1], lo[2]}, {hi[0], hi[1], lo[2]}, {lo[0], hi[1], lo[2]}},
      {{lo[0], lo[1], hi[2]}, {lo[0], hi[1], hi[2]}, {hi[0], hi[1], hi[2]}, {hi[0], lo[1], hi[2]}},
      {{lo[0], lo[1], lo[2]}, {lo[0], hi[1], lo[2]}, {lo[0], hi[1], hi[2]}, {lo[0], lo[1], hi[2]}},
      {{hi[0], lo[1], lo[2]}, {hi[0], lo[1], hi[2]}, {hi[0], hi[1], hi[2]}, {hi[0], hi[1], lo[2]}},
      {{lo[0], lo[1], lo[2]}, {lo[0], lo[1], hi[2]}, {hi[0], lo[1], hi[2]}, {hi[0], lo[1], lo[2]}},
      {{lo[0], hi[1], lo[2]}, {hi[0], hi[1], lo[2]}, {hi[0], hi[1], hi[2]}, {lo[0], hi[1], hi[2]}}};
  out->ntri = 12;
  out->xyz = (float *)std::malloc(sizeof(float) * 9 * 12);
  float *w = out->xyz;
  for (auto &q : quads) {
    const int seq[6] = {0, 1, 2, 0, 2, 3}; // addQuad: a, b, c, a, c, d
    for (int s : seq) {
      std::memcpy(w, q[s], 3 * sizeof(float));
      w += 3;
    }
  }
  for (int k = 0; k < 3; ++k) {
    out->bmin[k] = lo[k];
    out->bmax[k] = hi[k];
  }
  return 0;
}

int ib_gen_icosphere(int level, double radius, double jitter, uint64_t seed, ib_soup *out) {
  std::memset(out, 0, sizeof(*out));
  if (level < 0 || level > 10) return -1;
  IndexedMesh m = unit_icosphere(level);
  SplitMix64 r(seed);
  for (auto &p : m.v) p = p * (radius + (jitter > 0 ? r.uniform(-jitter, jitter) : 0.0));
  emit(m, out);
  set_bounds(out);
  return 0;
}

int ib_gen_displaced(int level, uint64_t seed, ib_soup *out) {
  std::memset(out, 0, sizeof(*out));
  if (level < 0 || level > 10) return -1;
  emit(displaced_icosphere(level, seed), out);
  set_bounds(out);
  return 0;
}

int ib_gen_decimated(int64_t target, uint64_t seed, ib_soup *out) {
  std::memset(out, 0, sizeof(*out));
  if (target < 20) return -1;
  target += target & 1; // closed triangle meshes of genus 0 have an even face count
  int level = 0;
  while ((int64_t)20 << (2 * level) < target) ++level;
  if (level > 10) return -1;
  IndexedMesh m = displaced_icosphere(level, seed);
  SplitMix64 r(seed ^ 0xD1CE5EEDULL);
  if ((int64_t)(m.f.size() / 3) > target) decimate(m, target, r);
  emit(m, out);
  set_bounds(out);
  return 0;
}

int ib_gen_stress(uint64_t seed, int ncomp, double remove_frac, ib_soup *out) {
  return ib_gen_stress_res(seed, ncomp, remove_frac, 64, 32, 32, out);
}

int ib_gen_stress_res(uint64_t seed, int ncomp, double remove_frac, int torus_nu, int torus_nv, int box_n,
                      ib_soup *out) {
  std::memset(out, 0, sizeof(*out));
  if (ncomp <= 0 || torus_nu < 3 || torus_nv < 3 || box_n < 1) return -1;
  SplitMix64 r(seed);
  IndexedMesh all;
  for (int c = 0; c < ncomp; ++c) {
    IndexedMesh part;
    if ((c & 1) == 0) {
      const double R = r.uniform(0.05, 0.2), rr = r.uniform(0.01, 0.05);
      part = torus(R, rr, torus_nu, torus_nv);
    } else {
      const double s[3] = {r.uniform(0.05, 0.3), r.uniform(0.05, 0.3), r.uniform(0.05, 0.3)};
      part = subdivided_box(s, box_n);
    }
    double Rm[9];
    random_rotation(r, Rm);
    const V3 t = {r.uniform(-0.7, 0.7), r.uniform(-0.7, 0.7), r.uniform(-0.7, 0.7)};
    const int base = (int)all.v.size();
    for (auto &p : part.v) all.v.push_back(xform(Rm, t, p));
    for (int id : part.f) all.f.push_back(base + id);
  }
  const int nf = (int)(all.f.size() / 3);
  const int nremove = (int)std::floor(remove_frac * nf);
  std::vector<int> perm(nf);
  for (int i = 0; i < nf; ++i) perm[i] = i;
  for (int i = 0; i < nremove; ++i) { // partial Fisher-Yates
    const int j = i + (int)(r.next() % (uint64_t)(nf - i));
    std::swap(perm[i], perm[j]);
  }
  std::vector<char> removed(nf, 0);
  for (int i = 0; i < nremove; ++i) removed[perm[i]] = 1;
  IndexedMesh kept;
  kept.v = all.v;
  for (int t = 0; t < nf; ++t)
    if (!removed[t]) kept.f.insert(kept.f.end(), {all.f[3 * t], all.f[3 * t + 1], all.f[3 * t + 2]});
  emit(kept, out);
  set_bounds(out);
  return 0;
}

int ib_gen_decimated_batch(int n, const int64_t *targets, const uint64_t *seeds, int threads, ib_soup *outs) {
  if (n < 0 || (n && (!targets || !seeds || !outs))) return -1;
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  threads = std::max(1, std::min(threads, n));
  std::atomic<int> next{0}, err{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&] {
      for (int i = next++; i < n; i = next++)
        if (ib_gen_decimated(targets[i], seeds[i], &outs[i]) != 0) err = 1;
    });
  for (auto &th : pool) th.join();
  return err.load() ? -1 : 0;
}

void ib_soup_free(ib_soup *s) {
  if (!s) return;
  std::free(s->xyz);
  s->xyz = nullptr;
  s->ntri = 0;
}

} // extern "C"
