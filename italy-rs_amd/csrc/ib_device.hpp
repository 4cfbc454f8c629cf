// Device-side geometry for the bake kernels (gfx950).  Every translation unit that includes this
// is compiled with -ffp-contract=off, so each a*b+c below is two correctly rounded operations and
// the triangle/box test is bit-identical to the host restatement of tri_box_overlap.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ib {

constexpr int kStack = 64;        // traversal stack entries (>= LBVH depth, checked at build)
constexpr int kStack4 = 128;      // brick query stack over the 4-wide view: <= 3 pushes per level
constexpr int kColCap = 32;       // crossings stored per +z column before the per-voxel fallback
constexpr uint32_t kSliver = 1u;  // triangle flag: distance evaluated in FP64
// leaf code ~L: first triangle L & kBlockMask, count (L >> kBlockShift) + 1 (meshes < 2^28 triangles)
constexpr int kBlockShift = 28;
constexpr int kBlockMask = (1 << kBlockShift) - 1;
constexpr int kMaxLeafBlock = 8;

// Internal LBVH node: both children's boxes inline so one 64-byte record decides both children.
// child >= 0: internal node index, child < 0: ~leaf (sorted triangle position).
struct alignas(16) Node {
  float lo0[3], hi0[3];
  float lo1[3], hi1[3];
  int32_t c0, c1, parent, pad;
};
static_assert(sizeof(Node) == 64, "node record is one 64-byte line");

// 4-wide view of the LBVH for the brick query: the record of internal node x holds x's
// grandchildren (an internal child is replaced by its two children; a leaf or leaf block stays one
// entry): up to 4 boxes and child codes (internal node index, whose own Node4 follows, or a leaf
// code < 0).  Entries past `count` are unused.
struct alignas(16) Node4 {
  float lo[4][3];
  float hi[4][3];
  int32_t child[4];
  int32_t count, pad[3];
};
static_assert(sizeof(Node4) == 128, "4-wide record is two 64-byte lines");

// Node of the hole-cap hierarchy (built on the host).  left >= 0: children left/right;
// left < 0: leaf holding triangles [-1 - left, -1 - left + right) of the reordered cap soup.
// (c, r) bounds every vertex below; area = sum of |triangle area|; dip = sum of area vectors.
struct alignas(16) CapNode {
  float c[3], r;
  float dip[3], area;
  float lo[3];
  int32_t left;
  float hi[3];
  int32_t right;
};
static_assert(sizeof(CapNode) == 64, "cap node record is one 64-byte line");
constexpr int kCapStack = 48;
// a cap node farther than this bound on |its solid angle| / 4pi is not evaluated, only bounded
constexpr float kCapTau = 1e-3f;

// ---------------------------------------------------------------------------------------------
// triangle / AABB separating axis test: same expression trees as src/convert/tri_box_overlap.h
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }

struct F3 {
  float x, y, z;
};

__device__ __forceinline__ bool sat_axis(float p0, float p1, float rad) {
  const float mn = smin(p0, p1), mx = smax(p0, p1);
  return !(mn > rad || mx < -rad);
}

// box centre c, half size bh (all components equal here), triangle t0,t1,t2
__device__ __forceinline__ bool tri_box(F3 c, float bh, F3 t0, F3 t1, F3 t2) {
  const F3 v0 = {t0.x - c.x, t0.y - c.y, t0.z - c.z};
  const F3 v1 = {t1.x - c.x, t1.y - c.y, t1.z - c.z};
  const F3 v2 = {t2.x - c.x, t2.y - c.y, t2.z - c.z};
  const F3 e0 = {v1.x - v0.x, v1.y - v0.y, v1.z - v0.z};
  const F3 e1 = {v2.x - v1.x, v2.y - v1.y, v2.z - v1.z};
  const F3 e2 = {v0.x - v2.x, v0.y - v2.y, v0.z - v2.z};
  float fex = fabsf(e0.x), fey = fabsf(e0.y), fez = fabsf(e0.z);
  // X01(a=e.z, b=e.y): p = a*v.y - b*v.z over v0,v2 ; rad = fa*h + fb*h
  if (!sat_axis(e0.z * v0.y - e0.y * v0.z, e0.z * v2.y - e0.y * v2.z, fez * bh + fey * bh)) return false;
  // Y02(a=e.z, b=e.x): p = -a*v.x + b*v.z over v0,v2
  if (!sat_axis(-e0.z * v0.x + e0.x * v0.z, -e0.z * v2.x + e0.x * v2.z, fez * bh + fex * bh)) return false;
  // Z12(a=e.y, b=e.x): p = a*v.x - b*v.y over v1,v2
  if (!sat_axis(e0.y * v1.x - e0.x * v1.y, e0.y * v2.x - e0.x * v2.y, fey * bh + fex * bh)) return false;
  fex = fabsf(e1.x);
  fey = fabsf(e1.y);
  fez = fabsf(e1.z);
  if (!sat_axis(e1.z * v0.y - e1.y * v0.z, e1.z * v2.y - e1.y * v2.z, fez * bh + fey * bh)) return false;
  if (!sat_axis(-e1.z * v0.x + e1.x * v0.z, -e1.z * v2.x + e1.x * v2.z, fez * bh + fex * bh)) return false;
  // Z0 over v0,v1
  if (!sat_axis(e1.y * v0.x - e1.x * v0.y, e1.y * v1.x - e1.x * v1.y, fey * bh + fex * bh)) return false;
  fex = fabsf(e2.x);
  fey = fabsf(e2.y);
  fez = fabsf(e2.z);
  // X2 over v0,v1
  if (!sat_axis(e2.z * v0.y - e2.y * v0.z, e2.z * v1.y - e2.y * v1.z, fez * bh + fey * bh)) return false;
  // Y1 over v0,v1
  if (!sat_axis(-e2.z * v0.x + e2.x * v0.z, -e2.z * v1.x + e2.x * v1.z, fez * bh + fex * bh)) return false;
  // Z12 over v1,v2
  if (!sat_axis(e2.y * v1.x - e2.x * v1.y, e2.y * v2.x - e2.x * v2.y, fey * bh + fex * bh)) return false;
  {
    const float mn = smin(smin(v0.x, v1.x), v2.x), mx = smax(smax(v0.x, v1.x), v2.x);
    if (mn > bh || mx < -bh) return false;
  }
  {
    const float mn = smin(smin(v0.y, v1.y), v2.y), mx = smax(smax(v0.y, v1.y), v2.y);
    if (mn > bh || mx < -bh) return false;
  }
  {
    const float mn = smin(smin(v0.z, v1.z), v2.z), mx = smax(smax(v0.z, v1.z), v2.z);
    if (mn > bh || mx < -bh) return false;
  }
  // plane test with n = cross(e0, e1) (glm order) and dot = (x + y) + z
  const F3 n = {e0.y * e1.z - e1.y * e0.z, e0.z * e1.x - e1.z * e0.x, e0.x * e1.y - e1.x * e0.y};
  F3 vmin, vmax;
  if (n.x > 0.0f) { vmin.x = -bh - v0.x; vmax.x = bh - v0.x; } else { vmin.x = bh - v0.x; vmax.x = -bh - v0.x; }
  if (n.y > 0.0f) { vmin.y = -bh - v0.y; vmax.y = bh - v0.y; } else { vmin.y = bh - v0.y; vmax.y = -bh - v0.y; }
  if (n.z > 0.0f) { vmin.z = -bh - v0.z; vmax.z = bh - v0.z; } else { vmin.z = bh - v0.z; vmax.z = -bh - v0.z; }
  {
    const float tx = n.x * vmin.x, ty = n.y * vmin.y, tz = n.z * vmin.z;
    if ((tx + ty) + tz > 0.0f) return false;
  }
  const float tx = n.x * vmax.x, ty = n.y * vmax.y, tz = n.z * vmax.z;
  return (tx + ty) + tz >= 0.0f;
}

// ---------------------------------------------------------------------------------------------
// closest point distance
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Squared distance from q to the box [lo, hi] (lo <= hi).  Per axis the gap is lo - q below the
// box, q - hi above it and 0 inside: max(lo - q, q - hi, 0) (one v_max3_f32).  The squares are
// summed as fma(z, z, fma(y, y, x * x)), the same form as the triangle distance's |r|^2, so a
// triangle vertex that is a box corner gives exactly its box's value.
__device__ __forceinline__ float sq3(float x, float y, float z) { return fmaf(z, z, fmaf(y, y, x * x)); }

__device__ __forceinline__ float box_d2(F3 q, const float *lo, const float *hi) {
  const float dx = fmaxf(fmaxf(lo[0] - q.x, q.x - hi[0]), 0.0f);
  const float dy = fmaxf(fmaxf(lo[1] - q.y, q.y - hi[1]), 0.0f);
  const float dz = fmaxf(fmaxf(lo[2] - q.z, q.z - hi[2]), 0.0f);
  return sq3(dx, dy, dz);
}

__device__ __forceinline__ double dot3d(double ax, double ay, double az, double bx, double by, double bz) {
  return ax * bx + ay * by + az * bz;
}

__device__ __forceinline__ double seg_d2(double ax, double ay, double az, double bx, double by, double bz) {
  const double ex = bx - ax, ey = by - ay, ez = bz - az;
  const double l = dot3d(ex, ey, ez, ex, ey, ez);
  double t = 0.0;
  if (l > 0.0) {
    t = -dot3d(ax, ay, az, ex, ey, ez) / l;
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
  }
  const double rx = ax + t * ex, ry = ay + t * ey, rz = az + t * ez;
  return dot3d(rx, ry, rz, rx, ry, rz);
}

// FP64 squared distance (degenerate-safe): min over the three edges and the interior projection
__device__ __noinline__ double tri_d2_f64(F3 q, F3 p0, F3 p1, F3 p2) {
  const double ax = (double)p0.x - q.x, ay = (double)p0.y - q.y, az = (double)p0.z - q.z;
  const double bx = (double)p1.x - q.x, by = (double)p1.y - q.y, bz = (double)p1.z - q.z;
  const double cx = (double)p2.x - q.x, cy = (double)p2.y - q.y, cz = (double)p2.z - q.z;
  double best = seg_d2(ax, ay, az, bx, by, bz);
  best = fmin(best, seg_d2(bx, by, bz, cx, cy, cz));
  best = fmin(best, seg_d2(cx, cy, cz, ax, ay, az));
  const double abx = bx - ax, aby = by - ay, abz = bz - az;
  const double acx = cx - ax, acy = cy - ay, acz = cz - az;
  const double nx = aby * acz - abz * acy, ny = abz * acx - abx * acz, nz = abx * acy - aby * acx;
  const double nn = dot3d(nx, ny, nz, nx, ny, nz);
  if (nn > 0.0) {
    const double s = dot3d(nx, ny, nz, ax, ay, az) / nn;
    const double px = s * nx, py = s * ny, pz = s * nz;
    const double pax = ax - px, pay = ay - py, paz = az - pz;
    const double pbx = bx - px, pby = by - py, pbz = bz - pz;
    const double pcx = cx - px, pcy = cy - py, pcz = cz - pz;
    const double u = dot3d(pay * pbz - paz * pby, paz * pbx - pax * pbz, pax * pby - pay * pbx, nx, ny, nz);
    const double v = dot3d(pby * pcz - pbz * pcy, pbz * pcx - pbx * pcz, pbx * pcy - pby * pcx, nx, ny, nz);
    const double w = dot3d(pcy * paz - pcz * pay, pcz * pax - pcx * paz, pcx * pay - pcy * pax, nx, ny, nz);
    if (u >= 0.0 && v >= 0.0 && w >= 0.0) {
      const double s2 = dot3d(nx, ny, nz, ax, ay, az);
      best = fmin(best, s2 * s2 / nn);
    }
  }
  return best;
}

// Distance record of one triangle (144 bytes), built once per mesh from the three vertices:
//   e_i = p_{i+1} - p_i                       edge vectors (p_3 = p_0)
//   n = cross(e_0, p_2 - p_0)                 plane normal (unnormalised), inv_nn = 1 / |n|^2
//   nie_i = -1 / |e_i|^2                      segment parameter scale
//   m_i = cross(n, e_i)                       in-plane normal of edge i, pointing into the triangle
// with cross(a, b)_x = fma(a_y, b_z, -(a_z * b_y)) (and cyclic) and dot(a, b) = fma(a_z, b_z,
// fma(a_y, b_y, a_x * b_x)).  The record is read at wave-uniform addresses (scalar loads).
struct alignas(16) TriRec {
  float p[3][3];
  float e[3][3];
  float n[3], inv_nn;
  float nie[3];
  float m[3][3];
  uint32_t flags;
};
static_assert(sizeof(TriRec) == 144, "distance record is 9 x 16 bytes");

__device__ __forceinline__ float fdot(F3 a, F3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
__device__ __forceinline__ F3 fcross(F3 a, F3 b) {
  return {fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x))};
}

__device__ __forceinline__ void make_trirec(F3 p0, F3 p1, F3 p2, uint32_t flags, TriRec &r) {
  const F3 P[3] = {p0, p1, p2};
  F3 E[3];
  for (int i = 0; i < 3; ++i) {
    const F3 a = P[i], b = P[i == 2 ? 0 : i + 1];
    E[i] = {b.x - a.x, b.y - a.y, b.z - a.z};
  }
  const F3 ac = {p2.x - p0.x, p2.y - p0.y, p2.z - p0.z};
  const F3 n = fcross(E[0], ac);
  for (int i = 0; i < 3; ++i) {
    r.p[i][0] = P[i].x; r.p[i][1] = P[i].y; r.p[i][2] = P[i].z;
    r.e[i][0] = E[i].x; r.e[i][1] = E[i].y; r.e[i][2] = E[i].z;
    r.nie[i] = -1.0f / fdot(E[i], E[i]);
    const F3 m = fcross(n, E[i]);
    r.m[i][0] = m.x; r.m[i][1] = m.y; r.m[i][2] = m.z;
  }
  r.n[0] = n.x; r.n[1] = n.y; r.n[2] = n.z;
  r.inv_nn = 1.0f / fdot(n, n);
  r.flags = flags;
}

// FP32 squared distance from q to a non-sliver triangle, branch free.  With a_i = p_i - q:
//   segment i: t = clamp(dot(a_i, e_i) * nie_i, 0, 1), r = a_i + t e_i (fma), d_i = |r|^2
//   face:      inside = max_i dot(m_i, a_i) <= 0, d_f = (dot(n, a_0)^2) * inv_nn
//   result     min(d_0, d_1, d_2, inside ? d_f : inf)
// |r|^2 = fma(r_z, r_z, fma(r_y, r_y, r_x * r_x)), the form of the box distance.
__device__ __forceinline__ float tri_d2_f32(F3 q, const TriRec &t) {
  F3 a[3];
  for (int i = 0; i < 3; ++i) a[i] = {t.p[i][0] - q.x, t.p[i][1] - q.y, t.p[i][2] - q.z};
  float d = INFINITY;
  for (int i = 0; i < 3; ++i) {
    const F3 e = {t.e[i][0], t.e[i][1], t.e[i][2]};
    const float u = fminf(fmaxf(fdot(a[i], e) * t.nie[i], 0.0f), 1.0f);
    const F3 r = {fmaf(u, e.x, a[i].x), fmaf(u, e.y, a[i].y), fmaf(u, e.z, a[i].z)};
    d = fminf(d, sq3(r.x, r.y, r.z));
  }
  const float k0 = fdot({t.m[0][0], t.m[0][1], t.m[0][2]}, a[0]);
  const float k1 = fdot({t.m[1][0], t.m[1][1], t.m[1][2]}, a[1]);
  const float k2 = fdot({t.m[2][0], t.m[2][1], t.m[2][2]}, a[2]);
  const float s = fdot({t.n[0], t.n[1], t.n[2]}, a[0]);
  const float df = (s * s) * t.inv_nn;
  return fmaxf(fmaxf(k0, k1), k2) <= 0.0f ? fminf(d, df) : d;
}

// SLIVERS = false is only instantiated for meshes without sliver triangles (the host counts them
// at build time), so that kernel carries no FP64 code.
template <bool SLIVERS>
__device__ __forceinline__ float tri_d2(F3 q, const TriRec &t) {
  if (SLIVERS && (t.flags & kSliver))
    return (float)tri_d2_f64(q, {t.p[0][0], t.p[0][1], t.p[0][2]}, {t.p[1][0], t.p[1][1], t.p[1][2]},
                             {t.p[2][0], t.p[2][1], t.p[2][2]});
  return tri_d2_f32(q, t);
}

// ---------------------------------------------------------------------------------------------
// watertight crossings (Woop, Benthin, Wald 2013) with exact FP64 edge signs and a half-open tie
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ int tie_pos(float dx, float dy) { return (dy > 0.0f || (dy == 0.0f && dx > 0.0f)) ? 1 : -1; }

// sign of px*qy - py*qx (exact: products of floats are exact in FP64)
__device__ __forceinline__ int edge_sign(float px, float py, float qx, float qy, double &e) {
  e = (double)px * (double)qy - (double)py * (double)qx;
  return e > 0.0 ? 1 : (e < 0.0 ? -1 : 0);
}

// 2D crossing test of a +z line at (X,Y) with triangle p0,p1,p2.  Returns 0 (miss) or the winding
// contribution (+1 where the triangle normal has +z component, i.e. the line leaves the solid);
// *zhit receives the z of the crossing.
__device__ __forceinline__ int column_cross(float X, float Y, F3 p0, F3 p1, F3 p2, float *zhit) {
  const float ax = p0.x - X, ay = p0.y - Y;
  const float bx = p1.x - X, by = p1.y - Y;
  const float cx = p2.x - X, cy = p2.y - Y;
  double U, V, W;
  int su = edge_sign(cx, cy, bx, by, U);  // edge B->C
  int sv = edge_sign(ax, ay, cx, cy, V);  // edge C->A
  int sw = edge_sign(bx, by, ax, ay, W);  // edge A->B
  if (su == 0) su = tie_pos(cx - bx, cy - by);
  if (sv == 0) sv = tie_pos(ax - cx, ay - cy);
  if (sw == 0) sw = tie_pos(bx - ax, by - ay);
  if (su != sv || sv != sw) return 0;
  const double det = U + V + W;
  if (det == 0.0) return 0;
  *zhit = (float)((U * (double)p0.z + V * (double)p1.z + W * (double)p2.z) / det);
  return -su;
}

// general-direction watertight ray test for the parity tie-break: 1 if hit at t > 0
struct Shear {
  int kx, ky, kz;
  float sx, sy, sz;
};

__device__ __forceinline__ Shear make_shear(float dx, float dy, float dz) {
  const float d[3] = {dx, dy, dz};
  int kz = 0;
  if (fabsf(d[1]) > fabsf(d[kz])) kz = 1;
  if (fabsf(d[2]) > fabsf(d[kz])) kz = 2;
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  if (d[kz] < 0.0f) {
    const int t = kx;
    kx = ky;
    ky = t;
  }
  Shear s;
  s.kx = kx;
  s.ky = ky;
  s.kz = kz;
  s.sx = d[kx] / d[kz];
  s.sy = d[ky] / d[kz];
  s.sz = 1.0f / d[kz];
  return s;
}

__device__ __forceinline__ float comp(F3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

__device__ __forceinline__ bool ray_hit(F3 o, const Shear &r, F3 p0, F3 p1, F3 p2) {
  const F3 A = {p0.x - o.x, p0.y - o.y, p0.z - o.z};
  const F3 B = {p1.x - o.x, p1.y - o.y, p1.z - o.z};
  const F3 C = {p2.x - o.x, p2.y - o.y, p2.z - o.z};
  const float ax = comp(A, r.kx) - r.sx * comp(A, r.kz), ay = comp(A, r.ky) - r.sy * comp(A, r.kz);
  const float bx = comp(B, r.kx) - r.sx * comp(B, r.kz), by = comp(B, r.ky) - r.sy * comp(B, r.kz);
  const float cx = comp(C, r.kx) - r.sx * comp(C, r.kz), cy = comp(C, r.ky) - r.sy * comp(C, r.kz);
  double U, V, W;
  int su = edge_sign(cx, cy, bx, by, U);
  int sv = edge_sign(ax, ay, cx, cy, V);
  int sw = edge_sign(bx, by, ax, ay, W);
  if (su == 0) su = tie_pos(cx - bx, cy - by);
  if (sv == 0) sv = tie_pos(ax - cx, ay - cy);
  if (sw == 0) sw = tie_pos(bx - ax, by - ay);
  if (su != sv || sv != sw) return false;
  const double det = U + V + W;
  if (det == 0.0) return false;
  const double az = (double)r.sz * comp(A, r.kz), bz = (double)r.sz * comp(B, r.kz), cz = (double)r.sz * comp(C, r.kz);
  const double t = (U * az + V * bz + W * cz) / det;
  return t > 0.0;
}

// the five parity directions: +z first, then four fixed generic unit vectors
__device__ __constant__ const float kParityDirs[15] = {
    0.0f,          0.0f,          1.0f,          0.5773502588f,  0.5773502588f,
    -0.5773502588f, -0.6666666865f, 0.3333333433f, 0.6666666865f,  0.2672612369f,
    -0.8017837405f, -0.5345224738f, -0.4850712717f, -0.7276068926f, 0.4850712717f};

// ordered-int encoding of floats for atomic min/max
__device__ __forceinline__ uint32_t f2o(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ __forceinline__ float o2f(uint32_t o) {
  const uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
  union {
    uint32_t u;
    float f;
  } c;
  c.u = u;
  return c.f;
}

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
  v &= 0x3ffu;
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

}  // namespace ib
