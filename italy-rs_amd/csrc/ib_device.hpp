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
// leaf code ~L: first triangle L & kBlockMask, count (L >> kBlockShift) + 1; blocks of more than one
// triangle are built only for meshes of at most 2^26 triangles, so that L stays non-negative
constexpr int kBlockShift = 26;
constexpr int kBlockMask = (1 << kBlockShift) - 1;
constexpr int kMaxLeafBlock = 32;

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
// code < 0).  Entries past `count` are unused.  Bit e of obb_mask: entry e is an internal node of at most
// kNodeObbMax triangles, whose oriented box is in the node-box array at its index.
struct alignas(16) Node4 {
  float lo[4][3];
  float hi[4][3];
  int32_t child[4];
  int32_t count, obb_mask, pad[2];
};
// internal nodes of at most this many triangles (and more than a leaf block) get an oriented box too
constexpr int kNodeObbMax = 1024;
static_assert(sizeof(Node4) == 128, "4-wide record is two 64-byte lines");

// Oriented box of a leaf block of two or more triangles: centre c, axes x, y, z (the FP32 roundings of an
// orthonormal FP64 frame whose z is the block's area-weighted normal) and half extents e (FP64 maxima of
// |axis . (v - c)| over the block's vertices, rounded up): every triangle of the block lies in
// { p : |axis_k . (p - c)| <= e_k }.  On a smooth surface a block is nearly planar, so the box is as thin
// as the patch's curvature, where its axis-aligned box is as thick as the patch is wide.
struct alignas(16) Obb {
  float c[3], ex;
  float x[3], ey;
  float y[3], ez;
  float z[3], pad;
};
static_assert(sizeof(Obb) == 64, "oriented box record is one 64-byte line");

// Node of the hole-cap hierarchy.  Leaf (left < 0 and right < 0): fans [-1 - left, -1 - left - right)
// of the sorted cap soup.  Inner node: left / right are child codes, a node index (>= 0) or a single fan
// ~index (< 0), at most one of them a fan.  (c, r) bounds every vertex below; lo / hi is the box in x and
// y; area = sum of |fan area|; dip = sum of area vectors; open2: a point closer to c than sqrt(open2) sees
// a far-field bound above kCapTauHi (cap_far), so the walks open the node without evaluating it.
struct alignas(16) CapNode {
  float c[3], r;
  float dip[3], area;
  float lo[2], open2;
  int32_t left;
  float hi[2], pad;
  int32_t right;
};
static_assert(sizeof(CapNode) == 64, "cap node record is one 64-byte line");
// leaves hold at most kCapLeaf fans; the tree's depth is at most 30 + ceil(log2 fans) (Karras tree over
// 30-bit codes split by sorted position), so a traversal that keeps one pending sibling per level fits 64
constexpr int kCapLeaf = 4;
constexpr int kCapStack = 64;
// a cap node whose far-field error bound (winding units) is at most kCapTau is not evaluated: its first-order
// term is added and its bound counted
constexpr float kCapTau = 2e-3f;
// ... and a node whose bound is at most kCapTauHi is not evaluated either while the bounds of the nodes not
// evaluated so far stay within kCapBudget (the sign needs |w - 1/2| >= 0.05 + bound; a voxel whose bound
// leaves it closer goes to the FP64 recheck, which the cap tree's cut makes cheap)
constexpr float kCapTauHi = 1e-2f;
constexpr float kCapBudget = 0.3f;
// the FP64 recheck sums the caps over a cut of their tree first: a node of the cut whose far-field error
// bound is at most kCapTau64 adds its first-order term and its bound; the voxel is decided there when |w|
// stays on one side of 1/2 and of the tie band's edges by more than the total bound, and every fan is
// summed otherwise
constexpr double kCapTau64 = 1e-4;
// the cut of the cap tree that the FP64 recheck tests: the nodes (or fans) kCapCutBits levels below the root
constexpr int kCapCutBits = 10;
constexpr int kCapCut = 1 << kCapCutBits;

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

// Distance record of one triangle (96 bytes), built once per mesh from the three vertices: an
// orthonormal frame at p0 in which the triangle is planar, with its corners p1 = (len, 0) and
// p2 = (c0, c1), c1 > 0:
//   e0 = p1 - p0, ac = p2 - p0, n = cross(e0, ac), len = sqrt(dot(e0, e0)), inv_len = 1 / len
//   X = e0 * inv_len, Z = n * (1 / sqrt(dot(n, n))), Y = cross(Z, X)
//   c = (dot(X, ac), dot(Y, ac)), e1 = (c0 - len, c1) = p2 - p1
//   inv_e1 = 1 / |e1|^2, nie2 = -1 / |c|^2
// with cross(a, b)_x = fma(a_y, b_z, -(a_z * b_y)) (and cyclic), dot(a, b) = fma(a_z, b_z,
// fma(a_y, b_y, a_x * b_x)) and |v|^2 = fma(v_1, v_1, v_0 * v_0) in 2D.  A sliver (flag kSliver)
// stores its three vertices in o, x, y instead, for the FP64 path.  The record is read at
// wave-uniform addresses (scalar loads).
struct alignas(16) TriRec {
  float o[3], x[3], y[3], z[3];
  float len, inv_len;
  float e1[2], inv_e1;
  float c[2], nie2;
  uint32_t flags;
  float pad[3];
};
static_assert(sizeof(TriRec) == 96, "distance record is 6 x 16 bytes");

__device__ __forceinline__ float fdot(F3 a, F3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
__device__ __forceinline__ F3 fcross(F3 a, F3 b) {
  return {fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x))};
}

__device__ __forceinline__ void make_trirec(F3 p0, F3 p1, F3 p2, uint32_t flags, TriRec &r) {
  r = TriRec{};
  r.flags = flags;
  if (flags & kSliver) {
    const F3 P[3] = {p0, p1, p2};
    float *dst[3] = {r.o, r.x, r.y};
    for (int i = 0; i < 3; ++i) dst[i][0] = P[i].x, dst[i][1] = P[i].y, dst[i][2] = P[i].z;
    return;
  }
  const F3 e0 = {p1.x - p0.x, p1.y - p0.y, p1.z - p0.z};
  const F3 ac = {p2.x - p0.x, p2.y - p0.y, p2.z - p0.z};
  const F3 n = fcross(e0, ac);
  r.len = sqrtf(fdot(e0, e0));
  r.inv_len = 1.0f / r.len;
  const float inv_n = 1.0f / sqrtf(fdot(n, n));
  const F3 X = {e0.x * r.inv_len, e0.y * r.inv_len, e0.z * r.inv_len};
  const F3 Z = {n.x * inv_n, n.y * inv_n, n.z * inv_n};
  const F3 Y = fcross(Z, X);
  r.o[0] = p0.x, r.o[1] = p0.y, r.o[2] = p0.z;
  r.x[0] = X.x, r.x[1] = X.y, r.x[2] = X.z;
  r.y[0] = Y.x, r.y[1] = Y.y, r.y[2] = Y.z;
  r.z[0] = Z.x, r.z[1] = Z.y, r.z[2] = Z.z;
  r.c[0] = fdot(X, ac);
  r.c[1] = fdot(Y, ac);
  r.e1[0] = r.c[0] - r.len;
  r.e1[1] = r.c[1];
  r.inv_e1 = 1.0f / fmaf(r.e1[1], r.e1[1], r.e1[0] * r.e1[0]);
  r.nie2 = -1.0f / fmaf(r.c[1], r.c[1], r.c[0] * r.c[0]);
}

__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

// Minimum of two squared distances (non-negative floats or +inf) through their bit patterns, which
// order like unsigned integers: the same value as fminf, without the canonicalisation fminf needs
// for an operand carried around a loop (one instruction per triangle evaluated).
__device__ __forceinline__ float min_nonneg(float a, float b) {
  return __uint_as_float(min(__float_as_uint(a), __float_as_uint(b)));
}

// FP32 squared distance from q to a non-sliver triangle, branch free.  With d = q - o and
// (u, v, w) = (dot(X, d), dot(Y, d), dot(Z, d)):
//   segment 0 (p0 p1): t = clamp(u * inv_len), r = fma(-t, len, u), d0 = fma(v, v, r * r)
//   segment 1 (p1 p2): a = (u - len, v), t = clamp(dot(a, e1) * inv_e1), r = fma(-t, e1, a), d1 = |r|^2
//   segment 2 (p2 p0): a = (u - c0, v - c1), t = clamp(dot(a, c) * nie2), r = fma(t, c, a), d2 = |r|^2
//   inside = max(-v, -k1, k2) <= 0, k1 = fma(e1_0, v, -(e1_1 * a1_0)), k2 = fma(c0, a2_1, -(c1 * a2_0))
//   result fma(w, w, inside ? 0 : min(d0, d1, d2))
// (2D dots as fma(a_1, b_1, a_0 * b_0)).
__device__ __forceinline__ float tri_d2_f32(F3 q, const TriRec &t) {
  const F3 d = {q.x - t.o[0], q.y - t.o[1], q.z - t.o[2]};
  const float u = fdot({t.x[0], t.x[1], t.x[2]}, d);
  const float v = fdot({t.y[0], t.y[1], t.y[2]}, d);
  const float w = fdot({t.z[0], t.z[1], t.z[2]}, d);
  const float t0 = clamp01(u * t.inv_len);
  const float r0 = fmaf(-t0, t.len, u);
  const float d0 = fmaf(v, v, r0 * r0);
  const float a1 = u - t.len;
  const float t1 = clamp01(fmaf(v, t.e1[1], a1 * t.e1[0]) * t.inv_e1);
  const float s1x = fmaf(-t1, t.e1[0], a1), s1y = fmaf(-t1, t.e1[1], v);
  const float d1 = fmaf(s1y, s1y, s1x * s1x);
  const float a2x = u - t.c[0], a2y = v - t.c[1];
  const float t2 = clamp01(fmaf(a2y, t.c[1], a2x * t.c[0]) * t.nie2);
  const float s2x = fmaf(t2, t.c[0], a2x), s2y = fmaf(t2, t.c[1], a2y);
  const float d2 = fmaf(s2y, s2y, s2x * s2x);
  const float k1 = fmaf(t.e1[0], v, -(t.e1[1] * a1));
  const float k2 = fmaf(t.c[0], a2y, -(t.c[1] * a2x));
  const bool inside = fmaxf(fmaxf(-v, -k1), k2) <= 0.0f;
  const float dmin = fminf(fminf(d0, d1), d2);
  return fmaf(w, w, inside ? 0.0f : dmin);
}

// Squared distance from q to the oriented box b, evaluated in FP32 in the box's frame.  Its error, from the
// rounding of q - c and of the dots and from the frame's departure from orthonormality, is below
// 2^-19 (Mq + Mp) for coordinates bounded by Mq (query) and Mp (vertices), a quarter of the pruning slack;
// the traversals therefore compare it against prune_bound(best, 1.25 * slack).
__device__ __forceinline__ float obb_d2(F3 q, const Obb &b);

// Pruning with FP32 distances.  |sqrt(tri_d2_f32(q, t)) - dist(q, t)| stays below
// 2^-17 (Mq + 3 Mp) when every coordinate of the query is at most Mq and every vertex coordinate
// at most Mp in magnitude (the worst case measured over 4e6 triangles and queries across six
// decades of scale and size is 4 u (Mq + 3 Mp), u = 2^-24).  A traversal enters a box when its
// squared distance is at most prune_bound(best, slack): every triangle whose FP32 distance could
// still beat `best` is then reached whatever the visiting order, so each traversal returns the
// minimum over all triangles and the query kernels, the brick shapes and the oracle agree exactly.
__device__ __forceinline__ float prune_bound(float best, float slack) {
  const float r = __builtin_amdgcn_sqrtf(best) + slack;  // hardware root, within 1 ulp
  return (r * r) * 1.0000077f;  // covers that root, these roundings and the box distance's
}
__device__ __forceinline__ float finite_abs(float x) { return fabsf(x) < INFINITY ? fabsf(x) : 0.0f; }
// slack for the queries of grid g over a mesh whose boxes `lo`/`hi` (n of them, empty boxes
// allowed) cover every vertex
__device__ __forceinline__ float dist_slack(float ox, float oy, float oz, float h, int nx, int ny, int nz,
                                            const float (*lo)[3], const float (*hi)[3], int n) {
  const float mq = fmaxf(fmaxf(fmaxf(fabsf(ox), fabsf(ox + (float)nx * h)), fmaxf(fabsf(oy), fabsf(oy + (float)ny * h))),
                         fmaxf(fabsf(oz), fabsf(oz + (float)nz * h)));
  float mp = 0.0f;
  for (int e = 0; e < n; ++e)
    for (int k = 0; k < 3; ++k) mp = fmaxf(mp, fmaxf(finite_abs(lo[e][k]), finite_abs(hi[e][k])));
  return (mq + 3.0f * mp) * 7.62939453125e-6f;  // 2^-17
}

__device__ __forceinline__ float obb_d2(F3 q, const Obb &b) {
  const F3 d = {q.x - b.c[0], q.y - b.c[1], q.z - b.c[2]};
  const float u = fdot({b.x[0], b.x[1], b.x[2]}, d);
  const float v = fdot({b.y[0], b.y[1], b.y[2]}, d);
  const float w = fdot({b.z[0], b.z[1], b.z[2]}, d);
  return sq3(fmaxf(fabsf(u) - b.ex, 0.0f), fmaxf(fabsf(v) - b.ey, 0.0f), fmaxf(fabsf(w) - b.ez, 0.0f));
}

// SLIVERS = false is only instantiated for meshes without sliver triangles (the host counts them
// at build time), so that kernel carries no FP64 code.
template <bool SLIVERS>
__device__ __forceinline__ float tri_d2(F3 q, const TriRec &t) {
  if (SLIVERS && (t.flags & kSliver))
    return (float)tri_d2_f64(q, {t.o[0], t.o[1], t.o[2]}, {t.x[0], t.x[1], t.x[2]}, {t.y[0], t.y[1], t.y[2]});
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
