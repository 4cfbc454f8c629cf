// Bake kernels for gfx950: mesh ingest (weld + closedness), LBVH build (Karras 2012), surface
// occupancy, +z column crossings and the per-voxel closest-triangle + winding-number query.
#include <cstring>

#include "ib_kernels.hpp"

#include <rocprim/rocprim.hpp>

namespace ib {

// ============================================================================================
// small utilities
// ============================================================================================
__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// several fills in one launch (the ranges must not overlap)
__global__ void k_fill_multi(FillList l) {
  for (int k = 0; k < l.count; ++k) {
    const FillRange f = l.r[k];
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < f.n; i += (uint64_t)gridDim.x * blockDim.x)
      f.p[i] = f.v;
  }
}

__device__ __forceinline__ uint32_t canon(float f) { return f == 0.0f ? 0u : __float_as_uint(f); }

__device__ __forceinline__ uint32_t hash3(uint32_t a, uint32_t b, uint32_t c) {
  uint64_t x = ((uint64_t)a << 32 | b) ^ ((uint64_t)c * 0x9E3779B97F4A7C15ULL);
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return (uint32_t)x;
}

__device__ __forceinline__ uint32_t hash64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return (uint32_t)x;
}

// ============================================================================================
// ingest: finite check + centroid bounds (block reduction, then 6 ordered-int atomics)
// ============================================================================================
__global__ __launch_bounds__(256) void k_prep(const float *__restrict__ xyz, int n, Scalars *sc) {
  __shared__ uint32_t red[6][256];
  uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
  bool finite = true;
  // a capped grid strides over the triangles, so that few blocks meet at the six global atomics
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const float *t = xyz + 9 * (size_t)i;
    for (int k = 0; k < 9; ++k) finite &= isfinite(t[k]);
    for (int k = 0; k < 3; ++k) {
      const uint32_t c = f2o(((t[k] + t[3 + k]) + t[6 + k]) / 3.0f);
      mn[k] = min(mn[k], c);
      mx[k] = max(mx[k], c);
    }
  }
  if (!finite) atomicOr(&sc->flags, kFlagNonFinite);
  for (int k = 0; k < 3; ++k) {
    red[k][threadIdx.x] = mn[k];
    red[3 + k][threadIdx.x] = mx[k];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = min(red[k][threadIdx.x], red[k][threadIdx.x + s]);
        red[3 + k][threadIdx.x] = max(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) {
      atomicMin(&sc->cmin[k], red[k][0]);
      atomicMax(&sc->cmax[k], red[3 + k][0]);
    }
}

// 30-bit Morton code of the centroid of triangle t (9 floats) over the centroid bounds (ordered-int encoded)
__device__ __forceinline__ uint32_t centroid_code(const float *t, const uint32_t *cmin, const uint32_t *cmax) {
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) {
    const float lo = o2f(cmin[k]), hi = o2f(cmax[k]);
    const float ext = hi - lo;
    const float c = ((t[k] + t[3 + k]) + t[6 + k]) / 3.0f;
    const float u = ext > 0.0f ? (c - lo) / ext : 0.5f;
    float s = u * 1024.0f;
    s = s < 0.0f ? 0.0f : (s > 1023.0f ? 1023.0f : s);
    q[k] = (uint32_t)s;
  }
  return (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
}

__global__ __launch_bounds__(256) void k_morton(const float *__restrict__ xyz, int n, const Scalars *sc,
                                                uint32_t *codes, int32_t *idx) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float *t = xyz + 9 * (size_t)i;
  codes[i] = centroid_code(t, sc->cmin, sc->cmax);
  idx[i] = i;
}

// ============================================================================================
// LBVH topology (Karras 2012, "Maximizing parallelism in the construction of BVHs, octrees and
// k-d trees"): keys are (code << 32 | sorted position) so equal codes still split uniquely.
// ============================================================================================
__device__ __forceinline__ int kdelta(const uint32_t *codes, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t a = ((uint64_t)codes[i] << 32) | (uint32_t)i;
  const uint64_t b = ((uint64_t)codes[j] << 32) | (uint32_t)j;
  return __clzll((long long)(a ^ b));
}

// children (an internal node index, or ~leaf) and covered range [lo, hi] of internal node i
__device__ __forceinline__ void karras_node(const uint32_t *__restrict__ codes, int n, int i, int &lc, int &rc,
                                            int &lo, int &hi) {
  const int d = (kdelta(codes, n, i, i + 1) - kdelta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = kdelta(codes, n, i, i - d);
  int lmax = 2;
  while (kdelta(codes, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (kdelta(codes, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = kdelta(codes, n, i, j);
  int s = 0, t = l;
  do {
    t = (t + 1) >> 1;
    if (kdelta(codes, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int gamma = i + s * d + (d < 0 ? d : 0);
  lo = min(i, j);
  hi = max(i, j);
  lc = (lo == gamma) ? ~gamma : gamma;
  rc = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
}

__global__ __launch_bounds__(256) void k_karras(const uint32_t *__restrict__ codes, int n, Node *nodes,
                                                int32_t *leaf_parent, int2 *ranges) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n - 1) return;
  int lc, rc, lo, hi;
  karras_node(codes, n, i, lc, rc, lo, hi);
  nodes[i].c0 = lc;
  nodes[i].c1 = rc;
  nodes[i].pad = 0;
  ranges[i] = make_int2(lo, hi);
  if (lc >= 0) nodes[lc].parent = i; else leaf_parent[~lc] = i;
  if (rc >= 0) nodes[rc].parent = i; else leaf_parent[~rc] = i;
  if (i == 0) nodes[0].parent = -1;
}

// sorted triangle records (3 x float4) with the sliver flag, and their distance records
__global__ __launch_bounds__(256) void k_gather(const float *__restrict__ xyz, const int32_t *__restrict__ order,
                                                int n, float4 *tris, TriRec *recs, Scalars *sc) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int o = order[i];
  const float *t = xyz + 9 * (size_t)o;
  const float abx = t[3] - t[0], aby = t[4] - t[1], abz = t[5] - t[2];
  const float acx = t[6] - t[0], acy = t[7] - t[1], acz = t[8] - t[2];
  const float nx = aby * acz - abz * acy, ny = abz * acx - abx * acz, nz = abx * acy - aby * acx;
  const float nn = nx * nx + ny * ny + nz * nz;
  const float l1 = abx * abx + aby * aby + abz * abz, l2 = acx * acx + acy * acy + acz * acz;
  const uint32_t flags = (nn > 1e-4f * (l1 * l2)) ? 0u : kSliver;
  if (flags) atomicAdd(&sc->slivers, 1u);
  tris[3 * i + 0] = make_float4(t[0], t[1], t[2], __uint_as_float(flags));
  tris[3 * i + 1] = make_float4(t[3], t[4], t[5], __int_as_float(o));
  tris[3 * i + 2] = make_float4(t[6], t[7], t[8], 0.0f);
  TriRec r;
  make_trirec({t[0], t[1], t[2]}, {t[3], t[4], t[5]}, {t[6], t[7], t[8]}, flags, r);
  recs[i] = r;
}

// Refit from the Karras ranges instead of bottom-up atomics: every internal node covers a contiguous
// range of sorted leaves, so each child slot's box is the union of the leaf boxes of the child's range.
// Min and max are exact, so any order of the unions gives the same box (up to the sign of a zero bound,
// which no distance or box test can see).  Pass 1 (one workgroup per 256-leaf block) builds a sparse
// table of the block's leaf boxes in LDS and answers the nodes whose range lies inside the block with
// two table reads per slot; it writes the leaf and block boxes and lists the other nodes.  Pass 2 gives
// each listed node one wave, which unions the partial blocks at the ends of each child's range from the
// leaf boxes and the whole blocks in between from the block boxes.  No workgroup waits for another.
constexpr int kRefitLevels = 9;  // windows of 1 .. 256 leaves

__device__ __forceinline__ void box_union(float lo[3], float hi[3], const float *b) {
  for (int k = 0; k < 3; ++k) {
    lo[k] = smin(lo[k], b[k]);
    hi[k] = smax(hi[k], b[3 + k]);
  }
}

__global__ __launch_bounds__(256) void k_refit_blocks(const float4 *__restrict__ tris, int n, Node *nodes,
                                                      const int2 *__restrict__ ranges, float *leafbox,
                                                      float *blockbox, float *root_box, uint32_t *span) {
  __shared__ float st[kRefitLevels][6][256];
  const int t = threadIdx.x, base = blockIdx.x * 256, i = base + t;
  float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};  // neutral past the last leaf
  if (i < n) {
    const float4 a = tris[3 * i], v = tris[3 * i + 1], c = tris[3 * i + 2];
    b[0] = smin(smin(a.x, v.x), c.x), b[1] = smin(smin(a.y, v.y), c.y), b[2] = smin(smin(a.z, v.z), c.z);
    b[3] = smax(smax(a.x, v.x), c.x), b[4] = smax(smax(a.y, v.y), c.y), b[5] = smax(smax(a.z, v.z), c.z);
    for (int k = 0; k < 6; ++k) leafbox[6 * (size_t)i + k] = b[k];
  }
  for (int k = 0; k < 6; ++k) st[0][k][t] = b[k];
  // st[l][.][t] = union of leaves t .. t + 2^l - 1 of the block, valid where t + 2^l <= 256
  for (int l = 1; l < kRefitLevels; ++l) {
    __syncthreads();
    const int h = 1 << (l - 1);
    if (t + 2 * h <= 256) {
      for (int k = 0; k < 3; ++k) st[l][k][t] = smin(st[l - 1][k][t], st[l - 1][k][t + h]);
      for (int k = 3; k < 6; ++k) st[l][k][t] = smax(st[l - 1][k][t], st[l - 1][k][t + h]);
    }
  }
  __syncthreads();
  if (t < 6) blockbox[6 * (size_t)blockIdx.x + t] = st[kRefitLevels - 1][t][0];
  if (i >= n - 1) return;
  const int2 r = ranges[i];
  if ((r.x >> 8) != (r.y >> 8)) {  // spans blocks: pass 2 (the count sits past the n-entry list)
    span[atomicAdd(&span[n], 1u)] = (uint32_t)i;
    return;
  }
  Node *nd = nodes + i;
  float rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int slot = 0; slot < 2; ++slot) {
    const int c = slot ? nd->c1 : nd->c0;
    const int2 cr = c < 0 ? make_int2(~c, ~c) : ranges[c];
    const int x = cr.x - base, y = cr.y - base;  // inside [0, 256): the child's range is inside the node's
    const int l = 31 - __clz(y - x + 1);
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = smin(st[l][k][x], st[l][k][y - (1 << l) + 1]);
      hi[k] = smax(st[l][3 + k][x], st[l][3 + k][y - (1 << l) + 1]);
    }
    for (int k = 0; k < 3; ++k) {
      (slot ? nd->lo1 : nd->lo0)[k] = lo[k];
      (slot ? nd->hi1 : nd->hi0)[k] = hi[k];
      rlo[k] = smin(rlo[k], lo[k]);
      rhi[k] = smax(rhi[k], hi[k]);
    }
  }
  if (i == 0)
    for (int k = 0; k < 3; ++k) root_box[k] = rlo[k], root_box[3 + k] = rhi[k];
}

__global__ __launch_bounds__(256) void k_refit_span(int n, Node *nodes, const int2 *__restrict__ ranges,
                                                    const float *__restrict__ leafbox,
                                                    const float *__restrict__ blockbox, float *root_box,
                                                    const uint32_t *__restrict__ span) {
  const int lane = threadIdx.x & 63;
  const uint32_t count = span[n];
  const uint32_t nwaves = gridDim.x * 4;
  for (uint32_t w = blockIdx.x * 4 + (threadIdx.x >> 6); w < count; w += nwaves) {
    const int i = (int)span[w];
    Node *nd = nodes + i;
    float rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int slot = 0; slot < 2; ++slot) {
      const int c = slot ? nd->c1 : nd->c0;
      const int2 cr = c < 0 ? make_int2(~c, ~c) : ranges[c];
      const int bx = cr.x >> 8, by = cr.y >> 8;
      float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      if (bx == by) {
        for (int j = cr.x + lane; j <= cr.y; j += 64) box_union(lo, hi, leafbox + 6 * (size_t)j);
      } else {
        for (int j = cr.x + lane; j < (bx + 1) * 256; j += 64) box_union(lo, hi, leafbox + 6 * (size_t)j);
        for (int j = bx + 1 + lane; j < by; j += 64) box_union(lo, hi, blockbox + 6 * (size_t)j);
        for (int j = by * 256 + lane; j <= cr.y; j += 64) box_union(lo, hi, leafbox + 6 * (size_t)j);
      }
      for (int off = 32; off >= 1; off >>= 1)
        for (int k = 0; k < 3; ++k) {
          lo[k] = smin(lo[k], __shfl_xor(lo[k], off));
          hi[k] = smax(hi[k], __shfl_xor(hi[k], off));
        }
      if (lane == 0)
        for (int k = 0; k < 3; ++k) {
          (slot ? nd->lo1 : nd->lo0)[k] = lo[k];
          (slot ? nd->hi1 : nd->hi0)[k] = hi[k];
        }
      for (int k = 0; k < 3; ++k) rlo[k] = smin(rlo[k], lo[k]), rhi[k] = smax(rhi[k], hi[k]);
    }
    if (i == 0 && lane == 0)
      for (int k = 0; k < 3; ++k) root_box[k] = rlo[k], root_box[3 + k] = rhi[k];
  }
}

// Collapse: a child slot whose node covers at most `block` leaves (a contiguous range [lo, hi] of
// sorted triangles, from the Karras ranges) becomes the leaf-block code ~(lo | (count - 1) << kBlockShift),
// so the traversals evaluate those triangles from the parent without visiting the nodes below
// (their union box is already in the parent's slot).  Only the immutable ranges are read, so the
// order in which threads rewrite slots does not matter.  Runs after the refit and the depth pass.
__global__ __launch_bounds__(256) void k_collapse_blocks(Node *nodes, const int2 *__restrict__ ranges, int ninternal,
                                                         int block) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ninternal) return;
  for (int k = 0; k < 2; ++k) {
    const int c = k ? nodes[i].c1 : nodes[i].c0;
    if (c < 0) continue;
    const int2 r = ranges[c];
    const int cnt = r.y - r.x + 1;
    if (cnt > block) continue;
    const int code = ~(r.x | ((cnt - 1) << kBlockShift));
    if (k) nodes[i].c1 = code;
    else nodes[i].c0 = code;
  }
}

// Oriented box of every leaf block of two or more triangles (the collapsed child slots of the live binary
// nodes, each block in exactly one slot), written at the block's first triangle.  Frame: z the normalised
// sum of the triangles' area vectors (the world z if that sum vanishes), x the first triangle's longest
// edge with its z part removed (a world axis if that vanishes), y = z x x, all in FP64 and rounded to FP32.
// The centre halves the vertices' span along the rounded axes; the half extents bound |axis . (v - c)| with
// the rounded axes and centre (from the span, with a margin for the FP64 roundings), rounded up to FP32.
__device__ void block_obb(const float4 *__restrict__ tris, int first, int count, Obb &b) {
  double n[3] = {0.0, 0.0, 0.0};
  for (int t = first; t < first + count; ++t) {
    const float4 a = tris[3 * t], p = tris[3 * t + 1], q = tris[3 * t + 2];
    const double e1[3] = {(double)p.x - a.x, (double)p.y - a.y, (double)p.z - a.z};
    const double e2[3] = {(double)q.x - a.x, (double)q.y - a.y, (double)q.z - a.z};
    n[0] += e1[1] * e2[2] - e1[2] * e2[1];
    n[1] += e1[2] * e2[0] - e1[0] * e2[2];
    n[2] += e1[0] * e2[1] - e1[1] * e2[0];
  }
  double z[3] = {0.0, 0.0, 1.0};
  const double nl = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (nl > 0.0 && nl < INFINITY) z[0] = n[0] / nl, z[1] = n[1] / nl, z[2] = n[2] / nl;
  double x[3];
  {
    const float4 a = tris[3 * first], p = tris[3 * first + 1], q = tris[3 * first + 2];
    const double e[3][3] = {{(double)p.x - a.x, (double)p.y - a.y, (double)p.z - a.z},
                            {(double)q.x - p.x, (double)q.y - p.y, (double)q.z - p.z},
                            {(double)a.x - q.x, (double)a.y - q.y, (double)a.z - q.z}};
    int k = 0;
    double best = -1.0;
    for (int i = 0; i < 3; ++i) {
      const double l = e[i][0] * e[i][0] + e[i][1] * e[i][1] + e[i][2] * e[i][2];
      if (l > best) best = l, k = i;
    }
    for (int i = 0; i < 3; ++i) x[i] = e[k][i];
  }
  for (int attempt = 0; attempt < 2; ++attempt) {
    const double dz = x[0] * z[0] + x[1] * z[1] + x[2] * z[2];
    for (int i = 0; i < 3; ++i) x[i] -= dz * z[i];
    const double xl = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    if (xl > 1e-12 && xl < INFINITY) {
      for (int i = 0; i < 3; ++i) x[i] /= xl;
      break;
    }
    x[0] = fabs(z[0]) < 0.5 ? 1.0 : 0.0;  // a world axis not along z
    x[1] = fabs(z[0]) < 0.5 ? 0.0 : 1.0;
    x[2] = 0.0;
  }
  const double y[3] = {z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2], z[0] * x[1] - z[1] * x[0]};
  const float ax[3][3] = {{(float)x[0], (float)x[1], (float)x[2]},
                          {(float)y[0], (float)y[1], (float)y[2]},
                          {(float)z[0], (float)z[1], (float)z[2]}};
  const float4 o = tris[3 * first];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int t = first; t < first + count; ++t)
    for (int v = 0; v < 3; ++v) {
      const float4 p = tris[3 * t + v];
      const double d[3] = {(double)p.x - o.x, (double)p.y - o.y, (double)p.z - o.z};
      for (int k = 0; k < 3; ++k) {
        const double s = ax[k][0] * d[0] + ax[k][1] * d[1] + ax[k][2] * d[2];
        lo[k] = fmin(lo[k], s);
        hi[k] = fmax(hi[k], s);
      }
    }
  double c[3] = {o.x, o.y, o.z};
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < 3; ++i) c[i] += 0.5 * (lo[k] + hi[k]) * ax[k][i];
  const float cf[3] = {(float)c[0], (float)c[1], (float)c[2]};
  // half extents about the rounded centre from the projections' range, as node_obb_lanes derives them
  double e[3], mag = 0.0;
  for (int k = 0; k < 3; ++k) mag += fmax(fabs(lo[k]), fabs(hi[k]));
  for (int k = 0; k < 3; ++k) {
    const double sc = ax[k][0] * ((double)cf[0] - o.x) + ax[k][1] * ((double)cf[1] - o.y) + ax[k][2] * ((double)cf[2] - o.z);
    e[k] = fmax(hi[k] - sc, sc - lo[k]) + 1e-12 * (mag + fabs(sc));
  }
  float ef[3];
  for (int k = 0; k < 3; ++k) {
    ef[k] = (float)e[k];
    if ((double)ef[k] < e[k]) ef[k] = nextafterf(ef[k], INFINITY);
  }
  b.c[0] = cf[0], b.c[1] = cf[1], b.c[2] = cf[2];
  b.x[0] = ax[0][0], b.x[1] = ax[0][1], b.x[2] = ax[0][2];
  b.y[0] = ax[1][0], b.y[1] = ax[1][1], b.y[2] = ax[1][2];
  b.z[0] = ax[2][0], b.z[1] = ax[2][1], b.z[2] = ax[2][2];
  b.ex = ef[0], b.ey = ef[1], b.ez = ef[2];
  b.pad = 0.0f;
}

// Only live nodes write: the root, and nodes covering more than `block` leaves.  A node covering fewer sits
// inside a collapsed subtree (its parent's slot became a leaf block), and its own smaller blocks may start
// at the same triangle as the live block around them.
__global__ __launch_bounds__(256) void k_block_obbs(const Node *__restrict__ nodes, const int2 *__restrict__ ranges,
                                                    const float4 *__restrict__ tris, int ninternal, int block,
                                                    Obb *obbs) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ninternal) return;
  if (i != 0 && ranges[i].y - ranges[i].x + 1 <= block) return;
  for (int k = 0; k < 2; ++k) {
    const int c = k ? nodes[i].c1 : nodes[i].c0;
    if (c >= 0) continue;
    const int code = ~c, first = code & kBlockMask, count = (code >> kBlockShift) + 1;
    if (count < 2) continue;
    Obb b;
    block_obb(tris, first, count, b);
    obbs[first] = b;
  }
}

// Wave reductions of doubles through DPP lane moves (no LDS round trips), for a wave whose 64 lanes are all
// active: pairs, quads (quad_perm), 8 and 16 lanes (half and full row mirrors), then rows 0+1 and 2+3
// (row_bcast:15) and all four rows (row_bcast:31); lane 63 ends with the total, which every lane reads.
// A lane the DPP move does not write keeps `id`, the operation's identity.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v, double id) {
  const long long a = __double_as_longlong(v), o = __double_as_longlong(id);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)a, CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(a >> 32), CTRL, ROWS, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <class Op>
__device__ __forceinline__ double wave_reduce(double v, double id, Op op) {
  v = op(v, dpp_move<0xB1, 0xF>(v, id));   // quad_perm [1, 0, 3, 2]
  v = op(v, dpp_move<0x4E, 0xF>(v, id));   // quad_perm [2, 3, 0, 1]
  v = op(v, dpp_move<0x141, 0xF>(v, id));  // row_half_mirror
  v = op(v, dpp_move<0x140, 0xF>(v, id));  // row_mirror
  v = op(v, dpp_move<0x142, 0xA>(v, id));  // row_bcast:15 into rows 1 and 3
  v = op(v, dpp_move<0x143, 0xC>(v, id));  // row_bcast:31 into rows 2 and 3
  const long long a = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)a, 63), hi = __builtin_amdgcn_readlane((int)(a >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double wave_sum(double v) {
  return wave_reduce(v, 0.0, [](double x, double y) { return x + y; });
}
__device__ __forceinline__ double wave_min(double v) {
  return wave_reduce(v, INFINITY, [](double x, double y) { return fmin(x, y); });
}
__device__ __forceinline__ double wave_max(double v) {
  return wave_reduce(v, -INFINITY, [](double x, double y) { return fmax(x, y); });
}

// Reduction over the L lanes of a node: the wave (L = 64, every lane receives the total) or one DPP row of
// 16 lanes (L = 16: pairs, quads, the half-row and row mirrors leave every lane with its row's total).
// Every lane of the wave must be active.
template <int L, class Op>
__device__ __forceinline__ double lanes_reduce(double v, double id, Op op) {
  if constexpr (L == 64) {
    return wave_reduce(v, id, op);
  } else {
    static_assert(L == 16, "a node takes the wave or one row");
    v = op(v, dpp_move<0xB1, 0xF>(v, id));
    v = op(v, dpp_move<0x4E, 0xF>(v, id));
    v = op(v, dpp_move<0x141, 0xF>(v, id));
    return op(v, dpp_move<0x140, 0xF>(v, id));
  }
}

// Oriented box of internal node i (block_obb's construction) by L lanes: the node's triangles dealt to the
// L lanes (the wave, or the 16 lanes of a DPP row, four nodes to a wave) and the sums, minima and maxima
// combined across them.  i < 0: the lanes box nothing (they still take part in the reductions).  Any frame
// gives a box that holds the triangles; the sums' order only moves the frame.
template <int L>
__device__ __noinline__ void node_obb_lanes(const int2 *__restrict__ ranges, const float4 *__restrict__ tris, int i,
                                            Obb *node_obbs) {
  const int lane = threadIdx.x & (L - 1);
  const auto sum = [](double a, double b) { return a + b; };
  const auto mn = [](double a, double b) { return fmin(a, b); };
  const auto mx = [](double a, double b) { return fmax(a, b); };
  const int first = i >= 0 ? ranges[i].x : 0, size = i >= 0 ? ranges[i].y - ranges[i].x + 1 : 0;
  double n[3] = {0.0, 0.0, 0.0};
  for (int t = first + lane; t < first + size; t += L) {
    const float4 a = tris[3 * t], p = tris[3 * t + 1], q = tris[3 * t + 2];
    const double e1[3] = {(double)p.x - a.x, (double)p.y - a.y, (double)p.z - a.z};
    const double e2[3] = {(double)q.x - a.x, (double)q.y - a.y, (double)q.z - a.z};
    n[0] += e1[1] * e2[2] - e1[2] * e2[1];
    n[1] += e1[2] * e2[0] - e1[0] * e2[2];
    n[2] += e1[0] * e2[1] - e1[1] * e2[0];
  }
  for (int k = 0; k < 3; ++k) n[k] = lanes_reduce<L>(n[k], 0.0, sum);
  double z[3] = {0.0, 0.0, 1.0};
  const double nl = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (nl > 0.0 && nl < INFINITY) z[0] = n[0] / nl, z[1] = n[1] / nl, z[2] = n[2] / nl;
  double x[3];
  {
    const float4 a = tris[3 * first], p = tris[3 * first + 1], q = tris[3 * first + 2];
    const double e[3][3] = {{(double)p.x - a.x, (double)p.y - a.y, (double)p.z - a.z},
                            {(double)q.x - p.x, (double)q.y - p.y, (double)q.z - p.z},
                            {(double)a.x - q.x, (double)a.y - q.y, (double)a.z - q.z}};
    int k = 0;
    double best = -1.0;
    for (int j = 0; j < 3; ++j) {
      const double l = e[j][0] * e[j][0] + e[j][1] * e[j][1] + e[j][2] * e[j][2];
      if (l > best) best = l, k = j;
    }
    for (int j = 0; j < 3; ++j) x[j] = e[k][j];
  }
  for (int attempt = 0; attempt < 2; ++attempt) {
    const double dz = x[0] * z[0] + x[1] * z[1] + x[2] * z[2];
    for (int j = 0; j < 3; ++j) x[j] -= dz * z[j];
    const double xl = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    if (xl > 1e-12 && xl < INFINITY) {
      for (int j = 0; j < 3; ++j) x[j] /= xl;
      break;
    }
    x[0] = fabs(z[0]) < 0.5 ? 1.0 : 0.0;
    x[1] = fabs(z[0]) < 0.5 ? 0.0 : 1.0;
    x[2] = 0.0;
  }
  const double y[3] = {z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2], z[0] * x[1] - z[1] * x[0]};
  const float ax[3][3] = {{(float)x[0], (float)x[1], (float)x[2]},
                          {(float)y[0], (float)y[1], (float)y[2]},
                          {(float)z[0], (float)z[1], (float)z[2]}};
  const float4 o = tris[3 * first];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int c = lane; c < 3 * size; c += L) {
    const float4 p = tris[3 * first + c];
    const double d[3] = {(double)p.x - o.x, (double)p.y - o.y, (double)p.z - o.z};
    for (int k = 0; k < 3; ++k) {
      const double s = ax[k][0] * d[0] + ax[k][1] * d[1] + ax[k][2] * d[2];
      lo[k] = fmin(lo[k], s);
      hi[k] = fmax(hi[k], s);
    }
  }
  for (int k = 0; k < 3; ++k) lo[k] = lanes_reduce<L>(lo[k], INFINITY, mn), hi[k] = lanes_reduce<L>(hi[k], -INFINITY, mx);
  if (lane != 0 || i < 0) return;
  double cc[3] = {o.x, o.y, o.z};
  for (int k = 0; k < 3; ++k)
    for (int j = 0; j < 3; ++j) cc[j] += 0.5 * (lo[k] + hi[k]) * ax[k][j];
  const float cf[3] = {(float)cc[0], (float)cc[1], (float)cc[2]};
  // half extents about the rounded centre from the projections' range: axis . (v - cf) is
  // axis . (v - o) - axis . (cf - o).  The margin covers the FP64 roundings of both sides, a few units in
  // the 16th digit of |v - o| (at most mag, the sum of the largest projections) and of |cf - o|.
  double e[3], mag = 0.0;
  for (int k = 0; k < 3; ++k) mag += fmax(fabs(lo[k]), fabs(hi[k]));
  for (int k = 0; k < 3; ++k) {
    const double sc = ax[k][0] * ((double)cf[0] - o.x) + ax[k][1] * ((double)cf[1] - o.y) + ax[k][2] * ((double)cf[2] - o.z);
    e[k] = fmax(hi[k] - sc, sc - lo[k]) + 1e-12 * (mag + fabs(sc));
  }
  Obb b;
  float ef[3];
  for (int k = 0; k < 3; ++k) {
    ef[k] = (float)e[k];
    if ((double)ef[k] < e[k]) ef[k] = nextafterf(ef[k], INFINITY);
  }
  b.c[0] = cf[0], b.c[1] = cf[1], b.c[2] = cf[2];
  b.x[0] = ax[0][0], b.x[1] = ax[0][1], b.x[2] = ax[0][2];
  b.y[0] = ax[1][0], b.y[1] = ax[1][1], b.y[2] = ax[1][2];
  b.z[0] = ax[2][0], b.z[1] = ax[2][1], b.z[2] = ax[2][2];
  b.ex = ef[0], b.ey = ef[1], b.ez = ef[2];
  b.pad = 0.0f;
  node_obbs[i] = b;
}

// Oriented box of every live internal node covering more than `block` and at most kNodeObbMax triangles,
// at the node's index (those are the internal entries k_build4 marks).  Each wave strides over groups of
// `group` (<= 64) consecutive nodes, finds the group's qualifying nodes with one ballot and boxes them one
// after another (most nodes qualify not, and a wave per node would spend the launch on them; a small mesh
// takes small groups, so that its few waves do not box many nodes each): a node of more than kNodeObbRow
// triangles with the whole wave, smaller ones (most of them) four at a time, one per row of 16 lanes.
constexpr int kNodeObbRow = 64;
__global__ __launch_bounds__(256) void k_node_obbs(const int2 *__restrict__ ranges, const float4 *__restrict__ tris,
                                                   int ninternal, int block, int group, Obb *node_obbs) {
  const int lane = threadIdx.x & 63, row = lane >> 4;
  for (int grp = blockIdx.x * 4 + (threadIdx.x >> 6); grp * group < ninternal; grp += gridDim.x * 4) {
    const int i = grp * group + lane;
    const int size = lane < group && i < ninternal ? ranges[i].y - ranges[i].x + 1 : 0;
    const bool live = size > block && size <= kNodeObbMax;
    for (uint64_t m = __ballot(live && size > kNodeObbRow); m; m &= m - 1)
      node_obb_lanes<64>(ranges, tris, grp * group + __builtin_ctzll(m), node_obbs);
    for (uint64_t m = __ballot(live && size <= kNodeObbRow); m;) {  // wave-uniform
      int pick = -1;  // row r takes the r-th of the remaining nodes
      for (int r = 0; r < 4 && m; ++r, m &= m - 1)
        if (r == row) pick = grp * group + __builtin_ctzll(m);
      node_obb_lanes<16>(ranges, tris, pick, node_obbs);
    }
  }
}

__global__ __launch_bounds__(256) void k_build4(const Node *__restrict__ nodes, const int2 *__restrict__ ranges,
                                                int ninternal, int block, Node4 *out) {
  const int x = blockIdx.x * 256 + threadIdx.x;
  if (x >= ninternal) return;
  Node4 r;
  int n = 0;
  auto put = [&](int code, const float *lo, const float *hi) {
    for (int k = 0; k < 3; ++k) {
      r.lo[n][k] = lo[k];
      r.hi[n][k] = hi[k];
    }
    r.child[n++] = code;
  };
  const Node &nd = nodes[x];
  for (int k = 0; k < 2; ++k) {
    const int c = k ? nd.c1 : nd.c0;
    if (c >= 0) {
      const Node &ch = nodes[c];
      put(ch.c0, ch.lo0, ch.hi0);
      put(ch.c1, ch.lo1, ch.hi1);
    } else {
      put(c, k ? nd.lo1 : nd.lo0, k ? nd.hi1 : nd.hi0);
    }
  }
  for (int e = n; e < 4; ++e) {
    for (int k = 0; k < 3; ++k) {
      r.lo[e][k] = INFINITY;
      r.hi[e][k] = -INFINITY;
    }
    r.child[e] = -1;
  }
  r.count = n;
  r.obb_mask = 0;
  if (block > 1)
    for (int e = 0; e < n; ++e)
      if (r.child[e] >= 0) {
        const int size = ranges[r.child[e]].y - ranges[r.child[e]].x + 1;
        if (size <= kNodeObbMax) r.obb_mask |= 1 << e;
      }
  r.pad[0] = r.pad[1] = 0;
  out[x] = r;
}

__global__ __launch_bounds__(256) void k_depth(const Node *__restrict__ nodes, const int32_t *__restrict__ leaf_parent,
                                               int n, Scalars *sc) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int d = 0;
  int p = leaf_parent[i];
  while (p >= 0 && d < kMaxDepth) {
    ++d;
    p = nodes[p].parent;
  }
  atomicMax(&sc->depth, (uint32_t)(p >= 0 ? kMaxDepth + 1 : d));
}

// one triangle: a single node whose second child is an empty box
__global__ void k_single(const float4 *__restrict__ tris, Node *nodes, float *root_box) {
  const float4 a = tris[0], b = tris[1], c = tris[2];
  Node nd;
  nd.lo0[0] = smin(smin(a.x, b.x), c.x);
  nd.lo0[1] = smin(smin(a.y, b.y), c.y);
  nd.lo0[2] = smin(smin(a.z, b.z), c.z);
  nd.hi0[0] = smax(smax(a.x, b.x), c.x);
  nd.hi0[1] = smax(smax(a.y, b.y), c.y);
  nd.hi0[2] = smax(smax(a.z, b.z), c.z);
  for (int k = 0; k < 3; ++k) {
    nd.lo1[k] = INFINITY;
    nd.hi1[k] = -INFINITY;
    root_box[k] = nd.lo0[k];
    root_box[3 + k] = nd.hi0[k];
  }
  nd.c0 = ~0;
  nd.c1 = ~0;
  nd.parent = -1;
  nd.pad = 0;
  nodes[0] = nd;
}

// ============================================================================================
// weld by exact position (-0 == +0) and count unmatched directed edges
// ============================================================================================
__global__ __launch_bounds__(256) void k_weld_insert(const float *__restrict__ xyz, int ncorner, int32_t *slots,
                                                     uint32_t mask, int32_t *cslot) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ncorner) return;
  const uint32_t a = canon(xyz[3 * i]), b = canon(xyz[3 * i + 1]), c = canon(xyz[3 * i + 2]);
  uint32_t s = hash3(a, b, c) & mask;
  for (;;) {
    int cur = slots[s];
    if (cur < 0) {
      cur = atomicCAS(&slots[s], -1, i);
      if (cur < 0) break;
    }
    if (canon(xyz[3 * cur]) == a && canon(xyz[3 * cur + 1]) == b && canon(xyz[3 * cur + 2]) == c) break;
    s = (s + 1) & mask;
  }
  cslot[i] = (int32_t)s;
}

__global__ __launch_bounds__(256) void k_weld_min(const int32_t *__restrict__ cslot, int ncorner, int32_t *minid) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ncorner) return;
  atomicMin(&minid[cslot[i]], i);
}

__global__ __launch_bounds__(256) void k_edge_insert(const int32_t *__restrict__ cslot,
                                                     const int32_t *__restrict__ minid, int ntri,
                                                     unsigned long long *keys, int32_t *net, uint32_t mask) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= ntri) return;
  int v[3];
  for (int k = 0; k < 3; ++k) v[k] = minid[cslot[3 * t + k]];
  for (int k = 0; k < 3; ++k) {
    const int a = v[k], b = v[k == 2 ? 0 : k + 1];
    if (a == b) continue;
    const uint32_t lo = (uint32_t)min(a, b), hi = (uint32_t)max(a, b);
    const unsigned long long key = ((unsigned long long)lo << 32) | hi;
    uint32_t s = hash64(key) & mask;
    for (;;) {
      unsigned long long cur = keys[s];
      if (cur == ~0ull) {
        cur = atomicCAS(&keys[s], ~0ull, key);
        if (cur == ~0ull) break;
      }
      if (cur == key) break;
      s = (s + 1) & mask;
    }
    atomicAdd(&net[s], a < b ? 1 : -1);
  }
}

// boundary edges: emit each unmatched directed edge (as corner indices = representative vertex
// ids) with its multiplicity.  Each block owns one contiguous range of slots: it counts its edges,
// reserves their places with one atomic, then writes them in slot order (block-wide prefix sums).
__device__ __forceinline__ uint32_t edge_mult(const unsigned long long *keys, const int32_t *net, uint32_t s,
                                              uint32_t end, int &m, int &lo, int &hi) {
  m = 0;
  if (s < end && keys[s] != ~0ull) {
    const unsigned long long key = keys[s];
    m = net[s];
    lo = (int)(key >> 32);
    hi = (int)(key & 0xffffffffu);
  }
  return (uint32_t)(m > 0 ? m : -m);
}

__global__ __launch_bounds__(256) void k_edge_collect(const unsigned long long *__restrict__ keys,
                                                      const int32_t *__restrict__ net, uint32_t cap, Scalars *sc,
                                                      int2 *out, uint32_t out_cap) {
  __shared__ uint32_t wsum[4], base;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t per = ((cap + gridDim.x - 1) / gridDim.x + 255) & ~255u;
  const uint32_t start = min(cap, blockIdx.x * per), end = min(cap, start + per);
  int m, lo, hi;
  uint32_t own = 0;
  for (uint32_t s = start + threadIdx.x; s < end; s += 256) own += edge_mult(keys, net, s, end, m, lo, hi);
  for (int off = 32; off >= 1; off >>= 1) own += __shfl_xor(own, off);
  if (lane == 0) wsum[wv] = own;
  __syncthreads();
  const uint32_t total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
  if (total == 0) return;  // block-uniform
  if (threadIdx.x == 0) base = atomicAdd(&sc->boundary, total);
  __syncthreads();
  uint32_t run = base;
  for (uint32_t it = start; it < end; it += 256) {
    const uint32_t cnt = edge_mult(keys, net, it + threadIdx.x, end, m, lo, hi);
    uint32_t incl = cnt;  // inclusive prefix over the wave
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t o = __shfl_up(incl, off);
      if (lane >= off) incl += o;
    }
    __syncthreads();  // the previous iteration's wave sums are read
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < 4; ++w) {
      before += w < wv ? wsum[w] : 0u;
      all += wsum[w];
    }
    const uint32_t pos = run + before + incl - cnt;
    for (uint32_t k = 0; k < cnt; ++k)
      if (pos + k < out_cap) out[pos + k] = m > 0 ? make_int2(lo, hi) : make_int2(hi, lo);
    run += all;
  }
}

// ============================================================================================
// Hole caps.  The unmatched directed edges (from, to) of an open mesh form closed loops: at every
// vertex as many of them leave as enter (each triangle enters and leaves each of its corners once, and
// a matched pair removes one leaving and one entering edge).  Each loop is closed by a fan from the mean
// of its start vertices; (mesh + reversed fans) is then a closed surface, and the mesh's winding number
// is the closed surface's plus the fans'.  The fans get a hierarchy (Karras tree over their centroids'
// Morton codes) with per-node bounding sphere, box, area and area vector.
// ============================================================================================
using u64 = unsigned long long;

__device__ __forceinline__ u64 shfl_down_u64(u64 v, int off) {
  const uint32_t lo = __shfl_down((uint32_t)v, off), hi = __shfl_down((uint32_t)(v >> 32), off);
  return ((u64)hi << 32) | lo;
}

// sort keys (from << 32 | to) of the unmatched edges
__global__ __launch_bounds__(256) void k_cap_keys(const int2 *__restrict__ bnd, uint32_t ne, u64 *keys) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ne) return;
  const int2 e = bnd[i];
  keys[i] = ((u64)(uint32_t)e.x << 32) | (uint32_t)e.y;
}

// for every edge, in (from, to) order: the key (to << 32 | from) and its position
__global__ __launch_bounds__(256) void k_cap_in_keys(const u64 *__restrict__ ek, uint32_t ne, u64 *ik, uint32_t *iv) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ne) return;
  const u64 k = ek[i];
  ik[i] = (k << 32) | (k >> 32);
  iv[i] = i;
}

__device__ __forceinline__ uint32_t lower_bound_u64(const u64 *__restrict__ a, uint32_t n, u64 key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// The successor of every edge along its loop: the r-th edge entering vertex v (in (to, from) order)
// continues with the r-th edge leaving v (in (from, to) order); any such pairing closes the loops.
// Starts the labelling: lab = least edge index of the edge and its successor, jmp = successor.
__global__ __launch_bounds__(256) void k_cap_succ(const u64 *__restrict__ ek, const u64 *__restrict__ ik,
                                                  const uint32_t *__restrict__ iv, uint32_t ne, uint32_t *lab,
                                                  uint32_t *jmp) {
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p >= ne) return;
  const uint32_t e = iv[p];
  const uint32_t v = (uint32_t)(ik[p] >> 32);
  const u64 key = (u64)v << 32;
  const uint32_t in0 = lower_bound_u64(ik, ne, key), out0 = lower_bound_u64(ek, ne, key);
  const uint32_t idx = out0 + (p - in0);
  // the balance above makes idx an edge leaving v; an edge without one would close on itself
  const uint32_t nxt = (idx < ne && (uint32_t)(ek[idx] >> 32) == v) ? idx : e;
  lab[e] = min(e, nxt);
  jmp[e] = nxt;
}

// Pointer jumping: after step t every edge holds the least index among the next 2^t + 1 edges of its loop
// (the loop's least edge index, its label, once 2^t + 1 reaches the loop's length).
__global__ __launch_bounds__(256) void k_cap_jump(const uint32_t *__restrict__ lab0, const uint32_t *__restrict__ jmp0,
                                                  uint32_t ne, uint32_t *lab1, uint32_t *jmp1) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ne) return;
  const uint32_t j = jmp0[i];
  lab1[i] = min(lab0[i], lab0[j]);
  jmp1[i] = jmp0[j];
}

__global__ __launch_bounds__(256) void k_iota(uint32_t *v, uint32_t n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) v[i] = i;
}

// bounds of the edges' start vertices, ordered-int encoded (bounds[0..2] min, [3..5] max; cleared to
// 0xffffffff and 0 first): the frame of the fixed-point sums
__global__ __launch_bounds__(256) void k_cap_bounds(const u64 *__restrict__ ek, uint32_t ne,
                                                    const float *__restrict__ xyz, uint32_t *bounds) {
  uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < ne; i += gridDim.x * 256) {
    const float *p = xyz + 3 * (size_t)(uint32_t)(ek[i] >> 32);
    for (int k = 0; k < 3; ++k) {
      const uint32_t o = f2o(p[k]);
      mn[k] = min(mn[k], o);
      mx[k] = max(mx[k], o);
    }
  }
  for (int off = 32; off >= 1; off >>= 1)
    for (int k = 0; k < 3; ++k) {
      mn[k] = min(mn[k], (uint32_t)__shfl_xor(mn[k], off));
      mx[k] = max(mx[k], (uint32_t)__shfl_xor(mx[k], off));
    }
  __shared__ uint32_t red[4][6];
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 3; ++k) red[wv][k] = mn[k], red[wv][3 + k] = mx[k];
  __syncthreads();
  if (threadIdx.x < 6) {  // one atomic per block and bound
    const int k = threadIdx.x;
    uint32_t v = red[0][k];
    for (int w = 1; w < 4; ++w) v = k < 3 ? min(v, red[w][k]) : max(v, red[w][k]);
    if (k < 3) atomicMin(&bounds[k], v);
    else atomicMax(&bounds[k], v);
  }
}

// fixed-point frame of axis k: x -> (x - lo) * 2^32 / (hi - lo), in [0, 2^32]
__device__ __forceinline__ void cap_frame(const uint32_t *bounds, int k, double &lo, double &scale) {
  lo = (double)o2f(bounds[k]);
  const double ext = (double)o2f(bounds[3 + k]) - lo;
  scale = ext > 0.0 ? 4294967296.0 / ext : 0.0;
}

// Per loop: the sums of its start vertices in that fixed point, and its edge count (acc[4 * label + 0..3]).
// Integer sums do not depend on the order of the atomics.  The edges come sorted by label, so each wave
// first adds up its runs of equal label (segmented suffix sums) and the first lane of each run adds it.
__global__ __launch_bounds__(256) void k_cap_sums(const uint32_t *__restrict__ slab, const uint32_t *__restrict__ sedge,
                                                  uint32_t ne, const u64 *__restrict__ ek,
                                                  const float *__restrict__ xyz, const uint32_t *__restrict__ bounds,
                                                  u64 *acc) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool valid = i < ne;
  const uint32_t L = valid ? slab[i] : 0xffffffffu;
  u64 v[4] = {0, 0, 0, 0};
  if (valid) {
    const float *p = xyz + 3 * (size_t)(uint32_t)(ek[sedge[i]] >> 32);
    for (int k = 0; k < 3; ++k) {
      double lo, scale;
      cap_frame(bounds, k, lo, scale);
      const double q = fmin(fmax(((double)p[k] - lo) * scale, 0.0), 4294967296.0);
      v[k] = (u64)rint(q);
    }
    v[3] = 1;
  }
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t Lo = __shfl_down(L, off);
    u64 o[4];
    for (int k = 0; k < 4; ++k) o[k] = shfl_down_u64(v[k], off);
    if (lane + off < 64 && Lo == L)
      for (int k = 0; k < 4; ++k) v[k] += o[k];
  }
  const uint32_t Lp = __shfl_up(L, 1);
  if (valid && (lane == 0 || Lp != L))
    for (int k = 0; k < 4; ++k) atomicAdd(&acc[4 * (size_t)L + k], v[k]);
}

// The fan triangle of every edge: its loop's apex (the mean start vertex), the edge's start, its end;
// and the Morton code of the triangle's centroid over the bounds of the boundary's vertices (every fan
// lies inside them), so that the caps need nothing from the mesh's own tree.
__global__ __launch_bounds__(256) void k_cap_tris(const u64 *__restrict__ ek, const uint32_t *__restrict__ lab,
                                                  uint32_t ne, const float *__restrict__ xyz,
                                                  const uint32_t *__restrict__ bounds, const u64 *__restrict__ acc,
                                                  float *tris, uint32_t *codes, int32_t *idx) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= ne) return;
  const uint32_t L = lab[e];
  const double cnt = (double)acc[4 * (size_t)L + 3];
  float t[9];
  for (int k = 0; k < 3; ++k) {
    double lo, scale;
    cap_frame(bounds, k, lo, scale);
    t[k] = scale > 0.0 ? (float)(lo + ((double)acc[4 * (size_t)L + k] / cnt) / scale) : (float)lo;
  }
  const u64 key = ek[e];
  const float *a = xyz + 3 * (size_t)(uint32_t)(key >> 32), *b = xyz + 3 * (size_t)(uint32_t)key;
  for (int k = 0; k < 3; ++k) {
    t[3 + k] = a[k];
    t[6 + k] = b[k];
  }
  for (int k = 0; k < 9; ++k) tris[9 * (size_t)e + k] = t[k];
  codes[e] = centroid_code(t, bounds, bounds + 3);
  idx[e] = (int32_t)e;
}

__global__ __launch_bounds__(256) void k_cap_gather(const float *__restrict__ tmp, const int32_t *__restrict__ order,
                                                    uint32_t ne, float *caps) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= ne) return;
  const float *src = tmp + 9 * (size_t)order[k];
  for (int j = 0; j < 9; ++j) caps[9 * (size_t)k + j] = src[j];
}

// Karras tree of the sorted fans: children codes (internal index, or ~triangle), parent (-1 at the root),
// covered range
__global__ __launch_bounds__(256) void k_cap_karras(const uint32_t *__restrict__ codes, int n, int2 *child,
                                                    int32_t *parent, int2 *ranges) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n - 1) return;
  int lc, rc, lo, hi;
  karras_node(codes, n, i, lc, rc, lo, hi);
  child[i] = make_int2(lc, rc);
  ranges[i] = make_int2(lo, hi);
  if (lc >= 0) parent[lc] = i;
  if (rc >= 0) parent[rc] = i;
  if (i == 0) parent[0] = -1;
}

// Per internal node the traversal reaches (the root and every node whose parent covers more than kCapLeaf
// fans): FP64 sums over the node's fans, bounds rounded outwards so that the FP32 far-field bound
// area / dmin^2 stays an upper bound.  A node of at most kCapLeaf fans becomes a leaf (left = -1 - first,
// right = -count); an inner node keeps its children's codes, of which at most one is a single fan (~index).
// With one fan, record 0 is a leaf of that fan.  W waves (thread tid of 64 W) take the node's fans in
// strides; with W > 1 their partial results combine through `red` in wave order, so the sums do not depend
// on timing.
constexpr int kCapBigFans = 1024;  // nodes above this many fans take a block of kCapBigWaves waves
constexpr int kCapBigWaves = 8;

template <int W>
__device__ __forceinline__ void cap_node(const float *__restrict__ caps, const int2 *__restrict__ child, int rec,
                                         int lo, int hi, int tid, double (*red)[13], CapNode *nodes) {
  const int lane = tid & 63, wv = tid >> 6;
  double area = 0.0, ac[3] = {0.0, 0.0, 0.0}, dip[3] = {0.0, 0.0, 0.0};
  double bl[3] = {INFINITY, INFINITY, INFINITY}, bh[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int t = lo + tid; t <= hi; t += 64 * W) {
    const float *v = caps + 9 * (size_t)t;
    double e1[3], e2[3], sum[3];
    for (int k = 0; k < 3; ++k) {
      sum[k] = ((double)v[k] + v[3 + k]) + v[6 + k];
      bl[k] = fmin(bl[k], fmin((double)v[k], fmin((double)v[3 + k], (double)v[6 + k])));
      bh[k] = fmax(bh[k], fmax((double)v[k], fmax((double)v[3 + k], (double)v[6 + k])));
      e1[k] = (double)v[3 + k] - v[k];
      e2[k] = (double)v[6 + k] - v[k];
    }
    const double nv[3] = {0.5 * (e1[1] * e2[2] - e1[2] * e2[1]), 0.5 * (e1[2] * e2[0] - e1[0] * e2[2]),
                          0.5 * (e1[0] * e2[1] - e1[1] * e2[0])};
    const double a = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    area += a;
    for (int k = 0; k < 3; ++k) {
      dip[k] += nv[k];
      ac[k] += a * sum[k] / 3.0;
    }
  }
  area = wave_sum(area);
  for (int k = 0; k < 3; ++k) {
    ac[k] = wave_sum(ac[k]);
    dip[k] = wave_sum(dip[k]);
    bl[k] = wave_min(bl[k]);
    bh[k] = wave_max(bh[k]);
  }
  if constexpr (W > 1) {
    if (lane == 0) {
      red[wv][0] = area;
      for (int k = 0; k < 3; ++k) {
        red[wv][1 + k] = ac[k];
        red[wv][4 + k] = dip[k];
        red[wv][7 + k] = bl[k];
        red[wv][10 + k] = bh[k];
      }
    }
    __syncthreads();
    area = red[0][0];
    for (int k = 0; k < 3; ++k) {
      ac[k] = red[0][1 + k];
      dip[k] = red[0][4 + k];
      bl[k] = red[0][7 + k];
      bh[k] = red[0][10 + k];
    }
    for (int w = 1; w < W; ++w) {
      area += red[w][0];
      for (int k = 0; k < 3; ++k) {
        ac[k] += red[w][1 + k];
        dip[k] += red[w][4 + k];
        bl[k] = fmin(bl[k], red[w][7 + k]);
        bh[k] = fmax(bh[k], red[w][10 + k]);
      }
    }
    __syncthreads();
  }
  double c[3];
  for (int k = 0; k < 3; ++k) c[k] = area > 0.0 ? ac[k] / area : 0.5 * (bl[k] + bh[k]);
  double r2 = 0.0;
  for (int t = lo + tid; t <= hi; t += 64 * W) {
    const float *v = caps + 9 * (size_t)t;
    for (int j = 0; j < 3; ++j) {
      const double dx = v[3 * j] - c[0], dy = v[3 * j + 1] - c[1], dz = v[3 * j + 2] - c[2];
      r2 = fmax(r2, dx * dx + dy * dy + dz * dz);
    }
  }
  r2 = wave_max(r2);
  if constexpr (W > 1) {
    if (lane == 0) red[wv][0] = r2;
    __syncthreads();
    for (int w = 0; w < W; ++w) r2 = fmax(r2, red[w][0]);
    __syncthreads();
  }
  if (tid != 0) return;
  CapNode nd;
  for (int k = 0; k < 3; ++k) {
    nd.c[k] = (float)c[k];
    nd.dip[k] = (float)dip[k];
  }
  for (int k = 0; k < 2; ++k) {
    nd.lo[k] = nextafterf((float)bl[k], -INFINITY);
    nd.hi[k] = nextafterf((float)bh[k], INFINITY);
  }
  nd.pad = 0.0f;
  // the float centre moved by at most one ulp per axis: widen r by that, then by a relative margin
  nd.r = (float)((sqrt(r2) + 1e-6 * (fabs(c[0]) + fabs(c[1]) + fabs(c[2]))) * (1.0 + 1e-5));
  nd.area = (float)(area * (1.0 + 1e-5));
  // cap_far's bound exceeds kCapTauHi while d - r < min(D0, D1), D0 = sqrt(area / (4 pi tau)) (zero order),
  // D1 = cbrt(2 area r / (4 pi tau)) (first order): below (r + min(D0, D1))^2, less a relative margin, no
  // voxel can accept the node
  const double tau = (double)kCapTauHi, ar = (double)nd.area, rr = (double)nd.r;
  const double d0 = sqrt(ar / (4.0 * M_PI * tau)), d1 = cbrt(2.0 * ar * rr / (4.0 * M_PI * tau));
  const double dopen = rr + fmin(d0, d1);
  nd.open2 = (float)(dopen * dopen * (1.0 - 1e-4));
  const int size = hi - lo + 1;
  if (size <= kCapLeaf) {
    nd.left = -1 - lo;
    nd.right = -size;
  } else {
    nd.left = child[rec].x;
    nd.right = child[rec].y;
  }
  nodes[rec] = nd;
}

// the node's fan range, or false when the traversal never reaches it (its parent is a leaf)
__device__ __forceinline__ bool cap_node_range(const int32_t *__restrict__ parent, const int2 *__restrict__ ranges,
                                               int n, int rec, int &lo, int &hi) {
  lo = hi = 0;
  if (n <= 1) return true;
  lo = ranges[rec].x;
  hi = ranges[rec].y;
  const int p = parent[rec];
  return !(p >= 0 && ranges[p].y - ranges[p].x + 1 <= kCapLeaf);
}

// one wave per node of at most kCapBigFans fans
__global__ __launch_bounds__(256) void k_cap_nodes(const float *__restrict__ caps, const int2 *__restrict__ child,
                                                   const int32_t *__restrict__ parent,
                                                   const int2 *__restrict__ ranges, int n, CapNode *nodes) {
  const int rec = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (rec >= (n > 1 ? n - 1 : 1)) return;  // wave-uniform
  int lo, hi;
  if (!cap_node_range(parent, ranges, n, rec, lo, hi) || hi - lo + 1 > kCapBigFans) return;
  cap_node<1>(caps, child, rec, lo, hi, threadIdx.x & 63, nullptr, nodes);
}

// one block per node of more than kCapBigFans fans (the few top levels), grid-stride over the nodes
__global__ __launch_bounds__(64 * kCapBigWaves) void k_cap_nodes_big(const float *__restrict__ caps,
                                                                     const int2 *__restrict__ child,
                                                                     const int32_t *__restrict__ parent,
                                                                     const int2 *__restrict__ ranges, int n,
                                                                     CapNode *nodes) {
  __shared__ double red[kCapBigWaves][13];
  for (int rec = blockIdx.x; rec < n - 1; rec += gridDim.x) {  // block-uniform
    int lo, hi;
    if (!cap_node_range(parent, ranges, n, rec, lo, hi) || hi - lo + 1 <= kCapBigFans) continue;
    cap_node<kCapBigWaves>(caps, child, rec, lo, hi, threadIdx.x, red, nodes);
  }
}

// The cut of the cap tree at depth kCapCutBits: slot s follows the path that its bits spell from the root
// (most significant first, 0 = left); a node or fan reached in fewer steps belongs to the slot whose
// remaining bits are zero, and the other slots on that path stay empty.  Each slot holds (code, first fan,
// fan count, 0): the slots partition the fans into contiguous ranges of the sorted cap soup.
__global__ __launch_bounds__(256) void k_cap_cut(const CapNode *__restrict__ nodes, const int2 *__restrict__ ranges,
                                                 int4 *cut) {
  const int slot = blockIdx.x * 256 + threadIdx.x;
  if (slot >= kCapCut) return;
  int code = 0;  // record 0 is the root (a leaf when the caps have at most kCapLeaf fans)
  bool mine = true;
  for (int k = kCapCutBits - 1; k >= 0; --k) {
    const int bit = (slot >> k) & 1;
    const bool inner = code >= 0 && !(nodes[code].left < 0 && nodes[code].right < 0);
    if (inner) code = bit ? nodes[code].right : nodes[code].left;
    else if (bit) mine = false;
  }
  int first = 0, count = 0;
  if (mine) {
    if (code < 0) {
      first = ~code;
      count = 1;
    } else if (nodes[code].left < 0 && nodes[code].right < 0) {
      first = -1 - nodes[code].left;
      count = -nodes[code].right;
    } else {
      first = ranges[code].x;
      count = ranges[code].y - ranges[code].x + 1;
    }
  }
  cut[slot] = make_int4(code, first, count, 0);
}

// ============================================================================================
// surface occupancy: voxelize.cpp:64-100 per triangle, bit-packed with atomicOr
// ============================================================================================
// inclusive cell range of a triangle's AABB: floor((min - origin) / h) clamped with
// glm::clamp(x, 0, n - 1) = min(max(x, 0), n - 1) on every axis (voxelize.cpp:70-77)
__device__ __forceinline__ void tri_cells(const float *p, const GridArgs &g, int lo[3], int hi[3]) {
  const int maxc[3] = {g.nx - 1, g.ny - 1, g.nz - 1};
  const float org[3] = {g.ox, g.oy, g.oz};
  for (int k = 0; k < 3; ++k) {
    const float mn = smin(smin(p[k], p[3 + k]), p[6 + k]);
    const float mx = smax(smax(p[k], p[3 + k]), p[6 + k]);
    int a = (int)floorf((mn - org[k]) / g.h);
    int b = (int)floorf((mx - org[k]) / g.h);
    a = a < 0 ? 0 : a;
    a = a > maxc[k] ? maxc[k] : a;
    b = b < 0 ? 0 : b;
    b = b > maxc[k] ? maxc[k] : b;
    lo[k] = a;
    hi[k] = b;
  }
}

// cell centre origin + (cell + 0.5) * h (voxelize.cpp:84), half size h * 0.5 (:85)
__device__ __forceinline__ F3 cell_centre(const GridArgs &g, int x, int y, int z) {
  return {g.ox + ((float)x + 0.5f) * g.h, g.oy + ((float)y + 0.5f) * g.h, g.oz + ((float)z + 0.5f) * g.h};
}

// One wave per triangle: the cells of the triangle's AABB are numbered x fastest and dealt to the
// 64 lanes (cell = lane + 64 i), so a triangle that spans thousands of cells (a coarse mesh on a fine
// grid) is tested 64 cells at a time instead of by one thread.  fn(x, y, z, hit) runs in every lane
// of every round, uniformly (hit = false for lanes past the last cell), so it may use wave ballots.
template <typename Fn>
__device__ __forceinline__ void tri_cells_wave(const float *p, const GridArgs &g, Fn &&fn) {
  const int lane = threadIdx.x & 63;
  const F3 t0 = {p[0], p[1], p[2]}, t1 = {p[3], p[4], p[5]}, t2 = {p[6], p[7], p[8]};
  int lo[3], hi[3];
  tri_cells(p, g, lo, hi);
  const float half = g.h * 0.5f;
  const uint32_t ex = (uint32_t)(hi[0] - lo[0] + 1), ey = (uint32_t)(hi[1] - lo[1] + 1);
  const uint64_t total = (uint64_t)ex * ey * (uint64_t)(hi[2] - lo[2] + 1);
  for (uint64_t base = 0; base < total; base += 64) {
    const uint64_t c = base + (uint64_t)lane;
    int x = 0, y = 0, z = 0;
    bool hit = false;
    if (c < total) {
      const uint64_t row = c / ex;
      x = lo[0] + (int)(c - row * ex);
      y = lo[1] + (int)(row % ey);
      z = lo[2] + (int)(row / ey);
      hit = tri_box(cell_centre(g, x, y, z), half, t0, t1, t2);
    }
    fn(x, y, z, hit);
  }
}

__global__ __launch_bounds__(256) void k_occupancy(const float *__restrict__ xyz, int n, GridArgs g,
                                                   uint32_t *words) {
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= n) return;
  const int wpr = (g.nx + 31) >> 5;
  tri_cells_wave(xyz + 9 * (size_t)t, g, [&](int x, int y, int z, bool hit) {
    if (hit) atomicOr(&words[((size_t)z * g.ny + y) * wpr + (x >> 5)], 1u << (x & 31));
  });
}


// per-cell attribute splat (sdf_baker.cpp:76-128, voxelize.cpp:94-104): every (triangle, cell)
// overlap pair, triangles in order (pairs of one triangle are contiguous, each cell once), stably
// sorted by cell, then summed sequentially per cell
__global__ __launch_bounds__(256) void k_splat_count(const float *__restrict__ xyz, int n, GridArgs g,
                                                     uint32_t *counts) {
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= n) return;
  uint32_t c = 0;
  tri_cells_wave(xyz + 9 * (size_t)t, g, [&](int, int, int, bool hit) {
    c += (uint32_t)__popcll(__ballot(hit));
  });
  if ((threadIdx.x & 63) == 0) counts[t] = c;
}

__global__ __launch_bounds__(256) void k_splat_emit(const float *__restrict__ xyz, int n, GridArgs g,
                                                    const uint32_t *__restrict__ offsets, uint32_t *cell,
                                                    int32_t *tri) {
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= n) return;
  const uint64_t below = (1ull << (threadIdx.x & 63)) - 1ull;
  uint32_t o = offsets[t];
  tri_cells_wave(xyz + 9 * (size_t)t, g, [&](int x, int y, int z, bool hit) {
    const uint64_t hits = __ballot(hit);
    if (hit) {
      const uint32_t at = o + (uint32_t)__popcll(hits & below);
      cell[at] = (uint32_t)(((uint64_t)z * g.ny + y) * g.nx + x);
      tri[at] = t;
    }
    o += (uint32_t)__popcll(hits);
  });
}

// out: per cell (k sums followed by the count as float), written only for touched cells
__global__ __launch_bounds__(256) void k_splat_reduce(const uint32_t *__restrict__ cell, const int32_t *__restrict__ tri,
                                                      uint32_t npairs, const float *__restrict__ attr, int k,
                                                      float *out, uint64_t ncell) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= npairs) return;
  const uint32_t c = cell[i];
  if (i > 0 && cell[i - 1] == c) return;  // not the first pair of its cell
  float sum[8];
  for (int j = 0; j < 8; ++j) sum[j] = 0.0f;
  uint32_t cnt = 0;
  for (uint32_t p = i; p < npairs && cell[p] == c; ++p) {
    const float *a = attr + (size_t)k * tri[p];
    for (int j = 0; j < k && j < 8; ++j) sum[j] += a[j];
    ++cnt;
  }
  const float fc = (float)cnt;
  // attributes 0..2 (the colour) interleaved per cell, as std::vector<glm::vec3> holds them; the others
  // in planes of ncell after them
  for (int j = 0; j < k && j < 8; ++j) out[j < 3 ? 3 * (size_t)c + j : (size_t)j * ncell + c] = sum[j] / fc;
}

// voxelizeMesh's cells from the (cell, triangle) pairs sorted by cell (triangles in order inside a cell):
// head[i] = 1 where a cell's run starts
__global__ __launch_bounds__(256) void k_vox_heads(const uint32_t *__restrict__ cell, uint32_t npairs, uint32_t *head) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < npairs) head[i] = (i == 0 || cell[i - 1] != cell[i]) ? 1u : 0u;
  if (i == npairs) head[i] = 0u;
}

// one thread per run (run index from the exclusive scan of the heads): the first-touch key
// (first triangle << 32 | cell), its position as the sort value, and the colour mean of voxelize.cpp:94-104
// (the sum over the cell's triangles in triangle order, divided by their count, per component)
__global__ __launch_bounds__(256) void k_vox_runs(const uint32_t *__restrict__ cell, const int32_t *__restrict__ tri,
                                                  uint32_t npairs, const uint32_t *__restrict__ run_of,
                                                  const float *__restrict__ tri_rgb, unsigned long long *keys,
                                                  uint32_t *vals, float *rgb) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= npairs || (i > 0 && cell[i - 1] == cell[i])) return;
  const uint32_t r = run_of[i], c = cell[i];
  keys[r] = ((unsigned long long)(uint32_t)tri[i] << 32) | c;
  vals[r] = r;
  if (!rgb) return;
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
  uint32_t n = 0;
  for (uint32_t p = i; p < npairs && cell[p] == c; ++p, ++n) {
    if (tri_rgb) {
      const float *a = tri_rgb + 3 * (size_t)tri[p];
      s0 += a[0], s1 += a[1], s2 += a[2];
    } else {
      s0 += 1.0f, s1 += 1.0f, s2 += 1.0f;
    }
  }
  const float fn = (float)n;
  rgb[3 * (size_t)r] = s0 / fn;
  rgb[3 * (size_t)r + 1] = s1 / fn;
  rgb[3 * (size_t)r + 2] = s2 / fn;
}

// the sorted runs as cells (x, y, z), their first triangles and their colours (gathered through the sort)
__global__ __launch_bounds__(256) void k_vox_out(const unsigned long long *__restrict__ keys,
                                                 const uint32_t *__restrict__ perm, uint32_t n, int nx, int ny,
                                                 const float *__restrict__ rgb, int32_t *cells, int32_t *first,
                                                 float *colors) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const unsigned long long key = keys[k];
  const uint32_t i = (uint32_t)(key & 0xffffffffu);
  if (cells) {
    cells[3 * k + 0] = (int32_t)(i % (uint32_t)nx);
    cells[3 * k + 1] = (int32_t)((i / (uint32_t)nx) % (uint32_t)ny);
    cells[3 * k + 2] = (int32_t)(i / ((uint32_t)nx * (uint32_t)ny));
  }
  if (first) first[k] = (int32_t)(key >> 32);
  if (colors) {
    const uint32_t r = perm[k];
    colors[3 * k] = rgb[3 * (size_t)r];
    colors[3 * k + 1] = rgb[3 * (size_t)r + 1];
    colors[3 * k + 2] = rgb[3 * (size_t)r + 2];
  }
}

// triangleMaterial (src/convert/sdf_baker.cpp:61-74) with sampleColorNearest (:44-53); the sRGB curve
// (:42) comes in as a table over the 256 byte values, made on the host with the reference's expression
__global__ __launch_bounds__(256) void k_tri_materials(uint32_t ntri, const float *__restrict__ uvs, uint64_t nuv,
                                                       const uint32_t *__restrict__ tri_mat, uint64_t ntm,
                                                       const MatDev *__restrict__ mats, uint32_t nmat,
                                                       const TexDev *__restrict__ texs, uint32_t ntex,
                                                       const uint8_t *__restrict__ pixels,
                                                       const float *__restrict__ lut, float *out) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= ntri) return;
  MatDev m = {{1.0f, 1.0f, 1.0f}, 0.0f, 1.0f, -1};  // MaterialAsset's defaults for an empty table
  if (nmat) {
    const uint32_t mi = t < ntm ? tri_mat[t] : 0u;
    m = mats[mi < nmat ? mi : 0u];
  }
  float col[3] = {m.base[0], m.base[1], m.base[2]};
  const uint64_t base = 3ull * t;
  if (m.tex >= 0 && (uint32_t)m.tex < ntex && base + 2 < nuv) {
    const TexDev tx = texs[m.tex];
    if (tx.w > 0 && tx.h > 0) {
      const float u = ((uvs[2 * base] + uvs[2 * base + 2]) + uvs[2 * base + 4]) / 3.0f;
      const float v = ((uvs[2 * base + 1] + uvs[2 * base + 3]) + uvs[2 * base + 5]) / 3.0f;
      const float wu = u - floorf(u), wv = v - floorf(v);
      const int px = min(max((int)(wu * (float)tx.w), 0), tx.w - 1);
      const int py = min(max((int)(wv * (float)tx.h), 0), tx.h - 1);
      const uint8_t *tp = pixels + tx.offset + ((uint64_t)py * (uint64_t)tx.w + (uint64_t)px) * 4;
      for (int k = 0; k < 3; ++k) {
        const float raw = tx.srgb ? lut[tp[k]] : (float)tp[k] / 255.0f;
        col[k] = raw * m.base[k];
      }
    }
  }
  float *o = out + 5ull * t;
  o[0] = col[0];
  o[1] = col[1];
  o[2] = col[2];
  o[3] = m.metallic;
  o[4] = m.roughness;
}



// ============================================================================================
// +z column crossings through the mesh and (reversed) hole caps
// ============================================================================================
__device__ __forceinline__ F3 f3(float4 v) { return {v.x, v.y, v.z}; }

struct ColumnSink {
  float2 *dst;
  int cnt;
  __device__ void add(float z, int s) {
    if (cnt < kColCap) dst[cnt] = make_float2(z, (float)s);
    ++cnt;
  }
};

// A leaf child code ~L names the triangles L & kBlockMask .. + (L >> kBlockShift): one triangle, or
// a collapsed block of up to kMaxLeafBlock consecutive ones.
template <typename F>
__device__ __forceinline__ void for_leaf_code(int code, F &&leaf) {
  const int t = code & kBlockMask, last = t + (code >> kBlockShift);
  for (int l = t; l <= last; ++l) leaf(l);
}

// a traversal push that found the stack full: impossible within kMaxTreeDepth (the host refuses deeper
// trees), counted so that the host reports IB_ERR_INTERNAL instead of a silently pruned search
__device__ __noinline__ void stack_overflow(Scalars *sc) { atomicAdd(&sc->stack_overflow, 1u); }

// visits every leaf whose box contains (X,Y) in xy
template <typename F>
__device__ __forceinline__ void for_column_leaves(const Node *__restrict__ nodes, int nnodes, float X, float Y,
                                                  Scalars *sc, F &&leaf) {
  int stack[kStack + 2];
  int sp = 0;
  int cur = 0;
  for (int visits = 0; visits <= nnodes; ++visits) {
    const Node nd = nodes[cur];
    const bool in0 = X >= nd.lo0[0] && X <= nd.hi0[0] && Y >= nd.lo0[1] && Y <= nd.hi0[1];
    const bool in1 = X >= nd.lo1[0] && X <= nd.hi1[0] && Y >= nd.lo1[1] && Y <= nd.hi1[1];
    int next = -1;
    if (in0) {
      if (nd.c0 < 0) for_leaf_code(~nd.c0, leaf); else next = nd.c0;
    }
    if (in1) {
      if (nd.c1 < 0) for_leaf_code(~nd.c1, leaf);
      else if (next < 0) next = nd.c1;
      else if (sp < kStack) stack[sp++] = nd.c1;
      else stack_overflow(sc);
    }
    if (next >= 0) {
      cur = next;
      continue;
    }
    if (sp == 0) break;
    cur = stack[--sp];
  }
}

// every cap triangle whose box contains (X, Y) in xy
template <typename F>
__device__ __forceinline__ void for_column_caps(const ColumnArgs &a, float X, float Y, F &&tri) {
  if (a.ncap == 0 || a.ncapnode == 0 || !a.cap_nodes) return;
  int stack[kCapStack];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const int code = stack[--sp];
    if (code < 0) {  // a single fan
      const float *t = a.caps + 9 * (size_t)~code;
      if (X >= fminf(fminf(t[0], t[3]), t[6]) && X <= fmaxf(fmaxf(t[0], t[3]), t[6]) &&
          Y >= fminf(fminf(t[1], t[4]), t[7]) && Y <= fmaxf(fmaxf(t[1], t[4]), t[7]))
        tri(t);
      continue;
    }
    const CapNode nd = a.cap_nodes[code];
    if (!(X >= nd.lo[0] && X <= nd.hi[0] && Y >= nd.lo[1] && Y <= nd.hi[1])) continue;
    if (nd.left < 0 && nd.right < 0) {
      const int first = -1 - nd.left;
      for (int k = 0; k < -nd.right; ++k) tri(a.caps + 9 * (size_t)(first + k));
    } else if (sp + 2 <= kCapStack) {
      stack[sp++] = nd.left;
      stack[sp++] = nd.right;
    } else {
      stack_overflow(a.sc);
    }
  }
}

// (X, Y) inside the closed xy bounding box of the triangle: a +z line can cross the triangle only then (the
// crossing test's half-open rule only decides between triangles that all contain the point in that box)
__device__ __forceinline__ bool column_in_xy_box(float X, float Y, const float4 *t) {
  return X >= fminf(fminf(t[0].x, t[1].x), t[2].x) && X <= fmaxf(fmaxf(t[0].x, t[1].x), t[2].x) &&
         Y >= fminf(fminf(t[0].y, t[1].y), t[2].y) && Y <= fmaxf(fmaxf(t[0].y, t[1].y), t[2].y);
}

// winding (integer part) of the closed surface mesh + reversed caps for a voxel whose column
// overflowed the stored crossings: recount crossings above q.z
__device__ __noinline__ int column_direct(const ColumnArgs &a, float X, float Y, float qz) {
  int w = 0;
  for_column_leaves(a.nodes, a.nnodes, X, Y, a.sc, [&](int leaf) {
    const float4 *t = a.tris + 3 * leaf;
    if (!column_in_xy_box(X, Y, t)) return;
    float z;
    const int s = column_cross(X, Y, f3(t[0]), f3(t[1]), f3(t[2]), &z);
    if (s && z > qz) w += s;
  });
  for_column_caps(a, X, Y, [&](const float *t) {
    float z;
    const int s = column_cross(X, Y, {t[0], t[1], t[2]}, {t[3], t[4], t[5]}, {t[6], t[7], t[8]}, &z);
    if (s && z > qz) w -= s;
  });
  return w;
}

// ============================================================================================
// hole caps: solid angles (Van Oosterom & Strackee 1983)
// ============================================================================================
// Omega = 2 atan2(det, den).  `singular` is set when q lies on or next to the triangle itself, where the
// solid angle jumps by 4 pi: den within 1e-3 la lb lc of zero or below (q's projection inside the
// triangle or at its edge) while q is within 1e-3 la lb lc of its plane by det, or within a few FP32 ulps
// of it along z (|det| / |n_z|, n = (b - a) x (c - a)), the resolution of the column's crossing z.
// There the FP32 sign, and the consistency of this solid angle with the column crossing of the same
// cap, are not reliable, and the voxel goes to the FP64 recheck.
__device__ __forceinline__ float solid_angle_f32(F3 q, const float *t, bool &singular) {
  const F3 a = {t[0] - q.x, t[1] - q.y, t[2] - q.z};
  const F3 b = {t[3] - q.x, t[4] - q.y, t[5] - q.z};
  const F3 c = {t[6] - q.x, t[7] - q.y, t[8] - q.z};
  const float la = sqrtf(dot3(a, a)), lb = sqrtf(dot3(b, b)), lc = sqrtf(dot3(c, c));
  const F3 bc = {b.y * c.z - b.z * c.y, b.z * c.x - b.x * c.z, b.x * c.y - b.y * c.x};
  const float det = dot3(a, bc);
  const float l3 = la * lb * lc;
  const float den = l3 + dot3(a, b) * lc + dot3(b, c) * la + dot3(c, a) * lb;
  const float tol = 1e-3f * l3;
  if (den <= tol) {
    const float nz = (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x);
    const float zeps = 4.8e-7f * (fabsf(q.z) + fabsf(t[2]) + fabsf(t[5]) + fabsf(t[8]));  // ~4 ulps of z
    singular = singular || fabsf(det) <= tol || fabsf(det) <= zeps * fabsf(nz);
  }
  return 2.0f * atan2f(det, den);
}

// Far field of a cap node seen from q: the bound in winding units, the term w1 in steradians.  The node's
// fans lie within r of c and have total area `area` and area vector `dip`.  First-order term w1 = dip . (c - q) /
// |c - q|^3: the integrand n . v / |v|^3 changes by at most 2 |dx| / |v|^3 per unit move dx of the point on the
// fan (the operator norm of the derivative of v / |v|^3), so |Omega - w1| <= 2 area r / (d - r)^3, d = |c - q|.
// Without the term, |Omega| <= area / (d - r)^2.  Returns the smaller bound (with w1 = 0 for the second);
// infinite when q lies within r of c.  Hardware reciprocal square root and reciprocal (1 ulp each) instead of
// correctly rounded sqrt and divides: d is rounded down by 2^-20 relative before r is subtracted, 1 / (d - r)
// up by 2^-20, and the bound up by 2^-18, which covers every rounding of the chain; the evaluation of w1 errs
// by less than 2^-21 |w1| <= 2^-21 area / (d - r)^2, inside the 1e-5 * area / (d - r)^2 added to its bound.
__device__ __forceinline__ float cap_far_at(const CapNode &nd, float dx, float dy, float dz, float d2, float &w1) {
  const float inv4pi = (float)(1.0 / (4.0 * M_PI));
  const float rd = __builtin_amdgcn_rsqf(d2);  // 1 / d (infinite at 0: then dmin is NaN)
  const float dmin = d2 * rd * (1.0f - 0x1p-20f) - nd.r;
  w1 = 0.0f;
  if (!(dmin > 0.0f)) return INFINITY;
  const float id = __builtin_amdgcn_rcpf(dmin) * (1.0f + 0x1p-20f);
  const float b0 = nd.area * inv4pi * id * id;
  const float b1 = b0 * (2.0f * nd.r * id);
  if (!(b1 < b0)) return b0 * (1.0f + 0x1p-18f);
  w1 = (nd.dip[0] * dx + nd.dip[1] * dy + nd.dip[2] * dz) * (rd * rd * rd);  // steradians, like the fans' sum
  return (b1 + 1e-5f * b0) * (1.0f + 0x1p-18f);
}
// the offset c - q and its squared length, as cap_far evaluates them
__device__ __forceinline__ float cap_d2(const CapNode &nd, F3 q, float &dx, float &dy, float &dz) {
  dx = nd.c[0] - q.x;
  dy = nd.c[1] - q.y;
  dz = nd.c[2] - q.z;
  return dx * dx + dy * dy + dz * dz;
}
__device__ __forceinline__ float cap_far(const CapNode &nd, F3 q, float &w1) {
  float dx, dy, dz;
  const float d2 = cap_d2(nd, q, dx, dy, dz);
  return cap_far_at(nd, dx, dy, dz, d2, w1);
}

// Cap winding with a rigorous error bound: caps near q are summed exactly (FP32 solid angles,
// Neumaier), every farther node contributes its first-order term and its error bound (cap_far) to
// `bound`, in winding units.
__device__ __forceinline__ void cap_winding_bounded(const ColumnArgs &a, F3 q, float &w, float &bound) {
  w = 0.0f;
  bound = 0.0f;
  // a batch runs the capped kernel for every job once one job has caps: closed meshes have none
  if (a.ncap == 0 || a.ncapnode == 0 || !a.cap_nodes) return;
  const float inv4pi = (float)(1.0 / (4.0 * M_PI));
  float sum = 0.0f, cmp = 0.0f, ub = 0.0f;
  bool singular = false;
  int stack[kCapStack];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const int code = stack[--sp];
    if (code < 0) {  // a single fan
      const float om = solid_angle_f32(q, a.caps + 9 * (size_t)~code, singular);
      const float t = sum + om;
      cmp += (fabsf(sum) >= fabsf(om)) ? (sum - t) + om : (om - t) + sum;
      sum = t;
      continue;
    }
    const CapNode nd = a.cap_nodes[code];
    float w1;
    const float b = cap_far(nd, q, w1);
    if (b <= kCapTau || (b <= kCapTauHi && ub + b <= kCapBudget)) {
      ub += b;
      const float t = sum + w1;
      cmp += (fabsf(sum) >= fabsf(w1)) ? (sum - t) + w1 : (w1 - t) + sum;
      sum = t;
      continue;
    }
    if (nd.left < 0 && nd.right < 0) {
      const int first = -1 - nd.left;
      for (int k = 0; k < -nd.right; ++k) {
        const float om = solid_angle_f32(q, a.caps + 9 * (size_t)(first + k), singular);
        const float t = sum + om;
        cmp += (fabsf(sum) >= fabsf(om)) ? (sum - t) + om : (om - t) + sum;
        sum = t;
      }
    } else if (sp + 2 <= kCapStack) {
      stack[sp++] = nd.left;
      stack[sp++] = nd.right;
    } else {
      ub += 1.0f;  // stack exhausted: give up the bound, the voxel goes to the FP64 recheck
    }
  }
  w = (sum + cmp) * inv4pi;
  bound = singular ? 1.0f : ub * 1.0001f + 1e-5f;  // + FP32 rounding of the bound and the sum
}

// FP64 solid angle of triangle t (9 floats) at q; `singular` as for the FP32 form, with 1e-9 in place of
// 1e-3 and the same z resolution (the column crossings are stored in FP32)
__device__ __forceinline__ double solid_angle_f64(F3 q, const float *t, bool &singular) {
  const double ax = (double)t[0] - q.x, ay = (double)t[1] - q.y, az = (double)t[2] - q.z;
  const double bx = (double)t[3] - q.x, by = (double)t[4] - q.y, bz = (double)t[5] - q.z;
  const double cx = (double)t[6] - q.x, cy = (double)t[7] - q.y, cz = (double)t[8] - q.z;
  const double la = sqrt(dot3d(ax, ay, az, ax, ay, az)), lb = sqrt(dot3d(bx, by, bz, bx, by, bz)),
               lc = sqrt(dot3d(cx, cy, cz, cx, cy, cz));
  const double det = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
  const double l3 = la * lb * lc;
  const double den = l3 + dot3d(ax, ay, az, bx, by, bz) * lc + dot3d(bx, by, bz, cx, cy, cz) * la +
                     dot3d(cx, cy, cz, ax, ay, az) * lb;
  if (den <= 1e-9 * l3) {
    const double nz = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
    const double zeps = 4.8e-7 * (fabs((double)q.z) + fabs((double)t[2]) + fabs((double)t[5]) + fabs((double)t[8]));
    singular = singular || fabs(det) <= 1e-9 * l3 || fabs(det) <= zeps * fabs(nz);
  }
  return 2.0 * atan2(det, den);
}

// Compensated sum of the 64 lanes' (sum, compensation) pairs by a butterfly of error-free two-sums;
// every lane returns the total, so the value is wave-uniform and the same from run to run.
__device__ __forceinline__ double wave_sum_f64(double sum, double comp) {
  double v = sum, e = comp;
  for (int off = 32; off >= 1; off >>= 1) {
    const double ov = __shfl_xor(v, off), oe = __shfl_xor(e, off);
    const double s = v + ov, bp = s - v;
    e = e + oe + ((v - (s - bp)) + (ov - bp));
    v = s;
  }
  return __shfl(v + e, 0);
}

// FP64 sum of the caps' solid angles [c0, c1) at q (steradians), computed by a whole wave: cap i goes to
// lane i % 64, each lane keeps a Neumaier-compensated sum, and the 64 partial sums are combined by
// wave_sum_f64.  `singular` (wave-uniform) is set when q lies on or next to one of those cap triangles.
__device__ __noinline__ double cap_solid_angle_f64_wave(F3 q, const float *caps, int c0, int c1, bool &singular) {
  const int lane = threadIdx.x & 63;
  double sum = 0.0, comp = 0.0;
  bool sing = false;
#pragma unroll 1
  for (int i = c0 + lane; i < c1; i += 64) {
    const double om = solid_angle_f64(q, caps + 9 * (size_t)i, sing);
    const double s = sum + om;
    comp += (fabs(sum) >= fabs(om)) ? (sum - s) + om : (om - s) + sum;
    sum = s;
  }
  singular = __ballot(sing) != 0;
  return wave_sum_f64(sum, comp);
}
__device__ __forceinline__ double cap_winding_f64_wave(F3 q, const float *caps, int ncap, bool &singular) {
  return cap_solid_angle_f64_wave(q, caps, 0, ncap, singular) * (1.0 / (4.0 * M_PI));
}

__device__ __forceinline__ void neumaier_add(double &sum, double &comp, double v) {
  const double t = sum + v;
  comp += (fabs(sum) >= fabs(v)) ? (sum - t) + v : (v - t) + sum;
  sum = t;
}

// Far field of a cap node seen from q in FP64, with cap_far's bounds: returns the error bound (winding
// units), w1 the first-order term (steradians); infinite when q lies within r of c.  The node's fields are
// floats, r and area rounded up; the area vector's rounding (at most 2^-24 |dip| <= 2^-24 area per
// component) moves w1 by less than 1.1e-7 of the zero-order bound, inside the 1e-6 of it added.
__device__ __forceinline__ double cap_far_f64(const CapNode &nd, F3 q, double &w1) {
  const double dx = (double)nd.c[0] - q.x, dy = (double)nd.c[1] - q.y, dz = (double)nd.c[2] - q.z;
  const double d = sqrt(dx * dx + dy * dy + dz * dz);
  const double dmin = d - (double)nd.r;
  w1 = 0.0;
  if (!(dmin > 0.0)) return INFINITY;
  const double b0 = (double)nd.area / (4.0 * M_PI * dmin * dmin);
  const double b1 = b0 * (2.0 * (double)nd.r / dmin);
  if (!(b1 < b0)) return b0 * (1.0 + 1e-12);
  w1 = ((double)nd.dip[0] * dx + (double)nd.dip[1] * dy + (double)nd.dip[2] * dz) / (d * d * d);
  return (b1 + 1e-6 * b0) * (1.0 + 1e-12);
}

// FP64 winding number of the caps at q with a rigorous error bound, by a whole wave, over the cut of the
// cap tree (k_cap_cut: kCapCut slots, each a node, a fan or nothing, partitioning the fans).  Lane l tests
// slots l * kCapCut / 64 ..: a node whose bound (cap_far_f64) is at most kCapTau64 adds its first-order
// term and its bound; the fans of every other slot are summed exactly by the whole wave, slot by slot in
// slot order.  Returns w_C; `bound` is the total error bound in winding units (the sums' FP64 rounding
// included); `singular` as for cap_winding_f64_wave: a fan inside a bounded node lies farther from q than
// the node's diameter (or its solid angle is below area / d^2 < pi), where the solid angle is far from
// singular.  Sums run in a fixed order: per lane, then wave_sum_f64.
constexpr int kCutPerLane = kCapCut / 64;
__device__ __noinline__ double cap_winding_f64_cut(F3 q, const ColumnArgs &a, double &bound, bool &singular) {
  const int lane = threadIdx.x & 63;
  double sum = 0.0, comp = 0.0, ub = 0.0;
  bool sing = false;
  static_assert(kCutPerLane <= 64, "one bit per slot of a lane");
  uint64_t near = 0;  // bit j: slot lane * kCutPerLane + j is summed exactly
  for (int j = 0; j < kCutPerLane; ++j) {
    const int4 e = a.cap_cut[lane * kCutPerLane + j];
    if (e.z == 0) continue;
    if (e.x >= 0) {
      double w1;
      const double b = cap_far_f64(a.cap_nodes[e.x], q, w1);
      if (b <= kCapTau64) {
        ub += b;
        neumaier_add(sum, comp, w1);
        continue;
      }
    }
    near |= 1ull << j;
  }
  for (uint64_t lanes = __ballot(near != 0); lanes; lanes &= lanes - 1) {
    const int l = __builtin_ctzll(lanes);
    for (uint64_t bits = __shfl((unsigned long long)near, l); bits; bits &= bits - 1) {
      const int4 e = a.cap_cut[l * kCutPerLane + __builtin_ctzll(bits)];
#pragma unroll 1
      for (int t = e.y + lane; t < e.y + e.z; t += 64)
        neumaier_add(sum, comp, solid_angle_f64(q, a.caps + 9 * (size_t)t, sing));
    }
  }
  singular = __ballot(sing) != 0;
  bound = wave_sum(ub) * (1.0 + 1e-9) + 1e-10;
  return wave_sum_f64(sum, comp) * (1.0 / (4.0 * M_PI));
}

// cap_winding_f64_cut by the 4 waves of a 256-thread block, for one q: thread t tests slots t, t + 256, ..
// (a q's near slots are neighbours in the cut's Morton order, so they spread over all four waves), each wave
// sums the fans of its threads' near slots, and the four waves' sums, bounds and flags are combined in wave
// order (the result is the same in every thread).  Every thread of the block calls it.
constexpr int kCutPerThread = kCapCut / 256;
static_assert(kCapCut % 256 == 0 && kCutPerThread <= 32, "whole slots per thread, one bit each");
__device__ __noinline__ double cap_winding_f64_cut_block(F3 q, const ColumnArgs &a, double &bound,
                                                         bool &singular) {
  __shared__ double s_sum[4], s_ub[4];
  __shared__ int s_sing[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double sum = 0.0, comp = 0.0, ub = 0.0;
  bool sing = false;
  uint32_t near = 0;  // bit j: slot threadIdx.x + 256 * j is summed exactly
  for (int j = 0; j < kCutPerThread; ++j) {
    const int4 e = a.cap_cut[threadIdx.x + 256 * j];
    if (e.z == 0) continue;
    if (e.x >= 0) {
      double w1;
      const double b = cap_far_f64(a.cap_nodes[e.x], q, w1);
      if (b <= kCapTau64) {
        ub += b;
        neumaier_add(sum, comp, w1);
        continue;
      }
    }
    near |= 1u << j;
  }
  for (uint64_t lanes = __ballot(near != 0); lanes; lanes &= lanes - 1) {
    const int l = __builtin_ctzll(lanes);
    for (uint32_t bits = (uint32_t)__shfl((int)near, l); bits; bits &= bits - 1) {
      const int4 e = a.cap_cut[wv * 64 + l + 256 * __builtin_ctz(bits)];
#pragma unroll 1
      for (int t = e.y + lane; t < e.y + e.z; t += 64)
        neumaier_add(sum, comp, solid_angle_f64(q, a.caps + 9 * (size_t)t, sing));
    }
  }
  const double ws = wave_sum_f64(sum, comp), wub = wave_sum(ub);
  const bool wsing = __ballot(sing) != 0;
  if (lane == 0) {
    s_sum[wv] = ws;
    s_ub[wv] = wub;
    s_sing[wv] = wsing ? 1 : 0;
  }
  __syncthreads();
  double t = 0.0, c = 0.0, u = 0.0;
  bool sg = false;
  for (int k = 0; k < 4; ++k) {
    neumaier_add(t, c, s_sum[k]);
    u += s_ub[k];
    sg = sg || s_sing[k] != 0;
  }
  __syncthreads();  // the next call rewrites the sums
  singular = sg;
  bound = u * (1.0 + 1e-9) + 1e-10;
  return (t + c) * (1.0 / (4.0 * M_PI));
}

// FP64 winding number of the mesh itself at q (no caps), by the whole wave over its ntri triangles: the
// direct form of the closure identity's sum, for a q that lies on a cap (where the identity does not hold)
// (the sum of solid angles of triangles [t0, t1), wave-uniform)
__device__ __noinline__ double mesh_solid_angle_f64_wave(F3 q, const float4 *__restrict__ tris, int t0, int t1) {
  const int lane = threadIdx.x & 63;
  double sum = 0.0, comp = 0.0;
  bool unused = false;
#pragma unroll 1
  for (int i = t0 + lane; i < t1; i += 64) {
    const float4 *t = tris + 3 * i;
    const float v[9] = {t[0].x, t[0].y, t[0].z, t[1].x, t[1].y, t[1].z, t[2].x, t[2].y, t[2].z};
    const double om = solid_angle_f64(q, v, unused);
    const double s = sum + om;
    comp += (fabs(sum) >= fabs(om)) ? (sum - s) + om : (om - s) + sum;
    sum = s;
  }
  return wave_sum_f64(sum, comp);
}
__device__ __forceinline__ double mesh_winding_f64_wave(F3 q, const float4 *__restrict__ tris, int ntri) {
  return mesh_solid_angle_f64_wave(q, tris, 0, ntri) * (1.0 / (4.0 * M_PI));
}

// conservative ray / box slab test for the parity tie-break (boxes inflated by a relative margin)
__device__ __forceinline__ bool ray_box(F3 o, F3 d, const float *lo, const float *hi) {
  float t0 = 0.0f, t1 = INFINITY;
  const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  for (int k = 0; k < 3; ++k) {
    const float m = 1e-5f * (fabsf(lo[k]) + fabsf(hi[k])) + 1e-30f;
    const float l = lo[k] - m, h = hi[k] + m;
    if (dd[k] == 0.0f) {
      if (oo[k] < l || oo[k] > h) return false;
      continue;
    }
    float a = (l - oo[k]) / dd[k], b = (h - oo[k]) / dd[k];
    if (a > b) {
      const float t = a;
      a = b;
      b = t;
    }
    t0 = fmaxf(t0, a * (1.0f - 1e-5f));
    t1 = fminf(t1, b * (1.0f + 1e-5f));
  }
  return t0 <= t1;
}

__device__ __noinline__ bool parity_inside(const Node *__restrict__ nodes, int nnodes, const float4 *__restrict__ tris,
                                           F3 q, Scalars *sc) {
  int votes = 0;
  for (int v = 0; v < 5; ++v) {
    const F3 d = {kParityDirs[3 * v], kParityDirs[3 * v + 1], kParityDirs[3 * v + 2]};
    const Shear sh = make_shear(d.x, d.y, d.z);
    int hits = 0;
    int stack[kStack + 2];
    int sp = 0, cur = 0;
    for (int visits = 0; visits <= nnodes; ++visits) {
      const Node nd = nodes[cur];
      int next = -1;
      if (ray_box(q, d, nd.lo0, nd.hi0)) {
        if (nd.c0 < 0) {
          for_leaf_code(~nd.c0, [&](int l) {
            const float4 *t = tris + 3 * l;
            hits += ray_hit(q, sh, f3(t[0]), f3(t[1]), f3(t[2])) ? 1 : 0;
          });
        } else {
          next = nd.c0;
        }
      }
      if (ray_box(q, d, nd.lo1, nd.hi1)) {
        if (nd.c1 < 0) {
          if (nd.c1 != nd.c0) {
            for_leaf_code(~nd.c1, [&](int l) {
              const float4 *t = tris + 3 * l;
              hits += ray_hit(q, sh, f3(t[0]), f3(t[1]), f3(t[2])) ? 1 : 0;
            });
          }
        } else if (next < 0) {
          next = nd.c1;
        } else if (sp < kStack) {
          stack[sp++] = nd.c1;
        } else {
          stack_overflow(sc);
        }
      }
      if (next >= 0) {
        cur = next;
        continue;
      }
      if (sp == 0) break;
      cur = stack[--sp];
    }
    if (hits & 1) ++votes;
  }
  return votes * 2 > 5;
}

// ============================================================================================
// per-voxel closest triangle + sign.  One wave = one 4x4x4 brick of voxels; a block of 4 waves
// covers an 8x8x4 region.
// ============================================================================================
struct WorkCount {
  uint32_t box = 0, tri = 0, obb = 0;
};


template <bool COUNT, bool SLIVERS>
__device__ __forceinline__ float closest_d2(const Node *__restrict__ nodes, int nnodes, const TriRec *__restrict__ recs,
                                            F3 q, float slack, Scalars *sc, WorkCount &wc) {
  float best = INFINITY, bound = INFINITY;
  int st[kStack];
  float sd[kStack];
  int sp = 0;
  int cur = 0;
  for (int visits = 0; visits <= nnodes; ++visits) {
    const Node nd = nodes[cur];
    const float d0 = box_d2(q, nd.lo0, nd.hi0);
    const float d1 = box_d2(q, nd.lo1, nd.hi1);
    if (COUNT) wc.box += 2;
    if (nd.c0 < 0 && d0 <= bound) {
      for_leaf_code(~nd.c0, [&](int l) {
        best = min_nonneg(best, tri_d2<SLIVERS>(q, recs[l]));
        if (COUNT) wc.tri++;
      });
      bound = prune_bound(best, slack);
    }
    if (nd.c1 < 0 && d1 <= bound) {
      for_leaf_code(~nd.c1, [&](int l) {
        best = min_nonneg(best, tri_d2<SLIVERS>(q, recs[l]));
        if (COUNT) wc.tri++;
      });
      bound = prune_bound(best, slack);
    }
    const bool g0 = nd.c0 >= 0 && d0 <= bound;
    const bool g1 = nd.c1 >= 0 && d1 <= bound;
    if (g0 && g1) {
      const bool first0 = d0 <= d1;
      if (sp < kStack) {
        st[sp] = first0 ? nd.c1 : nd.c0;
        sd[sp] = first0 ? d1 : d0;
        ++sp;
      } else {
        stack_overflow(sc);
      }
      cur = first0 ? nd.c0 : nd.c1;
      continue;
    }
    if (g0) {
      cur = nd.c0;
      continue;
    }
    if (g1) {
      cur = nd.c1;
      continue;
    }
    cur = -1;
    while (sp > 0) {
      --sp;
      if (sd[sp] <= bound) {
        cur = st[sp];
        break;
      }
    }
    if (cur < 0) break;
  }
  return best;
}

template <bool CAPS>
__device__ __forceinline__ void finish_voxel(const SdfArgs &a, F3 q, int x, int y, int z, float best);

// distance slack of a query grid over a mesh (dist_slack), from the root's two child boxes
__device__ __forceinline__ float query_slack(const ColumnArgs &c) {
  const Node root = c.nodes[0];
  const float lo[2][3] = {{root.lo0[0], root.lo0[1], root.lo0[2]}, {root.lo1[0], root.lo1[1], root.lo1[2]}};
  const float hi[2][3] = {{root.hi0[0], root.hi0[1], root.hi0[2]}, {root.hi1[0], root.hi1[1], root.hi1[2]}};
  return dist_slack(c.g.ox, c.g.oy, c.g.oz, c.g.h, c.g.nx, c.g.ny, c.g.nz, lo, hi, 2);
}

template <bool COUNT, bool SLIVERS, bool CAPS>
__global__ __launch_bounds__(256) void k_sdf(SdfArgs a) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int x = blockIdx.x * 8 + (wv & 1) * 4 + (lane & 3);
  const int y = blockIdx.y * 8 + (wv >> 1) * 4 + ((lane >> 2) & 3);
  const int z = blockIdx.z * 4 + (lane >> 4);
  const GridArgs &g = a.col.g;
  if (x >= g.nx || y >= g.ny || z >= g.nz) return;
  const F3 q = {g.ox + ((float)x + 0.5f) * g.h, g.oy + ((float)y + 0.5f) * g.h, g.oz + ((float)z + 0.5f) * g.h};
  WorkCount wc;
  const float best =
      closest_d2<COUNT, SLIVERS>(a.col.nodes, a.col.nnodes, a.col.recs, q, query_slack(a.col), a.col.sc, wc);
  if (COUNT) {
    atomicAdd(&a.col.sc->box_tests, (unsigned long long)wc.box);
    atomicAdd(&a.col.sc->tri_tests, (unsigned long long)wc.tri);
  }
  finish_voxel<CAPS>(a, q, x, y, z, best);
}

// Sign from the winding number and store.  w = (signed crossings above q of the closed surface
// mesh + reversed caps, from the column list) + (solid angle sum of the caps)/4pi.  Voxels whose
// column overflowed the stored crossings, or whose FP32 |w| lies within 0.05 of 0.5, are stored
// unsigned and listed for the resolve kernel.
// output address of voxel (x, y, z): dense index, or slot * 512 + offset inside its 8^3 brick
__device__ __forceinline__ size_t out_index(const SdfArgs &a, int x, int y, int z) {
  const GridArgs &g = a.col.g;
  if (!a.band_index) return ((size_t)z * g.ny + y) * g.nx + x;
  const size_t b = ((size_t)(z >> 3) * (size_t)(g.ny >> 3) + (size_t)(y >> 3)) * (size_t)(g.nx >> 3) + (x >> 3);
  return (size_t)a.band_index[b] * 512 + (size_t)((((z & 7) << 3) | (y & 7)) << 3 | (x & 7));
}

// store the voxel: its sign from the column crossings and, with caps, the caps' winding number wc (within
// cbound); voxels that cannot be decided here are stored unsigned and listed for the resolve kernel
template <bool CAPS>
__device__ __forceinline__ void store_voxel(const SdfArgs &a, F3 q, int x, int y, int z, float best, float wc,
                                            float cbound) {
  const GridArgs &g = a.col.g;
  const size_t vox = ((size_t)z * g.ny + y) * g.nx + x;
  const size_t col = (size_t)y * g.nx + x;
  const float d = sqrtf(best);
  const size_t oi = out_index(a, x, y, z);
  const int cnt = a.col.count[col];
  bool ambiguous = cnt > kColCap;
  bool inside = false;
  if (!ambiguous) {
    const float2 *cr = a.col.cross + col * kColCap;
    int wi = 0;
    for (int k = 0; k < cnt; ++k) {
      const float2 c = cr[k];
      if (c.x > q.z) wi += (int)c.y;
    }
    if (!CAPS) {
      inside = wi != 0;
    } else {
      const float w32 = (float)wi + wc;
      if (fabsf(fabsf(w32) - 0.5f) >= 0.05f + cbound) inside = fabsf(w32) > 0.5f;
      else ambiguous = true;
    }
  }
  if (ambiguous) {
    a.out[oi] = d;
    const uint32_t pos = atomicAdd(&a.col.sc->ambiguous, 1u);
    if (pos < a.amb_cap) a.amb_list[pos] = (uint32_t)vox;  // the host re-bakes on overflow
  } else {
    a.out[oi] = inside ? -d : d;
  }
}

// one voxel on its own (per-voxel kernel): its own walk of the cap tree
template <bool CAPS>
__device__ __forceinline__ void finish_voxel(const SdfArgs &a, F3 q, int x, int y, int z, float best) {
  float wc = 0.0f, cb = 0.0f;
  if (CAPS && a.col.count[(size_t)y * a.col.g.nx + x] <= kColCap)
    cap_winding_bounded(a.col, q, wc, cb);
  store_voxel<CAPS>(a, q, x, y, z, best, wc, cb);
}

// FP64 recheck / parity / overflow recount for the listed voxels: one wave per voxel (grid-stride
// over the count the query kernel left in the scalars).  The crossing count and the tie-break are
// wave-uniform; the FP64 cap sum is split over the lanes; lane 0 writes.
// sign from an FP64 winding number: |w| > 1/2, or the ray-parity vote when |w| is within kTieEps of 1/2
__device__ __forceinline__ bool decide_w64(const SdfArgs &a, double w64, F3 q) {
  if (fabs(fabs(w64) - 0.5) >= kTieEps) return fabs(w64) > 0.5;
  if ((threadIdx.x & 63) == 0) atomicAdd(&a.col.sc->ties, 1u);
  return parity_inside(a.col.nodes, a.col.nnodes, a.col.tris, q, a.col.sc);
}

// the signed count of +z crossings above q on its column (the closure identity's integer)
__device__ __forceinline__ int column_wi(const ColumnArgs &c, int x, int y, F3 q) {
  const size_t col = (size_t)y * c.g.nx + x;
  const int cnt = c.count[col];
  if (cnt > kColCap) return column_direct(c, q.x, q.y, q.z);
  int wi = 0;
  const float2 *cr = c.cross + col * kColCap;
  for (int k = 0; k < cnt; ++k)
    if (cr[k].x > q.z) wi += (int)cr[k].y;
  return wi;
}

__device__ __forceinline__ F3 voxel_centre(const GridArgs &g, uint32_t vox, int &x, int &y, int &z) {
  x = (int)(vox % (uint32_t)g.nx);
  y = (int)((vox / (uint32_t)g.nx) % (uint32_t)g.ny);
  z = (int)(vox / ((uint32_t)g.nx * (uint32_t)g.ny));
  return {g.ox + ((float)x + 0.5f) * g.h, g.oy + ((float)y + 0.5f) * g.h, g.oz + ((float)z + 0.5f) * g.h};
}

__device__ __forceinline__ void store_sign(const SdfArgs &a, int x, int y, int z, bool inside) {
  const size_t oi = out_index(a, x, y, z);
  const float d = a.out[oi];
  a.out[oi] = inside ? -d : d;
}

// FP64 sums handed to the chunked launches (k_singular_parts / k_singular_finish): kind 0, a voxel on a
// hole cap (the mesh's own winding number over its triangles); kind 1, a voxel the cut's bound left
// undecided (the column's integer plus every cap fan).  Returns false when the list is full.
__device__ __forceinline__ bool list_for_chunks(const SdfArgs &a, uint32_t vox, uint32_t kind) {
  uint32_t slot = 0;
  if ((threadIdx.x & 63) == 0) slot = atomicAdd(&a.col.sc->singular, 1u);
  slot = __shfl(slot, 0);
  if (slot >= a.sing_cap) return false;
  if ((threadIdx.x & 63) == 0) {
    a.sing[slot] = vox;
    a.sing_kind[slot] = kind;
  }
  return true;
}

// The voxels the query left ambiguous, by 256-thread blocks (grid-stride over blk).  Without hole caps the
// column's crossing count is the sign, one voxel per wave.  With caps, one voxel per block: the four waves
// test the cap tree's cut together (cap_winding_f64_cut_block) and the rest, rare, is wave 0's.
__device__ __forceinline__ void resolve_range_block(const SdfArgs &a, uint32_t blk, uint32_t nblk) {
  const uint32_t n = min(a.col.sc->ambiguous, a.amb_cap);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (a.col.ncap == 0) {
    for (uint32_t i = blk * 4 + wv; i < n; i += nblk * 4) {
      const uint32_t vox = a.amb_list[i];
      int x, y, z;
      const F3 q = voxel_centre(a.col.g, vox, x, y, z);
      const int wi = column_wi(a.col, x, y, q);
      if (lane == 0) store_sign(a, x, y, z, wi != 0);
    }
    return;
  }
  for (uint32_t i = blk; i < n; i += nblk) {  // the same voxels in every wave of the block
    const uint32_t vox = a.amb_list[i];
    int x, y, z;
    const F3 q = voxel_centre(a.col.g, vox, x, y, z);
    const int wi = column_wi(a.col, x, y, q);
    if (threadIdx.x == 0) atomicAdd(&a.col.sc->rechecked, 1u);
    // the cap tree's cut with its bound first: it decides when |w| stays on one side of 1/2 and of the
    // tie band's edges by more than the bound, and it finds every voxel on a cap; the others sum every
    // fan, in chunks across the chip (or here, when the chunk list is full or IB_OPT_RESOLVE_FULL is set)
    bool on_cap = false, decided = false;
    double w64 = 0.0;
    if (!a.resolve_full) {
      double bound;
      w64 = (double)wi + cap_winding_f64_cut_block(q, a.col, bound, on_cap);
      const double m = fabs(fabs(w64) - 0.5);
      decided = on_cap || m >= kTieEps + bound || m <= kTieEps - bound;
    }
    if (wv != 0) continue;
    if (!decided) {
      if (lane == 0) atomicAdd(&a.col.sc->rechecked_full, 1u);
      if (!a.resolve_full && list_for_chunks(a, vox, 1u)) continue;
      w64 = (double)wi + cap_winding_f64_wave(q, a.col.caps, a.col.ncap, on_cap);
    }
    if (on_cap) {
      if (lane == 0) atomicAdd(&a.col.sc->on_cap, 1u);
      if (list_for_chunks(a, vox, 0u)) continue;
      w64 = mesh_winding_f64_wave(q, a.col.tris, a.col.ntri);
    }
    const bool inside = decide_w64(a, w64, q);
    if (lane == 0) store_sign(a, x, y, z, inside);
  }
}

// resolve_range_block's work with one voxel per wave (wave = the global wave index): the form for batches
// of several jobs, whose resolve launches share the chip with each other and with other streams' work
__device__ __forceinline__ void resolve_range_wave(const SdfArgs &a, uint32_t wave, uint32_t nwaves) {
  const uint32_t n = min(a.col.sc->ambiguous, a.amb_cap);
  const int lane = threadIdx.x & 63;
  for (uint32_t i = wave; i < n; i += nwaves) {
    const uint32_t vox = a.amb_list[i];
    int x, y, z;
    const F3 q = voxel_centre(a.col.g, vox, x, y, z);
    const int wi = column_wi(a.col, x, y, q);
    bool inside;
    if (a.col.ncap == 0) {
      inside = wi != 0;
    } else {
      if (lane == 0) atomicAdd(&a.col.sc->rechecked, 1u);
      // the cap tree's cut with its bound first: it decides when |w| stays on one side of 1/2 and of the
      // tie band's edges by more than the bound, and it finds every voxel on a cap; the others sum every
      // fan, in chunks across the chip (or here, when the chunk list is full or IB_OPT_RESOLVE_FULL is set)
      bool on_cap = false, decided = false;
      double w64 = 0.0;
      if (!a.resolve_full) {
        double bound;
        w64 = (double)wi + cap_winding_f64_cut(q, a.col, bound, on_cap);
        const double m = fabs(fabs(w64) - 0.5);
        decided = on_cap || m >= kTieEps + bound || m <= kTieEps - bound;
      }
      if (!decided) {
        if (lane == 0) atomicAdd(&a.col.sc->rechecked_full, 1u);
        if (!a.resolve_full && list_for_chunks(a, vox, 1u)) continue;
        w64 = (double)wi + cap_winding_f64_wave(q, a.col.caps, a.col.ncap, on_cap);
      }
      if (on_cap) {
        if (lane == 0) atomicAdd(&a.col.sc->on_cap, 1u);
        if (list_for_chunks(a, vox, 0u)) continue;
        w64 = mesh_winding_f64_wave(q, a.col.tris, a.col.ntri);
      }
      inside = decide_w64(a, w64, q);
    }
    if (lane == 0) store_sign(a, x, y, z, inside);
  }
}

// The FP64 sums listed by resolve_range_block / resolve_range_wave, over chunks of kSingChunk triangles
// (kind 0) or cap fans (kind 1): one (voxel, chunk) item per wave (grid-stride), then per voxel the chunks'
// sum in chunk order and the decision.  The partials depend only on the item, not on the wave that sums it.
__device__ __forceinline__ void singular_parts(const SdfArgs &a, uint32_t wave, uint32_t nwaves) {
  const uint32_t ns = min(a.col.sc->singular, a.sing_cap), nc = a.sing_chunks;
  const uint32_t items = ns * nc;
  for (uint32_t it = wave; it < items; it += nwaves) {
    const uint32_t slot = it / nc, chunk = it % nc;
    int x, y, z;
    const F3 q = voxel_centre(a.col.g, a.sing[slot], x, y, z);
    const int t0 = (int)(chunk * kSingChunk);
    double part;
    if (a.sing_kind[slot] == 0) {
      part = mesh_solid_angle_f64_wave(q, a.col.tris, t0, min(a.col.ntri, t0 + kSingChunk));
    } else {
      bool unused = false;
      part = cap_solid_angle_f64_wave(q, a.col.caps, t0, min(a.col.ncap, t0 + kSingChunk), unused);
    }
    if ((threadIdx.x & 63) == 0) a.sing_part[(size_t)slot * nc + chunk] = part;
  }
}

__device__ __forceinline__ void singular_finish(const SdfArgs &a, uint32_t wave, uint32_t nwaves) {
  const uint32_t ns = min(a.col.sc->singular, a.sing_cap), nc = a.sing_chunks;
  const int lane = threadIdx.x & 63;
  for (uint32_t slot = wave; slot < ns; slot += nwaves) {
    int x, y, z;
    const F3 q = voxel_centre(a.col.g, a.sing[slot], x, y, z);
    double sum = 0.0, comp = 0.0;
    for (uint32_t k = lane; k < nc; k += 64) neumaier_add(sum, comp, a.sing_part[(size_t)slot * nc + k]);
    double w64 = wave_sum_f64(sum, comp) * (1.0 / (4.0 * M_PI));
    if (a.sing_kind[slot] != 0) w64 += (double)column_wi(a.col, x, y, q);
    const bool inside = decide_w64(a, w64, q);
    if (lane == 0) store_sign(a, x, y, z, inside);
  }
}

__global__ __launch_bounds__(256, 4) void k_resolve(SdfArgs a) {
  resolve_range_block(a, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(256, 4) void k_singular_parts(SdfArgs a) {
  singular_parts(a, blockIdx.x * 4 + (threadIdx.x >> 6), gridDim.x * 4);
}
__global__ __launch_bounds__(256) void k_singular_finish(SdfArgs a) {
  singular_finish(a, blockIdx.x * 4 + (threadIdx.x >> 6), gridDim.x * 4);
}

// --------------------------------------------------------------------------------------------
// Wave-coherent variant: the 64 lanes of a wave (one 4x4x4 brick) walk the tree together.  A
// child is entered if ANY lane's closest-so-far distance reaches its box; node and triangle
// records are loaded once per wave (uniform addresses -> scalar loads) and the traversal stack is
// wave-uniform in LDS.  Every lane evaluates every triangle the brick visits, so the per-voxel
// minimum is the same as the per-lane traversal's (an exact min over a superset of candidates).
// --------------------------------------------------------------------------------------------
// One wave = one brick of sx * sy * sz = 64 voxels with corner (bx, by, bz); lane l is voxel
// (l % sx, (l / sx) % sy, l / (sx * sy)) of the brick.  sstack is the wave's LDS stack.
// Node and triangle records read at wave-uniform addresses go through the constant address space,
// so they become scalar loads (s_load) into SGPRs instead of per-lane vector loads.  The records are
// read as plain 4-float vectors (no constructors, valid in either address space).
typedef float V4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) V4 CV4;

__device__ __forceinline__ Node load_node_uniform(const Node *p, int i) {
  const CV4 *r = (const CV4 *)(p + i);
  const V4 a = r[0], b = r[1], c = r[2], d = r[3];
  Node n;
  n.lo0[0] = a.x; n.lo0[1] = a.y; n.lo0[2] = a.z; n.hi0[0] = a.w;
  n.hi0[1] = b.x; n.hi0[2] = b.y; n.lo1[0] = b.z; n.lo1[1] = b.w;
  n.lo1[2] = c.x; n.hi1[0] = c.y; n.hi1[1] = c.z; n.hi1[2] = c.w;
  n.c0 = __float_as_int(d.x); n.c1 = __float_as_int(d.y); n.parent = __float_as_int(d.z); n.pad = 0;
  return n;
}

union Node4Bits {
  V4 v[8];
  Node4 r;
};

__device__ __forceinline__ Node4 load_node4_uniform(const Node4 *p, int i) {
  const CV4 *r = (const CV4 *)(p + i);
  Node4Bits u;
#pragma unroll
  for (int k = 0; k < 8; ++k) u.v[k] = r[k];
  return u.r;
}

// only the children, count and oriented-box mask of node i (its last 32 bytes); the boxes stay unread
__device__ __forceinline__ Node4 load_node4_links_uniform(const Node4 *p, int i) {
  const CV4 *r = (const CV4 *)(p + i);
  Node4Bits u;
  u.v[6] = r[6];
  u.v[7] = r[7];
  return u.r;
}

union ObbBits {
  V4 v[4];
  Obb r;
};

__device__ __forceinline__ Obb load_obb_uniform(const Obb *p, int i) {
  const CV4 *r = (const CV4 *)(p + i);
  ObbBits u;
#pragma unroll
  for (int k = 0; k < 4; ++k) u.v[k] = r[k];
  return u.r;
}

union CapNodeBits {
  V4 v[4];
  CapNode r;
};
__device__ __forceinline__ CapNode load_capnode_uniform(const CapNode *p, int i) {
  const CV4 *r = (const CV4 *)(p + i);
  CapNodeBits u;
#pragma unroll
  for (int k = 0; k < 4; ++k) u.v[k] = r[k];
  return u.r;
}

// The caps' winding number at every needed voxel of a wave's brick (V per lane), with cap_winding_bounded's
// bound, by one walk of the cap tree for the whole wave: a node is not entered when its far field (cap_far)
// is within the thresholds (kCapTau, kCapTauHi) for every needed voxel, else entered by all; fans are read once
// (scalar loads) and evaluated by every lane.  The stack is the wave's LDS stack, free after the
// closest-triangle walk.
template <int V>
__device__ __forceinline__ void cap_winding_wave(const ColumnArgs &a, const F3 *q, const bool *need, float *w,
                                                 float *bound, int *sstack) {
  const float inv4pi = (float)(1.0 / (4.0 * M_PI));
  float sum[V], cmp[V], ub[V];
  bool singular[V];
  bool any = false;
  for (int k = 0; k < V; ++k) {
    sum[k] = cmp[k] = ub[k] = 0.0f;
    singular[k] = false;
    w[k] = bound[k] = 0.0f;
    any = any || need[k];
  }
  // a batch runs the capped kernel for every job once one job has caps: closed meshes have none
  if (a.ncap == 0 || a.ncapnode == 0 || !a.cap_nodes || __ballot(any) == 0) return;
  const int lane = threadIdx.x & 63;
  typedef const __attribute__((address_space(4))) float CF;
  auto fan = [&](int i) {
    CF *p = (CF *)(a.caps + 9 * (size_t)i);
    float t[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) t[j] = p[j];
    for (int k = 0; k < V; ++k) {
      const float om = solid_angle_f32(q[k], t, singular[k]);
      const float s = sum[k] + om;
      cmp[k] += (fabsf(sum[k]) >= fabsf(om)) ? (sum[k] - s) + om : (om - s) + sum[k];
      sum[k] = s;
    }
  };
  int sp = 0, cur = 0;
  bool overflow = false;
  for (;;) {
    if (cur < 0) {
      fan(~cur);
    } else {
      const CapNode nd = load_capnode_uniform(a.cap_nodes, cur);
      float b[V], w1[V], dx[V], dy[V], dz[V], d2[V];
      // a needed voxel inside the node's opening radius opens it: no far field to evaluate
      bool near = false;
      for (int k = 0; k < V; ++k) {
        d2[k] = cap_d2(nd, q[k], dx[k], dy[k], dz[k]);
        near = near || (need[k] && d2[k] < nd.open2);
      }
      bool open = __ballot(near) != 0;
      if (!open) {
        for (int k = 0; k < V; ++k) {
          b[k] = cap_far_at(nd, dx[k], dy[k], dz[k], d2[k], w1[k]);
          open = open || (need[k] && !(b[k] <= kCapTau || (b[k] <= kCapTauHi && ub[k] + b[k] <= kCapBudget)));
        }
      }
      if (__ballot(open) == 0) {
        for (int k = 0; k < V; ++k) {
          ub[k] += b[k];
          const float s = sum[k] + w1[k];
          cmp[k] += (fabsf(sum[k]) >= fabsf(w1[k])) ? (sum[k] - s) + w1[k] : (w1[k] - s) + sum[k];
          sum[k] = s;
        }
      } else if (nd.left < 0 && nd.right < 0) {
        const int first = -1 - nd.left;
        for (int i = 0; i < -nd.right; ++i) fan(first + i);
      } else if (sp < kCapStack) {
        if (lane == 0) sstack[sp] = nd.right;
        ++sp;
        cur = nd.left;
        continue;
      } else {
        overflow = true;  // cannot happen within the depth bound; the voxels go to the FP64 recheck
      }
    }
    if (sp == 0) break;
    --sp;
    __builtin_amdgcn_wave_barrier();
    cur = __builtin_amdgcn_readfirstlane(sstack[sp]);
  }
  if (overflow && lane == 0) stack_overflow(a.sc);
  for (int k = 0; k < V; ++k) {
    w[k] = (sum[k] + cmp[k]) * inv4pi;
    bound[k] = (singular[k] || overflow) ? INFINITY : ub[k] * 1.0001f + 1e-5f;
  }
}

union TriRecBits {
  V4 v[6];
  TriRec r;
};

// Column crossings by rasterising the triangles: every mesh triangle and every cap fan (reversed) tests the
// column centres inside its xy box (one more column on each side for rounding) with the crossing test of the
// column walk, and appends each crossing to its column's list (an atomic slot; the list is a set: every use
// sums the crossings above a point, in any order).  A triangle over more than 16 columns is shared by the
// wave's lanes.  Counts start at zero.
__device__ __forceinline__ void column_append(const ColumnArgs &a, int x, int y, F3 p0, F3 p1, F3 p2, int sign) {
  const float X = a.g.ox + ((float)x + 0.5f) * a.g.h;
  const float Y = a.g.oy + ((float)y + 0.5f) * a.g.h;
  if (!(X >= fminf(fminf(p0.x, p1.x), p2.x) && X <= fmaxf(fmaxf(p0.x, p1.x), p2.x) &&
        Y >= fminf(fminf(p0.y, p1.y), p2.y) && Y <= fmaxf(fmaxf(p0.y, p1.y), p2.y)))
    return;
  float z;
  const int sg = column_cross(X, Y, p0, p1, p2, &z);
  if (!sg) return;
  const size_t col = (size_t)y * a.g.nx + x;
  const int k = atomicAdd(&a.count[col], 1);
  if (k < kColCap) a.cross[col * kColCap + k] = make_float2(z, (float)(sg * sign));
  else if (k == kColCap) atomicAdd(&a.sc->col_overflow, 1u);
}

// first and last column index whose centre can lie in [lo, hi] along an axis of origin o, spacing h, n columns
__device__ __forceinline__ void column_span(float lo, float hi, float o, float h, int n, int &a, int &b) {
  const float fa = fminf(fmaxf(floorf((lo - o) / h - 0.5f), -2.0f), (float)n + 2.0f);
  const float fb = fminf(fmaxf(ceilf((hi - o) / h - 0.5f), -2.0f), (float)n + 2.0f);
  a = max((int)fa - 1, 0);
  b = min((int)fb + 1, n - 1);
}

// triangle or fan i (i >= ntri: fan i - ntri) of the job, one per lane
__device__ __forceinline__ void column_splat(const ColumnArgs &a, int i) {
  const int lane = threadIdx.x & 63;
  const bool have = i < a.ntri + a.ncap;
  F3 p0 = {0.0f, 0.0f, 0.0f}, p1 = p0, p2 = p0;
  int sign = 1;
  if (have) {
    if (i < a.ntri) {
      p0 = f3(a.tris[3 * (size_t)i]);
      p1 = f3(a.tris[3 * (size_t)i + 1]);
      p2 = f3(a.tris[3 * (size_t)i + 2]);
    } else {
      const float *t = a.caps + 9 * (size_t)(i - a.ntri);
      p0 = {t[0], t[1], t[2]};
      p1 = {t[3], t[4], t[5]};
      p2 = {t[6], t[7], t[8]};
      sign = -1;  // caps enter the closed surface reversed
    }
  }
  int x0, x1, y0, y1;
  column_span(fminf(fminf(p0.x, p1.x), p2.x), fmaxf(fmaxf(p0.x, p1.x), p2.x), a.g.ox, a.g.h, a.g.nx, x0, x1);
  column_span(fminf(fminf(p0.y, p1.y), p2.y), fmaxf(fmaxf(p0.y, p1.y), p2.y), a.g.oy, a.g.h, a.g.ny, y0, y1);
  const int w = x1 - x0 + 1, hgt = y1 - y0 + 1;
  const int ncols = (have && w > 0 && hgt > 0) ? w * hgt : 0;
  const bool big = ncols > 16;
  if (!big)
    for (int c = 0; c < ncols; ++c) column_append(a, x0 + c % w, y0 + c / w, p0, p1, p2, sign);
  unsigned long long m = __ballot(big);
  while (m) {
    const int b = __ffsll((long long)m) - 1;
    m &= m - 1;
    const F3 q0 = {__shfl(p0.x, b), __shfl(p0.y, b), __shfl(p0.z, b)};
    const F3 q1 = {__shfl(p1.x, b), __shfl(p1.y, b), __shfl(p1.z, b)};
    const F3 q2 = {__shfl(p2.x, b), __shfl(p2.y, b), __shfl(p2.z, b)};
    const int bx = __shfl(x0, b), by = __shfl(y0, b), bw = __shfl(w, b), bn = __shfl(ncols, b), bs = __shfl(sign, b);
    for (int c = lane; c < bn; c += 64) column_append(a, bx + c % bw, by + c / bw, q0, q1, q2, bs);
  }
}

__global__ __launch_bounds__(256) void k_columns(ColumnArgs a) { column_splat(a, blockIdx.x * 256 + threadIdx.x); }

// the columns of every job of a batch in one launch: block b rasterises 256 triangles and fans of the job
// whose range of the block prefix holds b
__global__ __launch_bounds__(256) void k_columns_batch(const SdfArgs *__restrict__ jobs,
                                                       const uint32_t *__restrict__ prefix, uint32_t njobs) {
  const uint32_t b = blockIdx.x;
  int lo = 0, hi = (int)njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid; else hi = mid - 1;
  }
  column_splat(jobs[lo].col, (int)(b - prefix[lo]) * 256 + threadIdx.x);
}

// zero every job's column counts (blockIdx.y = job)
__global__ __launch_bounds__(256) void k_zero_counts_batch(const SdfArgs *__restrict__ jobs) {
  const ColumnArgs &a = jobs[blockIdx.y].col;
  const size_t ncol = (size_t)a.g.nx * a.g.ny;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < ncol; i += (size_t)gridDim.x * 256) a.count[i] = 0;
}




__device__ __forceinline__ TriRec load_trirec_uniform(const TriRec *p, int i) {
  const CV4 *r = (const CV4 *)(p + i);
  TriRecBits u;
#pragma unroll
  for (int k = 0; k < 6; ++k) u.v[k] = r[k];
  return u.r;
}

// Squared distance from a voxel to box e of the visited node, from the terms that lanes e * 16 + j hold
// (see wave_brick): the x-y term at byte address axy and the z term at byte address az of entry 0.
__device__ __forceinline__ float sep_box_d2(int axy, int az, int e, float fxy, float fz) {
  const float dxy = __int_as_float(__builtin_amdgcn_ds_bpermute(axy + 64 * e, __float_as_int(fxy)));
  const float dz = __int_as_float(__builtin_amdgcn_ds_bpermute(az + 64 * e, __float_as_int(fz)));
  return dxy + dz;
}

// One wave = one brick of sx * sy * sz voxels with corner (bx, by, bz), V voxels per lane: lane l
// holds voxels (l % sx, (l / sx) % sy, l / (sx * sy) + k * sz / V) for k < V.  sstack is the wave's
// LDS stack.  Every lane's voxels are queried against every triangle the brick visits.
// SEP (4 x 4 x sz bricks, sz <= 16): the squared distance from a voxel to an axis-aligned box is a sum of
// per-axis terms, and the brick's voxels take only 16 distinct (x, y) pairs and sz distinct z coordinates.
// At every node, lane e * 16 + j loads box e (a vector load) and evaluates its x-y term at the brick's
// j-th (x, y) pair and its z term at the j-th z coordinate; each lane then gathers its voxels' two terms
// per entry (ds_bpermute) and adds them.  One set of VALU instructions evaluates the terms of all four
// entries, where testing entry by entry repeats them per entry.
template <bool COUNT, bool SLIVERS, bool CAPS, int V, bool SEP>
__device__ __forceinline__ void wave_brick(const SdfArgs &a, int bx, int by, int bz, int sx, int sy, int *sstack) {
  const int lane = threadIdx.x & 63;
  const GridArgs &g = a.col.g;
  if (bx >= g.nx || by >= g.ny || bz >= g.nz) return;  // whole brick outside (wave-uniform)
  const int sz = 64 * V / (sx * sy), lz = sz / V;
  const int x = bx + lane % sx, y = by + (lane / sx) % sy;
  int z[V];
  bool valid[V];
  F3 q[V];
  float best[V], bound[V];
  for (int k = 0; k < V; ++k) {
    z[k] = bz + lane / (sx * sy) + k * lz;
    valid[k] = x < g.nx && y < g.ny && z[k] < g.nz;
    // lanes past the grid edge shadow the brick's first voxel so they never widen the search
    const int xs = valid[k] ? x : bx, ys = valid[k] ? y : by, zs = valid[k] ? z[k] : bz;
    q[k] = {g.ox + ((float)xs + 0.5f) * g.h, g.oy + ((float)ys + 0.5f) * g.h, g.oz + ((float)zs + 0.5f) * g.h};
    best[k] = INFINITY;
    bound[k] = INFINITY;
  }
  const float slack = query_slack(a.col), slack_obb = 1.25f * slack;
  float bound_obb[V];
  for (int k = 0; k < V; ++k) bound_obb[k] = INFINITY;
  const Obb *__restrict__ obbs = a.col.obbs;
  // the lane of voxel (sx/2, sy/2, lz/2) orders the children: its two box distances (first voxel)
  // are read with v_readlane and compared as integers (non-negative floats order like their bits)
  const int centre = __builtin_amdgcn_readfirstlane((sx >> 1) + (sy >> 1) * sx + (lz >> 1) * sx * sy);
  const Node4 *__restrict__ nodes4 = a.col.nodes4;
  const TriRec *__restrict__ recs = a.col.recs;
  // SEP roles: this lane's entry (lane >> 4), the coordinates of its terms (the brick's (x, y) pair and z
  // coordinate number lane & 15) and, per voxel, the byte addresses of the lanes holding that voxel's
  // terms of entry 0 (a voxel shadowing the first one reads the first pair and coordinate)
  int fbox = 0;
  float fcx = 0.0f, fcy = 0.0f, fcz = 0.0f;
  int gxy[V], gz[V];
  if (SEP) {
    const int fj = lane & 15;
    fbox = 3 * (lane >> 4);
    fcx = g.ox + ((float)(bx + (fj & 3)) + 0.5f) * g.h;
    fcy = g.oy + ((float)(by + (fj >> 2)) + 0.5f) * g.h;
    fcz = g.oz + ((float)(bz + fj) + 0.5f) * g.h;
    for (int k = 0; k < V; ++k) {
      gxy[k] = valid[k] ? 4 * (lane & 15) : 0;
      gz[k] = valid[k] ? 4 * (z[k] - bz) : 0;
    }
  }
  WorkCount wc;
  int sp = 0;
  int cur = 0;
  auto eval = [&](int l) {
    const TriRec t = load_trirec_uniform(recs, l);
    for (int k = 0; k < V; ++k) best[k] = min_nonneg(best[k], tri_d2<SLIVERS>(q[k], t));
    if (COUNT) wc.tri += V;
  };
  // compare-exchange of (key, child) pairs, ascending by key (wave-uniform values)
  auto cx = [](uint32_t &ka, int &ca, uint32_t &kb, int &cb) {
    if (kb < ka) {
      const uint32_t tk = ka;
      ka = kb;
      kb = tk;
      const int tc = ca;
      ca = cb;
      cb = tc;
    }
  };
  // SEP: this lane's bounds of the node visited next, and that node's links, loaded one visit ahead so
  // that the loads overlap the leaf blocks evaluated at the current node
  float plo[3] = {0.0f, 0.0f, 0.0f}, phi[3] = {0.0f, 0.0f, 0.0f};
  Node4 nd_next;
  auto fetch = [&](int node) {
    const float *nf = reinterpret_cast<const float *>(nodes4 + node) + fbox;
    for (int i = 0; i < 3; ++i) plo[i] = nf[i], phi[i] = nf[12 + i];
    nd_next = load_node4_links_uniform(nodes4, node);
  };
  if (SEP) fetch(0);
  for (int visits = 0; visits <= a.col.nnodes; ++visits) {
    const Node4 nd = SEP ? nd_next : load_node4_uniform(nodes4, cur);
    // SEP: the distances to all four boxes at once, so that the gathers overlap (an unused entry's
    // are not read); every lane takes part in them
    float dsep[4][V];
    if (SEP) {
      const float tx = fmaxf(fmaxf(plo[0] - fcx, fcx - phi[0]), 0.0f);
      const float ty = fmaxf(fmaxf(plo[1] - fcy, fcy - phi[1]), 0.0f);
      const float tz = fmaxf(fmaxf(plo[2] - fcz, fcz - phi[2]), 0.0f);
      const float fxy = fmaf(ty, ty, tx * tx), fz = tz * tz;
#pragma unroll
      for (int e = 0; e < 4; ++e)
        for (int k = 0; k < V; ++k) dsep[e][k] = sep_box_d2(gxy[k], gz[k], e, fxy, fz);
    }
    // per entry: does any lane reach its box, and the centre lane's distance to it (ordering key)
    uint32_t key[4];
    int ch[4];
    int lcode[4];  // leaf-block code of a leaf entry whose box some lane reaches, else -1
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      key[e] = 0xffffffffu;
      ch[e] = -1;
      lcode[e] = -1;
      if (e >= nd.count) continue;  // wave-uniform: unused entries cost no box test
      const float d = SEP ? dsep[e][0] : box_d2(q[0], nd.lo[e], nd.hi[e]);
      bool need = d <= bound[0];
      for (int k = 1; k < V; ++k)
        need = need || (SEP ? dsep[e][k] : box_d2(q[k], nd.lo[e], nd.hi[e])) <= bound[k];
      if (__ballot(need) == 0) continue;
      const int c = nd.child[e];
      if ((nd.obb_mask >> e) & 1) {  // a small subtree: its oriented box must reach some lane too
        const Obb b = load_obb_uniform(a.col.node_obbs, c);
        bool near = false;
        for (int k = 0; k < V; ++k) near = near || obb_d2(q[k], b) <= bound_obb[k];
        if (COUNT) wc.obb += V;
        if (__ballot(near) == 0) continue;
      }
      if (c < 0) {
        lcode[e] = ~c;
      } else {
        ch[e] = c;
        key[e] = (uint32_t)__builtin_amdgcn_readlane(__float_as_int(d), centre);
      }
    }
    if (COUNT) wc.box += nd.count * V;
    // nearest internal entry next, the others pushed farthest first (sorting network on 4)
    cx(key[0], ch[0], key[1], ch[1]);
    cx(key[2], ch[2], key[3], ch[3]);
    cx(key[0], ch[0], key[2], ch[2]);
    cx(key[1], ch[1], key[3], ch[3]);
    cx(key[1], ch[1], key[2], ch[2]);
#pragma unroll
    for (int e = 3; e >= 1; --e)
      if (ch[e] >= 0) {
        if (sp < kStack4) {
          if (lane == 0) sstack[sp] = ch[e];
          ++sp;
        } else {
          stack_overflow(a.col.sc);
        }
      }
    // the node visited next (-1: none), chosen before the leaf blocks are evaluated
    int next = ch[0];
    if (next < 0 && sp > 0) {
      --sp;
      __builtin_amdgcn_wave_barrier();
      next = __builtin_amdgcn_readfirstlane(sstack[sp]);
    }
    if (SEP && next >= 0) fetch(next);
    bool evaluated = false;
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (lcode[e] >= 0) {
        const int code = lcode[e];
        if (code >> kBlockShift) {  // a block of 2+ triangles: its oriented box must reach some lane too
          const Obb b = load_obb_uniform(obbs, code & kBlockMask);
          bool need = false;
          for (int k = 0; k < V; ++k) need = need || obb_d2(q[k], b) <= bound_obb[k];
          if (COUNT) wc.obb += V;
          if (__ballot(need) == 0) continue;
        }
        for_leaf_code(code, eval);
        evaluated = true;
        for (int k = 0; k < V; ++k) bound_obb[k] = prune_bound(best[k], slack_obb);
      }
    if (evaluated)
      for (int k = 0; k < V; ++k) bound[k] = prune_bound(best[k], slack);
    if (next < 0) break;
    cur = next;
  }
  if (COUNT) {
    int nv = 0;
    for (int k = 0; k < V; ++k) nv += valid[k] ? 1 : 0;
    if (nv) {
      atomicAdd(&a.col.sc->box_tests, (unsigned long long)wc.box * nv / V);
      atomicAdd(&a.col.sc->obb_tests, (unsigned long long)wc.obb * nv / V);
      atomicAdd(&a.col.sc->tri_tests, (unsigned long long)wc.tri * nv / V);
    }
  }
  float capw[V], capb[V];
  for (int k = 0; k < V; ++k) capw[k] = capb[k] = 0.0f;
  if (CAPS) {
    bool need[V];
    for (int k = 0; k < V; ++k) need[k] = valid[k] && a.col.count[(size_t)y * g.nx + x] <= kColCap;
    cap_winding_wave<V>(a.col, q, need, capw, capb, sstack);
  }
  for (int k = 0; k < V; ++k)
    if (valid[k]) store_voxel<CAPS>(a, q[k], x, y, z[k], best[k], capw[k], capb[k]);
}

// brick shapes a wave can take (sx, sy, voxels per lane; sz = 64 * V / (sx * sy)); the 4 waves of
// a block tile 2 x 2
constexpr int kNumShapes = 5;
__device__ __constant__ const int kShapes[kNumShapes][3] = {{4, 4, 1}, {8, 4, 1}, {8, 8, 1}, {16, 4, 1}, {4, 4, 2}};

// One block per block-tile (2 x 2 bricks); `order`, when given, lists the tiles in launch order
// (most expensive first, from k_tile_keys), so the long-running tiles do not form the tail.
template <bool COUNT, bool SLIVERS, bool CAPS, int V, bool SEP>
__global__ __launch_bounds__(256, 8) void k_sdf_wave(SdfArgs a, int shape, const uint32_t *__restrict__ order,
                                                  uint32_t tx_n, uint32_t ty_n) {
  __shared__ int sstack[4][kStack4];
  const int wv = threadIdx.x >> 6;
  const int sx = kShapes[shape][0], sy = kShapes[shape][1], sz = 64 * V / (sx * sy);
  const uint32_t tile = order ? order[blockIdx.x] : blockIdx.x;
  const int bx = (int)(tile % tx_n) * 2 * sx + (wv & 1) * sx;
  const int by = (int)((tile / tx_n) % ty_n) * 2 * sy + (wv >> 1) * sy;
  const int bz = (int)(tile / (tx_n * ty_n)) * sz;
  wave_brick<COUNT, SLIVERS, CAPS, V, SEP>(a, bx, by, bz, sx, sy, sstack[wv]);
}

// Launch-order key of every block-tile: an estimate of its query cost from the column crossings
// of the tile's centre column.  u = vertical distance from the tile centre to the nearest crossing
// (an upper bound of the distance there); voxels inside the surface see it in every direction and
// cost most, so their key is 4u.  A column that misses the mesh gets half the grid height, an
// overflowed column +inf.  mode 1: sort key ~bits(key), so an ascending sort puts the costliest
// tile first; mode 2: (31 - half-octave bucket of key / h) << 27 | Morton code of the tile.
__device__ __forceinline__ uint32_t tile_key(const ColumnArgs &a, int shape, uint32_t t, int mode) {
  const GridArgs &g = a.g;
  const uint32_t tx_n = (uint32_t)(g.nx + 2 * kShapes[shape][0] - 1) / (2 * kShapes[shape][0]);
  const uint32_t ty_n = (uint32_t)(g.ny + 2 * kShapes[shape][1] - 1) / (2 * kShapes[shape][1]);
  const int sx = kShapes[shape][0], sy = kShapes[shape][1], sz = 64 * kShapes[shape][2] / (sx * sy);
  const int xc = min((int)(t % tx_n) * 2 * sx + sx, g.nx - 1);
  const int yc = min((int)((t / tx_n) % ty_n) * 2 * sy + sy, g.ny - 1);
  const float zc = g.oz + ((float)((int)(t / (tx_n * ty_n)) * sz) + 0.5f * (float)sz) * g.h;
  const size_t col = (size_t)yc * g.nx + xc;
  const int cnt = a.count[col];
  float key;
  if (cnt > kColCap) {
    key = INFINITY;
  } else if (cnt == 0) {
    key = 0.5f * (float)g.nz * g.h;
  } else {
    float u = INFINITY;
    int w = 0;
    for (int k = 0; k < cnt; ++k) {
      const float2 c = a.cross[col * kColCap + k];
      u = fminf(u, fabsf(c.x - zc));
      if (c.x > zc) w += (int)c.y;
    }
    key = w != 0 ? 4.0f * u : u;
  }
  if (mode == 1) return ~__float_as_uint(key);
  // mode 2: half-octave cost buckets (costliest first), Morton order of the tile inside a bucket so
  // that blocks launched together work on neighbouring tiles and share cached tree records
  const float c = key / g.h;
  const int bucket = !(c < 1e30f) ? 31 : (c <= 1.0f ? 0 : min(31, (int)(2.0f * __log2f(c)) + 1));
  const uint32_t tx = t % tx_n, ty = (t / tx_n) % ty_n, tz = t / (tx_n * ty_n);
  const uint32_t m = expand10(tx & 511u) | (expand10(ty & 511u) << 1) | (expand10(tz & 511u) << 2);
  return ((uint32_t)(31 - bucket) << 27) | m;
}

__global__ __launch_bounds__(256) void k_tile_keys(ColumnArgs a, int shape, uint32_t ntiles, int mode, uint32_t *keys,
                                                   int32_t *ids) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= ntiles) return;
  keys[t] = tile_key(a, shape, t, mode);
  ids[t] = (int32_t)t;
}

__device__ __forceinline__ int job_of_tile(const BatchArgs &t, uint32_t tile) {
  int lo = 0, hi = (int)t.njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t.tile_prefix[mid] <= tile) lo = mid; else hi = mid - 1;
  }
  return lo;
}


// Persistent variant for a batch of meshes: blocks take block-tiles from one device work queue
// (tile_prefix[j] = first tile of job j) until the queue is empty.  SEP: the batch's brick shape is
// 4 x 4 x sz, and band jobs use the separable box tests too.
template <bool COUNT, bool SLIVERS, bool CAPS, int V, bool SEP>
__global__ __launch_bounds__(256, 8) void k_sdf_batch(BatchArgs t) {
  __shared__ int sstack[4][kStack4];
  __shared__ uint32_t s_tile;
  const int wv = threadIdx.x >> 6;
  const int sx = kShapes[t.shape][0], sy = kShapes[t.shape][1], sz = 64 * V / (sx * sy);
  const uint32_t total = t.tile_prefix[t.njobs];
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(t.head, 1u);
    __syncthreads();
    const uint32_t pos = s_tile;
    __syncthreads();
    if (pos >= total) break;
    const uint32_t tile = t.order ? t.order[pos] : pos;
    const int lo = job_of_tile(t, tile);
    const SdfArgs &a = t.jobs[lo];
    const uint32_t local = tile - t.tile_prefix[lo];
    const GridArgs &g = a.col.g;
    if (a.band_bricks) {  // narrow band: 2 block-tiles per 8^3 brick, 4x4x4 sub-brick per wave
      const uint32_t b = a.band_bricks[local >> 1];
      const uint32_t nbx = (uint32_t)g.nx >> 3, nby = (uint32_t)g.ny >> 3;
      const int sub = (int)((local & 1u) << 2) | wv;
      const int bx = (int)(b % nbx) * 8 + (sub & 1) * 4;
      const int by = (int)((b / nbx) % nby) * 8 + ((sub >> 1) & 1) * 4;
      const int bz = (int)(b / (nbx * nby)) * 8 + (sub >> 2) * 4;
      wave_brick<COUNT, SLIVERS, CAPS, 1, SEP>(a, bx, by, bz, 4, 4, sstack[wv]);
      continue;
    }
    const uint32_t tx = (g.nx + 2 * sx - 1) / (2 * sx), ty = (g.ny + 2 * sy - 1) / (2 * sy);
    const int bx = (int)(local % tx) * 2 * sx + (wv & 1) * sx;
    const int by = (int)((local / tx) % ty) * 2 * sy + (wv >> 1) * sy;
    const int bz = (int)(local / (tx * ty)) * sz;
    wave_brick<COUNT, SLIVERS, CAPS, V, SEP>(a, bx, by, bz, sx, sy, sstack[wv]);
  }
}

// resolve for every job of a batch: blockIdx.y = job
__global__ __launch_bounds__(256, 4) void k_resolve_batch(const SdfArgs *jobs, int per_voxel_blocks) {
  if (per_voxel_blocks)
    resolve_range_block(jobs[blockIdx.y], blockIdx.x, gridDim.x);
  else
    resolve_range_wave(jobs[blockIdx.y], blockIdx.x * 4 + (threadIdx.x >> 6), gridDim.x * 4);
}
__global__ __launch_bounds__(256, 4) void k_singular_parts_batch(const SdfArgs *jobs) {
  singular_parts(jobs[blockIdx.y], blockIdx.x * 4 + (threadIdx.x >> 6), gridDim.x * 4);
}
__global__ __launch_bounds__(256) void k_singular_finish_batch(const SdfArgs *jobs) {
  singular_finish(jobs[blockIdx.y], blockIdx.x * 4 + (threadIdx.x >> 6), gridDim.x * 4);
}

// ============================================================================================
// host-side launchers
// ============================================================================================
static inline unsigned nb(uint64_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

void fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s) {
  if (!n) return;
  const unsigned blocks = (unsigned)std::min<uint64_t>(nb(n), 65535u * 4);
  k_fill_u32<<<blocks, 256, 0, s>>>(p, v, n);
}

void launch_prep(const float *xyz, int n, Scalars *sc, hipStream_t s) {
  k_prep<<<std::min(nb(n), 512u), 256, 0, s>>>(xyz, n, sc);
}
void launch_morton(const float *xyz, int n, const Scalars *sc, uint32_t *codes, int32_t *idx, hipStream_t s) {
  k_morton<<<nb(n), 256, 0, s>>>(xyz, n, sc, codes, idx);
}

void fill_multi(const FillList &l, hipStream_t s) {
  uint64_t mx = 0;
  for (int k = 0; k < l.count; ++k) mx = std::max(mx, l.r[k].n);
  if (mx) k_fill_multi<<<std::min(nb(mx), 1024u), 256, 0, s>>>(l);
}
hipError_t sort_pairs(void *temp, size_t &temp_bytes, const uint32_t *kin, uint32_t *kout, const int32_t *vin,
                      int32_t *vout, int n, hipStream_t s) {
  return sort_pairs_u32(temp, temp_bytes, kin, kout, vin, vout, (uint32_t)n, 30, s);
}

void launch_karras(const uint32_t *codes, int n, Node *nodes, int32_t *leaf_parent, int2 *ranges, hipStream_t s) {
  if (n > 1) k_karras<<<nb(n - 1), 256, 0, s>>>(codes, n, nodes, leaf_parent, ranges);
}
void launch_gather(const float *xyz, const int32_t *order, int n, float4 *tris, TriRec *recs, Scalars *sc,
                   hipStream_t s) {
  k_gather<<<nb(n), 256, 0, s>>>(xyz, order, n, tris, recs, sc);
}
void launch_refit(const float4 *tris, int n, Node *nodes, const int32_t *leaf_parent, uint32_t *span,
                  float *leafbox, float *blockbox, float *root_box, Scalars *sc, const int2 *ranges, int block, Node4 *nodes4,
                  Obb *obbs, Obb *node_obbs, bool depth, hipStream_t s) {
  if (n == 1) {
    k_single<<<1, 1, 0, s>>>(tris, nodes, root_box);
    k_build4<<<1, 256, 0, s>>>(nodes, nullptr, 1, 1, nodes4);
    return;
  }
  const int nblk = (n + 255) / 256;
  k_refit_blocks<<<nblk, 256, 0, s>>>(tris, n, nodes, ranges, leafbox, blockbox, root_box, span);
  k_refit_span<<<std::min(2 * nblk, 1024), 256, 0, s>>>(n, nodes, ranges, leafbox, blockbox, root_box, span);
  if (depth) k_depth<<<nb(n), 256, 0, s>>>(nodes, leaf_parent, n, sc);
  if (block > 1) {
    k_collapse_blocks<<<nb(n - 1), 256, 0, s>>>(nodes, ranges, n - 1, block);
    k_block_obbs<<<nb(n - 1), 256, 0, s>>>(nodes, ranges, tris, n - 1, block, obbs);
    const int group = n >= (1 << 16) ? 64 : 8;  // groups of 8 nodes: 1.6 % off the 512-mesh batch
    k_node_obbs<<<std::min(nb(n - 1, 4 * group), 2048u), 256, 0, s>>>(ranges, tris, n - 1, block, group, node_obbs);
  }
  k_build4<<<nb(n - 1), 256, 0, s>>>(nodes, ranges, n - 1, block, nodes4);
}
void launch_depth(const Node *nodes, const int32_t *leaf_parent, int n, Scalars *sc, hipStream_t s) {
  if (n > 1) k_depth<<<nb(n), 256, 0, s>>>(nodes, leaf_parent, n, sc);
}
void launch_weld(const float *xyz, int ntri, int32_t *slots, int32_t *minid, uint32_t vmask, int32_t *cslot,
                 unsigned long long *ekeys, int32_t *enet, uint32_t emask, Scalars *sc, int2 *bnd, uint32_t bnd_cap,
                 hipStream_t s) {
  const int nc = 3 * ntri;
  k_weld_insert<<<nb(nc), 256, 0, s>>>(xyz, nc, slots, vmask, cslot);
  k_weld_min<<<nb(nc), 256, 0, s>>>(cslot, nc, minid);
  k_edge_insert<<<nb(ntri), 256, 0, s>>>(cslot, minid, ntri, ekeys, enet, emask);
  k_edge_collect<<<std::min(nb((uint64_t)emask + 1), 1024u), 256, 0, s>>>(ekeys, enet, emask + 1, sc, bnd, bnd_cap);
}
void launch_cap_keys(const int2 *bnd, uint32_t ne, unsigned long long *keys, hipStream_t s) {
  if (ne) k_cap_keys<<<nb(ne), 256, 0, s>>>(bnd, ne, keys);
}
hipError_t sort_keys_u64(void *temp, size_t &temp_bytes, const unsigned long long *kin, unsigned long long *kout,
                         uint32_t n, int end_bit, hipStream_t s) {
  return rocprim::radix_sort_keys(temp, temp_bytes, kin, kout, n, 0u, (unsigned)end_bit, s);
}
void launch_cap_in_keys(const unsigned long long *ek, uint32_t ne, unsigned long long *ik, uint32_t *iv,
                        hipStream_t s) {
  if (ne) k_cap_in_keys<<<nb(ne), 256, 0, s>>>(ek, ne, ik, iv);
}
void launch_cap_succ(const unsigned long long *ek, const unsigned long long *ik, const uint32_t *iv, uint32_t ne,
                     uint32_t *lab, uint32_t *jmp, hipStream_t s) {
  if (ne) k_cap_succ<<<nb(ne), 256, 0, s>>>(ek, ik, iv, ne, lab, jmp);
}
void launch_cap_jump(const uint32_t *lab0, const uint32_t *jmp0, uint32_t ne, uint32_t *lab1, uint32_t *jmp1,
                     hipStream_t s) {
  if (ne) k_cap_jump<<<nb(ne), 256, 0, s>>>(lab0, jmp0, ne, lab1, jmp1);
}
void launch_iota(uint32_t *v, uint32_t n, hipStream_t s) {
  if (n) k_iota<<<nb(n), 256, 0, s>>>(v, n);
}
void launch_cap_bounds(const unsigned long long *ek, uint32_t ne, const float *xyz, uint32_t *bounds, hipStream_t s) {
  if (ne) k_cap_bounds<<<std::min(nb(ne), 64u), 256, 0, s>>>(ek, ne, xyz, bounds);
}
void launch_cap_sums(const uint32_t *slab, const uint32_t *sedge, uint32_t ne, const unsigned long long *ek,
                     const float *xyz, const uint32_t *bounds, unsigned long long *acc, hipStream_t s) {
  if (ne) k_cap_sums<<<nb(ne), 256, 0, s>>>(slab, sedge, ne, ek, xyz, bounds, acc);
}
void launch_cap_tris(const unsigned long long *ek, const uint32_t *lab, uint32_t ne, const float *xyz,
                     const uint32_t *bounds, const unsigned long long *acc, float *tris, uint32_t *codes,
                     int32_t *idx, hipStream_t s) {
  if (ne) k_cap_tris<<<nb(ne), 256, 0, s>>>(ek, lab, ne, xyz, bounds, acc, tris, codes, idx);
}
void launch_cap_gather(const float *tmp, const int32_t *order, uint32_t ne, float *caps, hipStream_t s) {
  if (ne) k_cap_gather<<<nb(ne), 256, 0, s>>>(tmp, order, ne, caps);
}
void launch_cap_tree(const float *caps, const uint32_t *codes, int n, int2 *child, int32_t *parent, int2 *ranges,
                     CapNode *nodes, int4 *cut, hipStream_t s) {
  if (n <= 0) return;
  if (n > 1) k_cap_karras<<<nb(n - 1), 256, 0, s>>>(codes, n, child, parent, ranges);
  k_cap_nodes<<<nb(n > 1 ? n - 1 : 1, 4), 256, 0, s>>>(caps, child, parent, ranges, n, nodes);
  if (n > kCapBigFans)
    k_cap_nodes_big<<<(n - 1) < 256 ? (n - 1) : 256, 64 * kCapBigWaves, 0, s>>>(caps, child, parent, ranges, n, nodes);
  k_cap_cut<<<nb(kCapCut), 256, 0, s>>>(nodes, ranges, cut);
}
void launch_occupancy(const float *xyz, int n, const GridArgs &g, uint32_t *words, hipStream_t s) {
  k_occupancy<<<nb(n, 4), 256, 0, s>>>(xyz, n, g, words);
}
void launch_splat_count(const float *xyz, int n, const GridArgs &g, uint32_t *counts, hipStream_t s) {
  k_splat_count<<<nb(n, 4), 256, 0, s>>>(xyz, n, g, counts);
}
hipError_t exclusive_scan_u32(void *temp, size_t &temp_bytes, const uint32_t *in, uint32_t *out, uint32_t n,
                              hipStream_t s) {
  return rocprim::exclusive_scan(temp, temp_bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), s);
}
void launch_splat_emit(const float *xyz, int n, const GridArgs &g, const uint32_t *offsets, uint32_t *cell,
                       int32_t *tri, hipStream_t s) {
  k_splat_emit<<<nb(n, 4), 256, 0, s>>>(xyz, n, g, offsets, cell, tri);
}
// rocPRIM merge-sorts 2^17 .. 2^20 keys in about twenty launches; there the three digit passes of
// radix_sort_u32 take 0.6x its time (measured at 2^18 .. 2^19 keys).  Fewer keys sort in rocPRIM's few block
// launches, more in its one-sweep radix sort, both quicker than the digit passes.
constexpr uint32_t kRadixMin = 1u << 17, kRadixMax = 1u << 20;
hipError_t sort_pairs_u32(void *temp, size_t &temp_bytes, const uint32_t *kin, uint32_t *kout, const int32_t *vin,
                          int32_t *vout, uint32_t n, int end_bit, hipStream_t s) {
  if (n >= kRadixMin && n < kRadixMax) return radix_sort_u32(temp, temp_bytes, kin, kout, vin, vout, n, end_bit, s);
  return rocprim::radix_sort_pairs(temp, temp_bytes, kin, kout, vin, vout, n, 0u, (unsigned)end_bit, s);
}
void launch_splat_reduce(const uint32_t *cell, const int32_t *tri, uint32_t npairs, const float *attr, int k,
                         float *out, uint64_t ncell, hipStream_t s) {
  if (npairs) k_splat_reduce<<<nb(npairs), 256, 0, s>>>(cell, tri, npairs, attr, k, out, ncell);
}
void launch_tri_materials(uint32_t ntri, const float *uvs, uint64_t nuv, const uint32_t *tri_mat, uint64_t ntm,
                          const MatDev *mats, uint32_t nmat, const TexDev *texs, uint32_t ntex, const uint8_t *pixels,
                          const float *lut, float *out, hipStream_t s) {
  if (ntri) k_tri_materials<<<nb(ntri), 256, 0, s>>>(ntri, uvs, nuv, tri_mat, ntm, mats, nmat, texs, ntex, pixels, lut, out);
}
void launch_vox_heads(const uint32_t *cell, uint32_t npairs, uint32_t *head, hipStream_t s) {
  k_vox_heads<<<nb((uint64_t)npairs + 1), 256, 0, s>>>(cell, npairs, head);
}
void launch_vox_runs(const uint32_t *cell, const int32_t *tri, uint32_t npairs, const uint32_t *run_of,
                     const float *tri_rgb, unsigned long long *keys, uint32_t *vals, float *rgb, hipStream_t s) {
  if (npairs) k_vox_runs<<<nb(npairs), 256, 0, s>>>(cell, tri, npairs, run_of, tri_rgb, keys, vals, rgb);
}
hipError_t sort_pairs_u64(void *temp, size_t &temp_bytes, const unsigned long long *kin, unsigned long long *kout,
                          const uint32_t *vin, uint32_t *vout, uint32_t n, int end_bit, hipStream_t s) {
  return rocprim::radix_sort_pairs(temp, temp_bytes, kin, kout, vin, vout, n, 0u, (unsigned)end_bit, s);
}
void launch_vox_out(const unsigned long long *keys, const uint32_t *perm, uint32_t n, int nx, int ny, const float *rgb,
                    int32_t *cells, int32_t *first, float *colors, hipStream_t s) {
  if (n) k_vox_out<<<nb(n), 256, 0, s>>>(keys, perm, n, nx, ny, rgb, cells, first, colors);
}
void launch_columns_batch(const SdfArgs *jobs, const uint32_t *prefix, uint32_t njobs, uint32_t blocks,
                          hipStream_t s) {
  if (!njobs) return;
  k_zero_counts_batch<<<dim3(64, njobs), 256, 0, s>>>(jobs);
  if (blocks) k_columns_batch<<<blocks, 256, 0, s>>>(jobs, prefix, njobs);
}
void launch_columns(const ColumnArgs &a, hipStream_t s) {
  fill_u32(reinterpret_cast<uint32_t *>(a.count), 0u, (uint64_t)a.g.nx * a.g.ny, s);
  const uint64_t n = (uint64_t)a.ntri + (uint64_t)a.ncap;
  if (n) k_columns<<<nb(n), 256, 0, s>>>(a);
}
uint32_t column_blocks(const ColumnArgs &a) { return nb((uint64_t)a.ntri + (uint64_t)a.ncap); }
static const int kShapeHost[kNumShapes][3] = {{4, 4, 1}, {8, 4, 1}, {8, 8, 1}, {16, 4, 1}, {4, 4, 2}};

void tile_dims(const GridArgs &g, int shape, uint32_t d[3]) {
  const int sx = kShapeHost[shape][0], sy = kShapeHost[shape][1], sz = 64 * kShapeHost[shape][2] / (sx * sy);
  d[0] = nb(g.nx, 2 * sx);
  d[1] = nb(g.ny, 2 * sy);
  d[2] = nb(g.nz, sz);
}

void launch_tile_keys(const ColumnArgs &a, int shape, int mode, uint32_t *keys, int32_t *ids, hipStream_t s) {
  uint32_t d[3];
  tile_dims(a.g, shape, d);
  const uint32_t n = d[0] * d[1] * d[2];
  k_tile_keys<<<nb(n), 256, 0, s>>>(a, shape, n, mode, keys, ids);
}


template <bool COUNT, bool SLIVERS, bool CAPS>
static void launch_sdf_t(const SdfArgs &a, int variant, int shape, const uint32_t *order, hipStream_t s) {
  uint32_t d[3];
  tile_dims(a.col.g, shape, d);
  if (variant == 1) {
    // shape {4, 4, 2} is the only one with two voxels per lane.  4 x 4 shapes take the separable box tests
    // where the kernel keeps every value in registers (the FP64 sliver path, and the cap winding with two
    // voxels per lane, would spill them)
    const unsigned n = d[0] * d[1] * d[2];
    if (kShapeHost[shape][2] == 2) {
      if (!SLIVERS && !CAPS) k_sdf_wave<COUNT, SLIVERS, CAPS, 2, true><<<n, 256, 0, s>>>(a, shape, order, d[0], d[1]);
      else k_sdf_wave<COUNT, SLIVERS, CAPS, 2, false><<<n, 256, 0, s>>>(a, shape, order, d[0], d[1]);
    } else if (!SLIVERS && kShapeHost[shape][0] * kShapeHost[shape][1] == 16) {
      k_sdf_wave<COUNT, SLIVERS, CAPS, 1, true><<<n, 256, 0, s>>>(a, shape, order, d[0], d[1]);
    } else {
      k_sdf_wave<COUNT, SLIVERS, CAPS, 1, false><<<n, 256, 0, s>>>(a, shape, order, d[0], d[1]);
    }
    return;
  }
  dim3 grid(nb(a.col.g.nx, 8), nb(a.col.g.ny, 8), nb(a.col.g.nz, 4));
  k_sdf<COUNT, SLIVERS, CAPS><<<grid, 256, 0, s>>>(a);
}

#define IB_DISPATCH(fn, count, slivers, caps, ...)                              \
  switch (((count) ? 4 : 0) | ((slivers) ? 2 : 0) | ((caps) ? 1 : 0)) {        \
    case 0: fn<false, false, false>(__VA_ARGS__); break;                        \
    case 1: fn<false, false, true>(__VA_ARGS__); break;                         \
    case 2: fn<false, true, false>(__VA_ARGS__); break;                         \
    case 3: fn<false, true, true>(__VA_ARGS__); break;                          \
    case 4: fn<true, false, false>(__VA_ARGS__); break;                         \
    case 5: fn<true, false, true>(__VA_ARGS__); break;                          \
    case 6: fn<true, true, false>(__VA_ARGS__); break;                          \
    default: fn<true, true, true>(__VA_ARGS__); break;                          \
  }

void launch_sdf(const SdfArgs &a, bool count, int variant, bool slivers, int shape, const uint32_t *order,
                hipStream_t s) {
  IB_DISPATCH(launch_sdf_t, count, slivers, a.col.ncap > 0, a, variant, shape, order, s);
}

void launch_resolve(const SdfArgs &a, hipStream_t s) {
  k_resolve<<<2048, 256, 0, s>>>(a);  // one voxel per block on a mesh with caps
  if (a.sing_cap) {
    k_singular_parts<<<1024, 256, 0, s>>>(a);
    k_singular_finish<<<64, 256, 0, s>>>(a);
  }
}

template <bool COUNT, bool SLIVERS, bool CAPS>
static void launch_batch_t(const BatchArgs &t, uint32_t blocks, hipStream_t s) {
  if (kShapeHost[t.shape][2] == 2) k_sdf_batch<COUNT, SLIVERS, CAPS, 2, false><<<blocks, 256, 0, s>>>(t);
  else k_sdf_batch<COUNT, SLIVERS, CAPS, 1, false><<<blocks, 256, 0, s>>>(t);
}

void launch_sdf_batch(const BatchArgs &t, uint32_t blocks, bool count, bool slivers, bool caps, hipStream_t s) {
  IB_DISPATCH(launch_batch_t, count, slivers, caps, t, blocks, s);
}

// Brick selection by splatting: every triangle flags the bricks whose centre (a cell-centred grid of
// spacing H over the grid) lies within thresh of it, by the query's own squared distance (tri_d2 of the
// triangle's record), so a brick is flagged exactly when sqrt of the least squared distance over all
// triangles is at most thresh.  Candidates: the centres within thresh of the triangle's box along every
// axis, one more on each side for rounding.  A candidate farther than thresh + slack from the triangle's box
// is skipped before the exact test: the FP32 distance is within slack = 2^-17 (Mq + 3 Mp) of the exact one
// (tri_d2_f32's bound, for this triangle's vertices and these centres), so its root exceeds thresh too.
template <bool SLIVERS>
__global__ __launch_bounds__(256) void k_band_splat(const float4 *__restrict__ tris, const TriRec *__restrict__ recs,
                                                    int n, GridArgs cg, float thresh, uint32_t *flags) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const float4 v0 = tris[3 * t], v1 = tris[3 * t + 1], v2 = tris[3 * t + 2];
  const float lo[3] = {fminf(fminf(v0.x, v1.x), v2.x), fminf(fminf(v0.y, v1.y), v2.y), fminf(fminf(v0.z, v1.z), v2.z)};
  const float hi[3] = {fmaxf(fmaxf(v0.x, v1.x), v2.x), fmaxf(fmaxf(v0.y, v1.y), v2.y), fmaxf(fmaxf(v0.z, v1.z), v2.z)};
  const float org[3] = {cg.ox, cg.oy, cg.oz};
  const int dim[3] = {cg.nx, cg.ny, cg.nz};
  int a[3], b[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = max(0, (int)floorf((lo[k] - thresh - org[k]) / cg.h - 0.5f) - 1);
    b[k] = min(dim[k] - 1, (int)ceilf((hi[k] + thresh - org[k]) / cg.h - 0.5f) + 1);
  }
  float mq = 0.0f, mp = 0.0f;
  for (int k = 0; k < 3; ++k) {
    mq = fmaxf(mq, fmaxf(fabsf(org[k] + ((float)a[k] + 0.5f) * cg.h), fabsf(org[k] + ((float)b[k] + 0.5f) * cg.h)));
    mp = fmaxf(mp, fmaxf(fabsf(lo[k]), fabsf(hi[k])));
  }
  const float reach = thresh + (mq + 3.0f * mp) * 7.62939453125e-6f;  // 2^-17
  const float lim = reach * reach * 1.0000077f;  // covers the roundings of the box distance and of lim
  const TriRec r = recs[t];
  for (int z = a[2]; z <= b[2]; ++z) {
    const float qz = cg.oz + ((float)z + 0.5f) * cg.h;
    const float gz = fmaxf(fmaxf(lo[2] - qz, qz - hi[2]), 0.0f);
    for (int y = a[1]; y <= b[1]; ++y) {
      const float qy = cg.oy + ((float)y + 0.5f) * cg.h;
      const float gy = fmaxf(fmaxf(lo[1] - qy, qy - hi[1]), 0.0f);
      const float gyz = fmaf(gy, gy, gz * gz);
      if (gyz > lim) continue;
      for (int x = a[0]; x <= b[0]; ++x) {
        const float qx = cg.ox + ((float)x + 0.5f) * cg.h;
        const float gx = fmaxf(fmaxf(lo[0] - qx, qx - hi[0]), 0.0f);
        if (fmaf(gx, gx, gyz) > lim) continue;
        const size_t i = ((size_t)z * cg.ny + y) * cg.nx + x;
        if (flags[i]) continue;
        if (sqrtf(tri_d2<SLIVERS>({qx, qy, qz}, r)) <= thresh) flags[i] = 1u;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_band_scatter(const uint32_t *__restrict__ flags,
                                                      const uint32_t *__restrict__ slots, uint32_t n, int32_t *index,
                                                      uint32_t *bricks, uint32_t *count) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  index[i] = flags[i] ? (int32_t)slots[i] : -1;
  if (flags[i]) bricks[slots[i]] = i;
  if (i == n - 1) *count = slots[i] + flags[i];
}

void launch_band_splat(const float4 *tris, const TriRec *recs, int n, const GridArgs &cg, float thresh, bool slivers,
                       uint32_t *flags, hipStream_t s) {
  if (n <= 0) return;
  if (slivers) k_band_splat<true><<<nb(n), 256, 0, s>>>(tris, recs, n, cg, thresh, flags);
  else k_band_splat<false><<<nb(n), 256, 0, s>>>(tris, recs, n, cg, thresh, flags);
}

void launch_band_scatter(const uint32_t *flags, const uint32_t *slots, uint32_t nbrick, int32_t *index,
                         uint32_t *bricks, uint32_t *count, hipStream_t s) {
  if (nbrick) k_band_scatter<<<nb(nbrick), 256, 0, s>>>(flags, slots, nbrick, index, bricks, count);
}

void launch_resolve_batch(const SdfArgs *jobs, uint32_t njobs, bool singular, bool block_single, hipStream_t s) {
  if (!njobs) return;
  // a single job on an otherwise idle chip: one voxel per block on a mesh with caps; several jobs, or a job
  // beside other streams' work: one voxel per wave (measured: the served mix's capacity is higher so), a few
  // jobs the whole chip, a large batch a share per job
  if (njobs == 1 && block_single) {
    k_resolve_batch<<<dim3(2048, 1), 256, 0, s>>>(jobs, 1);
  } else {
    const unsigned per_job = njobs <= 4 ? 512u : 128u;
    k_resolve_batch<<<dim3(per_job, njobs), 256, 0, s>>>(jobs, 0);
  }
  if (singular) {
    k_singular_parts_batch<<<dim3(njobs <= 4 ? 1024u : 64u, njobs), 256, 0, s>>>(jobs);
    k_singular_finish_batch<<<dim3(16, njobs), 256, 0, s>>>(jobs);
  }
}

}  // namespace ib
