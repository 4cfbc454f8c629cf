/*
 * CPU oracle for the mesh -> SDF / occupancy bake path.  TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker, never by the product.  Build: oracle/Makefile (-O2 -ffp-contract=off).
 * Reference anchors are given per function; see ib_oracle.h for the list.
 */
#include "ib_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

/* std::min / std::max semantics for two floats (tri_box_overlap.h:32,39,...) */
static inline float smin(float a, float b) { return (b < a) ? b : a; }
static inline float smax(float a, float b) { return (a < b) ? b : a; }
/* glm::dot(vec3) = (x*x' + y*y') + z*z' */
static inline float gdot(const float a[3], const float b[3]) {
  const float tx = a[0] * b[0], ty = a[1] * b[1], tz = a[2] * b[2];
  return tx + ty + tz;
}

/* ============================================================================
 * a10: triBoxOverlap  (tri_box_overlap.h:12-114), expression trees kept literally
 * ========================================================================== */
static int plane_test(const float n[3], const float v[3], const float mb[3]) { /* :12-27 */
  float vmin[3], vmax[3];
  for (int q = 0; q < 3; ++q) {
    const float x = v[q];
    if (n[q] > 0.0f) {
      vmin[q] = -mb[q] - x;
      vmax[q] = mb[q] - x;
    } else {
      vmin[q] = mb[q] - x;
      vmax[q] = -mb[q] - x;
    }
  }
  if (gdot(n, vmin) > 0.0f) return 0;
  return gdot(n, vmax) >= 0.0f;
}
/* :29-35 */
static int ax_x01(float a, float b, float fa, float fb, const float v0[3], const float v2[3], const float bh[3]) {
  const float p0 = a * v0[1] - b * v0[2], p2 = a * v2[1] - b * v2[2];
  const float mn = smin(p0, p2), mx = smax(p0, p2);
  const float rad = fa * bh[1] + fb * bh[2];
  return !(mn > rad || mx < -rad);
}
/* :36-42 */
static int ax_x2(float a, float b, float fa, float fb, const float v0[3], const float v1[3], const float bh[3]) {
  const float p0 = a * v0[1] - b * v0[2], p1 = a * v1[1] - b * v1[2];
  const float mn = smin(p0, p1), mx = smax(p0, p1);
  const float rad = fa * bh[1] + fb * bh[2];
  return !(mn > rad || mx < -rad);
}
/* :43-49 */
static int ax_y02(float a, float b, float fa, float fb, const float v0[3], const float v2[3], const float bh[3]) {
  const float p0 = -a * v0[0] + b * v0[2], p2 = -a * v2[0] + b * v2[2];
  const float mn = smin(p0, p2), mx = smax(p0, p2);
  const float rad = fa * bh[0] + fb * bh[2];
  return !(mn > rad || mx < -rad);
}
/* :50-56 */
static int ax_y1(float a, float b, float fa, float fb, const float v0[3], const float v1[3], const float bh[3]) {
  const float p0 = -a * v0[0] + b * v0[2], p1 = -a * v1[0] + b * v1[2];
  const float mn = smin(p0, p1), mx = smax(p0, p1);
  const float rad = fa * bh[0] + fb * bh[2];
  return !(mn > rad || mx < -rad);
}
/* :57-63 */
static int ax_z12(float a, float b, float fa, float fb, const float v1[3], const float v2[3], const float bh[3]) {
  const float p1 = a * v1[0] - b * v1[1], p2 = a * v2[0] - b * v2[1];
  const float mn = smin(p1, p2), mx = smax(p1, p2);
  const float rad = fa * bh[0] + fb * bh[1];
  return !(mn > rad || mx < -rad);
}
/* :64-70 */
static int ax_z0(float a, float b, float fa, float fb, const float v0[3], const float v1[3], const float bh[3]) {
  const float p0 = a * v0[0] - b * v0[1], p1 = a * v1[0] - b * v1[1];
  const float mn = smin(p0, p1), mx = smax(p0, p1);
  const float rad = fa * bh[0] + fb * bh[1];
  return !(mn > rad || mx < -rad);
}

int orc_tri_box_overlap(const float c[3], const float bh[3], const float t0[3], const float t1[3], const float t2[3]) {
  float v0[3], v1[3], v2[3], e0[3], e1[3], e2[3];
  for (int k = 0; k < 3; ++k) { /* :74-75 */
    v0[k] = t0[k] - c[k];
    v1[k] = t1[k] - c[k];
    v2[k] = t2[k] - c[k];
  }
  for (int k = 0; k < 3; ++k) {
    e0[k] = v1[k] - v0[k];
    e1[k] = v2[k] - v1[k];
    e2[k] = v0[k] - v2[k];
  }
  float fex = fabsf(e0[0]), fey = fabsf(e0[1]), fez = fabsf(e0[2]); /* :77-83 */
  if (!ax_x01(e0[2], e0[1], fez, fey, v0, v2, bh)) return 0;
  if (!ax_y02(e0[2], e0[0], fez, fex, v0, v2, bh)) return 0;
  if (!ax_z12(e0[1], e0[0], fey, fex, v1, v2, bh)) return 0;
  fex = fabsf(e1[0]); /* :85-93 */
  fey = fabsf(e1[1]);
  fez = fabsf(e1[2]);
  if (!ax_x01(e1[2], e1[1], fez, fey, v0, v2, bh)) return 0;
  if (!ax_y02(e1[2], e1[0], fez, fex, v0, v2, bh)) return 0;
  if (!ax_z0(e1[1], e1[0], fey, fex, v0, v1, bh)) return 0;
  fex = fabsf(e2[0]); /* :95-103 */
  fey = fabsf(e2[1]);
  fez = fabsf(e2[2]);
  if (!ax_x2(e2[2], e2[1], fez, fey, v0, v1, bh)) return 0;
  if (!ax_y1(e2[2], e2[0], fez, fex, v0, v1, bh)) return 0;
  if (!ax_z12(e2[1], e2[0], fey, fex, v1, v2, bh)) return 0;
  for (int axis = 0; axis < 3; ++axis) { /* :105-110 */
    const float mn = smin(smin(v0[axis], v1[axis]), v2[axis]);
    const float mx = smax(smax(v0[axis], v1[axis]), v2[axis]);
    if (mn > bh[axis] || mx < -bh[axis]) return 0;
  }
  /* :112-113  glm::cross */
  const float n[3] = {e0[1] * e1[2] - e1[1] * e0[2], e0[2] * e1[0] - e1[2] * e0[0], e0[0] * e1[1] - e1[0] * e0[1]};
  return plane_test(n, v0, bh);
}

/* ============================================================================
 * framing
 * ========================================================================== */
static float longest_axis(const float bmin[3], const float bmax[3], float ext[3]) {
  for (int k = 0; k < 3; ++k) ext[k] = bmax[k] - bmin[k];
  /* std::max({ex, ey, ez, 1e-6f}) */
  float m = ext[0];
  if (m < ext[1]) m = ext[1];
  if (m < ext[2]) m = ext[2];
  if (m < 1e-6f) m = 1e-6f;
  return m;
}

void orc_sdf_framing(const float bmin[3], const float bmax[3], int res, float origin[3], float *h, int n[3]) {
  float ext[3];
  const float longest = longest_axis(bmin, bmax, ext); /* sdf_baker.cpp:139-140 */
  const float vs = longest / (float)(res > 1 ? res : 1); /* :141 */
  const float pad = vs * 2.0f;                           /* :143-144 kPad = 2 */
  const float pad2 = pad * 2.0f;                         /* :145 */
  for (int k = 0; k < 3; ++k) {
    origin[k] = bmin[k] - pad;
    const float padded = ext[k] + pad2;
    n[k] = (int)ceilf(padded / vs) + 1; /* :146-148 */
  }
  *h = vs;
}

void orc_vox_framing(const float bmin[3], const float bmax[3], int res, float origin[3], float *h) {
  float ext[3];
  const float longest = longest_axis(bmin, bmax, ext); /* voxelize.cpp:54-55 */
  *h = longest / (float)(res > 1 ? res : 1);           /* :56 */
  for (int k = 0; k < 3; ++k) origin[k] = bmin[k];     /* :57 */
}

/* per-triangle inclusive cell range, voxelize.cpp:70-77 / sdf_baker.cpp:95-101 */
static void cell_range(const float *t, const float origin[3], float h, const int maxc[3], int lo[3], int hi[3]) {
  for (int k = 0; k < 3; ++k) {
    const float mn = smin(smin(t[k], t[3 + k]), t[6 + k]);
    const float mx = smax(smax(t[k], t[3 + k]), t[6 + k]);
    int a = (int)floorf((mn - origin[k]) / h);
    int b = (int)floorf((mx - origin[k]) / h);
    /* glm::clamp(x, 0, m) = min(max(x, 0), m): res <= 0 gives -1, as the reference does */
    a = a < 0 ? 0 : a;
    a = a > maxc[k] ? maxc[k] : a;
    b = b < 0 ? 0 : b;
    b = b > maxc[k] ? maxc[k] : b;
    lo[k] = a;
    hi[k] = b;
  }
}

static void cell_center(const float origin[3], float h, int x, int y, int z, float c[3]) {
  c[0] = origin[0] + ((float)x + 0.5f) * h; /* voxelize.cpp:84 */
  c[1] = origin[1] + ((float)y + 0.5f) * h;
  c[2] = origin[2] + ((float)z + 0.5f) * h;
}

/* ============================================================================
 * a7: bakePerCellMaterial, sdf_baker.cpp:76-128
 * ========================================================================== */
/* Per triangle t in order, every cell of its clamped AABB range that triBoxOverlap accepts gets
 * mat[5t..5t+4] (r, g, b, metallic, roughness) added (FP32, starting from 0); a touched cell ends
 * as sum / count, an untouched one keeps (1, 1, 1, 0, 1) (:77-80, :121-127).  out: 5 planes of
 * nx*ny*nz floats, cell (z*ny + y)*nx + x. */
void orc_splat_materials(const float *xyz, size_t ntri, const float *mat, const float origin[3], float h, int nx,
                         int ny, int nz, float *out) {
  const size_t ncell = (size_t)nx * ny * nz;
  float *sum = (float *)calloc(5 * (ncell ? ncell : 1), sizeof(float));
  int *cnt = (int *)calloc(ncell ? ncell : 1, sizeof(int));
  const int maxc[3] = {nx - 1, ny - 1, nz - 1};
  const float bh[3] = {h * 0.5f, h * 0.5f, h * 0.5f};
  for (size_t t = 0; t < ntri; ++t) {
    const float *p = xyz + 9 * t;
    int lo[3], hi[3];
    cell_range(p, origin, h, maxc, lo, hi);
    for (int z = lo[2]; z <= hi[2]; ++z)
      for (int y = lo[1]; y <= hi[1]; ++y)
        for (int x = lo[0]; x <= hi[0]; ++x) {
          float c[3];
          cell_center(origin, h, x, y, z, c);
          if (!orc_tri_box_overlap(c, bh, p, p + 3, p + 6)) continue;
          const size_t i = ((size_t)z * ny + y) * nx + x;
          for (int k = 0; k < 5; ++k) sum[(size_t)k * ncell + i] += mat[5 * t + k];
          cnt[i] += 1;
        }
  }
  const float def[5] = {1.0f, 1.0f, 1.0f, 0.0f, 1.0f};
  for (size_t i = 0; i < ncell; ++i)
    for (int k = 0; k < 5; ++k)
      out[(size_t)k * ncell + i] = cnt[i] ? sum[(size_t)k * ncell + i] / (float)cnt[i] : def[k];
  free(sum);
  free(cnt);
}

/* ============================================================================
 * a9: voxelizeMesh (first-touch order), voxelize.cpp:52-107
 * ========================================================================== */
typedef struct {
  int64_t *keys; /* packed cell key, -1 empty */
  int64_t *vals;
  size_t cap;
} i64map;

static uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}
static void map_init(i64map *m, size_t cap) {
  size_t c = 16;
  while (c < cap * 2) c <<= 1;
  m->cap = c;
  m->keys = (int64_t *)malloc(c * sizeof(int64_t));
  m->vals = (int64_t *)malloc(c * sizeof(int64_t));
  for (size_t i = 0; i < c; ++i) m->keys[i] = -1;
}
static void map_free(i64map *m) {
  free(m->keys);
  free(m->vals);
}
static void map_grow(i64map *m);
/* returns pointer to value; *inserted set when new */
static int64_t *map_get(i64map *m, int64_t key, int *inserted, size_t *count) {
  if ((*count + 1) * 2 > m->cap) map_grow(m);
  size_t i = mix64((uint64_t)key) & (m->cap - 1);
  for (;;) {
    if (m->keys[i] == key) {
      *inserted = 0;
      return &m->vals[i];
    }
    if (m->keys[i] == -1) {
      m->keys[i] = key;
      *inserted = 1;
      (*count)++;
      return &m->vals[i];
    }
    i = (i + 1) & (m->cap - 1);
  }
}
static void map_grow(i64map *m) {
  i64map n;
  map_init(&n, m->cap);
  for (size_t i = 0; i < m->cap; ++i)
    if (m->keys[i] != -1) {
      size_t j = mix64((uint64_t)m->keys[i]) & (n.cap - 1);
      while (n.keys[j] != -1) j = (j + 1) & (n.cap - 1);
      n.keys[j] = m->keys[i];
      n.vals[j] = m->vals[i];
    }
  map_free(m);
  *m = n;
}

size_t orc_voxelize_first_touch(const float *xyz, size_t ntri, const float *colors, const float bmin[3],
                                const float bmax[3], int res, int32_t *cells, float *cell_colors, size_t cap) {
  float origin[3], h;
  orc_vox_framing(bmin, bmax, res, origin, &h);
  const int maxc[3] = {res - 1, res - 1, res - 1};
  const float half[3] = {h * 0.5f, h * 0.5f, h * 0.5f}; /* :85 */
  i64map map;
  map_init(&map, 1024);
  size_t count = 0;
  size_t ncells = 0;
  size_t scap = 1024;
  float *csum = (float *)malloc(scap * 3 * sizeof(float));
  int *ccnt = (int *)malloc(scap * sizeof(int));
  int32_t *tmpcells = (int32_t *)malloc(scap * 3 * sizeof(int32_t));
  for (size_t t = 0; t < ntri; ++t) {
    const float *tri = xyz + 9 * t;
    int lo[3], hi[3];
    cell_range(tri, origin, h, maxc, lo, hi);
    const float col[3] = {colors ? colors[3 * t] : 1.0f, colors ? colors[3 * t + 1] : 1.0f,
                          colors ? colors[3 * t + 2] : 1.0f};
    for (int z = lo[2]; z <= hi[2]; ++z)
      for (int y = lo[1]; y <= hi[1]; ++y)
        for (int x = lo[0]; x <= hi[0]; ++x) {
          float c[3];
          cell_center(origin, h, x, y, z, c);
          if (!orc_tri_box_overlap(c, half, tri, tri + 3, tri + 6)) continue;
          /* cells may be negative only when res <= 0; pack with an offset */
          const int64_t key = (((int64_t)z + 0x100000) << 42) | (((int64_t)y + 0x100000) << 21) | ((int64_t)x + 0x100000);
          int ins;
          int64_t *v = map_get(&map, key, &ins, &count);
          if (ins) {
            if (ncells == scap) {
              scap *= 2;
              csum = (float *)realloc(csum, scap * 3 * sizeof(float));
              ccnt = (int *)realloc(ccnt, scap * sizeof(int));
              tmpcells = (int32_t *)realloc(tmpcells, scap * 3 * sizeof(int32_t));
            }
            *v = (int64_t)ncells;
            tmpcells[3 * ncells] = x;
            tmpcells[3 * ncells + 1] = y;
            tmpcells[3 * ncells + 2] = z;
            csum[3 * ncells] = col[0];
            csum[3 * ncells + 1] = col[1];
            csum[3 * ncells + 2] = col[2];
            ccnt[ncells] = 1;
            ncells++;
          } else {
            const size_t i = (size_t)*v;
            csum[3 * i] += col[0];
            csum[3 * i + 1] += col[1];
            csum[3 * i + 2] += col[2];
            ccnt[i] += 1;
          }
        }
  }
  const size_t n = ncells < cap ? ncells : cap;
  for (size_t i = 0; i < n; ++i) {
    if (cells) {
      cells[3 * i] = tmpcells[3 * i];
      cells[3 * i + 1] = tmpcells[3 * i + 1];
      cells[3 * i + 2] = tmpcells[3 * i + 2];
    }
    if (cell_colors) { /* :102-104 */
      const float cnt = (float)ccnt[i];
      cell_colors[3 * i] = csum[3 * i] / cnt;
      cell_colors[3 * i + 1] = csum[3 * i + 1] / cnt;
      cell_colors[3 * i + 2] = csum[3 * i + 2] / cnt;
    }
  }
  free(csum);
  free(ccnt);
  free(tmpcells);
  map_free(&map);
  return ncells;
}

void orc_occupancy(const float *xyz, size_t ntri, const float origin[3], float h, int nx, int ny, int nz,
                   uint32_t *words) {
  const int maxc[3] = {nx - 1, ny - 1, nz - 1};
  const float half[3] = {h * 0.5f, h * 0.5f, h * 0.5f};
  const int wpr = (nx + 31) / 32;
  for (size_t t = 0; t < ntri; ++t) {
    const float *tri = xyz + 9 * t;
    int lo[3], hi[3];
    cell_range(tri, origin, h, maxc, lo, hi);
    for (int z = lo[2]; z <= hi[2]; ++z)
      for (int y = lo[1]; y <= hi[1]; ++y)
        for (int x = lo[0]; x <= hi[0]; ++x) {
          float c[3];
          cell_center(origin, h, x, y, z, c);
          if (orc_tri_box_overlap(c, half, tri, tri + 3, tri + 6))
            words[((size_t)z * ny + y) * wpr + (x >> 5)] |= 1u << (x & 31);
        }
  }
}

/* ============================================================================
 * distance
 * ========================================================================== */
int orc_tri_is_sliver(const float *t) {
  const float ab[3] = {t[3] - t[0], t[4] - t[1], t[5] - t[2]};
  const float ac[3] = {t[6] - t[0], t[7] - t[1], t[8] - t[2]};
  const float n[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
  const float nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  const float l1 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
  const float l2 = ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2];
  /* sin^2(angle at v0) < 1e-4, or any zero-length edge */
  return !(nn > 1e-4f * (l1 * l2));
}

static inline float dot3f(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline double dot3d(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Distance record of a triangle and the branch-free FP32 distance built on it.  This is the
 * closest-point query behind the bake's distances (the reference finds the closest triangle with
 * OptiX, src/convert/sdf_baker.cpp:150-305, and its shaders are not part of the tree); the result is
 * pinned against the FP64 distance (orc_tri_dist2_f64) and closed forms in the oracle tests.
 * Record: e0 = p1 - p0, ac = p2 - p0, n = cross(e0, ac), len = sqrt(dot(e0, e0)), inv_len = 1 / len,
 *   X = e0 * inv_len, Z = n * (1 / sqrt(dot(n, n))), Y = cross(Z, X), c = (dot(X, ac), dot(Y, ac)),
 *   e1 = (c0 - len, c1), inv_e1 = 1 / (e1_1^2 + e1_0^2 as fmaf(e1_1, e1_1, e1_0 * e1_0)),
 *   nie2 = -1 / fmaf(c1, c1, c0 * c0);
 *   cross(a, b)_x = fmaf(a_y, b_z, -(a_z * b_y)) (cyclic), dot(a, b) = fmaf(a_z, b_z, fmaf(a_y, b_y, a_x * b_x)).
 * Query: d = q - o, (u, v, w) = (dot(X, d), dot(Y, d), dot(Z, d)) in the triangle's plane frame;
 *   segment 0: t = clamp(u * inv_len), r = fmaf(-t, len, u), d0 = fmaf(v, v, r * r);
 *   segment 1: a = (u - len, v), t = clamp(fmaf(v, e1_1, a_0 * e1_0) * inv_e1), r = fmaf(-t, e1, a), d1 = |r|^2;
 *   segment 2: a = (u - c0, v - c1), t = clamp(fmaf(a_1, c1, a_0 * c0) * nie2), r = fmaf(t, c, a), d2 = |r|^2;
 *   inside = max(-v, -k1, k2) <= 0 with k1 = fmaf(e1_0, v, -(e1_1 * a1_0)), k2 = fmaf(c0, a2_1, -(c1 * a2_0));
 *   result fmaf(w, w, inside ? 0 : min(d0, d1, d2)).  |r|^2 in 2D is fmaf(r_1, r_1, r_0 * r_0). */
static inline float fdot3(const float a[3], const float b[3]) { return fmaf(a[2], b[2], fmaf(a[1], b[1], a[0] * b[0])); }
static inline void fcross3(const float a[3], const float b[3], float o[3]) {
  o[0] = fmaf(a[1], b[2], -(a[2] * b[1]));
  o[1] = fmaf(a[2], b[0], -(a[0] * b[2]));
  o[2] = fmaf(a[0], b[1], -(a[1] * b[0]));
}
static inline float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

void orc_trirec_make(const float *t, orc_trirec *r) {
  memset(r, 0, sizeof(*r));
  memcpy(r->v, t, 9 * sizeof(float));
  r->sliver = orc_tri_is_sliver(t);
  if (r->sliver) return;
  const float e0[3] = {t[3] - t[0], t[4] - t[1], t[5] - t[2]};
  const float ac[3] = {t[6] - t[0], t[7] - t[1], t[8] - t[2]};
  float n[3];
  fcross3(e0, ac, n);
  r->len = sqrtf(fdot3(e0, e0));
  r->inv_len = 1.0f / r->len;
  const float inv_n = 1.0f / sqrtf(fdot3(n, n));
  for (int k = 0; k < 3; ++k) {
    r->o[k] = t[k];
    r->x[k] = e0[k] * r->inv_len;
    r->z[k] = n[k] * inv_n;
  }
  fcross3(r->z, r->x, r->y);
  r->c[0] = fdot3(r->x, ac);
  r->c[1] = fdot3(r->y, ac);
  r->e1[0] = r->c[0] - r->len;
  r->e1[1] = r->c[1];
  r->inv_e1 = 1.0f / fmaf(r->e1[1], r->e1[1], r->e1[0] * r->e1[0]);
  r->nie2 = -1.0f / fmaf(r->c[1], r->c[1], r->c[0] * r->c[0]);
}

float orc_trirec_dist2(const float q[3], const orc_trirec *r) {
  if (r->sliver) {
    const double qd[3] = {q[0], q[1], q[2]};
    return (float)orc_tri_dist2_f64(qd, r->v);
  }
  const float d[3] = {q[0] - r->o[0], q[1] - r->o[1], q[2] - r->o[2]};
  const float u = fdot3(r->x, d), v = fdot3(r->y, d), w = fdot3(r->z, d);
  /* segment 0: p0 -> p1 along the X axis */
  const float t0 = clamp01(u * r->inv_len);
  const float r0 = fmaf(-t0, r->len, u);
  const float d0 = fmaf(v, v, r0 * r0);
  /* segment 1: p1 -> p2 */
  const float a1 = u - r->len;
  const float t1 = clamp01(fmaf(v, r->e1[1], a1 * r->e1[0]) * r->inv_e1);
  const float s1x = fmaf(-t1, r->e1[0], a1), s1y = fmaf(-t1, r->e1[1], v);
  const float d1 = fmaf(s1y, s1y, s1x * s1x);
  /* segment 2: p2 -> p0 */
  const float a2x = u - r->c[0], a2y = v - r->c[1];
  const float t2 = clamp01(fmaf(a2y, r->c[1], a2x * r->c[0]) * r->nie2);
  const float s2x = fmaf(t2, r->c[0], a2x), s2y = fmaf(t2, r->c[1], a2y);
  const float d2 = fmaf(s2y, s2y, s2x * s2x);
  const float k1 = fmaf(r->e1[0], v, -(r->e1[1] * a1));
  const float k2 = fmaf(r->c[0], a2y, -(r->c[1] * a2x));
  const int inside = fmaxf(fmaxf(-v, -k1), k2) <= 0.0f;
  const float dmin = fminf(fminf(d0, d1), d2);
  return fmaf(w, w, inside ? 0.0f : dmin);
}

float orc_tri_dist2_f32(const float q[3], const float *t) {
  orc_trirec r;
  orc_trirec_make(t, &r);
  return orc_trirec_dist2(q, &r);
}

static double seg_dist2_d(const double a[3], const double b[3]) {
  const double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double l = dot3d(ab, ab);
  double t = 0.0;
  if (l > 0.0) {
    t = -dot3d(a, ab) / l;
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
  }
  const double r[3] = {a[0] + t * ab[0], a[1] + t * ab[1], a[2] + t * ab[2]};
  return dot3d(r, r);
}

double orc_tri_dist2_f64(const double q[3], const float *t) {
  const double a[3] = {(double)t[0] - q[0], (double)t[1] - q[1], (double)t[2] - q[2]};
  const double b[3] = {(double)t[3] - q[0], (double)t[4] - q[1], (double)t[5] - q[2]};
  const double c[3] = {(double)t[6] - q[0], (double)t[7] - q[1], (double)t[8] - q[2]};
  double best = seg_dist2_d(a, b);
  double d = seg_dist2_d(b, c);
  if (d < best) best = d;
  d = seg_dist2_d(c, a);
  if (d < best) best = d;
  const double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double n[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
  const double nn = dot3d(n, n);
  if (nn > 0.0) {
    /* projection of the origin (the query) inside the triangle? */
    const double s = dot3d(n, a) / nn;
    const double p[3] = {s * n[0], s * n[1], s * n[2]};
    const double pa[3] = {a[0] - p[0], a[1] - p[1], a[2] - p[2]};
    const double pb[3] = {b[0] - p[0], b[1] - p[1], b[2] - p[2]};
    const double pc[3] = {c[0] - p[0], c[1] - p[1], c[2] - p[2]};
    double x[3];
    x[0] = pa[1] * pb[2] - pa[2] * pb[1];
    x[1] = pa[2] * pb[0] - pa[0] * pb[2];
    x[2] = pa[0] * pb[1] - pa[1] * pb[0];
    const double u = dot3d(x, n);
    x[0] = pb[1] * pc[2] - pb[2] * pc[1];
    x[1] = pb[2] * pc[0] - pb[0] * pc[2];
    x[2] = pb[0] * pc[1] - pb[1] * pc[0];
    const double v = dot3d(x, n);
    x[0] = pc[1] * pa[2] - pc[2] * pa[1];
    x[1] = pc[2] * pa[0] - pc[0] * pa[2];
    x[2] = pc[0] * pa[1] - pc[1] * pa[0];
    const double w = dot3d(x, n);
    if (u >= 0.0 && v >= 0.0 && w >= 0.0) {
      const double s2 = dot3d(n, a);
      const double pd = s2 * s2 / nn;
      if (pd < best) best = pd;
    }
  }
  return best;
}

/* ============================================================================
 * winding number / crossings / parity
 * ========================================================================== */
double orc_gwn_f64(const double q[3], const float *xyz, size_t ntri) {
  double sum = 0.0, comp = 0.0; /* Neumaier */
  for (size_t t = 0; t < ntri; ++t) {
    const float *v = xyz + 9 * t;
    const double a[3] = {(double)v[0] - q[0], (double)v[1] - q[1], (double)v[2] - q[2]};
    const double b[3] = {(double)v[3] - q[0], (double)v[4] - q[1], (double)v[5] - q[2]};
    const double c[3] = {(double)v[6] - q[0], (double)v[7] - q[1], (double)v[8] - q[2]};
    const double la = sqrt(dot3d(a, a)), lb = sqrt(dot3d(b, b)), lc = sqrt(dot3d(c, c));
    const double bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
    const double det = dot3d(a, bc);
    const double den = la * lb * lc + dot3d(a, b) * lc + dot3d(b, c) * la + dot3d(c, a) * lb;
    const double om = 2.0 * atan2(det, den);
    const double s = sum + om;
    if (fabs(sum) >= fabs(om))
      comp += (sum - s) + om;
    else
      comp += (om - s) + sum;
    sum = s;
  }
  return (sum + comp) / (4.0 * M_PI);
}

/* Half-open tie rule for an exactly-zero edge function: positive iff the edge direction D points
 * into the upper half plane (D.y > 0, or D.y == 0 and D.x > 0).  Antisymmetric in D, so the two
 * triangles sharing an edge never both claim a point on it. */
static inline int tie_pos(float dx, float dy) { return dy > 0.0f || (dy == 0.0f && dx > 0.0f); }
/* sign of e = px*qy - py*qx, evaluated in FP64: the float products are exact there, so the sign
 * is exact and swapping the edge's endpoints negates it exactly */
static inline int edge_sign(float px, float py, float qx, float qy, float dx, float dy) {
  const double ed = (double)px * (double)qy - (double)py * (double)qx;
  if (ed > 0.0) return 1;
  if (ed < 0.0) return -1;
  return tie_pos(dx, dy) ? 1 : -1;
}

/* Watertight ray / triangle (Woop, Benthin, Wald 2013).  Returns 0 for no hit, else the common
 * sign s of the edge functions; *t receives the hit distance (only meaningful if hit). */
typedef struct {
  int kx, ky, kz;
  float sx, sy, sz;
} ray_shear;

static void make_shear(const float d[3], ray_shear *r) {
  int kz = 0;
  if (fabsf(d[1]) > fabsf(d[kz])) kz = 1;
  if (fabsf(d[2]) > fabsf(d[kz])) kz = 2;
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  if (d[kz] < 0.0f) {
    const int tmp = kx;
    kx = ky;
    ky = tmp;
  }
  r->kx = kx;
  r->ky = ky;
  r->kz = kz;
  r->sx = d[kx] / d[kz];
  r->sy = d[ky] / d[kz];
  r->sz = 1.0f / d[kz];
}

static int ray_tri(const float o[3], const ray_shear *r, const float *v, double *tout) {
  float A[3], B[3], C[3];
  for (int k = 0; k < 3; ++k) {
    A[k] = v[k] - o[k];
    B[k] = v[3 + k] - o[k];
    C[k] = v[6 + k] - o[k];
  }
  const float ax = A[r->kx] - r->sx * A[r->kz], ay = A[r->ky] - r->sy * A[r->kz];
  const float bx = B[r->kx] - r->sx * B[r->kz], by = B[r->ky] - r->sy * B[r->kz];
  const float cx = C[r->kx] - r->sx * C[r->kz], cy = C[r->ky] - r->sy * C[r->kz];
  /* U: edge B->C, V: edge C->A, W: edge A->B */
  const int su = edge_sign(cx, cy, bx, by, cx - bx, cy - by);
  const int sv = edge_sign(ax, ay, cx, cy, ax - cx, ay - cy);
  const int sw = edge_sign(bx, by, ax, ay, bx - ax, by - ay);
  if (su != sv || sv != sw) return 0;
  const double U = (double)cx * by - (double)cy * bx;
  const double V = (double)ax * cy - (double)ay * cx;
  const double W = (double)bx * ay - (double)by * ax;
  const double det = U + V + W;
  if (det == 0.0) return 0;
  const double az = (double)r->sz * A[r->kz], bz = (double)r->sz * B[r->kz], cz = (double)r->sz * C[r->kz];
  *tout = (U * az + V * bz + W * cz) / det;
  return su;
}

int orc_crossings_up(const float q[3], const float *xyz, size_t ntri) {
  const float d[3] = {0.0f, 0.0f, 1.0f};
  ray_shear r;
  make_shear(d, &r);
  int w = 0;
  for (size_t t = 0; t < ntri; ++t) {
    double th;
    const int s = ray_tri(q, &r, xyz + 9 * t, &th);
    if (s != 0 && th > 0.0) w += -s; /* s<0 <=> normal.z>0 (exit) -> +1 */
  }
  return w;
}

void orc_parity_dirs(float dirs[15]) {
  /* +z first (sdf_bake.cu:56), then four fixed generic unit directions */
  static const float tab[15] = {0.0f,         0.0f,         1.0f,         0.5773502588f, 0.5773502588f,
                                -0.5773502588f, -0.6666666865f, 0.3333333433f, 0.6666666865f, 0.2672612369f,
                                -0.8017837405f, -0.5345224738f, -0.4850712717f, -0.7276068926f, 0.4850712717f};
  memcpy(dirs, tab, sizeof(tab));
}

void orc_parity_hits(const float q[3], const float *xyz, size_t ntri, int hits[5]) {
  float dirs[15];
  orc_parity_dirs(dirs);
  for (int v = 0; v < 5; ++v) {
    ray_shear r;
    make_shear(dirs + 3 * v, &r);
    hits[v] = 0;
    for (size_t t = 0; t < ntri; ++t) {
      double th;
      if (ray_tri(q, &r, xyz + 9 * t, &th) != 0 && th > 0.0) hits[v]++; /* every hit, sdf_bake.cu:58-65 */
    }
  }
}

int orc_parity_inside(const float q[3], const float *xyz, size_t ntri) {
  int hits[5], votes = 0;
  orc_parity_hits(q, xyz, ntri, hits);
  for (int v = 0; v < 5; ++v)
    if (hits[v] & 1) votes++; /* sdf_bake.cu:66-67 */
  return votes * 2 > 5;       /* sdf_bake.cu:69 */
}

int orc_gwn_sign(const float q[3], const float *xyz, size_t ntri, int *tie) {
  const double qd[3] = {q[0], q[1], q[2]};
  const double w = orc_gwn_f64(qd, xyz, ntri);
  const int t = fabs(fabs(w) - 0.5) < 1e-6;
  if (tie) *tie = t;
  return t ? orc_parity_inside(q, xyz, ntri) : fabs(w) > 0.5;
}

/* ============================================================================
 * LBVH (Karras 2012)
 * ========================================================================== */
static uint32_t expand10(uint32_t v) {
  v &= 0x3ffu;
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

static void centroid(const float *t, float c[3]) {
  for (int k = 0; k < 3; ++k) c[k] = ((t[k] + t[3 + k]) + t[6 + k]) / 3.0f;
}

void orc_morton_codes(const float *xyz, size_t n, uint32_t *codes) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < n; ++i) {
    float c[3];
    centroid(xyz + 9 * i, c);
    for (int k = 0; k < 3; ++k) {
      lo[k] = c[k] < lo[k] ? c[k] : lo[k];
      hi[k] = c[k] > hi[k] ? c[k] : hi[k];
    }
  }
  float ext[3];
  for (int k = 0; k < 3; ++k) ext[k] = hi[k] - lo[k];
  for (size_t i = 0; i < n; ++i) {
    float c[3];
    centroid(xyz + 9 * i, c);
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
      const float u = ext[k] > 0.0f ? (c[k] - lo[k]) / ext[k] : 0.5f;
      float s = u * 1024.0f;
      s = s < 0.0f ? 0.0f : (s > 1023.0f ? 1023.0f : s);
      q[k] = (uint32_t)s;
    }
    codes[i] = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
  }
}

static int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

typedef struct {
  const uint32_t *codes;
  int n;
} karras_ctx;

static int kdelta(const karras_ctx *k, int i, int j) {
  if (j < 0 || j >= k->n) return -1;
  const uint64_t a = ((uint64_t)k->codes[i] << 32) | (uint32_t)i;
  const uint64_t b = ((uint64_t)k->codes[j] << 32) | (uint32_t)j;
  return clz64(a ^ b);
}

static int cmp_pair(const void *pa, const void *pb) {
  const uint64_t a = *(const uint64_t *)pa, b = *(const uint64_t *)pb;
  return a < b ? -1 : (a > b ? 1 : 0);
}

int orc_bvh_build(const float *xyz, size_t ntri, orc_bvh *b) {
  memset(b, 0, sizeof(*b));
  const int n = (int)ntri;
  b->n = n;
  if (n <= 0) return -1;
  uint32_t *codes = (uint32_t *)malloc(sizeof(uint32_t) * n);
  orc_morton_codes(xyz, ntri, codes);
  /* stable sort by code: sort (code<<32 | index) */
  uint64_t *kv = (uint64_t *)malloc(sizeof(uint64_t) * n);
  for (int i = 0; i < n; ++i) kv[i] = ((uint64_t)codes[i] << 32) | (uint32_t)i;
  if (n > 1) qsort(kv, n, sizeof(uint64_t), cmp_pair);
  b->order = (int32_t *)malloc(sizeof(int32_t) * n);
  b->codes = (uint32_t *)malloc(sizeof(uint32_t) * n);
  for (int i = 0; i < n; ++i) {
    b->order[i] = (int32_t)(kv[i] & 0xffffffffu);
    b->codes[i] = (uint32_t)(kv[i] >> 32);
  }
  free(kv);
  free(codes);
  const int ni = n - 1;
  b->left = (int32_t *)malloc(sizeof(int32_t) * (ni > 0 ? ni : 1));
  b->right = (int32_t *)malloc(sizeof(int32_t) * (ni > 0 ? ni : 1));
  b->parent = (int32_t *)malloc(sizeof(int32_t) * (2 * n - 1));
  b->lo = (float *)malloc(sizeof(float) * 3 * (2 * n - 1));
  b->hi = (float *)malloc(sizeof(float) * 3 * (2 * n - 1));
  for (int i = 0; i < 2 * n - 1; ++i) b->parent[i] = -1;
  karras_ctx k = {b->codes, n};
  for (int i = 0; i < ni; ++i) {
    const int d = (kdelta(&k, i, i + 1) - kdelta(&k, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = kdelta(&k, i, i - d);
    int lmax = 2;
    while (kdelta(&k, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
      if (kdelta(&k, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = kdelta(&k, i, j);
    int s = 0, t = l;
    do {
      t = (t + 1) >> 1;
      if (kdelta(&k, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? d : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const int lc = (lo == gamma) ? ~gamma : gamma;
    const int rc = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
    b->left[i] = lc;
    b->right[i] = rc;
    b->parent[lc >= 0 ? lc : ni + ~lc] = i;
    b->parent[rc >= 0 ? rc : ni + ~rc] = i;
  }
  /* leaf boxes */
  for (int i = 0; i < n; ++i) {
    const float *t = xyz + 9 * (size_t)b->order[i];
    float *lo = b->lo + 3 * (ni + i), *hi = b->hi + 3 * (ni + i);
    for (int c = 0; c < 3; ++c) {
      lo[c] = smin(smin(t[c], t[3 + c]), t[6 + c]);
      hi[c] = smax(smax(t[c], t[3 + c]), t[6 + c]);
    }
  }
  /* refit: post-order by explicit stack */
  if (ni > 0) {
    int *st = (int *)malloc(sizeof(int) * (2 * n + 8));
    char *done = (char *)calloc(ni, 1);
    int sp = 0;
    st[sp++] = 0;
    while (sp) {
      const int x = st[sp - 1];
      const int c[2] = {b->left[x], b->right[x]};
      int pushed = 0;
      if (!done[x]) {
        done[x] = 1;
        for (int q = 0; q < 2; ++q)
          if (c[q] >= 0) {
            st[sp++] = c[q];
            pushed = 1;
          }
        if (pushed) continue;
      }
      sp--;
      const int ia = c[0] >= 0 ? c[0] : ni + ~c[0], ib = c[1] >= 0 ? c[1] : ni + ~c[1];
      for (int q = 0; q < 3; ++q) {
        b->lo[3 * x + q] = smin(b->lo[3 * ia + q], b->lo[3 * ib + q]);
        b->hi[3 * x + q] = smax(b->hi[3 * ia + q], b->hi[3 * ib + q]);
      }
    }
    free(st);
    free(done);
  }
  /* depth */
  int depth = 0;
  for (int i = 0; i < n; ++i) {
    int dd = 0, x = ni + i;
    while (b->parent[x] >= 0) {
      x = b->parent[x];
      dd++;
    }
    if (dd > depth) depth = dd;
  }
  b->depth = depth;
  return 0;
}

void orc_bvh_free(orc_bvh *b) {
  free(b->order);
  free(b->codes);
  free(b->left);
  free(b->right);
  free(b->parent);
  free(b->lo);
  free(b->hi);
  memset(b, 0, sizeof(*b));
}

/* ============================================================================
 * dense CPU bake
 * ========================================================================== */
/* squared point/box distance in the triangle distance's form fmaf(z, z, fmaf(y, y, x * x)) */
static inline float box_dist2(const float q[3], const float *lo, const float *hi) {
  float d[3];
  for (int k = 0; k < 3; ++k) d[k] = fmaxf(fmaxf(lo[k] - q[k], q[k] - hi[k]), 0.0f);
  return fmaf(d[2], d[2], fmaf(d[1], d[1], d[0] * d[0]));
}

typedef struct {
  const float *tris;       /* sorted soup, 9 floats */
  const orc_bvh *b;
  const orc_trirec *recs;  /* distance records of the sorted soup */
  float mp;                /* largest |vertex coordinate| */
} qctx;

static orc_trirec *make_recs(const float *sorted, size_t ntri) {
  orc_trirec *r = (orc_trirec *)malloc(sizeof(orc_trirec) * (ntri ? ntri : 1));
  for (size_t i = 0; i < ntri; ++i) orc_trirec_make(sorted + 9 * i, r + i);
  return r;
}

static float max_abs_coord(const float *xyz, size_t ntri) {
  float m = 0.0f;
  for (size_t i = 0; i < 9 * ntri; ++i) m = fmaxf(m, fabsf(xyz[i]));
  return m;
}

/* Pruning with FP32 distances: |sqrt(orc_trirec_dist2(q, t)) - dist(q, t)| < 2^-17 (Mq + 3 Mp) for
 * |q_k| <= Mq and |vertex_k| <= Mp (tests/test_oracle.py measures the worst case, about 4 u (Mq + 3 Mp)).
 * A box is entered when its squared distance is at most this bound on what can still beat `best`,
 * so the traversal's result is the minimum over all triangles, whatever the order. */
static inline float prune_bound(float best, const float q[3], float mp) {
  const float mq = fmaxf(fmaxf(fabsf(q[0]), fabsf(q[1])), fabsf(q[2]));
  const float r = sqrtf(best) + (mq + 3.0f * mp) * 7.62939453125e-6f;
  return (r * r) * 1.0000077f;
}

static float closest2(const qctx *c, const float q[3], uint64_t *nodes, uint64_t *tests) {
  const orc_bvh *b = c->b;
  const int ni = b->n - 1;
  float best = INFINITY;
  if (ni == 0) {
    (*tests)++;
    return orc_trirec_dist2(q, c->recs);
  }
  int stack[256];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const int x = stack[--sp];
    const int ch[2] = {b->left[x], b->right[x]};
    float dd[2];
    for (int k = 0; k < 2; ++k) {
      const int id = ch[k] >= 0 ? ch[k] : ni + ~ch[k];
      dd[k] = box_dist2(q, b->lo + 3 * id, b->hi + 3 * id);
      (*nodes)++;
    }
    const int first = dd[1] < dd[0] ? 1 : 0;
    const float bound = prune_bound(best, q, c->mp);
    for (int o = 1; o >= 0; --o) { /* push far first so near is popped first */
      const int k = o ? 1 - first : first;
      if (!(dd[k] <= bound)) continue;
      if (ch[k] < 0) {
        (*tests)++;
        const float d2 = orc_trirec_dist2(q, c->recs + (size_t)(~ch[k]));
        if (d2 < best) best = d2;
      } else {
        stack[sp++] = ch[k];
      }
    }
  }
  return best;
}

typedef struct {
  float z;
  int s;
} crossing;

static int cmp_cross(const void *a, const void *b) {
  const float za = ((const crossing *)a)->z, zb = ((const crossing *)b)->z;
  return za < zb ? -1 : (za > zb ? 1 : 0);
}

/* all crossings of the vertical line through (x,y): exact column z and sign */
static int column_crossings(const qctx *c, float x, float y, crossing **out, int *cap, uint64_t *tests) {
  const orc_bvh *b = c->b;
  const int ni = b->n - 1;
  const float o[3] = {x, y, 0.0f};
  const float d[3] = {0.0f, 0.0f, 1.0f};
  ray_shear r;
  make_shear(d, &r);
  int cnt = 0;
  int stack[256];
  int sp = 0;
  if (ni == 0) stack[sp++] = ~0;
  else stack[sp++] = 0;
  while (sp) {
    const int x0 = stack[--sp];
    const int id = x0 >= 0 ? x0 : ni + ~x0;
    const float *lo = b->lo + 3 * id, *hi = b->hi + 3 * id;
    if (!(x >= lo[0] && x <= hi[0] && y >= lo[1] && y <= hi[1])) continue;
    if (x0 >= 0) {
      stack[sp++] = b->left[x0];
      stack[sp++] = b->right[x0];
      continue;
    }
    (*tests)++;
    double th;
    const int s = ray_tri(o, &r, c->tris + 9 * (size_t)(~x0), &th);
    if (s == 0) continue;
    if (cnt == *cap) {
      *cap = *cap ? *cap * 2 : 16;
      *out = (crossing *)realloc(*out, sizeof(crossing) * (*cap));
    }
    (*out)[cnt].z = (float)th;
    (*out)[cnt].s = -s;
    cnt++;
  }
  if (cnt > 1) qsort(*out, (size_t)cnt, sizeof(crossing), cmp_cross);
  return cnt;
}

typedef struct {
  const qctx *c;
  const float *origin;
  float h;
  int nx, ny, nz, z0, z1, zstep;
  float *dist;
  int next_row;
  pthread_mutex_t *mu;
  uint64_t nodes, tests, ctests, inside;
} bake_job;

typedef struct {
  bake_job *job;
  uint64_t nodes, tests, ctests, inside;
} bake_thread;

static void *bake_thread_fn(void *arg) {
  bake_thread *th = (bake_thread *)arg;
  bake_job *j = th->job;
  crossing *cr = NULL;
  int cap = 0;
  for (;;) {
    pthread_mutex_lock(j->mu);
    const int y = j->next_row++;
    pthread_mutex_unlock(j->mu);
    if (y >= j->ny) break;
    const float py = j->origin[1] + ((float)y + 0.5f) * j->h;
    for (int x = 0; x < j->nx; ++x) {
      const float px = j->origin[0] + ((float)x + 0.5f) * j->h;
      const int nc = column_crossings(j->c, px, py, &cr, &cap, &th->ctests);
      for (int z = j->z0; z < j->z1; z += j->zstep) {
        const float q[3] = {px, py, j->origin[2] + ((float)z + 0.5f) * j->h};
        const float d2 = closest2(j->c, q, &th->nodes, &th->tests);
        int w = 0;
        for (int k = 0; k < nc; ++k)
          if (cr[k].z > q[2]) w += cr[k].s;
        const float d = sqrtf(d2);
        const int inside = w != 0; /* |w| > 0.5: orientation-agnostic, like the parity sign */
        th->inside += (uint64_t)inside;
        j->dist[((size_t)z * j->ny + y) * j->nx + x] = inside ? -d : d;
      }
    }
  }
  free(cr);
  return NULL;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int orc_bake_dense(const float *xyz, size_t ntri, const float origin[3], float h, int nx, int ny, int nz, int z0,
                   int z1, int threads, float *dist, orc_bake_stats *st) {
  return orc_bake_dense_strided(xyz, ntri, origin, h, nx, ny, nz, z0, z1, 1, threads, dist, st);
}

int orc_bake_dense_strided(const float *xyz, size_t ntri, const float origin[3], float h, int nx, int ny, int nz,
                           int z0, int z1, int zstep, int threads, float *dist, orc_bake_stats *st) {
  if (zstep < 1) zstep = 1;
  if (orc_boundary_edge_count(xyz, ntri) != 0) return -1;
  const double t0 = now_s();
  orc_bvh b;
  if (orc_bvh_build(xyz, ntri, &b) != 0) return -2;
  float *sorted = (float *)malloc(sizeof(float) * 9 * ntri);
  for (size_t i = 0; i < ntri; ++i) memcpy(sorted + 9 * i, xyz + 9 * (size_t)b.order[i], 9 * sizeof(float));
  orc_trirec *recs = make_recs(sorted, ntri);
  qctx c = {sorted, &b, recs, max_abs_coord(sorted, ntri)};
  if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  if (z0 < 0) z0 = 0;
  if (z1 > nz || z1 < 0) z1 = nz;
  pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
  bake_job job = {&c, origin, h, nx, ny, nz, z0, z1, zstep, dist, 0, &mu, 0, 0, 0, 0};
  pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * threads);
  bake_thread *ths = (bake_thread *)calloc(threads, sizeof(bake_thread));
  for (int i = 0; i < threads; ++i) {
    ths[i].job = &job;
    pthread_create(&tid[i], NULL, bake_thread_fn, &ths[i]);
  }
  uint64_t nodes = 0, tests = 0, ctests = 0, inside = 0;
  for (int i = 0; i < threads; ++i) {
    pthread_join(tid[i], NULL);
    nodes += ths[i].nodes;
    tests += ths[i].tests;
    ctests += ths[i].ctests;
    inside += ths[i].inside;
  }
  const double t1 = now_s();
  if (st) {
    st->voxels = (uint64_t)nx * ny * (uint64_t)((z1 - z0 + zstep - 1) / zstep);
    st->node_visits = nodes;
    st->tri_tests = tests;
    st->column_tri_tests = ctests;
    st->inside = inside;
    st->seconds = t1 - t0;
    st->threads = threads;
  }
  free(tid);
  free(ths);
  free(sorted);
  free(recs);
  orc_bvh_free(&b);
  return 0;
}

typedef struct {
  const float *xyz;
  size_t ntri;
  const double *pts;
  size_t npts;
  double *d, *w;
  size_t next;
  pthread_mutex_t *mu;
} f64_job;

static void *f64_fn(void *arg) {
  f64_job *j = (f64_job *)arg;
  for (;;) {
    pthread_mutex_lock(j->mu);
    const size_t i0 = j->next;
    j->next += 64;
    pthread_mutex_unlock(j->mu);
    if (i0 >= j->npts) break;
    const size_t i1 = i0 + 64 < j->npts ? i0 + 64 : j->npts;
    for (size_t i = i0; i < i1; ++i) {
      const double *q = j->pts + 3 * i;
      double best = INFINITY;
      for (size_t t = 0; t < j->ntri; ++t) {
        const double d2 = orc_tri_dist2_f64(q, j->xyz + 9 * t);
        if (d2 < best) best = d2;
      }
      const double w = orc_gwn_f64(q, j->xyz, j->ntri);
      j->w[i] = w;
      j->d[i] = fabs(w) > 0.5 ? -sqrt(best) : sqrt(best);
    }
  }
  return NULL;
}

void orc_sdf_f64_points(const float *xyz, size_t ntri, const double *pts, size_t npts, int threads, double *out_d,
                        double *out_w) {
  if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
  f64_job j = {xyz, ntri, pts, npts, out_d, out_w, 0, &mu};
  pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * threads);
  for (int i = 0; i < threads; ++i) pthread_create(&tid[i], NULL, f64_fn, &j);
  for (int i = 0; i < threads; ++i) pthread_join(tid[i], NULL);
  free(tid);
}

/* ============================================================================
 * closedness: weld by exact value (-0 == +0), cancel a->b against b->a
 * ========================================================================== */
static uint32_t canon_bits(float f) {
  if (f == 0.0f) f = 0.0f;
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

int64_t orc_boundary_edge_count(const float *xyz, size_t ntri) {
  /* weld: map position -> id */
  const size_t nv = 3 * ntri;
  size_t cap = 16;
  while (cap < nv * 2) cap <<= 1;
  uint32_t *kx = (uint32_t *)malloc(cap * 4), *ky = (uint32_t *)malloc(cap * 4), *kz = (uint32_t *)malloc(cap * 4);
  int64_t *kid = (int64_t *)malloc(cap * 8);
  for (size_t i = 0; i < cap; ++i) kid[i] = -1;
  int64_t *vid = (int64_t *)malloc(nv * 8);
  int64_t nextid = 0;
  for (size_t i = 0; i < nv; ++i) {
    const uint32_t a = canon_bits(xyz[3 * i]), b = canon_bits(xyz[3 * i + 1]), c = canon_bits(xyz[3 * i + 2]);
    size_t s = mix64(((uint64_t)a << 32 | b) ^ mix64(c)) & (cap - 1);
    for (;;) {
      if (kid[s] == -1) {
        kx[s] = a;
        ky[s] = b;
        kz[s] = c;
        kid[s] = nextid++;
        break;
      }
      if (kx[s] == a && ky[s] == b && kz[s] == c) break;
      s = (s + 1) & (cap - 1);
    }
    vid[i] = kid[s];
  }
  free(kx);
  free(ky);
  free(kz);
  free(kid);
  /* net directed count per undirected edge */
  i64map m;
  map_init(&m, nv);
  size_t count = 0;
  for (size_t t = 0; t < ntri; ++t)
    for (int k = 0; k < 3; ++k) {
      const int64_t a = vid[3 * t + k], b = vid[3 * t + (k + 1) % 3];
      if (a == b) continue;
      const int64_t lo = a < b ? a : b, hi = a < b ? b : a;
      int ins;
      int64_t *v = map_get(&m, (lo << 31) ^ hi, &ins, &count);
      if (ins) *v = 0;
      *v += (a < b) ? 1 : -1;
    }
  int64_t bnd = 0;
  for (size_t i = 0; i < m.cap; ++i)
    if (m.keys[i] != -1) bnd += m.vals[i] < 0 ? -m.vals[i] : m.vals[i];
  map_free(&m);
  free(vid);
  return bnd;
}

/* FP32 brute force unsigned squared distance: min over all triangles of orc_tri_dist2_f32 */
void orc_udf2_f32_brute(const float *xyz, size_t ntri, const float *pts, size_t npts, float *out) {
  for (size_t i = 0; i < npts; ++i) {
    float best = INFINITY;
    for (size_t t = 0; t < ntri; ++t) {
      const float d2 = orc_tri_dist2_f32(pts + 3 * i, xyz + 9 * t);
      if (d2 < best) best = d2;
    }
    out[i] = best;
  }
}

/* FP32 unsigned distance at arbitrary points through the LBVH (same per-triangle values as the
 * brute force; pruning never drops a candidate with box distance <= best) */
typedef struct {
  const qctx *c;
  const float *pts;
  size_t npts;
  float *out;
  size_t next;
  pthread_mutex_t *mu;
} udf_job;

static void *udf_fn(void *arg) {
  udf_job *j = (udf_job *)arg;
  uint64_t nodes = 0, tests = 0;
  for (;;) {
    pthread_mutex_lock(j->mu);
    const size_t i0 = j->next;
    j->next += 256;
    pthread_mutex_unlock(j->mu);
    if (i0 >= j->npts) break;
    const size_t i1 = i0 + 256 < j->npts ? i0 + 256 : j->npts;
    for (size_t i = i0; i < i1; ++i) j->out[i] = sqrtf(closest2(j->c, j->pts + 3 * i, &nodes, &tests));
  }
  return NULL;
}

int orc_udf_points_f32(const float *xyz, size_t ntri, const float *pts, size_t npts, int threads, float *out) {
  orc_bvh b;
  if (orc_bvh_build(xyz, ntri, &b) != 0) return -1;
  float *sorted = (float *)malloc(sizeof(float) * 9 * ntri);
  for (size_t i = 0; i < ntri; ++i) memcpy(sorted + 9 * i, xyz + 9 * (size_t)b.order[i], 9 * sizeof(float));
  orc_trirec *recs = make_recs(sorted, ntri);
  qctx c = {sorted, &b, recs, max_abs_coord(sorted, ntri)};
  if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
  udf_job j = {&c, pts, npts, out, 0, &mu};
  pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * threads);
  for (int i = 0; i < threads; ++i) pthread_create(&tid[i], NULL, udf_fn, &j);
  for (int i = 0; i < threads; ++i) pthread_join(tid[i], NULL);
  free(tid);
  free(sorted);
  free(recs);
  orc_bvh_free(&b);
  return 0;
}

typedef struct {
  const qctx *c;
  const float *xyz;
  size_t ntri;
  const float *pts;
  size_t npts;
  float *out;
  uint8_t *tie;
  size_t next;
  pthread_mutex_t *mu;
} signed_job;

static void *signed_fn(void *arg) {
  signed_job *j = (signed_job *)arg;
  uint64_t nodes = 0, tests = 0;
  for (;;) {
    pthread_mutex_lock(j->mu);
    const size_t i0 = j->next;
    j->next += 16;
    pthread_mutex_unlock(j->mu);
    if (i0 >= j->npts) break;
    const size_t i1 = i0 + 16 < j->npts ? i0 + 16 : j->npts;
    for (size_t i = i0; i < i1; ++i) {
      const float *q = j->pts + 3 * i;
      const float d = sqrtf(closest2(j->c, q, &nodes, &tests));
      int tie = 0;
      const int inside = orc_gwn_sign(q, j->xyz, j->ntri, &tie);
      j->out[i] = inside ? -d : d;
      if (j->tie) j->tie[i] = (uint8_t)tie;
    }
  }
  return NULL;
}

typedef struct {
  const qctx *c;
  const float *pts;
  size_t npts;
  float *out;
  size_t next;
  pthread_mutex_t *mu;
} closed_job;

static void *closed_fn(void *arg) {
  closed_job *j = (closed_job *)arg;
  uint64_t nodes = 0, tests = 0;
  crossing *cr = NULL;
  int cap = 0;
  for (;;) {
    pthread_mutex_lock(j->mu);
    const size_t i0 = j->next;
    j->next += 64;
    pthread_mutex_unlock(j->mu);
    if (i0 >= j->npts) break;
    const size_t i1 = i0 + 64 < j->npts ? i0 + 64 : j->npts;
    for (size_t i = i0; i < i1; ++i) {
      const float *q = j->pts + 3 * i;
      const float d = sqrtf(closest2(j->c, q, &nodes, &tests));
      const int nc = column_crossings(j->c, q[0], q[1], &cr, &cap, &tests);
      int w = 0;
      for (int k = 0; k < nc; ++k)
        if (cr[k].z > q[2]) w += cr[k].s;
      j->out[i] = w != 0 ? -d : d;
    }
  }
  free(cr);
  return NULL;
}

int orc_signed_points_closed(const float *xyz, size_t ntri, const float *pts, size_t npts, int threads, float *out) {
  if (ntri == 0) return -2;
  if (orc_boundary_edge_count(xyz, ntri) != 0) return -1;
  orc_bvh b;
  if (orc_bvh_build(xyz, ntri, &b) != 0) return -2;
  float *sorted = (float *)malloc(sizeof(float) * 9 * ntri);
  for (size_t i = 0; i < ntri; ++i) memcpy(sorted + 9 * i, xyz + 9 * (size_t)b.order[i], 9 * sizeof(float));
  orc_trirec *recs = make_recs(sorted, ntri);
  qctx c = {sorted, &b, recs, max_abs_coord(sorted, ntri)};
  if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
  closed_job j = {&c, pts, npts, out, 0, &mu};
  pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * threads);
  for (int i = 0; i < threads; ++i) pthread_create(&tid[i], NULL, closed_fn, &j);
  for (int i = 0; i < threads; ++i) pthread_join(tid[i], NULL);
  free(tid);
  free(sorted);
  free(recs);
  orc_bvh_free(&b);
  return 0;
}

int orc_signed_points_gwn(const float *xyz, size_t ntri, const float *pts, size_t npts, int threads, float *out,
                          uint8_t *tie) {
  orc_bvh b;
  if (ntri == 0 || orc_bvh_build(xyz, ntri, &b) != 0) return -1;
  float *sorted = (float *)malloc(sizeof(float) * 9 * ntri);
  for (size_t i = 0; i < ntri; ++i) memcpy(sorted + 9 * i, xyz + 9 * (size_t)b.order[i], 9 * sizeof(float));
  orc_trirec *recs = make_recs(sorted, ntri);
  qctx c = {sorted, &b, recs, max_abs_coord(sorted, ntri)};
  if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
  signed_job j = {&c, xyz, ntri, pts, npts, out, tie, 0, &mu};
  pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * threads);
  for (int i = 0; i < threads; ++i) pthread_create(&tid[i], NULL, signed_fn, &j);
  for (int i = 0; i < threads; ++i) pthread_join(tid[i], NULL);
  free(tid);
  free(sorted);
  free(recs);
  orc_bvh_free(&b);
  return 0;
}

/* ============================================================================
 * the grid consumer: pathtracer.cu:684-709 (sampleSdf, sdfNearestCellIndex), :798-846 (cubic
 * B-spline gradient normal), :587-613 + :740-792 (slab test + SDF march)
 * ========================================================================== */
typedef struct {
  float o[3], h;
  int n[3];
  const float *d;
} gview;

static void local_coords(const gview *g, const float p[3], float l[3]) {
  for (int k = 0; k < 3; ++k) l[k] = (p[k] - g->o[k]) / g->h - 0.5f; /* :685 */
}
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline float gat(const gview *g, int x, int y, int z) { return g->d[((size_t)z * g->n[1] + y) * g->n[0] + x]; }

static float sample_sdf(const gview *g, const float p[3]) { /* :684-701 */
  float l[3];
  local_coords(g, p, l);
  const int x0 = imax(0, imin(g->n[0] - 2, (int)floorf(l[0])));
  const int y0 = imax(0, imin(g->n[1] - 2, (int)floorf(l[1])));
  const int z0 = imax(0, imin(g->n[2] - 2, (int)floorf(l[2])));
  const float fx = fminf(fmaxf(l[0] - (float)x0, 0.0f), 1.0f);
  const float fy = fminf(fmaxf(l[1] - (float)y0, 0.0f), 1.0f);
  const float fz = fminf(fmaxf(l[2] - (float)z0, 0.0f), 1.0f);
  const float c00 = gat(g, x0, y0, z0) * (1.0f - fx) + gat(g, x0 + 1, y0, z0) * fx;
  const float c10 = gat(g, x0, y0 + 1, z0) * (1.0f - fx) + gat(g, x0 + 1, y0 + 1, z0) * fx;
  const float c01 = gat(g, x0, y0, z0 + 1) * (1.0f - fx) + gat(g, x0 + 1, y0, z0 + 1) * fx;
  const float c11 = gat(g, x0, y0 + 1, z0 + 1) * (1.0f - fx) + gat(g, x0 + 1, y0 + 1, z0 + 1) * fx;
  const float c0 = c00 * (1.0f - fy) + c10 * fy;
  const float c1 = c01 * (1.0f - fy) + c11 * fy;
  return c0 * (1.0f - fz) + c1 * fz;
}

static int nearest_cell(const gview *g, const float p[3]) { /* :703-709 */
  float l[3];
  local_coords(g, p, l);
  const int x = imax(0, imin(g->n[0] - 1, (int)lroundf(l[0])));
  const int y = imax(0, imin(g->n[1] - 1, (int)lroundf(l[1])));
  const int z = imax(0, imin(g->n[2] - 1, (int)lroundf(l[2])));
  return (z * g->n[1] + y) * g->n[0] + x;
}

static void bspline(float t, float w[4], float dw[4]) { /* :794-804 */
  const float t2 = t * t, t3 = t2 * t;
  w[0] = (1.0f - 3.0f * t + 3.0f * t2 - t3) * (1.0f / 6.0f);
  w[1] = (4.0f - 6.0f * t2 + 3.0f * t3) * (1.0f / 6.0f);
  w[2] = (1.0f + 3.0f * t + 3.0f * t2 - 3.0f * t3) * (1.0f / 6.0f);
  w[3] = t3 * (1.0f / 6.0f);
  dw[0] = (-1.0f + 2.0f * t - t2) * 0.5f;
  dw[1] = (-4.0f * t + 3.0f * t2) * 0.5f;
  dw[2] = (1.0f + 2.0f * t - 3.0f * t2) * 0.5f;
  dw[3] = t2 * 0.5f;
}

static void gradient_normal(const gview *g, const float p[3], float out[3]) { /* :806-846 */
  float l[3];
  local_coords(g, p, l);
  const int bx = (int)floorf(l[0]), by = (int)floorf(l[1]), bz = (int)floorf(l[2]);
  float wx[4], dwx[4], wy[4], dwy[4], wz[4], dwz[4];
  bspline(l[0] - (float)bx, wx, dwx);
  bspline(l[1] - (float)by, wy, dwy);
  bspline(l[2] - (float)bz, wz, dwz);
  int ix[4];
  for (int i = 0; i < 4; ++i) ix[i] = imax(0, imin(g->n[0] - 1, bx - 1 + i));
  float gr[3] = {0.0f, 0.0f, 0.0f};
  for (int k = 0; k < 4; ++k) {
    const int zrow = imax(0, imin(g->n[2] - 1, bz - 1 + k)) * g->n[1];
    float pl[3] = {0.0f, 0.0f, 0.0f};
    for (int j = 0; j < 4; ++j) {
      const size_t row = (size_t)(zrow + imax(0, imin(g->n[1] - 1, by - 1 + j))) * g->n[0];
      float v = 0.0f, dv = 0.0f;
      for (int i = 0; i < 4; ++i) {
        const float f = g->d[row + ix[i]];
        v += wx[i] * f;
        dv += dwx[i] * f;
      }
      pl[0] += dv * wy[j];
      pl[1] += v * dwy[j];
      pl[2] += v * wy[j];
    }
    gr[0] += pl[0] * wz[k];
    gr[1] += pl[1] * wz[k];
    gr[2] += pl[2] * dwz[k];
  }
  const float inv = 1.0f / sqrtf((gr[0] * gr[0] + gr[1] * gr[1]) + gr[2] * gr[2]);
  for (int k = 0; k < 3; ++k) out[k] = gr[k] * inv;
}

static float trace_grid(const gview *g, const float o[3], const float d[3], float t0, float t1) {
  float lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = g->o[k];
    hi[k] = g->o[k] + (float)g->n[k] * g->h; /* boundsMax = origin + n * voxelSize, :742 */
  }
  for (int a = 0; a < 3; ++a) { /* slabTest, :587-613 */
    const float inv = 1.0f / d[a];
    float tn = (lo[a] - o[a]) * inv, tf = (hi[a] - o[a]) * inv;
    if (tn > tf) {
      const float tmp = tn;
      tn = tf;
      tf = tmp;
    }
    if (tn > t0) t0 = tn;
    if (tf < t1) t1 = tf;
    if (t0 > t1) return -1.0f;
  }
  const float eps = g->h * 0.02f, min_step = g->h * 0.05f, min_adv = g->h * 0.3f; /* :757-759 */
  float t = t0, prev_t = t0, p[3];
  for (int k = 0; k < 3; ++k) p[k] = o[k] + t0 * d[k];
  float prev = sample_sdf(g, p);
  t += fmaxf(fabsf(prev), min_step);
  for (int i = 1; i < 512 && t <= t1; ++i) {
    for (int k = 0; k < 3; ++k) p[k] = o[k] + t * d[k];
    const float dist = sample_sdf(g, p);
    const int crossed = (prev > 0.0f) != (dist > 0.0f);
    if ((crossed || fabsf(dist) < eps) && (t - t0) > min_adv) {
      float th = t;
      if (crossed) { /* 8 bisections, then a secant step (:769-784) */
        float a = prev_t, b = t, fa = prev, fb = dist;
        for (int k = 0; k < 8; ++k) {
          const float m = 0.5f * (a + b);
          float pm[3];
          for (int c = 0; c < 3; ++c) pm[c] = o[c] + m * d[c];
          const float fm = sample_sdf(g, pm);
          if ((fa > 0.0f) != (fm > 0.0f)) {
            b = m;
            fb = fm;
          } else {
            a = m;
            fa = fm;
          }
        }
        const float den = fa - fb;
        th = fabsf(den) > 1e-9f ? a + (b - a) * (fa / den) : 0.5f * (a + b);
      }
      return fmaxf(th, t0);
    }
    prev_t = t;
    prev = dist;
    t += fmaxf(fabsf(dist) * 0.85f, min_step);
  }
  return -1.0f;
}

void orc_sample_grid(const float origin[3], float h, int nx, int ny, int nz, const float *dist, const float *pts,
                     size_t n, float *values, float *normals, int32_t *nearest) {
  const gview g = {{origin[0], origin[1], origin[2]}, h, {nx, ny, nz}, dist};
  for (size_t i = 0; i < n; ++i) {
    const float *p = pts + 3 * i;
    if (values) values[i] = sample_sdf(&g, p);
    if (normals) gradient_normal(&g, p, normals + 3 * i);
    if (nearest) nearest[i] = nearest_cell(&g, p);
  }
}

void orc_trace_grid(const float origin[3], float h, int nx, int ny, int nz, const float *dist, const float *org,
                    const float *dir, size_t n, float tmin, float tmax, float *t_hit) {
  const gview g = {{origin[0], origin[1], origin[2]}, h, {nx, ny, nz}, dist};
  for (size_t i = 0; i < n; ++i) t_hit[i] = trace_grid(&g, org + 3 * i, dir + 3 * i, tmin, tmax);
}

/* ---- analytic SDF bake (src/convert/sdf_procedural.cpp, sdf_primitives.h) ------------------------ */

static inline float glen3(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z); }
static inline float glen2(float x, float y) { return sqrtf(x * x + y * y); }

/* sdf_primitives.h:10 */
static float sd_sphere(float x, float y, float z, float r) { return glen3(x, y, z) - r; }
/* :12-15; std::max({qx, qy, qz}) keeps the first largest */
static float sd_box(float x, float y, float z, float bx, float by, float bz) {
  const float qx = fabsf(x) - bx, qy = fabsf(y) - by, qz = fabsf(z) - bz;
  float m = qx;
  if (m < qy) m = qy;
  if (m < qz) m = qz;
  return glen3(smax(qx, 0.0f), smax(qy, 0.0f), smax(qz, 0.0f)) + smin(m, 0.0f);
}
/* :17-20 */
static float sd_cylinder(float x, float y, float z, float r, float h) {
  const float dx = fabsf(glen2(x, z)) - r, dy = fabsf(y) - h;
  return smin(smax(dx, dy), 0.0f) + glen2(smax(dx, 0.0f), smax(dy, 0.0f));
}
/* :22-25 */
static float sd_torus(float x, float y, float z, float major, float minor) {
  return glen2(glen2(x, z) - major, y) - minor;
}
/* :31-34: glm::clamp = min(max(x, lo), hi); glm::mix(b, a, h) = b * (1 - h) + a * h */
static float op_smooth_union(float a, float b, float k) {
  const float h = smin(smax(0.5f + 0.5f * (b - a) / k, 0.0f), 1.0f);
  return (b * (1.0f - h) + a * h) - k * h * (1.0f - h);
}
/* :36 */
static float op_smooth_intersect(float a, float b, float k) { return -op_smooth_union(-a, -b, k); }

float orc_csg_eval(const orc_csg_op *prog, int nops, const float p[3], int *branch) {
  float sd[16];
  int sb[16];
  int sp = 0;
  for (int i = 0; i < nops; ++i) {
    const orc_csg_op *o = &prog[i];
    const float *a = o->a;
    const float x = p[0] - a[0], y = p[1] - a[1], z = p[2] - a[2];
    float d;
    switch (o->op) {
      case 1: d = sd_sphere(x, y, z, a[3]); break;
      case 2: d = sd_box(x, y, z, a[3], a[4], a[5]); break;
      case 3: d = sd_cylinder(x, y, z, a[3], a[4]); break;
      case 4: d = sd_torus(x, y, z, a[3], a[4]); break;
      case 5: d = ((a[3] * p[0] + a[4] * p[1]) + a[5] * p[2]) - a[6]; break;
      default: {
        if (sp < 2) return NAN;
        const float db = sd[--sp];
        const int bb = sb[sp];
        const float da = sd[sp - 1];
        const int ba = sb[sp - 1];
        float r = da;
        int rb = ba;
        switch (o->op) {
          case 10: if (db < da) { r = db; rb = bb; } break;         /* std::min (:27) */
          case 11: r = smax(da, -db); break;                          /* std::max(a, -b) (:29) */
          case 12: if (da < db) { r = db; rb = bb; } break;          /* std::max (:31) */
          case 13: r = op_smooth_union(da, db, a[0]); break;
          case 14: r = op_smooth_intersect(da, db, a[0]); break;
          case 15: if (!(da <= db)) { r = db; rb = bb; } break;      /* sdf_procedural.cpp:129 */
          default: return NAN;
        }
        sd[sp - 1] = r;
        sb[sp - 1] = rb;
        continue;
      }
    }
    if (sp >= 16) return NAN;
    sd[sp] = d;
    sb[sp] = o->branch;
    ++sp;
  }
  if (sp != 1) return NAN;
  if (branch) *branch = sb[0];
  return sd[0];
}

/* sdf_procedural.cpp:12-26 */
void orc_analytic_framing(const float bmin[3], const float bmax[3], int res, float origin[3], float *h, int32_t n[3]) {
  const float ex = bmax[0] - bmin[0], ey = bmax[1] - bmin[1], ez = bmax[2] - bmin[2];
  float longest = ex;
  if (longest < ey) longest = ey;
  if (longest < ez) longest = ez;
  const float vs = longest / (float)(res > 4 ? res : 4);
  *h = vs;
  origin[0] = bmin[0] - 2.0f * vs;
  origin[1] = bmin[1] - 2.0f * vs;
  origin[2] = bmin[2] - 2.0f * vs;
  n[0] = (int32_t)ceilf(ex / vs) + 4;
  n[1] = (int32_t)ceilf(ey / vs) + 4;
  n[2] = (int32_t)ceilf(ez / vs) + 4;
}

typedef struct {
  const orc_csg_op *prog;
  int nops;
  const float *origin;
  float h;
  int nx, ny, nz, z0, z1;
  float *dist;
  uint8_t *branch;
} csg_slab;

static void *csg_slab_run(void *arg) {
  const csg_slab *s = (const csg_slab *)arg;
  for (int z = s->z0; z < s->z1; ++z)
    for (int y = 0; y < s->ny; ++y)
      for (int x = 0; x < s->nx; ++x) {
        /* origin + (vec3(x, y, z) + 0.5) * voxelSize (:32-35) */
        const float p[3] = {s->origin[0] + ((float)x + 0.5f) * s->h, s->origin[1] + ((float)y + 0.5f) * s->h,
                            s->origin[2] + ((float)z + 0.5f) * s->h};
        int b = 0;
        const size_t i = ((size_t)z * s->ny + y) * s->nx + x;
        s->dist[i] = orc_csg_eval(s->prog, s->nops, p, &b);
        if (s->branch) s->branch[i] = (uint8_t)b;
      }
  return NULL;
}

void orc_bake_csg(const orc_csg_op *prog, int nops, const float origin[3], float h, int nx, int ny, int nz,
                  int threads, float *dist, uint8_t *branch) {
  if (threads < 1) threads = 1;
  if (threads > 64) threads = 64;
  pthread_t tid[64];
  csg_slab sl[64];
  for (int t = 0; t < threads; ++t) {
    sl[t] = (csg_slab){prog, nops, origin, h, nx, ny, nz, (int)((int64_t)nz * t / threads),
                       (int)((int64_t)nz * (t + 1) / threads), dist, branch};
    if (threads > 1) pthread_create(&tid[t], NULL, csg_slab_run, &sl[t]);
    else csg_slab_run(&sl[t]);
  }
  if (threads > 1)
    for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
}

/* makeCupSdf / makeGlassCupSdf (sdf_procedural.cpp:81-140), the lambdas written out per cell */
void orc_cup_cells(int glass, int res, float origin_y, float cx, float cz, float origin[3], float *h, int32_t n[3],
                   float *dist, uint8_t *branch) {
  const float R = 0.55f, wall = 0.09f, ri = R - wall, lift = 0.16f, rim_frac = 0.62f, k = 0.012f;
  const float oc[3] = {cx, origin_y + R, cz};
  const float ic[3] = {oc[0] + 0.0f, oc[1] + lift, oc[2] + 0.0f};
  const float rim_y = oc[1] + R * rim_frac;
  const float fill_y = ic[1] - ri + wall * 0.5f + 0.42f * (2.0f * ri);
  const float bmin[3] = {oc[0] - R, oc[1] - R, oc[2] - R};
  const float bmax[3] = {oc[0] + R, rim_y, oc[2] + R};
  orc_analytic_framing(bmin, bmax, res, origin, h, n);
  if (!dist) return;
  const float vs = *h;
  for (int z = 0; z < n[2]; ++z)
    for (int y = 0; y < n[1]; ++y)
      for (int x = 0; x < n[0]; ++x) {
        const float p[3] = {origin[0] + ((float)x + 0.5f) * vs, origin[1] + ((float)y + 0.5f) * vs,
                            origin[2] + ((float)z + 0.5f) * vs};
        const float d_outer = sd_sphere(p[0] - oc[0], p[1] - oc[1], p[2] - oc[2], R);
        const float d_inner = sd_sphere(p[0] - ic[0], p[1] - ic[1], p[2] - ic[2], ri);
        const float d_shell = smax(smax(d_outer, -d_inner), p[1] - rim_y);
        float d = d_shell;
        int b = 0;
        if (glass) {
          const float d_liquid = op_smooth_intersect(d_inner, p[1] - fill_y, k);
          if (!(d_shell <= d_liquid)) {
            d = d_liquid;
            b = 1;
          }
        }
        const size_t i = ((size_t)z * n[1] + y) * n[0] + x;
        dist[i] = d;
        if (branch) branch[i] = (uint8_t)b;
      }
}

/* Largest (exact - FP32) and (FP32 - exact) distance of orc_trirec_dist2 over n seeded random
 * non-sliver triangles and queries (coordinates across six decades of scale, triangle sizes across
 * four, queries far, near and very near), in units of u (Mq + 3 Mp), u = 2^-24: the constant behind
 * the traversal's pruning slack. */
void orc_f32_error_ratio(long n, uint64_t seed, double out[2]) {
  uint64_t s = seed;
  const double u = ldexp(1.0, -24);
  out[0] = out[1] = 0.0;
#define RND() ((s = s * 6364136223846793005ULL + 1442695040888963407ULL), (double)(s >> 11) * (1.0 / 9007199254740992.0))
  for (long it = 0; it < n; ++it) {
    const double S = pow(10.0, -3 + 6 * RND());
    const double size = S * pow(10.0, -4 + 4 * RND());
    double c[3];
    float t[9];
    for (int k = 0; k < 3; ++k) c[k] = S * (2 * RND() - 1);
    for (int v = 0; v < 3; ++v)
      for (int k = 0; k < 3; ++k) t[3 * v + k] = (float)(c[k] + size * (2 * RND() - 1));
    if (orc_tri_is_sliver(t)) continue;
    orc_trirec r;
    orc_trirec_make(t, &r);
    float q[3];
    const int mode = (int)(it % 3);
    for (int k = 0; k < 3; ++k)
      q[k] = (float)(mode == 0 ? S * (2 * RND() - 1) : c[k] + (mode == 1 ? size : size * 1e-3) * (2 * RND() - 1));
    const double qd[3] = {q[0], q[1], q[2]};
    const double d64 = sqrt(orc_tri_dist2_f64(qd, t));
    const double d32 = sqrt((double)orc_trirec_dist2(q, &r));
    double mq = fmax(fabs(q[0]), fmax(fabs(q[1]), fabs(q[2]))), mp = 0.0;
    for (int k = 0; k < 9; ++k) mp = fmax(mp, fabs(t[k]));
    const double scale = u * (mq + 3.0 * mp);
    out[0] = fmax(out[0], (d64 - d32) / scale);
    out[1] = fmax(out[1], (d32 - d64) / scale);
  }
#undef RND
}

/* ---- mesh ingestion gate --------------------------------------------------------------------------
 * weldPositions (src/io/mesh_weld.cpp:26-48): pitch cell = std::max(glm::length(bmax - bmin) * 1e-5f, 1e-6f)
 * with glm::length = sqrt((x*x + y*y) + z*z); key = std::lround(p / cell) per axis; the map assigns
 * weldMap.size() at the first insertion of a key, so indices count distinct keys in order of first
 * occurrence. */
static float weld_cell(const float bmin[3], const float bmax[3]) {
  const float ex = bmax[0] - bmin[0], ey = bmax[1] - bmin[1], ez = bmax[2] - bmin[2];
  const float diag = sqrtf((ex * ex + ey * ey) + ez * ez);
  const float a = diag * 1e-5f, b = 1e-6f;
  return a < b ? b : a; /* std::max(a, b) */
}

typedef struct {
  int64_t k[3];
} weld_key;

int64_t orc_weld_positions(const float *pos, size_t npos, const float bmin[3], const float bmax[3], uint32_t *welded,
                           float *unique) {
  const float cell = weld_cell(bmin, bmax);
  weld_key *keys = (weld_key *)malloc(sizeof(weld_key) * (npos ? npos : 1));
  for (size_t i = 0; i < npos; ++i)
    for (int a = 0; a < 3; ++a) {
      const float q = pos[3 * i + a] / cell;
      if (!(fabsf(q) < 0x1p63f)) { /* lround's result is unspecified here */
        free(keys);
        return -1;
      }
      keys[i].k[a] = (int64_t)lroundf(q);
    }
  size_t cap = 16;
  while (cap < 2 * npos) cap <<= 1;
  int64_t *slot = (int64_t *)malloc(sizeof(int64_t) * cap); /* index of the first position with the key */
  uint32_t *sid = (uint32_t *)malloc(sizeof(uint32_t) * cap);
  for (size_t s = 0; s < cap; ++s) slot[s] = -1;
  uint32_t next = 0;
  for (size_t i = 0; i < npos; ++i) {
    const weld_key *k = &keys[i];
    size_t s = mix64((uint64_t)k->k[0] ^ mix64((uint64_t)k->k[1] ^ mix64((uint64_t)k->k[2]))) & (cap - 1);
    for (;;) {
      if (slot[s] < 0) {
        slot[s] = (int64_t)i;
        sid[s] = next++;
        if (unique) memcpy(unique + 3 * (size_t)sid[s], pos + 3 * i, 3 * sizeof(float));
        break;
      }
      const weld_key *o = &keys[slot[s]];
      if (o->k[0] == k->k[0] && o->k[1] == k->k[1] && o->k[2] == k->k[2]) break;
      s = (s + 1) & (cap - 1);
    }
    welded[i] = sid[s];
  }
  free(slot);
  free(sid);
  free(keys);
  return (int64_t)next;
}

static int cmp_u64(const void *a, const void *b) {
  const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : x > y;
}

/* isWatertight (src/io/mesh_validate.cpp:21-59): undirected edges of the welded triangles, skipping
 * triangles with a repeated index; boundary = edges used once, nonmanifold = edges used more than twice. */
int orc_watertight(const float *xyz, size_t ntri, const float bmin[3], const float bmax[3], int64_t *boundary,
                   int64_t *nonmanifold) {
  *boundary = 0;
  *nonmanifold = 0;
  if (ntri == 0) return 0;
  uint32_t *w = (uint32_t *)malloc(sizeof(uint32_t) * 3 * ntri);
  if (orc_weld_positions(xyz, 3 * ntri, bmin, bmax, w, NULL) < 0) {
    free(w);
    return -1;
  }
  uint64_t *e = (uint64_t *)malloc(sizeof(uint64_t) * 3 * ntri);
  size_t ne = 0;
  for (size_t t = 0; t < ntri; ++t) {
    const uint32_t v[3] = {w[3 * t], w[3 * t + 1], w[3 * t + 2]};
    if (v[0] == v[1] || v[1] == v[2] || v[2] == v[0]) continue;
    for (int k = 0; k < 3; ++k) {
      const uint32_t a = v[k], b = v[(k + 1) % 3];
      e[ne++] = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
    }
  }
  if (ne > 1) qsort(e, ne, sizeof(uint64_t), cmp_u64);
  for (size_t i = 0; i < ne;) {
    size_t j = i;
    while (j < ne && e[j] == e[i]) ++j;
    if (j - i == 1) ++*boundary;
    else if (j - i > 2) ++*nonmanifold;
    i = j;
  }
  free(e);
  free(w);
  return (*boundary == 0 && *nonmanifold == 0) ? 1 : 0;
}

/* ---- FP64 unsigned distance through the LBVH (the FP64 oracle at sizes brute force cannot reach) ----
 * Same minimum as the brute force over every triangle: a subtree is skipped only when its box is farther
 * than the best FP64 distance so far by a relative margin (1e-9) far above the rounding of the FP64 box
 * distance, so no triangle that could hold the minimum is skipped. */
typedef struct {
  const float *sorted;
  const orc_bvh *b;
  const double *pts;
  size_t npts;
  double *out;
  size_t next;
  pthread_mutex_t *mu;
} f64bvh_job;

static double box_d2_f64(const double q[3], const float *lo, const float *hi) {
  double s = 0.0;
  for (int k = 0; k < 3; ++k) {
    double g = 0.0;
    if (q[k] < (double)lo[k]) g = (double)lo[k] - q[k];
    else if (q[k] > (double)hi[k]) g = q[k] - (double)hi[k];
    s += g * g;
  }
  return s;
}

static double closest2_f64(const float *sorted, const orc_bvh *b, const double q[3]) {
  const int n = b->n;
  if (n == 1) return orc_tri_dist2_f64(q, sorted);
  double best = INFINITY;
  int stack[256];
  int sp = 0;
  stack[sp++] = 0; /* internal node ids >= 0, leaves ~i */
  while (sp) {
    const int node = stack[--sp];
    if (node < 0) {
      const int leaf = ~node;
      const double d = orc_tri_dist2_f64(q, sorted + 9 * (size_t)leaf);
      if (d < best) best = d;
      continue;
    }
    const int ch[2] = {b->left[node], b->right[node]};
    double cd[2];
    for (int k = 0; k < 2; ++k) {
      const int box = ch[k] >= 0 ? ch[k] : (n - 1) + ~ch[k];
      cd[k] = box_d2_f64(q, b->lo + 3 * (size_t)box, b->hi + 3 * (size_t)box);
    }
    const int first = cd[1] < cd[0];
    for (int k = 1; k >= 0; --k) { /* push the farther child first */
      const int c = k ? first ^ 1 : first;
      if (cd[c] <= best * (1.0 + 1e-9) && sp < 256) stack[sp++] = ch[c];
    }
  }
  return best;
}

static void *f64bvh_fn(void *arg) {
  f64bvh_job *j = (f64bvh_job *)arg;
  for (;;) {
    pthread_mutex_lock(j->mu);
    const size_t i0 = j->next;
    j->next += 256;
    pthread_mutex_unlock(j->mu);
    if (i0 >= j->npts) break;
    const size_t i1 = i0 + 256 < j->npts ? i0 + 256 : j->npts;
    for (size_t i = i0; i < i1; ++i) j->out[i] = sqrt(closest2_f64(j->sorted, j->b, j->pts + 3 * i));
  }
  return NULL;
}

int orc_udf_points_f64(const float *xyz, size_t ntri, const double *pts, size_t npts, int threads, double *out) {
  orc_bvh b;
  if (ntri == 0 || orc_bvh_build(xyz, ntri, &b) != 0) return -1;
  if (b.depth >= 250) {
    orc_bvh_free(&b);
    return -2;
  }
  float *sorted = (float *)malloc(sizeof(float) * 9 * ntri);
  for (size_t i = 0; i < ntri; ++i) memcpy(sorted + 9 * i, xyz + 9 * (size_t)b.order[i], 9 * sizeof(float));
  if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
  f64bvh_job j = {sorted, &b, pts, npts, out, 0, &mu};
  pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * threads);
  for (int i = 0; i < threads; ++i) pthread_create(&tid[i], NULL, f64bvh_fn, &j);
  for (int i = 0; i < threads; ++i) pthread_join(tid[i], NULL);
  free(tid);
  free(sorted);
  orc_bvh_free(&b);
  return 0;
}

/* ---- per-triangle material (src/convert/sdf_baker.cpp:42-74; the same as triangleColor,
 * src/convert/voxelize.cpp:12-38, without metallic and roughness) ---------------------------------- */
static float srgb_to_linear(float c) { return c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f); }

void orc_triangle_materials(size_t ntri, const float *uvs, size_t nuv, const uint32_t *tri_material,
                            size_t ntri_material, const orc_material *materials, size_t nmat,
                            const orc_texture *textures, size_t ntex, float *out) {
  const orc_material def = {{1.0f, 1.0f, 1.0f}, 0.0f, 1.0f, -1};
  for (size_t t = 0; t < ntri; ++t) {
    /* materialForTriangle: mesh_asset.h:49-52 */
    const uint32_t mi = t < ntri_material ? tri_material[t] : 0u;
    const orc_material *mat = nmat ? &materials[mi < nmat ? mi : 0] : &def;
    float col[3] = {mat->base[0], mat->base[1], mat->base[2]};
    const size_t base = 3 * t;
    if (mat->tex >= 0 && (size_t)mat->tex < ntex && base + 2 < nuv) {
      const orc_texture *tx = &textures[mat->tex];
      if (tx->width > 0 && tx->height > 0) {
        /* (uv0 + uv1 + uv2) / 3.0f, component-wise */
        const float u = ((uvs[2 * base] + uvs[2 * base + 2]) + uvs[2 * base + 4]) / 3.0f;
        const float v = ((uvs[2 * base + 1] + uvs[2 * base + 3]) + uvs[2 * base + 5]) / 3.0f;
        const float wu = u - floorf(u), wv = v - floorf(v);
        int px = (int)(wu * (float)tx->width), py = (int)(wv * (float)tx->height);
        px = px < 0 ? 0 : (px > tx->width - 1 ? tx->width - 1 : px); /* std::clamp(x, 0, w - 1) */
        py = py < 0 ? 0 : (py > tx->height - 1 ? tx->height - 1 : py);
        const size_t idx = ((size_t)py * (size_t)tx->width + (size_t)px) * 4;
        for (int k = 0; k < 3; ++k) {
          const float raw = (float)tx->rgba[idx + k] / 255.0f;
          col[k] = (tx->srgb ? srgb_to_linear(raw) : raw) * mat->base[k];
        }
      }
    }
    out[5 * t] = col[0];
    out[5 * t + 1] = col[1];
    out[5 * t + 2] = col[2];
    out[5 * t + 3] = mat->metallic;
    out[5 * t + 4] = mat->roughness;
  }
}

/* ============================================================================
 * Narrow-band bake of any mesh (the precision-stress configuration's CPU baseline and checker).
 * Distance: the FP32 closest triangle through the LBVH (as orc_bake_dense).  Sign: the winding
 * number w = (signed +z crossings above q of the closed surface S + reversed cap) + w_cap(q), with
 * one cap for all holes: the cone from the mean of the boundary vertices P over every unmatched
 * directed edge u->v (triangle (P, u, v)), whose boundary is the mesh's (the spokes cancel along each
 * closed boundary loop).  w_cap in FP64: cap triangles near q summed exactly, farther LBVH nodes only
 * bounded (|Omega| <= area / dmin^2); when the bound leaves |w| within reach of 1/2 the whole cap is
 * summed; when q lies on the cap (where the identity does not hold) w is the mesh's own FP64 winding
 * number; |w| within 1e-6 of 1/2 is a tie, decided by orc_parity_inside.
 * ========================================================================== */
typedef struct {
  size_t n;          /* cap triangles */
  float *sorted;     /* 9 floats each, in the cap LBVH's leaf order */
  orc_bvh b;
  double *area;      /* per LBVH node (internal, then leaves): sum of triangle areas below */
  double *dip;       /* optional (triset_moments), 3 per node: sum of the area vectors below */
  double *ctr;       /* 3 per node: area-weighted centre of the triangles below */
  double *rad;       /* per node: largest distance from ctr to a vertex below */
} capset;

/* Unmatched directed edges (with multiplicity) after the exact-position weld, as (from, to) vertex ids
 * (2 per edge in *edges) and the position of each id (3 floats per id in *pos); returns the edge count. */
static size_t boundary_edges_ids(const float *xyz, size_t ntri, int64_t **edges, float **pos) {
  const size_t nv = 3 * ntri;
  size_t cap = 16;
  while (cap < nv * 2) cap <<= 1;
  uint32_t *kx = (uint32_t *)malloc(cap * 4), *ky = (uint32_t *)malloc(cap * 4), *kz = (uint32_t *)malloc(cap * 4);
  int64_t *kid = (int64_t *)malloc(cap * 8);
  for (size_t i = 0; i < cap; ++i) kid[i] = -1;
  int64_t *vid = (int64_t *)malloc((nv ? nv : 1) * 8);
  float *vp = (float *)malloc(sizeof(float) * 3 * (nv ? nv : 1));
  int64_t nextid = 0;
  for (size_t i = 0; i < nv; ++i) {
    const uint32_t a = canon_bits(xyz[3 * i]), b = canon_bits(xyz[3 * i + 1]), c = canon_bits(xyz[3 * i + 2]);
    size_t s = mix64(((uint64_t)a << 32 | b) ^ mix64(c)) & (cap - 1);
    for (;;) {
      if (kid[s] == -1) {
        kx[s] = a;
        ky[s] = b;
        kz[s] = c;
        memcpy(vp + 3 * nextid, xyz + 3 * i, 3 * sizeof(float));
        kid[s] = nextid++;
        break;
      }
      if (kx[s] == a && ky[s] == b && kz[s] == c) break;
      s = (s + 1) & (cap - 1);
    }
    vid[i] = kid[s];
  }
  free(kx);
  free(ky);
  free(kz);
  free(kid);
  i64map m;
  map_init(&m, nv);
  size_t count = 0;
  for (size_t t = 0; t < ntri; ++t)
    for (int k = 0; k < 3; ++k) {
      const int64_t a = vid[3 * t + k], b = vid[3 * t + (k + 1) % 3];
      if (a == b) continue;
      const int64_t lo = a < b ? a : b, hi = a < b ? b : a;
      int ins;
      int64_t *v = map_get(&m, (lo << 31) ^ hi, &ins, &count);
      if (ins) *v = 0;
      *v += (a < b) ? 1 : -1;
    }
  size_t ne = 0;
  for (size_t i = 0; i < m.cap; ++i)
    if (m.keys[i] != -1) ne += (size_t)(m.vals[i] < 0 ? -m.vals[i] : m.vals[i]);
  int64_t *e = (int64_t *)malloc(sizeof(int64_t) * 2 * (ne ? ne : 1));
  size_t k = 0;
  for (size_t i = 0; i < m.cap; ++i) {
    if (m.keys[i] == -1 || m.vals[i] == 0) continue;
    const int64_t key = m.keys[i], lo = key >> 31, hi = key & ((1ll << 31) - 1);
    const int64_t from = m.vals[i] > 0 ? lo : hi, to = m.vals[i] > 0 ? hi : lo;
    const int64_t mult = m.vals[i] < 0 ? -m.vals[i] : m.vals[i];
    for (int64_t r = 0; r < mult; ++r, ++k) {
      e[2 * k] = from;
      e[2 * k + 1] = to;
    }
  }
  map_free(&m);
  free(vid);
  *edges = e;
  *pos = vp;
  return ne;
}

static int cmp_edge(const void *pa, const void *pb) {
  const int64_t *a = (const int64_t *)pa, *b = (const int64_t *)pb;
  return a[0] != b[0] ? (a[0] < b[0] ? -1 : 1) : (a[1] < b[1] ? -1 : (a[1] > b[1] ? 1 : 0));
}

static double tri_area_f64(const float *t) {
  const double e1[3] = {(double)t[3] - t[0], (double)t[4] - t[1], (double)t[5] - t[2]};
  const double e2[3] = {(double)t[6] - t[0], (double)t[7] - t[1], (double)t[8] - t[2]};
  const double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  return 0.5 * sqrt(dot3d(c, c));
}

/* The cap: the unmatched edges decompose into closed loops (every vertex has as many unmatched edges
 * leaving as entering it); each loop, walked from an unused edge by any unused edge leaving the current
 * vertex until it returns to its start, is closed by a fan from the mean of its vertices: triangles
 * (P, u, v) for its edges u -> v, whose boundary is the loop. */
static int triset_build(const float *tris, size_t nc, capset *cs);
static int capset_build(const float *xyz, size_t ntri, capset *cs) {
  memset(cs, 0, sizeof(*cs));
  int64_t *e;
  float *vp;
  const size_t ne = boundary_edges_ids(xyz, ntri, &e, &vp);
  if (ne == 0) {
    free(e);
    free(vp);
    return 0;
  }
  qsort(e, ne, 2 * sizeof(int64_t), cmp_edge);
  char *used = (char *)calloc(ne, 1);
  size_t *loop = (size_t *)malloc(sizeof(size_t) * ne);
  float *caps = (float *)malloc(sizeof(float) * 9 * ne);
  size_t nc = 0;
  size_t *cursor = (size_t *)malloc(sizeof(size_t) * ne); /* next edge to try in each vertex's run */
  for (size_t i = 0; i < ne; ++i) cursor[i] = i;
  for (size_t s0 = 0; s0 < ne; ++s0) {
    if (used[s0]) continue;
    size_t len = 0, cur = s0;
    const int64_t start = e[2 * s0];
    for (;;) {
      used[cur] = 1;
      loop[len++] = cur;
      const int64_t v = e[2 * cur + 1];
      if (v == start) break;
      /* first edge of v's run, then the first unused one in it */
      size_t lo = 0, hi = ne;
      while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if (e[2 * mid] < v) lo = mid + 1;
        else hi = mid;
      }
      size_t nxt = ne;
      for (size_t q = cursor[lo < ne ? lo : 0]; lo < ne && q < ne && e[2 * q] == v; ++q)
        if (!used[q]) {
          nxt = q;
          cursor[lo] = q + 1;
          break;
        }
      if (nxt == ne) break; /* cannot happen for a boundary of closed loops; the fan closes it anyway */
      cur = nxt;
    }
    double c[3] = {0, 0, 0};
    for (size_t q = 0; q < len; ++q)
      for (int k = 0; k < 3; ++k) c[k] += vp[3 * e[2 * loop[q]] + k];
    const float apex[3] = {(float)(c[0] / len), (float)(c[1] / len), (float)(c[2] / len)};
    for (size_t q = 0; q < len; ++q, ++nc) {
      memcpy(caps + 9 * nc, apex, 3 * sizeof(float));
      memcpy(caps + 9 * nc + 3, vp + 3 * e[2 * loop[q]], 3 * sizeof(float));
      memcpy(caps + 9 * nc + 6, vp + 3 * e[2 * loop[q] + 1], 3 * sizeof(float));
    }
  }
  free(cursor);
  free(used);
  free(loop);
  free(e);
  free(vp);
  const int st = triset_build(caps, nc, cs); /* nc >= 1: every unmatched edge gives one fan */
  free(caps);
  return st;
}

/* A tree over any triangle set, in the form the bounded winding number walks: the LBVH, the triangles in
 * its leaf order, and every node's total area. */
static int triset_build(const float *tris, size_t nc, capset *cs) {
  if (orc_bvh_build(tris, nc, &cs->b) != 0) return -1;
  cs->n = nc;
  if (nc == 0) return 0;
  cs->sorted = (float *)malloc(sizeof(float) * 9 * nc);
  for (size_t i = 0; i < nc; ++i) memcpy(cs->sorted + 9 * i, tris + 9 * (size_t)cs->b.order[i], 9 * sizeof(float));
  const int ni = (int)nc - 1;
  cs->area = (double *)calloc(2 * nc - 1, sizeof(double));
  for (size_t i = 0; i < nc; ++i) {
    const double a = tri_area_f64(cs->sorted + 9 * i);
    for (int x = ni + (int)i; x >= 0; x = cs->b.parent[x]) cs->area[x] += a;
  }
  return 0;
}

static void capset_free(capset *cs) {
  if (cs->n) orc_bvh_free(&cs->b);
  free(cs->sorted);
  free(cs->area);
  free(cs->dip);
  free(cs->ctr);
  free(cs->rad);
}

/* First moments of every node of a triset: area vectors, area-weighted centres (the box centre for a node
 * of zero area) and the radius about that centre, each triangle carried up its parent chain. */
static void triset_moments(capset *cs) {
  const size_t n = cs->n, nn = 2 * n - 1;
  const int ni = (int)n - 1;
  const orc_bvh *b = &cs->b;
  cs->dip = (double *)calloc(3 * nn, sizeof(double));
  cs->ctr = (double *)calloc(3 * nn, sizeof(double));
  cs->rad = (double *)calloc(nn, sizeof(double));
  double *acc = (double *)calloc(3 * nn, sizeof(double));
  for (size_t i = 0; i < n; ++i) {
    const float *t = cs->sorted + 9 * i;
    const double e1[3] = {(double)t[3] - t[0], (double)t[4] - t[1], (double)t[5] - t[2]};
    const double e2[3] = {(double)t[6] - t[0], (double)t[7] - t[1], (double)t[8] - t[2]};
    const double nv[3] = {0.5 * (e1[1] * e2[2] - e1[2] * e2[1]), 0.5 * (e1[2] * e2[0] - e1[0] * e2[2]),
                          0.5 * (e1[0] * e2[1] - e1[1] * e2[0])};
    const double a = sqrt(dot3d(nv, nv));
    for (int x = ni + (int)i; x >= 0; x = b->parent[x])
      for (int k = 0; k < 3; ++k) {
        cs->dip[3 * x + k] += nv[k];
        acc[3 * x + k] += a * (((double)t[k] + t[3 + k]) + t[6 + k]) / 3.0;
      }
  }
  for (size_t x = 0; x < nn; ++x)
    for (int k = 0; k < 3; ++k)
      cs->ctr[3 * x + k] = cs->area[x] > 0.0 ? acc[3 * x + k] / cs->area[x]
                                             : 0.5 * ((double)b->lo[3 * x + k] + b->hi[3 * x + k]);
  free(acc);
  for (size_t i = 0; i < n; ++i) {
    const float *t = cs->sorted + 9 * i;
    for (int x = ni + (int)i; x >= 0; x = b->parent[x])
      for (int j = 0; j < 3; ++j) {
        const double d[3] = {(double)t[3 * j] - cs->ctr[3 * x], (double)t[3 * j + 1] - cs->ctr[3 * x + 1],
                             (double)t[3 * j + 2] - cs->ctr[3 * x + 2]};
        const double r = sqrt(dot3d(d, d));
        if (r > cs->rad[x]) cs->rad[x] = r;
      }
  }
  for (size_t x = 0; x < nn; ++x) cs->rad[x] *= 1.0 + 1e-12;
}

/* FP64 solid angle of t at q, and whether q lies on or next to t (as the device's singular test) */
static double solid_angle_chk(const double q[3], const float *t, int *singular) {
  const double a[3] = {(double)t[0] - q[0], (double)t[1] - q[1], (double)t[2] - q[2]};
  const double b[3] = {(double)t[3] - q[0], (double)t[4] - q[1], (double)t[5] - q[2]};
  const double c[3] = {(double)t[6] - q[0], (double)t[7] - q[1], (double)t[8] - q[2]};
  const double la = sqrt(dot3d(a, a)), lb = sqrt(dot3d(b, b)), lc = sqrt(dot3d(c, c));
  const double bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
  const double det = dot3d(a, bc);
  const double l3 = la * lb * lc;
  const double den = l3 + dot3d(a, b) * lc + dot3d(b, c) * la + dot3d(c, a) * lb;
  if (den <= 1e-9 * l3) {
    const double nz = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]);
    const double zeps = 4.8e-7 * (fabs(q[2]) + fabs((double)t[2]) + fabs((double)t[5]) + fabs((double)t[8]));
    if (fabs(det) <= 1e-9 * l3 || fabs(det) <= zeps * fabs(nz)) *singular = 1;
  }
  return 2.0 * atan2(det, den);
}

static inline void neumaier(double *sum, double *comp, double v) {
  const double s = *sum + v;
  *comp += (fabs(*sum) >= fabs(v)) ? (*sum - s) + v : (v - s) + *sum;
  *sum = s;
}

/* cap winding at q: exact near q, bounded far away (*bound, winding units); *singular if q is on the cap */
static double cap_winding_bounded_f64(const capset *cs, const double q[3], double *bound, int *singular) {
  const orc_bvh *b = &cs->b;
  const int ni = (int)cs->n - 1;
  double sum = 0.0, comp = 0.0, ub = 0.0;
  int stack[512];
  int sp = 0;
  stack[sp++] = ni > 0 ? 0 : ~0;
  while (sp) {
    const int x = stack[--sp];
    const int id = x >= 0 ? x : ni + ~x;
    const float *lo = b->lo + 3 * id, *hi = b->hi + 3 * id;
    double d2 = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double g = fmax(fmax((double)lo[k] - q[k], q[k] - (double)hi[k]), 0.0);
      d2 += g * g;
    }
    if (d2 > 0.0) {
      const double bnd = cs->area[id] / (4.0 * M_PI * d2);
      if (bnd <= 1e-3) {
        ub += bnd;
        continue;
      }
    }
    if (x < 0) {
      neumaier(&sum, &comp, solid_angle_chk(q, cs->sorted + 9 * (size_t)(~x), singular));
    } else if (sp + 2 <= 512) {
      stack[sp++] = b->left[x];
      stack[sp++] = b->right[x];
    } else {
      ub += 1.0;
    }
  }
  *bound = ub * (1.0 + 1e-9);
  return (sum + comp) / (4.0 * M_PI);
}

static double cap_winding_full_f64(const capset *cs, const double q[3], int *singular) {
  double sum = 0.0, comp = 0.0;
  for (size_t i = 0; i < cs->n; ++i) neumaier(&sum, &comp, solid_angle_chk(q, cs->sorted + 9 * i, singular));
  return (sum + comp) / (4.0 * M_PI);
}

/* FP64 winding numbers of a whole soup at many points, exact where the decision needs it: a tree over the
 * soup is walked with tri_winding_dipole_f64; when | |w| - 1/2 | exceeds its bound by more than the tie band
 * 1e-6, |w| > 1/2 and whether w lies within 1e-6 of 1/2 are those of the exact sum and the estimate is kept,
 * else the point takes the brute-force sum orc_gwn_f64 (exact[i] = 1), as does a point lying on a
 * triangle. */
/* Winding number of a triset with moments at q, with a first-order far field: the triangles of a node lie
 * within rad of ctr; seen from d = |ctr - q| > rad they subtend dip . (ctr - q) / d^3 steradians within
 * 2 area rad / (d - rad)^3 (the integrand n . v / |v|^3 changes by at most 2 |dx| / |v|^3 per unit move dx
 * of the point on the triangle), and at most area / (d - rad)^2 in all.  A node whose smaller bound is at
 * most 1e-4 in winding units adds that term (or nothing) and the bound; the other triangles are summed
 * exactly.  *bound: the total bound (FP64 rounding included); *singular as solid_angle_chk. */
static double tri_winding_dipole_f64(const capset *cs, const double q[3], double *bound, int *singular) {
  const orc_bvh *b = &cs->b;
  const int ni = (int)cs->n - 1;
  double sum = 0.0, comp = 0.0, ub = 0.0;
  int stack[512];
  int sp = 0;
  stack[sp++] = ni > 0 ? 0 : ~0;
  while (sp) {
    const int x = stack[--sp];
    if (x < 0) {
      neumaier(&sum, &comp, solid_angle_chk(q, cs->sorted + 9 * (size_t)(~x), singular));
      continue;
    }
    const double v[3] = {cs->ctr[3 * x] - q[0], cs->ctr[3 * x + 1] - q[1], cs->ctr[3 * x + 2] - q[2]};
    const double d = sqrt(dot3d(v, v)), dmin = d - cs->rad[x];
    if (dmin > 0.0) {
      const double b0 = cs->area[x] / (4.0 * M_PI * dmin * dmin);
      const double b1 = b0 * 2.0 * cs->rad[x] / dmin;
      if (b1 < b0 && b1 <= 1e-4) {
        ub += b1;
        neumaier(&sum, &comp, dot3d(cs->dip + 3 * x, v) / (d * d * d));
        continue;
      }
      if (b0 <= 1e-4) {
        ub += b0;
        continue;
      }
    }
    if (sp + 2 <= 512) {
      stack[sp++] = b->left[x];
      stack[sp++] = b->right[x];
    } else {
      ub += 1.0;
    }
  }
  *bound = ub * (1.0 + 1e-9) + 1e-12;
  return (sum + comp) / (4.0 * M_PI);
}

typedef struct {
  const capset *cs;
  const float *xyz;
  size_t ntri;
  const double *pts;
  size_t npts;
  double *w;
  uint8_t *exact;
  size_t next;
  pthread_mutex_t *mu;
} gwn_job;

static void *gwn_fn(void *arg) {
  gwn_job *j = (gwn_job *)arg;
  for (;;) {
    pthread_mutex_lock(j->mu);
    const size_t i0 = j->next;
    j->next += 64;
    pthread_mutex_unlock(j->mu);
    if (i0 >= j->npts) break;
    const size_t i1 = i0 + 64 < j->npts ? i0 + 64 : j->npts;
    for (size_t i = i0; i < i1; ++i) {
      const double *q = j->pts + 3 * i;
      double bound = INFINITY, w = 0.0;
      int sing = 0;
      if (j->cs->n) w = tri_winding_dipole_f64(j->cs, q, &bound, &sing);
      const int exact = sing || !(fabs(fabs(w) - 0.5) > bound + 1e-6 + 1e-12);
      j->w[i] = exact ? orc_gwn_f64(q, j->xyz, j->ntri) : w;
      j->exact[i] = (uint8_t)exact;
    }
  }
  return NULL;
}

int orc_gwn_f64_points(const float *xyz, size_t ntri, const double *pts, size_t npts, int threads, double *out_w,
                       uint8_t *out_exact) {
  capset cs;
  memset(&cs, 0, sizeof(cs));
  if (ntri && triset_build(xyz, ntri, &cs) != 0) return -1;
  if (ntri) triset_moments(&cs);
  if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
  gwn_job j = {&cs, xyz, ntri, pts, npts, out_w, out_exact, 0, &mu};
  pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * threads);
  for (int i = 0; i < threads; ++i) pthread_create(&tid[i], NULL, gwn_fn, &j);
  for (int i = 0; i < threads; ++i) pthread_join(tid[i], NULL);
  free(tid);
  capset_free(&cs);
  return 0;
}

/* +z crossings of the reversed cap on the column through (x, y): appended to *out with negated signs */
static int cap_column_crossings(const capset *cs, float x, float y, crossing **out, int *capn, int cnt) {
  const orc_bvh *b = &cs->b;
  const int ni = (int)cs->n - 1;
  const float o[3] = {x, y, 0.0f};
  const float d[3] = {0.0f, 0.0f, 1.0f};
  ray_shear r;
  make_shear(d, &r);
  int stack[512];
  int sp = 0;
  stack[sp++] = ni > 0 ? 0 : ~0;
  while (sp) {
    const int x0 = stack[--sp];
    const int id = x0 >= 0 ? x0 : ni + ~x0;
    const float *lo = b->lo + 3 * id, *hi = b->hi + 3 * id;
    if (!(x >= lo[0] && x <= hi[0] && y >= lo[1] && y <= hi[1])) continue;
    if (x0 >= 0) {
      stack[sp++] = b->left[x0];
      stack[sp++] = b->right[x0];
      continue;
    }
    double th;
    const int s = ray_tri(o, &r, cs->sorted + 9 * (size_t)(~x0), &th);
    if (s == 0) continue;
    if (cnt == *capn) {
      *capn = *capn ? *capn * 2 : 16;
      *out = (crossing *)realloc(*out, sizeof(crossing) * (*capn));
    }
    (*out)[cnt].z = (float)th;
    (*out)[cnt].s = s; /* reversed: -(-s) */
    cnt++;
  }
  return cnt;
}

typedef struct {
  const qctx *c;
  const capset *cs;
  const float *xyz;
  size_t ntri;
  const float *origin;
  float h;
  int nbx, nby;
  const uint32_t *bricks; /* linear brick index per task */
  const uint32_t *slots;  /* pool slot per task */
  size_t ntask;
  float *pool;
  size_t next;
  pthread_mutex_t *mu;
  uint64_t rechecked, full, ties, singular;
} band_job;

static void *band_fn(void *arg) {
  band_job *j = (band_job *)arg;
  crossing *cr = NULL;
  int capn = 0;
  uint64_t nodes = 0, tests = 0, rechecked = 0, full = 0, ties = 0, singular = 0;
  for (;;) {
    pthread_mutex_lock(j->mu);
    const size_t t = j->next++;
    pthread_mutex_unlock(j->mu);
    if (t >= j->ntask) break;
    const uint32_t bl = j->bricks[t];
    const int bx = (int)(bl % (uint32_t)j->nbx), by = (int)((bl / (uint32_t)j->nbx) % (uint32_t)j->nby),
              bz = (int)(bl / ((uint32_t)j->nbx * (uint32_t)j->nby));
    float *dst = j->pool + 512 * (size_t)j->slots[t];
    for (int ly = 0; ly < 8; ++ly)
      for (int lx = 0; lx < 8; ++lx) {
        const int x = bx * 8 + lx, y = by * 8 + ly;
        const float px = j->origin[0] + ((float)x + 0.5f) * j->h;
        const float py = j->origin[1] + ((float)y + 0.5f) * j->h;
        int nc = column_crossings(j->c, px, py, &cr, &capn, &tests);
        if (j->cs->n) nc = cap_column_crossings(j->cs, px, py, &cr, &capn, nc);
        for (int lz = 0; lz < 8; ++lz) {
          const int z = bz * 8 + lz;
          const float q[3] = {px, py, j->origin[2] + ((float)z + 0.5f) * j->h};
          const float d = sqrtf(closest2(j->c, q, &nodes, &tests));
          int wi = 0;
          for (int k = 0; k < nc; ++k)
            if (cr[k].z > q[2]) wi += cr[k].s;
          int inside;
          if (!j->cs->n) {
            inside = wi != 0;
          } else {
            const double qd[3] = {q[0], q[1], q[2]};
            double bound;
            int sing = 0;
            double w = wi + cap_winding_bounded_f64(j->cs, qd, &bound, &sing);
            if (sing || fabs(fabs(w) - 0.5) < bound + 1e-9) {
              ++rechecked;
              if (!sing) {
                ++full;
                w = wi + cap_winding_full_f64(j->cs, qd, &sing);
              }
              if (sing) {
                ++singular;
                w = orc_gwn_f64(qd, j->xyz, j->ntri);
              }
            }
            if (fabs(fabs(w) - 0.5) < 1e-6) {
              ++ties;
              inside = orc_parity_inside(q, j->xyz, j->ntri);
            } else {
              inside = fabs(w) > 0.5;
            }
          }
          dst[(lz * 8 + ly) * 8 + lx] = inside ? -d : d;
        }
      }
  }
  free(cr);
  pthread_mutex_lock(j->mu);
  j->rechecked += rechecked;
  j->full += full;
  j->ties += ties;
  j->singular += singular;
  pthread_mutex_unlock(j->mu);
  return NULL;
}

int orc_bake_band(const float *xyz, size_t ntri, const float origin[3], float h, int nx, int ny, int nz, int band,
                  int brick_stride, int threads, int32_t *brick_index, float *pool, size_t pool_bricks,
                  size_t *nbricks, orc_band_stats *st) {
  if (ntri == 0 || (nx | ny | nz) & 7 || band <= 0) return -2;
  if (brick_stride < 1) brick_stride = 1;
  const double t0 = now_s();
  orc_bvh b;
  if (orc_bvh_build(xyz, ntri, &b) != 0) return -2;
  float *sorted = (float *)malloc(sizeof(float) * 9 * ntri);
  for (size_t i = 0; i < ntri; ++i) memcpy(sorted + 9 * i, xyz + 9 * (size_t)b.order[i], 9 * sizeof(float));
  orc_trirec *recs = make_recs(sorted, ntri);
  qctx c = {sorted, &b, recs, max_abs_coord(sorted, ntri)};
  if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  /* bricks: coarse unsigned distance at the brick centres (cell centres of spacing 8h) against the
   * band threshold h * (band + 3.5 sqrt 3), as the bake computes them */
  const int nbx = nx / 8, nby = ny / 8, nbz = nz / 8;
  const size_t nb = (size_t)nbx * nby * nbz;
  float *cpts = (float *)malloc(sizeof(float) * 3 * nb);
  const float h8 = h * 8.0f;
  for (size_t i = 0; i < nb; ++i) {
    const int x = (int)(i % nbx), y = (int)((i / nbx) % nby), z = (int)(i / ((size_t)nbx * nby));
    cpts[3 * i] = origin[0] + ((float)x + 0.5f) * h8;
    cpts[3 * i + 1] = origin[1] + ((float)y + 0.5f) * h8;
    cpts[3 * i + 2] = origin[2] + ((float)z + 0.5f) * h8;
  }
  float *cd = (float *)malloc(sizeof(float) * nb);
  {
    pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    udf_job uj = {&c, cpts, nb, cd, 0, &mu};
    pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    for (int i = 0; i < threads; ++i) pthread_create(&tid[i], NULL, udf_fn, &uj);
    for (int i = 0; i < threads; ++i) pthread_join(tid[i], NULL);
    free(tid);
  }
  const float thresh = h * ((float)band + 6.0621778f);
  size_t kept = 0;
  for (size_t i = 0; i < nb; ++i) {
    const int in = cd[i] <= thresh;
    brick_index[i] = in ? (int32_t)kept : -1;
    kept += (size_t)in;
  }
  *nbricks = kept;
  free(cpts);
  free(cd);
  int rc = 0;
  if (kept > pool_bricks) {
    rc = -1;
  } else {
    uint32_t *bricks = (uint32_t *)malloc(sizeof(uint32_t) * (kept ? kept : 1));
    uint32_t *slots = (uint32_t *)malloc(sizeof(uint32_t) * (kept ? kept : 1));
    size_t ntask = 0;
    for (size_t i = 0; i < nb; ++i)
      if (brick_index[i] >= 0 && brick_index[i] % brick_stride == 0) {
        bricks[ntask] = (uint32_t)i;
        slots[ntask++] = (uint32_t)brick_index[i];
      }
    capset cs;
    capset_build(xyz, ntri, &cs);
    pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    band_job j = {&c, &cs, xyz, ntri, origin, h, nbx, nby, bricks, slots, ntask, pool, 0, &mu, 0, 0, 0, 0};
    pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    for (int i = 0; i < threads; ++i) pthread_create(&tid[i], NULL, band_fn, &j);
    for (int i = 0; i < threads; ++i) pthread_join(tid[i], NULL);
    free(tid);
    if (st) {
      st->voxels = 512 * (uint64_t)ntask;
      st->bricks_baked = ntask;
      st->cap_triangles = cs.n;
      st->rechecked = j.rechecked;
      st->full_cap_sums = j.full;
      st->on_cap = j.singular;
      st->ties = j.ties;
    }
    capset_free(&cs);
    free(bricks);
    free(slots);
  }
  if (st) {
    st->seconds = now_s() - t0;
    st->threads = threads;
  }
  free(sorted);
  free(recs);
  orc_bvh_free(&b);
  return rc;
}
