// The grid consumer on the device: what the renderer does with a baked SDF grid
// (src/render/kernels/pathtracer.cu:684-709 sampleSdf / sdfNearestCellIndex, :740-792 the SDF
// intersection program, :794-846 cubic B-spline gradient normals), run directly on device-resident
// bakes, both as a verifier of baked grids and as a measured consumer.  Same build flags as the
// bake kernels (-ffp-contract=off): every a*b+c is two roundings, as in the oracle's restatement.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ib_kernels.hpp"

namespace ib {
namespace {

struct V3 {
  float x, y, z;
};

__device__ __forceinline__ V3 local_coords(const GridView &g, V3 p) {
  // (p - origin) / voxelSize - 0.5 per component (pathtracer.cu:685)
  return {(p.x - g.ox) / g.h - 0.5f, (p.y - g.oy) / g.h - 0.5f, (p.z - g.oz) / g.h - 0.5f};
}

__device__ __forceinline__ float at(const GridView &g, int x, int y, int z) {
  return g.d[((size_t)z * g.ny + y) * g.nx + x];
}

// trilinear sample of the cell-centred grid, clamped to the last full cell (pathtracer.cu:684-701)
__device__ float sample_sdf(const GridView &g, V3 p) {
  const V3 l = local_coords(g, p);
  const int x0 = max(0, min(g.nx - 2, (int)floorf(l.x)));
  const int y0 = max(0, min(g.ny - 2, (int)floorf(l.y)));
  const int z0 = max(0, min(g.nz - 2, (int)floorf(l.z)));
  const float fx = fminf(fmaxf(l.x - (float)x0, 0.0f), 1.0f);
  const float fy = fminf(fmaxf(l.y - (float)y0, 0.0f), 1.0f);
  const float fz = fminf(fmaxf(l.z - (float)z0, 0.0f), 1.0f);
  const float c00 = at(g, x0, y0, z0) * (1.0f - fx) + at(g, x0 + 1, y0, z0) * fx;
  const float c10 = at(g, x0, y0 + 1, z0) * (1.0f - fx) + at(g, x0 + 1, y0 + 1, z0) * fx;
  const float c01 = at(g, x0, y0, z0 + 1) * (1.0f - fx) + at(g, x0 + 1, y0, z0 + 1) * fx;
  const float c11 = at(g, x0, y0 + 1, z0 + 1) * (1.0f - fx) + at(g, x0 + 1, y0 + 1, z0 + 1) * fx;
  const float c0 = c00 * (1.0f - fy) + c10 * fy;
  const float c1 = c01 * (1.0f - fy) + c11 * fy;
  return c0 * (1.0f - fz) + c1 * fz;
}

// nearest cell, rounding half away from zero, clamped (pathtracer.cu:703-709)
__device__ int nearest_cell(const GridView &g, V3 p) {
  const V3 l = local_coords(g, p);
  const int x = max(0, min(g.nx - 1, (int)lroundf(l.x)));
  const int y = max(0, min(g.ny - 1, (int)lroundf(l.y)));
  const int z = max(0, min(g.nz - 1, (int)lroundf(l.z)));
  return (z * g.ny + y) * g.nx + x;
}

// uniform cubic B-spline weights and their derivatives at fraction t (pathtracer.cu:794-804)
__device__ __forceinline__ void bspline(float t, float w[4], float dw[4]) {
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

// gradient of the cubic B-spline interpolant over the 4^3 neighbourhood, normalised
// (pathtracer.cu:806-846); normalise = v * (1 / sqrt((x*x + y*y) + z*z))
__device__ V3 gradient_normal(const GridView &g, V3 p) {
  const V3 l = local_coords(g, p);
  const int bx = (int)floorf(l.x), by = (int)floorf(l.y), bz = (int)floorf(l.z);
  float wx[4], dwx[4], wy[4], dwy[4], wz[4], dwz[4];
  bspline(l.x - (float)bx, wx, dwx);
  bspline(l.y - (float)by, wy, dwy);
  bspline(l.z - (float)bz, wz, dwz);
  int ix[4];
  for (int i = 0; i < 4; ++i) ix[i] = max(0, min(g.nx - 1, bx - 1 + i));
  V3 gr = {0.0f, 0.0f, 0.0f};
  for (int k = 0; k < 4; ++k) {
    const int zrow = max(0, min(g.nz - 1, bz - 1 + k)) * g.ny;
    V3 plane = {0.0f, 0.0f, 0.0f};
    for (int j = 0; j < 4; ++j) {
      const size_t row = (size_t)(zrow + max(0, min(g.ny - 1, by - 1 + j))) * g.nx;
      float v = 0.0f, dv = 0.0f;
      for (int i = 0; i < 4; ++i) {
        const float f = g.d[row + ix[i]];
        v += wx[i] * f;
        dv += dwx[i] * f;
      }
      plane.x += dv * wy[j];
      plane.y += v * dwy[j];
      plane.z += v * wy[j];
    }
    gr.x += plane.x * wz[k];
    gr.y += plane.y * wz[k];
    gr.z += plane.z * dwz[k];
  }
  const float inv = 1.0f / sqrtf((gr.x * gr.x + gr.y * gr.y) + gr.z * gr.z);
  return {gr.x * inv, gr.y * inv, gr.z * inv};
}

__device__ __forceinline__ V3 ray_at(V3 o, V3 d, float t) { return {o.x + t * d.x, o.y + t * d.y, o.z + t * d.z}; }

// the SDF intersection program (pathtracer.cu:587-613 slab test, :740-792 march): enter the grid
// box, step by max(0.85 |d|, 0.05 h), report a sign change (refined by 8 bisections and a secant
// step) or |d| < 0.02 h once past 0.3 h; -1 when the ray leaves the box or after 512 steps
__device__ float trace_grid(const GridView &g, V3 o, V3 d, float t0, float t1) {
  const float lo[3] = {g.ox, g.oy, g.oz};
  const float hi[3] = {g.ox + (float)g.nx * g.h, g.oy + (float)g.ny * g.h, g.oz + (float)g.nz * g.h};
  const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  for (int axis = 0; axis < 3; ++axis) {
    const float inv = 1.0f / dd[axis];
    float tn = (lo[axis] - oo[axis]) * inv, tf = (hi[axis] - oo[axis]) * inv;
    if (tn > tf) {
      const float tmp = tn;
      tn = tf;
      tf = tmp;
    }
    if (tn > t0) t0 = tn;
    if (tf < t1) t1 = tf;
    if (t0 > t1) return -1.0f;
  }
  const float eps = g.h * 0.02f, min_step = g.h * 0.05f, min_adv = g.h * 0.3f;
  float t = t0, prev_t = t0;
  float prev = sample_sdf(g, ray_at(o, d, t0));
  t += fmaxf(fabsf(prev), min_step);
  for (int i = 1; i < 512 && t <= t1; ++i) {
    const float dist = sample_sdf(g, ray_at(o, d, t));
    const bool crossed = (prev > 0.0f) != (dist > 0.0f);
    if ((crossed || fabsf(dist) < eps) && (t - t0) > min_adv) {
      float th = t;
      if (crossed) {
        float a = prev_t, b = t, fa = prev, fb = dist;
        for (int k = 0; k < 8; ++k) {
          const float m = 0.5f * (a + b);
          const float fm = sample_sdf(g, ray_at(o, d, m));
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

__global__ __launch_bounds__(256) void k_sample_grid(GridView g, const float *__restrict__ pts, uint64_t n,
                                                     float *values, float *normals, int32_t *nearest) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const V3 p = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
  if (values) values[i] = sample_sdf(g, p);
  if (normals) {
    const V3 nn = gradient_normal(g, p);
    normals[3 * i] = nn.x;
    normals[3 * i + 1] = nn.y;
    normals[3 * i + 2] = nn.z;
  }
  if (nearest) nearest[i] = nearest_cell(g, p);
}

__global__ __launch_bounds__(256) void k_trace_grid(GridView g, const float *__restrict__ org,
                                                    const float *__restrict__ dir, uint64_t n, float tmin, float tmax,
                                                    float *t_hit) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  t_hit[i] = trace_grid(g, {org[3 * i], org[3 * i + 1], org[3 * i + 2]}, {dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]},
                        tmin, tmax);
}

}  // namespace

void launch_sample_grid(const GridView &g, const float *pts, uint64_t n, float *values, float *normals,
                        int32_t *nearest, hipStream_t s) {
  if (n) k_sample_grid<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g, pts, n, values, normals, nearest);
}

void launch_trace_grid(const GridView &g, const float *org, const float *dir, uint64_t n, float tmin, float tmax,
                       float *t_hit, hipStream_t s) {
  if (n) k_trace_grid<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g, org, dir, n, tmin, tmax, t_hit);
}

}  // namespace ib
