// Analytic SDF bake on the device (src/convert/sdf_procedural.cpp:12-79): a CSG program instead of
// the reference's std::function, evaluated once per cell.  Primitives and operators follow
// src/convert/sdf_primitives.h expression by expression (glm's dot order (x*x + y*y) + z*z, its
// min(x, y) = y < x ? y : x and max(x, y) = x < y ? y : x, mix(x, y, a) = x*(1 - a) + y*a, and
// std::max of an initializer list keeping the first largest); built with -ffp-contract=off, so the
// values are those of the reference's host code.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ib_kernels.hpp"

namespace ib {
namespace {

__device__ __forceinline__ float gmin(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return x < y ? y : x; }
__device__ __forceinline__ float len3(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z); }
__device__ __forceinline__ float len2(float x, float y) { return sqrtf(x * x + y * y); }

__device__ float smooth_union(float a, float b, float k) {  // sdf_primitives.h:31-34
  const float h = gmin(gmax(0.5f + 0.5f * (b - a) / k, 0.0f), 1.0f);
  return (b * (1.0f - h) + a * h) - k * h * (1.0f - h);
}

// The program is wave-uniform (every lane runs the same instruction stream, read with scalar loads),
// so the stack depth is uniform too: the top of the stack lives in registers and the entries below
// it in LDS, one column per lane (conflict-free), instead of a per-lane scratch array.
__device__ __forceinline__ float prim(const CsgOp &op, float px, float py, float pz) {
  const float *a = op.a;
  const float x = px - a[0], y = py - a[1], z = pz - a[2];
  switch (op.op) {
    case kCsgSphere:  // sdf_primitives.h:10
      return len3(x, y, z) - a[3];
    case kCsgBox: {  // :12-15
      const float qx = fabsf(x) - a[3], qy = fabsf(y) - a[4], qz = fabsf(z) - a[5];
      float m = qx;
      if (m < qy) m = qy;
      if (m < qz) m = qz;
      return len3(gmax(qx, 0.0f), gmax(qy, 0.0f), gmax(qz, 0.0f)) + gmin(m, 0.0f);
    }
    case kCsgCylinder: {  // :17-20
      const float dx = fabsf(len2(x, z)) - a[3], dy = fabsf(y) - a[4];
      return gmin(gmax(dx, dy), 0.0f) + len2(gmax(dx, 0.0f), gmax(dy, 0.0f));
    }
    case kCsgTorus:  // :22-25
      return len2(len2(x, z) - a[3], y) - a[4];
    default:  // plane: dot(n, p) - offset, n = a[3..5], offset a[6] (a[0..2] unused)
      return ((a[3] * px + a[4] * py) + a[5] * pz) - a[6];
  }
}

struct ProgInline {  // the program in the kernel arguments (constant, read with scalar loads)
  CsgInline p;
  __device__ __forceinline__ const CsgOp &at(int k) const { return p.ops[k]; }
};
struct ProgGlobal {
  const CsgOp *__restrict__ p;
  __device__ __forceinline__ CsgOp at(int k) const { return p[k]; }
};

template <class Prog>
__global__ __launch_bounds__(256) void k_bake_csg(Prog prog, int nops, GridArgs g, float *dist, uint8_t *branch) {
  __shared__ float s_d[kCsgStack - 1][256];
  __shared__ uint8_t s_b[kCsgStack - 1][256];
  // a block is 64 x 4 cells of one z slice (a wave per row of 64 x), no index division
  const int tid = threadIdx.x;
  const int x = blockIdx.x * 64 + (tid & 63), y = blockIdx.y * 4 + (tid >> 6), z = blockIdx.z;
  if (x >= g.nx || y >= g.ny) return;
  const uint64_t i = ((uint64_t)z * g.ny + y) * g.nx + x;
  // origin + (vec3(x, y, z) + 0.5) * voxelSize (sdf_procedural.cpp:32-35)
  const float px = g.ox + ((float)x + 0.5f) * g.h, py = g.oy + ((float)y + 0.5f) * g.h,
              pz = g.oz + ((float)z + 0.5f) * g.h;
  float td = 0.0f;  // top of the stack
  int tb = 0;
  int depth = 0;  // values on the stack; depth - 1 of them in LDS (programs are validated on the host)
  for (int k = 0; k < nops; ++k) {
    const CsgOp op = prog.at(k);
    if (op.op <= kCsgPlane) {
      if (depth >= kCsgStack) continue;
      const float d = prim(op, px, py, pz);
      if (depth > 0) {
        s_d[depth - 1][tid] = td;
        s_b[depth - 1][tid] = (uint8_t)tb;
      }
      ++depth;
      td = d;
      tb = op.branch;
      continue;
    }
    if (depth < 2) continue;
    --depth;
    const float ad = s_d[depth - 1][tid];  // a = second from the top, b = the top
    const int ab = s_b[depth - 1][tid];
    const float bd = td;
    switch (op.op) {
      case kCsgUnion:  // std::min(a, b) (:27)
        if (!(bd < ad)) td = ad, tb = ab;
        break;
      case kCsgSubtract:  // std::max(a, -b) (:29)
        td = gmax(ad, -bd), tb = ab;
        break;
      case kCsgIntersect:  // std::max(a, b) (:31)
        if (!(ad < bd)) td = ad, tb = ab;
        break;
      case kCsgSmoothUnion:
        td = smooth_union(ad, bd, op.a[0]), tb = ab;
        break;
      case kCsgSmoothIntersect:  // -opSmoothUnion(-a, -b, k) (:36)
        td = -smooth_union(-ad, -bd, op.a[0]), tb = ab;
        break;
      default:  // select: a <= b ? a : b, with its branch (sdf_procedural.cpp:129)
        if (ad <= bd) td = ad, tb = ab;
        break;
    }
  }
  dist[i] = td;
  if (branch) branch[i] = (uint8_t)tb;
}

}  // namespace

void launch_bake_csg(const CsgOp *host_prog, const CsgOp *dev_prog, int nops, const GridArgs &g, float *dist,
                     uint8_t *branch, hipStream_t s) {
  if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0) return;
  const dim3 blocks((unsigned)((g.nx + 63) / 64), (unsigned)((g.ny + 3) / 4), (unsigned)g.nz);
  if (nops <= kCsgInline) {
    ProgInline p{};
    for (int k = 0; k < nops; ++k) p.p.ops[k] = host_prog[k];
    k_bake_csg<<<blocks, 256, 0, s>>>(p, nops, g, dist, branch);
  } else {
    k_bake_csg<<<blocks, 256, 0, s>>>(ProgGlobal{dev_prog}, nops, g, dist, branch);
  }
}

}  // namespace ib
