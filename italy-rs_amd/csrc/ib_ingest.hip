// The reference's mesh ingestion gate on the device:
//   weldPositions (src/io/mesh_weld.cpp:26-48): positions that round to the same point of a lattice of
//     pitch max(|bmax - bmin| * 1e-5, 1e-6) share an index, indices numbered in order of first
//     occurrence (the map's insertion order);
//   isWatertight (src/io/mesh_validate.cpp:21-59): every undirected edge of the welded, non-degenerate
//     triangles must be used exactly twice.
// After the lattice key (one correctly rounded FP32 divide and a round half away from zero per
// coordinate) everything is integer: the hash tables only decide where a key lives, never which index
// it gets, so the results equal the oracle's whatever the schedule.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ib_kernels.hpp"

namespace ib {
namespace {

__device__ __forceinline__ uint32_t mix_key(long long a, long long b, long long c) {
  uint64_t x = (uint64_t)a * 0x9E3779B97F4A7C15ULL;
  x ^= (uint64_t)b + 0x632BE59BD9B4E019ULL + (x << 6) + (x >> 2);
  x ^= (uint64_t)c + 0x8CB92BA72F3D8DD7ULL + (x << 6) + (x >> 2);
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return (uint32_t)x;
}

__device__ __forceinline__ uint32_t mix_edge(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return (uint32_t)k;
}

// std::lround(p / pitch) per coordinate (mesh_weld.cpp:40-42).  roundf rounds half away from zero,
// as lround does; a quotient that is not finite or reaches 2^63 (where lround is unspecified) sets *bad.
__global__ __launch_bounds__(256) void k_qweld_keys(const float *__restrict__ pos, uint32_t n, float pitch,
                                                    long long *keys, uint32_t *bad) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  for (int a = 0; a < 3; ++a) {
    const float q = pos[3 * (size_t)i + a] / pitch;
    long long k = 0;
    if (fabsf(q) < 0x1p63f)
      k = (long long)roundf(q);
    else
      atomicOr(bad, 1u);
    keys[3 * (size_t)i + a] = k;
  }
}

// Each key gets a slot holding the smallest position index with that key: the first insertion claims
// an empty slot (-1) with a compare-and-swap, later ones with the same key lower it with atomicMin.
// A claimed slot only ever holds indices of its own key, so a stale read still names the right key.
__global__ __launch_bounds__(256) void k_qweld_insert(const long long *__restrict__ keys, uint32_t n,
                                                      int32_t *slots, uint32_t mask, uint32_t *vslot) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const long long k0 = keys[3 * (size_t)i], k1 = keys[3 * (size_t)i + 1], k2 = keys[3 * (size_t)i + 2];
  uint32_t s = mix_key(k0, k1, k2) & mask;
  for (;;) {
    int cur = slots[s];
    if (cur < 0) {
      cur = atomicCAS(&slots[s], -1, (int)i);
      if (cur < 0) break;
    }
    if (keys[3 * (size_t)cur] == k0 && keys[3 * (size_t)cur + 1] == k1 && keys[3 * (size_t)cur + 2] == k2) {
      atomicMin(&slots[s], (int)i);
      break;
    }
    s = (s + 1) & mask;
  }
  vslot[i] = s;
}

// first[i] = first position with i's key; flag[i] = 1 where i is that position; flag[n] = 0 so that
// the exclusive scan's last entry is the number of indices
__global__ __launch_bounds__(256) void k_qweld_first(const int32_t *__restrict__ slots,
                                                     const uint32_t *__restrict__ vslot, uint32_t n, uint32_t *first,
                                                     uint32_t *flag) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i > n) return;
  if (i == n) {
    flag[n] = 0;
    return;
  }
  const uint32_t f = (uint32_t)slots[vslot[i]];
  first[i] = f;
  flag[i] = f == i ? 1u : 0u;
}

// welded[i] = rank of its first position among the first positions (= weldMap.size() at insertion)
__global__ __launch_bounds__(256) void k_qweld_out(const float *__restrict__ pos, const uint32_t *__restrict__ first,
                                                   const uint32_t *__restrict__ flag,
                                                   const uint32_t *__restrict__ rank, uint32_t n, uint32_t *welded,
                                                   float *unique) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  welded[i] = rank[first[i]];
  if (unique && flag[i])
    for (int a = 0; a < 3; ++a) unique[3 * (size_t)rank[i] + a] = pos[3 * (size_t)i + a];
}

// undirected edges (min << 32 | max) of every triangle without a repeated index (mesh_validate.cpp:30-39),
// counted per distinct edge
__global__ __launch_bounds__(256) void k_wt_edges(const uint32_t *__restrict__ welded, uint32_t ntri,
                                                  unsigned long long *ekeys, uint32_t *ecount, uint32_t mask) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= ntri) return;
  const uint32_t v[3] = {welded[3 * (size_t)t], welded[3 * (size_t)t + 1], welded[3 * (size_t)t + 2]};
  if (v[0] == v[1] || v[1] == v[2] || v[2] == v[0]) return;
  for (int k = 0; k < 3; ++k) {
    const uint32_t a = v[k], b = v[k == 2 ? 0 : k + 1];
    const unsigned long long key = a < b ? ((unsigned long long)a << 32 | b) : ((unsigned long long)b << 32 | a);
    uint32_t s = mix_edge(key) & mask;
    for (;;) {
      unsigned long long cur = ekeys[s];
      if (cur == ~0ull) {
        cur = atomicCAS(&ekeys[s], ~0ull, key);
        if (cur == ~0ull) break;
      }
      if (cur == key) break;
      s = (s + 1) & mask;
    }
    atomicAdd(&ecount[s], 1u);
  }
}

// out[0] += edges used once, out[1] += edges used more than twice (mesh_validate.cpp:41-47); one atomic
// per wave from ballots
__global__ __launch_bounds__(256) void k_wt_tally(const unsigned long long *__restrict__ ekeys,
                                                  const uint32_t *__restrict__ ecount, uint32_t cap,
                                                  unsigned long long *out) {
  const uint32_t s = blockIdx.x * 256 + threadIdx.x;
  bool once = false, many = false;
  if (s < cap && ekeys[s] != ~0ull) {
    const uint32_t c = ecount[s];
    once = c == 1;
    many = c > 2;
  }
  const unsigned long long b_once = __ballot(once), b_many = __ballot(many);
  if ((threadIdx.x & 63) == 0) {
    if (b_once) atomicAdd(&out[0], (unsigned long long)__popcll(b_once));
    if (b_many) atomicAdd(&out[1], (unsigned long long)__popcll(b_many));
  }
}

inline uint32_t blocks(uint64_t n) { return (uint32_t)((n + 255) / 256); }

}  // namespace

void launch_qweld(const float *pos, uint32_t n, float pitch, long long *keys, int32_t *slots, uint32_t mask,
                  uint32_t *vslot, uint32_t *first, uint32_t *flag, uint32_t *bad, hipStream_t s) {
  if (n == 0) return;
  k_qweld_keys<<<blocks(n), 256, 0, s>>>(pos, n, pitch, keys, bad);
  k_qweld_insert<<<blocks(n), 256, 0, s>>>(keys, n, slots, mask, vslot);
  k_qweld_first<<<blocks((uint64_t)n + 1), 256, 0, s>>>(slots, vslot, n, first, flag);
}

void launch_qweld_out(const float *pos, const uint32_t *first, const uint32_t *flag, const uint32_t *rank, uint32_t n,
                      uint32_t *welded, float *unique, hipStream_t s) {
  if (n) k_qweld_out<<<blocks(n), 256, 0, s>>>(pos, first, flag, rank, n, welded, unique);
}

void launch_watertight(const uint32_t *welded, uint32_t ntri, unsigned long long *ekeys, uint32_t *ecount,
                       uint32_t mask, unsigned long long *out, hipStream_t s) {
  if (ntri == 0) return;
  k_wt_edges<<<blocks(ntri), 256, 0, s>>>(welded, ntri, ekeys, ecount, mask);
  k_wt_tally<<<blocks((uint64_t)mask + 1), 256, 0, s>>>(ekeys, ecount, mask + 1, out);
}

}  // namespace ib
