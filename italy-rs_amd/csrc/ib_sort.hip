// Stable LSD radix sort of (uint32 key, int32 value) pairs for the build's sorts (Morton codes, tile launch
// keys, cell keys).  rocPRIM sorts fewer than 2^20 keys by merge sort, about twenty launches of a few
// microseconds each for the 2^18..2^20 keys of a large mesh; this sort takes three digit passes of at most
// 11 bits, each a count, a scan and a scatter launch over tiles of 2,048 keys.
//
// Pass p sorts by bits [p * D, min((p + 1) * D, end_bit)).  count: tile t's histogram of its digits into
// counts[d * ntiles + t].  rowscan: per digit, the exclusive scan over the tiles and the digit's total.
// scatter: the exclusive scan of the digit totals gives each digit's first position; each wave of the tile
// ranks its 512 keys in order (equal digits found with one ballot per digit bit, earlier keys counted in
// the wave's LDS counters), the waves' counts are prefix-summed per digit, and each key goes to its digit's
// first position plus the keys of that digit in earlier tiles and earlier waves plus its rank.  Keys keep
// their order within a digit, so the passes compose into a stable sort.
#include <hip/hip_runtime.h>

#include <cstring>

#include "ib_kernels.hpp"

namespace ib {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;                            // keys per thread
constexpr int kSortTile = kSortThreads * kSortItems;     // 2,048 keys per tile
constexpr int kWaveKeys = kSortTile / (kSortThreads / 64);  // 512 keys per wave
constexpr int kMaxDigitBits = 11;
constexpr int kMaxBins = 1 << kMaxDigitBits;

__global__ __launch_bounds__(kSortThreads) void k_sort_count(const uint32_t *__restrict__ keys, uint32_t n,
                                                              int shift, int bits, uint32_t ntiles,
                                                              uint32_t *__restrict__ counts) {
  __shared__ uint32_t hist[kMaxBins];
  const int bins = 1 << bits;
  for (int d = threadIdx.x; d < bins; d += kSortThreads) hist[d] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortTile;
  for (int i = 0; i < kSortItems; ++i) {
    const uint32_t e = base + i * kSortThreads + threadIdx.x;
    if (e < n) atomicAdd(&hist[(keys[e] >> shift) & (bins - 1)], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < bins; d += kSortThreads) counts[(size_t)d * ntiles + blockIdx.x] = hist[d];
}

// exclusive scan of v over the block's 256 threads (4 waves); returns the block total in `total`
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t &total, uint32_t *wsum) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
  for (int off = 1; off < 64; off <<= 1) {  // inclusive scan within the wave
    const uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t before = 0;
  for (int k = 0; k < w; ++k) before += wsum[k];
  total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
  __syncthreads();
  return before + x - v;
}

// per digit (one block each): exclusive scan of the digit's counts over the tiles in place, and the
// digit's total
__global__ __launch_bounds__(kSortThreads) void k_sort_rowscan(uint32_t *counts, uint32_t ntiles,
                                                                uint32_t *totals) {
  __shared__ uint32_t wsum[kSortThreads / 64];
  uint32_t *row = counts + (size_t)blockIdx.x * ntiles;
  uint32_t carry = 0;
  for (uint32_t c = 0; c < ntiles; c += kSortThreads) {
    const uint32_t i = c + threadIdx.x;
    const uint32_t v = i < ntiles ? row[i] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(v, total, wsum);
    if (i < ntiles) row[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(const uint32_t *__restrict__ kin,
                                                                const int32_t *__restrict__ vin, uint32_t n,
                                                                int shift, int bits, uint32_t ntiles,
                                                                const uint32_t *__restrict__ offsets,
                                                                const uint32_t *__restrict__ totals,
                                                                uint32_t *__restrict__ kout,
                                                                int32_t *__restrict__ vout) {
  __shared__ uint32_t wc[kSortThreads / 64][kMaxBins];  // per wave: keys of each digit so far, then offsets
  __shared__ uint32_t first[kMaxBins];                  // first output position of each digit
  __shared__ uint32_t wsum[kSortThreads / 64];
  const int bins = 1 << bits;
  {  // exclusive scan of the digit totals: thread t sums digits [t * per, (t + 1) * per)
    const int per = (bins + kSortThreads - 1) / kSortThreads, d0 = threadIdx.x * per;
    uint32_t sum = 0;
    for (int d = d0; d < d0 + per && d < bins; ++d) sum += totals[d];
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, total, wsum);
    for (int d = d0; d < d0 + per && d < bins; ++d) {
      first[d] = run;
      run += totals[d];
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int d = lane; d < bins; d += 64) wc[w][d] = 0;
  __builtin_amdgcn_wave_barrier();
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t base = blockIdx.x * kSortTile + w * kWaveKeys;
  uint32_t key[kSortItems], rank[kSortItems];
  int32_t val[kSortItems];
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const uint32_t e = base + i * 64 + lane;
    const bool have = e < n;
    key[i] = have ? kin[e] : 0u;
    val[i] = have ? vin[e] : 0;
    const uint32_t d = (key[i] >> shift) & (bins - 1);
    // lanes holding a key with this lane's digit (lanes past the end take part in no group)
    uint64_t same = __ballot(have);
    for (int b = 0; b < bits; ++b) {
      const uint64_t ones = __ballot(have && ((d >> b) & 1u));
      same &= ((d >> b) & 1u) ? ones : ~ones;
    }
    const uint32_t prior = have ? wc[w][d] : 0u;  // every lane reads before any lane adds
    __builtin_amdgcn_wave_barrier();
    const uint32_t before = (uint32_t)__popcll(same & below);
    rank[i] = prior + before;
    if (have && before == 0) wc[w][d] = prior + (uint32_t)__popcll(same);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // per digit: the tile's first output position (scanned counts) plus the keys of earlier waves
  for (int d = threadIdx.x; d < bins; d += kSortThreads) {
    uint32_t run = first[d] + offsets[(size_t)d * ntiles + blockIdx.x];
    for (int k = 0; k < kSortThreads / 64; ++k) {
      const uint32_t c = wc[k][d];
      wc[k][d] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const uint32_t e = base + i * 64 + lane;
    if (e >= n) continue;
    const uint32_t pos = wc[w][(key[i] >> shift) & (bins - 1)] + rank[i];
    kout[pos] = key[i];
    vout[pos] = val[i];
  }
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

hipError_t radix_sort_u32(void *temp, size_t &temp_bytes, const uint32_t *kin, uint32_t *kout, const int32_t *vin,
                          int32_t *vout, uint32_t n, int end_bit, hipStream_t s) {
  end_bit = end_bit < 1 ? 1 : (end_bit > 32 ? 32 : end_bit);
  const int passes = (end_bit + kMaxDigitBits - 1) / kMaxDigitBits;
  const int dbits = (end_bit + passes - 1) / passes;
  const uint32_t ntiles = n ? (n + kSortTile - 1) / kSortTile : 1;
  const size_t b_counts = align256(((size_t)ntiles << dbits) * 4), b_totals = align256((size_t)kMaxBins * 4),
               b_keys = align256((size_t)n * 4), b_vals = align256((size_t)n * 4);
  const size_t need = b_counts + b_totals + b_keys + b_vals;
  if (!temp) {
    temp_bytes = need;
    return hipSuccess;
  }
  if (temp_bytes < need) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  char *t = static_cast<char *>(temp);
  uint32_t *counts = reinterpret_cast<uint32_t *>(t);
  uint32_t *totals = reinterpret_cast<uint32_t *>(t + b_counts);
  uint32_t *tk = reinterpret_cast<uint32_t *>(t + b_counts + b_totals);
  int32_t *tv = reinterpret_cast<int32_t *>(t + b_counts + b_totals + b_keys);
  const uint32_t *sk = kin;
  const int32_t *sv = vin;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * dbits, bits = (shift + dbits <= end_bit ? dbits : end_bit - shift);
    // the last pass writes the output; earlier ones alternate so that the input is never written
    const bool to_out = ((passes - 1 - p) & 1) == 0;
    uint32_t *dk = to_out ? kout : tk;
    int32_t *dv = to_out ? vout : tv;
    k_sort_count<<<ntiles, kSortThreads, 0, s>>>(sk, n, shift, bits, ntiles, counts);
    k_sort_rowscan<<<1u << bits, kSortThreads, 0, s>>>(counts, ntiles, totals);
    k_sort_scatter<<<ntiles, kSortThreads, 0, s>>>(sk, sv, n, shift, bits, ntiles, counts, totals, dk, dv);
    sk = dk;
    sv = dv;
  }
  return hipGetLastError();
}

}  // namespace ib
