// C ABI of the bake (include/italy_bake.h): context, device buffers, bake orchestration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ib_kernels.hpp"
#include "italy_bake.h"

using namespace ib;

namespace {

thread_local std::string g_err;
thread_local uint64_t g_failures = 0;  // fail() calls on this thread: a call failed if it moved

ib_status fail(ib_status s, const std::string &msg) {
  g_err = msg;
  ++g_failures;
  return s;
}

#define IB_HIP(call)                                                                                  \
  do {                                                                                                \
    const hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) return fail(IB_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define IB_TRY(expr)                    \
  do {                                  \
    const ib_status s_ = (expr);        \
    if (s_ != IB_OK) return s_;         \
  } while (0)

struct Buf {
  void *p = nullptr;
  size_t bytes = 0;
};

}  // namespace

enum EventId { EV_START = 0, EV_INGEST, EV_BVH, EV_OCC, EV_COLS, EV_SDF, EV_OCC_SIDE, EV_Q0, EV_Q1, EV_CAP0, EV_CAP1, EV_COUNT };

struct ib_context {
  int device = 0;
  hipStream_t stream = nullptr;
  static constexpr int kSide = 4;   // streams a batch builds its meshes on, joined before the query
  hipStream_t side[kSide] = {};
  hipEvent_t fork = nullptr, join[kSide] = {};
  // end of the last call's device work (on whatever stream it ran): every call's stream waits for it
  // first, so calls on different streams never overlap on the context's shared scratch
  hipEvent_t tail = nullptr;
  bool tail_valid = false;
  uint64_t budget = 0;
  uint64_t allocated = 0;
  std::unordered_map<std::string, Buf> bufs;
  std::unordered_map<void *, size_t> user;
  Scalars *host_sc = nullptr;  // pinned
  Scalars *host_sc_early = nullptr;  // pinned: the weld's counters, read back before the tree is done
  char *staging = nullptr;     // pinned staging of large device -> host results (grown on demand)
  size_t staging_bytes = 0;
  hipEvent_t ev[EV_COUNT] = {};
  bool ev_valid[EV_COUNT] = {};
  ib_stats stats{};
  bool count_work = false;
  bool stats_in_sc = false;    // last bake was a plain bake: its device counters are at last_sc
  Scalars *last_sc = nullptr;
  int query_variant = 1;
  int brick_shape = 0;         // wave brick shape (autotuned for batches)
  int blocks_per_cu = 8;       // persistent blocks per CU (autotuned for batches)
  int leaf_block = 8;          // LBVH subtrees of at most this many triangles are evaluated as one leaf block
  int tile_order = 2;          // dense per-brick queries: 0 grid order, 1 costliest first, 2 cost buckets + Morton
  bool resolve_full = false;   // FP64 rechecks sum every cap fan (no cap-tree bound first)
  bool shared_chip = false;    // bakes share the GPU with other streams' work (IB_OPT_SHARED_CHIP)
  int cus = 0;
  Scalars *host_batch_sc = nullptr;  // pinned, batch_sc_cap entries
  uint32_t batch_sc_cap = 0;
  std::mutex mu;

  // freed blocks are kept per exact size and handed out again, so that repeated batches do not
  // call hipMalloc / hipFree (which can serialise work queued on other streams)
  std::unordered_map<size_t, std::vector<void *>> free_blocks;
  size_t cached = 0;
  ib_status alloc_raw(size_t bytes, void **p) {
    if (allocated + bytes > budget)
      return fail(IB_ERR_OUT_OF_BUDGET, "allocation of " + std::to_string(bytes) + " bytes exceeds the HBM budget (" +
                                            std::to_string(allocated) + " of " + std::to_string(budget) + " in use)");
    auto it = free_blocks.find(bytes);
    if (it != free_blocks.end() && !it->second.empty()) {
      *p = it->second.back();
      it->second.pop_back();
      cached -= bytes;
      allocated += bytes;
      return IB_OK;
    }
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess && cached) {  // out of device memory: return the cache and retry once
      release_cache();
      e = hipMalloc(p, bytes ? bytes : 1);
    }
    if (e != hipSuccess) return fail(IB_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocated += bytes;
    return IB_OK;
  }
  void free_raw(void *p, size_t bytes) {
    if (p) {
      free_blocks[bytes].push_back(p);
      cached += bytes;
      allocated -= bytes;
      if (cached > (budget >> 2)) release_cache();
    }
  }
  void release_cache() {
    for (auto &kv : free_blocks)
      for (void *q : kv.second) (void)hipFree(q);
    free_blocks.clear();
    cached = 0;
  }
  // Scratch-free calls: while fixed_base is set, named scratch is carved from that fixed device region
  // (bump allocation, 256-byte aligned; a name asked for again gets its earlier range when it fits)
  // instead of the allocating scratch below.  fixed_base is set for one call and cleared after it.
  char *fixed_base = nullptr;
  size_t fixed_cap = 0, fixed_used = 0;
  void *noalloc_base = nullptr;  // this device's region of ib_bake_dense_device_noalloc (set by ib_create)
  std::unordered_map<std::string, std::pair<size_t, size_t>> fixed_map;  // name -> (offset, bytes)
  void begin_fixed(void *base, size_t cap) {
    fixed_base = static_cast<char *>(base);
    fixed_cap = cap;
    fixed_used = 0;
    fixed_map.clear();
  }
  void end_fixed() { fixed_base = nullptr; }

  // named scratch that only grows; contents are not preserved across growth
  template <typename T>
  ib_status scratch(const char *name, size_t count, T **out) {
    const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    if (fixed_base) {
      auto it = fixed_map.find(name);
      if (it == fixed_map.end() || it->second.second < bytes) {
        const size_t off = (fixed_used + 255) & ~size_t(255);
        if (off + bytes > fixed_cap)
          return fail(IB_ERR_CAPACITY, std::string("the fixed scratch region (") + std::to_string(fixed_cap) +
                                           " bytes) cannot hold '" + name + "' (" + std::to_string(bytes) +
                                           " bytes after " + std::to_string(off) + ")");
        fixed_map[name] = {off, bytes};
        fixed_used = off + bytes;
        it = fixed_map.find(name);
      }
      *out = reinterpret_cast<T *>(fixed_base + it->second.first);
      return IB_OK;
    }
    Buf &b = bufs[name];
    if (b.bytes < bytes) {
      const size_t want = std::max(bytes + bytes / 4, 2 * b.bytes);  // geometric growth: few reallocations
      free_raw(b.p, b.bytes);
      b.p = nullptr;
      b.bytes = 0;
      void *p = nullptr;
      ib_status st = alloc_raw(want, &p);
      if (st != IB_OK) {
        st = alloc_raw(bytes, &p);
        if (st != IB_OK) return st;
        b.bytes = bytes;
      } else {
        b.bytes = want;
      }
      b.p = p;
    }
    *out = static_cast<T *>(b.p);
    return IB_OK;
  }
  void record(EventId e, hipStream_t s) {
    (void)hipEventRecord(ev[e], s);
    ev_valid[e] = true;
  }
};

namespace {

// Orders one entry point's device work after every earlier call on the context and records where it
// ends.  Built under the context lock once the call's stream is known.  On destruction (every return
// path) the side streams are joined into the call's stream, so work that a failed call left queued on
// them is ordered too, and the tail event is recorded there.  A call that failed (fail() ran on this
// thread meanwhile) also waits for its queued work before it returns, because its caller may free or
// reuse what that work reads and writes, and a batch's arena blocks go back to the context's free list.
struct Ordered {
  ib_context *c;
  hipStream_t s;
  uint64_t failures;
  Ordered(ib_context *ctx, hipStream_t st) : c(ctx), s(st), failures(g_failures) {
    if (c->tail_valid) (void)hipStreamWaitEvent(s, c->tail, 0);
  }
  ~Ordered() {
    for (int i = 0; i < ib_context::kSide; ++i) {
      (void)hipEventRecord(c->join[i], c->side[i]);
      (void)hipStreamWaitEvent(s, c->join[i], 0);
    }
    (void)hipEventRecord(c->tail, s);
    c->tail_valid = true;
    if (g_failures != failures) (void)hipStreamSynchronize(s);
  }
};

uint32_t pow2_at_least(uint64_t v) {
  uint64_t c = 1024;
  while (c < v) c <<= 1;
  return (uint32_t)c;
}

bool grid_ok(const ib_grid *g) {
  return g && g->nx > 0 && g->ny > 0 && g->nz > 0 && std::isfinite(g->h) && g->h > 0.0f &&
         std::isfinite(g->origin[0]) && std::isfinite(g->origin[1]) && std::isfinite(g->origin[2]);
}

GridArgs to_args(const ib_grid &g) { return GridArgs{g.origin[0], g.origin[1], g.origin[2], g.h, g.nx, g.ny, g.nz}; }

float ev_ms(ib_context *c, EventId a, EventId b) {
  if (!c->ev_valid[a] || !c->ev_valid[b]) return 0.0f;
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, c->ev[a], c->ev[b]) != hipSuccess) return 0.0f;
  return ms;
}

// Device allocations owned by one batch instead of the context's scratch: a bump allocator over
// chunks of 64 MiB, or 1/16 of a smaller budget (at least 4 MiB), so a batch of hundreds of meshes makes
// a handful of device allocations, the chunk sizes recur from batch to batch so the context's size-keyed
// cache serves them, and a small budget is not spent on one chunk.  Larger requests get a block of their
// own, rounded to 2 MiB.  Sub-allocations are 256-byte aligned.
struct Arena {
  ib_context *c;
  std::vector<std::pair<void *, size_t>> blocks;
  char *cur = nullptr;
  size_t left = 0;
  size_t chunk() const {
    const size_t k = std::min<size_t>(64ull << 20, std::max<size_t>(4ull << 20, c->budget / 16));
    return k & ~size_t((2u << 20) - 1);
  }
  template <typename T>
  ib_status get(size_t count, T **out) {
    const size_t bytes = (std::max<size_t>(count * sizeof(T), 256) + 255) & ~size_t(255);
    if (bytes > left) {
      const size_t kc = chunk();
      if (bytes > kc / 4) {  // big: a block of its own, the current chunk stays open
        const size_t bb = (bytes + (2u << 20) - 1) & ~size_t((2u << 20) - 1);
        void *p = nullptr;
        IB_TRY(c->alloc_raw(bb, &p));
        blocks.push_back({p, bb});
        *out = static_cast<T *>(p);
        return IB_OK;
      }
      void *p = nullptr;
      IB_TRY(c->alloc_raw(kc, &p));
      blocks.push_back({p, kc});
      cur = static_cast<char *>(p);
      left = kc;
    }
    *out = reinterpret_cast<T *>(cur);
    cur += bytes;
    left -= bytes;
    return IB_OK;
  }
  void release() {
    for (auto &b : blocks) c->free_raw(b.first, b.second);
    blocks.clear();
    cur = nullptr;
    left = 0;
  }
  ~Arena() { release(); }
};

struct MeshDev {
  int n = 0;
  Node *nodes = nullptr;
  float4 *tris = nullptr;
  TriRec *recs = nullptr;
  Node4 *nodes4 = nullptr;
  Obb *obbs = nullptr;
  Obb *node_obbs = nullptr;
  float *caps = nullptr;
  int ncap = 0;
  CapNode *cap_nodes = nullptr;
  int ncapnode = 0;
  int4 *cap_cut = nullptr;
  int2 *bnd = nullptr;  // the weld's unmatched edges, until the caps are built
  bool caps_built = false;  // the caps were built beside the tree (early_caps)
  int depth = 0;
  bool slivers = false;
};

template <typename T>
ib_status take(ib_context *c, Arena *arena, const char *name, size_t count, T **out) {
  return arena ? arena->get(count, out) : c->scratch(name, count, out);
}

// weld by exact position and count unmatched directed edges into sc; boundary edges land in the
// "e_bnd" scratch (valid until the next mesh is welded)
struct WeldBufs {
  uint32_t vcap = 0, ecap = 0;
  int32_t *slots = nullptr, *minid = nullptr, *cslot = nullptr, *enet = nullptr;
  unsigned long long *ekeys = nullptr;
  int2 *bnd = nullptr;
};

// the weld's tables, and the fills that clear them appended to `l`
ib_status take_weld(ib_context *c, int n, Arena *arena, WeldBufs &w, FillList &l) {
  w.vcap = pow2_at_least(2ull * 3 * n);
  w.ecap = pow2_at_least(2ull * 3 * n);
  IB_TRY(take(c, arena, "w_slots", w.vcap, &w.slots));
  IB_TRY(take(c, arena, "w_minid", w.vcap, &w.minid));
  IB_TRY(take(c, arena, "w_cslot", 3ull * n, &w.cslot));
  IB_TRY(take(c, arena, "e_keys", w.ecap, &w.ekeys));
  IB_TRY(take(c, arena, "e_net", w.ecap, &w.enet));
  IB_TRY(take(c, arena, "e_bnd", 3ull * n, &w.bnd));
  l.add(reinterpret_cast<uint32_t *>(w.slots), 0xffffffffu, w.vcap);
  l.add(reinterpret_cast<uint32_t *>(w.minid), 0x7fffffffu, w.vcap);
  l.add(reinterpret_cast<uint32_t *>(w.ekeys), 0xffffffffu, 2ull * w.ecap);
  l.add(reinterpret_cast<uint32_t *>(w.enet), 0u, w.ecap);
  return IB_OK;
}

// weld by exact position and count unmatched directed edges into sc; boundary edges land in the
// "e_bnd" scratch (valid until the next mesh is welded)
ib_status enqueue_weld(ib_context *c, const float *d_xyz, int n, Scalars *sc, int2 **bnd, hipStream_t s,
                       Arena *arena = nullptr) {
  WeldBufs w;
  FillList l;
  IB_TRY(take_weld(c, n, arena, w, l));
  fill_multi(l, s);
  launch_weld(d_xyz, n, w.slots, w.minid, w.vcap - 1, w.cslot, w.ekeys, w.enet, w.ecap - 1, sc, w.bnd, 3u * n, s);
  *bnd = w.bnd;
  return IB_OK;
}

ib_status enqueue_tree(ib_context *c, const float *d_xyz, int n, Scalars *sc, MeshDev &m, Arena *arena,
                       hipStream_t s, Arena *out, bool side_depth);

// ingest + LBVH of one mesh whose soup is on the device.  Outputs and temporaries come from `arena`
// when given (a batch builds its meshes on several streams at once, so each needs its own), else
// from the context scratch (reuse is ordered by the stream).  With `overlap` (a plain bake, or the only
// mesh of a batch) the weld and the depth walk run on side streams beside the tree build.  Nothing is
// read back here.
ib_status enqueue_build(ib_context *c, const float *d_xyz, int n, Scalars *sc, MeshDev &m, Arena *arena,
                        hipStream_t s, Arena *keep = nullptr, bool overlap = false) {
  Arena *out = keep ? keep : arena;  // the tree and records outlive the temporaries when `keep` is given
  // the scalars (cmin first, all ones; the rest zero) and the weld tables, cleared in one launch
  static_assert(offsetof(Scalars, cmin) == 0, "cmin leads the scalars");
  WeldBufs w;
  FillList l;
  l.add(sc->cmin, 0xffffffffu, 3);
  l.add(reinterpret_cast<uint32_t *>(sc) + 3, 0u, sizeof(Scalars) / 4 - 3);
  IB_TRY(take_weld(c, n, arena, w, l));
  fill_multi(l, s);
  // the weld runs on a side stream beside the tree build (they share only the cleared scalars); the
  // build's stream waits for the weld before it returns, on every path
  overlap = overlap || !arena;
  hipStream_t ws = s;
  if (overlap) {
    IB_HIP(hipEventRecord(c->fork, s));
    IB_HIP(hipStreamWaitEvent(c->side[1], c->fork, 0));
    ws = c->side[1];
  }
  launch_weld(d_xyz, n, w.slots, w.minid, w.vcap - 1, w.cslot, w.ekeys, w.enet, w.ecap - 1, sc, w.bnd, 3u * n, ws);
  m.bnd = w.bnd;
  ib_status st = IB_OK;
  if (overlap) {
    if (!arena) c->record(EV_INGEST, ws);
    IB_HIP(hipEventRecord(c->join[1], ws));
  }
  if (st == IB_OK) st = enqueue_tree(c, d_xyz, n, sc, m, arena, s, out, overlap);
  if (overlap) IB_HIP(hipStreamWaitEvent(s, c->join[1], 0));
  if (st == IB_OK && !arena) c->record(EV_BVH, s);
  return st;
}

// Morton codes, sort, Karras tree, records, refit, 4-wide view
ib_status enqueue_tree(ib_context *c, const float *d_xyz, int n, Scalars *sc, MeshDev &m, Arena *arena,
                       hipStream_t s, Arena *out, bool side_depth) {
  uint32_t *codes_in, *codes, *visits;
  int32_t *idx, *order, *leaf_parent;
  float *root_box;
  IB_TRY(take(c, arena, "codes_in", n, &codes_in));
  IB_TRY(take(c, arena, "codes", n, &codes));
  IB_TRY(take(c, arena, "idx", n, &idx));
  IB_TRY(take(c, arena, "order", n, &order));
  IB_TRY(take(c, arena, "leaf_parent", n, &leaf_parent));
  float *leafbox, *blockbox;
  IB_TRY(take(c, arena, "refit_span", (size_t)n + 1, &visits));
  IB_TRY(take(c, arena, "leafbox", 6ull * n, &leafbox));
  IB_TRY(take(c, arena, "blockbox", 6ull * ((n + 255) / 256), &blockbox));
  IB_TRY(take(c, arena, "root_box", 8, &root_box));
  IB_TRY(take(c, out, "nodes", std::max(n - 1, 1), &m.nodes));
  IB_TRY(take(c, out, "tris", 3ull * n, &m.tris));
  IB_TRY(take(c, out, "trirecs", (size_t)n, &m.recs));
  IB_TRY(take(c, out, "nodes4", std::max(n - 1, 1), &m.nodes4));
  IB_TRY(take(c, out, "obbs", (size_t)n, &m.obbs));
  IB_TRY(take(c, out, "node_obbs", (size_t)std::max(n - 1, 1), &m.node_obbs));
  launch_prep(d_xyz, n, sc, s);
  launch_morton(d_xyz, n, sc, codes_in, idx, s);
  size_t temp_bytes = 0;
  IB_HIP(sort_pairs(nullptr, temp_bytes, codes_in, codes, idx, order, n, s));
  char *temp;
  IB_TRY(take(c, arena, "sort_temp", temp_bytes, &temp));
  IB_HIP(sort_pairs(temp, temp_bytes, codes_in, codes, idx, order, n, s));
  int2 *ranges;
  IB_TRY(take(c, arena, "node_ranges", std::max(n - 1, 1), &ranges));
  launch_karras(codes, n, m.nodes, leaf_parent, ranges, s);
  // the parent chains are walked for the depth on a side stream beside the records and the refit
  if (side_depth) {
    IB_HIP(hipEventRecord(c->fork, s));
    IB_HIP(hipStreamWaitEvent(c->side[2], c->fork, 0));
    launch_depth(m.nodes, leaf_parent, n, sc, c->side[2]);
    IB_HIP(hipEventRecord(c->join[2], c->side[2]));
  }
  launch_gather(d_xyz, order, n, m.tris, m.recs, sc, s);
  fill_u32(visits + n, 0u, 1, s);  // the count of nodes that span 256-leaf blocks
  const int block = (uint64_t)n <= (uint64_t)kBlockMask + 1 ? c->leaf_block : 1;  // leaf codes index 2^26 triangles
  launch_refit(m.tris, n, m.nodes, leaf_parent, visits, leafbox, blockbox, root_box, sc, ranges, block, m.nodes4,
               m.obbs, m.node_obbs, !side_depth, s);
  if (side_depth) IB_HIP(hipStreamWaitEvent(s, c->join[2], 0));
  m.n = n;
  return IB_OK;
}

void enqueue_occupancy(const float *d_xyz, int n, const ib_grid &occ, uint32_t *d_occ, hipStream_t s) {
  const uint64_t words = (uint64_t)((occ.nx + 31) / 32) * occ.ny * occ.nz;
  fill_u32(d_occ, 0u, words, s);
  launch_occupancy(d_xyz, n, to_args(occ), d_occ, s);
}

// Hole caps of the ne unmatched edges `bnd` (vertex ids = first corner of each welded position), built on
// the stream without a host round trip: the loops are labelled on the device, each closed by a fan from its
// mean start vertex, and the fans get their tree (the kernels' comments give the steps).  The cap soup and
// tree come from `arena` (or the context's scratch); the temporaries from the context's scratch, one work
// region carved per step (offsets in units of the edge count rounded up to 64), free again once the stream
// has run the steps (every user of the scratch runs on the same stream or after it).
ib_status enqueue_caps(ib_context *c, const float *d_xyz, const int2 *bnd, uint32_t ne, MeshDev &m,
                       Arena *arena, hipStream_t s) {
  if (ne == 0) return IB_OK;
  const size_t st = ((size_t)ne + 63) & ~size_t(63);
  char *w;
  IB_TRY(c->scratch("cap_work", 108 * st, &w));
  auto u64at = [&](size_t off) { return reinterpret_cast<unsigned long long *>(w + off * st); };
  auto u32at = [&](size_t off) { return reinterpret_cast<uint32_t *>(w + off * st); };
  unsigned long long *ek0 = u64at(0), *ek1 = u64at(8), *ik0 = u64at(16), *ik1 = u64at(24);
  uint32_t *iv0 = u32at(32), *iv1 = u32at(36);
  unsigned long long *acc = u64at(40);
  float *tmp = reinterpret_cast<float *>(w + 72 * st);
  uint32_t *bounds;
  IB_TRY(c->scratch("cap_bounds", 8, &bounds));
  IB_TRY(take(c, arena, "caps", 9ull * ne, &m.caps));
  m.ncap = (int)ne;
  m.ncapnode = ne > 1 ? (int)ne - 1 : 1;
  IB_TRY(take(c, arena, "cap_nodes", (size_t)m.ncapnode, &m.cap_nodes));
  IB_TRY(take(c, arena, "cap_cut", (size_t)kCapCut, &m.cap_cut));
  // vertex ids are corner indices below 3n: keys of 32 + bits(3n - 1) bits
  int idbits = 1;
  while (idbits < 32 && ((uint64_t)3 * m.n - 1) >> idbits) ++idbits;
  int labbits = 1;
  while (labbits < 32 && ((uint64_t)ne - 1) >> labbits) ++labbits;
  // the largest sort space of the three sorts below
  size_t tb = 0, t1 = 0, t2 = 0, t3 = 0;
  IB_HIP(sort_keys_u64(nullptr, t1, ek0, ek1, ne, 32 + idbits, s));
  IB_HIP(sort_pairs_u64(nullptr, t2, ik0, ik1, iv0, iv1, ne, 32 + idbits, s));
  IB_HIP(sort_pairs_u32(nullptr, t3, u32at(16), u32at(24), reinterpret_cast<int32_t *>(iv0),
                        reinterpret_cast<int32_t *>(u32at(28)), ne, labbits, s));
  tb = std::max(t1, std::max(t2, t3));
  size_t t4 = 0;
  IB_HIP(sort_pairs(nullptr, t4, iv0, u32at(24), reinterpret_cast<int32_t *>(iv1),
                    reinterpret_cast<int32_t *>(u32at(28)), (int)ne, s));
  tb = std::max(tb, t4);
  char *temp;
  IB_TRY(c->scratch("cap_sort", tb, &temp));
  // 1. edges in (from, to) order; 2. the same edges in (to, from) order with their (from, to) positions
  launch_cap_keys(bnd, ne, ek0, s);
  size_t tbx = tb;
  IB_HIP(sort_keys_u64(temp, tbx, ek0, ek1, ne, 32 + idbits, s));
  launch_cap_in_keys(ek1, ne, ik0, iv0, s);
  tbx = tb;
  IB_HIP(sort_pairs_u64(temp, tbx, ik0, ik1, iv0, iv1, ne, 32 + idbits, s));
  // 3. successors and loop labels (ping-pong between the free halves of ik0 and ek0)
  uint32_t *lab[2] = {u32at(16), u32at(0)}, *jmp[2] = {u32at(20), u32at(4)};
  launch_cap_succ(ek1, ik1, iv1, ne, lab[0], jmp[0], s);
  int cur = 0;
  for (uint64_t span = 2; span < ne; span = 2 * span - 1) {  // labels cover span edges of each loop
    launch_cap_jump(lab[cur], jmp[cur], ne, lab[cur ^ 1], jmp[cur ^ 1], s);
    cur ^= 1;
  }
  const uint32_t *label = lab[cur];
  // 4. edges sorted by label (into ik1's space), the loops' fixed-point sums
  uint32_t *slab = u32at(24), *sedge = u32at(28);
  launch_iota(iv0, ne, s);
  tbx = tb;
  IB_HIP(sort_pairs_u32(temp, tbx, label, slab, reinterpret_cast<const int32_t *>(iv0),
                        reinterpret_cast<int32_t *>(sedge), ne, labbits, s));
  FillList fl{};
  fl.add(bounds, 0xffffffffu, 3);
  fl.add(bounds + 3, 0u, 3);
  fl.add(reinterpret_cast<uint32_t *>(acc), 0u, 8ull * ne);
  fill_multi(fl, s);
  launch_cap_bounds(ek1, ne, d_xyz, bounds, s);
  launch_cap_sums(slab, sedge, ne, ek1, d_xyz, bounds, acc, s);
  // 5. one fan triangle per edge, sorted by the Morton code of its centroid, into the mesh's cap soup
  uint32_t *codes = iv0;
  int32_t *idx = reinterpret_cast<int32_t *>(iv1);
  launch_cap_tris(ek1, label, ne, d_xyz, bounds, acc, tmp, codes, idx, s);
  uint32_t *codes_s = u32at(24);
  int32_t *order = reinterpret_cast<int32_t *>(u32at(28));
  tbx = tb;
  IB_HIP(sort_pairs(temp, tbx, codes, codes_s, idx, order, (int)ne, s));
  launch_cap_gather(tmp, order, ne, m.caps, s);
  // 6. the tree (children, parents and ranges in acc's space)
  int2 *child = reinterpret_cast<int2 *>(w + 40 * st), *ranges = reinterpret_cast<int2 *>(w + 48 * st);
  int32_t *parent = reinterpret_cast<int32_t *>(w + 56 * st);
  launch_cap_tree(m.caps, codes_s, (int)ne, child, parent, ranges, m.cap_nodes, m.cap_cut, s);
  return IB_OK;
}

// after the scalars are on the host: validity, depth, and hole caps for open meshes (from the unmatched
// edges the build's weld left in m.bnd: the context's scratch for a plain bake, the build's arena otherwise)
ib_status finish_build(ib_context *c, const float *d_xyz, int n, const Scalars &h, MeshDev &m, Arena *arena,
                       hipStream_t s) {
  if (h.flags & kFlagNonFinite) return fail(IB_ERR_INVALID_MESH, "mesh has non-finite vertex coordinates");
  m.depth = (int)h.depth;
  m.slivers = h.slivers > 0;
  if (m.depth > kMaxDepth) return fail(IB_ERR_INTERNAL, "BVH parent chain does not reach the root");
  if (m.depth > kMaxTreeDepth)
    return fail(IB_ERR_INTERNAL, "BVH depth " + std::to_string(m.depth) + " exceeds the bound of the key length (" +
                                     std::to_string(kMaxTreeDepth) + ")");
  if (m.caps_built) return IB_OK;
  m.ncap = 0;
  m.caps = nullptr;
  if (h.boundary == 0) return IB_OK;
  if (!c->ev_valid[EV_CAP0]) c->record(EV_CAP0, s);
  if (!m.bnd) return fail(IB_ERR_INTERNAL, "the weld's unmatched edges are gone");
  const uint32_t ne = std::min<uint32_t>(h.boundary, 3u * n);
  IB_TRY(enqueue_caps(c, d_xyz, m.bnd, ne, m, arena, s));
  m.bnd = nullptr;  // temporaries: not kept with a resident mesh
  c->record(EV_CAP1, s);
  return IB_OK;
}

// Hole caps of a mesh built alone (a plain bake, a batch of one, a resident mesh; enqueue_build with
// `overlap`, whose weld ran on side[1]): the weld's boundary count comes back as soon as the weld is done,
// and the caps are built on side[1] beside the rest of the tree build; `s` waits for them (join), or, with
// join false, the caller makes `s` wait for c->join[1] before the first use of the caps.  The caps need
// only the soup and the unmatched edges.  finish_build then leaves them as they are.
ib_status early_caps(ib_context *c, const float *d_xyz, int n, const Scalars *sc, MeshDev &m, Arena *arena,
                     hipStream_t s, bool join = true) {
  hipStream_t ws = c->side[1];
  IB_HIP(hipMemcpyAsync(c->host_sc_early, sc, sizeof(Scalars), hipMemcpyDeviceToHost, ws));
  IB_HIP(hipStreamSynchronize(ws));
  const uint32_t boundary = c->host_sc_early->boundary;
  m.caps_built = true;
  m.ncap = 0;
  m.caps = nullptr;
  if (boundary == 0) return IB_OK;
  if (!m.bnd) return fail(IB_ERR_INTERNAL, "the weld's unmatched edges are gone");
  if (!c->ev_valid[EV_CAP0]) c->record(EV_CAP0, ws);
  const uint32_t ne = std::min<uint32_t>(boundary, 3u * n);
  const ib_status st = enqueue_caps(c, d_xyz, m.bnd, ne, m, arena, ws);
  m.bnd = nullptr;
  if (st == IB_OK) c->record(EV_CAP1, ws);
  IB_HIP(hipEventRecord(c->join[1], ws));
  if (join) IB_HIP(hipStreamWaitEvent(s, c->join[1], 0));
  return st;
}

// brick selection: a brick is in the band when its centre is within band * h + 3.5 * sqrt(3) * h of the
// surface (1-Lipschitz distance: then every voxel within band * h of the surface lies in a kept brick)
float band_thresh(const ib_job &jb) { return jb.sdf.h * ((float)jb.band_voxels + 6.0621778f); }

// voxels on a hole cap per query whose mesh winding numbers are summed in chunks across the chip (more go
// to the wave that finds them)
constexpr uint32_t kSingCap = 1024;

// column crossings of one grid + the query argument block (buffers from arena or scratch)
ib_status prepare_query(ib_context *c, const MeshDev &m, const ib_grid &g, float *d_out, Scalars *sc, Arena *arena,
                        uint32_t amb_cap, SdfArgs &sa, hipStream_t s, bool columns = true) {
  const uint64_t ncol = (uint64_t)g.nx * g.ny;
  ColumnArgs ca{};
  ca.nodes = m.nodes;
  ca.nnodes = std::max(m.n - 1, 1);
  ca.ntri = m.n;
  ca.tris = m.tris;
  ca.recs = m.recs;
  ca.nodes4 = m.nodes4;
  ca.obbs = m.obbs;
  ca.node_obbs = m.node_obbs;
  ca.caps = m.caps;
  ca.ncap = m.ncap;
  ca.cap_nodes = m.cap_nodes;
  ca.cap_cut = m.cap_cut;
  ca.ncapnode = m.ncapnode;
  ca.g = to_args(g);
  ca.sc = sc;
  IB_TRY(take(c, arena, "col_cross", ncol * kColCap, &ca.cross));
  IB_TRY(take(c, arena, "col_count", ncol, &ca.count));
  if (columns) launch_columns(ca, s);
  sa = SdfArgs{};
  sa.col = ca;
  sa.out = d_out;
  sa.amb_cap = amb_cap;
  sa.resolve_full = c->resolve_full ? 1u : 0u;
  IB_TRY(take(c, arena, "amb_list", amb_cap, &sa.amb_list));
  if (m.ncap > 0) {  // only hole caps put a voxel on a cap or leave its FP64 recheck to the chunked sums
    sa.sing_cap = std::min<uint32_t>(amb_cap, kSingCap);
    sa.sing_chunks = (uint32_t)((std::max(m.n, m.ncap) + kSingChunk - 1) / kSingChunk);
    IB_TRY(take(c, arena, "sing_list", sa.sing_cap, &sa.sing));
    IB_TRY(take(c, arena, "sing_kind", sa.sing_cap, &sa.sing_kind));
    IB_TRY(take(c, arena, "sing_part", (size_t)sa.sing_cap * sa.sing_chunks, &sa.sing_part));
  }
  return IB_OK;
}

ib_status query_dense(ib_context *c, const MeshDev &m, const ib_grid &g, float *d_out, hipStream_t s) {
  Scalars *sc;
  IB_TRY(c->scratch("sc", 1, &sc));
  const uint64_t nvox = (uint64_t)g.nx * g.ny * g.nz;
  if (nvox >= (1ull << 32)) return fail(IB_ERR_UNSUPPORTED, "grid too large for 32-bit voxel indices");
  SdfArgs sa;
  IB_TRY(prepare_query(c, m, g, d_out, sc, nullptr, (uint32_t)nvox, sa, s));
  c->record(EV_COLS, s);
  const uint32_t *order = nullptr;
  if (c->query_variant == 1 && c->tile_order) {
    // costliest block-tiles first: sort the tiles by the column-crossing cost estimate
    uint32_t d[3];
    tile_dims(sa.col.g, c->brick_shape, d);
    const uint32_t nt = d[0] * d[1] * d[2];
    uint32_t *keys_in, *keys;
    int32_t *ids_in, *ids;
    IB_TRY(c->scratch("tile_keys_in", nt, &keys_in));
    IB_TRY(c->scratch("tile_keys", nt, &keys));
    IB_TRY(c->scratch("tile_ids_in", nt, &ids_in));
    IB_TRY(c->scratch("tile_ids", nt, &ids));
    launch_tile_keys(sa.col, c->brick_shape, c->tile_order, keys_in, ids_in, s);
    size_t temp_bytes = 0;
    IB_HIP(sort_pairs_u32(nullptr, temp_bytes, keys_in, keys, ids_in, ids, nt, 32, s));
    char *temp;
    IB_TRY(c->scratch("tile_sort_temp", temp_bytes, &temp));
    IB_HIP(sort_pairs_u32(temp, temp_bytes, keys_in, keys, ids_in, ids, nt, 32, s));
    order = reinterpret_cast<const uint32_t *>(ids);
  }
  c->record(EV_Q0, s);
  launch_sdf(sa, c->count_work, c->query_variant, m.slivers, c->brick_shape, order, s);
  c->record(EV_Q1, s);
  launch_resolve(sa, s);
  c->record(EV_SDF, s);
  return IB_OK;
}

// Device -> host copies of large results into the caller's (pageable) buffers: one DMA of all ranges into a
// pinned staging buffer the context keeps (grown to the largest result), then the host copies the staging
// out on several threads.  A pageable destination otherwise goes through the runtime's own small bounce
// buffers at a fraction of the link's rate.  Small results, and a failed staging allocation, copy directly.
// Waits for the stream.
struct HostCopy {
  void *dst;
  const void *src;
  size_t bytes;
};
ib_status download(ib_context *c, const HostCopy *cp, int n, hipStream_t s) {
  size_t total = 0;
  for (int i = 0; i < n; ++i) total += (cp[i].bytes + 255) & ~size_t(255);
  if (total >= (8u << 20) && c->staging_bytes < total) {
    if (c->staging) (void)hipHostFree(c->staging);
    c->staging = nullptr;
    c->staging_bytes = 0;
    void *p = nullptr;
    if (hipHostMalloc(&p, total, hipHostMallocDefault) == hipSuccess) {
      c->staging = static_cast<char *>(p);
      c->staging_bytes = total;
    }
  }
  if (total < (8u << 20) || !c->staging) {
    for (int i = 0; i < n; ++i)
      if (cp[i].bytes) IB_HIP(hipMemcpyAsync(cp[i].dst, cp[i].src, cp[i].bytes, hipMemcpyDeviceToHost, s));
    IB_HIP(hipStreamSynchronize(s));
    return IB_OK;
  }
  std::vector<size_t> off(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    off[i + 1] = off[i] + ((cp[i].bytes + 255) & ~size_t(255));
    if (cp[i].bytes)
      IB_HIP(hipMemcpyAsync(c->staging + off[i], cp[i].src, cp[i].bytes, hipMemcpyDeviceToHost, s));
  }
  IB_HIP(hipStreamSynchronize(s));
  // the bytes of all ranges, cut into equal slices, one per thread
  size_t sum = 0;
  for (int i = 0; i < n; ++i) sum += cp[i].bytes;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nt = (int)std::min<size_t>(std::min(8u, hw), std::max<size_t>(1, sum >> 23));
  auto work = [&](int t) {
    size_t lo = sum * t / nt, hi = sum * (t + 1) / nt, base = 0;
    for (int i = 0; i < n && lo < hi; ++i) {
      const size_t b0 = base, b1 = base + cp[i].bytes;
      base = b1;
      if (hi <= b0 || lo >= b1) continue;
      const size_t a = std::max(lo, b0), e = std::min(hi, b1);
      std::memcpy(static_cast<char *>(cp[i].dst) + (a - b0), c->staging + off[i] + (a - b0), e - a);
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
  work(0);
  for (auto &x : th) x.join();
  return IB_OK;
}

// empty mesh: no surface anywhere
ib_status fill_empty(ib_context *c, const ib_grid *sdf, float *d_dist, const ib_grid *occ, uint32_t *d_occ,
                     hipStream_t s) {
  (void)c;
  if (sdf && d_dist)
    fill_u32(reinterpret_cast<uint32_t *>(d_dist), 0x7f800000u, (uint64_t)sdf->nx * sdf->ny * sdf->nz, s);
  if (occ && d_occ) fill_u32(d_occ, 0u, (uint64_t)((occ->nx + 31) / 32) * occ->ny * occ->nz, s);
  return IB_OK;
}

void reset_stats(ib_context *c) {
  c->stats = ib_stats{};
  c->stats_in_sc = false;
  for (bool &v : c->ev_valid) v = false;
}

ib_status bake_device(ib_context *c, const float *d_xyz, uint64_t ntri, const ib_grid *sdf, float *d_dist,
                      const ib_grid *occ, uint32_t *d_occ, hipStream_t s) {
  if (sdf && !grid_ok(sdf)) return fail(IB_ERR_INVALID_ARGUMENT, "invalid SDF grid");
  if (occ && !grid_ok(occ)) return fail(IB_ERR_INVALID_ARGUMENT, "invalid occupancy grid");
  if ((sdf && !d_dist) || (occ && !d_occ)) return fail(IB_ERR_INVALID_ARGUMENT, "missing output buffer");
  if (ntri > (uint64_t)(INT_MAX / 9)) return fail(IB_ERR_INVALID_ARGUMENT, "too many triangles");
  reset_stats(c);
  c->stats.triangles = ntri;
  c->stats.voxels = sdf ? (uint64_t)sdf->nx * sdf->ny * sdf->nz : 0;
  c->stats.occ_cells = occ ? (uint64_t)occ->nx * occ->ny * occ->nz : 0;
  if (ntri == 0) return fill_empty(c, sdf, d_dist, occ, d_occ, s);
  if (!d_xyz) return fail(IB_ERR_INVALID_ARGUMENT, "null triangle buffer");
  MeshDev m;
  Scalars *sc;
  IB_TRY(c->scratch("sc", 1, &sc));
  c->stats_in_sc = true;
  c->last_sc = sc;
  c->record(EV_START, s);
  // occupancy needs only the soup: it runs on a side stream beside the build, the columns and the
  // query, and the caller's stream waits for it before anything after this call
  const bool side_occ = occ && d_occ;
  if (side_occ) {
    IB_HIP(hipEventRecord(c->fork, s));
    IB_HIP(hipStreamWaitEvent(c->side[0], c->fork, 0));
    enqueue_occupancy(d_xyz, (int)ntri, *occ, d_occ, c->side[0]);
    c->record(EV_OCC_SIDE, c->side[0]);
    IB_HIP(hipEventRecord(c->join[0], c->side[0]));
  }
  const ib_status st = [&]() -> ib_status {
    IB_TRY(enqueue_build(c, d_xyz, (int)ntri, sc, m, nullptr, s));
    IB_TRY(early_caps(c, d_xyz, (int)ntri, sc, m, nullptr, s));
    c->record(EV_OCC, s);
    IB_HIP(hipMemcpyAsync(c->host_sc, sc, sizeof(Scalars), hipMemcpyDeviceToHost, s));
    IB_HIP(hipStreamSynchronize(s));
    IB_TRY(finish_build(c, d_xyz, (int)ntri, *c->host_sc, m, nullptr, s));
    c->stats.bvh_depth = m.depth;
    c->stats.boundary_edges = c->host_sc->boundary;
    if (sdf) IB_TRY(query_dense(c, m, *sdf, d_dist, s));
    return IB_OK;
  }();
  if (side_occ) IB_HIP(hipStreamWaitEvent(s, c->join[0], 0));
  return st;
}

}  // namespace

// Scratch of the scratch-free bake (ib_bake_dense_device_noalloc): one fixed region per device, part of
// the library's code object, so every process that loads the library holds kNoAllocScratch bytes on each
// device it uses until it exits.  Host code never names the array except to look up its device address
// (noalloc_scratch_ptr); kernels get that address.  All calls on one device share the region.
constexpr size_t kNoAllocScratch = 128ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];
constexpr int kMaxDevices = 64;
static void *g_noalloc_ptr[kMaxDevices] = {};
static std::mutex g_noalloc_mu;

// the region's device address on `device`, looked up once per device with that device current
static ib_status noalloc_scratch_ptr(int device, void **out) {
  if (device < 0 || device >= kMaxDevices) return fail(IB_ERR_UNSUPPORTED, "device ordinal above 63");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(IB_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_noalloc_ptr[device] = p;
  }
  *out = g_noalloc_ptr[device];
  return IB_OK;
}

// A resident mesh: its soup copy, tree, records and caps live in `keep`, which belongs to context c.
struct ib_mesh {
  ib_context *c;
  Arena keep;
  MeshDev m;
  float *d_xyz = nullptr;
  uint64_t ntri = 0;
  uint32_t boundary = 0;
  uint64_t bytes = 0;
  explicit ib_mesh(ib_context *ctx) : c(ctx), keep{ctx} {}
};

// ============================================================================================
// C ABI
// ============================================================================================
extern "C" {

ib_status ib_mesh_create(ib_context *c, const float *d_xyz, uint64_t ntri, void *stream, ib_mesh **out) {
  if (!c || !out) return fail(IB_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (ntri == 0 || !d_xyz) return fail(IB_ERR_INVALID_ARGUMENT, "a resident mesh needs a device soup of triangles");
  if (ntri > (uint64_t)(INT_MAX / 9)) return fail(IB_ERR_INVALID_ARGUMENT, "too many triangles");
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
  std::unique_ptr<ib_mesh> mh(new ib_mesh(c));
  Arena temp{c};  // weld tables, Morton codes, sort space: released when the build is done
  Ordered ord(c, s);  // after the arenas: a failed build waits for its kernels before the blocks are released
  Scalars *sc;
  IB_TRY(temp.get(1, &sc));
  IB_TRY(mh->keep.get(9 * ntri, &mh->d_xyz));
  IB_HIP(hipMemcpyAsync(mh->d_xyz, d_xyz, 9 * ntri * sizeof(float), hipMemcpyDeviceToDevice, s));
  const int n = (int)ntri;
  IB_TRY(enqueue_build(c, mh->d_xyz, n, sc, mh->m, &temp, s, &mh->keep, true));
  IB_TRY(early_caps(c, mh->d_xyz, n, sc, mh->m, &mh->keep, s));
  IB_HIP(hipMemcpyAsync(c->host_sc, sc, sizeof(Scalars), hipMemcpyDeviceToHost, s));
  IB_HIP(hipStreamSynchronize(s));
  const Scalars h = *c->host_sc;
  IB_TRY(finish_build(c, mh->d_xyz, n, h, mh->m, &mh->keep, s));
  IB_HIP(hipStreamSynchronize(s));
  mh->ntri = ntri;
  mh->boundary = h.boundary;
  for (const auto &b : mh->keep.blocks) mh->bytes += b.second;
  *out = mh.release();
  return IB_OK;
}

void ib_mesh_destroy(ib_mesh *mh) {
  if (!mh) return;
  {
    std::lock_guard<std::mutex> lk(mh->c->mu);  // the blocks go back to the context's cache
    mh->keep.release();
  }
  delete mh;
}

uint64_t ib_mesh_bytes(const ib_mesh *mh) { return mh ? mh->bytes : 0; }
uint64_t ib_mesh_triangles(const ib_mesh *mh) { return mh ? mh->ntri : 0; }

const char *ib_last_error(void) { return g_err.c_str(); }
const char *ib_version(void) { return "italy-bake 0.1 gfx950"; }

ib_status ib_create(int device, uint64_t budget, ib_context **out) {
  if (!out) return fail(IB_ERR_INVALID_ARGUMENT, "null output");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(IB_ERR_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= count) return fail(IB_ERR_INVALID_ARGUMENT, "device ordinal out of range");
  IB_HIP(hipSetDevice(device));
  auto *c = new ib_context();
  c->device = device;
  size_t freeb = 0, total = 0;
  if (hipMemGetInfo(&freeb, &total) != hipSuccess) {
    delete c;
    return fail(IB_ERR_DEVICE, "hipMemGetInfo failed");
  }
  c->budget = budget ? budget : (uint64_t)(0.9 * (double)freeb);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void **)&c->host_sc, sizeof(Scalars), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&c->host_sc_early, sizeof(Scalars), hipHostMallocDefault) != hipSuccess) {
    delete c;
    return fail(IB_ERR_DEVICE, "stream / pinned allocation failed");
  }
  for (auto &e : c->ev) (void)hipEventCreate(&e);
  for (int i = 0; i < ib_context::kSide; ++i) {
    if (hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->join[i], hipEventDisableTiming) != hipSuccess) {
      ib_destroy(c);
      return fail(IB_ERR_DEVICE, "stream creation failed");
    }
  }
  if (hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->tail, hipEventDisableTiming) != hipSuccess) {
    ib_destroy(c);
    return fail(IB_ERR_DEVICE, "event creation failed");
  }
  hipDeviceProp_t prop;
  c->cus = hipGetDeviceProperties(&prop, device) == hipSuccess ? prop.multiProcessorCount : 256;
  if (noalloc_scratch_ptr(device, &c->noalloc_base) != IB_OK) {
    const std::string e = g_err;
    ib_destroy(c);
    return fail(IB_ERR_DEVICE, e);
  }
  *out = c;
  return IB_OK;
}

void ib_destroy(ib_context *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto &kv : c->bufs) (void)hipFree(kv.second.p);
  for (auto &kv : c->user) (void)hipFree(kv.first);
  c->release_cache();
  for (auto &e : c->ev) (void)hipEventDestroy(e);
  (void)hipHostFree(c->host_sc);
  if (c->host_sc_early) (void)hipHostFree(c->host_sc_early);
  if (c->host_batch_sc) (void)hipHostFree(c->host_batch_sc);
  if (c->staging) (void)hipHostFree(c->staging);
  for (int i = 0; i < ib_context::kSide; ++i) {
    if (c->side[i]) (void)hipStreamSynchronize(c->side[i]);
    if (c->side[i]) (void)hipStreamDestroy(c->side[i]);
    if (c->join[i]) (void)hipEventDestroy(c->join[i]);
  }
  if (c->fork) (void)hipEventDestroy(c->fork);
  if (c->tail) (void)hipEventDestroy(c->tail);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

ib_status ib_sdf_framing(const float bmin[3], const float bmax[3], int32_t res, ib_grid *g) {
  if (!bmin || !bmax || !g) return fail(IB_ERR_INVALID_ARGUMENT, "null argument");
  // sdf_baker.cpp:139-148
  float ext[3];
  for (int k = 0; k < 3; ++k) ext[k] = bmax[k] - bmin[k];
  const float longest = std::max({ext[0], ext[1], ext[2], 1e-6f});
  const float vs = longest / static_cast<float>(std::max(res, 1));
  const float pad = vs * 2.0f;
  const float pad2 = pad * 2.0f;
  for (int k = 0; k < 3; ++k) g->origin[k] = bmin[k] - pad;
  g->nx = static_cast<int>(std::ceil((ext[0] + pad2) / vs)) + 1;
  g->ny = static_cast<int>(std::ceil((ext[1] + pad2) / vs)) + 1;
  g->nz = static_cast<int>(std::ceil((ext[2] + pad2) / vs)) + 1;
  g->h = vs;
  g->reserved = 0;
  return IB_OK;
}

ib_status ib_vox_framing(const float bmin[3], const float bmax[3], int32_t res, float origin[3], float *vs) {
  if (!bmin || !bmax || !origin || !vs) return fail(IB_ERR_INVALID_ARGUMENT, "null argument");
  // voxelize.cpp:54-57
  float ext[3];
  for (int k = 0; k < 3; ++k) ext[k] = bmax[k] - bmin[k];
  const float longest = std::max({ext[0], ext[1], ext[2], 1e-6f});
  *vs = longest / static_cast<float>(std::max(res, 1));
  for (int k = 0; k < 3; ++k) origin[k] = bmin[k];
  return IB_OK;
}

ib_status ib_bake_dense_device(ib_context *c, const float *d_xyz, uint64_t ntri, const ib_grid *sdf, float *d_dist,
                               const ib_grid *occ, uint32_t *d_occ, void *stream) {
  if (!c) return fail(IB_ERR_INVALID_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
  Ordered ord(c, s);
  return bake_device(c, d_xyz, ntri, sdf, d_dist, occ, d_occ, s);
}

ib_status ib_bake_dense_device_noalloc(ib_context *c, const float *d_xyz, uint64_t ntri, const ib_grid *sdf,
                                       float *d_dist, const ib_grid *occ, uint32_t *d_occ, void *stream) {
  if (!c) return fail(IB_ERR_INVALID_ARGUMENT, "null context");
  if (ntri > IB_NOALLOC_MAX_TRIANGLES)
    return fail(IB_ERR_CAPACITY, "the scratch-free bake takes at most " + std::to_string(IB_NOALLOC_MAX_TRIANGLES) +
                                     " triangles (got " + std::to_string(ntri) + ")");
  if (sdf && ((uint64_t)sdf->nx * (uint64_t)sdf->ny > IB_NOALLOC_MAX_COLUMNS ||
              (uint64_t)sdf->nx * sdf->ny * sdf->nz > IB_NOALLOC_MAX_VOXELS))
    return fail(IB_ERR_CAPACITY, "the scratch-free bake takes SDF grids of at most " +
                                     std::to_string(IB_NOALLOC_MAX_COLUMNS) + " columns (nx * ny) and " +
                                     std::to_string(IB_NOALLOC_MAX_VOXELS) + " voxels");
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
  Ordered ord(c, s);
  // the region keeps the previous call's contents: the bake clears every table it reads before use
  c->begin_fixed(c->noalloc_base, kNoAllocScratch);
  const ib_status st = bake_device(c, d_xyz, ntri, sdf, d_dist, occ, d_occ, s);
  c->end_fixed();
  return st;
}

ib_status ib_bake_dense(ib_context *c, const float *xyz, uint64_t ntri, const ib_grid *sdf, float *dist,
                        const ib_grid *occ, uint32_t *occ_words) {
  if (!c) return fail(IB_ERR_INVALID_ARGUMENT, "null context");
  if (ntri && !xyz) return fail(IB_ERR_INVALID_ARGUMENT, "null triangle buffer");
  if ((sdf && !dist) || (occ && !occ_words)) return fail(IB_ERR_INVALID_ARGUMENT, "missing output buffer");
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  Ordered ord(c, s);
  float *d_xyz = nullptr, *d_dist = nullptr;
  uint32_t *d_occ = nullptr;
  const uint64_t nvox = sdf ? (uint64_t)sdf->nx * sdf->ny * sdf->nz : 0;
  const uint64_t nwords = occ ? (uint64_t)((occ->nx + 31) / 32) * occ->ny * occ->nz : 0;
  IB_TRY(c->scratch("in_xyz", std::max<uint64_t>(9 * ntri, 1), &d_xyz));
  if (sdf) IB_TRY(c->scratch("out_dist", nvox, &d_dist));
  if (occ) IB_TRY(c->scratch("out_occ", nwords, &d_occ));
  if (ntri) IB_HIP(hipMemcpyAsync(d_xyz, xyz, 9 * ntri * sizeof(float), hipMemcpyHostToDevice, s));
  IB_TRY(bake_device(c, d_xyz, ntri, sdf, d_dist, occ, d_occ, s));
  if (ntri && sdf) {
    Scalars *sc;
    IB_TRY(c->scratch("sc", 1, &sc));
    IB_HIP(hipMemcpyAsync(c->host_sc, sc, sizeof(Scalars), hipMemcpyDeviceToHost, s));
  }
  const HostCopy cp[2] = {{dist, d_dist, sdf ? nvox * sizeof(float) : 0},
                          {occ_words, d_occ, occ ? nwords * sizeof(uint32_t) : 0}};
  IB_TRY(download(c, cp, 2, s));
  if (ntri && sdf && c->host_sc->stack_overflow)
    return fail(IB_ERR_INTERNAL, "a traversal stack overflowed (the tree exceeds its depth bound)");
  return IB_OK;
}

ib_status ib_bake_sdf(ib_context *c, const float *xyz, uint64_t ntri, const float bmin[3], const float bmax[3],
                      int32_t res, ib_grid *grid, float *dist, uint64_t capacity) {
  if (!c || !grid) return fail(IB_ERR_INVALID_ARGUMENT, "null argument");
  IB_TRY(ib_sdf_framing(bmin, bmax, res, grid));
  if (!dist) return IB_OK;
  const uint64_t nvox = (uint64_t)grid->nx * grid->ny * grid->nz;
  if (capacity < nvox) return fail(IB_ERR_CAPACITY, "distance buffer holds " + std::to_string(capacity) +
                                                        " floats, the grid needs " + std::to_string(nvox));
  return ib_bake_dense(c, xyz, ntri, grid, dist, nullptr, nullptr);
}

// voxelizeMesh (src/convert/voxelize.cpp:52-107) without a dense grid: every (cell, triangle) overlap pair
// of the SAT test over each triangle's clamped cell range, in triangle order; a stable sort by cell puts each
// cell's triangles in order, so the first is the cell's first touch and the colour sum runs in triangle order
// (one thread per cell); the cells are then sorted by (first triangle, z, y, x), the reference's insertion
// order.  Memory and work scale with the pairs.  Two read-backs: the pair count, then the cell count.
static ib_status voxelize_sparse(ib_context *c, const float *xyz, uint64_t ntri, const float bmin[3],
                                 const float bmax[3], int32_t res, const float *tri_colors, int32_t *cells,
                                 int32_t *cell_tri, float *colors, bool want_colors, uint64_t capacity,
                                 uint64_t *count) {
  *count = 0;
  if (ntri == 0) return IB_OK;
  if (!xyz) return fail(IB_ERR_INVALID_ARGUMENT, "null triangle buffer");
  if (res <= 0)
    return fail(IB_ERR_INVALID_ARGUMENT, "resolution must be >= 1 (the reference clamps cells to resolution-1)");
  if ((uint64_t)res * res * res >= (1ull << 32)) return fail(IB_ERR_UNSUPPORTED, "resolution too large");
  if (ntri > (uint64_t)(INT_MAX / 9)) return fail(IB_ERR_INVALID_ARGUMENT, "too many triangles");
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  Ordered ord(c, s);
  ib_grid g{};
  IB_TRY(ib_vox_framing(bmin, bmax, res, g.origin, &g.h));
  g.nx = g.ny = g.nz = res;
  const uint64_t ncell = (uint64_t)res * res * res;
  const int n = (int)ntri;
  float *d_xyz, *d_rgb_in = nullptr;
  uint32_t *counts, *offsets;
  IB_TRY(c->scratch("in_xyz", 9 * ntri, &d_xyz));
  IB_TRY(c->scratch("sp_counts", ntri + 1, &counts));
  IB_TRY(c->scratch("sp_offsets", ntri + 1, &offsets));
  IB_HIP(hipMemcpyAsync(d_xyz, xyz, 9 * ntri * sizeof(float), hipMemcpyHostToDevice, s));
  if (want_colors && tri_colors) {
    IB_TRY(c->scratch("vx_rgb_in", 3 * ntri, &d_rgb_in));
    IB_HIP(hipMemcpyAsync(d_rgb_in, tri_colors, 3 * ntri * sizeof(float), hipMemcpyHostToDevice, s));
  }
  launch_splat_count(d_xyz, n, to_args(g), counts, s);
  fill_u32(counts + n, 0u, 1, s);
  size_t tb = 0;
  IB_HIP(exclusive_scan_u32(nullptr, tb, counts, offsets, (uint32_t)n + 1, s));
  char *temp;
  IB_TRY(c->scratch("sp_temp", tb, &temp));
  IB_HIP(exclusive_scan_u32(temp, tb, counts, offsets, (uint32_t)n + 1, s));
  uint32_t npairs = 0;
  IB_HIP(hipMemcpyAsync(&npairs, offsets + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  IB_HIP(hipStreamSynchronize(s));
  if (npairs == 0) return IB_OK;
  uint32_t *cell, *cell_s, *head, *run_of;
  int32_t *tri, *tri_s;
  IB_TRY(c->scratch("sp_cell", npairs, &cell));
  IB_TRY(c->scratch("sp_cell_s", npairs, &cell_s));
  IB_TRY(c->scratch("sp_tri", npairs, &tri));
  IB_TRY(c->scratch("sp_tri_s", npairs, &tri_s));
  IB_TRY(c->scratch("vx_head", (uint64_t)npairs + 1, &head));
  IB_TRY(c->scratch("vx_run", (uint64_t)npairs + 1, &run_of));
  launch_splat_emit(d_xyz, n, to_args(g), offsets, cell, tri, s);
  int bits = 1;
  while (bits < 32 && (1ull << bits) < ncell) ++bits;
  tb = 0;
  IB_HIP(sort_pairs_u32(nullptr, tb, cell, cell_s, tri, tri_s, npairs, bits, s));
  IB_TRY(c->scratch("sp_temp", tb, &temp));
  IB_HIP(sort_pairs_u32(temp, tb, cell, cell_s, tri, tri_s, npairs, bits, s));
  launch_vox_heads(cell_s, npairs, head, s);
  tb = 0;
  IB_HIP(exclusive_scan_u32(nullptr, tb, head, run_of, npairs + 1, s));
  IB_TRY(c->scratch("sp_temp", tb, &temp));
  IB_HIP(exclusive_scan_u32(temp, tb, head, run_of, npairs + 1, s));
  uint32_t nruns = 0;
  IB_HIP(hipMemcpyAsync(&nruns, run_of + npairs, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  IB_HIP(hipStreamSynchronize(s));
  *count = nruns;
  if (!cells && !cell_tri && !colors) return IB_OK;
  if (capacity < nruns) return fail(IB_ERR_CAPACITY, "cell buffer holds " + std::to_string(capacity) + " cells, " +
                                                         std::to_string(nruns) + " are occupied");
  unsigned long long *keys, *keys_s;
  uint32_t *vals, *perm;
  float *rgb = nullptr;
  IB_TRY(c->scratch("vx_keys", nruns, &keys));
  IB_TRY(c->scratch("vx_keys_s", nruns, &keys_s));
  IB_TRY(c->scratch("vx_vals", nruns, &vals));
  IB_TRY(c->scratch("vx_perm", nruns, &perm));
  if (want_colors) IB_TRY(c->scratch("vx_rgb", 3ull * nruns, &rgb));
  launch_vox_runs(cell_s, tri_s, npairs, run_of, d_rgb_in, keys, vals, rgb, s);
  tb = 0;
  IB_HIP(sort_pairs_u64(nullptr, tb, keys, keys_s, vals, perm, nruns, 64, s));
  IB_TRY(c->scratch("vx_temp", tb, &temp));
  IB_HIP(sort_pairs_u64(temp, tb, keys, keys_s, vals, perm, nruns, 64, s));
  int32_t *d_cells = nullptr, *d_tri = nullptr;
  float *d_colors = nullptr;
  if (cells) IB_TRY(c->scratch("vx_cells", 3ull * nruns, &d_cells));
  if (cell_tri) IB_TRY(c->scratch("vx_tri", nruns, &d_tri));
  if (colors) IB_TRY(c->scratch("vx_colors", 3ull * nruns, &d_colors));
  launch_vox_out(keys_s, perm, nruns, res, res, rgb, d_cells, d_tri, d_colors, s);
  if (cells) IB_HIP(hipMemcpyAsync(cells, d_cells, 3ull * nruns * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  if (cell_tri) IB_HIP(hipMemcpyAsync(cell_tri, d_tri, (uint64_t)nruns * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  if (colors) IB_HIP(hipMemcpyAsync(colors, d_colors, 3ull * nruns * sizeof(float), hipMemcpyDeviceToHost, s));
  IB_HIP(hipStreamSynchronize(s));
  return IB_OK;
}

ib_status ib_voxelize(ib_context *c, const float *xyz, uint64_t ntri, const float bmin[3], const float bmax[3],
                      int32_t res, int32_t *cells, int32_t *cell_tri, uint64_t capacity, uint64_t *count) {
  if (!c || !count) return fail(IB_ERR_INVALID_ARGUMENT, "null argument");
  return voxelize_sparse(c, xyz, ntri, bmin, bmax, res, nullptr, cells, cell_tri, nullptr, false, capacity, count);
}

ib_status ib_voxelize_colors(ib_context *c, const float *xyz, uint64_t ntri, const float bmin[3], const float bmax[3],
                             int32_t res, const float *tri_colors, int32_t *cells, float *colors, uint64_t capacity,
                             uint64_t *count) {
  if (!c || !count) return fail(IB_ERR_INVALID_ARGUMENT, "null argument");
  return voxelize_sparse(c, xyz, ntri, bmin, bmax, res, tri_colors, cells, nullptr, colors, colors != nullptr,
                         capacity, count);
}

static_assert(sizeof(ib_material) == sizeof(MatDev), "material layout");

// srgbToLinear (src/convert/sdf_baker.cpp:42) of every byte value b / 255, evaluated on the host with the
// reference's expression (float division, std::pow) so that the device's table lookup returns its values
static const float *srgb_table() {
  static float lut[256];
  static std::once_flag once;
  std::call_once(once, [] {
    for (int b = 0; b < 256; ++b) {
      const float c = (float)b / 255.0f;
      lut[b] = c <= 0.04045f ? c / 12.92f : std::pow((c + 0.055f) / 1.055f, 2.4f);
    }
  });
  return lut;
}

ib_status ib_triangle_materials(ib_context *c, uint64_t ntri, const float *uvs, uint64_t nuv,
                                const uint32_t *tri_material, uint64_t ntm, const ib_material *materials,
                                uint32_t nmat, const ib_texture *textures, uint32_t ntex, float *out) {
  if (!c || (ntri && !out)) return fail(IB_ERR_INVALID_ARGUMENT, "null argument");
  if ((nuv && !uvs) || (ntm && !tri_material) || (nmat && !materials) || (ntex && !textures))
    return fail(IB_ERR_INVALID_ARGUMENT, "null table with a non-zero size");
  if (ntri >= (1ull << 32) / 5) return fail(IB_ERR_INVALID_ARGUMENT, "too many triangles");
  if (ntri == 0) return IB_OK;
  std::vector<TexDev> tex(std::max<uint32_t>(ntex, 1));
  uint64_t bytes = 0;
  for (uint32_t i = 0; i < ntex; ++i) {
    const ib_texture &t = textures[i];
    tex[i] = TexDev{bytes, t.width, t.height, t.srgb ? 1 : 0, 0};
    if (t.width > 0 && t.height > 0) {
      if (!t.rgba) return fail(IB_ERR_INVALID_ARGUMENT, "texture " + std::to_string(i) + " has no texels");
      bytes += (uint64_t)t.width * (uint64_t)t.height * 4;
    }
  }
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  Ordered ord(c, s);
  float *d_uv = nullptr, *d_lut, *d_out;
  uint32_t *d_tm = nullptr;
  MatDev *d_mat = nullptr;
  TexDev *d_tex;
  uint8_t *d_px;
  if (nuv) IB_TRY(c->scratch("tm_uv", 2 * nuv, &d_uv));
  if (ntm) IB_TRY(c->scratch("tm_tri", ntm, &d_tm));
  if (nmat) IB_TRY(c->scratch("tm_mat", nmat, &d_mat));
  IB_TRY(c->scratch("tm_tex", tex.size(), &d_tex));
  IB_TRY(c->scratch("tm_px", std::max<uint64_t>(bytes, 4), &d_px));
  IB_TRY(c->scratch("tm_lut", 256, &d_lut));
  IB_TRY(c->scratch("tm_out", 5 * ntri, &d_out));
  if (nuv) IB_HIP(hipMemcpyAsync(d_uv, uvs, 2 * nuv * sizeof(float), hipMemcpyHostToDevice, s));
  if (ntm) IB_HIP(hipMemcpyAsync(d_tm, tri_material, ntm * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  if (nmat) IB_HIP(hipMemcpyAsync(d_mat, materials, nmat * sizeof(MatDev), hipMemcpyHostToDevice, s));
  IB_HIP(hipMemcpyAsync(d_tex, tex.data(), tex.size() * sizeof(TexDev), hipMemcpyHostToDevice, s));
  for (uint32_t i = 0; i < ntex; ++i)
    if (textures[i].width > 0 && textures[i].height > 0)
      IB_HIP(hipMemcpyAsync(d_px + tex[i].offset, textures[i].rgba,
                            (uint64_t)textures[i].width * textures[i].height * 4, hipMemcpyHostToDevice, s));
  IB_HIP(hipMemcpyAsync(d_lut, srgb_table(), 256 * sizeof(float), hipMemcpyHostToDevice, s));
  launch_tri_materials((uint32_t)ntri, d_uv, nuv, d_tm, ntm, d_mat, nmat, d_tex, ntex, d_px, d_lut, d_out, s);
  IB_HIP(hipMemcpyAsync(out, d_out, 5 * ntri * sizeof(float), hipMemcpyDeviceToHost, s));
  IB_HIP(hipStreamSynchronize(s));
  return IB_OK;
}

ib_status ib_splat_materials(ib_context *c, const float *xyz, uint64_t ntri, const float *tri_material,
                             const ib_grid *grid, float *base_color, float *metallic, float *roughness) {
  if (!c || !grid || !base_color || !metallic || !roughness) return fail(IB_ERR_INVALID_ARGUMENT, "null argument");
  if (ntri && (!xyz || !tri_material)) return fail(IB_ERR_INVALID_ARGUMENT, "null triangle buffer");
  if (!grid_ok(grid)) return fail(IB_ERR_INVALID_ARGUMENT, "invalid grid");
  const uint64_t ncell = (uint64_t)grid->nx * grid->ny * grid->nz;
  if (ncell >= (1ull << 32)) return fail(IB_ERR_UNSUPPORTED, "grid too large for 32-bit cell keys");
  if (ntri > (uint64_t)(INT_MAX / 9)) return fail(IB_ERR_INVALID_ARGUMENT, "too many triangles");
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  Ordered ord(c, s);
  constexpr int K = 5;
  float *d_out;
  IB_TRY(c->scratch("sp_out", (K + 0) * ncell, &d_out));
  // defaults of sdf_baker.cpp:77-80: colour 1 (3 interleaved floats per cell), metallic 0, roughness 1
  fill_u32(reinterpret_cast<uint32_t *>(d_out), 0x3f800000u, 3 * ncell, s);
  fill_u32(reinterpret_cast<uint32_t *>(d_out) + 3 * ncell, 0u, ncell, s);
  fill_u32(reinterpret_cast<uint32_t *>(d_out) + 4 * ncell, 0x3f800000u, ncell, s);
  if (ntri) {
    const int n = (int)ntri;
    float *d_xyz, *d_attr;
    uint32_t *counts, *offsets;
    IB_TRY(c->scratch("in_xyz", 9ull * n, &d_xyz));
    IB_TRY(c->scratch("sp_attr", (uint64_t)K * n, &d_attr));
    IB_TRY(c->scratch("sp_counts", (uint64_t)n + 1, &counts));
    IB_TRY(c->scratch("sp_offsets", (uint64_t)n + 1, &offsets));
    IB_HIP(hipMemcpyAsync(d_xyz, xyz, 9ull * n * sizeof(float), hipMemcpyHostToDevice, s));
    IB_HIP(hipMemcpyAsync(d_attr, tri_material, (uint64_t)K * n * sizeof(float), hipMemcpyHostToDevice, s));
    launch_splat_count(d_xyz, n, to_args(*grid), counts, s);
    fill_u32(counts + n, 0u, 1, s);
    size_t tb = 0;
    IB_HIP(exclusive_scan_u32(nullptr, tb, counts, offsets, (uint32_t)n + 1, s));
    char *temp;
    IB_TRY(c->scratch("sp_temp", tb, &temp));
    IB_HIP(exclusive_scan_u32(temp, tb, counts, offsets, (uint32_t)n + 1, s));
    uint32_t npairs = 0;
    IB_HIP(hipMemcpyAsync(&npairs, offsets + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    IB_HIP(hipStreamSynchronize(s));
    if (npairs) {
      uint32_t *cell, *cell_s;
      int32_t *tri, *tri_s;
      IB_TRY(c->scratch("sp_cell", npairs, &cell));
      IB_TRY(c->scratch("sp_cell_s", npairs, &cell_s));
      IB_TRY(c->scratch("sp_tri", npairs, &tri));
      IB_TRY(c->scratch("sp_tri_s", npairs, &tri_s));
      launch_splat_emit(d_xyz, n, to_args(*grid), offsets, cell, tri, s);
      int bits = 1;
      while (bits < 32 && (1ull << bits) < ncell) ++bits;
      tb = 0;
      IB_HIP(sort_pairs_u32(nullptr, tb, cell, cell_s, tri, tri_s, npairs, bits, s));
      IB_TRY(c->scratch("sp_temp", tb, &temp));
      IB_HIP(sort_pairs_u32(temp, tb, cell, cell_s, tri, tri_s, npairs, bits, s));
      launch_splat_reduce(cell_s, tri_s, npairs, d_attr, K, d_out, ncell, s);
    }
  }
  // the device buffer already has the host layout: rgb interleaved per cell (std::vector<glm::vec3>),
  // then the metallic and roughness planes
  const HostCopy cp[3] = {{base_color, d_out, 3 * ncell * sizeof(float)},
                          {metallic, d_out + 3 * ncell, ncell * sizeof(float)},
                          {roughness, d_out + 4 * ncell, ncell * sizeof(float)}};
  return download(c, cp, 3, s);
}

// ---- mesh ingestion gate (src/io/mesh_weld.cpp:26-48, src/io/mesh_validate.cpp:21-59) ---------------
// pitch = std::max(glm::length(bmax - bmin) * 1e-5f, 1e-6f), glm::length(v) = sqrt((x*x + y*y) + z*z)
// (mesh_weld.cpp:29-31); this file is compiled without FP contraction, so each product is rounded
static float weld_pitch(const float bmin[3], const float bmax[3]) {
  const float ex = bmax[0] - bmin[0], ey = bmax[1] - bmin[1], ez = bmax[2] - bmin[2];
  const float diag = std::sqrt((ex * ex + ey * ey) + ez * ez);
  return std::max(diag * 1e-5f, 1e-6f);
}

// Welds npos device positions into d_welded (device, npos entries) and, when d_unique is not null, the
// first position of each index (device, 3 * npos floats).  Waits for the result; *nunique = indices.
static ib_status weld_device(ib_context *c, const float *d_pos, uint64_t npos, const float bmin[3],
                             const float bmax[3], uint32_t *d_welded, float *d_unique, uint64_t *nunique,
                             hipStream_t s) {
  *nunique = 0;
  if (npos == 0) return IB_OK;
  const uint32_t n = (uint32_t)npos;
  const uint32_t cap = pow2_at_least(2ull * n);
  long long *keys;
  int32_t *slots;
  uint32_t *vslot, *first, *flag, *rank, *bad;
  IB_TRY(c->scratch("qw_keys", 3ull * n, &keys));
  IB_TRY(c->scratch("qw_slots", cap, &slots));
  IB_TRY(c->scratch("qw_vslot", n, &vslot));
  IB_TRY(c->scratch("qw_first", n, &first));
  IB_TRY(c->scratch("qw_flag", (uint64_t)n + 1, &flag));
  IB_TRY(c->scratch("qw_rank", (uint64_t)n + 1, &rank));
  IB_TRY(c->scratch("qw_bad", 2, &bad));
  fill_u32(reinterpret_cast<uint32_t *>(slots), 0xffffffffu, cap, s);
  fill_u32(bad, 0u, 1, s);
  launch_qweld(d_pos, n, weld_pitch(bmin, bmax), keys, slots, cap - 1, vslot, first, flag, bad, s);
  size_t tb = 0;
  IB_HIP(exclusive_scan_u32(nullptr, tb, flag, rank, n + 1, s));
  char *temp;
  IB_TRY(c->scratch("qw_temp", tb, &temp));
  IB_HIP(exclusive_scan_u32(temp, tb, flag, rank, n + 1, s));
  launch_qweld_out(d_pos, first, flag, rank, n, d_welded, d_unique, s);
  uint32_t h[2] = {0, 0};
  IB_HIP(hipMemcpyAsync(&h[0], bad, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  IB_HIP(hipMemcpyAsync(&h[1], rank + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  IB_HIP(hipStreamSynchronize(s));
  if (h[0])
    return fail(IB_ERR_INVALID_MESH, "a position divided by the weld pitch is not finite or exceeds 2^63 "
                                     "(std::lround is unspecified there)");
  *nunique = h[1];
  return IB_OK;
}

static ib_status weld_args_ok(uint64_t npos, const float bmin[3], const float bmax[3]) {
  if (!bmin || !bmax) return fail(IB_ERR_INVALID_ARGUMENT, "null bounds");
  if (npos >= (uint64_t)INT_MAX) return fail(IB_ERR_INVALID_ARGUMENT, "too many positions (indices are 32-bit)");
  return IB_OK;
}

ib_status ib_weld_positions_device(ib_context *c, const float *d_pos, uint64_t npos, const float bmin[3],
                                   const float bmax[3], uint32_t *d_welded, float *d_unique, uint64_t *nunique,
                                   void *stream) {
  if (!c || !nunique) return fail(IB_ERR_INVALID_ARGUMENT, "null argument");
  IB_TRY(weld_args_ok(npos, bmin, bmax));
  if (npos && (!d_pos || !d_welded)) return fail(IB_ERR_INVALID_ARGUMENT, "null position or index buffer");
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
  Ordered ord(c, s);
  reset_stats(c);
  c->record(EV_START, s);
  IB_TRY(weld_device(c, d_pos, npos, bmin, bmax, d_welded, d_unique, nunique, s));
  c->record(EV_INGEST, s);
  return IB_OK;
}

ib_status ib_weld_positions(ib_context *c, const float *pos, uint64_t npos, const float bmin[3], const float bmax[3],
                            uint32_t *welded, float *unique, uint64_t *nunique) {
  if (!c || !nunique) return fail(IB_ERR_INVALID_ARGUMENT, "null argument");
  IB_TRY(weld_args_ok(npos, bmin, bmax));
  *nunique = 0;
  if (npos == 0) return IB_OK;
  if (!pos || !welded) return fail(IB_ERR_INVALID_ARGUMENT, "null position or index buffer");
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  Ordered ord(c, s);
  reset_stats(c);
  float *d_pos, *d_unique = nullptr;
  uint32_t *d_welded;
  IB_TRY(c->scratch("in_xyz", 3 * npos, &d_pos));
  IB_TRY(c->scratch("qw_welded", npos, &d_welded));
  if (unique) IB_TRY(c->scratch("qw_unique", 3 * npos, &d_unique));
  IB_HIP(hipMemcpyAsync(d_pos, pos, 3 * npos * sizeof(float), hipMemcpyHostToDevice, s));
  c->record(EV_START, s);
  IB_TRY(weld_device(c, d_pos, npos, bmin, bmax, d_welded, d_unique, nunique, s));
  c->record(EV_INGEST, s);
  IB_HIP(hipMemcpyAsync(welded, d_welded, npos * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (unique) IB_HIP(hipMemcpyAsync(unique, d_unique, 3 * *nunique * sizeof(float), hipMemcpyDeviceToHost, s));
  IB_HIP(hipStreamSynchronize(s));
  return IB_OK;
}

// isWatertight's verdict and message (mesh_validate.cpp:22-58) from the edge counts
static void watertight_message(uint64_t ntri, uint64_t boundary, uint64_t nonmanifold, char *msg, uint64_t cap) {
  if (!msg || cap == 0) return;
  std::string e;
  if (ntri == 0) {
    e = "mesh has no triangles";
  } else if (boundary || nonmanifold) {
    e = "mesh is not watertight (" + std::to_string(boundary) + " boundary edge" + (boundary == 1 ? "" : "s");
    if (nonmanifold)
      e += ", " + std::to_string(nonmanifold) + " non-manifold edge" + (nonmanifold == 1 ? "" : "s");
    e += ") \u2014 fill holes and merge duplicate vertices in your DCC tool and re-export";
  }
  const size_t k = std::min<size_t>(e.size(), cap - 1);
  std::memcpy(msg, e.data(), k);
  msg[k] = '\0';
}

static ib_status watertight_device(ib_context *c, const float *d_xyz, uint64_t ntri, const float bmin[3],
                                   const float bmax[3], uint64_t *boundary, uint64_t *nonmanifold, hipStream_t s) {
  uint32_t *d_welded;
  IB_TRY(c->scratch("qw_welded", 3 * ntri, &d_welded));
  uint64_t nunique = 0;
  IB_TRY(weld_device(c, d_xyz, 3 * ntri, bmin, bmax, d_welded, nullptr, &nunique, s));
  const uint32_t cap = pow2_at_least(2ull * 3 * ntri);
  unsigned long long *ekeys, *out;
  uint32_t *ecount;
  IB_TRY(c->scratch("wt_keys", cap, &ekeys));
  IB_TRY(c->scratch("wt_count", cap, &ecount));
  IB_TRY(c->scratch("wt_out", 2, &out));
  fill_u32(reinterpret_cast<uint32_t *>(ekeys), 0xffffffffu, 2ull * cap, s);
  fill_u32(ecount, 0u, cap, s);
  fill_u32(reinterpret_cast<uint32_t *>(out), 0u, 4, s);
  launch_watertight(d_welded, (uint32_t)ntri, ekeys, ecount, cap - 1, out, s);
  c->record(EV_INGEST, s);
  unsigned long long h[2] = {0, 0};
  IB_HIP(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, s));
  IB_HIP(hipStreamSynchronize(s));
  *boundary = h[0];
  *nonmanifold = h[1];
  return IB_OK;
}

static ib_status watertight_entry(ib_context *c, const float *xyz, bool on_device, uint64_t ntri, const float bmin[3],
                                  const float bmax[3], int32_t *watertight, uint64_t *boundary, uint64_t *nonmanifold,
                                  char *msg, uint64_t msg_cap, void *stream) {
  if (!c || !watertight) return fail(IB_ERR_INVALID_ARGUMENT, "null argument");
  if (ntri > (uint64_t)(INT_MAX / 9)) return fail(IB_ERR_INVALID_ARGUMENT, "too many triangles");
  IB_TRY(weld_args_ok(3 * ntri, bmin, bmax));
  if (ntri && !xyz) return fail(IB_ERR_INVALID_ARGUMENT, "null triangle buffer");
  uint64_t b = 0, m = 0;
  if (ntri) {
    std::lock_guard<std::mutex> lk(c->mu);
    IB_HIP(hipSetDevice(c->device));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
    Ordered ord(c, s);
    reset_stats(c);
    c->stats.triangles = ntri;
    const float *d_xyz = xyz;
    if (!on_device) {
      float *d_in;
      IB_TRY(c->scratch("in_xyz", 9 * ntri, &d_in));
      IB_HIP(hipMemcpyAsync(d_in, xyz, 9 * ntri * sizeof(float), hipMemcpyHostToDevice, s));
      d_xyz = d_in;
    }
    c->record(EV_START, s);
    IB_TRY(watertight_device(c, d_xyz, ntri, bmin, bmax, &b, &m, s));
    c->stats.boundary_edges = (int64_t)b;
  }
  *watertight = (ntri > 0 && b == 0 && m == 0) ? 1 : 0;
  if (boundary) *boundary = b;
  if (nonmanifold) *nonmanifold = m;
  watertight_message(ntri, b, m, msg, msg_cap);
  return IB_OK;
}

ib_status ib_watertight(ib_context *c, const float *xyz, uint64_t ntri, const float bmin[3], const float bmax[3],
                        int32_t *watertight, uint64_t *boundary, uint64_t *nonmanifold, char *msg, uint64_t msg_cap) {
  return watertight_entry(c, xyz, false, ntri, bmin, bmax, watertight, boundary, nonmanifold, msg, msg_cap, nullptr);
}

ib_status ib_watertight_device(ib_context *c, const float *d_xyz, uint64_t ntri, const float bmin[3],
                               const float bmax[3], int32_t *watertight, uint64_t *boundary, uint64_t *nonmanifold,
                               char *msg, uint64_t msg_cap, void *stream) {
  return watertight_entry(c, d_xyz, true, ntri, bmin, bmax, watertight, boundary, nonmanifold, msg, msg_cap, stream);
}

ib_status ib_device_alloc(ib_context *c, uint64_t bytes, void **p) {
  if (!c || !p) return fail(IB_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  IB_TRY(c->alloc_raw(bytes, p));
  c->user[*p] = bytes;
  return IB_OK;
}

ib_status ib_device_free(ib_context *c, void *p) {
  if (!c) return fail(IB_ERR_INVALID_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  auto it = c->user.find(p);
  if (it == c->user.end()) return fail(IB_ERR_INVALID_ARGUMENT, "pointer was not allocated by ib_device_alloc");
  c->free_raw(p, it->second);
  c->user.erase(it);
  return IB_OK;
}

// copies on the context's stream after every earlier call's work; they return when the copy is done
static ib_status ordered_copy(ib_context *c, void *dst, const void *src, uint64_t bytes, hipMemcpyKind kind) {
  if (!c) return fail(IB_ERR_INVALID_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  Ordered ord(c, s);
  IB_HIP(hipMemcpyAsync(dst, src, bytes, kind, s));
  IB_HIP(hipStreamSynchronize(s));
  return IB_OK;
}

ib_status ib_memcpy_h2d(ib_context *c, void *dst, const void *src, uint64_t bytes) {
  return ordered_copy(c, dst, src, bytes, hipMemcpyHostToDevice);
}

ib_status ib_memcpy_d2h(ib_context *c, void *dst, const void *src, uint64_t bytes) {
  return ordered_copy(c, dst, src, bytes, hipMemcpyDeviceToHost);
}

static_assert(sizeof(ib_csg_op) == sizeof(CsgOp), "CSG instruction layout");

ib_status ib_analytic_framing(const float bmin[3], const float bmax[3], int32_t res, ib_grid *g) {
  if (!bmin || !bmax || !g) return fail(IB_ERR_INVALID_ARGUMENT, "null argument");
  // sdf_procedural.cpp:12-26
  float ext[3];
  for (int k = 0; k < 3; ++k) ext[k] = bmax[k] - bmin[k];
  const float longest = std::max({ext[0], ext[1], ext[2]});
  const float vs = longest / static_cast<float>(std::max(res, 4));
  if (!(vs > 0.0f) || !std::isfinite(vs)) return fail(IB_ERR_INVALID_ARGUMENT, "empty or non-finite bounds");
  for (int k = 0; k < 3; ++k) g->origin[k] = bmin[k] - 2.0f * vs;
  g->nx = static_cast<int>(std::ceil(ext[0] / vs)) + 4;
  g->ny = static_cast<int>(std::ceil(ext[1] / vs)) + 4;
  g->nz = static_cast<int>(std::ceil(ext[2] / vs)) + 4;
  g->h = vs;
  g->reserved = 0;
  return IB_OK;
}

// the analytic launch is a (x / 64, y / 4, z) grid of blocks
static bool analytic_grid_ok(const ib_grid *g) { return grid_ok(g) && g->ny <= 4 * 65535 && g->nz <= 65535; }

static ib_status check_program(const ib_csg_op *prog, uint32_t nops) {
  if (!prog || nops == 0) return fail(IB_ERR_INVALID_ARGUMENT, "empty CSG program");
  if (nops > 256) return fail(IB_ERR_UNSUPPORTED, "CSG program longer than 256 instructions");
  int depth = 0;
  for (uint32_t i = 0; i < nops; ++i) {
    const int op = prog[i].op;
    if (op >= IB_CSG_SPHERE && op <= IB_CSG_PLANE) {
      if (prog[i].branch < 0 || prog[i].branch > 255) return fail(IB_ERR_INVALID_ARGUMENT, "branch outside 0..255");
      if (++depth > kCsgStack) return fail(IB_ERR_UNSUPPORTED, "CSG stack deeper than 16");
    } else if (op >= IB_CSG_UNION && op <= IB_CSG_SELECT) {
      if (depth < 2) return fail(IB_ERR_INVALID_ARGUMENT, "CSG operator without two operands");
      --depth;
    } else {
      return fail(IB_ERR_INVALID_ARGUMENT, "unknown CSG opcode " + std::to_string(op));
    }
    for (float v : prog[i].a)
      if (!std::isfinite(v)) return fail(IB_ERR_INVALID_ARGUMENT, "non-finite CSG parameter");
  }
  if (depth != 1) return fail(IB_ERR_INVALID_ARGUMENT, "CSG program must leave exactly one value");
  return IB_OK;
}

// Programs of up to kCsgInline instructions go to the device inside the kernel arguments (the
// launch copies them), so the call is fully asynchronous; longer ones are uploaded to a scratch
// buffer and the call waits for that copy, since `prog` is pageable host memory.
static ib_status bake_analytic(ib_context *c, const ib_csg_op *prog, uint32_t nops, const ib_grid *grid, float *d_dist,
                               uint8_t *d_branch, hipStream_t s) {
  const CsgOp *host = reinterpret_cast<const CsgOp *>(prog);
  CsgOp *d_prog = nullptr;
  if (nops > (uint32_t)kCsgInline) {
    IB_TRY(c->scratch("csg_prog", nops, &d_prog));
    IB_HIP(hipMemcpyAsync(d_prog, host, nops * sizeof(CsgOp), hipMemcpyHostToDevice, s));
  }
  launch_bake_csg(host, d_prog, (int)nops, to_args(*grid), d_dist, d_branch, s);
  IB_HIP(hipGetLastError());
  if (d_prog) IB_HIP(hipStreamSynchronize(s));
  return IB_OK;
}

ib_status ib_bake_analytic_device(ib_context *c, const ib_csg_op *prog, uint32_t nops, const ib_grid *grid,
                                  float *d_dist, uint8_t *d_branch, void *stream) {
  if (!c) return fail(IB_ERR_INVALID_ARGUMENT, "null context");
  IB_TRY(check_program(prog, nops));
  if (!grid || !analytic_grid_ok(grid) || !d_dist) return fail(IB_ERR_INVALID_ARGUMENT, "invalid grid or output");
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
  Ordered ord(c, s);
  reset_stats(c);
  c->stats.voxels = (uint64_t)grid->nx * grid->ny * grid->nz;
  c->record(EV_START, s);
  IB_TRY(bake_analytic(c, prog, nops, grid, d_dist, d_branch, s));
  c->record(EV_SDF, s);
  return IB_OK;
}

ib_status ib_bake_analytic(ib_context *c, const ib_csg_op *prog, uint32_t nops, const ib_grid *grid, float *dist,
                           uint8_t *branch) {
  if (!c) return fail(IB_ERR_INVALID_ARGUMENT, "null context");
  IB_TRY(check_program(prog, nops));
  if (!grid || !analytic_grid_ok(grid) || !dist) return fail(IB_ERR_INVALID_ARGUMENT, "invalid grid or output");
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  Ordered ord(c, s);
  const uint64_t n = (uint64_t)grid->nx * grid->ny * grid->nz;
  float *d_dist;
  uint8_t *d_branch = nullptr;
  IB_TRY(c->scratch("an_dist", n, &d_dist));
  if (branch) IB_TRY(c->scratch("an_branch", n, &d_branch));
  IB_TRY(bake_analytic(c, prog, nops, grid, d_dist, d_branch, s));
  const HostCopy cp[2] = {{dist, d_dist, n * sizeof(float)}, {branch, d_branch, branch ? n : 0}};
  return download(c, cp, 2, s);
}

static ib_status grid_view(const ib_grid *grid, const float *d_distances, GridView &v) {
  if (!grid || !d_distances) return fail(IB_ERR_INVALID_ARGUMENT, "null grid or distances");
  if (!grid_ok(grid)) return fail(IB_ERR_INVALID_ARGUMENT, "invalid grid");
  if (grid->nx < 2 || grid->ny < 2 || grid->nz < 2)
    return fail(IB_ERR_INVALID_ARGUMENT, "trilinear sampling needs at least 2 cells per axis");
  v = GridView{grid->origin[0], grid->origin[1], grid->origin[2], grid->h, grid->nx, grid->ny, grid->nz, d_distances};
  return IB_OK;
}

ib_status ib_sample_grid_device(ib_context *c, const ib_grid *grid, const float *d_distances, const float *d_points,
                                uint64_t n, float *d_values, float *d_normals, int32_t *d_nearest, void *stream) {
  if (!c) return fail(IB_ERR_INVALID_ARGUMENT, "null context");
  GridView v;
  IB_TRY(grid_view(grid, d_distances, v));
  if (n && !d_points) return fail(IB_ERR_INVALID_ARGUMENT, "null points");
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
  Ordered ord(c, s);
  launch_sample_grid(v, d_points, n, d_values, d_normals, d_nearest, s);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

ib_status ib_trace_grid_device(ib_context *c, const ib_grid *grid, const float *d_distances, const float *d_origins,
                               const float *d_dirs, uint64_t n, float tmin, float tmax, float *d_t, void *stream) {
  if (!c) return fail(IB_ERR_INVALID_ARGUMENT, "null context");
  GridView v;
  IB_TRY(grid_view(grid, d_distances, v));
  if (n && (!d_origins || !d_dirs || !d_t)) return fail(IB_ERR_INVALID_ARGUMENT, "null ray buffer");
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
  Ordered ord(c, s);
  launch_trace_grid(v, d_origins, d_dirs, n, tmin, tmax, d_t, s);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

ib_status ib_memcpy_d2d(ib_context *c, void *dst, const void *src, uint64_t bytes) {
  return ordered_copy(c, dst, src, bytes, hipMemcpyDeviceToDevice);
}

ib_status ib_synchronize(ib_context *c) {
  if (!c) return fail(IB_ERR_INVALID_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  if (c->tail_valid) IB_HIP(hipEventSynchronize(c->tail));
  IB_HIP(hipStreamSynchronize(c->stream));
  return IB_OK;
}

ib_status ib_last_stats(ib_context *c, ib_stats *out) {
  if (!c || !out) return fail(IB_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  for (int e = 0; e < EV_COUNT; ++e)  // the events may sit on two streams
    if (c->ev_valid[e]) IB_HIP(hipEventSynchronize(c->ev[e]));
  ib_stats st = c->stats;
  st.ms_ingest = ev_ms(c, EV_START, EV_INGEST);
  st.ms_bvh = ev_ms(c, EV_INGEST, EV_BVH);
  st.ms_occupancy = c->ev_valid[EV_OCC_SIDE] ? ev_ms(c, EV_START, EV_OCC_SIDE) : ev_ms(c, EV_BVH, EV_OCC);
  st.ms_columns = ev_ms(c, EV_OCC, EV_COLS);
  st.ms_sdf = ev_ms(c, EV_COLS, EV_SDF);
  st.ms_query = ev_ms(c, EV_Q0, EV_Q1);
  st.ms_caps = ev_ms(c, EV_CAP0, EV_CAP1);
  st.ms_total = c->ev_valid[EV_SDF]   ? ev_ms(c, EV_START, EV_SDF)
                : c->ev_valid[EV_OCC] ? ev_ms(c, EV_START, EV_OCC)
                                      : ev_ms(c, EV_START, EV_INGEST);
  if (c->ev_valid[EV_OCC_SIDE]) st.ms_total = std::max(st.ms_total, (double)ev_ms(c, EV_START, EV_OCC_SIDE));
  // device-side counters of the last bake
  if (c->stats_in_sc && c->last_sc) {  // batched bakes accumulate their counters on the host
    const Scalars *sc = c->last_sc;
    Scalars h;
    IB_HIP(hipMemcpy(&h, sc, sizeof(Scalars), hipMemcpyDeviceToHost));
    st.column_overflows = h.col_overflow;
    st.rechecked = h.rechecked;
    st.ties = h.ties;
    st.box_tests = h.box_tests;
    st.tri_tests = h.tri_tests;
    st.stack_overflows = h.stack_overflow;
    st.obb_tests = h.obb_tests;
    st.on_cap = h.on_cap;
    st.rechecked_full = h.rechecked_full;
  }
  *out = st;
  return IB_OK;
}

// Batched bake (dense and narrow-band jobs): build every LBVH, one readback; narrow-band jobs get
// a coarse unsigned pass at brick centres and a brick compaction (second readback for the brick
// counts); then columns, ONE persistent query launch over all tiles, one resolve launch.
ib_status ib_bake_batch_device(ib_context *c, ib_job *jobs, uint32_t njobs, void *stream) {
  if (!c || (njobs && !jobs)) return fail(IB_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  IB_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
  reset_stats(c);
  if (njobs > 65535) return fail(IB_ERR_UNSUPPORTED, "at most 65535 jobs per batch");
  for (uint32_t j = 0; j < njobs; ++j) {
    ib_job &jb = jobs[j];
    const std::string tag = "job " + std::to_string(j) + ": ";
    if (jb.mesh) {  // a resident mesh brings its own soup
      jb.d_xyz = jb.mesh->d_xyz;
      jb.ntri = jb.mesh->ntri;
    }
    const bool has_sdf = jb.sdf.nx != 0, has_occ = jb.occ.nx != 0;
    // a narrow-band job may pass no pool with pool_bricks = 0 to learn the pool size (IB_ERR_CAPACITY)
    const bool sizing = jb.band_voxels > 0 && jb.pool_bricks == 0;
    if ((has_sdf && (!grid_ok(&jb.sdf) || (!jb.d_distances && !sizing))) ||
        (has_occ && (!grid_ok(&jb.occ) || !jb.d_occ_words)))
      return fail(IB_ERR_INVALID_ARGUMENT, tag + "invalid grid or missing output");
    if (jb.ntri > (uint64_t)(INT_MAX / 9) || (jb.ntri && !jb.d_xyz))
      return fail(IB_ERR_INVALID_ARGUMENT, tag + "invalid triangle buffer");
    if (has_sdf && (uint64_t)jb.sdf.nx * jb.sdf.ny * jb.sdf.nz >= (1ull << 32))
      return fail(IB_ERR_UNSUPPORTED, tag + "grid too large");
    if (jb.band_voxels < 0) return fail(IB_ERR_INVALID_ARGUMENT, tag + "negative band");
    if (jb.band_voxels > 0) {
      if (!has_sdf || (jb.sdf.nx | jb.sdf.ny | jb.sdf.nz) & 7)
        return fail(IB_ERR_INVALID_ARGUMENT, tag + "a narrow band needs an SDF grid with dims divisible by 8");
      if (!jb.d_brick_index) return fail(IB_ERR_INVALID_ARGUMENT, tag + "missing brick index");
    }
    jb.nbricks = 0;
  }
  if (c->batch_sc_cap < njobs) {
    if (c->host_batch_sc) (void)hipHostFree(c->host_batch_sc);
    c->host_batch_sc = nullptr;
    IB_HIP(hipHostMalloc((void **)&c->host_batch_sc, sizeof(Scalars) * std::max(njobs, 1u), hipHostMallocDefault));
    c->batch_sc_cap = njobs;
  }
  Arena arena{c};
  Ordered ord(c, s);  // after the arena: a failed batch waits for its kernels before the blocks are released
  Scalars *scs;
  IB_TRY(arena.get(std::max(njobs, 1u), &scs));
  std::vector<MeshDev> mesh(njobs);
  c->record(EV_START, s);
  // 1. ingest + LBVH of every mesh and occupancy, spread over the side streams (small meshes leave
  //    most of the GPU idle one at a time), joined into s; then one readback of all scalars
  IB_HIP(hipEventRecord(c->fork, s));
  for (int i = 0; i < ib_context::kSide; ++i) IB_HIP(hipStreamWaitEvent(c->side[i], c->fork, 0));
  // a batch of one mesh builds on s with the weld and the depth walk beside it, occupancy on side[0]
  const bool alone = njobs == 1;
  // narrow band: the bricks whose centre (a cell-centred grid of spacing 8h) lies within
  // band*h + 3.5*sqrt(3)*h of the surface (1-Lipschitz: every voxel of the band), flagged by the triangles
  std::vector<uint32_t> band_jobs;
  std::vector<uint32_t *> band_bricks(njobs, nullptr);
  std::vector<char> band_selected(njobs, 0);
  for (uint32_t j = 0; j < njobs; ++j)
    if (jobs[j].band_voxels > 0 && jobs[j].ntri) band_jobs.push_back(j);
  uint32_t *d_counts = nullptr;
  if (!band_jobs.empty()) IB_TRY(arena.get(band_jobs.size(), &d_counts));
  const auto select_band = [&](size_t k, bool slivers) -> ib_status {
    const uint32_t j = band_jobs[k];
    const ib_grid &g = jobs[j].sdf;
    ib_grid cg = g;
    cg.h = g.h * 8.0f;
    cg.nx = g.nx / 8;
    cg.ny = g.ny / 8;
    cg.nz = g.nz / 8;
    const uint32_t nbrick = (uint32_t)((g.nx / 8) * (g.ny / 8) * (g.nz / 8));
    uint32_t *flags, *slots;
    IB_TRY(arena.get(nbrick + 1, &flags));
    IB_TRY(arena.get(nbrick + 1, &slots));
    IB_TRY(arena.get(nbrick, &band_bricks[j]));
    fill_u32(flags, 0u, nbrick, s);
    launch_band_splat(mesh[j].tris, mesh[j].recs, mesh[j].n, to_args(cg), band_thresh(jobs[j]), slivers, flags, s);
    size_t tb = 0;
    IB_HIP(exclusive_scan_u32(nullptr, tb, flags, slots, nbrick, s));
    char *temp;
    IB_TRY(c->scratch("band_scan", tb, &temp));
    IB_HIP(exclusive_scan_u32(temp, tb, flags, slots, nbrick, s));
    launch_band_scatter(flags, slots, nbrick, jobs[j].d_brick_index, band_bricks[j], d_counts + k, s);
    band_selected[j] = 1;
    return IB_OK;
  };
  for (uint32_t j = 0; j < njobs; ++j) {
    const ib_job &jb = jobs[j];
    hipStream_t sj = alone ? s : c->side[j % ib_context::kSide];
    hipStream_t so = alone ? c->side[0] : sj;
    if (jb.ntri == 0 || jb.mesh) {  // nothing to build: the query counters still start at zero
      fill_u32(reinterpret_cast<uint32_t *>(scs + j), 0u, sizeof(Scalars) / 4, sj);
      if (jb.mesh) {
        mesh[j] = jb.mesh->m;
        if (jb.occ.nx) enqueue_occupancy(jb.d_xyz, (int)jb.ntri, jb.occ, jb.d_occ_words, so);
      }
      continue;
    }
    if (jb.occ.nx && alone) enqueue_occupancy(jb.d_xyz, (int)jb.ntri, jb.occ, jb.d_occ_words, so);
    IB_TRY(enqueue_build(c, jb.d_xyz, (int)jb.ntri, scs + j, mesh[j], &arena, sj, nullptr, alone));
    // a mesh built alone: its brick selection follows the tree on s before the host enqueues the caps (a
    // chain of small launches), whose kernels then run on side[1] beside it until the columns need them
    // (below).  The mesh's sliver count is not on the host yet: the splat takes the instance that tests
    // each triangle's sliver flag, which gives the same flags on a mesh without slivers.
    if (alone && !band_jobs.empty()) IB_TRY(select_band(0, true));
    if (alone) IB_TRY(early_caps(c, jb.d_xyz, (int)jb.ntri, scs + j, mesh[j], &arena, sj, false));
    if (jb.occ.nx && !alone) enqueue_occupancy(jb.d_xyz, (int)jb.ntri, jb.occ, jb.d_occ_words, sj);
  }
  const bool caps_side = alone && jobs[0].ntri && !jobs[0].mesh;  // early_caps recorded c->join[1]
  for (int i = 0; i < ib_context::kSide; ++i) {
    if (caps_side && i == 1) continue;
    IB_HIP(hipEventRecord(c->join[i], c->side[i]));
    IB_HIP(hipStreamWaitEvent(s, c->join[i], 0));
  }
  c->record(EV_BVH, s);
  IB_HIP(hipMemcpyAsync(c->host_batch_sc, scs, sizeof(Scalars) * njobs, hipMemcpyDeviceToHost, s));
  IB_HIP(hipStreamSynchronize(s));
  bool any_slivers = false, any_caps = false;
  for (uint32_t j = 0; j < njobs; ++j) {
    ib_job &jb = jobs[j];
    if (jb.ntri == 0) {
      if (jb.band_voxels > 0) {  // no surface: no bricks
        const ib_grid &g = jb.sdf;
        fill_u32(reinterpret_cast<uint32_t *>(jb.d_brick_index), 0xffffffffu,
                 (uint64_t)(g.nx / 8) * (g.ny / 8) * (g.nz / 8), s);
        if (jb.occ.nx) IB_TRY(fill_empty(c, nullptr, nullptr, &jb.occ, jb.d_occ_words, s));
      } else {
        IB_TRY(fill_empty(c, jb.sdf.nx ? &jb.sdf : nullptr, jb.d_distances, jb.occ.nx ? &jb.occ : nullptr,
                          jb.d_occ_words, s));
      }
      continue;
    }
    if (!jb.mesh) {
      const ib_status st = finish_build(c, jb.d_xyz, (int)jb.ntri, c->host_batch_sc[j], mesh[j], &arena, s);
      if (st != IB_OK) return fail(st, "job " + std::to_string(j) + ": " + g_err);
    }
    any_slivers |= mesh[j].slivers;
    any_caps |= mesh[j].ncap > 0;
    c->stats.triangles += jb.ntri;
    c->stats.boundary_edges += jb.mesh ? jb.mesh->boundary : c->host_batch_sc[j].boundary;
    c->stats.bvh_depth = std::max(c->stats.bvh_depth, mesh[j].depth);
  }
  c->record(EV_OCC, s);
  // 2. narrow band: the brick selection of every band job not selected beside its build, and the counts
  if (!band_jobs.empty()) {
    std::vector<uint32_t> host_counts(band_jobs.size());
    for (size_t k = 0; k < band_jobs.size(); ++k)
      if (!band_selected[band_jobs[k]]) IB_TRY(select_band(k, mesh[band_jobs[k]].slivers));
    IB_HIP(hipMemcpyAsync(host_counts.data(), d_counts, host_counts.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          s));
    IB_HIP(hipStreamSynchronize(s));
    bool short_pool = false;
    for (size_t k = 0; k < band_jobs.size(); ++k) {
      ib_job &jb = jobs[band_jobs[k]];
      jb.nbricks = host_counts[k];
      if (jb.nbricks > jb.pool_bricks) short_pool = true;
    }
    if (short_pool) return fail(IB_ERR_CAPACITY, "a brick pool is smaller than its narrow band (see nbricks)");
  }
  if (caps_side) IB_HIP(hipStreamWaitEvent(s, c->join[1], 0));  // the caps, for the columns and the query
  // 3. columns + per-job query blocks, tile prefix
  std::vector<SdfArgs> args;
  std::vector<uint32_t> owner, prefix(1, 0), col_prefix(1, 0);
  for (uint32_t j = 0; j < njobs; ++j) {
    const ib_job &jb = jobs[j];
    if (jb.ntri == 0 || jb.sdf.nx == 0) continue;
    const bool band = jb.band_voxels > 0;
    const uint64_t nvox = band ? jb.nbricks * 512ull : (uint64_t)jb.sdf.nx * jb.sdf.ny * jb.sdf.nz;
    if (band && jb.nbricks == 0) continue;
    SdfArgs sa;
    const uint32_t amb_cap = (uint32_t)std::min<uint64_t>(nvox, nvox / 16 + 4096);
    IB_TRY(prepare_query(c, mesh[j], jb.sdf, jb.d_distances, scs + j, &arena, amb_cap, sa, s, false));
    col_prefix.push_back(col_prefix.back() + column_blocks(sa.col));
    uint32_t tiles;
    if (band) {
      sa.band_index = jb.d_brick_index;
      sa.band_bricks = band_bricks[j];
      tiles = (uint32_t)(2 * jb.nbricks);
    } else {
      uint32_t d[3];
      tile_dims(to_args(jb.sdf), c->brick_shape, d);
      tiles = d[0] * d[1] * d[2];
    }
    args.push_back(sa);
    owner.push_back(j);
    prefix.push_back(prefix.back() + tiles);
    c->stats.voxels += nvox;
  }
  if (args.empty()) c->record(EV_COLS, s);
  if (!args.empty()) {
    SdfArgs *d_args;
    uint32_t *d_prefix, *d_head;
    IB_TRY(arena.get(args.size(), &d_args));
    IB_TRY(arena.get(prefix.size(), &d_prefix));
    IB_TRY(arena.get(1, &d_head));
    IB_HIP(hipMemcpyAsync(d_args, args.data(), args.size() * sizeof(SdfArgs), hipMemcpyHostToDevice, s));
    IB_HIP(hipMemcpyAsync(d_prefix, prefix.data(), prefix.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    // every job's column crossings in one launch
    uint32_t *d_col_prefix;
    IB_TRY(arena.get(col_prefix.size(), &d_col_prefix));
    IB_HIP(hipMemcpyAsync(d_col_prefix, col_prefix.data(), col_prefix.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice, s));
    launch_columns_batch(d_args, d_col_prefix, (uint32_t)args.size(), col_prefix.back(), s);
    c->record(EV_COLS, s);
    fill_u32(d_head, 0u, 1, s);
    BatchArgs t{d_args, d_prefix, (uint32_t)args.size(), d_head, c->brick_shape, nullptr};
    const uint32_t blocks = std::min<uint32_t>(prefix.back(), (uint32_t)(c->cus * c->blocks_per_cu));
    launch_sdf_batch(t, std::max(blocks, 1u), c->count_work, any_slivers, any_caps, s);
    launch_resolve_batch(d_args, (uint32_t)args.size(), any_caps, !c->shared_chip, s);
  }
  c->record(EV_SDF, s);
  IB_HIP(hipMemcpyAsync(c->host_batch_sc, scs, sizeof(Scalars) * njobs, hipMemcpyDeviceToHost, s));
  IB_HIP(hipStreamSynchronize(s));
  // 4. a job whose ambiguous-voxel list overflowed is queried again on its own with a full list
  for (size_t k = 0; k < args.size(); ++k) {
    const uint32_t j = owner[k];
    const Scalars &h = c->host_batch_sc[j];
    c->stats.rechecked += h.rechecked;
    c->stats.ties += h.ties;
    c->stats.column_overflows += h.col_overflow;
    c->stats.box_tests += h.box_tests;
    c->stats.tri_tests += h.tri_tests;
    c->stats.obb_tests += h.obb_tests;
    c->stats.on_cap += h.on_cap;
    c->stats.rechecked_full += h.rechecked_full;
    c->stats.stack_overflows += h.stack_overflow;
    if (h.stack_overflow)
      return fail(IB_ERR_INTERNAL, "job " + std::to_string(j) + ": a traversal stack overflowed");
    if (h.ambiguous > args[k].amb_cap) {
      SdfArgs sa = args[k];
      const ib_job &jb = jobs[j];
      const uint64_t nvox = jb.band_voxels > 0 ? jb.nbricks * 512ull : (uint64_t)jb.sdf.nx * jb.sdf.ny * jb.sdf.nz;
      fill_u32(reinterpret_cast<uint32_t *>(scs + j), 0u, sizeof(Scalars) / 4, s);
      sa.amb_cap = (uint32_t)nvox;
      IB_TRY(arena.get(nvox, &sa.amb_list));
      SdfArgs *d_one;
      uint32_t *d_pf, *d_hd;
      IB_TRY(arena.get(1, &d_one));
      IB_TRY(arena.get(2, &d_pf));
      IB_TRY(arena.get(1, &d_hd));
      const uint32_t pf[2] = {0, prefix[k + 1] - prefix[k]};
      IB_HIP(hipMemcpyAsync(d_one, &sa, sizeof(SdfArgs), hipMemcpyHostToDevice, s));
      IB_HIP(hipMemcpyAsync(d_pf, pf, sizeof(pf), hipMemcpyHostToDevice, s));
      fill_u32(d_hd, 0u, 1, s);
      BatchArgs t{d_one, d_pf, 1, d_hd, jb.band_voxels > 0 ? 0 : c->brick_shape};
      launch_sdf_batch(t, std::max(1u, std::min<uint32_t>(pf[1], (uint32_t)(c->cus * c->blocks_per_cu))), false,
                       any_slivers, any_caps, s);
      launch_resolve_batch(d_one, 1, any_caps, !c->shared_chip, s);
      IB_HIP(hipStreamSynchronize(s));
    }
  }
  return IB_OK;
}

ib_status ib_bake_band_device(ib_context *c, const float *d_xyz, uint64_t ntri, const ib_grid *grid,
                              int32_t band_voxels, int32_t *d_brick_index, float *d_pool, uint64_t pool_bricks,
                              uint64_t *nbricks, void *stream) {
  if (!grid || band_voxels <= 0 || !nbricks) return fail(IB_ERR_INVALID_ARGUMENT, "invalid narrow-band request");
  ib_job jb{};
  jb.d_xyz = d_xyz;
  jb.ntri = ntri;
  jb.sdf = *grid;
  jb.d_distances = d_pool;
  jb.band_voxels = band_voxels;
  jb.d_brick_index = d_brick_index;
  jb.pool_bricks = pool_bricks;
  const ib_status st = ib_bake_batch_device(c, &jb, 1, stream);
  *nbricks = jb.nbricks;
  return st;
}

ib_status ib_autotune_batch(ib_context *c, const ib_job *jobs, uint32_t njobs, uint32_t sample_stride,
                            int32_t *out_shape, int32_t *out_blocks, double *out_ms) {
  // the band pass uses fixed 4x4x4 sub-bricks; the shape and block count tune the dense tiles
  if (!c || (njobs && !jobs)) return fail(IB_ERR_INVALID_ARGUMENT, "null argument");
  if (sample_stride == 0) sample_stride = 1;
  // stratified sample: every sample_stride-th job in order of its voxel count, so that the sample
  // holds large and small grids in the batch's proportions
  std::vector<uint32_t> order(njobs);
  for (uint32_t j = 0; j < njobs; ++j) order[j] = j;
  const auto cells = [&](uint32_t j) {
    const ib_grid &g = jobs[j].sdf;
    return jobs[j].band_voxels > 0 ? 0ull : (uint64_t)g.nx * g.ny * g.nz;
  };
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cells(a) > cells(b); });
  std::vector<ib_job> sample;
  for (uint32_t k = sample_stride / 2; k < njobs; k += sample_stride) sample.push_back(jobs[order[k]]);
  if (sample.empty() && njobs) sample.push_back(jobs[order[0]]);
  static const int kBlocks[3] = {4, 8, 16};
  int best_shape = c->brick_shape, best_blocks = c->blocks_per_cu;
  double best_ms = 1e300;
  for (int shape = 0; shape < 5; ++shape)
    for (int b : kBlocks) {
      c->brick_shape = shape;
      c->blocks_per_cu = b;
      IB_TRY(ib_bake_batch_device(c, sample.data(), (uint32_t)sample.size(), nullptr));  // warm-up
      double ms = 0.0;
      {
        std::lock_guard<std::mutex> lk(c->mu);
        ms = ev_ms(c, EV_COLS, EV_SDF);  // the persistent query launch of the warm-up run
      }
      IB_TRY(ib_bake_batch_device(c, sample.data(), (uint32_t)sample.size(), nullptr));
      {
        std::lock_guard<std::mutex> lk(c->mu);
        ms = std::min(ms, (double)ev_ms(c, EV_COLS, EV_SDF));
      }
      if (ms < best_ms) {
        best_ms = ms;
        best_shape = shape;
        best_blocks = b;
      }
    }
  c->brick_shape = best_shape;
  c->blocks_per_cu = best_blocks;
  if (out_shape) *out_shape = best_shape;
  if (out_blocks) *out_blocks = best_blocks;
  if (out_ms) *out_ms = best_ms;
  return IB_OK;
}

ib_status ib_set_option(ib_context *c, int32_t option, int64_t value) {
  if (!c) return fail(IB_ERR_INVALID_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  switch (option) {
    case IB_OPT_COUNT_WORK:
      c->count_work = value != 0;
      return IB_OK;
    case IB_OPT_QUERY_VARIANT:
      if (value < 0 || value > 1) return fail(IB_ERR_INVALID_ARGUMENT, "query variant must be 0 or 1");
      c->query_variant = (int)value;
      return IB_OK;
    case IB_OPT_BRICK_SHAPE:
      if (value < 0 || value > 4) return fail(IB_ERR_INVALID_ARGUMENT, "brick shape must be 0..4");
      c->brick_shape = (int)value;
      return IB_OK;
    case IB_OPT_BLOCKS_PER_CU:
      if (value < 1 || value > 64) return fail(IB_ERR_INVALID_ARGUMENT, "blocks per CU must be 1..64");
      c->blocks_per_cu = (int)value;
      return IB_OK;
    case IB_OPT_LEAF_BLOCK:
      if (value < 1 || value > kMaxLeafBlock)
        return fail(IB_ERR_INVALID_ARGUMENT, "leaf block must be 1.." + std::to_string(kMaxLeafBlock));
      c->leaf_block = (int)value;
      return IB_OK;
    case IB_OPT_RESOLVE_FULL:
      c->resolve_full = value != 0;
      return IB_OK;
    case IB_OPT_SHARED_CHIP:
      c->shared_chip = value != 0;
      return IB_OK;
    case IB_OPT_TILE_ORDER:
      if (value < 0 || value > 2) return fail(IB_ERR_INVALID_ARGUMENT, "tile order must be 0..2");
      c->tile_order = (int)value;
      return IB_OK;
    default:
      return fail(IB_ERR_INVALID_ARGUMENT, "unknown option " + std::to_string(option));
  }
}

}  // extern "C"
