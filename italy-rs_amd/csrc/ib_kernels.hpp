// Kernel argument blocks and host launchers shared by ib_kernels.hip and ib_api.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "ib_device.hpp"

namespace ib {

// |w| within this of 0.5 after the FP64 recheck -> ray parity decides
constexpr double kTieEps = 1e-6;

constexpr uint32_t kFlagNonFinite = 1u;
// parent-chain walks stop here; a deeper chain means a malformed tree and is reported as depth
constexpr int kMaxDepth = 4096;
// Depth bound of the Karras tree: keys are (30-bit Morton code << 32 | sorted position), an internal
// node's split is at the first bit where its range's keys differ, and that prefix length grows strictly
// from a node to its internal children.  Prefix lengths lie in [2, 31] inside the code and in
// [64 - b, 63] inside a position of b = ceil(log2 n) bits, so a root-to-leaf path holds at most 30 + b
// internal nodes: 58 for the 2^28 triangles the entry points accept.  The binary traversals need one
// stack entry per level (kStack = 64); the 4-wide brick walk descends two levels per visit and pushes at
// most 3 entries per visit, at most 3 * 29 = 87 < kStack4.  A deeper tree is a construction bug.
constexpr int kMaxTreeDepth = 58;

// device-side scalars of one bake (reset before every bake)
struct Scalars {
  uint32_t cmin[3], cmax[3];  // centroid bounds, ordered-int encoded
  uint32_t flags;
  uint32_t depth;
  uint32_t boundary;      // unmatched directed edges (with multiplicity)
  uint32_t col_overflow;  // columns with more than kColCap crossings
  uint32_t rechecked;     // voxels re-evaluated in FP64
  uint32_t ties;          // voxels decided by ray parity
  uint32_t singular;      // entries claimed in the chunked FP64 list (voxels on a cap, undecided rechecks)
  uint32_t slivers;       // triangles taking the FP64 distance path
  uint32_t ambiguous;     // voxels handed to the resolve kernel (column overflow or |w| near 0.5)
  uint32_t stack_overflow;  // traversal pushes that found the stack full (0 by the depth bound; checked by the host)
  uint32_t rechecked_full;  // FP64 rechecks the cap tree's bound left undecided: every fan summed
  uint32_t on_cap;          // FP64 rechecks lying on a hole cap (decided by the mesh's own winding number)
  unsigned long long box_tests;  // counting mode only
  unsigned long long tri_tests;
  unsigned long long obb_tests;
};

struct GridArgs {
  float ox, oy, oz, h;
  int nx, ny, nz;
};

struct ColumnArgs {
  const Node *nodes;
  int nnodes;         // internal nodes (bounds every traversal loop)
  int ntri;           // triangles (tris / recs entries)
  const float4 *tris;
  const TriRec *recs;  // distance records, same order as tris
  const Node4 *nodes4; // 4-wide view of the nodes (brick query)
  const Obb *obbs;     // oriented box of every leaf block of 2+ triangles, at the block's first triangle
  const Obb *node_obbs;  // oriented box of every internal node of at most kNodeObbMax triangles, at its index
  const float *caps;  // 9 floats per cap triangle, in cap-hierarchy leaf order
  int ncap;
  const CapNode *cap_nodes;
  const int4 *cap_cut;  // the cap tree's cut (kCapCut slots: code, first fan, fan count) for the FP64 recheck
  int ncapnode;
  GridArgs g;
  float2 *cross;      // kColCap per column: (z, +-1)
  int32_t *count;     // per column
  Scalars *sc;
};

struct SdfArgs {
  ColumnArgs col;
  float *out;
  const int32_t *band_index;    // narrow band: slot of every 8^3 brick of the grid (-1: none); null: dense out
  const uint32_t *band_bricks;  // narrow band: brick linear index of every slot (2 block-tiles per brick)
  uint32_t *amb_list;  // voxels whose sign the resolve kernel decides
  uint32_t amb_cap;    // capacity of amb_list
  uint32_t sing_cap;     // voxels whose FP64 sum the chunked launches make (0: the resolving wave sums)
  uint32_t sing_chunks;  // chunks of kSingChunk triangles or cap fans (the larger count)
  uint32_t *sing;        // those voxels (sing_cap)
  uint32_t *sing_kind;   // 0: on a hole cap, the mesh's own winding number; 1: the column's integer + every cap fan
  double *sing_part;     // their chunk sums (sing_cap x sing_chunks)
  uint32_t resolve_full; // nonzero: the FP64 recheck sums every cap fan instead of walking the cap tree first
};
constexpr int kSingChunk = 2048;

void fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s);
struct FillRange {
  uint32_t *p;
  uint32_t v;
  uint64_t n;
};
constexpr int kMaxFill = 8;
struct FillList {
  FillRange r[kMaxFill];
  int count = 0;
  void add(uint32_t *p, uint32_t v, uint64_t n) { r[count++] = FillRange{p, v, n}; }
};
// the fills of `l` in one launch (ranges must not overlap)
void fill_multi(const FillList &l, hipStream_t s);
void launch_prep(const float *xyz, int n, Scalars *sc, hipStream_t s);
void launch_morton(const float *xyz, int n, const Scalars *sc, uint32_t *codes, int32_t *idx, hipStream_t s);
hipError_t sort_pairs(void *temp, size_t &temp_bytes, const uint32_t *kin, uint32_t *kout, const int32_t *vin,
                      int32_t *vout, int n, hipStream_t s);
void launch_karras(const uint32_t *codes, int n, Node *nodes, int32_t *leaf_parent, int2 *ranges, hipStream_t s);
void launch_gather(const float *xyz, const int32_t *order, int n, float4 *tris, TriRec *recs, Scalars *sc,
                   hipStream_t s);
// refit, depth, collapse of subtrees of at most `block` leaves into leaf-block codes, 4-wide view.
// Scratch: span n + 1 words (entry n zeroed by the caller), leafbox 6n floats, blockbox 6 per 256 leaves.
void launch_refit(const float4 *tris, int n, Node *nodes, const int32_t *leaf_parent, uint32_t *span,
                  float *leafbox, float *blockbox, float *root_box, Scalars *sc, const int2 *ranges, int block, Node4 *nodes4,
                  Obb *obbs, Obb *node_obbs, bool depth, hipStream_t s);
// the tree depth into sc->depth (parent chains of the Karras tree; launch_refit runs it when `depth`)
void launch_depth(const Node *nodes, const int32_t *leaf_parent, int n, Scalars *sc, hipStream_t s);
void launch_weld(const float *xyz, int ntri, int32_t *slots, int32_t *minid, uint32_t vmask, int32_t *cslot,
                 unsigned long long *ekeys, int32_t *enet, uint32_t emask, Scalars *sc, int2 *bnd, uint32_t bnd_cap,
                 hipStream_t s);
// positions of both endpoints (corner indices into the soup) of each boundary edge: 6 floats
// hole caps of the ne unmatched edges, all on the device: keys (from, to) sorted; keys (to, from) sorted with
// their positions; loop labels by pointer jumping; per-loop fixed-point vertex sums; one fan triangle per
// edge with its centroid's Morton code; the fans sorted by code; their tree
void launch_cap_keys(const int2 *bnd, uint32_t ne, unsigned long long *keys, hipStream_t s);
hipError_t sort_keys_u64(void *temp, size_t &temp_bytes, const unsigned long long *kin, unsigned long long *kout,
                         uint32_t n, int end_bit, hipStream_t s);
void launch_cap_in_keys(const unsigned long long *ek, uint32_t ne, unsigned long long *ik, uint32_t *iv,
                        hipStream_t s);
void launch_cap_succ(const unsigned long long *ek, const unsigned long long *ik, const uint32_t *iv, uint32_t ne,
                     uint32_t *lab, uint32_t *jmp, hipStream_t s);
void launch_cap_jump(const uint32_t *lab0, const uint32_t *jmp0, uint32_t ne, uint32_t *lab1, uint32_t *jmp1,
                     hipStream_t s);
void launch_iota(uint32_t *v, uint32_t n, hipStream_t s);
void launch_cap_bounds(const unsigned long long *ek, uint32_t ne, const float *xyz, uint32_t *bounds, hipStream_t s);
void launch_cap_sums(const uint32_t *slab, const uint32_t *sedge, uint32_t ne, const unsigned long long *ek,
                     const float *xyz, const uint32_t *bounds, unsigned long long *acc, hipStream_t s);
void launch_cap_tris(const unsigned long long *ek, const uint32_t *lab, uint32_t ne, const float *xyz,
                     const uint32_t *bounds, const unsigned long long *acc, float *tris, uint32_t *codes,
                     int32_t *idx, hipStream_t s);
void launch_cap_gather(const float *tmp, const int32_t *order, uint32_t ne, float *caps, hipStream_t s);
void launch_cap_tree(const float *caps, const uint32_t *codes, int n, int2 *child, int32_t *parent, int2 *ranges,
                     CapNode *nodes, int4 *cut, hipStream_t s);
void launch_occupancy(const float *xyz, int n, const GridArgs &g, uint32_t *words, hipStream_t s);
// 16 x 16 column tiles of a grid, and the columns of a whole batch in one launch (prefix: first
// column tile of each job, njobs + 1 entries on the device)
uint32_t column_blocks(const ColumnArgs &a);  // blocks of the column rasteriser for one job
void launch_columns_batch(const SdfArgs *jobs, const uint32_t *prefix, uint32_t njobs, uint32_t blocks, hipStream_t s);
void launch_columns(const ColumnArgs &a, hipStream_t s);
// a batch of query jobs drained by one persistent launch
struct BatchArgs {
  const SdfArgs *jobs;          // device array
  const uint32_t *tile_prefix;  // njobs + 1 (device)
  uint32_t njobs;
  uint32_t *head;               // work-queue counter (zeroed before the launch)
  int shape;                    // brick shape index
  const uint32_t *order;        // queue position -> global tile (null: identity, the batch order)
};

// block-tiles of a grid for a brick shape (x, y, z)
void tile_dims(const GridArgs &g, int shape, uint32_t d[3]);
// variant 0: per-lane traversal, 1: wave-coherent brick traversal; slivers: the mesh has triangles
// on the FP64 distance path; shape: brick shape of the wave (0: 4x4x4, 1: 8x4x2, 2: 8x8x1, 3: 16x4x1)
void launch_sdf(const SdfArgs &a, bool count, int variant, bool slivers, int shape, const uint32_t *order,
                hipStream_t s);
// launch-order keys (ascending = costliest first) and tile ids of the per-brick query's block-tiles
void launch_tile_keys(const ColumnArgs &a, int shape, int mode, uint32_t *keys, int32_t *ids, hipStream_t s);
void launch_sdf_batch(const BatchArgs &t, uint32_t blocks, bool count, bool slivers, bool caps, hipStream_t s);
// block_single: a single job's voxels take a 256-thread block each (else one wave each)
void launch_resolve_batch(const SdfArgs *jobs, uint32_t njobs, bool singular, bool block_single, hipStream_t s);
// narrow band: bricks whose centre distance (from the coarse pass) is <= thresh
// flags[b] = 1 for every brick centre of the coarse grid cg within thresh of some triangle (flags cleared first)
void launch_band_splat(const float4 *tris, const TriRec *recs, int n, const GridArgs &cg, float thresh, bool slivers,
                       uint32_t *flags, hipStream_t s);
void launch_band_scatter(const uint32_t *flags, const uint32_t *slots, uint32_t nbrick, int32_t *index,
                         uint32_t *bricks, uint32_t *count, hipStream_t s);
// FP64 winding recheck, parity tie-break and column-overflow recount for the listed voxels
void launch_resolve(const SdfArgs &a, hipStream_t s);

// sparse voxelizeMesh over the (cell, triangle) pairs sorted by cell: run heads (npairs + 1 flags), per run the
// first-touch key, its index and (rgb != null) the colour mean; then the runs sorted by key unpacked
void launch_vox_heads(const uint32_t *cell, uint32_t npairs, uint32_t *head, hipStream_t s);
void launch_vox_runs(const uint32_t *cell, const int32_t *tri, uint32_t npairs, const uint32_t *run_of,
                     const float *tri_rgb, unsigned long long *keys, uint32_t *vals, float *rgb, hipStream_t s);
hipError_t sort_pairs_u64(void *temp, size_t &temp_bytes, const unsigned long long *kin, unsigned long long *kout,
                          const uint32_t *vin, uint32_t *vout, uint32_t n, int end_bit, hipStream_t s);
void launch_vox_out(const unsigned long long *keys, const uint32_t *perm, uint32_t n, int nx, int ny, const float *rgb,
                    int32_t *cells, int32_t *first, float *colors, hipStream_t s);

// per-cell attribute splat
void launch_splat_count(const float *xyz, int n, const GridArgs &g, uint32_t *counts, hipStream_t s);
hipError_t exclusive_scan_u32(void *temp, size_t &temp_bytes, const uint32_t *in, uint32_t *out, uint32_t n,
                              hipStream_t s);
void launch_splat_emit(const float *xyz, int n, const GridArgs &g, const uint32_t *offsets, uint32_t *cell,
                       int32_t *tri, hipStream_t s);
// stable LSD radix sort of (key, value) pairs by the keys' low end_bit bits (ib_sort.hip); temp as rocPRIM's:
// a call with temp == nullptr returns the bytes needed in temp_bytes
hipError_t radix_sort_u32(void *temp, size_t &temp_bytes, const uint32_t *kin, uint32_t *kout, const int32_t *vin,
                          int32_t *vout, uint32_t n, int end_bit, hipStream_t s);
hipError_t sort_pairs_u32(void *temp, size_t &temp_bytes, const uint32_t *kin, uint32_t *kout, const int32_t *vin,
                          int32_t *vout, uint32_t n, int end_bit, hipStream_t s);
// per-cell means of the k attributes (k <= 8) of the (cell, triangle) pairs sorted by cell: attributes 0..2
// interleaved per cell at out[3 * cell + j], attribute j >= 3 at out[j * ncell + cell]
void launch_splat_reduce(const uint32_t *cell, const int32_t *tri, uint32_t npairs, const float *attr, int k,
                         float *out, uint64_t ncell, hipStream_t s);

// ---- per-triangle materials ----
struct MatDev {  // the layout of ib_material
  float base[3];
  float metallic, roughness;
  int32_t tex;
};
struct TexDev {
  uint64_t offset;  // first byte in the packed texel buffer
  int32_t w, h, srgb, pad;
};
// out: 5 floats per triangle; lut: srgbToLinear of byte / 255 for the 256 byte values
void launch_tri_materials(uint32_t ntri, const float *uvs, uint64_t nuv, const uint32_t *tri_mat, uint64_t ntm,
                          const MatDev *mats, uint32_t nmat, const TexDev *texs, uint32_t ntex, const uint8_t *pixels,
                          const float *lut, float *out, hipStream_t s);

// ---- grid consumer (ib_consumer.hip) ----
struct GridView {
  float ox, oy, oz, h;
  int nx, ny, nz;
  const float *d;  // distances, (z*ny + y)*nx + x
};
void launch_sample_grid(const GridView &g, const float *pts, uint64_t n, float *values, float *normals,
                        int32_t *nearest, hipStream_t s);
void launch_trace_grid(const GridView &g, const float *org, const float *dir, uint64_t n, float tmin, float tmax,
                       float *t_hit, hipStream_t s);

// ---- mesh ingestion gate (ib_ingest.hip) ----
// weldPositions: lattice keys, first-occurrence slots (slots: mask + 1 entries filled with -1), first
// position of every index and the first-occurrence flags (n + 1 entries; flag[n] = 0)
void launch_qweld(const float *pos, uint32_t n, float pitch, long long *keys, int32_t *slots, uint32_t mask,
                  uint32_t *vslot, uint32_t *first, uint32_t *flag, uint32_t *bad, hipStream_t s);
// welded indices from the exclusive scan `rank` of the flags; unique (may be null) gets the first positions
void launch_qweld_out(const float *pos, const uint32_t *first, const uint32_t *flag, const uint32_t *rank, uint32_t n,
                      uint32_t *welded, float *unique, hipStream_t s);
// isWatertight's edge counts: ekeys (mask + 1 entries, all bits set), ecount (zeroed); out[0] += edges
// used once, out[1] += edges used more than twice
void launch_watertight(const uint32_t *welded, uint32_t ntri, unsigned long long *ekeys, uint32_t *ecount,
                       uint32_t mask, unsigned long long *out, hipStream_t s);

// ---- analytic bake (ib_csg.hip) ----
// one instruction of a postfix CSG program (same layout as ib_csg_op of the C ABI)
struct CsgOp {
  int32_t op, branch;
  float a[8];
};
constexpr int kCsgSphere = 1, kCsgBox = 2, kCsgCylinder = 3, kCsgTorus = 4, kCsgPlane = 5;
constexpr int kCsgUnion = 10, kCsgSubtract = 11, kCsgIntersect = 12, kCsgSmoothUnion = 13,
              kCsgSmoothIntersect = 14, kCsgSelect = 15;
constexpr int kCsgStack = 16;
constexpr int kCsgInline = 64;  // programs up to this long travel in the kernel arguments
struct CsgInline {
  CsgOp ops[kCsgInline];
};
// `host_prog` is read at launch when nops <= kCsgInline (copied into the kernel arguments);
// longer programs are read from `dev_prog`, which must then hold a device copy.
void launch_bake_csg(const CsgOp *host_prog, const CsgOp *dev_prog, int nops, const GridArgs &g, float *dist,
                     uint8_t *branch, hipStream_t s);

}  // namespace ib
