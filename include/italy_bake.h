/*
 * italy_bake.h -- C ABI of the MI355X (gfx950) mesh -> SDF / occupancy bake.
 *
 * Drop-in boundary for the reference's src/convert bake path.  Plain pointers and sizes only,
 * no exceptions across the boundary: every entry point returns an ib_status and leaves a
 * message for ib_last_error().  The C++ wrappers in include/italy_bake.hpp rethrow non-zero
 * statuses as std::runtime_error, which is the reference's error convention
 * (src/render/optix_check.h:13-48).
 *
 * Conventions shared with the reference:
 *   - input is a triangle soup: 9 floats per triangle, 3 vertices in winding order
 *     (src/io/mesh_asset.h:34,46), with caller-supplied bounds (mesh_asset.h:43-44);
 *   - grids are cell-centred, value of cell (x,y,z) sampled at origin + (i + 0.5) * h and stored
 *     at index (z*ny + y)*nx + x (src/convert/sdf_grid.h:35, src/render/kernels/sdf_bake.cu:34-37,71);
 *   - distances are negative inside (sdf_bake.cu:69).  "Inside" means |generalized winding
 *     number| > 0.5, which agrees with the reference's ray-parity vote on closed meshes of either
 *     winding.
 *   - occupancy is the reference's surface voxelization (src/convert/voxelize.cpp:52-107 with the
 *     triangle/box test of src/convert/tri_box_overlap.h:72-114), bit-exact.
 */
#ifndef ITALY_BAKE_H
#define ITALY_BAKE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  IB_OK = 0,
  IB_ERR_INVALID_ARGUMENT = 1,
  IB_ERR_INVALID_MESH = 2,   /* non-finite coordinates */
  IB_ERR_OUT_OF_BUDGET = 3,  /* the request does not fit the context's HBM budget */
  IB_ERR_DEVICE = 4,         /* a HIP call failed; ib_last_error() names it */
  IB_ERR_NO_DEVICE = 5,
  IB_ERR_UNSUPPORTED = 6,
  IB_ERR_CAPACITY = 7,       /* caller buffer too small; the needed size was written back */
  IB_ERR_NOT_READY = 8,      /* served job still running */
  IB_ERR_INTERNAL = 9
} ib_status;

typedef struct ib_context ib_context;
typedef struct ib_mesh ib_mesh;

/* Cell-centred grid: cell (x,y,z) spans origin + [i, i+1) * h. */
typedef struct {
  float origin[3];
  float h;
  int32_t nx, ny, nz;
  int32_t reserved;
} ib_grid;

/* Per-bake counters and device timings (milliseconds, HIP events on the bake stream). */
typedef struct {
  double ms_total;       /* first kernel start .. last kernel end */
  double ms_ingest;      /* weld + closedness (a plain bake runs them on a side stream beside the LBVH build) */
  double ms_bvh;         /* LBVH build, from the end of the weld */
  double ms_occupancy;   /* a plain bake runs occupancy on a side stream beside the build, columns and query */
  double ms_columns;     /* +z column crossings */
  double ms_sdf;         /* closest-triangle + sign per voxel */
  uint64_t triangles;
  uint64_t voxels;       /* SDF voxel queries */
  uint64_t occ_cells;    /* occupancy cells covered by the grid */
  int32_t bvh_depth;     /* internal nodes on the longest root-to-leaf path (at most 30 + ceil(log2 ntri)) */
  int32_t pad0;
  int64_t boundary_edges;
  uint64_t column_overflows;
  uint64_t rechecked;    /* voxels whose winding number was re-evaluated in FP64 */
  uint64_t ties;         /* voxels decided by the ray-parity tie-break */
  uint64_t box_tests;    /* point/box distance tests of the closest-triangle traversal (counting mode) */
  uint64_t tri_tests;    /* point/triangle distance tests (counting mode) */
  uint64_t stack_overflows; /* traversal pushes that found the stack full: 0 unless the tree is malformed */
  double ms_caps;        /* device time of the hole caps of open meshes (loop labels, fans, cap tree), events */
  double ms_query;       /* the closest-triangle query kernel alone (dense bakes; HIP events around its launch) */
  uint64_t obb_tests;    /* point/oriented-box tests of leaf blocks in the per-brick query (counting mode) */
  uint64_t on_cap;       /* rechecked voxels lying on a hole cap, decided by the mesh's own FP64 winding number */
  uint64_t rechecked_full; /* rechecked voxels that the cap tree's FP64 bound left undecided: every cap fan summed */
} ib_stats;

/* Options for ib_set_option. */
typedef enum {
  IB_OPT_COUNT_WORK = 1, /* value != 0: bakes run the counting variant of the query kernel and fill
                            ib_stats.box_tests / tri_tests (slower; for roofline accounting) */
  IB_OPT_QUERY_VARIANT = 2, /* closest-triangle kernel: 0 = one traversal per voxel, 1 = one traversal per
                               brick shared by the wave (default); results are identical */
  IB_OPT_BRICK_SHAPE = 3,   /* voxels of one wave: 0 = 4x4x4 (default), 1 = 8x4x2, 2 = 8x8x1, 3 = 16x4x1,
                               4 = 4x4x8 (two voxels per lane) */
  IB_OPT_BLOCKS_PER_CU = 4, /* persistent blocks per compute unit in batched bakes (default 8) */
  IB_OPT_TILE_ORDER = 5,    /* launch order of a dense per-brick query's tiles: 0 grid order, 1 costliest
                               first (estimated from the column crossings), 2 (default) half-octave cost
                               buckets, costliest first, Morton order inside a bucket.  Speed only. */
  IB_OPT_LEAF_BLOCK = 6,    /* LBVH subtrees of at most this many triangles (1..32, default 8) are evaluated
                               as one block from their parent.  Speed only; applies to later builds. */
  IB_OPT_RESOLVE_FULL = 7,  /* value != 0: every FP64 recheck of an open mesh sums all hole-cap fans instead
                               of first bounding the far caps through the cap tree.  Signs are identical;
                               for tests and measurement. */
  IB_OPT_SHARED_CHIP = 8    /* value != 0: the context's bakes share the GPU with other streams' work (the bake
                               server sets it on its workers): the FP64 recheck of a single-job batch takes one
                               voxel per wave, as a batch of several jobs does, instead of one voxel per
                               256-thread block, which finishes sooner on an otherwise idle GPU.  Speed only;
                               default 0. */
} ib_option;

/* One mesh of a batch: device soup, optional SDF grid (sdf.nx == 0: none) and optional occupancy
 * grid (occ.nx == 0: none), device outputs laid out as for ib_bake_dense_device.  A job may name a
 * resident mesh (ib_mesh_create) instead of a soup: its LBVH is used as built, and d_xyz / ntri are
 * written back from the mesh.
 * Narrow band (band_voxels > 0, grid dims divisible by 8): only the 8^3 bricks that can hold a voxel
 * with |d| <= band_voxels * h are baked.  d_brick_index gets, per brick (x fastest), its slot in the
 * pool or -1; d_distances is the pool: slot * 512 + (z%8 * 8 + y%8) * 8 + x%8; nbricks is written
 * back (on IB_ERR_CAPACITY it is the size the pool needs). */
typedef struct {
  const float *d_xyz;
  uint64_t ntri;
  ib_grid sdf;
  float *d_distances;
  ib_grid occ;
  uint32_t *d_occ_words;
  int32_t band_voxels;
  int32_t reserved;
  int32_t *d_brick_index;
  uint64_t pool_bricks;
  uint64_t nbricks;
  const ib_mesh *mesh;   /* NULL: build from d_xyz / ntri */
} ib_job;

/* ---- context ------------------------------------------------------------------------------ */
/* device: HIP ordinal.  hbm_budget_bytes: cap on what the context may allocate (0 = 90% of free). */
ib_status ib_create(int device, uint64_t hbm_budget_bytes, ib_context **out);
void ib_destroy(ib_context *ctx);
const char *ib_last_error(void);
const char *ib_version(void);

/* ---- resident meshes ----------------------------------------------------------------------------
 * A mesh built once and kept on the device: a copy of its soup, its LBVH, distance records and hole caps.
 * Batched bakes take it in place of a soup (ib_job.mesh) and skip the ingest and LBVH build.  Built from
 * a device soup (copied, so the caller's buffer may be reused when the call returns) on `stream` (NULL =
 * the context's stream); returns when the mesh is complete.  Its memory counts against ctx's budget.
 * Destroy a mesh after the bakes that use it have returned and before its context.  No counterpart in
 * the reference, which builds its acceleration structure on every bakeSdf call
 * (src/convert/sdf_baker.cpp:170-193). */
ib_status ib_mesh_create(ib_context *ctx, const float *d_xyz, uint64_t ntri, void *stream, ib_mesh **out);
void ib_mesh_destroy(ib_mesh *mesh);
uint64_t ib_mesh_bytes(const ib_mesh *mesh);   /* device bytes the mesh holds */
uint64_t ib_mesh_triangles(const ib_mesh *mesh);

/* ---- framing (pure host arithmetic, identical to the reference) ----------------------------- */
/* grid of italy::bakeSdf(mesh, resolution): src/convert/sdf_baker.cpp:139-148 */
ib_status ib_sdf_framing(const float bmin[3], const float bmax[3], int32_t resolution, ib_grid *grid);
/* origin and voxel size of italy::voxelizeMesh(mesh, resolution): src/convert/voxelize.cpp:54-57 */
ib_status ib_vox_framing(const float bmin[3], const float bmax[3], int32_t resolution, float origin[3],
                         float *voxel_size);

/* ---- drop-in entry points (host buffers) ---------------------------------------------------- */
/* Replaces the distance part of italy::bakeSdf(const MeshAsset&, int) (src/convert/sdf_baker.h:8,
 * body sdf_baker.cpp:137-310): frames the grid as the reference does, writes nx*ny*nz signed
 * distances.  Pass distances = NULL to only receive the grid. */
ib_status ib_bake_sdf(ib_context *ctx, const float *xyz, uint64_t ntri, const float bmin[3], const float bmax[3],
                      int32_t resolution, ib_grid *grid, float *distances, uint64_t capacity);

/* Replaces italy::voxelizeMesh(const MeshAsset&, int) (src/convert/voxelize.h:8, body
 * voxelize.cpp:52-107): writes the occupied cells (x,y,z int triples) in the reference's first-touch
 * order (triangle index, then z, y, x).  cell_tri receives, per cell, the first touching triangle
 * (may be NULL).  On IB_ERR_CAPACITY *count holds the number of cells needed. */
ib_status ib_voxelize(ib_context *ctx, const float *xyz, uint64_t ntri, const float bmin[3], const float bmax[3],
                      int32_t resolution, int32_t *cells, int32_t *cell_tri, uint64_t capacity, uint64_t *count);

/* The whole of italy::voxelizeMesh (voxelize.cpp:52-107) in one call: the occupied cells in first-touch
 * order (as ib_voxelize) and VoxelGrid.colors, per cell the mean of its triangles' colours (tri_colors: 3
 * floats per triangle, the reference's triangleColor; NULL: white), summed in triangle order and divided by
 * the count per component.  Memory and work scale with the (cell, triangle) overlap pairs, not with the grid.
 * cells / colors (3 per cell each) may be NULL to only count; on IB_ERR_CAPACITY *count holds the number
 * of cells needed.  Host buffers. */
ib_status ib_voxelize_colors(ib_context *ctx, const float *xyz, uint64_t ntri, const float bmin[3],
                             const float bmax[3], int32_t resolution, const float *tri_colors, int32_t *cells,
                             float *colors, uint64_t capacity, uint64_t *count);

/* Per-triangle material as the bake and the voxelizer compute it: triangleMaterial
 * (src/convert/sdf_baker.cpp:61-74) and triangleColor (src/convert/voxelize.cpp:26-38), with
 * sampleColorNearest and srgbToLinear (sdf_baker.cpp:42-53).  Per triangle t: the material
 * tri_material[t] (0 past the table or out of range; MaterialAsset's defaults for an empty table,
 * src/io/mesh_asset.h:18-52); its base colour times the nearest texel at the centroid of the triangle's
 * three UVs (wrapped by x - floor(x)) when it names a texture with texels and the triangle has UVs, sRGB
 * textures converted to linear.  uvs: 2 floats per corner (ncorner_uv of them, 3 per triangle).  out: 5 floats
 * per triangle (r, g, b, metallic, roughness), the input of ib_splat_materials.  Host buffers. */
typedef struct {
  float base_color[3];          /* baseColorFactor */
  float metallic, roughness;
  int32_t base_color_texture;   /* index into the texture table, -1: none */
} ib_material;
typedef struct {
  const uint8_t *rgba;          /* width * height texels of 4 bytes, rows of width texels */
  int32_t width, height;
  int32_t srgb;                 /* != 0: colour bytes are sRGB encoded */
  int32_t reserved;
} ib_texture;
ib_status ib_triangle_materials(ib_context *ctx, uint64_t ntri, const float *uvs, uint64_t ncorner_uv,
                                const uint32_t *tri_material, uint64_t ntri_material, const ib_material *materials,
                                uint32_t nmaterials, const ib_texture *textures, uint32_t ntextures, float *out);

/* Per-cell material splat of bakeSdf (src/convert/sdf_baker.cpp:76-128): per-triangle material
 * (r,g,b,metallic,roughness: 5 floats each) averaged over every grid cell the triangle overlaps,
 * accumulated in triangle order; cells without triangles get (1,1,1,0,1). */
ib_status ib_splat_materials(ib_context *ctx, const float *xyz, uint64_t ntri, const float *tri_material,
                             const ib_grid *grid, float *base_color, float *metallic, float *roughness);

/* ---- mesh ingestion gate (src/io/mesh_weld.cpp, src/io/mesh_validate.cpp) ----------------------
 * The checks the reference runs on a loaded mesh before it bakes (src/app/main.cpp:227-241). */
/* Replaces italy::weldPositions(positions, boundsMin, boundsMax, &uniquePositions)
 * (src/io/mesh_weld.h:10-12, body mesh_weld.cpp:26-48): positions (3 floats each) that round to the same
 * point of the lattice of pitch max(|bmax - bmin| * 1e-5, 1e-6) share an index; indices count distinct
 * lattice points in order of first occurrence.  welded: npos indices; unique (may be NULL; room for
 * 3 * npos floats): the first position of each index; *nunique = weldedIndexCount(welded)
 * (mesh_weld.cpp:50-54).  A position whose quotient by the pitch is not finite or reaches 2^63 (where
 * std::lround is unspecified) is rejected with IB_ERR_INVALID_MESH.  Host buffers. */
ib_status ib_weld_positions(ib_context *ctx, const float *positions, uint64_t npos, const float bmin[3],
                            const float bmax[3], uint32_t *welded, float *unique, uint64_t *nunique);
/* Same with device buffers, on `stream` (NULL = the context's stream); returns when *nunique is known. */
ib_status ib_weld_positions_device(ib_context *ctx, const float *d_positions, uint64_t npos, const float bmin[3],
                                   const float bmax[3], uint32_t *d_welded, float *d_unique, uint64_t *nunique,
                                   void *stream);
/* Replaces italy::isWatertight(const MeshAsset&, std::string &err) (src/io/mesh_validate.h:9, body
 * mesh_validate.cpp:21-59) for a triangle soup with its bounds: welds the corners as above, skips
 * triangles with a repeated index, and counts per undirected edge how many triangles use it.
 * *watertight = 1 when ntri > 0 and every edge is used exactly twice.  boundary_edges (used once) and
 * nonmanifold_edges (used more than twice) may be NULL.  message (may be NULL) receives the reference's
 * err text, "" when watertight, truncated to message_cap - 1 bytes. */
ib_status ib_watertight(ib_context *ctx, const float *xyz, uint64_t ntri, const float bmin[3], const float bmax[3],
                        int32_t *watertight, uint64_t *boundary_edges, uint64_t *nonmanifold_edges, char *message,
                        uint64_t message_cap);
/* Same on a device soup, on `stream` (NULL = the context's stream); returns when the verdict is known. */
ib_status ib_watertight_device(ib_context *ctx, const float *d_xyz, uint64_t ntri, const float bmin[3],
                               const float bmax[3], int32_t *watertight, uint64_t *boundary_edges,
                               uint64_t *nonmanifold_edges, char *message, uint64_t message_cap, void *stream);

/* ---- explicit grids (the workload entry points) -------------------------------------------- */
/* Dense bake of one mesh: sdf grid (may be NULL) and occupancy grid (may be NULL).  Host buffers. */
ib_status ib_bake_dense(ib_context *ctx, const float *xyz, uint64_t ntri, const ib_grid *sdf, float *distances,
                        const ib_grid *occ, uint32_t *occ_words);
/* Same with device-resident input and outputs, enqueued on `stream` (hipStream_t, NULL = the
 * context's stream).  occ_words holds ny*nz*ceil(nx/32) words; bit x&31 of word
 * ((z*ny + y)*ceil(nx/32) + x/32). */
ib_status ib_bake_dense_device(ib_context *ctx, const float *d_xyz, uint64_t ntri, const ib_grid *sdf,
                               float *d_distances, const ib_grid *occ, uint32_t *d_occ_words, void *stream);

/* Scratch-free form of ib_bake_dense_device, for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference, whose bakeSdf allocates per call (src/convert/sdf_baker.cpp:165-305).  Same arguments and
 * results (bit for bit) as ib_bake_dense_device; it calls no hipMalloc, hipFree or other allocating
 * device routine: every table of the bake lives in one fixed 128 MiB region per device, which the library
 * holds from the first ib_create on that device until the process ends.  Limits (IB_ERR_CAPACITY before
 * any work beyond them): at most IB_NOALLOC_MAX_TRIANGLES triangles, SDF grids of at most
 * IB_NOALLOC_MAX_COLUMNS columns (nx * ny) and IB_NOALLOC_MAX_VOXELS voxels; the occupancy grid is
 * output only and has no limit.  Every call on one device shares the region, whatever its context, and
 * a call keeps it until the call's device work ends: calls that overlap on one device (two contexts, two
 * threads, or a later call before the earlier one's work has finished on another context) race. */
#define IB_NOALLOC_MAX_TRIANGLES 65536ull
#define IB_NOALLOC_MAX_COLUMNS 16384ull
#define IB_NOALLOC_MAX_VOXELS 2097152ull
ib_status ib_bake_dense_device_noalloc(ib_context *ctx, const float *d_xyz, uint64_t ntri, const ib_grid *sdf,
                                       float *d_distances, const ib_grid *occ, uint32_t *d_occ_words, void *stream);

/* Batched bake: every job's LBVH is built, then ONE persistent launch drains the brick tiles of
 * all jobs from a device work queue.  Synchronous: returns when every output is written.  Results
 * are identical to baking each job alone with ib_bake_dense_device (no counterpart in the
 * reference, which bakes one mesh per call: src/convert/sdf_baker.cpp:137-310). */
ib_status ib_bake_batch_device(ib_context *ctx, ib_job *jobs, uint32_t njobs, void *stream);

/* Narrow-band bake of one mesh (the configs' sparse-brick output): a batch of one band job. */
ib_status ib_bake_band_device(ib_context *ctx, const float *d_xyz, uint64_t ntri, const ib_grid *grid,
                              int32_t band_voxels, int32_t *d_brick_index, float *d_pool, uint64_t pool_bricks,
                              uint64_t *nbricks, void *stream);

/* Autotuner for batched bakes: bakes every sample_stride-th job of `jobs` (in order of voxel count, so that
 * the sample spans the batch's grid sizes) with each brick shape
 * (IB_OPT_BRICK_SHAPE) and persistent-block count (IB_OPT_BLOCKS_PER_CU), times the persistent
 * query launch with HIP events, and leaves the fastest setting on the context.  The choice changes
 * speed only, never results.  Outputs (may be NULL): the chosen shape, blocks per CU and time. */
ib_status ib_autotune_batch(ib_context *ctx, const ib_job *jobs, uint32_t njobs, uint32_t sample_stride,
                            int32_t *out_shape, int32_t *out_blocks, double *out_ms);

/* ---- served bakes ----------------------------------------------------------------------------
 * A server admits requests first-come-first-served while their device footprint fits its HBM budget,
 * bakes whatever is admitted in batches on `workers` contexts at once (each on its own stream; a
 * request that arrives while batches run joins the next free worker's batch), and keeps every
 * result resident on the device until ib_server_release.  Requests with a mesh_key also keep their
 * LBVH resident (a cache inside the same HBM budget; least recently used meshes are evicted when an
 * admission needs the memory), so a mesh seen before is not uploaded or built again.  The soup passed
 * to submit is read when the request starts and must stay valid until the ticket is done.  No
 * counterpart in the reference, whose bake is a synchronous call on the UI thread
 * (src/app/main.cpp:250-259). */
typedef struct ib_server ib_server;

typedef struct {
  const float *xyz;      /* host soup, 9 floats per triangle */
  uint64_t ntri;
  ib_grid sdf;           /* sdf.nx == 0: no SDF */
  int32_t band_voxels;   /* 0: dense SDF; > 0: narrow band in 8^3 bricks (grid dims divisible by 8) */
  int32_t reserved;
  ib_grid occ;           /* occ.nx == 0: no occupancy */
  uint64_t mesh_key;     /* != 0: the soup's identity; the server keeps the built LBVH resident under this key
                            and later requests with the same key (and the same soup) reuse it */
} ib_request;

typedef struct {
  int32_t state;         /* 0 queued, 1 running, 2 done, 3 failed */
  int32_t status;        /* ib_status of a failed request */
  double t_submit, t_admit, t_start, t_done;  /* seconds on the server's steady clock */
  uint64_t nbricks;      /* narrow band: bricks in the result */
  uint64_t resident_bytes;
  int32_t worker;
  int32_t reserved;
} ib_ticket_info;

typedef struct {
  uint64_t budget;
  uint64_t in_use;       /* reserved by admitted requests, kept by finished results and resident meshes */
  uint64_t cached;       /* staging blocks kept for reuse; counted against the budget at admission */
  uint64_t pending, ready, running_batches, batches, completed;
  uint64_t mesh_hits, mesh_builds, meshes_resident, mesh_bytes;  /* resident-mesh cache (ib_request.mesh_key) */
  uint64_t evictions;      /* resident meshes evicted to admit requests */
  uint64_t failed;         /* requests that ended in an error */
  uint64_t alloc_failures; /* staging allocations that failed even after the cache was freed */
} ib_server_counters;

ib_status ib_server_create(int device, uint64_t hbm_budget_bytes, int32_t workers, int32_t max_batch,
                           ib_server **out);
void ib_server_destroy(ib_server *srv);
const char *ib_server_last_error(void);
ib_status ib_server_submit(ib_server *srv, const ib_request *req, uint64_t *ticket);
ib_status ib_server_info(ib_server *srv, uint64_t ticket, ib_ticket_info *info);
ib_status ib_server_wait(ib_server *srv, uint64_t ticket, double timeout_s);
ib_status ib_server_drain(ib_server *srv, double timeout_s);
/* host copies of a finished result: distances (dense grid, or nbricks * 512 pool floats), the brick
 * index of a narrow band, the occupancy words; any pointer may be NULL */
ib_status ib_server_fetch(ib_server *srv, uint64_t ticket, float *distances, uint64_t capacity, int32_t *brick_index,
                          uint32_t *occ_words);
/* device pointers of a finished result (valid until release) */
ib_status ib_server_device_result(ib_server *srv, uint64_t ticket, const float **d_distances,
                                  const int32_t **d_brick_index, const uint32_t **d_occ_words);
/* frees a finished result; IB_ERR_NOT_READY while the ticket runs or another thread fetches it */
ib_status ib_server_release(ib_server *srv, uint64_t ticket);
ib_status ib_server_stats(ib_server *srv, ib_server_counters *out);

/* ---- analytic SDF bake (src/convert/sdf_procedural.cpp) -------------------------------------------
 * The reference bakes procedural content from a std::function (bakeAnalyticSdf / bakeAnalyticSdfPalette,
 * sdf_procedural.cpp:46-79); here the function is a postfix CSG program over a stack of
 * (distance, branch) values, evaluated on the device once per cell with the primitives and operators of
 * src/convert/sdf_primitives.h. */
typedef enum {
  IB_CSG_SPHERE = 1,            /* push sdSphere(p - c, r):            c = a[0..2], r = a[3]            */
  IB_CSG_BOX = 2,               /* push sdBox(p - c, b):               c = a[0..2], b = a[3..5]         */
  IB_CSG_CYLINDER = 3,          /* push sdCylinder(p - c, r, h):       c = a[0..2], r = a[3], h = a[4]  */
  IB_CSG_TORUS = 4,             /* push sdTorus(p - c, major, minor):  c = a[0..2], a[3], a[4]          */
  IB_CSG_PLANE = 5,             /* push dot(n, p) - d:                 n = a[3..5], d = a[6]            */
  IB_CSG_UNION = 10,            /* pop b, a; push std::min(a, b) (the operand chosen keeps its branch)  */
  IB_CSG_SUBTRACT = 11,         /* std::max(a, -b), branch of a                                          */
  IB_CSG_INTERSECT = 12,        /* std::max(a, b) (the operand chosen keeps its branch)                  */
  IB_CSG_SMOOTH_UNION = 13,     /* opSmoothUnion(a, b, k = a[0]), branch of a                           */
  IB_CSG_SMOOTH_INTERSECT = 14, /* opSmoothIntersect(a, b, k = a[0]), branch of a                       */
  IB_CSG_SELECT = 15            /* a <= b ? a : b with its branch (the two-branch pick of the glass cup) */
} ib_csg_opcode;

typedef struct {
  int32_t op;      /* ib_csg_opcode */
  int32_t branch;  /* branch id of a primitive (0..255) */
  float a[8];
} ib_csg_op;

/* Grid framing of the analytic bake (sdf_procedural.cpp:12-26): h = longest extent / max(res, 4),
 * origin = bmin - 2h, n = ceil(extent / h) + 4 per axis. */
ib_status ib_analytic_framing(const float bmin[3], const float bmax[3], int32_t res, ib_grid *out);
/* Evaluate a CSG program (at most 256 instructions, stack depth <= 16, exactly one value left) at
 * every cell centre of `grid`: d_distances (FP32) and, if not NULL, d_branch (one byte per cell).
 * Enqueued on `stream` (NULL: the context's stream); programs of up to 64 instructions travel in the
 * kernel arguments and the call returns without waiting, longer ones wait for their upload.
 * ib_last_stats then reports the launch's device time as ms_total (HIP events on `stream`).
 * Grids are limited to ny <= 262140 and nz <= 65535. */
ib_status ib_bake_analytic_device(ib_context *ctx, const ib_csg_op *program, uint32_t nops, const ib_grid *grid,
                                  float *d_distances, uint8_t *d_branch, void *stream);
/* Host-buffer form. */
ib_status ib_bake_analytic(ib_context *ctx, const ib_csg_op *program, uint32_t nops, const ib_grid *grid,
                           float *distances, uint8_t *branch);

/* ---- the grid consumer on the device ----------------------------------------------------------
 * What the renderer does with a baked SDF grid, run on device-resident bakes without a host round
 * trip (the renderer uploads SdfGrid to the device, src/render/optix_renderer.cpp:868-930).
 * Grids need at least 2 cells per axis.  Enqueued on `stream` (NULL = the context's stream). */
/* For n points (3 floats each): the trilinear sample (sampleSdf, src/render/kernels/pathtracer.cu:684-701),
 * the normalised cubic B-spline gradient (sdfGradientNormal, :806-846) and the nearest cell index
 * (sdfNearestCellIndex, :703-709).  Any output may be NULL. */
ib_status ib_sample_grid_device(ib_context *ctx, const ib_grid *grid, const float *d_distances, const float *d_points,
                                uint64_t n, float *d_values, float *d_normals, int32_t *d_nearest, void *stream);
/* Sphere-trace n rays (origins, directions: 3 floats each) through the grid exactly as the SDF
 * intersection program does (__intersection__sdf, pathtracer.cu:740-792, with the slab test of :587-613):
 * d_t receives the hit distance, or -1 for a miss, over [tmin, tmax]. */
ib_status ib_trace_grid_device(ib_context *ctx, const ib_grid *grid, const float *d_distances, const float *d_origins,
                               const float *d_dirs, uint64_t n, float tmin, float tmax, float *d_t, void *stream);

/* ---- device memory helpers (for callers without a GPU framework) --------------------------- */
ib_status ib_device_alloc(ib_context *ctx, uint64_t bytes, void **d_ptr);
ib_status ib_device_free(ib_context *ctx, void *d_ptr);
ib_status ib_memcpy_h2d(ib_context *ctx, void *d_dst, const void *src, uint64_t bytes);
ib_status ib_memcpy_d2h(ib_context *ctx, void *dst, const void *d_src, uint64_t bytes);
/* device-to-device copy on the context's stream (synchronous) */
ib_status ib_memcpy_d2d(ib_context *ctx, void *d_dst, const void *d_src, uint64_t bytes);
ib_status ib_synchronize(ib_context *ctx);
ib_status ib_last_stats(ib_context *ctx, ib_stats *out);
ib_status ib_set_option(ib_context *ctx, int32_t option, int64_t value);

#ifdef __cplusplus
}
#endif
#endif
