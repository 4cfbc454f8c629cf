/*
 * CPU oracle for the mesh -> SDF / occupancy bake path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (italy-rs_amd/) never links, loads or calls it.
 *
 * It restates, in plain C99 compiled with -ffp-contract=off:
 *   - triBoxOverlap          src/convert/tri_box_overlap.h:12-114 (literal expression trees)
 *   - voxelizeMesh           src/convert/voxelize.cpp:52-107 (framing, clamp, first-touch order)
 *   - bakeSdf grid framing   src/convert/sdf_baker.cpp:137-148
 *   - cell-centre sampling   src/render/kernels/sdf_bake.cu:34-37, index :71, sign :69
 *   - parity sign vote       src/render/kernels/sdf_bake.cu:53-69 (+z first, majority of 5)
 * and the exact quantities the new bake computes:
 *   - closest-point distance (FP32 in query-local coordinates, FP64 for slivers; FP64 reference)
 *   - generalized winding number (FP64 brute-force solid angles, Van Oosterom-Strackee 1983)
 *   - signed watertight crossing count along +z columns (Woop et al. 2013 edge functions)
 *   - LBVH (Karras 2012) over 30-bit Morton codes, same integer topology the device builds;
 *   - weldPositions / isWatertight  src/io/mesh_weld.cpp:26-48, src/io/mesh_validate.cpp:21-59.
 *
 * tests/test_oracle.py and tests/test_oracle_csg.py pin every function against known answers,
 * closed forms and independent brute force; a function without such a check is "parity unpinned".
 */
#ifndef IB_ORACLE_H
#define IB_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- a10: triangle / AABB separating-axis test ------------------------------ */
int orc_tri_box_overlap(const float center[3], const float half[3], const float t0[3], const float t1[3],
                        const float t2[3]);

/* ---- a1 / a9: framing -------------------------------------------------------- */
/* bakeSdf framing: sdf_baker.cpp:139-148 */
void orc_sdf_framing(const float bmin[3], const float bmax[3], int resolution, float origin[3], float *h, int n[3]);
/* voxelizeMesh framing: voxelize.cpp:54-57 */
void orc_vox_framing(const float bmin[3], const float bmax[3], int resolution, float origin[3], float *h);

/* ---- a9: voxelizeMesh, first-touch order --------------------------------------
 * xyz: 9 floats per triangle (soup).  colors: 3 floats per triangle (may be NULL -> 1,1,1).
 * Writes up to cap cells (3 ints each, x,y,z) and their averaged colours; returns the count.
 * res_clamp is resolution-1 on every axis, as voxelize.cpp:72-77. */
size_t orc_voxelize_first_touch(const float *xyz, size_t ntri, const float *colors, const float bmin[3],
                                const float bmax[3], int resolution, int32_t *cells, float *cell_colors, size_t cap);

/* Bit-packed occupancy over an explicit grid (origin, h, nx,ny,nz): same per-triangle cell range
 * (clamped to [0,n-1] per axis), same centre/half, same SAT.  Bit x&31 of word
 * ((z*ny + y)*((nx+31)/32) + x/32).  words must hold ny*nz*((nx+31)/32) zeroed words. */
void orc_occupancy(const float *xyz, size_t ntri, const float origin[3], float h, int nx, int ny, int nz,
                   uint32_t *words);

/* ---- distance ----------------------------------------------------------------- */
/* 1 if the triangle takes the FP64 path (near-degenerate: sin^2 of its corner angle < 1e-4). */
int orc_tri_is_sliver(const float *tri9);
/* distance record of one triangle: an orthonormal frame at p0 (X along p1 - p0, Z the unit normal,
 * Y = Z x X) with the triangle's corners in it, p1 = (len, 0), p2 = (c0, c1), and the reciprocals
 * the query uses; a sliver keeps its three vertices for the FP64 path instead. */
typedef struct {
  float o[3], x[3], y[3], z[3];
  float len, inv_len;
  float e1[2], inv_e1;
  float c[2], nie2;
  int sliver;
  float v[9]; /* the vertices (FP64 path) */
} orc_trirec;
void orc_trirec_make(const float *tri9, orc_trirec *r);
float orc_trirec_dist2(const float q[3], const orc_trirec *r);
/* bakePerCellMaterial (sdf_baker.cpp:76-128): 5 material planes (r, g, b, metallic, roughness) of
 * nx*ny*nz cells from per-triangle materials mat[5*ntri], summed in triangle order. */
void orc_splat_materials(const float *xyz, size_t ntri, const float *mat, const float origin[3], float h, int nx,
                         int ny, int nz, float *out);
/* squared distance point->triangle: branch-free FP32 over the distance record, FP64 for slivers */
float orc_tri_dist2_f32(const float q[3], const float *tri9);
/* squared distance point->triangle, all FP64 (min over edges + interior projection) */
double orc_tri_dist2_f64(const double q[3], const float *tri9);

/* ---- winding number ----------------------------------------------------------- */
/* FP64 brute force generalized winding number: sum of solid angles / 4pi */
double orc_gwn_f64(const double q[3], const float *xyz, size_t ntri);
/* FP64 winding numbers at npts points: a bounded hierarchical estimate where it decides |w| > 1/2 and the
 * 1e-6 tie band as the exact sum does, else orc_gwn_f64 (out_exact[i] = 1).  Returns 0, or -1 when the tree
 * cannot be built. */
int orc_gwn_f64_points(const float *xyz, size_t ntri, const double *pts, size_t npts, int threads, double *out_w,
                       uint8_t *out_exact);
/* Signed crossing count of the +z ray from (x,y,z) with the triangles (Woop edge functions with
 * exact zero handling and a half-open tie rule).  For a closed consistently oriented mesh this
 * equals the winding number of the point. */
int orc_crossings_up(const float q[3], const float *xyz, size_t ntri);
/* Parity vote (sdf_bake.cu:53-69 structure): 5 directions, dir 0 = +z, the rest from a fixed
 * table; each ray counts hits with t > 0; inside if the majority of rays see odd counts. */
int orc_parity_inside(const float q[3], const float *xyz, size_t ntri);
/* the 5 tie-break directions used by both the oracle and the device code */
void orc_parity_dirs(float dirs[15]);
/* hits (t > 0, every crossing counted once) of each of the 5 parity rays from q */
void orc_parity_hits(const float q[3], const float *xyz, size_t ntri, int hits[5]);
/* The bake's sign rule for any mesh, open or closed: inside iff |w| > 0.5 for the FP64 brute-force
 * winding number w at the FP32 point q, except that |w| within 1e-6 of 0.5 is a tie, decided by
 * orc_parity_inside (*tie = 1 then; tie may be NULL). */
int orc_gwn_sign(const float q[3], const float *xyz, size_t ntri, int *tie);
/* Signed distances at FP32 points of any mesh: the FP32 closest-triangle distance through the LBVH (equal
 * to the FP32 brute force) with the sign of orc_gwn_sign (negative inside); tie (may be NULL) gets 1 per
 * point decided by parity.  O(ntri) per point for the sign: for small meshes and samples.  -1: empty mesh. */
int orc_signed_points_gwn(const float *xyz, size_t ntri, const float *pts, size_t npts, int threads, float *out,
                          uint8_t *tie);
/* The dense bake at arbitrary FP32 points of a closed mesh: orc_bake_dense's distance and crossing sign
 * (the signed count of +z crossings above the point, through the LBVH), per point instead of per slab.
 * -1: the mesh has boundary edges, -2: empty. */
int orc_signed_points_closed(const float *xyz, size_t ntri, const float *pts, size_t npts, int threads, float *out);

/* ---- LBVH (Karras 2012) --------------------------------------------------------- */
typedef struct {
  int32_t n;          /* triangles */
  int32_t *order;     /* n: sorted position -> original triangle index */
  uint32_t *codes;    /* n: sorted Morton codes */
  int32_t *left;      /* n-1: child of internal node i; >=0 internal, <0 = ~leaf */
  int32_t *right;     /* n-1 */
  int32_t *parent;    /* 2n-1: parent of internal (0..n-2) and leaves (n-1+leaf) */
  float *lo, *hi;     /* (2n-1)*3: boxes of internal nodes then leaves */
  int32_t depth;      /* max leaf depth (root = 0) */
} orc_bvh;

int orc_bvh_build(const float *xyz, size_t ntri, orc_bvh *out);
void orc_bvh_free(orc_bvh *b);
/* Morton code of each triangle centroid (unsorted), same arithmetic as the device build */
void orc_morton_codes(const float *xyz, size_t ntri, uint32_t *codes);

/* ---- full CPU bake (the host-core baseline) ------------------------------------ */
typedef struct {
  uint64_t voxels;
  uint64_t node_visits;   /* distance traversal box tests */
  uint64_t tri_tests;     /* distance triangle tests */
  uint64_t column_tri_tests;
  uint64_t inside;
  double seconds;
  int threads;
} orc_bake_stats;

/* Dense SDF over an explicit cell-centred grid (index (z*ny+y)*nx+x, negative inside), using the
 * LBVH for the closest triangle and +z column crossings for the sign.  Only closed meshes are
 * signed exactly this way; the function returns -1 if the mesh has boundary edges.
 * z_begin/z_end restrict the slabs computed (for bounded samples); threads<=0 -> all cores. */
int orc_bake_dense(const float *xyz, size_t ntri, const float origin[3], float h, int nx, int ny, int nz,
                   int z_begin, int z_end, int threads, float *dist, orc_bake_stats *stats);

/* same, computing only every zstep-th slab of [z_begin, z_end) */
int orc_bake_dense_strided(const float *xyz, size_t ntri, const float origin[3], float h, int nx, int ny, int nz,
                           int z_begin, int z_end, int zstep, int threads, float *dist, orc_bake_stats *stats);

/* Narrow band of +-band voxels in 8^3 bricks (the bake's band output) for any mesh, open or closed.  Bricks
 * are kept by the bake's rule (FP32 unsigned distance at the brick centre, on the grid of spacing 8h,
 * <= h (band + 3.5 sqrt 3)); brick_index (nx/8 * ny/8 * nz/8, x fastest) gets each brick's slot in order of
 * its linear index, or -1, and *nbricks the count.  Every brick_stride-th slot is baked into pool
 * (slot * 512 + ((z%8)*8 + y%8)*8 + x%8; the other slots are left as they are): FP32 closest-triangle
 * distance, sign by the closure identity with a fan cap over each boundary loop and an FP64 cap winding
 * number (bounded far field, full sum near 1/2, the mesh's own winding number on the cap, parity on ties).
 * Returns 0, -1 if pool_bricks < *nbricks (nothing baked), -2 for an empty mesh or grid dims not
 * divisible by 8. */
typedef struct {
  uint64_t voxels, bricks_baked, cap_triangles, rechecked, full_cap_sums, on_cap, ties;
  double seconds;
  int threads;
} orc_band_stats;
int orc_bake_band(const float *xyz, size_t ntri, const float origin[3], float h, int nx, int ny, int nz, int band,
                  int brick_stride, int threads, int32_t *brick_index, float *pool, size_t pool_bricks,
                  size_t *nbricks, orc_band_stats *st);

/* FP64 brute-force signed distance at arbitrary points (inside where |GWN| > 0.5, FP64) */
void orc_sdf_f64_points(const float *xyz, size_t ntri, const double *pts, size_t npts, int threads, double *out_dist,
                        double *out_w);

/* FP32 brute force: out[i] = min over triangles of orc_tri_dist2_f32(pts[i], tri) */
void orc_udf2_f32_brute(const float *xyz, size_t ntri, const float *pts, size_t npts, float *out);

/* FP64 unsigned distance at points through the LBVH: the same minimum of orc_tri_dist2_f64 over every
 * triangle as the brute force (subtrees are skipped only beyond a 1e-9 relative margin); -1 for an empty
 * mesh, -2 for a tree deeper than 249 levels */
int orc_udf_points_f64(const float *xyz, size_t ntri, const double *pts, size_t npts, int threads, double *out);
/* FP32 unsigned distance at points through the LBVH (equals the FP32 brute force) */
int orc_udf_points_f32(const float *xyz, size_t ntri, const float *pts, size_t npts, int threads, float *out);

/* number of directed edges left unmatched after exact-position welding (0 => closed) */
int64_t orc_boundary_edge_count(const float *xyz, size_t ntri);

/* The renderer's grid consumer (src/render/kernels/pathtracer.cu): trilinear sampleSdf (:684-701),
 * normalised cubic B-spline gradient (:806-846) and nearest cell (:703-709) at n points; outputs may
 * be NULL.  orc_trace_grid: the SDF intersection program's march (:587-613, :740-792), -1 = miss. */
void orc_sample_grid(const float origin[3], float h, int nx, int ny, int nz, const float *dist, const float *pts,
                     size_t n, float *values, float *normals, int32_t *nearest);
void orc_trace_grid(const float origin[3], float h, int nx, int ny, int nz, const float *dist, const float *org,
                    const float *dir, size_t n, float tmin, float tmax, float *t_hit);

/* worst (exact - FP32), (FP32 - exact) distance of orc_trirec_dist2 in units of 2^-24 (Mq + 3 Mp) */
void orc_f32_error_ratio(long n, uint64_t seed, double out[2]);

/* Analytic SDF bake (src/convert/sdf_procedural.cpp).  A CSG instruction is {int32 op, int32 branch,
 * float a[8]} with the opcodes of include/italy_bake.h (ib_csg_opcode); orc_csg_eval evaluates a
 * postfix program at one point with the primitives of sdf_primitives.h:10-36, orc_bake_csg every
 * cell centre of a grid framed by orc_analytic_framing (sdf_procedural.cpp:12-26).
 * orc_cup_cells / orc_glass_cup_cells restate makeCupSdf / makeGlassCupSdf (:81-140) as written,
 * without a program: (grid, distances, branch), branch may be NULL. */
typedef struct {
  int32_t op, branch;
  float a[8];
} orc_csg_op;
float orc_csg_eval(const orc_csg_op *prog, int nops, const float p[3], int *branch);
void orc_analytic_framing(const float bmin[3], const float bmax[3], int res, float origin[3], float *h, int32_t n[3]);
void orc_bake_csg(const orc_csg_op *prog, int nops, const float origin[3], float h, int nx, int ny, int nz,
                  int threads, float *dist, uint8_t *branch);
void orc_cup_cells(int glass, int res, float origin_y, float cx, float cz, float origin[3], float *h, int32_t n[3],
                   float *dist, uint8_t *branch);

/* ---- mesh ingestion gate (src/io/mesh_weld.cpp:26-48, src/io/mesh_validate.cpp:21-59) ---------------
 * orc_weld_positions: positions rounding to the same point of a lattice of pitch
 * max(|bmax - bmin| * 1e-5, 1e-6) share an index; indices in order of first occurrence; unique (may be
 * NULL) gets the first position of each index.  Returns the number of indices, or -1 when p / pitch is
 * non-finite or beyond 2^63 (where lround is unspecified).
 * orc_watertight: 1 if ntri > 0 and every undirected edge of the welded, non-degenerate triangles is
 * used exactly twice, 0 if not, -1 as above; boundary = edges used once, nonmanifold = more than twice. */
int64_t orc_weld_positions(const float *pos, size_t npos, const float bmin[3], const float bmax[3], uint32_t *welded,
                           float *unique);
int orc_watertight(const float *xyz, size_t ntri, const float bmin[3], const float bmax[3], int64_t *boundary,
                   int64_t *nonmanifold);

/* ---- per-triangle material (src/convert/sdf_baker.cpp:42-74, voxelize.cpp:12-38) --------------------
 * (r, g, b, metallic, roughness) per triangle: the material of triangleMaterial / materialForTriangle
 * (mesh_asset.h:49-52; an empty table means the default material), its base colour times the nearest texel
 * at the centroid UV (sampleColorNearest, wrapped by x - floor(x)) when it has a texture and the triangle has
 * UVs, sRGB bytes through srgbToLinear.  uvs: 2 floats per corner (nuv corners). */
typedef struct {
  float base[3];
  float metallic, roughness;
  int32_t tex;
} orc_material;
typedef struct {
  const uint8_t *rgba;
  int32_t width, height, srgb;
} orc_texture;
void orc_triangle_materials(size_t ntri, const float *uvs, size_t nuv, const uint32_t *tri_material,
                            size_t ntri_material, const orc_material *materials, size_t nmat,
                            const orc_texture *textures, size_t ntex, float *out);

#ifdef __cplusplus
}
#endif
#endif
