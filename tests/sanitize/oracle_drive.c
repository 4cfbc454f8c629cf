/* Drives every entry point of the CPU oracle on small meshes, for the sanitizer builds (AddressSanitizer +
 * UndefinedBehaviorSanitizer, ThreadSanitizer for the threaded functions).  Checks only what the inputs fix
 * (closed-mesh signs, counts); the numerical pins live in tests/test_oracle*.py.  Exit status 0 = clean. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ib_oracle.h"

static int failures = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      ++failures;                                                  \
    }                                                              \
  } while (0)

int main(void) {
  float closed[12 * 9], open[12 * 9];
  size_t nc = 0, no = 0;
  /* the box [-0.5, 0.5]^3 wound outward, and a smaller one without its +z face */
  {
    static const int f[6][4][3] = {{{-1, -1, -1}, {-1, 1, -1}, {1, 1, -1}, {1, -1, -1}},
                                   {{-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1}},
                                   {{-1, -1, -1}, {1, -1, -1}, {1, -1, 1}, {-1, -1, 1}},
                                   {{-1, 1, -1}, {-1, 1, 1}, {1, 1, 1}, {1, 1, -1}},
                                   {{-1, -1, -1}, {-1, -1, 1}, {-1, 1, 1}, {-1, 1, -1}},
                                   {{1, -1, -1}, {1, 1, -1}, {1, 1, 1}, {1, -1, 1}}};
    const int q[2][3] = {{0, 1, 2}, {0, 2, 3}};
    for (int k = 0; k < 6; ++k)
      for (int t = 0; t < 2; ++t) {
        for (int v = 0; v < 3; ++v)
          for (int c = 0; c < 3; ++c) {
            closed[9 * nc + 3 * v + c] = 0.5f * (float)f[k][q[t][v]][c];
            if (k != 1) open[9 * no + 3 * v + c] = 0.37f * (float)f[k][q[t][v]][c] + 0.011f;
          }
        ++nc;
        if (k != 1) ++no;
      }
  }
  const float bmin[3] = {-0.5f, -0.5f, -0.5f}, bmax[3] = {0.5f, 0.5f, 0.5f};

  /* framing, occupancy, voxelize */
  float origin[3], h;
  int n[3];
  orc_sdf_framing(bmin, bmax, 16, origin, &h, n);
  CHECK(n[0] > 16 && h > 0.0f);
  float vo[3], vh;
  orc_vox_framing(bmin, bmax, 12, vo, &vh);
  const int nw = (n[0] + 31) / 32;
  uint32_t *words = (uint32_t *)calloc((size_t)nw * n[1] * n[2], 4);
  orc_occupancy(closed, nc, origin, h, n[0], n[1], n[2], words);
  free(words);
  float colors[12 * 3];
  for (size_t i = 0; i < 3 * nc; ++i) colors[i] = (float)i / 36.0f;
  const size_t ncell = orc_voxelize_first_touch(closed, nc, colors, bmin, bmax, 12, NULL, NULL, 0);
  int32_t *cells = (int32_t *)malloc(3 * ncell * sizeof(int32_t));
  float *ccol = (float *)malloc(3 * ncell * sizeof(float));
  CHECK(orc_voxelize_first_touch(closed, nc, colors, bmin, bmax, 12, cells, ccol, ncell) == ncell);
  free(cells);
  free(ccol);

  /* dense bake of the closed box (threads), and its refusal of the open one */
  const size_t nv = (size_t)n[0] * n[1] * n[2];
  float *dist = (float *)malloc(nv * sizeof(float));
  orc_bake_stats bs;
  CHECK(orc_bake_dense(closed, nc, origin, h, n[0], n[1], n[2], 0, n[2], 3, dist, &bs) == 0);
  const size_t centre = ((size_t)(n[2] / 2) * n[1] + n[1] / 2) * n[0] + n[0] / 2;
  CHECK(dist[centre] < 0.0f && dist[0] > 0.0f);
  CHECK(orc_bake_dense_strided(closed, nc, origin, h, n[0], n[1], n[2], 0, n[2], 3, 2, dist, &bs) == 0);
  CHECK(orc_bake_dense(open, no, origin, h, n[0], n[1], n[2], 0, n[2], 2, dist, &bs) == -1);
  free(dist);

  /* narrow band of the open box (threads) */
  {
    const int g = 32;
    const float org[3] = {-0.8f, -0.8f, -0.8f}, hh = 1.6f / g;
    const size_t nb = (size_t)(g / 8) * (g / 8) * (g / 8);
    int32_t *idx = (int32_t *)malloc(nb * sizeof(int32_t));
    float *pool = (float *)malloc(nb * 512 * sizeof(float));
    size_t nbricks = 0;
    orc_band_stats st;
    CHECK(orc_bake_band(open, no, org, hh, g, g, g, 2, 1, 3, idx, pool, nb, &nbricks, &st) == 0);
    CHECK(nbricks > 0 && st.cap_triangles > 0);
    free(idx);
    free(pool);
  }

  /* points: FP32/FP64 distances, winding numbers, parity */
  enum { NP = 64 };
  float pts[3 * NP];
  double dpts[3 * NP];
  for (int i = 0; i < NP; ++i)
    for (int c = 0; c < 3; ++c) {
      pts[3 * i + c] = -0.7f + 1.4f * (float)((i * 7 + c * 13) % 64) / 63.0f;
      dpts[3 * i + c] = pts[3 * i + c];
    }
  float fout[NP], f2[NP];
  double d64[NP], w64[NP];
  uint8_t tie[NP];
  CHECK(orc_signed_points_gwn(open, no, pts, NP, 3, fout, tie) == 0);
  CHECK(orc_signed_points_closed(closed, nc, pts, NP, 3, fout) == 0);
  CHECK(orc_signed_points_closed(open, no, pts, NP, 3, fout) == -1);
  CHECK(orc_udf_points_f64(closed, nc, dpts, NP, 3, d64) == 0);
  CHECK(orc_udf_points_f32(closed, nc, pts, NP, 3, fout) == 0);
  orc_udf2_f32_brute(closed, nc, pts, NP, f2);
  for (int i = 0; i < NP; ++i) CHECK(fout[i] * fout[i] >= 0.0f && f2[i] >= 0.0f);
  orc_sdf_f64_points(closed, nc, dpts, NP, 3, d64, w64);
  {
    double wd[NP];
    uint8_t ex[NP];
    CHECK(orc_gwn_f64_points(closed, nc, dpts, NP, 3, wd, ex) == 0);
    for (int i = 0; i < NP; ++i) CHECK((fabs(wd[i]) > 0.5) == (fabs(w64[i]) > 0.5));
    CHECK(orc_gwn_f64_points(open, no, dpts, NP, 3, wd, ex) == 0);
    CHECK(orc_gwn_f64_points(open, 0, dpts, NP, 1, wd, ex) == 0);
  }
  const double q0[3] = {0.01, 0.02, 0.03};
  CHECK(fabs(orc_gwn_f64(q0, closed, nc)) > 0.99);
  const float qf[3] = {0.01f, 0.02f, 0.03f};
  CHECK(orc_crossings_up(qf, closed, nc) != 0);
  CHECK(orc_parity_inside(qf, closed, nc) == 1);
  int hits[5], t1 = 0;
  orc_parity_hits(qf, closed, nc, hits);
  CHECK(orc_gwn_sign(qf, open, no, &t1) == 1 || t1 == 0);
  float dirs[15];
  orc_parity_dirs(dirs);
  CHECK(orc_boundary_edge_count(closed, nc) == 0 && orc_boundary_edge_count(open, no) == 4);

  /* records, LBVH */
  orc_trirec r;
  orc_trirec_make(closed, &r);
  CHECK(orc_trirec_dist2(qf, &r) >= 0.0f);
  CHECK(orc_tri_dist2_f32(qf, closed) >= 0.0f && orc_tri_dist2_f64(q0, closed) >= 0.0);
  CHECK(orc_tri_is_sliver(closed) == 0);
  orc_bvh b;
  CHECK(orc_bvh_build(closed, nc, &b) == 0);
  orc_bvh_free(&b);
  uint32_t codes[12];
  orc_morton_codes(closed, nc, codes);
  double ratio[2];
  orc_f32_error_ratio(2000, 7, ratio);

  /* grid consumer on a small analytic grid */
  {
    const int g = 12;
    const float org[3] = {-1.0f, -1.0f, -1.0f}, hh = 2.0f / g;
    float *grid = (float *)malloc((size_t)g * g * g * sizeof(float));
    for (int z = 0; z < g; ++z)
      for (int y = 0; y < g; ++y)
        for (int x = 0; x < g; ++x) {
          const float p[3] = {org[0] + (x + 0.5f) * hh, org[1] + (y + 0.5f) * hh, org[2] + (z + 0.5f) * hh};
          grid[((size_t)z * g + y) * g + x] = sqrtf(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) - 0.5f;
        }
    float vals[NP], nrm[3 * NP], th[NP];
    int32_t near[NP];
    orc_sample_grid(org, hh, g, g, g, grid, pts, NP, vals, nrm, near);
    float dirs2[3 * NP];
    for (int i = 0; i < NP; ++i) dirs2[3 * i] = 1.0f, dirs2[3 * i + 1] = 0.0f, dirs2[3 * i + 2] = 0.0f;
    orc_trace_grid(org, hh, g, g, g, grid, pts, dirs2, NP, 0.0f, 4.0f, th);
    free(grid);
  }

  /* analytic CSG: a sphere minus a box, baked on a small grid (threads), and the cup restatement */
  {
    orc_csg_op prog[3];
    memset(prog, 0, sizeof(prog));
    prog[0].op = 1;  /* sphere */
    prog[0].a[3] = 0.5f;
    prog[1].op = 2;  /* box */
    prog[1].a[3] = prog[1].a[4] = prog[1].a[5] = 0.3f;
    prog[2].op = 11; /* subtract */
    int br = 0;
    const float p[3] = {0.0f, 0.0f, 0.45f};
    (void)orc_csg_eval(prog, 3, p, &br);
    float og[3], hh;
    int32_t ng[3];
    orc_analytic_framing(bmin, bmax, 16, og, &hh, ng);
    float *cd = (float *)malloc((size_t)ng[0] * ng[1] * ng[2] * sizeof(float));
    uint8_t *cb = (uint8_t *)malloc((size_t)ng[0] * ng[1] * ng[2]);
    orc_bake_csg(prog, 3, og, hh, ng[0], ng[1], ng[2], 3, cd, cb);
    free(cd);
    free(cb);
    float co[3], ch;
    int32_t cn[3];
    orc_cup_cells(1, 24, 0.0f, 0.0f, 0.0f, co, &ch, cn, NULL, NULL);
    float *cupd = (float *)malloc((size_t)cn[0] * cn[1] * cn[2] * sizeof(float));
    uint8_t *cupb = (uint8_t *)malloc((size_t)cn[0] * cn[1] * cn[2]);
    orc_cup_cells(1, 24, 0.0f, 0.0f, 0.0f, co, &ch, cn, cupd, cupb);
    free(cupd);
    free(cupb);
  }

  /* ingestion gate */
  {
    uint32_t welded[12 * 3];
    float uniq[12 * 9];
    CHECK(orc_weld_positions(closed, 3 * nc, bmin, bmax, welded, uniq) == 8);
    int64_t bnd = 0, nm = 0;
    CHECK(orc_watertight(closed, nc, bmin, bmax, &bnd, &nm) == 1);
    CHECK(orc_watertight(open, no, bmin, bmax, &bnd, &nm) == 0 && bnd == 4);
  }

  /* per-triangle materials and the per-cell splat */
  {
    uint8_t texels[2 * 2 * 4];
    for (int i = 0; i < 16; ++i) texels[i] = (uint8_t)(17 * i);
    const orc_texture tex = {texels, 2, 2, 1};
    const orc_material mats[2] = {{{1.0f, 0.5f, 0.25f}, 0.1f, 0.9f, 0}, {{0.2f, 0.4f, 0.6f}, 0.0f, 0.5f, -1}};
    float uvs[12 * 3 * 2];
    for (int i = 0; i < 72; ++i) uvs[i] = (float)i * 0.03f - 0.4f;
    uint32_t tm[12];
    for (int i = 0; i < 12; ++i) tm[i] = (uint32_t)(i % 3);
    float mat[12 * 5];
    orc_triangle_materials(nc, uvs, 36, tm, 12, mats, 2, &tex, 1, mat);
    float *planes = (float *)malloc(5 * (size_t)n[0] * n[1] * n[2] * sizeof(float));
    orc_splat_materials(closed, nc, mat, origin, h, n[0], n[1], n[2], planes);
    free(planes);
  }
  if (failures) {
    fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  printf("oracle drive: ok\n");
  return 0;
}
