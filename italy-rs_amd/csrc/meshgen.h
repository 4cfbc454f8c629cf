/*
 * Seeded synthetic mesh generators (triangle soups, 9 floats per triangle) for the bake
 * workloads and for the two reference test meshes.  Host-only C ABI, no GPU code.
 *
 *   ib_gen_uv_sphere    the UV sphere of tests/sdf_sign_test.cpp:18-42 (same float expressions)
 *   ib_gen_cube_shell   the unit cube shell of tests/voxelize_test.cpp:11-33
 *   ib_gen_icosphere    icosphere, optional seeded radial jitter of the welded vertices
 *   ib_gen_displaced    icosphere with the 8-mode seeded radial displacement
 *   ib_gen_decimated    displaced icosphere decimated by random edge collapses to a target count
 *   ib_gen_stress       union of seeded tori and subdivided boxes with random triangles removed
 */
#ifndef ITALY_MESHGEN_H
#define ITALY_MESHGEN_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  float *xyz;    /* 9 * ntri floats, owned by the library */
  int64_t ntri;
  float bmin[3]; /* exact bounds of the soup */
  float bmax[3];
} ib_soup;

int ib_gen_uv_sphere(float r, int lon_segs, int lat_segs, ib_soup *out);
int ib_gen_cube_shell(ib_soup *out);
int ib_gen_icosphere(int level, double radius, double jitter, uint64_t seed, ib_soup *out);
int ib_gen_displaced(int level, uint64_t seed, ib_soup *out);
int ib_gen_decimated(int64_t target_tris, uint64_t seed, ib_soup *out);
/* ncomp components (even index torus, odd index box), remove_frac of all triangles removed */
int ib_gen_stress(uint64_t seed, int ncomp, double remove_frac, ib_soup *out);
/* same with the tori's segment counts (nu around the ring, nv around the tube) and the boxes' quads per face edge
 * given (ib_gen_stress uses 64, 32 and 32) */
int ib_gen_stress_res(uint64_t seed, int ncomp, double remove_frac, int torus_nu, int torus_nv, int box_n,
                      ib_soup *out);
/* n decimated meshes generated on `threads` host threads (<= 0: all cores) */
int ib_gen_decimated_batch(int n, const int64_t *targets, const uint64_t *seeds, int threads, ib_soup *outs);
void ib_soup_free(ib_soup *s);

#ifdef __cplusplus
}
#endif
#endif
