/* Seeded synthetic inputs for the GMorpho closing workloads (host only).
 *
 * These have no counterpart in the reference, which loads meshes from OFF files
 * (Common/Mesh.h:602) inside its Qt viewer. They generate the meshes and
 * ScaleField lattices the benchmarks are quoted on; buffers are caller-owned
 * and sized with gmg_icosphere_counts (a soup of n spheres has n times the
 * icosphere counts). All functions return 0 on success, -1 on bad arguments.
 */
#ifndef GMORPHO_MESHGEN_H
#define GMORPHO_MESHGEN_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

int gmg_icosphere_counts(int subdiv, int* nv, int* nf);
/* Unit icosphere, outward counter-clockwise faces. */
int gmg_icosphere(int subdiv, float* P, uint32_t* T);
/* Icosphere radially displaced by amp * sum_{l<=lmax,|m|<=l} a_lm Y_lm with
 * a_lm ~ N(0, 1/(l+1)^2), then scaled into the unit cube. */
int gmg_sh_blob(int subdiv, uint64_t seed, int lmax, double amp, float* P, uint32_t* T);
/* nspheres icospheres, centres U[cmin,cmax]^3, radii U[rmin,rmax], then scaled
 * into the unit cube. Faces are concatenated: overlapping spheres are kept as
 * separate shells. */
int gmg_sphere_soup(int nspheres, int subdiv, uint64_t seed, double cmin, double cmax, double rmin,
                    double rmax, float* P, uint32_t* T);
/* L^3 FP32 lattice (x fastest) of U[lo,hi] values, 3x3x3 box-smoothed
 * smooth_passes times with clamped borders. */
int gmg_scale_lattice(int L, uint64_t seed, double lo, double hi, int smooth_passes, float* out);

#ifdef __cplusplus
}
#endif
#endif
