/* Proxy decimation of a closing mesh: C-ABI of libgmorpho_decimate.so (host).
 *
 * Replaces the reference's Decimator (Decimator/Decimator.h:36-78,
 * Decimator.cpp): HC Laplacian smoothing, then quadric-error edge collapses
 * whose new vertex minimizes the summed plane quadric under linear
 * constraints that keep it on the outer side of every face around the edge
 * (the bounding proxy), with a normal-flip check, a triangle-quality penalty
 * and a maximum edge length read from the ScaleField grid. The reference
 * builds this on VCG's LocalOptimization and a Fortran QP solver, neither of
 * which is in its tree; here the collapse loop, the quadrics, the
 * eigen-regularized placement and the dual active-set QP are written from the
 * published methods (Garland-Heckbert quadrics, Goldfarb-Idnani QP, Vollmer
 * et al. HC smoothing).
 *
 * Conventions: 0 = success, otherwise a GM_ERR_* value of gmorpho_c.h;
 * gmd_last_error gives the message. Host pointers are borrowed for the call.
 */
#ifndef GMORPHO_DECIMATE_H
#define GMORPHO_DECIMATE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct gmd_session gmd_session;

/* Decimator::Decimator (Decimator.cpp:1268-1290; session set-up 1096-1194):
 * copies the mesh (P float[3*nv], T uint32[3*nt], a closed triangle mesh),
 * smooths it with num_hc_iters HC iterations and fills the collapse heap.
 * Each edge's collapse is penalized (priority 1e11) when its length exceeds
 * max_edge_length * min(s(v0), s(v1)) * cell_size, s(v) the scale-grid value
 * of the voxel holding v (grid of res[0] x res[1] x res[2] signed chars, x
 * fastest, voxel i at bbox_min + cell_size * i; NULL: no length limit).
 * use_linear_constraints: the bounding constraints of the placement. */
int gmd_create(const float* P, int nv, const uint32_t* T, int nt, int target_num_faces, float target_error,
               int use_linear_constraints, float max_edge_length, const float bbox_min[3], const int res[3],
               float cell_size, const int8_t* scale_grid, int num_hc_iters, gmd_session** out);
int gmd_destroy(gmd_session* s);
const char* gmd_last_error(const gmd_session* s);
/* Decimator::Optimize (Decimator.cpp:1292-1297, 1196-1201): collapses the
 * cheapest feasible edges until the face count reaches the target, the next
 * priority exceeds the target error, the heap is empty or time_budget_s
 * seconds have passed (<= 0: no limit). heap_not_empty (optional) tells
 * whether more collapses remain. */
int gmd_optimize(gmd_session* s, double time_budget_s, int* heap_not_empty);
/* num_faces(), error() (Decimator.h:56-57) and the vertex count. */
int gmd_state(const gmd_session* s, int* num_faces, int* num_vertices, float* error);
/* Decimator::GetMesh (Decimator.cpp:1203-1239): the live vertices in order
 * and the live faces re-indexed; P float[3*num_vertices], T uint32[3*num_faces]. */
int gmd_get_mesh(const gmd_session* s, float* P, uint32_t* T);

#ifdef __cplusplus
}
#endif
#endif
