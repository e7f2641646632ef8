/* GMorpho closing path on MI355X: C-ABI boundary of libgmorpho_hip.so.
 *
 * Each entry point replaces a method of the reference's C++ classes
 * (MorphoGraphics::GMorpho, Grid, Voxelizer, ScaleField, MarchingCubesMesher);
 * the reference file:line it stands for is given beside it. The C++ facade in
 * include/gmorpho.hpp wraps these calls under the reference's class and method
 * names.
 *
 * Conventions
 *  - Every function returns GM_OK (0) or a GM_ERR_* code; gm_last_error()
 *    gives the message ("GMorpho Error: ..." as GMorpho::Exception formats it,
 *    GMorpho.h:52). No C++ exception crosses this boundary.
 *  - Host pointers are borrowed for the duration of the call only. Device
 *    memory belongs to the context and is released by gm_destroy.
 *  - One context = one HIP stream on one device. Calls are synchronous with
 *    respect to the host unless noted ("enqueue"); contexts are
 *    thread-compatible, not thread-safe.
 *  - Grids cross the boundary in the reference's cell-byte layout: R^3 bytes,
 *    index x + R*(y + R*z), bit k = fine voxel (2x+(k&1), 2y+((k>>1)&1),
 *    2z+((k>>2)&1)) (Grid.h:61-68). N = 2R fine voxels per axis; R must be a
 *    multiple of 16 (16 <= R <= 1024).
 *  - Radii are in fine voxels.
 */
#ifndef GMORPHO_C_H
#define GMORPHO_C_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct gm_ctx gm_ctx;

enum {
  GM_OK = 0,
  GM_ERR_ARG = 1,         /* invalid argument */
  GM_ERR_STATE = 2,       /* call out of order (e.g. no grid loaded) */
  GM_ERR_HIP = 3,         /* HIP runtime failure */
  GM_ERR_CAPACITY = 4,    /* output does not fit the caller's limits */
  GM_ERR_UNSUPPORTED = 5  /* parameter outside what this build supports */
};

/* Grids owned by a context (GMorpho.h:138-141). */
enum { GM_GRID_INPUT = 0, GM_GRID_DILATION = 1, GM_GRID_CONTOUR = 2, GM_GRID_CLOSING = 3 };
/* Structuring element of the dilation (use_asymmetric_closing_, GMorpho.h:69). */
enum { GM_SE_SPHERE = 0, GM_SE_CUBE = 1 };

/* GMorpho::GMorpho / ~GMorpho (GMorpho.cu:76-101). */
int gm_create(int device, gm_ctx** out);
int gm_destroy(gm_ctx* ctx);
/* Message of the last failing call on ctx ("" if none). */
const char* gm_last_error(const gm_ctx* ctx);

/* GMorpho::Load (GMorpho.cu:7731-7773, GMorpho.h:76-78): copy the mesh,
 * voxelize it at base_res cells per axis with margin se_size + 1/base_res,
 * and set up the grids. N (normals) may be NULL; it is not used. */
int gm_load(gm_ctx* ctx, const float* P, const float* N, int num_of_vertices, const uint32_t* T,
            int num_of_faces, int base_res, float se_size);
/* Voxelizer::Load (Voxelizer.cu:74-102). */
int gm_load_mesh(gm_ctx* ctx, const float* P, const float* N, int num_of_vertices, const uint32_t* T,
                 int num_of_faces);
/* Voxelizer::VoxelizeConservative (Voxelizer.cu:724-796) with
 * ComputeGridAttributes (431-455): fills GM_GRID_INPUT. */
int gm_voxelize(gm_ctx* ctx, int base_res, float margin);
/* Voxelizer::VoxelizeByMultiSlicing (Voxelizer.cu:375-400, kernel 188-331;
 * the alternative GMorpho::Load leaves commented out, GMorpho.cu:7741): the
 * same grid attributes, GM_GRID_INPUT from the parity of rays along x
 * through the fine column centres, without the conservative shell. */
int gm_voxelize_multislice(gm_ctx* ctx, int base_res, float margin);

/* Grid::Init (Grid.cu:107-131) for a caller-supplied input grid: bounding
 * box, base resolution (cells per axis), data resolution (cells of the data
 * region; coarse passes only visit cells below it) and cell size. */
int gm_grid_init(gm_ctx* ctx, const float bbox_min[3], const float bbox_max[3], int base_res,
                 const int data_res[3], float cell_size);
/* Grid::CopyVoxelsFrom / voxels() transfer (Grid.cu:139-151): cell bytes in/out. */
int gm_grid_set_cells(gm_ctx* ctx, int which, const uint8_t* cells);
int gm_grid_get_cells(gm_ctx* ctx, int which, uint8_t* cells);
/* GMorpho::bbox() / res() / cell_size() (GMorpho.h:63-65) and Grid::data_res(). */
int gm_grid_info(const gm_ctx* ctx, float bbox_min[3], float bbox_max[3], int res[3], int data_res[3],
                 float* cell_size);

/* ScaleField (ScaleField.h:33-140). Per-voxel radius r(p) in fine voxels, as
 * the max-mipmap level 0 stores it (Grid.cu:957-998):
 *  - uniform: every voxel has radius r (ScaleField::UpdateGridGlobalScale,
 *    ScaleField.cu:143; the fixed-size path of Grid.cu:725). r <= 1 makes
 *    gm_close skip the morphology (GMorpho.cu:7784, 7803).
 *  - int8 grid: N^3 signed chars (ScaleField::scale_grid(), x fastest),
 *    clamped to [2,127] per voxel.
 *  - FP32 lattice: L^3 values (x fastest), N/L a power of two; node i sits on
 *    fine voxel i*N/L; values are rounded to 1/256, sampled trilinearly in
 *    fixed point with clamped borders, floored, then clamped to [2,127]. */
int gm_scale_uniform(gm_ctx* ctx, int r_fine);
int gm_scale_grid_i8(gm_ctx* ctx, const int8_t* grid);
int gm_scale_lattice_f32(gm_ctx* ctx, const float* lattice, int L);

/* ScaleField editing on the device grid of signed chars (the reference's
 * scale_grid_, ScaleField.h:135). The scale field's voxel i has its sample
 * point at bbox_min + (cell_size / 2) * i (ScaleField::Init with twice the
 * grid's resolution and half its cell, BroxyViewer.cpp:133-136).
 *  - fill: ScaleField::UpdateGridGlobalScale (ScaleField.cu:143-146), every
 *    voxel set to value (the reference memsets floor(global_scale /
 *    cell_size)); the grid becomes the context's radius field.
 *  - brush: ScaleField::AddPoint + UpdateGrid (ScaleField.h:102-109,
 *    ScaleField.cu:148-192) for one point, kernel UpdateScaleGrid
 *    (ScaleField.cu:81-141) with the tukey falloff (cuda_math.h:696-704):
 *    radius `scale` and support `support` in world units, edit mode
 *    GM_BRUSH_* (EditMode, ScaleField.h:32). Only voxels closer than
 *    scale + h are visited; the radius planes of a per-voxel closing are
 *    rebuilt for the touched bricks only. A context holding a uniform radius
 *    is first filled with it.
 *  - get: the grid back to the host (ScaleField::GetScaleGridCPU,
 *    ScaleField.cu:74-79). */
enum { GM_BRUSH_NONE = 0, GM_BRUSH_MAX = 1, GM_BRUSH_MIN = 2, GM_BRUSH_ZERO = 3 };
int gm_scale_grid_fill(gm_ctx* ctx, int value);
int gm_scale_brush(gm_ctx* ctx, const float position[3], float scale, float support, float global_scale, int mode);
int gm_scale_grid_get(gm_ctx* ctx, int8_t* grid);

/* GMorpho::DilateBySphereMipmap / DilateByCubeMipmap (GMorpho.cu:3538, 3231):
 * GM_GRID_DILATION from GM_GRID_INPUT. Ball radii up to 127 (radii above 31
 * run separable distance-transform passes); cube radii up to 31. */
int gm_dilate(gm_ctx* ctx, int se);
/* GMorpho::ErodeBySphereMipmap (GMorpho.cu:6627): GM_GRID_CONTOUR and
 * GM_GRID_CLOSING from GM_GRID_DILATION. */
int gm_erode(gm_ctx* ctx);
/* The morphology of GMorpho::Update (GMorpho.cu:7775-7806): dilate then
 * erode, both skipped for a uniform radius <= 1 (closing = input). */
int gm_close(gm_ctx* ctx, int se);
/* GMorpho::ExtractClosingMesh (GMorpho.cu:6814-7001) on GM_GRID_CLOSING:
 * vertices, area-weighted unit normals and triangles stay on the device.
 * Enqueue only; gm_mesh_counts waits for it. */
int gm_extract_mesh(gm_ctx* ctx);
/* GMorpho::Update (GMorpho.cu:7775-7817) up to the mesh: gm_close(ctx, se)
 * then gm_extract_mesh(ctx), with identical results, enqueued as one
 * sequence (with graphs on, one captured launch per call instead of two).
 * Enqueue only; gm_mesh_counts waits for it. The normal filtering the
 * reference ends with is gm_filter_normals. */
int gm_update(gm_ctx* ctx, int se);
/* Sizes of the extracted mesh (waits for the extraction). */
int gm_mesh_counts(gm_ctx* ctx, uint32_t* num_vertices, uint32_t* num_triangles);
/* Copy the mesh out (GMorpho.cu:6985-6995): V, Nrm float[3*nv], T uint32[3*nt]. */
int gm_mesh_copy(gm_ctx* ctx, float* V, float* Nrm, uint32_t* T);

/* The bilateral normal filtering of GMorpho::ExtractClosingMesh
 * (GMorpho.cu:6946-6978, kernels BilateralMeshFiltering and
 * BilateralMeshFilteringNormalize 6539-6620): num_iters iterations on the
 * extracted mesh's unit normals, sigma_p = 1, sigma_n = 1e10 for the first
 * two iterations and 0.2 after (GMorpho::Update passes 200 iterations for the
 * asymmetric closing and 20 otherwise, GMorpho.cu:7809-7812). Waits for the
 * extraction. Each vertex adds its triangles' terms in triangle order, so the
 * result is the same every run (the reference's float atomics are not). */
int gm_filter_normals(gm_ctx* ctx, int num_iters);

/* Grid::BuildTexMipmaps (Grid.cu:1000-1154) of grid `which` with the current
 * radius field: levels 0..log2(R), level L = (R>>L)^3 64-bit texels (x
 * fastest) concatenated. R must be a power of two. */
int gm_build_pyramid(gm_ctx* ctx, int which);
int gm_pyramid_size(const gm_ctx* ctx, int* levels, uint64_t* texels);
int gm_pyramid_get(gm_ctx* ctx, uint64_t* texels);

/* Launch-shape tuning of the streaming dilation and erosion passes. No
 * counterpart in the reference (its kernels have fixed block shapes).
 * gm_autotune times gm_close on the context's current grid and radius with
 * each candidate shape (planes per tile 32 or 16; at most 8, 4 or 2 waves per
 * workgroup), checks that every shape gives the same closing, and records the
 * fastest for (base resolution, uniform radius or -1 for per-voxel radii,
 * se) in a table of the process, which every context's gm_close / gm_dilate
 * / gm_erode then use; with path it also writes the table (one JSON object a
 * line). gm_tune_load reads such a file into the table. Batched closings use
 * the default shape. */
int gm_autotune(gm_ctx* ctx, int se, int reps, const char* path, int* tile_planes, int* waves);
int gm_tune_load(const char* path);

/* Per-stage device timings of the last gm_close / gm_extract_mesh, in ms,
 * from HIP events on the context stream (the reference prints GET_TIME
 * brackets, GMorpho.cu:7796-7815). 16 stages: brick summary (with the
 * radius planes on the first closing after a per-voxel field is set),
 * dilate coarse, tight contour, dilate fine, contour K, erode coarse, base
 * contour, erode fine, mc count, mc scan (one small kernel), mc vertices and
 * triangles, mc finish (late triangles and normalization when a cube lists
 * more triangles than it counts; the emit normalizes otherwise); then the
 * streaming launches of the four
 * passes alone (dilate coarse, dilate fine, erode coarse, erode fine). A
 * pass stage holds its tile selection and its streaming launches (one per
 * radius value with per-voxel radii). Timing is off unless enabled. */
int gm_set_timing(gm_ctx* ctx, int enable);
/* gm_close and gm_extract_mesh replay their launch sequence as a captured
 * HIP graph once it was captured for the context's current buffers, sizes,
 * radius field and grid attributes. mode 1 (the default) captures on the
 * first call of a sequence; mode 2 runs a sequence eagerly the first time
 * and captures it when it comes again unchanged (for contexts that run a
 * different job every call, where a capture would never be replayed);
 * mode 0 runs every launch eagerly. Per-stage timing also runs eagerly.
 * No counterpart in the reference, which launches every kernel from the
 * host. */
int gm_set_graphs(gm_ctx* ctx, int mode);
int gm_stage_times(gm_ctx* ctx, float* ms, int n);
/* Wait for all work enqueued on the context. */
int gm_sync(gm_ctx* ctx);
/* The context's HIP stream (hipStream_t), for callers that enqueue around it. */
int gm_get_stream(gm_ctx* ctx, void** stream);

/* Diagnostics, no counterpart in the reference: `reps` device copies of
 * `bytes` (a multiple of 16) between two fresh buffers by a plain 16-byte
 * per lane copy kernel, the known-size reference for HBM byte counters. */
int gm_diag_copy_probe(int device, uint64_t bytes, int reps);

/* The marching-cubes tables in use: per case the counted triangles, the listed
 * triangles and the packed edge ids (nibble 3q+i = edge i of triangle q, 0xf
 * terminates). Host only. */
int gm_mc_tables(uint32_t* ntri256, uint32_t* len256, uint64_t* tri256);

/* ---- Batched closing. No counterpart in the reference (one closing per
 * GMorpho::Update, GMorpho.cu:7775); for callers with many grids at once.
 * gm_batch_close_mesh does, for each context, what gm_close(ctx, se)
 * followed by gm_extract_mesh(ctx) does, with identical results: contexts
 * with a uniform radius in [2, 31] and the sphere run together, one kernel
 * launch per step for all of them on the batch's stream; the others run on
 * their own streams. The contexts must be on the batch's device and hold a
 * grid and a radius; it returns when every mesh is ready for gm_mesh_counts
 * and gm_mesh_copy. One batch runs one call at a time. */
typedef struct gm_batch gm_batch;
int gm_batch_create(int device, gm_batch** out);
int gm_batch_destroy(gm_batch* b);
const char* gm_batch_last_error(const gm_batch* b);
int gm_batch_close_mesh(gm_batch* b, gm_ctx* const* ctxs, int n, int se);
/* The batched launches are captured as a HIP graph and replayed when the same
 * contexts (with unchanged buffers, sizes and radii) are closed again; the
 * modes are those of gm_set_graphs (1 by default). */
int gm_batch_set_graphs(gm_batch* b, int mode);

/* ---- Allocation-free closing. No counterpart in the reference; for callers
 * that cannot allocate device memory: code under HIP graph capture, serving
 * loops and harnesses that forbid hipMalloc. gm_close_cells_noalloc does what
 * gm_close does for a uniform ball of radius `radius` (GMorpho.cu:7775-7806;
 * radius <= 1 copies the input), on grids the caller owns: cells_in and
 * cells_out are device buffers of base_res^3 bytes in the cell-byte layout
 * above (16-byte aligned; they may not overlap), data_res as in
 * gm_grid_init. It enqueues on `stream` (a hipStream_t, NULL for the null
 * stream) and returns; the result is ready when the stream reaches it. Every
 * intermediate buffer lives in one static device array of
 * GM_NOALLOC_SCRATCH_MIB MiB per device, held until the process ends, so
 * base_res is limited to GM_NOALLOC_MAX_RES (GM_ERR_CAPACITY beyond) and
 * radius to [0, 31]. Calls on one device share that array: calls that overlap
 * on one device (on different streams or host threads) race and must be
 * serialized by the caller. gm_noalloc_init takes the array's device address
 * (optional: the first call on a device does it); gm_noalloc_last_error
 * gives the message of the calling thread's last failure. */
#define GM_NOALLOC_MAX_RES 128
#define GM_NOALLOC_SCRATCH_MIB 10
int gm_noalloc_init(int device);
int gm_close_cells_noalloc(int device, const uint8_t* cells_in, uint8_t* cells_out, int base_res,
                           const int data_res[3], int radius, void* stream);
const char* gm_noalloc_last_error(void);

/* ---- Job engine. No counterpart in the reference, which runs one closing
 * at a time from its viewer (GMorpho::Update, GMorpho.cu:7775); it is for
 * callers with many jobs: batches of uneven grids and radii, and a serving
 * loop. A pool of `slots` contexts (one HIP stream and one host thread each)
 * loads (voxelizes) jobs concurrently on one device; the closings and meshes
 * of the loaded small grids (base_res <= 192) run together through
 * gm_batch_close_mesh (continuous batching: jobs loaded while a batch runs
 * join the next one), larger grids on their own slot's stream.
 * A job is admitted when a slot is free and the device memory the slot would
 * then hold fits hbm_budget_bytes; smaller jobs may pass one that does not
 * fit yet. Slots keep their largest buffers. Every job's mesh is identical to
 * running it alone. */
typedef struct gm_engine gm_engine;
typedef struct {
  const float* P;          /* float[3*num_of_vertices] */
  int num_of_vertices;
  const uint32_t* T;       /* uint32[3*num_of_faces] */
  int num_of_faces;
  int base_res;            /* cells per axis (multiple of 16, 16..1024) */
  float se_size;           /* the margin parameter of gm_load */
  int radius;              /* uniform radius in fine voxels (lattice == NULL) */
  const float* lattice;    /* optional FP32 radius lattice, lattice_res^3 */
  int lattice_res;
  int se;                  /* GM_SE_SPHERE or GM_SE_CUBE */
} gm_job;
int gm_engine_create(int device, int slots, uint64_t hbm_budget_bytes, gm_engine** out);
/* Running jobs finish; queued ones are dropped. */
int gm_engine_destroy(gm_engine* e);
const char* gm_engine_last_error(const gm_engine* e);
/* Queue a job; returns at once with its id. The inputs are copied. */
int gm_engine_submit(gm_engine* e, const gm_job* job, uint64_t* id);
/* Queue a job without copying its inputs: P, T and lattice must stay valid
 * and unchanged until gm_engine_wait, gm_engine_view or gm_engine_take
 * returns for the job (or gm_engine_destroy does). */
int gm_engine_submit_borrowed(gm_engine* e, const gm_job* job, uint64_t* id);
/* Wait for a job; returns its status and mesh sizes. */
int gm_engine_wait(gm_engine* e, uint64_t id, uint32_t* num_vertices, uint32_t* num_triangles);
/* Copy a finished job's mesh out (V, Nrm float[3*nv], T uint32[3*nt]) and
 * release the job. */
int gm_engine_take(gm_engine* e, uint64_t id, float* V, float* Nrm, uint32_t* T);
/* Wait for a job and point at its mesh where the engine holds it (pinned
 * host memory the device copied it into, so no host copy is made); the
 * arrays stay valid until gm_engine_release(e, id). */
int gm_engine_view(gm_engine* e, uint64_t id, const float** V, const float** Nrm, const uint32_t** T,
                   uint32_t* num_vertices, uint32_t* num_triangles);
/* Release a job (waits for it first); its id is unknown afterwards. */
int gm_engine_release(gm_engine* e, uint64_t id);
/* Submit, start and completion time of a job (seconds on the engine's clock). */
int gm_engine_job_times(gm_engine* e, uint64_t id, double t[3]);
/* Submit, start, loaded (voxelized, radius set), meshed (counts known) and
 * done (mesh on the host) times of a job, seconds on the engine's clock. */
int gm_engine_job_phases(gm_engine* e, uint64_t id, double t[5]);
/* The engine's clock now (seconds). */
int gm_engine_now(gm_engine* e, double* t);
/* Device bytes reserved by the slots, jobs running, jobs queued. */
int gm_engine_stats(gm_engine* e, uint64_t* reserved_bytes, int* running, int* queued);

#ifdef __cplusplus
}
#endif
#endif
