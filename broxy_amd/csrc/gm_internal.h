// Internal declarations shared by the GMorpho HIP sources.
//
// Device layouts (N = 2R fine voxels per axis, R a multiple of 16):
//  * Bit grid ("rows"): fine voxel (x,y,z) is bit x%32 of u32 word
//    ((z*N + y)*W + x/32), W = N/32. One u32 row word spans 16 cells of the
//    cell row (y/2, z/2) with sub-voxel bits (y&1, z&1).
//  * Cell flags: u16 per (cell row, fine word): ((cz*R + cy)*W + wx), bit j =
//    cell (16*wx + j, cy, cz). Used for the coarse-pass results and for the
//    contour tags that gate the fine passes.
//  * The reference's cell-byte layout (R^3 bytes, index x + R*(y + R*z), bit k =
//    sub-voxel (k&1, (k>>1)&1, (k>>2)&1)) only exists at the C-ABI boundary.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

namespace gm {

typedef uint32_t u32;
typedef uint64_t u64;
typedef uint16_t u16;

// ---------------------------------------------------------------------------
// Radius field. mode 0: uniform r; 1: signed-char grid N^3 (x fastest),
// clamped to [2,127]; 2: fixed-point lattice (L^3 int32 values in 1/256 units,
// node i on fine voxel i*s), trilinear, floor, clamped to [2,127].
struct RadiusField {
  int mode;
  int r;
  int N;
  const int8_t* grid;
  const int32_t* lat;
  int L;
  int s;      // N / L (power of two)
  int lshift; // log2(s)
  int narrow; // s^3 * max |lat| < 2^30: lattice sums fit 32-bit integers
};

__device__ __forceinline__ int radius_at(const RadiusField& f, int x, int y, int z) {
  if (f.mode == 0) return f.r & 0x7f;
  if (f.mode == 1) {
    int v = f.grid[(size_t)x + (size_t)f.N * ((size_t)y + (size_t)f.N * (size_t)z)];
    return v < 2 ? 2 : (v > 127 ? 127 : v);
  }
  int ix = x >> f.lshift, iy = y >> f.lshift, iz = z >> f.lshift;
  long long fx = x - (ix << f.lshift), fy = y - (iy << f.lshift), fz = z - (iz << f.lshift);
  long long s = f.s;
  long long acc = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    int xx = ix + (k & 1), yy = iy + ((k >> 1) & 1), zz = iz + ((k >> 2) & 1);
    xx = xx < f.L - 1 ? xx : f.L - 1;
    yy = yy < f.L - 1 ? yy : f.L - 1;
    zz = zz < f.L - 1 ? zz : f.L - 1;
    long long w = ((k & 1) ? fx : s - fx) * ((k & 2) ? fy : s - fy) * ((k & 4) ? fz : s - fz);
    acc += w * (long long)f.lat[(size_t)xx + (size_t)f.L * ((size_t)yy + (size_t)f.L * (size_t)zz)];
  }
  long long den = s * s * s * 256;
  long long v = acc >= 0 ? acc / den : -((-acc + den - 1) / den);
  v = v < -128 ? -128 : (v > 127 ? 127 : v);
  return v < 2 ? 2 : (int)v;
}

// Many jobs per launch: unit u (a tile, workgroup or list entry) belongs to
// the job j with pre[j] <= u < pre[j+1] (pre: prefix of the per-job counts;
// u is uniform over the caller's wave).
__device__ __forceinline__ int batch_job(const u32* __restrict__ pre, int n, u32 u) {
  int lo = 0, hi = n - 1;  // largest j with pre[j] <= u
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= u) lo = mid;
    else hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// ---------------------------------------------------------------------------
// Structuring-element chord tables. A table lists, per "dz class", the dz
// values it stands for and the (dy class, x half-width m) chords. Two kinds:
//  corner (query at a lattice point P; fine voxel q is at distance
//          delta(q-P) per axis with delta(t) = t (t>=0), -t-1 (t<0)): a class
//          value b stands for offsets {b, -b-1}; window along x is [-(m+1), m];
//  centre (query at a voxel centre; per-axis distance |t|-1/2 for t != 0):
//          class b stands for {b, -b} (just {0} for b = 0); window [-m, m].
// Layout (u32): [0] kind, [1] nclass, [2] max m, [3] max |offset| in y/z,
// then per class: header (b | npairs<<16), npairs words (a | m<<16).
// SE_BOX: the cube's chords (every chord of half-width rho in a
// (2 rho + 2)^3 corner-geometry window), for the compiled-table kernels only;
// the runtime cube tables are SE_CORNER tables built with cube = true.
enum { SE_CORNER = 0, SE_CENTRE = 1, SE_BOX = 2 };
constexpr int kMaxRhoFast = 32;  // streaming kernel supports m, |offset| <= 32

// Launch configuration of the streaming morphology kernels.
constexpr int kTileRows = 64;  // lanes per workgroup (one output row each)
constexpr int kTileZ = 32;     // planes per tile (default)
constexpr int kTileZMin = 16;  // the smallest tile depth a launch may use (queues are sized for it)

struct MorphDims {
  int R, N, W;
  int dR[3];  // data region in cells (clamped to R)
};

// Error plumbing: every host-side helper returns hipError_t.
#define GM_CHECK(x)                          \
  do {                                       \
    hipError_t e_ = (x);                     \
    if (e_ != hipSuccess) return e_;         \
  } while (0)

// ---- layout kernels (gm_layout.hip)
hipError_t launch_cells_to_rows(const uint8_t* cells, u32* rows, int R, hipStream_t s);
hipError_t launch_rows_to_cells(const u32* rows, uint8_t* cells, int R, hipStream_t s);
// bytes: a multiple of 16
hipError_t launch_copy_probe(const void* src, void* dst, size_t bytes, hipStream_t s);

// ---- voxelizer (gm_voxelize.hip)
// Index check and bounding box of a device mesh: out (8 u32, initialised to
// 0, 0, ~0, ~0, ~0, 0, 0, 0) receives the bad-index flag, the non-finite
// flag and order-preserving keys of the per-axis min and max coordinates.
// out: mesh_check_words() u32 (the 8 results first, then per-workgroup parts).
size_t mesh_check_words();
hipError_t launch_mesh_check(const float* P, int nv, const u32* T, int nt, u32* out, hipStream_t s);
// The float a launch_mesh_check key stands for.
inline float mesh_key_float(u32 k) {
  const u32 u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
hipError_t launch_voxelize(const float* P, int nv, const u32* T, int nt, const float* bmin3, double scale,
                           int R, int32_t* snap, u32* mark, u32* shell, u32* rows, hipStream_t s, bool clear);
// MultiSlicing: parity along x only (the marks are cleared by its fill too)
hipError_t launch_voxelize_multislice(const float* P, int nv, const u32* T, int nt, const float* bmin3, double scale,
                                      int R, int32_t* snap, u32* mark, u32* rows, hipStream_t s, bool clear);

// ---- morphology (gm_morph.hip)
struct MorphBuffers {
  u32* I;
  u32* D;
  u32* K;
  u32* C;
  u16* DF;  // coarse dilation: cell full
  u16* TG;  // tight contour of DF
  u16* CL;  // coarse erosion: cell cleared
  u16* UG;  // base contour of the coarse erosion
  u16* IFL; // cells full in I
  u16* DAL; // cells with a voxel of D
  u32* tables;       // device SE tables (concatenated)
  uint8_t* brickI;   // per 1 word x 8 rows x 8 planes brick of I: bit0 any, bit1 all
  uint8_t* brickK;   // same for K (bit0 only)
  u32* queue;        // active-tile queues of the streaming passes
  int queueCap;      // tiles per queue list (the most tiles of any pass)
  int* qsel;         // counter block of the next pass (flips every pass)
  // Per-voxel radii with the sphere (null otherwise): the radius value v of
  // every voxel as 5 bit planes (plane k = bit k, planes one grid apart,
  // each brick-major: the 64 words of a 1 word x 8 rows x 8 planes brick
  // contiguous), the set of values present per brick (u32, bit v), the
  // source grid of one erosion radius (K masked to r = v) and the bricks
  // holding K ([0] count, then brick ids).
  u32* rp;
  u32* rvset;
  u32* kv;
  u32* kbl;
  int rpBuild;  // rebuild rp and rvset from the radius field first
  // launch shape of the streaming passes (0: defaults): planes per tile
  // (kTileZMin or kTileZ) and the most waves per workgroup
  int tileZ;
  int waveCap;
};
// Active-tile queues for base resolution R (u32 words): two blocks of 16
// counters (uniform radius) at word 0, two blocks of 32 x 8 counters (one
// per radius value and XCD queue) at kVarCounters, then from kQueueLists the
// 8 lists of every tile (uniform radius) and the 32 x 8 lists of
// var_queue_cap tiles (per radius value). The counters do not move with R.
constexpr int kVarCounters = 32;
constexpr int kQueueLists = 544;
inline size_t tile_count_max(int R) {
  const size_t N = 2 * (size_t)R, W = N / 32;
  return W * ((N + kTileRows - 1) / kTileRows) * ((N + kTileZMin - 1) / kTileZMin);
}
inline size_t var_queue_base(int R) { return kQueueLists + 8 * tile_count_max(R); }
// Tiles of one per-XCD queue: each run of 8 slabs puts one slab in every queue.
inline size_t var_queue_cap(int R) {
  const size_t N = 2 * (size_t)R, W = N / 32;
  const size_t slabs = ((N + kTileRows - 1) / kTileRows) * ((N + kTileZMin - 1) / kTileZMin);
  return (slabs + 7) / 8 * W;
}
inline size_t tile_queue_words(int R) { return var_queue_base(R) + 256 * var_queue_cap(R); }
// Device words of the per-voxel radius buffers of base resolution R.
inline size_t radius_plane_words(int R) { return 5 * (size_t)R * R * R / 4; }
inline size_t brick_words(int R) { return (size_t)R * R * R / 256; }  // (N/8)^2 * W bricks
struct SETableSet {
  // offsets (in u32) into MorphBuffers::tables for each rho, -1 if absent
  int dil[129];
  int ero_corner[129];
  int ero_centre[129];
};
// ev (optional): 5 stage events; evs (optional): 2 events recorded just
// before the coarse and the fine stream kernels (after their tile selection).
hipError_t launch_dilate(const MorphDims& d, const MorphBuffers& b, const RadiusField& rf, int se,
                         const SETableSet& tabs, int rho_lo, int rho_hi, int tabWords, hipStream_t s,
                         hipEvent_t* ev, hipEvent_t* evs);
// Brick summary of a bit grid of edge N: one byte per 1 word x 8 rows x 8
// planes, bit0 any voxel set, bit1 all set. clear (optional): a u32 counter
// set to zero by the same launch.
hipError_t launch_brick_summary(const u32* G, int N, uint8_t* out, u32* clear, hipStream_t s);
hipError_t launch_erode(const MorphDims& d, const MorphBuffers& b, const RadiusField& rf,
                        const SETableSet& tabs, int rho_lo, int rho_hi, int tabWords, hipStream_t s,
                        hipEvent_t* ev, hipEvent_t* evs);
// Batched closing of uniform-radius sphere jobs (radius r in [2, 31]), all
// jobs on stream s. The caller provides a pinned host blob and a device blob
// of batch_morph_bytes(n) (per-job argument tables and work prefixes), and a
// queue of kQueueLists + 256 * cap words, cap >= batch_queue_cap, whose
// counters start at zero; qsel is the batch's counter-block toggle.
struct BatchMorphJob {
  MorphDims d;
  MorphBuffers b;
  int r;
};
size_t batch_morph_bytes(int n);
// Brick summaries of many grids in one launch (dPre: prefix of (N/8)^2 per
// job); clear0 / clear1 (optional) are zeroed by the job's first workgroup.
struct BrickJob {
  const u32* G;
  uint8_t* out;
  u32* clear0;
  u32* clear1;
  int N;
};
hipError_t launch_brick_summary_batched(const BrickJob* dJobs, const u32* dPre, int n, u32 blocks, int wmax,
                                        hipStream_t s);
size_t batch_queue_cap(const BatchMorphJob* jobs, int n);
// ---- radii above the chord tables (gm_edt.hip, orchestration in gm_morph.hip)
size_t edt_scratch_ints(int N);
hipError_t launch_edt(const u32* src, int N, int s, int a, int b, int nq, int reach, bool weighted, bool cheb,
                      const RadiusField& rf, int* tmp, int* out, hipStream_t st);
hipError_t launch_edt_dil_coarse(const int* dt, const MorphDims& d, const RadiusField& rf, bool cube, const u16* ifl,
                                 u16* df, hipStream_t st);
hipError_t launch_edt_dil_fine(const int* dt, const MorphDims& d, const RadiusField& rf, bool cube, const u16* tg,
                               u32* D, hipStream_t st);
hipError_t launch_edt_ero_coarse(const int* et, int thr, const MorphDims& d, const u16* dal, const u16* ifl, u16* cl,
                                 hipStream_t st);
hipError_t launch_edt_ero_fine(const int* et, int thr, const MorphDims& d, const u16* ug, u32* C, hipStream_t st);
hipError_t launch_dilate_edt(const MorphDims& d, const MorphBuffers& b, const RadiusField& rf, int se, int rho_hi,
                             int* tmp, int* out, hipStream_t s, hipEvent_t* ev, hipEvent_t* evs);
hipError_t launch_erode_edt(const MorphDims& d, const MorphBuffers& b, const RadiusField& rf, int rho_hi, int* tmp,
                            int* out, hipStream_t s, hipEvent_t* ev, hipEvent_t* evs);

// Side streams of the batched closing: the per-radius launches of a pass are
// independent (each job has one radius and its own buffers), so they are
// spread over s and these streams, forked from and joined back into s with
// the events (which keeps the sequence capturable as a graph).
struct SideStreams {
  int n = 0;
  hipStream_t s[3] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[3] = {};
};
hipError_t launch_close_batched(const BatchMorphJob* jobs, int n, void* hBlob, void* dBlob, u32* queue, int cap,
                                int* qsel, hipStream_t s, const SideStreams& side);
// Radius planes and per-brick value sets of a per-voxel radius field
// (values clamped to [2,31]; the caller has checked the range). brickBox
// (optional): {w0, by0, bz0, nw, nby, nbz}, the bricks (1 word x 8 rows x 8
// planes) to rebuild; all of them when null.
hipError_t launch_radius_planes(const RadiusField& rf, int N, u32* rp, u32* rvset, hipStream_t s,
                                const int* brickBox = nullptr);

// ---- ScaleField editing (gm_scale.hip)
// One brush point (ScaleField::UpdateScaleGrid): fine voxel i samples at
// bmin + cell * i; h = (support - scale) / sqrt(1 - sqrt(0.5)); mode 1 MAX,
// 2 MIN, 3 ZERO.
struct ScaleBrush {
  float bmin[3];
  float cell;
  float pos[3];
  float scale, h, global_scale;
  int mode;
};
hipError_t launch_scale_fill(int8_t* grid, size_t n, int8_t value, hipStream_t s);
// box: {x0, y0, z0, nx, ny, nz} of the voxels that can change; hist: 128
// counts of clamped radius values, updated by the changed voxels.
hipError_t launch_scale_brush(int8_t* grid, int N, const ScaleBrush& br, const int box[6],
                              unsigned long long* hist, hipStream_t s);

// Host check that the chord tables compiled into the streaming kernels equal
// the ones build_se_table produces.
bool ct_tables_ok();

// Streaming kernels compiled for one uniform radius (gm_morph_ct.hip, built
// as kCtParts objects, radius rho in part rho % kCtParts): fn[pass][rho] is
// the kernel of pass 0..3 (coarse dilation, fine dilation, coarse erosion,
// fine erosion) for chord-table radius rho, null where none is compiled.
constexpr int kCtMaxRho = 32;
constexpr int kCtParts = 8;
struct CtKernels {
  const void* fn[4][kCtMaxRho + 1];
  const void* fnB[4][kCtMaxRho + 1];  // batched forms (argument table + job/tile queue)
  const void* fnBox[2][kCtMaxRho + 1];  // cube dilation, coarse and fine (uniform radius rho)
};
void ct_register_part0(CtKernels&);
void ct_register_part1(CtKernels&);
void ct_register_part2(CtKernels&);
void ct_register_part3(CtKernels&);
void ct_register_part4(CtKernels&);
void ct_register_part5(CtKernels&);
void ct_register_part6(CtKernels&);
void ct_register_part7(CtKernels&);

// ---- marching cubes (gm_mc.hip)
struct MCBuffers {
  u32* blockV;   // per 8^3 Morton block: vertex count, then exclusive offset
  u32* blockT;
  u32* cellV;    // per listed block (list order) and cell (local Morton index): block-local vertex base
  unsigned long long* cellCase;  // per listed block (list order) and cell: its 8 cube cases
  u32* posOf;    // per block: its index in list, ~0u when not listed
  u32* order;    // emit order of the listed blocks: 4 class counts, then 4 x nblocks list positions
  float* V;
  float* Nrm;
  u32* T;
  u32* late;     // [0] count, capLate (slot, a, b, c) quadruples, then their copy in slot order
  u32* totals;   // [0] nV [1] nT [2] overflow (1 output, 2 late list, 4 block list) [3] border flag
                 // [4] cubes listing more triangles than they count [5] listed blocks (8 words)
  u32* scanPart; // per 1024 blocks: vertex and triangle totals (mc_scan_part_bytes)
  uint8_t* brick; // brick summary of the closing grid (1 word x 8 rows x 8 planes; batched extraction)
  u32* list;      // [0] number of active blocks (zero between extractions), then their ids
  u32 capV, capT, capLate;
  u32 capList;   // listed blocks cellV and cellCase hold
  int nblocks;   // (P/8)^3
  int P;         // padded power-of-two cells per axis
};
// One-time set-up of the marching-cubes tables (call before a capture).
hipError_t mc_prepare();
hipError_t launch_mc(const u32* C, int R, const float* bmin3, float cell, const MCBuffers& m, hipStream_t s,
                     hipEvent_t* ev);
size_t mc_scan_parts(int nblocks);
size_t mc_scan_part_bytes(int nblocks);  // device bytes of MCBuffers::scanPart
hipError_t launch_mc_emit(const u32* C, int R, const float* bmin3, float cell, const MCBuffers& m,
                          hipStream_t s, hipEvent_t* ev);
// Marching cubes of many closed grids, all on stream s: the same kernels per
// job, one launch per step. hBlob (pinned) and dBlob hold batch_mc_bytes(n);
// hTotals (pinned, 4 u32 per job) receives each job's totals.
struct BatchMCJob {
  const u32* C;
  int R;
  float bmin[3];
  float cell;
  MCBuffers m;
};
size_t batch_mc_bytes(int n);
hipError_t launch_mc_batched(const BatchMCJob* jobs, int n, void* hBlob, void* dBlob, u32* hTotals, hipStream_t s);

// ---- bilateral normal filtering (gm_normals.hip): iters iterations on the
// unit normals Nrm of a mesh (V, T) with nv vertices and nt triangles.
// scratch holds bilateral_scratch_words(nv, nt) words, Ntmp nv * 3 floats.
size_t bilateral_scratch_words(u32 nv, u32 nt);
hipError_t launch_bilateral(const float* V, float* Nrm, const u32* T, u32 nv, u32 nt, int iters, u32* scratch,
                            float* Ntmp, hipStream_t s);

// ---- reference-format max-mipmap (gm_pyramid.hip)
hipError_t launch_pyramid(const u32* rows, int R, const RadiusField& rf, u64* levels, hipStream_t s);

}  // namespace gm
