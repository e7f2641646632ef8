// C-ABI of the GMorpho closing path: context, device buffers and the
// orchestration of GMorpho::Load / Update / ExtractClosingMesh
// (GMorpho.cu:7731-7845, 6814-7001) on one HIP stream.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/gmorpho_c.h"
#include "gm_internal.h"

namespace gm {
std::vector<u32> build_se_table(int kind, int rho, bool cube);
hipError_t mc_tables_host(u32* ntri, u32* len, unsigned long long* tri);
}  // namespace gm

using namespace gm;

struct gm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // grid
  bool has_grid = false;
  int R = 0, N = 0, W = 0;
  int capR = 0;  // grid buffers are sized for capR; smaller grids reuse them
  float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0}, cell = 0.f;
  int data_res[3] = {0, 0, 0};
  u32 *I = nullptr, *D = nullptr, *K = nullptr, *C = nullptr;
  u16 *DF = nullptr, *TG = nullptr, *CL = nullptr, *UG = nullptr, *IFL = nullptr, *DAL = nullptr;
  uint8_t *brickI = nullptr, *brickK = nullptr;
  u32* queue = nullptr;  // active-tile queues of the streaming passes
  int qsel = 0;          // counter block of the next streaming pass
  uint8_t* stage = nullptr;  // R^3 bytes for layout conversions
  // mesh input
  float* dP = nullptr;
  u32* dT = nullptr;
  int32_t* dSnap = nullptr;
  int nv = 0, nf = 0;
  size_t capP = 0, capTri = 0;
  u32 *mark = nullptr, *shell = nullptr;
  bool voxDirty = true;  // mark / shell may hold bits (fresh or interrupted)
  u32* dCheck = nullptr;          // mesh check results (8 u32, then the per-workgroup parts)
  uint8_t* hStage = nullptr;      // pinned upload staging of the mesh
  size_t capStage = 0;
  // radius field
  int rmode = 0, r = 0, rlo = 0, rhi = 0, L = 0;
  int8_t* dGrid = nullptr;
  size_t capGrid = 0;  // bytes of dGrid
  // counts of the clamped radius values of dGrid (device and host copies),
  // kept by the fill and the brush so the range needs no grid scan
  unsigned long long* dHist = nullptr;
  unsigned long long hHist[128] = {};
  int32_t* dLat = nullptr;
  long long latMaxAbs = 0;  // largest |fixed-point lattice value|
  // per-voxel radii by value: radius planes, per-brick value sets, masked
  // erosion source, K brick list (sized for capVarR)
  u32 *rp = nullptr, *rvset = nullptr, *kv = nullptr, *kbl = nullptr;
  int capVarR = 0;
  bool rpValid = false;  // rp/rvset hold the current radius field
  // radii above the chord tables: distance-transform scratch (N^3 ints each)
  int *edtTmp = nullptr, *edtOut = nullptr;
  size_t capEdt = 0;
  // structuring-element tables
  u32* dTables = nullptr;
  size_t tablesCap = 0;
  SETableSet tabs;
  int tabLo = -1, tabHi = -1;
  int tabWords = 0;  // longest table of the current set
  u32* dCube = nullptr;  // cube dilation tables
  size_t cubeCap = 0;
  SETableSet cubeTabs;
  int cubeLo = -1, cubeHi = -1, cubeWords = 0;
  // marching cubes
  MCBuffers mc{};
  u32* hTotals = nullptr;
  bool meshPending = false;
  int Pmc = 0;
  // bilateral normal filtering: incidence lists and a second normal buffer
  u32* bilScratch = nullptr;
  size_t bilWords = 0;
  float* bilN = nullptr;
  size_t bilNCap = 0;  // vertices of bilN
  // pyramid
  u64* dPyr = nullptr;
  uint64_t capPyr = 0;  // texels of dPyr
  int pyrLevels = 0;
  uint64_t pyrTexels = 0;
  // Captured launch sequences (hipGraph) of gm_close (per structuring
  // element) and gm_extract_mesh, replayed while their key holds. graphGen
  // changes whenever a launch argument could (buffers, sizes, radius field,
  // tables, grid attributes).
  struct Graph {
    hipGraphExec_t exec = nullptr;
    uint64_t key[4] = {0, 0, 0, 0};
    uint64_t seen[4] = {0, 0, 0, 0};  // key of the last eager run (repeat mode)
  };
  Graph gClose[2], gMesh, gUpdate[2];
  uint64_t graphGen = 1;
  // 0: eager; 1: capture at once; 2: capture when a key repeats (a context
  // that runs a different job every time never pays for a capture)
  int graphs = 1;
  // launch shape forced while gm_autotune measures (tileZ 0: the table's)
  int forceTileZ = 0, forceWaveCap = 0;
  int tuneSe = 0;  // structuring element of the closing being launched (tuning-table key)
  // device bytes held by the context's own allocations
  std::unordered_map<void*, size_t> allocBytes;
  uint64_t devBytes = 0;
  // timing
  bool timing = false;
  hipEvent_t ev[20] = {};
  int evValid = 0;
};

namespace {

// Launch shapes chosen by gm_autotune, per (base resolution, uniform radius
// or -1 for per-voxel radii, structuring element); shared by the contexts of
// the process.
struct TuneKey {
  int R, r, se;
  bool operator<(const TuneKey& o) const { return std::tie(R, r, se) < std::tie(o.R, o.r, o.se); }
};
struct TuneShape {
  int tileZ = 0, waveCap = 0;
  double ms = 0.0;
};
std::mutex g_tuneMu;
std::map<TuneKey, TuneShape> g_tune;

int fail(gm_ctx* c, int code, const std::string& msg) {
  if (c) c->err = "GMorpho Error: " + msg;
  return code;
}

#define HIPCALL(ctx, x)                                                                       \
  do {                                                                                        \
    hipError_t e_ = (x);                                                                      \
    if (e_ != hipSuccess) return fail(ctx, GM_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
void dfree(T*& p) {
  if (p) hipFree(p);
  p = nullptr;
}

// Context allocations, counted so that a job engine can charge a slot for
// what its context really holds after buffers grew.
template <class T>
hipError_t calloc_dev(gm_ctx* c, T** p, size_t bytes) {
  const hipError_t e = hipMalloc((void**)p, bytes);
  if (e == hipSuccess) {
    c->allocBytes[(void*)*p] = bytes;
    c->devBytes += bytes;
  }
  return e;
}
template <class T>
void cfree(gm_ctx* c, T*& p) {
  if (!p) return;
  auto it = c->allocBytes.find((void*)p);
  if (it != c->allocBytes.end()) {
    c->devBytes -= it->second;
    c->allocBytes.erase(it);
  }
  hipFree(p);
  p = nullptr;
}

// Launch arguments may have changed: captured sequences are out of date.
void touch(gm_ctx* c) { ++c->graphGen; }

void drop_graph(gm_ctx::Graph& g) {
  if (g.exec) hipGraphExecDestroy(g.exec);
  g.exec = nullptr;
}

// Runs enqueue() (kernel launches and async copies on c->stream only) as a
// captured graph: replayed when the key matches the last capture, else
// captured again. Without graphs (or with per-stage timing) it runs eagerly.
template <class F>
int run_captured(gm_ctx* c, gm_ctx::Graph& g, const uint64_t key[4], F enqueue) {
  if (!c->graphs || c->timing) return enqueue();
  if (g.exec && std::equal(key, key + 4, g.key)) {
    HIPCALL(c, hipGraphLaunch(g.exec, c->stream));
    return GM_OK;
  }
  drop_graph(g);
  if (c->graphs == 2 && !std::equal(key, key + 4, g.seen)) {
    std::copy(key, key + 4, g.seen);
    return enqueue();
  }
  HIPCALL(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  const int rc = enqueue();
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &graph);
  if (rc != GM_OK || e != hipSuccess) {
    if (graph) hipGraphDestroy(graph);
    if (rc != GM_OK) return rc;
    return fail(c, GM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    g.exec = nullptr;
    return fail(c, GM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  std::copy(key, key + 4, g.key);
  HIPCALL(c, hipGraphLaunch(g.exec, c->stream));
  return GM_OK;
}

void free_grid(gm_ctx* c) {
  touch(c);
  cfree(c, c->I); cfree(c, c->D); cfree(c, c->K); cfree(c, c->C);
  cfree(c, c->DF); cfree(c, c->TG); cfree(c, c->CL); cfree(c, c->UG); cfree(c, c->IFL); cfree(c, c->DAL); cfree(c, c->brickI); cfree(c, c->brickK); cfree(c, c->queue);
  cfree(c, c->stage); cfree(c, c->mark); cfree(c, c->shell);
  cfree(c, c->mc.blockV); cfree(c, c->mc.blockT); cfree(c, c->mc.cellV); cfree(c, c->mc.cellCase); cfree(c, c->mc.posOf); cfree(c, c->mc.order);
  cfree(c, c->mc.V); cfree(c, c->mc.Nrm); cfree(c, c->mc.T); cfree(c, c->mc.late); cfree(c, c->mc.totals); cfree(c, c->mc.scanPart);
  cfree(c, c->mc.brick); cfree(c, c->mc.list);
  cfree(c, c->dPyr);
  c->capPyr = 0;
  cfree(c, c->dGrid);
  c->capGrid = 0;
  cfree(c, c->rp); cfree(c, c->rvset); cfree(c, c->kv); cfree(c, c->kbl);
  c->capVarR = 0;
  c->rpValid = false;
  cfree(c, c->edtTmp); cfree(c, c->edtOut);
  c->capEdt = 0;
  c->has_grid = false;
  c->capR = 0;
  c->meshPending = false;
  c->pyrLevels = 0;
}

int alloc_mesh_out(gm_ctx* c, u32 capV, u32 capT, u32 capLate) {
  if (capV > c->mc.capV || capT > c->mc.capT || capLate > c->mc.capLate || !c->mc.late) touch(c);
  if (capV > c->mc.capV) {
    cfree(c, c->mc.V);
    cfree(c, c->mc.Nrm);
    HIPCALL(c, calloc_dev(c, &c->mc.V, (size_t)capV * 12));
    HIPCALL(c, calloc_dev(c, &c->mc.Nrm, (size_t)capV * 12));
    c->mc.capV = capV;
  }
  if (capT > c->mc.capT) {
    cfree(c, c->mc.T);
    HIPCALL(c, calloc_dev(c, &c->mc.T, (size_t)capT * 12));
    c->mc.capT = capT;
  }
  if (capLate > c->mc.capLate || !c->mc.late) {
    cfree(c, c->mc.late);
    // entries, then their copy in slot order (the late-triangle fixup)
    HIPCALL(c, calloc_dev(c, &c->mc.late, (size_t)(1 + 8 * capLate) * 4));
    c->mc.capLate = capLate;
  }
  return GM_OK;
}

// Capacity after an overflow: the need plus a quarter, and at least half as
// much again as before, so a context that runs jobs of varying output sizes
// regrows (and frees, which waits for the whole device) only a few times.
u32 grow_cap(u32 cap, u32 need) {
  const uint64_t a = (uint64_t)need + need / 4 + 1024, b = (uint64_t)cap + cap / 2;
  return (u32)std::min<uint64_t>(std::max(a, b), 0xFFFFFFF0u);
}

// Per-listed-block marching-cubes buffers (cube cases, vertex bases) for cap
// listed blocks.
int alloc_mc_list(gm_ctx* c, u32 cap) {
  if (cap <= c->mc.capList && c->mc.cellCase) return GM_OK;
  touch(c);
  cfree(c, c->mc.cellV);
  cfree(c, c->mc.cellCase);
  c->mc.capList = 0;
  HIPCALL(c, calloc_dev(c, &c->mc.cellV, (size_t)cap * 512 * 4));
  HIPCALL(c, calloc_dev(c, &c->mc.cellCase, (size_t)cap * 512 * 8));
  c->mc.capList = cap;
  return GM_OK;
}

// Grid dimensions of base resolution R inside buffers sized for c->capR.
void set_dims(gm_ctx* c, int R) {
  touch(c);
  if (c->R != R && c->rmode != 0) {
    // a per-voxel radius field belongs to the previous grid size
    c->rmode = 0;
    c->rpValid = false;
    c->r = 0;
    c->rlo = c->rhi = 0;
  }
  c->R = R;
  c->N = 2 * R;
  c->W = c->N / 32;
  int P = 1;
  while (P < R) P <<= 1;
  c->Pmc = P;
  c->mc.P = P;
  c->mc.nblocks = (P / 8) * (P / 8) * (P / 8);
}

// Buffers for base resolution R. A context keeps the largest buffers it has
// allocated and reuses them for smaller grids, so a sequence of jobs of mixed
// sizes allocates (and frees, which waits for the device) only when a new
// largest size comes.
int alloc_grid(gm_ctx* c, int R) {
  if (R < 16 || R > 1024 || (R % 16) != 0)
    return fail(c, GM_ERR_UNSUPPORTED, "base resolution must be a multiple of 16 in [16,1024]");
  if (c->capR > 0 && R <= c->capR) {
    set_dims(c, R);
    c->has_grid = true;
    return GM_OK;
  }
  free_grid(c);
  c->R = R;
  c->N = 2 * R;
  c->W = c->N / 32;
  size_t rowBytes = (size_t)c->N * c->N * c->W * 4;
  size_t flagBytes = (size_t)R * R * c->W * 2;
  HIPCALL(c, calloc_dev(c, &c->I, rowBytes));
  HIPCALL(c, calloc_dev(c, &c->D, rowBytes));
  HIPCALL(c, calloc_dev(c, &c->K, rowBytes));
  HIPCALL(c, calloc_dev(c, &c->C, rowBytes));
  HIPCALL(c, calloc_dev(c, &c->mark, rowBytes));
  HIPCALL(c, calloc_dev(c, &c->shell, rowBytes));
  c->voxDirty = true;
  HIPCALL(c, calloc_dev(c, &c->DF, flagBytes));
  HIPCALL(c, calloc_dev(c, &c->TG, flagBytes));
  HIPCALL(c, calloc_dev(c, &c->CL, flagBytes));
  HIPCALL(c, calloc_dev(c, &c->UG, flagBytes));
  HIPCALL(c, calloc_dev(c, &c->IFL, flagBytes));
  HIPCALL(c, calloc_dev(c, &c->DAL, flagBytes));
  size_t brickBytes = (size_t)(c->N / 8) * (c->N / 8) * c->W;
  HIPCALL(c, calloc_dev(c, &c->brickI, brickBytes));
  HIPCALL(c, calloc_dev(c, &c->brickK, brickBytes));
  HIPCALL(c, calloc_dev(c, &c->queue, tile_queue_words(R) * 4));
  // both counter blocks start at zero (each pass clears the other one)
  HIPCALL(c, hipMemsetAsync(c->queue, 0, kQueueLists * 4, c->stream));
  c->qsel = 0;
  HIPCALL(c, calloc_dev(c, &c->stage, (size_t)R * R * R));
  HIPCALL(c, hipMemsetAsync(c->I, 0, rowBytes, c->stream));
  HIPCALL(c, hipMemsetAsync(c->D, 0, rowBytes, c->stream));
  HIPCALL(c, hipMemsetAsync(c->K, 0, rowBytes, c->stream));
  HIPCALL(c, hipMemsetAsync(c->C, 0, rowBytes, c->stream));
  int P = 1;
  while (P < R) P <<= 1;
  c->Pmc = P;
  c->mc.P = P;
  c->mc.nblocks = (P / 8) * (P / 8) * (P / 8);
  HIPCALL(c, calloc_dev(c, &c->mc.blockV, (size_t)c->mc.nblocks * 4));
  HIPCALL(c, calloc_dev(c, &c->mc.blockT, (size_t)c->mc.nblocks * 4));
  c->mc.capList = 0;
  {
    // a closed surface lists on the order of 6 (P/8)^2 blocks; more grow the
    // per-block buffers when the mesh is read
    const size_t B = (size_t)P / 8;
    int rcl = alloc_mc_list(c, (u32)std::min<size_t>((size_t)c->mc.nblocks, std::max<size_t>(64, 8 * B * B)));
    if (rcl) return rcl;
  }
  HIPCALL(c, calloc_dev(c, &c->mc.posOf, (size_t)c->mc.nblocks * 4));
  HIPCALL(c, calloc_dev(c, &c->mc.order, ((size_t)4 * c->mc.nblocks + 4) * 4));
  HIPCALL(c, calloc_dev(c, &c->mc.totals, 32));  // nV, nT, overflow, border, late cubes, listed blocks, spare
  HIPCALL(c, calloc_dev(c, &c->mc.scanPart, mc_scan_part_bytes(c->mc.nblocks)));
  HIPCALL(c, calloc_dev(c, &c->mc.brick, brickBytes));
  HIPCALL(c, calloc_dev(c, &c->mc.list, ((size_t)c->mc.nblocks + 1) * 4));
  // the block-list count starts from zero (each extraction's last kernel
  // clears it for the next)
  HIPCALL(c, hipMemsetAsync(c->mc.list, 0, 4, c->stream));
  c->mc.capV = c->mc.capT = c->mc.capLate = 0;
  int rc = alloc_mesh_out(c, (u32)std::min<size_t>((size_t)4 * c->N * c->N, 1u << 30),
                          (u32)std::min<size_t>((size_t)8 * c->N * c->N, 1u << 30), 4096);
  if (rc) return rc;
  c->has_grid = true;
  c->capR = R;
  // default radius: uniform 0 (closing = input) until a ScaleField is set
  if (c->rmode != 0) {
    c->rmode = 0;
    c->r = 0;
    c->rlo = c->rhi = 0;
  }
  return GM_OK;
}

int need_grid(gm_ctx* c) {
  if (!c) return GM_ERR_ARG;
  if (!c->has_grid) return fail(c, GM_ERR_STATE, "no grid: call gm_load, gm_voxelize or gm_grid_init first");
  return GM_OK;
}

u32* grid_rows(gm_ctx* c, int which) {
  switch (which) {
    case GM_GRID_INPUT: return c->I;
    case GM_GRID_DILATION: return c->D;
    case GM_GRID_CONTOUR: return c->K;
    case GM_GRID_CLOSING: return c->C;
    default: return nullptr;
  }
}

RadiusField radius_field(const gm_ctx* c) {
  RadiusField f{};
  f.mode = c->rmode;
  f.r = c->r;
  f.N = c->N;
  f.grid = c->dGrid;
  f.lat = c->dLat;
  f.L = c->L;
  f.s = c->L ? c->N / c->L : 1;
  f.lshift = 0;
  while ((1 << f.lshift) < f.s) ++f.lshift;
  f.narrow = (double)f.s * f.s * f.s * (double)c->latMaxAbs < (double)(1 << 30) ? 1 : 0;
  return f;
}

int ensure_tables(gm_ctx* c, int rlo, int rhi) {
  if (c->tabLo == rlo && c->tabHi == rhi && c->dTables) return GM_OK;
  static const bool ctOk = ct_tables_ok();
  if (!ctOk) return fail(c, GM_ERR_UNSUPPORTED, "compiled chord tables differ from the structuring element");
  std::vector<u32> all;
  SETableSet t;
  for (int i = 0; i <= 128; ++i) t.dil[i] = t.ero_corner[i] = t.ero_centre[i] = -1;
  int longest = 0;
  auto add = [&](const std::vector<u32>& v) {
    int off = (int)all.size();
    all.insert(all.end(), v.begin(), v.end());
    longest = std::max(longest, (int)v.size());
    return off;
  };
  // sphere dilation uses rho = r+1, cube uses rho = r; both are corner tables,
  // so dil[] holds sphere tables at r+1 and cube tables are built separately.
  for (int r = rlo; r <= rhi; ++r) t.dil[r + 1] = add(build_se_table(SE_CORNER, r + 1, false));
  for (int r = rlo; r <= rhi; ++r) {
    int rho = r - 1;
    if (rho < 1) continue;
    t.ero_corner[rho] = add(build_se_table(SE_CORNER, rho, false));
    t.ero_centre[rho] = add(build_se_table(SE_CENTRE, rho, false));
  }
  c->tabs = t;
  size_t bytes = all.size() * 4;
  if (bytes > c->tablesCap) {
    cfree(c, c->dTables);
    HIPCALL(c, calloc_dev(c, &c->dTables, bytes));
    c->tablesCap = bytes;
  }
  HIPCALL(c, hipMemcpyAsync(c->dTables, all.data(), bytes, hipMemcpyHostToDevice, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  touch(c);
  c->tabLo = rlo;
  c->tabHi = rhi;
  c->tabWords = longest;
  return GM_OK;
}

// Cube dilation tables live in a second table set (dil[] indexed by r),
// kept by the context like the sphere tables.
int cube_tables(gm_ctx* c, int rlo, int rhi) {
  if (c->cubeLo == rlo && c->cubeHi == rhi && c->dCube) return GM_OK;
  std::vector<u32> all;
  SETableSet& t = c->cubeTabs;
  int longest = 0;
  for (int i = 0; i <= 128; ++i) t.dil[i] = t.ero_corner[i] = t.ero_centre[i] = -1;
  for (int r = rlo; r <= rhi; ++r) {
    int off = (int)all.size();
    std::vector<u32> v = build_se_table(SE_CORNER, r, true);
    all.insert(all.end(), v.begin(), v.end());
    longest = std::max(longest, (int)v.size());
    t.dil[r] = off;
  }
  const size_t bytes = all.size() * 4;
  if (bytes > c->cubeCap) {
    cfree(c, c->dCube);
    c->cubeCap = 0;
    HIPCALL(c, calloc_dev(c, &c->dCube, bytes));
    c->cubeCap = bytes;
  }
  HIPCALL(c, hipMemcpyAsync(c->dCube, all.data(), bytes, hipMemcpyHostToDevice, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  touch(c);
  c->cubeLo = rlo;
  c->cubeHi = rhi;
  c->cubeWords = longest;
  return GM_OK;
}

MorphBuffers morph_buffers(gm_ctx* c) {
  MorphBuffers b{};
  b.I = c->I;
  b.D = c->D;
  b.K = c->K;
  b.C = c->C;
  b.DF = c->DF;
  b.TG = c->TG;
  b.CL = c->CL;
  b.UG = c->UG;
  b.IFL = c->IFL;
  b.DAL = c->DAL;
  b.tables = c->dTables;
  b.brickI = c->brickI;
  b.brickK = c->brickK;
  b.queue = c->queue;
  b.queueCap = (int)tile_count_max(c->R);
  b.qsel = &c->qsel;
  if (c->forceTileZ) {
    b.tileZ = c->forceTileZ;
    b.waveCap = c->forceWaveCap;
  } else {
    std::lock_guard<std::mutex> lk(g_tuneMu);
    const int r = c->rmode == 0 ? c->r : -1;
    auto it = g_tune.find(TuneKey{c->R, r, c->tuneSe});
    if (it != g_tune.end()) {
      b.tileZ = it->second.tileZ;
      b.waveCap = it->second.waveCap;
    }
  }
  if (c->rmode != 0 && c->capVarR >= c->R) {
    b.rp = c->rp;
    b.rvset = c->rvset;
    b.kv = c->kv;
    b.kbl = c->kbl;
    b.rpBuild = c->rpValid ? 0 : 1;
  }
  return b;
}

MorphDims morph_dims(gm_ctx* c) {
  MorphDims d{};
  d.R = c->R;
  d.N = c->N;
  d.W = c->W;
  for (int a = 0; a < 3; ++a) d.dR[a] = std::max(0, std::min(c->data_res[a], c->R));
  return d;
}

// Radii above 31 fine voxels (the chord tables' range) go through the
// distance-transform passes.
bool use_edt(const gm_ctx* c) { return c->rhi > 31; }
int check_radius(gm_ctx*, int) { return GM_OK; }
int ensure_edt_buffers(gm_ctx* c) {
  const size_t n = edt_scratch_ints(c->N);
  if (n <= c->capEdt) return GM_OK;
  cfree(c, c->edtTmp);
  cfree(c, c->edtOut);
  c->capEdt = 0;
  HIPCALL(c, calloc_dev(c, &c->edtTmp, n * 4));
  HIPCALL(c, calloc_dev(c, &c->edtOut, n * 4));
  c->capEdt = n;
  touch(c);
  return GM_OK;
}

// Buffers of the per-value passes of a per-voxel radius field (kept at the
// largest size used, like the grid buffers).
int ensure_var_buffers(gm_ctx* c) {
  if (c->rmode == 0 || c->capVarR >= c->R) return GM_OK;
  cfree(c, c->rp); cfree(c, c->rvset); cfree(c, c->kv); cfree(c, c->kbl);
  c->capVarR = 0;
  c->rpValid = false;
  const int R = c->R;
  HIPCALL(c, calloc_dev(c, &c->rp, radius_plane_words(R) * 4));
  HIPCALL(c, calloc_dev(c, &c->rvset, brick_words(R) * 4));
  HIPCALL(c, calloc_dev(c, &c->kv, (size_t)R * R * R / 4 * 4));
  HIPCALL(c, calloc_dev(c, &c->kbl, (brick_words(R) + 1) * 4));
  c->capVarR = R;
  touch(c);
  return GM_OK;
}

// Device radius grid (N^3 signed chars) and its value histogram.
int ensure_scale_grid(gm_ctx* c) {
  const size_t n = (size_t)c->N * c->N * c->N;
  if (n > c->capGrid) {
    cfree(c, c->dGrid);
    c->capGrid = 0;
    HIPCALL(c, calloc_dev(c, &c->dGrid, n));
    c->capGrid = n;
  }
  if (!c->dHist) HIPCALL(c, calloc_dev(c, &c->dHist, sizeof(c->hHist)));
  return GM_OK;
}
int upload_hist(gm_ctx* c) {
  HIPCALL(c, hipMemcpyAsync(c->dHist, c->hHist, sizeof(c->hHist), hipMemcpyHostToDevice, c->stream));
  return GM_OK;
}
void set_range_from_hist(gm_ctx* c) {
  int lo = 127, hi = 2;
  for (int v = 2; v < 128; ++v)
    if (c->hHist[v]) {
      lo = std::min(lo, v);
      hi = std::max(hi, v);
    }
  c->rlo = std::min(lo, hi);
  c->rhi = hi;
}

// Host-synchronous set-up of a dilation (tables, per-value buffers): done
// before any capture.
int prep_dilate(gm_ctx* c, int se) {
  int rc = check_radius(c, se);
  if (rc) return rc;
  if (use_edt(c)) return ensure_edt_buffers(c);
  const int lo = std::max(c->rlo, 0), hi = c->rhi;
  if (se == GM_SE_CUBE) return cube_tables(c, lo, hi);
  rc = ensure_tables(c, lo, hi);
  if (rc) return rc;
  return ensure_var_buffers(c);
}
int prep_erode(gm_ctx* c) {
  int rc = check_radius(c, GM_SE_SPHERE);
  if (rc) return rc;
  if (use_edt(c)) return ensure_edt_buffers(c);
  rc = ensure_tables(c, std::max(c->rlo, 0), c->rhi);
  if (rc) return rc;
  return ensure_var_buffers(c);
}
// Launches only (capturable).
int enq_dilate(gm_ctx* c, int se) {
  const int lo = std::max(c->rlo, 0), hi = c->rhi;
  hipError_t e;
  if (use_edt(c)) {
    e = launch_dilate_edt(morph_dims(c), morph_buffers(c), radius_field(c), se == GM_SE_CUBE ? 1 : 0, hi, c->edtTmp,
                          c->edtOut, c->stream, c->timing ? &c->ev[0] : nullptr, c->timing ? &c->ev[15] : nullptr);
  } else if (se == GM_SE_CUBE) {
    MorphBuffers b = morph_buffers(c);
    b.tables = c->dCube;
    e = launch_dilate(morph_dims(c), b, radius_field(c), 1, c->cubeTabs, lo, hi, c->cubeWords, c->stream,
                      c->timing ? &c->ev[0] : nullptr, c->timing ? &c->ev[15] : nullptr);
  } else {
    e = launch_dilate(morph_dims(c), morph_buffers(c), radius_field(c), 0, c->tabs, lo, hi, c->tabWords, c->stream,
                      c->timing ? &c->ev[0] : nullptr, c->timing ? &c->ev[15] : nullptr);
  }
  if (e != hipSuccess) return fail(c, GM_ERR_HIP, std::string("dilation: ") + hipGetErrorString(e));
  return GM_OK;
}
int enq_erode(gm_ctx* c) {
  hipError_t e =
      use_edt(c) ? launch_erode_edt(morph_dims(c), morph_buffers(c), radius_field(c), c->rhi, c->edtTmp, c->edtOut,
                                    c->stream, c->timing ? &c->ev[5] : nullptr, c->timing ? &c->ev[17] : nullptr)
                 : launch_erode(morph_dims(c), morph_buffers(c), radius_field(c), c->tabs, c->rlo, c->rhi,
                                c->tabWords, c->stream, c->timing ? &c->ev[5] : nullptr,
                                c->timing ? &c->ev[17] : nullptr);
  if (e != hipSuccess) return fail(c, GM_ERR_HIP, std::string("erosion: ") + hipGetErrorString(e));
  return GM_OK;
}
// State after a dilation or erosion was enqueued: the sphere's per-value
// passes built the radius planes.
void after_morph(gm_ctx* c, bool planes) {
  if (planes && !use_edt(c) && c->rmode != 0 && c->capVarR >= c->R) c->rpValid = true;
  c->meshPending = false;
}

int do_dilate(gm_ctx* c, int se) {
  int rc = prep_dilate(c, se);
  if (rc == GM_OK) rc = enq_dilate(c, se);
  if (rc == GM_OK) after_morph(c, se == GM_SE_SPHERE);
  return rc;
}

int do_erode(gm_ctx* c) {
  int rc = prep_erode(c);
  if (rc == GM_OK) rc = enq_erode(c);
  if (rc == GM_OK) after_morph(c, true);
  return rc;
}

}  // namespace

extern "C" {

int gm_create(int device, gm_ctx** out) {
  if (!out) return GM_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return GM_ERR_HIP;
  if (device < 0 || device >= n) return GM_ERR_ARG;
  gm_ctx* c = new gm_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return GM_ERR_HIP;
  }
  for (auto& e : c->ev) hipEventCreate(&e);
  // pinned: 4 mesh totals, then 8 words of the mesh check
  hipHostMalloc((void**)&c->hTotals, 48, hipHostMallocDefault);
  *out = c;
  return GM_OK;
}

int gm_destroy(gm_ctx* c) {
  if (!c) return GM_ERR_ARG;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  drop_graph(c->gClose[0]);
  drop_graph(c->gClose[1]);
  drop_graph(c->gMesh);
  drop_graph(c->gUpdate[0]);
  drop_graph(c->gUpdate[1]);
  free_grid(c);
  cfree(c, c->dP);
  cfree(c, c->dT);
  cfree(c, c->dSnap);
  cfree(c, c->dCheck);
  if (c->hStage) hipHostFree(c->hStage);
  cfree(c, c->dLat);
  cfree(c, c->dTables);
  cfree(c, c->dCube);
  cfree(c, c->bilScratch);
  cfree(c, c->bilN);
  cfree(c, c->dHist);
  for (auto& e : c->ev) hipEventDestroy(e);
  if (c->hTotals) hipHostFree(c->hTotals);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return GM_OK;
}

const char* gm_last_error(const gm_ctx* c) { return c ? c->err.c_str() : "GMorpho Error: null context"; }

int gm_load_mesh(gm_ctx* c, const float* P, const float* N, int nv, const uint32_t* T, int nf) {
  (void)N;
  if (!c) return GM_ERR_ARG;
  if (!P || !T || nv <= 0 || nf <= 0) return fail(c, GM_ERR_ARG, "empty mesh");
  hipSetDevice(c->device);
  if ((size_t)nv > c->capP) {
    cfree(c, c->dP);
    cfree(c, c->dSnap);
    c->capP = 0;
    HIPCALL(c, calloc_dev(c, &c->dP, (size_t)nv * 12));
    HIPCALL(c, calloc_dev(c, &c->dSnap, (size_t)nv * 12));
    c->capP = nv;
  }
  if ((size_t)nf > c->capTri) {
    cfree(c, c->dT);
    c->capTri = 0;
    HIPCALL(c, calloc_dev(c, &c->dT, (size_t)nf * 12));
    c->capTri = nf;
  }
  if (!c->dCheck) HIPCALL(c, calloc_dev(c, &c->dCheck, mesh_check_words() * 4));
  // through pinned staging: one host copy, then asynchronous DMA
  const size_t bytes = (size_t)nv * 12 + (size_t)nf * 12;
  if (bytes > c->capStage) {
    if (c->hStage) hipHostFree(c->hStage);
    c->hStage = nullptr;
    c->capStage = 0;
    HIPCALL(c, hipHostMalloc((void**)&c->hStage, bytes, hipHostMallocDefault));
    c->capStage = bytes;
  }
  // the previous upload from the staging buffer has finished (every call
  // below ends with a synchronization)
  std::memcpy(c->hStage, P, (size_t)nv * 12);
  std::memcpy(c->hStage + (size_t)nv * 12, T, (size_t)nf * 12);
  HIPCALL(c, hipMemcpyAsync(c->dP, c->hStage, (size_t)nv * 12, hipMemcpyHostToDevice, c->stream));
  HIPCALL(c, hipMemcpyAsync(c->dT, c->hStage + (size_t)nv * 12, (size_t)nf * 12, hipMemcpyHostToDevice, c->stream));
  // index check and bounding box on the device (kept for ComputeGridAttributes)
  HIPCALL(c, launch_mesh_check(c->dP, nv, c->dT, nf, c->dCheck, c->stream));
  HIPCALL(c, hipMemcpyAsync(c->hTotals + 4, c->dCheck, 32, hipMemcpyDeviceToHost, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  const u32* chk = c->hTotals + 4;
  if (chk[0]) return fail(c, GM_ERR_ARG, "triangle index out of range");
  if (chk[1]) return fail(c, GM_ERR_ARG, "vertex coordinate not finite");
  c->nv = nv;
  c->nf = nf;
  for (int a = 0; a < 3; ++a) {
    c->bmin[a] = mesh_key_float(chk[2 + a]);
    c->bmax[a] = mesh_key_float(chk[5 + a]);
  }
  c->cell = -1.f;  // grid attributes not computed yet
  return GM_OK;
}

}  // extern "C"

namespace {
// GM_GRID_INPUT from the loaded mesh: the conservative voxelization (parity
// fill along z plus the surface shell), or MultiSlicing (parity along x).
int voxelize_impl(gm_ctx* c, int base_res, float margin, bool multislice) {
  if (!c) return GM_ERR_ARG;
  touch(c);
  if (c->nv <= 0) return fail(c, GM_ERR_STATE, "no mesh: call gm_load_mesh first");
  hipSetDevice(c->device);
  // Voxelizer::ComputeGridAttributes (Voxelizer.cu:431-455), in float.
  float mn[3], mx[3];
  if (c->cell < 0.f) {
    std::memcpy(mn, c->bmin, sizeof(mn));
    std::memcpy(mx, c->bmax, sizeof(mx));
  } else {
    return fail(c, GM_ERR_STATE, "grid already voxelized for this mesh: call gm_load_mesh again");
  }
  int rc = alloc_grid(c, base_res);
  if (rc) return rc;
  float thick = std::max(std::fabs(mx[0] - mn[0]), std::max(std::fabs(mx[1] - mn[1]), std::fabs(mx[2] - mn[2]))) /
                (float)base_res;
  float off = 2 * thick + margin;
  for (int a = 0; a < 3; ++a) {
    mn[a] -= off;
    mx[a] += off;
  }
  float cs = std::max(std::fabs(mx[0] - mn[0]), std::max(std::fabs(mx[1] - mn[1]), std::fabs(mx[2] - mn[2]))) /
             (float)base_res;
  for (int a = 0; a < 3; ++a) {
    c->bmin[a] = mn[a];
    c->bmax[a] = mx[a];
    c->data_res[a] = (int)std::ceil((mx[a] - mn[a]) / cs);
  }
  c->cell = cs;
  float maxext = std::max(std::fabs(mx[0] - mn[0]), std::max(std::fabs(mx[1] - mn[1]), std::fabs(mx[2] - mn[2])));
  double scale = (double)c->N / (double)maxext;
  // the fill leaves the marks and the shell zero; they are cleared here only
  // after allocation or a voxelization that did not complete its launches
  const bool clear = c->voxDirty;
  c->voxDirty = true;
  hipError_t e = multislice ? launch_voxelize_multislice(c->dP, c->nv, c->dT, c->nf, c->bmin, scale, c->R, c->dSnap,
                                                         c->mark, c->I, c->stream, clear)
                            : launch_voxelize(c->dP, c->nv, c->dT, c->nf, c->bmin, scale, c->R, c->dSnap, c->mark,
                                              c->shell, c->I, c->stream, clear);
  if (e != hipSuccess) return fail(c, GM_ERR_HIP, std::string("voxelize: ") + hipGetErrorString(e));
  c->voxDirty = false;
  HIPCALL(c, hipStreamSynchronize(c->stream));
  c->meshPending = false;
  return GM_OK;
}
}  // namespace

extern "C" {

int gm_voxelize(gm_ctx* c, int base_res, float margin) { return voxelize_impl(c, base_res, margin, false); }

int gm_voxelize_multislice(gm_ctx* c, int base_res, float margin) { return voxelize_impl(c, base_res, margin, true); }

int gm_load(gm_ctx* c, const float* P, const float* N, int nv, const uint32_t* T, int nf, int base_res,
            float se_size) {
  int rc = gm_load_mesh(c, P, N, nv, T, nf);
  if (rc) return rc;
  // margin of GMorpho::Load (GMorpho.cu:7737)
  float margin = se_size + 2.f / (2 * ((float)base_res));
  return gm_voxelize(c, base_res, margin);
}

int gm_grid_init(gm_ctx* c, const float bmin[3], const float bmax[3], int base_res, const int data_res[3],
                 float cell) {
  if (!c || !bmin || !bmax || !data_res) return GM_ERR_ARG;
  touch(c);
  hipSetDevice(c->device);
  int rc = alloc_grid(c, base_res);
  if (rc) return rc;
  for (int a = 0; a < 3; ++a) {
    c->bmin[a] = bmin[a];
    c->bmax[a] = bmax[a];
    c->data_res[a] = data_res[a];
  }
  c->cell = cell;
  size_t rowBytes = (size_t)c->N * c->N * c->W * 4;
  HIPCALL(c, hipMemsetAsync(c->I, 0, rowBytes, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  return GM_OK;
}

int gm_grid_set_cells(gm_ctx* c, int which, const uint8_t* cells) {
  int rc = need_grid(c);
  if (rc) return rc;
  u32* rows = grid_rows(c, which);
  if (!rows || !cells) return fail(c, GM_ERR_ARG, "bad grid id or null buffer");
  hipSetDevice(c->device);
  HIPCALL(c, hipMemcpyAsync(c->stage, cells, (size_t)c->R * c->R * c->R, hipMemcpyHostToDevice, c->stream));
  HIPCALL(c, launch_cells_to_rows(c->stage, rows, c->R, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  c->meshPending = false;
  return GM_OK;
}

int gm_grid_get_cells(gm_ctx* c, int which, uint8_t* cells) {
  int rc = need_grid(c);
  if (rc) return rc;
  u32* rows = grid_rows(c, which);
  if (!rows || !cells) return fail(c, GM_ERR_ARG, "bad grid id or null buffer");
  hipSetDevice(c->device);
  HIPCALL(c, launch_rows_to_cells(rows, c->stage, c->R, c->stream));
  HIPCALL(c, hipMemcpyAsync(cells, c->stage, (size_t)c->R * c->R * c->R, hipMemcpyDeviceToHost, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  return GM_OK;
}

int gm_grid_info(const gm_ctx* c, float bmin[3], float bmax[3], int res[3], int data_res[3], float* cell) {
  if (!c) return GM_ERR_ARG;
  if (!c->has_grid) return GM_ERR_STATE;
  for (int a = 0; a < 3; ++a) {
    if (bmin) bmin[a] = c->bmin[a];
    if (bmax) bmax[a] = c->bmax[a];
    if (res) res[a] = c->R;
    if (data_res) data_res[a] = c->data_res[a];
  }
  if (cell) *cell = c->cell;
  return GM_OK;
}

int gm_scale_uniform(gm_ctx* c, int r) {
  int rc = need_grid(c);
  if (rc) return rc;
  touch(c);
  if (r < 0 || r > 127) return fail(c, GM_ERR_ARG, "uniform radius must be in [0,127]");
  c->rmode = 0;
  c->r = r;
  c->rlo = c->rhi = r;
  c->meshPending = false;
  c->rpValid = false;
  return GM_OK;
}

int gm_scale_grid_i8(gm_ctx* c, const int8_t* grid) {
  int rc = need_grid(c);
  if (rc) return rc;
  touch(c);
  if (!grid) return fail(c, GM_ERR_ARG, "null scale grid");
  hipSetDevice(c->device);
  size_t n = (size_t)c->N * c->N * c->N;
  rc = ensure_scale_grid(c);
  if (rc) return rc;
  HIPCALL(c, hipMemcpyAsync(c->dGrid, grid, n, hipMemcpyHostToDevice, c->stream));
  std::memset(c->hHist, 0, sizeof(c->hHist));
  for (size_t i = 0; i < n; ++i) {
    int v = grid[i];
    ++c->hHist[v < 2 ? 2 : (v > 127 ? 127 : v)];
  }
  rc = upload_hist(c);
  if (rc) return rc;
  HIPCALL(c, hipStreamSynchronize(c->stream));
  c->rmode = 1;
  set_range_from_hist(c);
  c->meshPending = false;
  c->rpValid = false;
  return GM_OK;
}

int gm_scale_grid_fill(gm_ctx* c, int value) {
  int rc = need_grid(c);
  if (rc) return rc;
  touch(c);
  if (value < -128 || value > 127) return fail(c, GM_ERR_ARG, "radius grid value must fit a signed char");
  hipSetDevice(c->device);
  rc = ensure_scale_grid(c);
  if (rc) return rc;
  const size_t n = (size_t)c->N * c->N * c->N;
  HIPCALL(c, launch_scale_fill(c->dGrid, n, (int8_t)value, c->stream));
  std::memset(c->hHist, 0, sizeof(c->hHist));
  c->hHist[value < 2 ? 2 : (value > 127 ? 127 : value)] = n;
  rc = upload_hist(c);
  if (rc) return rc;
  HIPCALL(c, hipStreamSynchronize(c->stream));
  c->rmode = 1;
  set_range_from_hist(c);
  c->meshPending = false;
  c->rpValid = false;
  return GM_OK;
}

int gm_scale_brush(gm_ctx* c, const float position[3], float scale, float support, float global_scale, int mode) {
  int rc = need_grid(c);
  if (rc) return rc;
  touch(c);
  if (!position) return fail(c, GM_ERR_ARG, "null brush position");
  if (mode < GM_BRUSH_NONE || mode > GM_BRUSH_ZERO) return fail(c, GM_ERR_ARG, "unknown brush mode");
  if (!(c->cell > 0.f)) return fail(c, GM_ERR_STATE, "grid attributes not set");
  hipSetDevice(c->device);
  if (c->rmode != 1) {
    // the reference's scale grid always exists: a uniform radius is a filled grid
    rc = gm_scale_grid_fill(c, c->rmode == 0 ? std::min(c->r, 127) : 0);
    if (rc) return rc;
  }
  if (mode == GM_BRUSH_NONE) return GM_OK;
  ScaleBrush br{};
  const float fc = 0.5f * c->cell;  // the scale field's cell: half a grid cell
  for (int a = 0; a < 3; ++a) {
    br.bmin[a] = c->bmin[a];
    br.pos[a] = position[a];
  }
  br.cell = fc;
  br.scale = scale;
  volatile float rt = std::sqrt(1.f - std::sqrt(0.5f));
  br.h = (support - scale) / (float)rt;
  br.global_scale = global_scale;
  br.mode = mode;
  // voxels whose sample point can lie closer than scale + h
  const double reach = (double)scale + (double)br.h;
  int box[6] = {0, 0, 0, 0, 0, 0};
  bool empty = !(reach > 0.0);
  for (int a = 0; a < 3 && !empty; ++a) {
    const double lo = std::ceil(((double)position[a] - c->bmin[a] - reach) / fc) - 1;
    const double hi = std::floor(((double)position[a] - c->bmin[a] + reach) / fc) + 1;
    const int l = (int)std::max(0.0, std::min(lo, (double)c->N));
    const int h = (int)std::max(-1.0, std::min(hi, (double)c->N - 1));
    if (h < l) empty = true;
    box[a] = l;
    box[3 + a] = h - l + 1;
  }
  if (empty) return GM_OK;
  HIPCALL(c, launch_scale_brush(c->dGrid, c->N, br, box, c->dHist, c->stream));
  HIPCALL(c, hipMemcpyAsync(c->hHist, c->dHist, sizeof(c->hHist), hipMemcpyDeviceToHost, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  const int oldHi = c->rhi;
  set_range_from_hist(c);
  c->meshPending = false;
  if (c->rpValid && c->capVarR >= c->R && c->rhi <= 31 && oldHi <= 31) {
    // rebuild the radius planes of the bricks the brush touched only
    const int bb[6] = {box[0] / 32, box[1] / 8, box[2] / 8, (box[0] + box[3] - 1) / 32 - box[0] / 32 + 1,
                       (box[1] + box[4] - 1) / 8 - box[1] / 8 + 1, (box[2] + box[5] - 1) / 8 - box[2] / 8 + 1};
    HIPCALL(c, launch_radius_planes(radius_field(c), c->N, c->rp, c->rvset, c->stream, bb));
  } else {
    c->rpValid = false;
  }
  return GM_OK;
}

int gm_scale_grid_get(gm_ctx* c, int8_t* grid) {
  int rc = need_grid(c);
  if (rc) return rc;
  if (!grid) return fail(c, GM_ERR_ARG, "null buffer");
  if (c->rmode != 1) return fail(c, GM_ERR_STATE, "the radius field is not a scale grid");
  hipSetDevice(c->device);
  HIPCALL(c, hipMemcpyAsync(grid, c->dGrid, (size_t)c->N * c->N * c->N, hipMemcpyDeviceToHost, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  return GM_OK;
}

int gm_scale_lattice_f32(gm_ctx* c, const float* lat, int L) {
  int rc = need_grid(c);
  if (rc) return rc;
  // a new field of the same lattice size, radius range and arithmetic width
  // keeps the launch arguments (the kernels read the lattice through its
  // device pointer): a captured closing is then replayed, not captured again
  const RadiusField before = radius_field(c);
  const int beforeLo = c->rlo, beforeHi = c->rhi;
  if (!lat || L < 2 || L > c->N || (c->N % L) != 0 || ((c->N / L) & (c->N / L - 1)) != 0)
    return fail(c, GM_ERR_ARG, "lattice size must divide N with a power-of-two ratio");
  hipSetDevice(c->device);
  size_t n = (size_t)L * L * L;
  std::vector<int32_t> q(n);
  long long mn = (1LL << 40), mx = -(1LL << 40);
  for (size_t i = 0; i < n; ++i) {
    double v = std::floor((double)lat[i] * 256.0 + 0.5);
    if (v > 2e9 || v < -2e9) return fail(c, GM_ERR_ARG, "lattice value out of range");
    q[i] = (int32_t)v;
    mn = std::min<long long>(mn, q[i]);
    mx = std::max<long long>(mx, q[i]);
  }
  if (c->L != L) {
    touch(c);  // the lattice moves
    cfree(c, c->dLat);
    HIPCALL(c, calloc_dev(c, &c->dLat, n * 4));
  }
  HIPCALL(c, hipMemcpyAsync(c->dLat, q.data(), n * 4, hipMemcpyHostToDevice, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  auto fl = [](long long v) { return v >= 0 ? v / 256 : -((-v + 255) / 256); };
  c->L = L;
  c->latMaxAbs = std::max(std::llabs(mn), std::llabs(mx));
  c->rmode = 2;
  c->rpValid = false;
  c->rlo = (int)std::max<long long>(2, std::min<long long>(127, fl(mn)));
  c->rhi = (int)std::max<long long>(2, std::min<long long>(127, fl(mx)));
  c->meshPending = false;
  const RadiusField after = radius_field(c);
  if (before.mode != after.mode || before.L != after.L || before.lat != after.lat || before.narrow != after.narrow ||
      beforeLo != c->rlo || beforeHi != c->rhi)
    touch(c);
  return GM_OK;
}

int gm_dilate(gm_ctx* c, int se) {
  int rc = need_grid(c);
  if (rc) return rc;
  if (se != GM_SE_SPHERE && se != GM_SE_CUBE) return fail(c, GM_ERR_ARG, "unknown structuring element");
  hipSetDevice(c->device);
  c->tuneSe = se;
  c->evValid = 0;
  rc = do_dilate(c, se);
  if (rc) return rc;
  HIPCALL(c, hipStreamSynchronize(c->stream));
  return GM_OK;
}

int gm_erode(gm_ctx* c) {
  int rc = need_grid(c);
  if (rc) return rc;
  hipSetDevice(c->device);
  c->evValid = 0;
  rc = do_erode(c);
  if (rc) return rc;
  HIPCALL(c, hipStreamSynchronize(c->stream));
  return GM_OK;
}

// The closing's set-up shared by gm_close and gm_update: *copied = true when
// the radius skips the morphology (the closing is the input, copied here);
// otherwise the passes are prepared and key is the captured sequence's key.
static int close_setup(gm_ctx* c, int se, bool* copied, uint64_t key[4]) {
  int rc = need_grid(c);
  if (rc) return rc;
  if (se != GM_SE_SPHERE && se != GM_SE_CUBE) return fail(c, GM_ERR_ARG, "unknown structuring element");
  hipSetDevice(c->device);
  c->tuneSe = se;
  c->evValid = 0;
  *copied = false;
  size_t rowBytes = (size_t)c->N * c->N * c->W * 4;
  if (c->rmode == 0 && c->r <= 1) {
    // GMorpho.cu:7784, 7803: no morphology below two fine voxels.
    HIPCALL(c, hipMemcpyAsync(c->D, c->I, rowBytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCALL(c, hipMemsetAsync(c->K, 0, rowBytes, c->stream));
    HIPCALL(c, hipMemcpyAsync(c->C, c->I, rowBytes, hipMemcpyDeviceToDevice, c->stream));
    c->meshPending = false;
    *copied = true;
    return GM_OK;
  }
  rc = prep_dilate(c, se);
  if (rc == GM_OK) rc = prep_erode(c);
  if (rc) return rc;
  // the tile-queue toggle returns to its value after the four passes, so a
  // captured sequence stays valid while it starts from the same value
  const MorphBuffers shape = morph_buffers(c);
  key[0] = c->graphGen;
  key[1] = (uint64_t)se | ((uint64_t)shape.tileZ << 8) | ((uint64_t)shape.waveCap << 16);
  key[2] = (uint64_t)c->qsel;
  key[3] = (uint64_t)(c->rpValid ? 1 : 0);
  return GM_OK;
}

// Marching cubes of C and the copy of its totals, enqueued (GMorpho.cu:6814-7001).
static int enq_mesh(gm_ctx* c) {
  hipError_t e = launch_mc(c->C, c->R, c->bmin, c->cell, c->mc, c->stream, c->timing ? &c->ev[10] : nullptr);
  if (e != hipSuccess) return fail(c, GM_ERR_HIP, std::string("mesh extraction: ") + hipGetErrorString(e));
  HIPCALL(c, hipMemcpyAsync(c->hTotals, c->mc.totals, 16, hipMemcpyDeviceToHost, c->stream));
  return GM_OK;
}

int gm_close(gm_ctx* c, int se) {
  bool copied = false;
  uint64_t key[4] = {0, 0, 0, 0};
  int rc = close_setup(c, se, &copied, key);
  if (rc || copied) return rc;
  rc = run_captured(c, c->gClose[se], key, [&]() -> int {
    int r = enq_dilate(c, se);
    return r == GM_OK ? enq_erode(c) : r;
  });
  if (rc) return rc;
  after_morph(c, true);
  if (c->timing) c->evValid |= 1;
  return GM_OK;
}

int gm_extract_mesh(gm_ctx* c) {
  int rc = need_grid(c);
  if (rc) return rc;
  hipSetDevice(c->device);
  HIPCALL(c, mc_prepare());
  const uint64_t key[4] = {c->graphGen, 0, 0, 0};
  rc = run_captured(c, c->gMesh, key, [&]() -> int { return enq_mesh(c); });
  if (rc) return rc;
  c->meshPending = true;
  if (c->timing) c->evValid |= 2;
  return GM_OK;
}

int gm_update(gm_ctx* c, int se) {
  bool copied = false;
  uint64_t key[4] = {0, 0, 0, 0};
  int rc = close_setup(c, se, &copied, key);
  if (rc) return rc;
  if (copied) return gm_extract_mesh(c);
  HIPCALL(c, mc_prepare());
  rc = run_captured(c, c->gUpdate[se], key, [&]() -> int {
    int r = enq_dilate(c, se);
    if (r == GM_OK) r = enq_erode(c);
    return r == GM_OK ? enq_mesh(c) : r;
  });
  if (rc) return rc;
  after_morph(c, true);
  c->meshPending = true;
  if (c->timing) c->evValid |= 3;
  return GM_OK;
}

int gm_mesh_counts(gm_ctx* c, uint32_t* nv, uint32_t* nt) {
  int rc = need_grid(c);
  if (rc) return rc;
  if (!c->meshPending) return fail(c, GM_ERR_STATE, "no mesh extracted since the last grid change");
  hipSetDevice(c->device);
  HIPCALL(c, hipStreamSynchronize(c->stream));
  if (c->hTotals[2] & 4u) {
    // more listed blocks than the per-block buffers hold: grow them and
    // extract again (nothing was stored)
    u32 listed = 0;
    HIPCALL(c, hipMemcpyAsync(&listed, c->mc.totals + 5, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    rc = alloc_mc_list(c, std::min<u32>((u32)c->mc.nblocks, grow_cap(c->mc.capList, listed)));
    if (rc) return rc;
    hipError_t e = launch_mc(c->C, c->R, c->bmin, c->cell, c->mc, c->stream, nullptr);
    if (e != hipSuccess) return fail(c, GM_ERR_HIP, std::string("mesh extraction: ") + hipGetErrorString(e));
    HIPCALL(c, hipMemcpyAsync(c->hTotals, c->mc.totals, 16, hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    if (c->hTotals[2] & 4u) return fail(c, GM_ERR_CAPACITY, "marching-cubes block list overflow");
  }
  // grow and re-emit when the output did not fit
  for (int attempt = 0; attempt < 3 && c->hTotals[2] != 0; ++attempt) {
    u32 lateN = 0;
    HIPCALL(c, hipMemcpyAsync(&lateN, c->mc.late, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    rc = alloc_mesh_out(c, std::max(c->mc.capV, grow_cap(c->mc.capV, c->hTotals[0])),
                        std::max(c->mc.capT, grow_cap(c->mc.capT, c->hTotals[1])),
                        std::max(c->mc.capLate, grow_cap(c->mc.capLate, lateN)));
    if (rc) return rc;
    u32 tot[4] = {c->hTotals[0], c->hTotals[1], 0u, 0u};
    HIPCALL(c, hipMemcpyAsync(c->mc.totals, tot, 16, hipMemcpyHostToDevice, c->stream));
    // the emit appends to the late list from zero (the count kernel cleared
    // it before the first emit; a regrown list is fresh memory)
    const u32 zero = 0u;
    HIPCALL(c, hipMemcpyAsync(c->mc.late, &zero, 4, hipMemcpyHostToDevice, c->stream));
    hipError_t e = launch_mc_emit(c->C, c->R, c->bmin, c->cell, c->mc, c->stream, nullptr);
    if (e != hipSuccess) return fail(c, GM_ERR_HIP, std::string("mesh extraction: ") + hipGetErrorString(e));
    HIPCALL(c, hipMemcpyAsync(c->hTotals, c->mc.totals, 16, hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
  }
  if (c->hTotals[2] != 0) return fail(c, GM_ERR_CAPACITY, "mesh output overflow");
  if (nv) *nv = c->hTotals[0];
  if (nt) *nt = c->hTotals[1];
  return GM_OK;
}

int gm_mesh_copy(gm_ctx* c, float* V, float* Nrm, uint32_t* T) {
  uint32_t nv = 0, nt = 0;
  int rc = gm_mesh_counts(c, &nv, &nt);
  if (rc) return rc;
  if (V && nv) HIPCALL(c, hipMemcpyAsync(V, c->mc.V, (size_t)nv * 12, hipMemcpyDeviceToHost, c->stream));
  if (Nrm && nv) HIPCALL(c, hipMemcpyAsync(Nrm, c->mc.Nrm, (size_t)nv * 12, hipMemcpyDeviceToHost, c->stream));
  if (T && nt) HIPCALL(c, hipMemcpyAsync(T, c->mc.T, (size_t)nt * 12, hipMemcpyDeviceToHost, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  return GM_OK;
}

int gm_filter_normals(gm_ctx* c, int iters) {
  if (!c) return GM_ERR_ARG;
  if (iters < 0) return fail(c, GM_ERR_ARG, "negative iteration count");
  uint32_t nv = 0, nt = 0;
  int rc = gm_mesh_counts(c, &nv, &nt);
  if (rc) return rc;
  if (iters == 0 || nv == 0) return GM_OK;
  const size_t words = bilateral_scratch_words(nv, nt);
  if (words > c->bilWords) {
    cfree(c, c->bilScratch);
    c->bilWords = 0;
    HIPCALL(c, calloc_dev(c, &c->bilScratch, words * 4));
    c->bilWords = words;
  }
  if (nv > c->bilNCap) {
    cfree(c, c->bilN);
    c->bilNCap = 0;
    HIPCALL(c, calloc_dev(c, &c->bilN, (size_t)nv * 12));
    c->bilNCap = nv;
  }
  hipError_t e = launch_bilateral(c->mc.V, c->mc.Nrm, c->mc.T, nv, nt, iters, c->bilScratch, c->bilN, c->stream);
  if (e != hipSuccess) return fail(c, GM_ERR_HIP, std::string("normal filtering: ") + hipGetErrorString(e));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  return GM_OK;
}

int gm_build_pyramid(gm_ctx* c, int which) {
  int rc = need_grid(c);
  if (rc) return rc;
  u32* rows = grid_rows(c, which);
  if (!rows) return fail(c, GM_ERR_ARG, "bad grid id");
  if (c->R & (c->R - 1)) return fail(c, GM_ERR_UNSUPPORTED, "the max-mipmap needs a power-of-two resolution");
  hipSetDevice(c->device);
  int levels = 0;
  uint64_t tex = 0;
  for (int s = c->R; s >= 1; s >>= 1) {
    ++levels;
    tex += (uint64_t)s * s * s;
  }
  if (tex > c->capPyr) {
    cfree(c, c->dPyr);
    c->capPyr = 0;
    HIPCALL(c, calloc_dev(c, &c->dPyr, tex * 8));
    c->capPyr = tex;
  }
  HIPCALL(c, launch_pyramid(rows, c->R, radius_field(c), c->dPyr, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  c->pyrLevels = levels;
  c->pyrTexels = tex;
  return GM_OK;
}

int gm_pyramid_size(const gm_ctx* c, int* levels, uint64_t* texels) {
  if (!c) return GM_ERR_ARG;
  if (!c->pyrLevels) return GM_ERR_STATE;
  if (levels) *levels = c->pyrLevels;
  if (texels) *texels = c->pyrTexels;
  return GM_OK;
}

int gm_pyramid_get(gm_ctx* c, uint64_t* out) {
  if (!c || !out) return GM_ERR_ARG;
  if (!c->pyrLevels) return fail(c, GM_ERR_STATE, "no pyramid built");
  HIPCALL(c, hipMemcpyAsync(out, c->dPyr, c->pyrTexels * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  return GM_OK;
}

uint64_t gm_ctx_device_bytes(const gm_ctx* c) { return c ? c->devBytes : 0; }

// Frees the context's grid buffers (a job engine releasing an idle
// context's memory for a job that would not fit the budget otherwise).
int gm_ctx_release(gm_ctx* c) {
  if (!c) return GM_ERR_ARG;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  free_grid(c);
  return GM_OK;
}

// Buffers for grids up to base resolution R, allocated now (a job engine
// sizes a context once for the largest job it will batch, so that jobs of
// mixed sizes never reallocate: every hipFree waits for the whole device).
// The context holds no grid afterwards.
int gm_ctx_reserve(gm_ctx* c, int R) {
  if (!c) return GM_ERR_ARG;
  if (R <= c->capR) return GM_OK;
  hipSetDevice(c->device);
  const int rc = alloc_grid(c, R);
  c->has_grid = false;
  return rc;
}

int gm_set_graphs(gm_ctx* c, int enable) {
  if (!c || enable < 0 || enable > 2) return GM_ERR_ARG;
  c->graphs = enable;
  return GM_OK;
}

int gm_set_timing(gm_ctx* c, int enable) {
  if (!c) return GM_ERR_ARG;
  c->timing = enable != 0;
  c->evValid = 0;
  return GM_OK;
}

constexpr int kNumStages = 16;
int gm_stage_times(gm_ctx* c, float* ms, int n) {
  if (!c || !ms) return GM_ERR_ARG;
  HIPCALL(c, hipStreamSynchronize(c->stream));
  // events: dilate 0..4, erode 5..9, mc 10..14 (one stage per kernel or
  // scan); then the four stream kernels alone (from 15..18, recorded after
  // each pass's tile selection, to the end of the pass)
  const int pairs[kNumStages][2] = {{0, 1},   {1, 2},   {2, 3},   {3, 4},  {5, 6},  {6, 7},  {7, 8},  {8, 9},
                                    {10, 11}, {11, 12}, {12, 13}, {13, 14}, {15, 2}, {16, 4}, {17, 7}, {18, 9}};
  for (int i = 0; i < n && i < kNumStages; ++i) {
    ms[i] = -1.f;
    bool ok = (i < 8 || i >= 12) ? (c->evValid & 1) : (c->evValid & 2);
    if (!ok) continue;
    float t = 0.f;
    if (hipEventElapsedTime(&t, c->ev[pairs[i][0]], c->ev[pairs[i][1]]) == hipSuccess) ms[i] = t;
  }
  return GM_OK;
}

// ---- Launch-shape tuning of the streaming passes.
namespace {
const int kShapes[][2] = {{kTileZ, 8}, {kTileZ, 4}, {kTileZ, 2}, {kTileZMin, 8}, {kTileZMin, 4}, {kTileZMin, 2}};

uint64_t grid_hash(gm_ctx* c, u32* rows) {
  // FNV-1a over the closing grid (device -> pinned staging of the mesh check)
  const size_t words = (size_t)c->N * c->N * c->W;
  std::vector<u32> h(words);
  if (hipMemcpyAsync(h.data(), rows, words * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return 0;
  uint64_t x = 1469598103934665603ull;
  for (u32 w : h) x = (x ^ w) * 1099511628211ull;
  return x;
}

void write_tune_file(const char* path) {
  std::lock_guard<std::mutex> lk(g_tuneMu);
  FILE* f = std::fopen(path, "w");
  if (!f) return;
  for (const auto& kv : g_tune)
    std::fprintf(f, "{\"base_res\": %d, \"radius\": %d, \"se\": %d, \"tile_planes\": %d, \"waves\": %d, \"ms\": %.5f}\n",
                 kv.first.R, kv.first.r, kv.first.se, kv.second.tileZ, kv.second.waveCap, kv.second.ms);
  std::fclose(f);
}
}  // namespace

int gm_autotune(gm_ctx* c, int se, int reps, const char* path, int* tile_planes, int* waves) {
  int rc = need_grid(c);
  if (rc) return rc;
  if (se != GM_SE_SPHERE && se != GM_SE_CUBE) return fail(c, GM_ERR_ARG, "unknown structuring element");
  if (c->rmode == 0 && c->r <= 1) return fail(c, GM_ERR_STATE, "radius <= 1: no morphology to tune");
  if (use_edt(c)) return fail(c, GM_ERR_STATE, "radii above 31 run the distance-transform passes: no launch shape to tune");
  if (reps < 1) reps = 1;
  hipSetDevice(c->device);
  const bool timing = c->timing;
  c->timing = false;
  double best = 1e30;
  int bestZ = 0, bestW = 0;
  uint64_t ref = 0;
  for (const auto& shp : kShapes) {
    c->forceTileZ = shp[0];
    c->forceWaveCap = shp[1];
    touch(c);
    rc = gm_close(c, se);  // warm (capture, code objects)
    if (rc == GM_OK) rc = gm_sync(c);
    if (rc) break;
    const uint64_t hsh = grid_hash(c, c->C);
    if (!ref) ref = hsh;
    if (hsh != ref) {  // every shape must give the same closing
      rc = fail(c, GM_ERR_STATE, "launch shape changed the closing");
      break;
    }
    std::vector<double> t;
    for (int k = 0; k < reps && rc == GM_OK; ++k) {
      const auto t0 = std::chrono::steady_clock::now();
      rc = gm_close(c, se);
      if (rc == GM_OK) rc = gm_sync(c);
      t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    if (rc) break;
    std::sort(t.begin(), t.end());
    const double ms = t[t.size() / 2];
    if (ms < best) {
      best = ms;
      bestZ = shp[0];
      bestW = shp[1];
    }
  }
  c->forceTileZ = c->forceWaveCap = 0;
  c->timing = timing;
  touch(c);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(g_tuneMu);
    g_tune[TuneKey{c->R, c->rmode == 0 ? c->r : -1, se}] = TuneShape{bestZ, bestW, best};
  }
  if (path && *path) write_tune_file(path);
  if (tile_planes) *tile_planes = bestZ;
  if (waves) *waves = bestW;
  return GM_OK;
}

int gm_tune_load(const char* path) {
  if (!path) return GM_ERR_ARG;
  FILE* f = std::fopen(path, "r");
  if (!f) return GM_ERR_ARG;
  char line[512];
  int n = 0;
  while (std::fgets(line, sizeof(line), f)) {
    int R = 0, r = 0, se = 0, z = 0, w = 0;
    double ms = 0.0;
    if (std::sscanf(line, " {\"base_res\": %d, \"radius\": %d, \"se\": %d, \"tile_planes\": %d, \"waves\": %d, \"ms\": %lf",
                    &R, &r, &se, &z, &w, &ms) == 6 &&
        (z == kTileZ || z == kTileZMin) && w >= 1 && w <= 8 && (se == 0 || se == 1)) {
      std::lock_guard<std::mutex> lk(g_tuneMu);
      g_tune[TuneKey{R, r, se}] = TuneShape{z, w, ms};
      ++n;
    }
  }
  std::fclose(f);
  return n > 0 ? GM_OK : GM_ERR_ARG;
}

// ---- Batched closing + mesh of many contexts (one launch per step).
struct gm_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  void *hMorph = nullptr, *dMorph = nullptr;
  size_t capMorph = 0;
  void *hMc = nullptr, *dMc = nullptr;
  size_t capMc = 0;
  u32* hTot = nullptr;
  size_t capTot = 0;
  u32* queue = nullptr;
  size_t capQueue = 0;  // words
  int qsel = 0;
  // the captured launch sequence of the last batch and what it was built for
  // (the contexts, their launch-argument generations, the element, the
  // queue toggle): replayed when the same batch is closed again
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> key, seen;
  int graphs = 1;  // as gm_ctx::graphs
  SideStreams side;
};

namespace {
int bfail(gm_batch* b, int code, const std::string& msg) {
  b->err = "GMorpho Error: " + msg;
  return code;
}
#define BHIP(b, x)                                                                                  \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess) return bfail(b, GM_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
// Pinned host + device pair of at least `bytes`, kept at the largest size.
int batch_blobs(gm_batch* b, void*& h, void*& d, size_t& cap, size_t bytes) {
  if (bytes <= cap) return GM_OK;
  if (h) hipHostFree(h);
  if (d) hipFree(d);
  h = d = nullptr;
  cap = 0;
  const size_t want = bytes + bytes / 4;
  BHIP(b, hipHostMalloc(&h, want, hipHostMallocDefault));
  BHIP(b, hipMalloc(&d, want));
  cap = want;
  return GM_OK;
}
}  // namespace

int gm_sync(gm_ctx* c) {
  if (!c) return GM_ERR_ARG;
  HIPCALL(c, hipStreamSynchronize(c->stream));
  return GM_OK;
}

int gm_get_stream(gm_ctx* c, void** s) {
  if (!c || !s) return GM_ERR_ARG;
  *s = (void*)c->stream;
  return GM_OK;
}

int gm_diag_copy_probe(int device, uint64_t bytes, int reps) {
  if (bytes == 0 || (bytes % 16) != 0 || reps < 1) return GM_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return GM_ERR_HIP;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) return GM_ERR_HIP;
  if (hipMalloc(&b, bytes) != hipSuccess) {
    hipFree(a);
    return GM_ERR_HIP;
  }
  hipError_t e = hipMemset(a, 1, bytes);
  for (int k = 0; k < reps && e == hipSuccess; ++k) e = launch_copy_probe(a, b, bytes, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  hipFree(a);
  hipFree(b);
  return e == hipSuccess ? GM_OK : GM_ERR_HIP;
}

int gm_mc_tables(uint32_t* ntri, uint32_t* len, uint64_t* tri) {
  if (!ntri || !len || !tri) return GM_ERR_ARG;
  mc_tables_host(ntri, len, (unsigned long long*)tri);
  return GM_OK;
}


int gm_batch_create(int device, gm_batch** out) {
  if (!out) return GM_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return GM_ERR_HIP;
  if (device < 0 || device >= n) return GM_ERR_ARG;
  gm_batch* b = new gm_batch();
  b->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return GM_ERR_HIP;
  }
  bool ok = hipEventCreateWithFlags(&b->side.fork, hipEventDisableTiming) == hipSuccess;
  for (int k = 0; k < 3 && ok; ++k) {
    ok = hipStreamCreateWithFlags(&b->side.s[k], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&b->side.join[k], hipEventDisableTiming) == hipSuccess;
    if (ok) b->side.n = k + 1;
  }
  if (!ok) {
    gm_batch_destroy(b);
    return GM_ERR_HIP;
  }
  *out = b;
  return GM_OK;
}

int gm_batch_destroy(gm_batch* b) {
  if (!b) return GM_ERR_ARG;
  hipSetDevice(b->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->hMorph) hipHostFree(b->hMorph);
  if (b->hMc) hipHostFree(b->hMc);
  if (b->hTot) hipHostFree(b->hTot);
  dfree(b->dMorph);
  dfree(b->dMc);
  dfree(b->queue);
  if (b->exec) hipGraphExecDestroy(b->exec);
  for (int k = 0; k < 3; ++k) {
    if (b->side.s[k]) {
      hipStreamSynchronize(b->side.s[k]);
      hipStreamDestroy(b->side.s[k]);
    }
    if (b->side.join[k]) hipEventDestroy(b->side.join[k]);
  }
  if (b->side.fork) hipEventDestroy(b->side.fork);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return GM_OK;
}

const char* gm_batch_last_error(const gm_batch* b) { return b ? b->err.c_str() : ""; }

int gm_batch_set_graphs(gm_batch* b, int enable) {
  if (!b || enable < 0 || enable > 2) return GM_ERR_ARG;
  b->graphs = enable;
  return GM_OK;
}

int gm_batch_close_mesh(gm_batch* b, gm_ctx* const* ctxs, int n, int se) {
  if (!b || (!ctxs && n > 0) || n < 0) return GM_ERR_ARG;
  if (se != GM_SE_SPHERE && se != GM_SE_CUBE) return bfail(b, GM_ERR_ARG, "unknown structuring element");
  if (n > (1 << 12)) return bfail(b, GM_ERR_UNSUPPORTED, "more than 4096 contexts in one batch");
  hipSetDevice(b->device);
  static const bool ctOk = ct_tables_ok();
  if (!ctOk) return bfail(b, GM_ERR_UNSUPPORTED, "compiled chord tables differ from the structuring element");
  std::vector<gm_ctx*> grouped, single;
  for (int i = 0; i < n; ++i) {
    gm_ctx* c = ctxs[i];
    if (!c || !c->has_grid) return bfail(b, GM_ERR_STATE, "context without a grid in the batch");
    if (c->device != b->device) return bfail(b, GM_ERR_ARG, "context on another device");
    c->evValid = 0;
    if (se == GM_SE_SPHERE && c->rmode == 0 && c->r >= 2 && c->r <= 31) grouped.push_back(c);
    else single.push_back(c);
  }
  // the contexts' own streams are idle before the batch stream touches their buffers
  for (int i = 0; i < n; ++i) BHIP(b, hipStreamSynchronize(ctxs[i]->stream));
  for (gm_ctx* c : single) {
    const int rc = gm_update(c, se);
    if (rc != GM_OK) return bfail(b, rc, c->err);
  }
  const int m = (int)grouped.size();
  if (m > 0) {
    std::vector<BatchMorphJob> mj(m);
    std::vector<BatchMCJob> cj(m);
    for (int k = 0; k < m; ++k) {
      gm_ctx* c = grouped[k];
      mj[k].d = morph_dims(c);
      mj[k].b = morph_buffers(c);
      mj[k].b.tileZ = 0;  // the batched launch counts tiles of the default depth
      mj[k].b.waveCap = 0;
      mj[k].r = c->r;
      cj[k].C = c->C;
      cj[k].R = c->R;
      for (int a = 0; a < 3; ++a) cj[k].bmin[a] = c->bmin[a];
      cj[k].cell = c->cell;
      cj[k].m = c->mc;
    }
    int rc = batch_blobs(b, b->hMorph, b->dMorph, b->capMorph, batch_morph_bytes(m));
    if (rc) return rc;
    rc = batch_blobs(b, b->hMc, b->dMc, b->capMc, batch_mc_bytes(m));
    if (rc) return rc;
    if ((size_t)m * 4 > b->capTot) {
      if (b->hTot) hipHostFree(b->hTot);
      b->hTot = nullptr;
      b->capTot = 0;
      BHIP(b, hipHostMalloc((void**)&b->hTot, (size_t)m * 16, hipHostMallocDefault));
      b->capTot = (size_t)m * 4;
    }
    BHIP(b, mc_prepare());
    const size_t cap = batch_queue_cap(mj.data(), m);
    const size_t words = kQueueLists + 256 * cap;
    if (words > b->capQueue) {
      dfree(b->queue);
      b->capQueue = 0;
      BHIP(b, hipMalloc(&b->queue, words * 4));
      // both counter blocks start at zero (each selection clears the other)
      BHIP(b, hipMemsetAsync(b->queue, 0, kQueueLists * 4, b->stream));
      b->capQueue = words;
      b->qsel = 0;
    }
    std::vector<uint64_t> key = {(uint64_t)m, (uint64_t)se, (uint64_t)b->qsel, (uint64_t)(uintptr_t)b->queue,
                                 (uint64_t)(uintptr_t)b->dMorph, (uint64_t)(uintptr_t)b->dMc};
    for (gm_ctx* c : grouped) {
      key.push_back((uint64_t)(uintptr_t)c);
      key.push_back(c->graphGen);
    }
    auto enqueue = [&]() -> int {
      hipError_t e =
          launch_close_batched(mj.data(), m, b->hMorph, b->dMorph, b->queue, (int)cap, &b->qsel, b->stream, b->side);
      if (e != hipSuccess) return bfail(b, GM_ERR_HIP, std::string("batched closing: ") + hipGetErrorString(e));
      e = launch_mc_batched(cj.data(), m, b->hMc, b->dMc, b->hTot, b->stream);
      if (e != hipSuccess) return bfail(b, GM_ERR_HIP, std::string("batched mesh extraction: ") + hipGetErrorString(e));
      return GM_OK;
    };
    const bool firstSeen = b->graphs == 2 && !(b->exec && key == b->key) && key != b->seen;
    if (firstSeen) {
      // the eager launches below rewrite the pinned tables a captured batch
      // would copy from: that capture is stale from here on
      b->seen = key;
      if (b->exec) hipGraphExecDestroy(b->exec);
      b->exec = nullptr;
      b->key.clear();
    }
    if (!b->graphs || firstSeen) {
      int rc2 = enqueue();
      if (rc2) return rc2;
    } else if (b->exec && key == b->key) {
      // the host tables of this batch are still in the pinned blobs the
      // captured copies read
      BHIP(b, hipGraphLaunch(b->exec, b->stream));
    } else {
      if (b->exec) hipGraphExecDestroy(b->exec);
      b->exec = nullptr;
      b->key.clear();
      BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
      const int rc2 = enqueue();
      hipGraph_t graph = nullptr;
      const hipError_t ec = hipStreamEndCapture(b->stream, &graph);
      if (rc2 != GM_OK || ec != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        if (rc2) return rc2;
        return bfail(b, GM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
      }
      const hipError_t ei = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        b->exec = nullptr;
        return bfail(b, GM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
      }
      b->key = key;
      BHIP(b, hipGraphLaunch(b->exec, b->stream));
    }
    BHIP(b, hipStreamSynchronize(b->stream));
    for (int k = 0; k < m; ++k) {
      std::memcpy(grouped[k]->hTotals, b->hTot + 4 * k, 16);
      grouped[k]->meshPending = true;
    }
  }
  for (gm_ctx* c : single) BHIP(b, hipStreamSynchronize(c->stream));
  return GM_OK;
}

}  // extern "C"
