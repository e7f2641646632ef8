// Marching cubes on the closed bit grid (ExtractClosingMesh, GMorpho.cu:6814-7001;
// MarchingCubesMesher::createMesh3D, MarchingCubesMesher.cu:1297-1618).
//
// Output order is the reference's: cells of the MCM contour (cells in
// [1,R-2]^3 with a mixed fine cube, Grid.cu:1329-1462) sorted by Morton code,
// the 8 fine cubes of a cell in sub-voxel order, mixed cubes only; each cube
// owns the vertices on its edges 0, 3, 8 (in that order) and emits its
// triangles from the triangulation table with reversed winding.
//
// MI355X design: the Morton order is walked directly. A workgroup owns an
// aligned 8x8x8 block of cells (512 consecutive Morton codes), so the
// reference's tag / iota / copy_if / sort_by_key / scan chain becomes
//   mc_count  (per block vertex/triangle totals)
//   mc_scan   (exclusive scan of the block totals, one workgroup)
//   mc_verts  (block-local scan, vertex positions, per-cell vertex base)
//   mc_tris   (triangles; neighbour vertex ids from the per-cell base)
//   mc_normals(per vertex gather over the <= 4 cubes sharing its edge)
//   mc_fixup  (triangles the count table does not cover), mc_normalize.
// The per-cell vertex base map is dense (R^3 u32) but only written and read
// at surface cells, replacing the reference's 4*N^3-byte memset + splat.
#include "gm_internal.h"

namespace gm {

namespace {

#include "mc_tables.inc"

struct McConst {
  u32 ntri[256];  // triangles counted per case (numTriTable semantics)
  u32 len[256];   // triangles listed per case
};
__constant__ McConst g_mc;
__constant__ unsigned long long g_tri[256];

constexpr int kMcThreads = 256;
constexpr int kBox = 24;  // fine voxels per axis held in LDS (block 16 + halo)
constexpr int kLo = 2;    // halo below the block

__device__ __forceinline__ u32 part1by2(u32 x) {
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0xff0000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}
__device__ __forceinline__ u32 compact1by2(u32 x) {
  x &= 0x09249249u;
  x = (x | (x >> 2)) & 0x030c30c3u;
  x = (x | (x >> 4)) & 0x0300f00fu;
  x = (x | (x >> 8)) & 0xff0000ffu;
  x = (x | (x >> 16)) & 0x000003ffu;
  return x;
}

struct Box {
  u32 rows[kBox * kBox];  // bit i = fine x0+i of row (y0+j, z0+k)
  int x0, y0, z0;
  __device__ __forceinline__ int at(int x, int y, int z) const {
    return (rows[(y - y0) + kBox * (z - z0)] >> (x - x0)) & 1;
  }
};

__device__ void load_box(Box& bx, const u32* __restrict__ C, int N, int fx0, int fy0, int fz0) {
  const int W = N / 32;
  if (threadIdx.x == 0) {
    bx.x0 = fx0;
    bx.y0 = fy0;
    bx.z0 = fz0;
  }
  int w0 = (fx0 >= 0) ? (fx0 >> 5) : -((31 - fx0) >> 5);
  int o = fx0 - 32 * w0;
  for (int i = threadIdx.x; i < kBox * kBox; i += blockDim.x) {
    int y = fy0 + (i % kBox), z = fz0 + (i / kBox);
    u32 a = 0, b = 0;
    if ((unsigned)y < (unsigned)N && (unsigned)z < (unsigned)N) {
      size_t r = ((size_t)z * N + y) * W;
      if ((unsigned)w0 < (unsigned)W) a = C[r + w0];
      if ((unsigned)(w0 + 1) < (unsigned)W) b = C[r + w0 + 1];
    }
    u64 v = ((u64)b << 32) | a;
    bx.rows[i] = (u32)(v >> o) & ((1u << kBox) - 1u);
  }
  __syncthreads();
}

// Corner k of the cube at f: (k&1 ^ (k>>1)&1, (k>>1)&1, (k>>2)&1) in
// MarchingCubesMesher.h:39-46 order (0,0,0),(1,0,0),(1,1,0),(0,1,0),...
__device__ __forceinline__ int cube_case(const Box& b, int fx, int fy, int fz) {
  int v = 0;
  v |= b.at(fx, fy, fz) << 0;
  v |= b.at(fx + 1, fy, fz) << 1;
  v |= b.at(fx + 1, fy + 1, fz) << 2;
  v |= b.at(fx, fy + 1, fz) << 3;
  v |= b.at(fx, fy, fz + 1) << 4;
  v |= b.at(fx + 1, fy, fz + 1) << 5;
  v |= b.at(fx + 1, fy + 1, fz + 1) << 6;
  v |= b.at(fx, fy + 1, fz + 1) << 7;
  return v;
}
__device__ __forceinline__ int own_verts(int v) {
  int c0 = v & 1;
  return (c0 ^ ((v >> 1) & 1)) + (c0 ^ ((v >> 3) & 1)) + (c0 ^ ((v >> 4) & 1));
}
__device__ __forceinline__ bool mixed(int v) { return v != 0 && v != 255; }
__device__ __forceinline__ bool cell_in_m_range(int cx, int cy, int cz, int R) {
  return cx >= 1 && cy >= 1 && cz >= 1 && cx <= R - 2 && cy <= R - 2 && cz <= R - 2;
}

// Per-cell vertex / triangle counts (cells outside [1,R-2]^3 contribute 0).
__device__ __forceinline__ void cell_counts(const Box& b, int cx, int cy, int cz, int R, u32& nv, u32& nt) {
  nv = 0;
  nt = 0;
  if (!cell_in_m_range(cx, cy, cz, R)) return;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    int v = cube_case(b, 2 * cx + (j & 1), 2 * cy + ((j >> 1) & 1), 2 * cz + ((j >> 2) & 1));
    nv += own_verts(v);
    nt += g_mc.ntri[v];
  }
}

// Block index -> cell coordinates of local Morton index l.
__device__ __forceinline__ void block_cell(int blk, int l, int& cx, int& cy, int& cz) {
  u32 code = ((u32)blk << 9) | (u32)l;
  cx = (int)compact1by2(code);
  cy = (int)compact1by2(code >> 1);
  cz = (int)compact1by2(code >> 2);
}

// Exclusive scan of a (two values per thread) over the workgroup in thread
// order; returns the exclusive prefix of the thread's first value and writes
// the block total to *total.
__device__ u32 block_exscan(u32 v, u32* lds8, u32* total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  u32 x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    u32 y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) lds8[wv] = x;
  __syncthreads();
  u32 off = 0, tot = 0;
  const int nw = blockDim.x >> 6;
  for (int i = 0; i < nw; ++i) {
    u32 s = lds8[i];
    if (i < wv) off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + x - v;
}

__global__ void __launch_bounds__(kMcThreads) k_mc_count(const u32* __restrict__ C, int R, int P, u32* blockV,
                                                          u32* blockT) {
  __shared__ Box bx;
  __shared__ u32 red[8];
  const int blk = blockIdx.x;
  int c0x, c0y, c0z;
  block_cell(blk, 0, c0x, c0y, c0z);
  if (c0x >= R || c0y >= R || c0z >= R) {
    if (threadIdx.x == 0) {
      blockV[blk] = 0;
      blockT[blk] = 0;
    }
    return;
  }
  load_box(bx, C, 2 * R, 2 * c0x - kLo, 2 * c0y - kLo, 2 * c0z - kLo);
  u32 nv = 0, nt = 0;
  for (int h = 0; h < 2; ++h) {
    int cx, cy, cz;
    block_cell(blk, 2 * threadIdx.x + h, cx, cy, cz);
    u32 a, b;
    cell_counts(bx, cx, cy, cz, R, a, b);
    nv += a;
    nt += b;
  }
  u32 tv, tt;
  block_exscan(nv, red, &tv);
  block_exscan(nt, red, &tt);
  if (threadIdx.x == 0) {
    blockV[blk] = tv;
    blockT[blk] = tt;
  }
}

// Exclusive scan of the block totals (single workgroup of 1024 threads).
__global__ void __launch_bounds__(1024) k_mc_scan(u32* blockV, u32* blockT, int nblocks, u32* totals, u32 capV,
                                                   u32 capT) {
  __shared__ u32 red[16];
  const int t = threadIdx.x;
  int per = (nblocks + 1023) / 1024;
  int b0 = t * per, b1 = min(b0 + per, nblocks);
  u32 sv = 0, st = 0;
  for (int b = b0; b < b1; ++b) {
    sv += blockV[b];
    st += blockT[b];
  }
  u32 tv, tt;
  u32 ov = block_exscan(sv, red, &tv);
  u32 ot = block_exscan(st, red, &tt);
  for (int b = b0; b < b1; ++b) {
    u32 v = blockV[b], w = blockT[b];
    blockV[b] = ov;
    blockT[b] = ot;
    ov += v;
    ot += w;
  }
  if (t == 0) {
    totals[0] = tv;
    totals[1] = tt;
    totals[2] = (tv > capV || tt > capT) ? 1u : 0u;
    totals[3] = 0u;
  }
}

// Position of the vertex on owned edge e (0: x, 3: y, 8: z) of cube f:
// findIntersection (MarchingCubesMesher.cu:98-100) with {0,1} corner values.
__device__ __forceinline__ float lerp_edge(float p1, float p2, float v1, float v2) {
  float d = __fsub_rn(p2, p1);
  float t = __fsub_rn(0.95f, v1);
  float m = __fmul_rn(t, d);
  float q = __fdiv_rn(m, __fsub_rn(v2, v1));
  return __fadd_rn(p1, q);
}
struct Geo {
  float bx, by, bz, h;
};
__device__ __forceinline__ void edge_pos(const Geo& g, const Box& b, int fx, int fy, int fz, int e, float* o) {
  float px = fmaf((float)fx, g.h, g.bx);
  float py = fmaf((float)fy, g.h, g.by);
  float pz = fmaf((float)fz, g.h, g.bz);
  float v0 = (float)b.at(fx, fy, fz);
  o[0] = px;
  o[1] = py;
  o[2] = pz;
  if (e == 0) o[0] = lerp_edge(px, __fadd_rn(px, g.h), v0, (float)b.at(fx + 1, fy, fz));
  else if (e == 3) o[1] = lerp_edge(__fadd_rn(py, g.h), py, (float)b.at(fx, fy + 1, fz), v0);
  else o[2] = lerp_edge(pz, __fadd_rn(pz, g.h), v0, (float)b.at(fx, fy, fz + 1));
}

// Owner of local edge e of a cube: offset of the owning cube and its edge.
__constant__ int8_t g_own[12][4] = {{0, 0, 0, 0}, {1, 0, 0, 3}, {0, 1, 0, 0}, {0, 0, 0, 3},
                                    {0, 0, 1, 0}, {1, 0, 1, 3}, {0, 1, 1, 0}, {0, 0, 1, 3},
                                    {0, 0, 0, 8}, {1, 0, 0, 8}, {1, 1, 0, 8}, {0, 1, 0, 8}};

__device__ __forceinline__ int rank_of(int v, int e) {  // index of owned edge e among the cube's own vertices
  int c0 = v & 1;
  int e0 = c0 ^ ((v >> 1) & 1), e3 = c0 ^ ((v >> 3) & 1);
  return e == 0 ? 0 : (e == 3 ? e0 : e0 + e3);
}
__device__ __forceinline__ int cell_cube_prefix(const Box& b, int cx, int cy, int cz, int jlim) {
  int s = 0;
  for (int j = 0; j < jlim; ++j)
    s += own_verts(cube_case(b, 2 * cx + (j & 1), 2 * cy + ((j >> 1) & 1), 2 * cz + ((j >> 2) & 1)));
  return s;
}

__global__ void __launch_bounds__(kMcThreads) k_mc_verts(const u32* __restrict__ C, int R, Geo geo,
                                                          const u32* __restrict__ blockV, u32* cellV, float* V,
                                                          u32 capV, u32* totals) {
  __shared__ Box bx;
  __shared__ u32 red[8];
  const int blk = blockIdx.x;
  int c0x, c0y, c0z;
  block_cell(blk, 0, c0x, c0y, c0z);
  if (c0x >= R || c0y >= R || c0z >= R) return;
  u32 base = blockV[blk];
  u32 next = (blk + 1 < (int)gridDim.x) ? blockV[blk + 1] : totals[0];
  if (next == base) return;  // no vertex in this block (uniform across the workgroup)
  load_box(bx, C, 2 * R, 2 * c0x - kLo, 2 * c0y - kLo, 2 * c0z - kLo);
  int cxs[2], cys[2], czs[2];
  u32 nv[2];
  for (int h = 0; h < 2; ++h) {
    block_cell(blk, 2 * threadIdx.x + h, cxs[h], cys[h], czs[h]);
    u32 t;
    cell_counts(bx, cxs[h], cys[h], czs[h], R, nv[h], t);
  }
  u32 tot;
  u32 off = block_exscan(nv[0] + nv[1], red, &tot) + base;
  for (int h = 0; h < 2; ++h) {
    if (nv[h] == 0) continue;
    int cx = cxs[h], cy = cys[h], cz = czs[h];
    cellV[(size_t)cx + (size_t)R * ((size_t)cy + (size_t)R * cz)] = off;
    u32 vid = off;
    for (int j = 0; j < 8; ++j) {
      int fx = 2 * cx + (j & 1), fy = 2 * cy + ((j >> 1) & 1), fz = 2 * cz + ((j >> 2) & 1);
      int v = cube_case(bx, fx, fy, fz);
      int c0 = v & 1;
      const int es[3] = {0, 3, 8};
      const int on[3] = {c0 ^ ((v >> 1) & 1), c0 ^ ((v >> 3) & 1), c0 ^ ((v >> 4) & 1)};
      for (int q = 0; q < 3; ++q) {
        if (!on[q]) continue;
        if (vid < capV) {
          float p[3];
          edge_pos(geo, bx, fx, fy, fz, es[q], p);
          V[3 * (size_t)vid] = p[0];
          V[3 * (size_t)vid + 1] = p[1];
          V[3 * (size_t)vid + 2] = p[2];
        }
        ++vid;
      }
    }
    off += nv[h];
  }
}

// Vertex id of local edge e of the cube at fine (fx,fy,fz): 0 when the owning
// cube's cell is outside [1,R-2]^3 (the reference's absent neighbour).
__device__ __forceinline__ u32 edge_vid(const Box& b, const u32* __restrict__ cellV, int R, int fx, int fy, int fz,
                                        int e, u32* border) {
  int gx = fx + g_own[e][0], gy = fy + g_own[e][1], gz = fz + g_own[e][2], oe = g_own[e][3];
  int cx = gx >> 1, cy = gy >> 1, cz = gz >> 1;
  if (!cell_in_m_range(cx, cy, cz, R)) {
    *border = 1u;
    return 0u;
  }
  int j = (gx & 1) | ((gy & 1) << 1) | ((gz & 1) << 2);
  u32 base = cellV[(size_t)cx + (size_t)R * ((size_t)cy + (size_t)R * cz)];
  return base + (u32)cell_cube_prefix(b, cx, cy, cz, j) + (u32)rank_of(cube_case(b, gx, gy, gz), oe);
}

__global__ void __launch_bounds__(kMcThreads) k_mc_tris(const u32* __restrict__ C, int R,
                                                         const u32* __restrict__ blockT,
                                                         const u32* __restrict__ cellV, u32* T, u32 capT,
                                                         u32* late, u32 capLate, u32* totals) {
  __shared__ Box bx;
  __shared__ u32 red[8];
  const int blk = blockIdx.x;
  int c0x, c0y, c0z;
  block_cell(blk, 0, c0x, c0y, c0z);
  if (c0x >= R || c0y >= R || c0z >= R) return;
  u32 base = blockT[blk];
  u32 next = (blk + 1 < (int)gridDim.x) ? blockT[blk + 1] : totals[1];
  if (next == base) return;
  load_box(bx, C, 2 * R, 2 * c0x - kLo, 2 * c0y - kLo, 2 * c0z - kLo);
  int cxs[2], cys[2], czs[2];
  u32 nt[2];
  for (int h = 0; h < 2; ++h) {
    block_cell(blk, 2 * threadIdx.x + h, cxs[h], cys[h], czs[h]);
    u32 t;
    cell_counts(bx, cxs[h], cys[h], czs[h], R, t, nt[h]);
  }
  u32 tot;
  u32 off = block_exscan(nt[0] + nt[1], red, &tot) + base;
  u32 border = 0u;
  for (int h = 0; h < 2; ++h) {
    if (nt[h] == 0) continue;
    int cx = cxs[h], cy = cys[h], cz = czs[h];
    u32 slot = off;
    for (int j = 0; j < 8; ++j) {
      int fx = 2 * cx + (j & 1), fy = 2 * cy + ((j >> 1) & 1), fz = 2 * cz + ((j >> 2) & 1);
      int v = cube_case(bx, fx, fy, fz);
      if (!mixed(v)) continue;
      unsigned long long tri = g_tri[v];
      u32 cnt = g_mc.ntri[v], len = g_mc.len[v];
      for (u32 q = 0; q < len; ++q) {
        int e0 = (int)((tri >> (12 * q)) & 0xf), e1 = (int)((tri >> (12 * q + 4)) & 0xf),
            e2 = (int)((tri >> (12 * q + 8)) & 0xf);
        u32 a = edge_vid(bx, cellV, R, fx, fy, fz, e2, &border);
        u32 b = edge_vid(bx, cellV, R, fx, fy, fz, e1, &border);
        u32 c = edge_vid(bx, cellV, R, fx, fy, fz, e0, &border);
        u32 s = slot + q;
        if (q < cnt) {
          if (s < capT) {
            T[3 * (size_t)s] = a;
            T[3 * (size_t)s + 1] = b;
            T[3 * (size_t)s + 2] = c;
          }
        } else {
          // listed but not counted: overwrites slot s after all counted writes
          u32 k = atomicAdd(&late[0], 1u);
          if (k < capLate) {
            late[1 + 4 * k] = s;
            late[2 + 4 * k] = a;
            late[3 + 4 * k] = b;
            late[4 + 4 * k] = c;
          } else {
            atomicOr(&totals[2], 2u);
          }
        }
      }
      slot += cnt;
    }
    off += nt[h];
  }
  if (border) atomicOr(&totals[3], 1u);
}

// Area-weighted normal sums (GMorpho.cu:6498-6524) gathered per vertex: the
// triangles that use the vertex's edge are those of the <= 4 cubes around it.
__constant__ int8_t g_share[3][4][4] = {
    // owned edge 0 (x): cubes f, f-y, f-z, f-y-z with local edges 0, 2, 4, 6
    {{0, 0, 0, 0}, {0, -1, 0, 2}, {0, 0, -1, 4}, {0, -1, -1, 6}},
    // owned edge 3 (y): cubes f, f-x, f-z, f-x-z with local edges 3, 1, 7, 5
    {{0, 0, 0, 3}, {-1, 0, 0, 1}, {0, 0, -1, 7}, {-1, 0, -1, 5}},
    // owned edge 8 (z): cubes f, f-x, f-y, f-x-y with local edges 8, 9, 11, 10
    {{0, 0, 0, 8}, {-1, 0, 0, 9}, {0, -1, 0, 11}, {-1, -1, 0, 10}}};

__device__ __forceinline__ void local_edge_pos(const Geo& g, const Box& b, int fx, int fy, int fz, int e, float* p) {
  edge_pos(g, b, fx + g_own[e][0], fy + g_own[e][1], fz + g_own[e][2], g_own[e][3], p);
}

__global__ void __launch_bounds__(kMcThreads) k_mc_normals(const u32* __restrict__ C, int R, Geo geo,
                                                            const u32* __restrict__ blockV,
                                                            const u32* __restrict__ cellV, float* Nrm, u32 capV,
                                                            u32* totals) {
  __shared__ Box bx;
  const int blk = blockIdx.x;
  int c0x, c0y, c0z;
  block_cell(blk, 0, c0x, c0y, c0z);
  if (c0x >= R || c0y >= R || c0z >= R) return;
  u32 base = blockV[blk];
  u32 next = (blk + 1 < (int)gridDim.x) ? blockV[blk + 1] : totals[0];
  if (next == base) return;
  load_box(bx, C, 2 * R, 2 * c0x - kLo, 2 * c0y - kLo, 2 * c0z - kLo);
  for (int h = 0; h < 2; ++h) {
    int cx, cy, cz;
    block_cell(blk, 2 * threadIdx.x + h, cx, cy, cz);
    if (!cell_in_m_range(cx, cy, cz, R)) continue;
    u32 nvc, ntc;
    cell_counts(bx, cx, cy, cz, R, nvc, ntc);
    if (nvc == 0) continue;
    u32 vid = cellV[(size_t)cx + (size_t)R * ((size_t)cy + (size_t)R * cz)];
    for (int j = 0; j < 8; ++j) {
      int fx = 2 * cx + (j & 1), fy = 2 * cy + ((j >> 1) & 1), fz = 2 * cz + ((j >> 2) & 1);
      int v = cube_case(bx, fx, fy, fz);
      int c0 = v & 1;
      const int on[3] = {c0 ^ ((v >> 1) & 1), c0 ^ ((v >> 3) & 1), c0 ^ ((v >> 4) & 1)};
      for (int q = 0; q < 3; ++q) {
        if (!on[q]) continue;
        float n[3] = {0.f, 0.f, 0.f};
        for (int s = 0; s < 4; ++s) {
          int gx = fx + g_share[q][s][0], gy = fy + g_share[q][s][1], gz = fz + g_share[q][s][2];
          int le = g_share[q][s][3];
          if (!cell_in_m_range(gx >> 1, gy >> 1, gz >> 1, R)) continue;
          int gv = cube_case(bx, gx, gy, gz);
          if (!mixed(gv)) continue;
          unsigned long long tri = g_tri[gv];
          u32 cnt = g_mc.ntri[gv];
          for (u32 t = 0; t < cnt; ++t) {
            int e0 = (int)((tri >> (12 * t)) & 0xf), e1 = (int)((tri >> (12 * t + 4)) & 0xf),
                e2 = (int)((tri >> (12 * t + 8)) & 0xf);
            if (e0 != le && e1 != le && e2 != le) continue;
            float p0[3], p1[3], p2[3];
            local_edge_pos(geo, bx, gx, gy, gz, e2, p0);
            local_edge_pos(geo, bx, gx, gy, gz, e1, p1);
            local_edge_pos(geo, bx, gx, gy, gz, e0, p2);
            float a0 = __fsub_rn(p1[0], p0[0]), a1 = __fsub_rn(p1[1], p0[1]), a2 = __fsub_rn(p1[2], p0[2]);
            float b0 = __fsub_rn(p2[0], p0[0]), b1 = __fsub_rn(p2[1], p0[1]), b2 = __fsub_rn(p2[2], p0[2]);
            n[0] = __fadd_rn(n[0], __fsub_rn(__fmul_rn(a1, b2), __fmul_rn(a2, b1)));
            n[1] = __fadd_rn(n[1], __fsub_rn(__fmul_rn(a2, b0), __fmul_rn(a0, b2)));
            n[2] = __fadd_rn(n[2], __fsub_rn(__fmul_rn(a0, b1), __fmul_rn(a1, b0)));
          }
        }
        if (vid < capV) {
          Nrm[3 * (size_t)vid] = n[0];
          Nrm[3 * (size_t)vid + 1] = n[1];
          Nrm[3 * (size_t)vid + 2] = n[2];
        }
        ++vid;
      }
    }
  }
}

__device__ __forceinline__ void face(const float* V, const u32* t, float* n) {
  const float* p0 = V + 3 * (size_t)t[0];
  const float* p1 = V + 3 * (size_t)t[1];
  const float* p2 = V + 3 * (size_t)t[2];
  float a0 = __fsub_rn(p1[0], p0[0]), a1 = __fsub_rn(p1[1], p0[1]), a2 = __fsub_rn(p1[2], p0[2]);
  float b0 = __fsub_rn(p2[0], p0[0]), b1 = __fsub_rn(p2[1], p0[1]), b2 = __fsub_rn(p2[2], p0[2]);
  n[0] = __fsub_rn(__fmul_rn(a1, b2), __fmul_rn(a2, b1));
  n[1] = __fsub_rn(__fmul_rn(a2, b0), __fmul_rn(a0, b2));
  n[2] = __fsub_rn(__fmul_rn(a0, b1), __fmul_rn(a1, b0));
}

// Applies the listed-but-uncounted triangles in slot order (single thread):
// each replaces the triangle in its slot, and the normal sums follow.
__global__ void k_mc_fixup(const float* V, float* Nrm, u32* T, u32* late, u32 capLate, u32* totals, u32 capV,
                           u32 capT) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  u32 n = late[0];
  if (n > capLate) n = capLate;
  if (n == 0) return;
  // insertion sort by slot (the list is tiny)
  for (u32 i = 1; i < n; ++i) {
    u32 e[4] = {late[1 + 4 * i], late[2 + 4 * i], late[3 + 4 * i], late[4 + 4 * i]};
    u32 j = i;
    while (j > 0 && late[1 + 4 * (j - 1)] > e[0]) {
      for (int q = 0; q < 4; ++q) late[1 + 4 * j + q] = late[1 + 4 * (j - 1) + q];
      --j;
    }
    for (int q = 0; q < 4; ++q) late[1 + 4 * j + q] = e[q];
  }
  const u32 nT = totals[1], nV = totals[0];
  if (nT > capT || nV > capV) return;
  for (u32 i = 0; i < n; ++i) {
    u32 s = late[1 + 4 * i];
    if (s >= nT) continue;
    u32* t = T + 3 * (size_t)s;
    float f[3];
    face(V, t, f);
    for (int k = 0; k < 3; ++k)
      for (int q = 0; q < 3; ++q) Nrm[3 * (size_t)t[k] + q] = __fsub_rn(Nrm[3 * (size_t)t[k] + q], f[q]);
    t[0] = late[2 + 4 * i];
    t[1] = late[3 + 4 * i];
    t[2] = late[4 + 4 * i];
    face(V, t, f);
    for (int k = 0; k < 3; ++k)
      for (int q = 0; q < 3; ++q) Nrm[3 * (size_t)t[k] + q] = __fadd_rn(Nrm[3 * (size_t)t[k] + q], f[q]);
  }
}

// NormalizeMeshNormals (GMorpho.cu:6526-6537).
__global__ void k_mc_normalize(float* Nrm, const u32* totals, u32 capV) {
  u32 nV = totals[0];
  if (nV > capV) nV = capV;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nV) return;
  float x = Nrm[3 * i], y = Nrm[3 * i + 1], z = Nrm[3 * i + 2];
  float l = sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(x, x), __fmul_rn(y, y)), __fmul_rn(z, z)));
  if (l > 0.f) {
    Nrm[3 * i] = __fdiv_rn(x, l);
    Nrm[3 * i + 1] = __fdiv_rn(y, l);
    Nrm[3 * i + 2] = __fdiv_rn(z, l);
  }
}

bool g_tables_uploaded = false;

hipError_t upload_tables() {
  if (g_tables_uploaded) return hipSuccess;
  McConst h{};
  for (int c = 0; c < 256; ++c) {
    int n = 0;
    while (n < 16 && ((kTriPacked[c] >> (4 * n)) & 0xf) != 0xf) ++n;
    h.len[c] = (u32)(n / 3);
    h.ntri[c] = h.len[c];
  }
  // The reference's per-case triangle count (tables.h:61-318) lists 4 for
  // case 91, whose triangulation has 5 triangles.
  h.ntri[91] = 4;
  GM_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_mc), &h, sizeof(h)));
  GM_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_tri), kTriPacked, sizeof(kTriPacked)));
  g_tables_uploaded = true;
  return hipSuccess;
}

}  // namespace

hipError_t mc_tables_host(u32* ntri, u32* len, unsigned long long* tri) {
  for (int c = 0; c < 256; ++c) {
    int n = 0;
    while (n < 16 && ((kTriPacked[c] >> (4 * n)) & 0xf) != 0xf) ++n;
    len[c] = (u32)(n / 3);
    ntri[c] = len[c];
    tri[c] = kTriPacked[c];
  }
  ntri[91] = 4;
  return hipSuccess;
}

hipError_t launch_mc(const u32* C, int R, const float* bmin3, float cell, const MCBuffers& m, hipStream_t s,
                     hipEvent_t* ev) {
  GM_CHECK(upload_tables());
  if (ev) GM_CHECK(hipEventRecord(ev[0], s));
  hipLaunchKernelGGL(k_mc_count, dim3(m.nblocks), dim3(kMcThreads), 0, s, C, R, m.P, m.blockV, m.blockT);
  GM_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_mc_scan, dim3(1), dim3(1024), 0, s, m.blockV, m.blockT, m.nblocks, m.totals, m.capV, m.capT);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[1], s));
  return launch_mc_emit(C, R, bmin3, cell, m, s, ev ? ev + 1 : nullptr);
}

hipError_t launch_mc_emit(const u32* C, int R, const float* bmin3, float cell, const MCBuffers& m, hipStream_t s,
                          hipEvent_t* ev) {
  GM_CHECK(upload_tables());
  Geo geo{bmin3[0], bmin3[1], bmin3[2], 0.5f * cell};
  GM_CHECK(hipMemsetAsync(m.late, 0, sizeof(u32), s));
  hipLaunchKernelGGL(k_mc_verts, dim3(m.nblocks), dim3(kMcThreads), 0, s, C, R, geo, m.blockV, m.cellV, m.V, m.capV,
                     m.totals);
  GM_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_mc_tris, dim3(m.nblocks), dim3(kMcThreads), 0, s, C, R, m.blockT, m.cellV, m.T, m.capT,
                     m.late, m.capLate, m.totals);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[1], s));
  hipLaunchKernelGGL(k_mc_normals, dim3(m.nblocks), dim3(kMcThreads), 0, s, C, R, geo, m.blockV, m.cellV, m.Nrm,
                     m.capV, m.totals);
  GM_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_mc_fixup, dim3(1), dim3(64), 0, s, m.V, m.Nrm, m.T, m.late, m.capLate, m.totals, m.capV,
                     m.capT);
  GM_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_mc_normalize, dim3((m.capV + 255) / 256), dim3(256), 0, s, m.Nrm, m.totals, m.capV);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[2], s));
  return hipSuccess;
}

}  // namespace gm
