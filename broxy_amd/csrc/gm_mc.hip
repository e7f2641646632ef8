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
//   mc_blocks  list of the blocks that can hold a mixed cube
//   mc_count   the block's voxels in LDS: the 8 cube cases of every cell
//              (stored for mc_emit), per block vertex/triangle totals, and
//              the block-local vertex base of every cell with vertices
//   scan       exclusive scan of the block totals (chunk totals summed by
//              mc_count, then one down-sweep kernel, which also sorts the
//              blocks into work classes, heaviest first)
//   mc_emit    the stored cases of the block and of the cells around it in
//              LDS, a block scan over cells, then one thread per triangle
//              (vertex ids of other blocks' edges from their block base +
//              cell base; face normals kept in LDS), then one thread per
//              vertex (position, area-weighted normal sum gathered from the
//              <= 4 cubes around its edge)
//   mc_fixup   triangles the count table does not cover, mc_normalize.
// The cube cases and block-local vertex bases are stored per listed block
// (512 cells each, by list position), replacing the reference's 4*N^3-byte
// memset + splat; their capacity grows when more blocks are listed.
#include <algorithm>
#include <mutex>

#include "gm_internal.h"

namespace gm {

namespace {

#include "mc_tables.inc"

struct McConst {
  u32 ntri[256];          // triangles counted per case (numTriTable semantics)
  u32 len[256];           // triangles listed per case
  uint8_t use[256][12];   // bit t: counted triangle t of the case uses edge e
};
__constant__ McConst g_mc;
__constant__ unsigned long long g_tri[256];

constexpr int kMcThreads = 256;

// Per-workgroup LDS copy of the case tables: the lookups are per lane, so
// they would otherwise be divergent constant-memory loads.
struct McTab {
  unsigned long long tri[256];
  uint8_t ntri[256];
  uint8_t len[256];
  uint8_t use[256][12];
};
__device__ __forceinline__ void load_mc_tab(McTab& t) {
  for (int i = threadIdx.x; i < 256; i += blockDim.x) {
    t.ntri[i] = (uint8_t)g_mc.ntri[i];
    t.len[i] = (uint8_t)g_mc.len[i];
    t.tri[i] = g_tri[i];
#pragma unroll
    for (int e = 0; e < 12; ++e) t.use[i][e] = g_mc.use[i][e];
  }
}
// Voxel box of one Morton block in LDS (count kernel): the 17 x 17 rows of
// fine voxels [16b, 16b + 17) per axis, 24 bits of x each.
constexpr int kBoxRows = 17;
constexpr int kBoxBits = 24;

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
  u32 rows[kBoxRows * kBoxRows];  // bit i = fine x0+i of row (y0+j, z0+k)
  int x0, y0, z0;
  // per wave: bit 0 = some row of its 17-bit cube span has a set voxel,
  // bit 1 = some row has a clear one
  u32 uni[16];
  // three consecutive x bits starting at x
  __device__ __forceinline__ u32 at3(int x, int y, int z) const {
    return (rows[(y - y0) + kBoxRows * (z - z0)] >> (x - x0)) & 7u;
  }
};

// Loads rows [fy0, fy0+ny) x [fz0, fz0+nz) of the box (x range kBoxBits bits
// from fx0; ny, nz <= kBoxRows).
__device__ void load_box(Box& bx, const u32* __restrict__ C, int N, int fx0, int fy0, int fz0, int ny, int nz) {
  const int W = N / 32;
  if (threadIdx.x == 0) {
    bx.x0 = fx0;
    bx.y0 = fy0;
    bx.z0 = fz0;
  }
  int w0 = (fx0 >= 0) ? (fx0 >> 5) : -((31 - fx0) >> 5);
  int o = fx0 - 32 * w0;
  auto row = [&](int i, u32& a, u32& b) {
    const int y = fy0 + i % ny, z = fz0 + i / ny;
    a = b = 0u;
    if ((unsigned)y < (unsigned)N && (unsigned)z < (unsigned)N) {
      const size_t r = ((size_t)z * N + y) * W;
      if ((unsigned)w0 < (unsigned)W) a = C[r + w0];
      if ((unsigned)(w0 + 1) < (unsigned)W) b = C[r + w0 + 1];
    }
  };
  u32 all = 0x1ffffu, any = 0u;  // over the 17 bits the cubes of a block read
  auto put = [&](int i, u32 a, u32 b) {
    const u64 v = ((u64)b << 32) | a;
    const u32 row = (u32)(v >> o) & ((1u << kBoxBits) - 1u);
    bx.rows[i % ny + kBoxRows * (i / ny)] = row;
    all &= row;
    any |= row & 0x1ffffu;
  };
  // two rows per thread in flight at once (the 17x17 rows of the count
  // kernel's box take one round on 256 threads instead of two)
  for (int i = threadIdx.x; i < ny * nz; i += 2 * blockDim.x) {
    const int i2 = i + (int)blockDim.x;
    u32 a, b, a2 = 0u, b2 = 0u;
    row(i, a, b);
    if (i2 < ny * nz) row(i2, a2, b2);
    put(i, a, b);
    if (i2 < ny * nz) put(i2, a2, b2);
  }
  {
    const bool one = __ballot(any != 0u) != 0ull, zero = __ballot((all & 0x1ffffu) != 0x1ffffu) != 0ull;
    if ((threadIdx.x & 63u) == 0u) bx.uni[threadIdx.x >> 6] = (one ? 1u : 0u) | (zero ? 2u : 0u);
  }
  __syncthreads();
}

// A cube's case: bit k = voxel at corner k, corners in MarchingCubesMesher.h:39-46
// order (0,0,0),(1,0,0),(1,1,0),(0,1,0),(0,0,1),(1,0,1),(1,1,1),(0,1,1). It
// owns the vertices on its edges 0, 3, 8 (x, y, z from corner 0) that are cut.
__device__ __forceinline__ bool mixed(int v) { return v != 0 && v != 255; }
__device__ __forceinline__ bool cell_in_m_range(int cx, int cy, int cz, int R) {
  return cx >= 1 && cy >= 1 && cz >= 1 && cx <= R - 2 && cy <= R - 2 && cz <= R - 2;
}

// swap of the two low bits of p (p < 4)
__device__ __forceinline__ u32 swap2(u32 p) { return ((p * 5u) >> 1) & 3u; }

// The 8 cube cases of cell (cx,cy,cz) as the bytes of one word (byte j =
// cube (j&1, j>>1&1, j>>2) of the cell, corner order above), 0 when the
// cell's 3x3x3 fine voxels are uniform: the 9 rows of 3 bits are read once
// and each case takes two bits from four of them.
__device__ __forceinline__ unsigned long long cell_cases_box(const Box& b, int cx, int cy, int cz) {
  u32 r[3][3];
  u32 a = 7u, o = 0u;
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      r[j][k] = b.at3(2 * cx, 2 * cy + j, 2 * cz + k);
      a &= r[j][k];
      o |= r[j][k];
    }
  if (a == 7u || o == 0u) return 0ull;
  unsigned long long w = 0ull;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int x = j & 1, y = (j >> 1) & 1, z = j >> 2;
    const u32 c = ((r[y][z] >> x) & 3u) | (swap2((r[y + 1][z] >> x) & 3u) << 2) | (((r[y][z + 1] >> x) & 3u) << 4) |
                  (swap2((r[y + 1][z + 1] >> x) & 3u) << 6);
    w |= (unsigned long long)c << (8 * j);
  }
  return w;
}

// Local Morton index l in [0,512) <-> local cell coordinates in [0,8)^3.
__device__ __forceinline__ void local_cell(int l, int& lx, int& ly, int& lz) {
  lx = (l & 1) | ((l >> 2) & 2) | ((l >> 4) & 4);
  ly = ((l >> 1) & 1) | ((l >> 3) & 2) | ((l >> 5) & 4);
  lz = ((l >> 2) & 1) | ((l >> 4) & 2) | ((l >> 6) & 4);
}
__device__ __forceinline__ int spread3(int x) { return (x & 1) | ((x & 2) << 2) | ((x & 4) << 4); }
__device__ __forceinline__ int local_morton(int lx, int ly, int lz) {
  return spread3(lx) | (spread3(ly) << 1) | (spread3(lz) << 2);
}

// Block index -> cell coordinates of local Morton index l.
__device__ __forceinline__ void block_cell(int blk, int l, int& cx, int& cy, int& cz) {
  u32 code = ((u32)blk << 9) | (u32)l;
  cx = (int)compact1by2(code);
  cy = (int)compact1by2(code >> 1);
  cz = (int)compact1by2(code >> 2);
}

// Exclusive scan over the workgroup in thread order of one value per thread;
// writes the workgroup total to *total. trailing = false leaves out the
// closing barrier, for callers that pass another barrier before lds16 is
// written again.
__device__ u32 block_exscan(u32 v, u32* lds16, u32* total, bool trailing = true) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  u32 x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    u32 y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) lds16[wv] = x;
  __syncthreads();
  u32 off = 0, tot = 0;
  const int nw = blockDim.x >> 6;
  for (int i = 0; i < nw; ++i) {
    u32 s = lds16[i];
    if (i < wv) off += s;
    tot += s;
  }
  if (trailing) __syncthreads();
  *total = tot;
  return off + x - v;
}

constexpr int kListCap = 1024;  // work-list window (vertices or triangles of a block)
constexpr int kFaceCap = 1024;  // face normals kept in LDS (blocks with at most this many triangles)
static_assert(kFaceCap <= kListCap, "the triangles whose face normals are kept fit one window");

constexpr int kNbCells = 9 * 9 * 9 - 8 * 8 * 8;
static_assert(kNbCells <= kMcThreads, "one thread per cell just above a block");

// Index of a cell just above a block from its local coordinates in [0,8]^3
// with at least one equal to 8: the x = 8 face (81), then y = 8 with x < 8
// (72), then z = 8 with x, y < 8 (64).
__device__ __forceinline__ int nb_index(int lx, int ly, int lz) {
  return lx == 8 ? ly + 9 * lz : (ly == 8 ? 81 + lx + 8 * lz : 153 + lx + 8 * ly);
}
__device__ __forceinline__ void nb_cell(int k, int& lx, int& ly, int& lz) {
  if (k < 81) {
    lx = 8;
    ly = k % 9;
    lz = k / 9;
  } else if (k < 153) {
    lx = (k - 81) & 7;
    ly = 8;
    lz = (k - 81) >> 3;
  } else {
    lx = (k - 153) & 7;
    ly = (k - 153) >> 3;
    lz = 8;
  }
}

// Axes (bit 0/1/2 = x/y/z) in which the cell at local coordinates 7 - l lies
// below the block, for l = nb_cell of a cell just above it.
__device__ __forceinline__ int shell_axes(int lx, int ly, int lz) {
  return (lx == 8 ? 1 : 0) | (ly == 8 ? 2 : 0) | (lz == 8 ? 4 : 0);
}

// Vertices owned by the cubes before cube j of a cell whose 8 cube cases are
// the bytes of w (cut edges among 0, 3, 8 per byte, summed over bytes < j).
__device__ __forceinline__ u32 own_prefix8(unsigned long long w, int j) {
  constexpr unsigned long long m1 = 0x0101010101010101ull;
  const unsigned long long c0 = w & m1;
  const unsigned long long x = (c0 ^ ((w >> 1) & m1)) + (c0 ^ ((w >> 3) & m1)) + (c0 ^ ((w >> 4) & m1));
  const unsigned long long keep = j ? (~0ull >> (64 - 8 * j)) : 0ull;
  return (u32)(((x & keep) * m1) >> 56);
}

struct BlockLds {
  McTab tab;
  alignas(8) uint8_t cases[4096];  // [local cell * 8 + cube], 0 for cells outside M range
  alignas(8) uint8_t tpre[4096];   // counted triangles of the cell's earlier cubes
  // the 217 cells just above the block (one local coordinate 8, index
  // nb_index) that own vertices of the block's triangles: their 8 cube cases
  // and global vertex base; and the cells just below it (local coordinate
  // -1, index nb_index of 7 - l) whose cubes share vertices of the block.
  // Cases are 0 for cells outside [1,R-2]^3.
  unsigned long long nbAbove[kNbCells];
  unsigned long long nbBelow[kNbCells];
  u32 nbBase[kNbCells];
  // Shell: the cubes of the cells below the block that share vertices with
  // it (fine coordinate one below the block in every axis where the cell is
  // below it). Their counted triangles are numbered after the block's own
  // (shell base of the cell, then byte j = in-cell prefix of cube j), so
  // their face normals can follow the block's in fn.
  u32 nbTB[kNbCells];
  unsigned long long nbTP[kNbCells];
  u32 nshell;
  u32 vtOff[512];  // per cell with a mixed cube: block-local vertex base (low 16 bits), triangle base (high 16)
  u32 red[16];
  uint16_t cl[512];  // the cells with a mixed cube (local Morton index), S.ncl of them
  u32 ncl;
  uint16_t list[kListCap];  // cube index << 3 | item, by local output index (shell cubes: 4096 + 8 k + j)
  float fn[kFaceCap][3];     // face normal of each local triangle (when it fits)
};

// Cube cases of a listed neighbour cell (0 outside [1,R-2]^3 or in a block
// without mixed cubes); *blkOut = its block.
__device__ __forceinline__ unsigned long long listed_cell_cases(const unsigned long long* __restrict__ cellCase,
                                                                const u32* __restrict__ posOf, int R, int cx, int cy,
                                                                int cz, u32* code, u32* pos) {
  if (!cell_in_m_range(cx, cy, cz, R)) return 0ull;
  *code = part1by2((u32)cx) | (part1by2((u32)cy) << 1) | (part1by2((u32)cz) << 2);
  *pos = posOf[*code >> 9];
  return *pos == ~0u ? 0ull : cellCase[(size_t)*pos * 512 + (*code & 511u)];
}

// Stages listed block `it` (origin cell c0): the cube cases of its 512 cells
// (written by the count kernel), the block-local vertex and triangle bases
// of the cells with a mixed cube (also from the count kernel) and the list
// of those cells (S.cl, S.ncl; the later phases walk it instead of all 512
// cells), and the cells just above and below the block. Two cells per
// thread (kMcThreads = 256).
static_assert(kMcThreads == 256, "stage_block takes cells t and t + 256");
__device__ void stage_block(BlockLds& S, const unsigned long long* __restrict__ cellCase,
                            const u32* __restrict__ posOf, const u32* __restrict__ blockV,
                            const u32* __restrict__ cellL, int R, u32 it, int c0x, int c0y, int c0z) {
  const int t = (int)threadIdx.x;
  unsigned long long above = 0ull, below = 0ull, shellPre = 0ull;
  u32 nbB = 0u, shellCnt = 0u;
  if (t < kNbCells) {
    int lx, ly, lz;
    nb_cell(t, lx, ly, lz);
    u32 code = 0u, pos = 0u;
    above = listed_cell_cases(cellCase, posOf, R, c0x + lx, c0y + ly, c0z + lz, &code, &pos);
    if (above) nbB = blockV[code >> 9] + (cellL[(size_t)pos * 512 + (code & 511u)] & 0xffffu);
    below = listed_cell_cases(cellCase, posOf, R, c0x + 7 - lx, c0y + 7 - ly, c0z + 7 - lz, &code, &pos);
    // the below cell's shell cubes: sub-position bit set in each axis where
    // the cell is below the block
    const int need = shell_axes(lx, ly, lz);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      shellPre |= (unsigned long long)shellCnt << (8 * j);
      if ((j & need) == need) shellCnt += S.tab.ntri[(below >> (8 * j)) & 0xffull];
    }
  }
  const unsigned long long* own = cellCase + (size_t)it * 512;
  const u32* ownL = cellL + (size_t)it * 512;
  unsigned long long w[2];
  u32 base[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    w[h] = own[t + 256 * h];
    base[h] = ownL[t + 256 * h];  // written only for cells with a mixed cube
  }
  const int lane = t & 63, wv = t >> 6;
  unsigned long long m[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int l = t + 256 * h;
    *reinterpret_cast<unsigned long long*>(&S.cases[8 * l]) = w[h];
    if (w[h]) S.vtOff[l] = base[h];
    m[h] = __ballot(w[h] != 0ull);
    if (lane == 0) S.red[4 * h + wv] = (u32)__popcll(m[h]);
  }
  // shell triangles: inclusive scan over the wave, wave totals in red[8..11]
  u32 shellInc = shellCnt;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const u32 y = __shfl_up(shellInc, d, 64);
    if (lane >= d) shellInc += y;
  }
  if (lane == 63) S.red[8 + wv] = shellInc;
  if (t < kNbCells) {
    S.nbAbove[t] = above;
    S.nbBelow[t] = below;
    S.nbBase[t] = nbB;
    S.nbTP[t] = shellPre;
  }
  __syncthreads();
  {
    u32 off = shellInc - shellCnt, tot = 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const u32 c = S.red[8 + k];
      off += k < wv ? c : 0u;
      tot += c;
    }
    if (t < kNbCells) S.nbTB[t] = off;
    if (t == 0) S.nshell = tot;
  }
  // list position: cells with a mixed cube of the earlier waves, then of
  // this wave's earlier lanes
  u32 pre[2] = {0u, 0u}, n = 0u;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const u32 c = S.red[k];
    pre[0] += k < wv ? c : 0u;
    pre[1] += k < 4 + wv ? c : 0u;
    n += c;
  }
  const unsigned long long below_lane = (1ull << lane) - 1ull;
#pragma unroll
  for (int h = 0; h < 2; ++h)
    if (w[h]) S.cl[pre[h] + (u32)__popcll(m[h] & below_lane)] = (uint16_t)(t + 256 * h);
  if (t == 0) S.ncl = n;
  __syncthreads();
}

__device__ __forceinline__ bool block_outside(int blk, int R, int& c0x, int& c0y, int& c0z) {
  block_cell(blk, 0, c0x, c0y, c0z);
  return c0x >= R || c0y >= R || c0z >= R;
}

// Emit order: the blocks with output in kOrderClasses classes by work
// (triangles + 2 x vertices), heaviest class first, so that the longest
// blocks start early. order[0..kOrderClasses) count the members of each
// class; class c lists its members' list positions at
// order[kOrderClasses + c * nblocks ...]. Built by the scan's down-sweep.
// Device scan of the block totals: chunks of kChunk blocks.
constexpr int kChunk = 1024;
constexpr int kPartStride = 32;  // words per chunk total pair in the single-grid scan (one cache line)
constexpr int kOrderClasses = 4;
__device__ __forceinline__ int order_class(u32 work) {
  return work >= 2048u ? 0 : (work >= 1024u ? 1 : (work >= 512u ? 2 : 3));
}
// List position of the i-th block in emit order.
__device__ __forceinline__ u32 order_at(const u32* __restrict__ order, int nblocks, u32 i) {
  u32 c = 0;
  while (c + 1 < (u32)kOrderClasses && i >= order[c]) i -= order[c++];
  return order[kOrderClasses + (size_t)c * nblocks + i];
}

// Morton blocks that can hold a mixed cube: the bricks (1 word x 8 rows x 8
// planes of C) around the block's 17^3 fine voxels are not all empty and not
// all full. Active blocks are appended to list (list[0] = count, then block
// ids, unordered); the others get zero counts. One thread per block.
__device__ __forceinline__ void mc_blocks_body(const uint8_t* __restrict__ brick, int R, int nblocks, u32* blockV,
                                               u32* blockT, u32* list, u32* posOf, int blk) {
  bool act = false;
  if (blk < nblocks) {
    int c0x, c0y, c0z;
    if (!block_outside(blk, R, c0x, c0y, c0z)) {
      const int N = 2 * R, W = N / 32, NB = N / 8;
      const int fx0 = 2 * c0x, fy0 = 2 * c0y, fz0 = 2 * c0z;
      const int w0 = fx0 >> 5, w1 = min((fx0 + 16) >> 5, W - 1);
      const int y0 = fy0 >> 3, y1 = min((fy0 + 16) >> 3, NB - 1);
      const int z0 = fz0 >> 3, z1 = min((fz0 + 16) >> 3, NB - 1);
      u32 any = 0u, all = 2u;
      for (int z = z0; z <= z1; ++z)
        for (int y = y0; y <= y1; ++y)
          for (int w = w0; w <= w1; ++w) {
            const u32 v = brick[((size_t)z * NB + y) * W + w];
            any |= v & 1u;
            all &= v & 2u;
          }
      act = any != 0u && all == 0u;
    }
    if (!act) {
      blockV[blk] = 0u;
      blockT[blk] = 0u;
      posOf[blk] = ~0u;
    }
  }
  const unsigned long long m = __ballot(act);
  if (m) {
    const int lane = (int)(threadIdx.x & 63);
    const int leader = __ffsll((long long)m) - 1;
    u32 base = 0u;
    if (lane == leader) base = atomicAdd(&list[0], (u32)__popcll(m));
    base = __shfl(base, leader, 64);
    if (act) {
      const u32 pos = base + (u32)__popcll(m & ((1ull << lane) - 1ull));
      list[1 + pos] = (u32)blk;
      posOf[blk] = pos;
    }
  }
}

// Active Morton blocks straight from C (one launch instead of a brick
// summary and a block pass): workgroup (BY, BZ) takes the blocks (bx, BY, BZ)
// of every bx. A block is active when its 17^3 fine voxels (the corners of
// its 16^3 cubes, clamped to the grid) are neither all clear nor all set; a
// block with no mixed cube has no vertex, so listing exactly these blocks
// gives the same mesh as any superset. Thread (word w, row y) walks the 17
// planes and folds its word into the blocks it touches: 2w (bits 0..16),
// 2w + 1 (bits 16..31) and 2w - 1 (bit 0). list[0] is zero on entry (the
// finish kernel of the previous extraction clears it).
constexpr int kMaxBlocksPerAxis = 128;
__global__ void __launch_bounds__(512) k_mc_blocks_c(const u32* __restrict__ C, int R, int nb, int nblocks,
                                                      u32* blockV, u32* blockT, u32* list, u32* posOf, u32* order,
                                                      u32* part, u32* totals) {
  __shared__ u32 anyB[kMaxBlocksPerAxis], allB[kMaxBlocksPerAxis];
  if (blockIdx.x == 0 && blockIdx.y == 0) {
    if (threadIdx.x < kOrderClasses) order[threadIdx.x] = 0u;
    if (threadIdx.x == 0) totals[4] = 0u;  // late cubes, tallied by the count kernel
    for (int i = threadIdx.x; i < (nblocks + kChunk - 1) / kChunk; i += blockDim.x) {
      part[kPartStride * i] = 0u;
      part[kPartStride * i + 1] = 0u;
    }
  }
  const int N = 2 * R, W = N / 32;
  const int BY = blockIdx.x, BZ = blockIdx.y;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    anyB[i] = 0u;
    allB[i] = 1u;
  }
  __syncthreads();
  const int fy0 = 16 * BY, fz0 = 16 * BZ;
  if (fy0 < N && fz0 < N) {
    const int ny = min(17, N - fy0), nz = min(17, N - fz0);
    for (int p = threadIdx.x; p < ny * W; p += blockDim.x) {
      const int w = p % W, y = fy0 + p / W;
      const u32* g = C + ((size_t)fz0 * N + y) * W + w;
      u32 v[17];  // the 17 loads in flight together
#pragma unroll
      for (int z = 0; z < 17; ++z) v[z] = z < nz ? g[(size_t)z * N * W] : v[0];
      u32 any = 0u, all = ~0u;
#pragma unroll
      for (int z = 0; z < 17; ++z) {
        any |= v[z];
        all &= v[z];
      }
      // block 2w: bits 0..16; 2w + 1: bits 16..31 (and bit 0 of w + 1); 2w - 1: bit 0
      constexpr u32 kLo = 0x1ffffu, kHi = 0xffff0000u;
      if (any & kLo) atomicOr(&anyB[2 * w], 1u);
      if ((all & kLo) != kLo) atomicAnd(&allB[2 * w], 0u);
      if (any & kHi) atomicOr(&anyB[2 * w + 1], 1u);
      if ((all & kHi) != kHi) atomicAnd(&allB[2 * w + 1], 0u);
      if (w > 0) {
        if (any & 1u) atomicOr(&anyB[2 * w - 1], 1u);
        if (!(all & 1u)) atomicAnd(&allB[2 * w - 1], 0u);
      }
    }
  }
  __syncthreads();
  // one wave lists the row of blocks (nb <= 128: two rounds at most)
  if (threadIdx.x >= 64) return;
  const int lane = (int)threadIdx.x;
  for (int bx0 = 0; bx0 < nb; bx0 += 64) {
    const int bx = bx0 + lane;
    bool act = false;
    u32 blk = 0u;
    if (bx < nb) {
      blk = part1by2((u32)bx) | (part1by2((u32)BY) << 1) | (part1by2((u32)BZ) << 2);
      act = 16 * bx < N && fy0 < N && fz0 < N && anyB[bx] != 0u && allB[bx] == 0u;
      if (!act) {
        blockV[blk] = 0u;
        blockT[blk] = 0u;
        posOf[blk] = ~0u;
      }
    }
    const unsigned long long m = __ballot(act);
    if (m) {
      const int leader = __ffsll((long long)m) - 1;
      u32 base = 0u;
      if (lane == leader) base = atomicAdd(&list[0], (u32)__popcll(m));
      base = __shfl(base, leader, 64);
      if (act) {
        const u32 pos = base + (u32)__popcll(m & ((1ull << lane) - 1ull));
        list[1 + pos] = blk;
        posOf[blk] = pos;
      }
    }
  }
}

// Count kernels: threads per listed block and consecutive cells per thread
// (128 threads with 4 cells each measured slower).
constexpr int kCountThreads = 256;
constexpr int kCountCells = 512 / kCountThreads;

// Triangle counts per case and the listed triangles beyond them (case 91),
// the only tables the count kernels read.
__device__ __forceinline__ void load_ntri(uint8_t* ntri, uint8_t* xtra) {
  for (int i = threadIdx.x; i < 256; i += blockDim.x) {
    ntri[i] = (uint8_t)g_mc.ntri[i];
    xtra[i] = (uint8_t)(g_mc.len[i] - g_mc.ntri[i]);
  }
}

// Vertex and triangle counts of listed block `it`, the block-local vertex
// base of each of its cells with vertices, and the cube cases of its cells
// (for the emit kernel). kCountThreads threads, kCountCells consecutive
// Morton cells each. part (optional): the block's totals are added to its
// scan chunk's totals (kPartStride words per chunk, one cache line each).
__device__ __forceinline__ void mc_count_block(const u32* __restrict__ C, int R, u32 it, int blk, u32* blockV,
                                               u32* blockT, u32* cellL, unsigned long long* cellCase, Box& bx,
                                               u32* red, const uint8_t* ntri, const uint8_t* xtra, u32* totals,
                                               u32* part = nullptr) {
  int c0x, c0y, c0z;
  block_outside(blk, R, c0x, c0y, c0z);
  // cells of the block need fine voxels [16b, 16b+17) per axis
  load_box(bx, C, 2 * R, 2 * c0x, 2 * c0y, 2 * c0z, kBoxRows, kBoxRows);
  // uniform block: no mixed cube (a set and a clear voxel somewhere in the
  // 17^3 voxels its cubes read)
  u32 u = 0u;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) u |= bx.uni[w];
  const bool nonuni = u == 3u;
  u32 nvh[kCountCells] = {}, nth[kCountCells] = {}, nv = 0, nt = 0, nx = 0;
  unsigned long long w[kCountCells] = {};
  if (nonuni) {
#pragma unroll
    for (int h = 0; h < kCountCells; ++h) {
      int cx, cy, cz;
      block_cell(blk, kCountCells * threadIdx.x + h, cx, cy, cz);
      if (!cell_in_m_range(cx, cy, cz, R)) continue;
      w[h] = cell_cases_box(bx, cx, cy, cz);
      nvh[h] = own_prefix8(w[h], 8);
      nv += nvh[h];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        nth[h] += ntri[(w[h] >> (8 * j)) & 0xffu];
        nx += xtra[(w[h] >> (8 * j)) & 0xffu];
      }
      nt += nth[h];
    }
  }
  // listed-but-uncounted triangles anywhere: the emit then leaves the normal
  // sums for the fixup (rare; one atomic per thread that has any)
  if (nx) atomicAdd(&totals[4], nx);
  {
    unsigned long long* dst = cellCase + (size_t)it * 512 + kCountCells * threadIdx.x;
#pragma unroll
    for (int h = 0; h < kCountCells; ++h) dst[h] = w[h];
  }
  // one scan of (vertices | triangles << 16): a block's 4096 cubes own at
  // most 3 vertices and count at most 5 triangles each (12288, 20480), so
  // neither field carries
  u32 tot;
  // (no closing barrier: the caller's loop ends with one)
  u32 off = block_exscan(nv | (nt << 16), red, &tot, false);
  const u32 tv = tot & 0xffffu, tt = tot >> 16;
  // block-local vertex and triangle bases (vertex | triangle << 16) of the
  // cells with a mixed cube (cells are in Morton order, consecutive per
  // thread; stored by list position and local Morton index): the emit of
  // this block stages its cells from them, and the emit of other blocks reads
  // the vertex base for the edges owned by this block's cubes
#pragma unroll
  for (int h = 0; h < kCountCells; ++h) {
    if (w[h]) cellL[(size_t)it * 512 + kCountCells * threadIdx.x + h] = off;
    off += nvh[h] | (nth[h] << 16);
  }
  if (threadIdx.x == 0) {
    blockV[blk] = tv;
    blockT[blk] = tt;
    if (part && (tv | tt)) {
      u32* pc = part + kPartStride * (blk / kChunk);
      atomicAdd(&pc[0], tv);
      atomicAdd(&pc[1], tt);
    }
  }
}

// Vertex and triangle counts of the active blocks; the workgroups walk the
// list (each ends when the list does). 8 waves per SIMD (registers capped to
// fit), so that the 2048 workgroups of a full list are resident at once.
__global__ void __launch_bounds__(kCountThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) k_mc_count(const u32* __restrict__ C, int R,
                                                          const u32* __restrict__ list, u32* blockV,
                                                          u32* blockT, u32* cellL, unsigned long long* cellCase,
                                                          u32* late, u32* part, u32 capList, u32* totals) {
  __shared__ Box bx;
  __shared__ u32 red[16];
  __shared__ uint8_t ntri[256], xtra[256];
  // the late-triangle count starts from zero in the emit kernel that follows
  if (blockIdx.x == 0 && threadIdx.x == 0) late[0] = 0u;
  const u32 n = list[0];
  // more listed blocks than the per-block buffers hold: nothing is stored,
  // the scan flags it and the host grows them and extracts again
  if (blockIdx.x >= n || n > capList) return;
  load_ntri(ntri, xtra);
  for (u32 it = blockIdx.x; it < n; it += gridDim.x) {
    mc_count_block(C, R, it, (int)list[1 + it], blockV, blockT, cellL, cellCase, bx, red, ntri, xtra, totals, part);
    __syncthreads();  // the box is reloaded by the next iteration
  }
}

// ---- device-wide exclusive scan of (blockV, blockT), 1024 entries per chunk

__device__ __forceinline__ void scan_reduce_body(const u32* __restrict__ a, const u32* __restrict__ b, int n, u32* part,
                                                 int chunk) {
  __shared__ u32 red[16];
  int base = chunk * kChunk;
  u32 sa = 0, sb = 0;
  for (int i = base + threadIdx.x; i < min(base + kChunk, n); i += 256) {
    sa += a[i];
    sb += b[i];
  }
  u32 ta, tb;
  block_exscan(sa, red, &ta);
  block_exscan(sb, red, &tb);
  if (threadIdx.x == 0) {
    part[2 * chunk] = ta;
    part[2 * chunk + 1] = tb;
  }
}

// Down-sweep of one chunk: the chunk's offset is the sum of the chunk totals
// before it (part holds the totals of all chunks); chunk 0 also writes the
// overall totals and the overflow flag.
__device__ __forceinline__ void scan_down_body(u32* a, u32* b, int n, const u32* __restrict__ part, int pstride,
                                               int nparts, int chunk, u32* totals, u32 capV, u32 capT,
                                               const u32* __restrict__ posOf = nullptr, u32* order = nullptr,
                                               u32* list = nullptr, u32 capList = 0, bool clearList = false) {
  __shared__ u32 red[16];
  __shared__ u32 cls[2 * kOrderClasses];
  int base = chunk * kChunk + 4 * threadIdx.x;
  u32 va[4], vb[4];
  u32 sa = 0, sb = 0;
  for (int k = 0; k < 4; ++k) {
    int i = base + k;
    va[k] = i < n ? a[i] : 0u;
    vb[k] = i < n ? b[i] : 0u;
    sa += va[k];
    sb += vb[k];
  }
  if (order && threadIdx.x < kOrderClasses) cls[threadIdx.x] = 0u;
  const int lim = chunk == 0 ? nparts : chunk;  // uniform
  u32 pa = 0, pb = 0;
  for (int p = threadIdx.x; p < lim; p += blockDim.x) {
    pa += part[pstride * p];
    pb += part[pstride * p + 1];
  }
  u32 offA, offB;
  block_exscan(pa, red, &offA);
  block_exscan(pb, red, &offB);
  if (chunk == 0) {
    if (threadIdx.x == 0) {
      const u32 listed = list ? list[0] : 0u;
      totals[0] = offA;
      totals[1] = offB;
      totals[2] = ((offA > capV || offB > capT) ? 1u : 0u) | (listed > capList ? 4u : 0u);
      totals[3] = 0u;
      if (clearList) {
        // the count kernel has read it: kept for the host, cleared for the
        // next extraction; the emit's completion counter starts at zero
        totals[5] = listed;
        totals[6] = 0u;
        list[0] = 0u;
      }
    }
    offA = offB = 0u;
  }
  u32 ta, tb;
  u32 oa = block_exscan(sa, red, &ta) + offA;
  u32 ob = block_exscan(sb, red, &tb) + offB;
  for (int k = 0; k < 4; ++k) {
    int i = base + k;
    if (i < n) {
      a[i] = oa;
      b[i] = ob;
    }
    oa += va[k];
    ob += vb[k];
  }
  if (order) {
    // the chunk's blocks with output join their emit-order class: slots
    // within the chunk from LDS counters, one global reservation per class
    u32 slot[4];
    for (int k = 0; k < 4; ++k) {
      const u32 work = vb[k] + 2u * va[k];
      slot[k] = work ? atomicAdd(&cls[order_class(work)], 1u) : 0u;
    }
    __syncthreads();
    if (threadIdx.x < kOrderClasses) {
      const u32 c = cls[threadIdx.x];
      cls[kOrderClasses + threadIdx.x] = c ? atomicAdd(&order[threadIdx.x], c) : 0u;
    }
    __syncthreads();
    for (int k = 0; k < 4; ++k) {
      const u32 work = vb[k] + 2u * va[k];
      if (!work) continue;
      const int c = order_class(work);
      order[kOrderClasses + (size_t)c * n + cls[kOrderClasses + c] + slot[k]] = posOf[base + k];
    }
  }
}
__global__ void __launch_bounds__(256) k_scan_down(u32* a, u32* b, int n, const u32* __restrict__ part, int nparts,
                                                   u32* totals, u32 capV, u32 capT, const u32* __restrict__ posOf,
                                                   u32* order, u32* list, u32 capList) {
  scan_down_body(a, b, n, part, kPartStride, nparts, blockIdx.x, totals, capV, capT, posOf, order, list, capList,
                 true);
}

// Position of the vertex on owned edge e (0: x, 3: y, 8: z) of cube f:
// findIntersection (MarchingCubesMesher.cu:98-100) with {0,1} corner values.
// The division by v2 - v1 = +-1 is exact, so it is applied as a sign.
__device__ __forceinline__ float lerp_edge(float p1, float p2, float v1, float v2) {
  float d = __fsub_rn(p2, p1);
  float t = __fsub_rn(0.95f, v1);
  float m = __fmul_rn(t, d);
  float q = (v2 > v1) ? m : -m;
  return __fadd_rn(p1, q);
}
struct Geo {
  float bx, by, bz, h;
};

// Owner of local edge e of a cube: offset (dx,dy,dz) of the owning cube and
// which of its own edges (0, 3, 8) it is, packed 5 bits per edge:
// dx | dy<<1 | dz<<2 | k<<3 with k = 0, 1, 2 for edges 0, 3, 8.
constexpr unsigned long long kOwnPacked =
    (0ull << 0) | ((1ull | (1ull << 3)) << 5) | (2ull << 10) | ((1ull << 3) << 15) | (4ull << 20) |
    ((5ull | (1ull << 3)) << 25) | (6ull << 30) | ((4ull | (1ull << 3)) << 35) | ((2ull << 3) << 40) |
    ((1ull | (2ull << 3)) << 45) | ((3ull | (2ull << 3)) << 50) | ((2ull | (2ull << 3)) << 55);
__device__ __forceinline__ void own_of(int e, int& dx, int& dy, int& dz, int& oe) {
  const int v = (int)((kOwnPacked >> (5 * e)) & 31ull);
  dx = v & 1;
  dy = (v >> 1) & 1;
  dz = (v >> 2) & 1;
  const int k = v >> 3;
  oe = k == 0 ? 0 : (k == 1 ? 3 : 8);
}

__device__ __forceinline__ int rank_of(int v, int e) {  // index of owned edge e among the cube's own vertices
  int c0 = v & 1;
  int e0 = c0 ^ ((v >> 1) & 1), e3 = c0 ^ ((v >> 3) & 1);
  return e == 0 ? 0 : (e == 3 ? e0 : e0 + e3);
}

// Area-weighted normal sums (GMorpho.cu:6498-6524) gathered per vertex: the
// triangles that use the vertex's edge are those of the <= 4 cubes around it.
struct ShareCube {
  int dx, dy, dz, le;
};
constexpr ShareCube kShare[3][4] = {
    // owned edge 0 (x): cubes f, f-y, f-z, f-y-z with local edges 0, 2, 4, 6
    {{0, 0, 0, 0}, {0, -1, 0, 2}, {0, 0, -1, 4}, {0, -1, -1, 6}},
    // owned edge 3 (y): cubes f, f-x, f-z, f-x-z with local edges 3, 1, 7, 5
    {{0, 0, 0, 3}, {-1, 0, 0, 1}, {0, 0, -1, 7}, {-1, 0, -1, 5}},
    // owned edge 8 (z): cubes f, f-x, f-y, f-x-y with local edges 8, 9, 11, 10
    {{0, 0, 0, 8}, {-1, 0, 0, 9}, {0, -1, 0, 11}, {-1, -1, 0, 10}}};
// The 4 cubes of owned edge q packed one byte each: bit0/1/2 = offset -1 in
// x/y/z, bits 3..6 = the edge's index in that cube. Selected by value (q is
// per lane), so no table is read from memory.
constexpr u32 share_pack(int q) {
  u32 v = 0;
  for (int s = 0; s < 4; ++s) {
    const ShareCube& c = kShare[q][s];
    v |= (u32)((c.dx < 0 ? 1 : 0) | (c.dy < 0 ? 2 : 0) | (c.dz < 0 ? 4 : 0) | (c.le << 3)) << (8 * s);
  }
  return v;
}
__device__ __forceinline__ u32 share_of(int q) {
  return q == 0 ? share_pack(0) : (q == 1 ? share_pack(1) : share_pack(2));
}

// edge_pos with the two voxel values given (v0 at (fx,fy,fz), v1 one step
// along the edge's axis): the same float operations.
__device__ __forceinline__ void edge_pos_v(const Geo& g, int fx, int fy, int fz, int e, u32 v0, u32 v1, float* o) {
  float px = fmaf((float)fx, g.h, g.bx);
  float py = fmaf((float)fy, g.h, g.by);
  float pz = fmaf((float)fz, g.h, g.bz);
  o[0] = px;
  o[1] = py;
  o[2] = pz;
  if (e == 0) o[0] = lerp_edge(px, __fadd_rn(px, g.h), (float)v0, (float)v1);
  else if (e == 3) o[1] = lerp_edge(__fadd_rn(py, g.h), py, (float)v1, (float)v0);
  else o[2] = lerp_edge(pz, __fadd_rn(pz, g.h), (float)v0, (float)v1);
}
// Corner index (case bit order) of the corner at offset (ox, oy, oz).
__device__ __forceinline__ int corner_of(int ox, int oy, int oz) { return 4 * oz + (oy ? (ox ? 2 : 3) : ox); }
// Position of the vertex on local edge e of cube (fx,fy,fz) with case v: both
// ends of the edge are corners of the cube, so their values are bits of v.
__device__ __forceinline__ void local_edge_pos(const Geo& g, int fx, int fy, int fz, int v, int e, float* p) {
  int dx, dy, dz, oe;
  own_of(e, dx, dy, dz, oe);
  const int ca = corner_of(dx, dy, dz);
  const int cb = corner_of(dx + (oe == 0), dy + (oe == 3), dz + (oe == 8));
  edge_pos_v(g, fx + dx, fy + dy, fz + dz, oe, (u32)(v >> ca) & 1u, (u32)(v >> cb) & 1u, p);
}

// Face normal of a triangle from its three corner positions (the reference's
// cross product, GMorpho.cu:6498-6524).
__device__ __forceinline__ void face_cross(const float* p0, const float* p1, const float* p2, float* f) {
  float a0 = __fsub_rn(p1[0], p0[0]), a1 = __fsub_rn(p1[1], p0[1]), a2 = __fsub_rn(p1[2], p0[2]);
  float b0 = __fsub_rn(p2[0], p0[0]), b1 = __fsub_rn(p2[1], p0[1]), b2 = __fsub_rn(p2[2], p0[2]);
  f[0] = __fsub_rn(__fmul_rn(a1, b2), __fmul_rn(a2, b1));
  f[1] = __fsub_rn(__fmul_rn(a2, b0), __fmul_rn(a0, b2));
  f[2] = __fsub_rn(__fmul_rn(a0, b1), __fmul_rn(a1, b0));
}

__device__ __forceinline__ u32 prefix_ntri(const BlockLds& S, int l, int j) { return S.tpre[8 * l + j]; }

// v0: the case of cube f; pv: the vertex's own position (the same float
// values local_edge_pos gives for its edge in every sharing cube). With
// useFn, the face normals of the block's own cubes come from S.fn (written
// by the triangle phase from the same positions), and with useShell those
// of the shell cubes too, from shellBase on; otherwise cubes of other blocks
// are computed here. The sum runs in the same order either way.
__device__ void normal_sum(const BlockLds& S, const Geo& geo, int R, int c0x, int c0y, int c0z, bool useFn,
                           bool useShell, u32 shellBase, int fx, int fy, int fz, int q, int v0, const float* pv,
                           float* n) {
  n[0] = n[1] = n[2] = 0.f;
  const u32 sp = share_of(q);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const u32 e = (sp >> (8 * s)) & 0xffu;
    const int gx = fx - (int)(e & 1u), gy = fy - (int)((e >> 1) & 1u), gz = fz - (int)((e >> 2) & 1u);
    const int le = (int)(e >> 3);
    const int cx = gx >> 1, cy = gy >> 1, cz = gz >> 1;
    if (s > 0 && !cell_in_m_range(cx, cy, cz, R)) continue;
    const int lx = cx - c0x, ly = cy - c0y, lz = cz - c0z;
    const int j = (gx & 1) | ((gy & 1) << 1) | ((gz & 1) << 2);
    const bool inBlock = (unsigned)lx < 8u && (unsigned)ly < 8u && (unsigned)lz < 8u;
    const int l = inBlock ? local_morton(lx, ly, lz) : 0;
    // a cube of the block, or of a cell just below it
    const int gv = s == 0 ? v0
                          : (inBlock ? (int)S.cases[8 * l + j]
                                     : (int)((S.nbBelow[nb_index(7 - lx, 7 - ly, 7 - lz)] >> (8 * j)) & 0xffu));
    if (!mixed(gv)) continue;
    if (inBlock ? useFn : useShell) {
      u32 lt;
      if (inBlock) {
        lt = (S.vtOff[l] >> 16) + prefix_ntri(S, l, j);
      } else {
        const int k = nb_index(7 - lx, 7 - ly, 7 - lz);
        lt = shellBase + S.nbTB[k] + (u32)((S.nbTP[k] >> (8 * j)) & 0xffull);
      }
      u32 use = S.tab.use[gv][le];
      while (use) {
        const u32 t = (u32)__builtin_ctz(use);
        use &= use - 1u;
        const float* f = S.fn[lt + t];
        n[0] = __fadd_rn(n[0], f[0]);
        n[1] = __fadd_rn(n[1], f[1]);
        n[2] = __fadd_rn(n[2], f[2]);
      }
      continue;
    }
    unsigned long long tri = S.tab.tri[gv];
    // the cube's counted triangles that use edge le, in order
    u32 use = S.tab.use[gv][le];
    while (use) {
      const u32 t = (u32)__builtin_ctz(use);
      use &= use - 1u;
      int e0 = (int)((tri >> (12 * t)) & 0xf), e1 = (int)((tri >> (12 * t + 4)) & 0xf),
          e2 = (int)((tri >> (12 * t + 8)) & 0xf);
      float p0[3], p1[3], p2[3];
      auto pos = [&](int e, float* o) {
        if (e == le) {
          o[0] = pv[0];
          o[1] = pv[1];
          o[2] = pv[2];
        } else {
          local_edge_pos(geo, gx, gy, gz, gv, e, o);
        }
      };
      pos(e2, p0);
      pos(e1, p1);
      pos(e0, p2);
      float f[3];
      face_cross(p0, p1, p2, f);
      n[0] = __fadd_rn(n[0], f[0]);
      n[1] = __fadd_rn(n[1], f[1]);
      n[2] = __fadd_rn(n[2], f[2]);
    }
  }
}

__device__ __forceinline__ unsigned long long cell_cases(const BlockLds& S, int l) {
  return *reinterpret_cast<const unsigned long long*>(&S.cases[8 * l]);
}

// Vertices of a staged block, in two phases per window of kListCap local
// vertex ids: the mixed cubes write (cube, owned edge) at their local id,
// then every thread takes consecutive ids (position + gathered normal sum).
__device__ __forceinline__ void unit3(float* n) {
  const float l = sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(n[0], n[0]), __fmul_rn(n[1], n[1])), __fmul_rn(n[2], n[2])));
  if (l > 0.f) {
    n[0] = __fdiv_rn(n[0], l);
    n[1] = __fdiv_rn(n[1], l);
    n[2] = __fdiv_rn(n[2], l);
  }
}

// unit: the sums are final (no listed-but-uncounted triangle anywhere), so
// they are stored normalized (NormalizeMeshNormals, GMorpho.cu:6526-6537).
__device__ void mc_verts_block(BlockLds& S, int R, Geo geo, int blk, int c0x, int c0y, int c0z, bool useFn,
                               bool useShell, u32 shellBase, u32 base, u32 nverts, float* V, float* Nrm, u32 capV,
                               bool unit) {
  for (u32 w0 = 0; w0 < nverts; w0 += kListCap) {
    for (u32 i = threadIdx.x; i < S.ncl; i += blockDim.x) {
      const int l = S.cl[i];
      const unsigned long long w = cell_cases(S, l);
      // the cell's vertices, in cube and edge order, from its local base
      u32 lv = (S.vtOff[l] & 0xffffu) - w0;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int v = (int)((w >> (8 * j)) & 0xffu);
        const int c0 = v & 1;
        const int on[3] = {c0 ^ ((v >> 1) & 1), c0 ^ ((v >> 3) & 1), c0 ^ ((v >> 4) & 1)};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          if (!on[q]) continue;
          if (lv < (u32)kListCap) S.list[lv] = (uint16_t)(((8 * l + j) << 3) | q);
          ++lv;
        }
      }
    }
    __syncthreads();
    const u32 n = min((u32)kListCap, nverts - w0);
    for (u32 i = threadIdx.x; i < n; i += blockDim.x) {
      const u32 vid = base + w0 + i;
      if (vid >= capV) break;
      const int e = S.list[i];
      const int s = e >> 3, q = e & 7, l = s >> 3, j = s & 7;
      int cx, cy, cz;
      local_cell(l, cx, cy, cz);
      const int fx = 2 * (c0x + cx) + (j & 1), fy = 2 * (c0y + cy) + ((j >> 1) & 1), fz = 2 * (c0z + cz) + ((j >> 2) & 1);
      float p[3], nn[3];
      const int v = S.cases[s];
      // own edge q runs from corner 0 to corner 1 (x), 3 (y) or 4 (z)
      edge_pos_v(geo, fx, fy, fz, q == 0 ? 0 : (q == 1 ? 3 : 8), (u32)v & 1u,
                 (u32)(v >> (q == 0 ? 1 : (q == 1 ? 3 : 4))) & 1u, p);
      normal_sum(S, geo, R, c0x, c0y, c0z, useFn, useShell, shellBase, fx, fy, fz, q, v, p, nn);
      if (unit) unit3(nn);
      V[3 * (size_t)vid] = p[0];
      V[3 * (size_t)vid + 1] = p[1];
      V[3 * (size_t)vid + 2] = p[2];
      Nrm[3 * (size_t)vid] = nn[0];
      Nrm[3 * (size_t)vid + 1] = nn[1];
      Nrm[3 * (size_t)vid + 2] = nn[2];
    }
    __syncthreads();
  }
}

// Vertex id of local edge e of the cube at fine (fx,fy,fz): 0 when the owning
// cube's cell is outside [1,R-2]^3 (the reference's absent neighbour).
__device__ __forceinline__ u32 edge_vid(const BlockLds& S, int blk, int c0x, int c0y, int c0z, u32 base,
                                        const u32* __restrict__ blockV, const u32* __restrict__ cellL, int R, int fx,
                                        int fy, int fz, int e, u32* border) {
  int dx, dy, dz, oe;
  own_of(e, dx, dy, dz, oe);
  int gx = fx + dx, gy = fy + dy, gz = fz + dz;
  int cx = gx >> 1, cy = gy >> 1, cz = gz >> 1;
  if (!cell_in_m_range(cx, cy, cz, R)) {
    *border = 1u;
    return 0u;
  }
  int j = (gx & 1) | ((gy & 1) << 1) | ((gz & 1) << 2);
  int lx = cx - c0x, ly = cy - c0y, lz = cz - c0z;
  if (lx < 8 && ly < 8 && lz < 8) {
    const int l = local_morton(lx, ly, lz);
    const unsigned long long w = cell_cases(S, l);
    return base + (S.vtOff[l] & 0xffffu) + own_prefix8(w, j) + (u32)rank_of((int)((w >> (8 * j)) & 0xffu), oe);
  }
  // a cell just above the block (staged): its global base + in-cell prefix
  const int k = nb_index(lx, ly, lz);
  const unsigned long long w = S.nbAbove[k];
  return S.nbBase[k] + own_prefix8(w, j) + (u32)rank_of((int)((w >> (8 * j)) & 0xffu), oe);
}

// Triangles of a staged block, in the same two phases over local triangle
// slots. Listed-but-uncounted triangles (len > ntri) go to the late list
// from the fill phase of the first window. With useFn (one window) each
// counted triangle's face normal goes to S.fn at its local index; with
// useShell the counted triangles of the shell cubes follow as extra items
// (face normal only, at ntris + their shell index).
__device__ void mc_tris_block(BlockLds& S, int R, Geo geo, int blk, int c0x, int c0y, int c0z, bool useFn,
                              bool useShell, u32 vbase, u32 tbase, u32 ntris, const u32* __restrict__ blockV,
                              const u32* __restrict__ cellL, u32* T, u32 capT, u32* late, u32 capLate, u32* totals) {
  u32 border = 0u;
  for (u32 w0 = 0; w0 < ntris || w0 == 0; w0 += kListCap) {
    if (useShell)  // one window: ntris + nshell <= kFaceCap <= kListCap
      for (u32 k = threadIdx.x; k < (u32)kNbCells; k += blockDim.x) {
        const unsigned long long wb = S.nbBelow[k];
        if (!wb) continue;
        int lx, ly, lz;
        nb_cell((int)k, lx, ly, lz);
        const int need = shell_axes(lx, ly, lz);
        u32 lt = ntris + S.nbTB[k];
        for (int j = 0; j < 8; ++j) {
          if ((j & need) != need) continue;
          const u32 cnt = S.tab.ntri[(wb >> (8 * j)) & 0xffull];
          for (u32 q = 0; q < cnt; ++q) S.list[lt + q] = (uint16_t)(((4096u + 8u * k + (u32)j) << 3) | q);
          lt += cnt;
        }
      }
    for (u32 i = threadIdx.x; i < S.ncl; i += blockDim.x) {
      const int l = S.cl[i];
      const unsigned long long w = cell_cases(S, l);
      // the cell's counted triangles, in cube order, from its local base;
      // the first window also stores the in-cell prefix per cube
      const u32 lt0 = S.vtOff[l] >> 16;
      u32 lt = lt0;
      unsigned long long tp = 0ull;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int v = (int)((w >> (8 * j)) & 0xffu);
        const int s = 8 * l + j;
        const u32 cnt = S.tab.ntri[v], len = S.tab.len[v];
        tp |= (unsigned long long)(lt - lt0) << (8 * j);
        for (u32 q = 0; q < cnt; ++q)
          if (lt + q - w0 < (u32)kListCap) S.list[lt + q - w0] = (uint16_t)((s << 3) | q);
        if (w0 == 0 && len > cnt) {
          int cx, cy, cz;
          local_cell(l, cx, cy, cz);
          const int fx = 2 * (c0x + cx) + (j & 1), fy = 2 * (c0y + cy) + ((j >> 1) & 1),
                    fz = 2 * (c0z + cz) + ((j >> 2) & 1);
          const unsigned long long tri = S.tab.tri[v];
          for (u32 q = cnt; q < len; ++q) {
            const int e0 = (int)((tri >> (12 * q)) & 0xf), e1 = (int)((tri >> (12 * q + 4)) & 0xf),
                      e2 = (int)((tri >> (12 * q + 8)) & 0xf);
            const u32 a = edge_vid(S, blk, c0x, c0y, c0z, vbase, blockV, cellL, R, fx, fy, fz, e2, &border);
            const u32 b = edge_vid(S, blk, c0x, c0y, c0z, vbase, blockV, cellL, R, fx, fy, fz, e1, &border);
            const u32 c = edge_vid(S, blk, c0x, c0y, c0z, vbase, blockV, cellL, R, fx, fy, fz, e0, &border);
            // listed but not counted: overwrites slot sl after all counted writes
            const u32 k = atomicAdd(&late[0], 1u);
            if (k < capLate) {
              late[1 + 4 * k] = tbase + lt + q;
              late[2 + 4 * k] = a;
              late[3 + 4 * k] = b;
              late[4 + 4 * k] = c;
            } else {
              atomicOr(&totals[2], 2u);
            }
          }
        }
        lt += cnt;
      }
      if (w0 == 0) *reinterpret_cast<unsigned long long*>(&S.tpre[8 * l]) = tp;
    }
    __syncthreads();
    const u32 n = min((u32)kListCap, ntris - w0);
    const u32 nx = n + (useShell ? S.nshell : 0u);
    for (u32 i = threadIdx.x; i < nx; i += blockDim.x) {
      const int e = S.list[i];
      if (i >= n) {
        // a shell cube's triangle: its face normal only
        const int s = (e >> 3) - 4096, q = e & 7, k = s >> 3, j = s & 7;
        int lx, ly, lz;
        nb_cell(k, lx, ly, lz);
        const int fx = 2 * (c0x + 7 - lx) + (j & 1), fy = 2 * (c0y + 7 - ly) + ((j >> 1) & 1),
                  fz = 2 * (c0z + 7 - lz) + ((j >> 2) & 1);
        const int v = (int)((S.nbBelow[k] >> (8 * j)) & 0xffull);
        const unsigned long long tri = S.tab.tri[v];
        float p0[3], p1[3], p2[3];
        local_edge_pos(geo, fx, fy, fz, v, (int)((tri >> (12 * q + 8)) & 0xf), p0);
        local_edge_pos(geo, fx, fy, fz, v, (int)((tri >> (12 * q + 4)) & 0xf), p1);
        local_edge_pos(geo, fx, fy, fz, v, (int)((tri >> (12 * q)) & 0xf), p2);
        face_cross(p0, p1, p2, S.fn[i]);
        continue;
      }
      const u32 sl = tbase + w0 + i;
      if (sl >= capT) continue;
      const int s = e >> 3, q = e & 7, l = s >> 3, j = s & 7;
      int cx, cy, cz;
      local_cell(l, cx, cy, cz);
      const int fx = 2 * (c0x + cx) + (j & 1), fy = 2 * (c0y + cy) + ((j >> 1) & 1), fz = 2 * (c0z + cz) + ((j >> 2) & 1);
      const unsigned long long tri = S.tab.tri[S.cases[s]];
      const int e0 = (int)((tri >> (12 * q)) & 0xf), e1 = (int)((tri >> (12 * q + 4)) & 0xf),
                e2 = (int)((tri >> (12 * q + 8)) & 0xf);
      T[3 * (size_t)sl] = edge_vid(S, blk, c0x, c0y, c0z, vbase, blockV, cellL, R, fx, fy, fz, e2, &border);
      T[3 * (size_t)sl + 1] = edge_vid(S, blk, c0x, c0y, c0z, vbase, blockV, cellL, R, fx, fy, fz, e1, &border);
      T[3 * (size_t)sl + 2] = edge_vid(S, blk, c0x, c0y, c0z, vbase, blockV, cellL, R, fx, fy, fz, e0, &border);
      if (useFn) {
        const int v = S.cases[s];
        float p0[3], p1[3], p2[3];
        local_edge_pos(geo, fx, fy, fz, v, e2, p0);
        local_edge_pos(geo, fx, fy, fz, v, e1, p1);
        local_edge_pos(geo, fx, fy, fz, v, e0, p2);
        face_cross(p0, p1, p2, S.fn[i]);
      }
    }
    __syncthreads();
  }
  if (border) atomicOr(&totals[3], 1u);
}

// Listed block `it` (id blk): staged once, then its triangles and its
// vertices.
__device__ __forceinline__ void mc_emit_block(BlockLds& S, int R, Geo geo, u32 it, int blk, int nblocks,
                                              const u32* __restrict__ blockV, const u32* __restrict__ blockT,
                                              const u32* __restrict__ cellL,
                                              const unsigned long long* __restrict__ cellCase,
                                              const u32* __restrict__ posOf, float* V, float* Nrm, u32 capV, u32* T,
                                              u32 capT, u32* late, u32 capLate, u32* totals) {
  int c0x, c0y, c0z;
  block_outside(blk, R, c0x, c0y, c0z);
  const u32 base = blockV[blk];
  const u32 next = (blk + 1 < nblocks) ? blockV[blk + 1] : totals[0];
  const u32 tbase = blockT[blk];
  const u32 tnext = (blk + 1 < nblocks) ? blockT[blk + 1] : totals[1];
  if (next == base && tnext == tbase) return;  // uniform across the workgroup
  stage_block(S, cellCase, posOf, blockV, cellL, R, it, c0x, c0y, c0z);
  // triangles first: their face normals (when they fit in LDS) are then
  // gathered by the vertices
  const u32 ntris = tnext - tbase;
  const bool useFn = ntris <= (u32)kFaceCap && tnext <= capT;
  // the shell's face normals too, when they fit after the block's
  const bool useShell = useFn && ntris + S.nshell <= (u32)kFaceCap;
  mc_tris_block(S, R, geo, blk, c0x, c0y, c0z, useFn, useShell, base, tbase, ntris, blockV, cellL, T, capT, late,
                capLate, totals);
  mc_verts_block(S, R, geo, blk, c0x, c0y, c0z, useFn, useShell, ntris, base, next - base, V, Nrm, capV,
                 totals[4] == 0u);
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

// Applies the listed-but-uncounted triangles in slot order: each replaces the
// triangle in its slot, and the normal sums follow. The workgroup first ranks
// the entries by slot (slots are unique: each is the first slot of the next
// cube with triangles) into the list's second half, then one lane applies
// them in that order, so the float sums are deterministic. Entries whose slot
// or vertex ids fall outside the mesh are skipped.
__device__ __forceinline__ void mc_fixup_body(const float* V, float* Nrm, u32* T, u32* late, u32 capLate,
                                              u32* totals, u32 capV, u32 capT) {
  u32 n = late[0];
  if (n > capLate) n = capLate;
  if (n == 0) return;  // uniform over the workgroup
  u32* sorted = late + 1 + 4 * (size_t)capLate;
  for (u32 i = threadIdx.x; i < n; i += blockDim.x) {
    const u32 si = late[1 + 4 * i];
    u32 rank = 0;
    for (u32 j = 0; j < n; ++j) {
      const u32 sj = late[1 + 4 * j];
      rank += (sj < si || (sj == si && j < i)) ? 1u : 0u;
    }
    for (int q = 0; q < 4; ++q) sorted[4 * rank + q] = late[1 + 4 * i + q];
  }
  __threadfence_block();
  __syncthreads();
  const u32 nT = totals[1], nV = totals[0];
  if (threadIdx.x != 0 || nT > capT || nV > capV) return;
  for (u32 i = 0; i < n; ++i) {
    const u32 s = sorted[4 * i];
    const u32 a = sorted[4 * i + 1], b = sorted[4 * i + 2], c = sorted[4 * i + 3];
    if (s >= nT || a >= nV || b >= nV || c >= nV) continue;
    u32* t = T + 3 * (size_t)s;
    if (t[0] >= nV || t[1] >= nV || t[2] >= nV) continue;
    float f[3];
    face(V, t, f);
    for (int k = 0; k < 3; ++k)
      for (int q = 0; q < 3; ++q) Nrm[3 * (size_t)t[k] + q] = __fsub_rn(Nrm[3 * (size_t)t[k] + q], f[q]);
    t[0] = a;
    t[1] = b;
    t[2] = c;
    face(V, t, f);
    for (int k = 0; k < 3; ++k)
      for (int q = 0; q < 3; ++q) Nrm[3 * (size_t)t[k] + q] = __fadd_rn(Nrm[3 * (size_t)t[k] + q], f[q]);
  }
}
// NormalizeMeshNormals (GMorpho.cu:6526-6537).
__device__ __forceinline__ void mc_normalize_body(float* Nrm, const u32* totals, u32 capV, size_t first,
                                                  size_t stride) {
  u32 nV = totals[0];
  if (nV > capV) nV = capV;
  for (size_t i = first; i < nV; i += stride) {
    float n[3] = {Nrm[3 * i], Nrm[3 * i + 1], Nrm[3 * i + 2]};
    unit3(n);
    Nrm[3 * i] = n[0];
    Nrm[3 * i + 1] = n[1];
    Nrm[3 * i + 2] = n[2];
  }
}

// The mesh's last step: nothing when no cube lists more triangles than it
// counts (the emit stored unit normals); otherwise the late triangles are
// applied, then every normal is normalized, by one workgroup (rare). With a
// list (batched extraction) it also clears the block-list count for the
// next extraction, keeping it in totals[5] for the host (read when the list
// overflowed); a single extraction's scan does that.
constexpr int kFinishThreads = 1024;
__device__ __forceinline__ void mc_finish_body(const float* V, float* Nrm, u32* T, u32* late, u32 capLate,
                                               u32* totals, u32 capV, u32 capT, u32* list) {
  if (list && threadIdx.x == 0) {
    const u32 listed = list[0];
    if (listed) {
      totals[5] = listed;
      list[0] = 0u;
    }
  }
  if (totals[4] == 0u) return;
  mc_fixup_body(V, Nrm, T, late, capLate, totals, capV, capT);
  __threadfence_block();
  __syncthreads();
  mc_normalize_body(Nrm, totals, capV, threadIdx.x, blockDim.x);
}

// Vertices and triangles of the active blocks: each listed block is staged
// once (voxel box, cube cases, per-cell prefix sums), then its vertices and
// its triangles are written. Vertex ids on edges owned by another block's
// cube come from that block's base and the per-cell local bases written by
// k_mc_count, so no block waits for another.
__global__ void __launch_bounds__(kMcThreads) __attribute__((amdgpu_waves_per_eu(4))) k_mc_emit(int R, Geo geo, const u32* __restrict__ list, int nblocks,
                                                         const u32* __restrict__ blockV,
                                                         const u32* __restrict__ blockT,
                                                         const u32* __restrict__ cellL,
                                                         const unsigned long long* __restrict__ cellCase,
                                                         const u32* __restrict__ posOf,
                                                         const u32* __restrict__ order, float* V, float* Nrm,
                                                         u32 capV, u32* T, u32 capT, u32* late, u32 capLate,
                                                         u32* totals, u32 capList) {
  __shared__ BlockLds S;
  __shared__ u32 lastGroup;
  if (totals[2] & 4u) return;  // the per-block buffers overflowed: extracted again
  u32 nwork = 0u;  // blocks with output
  for (int c = 0; c < kOrderClasses; ++c) nwork += order[c];
  if (blockIdx.x >= nwork) return;  // the grid is sized for every block of the grid
  load_mc_tab(S.tab);
  for (u32 i = blockIdx.x; i < nwork; i += gridDim.x) {
    __syncthreads();  // LDS of the previous block is no longer read
    const u32 it = order_at(order, nblocks, i);
    mc_emit_block(S, R, geo, it, (int)list[1 + it], nblocks, blockV, blockT, cellL, cellCase, posOf, V, Nrm, capV, T,
                  capT, late, capLate, totals);
  }
  // Cubes listing more triangles than they count (rare; the count kernel
  // tallied them): every block stored raw normal sums, and the last
  // workgroup to finish applies the late triangles and normalizes. Only
  // then do the workgroups publish their writes (release, then the
  // completion count in totals[6], zeroed by the scan; acquire in the last).
  if (totals[4] == 0u) return;  // uniform: unit normals are already stored
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const u32 done = atomicAdd(&totals[6], 1u) + 1u;
    lastGroup = done == min(gridDim.x, nwork) ? 1u : 0u;
  }
  __syncthreads();
  if (!lastGroup) return;
  __threadfence();
  mc_finish_body(V, Nrm, T, late, capLate, totals, capV, capT, nullptr);
  if (threadIdx.x == 0) totals[6] = 0u;  // a re-emit into grown buffers counts again
}

// ---- Batched forms: many closed grids in one launch per step.
struct MCJob {
  const u32* C;
  int R;
  Geo geo;
  MCBuffers m;
};
__global__ void __launch_bounds__(256) k_mc_blocks_b(const MCJob* __restrict__ jobs, const u32* __restrict__ pre,
                                                     int n) {
  const int j = batch_job(pre, n, blockIdx.x);
  const MCJob& J = jobs[j];
  // each job's emit order is built too (a mesh that overflows its output is
  // emitted again by the single-job emit, which walks that order)
  if (blockIdx.x == pre[j] && threadIdx.x < kOrderClasses) J.m.order[threadIdx.x] = 0u;
  if (blockIdx.x == pre[j] && threadIdx.x == 0) J.m.totals[4] = 0u;
  mc_blocks_body(J.m.brick, J.R, J.m.nblocks, J.m.blockV, J.m.blockT, J.m.list, J.m.posOf,
                 (int)(blockIdx.x - pre[j]) * 256 + (int)threadIdx.x);
}
// Prefix of the jobs' active-block list lengths (one workgroup).
__global__ void k_list_prefix(const MCJob* __restrict__ jobs, int n, u32* lpre) {
  if (threadIdx.x != 0) return;
  u32 acc = 0;
  for (int j = 0; j < n; ++j) {
    lpre[j] = acc;
    acc += jobs[j].m.list[0];
  }
  lpre[n] = acc;
}
__global__ void __launch_bounds__(kCountThreads) k_mc_count_b(const MCJob* __restrict__ jobs, const u32* __restrict__ lpre,
                                                            int n) {
  __shared__ Box bx;
  __shared__ u32 red[16];
  __shared__ uint8_t ntri[256], xtra[256];
  if (blockIdx.x >= lpre[n]) return;
  load_ntri(ntri, xtra);
  for (u32 u = blockIdx.x; u < lpre[n]; u += gridDim.x) {
    const int j = batch_job(lpre, n, u);
    const MCJob& J = jobs[j];
    if (J.m.list[0] > J.m.capList) continue;  // uniform per workgroup; flagged by the scan
    mc_count_block(J.C, J.R, u - lpre[j], (int)J.m.list[1 + u - lpre[j]], J.m.blockV, J.m.blockT, J.m.cellV,
                   J.m.cellCase, bx, red, ntri, xtra, J.m.totals);
    __syncthreads();
  }
}
__global__ void __launch_bounds__(256) k_scan_reduce_b(const MCJob* __restrict__ jobs, const u32* __restrict__ pre,
                                                       int n) {
  const int j = batch_job(pre, n, blockIdx.x);
  const MCJob& J = jobs[j];
  scan_reduce_body(J.m.blockV, J.m.blockT, J.m.nblocks, J.m.scanPart, (int)(blockIdx.x - pre[j]));
}
__global__ void __launch_bounds__(256) k_scan_down_b(const MCJob* __restrict__ jobs, const u32* __restrict__ pre,
                                                     int n) {
  const int j = batch_job(pre, n, blockIdx.x);
  const MCJob& J = jobs[j];
  scan_down_body(J.m.blockV, J.m.blockT, J.m.nblocks, J.m.scanPart, 2, (J.m.nblocks + kChunk - 1) / kChunk,
                 (int)(blockIdx.x - pre[j]), J.m.totals, J.m.capV, J.m.capT, J.m.posOf, J.m.order, J.m.list,
                 J.m.capList);
}
__global__ void __launch_bounds__(kMcThreads) __attribute__((amdgpu_waves_per_eu(4))) k_mc_emit_b(const MCJob* __restrict__ jobs, const u32* __restrict__ lpre,
                                                           int n) {
  __shared__ BlockLds S;
  if (blockIdx.x >= lpre[n]) return;
  load_mc_tab(S.tab);
  for (u32 u = blockIdx.x; u < lpre[n]; u += gridDim.x) {
    __syncthreads();  // LDS of the previous block is no longer read
    const int j = batch_job(lpre, n, u);
    const MCJob& J = jobs[j];
    if (J.m.list[0] > J.m.capList) continue;  // uniform per workgroup; flagged by the scan
    mc_emit_block(S, J.R, J.geo, u - lpre[j], (int)J.m.list[1 + u - lpre[j]], J.m.nblocks, J.m.blockV, J.m.blockT,
                  J.m.cellV, J.m.cellCase, J.m.posOf, J.m.V, J.m.Nrm, J.m.capV, J.m.T, J.m.capT, J.m.late,
                  J.m.capLate, J.m.totals);
  }
}
__global__ void __launch_bounds__(kFinishThreads) k_mc_finish_b(const MCJob* __restrict__ jobs) {
  const MCJob& J = jobs[blockIdx.x];
  mc_finish_body(J.m.V, J.m.Nrm, J.m.T, J.m.late, J.m.capLate, J.m.totals, J.m.capV, J.m.capT, J.m.list);
}
__global__ void k_totals_b(const MCJob* __restrict__ jobs, int n, u32* out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  for (int k = 0; k < 4; ++k) out[4 * j + k] = jobs[j].m.totals[k];
}

hipError_t upload_tables_once() {
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
  for (int c = 0; c < 256; ++c)
    for (u32 t = 0; t < h.ntri[c]; ++t)
      for (int k = 0; k < 3; ++k) h.use[c][(kTriPacked[c] >> (12 * t + 4 * k)) & 0xf] |= (uint8_t)(1u << t);
  GM_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_mc), &h, sizeof(h)));
  GM_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_tri), kTriPacked, sizeof(kTriPacked)));
  return hipSuccess;
}
// Once per process (contexts on several host threads share the tables).
hipError_t upload_tables() {
  static std::once_flag once;
  static hipError_t status = hipSuccess;
  std::call_once(once, [] { status = upload_tables_once(); });
  return status;
}

}  // namespace

hipError_t mc_prepare() { return upload_tables(); }

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

size_t mc_scan_parts(int nblocks) { return (size_t)((nblocks + kChunk - 1) / kChunk); }
size_t mc_scan_part_bytes(int nblocks) { return mc_scan_parts(nblocks) * kPartStride * 4; }

// Workgroups of the kernels that walk the active-block list: enough to fill
// the chip several times over, never more than there are blocks. The emit
// kernels are long per block, so they get one workgroup per active block up
// to a larger cap.
int mc_list_grid(int nblocks, int cap) { return nblocks < cap ? nblocks : cap; }

hipError_t launch_mc(const u32* C, int R, const float* bmin3, float cell, const MCBuffers& m, hipStream_t s,
                     hipEvent_t* ev) {
  GM_CHECK(upload_tables());
  if (ev) GM_CHECK(hipEventRecord(ev[0], s));
  // active Morton blocks, straight from C
  const int nb = m.P / 8;
  if (nb > kMaxBlocksPerAxis) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mc_blocks_c, dim3(nb, nb), dim3(512), 0, s, C, R, nb, m.nblocks, m.blockV, m.blockT, m.list,
                     m.posOf, m.order, m.scanPart, m.totals);
  GM_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_mc_count, dim3(mc_list_grid(m.nblocks, 2048)), dim3(kCountThreads), 0, s, C, R, m.list,
                     m.blockV, m.blockT, m.cellV, m.cellCase, m.late, m.scanPart, m.capList, m.totals);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[1], s));
  // the chunk totals were summed by the count kernel
  int nparts = (int)mc_scan_parts(m.nblocks);
  hipLaunchKernelGGL(k_scan_down, dim3(nparts), dim3(256), 0, s, m.blockV, m.blockT, m.nblocks, m.scanPart, nparts,
                     m.totals, m.capV, m.capT, m.posOf, m.order, m.list, m.capList);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[2], s));
  return launch_mc_emit(C, R, bmin3, cell, m, s, ev ? ev + 2 : nullptr);
}

namespace {
size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }
struct McBlobLayout {
  size_t jobs, bricks, preB, preK, preC, up, lpre, totals, all;
  explicit McBlobLayout(int n) {
    jobs = 0;
    bricks = align16(jobs + (size_t)n * sizeof(MCJob));
    preB = align16(bricks + (size_t)n * sizeof(BrickJob));
    preK = align16(preB + (size_t)(n + 1) * 4);
    preC = align16(preK + (size_t)(n + 1) * 4);
    up = align16(preC + (size_t)(n + 1) * 4);  // bytes uploaded from the host blob
    lpre = up;
    totals = align16(lpre + (size_t)(n + 1) * 4);
    all = align16(totals + (size_t)n * 16);
  }
};
}  // namespace

size_t batch_mc_bytes(int n) { return McBlobLayout(n).all; }

hipError_t launch_mc_batched(const BatchMCJob* in, int n, void* hBlob, void* dBlob, u32* hTotals, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  GM_CHECK(upload_tables());
  const McBlobLayout L(n);
  char* h = static_cast<char*>(hBlob);
  char* d = static_cast<char*>(dBlob);
  MCJob* hj = reinterpret_cast<MCJob*>(h + L.jobs);
  BrickJob* hb = reinterpret_cast<BrickJob*>(h + L.bricks);
  u32* preB = reinterpret_cast<u32*>(h + L.preB);
  u32* preK = reinterpret_cast<u32*>(h + L.preK);
  u32* preC = reinterpret_cast<u32*>(h + L.preC);
  preB[0] = preK[0] = preC[0] = 0;
  long long blocksAll = 0;
  int wmax = 1;
  for (int j = 0; j < n; ++j) {
    const BatchMCJob& J = in[j];
    const int N = 2 * J.R;
    hj[j].C = J.C;
    hj[j].R = J.R;
    hj[j].geo = Geo{J.bmin[0], J.bmin[1], J.bmin[2], 0.5f * J.cell};
    hj[j].m = J.m;
    hb[j] = BrickJob{J.C, J.m.brick, J.m.list, J.m.late, N};
    preB[j + 1] = preB[j] + (u32)((N / 8) * (N / 8));
    preK[j + 1] = preK[j] + (u32)((J.m.nblocks + 255) / 256);
    preC[j + 1] = preC[j] + (u32)mc_scan_parts(J.m.nblocks);
    blocksAll += J.m.nblocks;
    wmax = std::max(wmax, N / 32);
  }
  GM_CHECK(hipMemcpyAsync(d, h, L.up, hipMemcpyHostToDevice, s));
  const MCJob* dj = reinterpret_cast<const MCJob*>(d + L.jobs);
  const u32* dPreK = reinterpret_cast<const u32*>(d + L.preK);
  const u32* dPreC = reinterpret_cast<const u32*>(d + L.preC);
  u32* lpre = reinterpret_cast<u32*>(d + L.lpre);
  u32* tot = reinterpret_cast<u32*>(d + L.totals);
  // active Morton blocks from brick summaries of the closed grids (their
  // list and late counters cleared there)
  GM_CHECK(launch_brick_summary_batched(reinterpret_cast<const BrickJob*>(d + L.bricks),
                                        reinterpret_cast<const u32*>(d + L.preB), n, preB[n], wmax, s));
  hipLaunchKernelGGL(k_mc_blocks_b, dim3(preK[n]), dim3(256), 0, s, dj, dPreK, n);
  GM_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_list_prefix, dim3(1), dim3(64), 0, s, dj, n, lpre);
  GM_CHECK(hipGetLastError());
  const unsigned listGrid = (unsigned)std::max<long long>(1, std::min<long long>(blocksAll, 16384));
  hipLaunchKernelGGL(k_mc_count_b, dim3(std::min(listGrid, 4096u)), dim3(kCountThreads), 0, s, dj, lpre, n);
  GM_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_scan_reduce_b, dim3(preC[n]), dim3(256), 0, s, dj, dPreC, n);
  GM_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_scan_down_b, dim3(preC[n]), dim3(256), 0, s, dj, dPreC, n);
  GM_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_mc_emit_b, dim3(listGrid), dim3(kMcThreads), 0, s, dj, lpre, n);
  GM_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_mc_finish_b, dim3(n), dim3(kFinishThreads), 0, s, dj);
  GM_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_totals_b, dim3((n + 255) / 256), dim3(256), 0, s, dj, n, tot);
  GM_CHECK(hipGetLastError());
  GM_CHECK(hipMemcpyAsync(hTotals, tot, (size_t)n * 16, hipMemcpyDeviceToHost, s));
  return hipSuccess;
}

hipError_t launch_mc_emit(const u32* C, int R, const float* bmin3, float cell, const MCBuffers& m, hipStream_t s,
                          hipEvent_t* ev) {
  GM_CHECK(upload_tables());
  Geo geo{bmin3[0], bmin3[1], bmin3[2], 0.5f * cell};
  hipLaunchKernelGGL(k_mc_emit, dim3(mc_list_grid(m.nblocks, 16384)), dim3(kMcThreads), 0, s, R, geo, m.list,
                     m.nblocks, m.blockV, m.blockT, m.cellV, m.cellCase, m.posOf, m.order, m.V, m.Nrm, m.capV, m.T,
                     m.capT, m.late, m.capLate, m.totals, m.capList);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[1], s));
  // the late triangles and normalization, when any, ran in the emit's last
  // workgroup: the stage after it is empty
  if (ev) GM_CHECK(hipEventRecord(ev[2], s));
  return hipSuccess;
}

}  // namespace gm
