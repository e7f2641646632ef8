// Ball dilation and erosion on bit-row grids (the GMorpho two-pass closing).
//
// Semantics (per fine voxel, restating the reference):
//  dilation (DilateBySphereMipmap, GMorpho.cu:3538-3596):
//    coarse  DF(c) = I(c) full, or c < dR and F(2c)           (2828-2928)
//    tight   TG(c) = c in [0,R-2]^3 and DF mixed on c+{0,1}^3 (Grid.cu:1572)
//    fine    D(c)  = TG(c) ? F(2c+k) per sub-voxel : (DF(c) ? 0xff : I(c))  (2930-3056)
//    F(p) = exists q in I with SqDist(p, [q,q+1]^3) < (r(p)+1)^2   (sphere)
//           or q in [p-r-1, p+r]^3                                  (cube, 1838/2603)
//  erosion (ErodeBySphereMipmap, GMorpho.cu:6627-6702):
//    K     = empty voxels of non-full cells in [1,R-2]^3 with an occupied
//            6-neighbour (Grid.cu:1230-1327)
//    coarse CL(c) = c < dR, D(c) != 0, I(c) != 0xff and Ec(2c+1)   (4349-4648)
//    base   UG(c) = c in [1,R-2]^3 with a 6-neighbour cell C'(c') != C'(c),
//           C'(c) = CL(c) ? 0 : D(c)                               (Grid.cu:1499)
//    fine   C(c)  = UG(c) ? ~Ef(2c+k) per sub-voxel : C'(c)        (5041-5374)
//    Ec(P) = exists q in K with SqDist(P, [q,q+1]^3) < (r(q)-1)^2
//    Ef(p) = exists q in K with SqDist(p+1/2, [q,q+1]^3) < (r(q)-1)^2
//
// Kernel design (MI355X): instead of one DFS per voxel over a max-mipmap, each
// predicate is an OR of x-runs: for every (dy, dz) of the structuring element
// the x-extent is an interval, so F = OR_{dy,dz} G_m(src)(y+dy, z+dz) with G_m
// the x-window dilation of a 32-bit row word. A 64-lane workgroup owns one
// row-word column x 64 output rows and streams the input planes along z:
//  * the chord table of the structuring element sits in LDS (uniform reads);
//  * per plane, the rows the tile reads are prefetched into registers one
//    plane ahead, and G_0..G_mmax of every row is built in LDS;
//  * every lane ORs the chords of each dz class (classes b and -b-1 share
//    their chords) into a power-of-two ring of output-plane accumulators;
//  * a plane is handed to the pass's epilogue as soon as it is complete.
// The coarse passes only need even (dilation) or odd (erosion) rows and
// planes, so their lanes own every other row. Tiles whose source region is
// empty (per-brick occupancy summaries) or whose fine pass has no tagged
// cell skip the streaming and run the epilogue on F = 0.
// Radii that vary per voxel are handled by one pass per distinct radius
// found in the tile: gather radius (dilation) masks the output voxels,
// scatter radius (erosion) masks the source voxels.
#include <algorithm>
#include <vector>

#include "gm_morph_dev.h"

namespace gm {

namespace {

// Tile selection of each pass (one wave per tile). A workgroup of 8 waves
// covers half a slab of 16 word tiles at N = 512, so it reserves its queue
// slots with one or two device atomics: those atomics, from every workgroup
// on one cache line, serialize across the XCDs. With 4 waves per workgroup
// (twice the atomics again) they bounded the selection; with 16, the 2048
// tiles of N = 512 fill only 128 workgroups, half the CUs.
constexpr int kSelectWaves = 8;
template <int PASS>
__global__ void __launch_bounds__(64 * kSelectWaves) k_select(MorphArgs a) {
  select_tiles<PASS>(a);
}

// Runtime-table kernels (any radius field, sphere or cube).
__global__ void __launch_bounds__(64 * kMaxWaves) k_dil_coarse(MorphArgs a) { run_pass<0, 0>(a); }
__global__ void __launch_bounds__(64 * kMaxWaves) k_dil_fine(MorphArgs a) { run_pass<1, 0>(a); }
__global__ void __launch_bounds__(64 * kMaxWaves) k_ero_coarse(MorphArgs a) { run_pass<2, 0>(a); }
__global__ void __launch_bounds__(64 * kMaxWaves) k_ero_fine(MorphArgs a) { run_pass<3, 0>(a); }

// (wx, cy, cz) of cell-row word g of an R x R x W flag grid. R*R*W <= 2^26,
// so 32-bit divisions do (64-bit ones are expanded into long VALU
// sequences and made these kernels VALU-bound).
__device__ __forceinline__ void cell_word_at(u32 g, u32 R, u32 W, int& wx, int& cy, int& cz) {
  const u32 t = g / W;
  const u32 z = t / R;
  wx = (int)(g - t * W);
  cy = (int)(t - z * R);
  cz = (int)z;
}

// Tight contour of DF -> TG flags (one thread per cell-row word).
__device__ __forceinline__ void tight_body(const MorphArgs& a, long long gid) {
  long long total = (long long)a.R * a.R * a.W;
  if (gid >= total) return;
  int wx, cy, cz;
  cell_word_at((u32)gid, (u32)a.R, (u32)a.W, wx, cy, cz);
  // flag words (cy + (d & 1), cz + (d >> 1)) at wx and wx + 1: the cell
  // rows at x and y distance 1 are gids +1 and +W, consecutive over the
  // lanes of a workgroup (a multiple of 64 threads), so those inside the
  // wave come from the lanes that hold them
  const int lane = (int)(threadIdx.x & 63u);
  u32 v[4], nx[4];
  v[0] = flag_word(a.DF, a.R, a.W, cy, cz, wx);
  v[2] = flag_word(a.DF, a.R, a.W, cy, cz + 1, wx);
  v[1] = __shfl(v[0], min(lane + a.W, 63));
  v[3] = __shfl(v[2], min(lane + a.W, 63));
  if (cy + 1 >= a.R) {
    v[1] = v[3] = 0u;
  } else if (lane + a.W > 63) {
    v[1] = flag_word(a.DF, a.R, a.W, cy + 1, cz, wx);
    v[3] = flag_word(a.DF, a.R, a.W, cy + 1, cz + 1, wx);
  }
#pragma unroll
  for (int d = 0; d < 4; ++d) nx[d] = __shfl(v[d], min(lane + 1, 63));
  if (wx + 1 >= a.W) {
#pragma unroll
    for (int d = 0; d < 4; ++d) nx[d] = 0u;
  } else if (lane == 63) {
#pragma unroll
    for (int d = 0; d < 4; ++d) nx[d] = flag_word(a.DF, a.R, a.W, cy + (d & 1), cz + (d >> 1), wx + 1);
  }
  u32 all = 0xffffu, any = 0u;
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    u32 sh = ((v[d] >> 1) | (nx[d] << 15)) & 0xffffu;
    all &= v[d] & sh;
    any |= v[d] | sh;
  }
  u32 valid = cell_range_mask(wx, 0, a.R - 1);
  if (cy > a.R - 2 || cz > a.R - 2) valid = 0;
  const u32 tg = any & ~all & valid;
  a.outFlags[gid] = (u16)tg;
  // D outside the tight contour: DF ? 1 : I; on it: I (the fine pass ORs in
  // F, which contains I)
  const u32 df = expand2(v[0] & ~tg);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const size_t o = ((size_t)(2 * cz + (s >> 1)) * a.N + 2 * cy + (s & 1)) * a.W + wx;
    a.outRows[o] = df | a.I[o];
  }
}
__global__ void k_tight(MorphArgs a) { tight_body(a, (long long)blockIdx.x * blockDim.x + threadIdx.x); }

// Brick summary of a grid (brick = 1 word x 8 rows x 8 planes; bit0 = any
// voxel set, bit1 = all set). Workgroup = one 8x8 (row, plane) brick column
// over every word: thread (w, row) walks the 8 planes, so each plane step is
// one contiguous read of 8 rows; the 8 rows of a word meet in LDS.
// clear (optional): a counter the next kernel starts from, zeroed here so
// that no separate memset is needed.
// Workgroups may have more than 8*W threads (batched launches size them for
// the widest job); the extra threads only meet the barriers.
// ifl (optional): the cell flags of the full cells of G (all 8 fine voxels
// set), for the 4x4 cells of the brick column (u16 per cell row word).
__device__ __forceinline__ void brick_summary_body(const u32* __restrict__ G, int N, int W, uint8_t* __restrict__ out,
                                                   int by, int bz, u16* __restrict__ ifl = nullptr) {
  __shared__ u32 anyW[64], allW[64];
  __shared__ u32 pairW[8][4][64];  // row yl, plane pair k, word: AND of planes 2k, 2k+1
  const int NB = N / 8;
  const bool act = threadIdx.x < 8u * (unsigned)W;
  const int w = threadIdx.x % W, yl = threadIdx.x / W;
  if (threadIdx.x < W) {
    anyW[threadIdx.x] = 0u;
    allW[threadIdx.x] = 1u;
  }
  __syncthreads();
  if (act) {
    const int y = 8 * by + yl;
    u32 any = 0u, all = 0xffffffffu, pr[4];
    const u32* g = G + ((size_t)(8 * bz) * N + y) * W + w;
#pragma unroll
    for (int zl = 0; zl < 8; ++zl) {
      u32 v = g[(size_t)zl * N * W];
      any |= v;
      all &= v;
      pr[zl >> 1] = (zl & 1) ? (pr[zl >> 1] & v) : v;
    }
    if (any) atomicOr(&anyW[w], 1u);
    if (all != 0xffffffffu) atomicAnd(&allW[w], 0u);
    if (ifl)
#pragma unroll
      for (int k = 0; k < 4; ++k) pairW[yl][k][w] = pr[k];
  }
  __syncthreads();
  if (threadIdx.x < W)
    out[((size_t)bz * NB + by) * W + threadIdx.x] = (uint8_t)(anyW[threadIdx.x] | (allW[threadIdx.x] << 1));
  if (ifl && act && !(yl & 1)) {
    // cell (4by + yl/2, 4bz + k): rows yl, yl+1 of planes 2k, 2k+1
    const int R = N / 2;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const u32 f = pairW[yl][k][w] & pairW[yl + 1][k][w];
      ifl[((size_t)(4 * bz + k) * R + (4 * by + (yl >> 1))) * W + w] = (u16)gather2(f & (f >> 1));
    }
  }
}
__global__ void __launch_bounds__(512) k_brick_summary(const u32* __restrict__ G, int N, int W,
                                                        uint8_t* __restrict__ out, u32* clear,
                                                        u16* __restrict__ ifl = nullptr) {
  if (clear && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *clear = 0u;
  brick_summary_body(G, N, W, out, blockIdx.x, blockIdx.y, ifl);
}

// Outer contour K of D (Grid.cu:1230-1327) and its brick summary: voxels
// outside D, 6-adjacent to D, in a cell that is not full, away from the data
// border. Same workgroup mapping as k_brick_summary.
// kbl (optional): the bricks holding K are appended to it ([0] count).
__device__ __forceinline__ void contour_K_rows(const MorphArgs& a, int wx, int yl, int by, int bz, u32* anyW);
__device__ __forceinline__ void contour_K_body(const MorphArgs& a, uint8_t* __restrict__ brickK, u32* kbl, int by,
                                               int bz) {
  __shared__ u32 anyW[64];
  const int N = a.N, W = a.W, NB = N / 8;
  const int wx = threadIdx.x % W, yl = threadIdx.x / W;
  if (threadIdx.x < W) anyW[threadIdx.x] = 0u;
  __syncthreads();
  if (threadIdx.x < 8u * (unsigned)W) contour_K_rows(a, wx, yl, by, bz, anyW);
  __syncthreads();
  if (threadIdx.x < 64) {  // W <= 64: the first wave holds every word of the row
    const bool in = (int)threadIdx.x < W;
    const bool has = in && anyW[threadIdx.x] != 0u;
    const size_t b = ((size_t)bz * NB + by) * W + threadIdx.x;
    if (in) brickK[b] = (uint8_t)(has ? 1u : 0u);
    if (kbl) {
      // one reservation for the row's bricks holding K
      const unsigned long long m = __ballot(has);
      if (m) {
        const int lane = (int)threadIdx.x;
        u32 base = 0u;
        if (lane == 0) base = atomicAdd(kbl, (u32)__popcll(m));
        base = __shfl(base, 0, 64);
        if (has) kbl[1 + base + (u32)__popcll(m & ((1ull << lane) - 1ull))] = (u32)b;
      }
    }
  }
}
__global__ void __launch_bounds__(512) k_contour_K(MorphArgs a, uint8_t* __restrict__ brickK, u32* kbl) {
  contour_K_body(a, brickK, kbl, blockIdx.x, blockIdx.y);
}

// The rows of thread (wx, yl) of brick row (by, bz) for contour_K_body.
__device__ __forceinline__ void contour_K_rows(const MorphArgs& a, int wx, int yl, int by, int bz, u32* anyW) {
  const int N = a.N, W = a.W;
  const u32* D = a.D;
  const int y = 8 * by + yl;
  const int cy = y >> 1;
  const bool yin = cy >= 1 && cy <= a.R - 2;
  const u32 xin = expand2(cell_range_mask(wx, 1, a.R - 1));
  u32 any = 0u;
  u32 zm = row_word(D, N, W, y, 8 * bz - 1, wx);
  u32 c = row_word(D, N, W, y, 8 * bz, wx);
  // Thread t holds (wx, yl) = (t % W, t / W): the words at x distance 1 are
  // threads t -+ 1 and those at y distance 1 threads t -+ W; inside the wave
  // they come from the lanes that hold them, the others are loaded.
  const int lane = (int)(threadIdx.x & 63u);
  const int sl = max(lane - 1, 0), sr = min(lane + 1, 63), sym = max(lane - W, 0), syp = min(lane + W, 63);
  const bool ymIn = yl >= 1 && lane >= W, ypIn = yl <= 6 && lane + W <= 63;
#pragma unroll 2
  for (int zl = 0; zl < 8; ++zl) {
    const int z = 8 * bz + zl;
    const u32 zp = row_word(D, N, W, y, z + 1, wx);
    const u32 diag = row_word(D, N, W, y ^ 1, z ^ 1, wx);
    u32 l = __shfl(c, sl), r = __shfl(c, sr), ym = __shfl(c, sym), yp = __shfl(c, syp);
    if (wx == 0)
      l = 0u;
    else if (lane == 0)
      l = row_word(D, N, W, y, z, wx - 1);
    if (wx == W - 1)
      r = 0u;
    else if (lane == 63)
      r = row_word(D, N, W, y, z, wx + 1);
    if (!ymIn) ym = row_word(D, N, W, y - 1, z, wx);
    if (!ypIn) yp = row_word(D, N, W, y + 1, z, wx);
    const u32 xn = (c << 1) | (l >> 31) | (c >> 1) | (r << 31);
    const u32 n6 = xn | ym | yp | zm | zp;
    // the 2x2 fine rows of this voxel's cell row
    const u32 ysib = (y & 1) ? ym : yp;
    const u32 zsib = (z & 1) ? zm : zp;
    const u32 f = c & ysib & zsib & diag;
    if (!(y & 1) && !(z & 1)) {
      // the cells with a voxel of D (read by the coarse erosion)
      const u32 o = c | ysib | zsib | diag;
      a.DAL[((size_t)(z >> 1) * a.R + (y >> 1)) * W + wx] = (u16)gather2(o | (o >> 1));
    }
    const u32 cfull = expand2(gather2(f & (f >> 1)));
    const int cz = z >> 1;
    const u32 inner = (yin && cz >= 1 && cz <= a.R - 2) ? xin : 0u;
    const u32 k = ~c & n6 & ~cfull & inner;
    a.outRows[((size_t)z * N + y) * W + wx] = k;
    any |= k;
    zm = c;
    c = zp;
  }
  if (any) atomicOr(&anyW[wx], 1u);
}

// Radius values of the 32 voxels of row word (y, z, w) of a lattice field
// into the bit planes p and the value set: T is int when s^3 times the
// largest lattice magnitude fits 31 bits (rf.narrow), else long long.
template <class T>
__device__ __forceinline__ void lattice_word(const RadiusField& rf, int y, int z, int w, u32 p[5], u32& set) {
  const int ls = rf.lshift, L = rf.L;
  const T s = (T)rf.s;
  const int iy = y >> ls, iz = z >> ls;
  const T fy = (T)(y - (iy << ls)), fz = (T)(z - (iz << ls));
  const int y0 = min(iy, L - 1), y1 = min(iy + 1, L - 1), z0 = min(iz, L - 1), z1 = min(iz + 1, L - 1);
  const T w00 = (s - fy) * (s - fz), w10 = fy * (s - fz), w01 = (s - fy) * fz, w11 = fy * fz;
  const int32_t* l00 = rf.lat + (size_t)L * (y0 + (size_t)L * z0);
  const int32_t* l10 = rf.lat + (size_t)L * (y1 + (size_t)L * z0);
  const int32_t* l01 = rf.lat + (size_t)L * (y0 + (size_t)L * z1);
  const int32_t* l11 = rf.lat + (size_t)L * (y1 + (size_t)L * z1);
  auto colsum = [&](int ix) {
    const int xx = min(ix, L - 1);
    return w00 * (T)l00[xx] + w10 * (T)l10[xx] + w01 * (T)l01[xx] + w11 * (T)l11[xx];
  };
  const int sh = 3 * ls + 8;
  int cur = -1;
  T acc = 0, dA = 0;
#pragma unroll
  for (int b = 0; b < 32; ++b) {
    const int x = 32 * w + b;
    const int ix = x >> ls;
    if (ix != cur) {  // same b for every lane: 32w is a multiple of s or s > 32
      const T A0 = colsum(ix), A1 = colsum(ix + 1);
      const T fx = (T)(x - (ix << ls));
      acc = (s - fx) * A0 + fx * A1;  // s A0 + fx (A1 - A0), then + (A1 - A0) per voxel
      dA = A1 - A0;
      cur = ix;
    } else {
      acc += dA;
    }
    int v = (int)(acc >> sh);
    v = min(max(v, 2), 31);
    set |= 1u << v;
#pragma unroll
    for (int k = 0; k < 5; ++k) p[k] |= (u32)((v >> k) & 1) << b;
  }
}

// Radius value v in [2,31] of every voxel as 5 bit planes (plane k at
// rp + k*N*N*W holds bit k, brick-major layout) and the set of values of
// each brick (vset, bit v). Lattice fields follow radius_at: trilinear in
// fixed point, floor, clamped; the interpolant along x is
// (s - fx) A(ix) + fx A(ix+1), A the y/z-weighted column sums, and the
// division by s^3 * 256 (a power of two) is an arithmetic shift.
__global__ void __launch_bounds__(256) k_radius_planes(RadiusField rf, int N, int W, u32* __restrict__ rp,
                                                       u32* __restrict__ vset, int bw0, int bnw, int by0, int bnby,
                                                       int bz0, int bnbz) {
  const int NB = N / 8;
  // one wave per brick of the box (bricks bw0.., by0.., bz0..), lane = row +
  // 8 * plane: the 5 plane stores of a wave are each one contiguous 256-byte
  // brick
  const size_t i = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= (size_t)bnw * bnby * bnbz) return;
  const int lane = lane_id();
  const int w = bw0 + (int)(i % bnw);
  const size_t t = i / bnw;
  const int by = by0 + (int)(t % bnby), bz = bz0 + (int)(t / bnby);
  const size_t b = ((size_t)bz * NB + by) * W + w;
  const int y = 8 * by + (lane & 7), z = 8 * bz + (lane >> 3);
  const size_t plane = (size_t)N * N * W;
  u32 set = 0u;
  u32 p[5] = {0u, 0u, 0u, 0u, 0u};
  if (rf.mode == 1) {
    const int8_t* row = rf.grid + ((size_t)z * N + y) * (size_t)N + 32 * w;
    const int4 q0 = reinterpret_cast<const int4*>(row)[0];
    const int4 q1 = reinterpret_cast<const int4*>(row)[1];
    const int qw[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      int v = (int)(int8_t)((u32)qw[k >> 2] >> (8 * (k & 3)));
      v = v < 2 ? 2 : (v > 31 ? 31 : v);
      set |= 1u << v;
#pragma unroll
      for (int j = 0; j < 5; ++j) p[j] |= (u32)((v >> j) & 1) << k;
    }
  } else if (rf.narrow) {
    lattice_word<int>(rf, y, z, w, p, set);
  } else {
    lattice_word<long long>(rf, y, z, w, p, set);
  }
#pragma unroll
  for (int j = 0; j < 5; ++j) rp[j * plane + b * 64 + lane] = p[j];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) set |= (u32)__shfl_xor((int)set, off);
  if (lane == 0) vset[b] = set;
}

// DF = I-full for every cell word (the per-value coarse dilation ORs its
// launches into it): one thread per cell-row word.
__global__ void k_ifull(const u32* __restrict__ I, int R, int N, int W, u16* __restrict__ DF) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)R * R * W) return;
  int wx, cy, cz;
  cell_word_at((u32)gid, (u32)R, (u32)W, wx, cy, cz);
  const u32* r = I + ((size_t)(2 * cz) * N + 2 * cy) * W + wx;
  const size_t pz = (size_t)N * W;
  const u32 f = r[0] & r[W] & r[pz] & r[pz + W];
  DF[gid] = (u16)gather2(f & (f >> 1));
}

// Clears the masked-source grid of the per-value erosion and the K brick list.
__global__ void k_kv_clear(uint4* __restrict__ kv, size_t n16, u32* kbl) {
  if (blockIdx.x == 0 && threadIdx.x == 0) kbl[0] = 0u;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    kv[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Source of the erosion launch of radius value v: kv = K & (r == v) on the
// bricks of K whose value set holds v, or held vprev (the previous launch's
// value, whose bits are cleared here). Every other word of kv is zero. One
// wave per listed brick (lane = row + 8 * plane); a grid of about one wave
// per brick keeps the loads of all bricks in flight at once.
__global__ void __launch_bounds__(256) k_kv(const u32* __restrict__ K, const u32* __restrict__ rp,
                                            const u32* __restrict__ vset, const u32* __restrict__ kbl,
                                            u32* __restrict__ kv, int N, int W, int v, int vprev) {
  const int NB = N / 8;
  const int lane = lane_id();
  const u32 n = kbl[0];
  const u32 want = (1u << v) | (vprev >= 0 ? 1u << vprev : 0u);
  for (u32 i = blockIdx.x * 4u + (threadIdx.x >> 6); i < n; i += gridDim.x * 4u) {
    const u32 b = kbl[1 + i];
    if (!(vset[b] & want)) continue;
    const int w = (int)(b % (u32)W), t = (int)(b / (u32)W);
    const int y = 8 * (t % NB) + (lane & 7), z = 8 * (t / NB) + (lane >> 3);
    const size_t o = ((size_t)z * N + y) * W + w;
    kv[o] = K[o] & radius_eq(rp, N, W, y, z, w, v);
  }
}

// C' rows of a cell row (4 fine rows), masked by CL.
__device__ __forceinline__ void cprime_rows(const MorphArgs& a, int cy, int cz, int w, u32 r[4]) {
  u32 cl = expand2(flag_word(a.CL, a.R, a.W, cy, cz, w));
#pragma unroll
  for (int s = 0; s < 4; ++s) r[s] = row_word(a.D, a.N, a.W, 2 * cy + (s & 1), 2 * cz + (s >> 1), w) & ~cl;
}

// Base (6-connected) contour of C' -> UG flags, one thread per cell-row word.
__device__ __forceinline__ void base_contour_body(const MorphArgs& a, long long gid) {
  long long total = (long long)a.R * a.R * a.W;
  if (gid >= total) return;
  int wx, cy, cz;
  cell_word_at((u32)gid, (u32)a.R, (u32)a.W, wx, cy, cz);
  u32 valid = cell_range_mask(wx, 1, a.R - 1);
  if (cy < 1 || cy > a.R - 2 || cz < 1 || cz > a.R - 2) valid = 0;
  u32 c[4], l[4], r[4], yp[4], ym[4], o[4];
  cprime_rows(a, cy, cz, wx, c);
  // The cell rows at x and y distance 1 are gids +-1 and +-W. Gids are
  // consecutive over the lanes of a workgroup (a multiple of 64 threads),
  // so a neighbour inside the wave comes from the lane that holds it.
  const int lane = (int)(threadIdx.x & 63u);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    l[s] = __shfl(c[s], max(lane - 1, 0));
    r[s] = __shfl(c[s], min(lane + 1, 63));
    yp[s] = __shfl(c[s], min(lane + a.W, 63));
    ym[s] = __shfl(c[s], max(lane - a.W, 0));
  }
  // C outside the base contour is C'; on it all ones (the fine pass clears Ef)
  auto put = [&](u32 ug) {
    a.outFlags[gid] = (u16)ug;
    const u32 ones = expand2(ug);
#pragma unroll
    for (int s = 0; s < 4; ++s)
      a.outRows[((size_t)(2 * cz + (s >> 1)) * a.N + 2 * cy + (s & 1)) * a.W + wx] = c[s] | ones;
  };
  if (!valid) {
    put(0u);
    return;
  }
  // words outside the grid are empty; valid cells have 1 <= cy <= R-2
  if (wx == 0) {
#pragma unroll
    for (int s = 0; s < 4; ++s) l[s] = 0u;
  } else if (lane == 0) {
    cprime_rows(a, cy, cz, wx - 1, l);
  }
  if (wx == a.W - 1) {
#pragma unroll
    for (int s = 0; s < 4; ++s) r[s] = 0u;
  } else if (lane == 63) {
    cprime_rows(a, cy, cz, wx + 1, r);
  }
  {
    // the wave's first row loads y - 1 and its last row y + 1 (W = 64: one
    // row, both loads); a lane that needs one of them issues one load
    const bool needP = lane + a.W > 63, needM = lane < a.W;
    if (needP && needM) {
      cprime_rows(a, cy + 1, cz, wx, yp);
      cprime_rows(a, cy - 1, cz, wx, ym);
    } else if (needP || needM) {
      u32 t[4];
      cprime_rows(a, needP ? cy + 1 : cy - 1, cz, wx, t);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        yp[s] = needP ? t[s] : yp[s];
        ym[s] = needP ? ym[s] : t[s];
      }
    }
  }
  u32 diff = 0u;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    u32 xp = (c[s] >> 2) | (r[s] << 30);  // cell j+1 aligned on cell j
    u32 xm = (c[s] << 2) | (l[s] >> 30);  // cell j-1
    diff |= (c[s] ^ xp) | (c[s] ^ xm) | (c[s] ^ yp[s]) | (c[s] ^ ym[s]);
  }
#pragma unroll
  for (int q = -1; q <= 1; q += 2) {
    cprime_rows(a, cy, cz + q, wx, o);
#pragma unroll
    for (int s = 0; s < 4; ++s) diff |= c[s] ^ o[s];
  }
  put(gather2(diff | (diff >> 1)) & valid);
}
__global__ void k_base_contour(MorphArgs a) {
  base_contour_body(a, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- Batched forms (many jobs per launch): the job of a workgroup or wave
// comes from the prefix array of the per-job counts (batch_job).
__global__ void __launch_bounds__(512) k_brick_summary_b(const MorphArgs* __restrict__ tab,
                                                          const u32* __restrict__ pre, int n) {
  const int j = batch_job(pre, n, blockIdx.x);
  const MorphArgs& a = tab[j];
  const int NB = a.N / 8, l = (int)(blockIdx.x - pre[j]);
  brick_summary_body(a.I, a.N, a.W, const_cast<uint8_t*>(a.brickI), l % NB, l / NB, a.IFL);
}
__global__ void __launch_bounds__(512) k_brick_summary_gb(const BrickJob* __restrict__ jobs,
                                                           const u32* __restrict__ pre, int n) {
  const int j = batch_job(pre, n, blockIdx.x);
  const BrickJob& J = jobs[j];
  const int NB = J.N / 8, l = (int)(blockIdx.x - pre[j]);
  if (l == 0 && threadIdx.x == 0) {
    if (J.clear0) *J.clear0 = 0u;
    if (J.clear1) *J.clear1 = 0u;
  }
  brick_summary_body(J.G, J.N, J.N / 32, J.out, l % NB, l / NB);
}
__global__ void __launch_bounds__(512) k_contour_K_b(const MorphArgs* __restrict__ tab, const u32* __restrict__ pre,
                                                      int n) {
  const int j = batch_job(pre, n, blockIdx.x);
  const MorphArgs& a = tab[j];
  const int NB = a.N / 8, l = (int)(blockIdx.x - pre[j]);
  contour_K_body(a, const_cast<uint8_t*>(a.brickK), nullptr, l % NB, l / NB);
}
__global__ void __launch_bounds__(256) k_tight_b(const MorphArgs* __restrict__ tab, const u32* __restrict__ pre,
                                                  int n) {
  const int j = batch_job(pre, n, blockIdx.x);
  tight_body(tab[j], (long long)(blockIdx.x - pre[j]) * blockDim.x + threadIdx.x);
}
__global__ void __launch_bounds__(256) k_base_contour_b(const MorphArgs* __restrict__ tab,
                                                         const u32* __restrict__ pre, int n) {
  const int j = batch_job(pre, n, blockIdx.x);
  base_contour_body(tab[j], (long long)(blockIdx.x - pre[j]) * blockDim.x + threadIdx.x);
}
template <int PASS>
__global__ void __launch_bounds__(256) k_select_b(const MorphArgs* __restrict__ tab, const u32* __restrict__ pre, int n,
                                                   const int* __restrict__ rad, u32* queue, int qsel, int listBase,
                                                   int cap) {
  select_tiles_batched<PASS>(tab, pre, n, rad, queue, qsel, listBase, cap);
}

// LDS bytes of a streaming launch with `waves` waves per workgroup.
size_t lds_bytes(const MorphArgs& a, bool var, int waves) {
  const int nplanes = a.ystep == 2 ? a.Zt / 2 : a.Zt;
  size_t words = (size_t)a.tabWords + (size_t)ct_slice_words(a.maxm, a.maxoff, a.ystep) * waves +
                 (size_t)nplanes * kTileRows;
  if (var) words += (size_t)nplanes * kTileRows;
  return words * 4;
}
// Waves per workgroup: as many as fit a 64 KiB LDS budget, at most 8 (or the
// launch shape's cap).
int stream_waves(const MorphArgs& a, bool var) {
  int w = a.waveCap > 0 ? std::min(a.waveCap, kMaxWaves) : kMaxWaves;
  while (w > 1 && lds_bytes(a, var, w) > 64 * 1024) --w;
  return w;
}

MorphArgs base_args(const MorphDims& d, const MorphBuffers& b, const RadiusField& rf, const SETableSet& tabs) {
  MorphArgs a{};
  a.R = d.R;
  a.N = d.N;
  a.W = d.W;
  a.dR0 = d.dR[0];
  a.dR1 = d.dR[1];
  a.dR2 = d.dR[2];
  a.Zt = b.tileZ == kTileZMin ? kTileZMin : kTileZ;
  a.waveCap = b.waveCap;
  a.rf = rf;
  a.tables = b.tables;
  a.tabs = tabs;
  a.I = b.I;
  a.D = b.D;
  a.K = b.K;
  a.DF = b.DF;
  a.TG = b.TG;
  a.CL = b.CL;
  a.UG = b.UG;
  a.IFL = b.IFL;
  a.DAL = b.DAL;
  a.brickI = b.brickI;
  a.brickK = b.brickK;
  a.queue = b.queue;
  a.queueCap = b.queueCap;
  a.vqBase = (int)var_queue_base(d.R);
  a.vqCap = (int)var_queue_cap(d.R);
  return a;
}

dim3 stream_grid(const MorphArgs& a) {
  const int rows = a.ystep * kTileRows;
  return dim3((unsigned)(a.W * ((a.N + rows - 1) / rows) * ((a.N + a.Zt - 1) / a.Zt)));
}

const CtKernels& ct_kernels() {
  static const CtKernels k = [] {
    CtKernels t{};
    ct_register_part0(t);
    ct_register_part1(t);
    ct_register_part2(t);
    ct_register_part3(t);
    ct_register_part4(t);
    ct_register_part5(t);
    ct_register_part6(t);
    ct_register_part7(t);
    return t;
  }();
  return k;
}

int device_cus() {
  static int cus = [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1)
      n = 256;
    return n;
  }();
  return cus;
}

// Launches streaming pass `pass` (0..3): tile selection (skipped tiles are
// finished there, active ones queued), then persistent workgroups of the
// kernel compiled for chord-table radius rho when there is one (fixedOk:
// uniform radius, sphere), else of the runtime-table kernel `generic`.
// evStream (optional) is recorded between the tile selection and the stream
// kernel, so that the stream kernel can be timed alone.
hipError_t launch_pass(int pass, const void* generic, MorphArgs a, int* qsel, bool var, bool fixedOk, int rho,
                       hipStream_t s, hipEvent_t evStream, bool box = false) {
  a.qsel = *qsel;
  *qsel = 1 - *qsel;
  const void* fn = generic;
  const void* compiled = nullptr;
  if (fixedOk && rho >= 1 && rho <= kCtMaxRho)
    compiled = box ? (pass < 2 ? ct_kernels().fnBox[pass][rho] : nullptr) : ct_kernels().fn[pass][rho];
  if (compiled) {
    fn = compiled;
    a.tabWords = 0;  // the chords are compiled in: no table in LDS
  }
  const void* sel[4] = {reinterpret_cast<const void*>(&k_select<0>), reinterpret_cast<const void*>(&k_select<1>),
                        reinterpret_cast<const void*>(&k_select<2>), reinterpret_cast<const void*>(&k_select<3>)};
  const unsigned tiles = stream_grid(a).x;
  a.rvar = 0;
  a.qcOff = 16 * a.qsel;
  a.qlOff = kQueueList;
  a.qCap = a.queueCap;
  a.qBack = 1;
  void* args[] = {&a};
  GM_CHECK(hipLaunchKernel(sel[pass], dim3((tiles + kSelectWaves - 1) / kSelectWaves), dim3(64 * kSelectWaves),
                            args, 0, s));
  if (evStream) GM_CHECK(hipEventRecord(evStream, s));
  const int nw = stream_waves(a, var);
  const size_t lds = lds_bytes(a, var, nw);
  // one workgroup per tile at most: workgroup i takes entries i/8, i/8 + G/8,
  // ... of queue i % 8 (G, the grid, is a multiple of 8)
  unsigned grid = (tiles + 7u) & ~7u;
  GM_CHECK(hipLaunchKernel(fn, dim3(grid), dim3(64 * nw), args, lds, s));
  return hipGetLastError();
}

// Streaming pass `pass` with per-voxel radii (sphere): one tile selection
// queues every tile under each radius value that can reach it, then one
// launch per value v in [vlo, vhi] of the kernel compiled for the chord-table
// radius of v (dilation v + 1, erosion v - 1). Erosion launches read the
// source kv, rebuilt for each value first (*vprev: the value kv holds).
hipError_t launch_pass_var(int pass, MorphArgs a, const MorphBuffers& b, int vlo, int vhi, int* vprev,
                           hipStream_t s, hipEvent_t evStream) {
  a.qsel = *b.qsel;
  *b.qsel = 1 - *b.qsel;
  a.rvar = 1;
  a.rp = b.rp;
  a.rvset = b.rvset;
  const void* sel[4] = {reinterpret_cast<const void*>(&k_select<0>), reinterpret_cast<const void*>(&k_select<1>),
                        reinterpret_cast<const void*>(&k_select<2>), reinterpret_cast<const void*>(&k_select<3>)};
  const unsigned tiles = stream_grid(a).x;
  {
    void* args[] = {&a};
    GM_CHECK(hipLaunchKernel(sel[pass], dim3((tiles + kSelectWaves - 1) / kSelectWaves), dim3(64 * kSelectWaves),
                            args, 0, s));
  }
  if (evStream) GM_CHECK(hipEventRecord(evStream, s));
  const bool ero = pass >= 2;
  // k_kv: 4 waves per workgroup, about one wave per brick of K (bounded)
  const unsigned bricks = (unsigned)((a.N / 8) * (a.N / 8) * a.W);
  const unsigned kvGrid = std::max(64u, std::min(bricks / 32u, 16384u));
  // a value reaches few tiles: fewer persistent workgroups than tiles
  const unsigned grid = std::min((tiles + 7u) & ~7u, 2048u);
  for (int v = vlo; v <= vhi; ++v) {
    const int rho = ero ? v - 1 : v + 1;
    if (rho < 1 || rho > kCtMaxRho || !ct_kernels().fn[pass][rho]) return hipErrorInvalidValue;
    MorphArgs c = a;
    c.vcur = v;
    c.qcOff = kVarCounters + 256 * a.qsel + 8 * v;
    c.qlOff = a.vqBase + 8 * v * a.vqCap;
    c.qCap = a.vqCap;
    c.tabWords = 0;
    c.maxm = ero ? rho : rho - 1;
    c.maxoff = rho;
    if (ero) {
      hipLaunchKernelGGL(k_kv, dim3(kvGrid), dim3(256), 0, s, b.K, b.rp, b.rvset, b.kbl, b.kv, a.N, a.W, v, *vprev);
      GM_CHECK(hipGetLastError());
      *vprev = v;
      c.K = b.kv;
    }
    const int nw = stream_waves(c, false);
    void* args[] = {&c};
    GM_CHECK(hipLaunchKernel(ct_kernels().fn[pass][rho], dim3(grid), dim3(64 * nw), args, lds_bytes(c, false, nw),
                             s));
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_radius_planes(const RadiusField& rf, int N, u32* rp, u32* rvset, hipStream_t s,
                                const int* brickBox) {
  if (rf.mode != 1 && rf.mode != 2) return hipErrorInvalidValue;
  const int NB = N / 8, W = N / 32;
  int bx[6] = {0, 0, 0, W, NB, NB};
  if (brickBox)
    for (int k = 0; k < 6; ++k) bx[k] = brickBox[k];
  const size_t bricks = (size_t)bx[3] * bx[4] * bx[5];
  if (bricks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_radius_planes, dim3((unsigned)((bricks + 3) / 4)), dim3(256), 0, s, rf, N, W, rp, rvset, bx[0],
                     bx[3], bx[1], bx[4], bx[2], bx[5]);
  return hipGetLastError();
}

hipError_t launch_brick_summary_batched(const BrickJob* dJobs, const u32* dPre, int n, u32 blocks, int wmax,
                                        hipStream_t s) {
  if (n <= 0 || blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_brick_summary_gb, dim3(blocks), dim3(8 * wmax), 0, s, dJobs, dPre, n);
  return hipGetLastError();
}

hipError_t launch_brick_summary(const u32* G, int N, uint8_t* out, u32* clear, hipStream_t s) {
  hipLaunchKernelGGL(k_brick_summary, dim3(N / 8, N / 8), dim3(8 * (N / 32)), 0, s, G, N, N / 32, out, clear);
  return hipGetLastError();
}

// rho_lo/rho_hi: range of the per-voxel radius r over the field; tabWords:
// longest SE table of the launch. A per-voxel radius field with the sphere
// (b.rp set) runs one launch per radius value; the cube, with any field, runs
// the runtime-table kernels.
hipError_t launch_dilate(const MorphDims& d, const MorphBuffers& b, const RadiusField& rf, int se,
                         const SETableSet& tabs, int rho_lo, int rho_hi, int tabWords, hipStream_t s,
                         hipEvent_t* ev, hipEvent_t* evs) {
  MorphArgs a = base_args(d, b, rf, tabs);
  a.se = se;
  const int rmax = rho_hi + (se == 0 ? 1 : 0);
  // corner tables: m <= rmax-1 (sphere) or r (cube); offsets <= rmax (+1 for cube)
  a.maxm = se == 0 ? rmax - 1 : rho_hi;
  a.maxoff = se == 0 ? rmax : rho_hi + 1;
  a.tabWords = tabWords;
  const bool var = rf.mode != 0;
  const bool perValue = var && se == 0 && b.rp;
  long long cellwords = (long long)d.R * d.R * d.W;
  if (ev) GM_CHECK(hipEventRecord(ev[0], s));
  if (perValue && b.rpBuild) GM_CHECK(launch_radius_planes(rf, d.N, b.rp, b.rvset, s));
  hipLaunchKernelGGL(k_brick_summary, dim3(d.N / 8, d.N / 8), dim3(8 * d.W), 0, s, b.I, d.N, d.W, b.brickI,
                     nullptr, b.IFL);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[1], s));
  // uniform radius: the compiled chord tables (ball rho = r + 1, cube rho = r)
  const bool fixedOk = !var;
  const int rho = (rf.r & 0x7f) + (se == 0 ? 1 : 0);
  const bool box = se != 0;
  const int vlo = std::max(rho_lo, 2), vhi = rho_hi;
  int vprev = -1;
  a.ystep = 2;
  a.outFlags = b.DF;
  if (perValue) {
    hipLaunchKernelGGL(k_ifull, dim3((unsigned)((cellwords + 255) / 256)), dim3(256), 0, s, b.I, d.R, d.N, d.W,
                       b.DF);
    GM_CHECK(hipGetLastError());
  }
  if (perValue) GM_CHECK(launch_pass_var(0, a, b, vlo, vhi, &vprev, s, evs ? evs[0] : nullptr));
  else
    GM_CHECK(launch_pass(0, reinterpret_cast<const void*>(&k_dil_coarse), a, b.qsel, var, fixedOk, rho, s,
                         evs ? evs[0] : nullptr, box));
  if (ev) GM_CHECK(hipEventRecord(ev[2], s));
  a.outFlags = b.TG;
  a.outRows = b.D;
  hipLaunchKernelGGL(k_tight, dim3((unsigned)((cellwords + 255) / 256)), dim3(256), 0, s, a);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[3], s));
  a.ystep = 1;
  a.outRows = b.D;
  if (perValue) GM_CHECK(launch_pass_var(1, a, b, vlo, vhi, &vprev, s, evs ? evs[1] : nullptr));
  else
    GM_CHECK(launch_pass(1, reinterpret_cast<const void*>(&k_dil_fine), a, b.qsel, var, fixedOk, rho, s,
                         evs ? evs[1] : nullptr, box));
  if (ev) GM_CHECK(hipEventRecord(ev[4], s));
  return hipSuccess;
}

hipError_t launch_erode(const MorphDims& d, const MorphBuffers& b, const RadiusField& rf, const SETableSet& tabs,
                        int rho_lo, int rho_hi, int tabWords, hipStream_t s, hipEvent_t* ev, hipEvent_t* evs) {
  MorphArgs a = base_args(d, b, rf, tabs);
  int rmax = rho_hi - 1;  // erosion radius r-1
  if (rmax < 1) rmax = 1;
  // corner: m <= rmax-1, offsets <= rmax; centre: m <= rmax, offsets <= rmax
  a.maxm = rmax;
  a.maxoff = rmax;
  a.tabWords = tabWords;
  const bool var = rf.mode != 0;
  const bool perValue = var && b.rp;
  long long cellwords = (long long)d.R * d.R * d.W;
  if (ev) GM_CHECK(hipEventRecord(ev[0], s));
  if (perValue) {
    if (b.rpBuild) GM_CHECK(launch_radius_planes(rf, d.N, b.rp, b.rvset, s));
    const size_t n16 = (size_t)d.N * d.N * d.W / 4;
    hipLaunchKernelGGL(k_kv_clear, dim3(2048), dim3(256), 0, s, reinterpret_cast<uint4*>(b.kv), n16, b.kbl);
    GM_CHECK(hipGetLastError());
  }
  a.outRows = b.K;
  hipLaunchKernelGGL(k_contour_K, dim3(d.N / 8, d.N / 8), dim3(8 * d.W), 0, s, a, b.brickK,
                     perValue ? b.kbl : nullptr);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[1], s));
  const bool fixedOk = !var;
  const int rho = (rf.r & 0x7f) - 1;
  const int vlo = std::max(rho_lo, 2), vhi = rho_hi;
  int vprev = -1;
  a.ystep = 2;
  a.outFlags = b.CL;
  if (perValue) GM_CHECK(launch_pass_var(2, a, b, vlo, vhi, &vprev, s, evs ? evs[0] : nullptr));
  else
    GM_CHECK(launch_pass(2, reinterpret_cast<const void*>(&k_ero_coarse), a, b.qsel, var, fixedOk, rho, s,
                         evs ? evs[0] : nullptr));
  if (ev) GM_CHECK(hipEventRecord(ev[2], s));
  a.outFlags = b.UG;
  a.outRows = b.C;
  hipLaunchKernelGGL(k_base_contour, dim3((unsigned)((cellwords + 255) / 256)), dim3(256), 0, s, a);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[3], s));
  a.ystep = 1;
  a.outRows = b.C;
  if (perValue) GM_CHECK(launch_pass_var(3, a, b, vlo, vhi, &vprev, s, evs ? evs[1] : nullptr));
  else
    GM_CHECK(launch_pass(3, reinterpret_cast<const void*>(&k_ero_fine), a, b.qsel, var, fixedOk, rho, s,
                         evs ? evs[1] : nullptr));
  if (ev) GM_CHECK(hipEventRecord(ev[4], s));
  return hipSuccess;
}

// Ball dilation and erosion for radii above the chord tables' range, through
// the separable distance transforms of gm_edt.hip; the elementwise passes
// (brick summary, tight contour, contour K, base contour) are the same as
// above, and each transform pass writes what the chord pass it replaces
// writes (DF, D bits on TG cells, CL, C bits on UG cells). tmp and out hold
// N^3 ints each.
// ev / evs (optional): the stage events of launch_dilate / launch_erode.
hipError_t launch_dilate_edt(const MorphDims& d, const MorphBuffers& b, const RadiusField& rf, int se, int rho_hi,
                             int* tmp, int* out, hipStream_t s, hipEvent_t* ev, hipEvent_t* evs) {
  const bool cube = se != 0;
  MorphArgs a = base_args(d, b, rf, SETableSet{});
  const long long cellwords = (long long)d.R * d.R * d.W;
  if (ev) GM_CHECK(hipEventRecord(ev[0], s));
  hipLaunchKernelGGL(k_brick_summary, dim3(d.N / 8, d.N / 8), dim3(8 * d.W), 0, s, b.I, d.N, d.W, b.brickI,
                     nullptr, b.IFL);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[1], s));
  if (evs) GM_CHECK(hipEventRecord(evs[0], s));
  // fine units, corner geometry: a source reaches at most r + 1 voxels
  GM_CHECK(launch_edt(b.I, d.N, 1, 1, 0, d.N, rho_hi + 2, false, cube, rf, tmp, out, s));
  GM_CHECK(launch_edt_dil_coarse(out, d, rf, cube, b.IFL, b.DF, s));
  if (ev) GM_CHECK(hipEventRecord(ev[2], s));
  a.outFlags = b.TG;
  a.outRows = b.D;
  hipLaunchKernelGGL(k_tight, dim3((unsigned)((cellwords + 255) / 256)), dim3(256), 0, s, a);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[3], s));
  if (evs) GM_CHECK(hipEventRecord(evs[1], s));
  GM_CHECK(launch_edt_dil_fine(out, d, rf, cube, b.TG, b.D, s));
  if (ev) GM_CHECK(hipEventRecord(ev[4], s));
  return hipSuccess;
}

hipError_t launch_erode_edt(const MorphDims& d, const MorphBuffers& b, const RadiusField& rf, int rho_hi, int* tmp,
                            int* out, hipStream_t s, hipEvent_t* ev, hipEvent_t* evs) {
  MorphArgs a = base_args(d, b, rf, SETableSet{});
  const long long cellwords = (long long)d.R * d.R * d.W;
  if (ev) GM_CHECK(hipEventRecord(ev[0], s));
  a.outRows = b.K;
  hipLaunchKernelGGL(k_contour_K, dim3(d.N / 8, d.N / 8), dim3(8 * d.W), 0, s, a, b.brickK, nullptr);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[1], s));
  if (evs) GM_CHECK(hipEventRecord(evs[0], s));
  // doubled units: a source of radius r reaches 2 (r - 1); with a uniform
  // radius the sources are binary and the threshold is 4 (r - 1)^2, with
  // per-voxel radii each source carries its weight and the threshold is 0
  const bool weighted = rf.mode != 0;
  const int rr = (rf.r & 0x7f) - 1;
  const int thr = weighted ? 0 : 4 * rr * rr;
  const int reach = rho_hi + 2;
  GM_CHECK(launch_edt(b.K, d.N, 2, 4, 2, d.R, reach, weighted, false, rf, tmp, out, s));
  GM_CHECK(launch_edt_ero_coarse(out, thr, d, b.DAL, b.IFL, b.CL, s));
  if (ev) GM_CHECK(hipEventRecord(ev[2], s));
  a.outFlags = b.UG;
  a.outRows = b.C;
  hipLaunchKernelGGL(k_base_contour, dim3((unsigned)((cellwords + 255) / 256)), dim3(256), 0, s, a);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[3], s));
  if (evs) GM_CHECK(hipEventRecord(evs[1], s));
  GM_CHECK(launch_edt(b.K, d.N, 2, 2, 1, d.N, reach, weighted, false, rf, tmp, out, s));
  GM_CHECK(launch_edt_ero_fine(out, thr, d, b.UG, b.C, s));
  if (ev) GM_CHECK(hipEventRecord(ev[4], s));
  return hipSuccess;
}

// ---------------------------------------------------------------------------
// Batched closing of many uniform-radius sphere jobs: every step is one
// launch over all jobs (per radius value for the streaming passes), with the
// per-job arguments built exactly as launch_dilate / launch_erode build them.
size_t batch_morph_bytes(int n) {
  return (size_t)7 * n * sizeof(MorphArgs) + (size_t)4 * (n + 1) * sizeof(u32) + (size_t)n * sizeof(int);
}
size_t batch_queue_cap(const BatchMorphJob* jobs, int n) {
  size_t cap = 0;
  for (int j = 0; j < n; ++j) cap += var_queue_cap(jobs[j].d.R);
  return cap;
}

hipError_t launch_close_batched(const BatchMorphJob* jobs, int n, void* hBlob, void* dBlob, u32* queue, int cap,
                                int* qsel, hipStream_t s, const SideStreams& side) {
  if (n <= 0) return hipSuccess;
  MorphArgs* hT = static_cast<MorphArgs*>(hBlob);
  u32* hP = reinterpret_cast<u32*>(hT + 7 * n);
  int* hR = reinterpret_cast<int*>(hP + 4 * (n + 1));
  MorphArgs* dT = static_cast<MorphArgs*>(dBlob);
  u32* dP = reinterpret_cast<u32*>(dT + 7 * n);
  int* dR = reinterpret_cast<int*>(dP + 4 * (n + 1));
  u32* preB = hP;                  // brick rows (N/8)^2 per job
  u32* preW = hP + (n + 1);        // cell-row words / 256
  u32* preT2 = hP + 2 * (n + 1);   // tiles of the coarse passes (ystep 2)
  u32* preT1 = hP + 3 * (n + 1);   // tiles of the fine passes
  long long tilesOf[2][kCtMaxRho + 1] = {};
  bool present[kCtMaxRho + 1] = {};
  int wmax = 1;
  preB[0] = preW[0] = preT2[0] = preT1[0] = 0;
  const SETableSet noTabs{};
  for (int j = 0; j < n; ++j) {
    const BatchMorphJob& J = jobs[j];
    const int r = J.r;
    if (r < 2 || r > kCtMaxRho - 1) return hipErrorInvalidValue;
    RadiusField rf{};
    rf.mode = 0;
    rf.r = r;
    rf.N = J.d.N;
    MorphArgs a = base_args(J.d, J.b, rf, noTabs);
    a.se = 0;
    a.tabWords = 0;
    a.maxm = r;        // dilation rho = r + 1: corner m <= rho - 1, offsets <= rho
    a.maxoff = r + 1;
    MorphArgs e = a;   // erosion rho = r - 1
    e.maxm = e.maxoff = r - 1;
    MorphArgs t[7] = {a, a, a, e, e, e, e};
    t[0].ystep = 2;
    t[0].outFlags = J.b.DF;
    t[1].outFlags = J.b.TG;
    t[1].outRows = J.b.D;
    t[2].ystep = 1;
    t[2].outFlags = J.b.TG;
    t[2].outRows = J.b.D;
    t[3].outRows = J.b.K;
    t[4].ystep = 2;
    t[4].outFlags = J.b.CL;
    t[5].outFlags = J.b.UG;
    t[5].outRows = J.b.C;
    t[6].ystep = 1;
    t[6].outFlags = J.b.UG;
    t[6].outRows = J.b.C;
    for (int k = 0; k < 7; ++k) hT[k * n + j] = t[k];
    const int N = J.d.N, W = J.d.W, R = J.d.R;
    const u32 t2 = (u32)(W * ((N + 2 * kTileRows - 1) / (2 * kTileRows)) * ((N + kTileZ - 1) / kTileZ));
    const u32 t1 = (u32)tile_count_max(R);
    preB[j + 1] = preB[j] + (u32)((N / 8) * (N / 8));
    preW[j + 1] = preW[j] + (u32)(((long long)R * R * W + 255) / 256);
    preT2[j + 1] = preT2[j] + t2;
    preT1[j + 1] = preT1[j] + t1;
    hR[j] = r;
    present[r] = true;
    tilesOf[0][r] += t2;
    tilesOf[1][r] += t1;
    wmax = std::max(wmax, W);
  }
  GM_CHECK(hipMemcpyAsync(dBlob, hBlob, batch_morph_bytes(n), hipMemcpyHostToDevice, s));
  const CtKernels& ct = ct_kernels();
  // one streaming pass: selection into per-radius queues, then one launch
  // per radius present
  auto pass = [&](int p, const MorphArgs* tab, const u32* pre, u32 units) -> hipError_t {
    const int q = *qsel;
    *qsel = 1 - q;
    const void* sel[4] = {reinterpret_cast<const void*>(&k_select_b<0>), reinterpret_cast<const void*>(&k_select_b<1>),
                          reinterpret_cast<const void*>(&k_select_b<2>), reinterpret_cast<const void*>(&k_select_b<3>)};
    int nn = n, qq = q, lb = kQueueLists, cp = cap;
    void* sargs[] = {&tab, &pre, &nn, &dR, &queue, &qq, &lb, &cp};
    GM_CHECK(hipLaunchKernel(sel[p], dim3((units + 3) / 4), dim3(256), sargs, 0, s));
    // a launch per radius present, round robin over s and the side streams
    int nr = 0;
    for (int r = 2; r <= kCtMaxRho - 1; ++r) nr += present[r] ? 1 : 0;
    const int ns = std::min(nr - 1, side.n);
    if (ns > 0) {
      GM_CHECK(hipEventRecord(side.fork, s));
      for (int k = 0; k < ns; ++k) GM_CHECK(hipStreamWaitEvent(side.s[k], side.fork, 0));
    }
    int slot = 0;
    for (int r = 2; r <= kCtMaxRho - 1; ++r) {
      if (!present[r]) continue;
      const hipStream_t sr = slot == 0 ? s : side.s[slot - 1];
      slot = slot == ns ? 0 : slot + 1;
      const int rho = p < 2 ? r + 1 : r - 1;
      const void* fn = ct.fnB[p][rho];
      if (!fn) return hipErrorInvalidValue;
      MorphArgs sz{};
      sz.Zt = kTileZ;
      sz.ystep = (p == 0 || p == 2) ? 2 : 1;
      sz.maxm = p < 2 ? rho - 1 : rho;
      sz.maxoff = rho;
      sz.tabWords = 0;
      const int nw = stream_waves(sz, false);
      const size_t lds = lds_bytes(sz, false, nw);
      const long long tl = tilesOf[(p == 0 || p == 2) ? 0 : 1][r];
      const unsigned grid = (unsigned)std::min<long long>((tl + 7) & ~7LL, 8192);
      int qc = kVarCounters + 256 * q + 8 * r, ql = kQueueLists + 8 * r * cap, qcap = cap;
      const u32* cq = queue;
      void* args[] = {&tab, &cq, &qc, &ql, &qcap};
      GM_CHECK(hipLaunchKernel(fn, dim3(grid), dim3(64 * nw), args, lds, sr));
    }
    for (int k = 0; k < ns; ++k) {
      GM_CHECK(hipEventRecord(side.join[k], side.s[k]));
      GM_CHECK(hipStreamWaitEvent(s, side.join[k], 0));
    }
    return hipGetLastError();
  };
  hipLaunchKernelGGL(k_brick_summary_b, dim3(preB[n]), dim3(8 * wmax), 0, s, dT, dP, n);
  GM_CHECK(hipGetLastError());
  GM_CHECK(pass(0, dT, dP + 2 * (n + 1), preT2[n]));
  hipLaunchKernelGGL(k_tight_b, dim3(preW[n]), dim3(256), 0, s, dT + n, dP + (n + 1), n);
  GM_CHECK(hipGetLastError());
  GM_CHECK(pass(1, dT + 2 * n, dP + 3 * (n + 1), preT1[n]));
  hipLaunchKernelGGL(k_contour_K_b, dim3(preB[n]), dim3(8 * wmax), 0, s, dT + 3 * n, dP, n);
  GM_CHECK(hipGetLastError());
  GM_CHECK(pass(2, dT + 4 * n, dP + 2 * (n + 1), preT2[n]));
  hipLaunchKernelGGL(k_base_contour_b, dim3(preW[n]), dim3(256), 0, s, dT + 5 * n, dP + (n + 1), n);
  GM_CHECK(hipGetLastError());
  GM_CHECK(pass(3, dT + 6 * n, dP + 3 * (n + 1), preT1[n]));
  return hipSuccess;
}

// ---------------------------------------------------------------------------
// Host-side structuring-element tables.
static int isq_lt(long long X) {  // max m >= 0 with m*m < X (X >= 1)
  long long m = 0;
  while ((m + 1) * (m + 1) < X) ++m;
  return (int)m;
}
static long long ecen(long long t) { return t == 0 ? 0 : (2 * t - 1) * (2 * t - 1); }

// Layout (u32): [0] kind, [1] nclass, [2] max m, [3] max |offset| (HY),
// [4] total length in words, then per class: header (b | npairs<<16) and
// npairs chords (a | m<<16).
std::vector<u32> build_se_table(int kind, int rho, bool cube) {
  std::vector<u32> out = {(u32)kind, 0, 0, 0, 0};
  int nclass = 0, maxm = 0, maxab = 0;
  if (kind == SE_CORNER) {
    int bmax = cube ? rho : rho - 1;
    for (int b = 0; b <= bmax; ++b) {
      std::vector<u32> ent;
      int amax = cube ? rho : -1;
      if (!cube)
        for (int a = 0; (long long)a * a + (long long)b * b < (long long)rho * rho; ++a) amax = a;
      for (int a = 0; a <= amax; ++a) {
        int m = cube ? rho : isq_lt((long long)rho * rho - (long long)a * a - (long long)b * b);
        ent.push_back((u32)a | ((u32)m << 16));
        maxm = m > maxm ? m : maxm;
        maxab = a > maxab ? a : maxab;
      }
      if (ent.empty()) continue;
      maxab = b > maxab ? b : maxab;
      out.push_back((u32)b | ((u32)ent.size() << 16));
      out.insert(out.end(), ent.begin(), ent.end());
      ++nclass;
    }
    out[3] = (u32)(maxab + 1);  // offsets span [-(maxab+1), maxab]
  } else {
    long long lim = 4LL * rho * rho;
    for (int b = 0; ecen(b) < lim; ++b) {
      std::vector<u32> ent;
      for (int a = 0; ecen(a) + ecen(b) < lim; ++a) {
        long long rem = lim - ecen(a) - ecen(b);
        int m = 0;
        while (ecen(m + 1) < rem) ++m;
        ent.push_back((u32)a | ((u32)m << 16));
        maxm = m > maxm ? m : maxm;
        maxab = a > maxab ? a : maxab;
      }
      maxab = b > maxab ? b : maxab;
      out.push_back((u32)b | ((u32)ent.size() << 16));
      out.insert(out.end(), ent.begin(), ent.end());
      ++nclass;
    }
    out[3] = (u32)maxab;
  }
  out[1] = (u32)nclass;
  out[2] = (u32)maxm;
  out[4] = (u32)out.size();
  return out;
}

// The compiled chord tables must list exactly the chords of build_se_table.
template <int KIND, int RHO>
static bool ct_matches() {
  constexpr SETabCT<KIND, RHO> T{};
  // the box is the runtime cube table (corner geometry, cube = true)
  const std::vector<u32> h = build_se_table(KIND == SE_BOX ? SE_CORNER : KIND, RHO, KIND == SE_BOX);
  if ((int)h[2] != T.maxm || (int)h[3] != T.hy) return false;
  size_t p = 5;
  int seen = 0;
  for (u32 c = 0; c < h[1]; ++c) {
    const int b = (int)(h[p] & 0xffffu), np = (int)(h[p] >> 16);
    ++p;
    if (b >= SETabCT<KIND, RHO>::NMAX || T.np[b] != np) return false;
    for (int e = 0; e < np; ++e)
      if ((int)(h[p + e] & 0xffffu) != e || T.m[b][e] != (int)(h[p + e] >> 16)) return false;
    p += np;
    ++seen;
  }
  int listed = 0;
  for (int b = 0; b < SETabCT<KIND, RHO>::NMAX; ++b) listed += T.np[b] > 0;
  return listed == seen && (int)p == (int)h[4];
}
template <int... Rs>
static bool ct_matches_all(std::integer_sequence<int, Rs...>) {
  return ((ct_matches<SE_CORNER, Rs + 1>() && ct_matches<SE_CENTRE, Rs + 1>() && ct_matches<SE_BOX, Rs + 1>()) &&
          ...);
}
bool ct_tables_ok() { return ct_matches_all(std::make_integer_sequence<int, kCtMaxRho>{}); }

}  // namespace gm
