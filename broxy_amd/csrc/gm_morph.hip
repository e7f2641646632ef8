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
#include <vector>

#include "gm_internal.h"

namespace gm {

namespace {

__device__ __forceinline__ u32 spread2(u32 v) {
  v &= 0xffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}
__device__ __forceinline__ u32 gather2(u32 v) {
  v &= 0x55555555u;
  v = (v | (v >> 1)) & 0x33333333u;
  v = (v | (v >> 2)) & 0x0f0f0f0fu;
  v = (v | (v >> 4)) & 0x00ff00ffu;
  v = (v | (v >> 8)) & 0x0000ffffu;
  return v;
}
// 16 cell flags -> 32 fine bits (both x sub-voxels of each cell)
__device__ __forceinline__ u32 expand2(u32 v) {
  u32 s = spread2(v);
  return s | (s << 1);
}

__device__ __forceinline__ u32 row_word(const u32* __restrict__ g, int N, int W, int y, int z, int w) {
  if ((unsigned)y >= (unsigned)N || (unsigned)z >= (unsigned)N || (unsigned)w >= (unsigned)W) return 0u;
  return g[((size_t)z * N + y) * W + w];
}
__device__ __forceinline__ u32 flag_word(const u16* __restrict__ g, int R, int W, int cy, int cz, int w) {
  if ((unsigned)cy >= (unsigned)R || (unsigned)cz >= (unsigned)R || (unsigned)w >= (unsigned)W) return 0u;
  return g[((size_t)cz * R + cy) * W + w];
}
// Mask of cells 16w..16w+15 whose x index lies in [lo, hi).
__device__ __forceinline__ u32 cell_range_mask(int w, int lo, int hi) {
  int a = lo - 16 * w, b = hi - 16 * w;
  a = a < 0 ? 0 : (a > 16 ? 16 : a);
  b = b < 0 ? 0 : (b > 16 ? 16 : b);
  if (b <= a) return 0u;
  u32 m = (b >= 16) ? 0xffffu : ((1u << b) - 1u);
  return m & ~((1u << a) - 1u);
}

// shr(k): bit x reads source x+k; shl(k): bit x reads source x-k (1 <= k <= 32)
__device__ __forceinline__ u32 shr_k(u32 C, u32 Rw, int k) {
  return k >= 32 ? Rw : __builtin_amdgcn_alignbit(Rw, C, (u32)k);
}
__device__ __forceinline__ u32 shl_k(u32 L, u32 C, int k) {
  return k >= 32 ? L : __builtin_amdgcn_alignbit(C, L, (u32)(32 - k));
}

// Output geometry of one streaming tile: lane i of every wave owns row
// yfirst + ystep*i; output planes are zfirst, zfirst + zstep, ... <= zlast.
struct TileGeo {
  int wx;
  int yfirst, ystep;
  int tz0, zfirst, zlast, zstep;
  __device__ int nout() const { return zlast < zfirst ? 0 : (zlast - zfirst) / zstep + 1; }
};

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }
__device__ __forceinline__ int wave_id() { return (int)(threadIdx.x >> 6); }
__device__ __forceinline__ int num_waves() { return (int)(blockDim.x >> 6); }

// One streaming pass of the table tl (in LDS) over the source src, ORed into
// acc[plane index][lane] (plane index = (zo - zfirst) / zstep). The waves of
// the workgroup take the input planes round robin; each builds the G table of
// its plane in its own LDS slice (gtAll + wave * gtStride) and accumulates
// with LDS atomics, so the plane loop needs no workgroup barrier.
template <class Src>
__device__ void stream(const u32* tl, const Src& src, u32* gtAll, int gtStride, u32* acc, const TileGeo& g) {
  const int lane = lane_id(), wv = wave_id(), nw = num_waves();
  const int kind = (int)tl[0], nclass = (int)tl[1], maxm = (int)tl[2], HY = (int)tl[3];
  const int dzl = -HY, dzh = (kind == SE_CORNER) ? HY - 1 : HY;
  const int GR = g.ystep * (kTileRows - 1) + 1 + 2 * HY;
  const int ybase = g.yfirst - HY;
  u32* gt = gtAll + wv * gtStride;
  const int zbeg = g.zfirst + dzl, zend = g.zlast + dzh;
  const int zshift = g.zstep == 2 ? 1 : 0;
  const int base = g.ystep * lane + HY;
  u32 pre[4][3];
  int zi = zbeg + wv;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int gr = lane + kTileRows * k;
    pre[k][0] = pre[k][1] = pre[k][2] = 0u;
    if (gr < GR && zi <= zend) {
      pre[k][0] = src(ybase + gr, zi, g.wx - 1);
      pre[k][1] = src(ybase + gr, zi, g.wx);
      pre[k][2] = src(ybase + gr, zi, g.wx + 1);
    }
  }
  for (; zi <= zend; zi += nw) {
    // stage 1: x-window dilations of this plane's rows into the wave's slice
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int gr = lane + kTileRows * k;
      if (gr < GR) {
        const u32 Lw = pre[k][0], Cw = pre[k][1], Rw = pre[k][2];
        u32 gg;
        if (kind == SE_CORNER) {
          gg = Cw | shl_k(Lw, Cw, 1);
          gt[gr] = gg;
          for (int m = 1; m <= maxm; ++m) {
            gg |= shr_k(Cw, Rw, m) | shl_k(Lw, Cw, m + 1);
            gt[m * GR + gr] = gg;
          }
        } else {
          gg = Cw;
          gt[gr] = gg;
          for (int m = 1; m <= maxm; ++m) {
            gg |= shr_k(Cw, Rw, m) | shl_k(Lw, Cw, m);
            gt[m * GR + gr] = gg;
          }
        }
      }
    }
    // prefetch this wave's next plane while this one is reduced
    const int zn = zi + nw;
    if (zn <= zend) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        int gr = lane + kTileRows * k;
        if (gr < GR) {
          pre[k][0] = src(ybase + gr, zn, g.wx - 1);
          pre[k][1] = src(ybase + gr, zn, g.wx);
          pre[k][2] = src(ybase + gr, zn, g.wx + 1);
        }
      }
    }
    // the slice is written and read by this wave only: in-order LDS suffices
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // stage 2: chords of each dz class
    const u32* p = tl + 5;
    for (int c = 0; c < nclass; ++c) {
      const u32 hw = p[0];
      const int b = (int)(hw & 0xffffu), np = (int)(hw >> 16);
      ++p;
      u32 disk = 0u;
      if (kind == SE_CORNER) {
#pragma unroll 4
        for (int e = 0; e < np; ++e) {
          const u32 ew = p[e];
          const int a = (int)(ew & 0xffffu);
          const u32* gm = gt + (ew >> 16) * GR;
          disk |= gm[base + a] | gm[base - a - 1];
        }
      } else {
#pragma unroll 4
        for (int e = 0; e < np; ++e) {
          const u32 ew = p[e];
          const int a = (int)(ew & 0xffffu);
          const u32* gm = gt + (ew >> 16) * GR;
          disk |= gm[base + a] | gm[base - a];
        }
      }
      p += np;
      if (!disk) continue;
      // output planes zo = zi - dz for the dz values of the class
      int zo = zi - b;
      if (zo >= g.zfirst && zo <= g.zlast && ((zo - g.zfirst) & (g.zstep - 1)) == 0)
        atomicOr(&acc[((zo - g.zfirst) >> zshift) * kTileRows + lane], disk);
      const int dz1 = (kind == SE_CORNER) ? -b - 1 : -b;
      if (dz1 != b) {
        zo = zi - dz1;
        if (zo >= g.zfirst && zo <= g.zlast && ((zo - g.zfirst) & (g.zstep - 1)) == 0)
          atomicOr(&acc[((zo - g.zfirst) >> zshift) * kTileRows + lane], disk);
      }
    }
    // keep the next iteration's stage 1 behind this stage 2 in the slice
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

struct GridSrc {
  const u32* g;
  int N, W;
  __device__ __forceinline__ u32 operator()(int y, int z, int w) const { return row_word(g, N, W, y, z, w); }
};

// Source voxels whose erosion radius r(q)-1 equals rho.
struct MaskedSrc {
  const u32* g;
  int N, W;
  RadiusField rf;
  int rho;
  __device__ __forceinline__ u32 operator()(int y, int z, int w) const {
    u32 v = row_word(g, N, W, y, z, w);
    u32 out = 0u;
    while (v) {
      int b = __builtin_ctz(v);
      v &= v - 1u;
      if (radius_at(rf, 32 * w + b, y, z) - 1 == rho) out |= 1u << b;
    }
    return out;
  }
};

// ---------------------------------------------------------------------------
// Kernel parameters shared by the streaming kernels.
struct MorphArgs {
  int R, N, W;
  int dR0, dR1, dR2;
  int Zt;
  RadiusField rf;
  int se;  // 0 sphere, 1 cube (dilation)
  const u32* tables;
  SETableSet tabs;
  int maxm, maxoff, tabWords;  // LDS carve sizes
  int ystep;                   // row step of this launch's lanes
  const u32* I;
  const u32* D;
  const u32* K;
  u32* outRows;
  const u16* DF;
  const u16* TG;
  const u16* CL;
  const u16* UG;
  const uint8_t* brickI;  // per brick (1 word x 8 rows x 8 planes): bit0 any, bit1 all
  const uint8_t* brickK;
  u16* outFlags;
};

struct Lds {
  u32* tl;
  u32* gt;
  int gtStride;
  u32* out;  // output planes x 64
  u32* acc;  // per-radius pass accumulator (per-voxel radii only)
};
__device__ __forceinline__ int gt_stride(const MorphArgs& a) {
  return (a.maxm + 1) * (a.ystep * (kTileRows - 1) + 1 + 2 * a.maxoff);
}
__device__ __forceinline__ Lds carve(u32* smem, const MorphArgs& a, int nplanes) {
  Lds L;
  L.tl = smem;
  L.gt = L.tl + a.tabWords;
  L.gtStride = gt_stride(a);
  L.out = L.gt + L.gtStride * num_waves();
  L.acc = L.out + nplanes * kTileRows;
  return L;
}

__device__ __forceinline__ void load_table(u32* tl, const u32* src) {
  __syncthreads();
  const int n = (int)src[4];
  for (int i = threadIdx.x; i < n; i += blockDim.x) tl[i] = src[i];
  __syncthreads();
}

__device__ __forceinline__ int table_off(const MorphArgs& a, int which, int rho) {
  if (rho < 0 || rho > 128) return -1;
  return which == 0 ? a.tabs.dil[rho] : (which == 1 ? a.tabs.ero_corner[rho] : a.tabs.ero_centre[rho]);
}

// Any brick flag (bit) set in the brick region covering rows [y0,y1),
// planes [z0,z1), words [w0,w1] (clamped to the grid).
__device__ bool bricks_any(const uint8_t* br, int N, int W, int y0, int y1, int z0, int z1, int w0, int w1,
                           int bit) {
  const int NB = N / 8;
  int by0 = max(y0, 0) >> 3, by1 = (min(y1, N) - 1) >> 3;
  int bz0 = max(z0, 0) >> 3, bz1 = (min(z1, N) - 1) >> 3;
  w0 = max(w0, 0);
  w1 = min(w1, W - 1);
  int ny = by1 - by0 + 1, nz = bz1 - bz0 + 1, nw = w1 - w0 + 1;
  int found = 0;
  if (ny > 0 && nz > 0 && nw > 0) {
    int n = ny * nz * nw;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      int w = w0 + i % nw, t = i / nw;
      int by = by0 + t % ny, bz = bz0 + t / ny;
      if ((br[((size_t)bz * NB + by) * W + w] >> bit) & 1) found = 1;
    }
  }
  return __syncthreads_or(found) != 0;
}
// Every brick of the region has the "all" bit.
__device__ bool bricks_all(const uint8_t* br, int N, int W, int y0, int y1, int z0, int z1, int w) {
  const int NB = N / 8;
  int by0 = max(y0, 0) >> 3, by1 = (min(y1, N) - 1) >> 3;
  int bz0 = max(z0, 0) >> 3, bz1 = (min(z1, N) - 1) >> 3;
  int ny = by1 - by0 + 1, nz = bz1 - bz0 + 1;
  int miss = 0;
  for (int i = threadIdx.x; i < ny * nz; i += blockDim.x) {
    int by = by0 + i % ny, bz = bz0 + i / ny;
    if (!((br[((size_t)bz * NB + by) * W + w] >> 1) & 1)) miss = 1;
  }
  return __syncthreads_or(miss) == 0;
}

__device__ __forceinline__ bool tile_has_flags(const u16* F, int R, int W, int wx, int y0, int y1, int z0, int z1) {
  int cy0 = y0 >> 1, cy1 = min((y1 + 1) >> 1, R);
  int cz0 = z0 >> 1, cz1 = min((z1 + 1) >> 1, R);
  int ny = cy1 - cy0, nz = cz1 - cz0;
  u32 acc = 0u;
  if (ny > 0 && nz > 0)
    for (int i = threadIdx.x; i < ny * nz; i += blockDim.x) acc |= flag_word(F, R, W, cy0 + i % ny, cz0 + i / ny, wx);
  return __syncthreads_or(acc != 0u) != 0;
}

// Distinct radii of a tile (128-bit set in LDS). Dilation: radius+add of the
// output voxels; erosion: radius-1 of the K voxels within the halo.
__device__ void rho_set_dilate(const RadiusField& rf, int add, const TileGeo& g, int N, u32* set4) {
  if (threadIdx.x < 4) set4[threadIdx.x] = 0u;
  __syncthreads();
  const int y = g.yfirst + g.ystep * lane_id();
  u32 loc[4] = {0, 0, 0, 0};
  if (y < N)
    for (int p = wave_id(); p < g.nout(); p += num_waves()) {
      const int z = g.zfirst + p * g.zstep;
      for (int b = 0; b < 32; ++b) {
        int r = radius_at(rf, 32 * g.wx + b, y, z) + add;
        loc[r >> 5] |= 1u << (r & 31);
      }
    }
  for (int i = 0; i < 4; ++i)
    if (loc[i]) atomicOr(&set4[i], loc[i]);
  __syncthreads();
}
__device__ void rho_set_erode(const RadiusField& rf, const u32* K, int N, int W, const TileGeo& g, int halo,
                              u32* set4) {
  if (threadIdx.x < 4) set4[threadIdx.x] = 0u;
  __syncthreads();
  u32 loc[4] = {0, 0, 0, 0};
  const int y0 = g.yfirst - halo, y1 = g.yfirst + g.ystep * (kTileRows - 1) + halo;
  const int z0 = g.zfirst - halo, z1 = g.zlast + halo;
  const int ny = y1 - y0 + 1;
  for (int i = threadIdx.x; i < ny * (z1 - z0 + 1); i += blockDim.x) {
    const int y = y0 + i % ny, z = z0 + i / ny;
    for (int w = g.wx - 1; w <= g.wx + 1; ++w) {
      u32 v = row_word(K, N, W, y, z, w);
      while (v) {
        int b = __builtin_ctz(v);
        v &= v - 1u;
        int r = radius_at(rf, 32 * w + b, y, z) - 1;
        r = r < 0 ? 0 : r;
        loc[r >> 5] |= 1u << (r & 31);
      }
    }
  }
  for (int i = 0; i < 4; ++i)
    if (loc[i]) atomicOr(&set4[i], loc[i]);
  __syncthreads();
}

// Runs the predicate of a pass over the tile and calls fin(zo, F) for every
// output plane of the thread's row (F = 0 when skip is set). which: 0
// dilation table (radius+add), 1 erosion corner, 2 erosion centre (radius-1).
template <class Fin>
__device__ void run_tile(const MorphArgs& a, Lds L, u32* set4, const TileGeo& g, const u32* src, int which,
                         bool skip, Fin& fin) {
  const int lane = lane_id(), wv = wave_id(), nw = num_waves();
  const int nout = g.nout();
  const int add = which == 0 ? (a.se == 0 ? 1 : 0) : -1;
  if (!skip) {
    for (int i = threadIdx.x; i < nout * kTileRows; i += blockDim.x) L.out[i] = 0u;
    if (a.rf.mode == 0) {
      const int rho = (a.rf.r & 0x7f) + add;
      const int off = table_off(a, which, rho);
      if (off >= 0 && rho > 0) {
        load_table(L.tl, a.tables + off);
        stream(L.tl, GridSrc{src, a.N, a.W}, L.gt, L.gtStride, L.out, g);
      }
    } else {
      // per-voxel radii: one pass per distinct radius, masked into out
      const int y = g.yfirst + g.ystep * lane;
      if (which == 0) rho_set_dilate(a.rf, add, g, a.N, set4);
      else rho_set_erode(a.rf, src, a.N, a.W, g, a.maxoff, set4);
      for (int i = 0; i < 4; ++i) {
        u32 m = set4[i];
        while (m) {
          const int bit = __builtin_ctz(m);
          m &= m - 1u;
          const int rho = 32 * i + bit;
          const int off = table_off(a, which, rho);
          if (off < 0 || rho <= 0) continue;
          load_table(L.tl, a.tables + off);
          for (int k = threadIdx.x; k < nout * kTileRows; k += blockDim.x) L.acc[k] = 0u;
          __syncthreads();
          if (which == 0) stream(L.tl, GridSrc{src, a.N, a.W}, L.gt, L.gtStride, L.acc, g);
          else stream(L.tl, MaskedSrc{src, a.N, a.W, a.rf, rho}, L.gt, L.gtStride, L.acc, g);
          __syncthreads();
          for (int p = wv; p < nout; p += nw) {
            u32 v = L.acc[p * kTileRows + lane];
            if (!v) continue;
            if (which == 0) {
              const int z = g.zfirst + p * g.zstep;
              u32 mk = 0u;
              if (y < a.N)
                for (int b = 0; b < 32; ++b)
                  if (radius_at(a.rf, 32 * g.wx + b, y, z) + add == rho) mk |= 1u << b;
              v &= mk;
            }
            L.out[p * kTileRows + lane] |= v;
          }
        }
      }
    }
    __syncthreads();
  }
  for (int p = wv; p < nout; p += nw) fin(g.zfirst + p * g.zstep, skip ? 0u : L.out[p * kTileRows + lane]);
}

// Tiles are numbered slab-major (word fastest, then row tile, then plane
// tile). Workgroups are dealt round-robin to the 8 XCDs, so workgroup i is
// given tile (i % 8) * (T / 8) + i / 8: each XCD walks a contiguous run of
// tiles, and the W word tiles of one slab, which share every 64-byte row
// segment, meet in the same L2 at about the same time.
__device__ __forceinline__ int xcd_tile(int i, int T) {
  const int per = T >> 3;
  return i < (per << 3) ? (i & 7) * per + (i >> 3) : i;
}

__device__ __forceinline__ TileGeo tile_geo(const MorphArgs& a, int yoff, int zoff, int zstep) {
  TileGeo g;
  const int rows = a.ystep * kTileRows;
  const int nty = (a.N + rows - 1) / rows, ntz = (a.N + a.Zt - 1) / a.Zt;
  const int j = xcd_tile(blockIdx.x, a.W * nty * ntz);
  const int slab = j / a.W;
  g.wx = j - slab * a.W;
  const int ty = slab % nty, tz = slab / nty;
  g.yfirst = ty * rows + yoff;
  g.ystep = a.ystep;
  g.tz0 = tz * a.Zt;
  g.zfirst = g.tz0 + zoff;
  int zl = min(g.tz0 + a.Zt, a.N) - 1;
  g.zlast = zl < g.zfirst ? g.zfirst - zstep : g.zfirst + ((zl - g.zfirst) / zstep) * zstep;
  g.zstep = zstep;
  return g;
}

// ---------------------------------------------------------------------------
// Epilogues.
struct DilCoarseFin {  // DF(c) = I(c) full | (c < dR & F(2c))
  const MorphArgs* a;
  int w;
  int y;
  u32 xin;
  __device__ void operator()(int z, u32 f) const {
    if (y >= a->N) return;
    const int cy = y >> 1, cz = z >> 1;
    u32 i0 = row_word(a->I, a->N, a->W, y, z, w), i1 = row_word(a->I, a->N, a->W, y + 1, z, w);
    u32 i2 = row_word(a->I, a->N, a->W, y, z + 1, w);
    u32 i3 = row_word(a->I, a->N, a->W, y + 1, z + 1, w);
    u32 full = i0 & i1 & i2 & i3;
    u32 ifull = gather2(full & (full >> 1));
    u32 inr = (cy < a->dR1 && cz < a->dR2) ? xin : 0u;
    a->outFlags[((size_t)cz * a->R + cy) * a->W + w] = (u16)(ifull | (gather2(f) & inr));
  }
};
struct DilFineFin {  // D = TG ? F : (DF ? 1 : I)
  const MorphArgs* a;
  int w;
  int y;
  __device__ void operator()(int z, u32 f) const {
    if (y >= a->N) return;
    const int cy = y >> 1, cz = z >> 1;
    u32 df = expand2(flag_word(a->DF, a->R, a->W, cy, cz, w));
    u32 tg = expand2(flag_word(a->TG, a->R, a->W, cy, cz, w));
    u32 d = df | row_word(a->I, a->N, a->W, y, z, w);
    a->outRows[((size_t)z * a->N + y) * a->W + w] = (tg & f) | (~tg & d);
  }
};
struct EroCoarseFin {  // CL(c) = c < dR & D(c) != 0 & I(c) != 0xff & Ec(2c+1)
  const MorphArgs* a;
  int w;
  int y;  // odd row 2cy+1
  u32 xin;
  __device__ void operator()(int z, u32 e) const {
    if (y >= a->N) return;
    const int cy = y >> 1, cz = z >> 1, y0 = y - 1, z0 = z - 1;
    u32 d0 = row_word(a->D, a->N, a->W, y0, z0, w), d1 = row_word(a->D, a->N, a->W, y0 + 1, z0, w);
    u32 d2 = row_word(a->D, a->N, a->W, y0, z0 + 1, w), d3 = row_word(a->D, a->N, a->W, y0 + 1, z0 + 1, w);
    u32 dany = d0 | d1 | d2 | d3;
    u32 i0 = row_word(a->I, a->N, a->W, y0, z0, w), i1 = row_word(a->I, a->N, a->W, y0 + 1, z0, w);
    u32 i2 = row_word(a->I, a->N, a->W, y0, z0 + 1, w), i3 = row_word(a->I, a->N, a->W, y0 + 1, z0 + 1, w);
    u32 full = i0 & i1 & i2 & i3;
    u32 inr = (cy < a->dR1 && cz < a->dR2) ? xin : 0u;
    a->outFlags[((size_t)cz * a->R + cy) * a->W + w] =
        (u16)(gather2(e >> 1) & gather2(dany | (dany >> 1)) & ~gather2(full & (full >> 1)) & inr);
  }
};
struct EroFineFin {  // C = UG ? ~Ef : C'
  const MorphArgs* a;
  int w;
  int y;
  __device__ void operator()(int z, u32 e) const {
    if (y >= a->N) return;
    const int cy = y >> 1, cz = z >> 1;
    u32 cl = expand2(flag_word(a->CL, a->R, a->W, cy, cz, w));
    u32 ug = expand2(flag_word(a->UG, a->R, a->W, cy, cz, w));
    u32 cw = row_word(a->D, a->N, a->W, y, z, w) & ~cl;
    a->outRows[((size_t)z * a->N + y) * a->W + w] = (ug & ~e) | (~ug & cw);
  }
};

constexpr int kMaxWaves = 8;

// Coarse dilation -> DF flags (lanes own the even rows, even planes).
__global__ void __launch_bounds__(64 * kMaxWaves) k_dil_coarse(MorphArgs a) {
  extern __shared__ u32 smem[];
  __shared__ u32 set4[4];
  TileGeo g = tile_geo(a, 0, 0, 2);
  Lds L = carve(smem, a, g.nout());
  const int y1 = g.yfirst + a.ystep * kTileRows;
  // F = 0 when no occupied voxel is in reach; F unused when every cell is full
  bool skip = !bricks_any(a.brickI, a.N, a.W, g.yfirst - a.maxoff, y1 + a.maxoff, g.zfirst - a.maxoff,
                          g.zlast + 2 + a.maxoff, g.wx - 1, g.wx + 1, 0) ||
              bricks_all(a.brickI, a.N, a.W, g.yfirst, y1, g.zfirst, g.zlast + 2, g.wx);
  DilCoarseFin fin{&a, g.wx, g.yfirst + 2 * lane_id(), cell_range_mask(g.wx, 0, a.dR0)};
  run_tile(a, L, set4, g, a.I, 0, skip, fin);
}

// Fine dilation -> D rows.
__global__ void __launch_bounds__(64 * kMaxWaves) k_dil_fine(MorphArgs a) {
  extern __shared__ u32 smem[];
  __shared__ u32 set4[4];
  TileGeo g = tile_geo(a, 0, 0, 1);
  Lds L = carve(smem, a, g.nout());
  bool skip = !tile_has_flags(a.TG, a.R, a.W, g.wx, g.yfirst, g.yfirst + kTileRows, g.zfirst, g.zlast + 1);
  DilFineFin fin{&a, g.wx, g.yfirst + lane_id()};
  run_tile(a, L, set4, g, a.I, 0, skip, fin);
}

// Coarse erosion -> CL flags (lanes own the odd rows, odd planes).
__global__ void __launch_bounds__(64 * kMaxWaves) k_ero_coarse(MorphArgs a) {
  extern __shared__ u32 smem[];
  __shared__ u32 set4[4];
  TileGeo g = tile_geo(a, 1, 1, 2);
  Lds L = carve(smem, a, g.nout());
  const int y1 = g.yfirst + a.ystep * kTileRows;
  bool skip = !bricks_any(a.brickK, a.N, a.W, g.yfirst - a.maxoff, y1 + a.maxoff, g.zfirst - a.maxoff,
                          g.zlast + 1 + a.maxoff, g.wx - 1, g.wx + 1, 0);
  EroCoarseFin fin{&a, g.wx, g.yfirst + 2 * lane_id(), cell_range_mask(g.wx, 0, a.dR0)};
  run_tile(a, L, set4, g, a.K, 1, skip, fin);
}

// Fine erosion -> C rows.
__global__ void __launch_bounds__(64 * kMaxWaves) k_ero_fine(MorphArgs a) {
  extern __shared__ u32 smem[];
  __shared__ u32 set4[4];
  TileGeo g = tile_geo(a, 0, 0, 1);
  Lds L = carve(smem, a, g.nout());
  bool skip = !tile_has_flags(a.UG, a.R, a.W, g.wx, g.yfirst, g.yfirst + kTileRows, g.zfirst, g.zlast + 1);
  EroFineFin fin{&a, g.wx, g.yfirst + lane_id()};
  run_tile(a, L, set4, g, a.K, 2, skip, fin);
}

// Tight contour of DF -> TG flags (one thread per cell-row word).
__global__ void k_tight(MorphArgs a) {
  long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long total = (long long)a.R * a.R * a.W;
  if (gid >= total) return;
  int wx = (int)(gid % a.W);
  long long t = gid / a.W;
  int cy = (int)(t % a.R), cz = (int)(t / a.R);
  u32 all = 0xffffu, any = 0u;
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    int yy = cy + (d & 1), zz = cz + (d >> 1);
    u32 v = flag_word(a.DF, a.R, a.W, yy, zz, wx);
    u32 nx = flag_word(a.DF, a.R, a.W, yy, zz, wx + 1);
    u32 sh = ((v >> 1) | (nx << 15)) & 0xffffu;
    all &= v & sh;
    any |= v | sh;
  }
  u32 valid = cell_range_mask(wx, 0, a.R - 1);
  if (cy > a.R - 2 || cz > a.R - 2) valid = 0;
  a.outFlags[gid] = (u16)(any & ~all & valid);
}

// Brick summary of a grid (brick = 1 word x 8 rows x 8 planes; bit0 = any
// voxel set, bit1 = all set). Workgroup = one 8x8 (row, plane) brick column
// over every word: thread (w, row) walks the 8 planes, so each plane step is
// one contiguous read of 8 rows; the 8 rows of a word meet in LDS.
__global__ void __launch_bounds__(512) k_brick_summary(const u32* __restrict__ G, int N, int W,
                                                        uint8_t* __restrict__ out) {
  __shared__ u32 anyW[64], allW[64];
  const int NB = N / 8;
  const int by = blockIdx.x, bz = blockIdx.y;
  const int w = threadIdx.x % W, yl = threadIdx.x / W;
  if (threadIdx.x < W) {
    anyW[threadIdx.x] = 0u;
    allW[threadIdx.x] = 1u;
  }
  __syncthreads();
  const int y = 8 * by + yl;
  u32 any = 0u, all = 0xffffffffu;
  const u32* g = G + ((size_t)(8 * bz) * N + y) * W + w;
#pragma unroll
  for (int zl = 0; zl < 8; ++zl) {
    u32 v = g[(size_t)zl * N * W];
    any |= v;
    all &= v;
  }
  if (any) atomicOr(&anyW[w], 1u);
  if (all != 0xffffffffu) atomicAnd(&allW[w], 0u);
  __syncthreads();
  if (threadIdx.x < W)
    out[((size_t)bz * NB + by) * W + threadIdx.x] = (uint8_t)(anyW[threadIdx.x] | (allW[threadIdx.x] << 1));
}

// Outer contour K of D (Grid.cu:1230-1327) and its brick summary: voxels
// outside D, 6-adjacent to D, in a cell that is not full, away from the data
// border. Same workgroup mapping as k_brick_summary.
__global__ void __launch_bounds__(512) k_contour_K(MorphArgs a, uint8_t* __restrict__ brickK) {
  __shared__ u32 anyW[64];
  const int N = a.N, W = a.W, NB = N / 8;
  const int by = blockIdx.x, bz = blockIdx.y;
  const int wx = threadIdx.x % W, yl = threadIdx.x / W;
  if (threadIdx.x < W) anyW[threadIdx.x] = 0u;
  __syncthreads();
  const u32* D = a.D;
  const int y = 8 * by + yl;
  const int cy = y >> 1;
  const bool yin = cy >= 1 && cy <= a.R - 2;
  const u32 xin = expand2(cell_range_mask(wx, 1, a.R - 1));
  u32 any = 0u;
  u32 zm = row_word(D, N, W, y, 8 * bz - 1, wx);
  u32 c = row_word(D, N, W, y, 8 * bz, wx);
#pragma unroll 2
  for (int zl = 0; zl < 8; ++zl) {
    const int z = 8 * bz + zl;
    const u32 zp = row_word(D, N, W, y, z + 1, wx);
    const u32 l = row_word(D, N, W, y, z, wx - 1), r = row_word(D, N, W, y, z, wx + 1);
    const u32 xn = (c << 1) | (l >> 31) | (c >> 1) | (r << 31);
    const u32 ym = row_word(D, N, W, y - 1, z, wx), yp = row_word(D, N, W, y + 1, z, wx);
    const u32 n6 = xn | ym | yp | zm | zp;
    // the 2x2 fine rows of this voxel's cell row
    const u32 ysib = (y & 1) ? ym : yp;
    const u32 zsib = (z & 1) ? zm : zp;
    const u32 diag = row_word(D, N, W, y ^ 1, z ^ 1, wx);
    const u32 f = c & ysib & zsib & diag;
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
  __syncthreads();
  if (threadIdx.x < W) brickK[((size_t)bz * NB + by) * W + threadIdx.x] = (uint8_t)anyW[threadIdx.x];
}

// C' rows of a cell row (4 fine rows), masked by CL.
__device__ __forceinline__ void cprime_rows(const MorphArgs& a, int cy, int cz, int w, u32 r[4]) {
  u32 cl = expand2(flag_word(a.CL, a.R, a.W, cy, cz, w));
#pragma unroll
  for (int s = 0; s < 4; ++s) r[s] = row_word(a.D, a.N, a.W, 2 * cy + (s & 1), 2 * cz + (s >> 1), w) & ~cl;
}

// Base (6-connected) contour of C' -> UG flags, one thread per cell-row word.
__global__ void k_base_contour(MorphArgs a) {
  long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long total = (long long)a.R * a.R * a.W;
  if (gid >= total) return;
  int wx = (int)(gid % a.W);
  long long t = gid / a.W;
  int cy = (int)(t % a.R), cz = (int)(t / a.R);
  u32 valid = cell_range_mask(wx, 1, a.R - 1);
  if (cy < 1 || cy > a.R - 2 || cz < 1 || cz > a.R - 2) valid = 0;
  if (!valid) {
    a.outFlags[gid] = 0;
    return;
  }
  u32 c[4], l[4], r[4], o[4];
  cprime_rows(a, cy, cz, wx, c);
  cprime_rows(a, cy, cz, wx - 1, l);
  cprime_rows(a, cy, cz, wx + 1, r);
  u32 diff = 0u;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    u32 xp = (c[s] >> 2) | (r[s] << 30);  // cell j+1 aligned on cell j
    u32 xm = (c[s] << 2) | (l[s] >> 30);  // cell j-1
    diff |= (c[s] ^ xp) | (c[s] ^ xm);
  }
  const int nb[4][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    cprime_rows(a, cy + nb[q][0], cz + nb[q][1], wx, o);
#pragma unroll
    for (int s = 0; s < 4; ++s) diff |= c[s] ^ o[s];
  }
  a.outFlags[gid] = (u16)(gather2(diff | (diff >> 1)) & valid);
}

// LDS bytes of a streaming launch with `waves` waves per workgroup.
size_t lds_bytes(const MorphArgs& a, bool var, int waves) {
  const int GR = a.ystep * (kTileRows - 1) + 1 + 2 * a.maxoff;
  const int nplanes = a.ystep == 2 ? a.Zt / 2 : a.Zt;
  size_t words = (size_t)a.tabWords + (size_t)(a.maxm + 1) * GR * waves + (size_t)nplanes * kTileRows;
  if (var) words += (size_t)nplanes * kTileRows;
  return words * 4;
}
// Waves per workgroup: as many as fit a 64 KiB LDS budget, at most 8.
int stream_waves(const MorphArgs& a, bool var) {
  int w = kMaxWaves;
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
  a.Zt = kTileZ;
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
  a.brickI = b.brickI;
  a.brickK = b.brickK;
  return a;
}

dim3 stream_grid(const MorphArgs& a) {
  const int rows = a.ystep * kTileRows;
  return dim3((unsigned)(a.W * ((a.N + rows - 1) / rows) * ((a.N + a.Zt - 1) / a.Zt)));
}

}  // namespace

// rho_lo/rho_hi: range of the per-voxel radius r over the field; tabWords:
// longest SE table of the launch.
hipError_t launch_dilate(const MorphDims& d, const MorphBuffers& b, const RadiusField& rf, int se,
                         const SETableSet& tabs, int rho_lo, int rho_hi, int tabWords, hipStream_t s,
                         hipEvent_t* ev) {
  (void)rho_lo;
  MorphArgs a = base_args(d, b, rf, tabs);
  a.se = se;
  const int rmax = rho_hi + (se == 0 ? 1 : 0);
  // corner tables: m <= rmax-1 (sphere) or r (cube); offsets <= rmax (+1 for cube)
  a.maxm = se == 0 ? rmax - 1 : rho_hi;
  a.maxoff = se == 0 ? rmax : rho_hi + 1;
  a.tabWords = tabWords;
  const bool var = rf.mode != 0;
  long long cellwords = (long long)d.R * d.R * d.W;
  if (ev) GM_CHECK(hipEventRecord(ev[0], s));
  hipLaunchKernelGGL(k_brick_summary, dim3(d.N / 8, d.N / 8), dim3(8 * d.W), 0, s, b.I, d.N, d.W, b.brickI);
  GM_CHECK(hipGetLastError());
  a.ystep = 2;
  a.outFlags = b.DF;
  int nw = stream_waves(a, var);
  hipLaunchKernelGGL(k_dil_coarse, stream_grid(a), dim3(64 * nw), lds_bytes(a, var, nw), s, a);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[1], s));
  a.outFlags = b.TG;
  hipLaunchKernelGGL(k_tight, dim3((unsigned)((cellwords + 255) / 256)), dim3(256), 0, s, a);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[2], s));
  a.ystep = 1;
  a.outRows = b.D;
  nw = stream_waves(a, var);
  hipLaunchKernelGGL(k_dil_fine, stream_grid(a), dim3(64 * nw), lds_bytes(a, var, nw), s, a);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[3], s));
  return hipSuccess;
}

hipError_t launch_erode(const MorphDims& d, const MorphBuffers& b, const RadiusField& rf, const SETableSet& tabs,
                        int rho_lo, int rho_hi, int tabWords, hipStream_t s, hipEvent_t* ev) {
  (void)rho_lo;
  MorphArgs a = base_args(d, b, rf, tabs);
  int rmax = rho_hi - 1;  // erosion radius r-1
  if (rmax < 1) rmax = 1;
  // corner: m <= rmax-1, offsets <= rmax; centre: m <= rmax, offsets <= rmax
  a.maxm = rmax;
  a.maxoff = rmax;
  a.tabWords = tabWords;
  const bool var = rf.mode != 0;
  long long cellwords = (long long)d.R * d.R * d.W;
  if (ev) GM_CHECK(hipEventRecord(ev[0], s));
  a.outRows = b.K;
  hipLaunchKernelGGL(k_contour_K, dim3(d.N / 8, d.N / 8), dim3(8 * d.W), 0, s, a, b.brickK);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[1], s));
  a.ystep = 2;
  a.outFlags = b.CL;
  int nw = stream_waves(a, var);
  hipLaunchKernelGGL(k_ero_coarse, stream_grid(a), dim3(64 * nw), lds_bytes(a, var, nw), s, a);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[2], s));
  a.outFlags = b.UG;
  hipLaunchKernelGGL(k_base_contour, dim3((unsigned)((cellwords + 255) / 256)), dim3(256), 0, s, a);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[3], s));
  a.ystep = 1;
  a.outRows = b.C;
  nw = stream_waves(a, var);
  hipLaunchKernelGGL(k_ero_fine, stream_grid(a), dim3(64 * nw), lds_bytes(a, var, nw), s, a);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[4], s));
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

}  // namespace gm
