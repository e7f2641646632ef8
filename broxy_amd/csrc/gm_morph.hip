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
// row word column x 64 rows and streams the input planes along z: per plane
// it builds G_0..G_mmax for all rows it needs in LDS, then every lane ORs the
// chords of each dz class into a per-lane ring of output-plane accumulators.
// The chords are read from a table in uniform (scalar) memory. Radii that vary
// per voxel are handled by one pass per distinct radius found in the tile:
// gather radius (dilation) masks the output voxels, scatter radius (erosion)
// masks the source voxels.
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
  u32 m = 0;
  int base = 16 * w;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    int cx = base + j;
    if (cx >= lo && cx < hi) m |= 1u << j;
  }
  return m;
}

// shr(k): bit x reads source x+k; shl(k): bit x reads source x-k (1 <= k <= 32)
__device__ __forceinline__ u32 shr_k(u32 L, u32 C, u32 Rw, int k) {
  (void)L;
  return k >= 32 ? Rw : __builtin_amdgcn_alignbit(Rw, C, (u32)k);
}
__device__ __forceinline__ u32 shl_k(u32 L, u32 C, u32 Rw, int k) {
  (void)Rw;
  return k >= 32 ? L : __builtin_amdgcn_alignbit(C, L, (u32)(32 - k));
}

struct TabHdr {
  int kind, nclass, maxm, maxoff;
};
__device__ __forceinline__ TabHdr tab_hdr(const u32* tab) {
  return TabHdr{(int)tab[0], (int)tab[1], (int)tab[2], (int)tab[3]};
}
__device__ __forceinline__ int dz_lo(const TabHdr& h) { return h.kind == SE_CORNER ? -h.maxoff : -h.maxoff; }
__device__ __forceinline__ int dz_hi(const TabHdr& h) { return h.kind == SE_CORNER ? h.maxoff - 1 : h.maxoff; }

// LDS carve-up for one workgroup (sizes in u32).
struct Lds {
  u32* gt;   // (maxm+1) * GR
  u32* acc;  // ZR * 64
  u32* out;  // Zt * 64
  int GR, HY, ZR;
};

// One streaming pass: out[zo][lane] |= OR over the SE of the source around
// (row ty0+lane, plane zo) (masked by omask when given). src(y, z, w) returns
// the source word (0 outside the grid).
template <class Src, class OMask>
__device__ void stream_pass(const u32* __restrict__ tab, const Src& src, const OMask& omask, Lds L, int wx,
                            int ty0, int tz0, int Zt) {
  const int lane = threadIdx.x;
  const TabHdr h = tab_hdr(tab);
  const int dzl = dz_lo(h), dzh = dz_hi(h);
  const int ZR = dzh - dzl + 1;
  const int HY = h.maxoff;
  const int GR = kTileRows + 2 * HY;
  for (int s = 0; s < ZR; ++s) L.acc[s * kTileRows + lane] = 0u;
  const int zbeg = tz0 + dzl, zend = tz0 + Zt - 1 + dzh;
  for (int zi = zbeg; zi <= zend; ++zi) {
    __syncthreads();
    // stage 1: x-window dilations G_0..G_maxm of every row this tile reads
    for (int gr = lane; gr < GR; gr += kTileRows) {
      int y = ty0 - HY + gr;
      u32 Lw = src(y, zi, wx - 1), Cw = src(y, zi, wx), Rw = src(y, zi, wx + 1);
      u32 g;
      if (h.kind == SE_CORNER) {
        g = Cw | shl_k(Lw, Cw, Rw, 1);
        L.gt[gr] = g;
        for (int m = 1; m <= h.maxm; ++m) {
          g |= shr_k(Lw, Cw, Rw, m) | shl_k(Lw, Cw, Rw, m + 1);
          L.gt[m * GR + gr] = g;
        }
      } else {
        g = Cw;
        L.gt[gr] = g;
        for (int m = 1; m <= h.maxm; ++m) {
          g |= shr_k(Lw, Cw, Rw, m) | shl_k(Lw, Cw, Rw, m);
          L.gt[m * GR + gr] = g;
        }
      }
    }
    __syncthreads();
    // stage 2: chords of each dz class
    const u32* p = tab + 4;
    const int base = lane + HY;
    for (int c = 0; c < h.nclass; ++c) {
      u32 hw = p[0];
      int b = (int)(hw & 0xffffu), np = (int)(hw >> 16);
      ++p;
      u32 disk = 0u;
      if (h.kind == SE_CORNER) {
        for (int e = 0; e < np; ++e) {
          u32 ew = p[e];
          int a = (int)(ew & 0xffffu), m = (int)(ew >> 16);
          const u32* gm = L.gt + m * GR;
          disk |= gm[base + a] | gm[base - a - 1];
        }
      } else {
        for (int e = 0; e < np; ++e) {
          u32 ew = p[e];
          int a = (int)(ew & 0xffffu), m = (int)(ew >> 16);
          const u32* gm = L.gt + m * GR;
          disk |= gm[base + a] | gm[base - a];
        }
      }
      p += np;
      // dz values of the class; output plane zo = zi - dz
      int dz0 = b, dz1 = (h.kind == SE_CORNER) ? -b - 1 : -b;
      int zo = zi - dz0;
      if (zo >= tz0 && zo < tz0 + Zt) L.acc[((zo - tz0) % ZR) * kTileRows + lane] |= disk;
      if (dz1 != dz0) {
        zo = zi - dz1;
        if (zo >= tz0 && zo < tz0 + Zt) L.acc[((zo - tz0) % ZR) * kTileRows + lane] |= disk;
      }
    }
    // stage 3: plane zi - dzh has received all contributions
    int zd = zi - dzh;
    if (zd >= tz0 && zd < tz0 + Zt) {
      int s = (zd - tz0) % ZR;
      u32 v = L.acc[s * kTileRows + lane];
      L.acc[s * kTileRows + lane] = 0u;
      L.out[(zd - tz0) * kTileRows + lane] |= v & omask(ty0 + lane, zd);
    }
  }
}

struct AllOnes {
  __device__ __forceinline__ u32 operator()(int, int) const { return 0xffffffffu; }
};

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

// Output voxels whose dilation radius (r(p)+1 sphere, r(p) cube) equals rho.
struct RhoOMask {
  RadiusField rf;
  int wx, N, rho, add, ystep, yoff, zstep, zoff;
  __device__ __forceinline__ u32 operator()(int y, int z) const {
    if (y >= N || z >= N) return 0u;
    if (((y - yoff) % ystep) != 0 || ((z - zoff) % zstep) != 0) return 0u;
    u32 m = 0u;
    for (int b = 0; b < 32; ++b)
      if (radius_at(rf, 32 * wx + b, y, z) + add == rho) m |= 1u << b;
    return m;
  }
};

__device__ __forceinline__ Lds carve(u32* smem, int maxm, int maxoff, int Zt) {
  Lds L;
  L.HY = maxoff;
  L.GR = kTileRows + 2 * maxoff;
  L.ZR = 2 * maxoff + 1;
  L.gt = smem;
  L.acc = L.gt + (maxm + 1) * L.GR;
  L.out = L.acc + L.ZR * kTileRows;
  return L;
}

__device__ __forceinline__ void clear_out(Lds L, int Zt) {
  for (int i = threadIdx.x; i < Zt * kTileRows; i += kTileRows) L.out[i] = 0u;
}

// Collect the distinct radii relevant to a tile into a 128-bit set in LDS.
// Dilation: radii of the tile's output voxels (+add). Erosion: radii-1 of the
// source voxels of K within the tile's input window (halo h).
__device__ void rho_set_dilate(const RadiusField& rf, int add, int wx, int ty0, int tz0, int Zt, int N, int ystep,
                               int yoff, int zstep, int zoff, u32* set4) {
  const int lane = threadIdx.x;
  if (lane < 4) set4[lane] = 0u;
  __syncthreads();
  int y = ty0 + lane;
  u32 loc[4] = {0, 0, 0, 0};
  if (y < N && ((y - yoff) % ystep) == 0)
    for (int z = tz0; z < tz0 + Zt && z < N; ++z) {
      if (((z - zoff) % zstep) != 0) continue;
      for (int b = 0; b < 32; ++b) {
        int r = radius_at(rf, 32 * wx + b, y, z) + add;
        loc[r >> 5] |= 1u << (r & 31);
      }
    }
  for (int i = 0; i < 4; ++i)
    if (loc[i]) atomicOr(&set4[i], loc[i]);
  __syncthreads();
}

__device__ void rho_set_erode(const RadiusField& rf, const u32* K, int N, int W, int wx, int ty0, int tz0, int Zt,
                              int halo, u32* set4) {
  const int lane = threadIdx.x;
  if (lane < 4) set4[lane] = 0u;
  __syncthreads();
  u32 loc[4] = {0, 0, 0, 0};
  for (int y = ty0 - halo + lane; y < ty0 + kTileRows + halo; y += kTileRows)
    for (int z = tz0 - halo; z < tz0 + Zt + halo; ++z)
      for (int w = wx - 1; w <= wx + 1; ++w) {
        u32 v = row_word(K, N, W, y, z, w);
        while (v) {
          int b = __builtin_ctz(v);
          v &= v - 1u;
          int r = radius_at(rf, 32 * w + b, y, z) - 1;
          r = r < 0 ? 0 : r;
          loc[r >> 5] |= 1u << (r & 31);
        }
      }
  for (int i = 0; i < 4; ++i)
    if (loc[i]) atomicOr(&set4[i], loc[i]);
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Kernel parameters shared by the four streaming kernels.
struct MorphArgs {
  int R, N, W;
  int dR0, dR1, dR2;
  int Zt;
  RadiusField rf;
  int se;  // 0 sphere, 1 cube (dilation)
  const u32* tables;
  SETableSet tabs;
  int rho_lo, rho_hi;  // radius range over the whole grid (for LDS sizing and uniform fast path)
  int maxm, maxoff;    // LDS carve sizes
  const u32* I;
  const u32* D;
  const u32* K;
  u32* outRows;
  const u16* DF;
  const u16* TG;
  const u16* CL;
  const u16* UG;
  u16* outFlags;
};

__device__ __forceinline__ const u32* dil_table(const MorphArgs& a, int rho) { return a.tables + a.tabs.dil[rho]; }

// Runs the dilation predicate F over the tile into L.out. ystep/zstep select
// the output rows/planes that matter (coarse pass: even only).
__device__ void dilate_tile(const MorphArgs& a, Lds L, u32* set4, int wx, int ty0, int tz0, int ystep, int zstep) {
  GridSrc src{a.I, a.N, a.W};
  clear_out(L, a.Zt);
  const int add = a.se == 0 ? 1 : 0;
  if (a.rf.mode == 0) {
    int rho = (a.rf.r & 0x7f) + add;
    if (rho > 128 || a.tabs.dil[rho] < 0) return;
    stream_pass(dil_table(a, rho), src, AllOnes(), L, wx, ty0, tz0, a.Zt);
    return;
  }
  rho_set_dilate(a.rf, add, wx, ty0, tz0, a.Zt, a.N, ystep, 0, zstep, 0, set4);
  for (int i = 0; i < 4; ++i) {
    u32 m = set4[i];
    while (m) {
      int bit = __builtin_ctz(m);
      m &= m - 1u;
      int rho = 32 * i + bit;
      if (rho > 128 || a.tabs.dil[rho] < 0) continue;
      RhoOMask om{a.rf, wx, a.N, rho, add, ystep, 0, zstep, 0};
      stream_pass(dil_table(a, rho), src, om, L, wx, ty0, tz0, a.Zt);
    }
  }
}

template <bool CENTRE>
__device__ void erode_tile(const MorphArgs& a, Lds L, u32* set4, int wx, int ty0, int tz0) {
  clear_out(L, a.Zt);
  if (a.rf.mode == 0) {
    int rho = (a.rf.r & 0x7f) - 1;
    if (rho <= 0 || (CENTRE ? a.tabs.ero_centre[rho] : a.tabs.ero_corner[rho]) < 0) return;
    const u32* tab = a.tables + (CENTRE ? a.tabs.ero_centre[rho] : a.tabs.ero_corner[rho]);
    stream_pass(tab, GridSrc{a.K, a.N, a.W}, AllOnes(), L, wx, ty0, tz0, a.Zt);
    return;
  }
  rho_set_erode(a.rf, a.K, a.N, a.W, wx, ty0, tz0, a.Zt, a.maxoff, set4);
  for (int i = 0; i < 4; ++i) {
    u32 m = set4[i];
    while (m) {
      int bit = __builtin_ctz(m);
      m &= m - 1u;
      int rho = 32 * i + bit;
      if (rho <= 0 || (CENTRE ? a.tabs.ero_centre[rho] : a.tabs.ero_corner[rho]) < 0) continue;
      const u32* tab = a.tables + (CENTRE ? a.tabs.ero_centre[rho] : a.tabs.ero_corner[rho]);
      MaskedSrc src{a.K, a.N, a.W, a.rf, rho};
      stream_pass(tab, src, AllOnes(), L, wx, ty0, tz0, a.Zt);
    }
  }
}

// ---------------------------------------------------------------------------
// Coarse dilation -> DF flags.
__global__ void __launch_bounds__(kTileRows) k_dil_coarse(MorphArgs a) {
  extern __shared__ u32 smem[];
  __shared__ u32 set4[4];
  const int wx = blockIdx.x, ty0 = blockIdx.y * kTileRows, tz0 = blockIdx.z * a.Zt;
  Lds L = carve(smem, a.maxm, a.maxoff, a.Zt);
  dilate_tile(a, L, set4, wx, ty0, tz0, 2, 2);
  __syncthreads();
  const int lane = threadIdx.x;
  const int y = ty0 + lane;
  if (y >= a.N || (y & 1)) return;
  const int cy = y >> 1;
  const u32 xin = cell_range_mask(wx, 0, a.dR0);
  for (int z = tz0; z < tz0 + a.Zt && z < a.N; z += 2) {
    const int cz = z >> 1;
    u32 f = gather2(L.out[(z - tz0) * kTileRows + lane]);  // F at even x
    u32 i0 = row_word(a.I, a.N, a.W, y, z, wx), i1 = row_word(a.I, a.N, a.W, y + 1, z, wx);
    u32 i2 = row_word(a.I, a.N, a.W, y, z + 1, wx), i3 = row_word(a.I, a.N, a.W, y + 1, z + 1, wx);
    u32 full = i0 & i1 & i2 & i3;
    u32 ifull = gather2(full & (full >> 1));
    u32 inr = (cy < a.dR1 && cz < a.dR2) ? xin : 0u;
    a.outFlags[((size_t)cz * a.R + cy) * a.W + wx] = (u16)(ifull | (f & inr));
  }
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

__device__ __forceinline__ bool tile_has_flags(const u16* F, int R, int W, int wx, int ty0, int tz0, int Zt, int N) {
  // cells covered by the tile's rows/planes
  int cy0 = ty0 >> 1, cy1 = min((ty0 + kTileRows) >> 1, R);
  int cz0 = tz0 >> 1, cz1 = min((tz0 + Zt + 1) >> 1, R);
  u32 acc = 0u;
  for (int cz = cz0; cz < cz1; ++cz)
    for (int cy = cy0 + (int)threadIdx.x; cy < cy1; cy += kTileRows) acc |= flag_word(F, R, W, cy, cz, wx);
  return __syncthreads_or(acc != 0u) != 0;
}

// Fine dilation -> D rows.
__global__ void __launch_bounds__(kTileRows) k_dil_fine(MorphArgs a) {
  extern __shared__ u32 smem[];
  __shared__ u32 set4[4];
  const int wx = blockIdx.x, ty0 = blockIdx.y * kTileRows, tz0 = blockIdx.z * a.Zt;
  const int lane = threadIdx.x;
  const int y = ty0 + lane;
  Lds L = carve(smem, a.maxm, a.maxoff, a.Zt);
  bool work = tile_has_flags(a.TG, a.R, a.W, wx, ty0, tz0, a.Zt, a.N);
  if (work) {
    dilate_tile(a, L, set4, wx, ty0, tz0, 1, 1);
    __syncthreads();
  }
  if (y >= a.N) return;
  const int cy = y >> 1;
  for (int z = tz0; z < tz0 + a.Zt && z < a.N; ++z) {
    const int cz = z >> 1;
    u32 df = expand2(flag_word(a.DF, a.R, a.W, cy, cz, wx));
    u32 iw = row_word(a.I, a.N, a.W, y, z, wx);
    u32 d = df | iw;
    if (work) {
      u32 tg = expand2(flag_word(a.TG, a.R, a.W, cy, cz, wx));
      d = (tg & L.out[(z - tz0) * kTileRows + lane]) | (~tg & d);
    }
    a.outRows[((size_t)z * a.N + y) * a.W + wx] = d;
  }
}

// Outer contour K of D (Grid.cu:1230-1327), one thread per row word.
__global__ void k_contour_K(MorphArgs a) {
  long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long total = (long long)a.N * a.N * a.W;
  if (gid >= total) return;
  int wx = (int)(gid % a.W);
  long long t = gid / a.W;
  int y = (int)(t % a.N), z = (int)(t / a.N);
  const u32* D = a.D;
  int N = a.N, W = a.W;
  u32 c = row_word(D, N, W, y, z, wx);
  u32 l = row_word(D, N, W, y, z, wx - 1), r = row_word(D, N, W, y, z, wx + 1);
  u32 xn = (c << 1) | (l >> 31) | (c >> 1) | (r << 31);
  u32 n6 = xn | row_word(D, N, W, y - 1, z, wx) | row_word(D, N, W, y + 1, z, wx) |
           row_word(D, N, W, y, z - 1, wx) | row_word(D, N, W, y, z + 1, wx);
  // cell fullness over the 4 rows of this cell row
  int y0 = y & ~1, z0 = z & ~1;
  u32 f = row_word(D, N, W, y0, z0, wx) & row_word(D, N, W, y0 + 1, z0, wx) & row_word(D, N, W, y0, z0 + 1, wx) &
          row_word(D, N, W, y0 + 1, z0 + 1, wx);
  u32 cfull = expand2(gather2(f & (f >> 1)));
  int cy = y >> 1, cz = z >> 1;
  u32 inner = (cy >= 1 && cy <= a.R - 2 && cz >= 1 && cz <= a.R - 2) ? expand2(cell_range_mask(wx, 1, a.R - 1)) : 0u;
  a.outRows[gid] = ~c & n6 & ~cfull & inner;
}

// Coarse erosion -> CL flags.
__global__ void __launch_bounds__(kTileRows) k_ero_coarse(MorphArgs a) {
  extern __shared__ u32 smem[];
  __shared__ u32 set4[4];
  const int wx = blockIdx.x, ty0 = blockIdx.y * kTileRows, tz0 = blockIdx.z * a.Zt;
  Lds L = carve(smem, a.maxm, a.maxoff, a.Zt);
  erode_tile<false>(a, L, set4, wx, ty0, tz0);
  __syncthreads();
  const int lane = threadIdx.x;
  const int y = ty0 + lane;
  if (y >= a.N || !(y & 1)) return;
  const int cy = y >> 1;
  const u32 xin = cell_range_mask(wx, 0, a.dR0);
  for (int z = tz0 + ((tz0 & 1) ? 0 : 1); z < tz0 + a.Zt && z < a.N; z += 2) {
    const int cz = z >> 1;
    u32 e = gather2(L.out[(z - tz0) * kTileRows + lane] >> 1);  // Ec at odd x = 2c+1
    int y0 = y - 1, z0 = z - 1;
    u32 d0 = row_word(a.D, a.N, a.W, y0, z0, wx), d1 = row_word(a.D, a.N, a.W, y0 + 1, z0, wx);
    u32 d2 = row_word(a.D, a.N, a.W, y0, z0 + 1, wx), d3 = row_word(a.D, a.N, a.W, y0 + 1, z0 + 1, wx);
    u32 dany = d0 | d1 | d2 | d3;
    u32 dnz = gather2(dany | (dany >> 1));
    u32 i0 = row_word(a.I, a.N, a.W, y0, z0, wx), i1 = row_word(a.I, a.N, a.W, y0 + 1, z0, wx);
    u32 i2 = row_word(a.I, a.N, a.W, y0, z0 + 1, wx), i3 = row_word(a.I, a.N, a.W, y0 + 1, z0 + 1, wx);
    u32 full = i0 & i1 & i2 & i3;
    u32 ifull = gather2(full & (full >> 1));
    u32 inr = (cy < a.dR1 && cz < a.dR2) ? xin : 0u;
    a.outFlags[((size_t)cz * a.R + cy) * a.W + wx] = (u16)(e & dnz & ~ifull & inr);
  }
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

// Fine erosion -> C rows.
__global__ void __launch_bounds__(kTileRows) k_ero_fine(MorphArgs a) {
  extern __shared__ u32 smem[];
  __shared__ u32 set4[4];
  const int wx = blockIdx.x, ty0 = blockIdx.y * kTileRows, tz0 = blockIdx.z * a.Zt;
  const int lane = threadIdx.x;
  const int y = ty0 + lane;
  Lds L = carve(smem, a.maxm, a.maxoff, a.Zt);
  bool work = tile_has_flags(a.UG, a.R, a.W, wx, ty0, tz0, a.Zt, a.N);
  if (work) {
    erode_tile<true>(a, L, set4, wx, ty0, tz0);
    __syncthreads();
  }
  if (y >= a.N) return;
  const int cy = y >> 1;
  for (int z = tz0; z < tz0 + a.Zt && z < a.N; ++z) {
    const int cz = z >> 1;
    u32 cl = expand2(flag_word(a.CL, a.R, a.W, cy, cz, wx));
    u32 cw = row_word(a.D, a.N, a.W, y, z, wx) & ~cl;
    if (work) {
      u32 ug = expand2(flag_word(a.UG, a.R, a.W, cy, cz, wx));
      cw = (ug & ~L.out[(z - tz0) * kTileRows + lane]) | (~ug & cw);
    }
    a.outRows[((size_t)z * a.N + y) * a.W + wx] = cw;
  }
}

size_t lds_bytes(int maxm, int maxoff, int Zt) {
  int GR = kTileRows + 2 * maxoff;
  int ZR = 2 * maxoff + 1;
  return (size_t)4 * ((size_t)(maxm + 1) * GR + (size_t)ZR * kTileRows + (size_t)Zt * kTileRows);
}

MorphArgs base_args(const MorphDims& d, const MorphBuffers& b, const RadiusField& rf, const SETableSet& tabs,
                    int rho_lo, int rho_hi) {
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
  a.rho_lo = rho_lo;
  a.rho_hi = rho_hi;
  a.I = b.I;
  a.D = b.D;
  a.K = b.K;
  a.DF = b.DF;
  a.TG = b.TG;
  a.CL = b.CL;
  a.UG = b.UG;
  return a;
}

}  // namespace

// rho_lo/rho_hi: range of the per-voxel radius r over the field.
hipError_t launch_dilate(const MorphDims& d, const MorphBuffers& b, const RadiusField& rf, int se,
                         const SETableSet& tabs, int rho_lo, int rho_hi, hipStream_t s, hipEvent_t* ev) {
  MorphArgs a = base_args(d, b, rf, tabs, rho_lo, rho_hi);
  a.se = se;
  int rmax = rho_hi + (se == 0 ? 1 : 0);
  // corner tables: maxm <= rmax-1 (sphere) or r (cube); offsets <= rmax (+1 for cube)
  a.maxm = se == 0 ? rmax - 1 : rho_hi;
  a.maxoff = se == 0 ? rmax : rho_hi + 1;
  size_t lds = lds_bytes(a.maxm, a.maxoff, a.Zt);
  dim3 tiles(d.W, (d.N + kTileRows - 1) / kTileRows, (d.N + a.Zt - 1) / a.Zt);
  long long cellwords = (long long)d.R * d.R * d.W;
  dim3 ewg((unsigned)((cellwords + 255) / 256));
  if (ev) GM_CHECK(hipEventRecord(ev[0], s));
  a.outFlags = b.DF;
  hipLaunchKernelGGL(k_dil_coarse, tiles, dim3(kTileRows), lds, s, a);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[1], s));
  a.outFlags = b.TG;
  hipLaunchKernelGGL(k_tight, ewg, dim3(256), 0, s, a);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[2], s));
  a.outRows = b.D;
  hipLaunchKernelGGL(k_dil_fine, tiles, dim3(kTileRows), lds, s, a);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[3], s));
  return hipSuccess;
}

hipError_t launch_erode(const MorphDims& d, const MorphBuffers& b, const RadiusField& rf,
                        const SETableSet& tabs, int rho_lo, int rho_hi, hipStream_t s, hipEvent_t* ev) {
  MorphArgs a = base_args(d, b, rf, tabs, rho_lo, rho_hi);
  int rmax = rho_hi - 1;  // erosion radius r-1
  if (rmax < 1) rmax = 1;
  // corner: m <= rmax-1, offsets <= rmax; centre: m <= rmax, offsets <= rmax
  a.maxm = rmax;
  a.maxoff = rmax;
  size_t lds = lds_bytes(a.maxm, a.maxoff, a.Zt);
  dim3 tiles(d.W, (d.N + kTileRows - 1) / kTileRows, (d.N + a.Zt - 1) / a.Zt);
  long long rowwords = (long long)d.N * d.N * d.W;
  long long cellwords = (long long)d.R * d.R * d.W;
  if (ev) GM_CHECK(hipEventRecord(ev[0], s));
  a.outRows = b.K;
  hipLaunchKernelGGL(k_contour_K, dim3((unsigned)((rowwords + 255) / 256)), dim3(256), 0, s, a);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[1], s));
  a.outFlags = b.CL;
  hipLaunchKernelGGL(k_ero_coarse, tiles, dim3(kTileRows), lds, s, a);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[2], s));
  a.outFlags = b.UG;
  hipLaunchKernelGGL(k_base_contour, dim3((unsigned)((cellwords + 255) / 256)), dim3(256), 0, s, a);
  GM_CHECK(hipGetLastError());
  if (ev) GM_CHECK(hipEventRecord(ev[3], s));
  a.outRows = b.C;
  hipLaunchKernelGGL(k_ero_fine, tiles, dim3(kTileRows), lds, s, a);
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

std::vector<u32> build_se_table(int kind, int rho, bool cube) {
  std::vector<u32> out = {(u32)kind, 0, 0, 0};
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
  return out;
}

}  // namespace gm
