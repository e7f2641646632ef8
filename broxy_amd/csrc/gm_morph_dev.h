// Device code shared by the GMorpho morphology kernels: bit-row helpers, the
// streaming chord engine (runtime and compiled tables), tile geometry, skip
// tests, epilogues and the pass bodies. Included by gm_morph.hip (runtime-
// table kernels, host launchers) and gm_morph_ct.hip (one kernel per pass and
// compiled radius). Everything here has internal linkage.
#pragma once
#include <utility>

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
// Branch-free form of flag_word (clamped address, masked value): a run of
// them issues its loads back to back.
__device__ __forceinline__ u32 ld_flag(const u16* __restrict__ g, int R, int W, int cy, int cz, int w) {
  const bool in = (unsigned)cy < (unsigned)R && (unsigned)cz < (unsigned)R && (unsigned)w < (unsigned)W;
  const int yc = min(max(cy, 0), R - 1), zc = min(max(cz, 0), R - 1), wc = min(max(w, 0), W - 1);
  const u32 v = g[((size_t)zc * R + yc) * W + wc];
  return in ? v : 0u;
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
// [rlo, rhi]: the planes that can contribute, of the output (rout != 0: a
// mask keeps outputs of those planes only) or of the input (the source is
// zero elsewhere); the compiled streams skip every other input plane.
struct TileGeo {
  int wx;
  int yfirst, ystep;
  int tz0, zfirst, zlast, zstep;
  int rlo, rhi, rout;
  __device__ int nout() const { return zlast < zfirst ? 0 : (zlast - zfirst) / zstep + 1; }
};

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }
// uniform per wave; readfirstlane lets the compiler keep it in an SGPR
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }
__device__ __forceinline__ int num_waves() { return (int)(blockDim.x >> 6); }

// Buffer resource of a bit grid (N x N rows of W words): loads through it
// return 0 at byte offsets at or past the grid's end, so an offset of
// kOutOfGrid reads a zero word without a branch. Grids are at most 2048^3
// bits (1 GiB), below kOutOfGrid.
constexpr int kOutOfGrid = 0x7fff0000;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t grid_rsrc(const u32* G, int N, int W) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<u32*>(G), (short)0, N * N * W * 4, 0x00020000);
}
// Words wx-1, wx, wx+1 of row (y, z) of a bit grid, zero outside the grid,
// in two steps: row_window issues one 12-byte load of the words from wx - 1
// (from wx when wx = 0, a wave-uniform choice, so that the window never
// starts before the grid), and pick3 selects the three words where they are
// used. Words past the grid's end read as 0 (the buffer range check is per
// dword) and are masked anyway. Lanes own rows one row (W words) apart, so
// each lane touches its own cache line: one address and one tag lookup per
// lane instead of three. The load is a buffer-load intrinsic with an
// out-of-grid offset for rows outside the grid, and nothing selects on its
// result before the use: with a plain load the compiler turns
// "in ? load : 0" into a branch around the load and waits for it at once,
// which serializes a plane's loads and defeats the prefetch of the planes
// ahead.
typedef u32 u32x3 __attribute__((ext_vector_type(3)));
// Planes whose row loads a wave of a compiled stream keeps in flight.
#ifndef GM_STREAM_DEPTH
#define GM_STREAM_DEPTH 2
#endif
constexpr int kStreamDepth = GM_STREAM_DEPTH;
__device__ __forceinline__ void row_window(__amdgpu_buffer_rsrc_t G, int N, int W, int y, int z, int wx, u32x3& v) {
  const bool in = (unsigned)y < (unsigned)N && (unsigned)z < (unsigned)N;
  const int st = (z * N + y) * W + (wx > 0 ? wx - 1 : 0);
  v = __builtin_amdgcn_raw_buffer_load_b96(G, in ? 4 * st : kOutOfGrid, 0, 0);
}
__device__ __forceinline__ void pick3(const u32x3& v, int wx, int W, u32& l, u32& c, u32& r) {
  l = wx > 0 ? v.x : 0u;
  c = wx > 0 ? v.y : v.x;
  r = wx + 1 < W ? (wx > 0 ? v.z : v.y) : 0u;
}

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
  // row windows of the wave's next plane (rows past the window and planes
  // past zend read as rows outside the grid: zero, and no branch)
  typename Src::Win pre[4];
  const __amdgpu_buffer_rsrc_t rs = grid_rsrc(src.g, src.N, src.W);
  int zi = zbeg + wv;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int gr = lane + kTileRows * k;
    src.window(rs, (gr < GR && zi <= zend) ? ybase + gr : -1, zi, g.wx, pre[k]);
  }
  for (; zi <= zend; zi += nw) {
    // stage 1: x-window dilations of this plane's rows into the wave's slice
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int gr = lane + kTileRows * k;
      if (gr < GR) {
        u32 Lw, Cw, Rw;
        src.pick(pre[k], g.wx, Lw, Cw, Rw);
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
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int gr = lane + kTileRows * k;
      src.window(rs, (gr < GR && zn <= zend) ? ybase + gr : -1, zn, g.wx, pre[k]);
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

// Sources of a stream: window() starts the loads of words wx-1..wx+1 of row
// (y, z) (zero outside the grid), pick() gives them once they are used.
struct GridSrc {
  const u32* g;
  int N, W;
  typedef u32x3 Win;
  __device__ __forceinline__ void window(__amdgpu_buffer_rsrc_t rs, int y, int z, int wx, Win& v) const {
    row_window(rs, N, W, y, z, wx, v);
  }
  __device__ __forceinline__ void pick(const Win& v, int wx, u32& l, u32& c, u32& r) const { pick3(v, wx, W, l, c, r); }
};

// ---------------------------------------------------------------------------
// Compile-time chord tables. The same chords as build_se_table (host), built
// by a constexpr constructor so that the chord loop of a fixed radius unrolls
// into LDS reads at immediate offsets (ds_read2_b32 pairs).
constexpr long long ct_ecen(long long t) { return t == 0 ? 0 : (2 * t - 1) * (2 * t - 1); }
constexpr int ct_isq_lt(long long X) {  // max m >= 0 with m*m < X (X >= 1)
  int m = 0;
  while ((long long)(m + 1) * (m + 1) < X) ++m;
  return m;
}
template <int KIND, int RHO>
struct SETabCT {
  static constexpr int NMAX = RHO + 2;
  int m[NMAX][NMAX];  // chord half-width of (class b, dy class a); -1: no chord
  int np[NMAX];
  int nb, hy, maxm;
  constexpr SETabCT() : m{}, np{}, nb(0), hy(0), maxm(0) {
    int maxab = 0;
    for (int b = 0; b < NMAX; ++b) {
      np[b] = 0;
      for (int a = 0; a < NMAX; ++a) m[b][a] = -1;
    }
    if (KIND == SE_BOX) {
      for (int b = 0; b <= RHO; ++b) {
        for (int a = 0; a <= RHO; ++a) m[b][a] = RHO;
        np[b] = RHO + 1;
      }
      nb = RHO + 1;
      maxm = RHO;
      hy = RHO + 1;
    } else if (KIND == SE_CORNER) {
      const long long r2 = (long long)RHO * RHO;
      for (int b = 0; b <= RHO - 1; ++b) {
        for (int a = 0; (long long)a * a + (long long)b * b < r2; ++a) {
          m[b][a] = ct_isq_lt(r2 - (long long)a * a - (long long)b * b);
          np[b] = a + 1;
          maxm = m[b][a] > maxm ? m[b][a] : maxm;
          maxab = a > maxab ? a : maxab;
        }
        if (np[b]) {
          maxab = b > maxab ? b : maxab;
          nb = b + 1;
        }
      }
      hy = maxab + 1;
    } else {
      const long long lim = 4LL * RHO * RHO;
      for (int b = 0; ct_ecen(b) < lim; ++b) {
        for (int a = 0; ct_ecen(a) + ct_ecen(b) < lim; ++a) {
          const long long rem = lim - ct_ecen(a) - ct_ecen(b);
          int mm = 0;
          while (ct_ecen(mm + 1) < rem) ++mm;
          m[b][a] = mm;
          np[b] = a + 1;
          maxm = mm > maxm ? mm : maxm;
          maxab = a > maxab ? a : maxab;
        }
        maxab = b > maxab ? b : maxab;
        nb = b + 1;
      }
      hy = maxab;
    }
  }
};

// Rows of one parity half of a ystep-2 G table (GR rows): padded so that the
// two halves start 16 mod 32 rows apart. The build stores of a wave write
// lanes 2j and 2j + 1 to rows j of the two halves; with that offset (and an
// odd row stride) the 32 lanes of each half-wave store to 32 distinct banks.
#ifndef GM_GRH_ALIGN
#define GM_GRH_ALIGN 0
#endif
__host__ __device__ constexpr int ct_grh(int GR) {
  const int h = (GR + 1) / 2;
  return GM_GRH_ALIGN == 1 ? h + ((16 - h % 32) + 32) % 32 : (GM_GRH_ALIGN == 2 ? h + ((32 - h % 64) + 64) % 64 : h);
}
// G-table geometry of a fixed-radius stream. Rows of the tile window are
// r = 0..GR-1 (row yfirst - HY + r); row r holds G_0..G_MAXM in consecutive
// words (stride M1, odd, so that lanes reading rows lane + c hit distinct
// banks). ystep 1: row r at r*M1. ystep 2: rows are split by parity (r even
// at (r/2)*M1, odd at (GRH + r/2)*M1) so that the lanes, which own rows
// 2*lane, again read rows one stride apart. Every chord of a ball then lies
// within (2*HY+1)*M1 words of the lane's base.
template <int KIND, int RHO, int YS>
struct StreamCT {
  static constexpr SETabCT<KIND, RHO> T{};
  static constexpr int HY = T.hy, MAXM = T.maxm, NB = T.nb, NA = SETabCT<KIND, RHO>::NMAX;
  static constexpr int M1 = (MAXM + 1) | 1;
  static constexpr int GR = YS * (kTileRows - 1) + 1 + 2 * HY;
  static constexpr int GRH = ct_grh(GR);
  static constexpr int NK = (GR + kTileRows - 1) / kTileRows;
  static constexpr int row_at(int r) { return YS == 1 ? r * M1 : ((r & 1) * GRH + (r >> 1)) * M1; }
  // word of G_m at the row offset d from the lane's row, relative to gt + lane*M1
  static constexpr int off(int m, int d) {
    return YS == 1 ? (HY + d) * M1 + m : (((HY + d) & 1) * GRH + ((HY + d) >> 1)) * M1 + m;
  }
};
// Words per wave slice of a fixed-radius stream (upper bound, any kind).
__host__ __device__ constexpr int ct_slice_words(int maxm, int maxoff, int ystep) {
  return (maxm + 2) * (ystep == 2 ? 2 * ct_grh(ystep * (kTileRows - 1) + 1 + 2 * maxoff)
                                  : (((ystep * (kTileRows - 1) + 1 + 2 * maxoff) + 1) & ~1));
}

template <class C, int KIND, int B, int A>
__device__ __forceinline__ u32 ct_pair(const u32* gl) {
  constexpr int m = C::T.m[B][A];
  if constexpr (m < 0) {
    return 0u;
  } else if constexpr (KIND == SE_CENTRE && A == 0) {
    return gl[C::off(m, 0)];
  } else {
    constexpr int d2 = KIND != SE_CENTRE ? -A - 1 : -A;
    return gl[C::off(m, A)] | gl[C::off(m, d2)];
  }
}
template <class C, int KIND, int B, int... As>
__device__ __forceinline__ u32 ct_disk(const u32* gl, std::integer_sequence<int, As...>) {
  return (0u | ... | ct_pair<C, KIND, B, As>(gl));
}
// Chord ORs of every dz class into disk[] (all LDS reads independent).
// Chords of class B that class B+1 does not already cover. The half-width
// of a dy class never grows with |dz| (m[b][a] >= m[b+1][a], checked when
// the table is compiled), and G_m grows with m, so disk(B) = disk(B+1) | the
// chords of B whose half-width differs from B+1's.
template <class C, int B, int A>
constexpr bool ct_new_chord() {
  return C::T.m[B][A] >= 0 && (B + 1 >= C::NB || C::T.m[B + 1][A] != C::T.m[B][A]);
}
template <class C, int KIND, int B, int... As>
__device__ __forceinline__ u32 ct_new(const u32* gl, std::integer_sequence<int, As...>) {
  return (0u | ... | (ct_new_chord<C, B, As>() ? ct_pair<C, KIND, B, As>(gl) : 0u));
}
template <class C, int KIND, int B>
__device__ __forceinline__ void ct_disks_from(const u32* gl, u32* disk) {
  u32 d = ct_new<C, KIND, B>(gl, std::make_integer_sequence<int, C::NA>{});
  if constexpr (B + 1 < C::NB) d |= disk[B + 1];
  disk[B] = d;
  if constexpr (B > 0) ct_disks_from<C, KIND, B - 1>(gl, disk);
}
template <class C>
constexpr bool ct_monotone() {
  for (int b = 0; b + 1 < C::NB; ++b)
    for (int a = 0; a < C::NA; ++a)
      if (C::T.m[b + 1][a] >= 0 && C::T.m[b][a] < C::T.m[b + 1][a]) return false;
  for (int b = 0; b < C::NB; ++b)
    if (C::T.np[b] == 0) return false;
  return true;
}
// OR of class B's disk into the output planes zi - dz for its dz values; the
// plane tests are uniform over the wave.
template <class C, int KIND, int B>
__device__ __forceinline__ void ct_scatter(u32 disk, int zi, const TileGeo& g, u32* acc, int lane, int zshift) {
  if constexpr (C::T.np[B] > 0) {
    int zo = zi - B;
    if (zo >= g.zfirst && zo <= g.zlast && ((zo - g.zfirst) & (g.zstep - 1)) == 0)
      atomicOr(&acc[((zo - g.zfirst) >> zshift) * kTileRows + lane], disk);
    constexpr int dz1 = KIND != SE_CENTRE ? -B - 1 : -B;
    if constexpr (dz1 != B) {
      zo = zi - dz1;
      if (zo >= g.zfirst && zo <= g.zlast && ((zo - g.zfirst) & (g.zstep - 1)) == 0)
        atomicOr(&acc[((zo - g.zfirst) >> zshift) * kTileRows + lane], disk);
    }
  }
}
template <class C, int KIND, int... Bs>
__device__ __forceinline__ void ct_scatter_all(const u32* disk, int zi, const TileGeo& g, u32* acc, int lane,
                                               int zshift, std::integer_sequence<int, Bs...>) {
  (ct_scatter<C, KIND, Bs>(disk[Bs], zi, g, acc, lane, zshift), ...);
}


// Fixed-radius version of stream(): same plane schedule and accumulation,
// chords and G-table geometry fixed at compile time.
template <int KIND, int RHO, int YS>
__device__ void stream_ct(const GridSrc& src, u32* gtAll, int gtStride, u32* acc, const TileGeo& g) {
  using C = StreamCT<KIND, RHO, YS>;
  const int lane = lane_id(), wv = wave_id(), nw = num_waves();
  constexpr int HY = C::HY;
  constexpr int dzl = -HY, dzh = (KIND != SE_CENTRE) ? HY - 1 : HY;
  const int ybase = g.yfirst - HY;
  u32* gt = gtAll + wv * gtStride;
  const u32* gl = gt + lane * C::M1;
  const int zbeg = max(g.zfirst + dzl, g.rout ? g.rlo + dzl : g.rlo);
  const int zend = min(g.zlast + dzh, g.rout ? g.rhi + dzh : g.rhi);
  const int zshift = YS == 2 ? 1 : 0;
  static_assert(ct_monotone<C>(), "chord half-widths must not grow with |dz|");
  static_assert((size_t)C::M1 * (YS == 1 ? C::GR : 2 * C::GRH) <= (size_t)ct_slice_words(C::MAXM, C::HY, YS),
                "slice too small");
  // the rows of the wave's next kStreamDepth planes are in flight while a
  // plane is reduced: buffer s holds the row windows of plane
  // zi + ((s - i) mod kStreamDepth) nw at step i
  u32x3 buf[kStreamDepth][C::NK];
  int zi = zbeg + wv;
  const __amdgpu_buffer_rsrc_t rs = grid_rsrc(src.g, src.N, src.W);
  auto load = [&](u32x3 (&buf)[C::NK], int z) {
#pragma unroll
    for (int k = 0; k < C::NK; ++k) row_window(rs, src.N, src.W, ybase + lane + kTileRows * k, z, g.wx, buf[k]);
  };
  // x-window dilations of one plane's rows into the wave's slice
  auto build = [&](const u32x3 (&buf)[C::NK]) {
#pragma unroll
    for (int k = 0; k < C::NK; ++k) {
      const int gr = lane + kTileRows * k;
      if (k + 1 < C::NK || gr < C::GR) {
        u32 Lw, Cw, Rw;
        pick3(buf[k], g.wx, src.W, Lw, Cw, Rw);
        u32* gw = gt + (YS == 1 ? gr * C::M1 : ((gr & 1) * C::GRH + (gr >> 1)) * C::M1);
        u32 gg = KIND != SE_CENTRE ? (Cw | shl_k(Lw, Cw, 1)) : Cw;
        gw[0] = gg;
#pragma unroll
        for (int m = 1; m <= C::MAXM; ++m) {
          gg |= shr_k(Cw, Rw, m) | shl_k(Lw, Cw, KIND != SE_CENTRE ? m + 1 : m);
          gw[m] = gg;
        }
      }
    }
  };
  // chords of every class, ORed into the output planes
  auto reduce = [&](int z) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    u32 disk[C::NB];
    ct_disks_from<C, KIND, C::NB - 1>(gl, disk);
    ct_scatter_all<C, KIND>(disk, z, g, acc, lane, zshift, std::make_integer_sequence<int, C::NB>{});
    // the next build writes the slice only after these reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  // planes past zend are loaded (in bounds, masked) and not used. The
  // scheduling barriers keep each buffer's loads together and in buffer
  // order, so the wait before a build covers that buffer's loads only.
#pragma unroll
  for (int s = 0; s < kStreamDepth; ++s) {
    load(buf[s], zi + s * nw);
    __builtin_amdgcn_sched_barrier(0);
  }
  // Whole rounds of kStreamDepth steps: every step issues its loads (a
  // step past zend loads planes it never uses), so the number of loads
  // issued after a buffer's is the same on every path and each wait is
  // exact; only the build and the reduction of a plane past zend are
  // skipped (a wave-uniform test).
  zi = __builtin_amdgcn_readfirstlane(zi);
  while (zi <= zend) {
#pragma unroll
    for (int s = 0; s < kStreamDepth; ++s) {
      const bool live = zi <= zend;
      if (live) build(buf[s]);
      load(buf[s], zi + kStreamDepth * nw);
      __builtin_amdgcn_sched_barrier(0);
      if (live) reduce(zi);
      zi += nw;
    }
  }
}

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
  typedef u32x3 Win;  // the three masked words, computed at once
  __device__ __forceinline__ void window(__amdgpu_buffer_rsrc_t, int y, int z, int wx, Win& v) const {
    v.x = (*this)(y, z, wx - 1);
    v.y = (*this)(y, z, wx);
    v.z = (*this)(y, z, wx + 1);
  }
  __device__ __forceinline__ void pick(const Win& v, int, u32& l, u32& c, u32& r) const {
    l = v.x;
    c = v.y;
    r = v.z;
  }
};

// ---------------------------------------------------------------------------
// Kernel parameters shared by the streaming kernels.
struct MorphArgs {
  int R, N, W;
  int dR0, dR1, dR2;
  int Zt;
  int waveCap;  // most waves per streaming workgroup (host-side launch choice)
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
  u16* IFL;  // cells full in I (written by the brick summary of I)
  u16* DAL;  // cells with a voxel of D (written by the contour-K kernel)
  const uint8_t* brickI;  // per brick (1 word x 8 rows x 8 planes): bit0 any, bit1 all
  const uint8_t* brickK;
  u16* outFlags;
  // active-tile queues: two blocks of counters ([0..7] tiles at the front
  // of each queue list, [8..15] tiles at its back) at words 0 and 16, then
  // 8 lists of queueCap tile ids. A launch uses counter block qsel; its tile
  // selection clears the other block for the next launch.
  u32* queue;
  int queueCap;
  int qsel;
  // queue a streaming launch drains: 8 counters at word qcOff, 8 lists of
  // qCap entries from word qlOff; qBack != 0: 8 more counters at qcOff + 8
  // count entries stored from the back of each list (taken after the front)
  int qcOff, qlOff, qCap, qBack;
  // per-value queues: counter blocks of 32 values x 8 queues at
  // kVarCounters + 256*qsel, lists of vqCap entries at vqBase + (8v + x)*vqCap
  int vqBase, vqCap;
  // per-voxel radii by radius value (rvar != 0): radius planes, per-brick
  // value sets, and the value vcur a streaming launch serves
  const u32* rp;
  const u32* rvset;
  int rvar;
  int vcur;
};

struct Lds {
  u32* tl;
  u32* gt;
  int gtStride;
  u32* out;  // output planes x 64
  u32* acc;  // per-radius pass accumulator (per-voxel radii only)
};
__device__ __forceinline__ int gt_stride(const MorphArgs& a) { return ct_slice_words(a.maxm, a.maxoff, a.ystep); }
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

// Skip tests of one tile, evaluated by one wave (the result is uniform over
// the wave).
// Whether a brick flag (bit) is set in the brick region covering rows
// [y0,y1), planes [z0,z1), words [w0,w1] (clamped to the grid), and the
// first and last brick plane (bz) holding one (lo > hi when none does;
// uniform over the wave).
__device__ bool wave_bricks_zrange(const uint8_t* br, int N, int W, int y0, int y1, int z0, int z1, int w0, int w1,
                                   int bit, int& lo, int& hi) {
  const int NB = N / 8;
  const int by0 = max(y0, 0) >> 3, by1 = (min(y1, N) - 1) >> 3;
  const int bz0 = max(z0, 0) >> 3, bz1 = (min(z1, N) - 1) >> 3;
  w0 = max(w0, 0);
  w1 = min(w1, W - 1);
  const int ny = by1 - by0 + 1, nz = bz1 - bz0 + 1, nw = w1 - w0 + 1;
  int llo = 1 << 20, lhi = -1;
  if (ny > 0 && nz > 0 && nw > 0) {
    const int n = ny * nz * nw;
    for (int i0 = 0; i0 < n; i0 += 256) {
      u32 v[4];
      int zb[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = min(i0 + 64 * q + lane_id(), n - 1);
        const int w = w0 + i % nw, t = i / nw;
        const int by = by0 + t % ny, bz = bz0 + t / ny;
        v[q] = br[((size_t)bz * NB + by) * W + w];
        zb[q] = bz;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if ((v[q] >> bit) & 1u) {
          llo = min(llo, zb[q]);
          lhi = max(lhi, zb[q]);
        }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    llo = min(llo, __shfl_xor(llo, off));
    lhi = max(lhi, __shfl_xor(lhi, off));
  }
  lo = __builtin_amdgcn_readfirstlane(llo);
  hi = __builtin_amdgcn_readfirstlane(lhi);
  return hi >= lo;
}
// wave_bricks_zrange of bit 0 over the region, and at the same time whether
// every brick of the inner region (rows [ay0,ay1), planes [az0,az1), word aw;
// inside the region) has bit 1 (all): one pass over the bricks, 8 loads in
// flight per lane, unsigned index arithmetic.
__device__ bool wave_bricks_zrange_all(const uint8_t* br, int N, int W, int y0, int y1, int z0, int z1, int w0,
                                       int w1, int ay0, int ay1, int az0, int az1, int aw, int& lo, int& hi,
                                       bool& all) {
  const int NB = N / 8;
  const int by0 = max(y0, 0) >> 3, by1 = (min(y1, N) - 1) >> 3;
  const int bz0 = max(z0, 0) >> 3, bz1 = (min(z1, N) - 1) >> 3;
  const int aby0 = max(ay0, 0) >> 3, aby1 = (min(ay1, N) - 1) >> 3;
  const int abz0 = max(az0, 0) >> 3, abz1 = (min(az1, N) - 1) >> 3;
  w0 = max(w0, 0);
  w1 = min(w1, W - 1);
  const int ny = by1 - by0 + 1, nz = bz1 - bz0 + 1, nw = w1 - w0 + 1;
  int llo = 1 << 20, lhi = -1;
  u32 inner = 2u;
  if (ny > 0 && nz > 0 && nw > 0) {
    const u32 n = (u32)(ny * nz * nw);
    for (u32 i0 = 0; i0 < n; i0 += 512) {
      u32 v[8];
      int zb[8], in[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const u32 i = min(i0 + 64u * q + (u32)lane_id(), n - 1u);
        const u32 t = i / (u32)nw;
        const int w = w0 + (int)(i - t * (u32)nw);
        const u32 tz = t / (u32)ny;
        const int by = by0 + (int)(t - tz * (u32)ny), bz = bz0 + (int)tz;
        v[q] = br[((size_t)bz * NB + by) * W + w];
        zb[q] = bz;
        in[q] = w == aw && by >= aby0 && by <= aby1 && bz >= abz0 && bz <= abz1;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if (v[q] & 1u) {
          llo = min(llo, zb[q]);
          lhi = max(lhi, zb[q]);
        }
        if (in[q]) inner &= v[q];
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    llo = min(llo, __shfl_xor(llo, off));
    lhi = max(lhi, __shfl_xor(lhi, off));
  }
  lo = __builtin_amdgcn_readfirstlane(llo);
  hi = __builtin_amdgcn_readfirstlane(lhi);
  all = __any((inner & 2u) == 0u) == 0;
  return hi >= lo;
}

// Some cell-row flag of word column wx set in rows [y0,y1), planes [z0,z1),
// and the fine planes [olo, ohi] of the cell planes holding one
// (wave-uniform; olo > ohi when none does).
__device__ bool wave_tile_flag_planes(const u16* F, int R, int W, int wx, int y0, int y1, int z0, int z1, int& olo,
                                      int& ohi) {
  const int cy0 = y0 >> 1, cy1 = min((y1 + 1) >> 1, R);
  const int cz0 = z0 >> 1, cz1 = min((z1 + 1) >> 1, R);
  const int ny = cy1 - cy0, nz = cz1 - cz0;
  int clo = 1 << 20, chi = -1;
  if (ny > 0 && nz > 0) {
    const int n = ny * nz;
    for (int i0 = 0; i0 < n; i0 += 256) {
      u32 v[4];
      int cz[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = min(i0 + 64 * q + lane_id(), n - 1);
        cz[q] = cz0 + i / ny;
        v[q] = ld_flag(F, R, W, cy0 + i % ny, cz[q], wx);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (v[q]) {
          clo = min(clo, cz[q]);
          chi = max(chi, cz[q]);
        }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    clo = min(clo, __shfl_xor(clo, off));
    chi = max(chi, __shfl_xor(chi, off));
  }
  olo = __builtin_amdgcn_readfirstlane(2 * clo);
  ohi = __builtin_amdgcn_readfirstlane(2 * chi + 1);
  return chi >= 0;
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

// Word of row (y, z, w) in the brick-major layout of the radius planes: the
// 64 words of a brick (1 word x 8 rows x 8 planes) are contiguous, row
// fastest, bricks in grid order.
__device__ __forceinline__ size_t brick_major(int N, int W, int y, int z, int w) {
  return (((size_t)(z >> 3) * (N >> 3) + (y >> 3)) * W + w) * 64 + (z & 7) * 8 + (y & 7);
}
// Bits of row word (y, z, w) whose radius value is v (radius planes rp; the
// address is clamped, callers mask rows outside the grid).
__device__ __forceinline__ u32 radius_eq(const u32* __restrict__ rp, int N, int W, int y, int z, int w, int v) {
  const int yc = min(max(y, 0), N - 1), zc = min(max(z, 0), N - 1);
  const size_t plane = (size_t)N * N * W;
  const u32* p = rp + brick_major(N, W, yc, zc, w);
  u32 b[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) b[k] = p[k * plane];
  u32 m = ~0u;
#pragma unroll
  for (int k = 0; k < 5; ++k) m &= ((v >> k) & 1) ? b[k] : ~b[k];
  return m;
}


// Radius values present in the bricks of rows [y0,y1), planes [z0,z1), words
// [w0,w1] (clamped to the grid; with gate, only bricks whose gate bit0 is
// set), and in lane v (< 32) the first and last brick plane (bz) holding
// value v (lo > hi when none does).
__device__ u32 wave_vset_planes(const u32* vset, const uint8_t* gate, int N, int W, int y0, int y1, int z0, int z1,
                                int w0, int w1, int& lo, int& hi) {
  const int NB = N / 8;
  const int by0 = max(y0, 0) >> 3, by1 = (min(y1, N) - 1) >> 3;
  const int bz0 = max(z0, 0) >> 3, bz1 = (min(z1, N) - 1) >> 3;
  w0 = max(w0, 0);
  w1 = min(w1, W - 1);
  const int ny = by1 - by0 + 1, nw = w1 - w0 + 1;
  const int lane = lane_id();
  lo = 1 << 20;
  hi = -1;
  u32 all = 0u;
  if (ny <= 0 || nw <= 0) return 0u;
  const int n = ny * nw;
  for (int bz = bz0; bz <= bz1; ++bz) {
    u32 found = 0u;
    for (int i0 = 0; i0 < n; i0 += 64) {
      const int i = min(i0 + lane, n - 1);
      const int w = w0 + i % nw, by = by0 + i / nw;
      const size_t b = ((size_t)bz * NB + by) * W + w;
      const u32 v = vset[b];
      found |= (gate == nullptr || (gate[b] & 1u)) ? v : 0u;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) found |= (u32)__shfl_xor((int)found, off);
    all |= found;
    if ((found >> (lane & 31)) & 1u) {
      lo = min(lo, bz);
      hi = max(hi, bz);
    }
  }
  return all;
}

// Runs the predicate of a pass over the tile and calls fin(zo, F) for every
// output plane of the thread's row (F = 0 when skip is set). which: 0
// dilation table (radius+add), 1 erosion corner, 2 erosion centre (radius-1).
template <int WHICH, int YS, int RHO, int KIND = -1, class Fin>
__device__ void run_tile(const MorphArgs& a, Lds L, u32* set4, const TileGeo& g, const u32* src, bool skip,
                         Fin& fin) {
  constexpr int which = WHICH;
  const int lane = lane_id(), wv = wave_id(), nw = num_waves();
  const int nout = g.nout();
  const int add = which == 0 ? (a.se == 0 ? 1 : 0) : -1;
  if (!skip) {
    for (int i = threadIdx.x; i < nout * kTileRows; i += blockDim.x) L.out[i] = 0u;
    if constexpr (RHO > 0) {
      // uniform radius with a compiled chord table (the host picks the kernel)
      __syncthreads();
      // (KIND >= 0: the compiled table of that kind, the cube's for the box dilations)
      stream_ct<KIND >= 0 ? KIND : (WHICH == 2 ? SE_CENTRE : SE_CORNER), RHO, YS>(GridSrc{src, a.N, a.W}, L.gt,
                                                                                   L.gtStride, L.out, g);
    } else if (a.rf.mode == 0) {
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
  if (skip) {
    for (int p = wv; p < nout; p += nw) fin.skipped(g.zfirst + p * g.zstep);
    return;
  }
  // the wave's output planes, 4 at a time: all loads first, then the stores
  for (int p0 = wv; p0 < nout; p0 += 4 * nw) {
    typename Fin::In in[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) fin.load(g.zfirst + min(p0 + k * nw, nout - 1) * g.zstep, in[k]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int p = p0 + k * nw;
      if (p < nout) fin.store(g.zfirst + p * g.zstep, L.out[p * kTileRows + lane], in[k]);
    }
  }
}

constexpr int kQueueList = kQueueLists;  // first word of the uniform-radius lists

__device__ __forceinline__ int tile_count(const MorphArgs& a) {
  const int rows = a.ystep * kTileRows;
  return a.W * ((a.N + rows - 1) / rows) * ((a.N + a.Zt - 1) / a.Zt);
}
// Queue (one per XCD) of the tiles of slab s: slab 8q + j goes to (j + q) % 8,
// so that the word tiles of a slab, which share every 64-byte row segment,
// run in one L2, and a row band or plane band spreads over all XCDs.
__device__ __forceinline__ int slab_queue(int slab) { return ((slab & 7) + (slab >> 3)) & 7; }

// Geometry of tile j (word fastest, then row tile, then plane tile).
__device__ __forceinline__ TileGeo tile_at(const MorphArgs& a, int j, int yoff, int zoff, int zstep) {
  TileGeo g;
  const int rows = a.ystep * kTileRows;
  const int nty = (a.N + rows - 1) / rows;
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
  g.rlo = -(1 << 20);
  g.rhi = 1 << 20;
  g.rout = 0;
  return g;
}

// Queue entries: tile id in the low kEntryTileBits bits, then the plane range
// [rlo, rhi] that can contribute, 7 bits each, relative to tz0 - kEntryZBase
// (clamped, which only widens the range).
constexpr int kEntryTileBits = 18;
constexpr int kEntryZBase = 40;
__device__ __forceinline__ u32 entry_pack(int t, int tz0, int rlo, int rhi) {
  const int lo = min(max(rlo - tz0 + kEntryZBase, 0), 127), hi = min(max(rhi - tz0 + kEntryZBase, 0), 127);
  return (u32)t | ((u32)lo << kEntryTileBits) | ((u32)hi << (kEntryTileBits + 7));
}
__device__ __forceinline__ int entry_tile(u32 e) { return (int)(e & ((1u << kEntryTileBits) - 1u)); }
__device__ __forceinline__ void entry_range(u32 e, TileGeo& g, int rout) {
  g.rlo = (int)((e >> kEntryTileBits) & 127u) + g.tz0 - kEntryZBase;
  g.rhi = (int)((e >> (kEntryTileBits + 7)) & 127u) + g.tz0 - kEntryZBase;
  g.rout = rout;
}

// ---------------------------------------------------------------------------
// Epilogues.
// Each epilogue loads what its output word needs (load), then combines it
// with the tile result (store); skipped() writes the output of a tile whose
// result is known without the predicate (nothing to load).
// With per-voxel radii (a->rvar) one launch serves one radius value: the
// coarse outputs were initialised before the launches and every launch ORs
// in its bits (dilation: output voxels of value vcur; erosion: the source is
// K restricted to that value); the fine outputs start from their default
// (D: I on tagged cells; C: all ones on tagged cells) and accumulate the same
// way. A uniform radius is the one-launch case of the same formulas.
struct DilCoarseFin {  // DF(c) = I(c) full | (c < dR & F(2c))
  const MorphArgs* a;
  int w;
  int y;
  u32 xin;
  u16 skipValue;  // 0: no voxel in reach (I empty here); 0xffff: every cell full
  struct In {
    u32 full;  // uniform: I-full bits of the 4 rows; per value: DF so far
    u32 m;     // per value: output voxels of value vcur
  };
  __device__ size_t at(int z) const { return ((size_t)(z >> 1) * a->R + (y >> 1)) * a->W + w; }
  __device__ void load(int z, In& in) const {
    if (a->rvar) {
      in.full = ld_flag(a->outFlags, a->R, a->W, y >> 1, z >> 1, w);
      in.m = radius_eq(a->rp, a->N, a->W, y, z, w, a->vcur);
    } else {
      in.full = ld_flag(a->IFL, a->R, a->W, y >> 1, z >> 1, w);
      in.m = ~0u;
    }
  }
  __device__ void store(int z, u32 f, const In& in) const {
    if (y >= a->N) return;
    const int cy = y >> 1, cz = z >> 1;
    const u32 inr = (cy < a->dR1 && cz < a->dR2) ? xin : 0u;
    a->outFlags[at(z)] = (u16)(in.full | (gather2(f & in.m) & inr));
  }
  __device__ void skipped(int z) const {
    if (y >= a->N) return;
    a->outFlags[at(z)] = skipValue;
  }
};
// The fine epilogues change only the bits they set or clear, with a device
// atomic on the output word (each word belongs to one tile), so D and C are
// not read back.
struct DilFineFin {  // D |= TG & F (& value mask); D holds TG ? I : (DF ? 1 : I)
  const MorphArgs* a;
  int w;
  int y;
  struct In {
    u32 tg, m;
  };
  __device__ void load(int z, In& in) const {
    in.tg = ld_flag(a->TG, a->R, a->W, y >> 1, z >> 1, w);
    in.m = a->rvar ? radius_eq(a->rp, a->N, a->W, y, z, w, a->vcur) : ~0u;
  }
  __device__ void store(int z, u32 f, const In& in) const {
    if (y >= a->N) return;
    const u32 v = expand2(in.tg) & f & in.m;
    if (v) atomicOr(&a->outRows[((size_t)z * a->N + y) * a->W + w], v);
  }
  __device__ void skipped(int) const {}
};
struct EroCoarseFin {  // CL(c) = c < dR & D(c) != 0 & I(c) != 0xff & Ec(2c+1)
  const MorphArgs* a;
  int w;
  int y;  // odd row 2cy+1
  u32 xin;
  struct In {
    u32 dany, full, old;  // cells: with a voxel of D, full in I; CL so far
  };
  __device__ size_t at(int z) const { return ((size_t)(z >> 1) * a->R + (y >> 1)) * a->W + w; }
  __device__ void load(int z, In& in) const {
    const int cy = y >> 1, cz = z >> 1;
    in.dany = ld_flag(a->DAL, a->R, a->W, cy, cz, w);
    in.full = ld_flag(a->IFL, a->R, a->W, cy, cz, w);
    in.old = a->rvar ? ld_flag(a->outFlags, a->R, a->W, cy, cz, w) : 0u;
  }
  __device__ void store(int z, u32 e, const In& in) const {
    if (y >= a->N) return;
    const int cy = y >> 1, cz = z >> 1;
    const u32 inr = (cy < a->dR1 && cz < a->dR2) ? xin : 0u;
    a->outFlags[at(z)] = (u16)(in.old | (gather2(e >> 1) & in.dany & ~in.full & inr));
  }
  __device__ void skipped(int z) const {  // no K in reach: E = 0
    if (y >= a->N) return;
    a->outFlags[at(z)] = 0;
  }
};
struct EroFineFin {  // C &= ~(UG & Ef); C holds C' = D & ~CL, all ones on UG cells
  const MorphArgs* a;
  int w;
  int y;
  struct In {
    u32 ug;
  };
  __device__ void load(int z, In& in) const { in.ug = ld_flag(a->UG, a->R, a->W, y >> 1, z >> 1, w); }
  __device__ void store(int z, u32 e, const In& in) const {
    if (y >= a->N) return;
    const u32 v = expand2(in.ug) & e;
    if (v) atomicAnd(&a->outRows[((size_t)z * a->N + y) * a->W + w], ~v);
  }
  __device__ void skipped(int) const {}
};

constexpr int kMaxWaves = 8;

// Pass geometry. PASS 0: coarse dilation -> DF flags (lanes own the even
// rows, even planes); 1: fine dilation -> D rows; 2: coarse erosion -> CL
// flags (odd rows, odd planes); 3: fine erosion -> C rows.
template <int PASS>
__device__ __forceinline__ TileGeo pass_tile(const MorphArgs& a, int t) {
  if constexpr (PASS == 0) return tile_at(a, t, 0, 0, 2);
  else if constexpr (PASS == 2) return tile_at(a, t, 1, 1, 2);
  else return tile_at(a, t, 0, 0, 1);
}

// Tile selection, one wave per tile: a tile whose result is known without the
// predicate is finished here (coarse passes store their constant; the fine
// passes' default output is already written), every other tile is appended to
// the queue of its slab. With per-voxel radii the tile goes to the queue of
// every radius value that can reach it (per-brick value sets; erosion: of the
// bricks holding K); the coarse erosion output of every tile is cleared here
// for the per-value launches to OR into (the coarse dilation output is set
// to I-full beforehand).
// Skip test of tile t for pass PASS (one wave): finishes a skipped tile (a
// coarse pass stores its constant), returns whether the tile is active. With
// per-voxel radii, vals gets the radius values that reach the tile and lane v
// the brick planes [blo, bhi] of value v.
// olo, ohi: the output planes that need the predicate (fine passes: the
// planes holding contour tags; coarse passes: all).
template <int PASS>
__device__ bool select_tile(const MorphArgs& a, int t, const TileGeo& g, u32& vals, int& blo, int& bhi, int& slo,
                            int& shi, int& olo, int& ohi) {
  const int lane = lane_id();
  const int y1 = g.yfirst + a.ystep * kTileRows;
  bool skip;
  vals = 0u;
  blo = 0;
  bhi = -1;
  slo = 0;
  shi = -1;
  olo = g.zfirst;
  ohi = g.zlast;
  if constexpr (PASS == 0) {
    // F = 0 when no occupied voxel is in reach; F unused when every cell is full
    bool allFull;
    const bool empty = !wave_bricks_zrange_all(a.brickI, a.N, a.W, g.yfirst - a.maxoff, y1 + a.maxoff,
                                               g.zfirst - a.maxoff, g.zlast + 2 + a.maxoff, g.wx - 1, g.wx + 1,
                                               g.yfirst, y1, g.zfirst, g.zlast + 2, g.wx, slo, shi, allFull);
    const bool full = !empty && allFull;
    skip = empty || full;
    if (a.rvar) {
      // DF already holds I-full, which is also the value of a skipped tile
      if (!skip)
        vals = wave_vset_planes(a.rvset, nullptr, a.N, a.W, g.yfirst, y1, g.zfirst, g.zlast + 1, g.wx, g.wx, blo, bhi);
    } else if (skip) {
      DilCoarseFin fin{&a, g.wx, g.yfirst + 2 * lane, 0u, (u16)(full ? 0xffffu : 0u)};
      for (int p = 0; p < g.nout(); ++p) fin.skipped(g.zfirst + p * g.zstep);
    }
  } else if constexpr (PASS == 2) {
    skip = !wave_bricks_zrange(a.brickK, a.N, a.W, g.yfirst - a.maxoff, y1 + a.maxoff, g.zfirst - a.maxoff,
                               g.zlast + 1 + a.maxoff, g.wx - 1, g.wx + 1, 0, slo, shi);
    if (skip || a.rvar) {
      EroCoarseFin fin{&a, g.wx, g.yfirst + 2 * lane, 0u};
      for (int p = 0; p < g.nout(); ++p) fin.skipped(g.zfirst + p * g.zstep);
    }
    if (!skip && a.rvar)
      vals = wave_vset_planes(a.rvset, a.brickK, a.N, a.W, g.yfirst - a.maxoff, y1 + a.maxoff, g.zfirst - a.maxoff,
                              g.zlast + 1 + a.maxoff, g.wx - 1, g.wx + 1, blo, bhi);
  } else {
    skip = !wave_tile_flag_planes(PASS == 1 ? a.TG : a.UG, a.R, a.W, g.wx, g.yfirst, g.yfirst + kTileRows,
                                  g.zfirst, g.zlast + 1, olo, ohi);
    olo = max(olo, g.zfirst);
    ohi = min(ohi, g.zlast);
    if (!skip && !a.rvar)  // the source planes in reach (I, or K)
      wave_bricks_zrange(PASS == 1 ? a.brickI : a.brickK, a.N, a.W, g.yfirst - a.maxoff, y1 + a.maxoff,
                         g.zfirst - a.maxoff, g.zlast + 1 + a.maxoff, g.wx - 1, g.wx + 1, 0, slo, shi);
    if (!skip && a.rvar) {
      if constexpr (PASS == 1)
        vals = wave_vset_planes(a.rvset, nullptr, a.N, a.W, g.yfirst, y1, g.zfirst, g.zlast + 1, g.wx, g.wx, blo, bhi);
      else
        vals = wave_vset_planes(a.rvset, a.brickK, a.N, a.W, g.yfirst - a.maxoff, y1 + a.maxoff,
                                g.zfirst - a.maxoff, g.zlast + 1 + a.maxoff, g.wx - 1, g.wx + 1, blo, bhi);
    }
  }
  return !skip;
}

// Tile selection, one wave per tile: a tile whose result is known without the
// predicate is finished here (coarse passes store their constant; the fine
// passes' default output is already written), every other tile is appended to
// the queue of its slab. With per-voxel radii the tile goes to the queue of
// every radius value that can reach it (per-brick value sets; erosion: of the
// bricks holding K); the coarse erosion output of every tile is cleared here
// for the per-value launches to OR into (the coarse dilation output is set
// to I-full beforehand).
template <int PASS>
__device__ void select_tiles(const MorphArgs& a) {
  // queue reservations per workgroup: its tiles (consecutive, mostly of one
  // slab) take one atomic per queue list they use, so the queue counters see
  // few atomics. Lists: uniform radius [x], per-voxel radii [8v + x].
  __shared__ u32 wgCount[256], wgBase[256];
  if (blockIdx.x == 0) {
    // the counter blocks of the next pass (the previous pass has finished)
    if (threadIdx.x < 16) a.queue[16 * (1 - a.qsel) + threadIdx.x] = 0u;
    for (int i = threadIdx.x; i < 256; i += blockDim.x) a.queue[kVarCounters + 256 * (1 - a.qsel) + i] = 0u;
  }
  for (int i = threadIdx.x; i < 256; i += blockDim.x) wgCount[i] = 0u;
  __syncthreads();
  const int t = blockIdx.x * (int)(blockDim.x >> 6) + wave_id();
  const int lane = lane_id();
  int list = -1;  // this lane's queue list, or none
  u32 slot = 0u, entry = 0u;
  if (t < tile_count(a)) {
    const TileGeo g = pass_tile<PASS>(a, t);
    u32 vals;
    int blo, bhi;  // lane v: brick planes holding value v
    int slo, shi;  // brick planes holding source voxels in reach (uniform radius)
    int olo, ohi;  // output planes that need the predicate
    if (select_tile<PASS>(a, t, g, vals, blo, bhi, slo, shi, olo, ohi)) {
      const int x = slab_queue(t / a.W);
      if (a.rvar) {
        // lane v queues the tile for radius value v, with the planes that can
        // contribute: the output planes of value v (dilation, outputs masked
        // by value) or the input planes of K of value v (erosion, source
        // masked)
        if (lane < 32 && ((vals >> lane) & 1u)) {
          int rlo = 8 * blo, rhi = 8 * bhi + 7;
          if constexpr (PASS < 2) {
            rlo = max(rlo, olo);
            rhi = min(rhi, ohi);
          } else {
            rlo = max(rlo, olo - a.maxoff);
            rhi = min(rhi, ohi + a.maxoff);
          }
          list = lane * 8 + x;
          entry = entry_pack(t, g.tz0, rlo, rhi);
        }
      } else if (lane == 0) {
        // the stream reads only the input planes holding source voxels
        // within reach of the output planes that need the predicate; tiles
        // with most of their planes in reach go to the front of the list
        // (they start first, mostly one per CU), the others to the back
        const int lo = max(8 * slo, olo - a.maxoff), hi = min(8 * shi + 7, ohi + a.maxoff);
        const int span = g.zlast - g.zfirst + 1 + 2 * a.maxoff;
        list = (4 * (hi - lo + 1) >= 3 * span) ? x : 8 + x;
        entry = entry_pack(t, g.tz0, lo, hi);
      }
    }
  }
  if (list >= 0) slot = atomicAdd(&wgCount[list], 1u);
  __syncthreads();
  u32* qc = a.rvar ? a.queue + kVarCounters + 256 * a.qsel : a.queue + 16 * a.qsel;
  for (int i = threadIdx.x; i < 256; i += blockDim.x)
    if (wgCount[i]) wgBase[i] = atomicAdd(&qc[i], wgCount[i]);
  __syncthreads();
  if (list >= 0) {
    if (a.rvar) {
      a.queue[a.vqBase + (size_t)list * a.vqCap + wgBase[list] + slot] = entry;
    } else {
      const int x = list & 7;
      const u32 pos = list < 8 ? wgBase[list] + slot : (u32)a.queueCap - 1u - (wgBase[list] + slot);
      a.queue[kQueueList + x * a.queueCap + pos] = entry;
    }
  }
}

// ---------------------------------------------------------------------------
// Many jobs in one launch. A batched kernel gets the per-job arguments as a
// table in device memory (a job's entry is exactly its single-job launch
// argument) and a prefix array pre[0..n] of the per-job work units (tiles,
// blocks, words); unit u belongs to the job j with pre[j] <= u < pre[j+1].

// Batched queue entries: job in the high kBatchJobBits bits, tile below.
constexpr int kBatchTileBits = 20;
constexpr int kBatchJobBits = 12;

// Batched tile selection (uniform radius per job): the tile of job j is
// queued under the job's radius value v = rad[j] (per-value counters at
// kVarCounters + 256*qsel + 8v + x, lists at listBase + (8v + x)*cap).
template <int PASS>
__device__ void select_tiles_batched(const MorphArgs* __restrict__ tab, const u32* __restrict__ pre, int n,
                                     const int* __restrict__ rad, u32* queue, int qsel, int listBase, int cap) {
  // one reservation per workgroup and list, as in select_tiles
  __shared__ u32 wgCount[256], wgBase[256];
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < 256; i += blockDim.x) queue[kVarCounters + 256 * (1 - qsel) + i] = 0u;
  for (int i = threadIdx.x; i < 256; i += blockDim.x) wgCount[i] = 0u;
  __syncthreads();
  const u32 u = blockIdx.x * (blockDim.x >> 6) + (u32)wave_id();
  int list = -1;
  u32 slot = 0u, entry = 0u;
  if (u < pre[n]) {
    const int j = batch_job(pre, n, u);
    const MorphArgs& a = tab[j];
    const int t = (int)(u - pre[j]);
    const TileGeo g = pass_tile<PASS>(a, t);
    u32 vals;
    int blo, bhi, slo, shi, olo, ohi;
    if (select_tile<PASS>(a, t, g, vals, blo, bhi, slo, shi, olo, ohi) && lane_id() == 0) {
      list = rad[j] * 8 + slab_queue(t / a.W);
      entry = ((u32)j << kBatchTileBits) | (u32)t;
    }
  }
  if (list >= 0) slot = atomicAdd(&wgCount[list], 1u);
  __syncthreads();
  u32* qc = queue + kVarCounters + 256 * qsel;
  for (int i = threadIdx.x; i < 256; i += blockDim.x)
    if (wgCount[i]) wgBase[i] = atomicAdd(&qc[i], wgCount[i]);
  __syncthreads();
  if (list >= 0) queue[listBase + (size_t)list * cap + wgBase[list] + slot] = entry;
}

// The predicate of one active tile and its epilogue. RHO > 0: uniform radius
// with the compiled chord table of that radius; RHO = 0: runtime tables
// (uniform or per-voxel radii, sphere or cube). PASS 4 and 5 are passes 0 and
// 1 (dilation) with the cube's compiled table.
template <int PASS, int RHO>
__device__ __forceinline__ void run_pass_tile(const MorphArgs& a, u32 entry, u32* smem, u32* set4) {
  constexpr int P = PASS & 3;
  constexpr int KIND = PASS >= 4 ? SE_BOX : -1;
  TileGeo g = pass_tile<P>(a, entry_tile(entry));
  // per-voxel radii: a dilation entry holds the output planes of its value;
  // otherwise entries hold the input planes with source voxels
  entry_range(entry, g, (P < 2 && a.rvar) ? 1 : 0);
  Lds L = carve(smem, a, g.nout());
  if constexpr (P == 0) {
    DilCoarseFin fin{&a, g.wx, g.yfirst + 2 * lane_id(), cell_range_mask(g.wx, 0, a.dR0), 0};
    run_tile<0, 2, RHO, KIND>(a, L, set4, g, a.I, false, fin);
  } else if constexpr (P == 1) {
    DilFineFin fin{&a, g.wx, g.yfirst + lane_id()};
    run_tile<0, 1, RHO, KIND>(a, L, set4, g, a.I, false, fin);
  } else if constexpr (P == 2) {
    EroCoarseFin fin{&a, g.wx, g.yfirst + 2 * lane_id(), cell_range_mask(g.wx, 0, a.dR0)};
    run_tile<1, 2, RHO>(a, L, set4, g, a.K, false, fin);
  } else {
    EroFineFin fin{&a, g.wx, g.yfirst + lane_id()};
    run_tile<2, 1, RHO>(a, L, set4, g, a.K, false, fin);
  }
}

// Persistent workgroups over the active tiles. Workgroup i serves queue
// i % 8 (the grid is a multiple of 8, so every queue has workgroups and is
// drained); workgroups are dealt round-robin to the 8 XCDs, so a queue's
// tiles stay in one L2. Tiles are not taken from other queues: the word
// tiles of a slab write 4-byte pieces of the same 64-byte row segments and
// run several times slower when their slab is spread over XCDs. A workgroup
// ends when its queue is drained, so every wave reaches the exit.
template <int PASS, int RHO>
__device__ void run_pass(const MorphArgs& a) {
  extern __shared__ u32 smem[];
  __shared__ u32 set4[4];
  const int x = (int)(blockIdx.x & 7u);
  const u32* qc = a.queue + a.qcOff;
  const u32* ql = a.queue + a.qlOff + (size_t)x * a.qCap;
  const u32 front = qc[x], n = front + (a.qBack ? qc[8 + x] : 0u);
  for (u32 k = blockIdx.x >> 3; k < n; k += gridDim.x >> 3) {
    run_pass_tile<PASS, RHO>(a, k < front ? ql[k] : ql[a.qCap - 1 - (k - front)], smem, set4);
    __syncthreads();  // LDS is reused by the next tile
  }
}

// Batched form of run_pass: the entries name (job, tile) and the job's
// arguments come from the table; same queue walk.
template <int PASS, int RHO>
__device__ void run_pass_batched(const MorphArgs* __restrict__ tab, const u32* __restrict__ queue, int qcOff,
                                 int qlOff, int qCap) {
  extern __shared__ u32 smem[];
  __shared__ u32 set4[4];
  const int x = (int)(blockIdx.x & 7u);
  const u32* qc = queue + qcOff;
  const u32* ql = queue + qlOff + (size_t)x * qCap;
  for (u32 k = blockIdx.x >> 3; k < qc[x]; k += gridDim.x >> 3) {
    const u32 e = __builtin_amdgcn_readfirstlane(ql[k]);
    const MorphArgs& a = tab[e >> kBatchTileBits];
    const int t = (int)(e & ((1u << kBatchTileBits) - 1u));
    run_pass_tile<PASS, RHO>(a, entry_pack(t, pass_tile<PASS>(a, t).tz0, -(1 << 20), 1 << 20), smem, set4);
    __syncthreads();
  }
}

}  // namespace
}  // namespace gm
