// Ball morphology for radii beyond the chord tables (32..127 fine voxels).
//
// The chord engine (gm_morph_dev.h) reads three words per row and keeps the
// x-windows of up to 32 voxels in LDS, so it stops at radius 31. The
// reference's predicates (GMorpho.cu:2828-3056 dilation, 4349-4648 and
// 5041-5374 erosion; integer form in the oracle) are sums of per-axis
// squared distances from a query point to the box of a source voxel, so they
// separate by axis: a distance transform computed axis by axis gives, at
// every query point, the smallest value over all sources, and each pass
// compares it with its threshold. Cost does not depend on the radius beyond
// the window each axis scans (the largest radius present).
//
//   dilation (gather, radius of the query point p, fine units):
//     F(p) = min_{q in I} sum_a dist(p_a, [q_a, q_a + 1])^2 < (r(p) + 1)^2
//     and, for the cube (GMorpho.cu:1838-1934), with the largest per-axis
//     distance in place of the sum: F(p) = min_q max_a dist(...) <= r(p)
//   erosion (scatter, radius of the source, doubled units so that the
//   sub-voxel centres are integers):
//     E(Q) = min_{q in K} [sum_a dist(Q_a, [2 q_a, 2 q_a + 2])^2 - 4 (r(q) - 1)^2] < 0
//     at Q = 2(2c + 1) (coarse, cell centres) and Q = 2p + 1 (fine, sub-voxel
//     centres).
//
// Each axis stage takes, for every query position X_i = a i + b and every
// source index j within reach, dist(X_i, [s j, s j + s])^2 plus the previous
// stage's value at j (s = 1 for the dilation's corner geometry, 2 for the
// erosion's doubled units). The first stage reads the bit grid: the nearest
// source on each side (binary sources), or every source in reach with its
// own weight (erosion with per-voxel radii). Values are exact integers.
#include "gm_internal.h"

namespace gm {

namespace {

constexpr int kInf = 1 << 29;

// dist(X, [s j, s j + s])
__device__ __forceinline__ int seg_dist(int X, int j, int s) {
  const int lo = s * j, hi = lo + s;
  return X < lo ? lo - X : (X > hi ? X - hi : 0);
}

// Highest set bit index <= j in the row (or -1), looking back at most `reach`
// positions.
__device__ __forceinline__ int last_set_at_or_below(const u32* __restrict__ row, int j, int reach) {
  const int stop = j - reach;
  int w = j >> 5;
  u32 m = row[w] & (0xffffffffu >> (31 - (j & 31)));
  while (true) {
    if (m) {
      const int k = 32 * w + 31 - __clz((int)m);
      return k >= stop ? k : -1;
    }
    if (--w < 0 || 32 * w + 31 < stop) return -1;
    m = row[w];
  }
}
// Lowest set bit index >= j (j may be N) within `reach`, or -1.
__device__ __forceinline__ int first_set_at_or_above(const u32* __restrict__ row, int W, int j, int reach) {
  const int stop = j + reach;
  int w = j >> 5;
  if (w >= W) return -1;
  u32 m = row[w] & (0xffffffffu << (j & 31));
  while (true) {
    if (m) {
      const int k = 32 * w + __ffs((int)m) - 1;
      return k <= stop ? k : -1;
    }
    if (++w >= W || 32 * w > stop) return -1;
    m = row[w];
  }
}

// Stage x over the bit grid: out[(z N + y) nq + i] for every fine row (y, z)
// and query i. WEIGHTED: min over the row's sources in reach of
// dist^2 - 4 (r(q) - 1)^2 (erosion with per-voxel radii); otherwise the
// squared distance to the nearest source.
template <bool WEIGHTED, bool CHEB>
__global__ void __launch_bounds__(256) k_edt_x(const u32* __restrict__ src, int N, int W, int s, int a, int b,
                                               int nq, int reach, RadiusField rf, int* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)N * N * nq;
  if (t >= total) return;
  const int i = (int)(t % (size_t)nq);
  const size_t rowid = t / (size_t)nq;  // z * N + y
  const u32* row = src + rowid * (size_t)W;
  const int X = a * i + b;
  // sources j whose segment can be within reach (in source units)
  const int jx = X / s;  // X >= 0: the segment starting at or before X
  int best = kInf;
  if constexpr (!WEIGHTED) {
    const int jl = last_set_at_or_below(row, min(jx, N - 1), reach + 1);
    if (jl >= 0) {
      const int d = seg_dist(X, jl, s);
      best = min(best, CHEB ? d : d * d);
    }
    const int jr = first_set_at_or_above(row, W, jx + 1, reach + 1);
    if (jr >= 0) {
      const int d = seg_dist(X, jr, s);
      best = min(best, CHEB ? d : d * d);
    }
  } else {
    const int y = (int)(rowid % (size_t)N), z = (int)(rowid / (size_t)N);
    const int j0 = max(0, jx - reach - 1), j1 = min(N - 1, jx + reach + 1);
    for (int w = j0 >> 5; w <= (j1 >> 5); ++w) {
      u32 m = row[w];
      if (32 * w < j0) m &= 0xffffffffu << (j0 - 32 * w);
      if (32 * w + 31 > j1) m &= 0xffffffffu >> (32 * w + 31 - j1);
      while (m) {
        const int j = 32 * w + __ffs((int)m) - 1;
        m &= m - 1u;
        const int d = seg_dist(X, j, s);
        const int rr = radius_at(rf, j, y, z) - 1;
        best = min(best, d * d - 4 * rr * rr);
      }
    }
  }
  out[t] = best;
}

// Stage y or z: out[(o nq + k) inner + ii] = min over sources j in reach of
// in[(o nsrc + j) inner + ii] + dist(a k + b, [s j, s j + s])^2 (CHEB: the
// larger of the two, the cube's per-axis distance bound). A thread takes
// kEdtBlock consecutive queries k of one (o, ii): each source value is loaded
// once for all of them (loads per query fall by the block size; the
// minimum-plus work per query is unchanged).
constexpr int kEdtBlock = 8;
template <bool CHEB>
__global__ void __launch_bounds__(256) k_edt_axis(const int* __restrict__ in, int* __restrict__ out, int nouter,
                                                  int nsrc, int nq, long long inner, int s, int a, int b,
                                                  int reach) {
  const int nkb = (nq + kEdtBlock - 1) / kEdtBlock;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)nouter * nkb * (size_t)inner;
  if (t >= total) return;
  const size_t ii = t % (size_t)inner;
  const size_t ok = t / (size_t)inner;
  const int k0 = (int)(ok % (size_t)nkb) * kEdtBlock, o = (int)(ok / (size_t)nkb);
  const int k1 = min(nq - 1, k0 + kEdtBlock - 1);
  const int j0 = max(0, (a * k0 + b) / s - reach - 1), j1 = min(nsrc - 1, (a * k1 + b) / s + reach + 1);
  const int* p = in + ((size_t)o * nsrc) * (size_t)inner + ii;
  int best[kEdtBlock];
#pragma unroll
  for (int q = 0; q < kEdtBlock; ++q) best[q] = kInf;
  for (int j = j0; j <= j1; ++j) {
    const int v = p[(size_t)j * (size_t)inner];
    if (v >= kInf) continue;
#pragma unroll
    for (int q = 0; q < kEdtBlock; ++q) {
      const int d = seg_dist(a * (k0 + q) + b, j, s);
      best[q] = min(best[q], CHEB ? max(v, d) : v + d * d);
    }
  }
  int* po = out + ((size_t)o * nq + k0) * (size_t)inner + ii;
#pragma unroll
  for (int q = 0; q < kEdtBlock; ++q)
    if (k0 + q <= k1) po[(size_t)q * (size_t)inner] = best[q];
}

// ---- pass outputs (the same contracts as the chord kernels' epilogues)

// DF(c) = I-full(c) | (c in the data region & F(2c)); one thread per
// cell-row word (16 cells).
// cube: dt is the per-axis (Chebyshev) distance, F(p) = dt <= r(p)
__global__ void __launch_bounds__(256) k_edt_dil_coarse(const int* __restrict__ dt, MorphDims d, RadiusField rf,
                                                        int cube, const u16* __restrict__ ifl, u16* __restrict__ df) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)d.R * d.R * d.W;
  if (t >= total) return;
  const int w = (int)(t % d.W);
  const int cy = (int)((t / d.W) % d.R), cz = (int)(t / ((size_t)d.W * d.R));
  u32 f = 0u;
  if (cy < d.dR[1] && cz < d.dR[2]) {
    for (int j = 0; j < 16; ++j) {
      const int cx = 16 * w + j;
      if (cx >= d.dR[0]) break;
      const int px = 2 * cx, py = 2 * cy, pz = 2 * cz;
      const int r = radius_at(rf, px, py, pz), v = dt[((size_t)pz * d.N + py) * d.N + px];
      if (cube ? v <= r : v < (r + 1) * (r + 1)) f |= 1u << j;
    }
  }
  df[t] = (u16)(ifl[t] | f);
}

// D |= expand(TG) & F on the fine voxels of tagged cells; one thread per
// fine row word.
__global__ void __launch_bounds__(256) k_edt_dil_fine(const int* __restrict__ dt, MorphDims d, RadiusField rf,
                                                      int cube, const u16* __restrict__ tg, u32* __restrict__ D) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)d.N * d.N * d.W;
  if (t >= total) return;
  const int w = (int)(t % d.W);
  const int y = (int)((t / d.W) % d.N), z = (int)(t / ((size_t)d.W * d.N));
  const u32 tags = tg[((size_t)(z >> 1) * d.R + (y >> 1)) * d.W + w];
  if (!tags) return;
  u32 v = 0u;
  for (int j = 0; j < 32; ++j) {
    if (!((tags >> (j >> 1)) & 1u)) continue;
    const int x = 32 * w + j;
    const int r = radius_at(rf, x, y, z), e = dt[((size_t)z * d.N + y) * d.N + x];
    if (cube ? e <= r : e < (r + 1) * (r + 1)) v |= 1u << j;
  }
  if (v) D[t] |= v;
}

// CL(c) = c in the data region & D-any(c) & !I-full(c) & E(2(2c+1)); et is
// the transform at the cell centres (R^3, x fastest), thr its threshold
// (0 for weighted sources).
__global__ void __launch_bounds__(256) k_edt_ero_coarse(const int* __restrict__ et, int thr, MorphDims d,
                                                        const u16* __restrict__ dal, const u16* __restrict__ ifl,
                                                        u16* __restrict__ cl) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)d.R * d.R * d.W;
  if (t >= total) return;
  const int w = (int)(t % d.W);
  const int cy = (int)((t / d.W) % d.R), cz = (int)(t / ((size_t)d.W * d.R));
  u32 e = 0u;
  if (cy < d.dR[1] && cz < d.dR[2]) {
    for (int j = 0; j < 16; ++j) {
      const int cx = 16 * w + j;
      if (cx >= d.dR[0]) break;
      if (et[((size_t)cz * d.R + cy) * d.R + cx] < thr) e |= 1u << j;
    }
  }
  cl[t] = (u16)(e & dal[t] & ~ifl[t]);
}

// C &= ~(expand(UG) & E) at the sub-voxel centres (et: N^3, x fastest).
__global__ void __launch_bounds__(256) k_edt_ero_fine(const int* __restrict__ et, int thr, MorphDims d,
                                                      const u16* __restrict__ ug, u32* __restrict__ C) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)d.N * d.N * d.W;
  if (t >= total) return;
  const int w = (int)(t % d.W);
  const int y = (int)((t / d.W) % d.N), z = (int)(t / ((size_t)d.W * d.N));
  const u32 tags = ug[((size_t)(z >> 1) * d.R + (y >> 1)) * d.W + w];
  if (!tags) return;
  u32 v = 0u;
  for (int j = 0; j < 32; ++j) {
    if (!((tags >> (j >> 1)) & 1u)) continue;
    if (et[((size_t)z * d.N + y) * d.N + 32 * w + j] < thr) v |= 1u << j;
  }
  if (v) C[t] &= ~v;
}

unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

size_t edt_scratch_ints(int N) { return (size_t)N * N * N; }

// Transform of the bit grid src at the query lattice X = a i + b (nq per
// axis) into out (nq^3, x fastest); tmp holds N^2 nq ints. weighted: per-voxel
// erosion weights from rf.
hipError_t launch_edt(const u32* src, int N, int s, int a, int b, int nq, int reach, bool weighted, bool cheb,
                      const RadiusField& rf, int* tmp, int* out, hipStream_t st) {
  const int W = N / 32;
  const size_t n1 = (size_t)N * N * nq;
  if (weighted)
    hipLaunchKernelGGL((k_edt_x<true, false>), dim3(blocks_for(n1)), dim3(256), 0, st, src, N, W, s, a, b, nq, reach,
                       rf, out);
  else if (cheb)
    hipLaunchKernelGGL((k_edt_x<false, true>), dim3(blocks_for(n1)), dim3(256), 0, st, src, N, W, s, a, b, nq, reach,
                       rf, out);
  else
    hipLaunchKernelGGL((k_edt_x<false, false>), dim3(blocks_for(n1)), dim3(256), 0, st, src, N, W, s, a, b, nq,
                       reach, rf, out);
  GM_CHECK(hipGetLastError());
  // y: [z][j][i] -> [z][k][i], kEdtBlock queries k per thread
  const size_t nkb = (size_t)((nq + kEdtBlock - 1) / kEdtBlock);
  const size_t n2 = (size_t)N * nkb * nq;
  if (cheb)
    hipLaunchKernelGGL(k_edt_axis<true>, dim3(blocks_for(n2)), dim3(256), 0, st, out, tmp, N, N, nq, (long long)nq, s,
                       a, b, reach);
  else
    hipLaunchKernelGGL(k_edt_axis<false>, dim3(blocks_for(n2)), dim3(256), 0, st, out, tmp, N, N, nq, (long long)nq,
                       s, a, b, reach);
  GM_CHECK(hipGetLastError());
  // z: [j][k][i] -> [k'][k][i]
  const size_t n3 = nkb * (size_t)nq * nq;
  if (cheb)
    hipLaunchKernelGGL(k_edt_axis<true>, dim3(blocks_for(n3)), dim3(256), 0, st, tmp, out, 1, N, nq,
                       (long long)nq * nq, s, a, b, reach);
  else
    hipLaunchKernelGGL(k_edt_axis<false>, dim3(blocks_for(n3)), dim3(256), 0, st, tmp, out, 1, N, nq,
                       (long long)nq * nq, s, a, b, reach);
  return hipGetLastError();
}

hipError_t launch_edt_dil_coarse(const int* dt, const MorphDims& d, const RadiusField& rf, bool cube, const u16* ifl,
                                 u16* df, hipStream_t st) {
  hipLaunchKernelGGL(k_edt_dil_coarse, dim3(blocks_for((size_t)d.R * d.R * d.W)), dim3(256), 0, st, dt, d, rf,
                     cube ? 1 : 0, ifl, df);
  return hipGetLastError();
}
hipError_t launch_edt_dil_fine(const int* dt, const MorphDims& d, const RadiusField& rf, bool cube, const u16* tg,
                               u32* D, hipStream_t st) {
  hipLaunchKernelGGL(k_edt_dil_fine, dim3(blocks_for((size_t)d.N * d.N * d.W)), dim3(256), 0, st, dt, d, rf,
                     cube ? 1 : 0, tg, D);
  return hipGetLastError();
}
hipError_t launch_edt_ero_coarse(const int* et, int thr, const MorphDims& d, const u16* dal, const u16* ifl, u16* cl,
                                 hipStream_t st) {
  hipLaunchKernelGGL(k_edt_ero_coarse, dim3(blocks_for((size_t)d.R * d.R * d.W)), dim3(256), 0, st, et, thr, d, dal,
                     ifl, cl);
  return hipGetLastError();
}
hipError_t launch_edt_ero_fine(const int* et, int thr, const MorphDims& d, const u16* ug, u32* C, hipStream_t st) {
  hipLaunchKernelGGL(k_edt_ero_fine, dim3(blocks_for((size_t)d.N * d.N * d.W)), dim3(256), 0, st, et, thr, d, ug, C);
  return hipGetLastError();
}

}  // namespace gm
