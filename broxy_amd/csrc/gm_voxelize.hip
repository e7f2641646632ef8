// Voxelizer: triangle mesh -> bit grid at N = 2R fine voxels per axis.
//
// Restates Voxelizer::VoxelizeConservative (Voxelizer.cu:724-796) without
// OpenGL. The reference draws the mesh twice into an N x N x N/32 image:
//  * volume pass (vol_vox.vert, vol_vox.geo, vol_vox_bit.frag:35-52):
//    orthographic along z; every pixel centre a triangle covers XORs the
//    column of voxels at or below the triangle's depth there -> parity fill;
//  * surface pass (surf_vox_conserv.geo:41-122, surf_vox_conserv_bit.frag:35-47):
//    each triangle whose squared aspect ratio is below 50^2 is rasterised
//    conservatively on its dominant plane, and the voxel under the plane's
//    depth at each touched pixel centre is set.
// Rasterisation here is exact integer arithmetic on vertices snapped to
// 1/256 voxel: pixel centres use a top-left tie rule, conservative coverage
// is the closed triangle/pixel-square overlap, depths are floored plane
// values. The parity pass writes one toggle bit per covered pixel at the
// triangle's depth; a suffix XOR along z then yields the filled columns.
#include <algorithm>

#include "gm_internal.h"

namespace gm {

namespace {

typedef long long i64;

// floor(a / 256): an arithmetic shift (the snap grid has 256 steps per voxel).
__device__ __forceinline__ i64 fdiv256(i64 a) { return a >> 8; }

// floor(a / b) for b > 0, with rb = 1 / b in double: the rounded quotient is
// corrected against the exact remainder, so the result is the integer floor.
// (A 64-bit integer division is a long software sequence on the GPU; here it
// ran once per covered pixel.) The quotients are voxel depths (|q| < 2^13)
// and |a| < 2^62, so the double estimate is off by less than one and the
// loops run at most once.
__device__ __forceinline__ i64 fdiv_r(i64 a, i64 b, double rb) {
  i64 q = (i64)floor((double)a * rb);
  i64 r = a - q * b;
  while (r < 0) {
    --q;
    r += b;
  }
  while (r >= b) {
    ++q;
    r -= b;
  }
  return q;
}

__global__ void k_snap(const float* __restrict__ P, int nv, float bx, float by, float bz, double sc,
                       int32_t* __restrict__ X) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const float b[3] = {bx, by, bz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double v = __dmul_rn(__dsub_rn((double)P[3 * i + a], (double)b[a]), sc);
    X[3 * i + a] = (int32_t)floor(__dadd_rn(__dmul_rn(v, 256.0), 0.5));
  }
}

__device__ __forceinline__ void toggle(u32* g, int N, int W, int x, int y, int z) {
  atomicXor(&g[((size_t)z * N + y) * W + (x >> 5)], 1u << (x & 31));
}
__device__ __forceinline__ void setbit(u32* g, int N, int W, int x, int y, int z) {
  atomicOr(&g[((size_t)z * N + y) * W + (x >> 5)], 1u << (x & 31));
}

__global__ void k_vox_tris(const int32_t* __restrict__ X, const u32* __restrict__ T, int nt, int N,
                           u32* __restrict__ mark, u32* __restrict__ shell) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const i64 S = 256;
  const int W = N / 32;
  i64 A[3], B[3], C[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    A[a] = X[3 * (size_t)T[3 * (size_t)t] + a];
    B[a] = X[3 * (size_t)T[3 * (size_t)t + 1] + a];
    C[a] = X[3 * (size_t)T[3 * (size_t)t + 2] + a];
  }
  // ---- volume pass
  {
    i64 a0 = A[0], a1 = A[1], a2 = A[2];
    i64 b0 = B[0], b1 = B[1], b2 = B[2];
    i64 c0 = C[0], c1 = C[1], c2 = C[2];
    i64 area = (b0 - a0) * (c1 - a1) - (b1 - a1) * (c0 - a0);
    if (area != 0) {
      if (area < 0) {
        i64 t0 = b0, t1 = b1, t2 = b2;
        b0 = c0; b1 = c1; b2 = c2;
        c0 = t0; c1 = t1; c2 = t2;
        area = -area;
      }
      i64 xmn = min(a0, min(b0, c0)), xmx = max(a0, max(b0, c0));
      i64 ymn = min(a1, min(b1, c1)), ymx = max(a1, max(b1, c1));
      i64 i0 = max((i64)0, fdiv256(xmn - S / 2 + S - 1)), i1 = min((i64)N - 1, fdiv256(xmx - S / 2));
      i64 j0 = max((i64)0, fdiv256(ymn - S / 2 + S - 1)), j1 = min((i64)N - 1, fdiv256(ymx - S / 2));
      const i64 dz = area * S;
      const double rdz = 1.0 / (double)dz;
      const i64 vx[3] = {a0, b0, c0}, vy[3] = {a1, b1, c1}, vz[3] = {a2, b2, c2};
      for (i64 j = j0; j <= j1; ++j)
        for (i64 i = i0; i <= i1; ++i) {
          i64 px = i * S + S / 2, py = j * S + S / 2;
          i64 w[3];
          bool in = true;
#pragma unroll
          for (int e = 0; e < 3; ++e) {
            int e1 = e == 2 ? 0 : e + 1;
            i64 ex = vx[e1] - vx[e], ey = vy[e1] - vy[e];
            i64 f = ex * (py - vy[e]) - ey * (px - vx[e]);
            bool tl = (ey < 0) || (ey == 0 && ex < 0);
            if (f < 0 || (f == 0 && !tl)) in = false;
            w[e == 0 ? 2 : e - 1] = f;
          }
          if (!in) continue;
          i64 num = w[0] * vz[0] + w[1] * vz[1] + w[2] * vz[2];
          i64 zc = fdiv_r(num, dz, rdz);
          if (zc < 0) continue;
          if (zc > N - 1) zc = N - 1;
          toggle(mark, N, W, (int)i, (int)j, (int)zc);
        }
    }
  }
  // ---- surface pass
  {
    i64 e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    i64 e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
    i64 n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    typedef unsigned __int128 u128;
    u128 a2 = (u128)((__int128)n[0] * n[0]) + (u128)((__int128)n[1] * n[1]) + (u128)((__int128)n[2] * n[2]);
    if (a2 == 0) return;
    i64 d0[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
    i64 d1[3] = {B[0] - C[0], B[1] - C[1], B[2] - C[2]};
    i64 l0 = d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2];
    i64 l1 = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
    i64 l2 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
    u128 ml = (u128)max(l0, max(l1, l2));
    if (!(ml * ml < (u128)2500 * a2)) return;
    i64 an0 = n[0] < 0 ? -n[0] : n[0], an1 = n[1] < 0 ? -n[1] : n[1], an2 = n[2] < 0 ? -n[2] : n[2];
    i64 mxn = max(an0, max(an1, an2));
    int ax = an0 == mxn ? 0 : (an1 == mxn ? 1 : 2);
    int ua = ax == 0 ? 1 : (ax == 1 ? 2 : 0);
    int va = ax == 0 ? 2 : (ax == 1 ? 0 : 1);
    i64 U[3] = {A[ua], B[ua], C[ua]}, Vv[3] = {A[va], B[va], C[va]};
    i64 area = (U[1] - U[0]) * (Vv[2] - Vv[0]) - (Vv[1] - Vv[0]) * (U[2] - U[0]);
    if (area < 0) {
      i64 t0 = U[1], t1 = Vv[1];
      U[1] = U[2]; Vv[1] = Vv[2];
      U[2] = t0; Vv[2] = t1;
    }
    i64 umn = min(U[0], min(U[1], U[2])), umx = max(U[0], max(U[1], U[2]));
    i64 vmn = min(Vv[0], min(Vv[1], Vv[2])), vmx = max(Vv[0], max(Vv[1], Vv[2]));
    i64 iu0 = max((i64)0, fdiv256(umn - 1)), iu1 = min((i64)N - 1, fdiv256(umx));
    i64 iv0 = max((i64)0, fdiv256(vmn - 1)), iv1 = min((i64)N - 1, fdiv256(vmx));
    i64 nd = n[ax], nu = n[ua], nvv = n[va];
    i64 Ad = A[ax], Au = A[ua], Av = A[va];
    // depth q = floor(num / (nd S)), taken with a positive divisor
    const i64 sgn = nd > 0 ? 1 : -1, dq = sgn * nd * S;
    const double rdq = 1.0 / (double)dq;
    for (i64 iv = iv0; iv <= iv1; ++iv)
      for (i64 iu = iu0; iu <= iu1; ++iu) {
        i64 ulo = iu * S, uhi = ulo + S, vlo = iv * S, vhi = vlo + S;
        bool ok = true;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
          int e1i = e == 2 ? 0 : e + 1;
          i64 eu = U[e1i] - U[e], ev = Vv[e1i] - Vv[e];
          i64 pu = (-ev > 0) ? uhi : ulo;
          i64 pv = (eu > 0) ? vhi : vlo;
          i64 f = eu * (pv - Vv[e]) - ev * (pu - U[e]);
          if (f < 0) ok = false;
        }
        if (!ok) continue;
        i64 pu = ulo + S / 2, pv = vlo + S / 2;
        i64 num = Ad * nd - (nu * (pu - Au) + nvv * (pv - Av));
        i64 q = fdiv_r(sgn * num, dq, rdq);
        if (q < 0 || q >= N) continue;
        int c3[3];
        c3[ua] = (int)iu;
        c3[va] = (int)iv;
        c3[ax] = (int)q;
        setbit(shell, N, W, c3[0], c3[1], c3[2]);
      }
  }
}

// MultiSlicing (Voxelizer::VoxelizeByMultiSlicing, Voxelizer.cu:188-331):
// rays along +x through the fine column centres (y + 1/2, z + 1/2); a
// triangle covering a centre under the top-left rule toggles the column from
// the voxel holding its crossing on, as one toggle bit at that voxel (the fill
// below takes the running XOR along x). No surface shell. Same snapped
// vertices and exact integer rule as the conservative path.
__global__ void k_vox_ms_tris(const int32_t* __restrict__ X, const u32* __restrict__ T, int nt, int N,
                              u32* __restrict__ mark) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const i64 S = 256;
  const int W = N / 32;
  i64 A[3], B[3], C[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    A[a] = X[3 * (size_t)T[3 * (size_t)t] + a];
    B[a] = X[3 * (size_t)T[3 * (size_t)t + 1] + a];
    C[a] = X[3 * (size_t)T[3 * (size_t)t + 2] + a];
  }
  i64 area = (B[1] - A[1]) * (C[2] - A[2]) - (B[2] - A[2]) * (C[1] - A[1]);
  if (area == 0) return;
  i64 vy[3] = {A[1], B[1], C[1]}, vz[3] = {A[2], B[2], C[2]}, vx[3] = {A[0], B[0], C[0]};
  if (area < 0) {  // counter-clockwise in (y, z)
    i64 t0 = vy[1], t1 = vz[1], t2 = vx[1];
    vy[1] = vy[2]; vz[1] = vz[2]; vx[1] = vx[2];
    vy[2] = t0; vz[2] = t1; vx[2] = t2;
    area = -area;
  }
  const i64 ymn = min(vy[0], min(vy[1], vy[2])), ymx = max(vy[0], max(vy[1], vy[2]));
  const i64 zmn = min(vz[0], min(vz[1], vz[2])), zmx = max(vz[0], max(vz[1], vz[2]));
  const i64 j0 = max((i64)0, fdiv256(ymn - S / 2 + S - 1)), j1 = min((i64)N - 1, fdiv256(ymx - S / 2));
  const i64 k0 = max((i64)0, fdiv256(zmn - S / 2 + S - 1)), k1 = min((i64)N - 1, fdiv256(zmx - S / 2));
  const i64 dx = area * S;
  const double rdx = 1.0 / (double)dx;
  for (i64 k = k0; k <= k1; ++k)
    for (i64 j = j0; j <= j1; ++j) {
      const i64 py = j * S + S / 2, pz = k * S + S / 2;
      i64 w[3];
      bool in = true;
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        const int e1 = e == 2 ? 0 : e + 1;
        const i64 ey = vy[e1] - vy[e], ez = vz[e1] - vz[e];
        const i64 f = ey * (pz - vz[e]) - ez * (py - vy[e]);
        const bool tl = (ez < 0) || (ez == 0 && ey < 0);
        if (f < 0 || (f == 0 && !tl)) in = false;
        w[e == 0 ? 2 : e - 1] = f;
      }
      if (!in) continue;
      i64 xq = fdiv_r(w[0] * vx[0] + w[1] * vx[1] + w[2] * vx[2], dx, rdx);
      if (xq >= N) continue;
      if (xq < 0) xq = 0;
      atomicXor(&mark[((size_t)k * N + (size_t)j) * W + (size_t)(xq >> 5)], 1u << (xq & 31));
    }
}

// Running XOR of the toggle bits along x, one thread per row (prefix XOR of
// the bits inside a word, then the parity carried from the words before);
// clears the toggles behind it.
__global__ void k_vox_ms_fill(u32* __restrict__ mark, u32* __restrict__ rows, int N) {
  const size_t row = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= (size_t)N * N) return;
  const int W = N / 32;
  u32 carry = 0u;
  for (int w = 0; w < W; ++w) {
    const size_t i = row * W + w;
    u32 v = mark[i];
    mark[i] = 0u;
    v ^= v << 1;
    v ^= v << 2;
    v ^= v << 4;
    v ^= v << 8;
    v ^= v << 16;
    if (carry) v = ~v;
    rows[i] = v;
    carry = v >> 31;
  }
}

// Suffix XOR along z per (row y, word wx) column, then OR with the shell.
// A workgroup owns 64 consecutive columns (one per lane, coalesced) and
// splits z into kFillChunks chunks, one wave each: each wave XORs its chunk,
// the chunk parities give every chunk the XOR of the chunks above it (LDS),
// then each wave walks its chunk down from that value. The walk clears the
// marks and the shell behind it, so they are zero for the next voxelization
// without a separate clearing pass.
constexpr int kFillChunks = 16;
__global__ void __launch_bounds__(64 * kFillChunks) k_vox_fill(u32* __restrict__ mark,
                                                               u32* __restrict__ shell, u32* __restrict__ rows,
                                                               int N) {
  __shared__ u32 par[kFillChunks][64];
  const int W = N / 32;
  const int lane = (int)(threadIdx.x & 63), chunk = (int)(threadIdx.x >> 6);
  const size_t col = (size_t)blockIdx.x * 64 + lane;  // = y*W + wx
  const bool ok = col < (size_t)N * W;
  const size_t plane = (size_t)N * W;
  const int zc = N / kFillChunks, z0 = chunk * zc;
  u32 x = 0u;
  if (ok)
    for (int z = z0; z < z0 + zc; ++z) x ^= mark[(size_t)z * plane + col];
  par[chunk][lane] = x;
  __syncthreads();
  u32 acc = 0u;  // XOR of the chunks above this one
  for (int c = chunk + 1; c < kFillChunks; ++c) acc ^= par[c][lane];
  if (!ok) return;
  for (int z = z0 + zc - 1; z >= z0; --z) {
    const size_t i = (size_t)z * plane + col;
    acc ^= mark[i];
    rows[i] = acc | shell[i];
    mark[i] = 0u;
    shell[i] = 0u;
  }
}

// Float keys that order like the floats (for min / max by integer atomics).
__device__ __forceinline__ u32 fkey(float f) {
  const u32 u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Mesh check before voxelization, in two kernels without atomics (device
// atomics on one cache line from every workgroup serialize across the XCDs:
// 154 us for config 1 with one set per workgroup). k_mesh_check: each of
// kCheckBlocks workgroups reduces its share (wave shuffles, then LDS) into 8
// words at part + 8 * block: [0] some triangle index >= nv, [1] some
// coordinate not finite, [2..4] / [5..7] keys of the per-axis minimum /
// maximum coordinate. k_mesh_check_final reduces the parts into out[0..7].
constexpr int kCheckBlocks = 256;
__device__ __forceinline__ u32 wave_or(u32 v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= (u32)__shfl_xor((int)v, o);
  return v;
}
__device__ __forceinline__ u32 wave_min(u32 v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, (u32)__shfl_xor((int)v, o));
  return v;
}
__device__ __forceinline__ u32 wave_max(u32 v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (u32)__shfl_xor((int)v, o));
  return v;
}
// The 8 check words of this thread reduced over the workgroup (256 threads):
// the result in w[] of thread 0.
__device__ __forceinline__ void block_check_reduce(u32 (&w)[8]) {
  __shared__ u32 red[4][8];
  const int lane = (int)(threadIdx.x & 63), wv = (int)(threadIdx.x >> 6);
  w[0] = wave_or(w[0]);
  w[1] = wave_or(w[1]);
  for (int a = 0; a < 3; ++a) {
    w[2 + a] = wave_min(w[2 + a]);
    w[5 + a] = wave_max(w[5 + a]);
  }
  if (lane == 0)
    for (int k = 0; k < 8; ++k) red[wv][k] = w[k];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 1; q < 4; ++q) {
      w[0] |= red[q][0];
      w[1] |= red[q][1];
      for (int a = 0; a < 3; ++a) {
        w[2 + a] = min(w[2 + a], red[q][2 + a]);
        w[5 + a] = max(w[5 + a], red[q][5 + a]);
      }
    }
}
__global__ void __launch_bounds__(256) k_mesh_check(const float* __restrict__ P, int nv, const u32* __restrict__ T,
                                                     int nt, u32* __restrict__ part) {
  u32 w[8] = {0u, 0u, ~0u, ~0u, ~0u, 0u, 0u, 0u};
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < 3 * (size_t)nt; i += stride)
    w[0] |= T[i] >= (u32)nv ? 1u : 0u;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (size_t)nv; i += stride)
    for (int a = 0; a < 3; ++a) {
      const float v = P[3 * i + a];
      if (!isfinite(v)) w[1] = 1u;
      const u32 k = fkey(v);
      w[2 + a] = min(w[2 + a], k);
      w[5 + a] = max(w[5 + a], k);
    }
  block_check_reduce(w);
  if (threadIdx.x == 0)
    for (int k = 0; k < 8; ++k) part[8 * blockIdx.x + k] = w[k];
}
__global__ void __launch_bounds__(256) k_mesh_check_final(const u32* __restrict__ part, int nparts, u32* out) {
  u32 w[8] = {0u, 0u, ~0u, ~0u, ~0u, 0u, 0u, 0u};
  for (int b = (int)threadIdx.x; b < nparts; b += blockDim.x) {
    const u32* p = part + 8 * b;
    w[0] |= p[0];
    w[1] |= p[1];
    for (int a = 0; a < 3; ++a) {
      w[2 + a] = min(w[2 + a], p[2 + a]);
      w[5 + a] = max(w[5 + a], p[5 + a]);
    }
  }
  block_check_reduce(w);
  if (threadIdx.x == 0)
    for (int k = 0; k < 8; ++k) out[k] = w[k];
}

}  // namespace

size_t mesh_check_words() { return 8 + 8 * (size_t)kCheckBlocks; }

hipError_t launch_voxelize_multislice(const float* P, int nv, const u32* T, int nt, const float* bmin3, double scale,
                                      int R, int32_t* snap, u32* mark, u32* rows, hipStream_t s, bool clear) {
  const int N = 2 * R;
  const size_t words = (size_t)N * N * (N / 32);
  if (clear) GM_CHECK(hipMemsetAsync(mark, 0, words * 4, s));
  if (nv > 0)
    hipLaunchKernelGGL(k_snap, dim3((nv + 255) / 256), dim3(256), 0, s, P, nv, bmin3[0], bmin3[1], bmin3[2],
                       scale, snap);
  GM_CHECK(hipGetLastError());
  if (nt > 0) hipLaunchKernelGGL(k_vox_ms_tris, dim3((nt + 127) / 128), dim3(128), 0, s, snap, T, nt, N, mark);
  GM_CHECK(hipGetLastError());
  const size_t nrows = (size_t)N * N;
  hipLaunchKernelGGL(k_vox_ms_fill, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, s, mark, rows, N);
  return hipGetLastError();
}

hipError_t launch_mesh_check(const float* P, int nv, const u32* T, int nt, u32* out, hipStream_t s) {
  const size_t work = std::max<size_t>(3 * (size_t)nt, (size_t)nv);
  const int blocks = (int)std::max<size_t>(1, std::min<size_t>((work + 255) / 256, kCheckBlocks));
  hipLaunchKernelGGL(k_mesh_check, dim3(blocks), dim3(256), 0, s, P, nv, T, nt, out + 8);
  GM_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_mesh_check_final, dim3(1), dim3(256), 0, s, out + 8, blocks, out);
  return hipGetLastError();
}

hipError_t launch_voxelize(const float* P, int nv, const u32* T, int nt, const float* bmin3, double scale,
                           int R, int32_t* snap, u32* mark, u32* shell, u32* rows, hipStream_t s, bool clear) {
  const int N = 2 * R, W = N / 32;
  size_t words = (size_t)N * N * W;
  if (clear) {
    GM_CHECK(hipMemsetAsync(mark, 0, words * 4, s));
    GM_CHECK(hipMemsetAsync(shell, 0, words * 4, s));
  }
  if (nv > 0)
    hipLaunchKernelGGL(k_snap, dim3((nv + 255) / 256), dim3(256), 0, s, P, nv, bmin3[0], bmin3[1], bmin3[2],
                       scale, snap);
  GM_CHECK(hipGetLastError());
  if (nt > 0) hipLaunchKernelGGL(k_vox_tris, dim3((nt + 127) / 128), dim3(128), 0, s, snap, T, nt, N, mark, shell);
  GM_CHECK(hipGetLastError());
  // N is a multiple of 32, so each of the kFillChunks chunks has N / 16 planes
  hipLaunchKernelGGL(k_vox_fill, dim3((unsigned)((words / N + 63) / 64)), dim3(64 * kFillChunks), 0, s, mark, shell,
                     rows, N);
  return hipGetLastError();
}

}  // namespace gm
