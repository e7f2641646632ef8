// Bilateral normal filtering of the extracted mesh: the iteration loop of
// GMorpho::ExtractClosingMesh (GMorpho.cu:6946-6978) over the kernels
// BilateralMeshFiltering and BilateralMeshFilteringNormalize
// (GMorpho.cu:6539-6620).
//
// Per iteration and triangle corner v with other corners a, b (corner
// order), the reference adds w(v,a) n_a + w(v,b) n_b to v's sum with float
// atomics, w(v,u) = exp(-(|p_v - p_u|^2 / sigma_p^2 + |n_v - n_u|^2 /
// sigma_n^2)), then normalizes every sum. Atomics make the order of the
// additions, and so the low bits, vary from run to run. Here each vertex
// gathers instead: the corners incident to a vertex are listed once per mesh
// (a counting sort of the 3 |T| corners by vertex, each list then sorted by
// corner id 3 t + c), and one thread per vertex adds its corners' sums in
// triangle order and normalizes. The result is the reference's arithmetic
// with its atomics taken in triangle order, the same every run. The weights
// use an exp of fixed operations (exp_weight) rather than the platform's,
// because the iterations with sigma_n = 0.2 amplify any difference in the
// last bit: with it, the host restatement gives the same bits.
#include <algorithm>

#include "gm_internal.h"

namespace gm {

namespace {

constexpr int kScanBlock = 1024;  // entries per scan workgroup (256 threads x 4)

__global__ void __launch_bounds__(256) k_zero_u32(u32* p, u32 n) {
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// Corners per vertex (deg zeroed; one thread per triangle).
__global__ void __launch_bounds__(256) k_corner_degree(const u32* __restrict__ T, u32 nt, u32 nv, u32* deg) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  for (int c = 0; c < 3; ++c) {
    const u32 v = T[3 * (size_t)t + c];
    if (v < nv) atomicAdd(&deg[v], 1u);
  }
}

// Exclusive scan of n entries in place: per-workgroup totals, one workgroup
// scanning the totals, then each workgroup scanning its entries from its
// offset.
__device__ __forceinline__ u32 wg_exscan(u32 v, u32* lds, u32* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  u32 x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const u32 y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) lds[wave] = x;
  __syncthreads();
  u32 before = 0u, all = 0u;
  for (int w = 0; w < nw; ++w) {
    const u32 s = lds[w];
    if (w < wave) before += s;
    all += s;
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}
__global__ void __launch_bounds__(256) k_scan_totals(const u32* __restrict__ a, u32 n, u32* part) {
  __shared__ u32 lds[4];
  const size_t base = (size_t)blockIdx.x * kScanBlock + 4 * threadIdx.x;
  u32 s = 0u;
  for (int k = 0; k < 4; ++k) s += base + k < n ? a[base + k] : 0u;
  u32 tot;
  wg_exscan(s, lds, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}
__global__ void __launch_bounds__(256) k_scan_parts(u32* part, u32 nparts) {
  __shared__ u32 lds[4];
  u32 carry = 0u;
  for (u32 b = 0; b < nparts; b += blockDim.x) {
    const u32 i = b + threadIdx.x;
    const u32 v = i < nparts ? part[i] : 0u;
    u32 tot;
    const u32 ex = wg_exscan(v, lds, &tot);
    if (i < nparts) part[i] = carry + ex;
    carry += tot;
  }
}
__global__ void __launch_bounds__(256) k_scan_apply(u32* a, u32 n, const u32* __restrict__ part) {
  __shared__ u32 lds[4];
  const size_t base = (size_t)blockIdx.x * kScanBlock + 4 * threadIdx.x;
  u32 v[4], s = 0u;
  for (int k = 0; k < 4; ++k) {
    v[k] = base + k < n ? a[base + k] : 0u;
    s += v[k];
  }
  u32 tot;
  u32 off = wg_exscan(s, lds, &tot) + part[blockIdx.x];
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) a[base + k] = off;
    off += v[k];
  }
}

// Corner ids (3 t + c) into their vertex's list (cursor zeroed).
__global__ void __launch_bounds__(256) k_corner_fill(const u32* __restrict__ T, u32 nt, u32 nv,
                                                     const u32* __restrict__ off, u32* cursor, u32* ent) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  for (int c = 0; c < 3; ++c) {
    const u32 v = T[3 * (size_t)t + c];
    if (v < nv) ent[off[v] + atomicAdd(&cursor[v], 1u)] = 3u * t + (u32)c;
  }
}

// Each vertex's list in corner order (lists are short: insertion sort).
__global__ void __launch_bounds__(256) k_corner_sort(const u32* __restrict__ off, u32 nv, u32* ent) {
  const u32 v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const u32 b = off[v], e = off[v + 1];
  for (u32 i = b + 1; i < e; ++i) {
    const u32 x = ent[i];
    u32 j = i;
    while (j > b && ent[j - 1] > x) {
      ent[j] = ent[j - 1];
      --j;
    }
    ent[j] = x;
  }
}

// exp(x) of a weight (x <= 0) in float: Cody-Waite reduction x = k ln2 + r
// with ln2 split as 355/512 + (ln2 - 355/512) (k ln2_hi is exact), exp(r) by
// its Taylor series to degree 7 in Horner form (|r| <= 0.35), times 2^k; 0
// below -87, where the result would leave the normal range. Explicitly
// rounded operations in a fixed order, so any IEEE float unit gives the same
// bits (the test suite holds the host copy to exp within 3 ulp).
__device__ __forceinline__ float exp_weight(float x) {
  if (!(x >= -87.f)) return 0.f;
  x = fminf(x, 0.f);
  const float k = rintf(__fmul_rn(x, 1.44269504f));
  const float r = __fsub_rn(__fsub_rn(x, __fmul_rn(k, 0.693359375f)), __fmul_rn(k, -2.12194440e-4f));
  const float c[8] = {1.f / 5040.f, 1.f / 720.f, 1.f / 120.f, 1.f / 24.f, 1.f / 6.f, 0.5f, 1.f, 1.f};
  float p = c[0];
#pragma unroll
  for (int i = 1; i < 8; ++i) p = __fadd_rn(__fmul_rn(p, r), c[i]);
  return ldexpf(p, (int)k);
}

__device__ __forceinline__ float sqdist(const float* a, const float* b) {
  const float x = __fsub_rn(a[0], b[0]), y = __fsub_rn(a[1], b[1]), z = __fsub_rn(a[2], b[2]);
  return __fadd_rn(__fadd_rn(__fmul_rn(x, x), __fmul_rn(y, y)), __fmul_rn(z, z));
}

// One iteration: Nout[v] = normalize(sum over v's corners, in triangle order,
// of w(v,a) n_a + w(v,b) n_b). A zero sum stays zero. The same operations in
// the same order as the oracle's restatement, so the results are identical.
__global__ void __launch_bounds__(256) k_bilateral(const float* __restrict__ V, const float* __restrict__ Nin,
                                                   float* __restrict__ Nout, const u32* __restrict__ T,
                                                   const u32* __restrict__ off, const u32* __restrict__ ent,
                                                   u32 nv, float sp2, float sn2) {
  const u32 v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const float pv[3] = {V[3 * (size_t)v], V[3 * (size_t)v + 1], V[3 * (size_t)v + 2]};
  const float nvv[3] = {Nin[3 * (size_t)v], Nin[3 * (size_t)v + 1], Nin[3 * (size_t)v + 2]};
  float acc[3] = {0.f, 0.f, 0.f};
  const u32 e = off[v + 1];
  for (u32 i = off[v]; i < e; ++i) {
    const u32 id = ent[i], t = id / 3u, c = id - 3u * t;
    const u32 a = T[3 * (size_t)t + (c == 0 ? 1 : 0)], b = T[3 * (size_t)t + (c == 2 ? 1 : 2)];
    if (a >= nv || b >= nv) continue;
    const float pa[3] = {V[3 * (size_t)a], V[3 * (size_t)a + 1], V[3 * (size_t)a + 2]};
    const float pb[3] = {V[3 * (size_t)b], V[3 * (size_t)b + 1], V[3 * (size_t)b + 2]};
    const float na[3] = {Nin[3 * (size_t)a], Nin[3 * (size_t)a + 1], Nin[3 * (size_t)a + 2]};
    const float nb[3] = {Nin[3 * (size_t)b], Nin[3 * (size_t)b + 1], Nin[3 * (size_t)b + 2]};
    const float ea = __fadd_rn(__fdiv_rn(sqdist(pv, pa), sp2), __fdiv_rn(sqdist(nvv, na), sn2));
    const float eb = __fadd_rn(__fdiv_rn(sqdist(pv, pb), sp2), __fdiv_rn(sqdist(nvv, nb), sn2));
    const float wa = exp_weight(-ea), wb = exp_weight(-eb);
#pragma unroll
    for (int q = 0; q < 3; ++q) acc[q] = __fadd_rn(acc[q], __fadd_rn(__fmul_rn(wa, na[q]), __fmul_rn(wb, nb[q])));
  }
  // sqrtf is correctly rounded here (the __fsqrt_rn builtin was measured not
  // to be on this toolchain: 15% of a million inputs off by an ulp)
  const float l = sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(acc[0], acc[0]), __fmul_rn(acc[1], acc[1])),
                                  __fmul_rn(acc[2], acc[2])));
#pragma unroll
  for (int q = 0; q < 3; ++q) Nout[3 * (size_t)v + q] = l > 0.f ? __fdiv_rn(acc[q], l) : acc[q];
}

inline unsigned blocks_for(size_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

size_t bilateral_scratch_words(u32 nv, u32 nt) {
  // off (nv + 1), cursor (nv), entries (3 nt), scan totals
  return (size_t)nv + 1 + nv + 3 * (size_t)nt + blocks_for((size_t)nv + 1, kScanBlock);
}

hipError_t launch_bilateral(const float* V, float* Nrm, const u32* T, u32 nv, u32 nt, int iters, u32* scratch,
                            float* Ntmp, hipStream_t s) {
  if (nv == 0 || iters <= 0) return hipSuccess;
  u32* off = scratch;
  u32* cursor = off + nv + 1;
  u32* ent = cursor + nv;
  u32* part = ent + 3 * (size_t)nt;
  const u32 n = nv + 1;
  const unsigned nparts = blocks_for(n, kScanBlock);
  hipLaunchKernelGGL(k_zero_u32, dim3(std::min(blocks_for(2 * (size_t)nv + 1, 256), 4096u)), dim3(256), 0, s, off,
                     2 * nv + 1);
  hipLaunchKernelGGL(k_corner_degree, dim3(blocks_for(nt, 256)), dim3(256), 0, s, T, nt, nv, off);
  hipLaunchKernelGGL(k_scan_totals, dim3(nparts), dim3(256), 0, s, off, n, part);
  hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(256), 0, s, part, nparts);
  hipLaunchKernelGGL(k_scan_apply, dim3(nparts), dim3(256), 0, s, off, n, part);
  hipLaunchKernelGGL(k_corner_fill, dim3(blocks_for(nt, 256)), dim3(256), 0, s, T, nt, nv, off, cursor, ent);
  hipLaunchKernelGGL(k_corner_sort, dim3(blocks_for(nv, 256)), dim3(256), 0, s, off, nv, ent);
  GM_CHECK(hipGetLastError());
  // the reference's schedule: sigma_p = 1; sigma_n = 1e10 for the first two
  // iterations, 0.2 after (squares taken in float, as its kernel does)
  const float sp = 1.f, sp2 = sp * sp;
  const float* src = Nrm;
  for (int it = 0; it < iters; ++it) {
    const float sn = it < 2 ? 1e10f : 0.2f, sn2 = sn * sn;
    float* dst = (it & 1) ? Nrm : Ntmp;
    hipLaunchKernelGGL(k_bilateral, dim3(blocks_for(nv, 256)), dim3(256), 0, s, V, src, dst, T, off, ent, nv, sp2,
                       sn2);
    src = dst;
  }
  GM_CHECK(hipGetLastError());
  if (src != Nrm) GM_CHECK(hipMemcpyAsync(Nrm, src, (size_t)nv * 12, hipMemcpyDeviceToDevice, s));
  return hipSuccess;
}

}  // namespace gm
