// Conversions between the reference's cell-byte grid (the C-ABI format, as
// Grid::voxels holds it, Grid.h:42-59) and the bit-row layout the kernels use.
// One thread moves 16 cells (one 16-byte vector) <-> the four 32-bit row words
// (two fine rows x two fine planes) that hold the same 128 voxels.
#include "gm_internal.h"

namespace gm {

// Spread the low 16 bits of v so bit j lands on bit 2j.
__device__ __forceinline__ u32 spread2(u32 v) {
  v &= 0xffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}
// Inverse of spread2 on the even bits.
__device__ __forceinline__ u32 gather2(u32 v) {
  v &= 0x55555555u;
  v = (v | (v >> 1)) & 0x33333333u;
  v = (v | (v >> 2)) & 0x0f0f0f0fu;
  v = (v | (v >> 4)) & 0x00ff00ffu;
  v = (v | (v >> 8)) & 0x0000ffffu;
  return v;
}

__global__ void k_cells_to_rows(const uint8_t* __restrict__ cells, u32* __restrict__ rows, int R) {
  const int N = 2 * R, W = N / 32;
  long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long total = (long long)R * R * W;
  if (gid >= total) return;
  int wx = (int)(gid % W);
  long long t = gid / W;
  int cy = (int)(t % R);
  int cz = (int)(t / R);
  const uint4 v = *reinterpret_cast<const uint4*>(cells + (size_t)16 * wx + (size_t)R * ((size_t)cy + (size_t)R * cz));
  u32 w[4] = {v.x, v.y, v.z, v.w};
  // out[s]: fine row (2cy + (s&1), 2cz + (s>>1)); cell bit pair k = 2s, 2s+1
  u32 out[4] = {0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    u32 lo = 0, hi = 0;  // bit j = sub-voxel x=0 / x=1 of cell j
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      u32 byte = (w[j >> 2] >> (8 * (j & 3))) & 0xffu;
      lo |= ((byte >> (2 * s)) & 1u) << j;
      hi |= ((byte >> (2 * s + 1)) & 1u) << j;
    }
    out[s] = spread2(lo) | (spread2(hi) << 1);
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    int y = 2 * cy + (s & 1), z = 2 * cz + (s >> 1);
    rows[((size_t)z * N + y) * W + wx] = out[s];
  }
}

__global__ void k_rows_to_cells(const u32* __restrict__ rows, uint8_t* __restrict__ cells, int R) {
  const int N = 2 * R, W = N / 32;
  long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long total = (long long)R * R * W;
  if (gid >= total) return;
  int wx = (int)(gid % W);
  long long t = gid / W;
  int cy = (int)(t % R);
  int cz = (int)(t / R);
  u32 lo[4], hi[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    int y = 2 * cy + (s & 1), z = 2 * cz + (s >> 1);
    u32 r = rows[((size_t)z * N + y) * W + wx];
    lo[s] = gather2(r);
    hi[s] = gather2(r >> 1);
  }
  u32 w[4] = {0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    u32 byte = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s) byte |= (((lo[s] >> j) & 1u) << (2 * s)) | (((hi[s] >> j) & 1u) << (2 * s + 1));
    w[j >> 2] |= byte << (8 * (j & 3));
  }
  *reinterpret_cast<uint4*>(cells + (size_t)16 * wx + (size_t)R * ((size_t)cy + (size_t)R * cz)) =
      make_uint4(w[0], w[1], w[2], w[3]);
}

hipError_t launch_cells_to_rows(const uint8_t* cells, u32* rows, int R, hipStream_t s) {
  long long total = (long long)R * R * (R / 16);
  int bs = 256;
  hipLaunchKernelGGL(k_cells_to_rows, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0, s, cells, rows, R);
  return hipGetLastError();
}

// A plain device copy of whole 16-byte words, one per lane per step: the
// reference point for the HBM byte counters (its bytes are known exactly).
__global__ void __launch_bounds__(256) k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                    size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy_probe(const void* src, void* dst, size_t bytes, hipStream_t s) {
  const size_t n = bytes / 16;
  hipLaunchKernelGGL(k_copy_probe, dim3(8192), dim3(256), 0, s, static_cast<const uint4*>(src),
                     static_cast<uint4*>(dst), n);
  return hipGetLastError();
}

hipError_t launch_rows_to_cells(const u32* rows, uint8_t* cells, int R, hipStream_t s) {
  long long total = (long long)R * R * (R / 16);
  int bs = 256;
  hipLaunchKernelGGL(k_rows_to_cells, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0, s, rows, cells, R);
  return hipGetLastError();
}

}  // namespace gm
