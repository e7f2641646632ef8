// ScaleField editing on the device: the global-scale fill
// (ScaleField::UpdateGridGlobalScale, ScaleField.cu:143-146) and the brush
// (ScaleField::UpdateGrid / UpdateScaleGrid, ScaleField.cu:81-192, tukey
// falloff cuda_math.h:696-704, edit modes ScaleField.h:32) on the N^3
// signed-char radius grid.
//
// The reference runs the brush over every voxel of the grid; only voxels
// closer than scale + h to the brush point can change, so here the launch
// covers the box around that sphere. Each changed voxel also moves one count
// between the bins of a 128-bin histogram of the clamped radius values
// (counts per workgroup in LDS, then one atomic per non-zero bin), which
// gives the grid's exact radius range afterwards without reading the grid.
#include <algorithm>

#include "gm_internal.h"

namespace gm {

namespace {

__device__ __forceinline__ int clamp_radius(int v) { return v < 2 ? 2 : (v > 127 ? 127 : v); }

__global__ void __launch_bounds__(256) k_scale_fill(int8_t* __restrict__ grid, size_t n, int8_t value) {
  const size_t n16 = n / 16;
  const int b = (int)(uint8_t)value;
  const u32 w = (u32)b * 0x01010101u;
  uint4* g16 = reinterpret_cast<uint4*>(grid);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    g16[i] = make_uint4(w, w, w, w);
  for (size_t i = 16 * n16 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    grid[i] = value;
}

// One thread per voxel of the box [x0, x0+nx) x [y0, y0+ny) x [z0, z0+nz).
__global__ void __launch_bounds__(256) k_scale_brush(int8_t* __restrict__ grid, int N, ScaleBrush br, int x0, int y0,
                                                     int z0, int nx, int ny, int nz,
                                                     unsigned long long* __restrict__ hist) {
  __shared__ int local[128];
  for (int i = threadIdx.x; i < 128; i += blockDim.x) local[i] = 0;
  __syncthreads();
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid < (long long)nx * ny * nz) {
    const int x = x0 + (int)(gid % nx);
    const long long t = gid / nx;
    const int y = y0 + (int)(t % ny), z = z0 + (int)(t / ny);
    const int id[3] = {x, y, z};
    float d2 = 0.f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float u = __fadd_rn(__fmul_rn(br.cell, (float)id[a]), br.bmin[a]);
      const float r = __fsub_rn(u, br.pos[a]);
      d2 = __fadd_rn(d2, __fmul_rn(r, r));
    }
    const float dist = sqrtf(d2);
    const float v = br.scale;
    if (dist < __fadd_rn(v, br.h)) {
      float distrib = 1.f;
      if (!(dist < v)) {
        const float q = __fdiv_rn(__fsub_rn(dist, v), br.h);
        const float o = __fsub_rn(1.f, __fmul_rn(q, q));
        distrib = __fmul_rn(o, o);
      }
      int8_t* g = grid + (size_t)x + (size_t)N * ((size_t)y + (size_t)N * z);
      const int8_t old = *g;
      const float prev = (float)old;
      const float target = br.mode == 3 ? __fdiv_rn(br.global_scale, br.cell) : __fdiv_rn(v, br.cell);
      float d = floorf(__fadd_rn(__fmul_rn(prev, __fsub_rn(1.f, distrib)), __fmul_rn(distrib, target)));
      d = fmaxf(fminf(d, 127.f), 0.f);
      int8_t nw = old;
      if (br.mode == 1) nw = (int8_t)floorf(fmaxf(d, prev));
      else if (br.mode == 2) nw = (int8_t)floorf(fminf(d, prev));
      else if (br.mode == 3) nw = (int8_t)d;
      if (nw != old) {
        *g = nw;
        atomicSub(&local[clamp_radius(old)], 1);
        atomicAdd(&local[clamp_radius(nw)], 1);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 128; i += blockDim.x)
    if (local[i]) atomicAdd(&hist[i], (unsigned long long)(long long)local[i]);
}

}  // namespace

hipError_t launch_scale_fill(int8_t* grid, size_t n, int8_t value, hipStream_t s) {
  const unsigned blocks = (unsigned)std::min<size_t>((n / 16 + 255) / 256 + 1, 8192);
  hipLaunchKernelGGL(k_scale_fill, dim3(blocks), dim3(256), 0, s, grid, n, value);
  return hipGetLastError();
}

hipError_t launch_scale_brush(int8_t* grid, int N, const ScaleBrush& br, const int box[6],
                              unsigned long long* hist, hipStream_t s) {
  const long long n = (long long)box[3] * box[4] * box[5];
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scale_brush, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, grid, N, br, box[0], box[1],
                     box[2], box[3], box[4], box[5], hist);
  return hipGetLastError();
}

}  // namespace gm
