// Grid::BuildTexMipmaps (Grid.cu:643-1154): the occupancy + radius max-mipmap
// in the reference's texel format. Level 0 has one 64-bit texel per cell:
// bits 7k..7k+6 = radius of sub-voxel k (FillLevel0, Grid.cu:957-998 / 643-731),
// bits 56+k = occupancy of sub-voxel k. Level L+1 keeps, per child l, the max
// of the child's 7-bit radii and whether the child texel has any occupancy bit
// (BuildMipmapFromTo, Grid.cu:733-805). Levels are stored as flat arrays
// (x fastest), level L at offset sum_{i<L} (R>>i)^3. The closing kernels do not
// traverse this pyramid; it is built for callers of the Grid interface.
#include "gm_internal.h"

namespace gm {

namespace {

__global__ void k_level0(const u32* __restrict__ rows, int R, RadiusField rf, u64* __restrict__ out) {
  const int N = 2 * R, W = N / 32;
  long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long total = (long long)R * R * R;
  if (gid >= total) return;
  int cx = (int)(gid % R);
  long long t = gid / R;
  int cy = (int)(t % R), cz = (int)(t / R);
  u64 texel = 0;
  for (int k = 0; k < 8; ++k) {
    int x = 2 * cx + (k & 1), y = 2 * cy + ((k >> 1) & 1), z = 2 * cz + ((k >> 2) & 1);
    u32 w = rows[((size_t)z * N + y) * W + (x >> 5)];
    texel |= (u64)((w >> (x & 31)) & 1u) << (56 + k);
    texel |= (u64)(radius_at(rf, x, y, z) & 0x7f) << (7 * k);
  }
  out[gid] = texel;
}

__global__ void k_level_up(const u64* __restrict__ from, int sf, u64* __restrict__ to) {
  const int st = sf / 2;
  long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long total = (long long)st * st * st;
  if (gid >= total) return;
  int x = (int)(gid % st);
  long long t = gid / st;
  int y = (int)(t % st), z = (int)(t / st);
  u64 texel = 0;
  for (int l = 0; l < 8; ++l) {
    int cx = 2 * x + (l & 1), cy = 2 * y + ((l >> 1) & 1), cz = 2 * z + ((l >> 2) & 1);
    u64 c = from[(size_t)cx + (size_t)sf * ((size_t)cy + (size_t)sf * cz)];
    u64 m = 0;
    for (int k = 0; k < 8; ++k) {
      u64 r = (c >> (7 * k)) & 0x7f;
      m = r > m ? r : m;
    }
    texel |= m << (7 * l);
    if ((c >> 56) != 0) texel |= 1ull << (56 + l);
  }
  to[gid] = texel;
}

}  // namespace

hipError_t launch_pyramid(const u32* rows, int R, const RadiusField& rf, u64* levels, hipStream_t s) {
  long long n0 = (long long)R * R * R;
  hipLaunchKernelGGL(k_level0, dim3((unsigned)((n0 + 255) / 256)), dim3(256), 0, s, rows, R, rf, levels);
  GM_CHECK(hipGetLastError());
  u64* cur = levels;
  int sz = R;
  while (sz > 1) {
    u64* nxt = cur + (size_t)sz * sz * sz;
    long long n = (long long)(sz / 2) * (sz / 2) * (sz / 2);
    hipLaunchKernelGGL(k_level_up, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, cur, sz, nxt);
    GM_CHECK(hipGetLastError());
    cur = nxt;
    sz /= 2;
  }
  return hipSuccess;
}

}  // namespace gm
