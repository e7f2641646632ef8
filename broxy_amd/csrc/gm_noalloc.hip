// Allocation-free closing: gm_close's kernels on caller-owned device grids,
// with every intermediate buffer in one file-scope device array.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, serving loops that forbid hipMalloc). The grids come in and go out
// in the reference's cell-byte layout (Grid.h:61-68); the bit-row grids, cell
// flags, brick summaries and tile queues of the closing live in `scratch`, a
// static __device__ array of kNoAllocScratchBytes that each device holds for
// the life of the process. Its device address is taken once per device with
// hipGetSymbolAddress and only that address is handed to kernels.
#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/gmorpho_c.h"
#include "gm_internal.h"

namespace gm {

namespace {

constexpr int kNoAllocMaxRes = GM_NOALLOC_MAX_RES;

// Byte offsets of the closing's buffers for base resolution R (each 256-byte
// aligned): I, D, K, C rows; DF, TG, CL, UG, I-full, D-any flags; the two
// brick summaries; the tile queues.
struct NoAllocLayout {
  size_t rows[4], flags[6], brickI, brickK, queue, end;
  explicit NoAllocLayout(int R) {
    const size_t N = 2 * (size_t)R, W = N / 32;
    const size_t rowBytes = N * N * W * 4, flagBytes = (size_t)R * R * W * 2, brickBytes = (N / 8) * (N / 8) * W;
    size_t o = 0;
    auto take = [&o](size_t n) {
      const size_t at = o;
      o = (o + n + 255) & ~(size_t)255;
      return at;
    };
    for (auto& r : rows) r = take(rowBytes);
    for (auto& f : flags) f = take(flagBytes);
    brickI = take(brickBytes);
    brickK = take(brickBytes);
    queue = take(tile_queue_words(R) * 4);
    end = o;
  }
};

}  // namespace

// The closing's scratch at the largest supported resolution.
constexpr size_t kNoAllocScratchBytes = (size_t)GM_NOALLOC_SCRATCH_MIB << 20;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// Zeroes both counter blocks of the tile queues (kQueueLists words): the
// selection of each pass clears the block the next pass uses, so only the
// first pass needs this.
__global__ void k_noalloc_clear(u32* queue) {
  for (int i = threadIdx.x; i < kQueueLists; i += blockDim.x) queue[i] = 0u;
}

__global__ void k_copy16(const uint4* __restrict__ in, uint4* __restrict__ out, size_t n16) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};  // device address of scratch, per device
thread_local std::string g_err;

int nfail(int code, const std::string& msg) {
  g_err = "GMorpho Error: " + msg;
  return code;
}

// The device address of scratch on the current device (taken once).
hipError_t scratch_ptr(int dev, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_ptr[dev];
  return hipSuccess;
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) hipSetDevice(prev);
  }
};

}  // namespace

}  // namespace gm

using namespace gm;

extern "C" {

const char* gm_noalloc_last_error(void) { return g_err.c_str(); }

int gm_noalloc_init(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n || device >= kMaxDevices)
    return nfail(GM_ERR_ARG, "no such device");
  DeviceGuard g(device);
  unsigned char* p = nullptr;
  const hipError_t e = scratch_ptr(device, &p);
  if (e != hipSuccess) return nfail(GM_ERR_HIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
  return GM_OK;
}

int gm_close_cells_noalloc(int device, const uint8_t* cells_in, uint8_t* cells_out, int base_res,
                           const int data_res[3], int radius, void* stream) {
  const int R = base_res;
  if (!cells_in || !cells_out || !data_res) return nfail(GM_ERR_ARG, "null argument");
  if (((uintptr_t)cells_in | (uintptr_t)cells_out) & 15u) return nfail(GM_ERR_ARG, "grids must be 16-byte aligned");
  if (R < 16 || (R % 16) != 0) return nfail(GM_ERR_UNSUPPORTED, "base resolution must be a multiple of 16");
  if (R > kNoAllocMaxRes)
    return nfail(GM_ERR_CAPACITY, "base resolution " + std::to_string(R) + " exceeds the allocation-free limit of " +
                                      std::to_string(kNoAllocMaxRes));
  if (radius < 0 || radius > 31) return nfail(GM_ERR_UNSUPPORTED, "radius must be in [0, 31]");
  const NoAllocLayout L(R);
  if (L.end > kNoAllocScratchBytes) return nfail(GM_ERR_CAPACITY, "scratch too small");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n || device >= kMaxDevices)
    return nfail(GM_ERR_ARG, "no such device");
  DeviceGuard g(device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  unsigned char* base = nullptr;
  hipError_t e = scratch_ptr(device, &base);
  if (e != hipSuccess) return nfail(GM_ERR_HIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
  if (radius <= 1) {
    // GMorpho.cu:7784, 7803: no morphology below two fine voxels
    const size_t n16 = (size_t)R * R * R / 16;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)std::min<size_t>((n16 + 255) / 256, 4096)), dim3(256), 0, s,
                       reinterpret_cast<const uint4*>(cells_in), reinterpret_cast<uint4*>(cells_out), n16);
    e = hipGetLastError();
    return e == hipSuccess ? GM_OK : nfail(GM_ERR_HIP, hipGetErrorString(e));
  }
  MorphDims d{};
  d.R = R;
  d.N = 2 * R;
  d.W = d.N / 32;
  for (int a = 0; a < 3; ++a) d.dR[a] = std::max(0, std::min(data_res[a], R));
  MorphBuffers b{};
  u32* rows[4];
  for (int k = 0; k < 4; ++k) rows[k] = reinterpret_cast<u32*>(base + L.rows[k]);
  u16* fl[6];
  for (int k = 0; k < 6; ++k) fl[k] = reinterpret_cast<u16*>(base + L.flags[k]);
  b.I = rows[0];
  b.D = rows[1];
  b.K = rows[2];
  b.C = rows[3];
  b.DF = fl[0];
  b.TG = fl[1];
  b.CL = fl[2];
  b.UG = fl[3];
  b.IFL = fl[4];
  b.DAL = fl[5];
  b.brickI = base + L.brickI;
  b.brickK = base + L.brickK;
  b.queue = reinterpret_cast<u32*>(base + L.queue);
  b.queueCap = (int)tile_count_max(R);
  int qsel = 0;
  b.qsel = &qsel;
  // the uniform ball runs the kernels compiled for its radius: no tables
  b.tables = nullptr;
  SETableSet tabs;
  for (int i = 0; i <= 128; ++i) tabs.dil[i] = tabs.ero_corner[i] = tabs.ero_centre[i] = -1;
  RadiusField rf{};
  rf.mode = 0;
  rf.r = radius;
  rf.N = d.N;
  rf.s = 1;
  hipLaunchKernelGGL(k_noalloc_clear, dim3(1), dim3(256), 0, s, b.queue);
  e = hipGetLastError();
  if (e == hipSuccess) e = launch_cells_to_rows(cells_in, b.I, R, s);
  if (e == hipSuccess) e = launch_dilate(d, b, rf, GM_SE_SPHERE, tabs, radius, radius, 0, s, nullptr, nullptr);
  if (e == hipSuccess) e = launch_erode(d, b, rf, tabs, radius, radius, 0, s, nullptr, nullptr);
  if (e == hipSuccess) e = launch_rows_to_cells(b.C, cells_out, R, s);
  return e == hipSuccess ? GM_OK : nfail(GM_ERR_HIP, std::string("allocation-free closing: ") + hipGetErrorString(e));
}

}  // extern "C"
