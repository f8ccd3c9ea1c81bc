// Host runtime shared by the traversal entry points (see runtime.h).

#include "runtime.h"

#include <cstdlib>

namespace lsg {

static std::mutex g_ws_mu;
static std::vector<Workspace *> g_ws;

int get_workspace(int device, Workspace **out) {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  for (auto *w : g_ws)
    if (w->device == device) {
      *out = w;
      return LSG_OK;
    }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(LSG_ERR_NODEV, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(LSG_ERR_ARG, "device %d out of range (%d devices)", device, ndev);
  HIPCHK(hipSetDevice(device));
  auto *w = new Workspace;
  w->device = device;
  HIPCHK(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&w->ev0));
  HIPCHK(hipEventCreate(&w->ev1));
  HIPCHK(hipDeviceGetAttribute(&w->cus, hipDeviceAttributeMultiprocessorCount, device));
  int khz = 0;
  HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
  w->wall_khz = khz > 0 ? khz : 100000.0;
  HIPCHK(hipMalloc(&w->ctrl, sizeof(Ctrl)));
  HIPCHK(hipHostMalloc((void **)&w->pinned, 256, hipHostMallocDefault));
  g_ws.push_back(w);
  *out = w;
  return LSG_OK;
}

int ws_buffer(Workspace *w, int slot, size_t bytes, void **out) {
  if (bytes == 0) bytes = 256;
  if (w->buf_bytes[slot] < bytes) {
    if (w->buf[slot]) HIPCHK(hipFree(w->buf[slot]));
    w->buf[slot] = nullptr;
    w->buf_bytes[slot] = 0;
    if (hipMalloc(&w->buf[slot], bytes) != hipSuccess) {
      w->buf[slot] = nullptr;
      return fail(LSG_ERR_NOMEM, "device workspace: %zu bytes unavailable", bytes);
    }
    w->buf_bytes[slot] = bytes;
  }
  *out = w->buf[slot];
  return LSG_OK;
}

uint64_t spin_limit_ticks(const Workspace *w) {
  double sec = 20.0;
  if (const char *e = getenv("LSG_BARRIER_TIMEOUT_S")) sec = atof(e);
  if (sec < 0) sec = 0;
  return (uint64_t)(sec * w->wall_khz * 1000.0);
}

// Blocks of one launch begin within tens of microseconds, so after 5 ms the
// missing blocks are waiting for a slot.  (A process pays this once per
// persistent kernel whose occupancy figure is too high; 50 ms made a CLI
// run's first traversal 50 ms long.)
uint64_t residency_limit_ticks(const Workspace *w) { return (uint64_t)(0.005 * w->wall_khz * 1000.0); }

int resident_grid(Workspace *w, const void *fn, int bs, int *grid) {
  for (auto &e : w->grids)
    if (e.fn == fn && e.bs == bs) {
      *grid = e.grid;
      return LSG_OK;
    }
  int per_cu = 0;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, bs, 0));
  if (per_cu <= 0) return fail(LSG_ERR_HIP, "persistent kernel cannot be resident (block %d)", bs);
  *grid = std::min<int>(per_cu * w->cus, (int)kMaxGridBlocks);  // the barrier tree's capacity
  w->grids.push_back({fn, bs, *grid});
  return LSG_OK;
}

void record_resident_grid(Workspace *w, const void *fn, int bs, int grid) {
  for (auto &e : w->grids)
    if (e.fn == fn && e.bs == bs) e.grid = grid;
}

}  // namespace lsg

namespace lsg {

// Plain 16-byte-per-lane copy, grid-stride: the HBM reference figure and the
// known-size kernel that PMC byte counters are calibrated against.
__global__ void stream_copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x, st = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) dst[i] = src[i];
}

}  // namespace lsg

using namespace lsg;

extern "C" {

int lsg_stream_copy(int device, size_t bytes, int iters, double *ms_per_copy) {
  if (bytes < 16 || iters <= 0 || !ms_per_copy) return fail(LSG_ERR_ARG, "stream_copy: bad arguments");
  Workspace *w;
  int rc = get_workspace(device, &w);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(w->mu);
  HIPCHK(hipSetDevice(device));
  size_t n = bytes / 16;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
    if (a) (void)hipFree(a);
    return fail(LSG_ERR_NOMEM, "stream_copy: %zu bytes unavailable", 2 * n * 16);
  }
  HIPCHK(hipMemsetAsync(a, 1, n * 16, w->stream));
  uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, (size_t)w->cus * 8);
  hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, w->stream, (const uint4 *)a, (uint4 *)b, n);
  HIPCHK(hipEventRecord(w->ev0, w->stream));
  for (int i = 0; i < iters; ++i)
    hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, w->stream, (const uint4 *)a, (uint4 *)b, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(w->ev1, w->stream));
  HIPCHK(hipEventSynchronize(w->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, w->ev0, w->ev1));
  *ms_per_copy = ms / iters;
  HIPCHK(hipFree(a));
  HIPCHK(hipFree(b));
  return LSG_OK;
}

int lsg_workspace_release(int device) {
  Workspace *w;
  int rc = get_workspace(device, &w);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(w->mu);
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamSynchronize(w->stream));
  for (auto &c : w->captured) (void)hipGraphExecDestroy(c.exec);
  w->captured.clear();
  for (int i = 0; i < kBufCount; ++i) {
    if (w->buf[i]) HIPCHK(hipFree(w->buf[i]));
    w->buf[i] = nullptr;
    w->buf_bytes[i] = 0;
  }
  return LSG_OK;
}

int lsg_dev_malloc(int device, size_t bytes, void **ptr) {
  if (!ptr) return fail(LSG_ERR_ARG, "dev_malloc: NULL out pointer");
  *ptr = nullptr;
  HIPCHK(hipSetDevice(device));
  if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
    *ptr = nullptr;
    return fail(LSG_ERR_NOMEM, "dev_malloc: %zu bytes unavailable on device %d", bytes, device);
  }
  return LSG_OK;
}

int lsg_dev_free(int device, void *ptr) {
  HIPCHK(hipSetDevice(device));
  if (ptr) HIPCHK(hipFree(ptr));
  return LSG_OK;
}

int lsg_dev_copy(int device, void *dst, const void *src, size_t bytes, int kind) {
  if (bytes == 0) return LSG_OK;
  if (!dst || !src) return fail(LSG_ERR_ARG, "dev_copy: NULL pointer");
  hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  if (kind < 0 || kind > 2) return fail(LSG_ERR_ARG, "dev_copy: bad kind %d", kind);
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipMemcpy(dst, src, bytes, k));
  return LSG_OK;
}

}  // extern "C"
