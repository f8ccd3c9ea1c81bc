// Host runtime shared by the traversal entry points: per-device workspace
// (stream, events, control block, grow-only scratch buffers), persistent-grid
// calibration and spin bounds.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "engine.h"
#include "lsg_internal.h"

#define HIPCHK(expr)                                                                                   \
  do {                                                                                                 \
    hipError_t _e = (expr);                                                                            \
    if (_e != hipSuccess) return ::lsg::fail(LSG_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

namespace lsg {

enum BufSlot {
  kBufQueues = 0,
  kBufHeavy,
  kBufStamps,
  kBufMasks,
  kBufTrace,
  kBufTranspose,
  kBufSeeds,
  kBufSeedList,
  kBufUnion,
  kBufLb,
  kBufValidate,
  kBufGraphPiles,
  kBufCount
};

struct Workspace {
  int device = -1;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int cus = 0;
  double wall_khz = 0;  // device wall clock rate (wall_clock64 ticks per ms)
  Ctrl *ctrl = nullptr;
  uint32_t *pinned = nullptr;  // 256 bytes of pinned host memory for small readbacks
  void *buf[kBufCount] = {};
  size_t buf_bytes[kBufCount] = {};
  struct GridEntry {
    const void *fn;
    int bs, grid;
  };
  std::vector<GridEntry> grids;             // calibrated persistent grid per kernel
  std::vector<lsg_round_trace> last_trace;  // host copy of the last traced call
  std::vector<float> last_feat;             // per-phase feature records of the last traced batched call
  struct Captured {
    std::vector<uint8_t> key;  // exact bytes of the launch arguments
    hipGraphExec_t exec;
  };
  std::vector<Captured> captured;           // hipGraph launch sequences, most recent last
  std::mutex mu;                            // one traversal at a time per device
};

// Launches the kernel sequence that `enqueue` puts on w->stream as one
// hipGraph: captured and instantiated the first time a key is seen, replayed
// afterwards.  `key` must hold every byte the kernels' arguments depend on.
template <class F>
int launch_captured(Workspace *w, const std::vector<uint8_t> &key, F enqueue) {
  for (auto &c : w->captured)
    if (c.key == key) {
      HIPCHK(hipGraphLaunch(c.exec, w->stream));
      return LSG_OK;
    }
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue();
  hipError_t e = hipStreamEndCapture(w->stream, &graph);
  if (rc) {
    if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(LSG_ERR_HIP, "stream capture failed: %s", hipGetErrorString(e));
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(LSG_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
  if (w->captured.size() >= 16) {
    (void)hipGraphExecDestroy(w->captured.front().exec);
    w->captured.erase(w->captured.begin());
  }
  w->captured.push_back({key, exec});
  HIPCHK(hipGraphLaunch(exec, w->stream));
  return LSG_OK;
}

template <class T>
void key_append(std::vector<uint8_t> &k, const T &v) {
  const uint8_t *p = reinterpret_cast<const uint8_t *>(&v);
  k.insert(k.end(), p, p + sizeof(T));
}

// Returns the workspace of `device`, creating it on first use.
int get_workspace(int device, Workspace **out);
// Grow-only device buffer in `slot` of at least `bytes`; contents undefined.
int ws_buffer(Workspace *w, int slot, size_t bytes, void **out);
// Grid barrier spin bound (LSG_BARRIER_TIMEOUT_S, default 20 s) in wall-clock ticks.
uint64_t spin_limit_ticks(const Workspace *w);
// Start-up residency handshake bound (5 ms) in wall-clock ticks.
uint64_t residency_limit_ticks(const Workspace *w);
// Largest grid believed co-resident for (kernel, block size).
int resident_grid(Workspace *w, const void *fn, int bs, int *grid);
void record_resident_grid(Workspace *w, const void *fn, int bs, int grid);

}  // namespace lsg
