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

enum BufSlot { kBufQueues = 0, kBufHeavy, kBufStamps, kBufMasks, kBufTrace, kBufTranspose, kBufSeeds, kBufCount };

struct Workspace {
  int device = -1;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int cus = 0;
  double wall_khz = 0;  // device wall clock rate (wall_clock64 ticks per ms)
  Ctrl *ctrl = nullptr;
  void *buf[kBufCount] = {};
  size_t buf_bytes[kBufCount] = {};
  struct GridEntry {
    const void *fn;
    int bs, grid;
  };
  std::vector<GridEntry> grids;             // calibrated persistent grid per kernel
  std::vector<lsg_round_trace> last_trace;  // host copy of the last traced call
  std::vector<float> last_feat;             // per-phase feature records of the last traced batched call
  std::mutex mu;                            // one traversal at a time per device
};

// Returns the workspace of `device`, creating it on first use.
int get_workspace(int device, Workspace **out);
// Grow-only device buffer in `slot` of at least `bytes`; contents undefined.
int ws_buffer(Workspace *w, int slot, size_t bytes, void **out);
// Grid barrier spin bound (LSG_BARRIER_TIMEOUT_S, default 20 s) in wall-clock ticks.
uint64_t spin_limit_ticks(const Workspace *w);
// Start-up residency handshake bound (50 ms) in wall-clock ticks.
uint64_t residency_limit_ticks(const Workspace *w);
// Largest grid believed co-resident for (kernel, block size).
int resident_grid(Workspace *w, const void *fn, int bs, int *grid);
void record_resident_grid(Workspace *w, const void *fn, int bs, int grid);

}  // namespace lsg
