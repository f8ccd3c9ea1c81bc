// Connected components on gfx950: hooking and pointer jumping over a device CSR.
//
// Reference: apps/cc/kernel.cu -- init (:18-29) labels every node with its
// own id, hook_init/hook_high_to_low/hook_low_to_high (:249-336) point the
// root of one endpoint of every edge at the other's, p_jump (:338-369)
// shortcuts parents, count_components (:449-462) counts roots, and gg_main
// (:630-723) iterates until no hook succeeds.  Edges are taken without
// direction (weakly connected components of a directed graph).  The
// reference's hooks are plain racing stores, so which id ends up as a
// component's label depends on the schedule; here a hook is an atomicMin of
// the larger root's parent to the smaller root, so every vertex ends labelled
// with the smallest id of its component -- the same partition, with one
// deterministic label per component.
//
// Layout: label u32[V] (the parent pointer, node_data); four control words
// (changed flag, component count, done flag, passes).  One iteration is one
// hook pass over all edges (16 consecutive edges per thread, their source
// found once by binary search) followed by pointer jumping to the root for
// every vertex.
//
// All passes run in one persistent launch on the runtime the traversals use
// (residency handshake, leader grid barrier): a hook pass, a barrier, a
// jump pass, and a barrier whose leader ends the loop when the pass hooked
// nothing (the reference's gg_main_pipe_4 loop, apps/cc/kernel.cu:465-707,
// which copies its flag back to the host after every pass instead).  Every
// label access within the launch is a coherent load or an atomic, so blocks
// on different XCDs see each other's hooks without cache maintenance.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>

#include "engine.h"
#include "lsg_internal.h"
#include "runtime.h"

namespace lsg {

constexpr uint32_t kHookRun = 16;  // edges per thread in a hook pass

// control word of the component count
constexpr int kCtlCount = 1;

__global__ void cc_init_kernel(uint32_t *label, uint32_t V, uint32_t *ctl) {
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < V; v += (size_t)gridDim.x * blockDim.x)
    label[v] = (uint32_t)v;
  if (blockIdx.x == 0 && threadIdx.x < 4) ctl[threadIdx.x] = 0;
}

// Root of v: follows parents (every parent id is <= the child's, so the walk
// ends at a vertex labelled with itself).
__device__ __forceinline__ uint32_t find_root(const uint32_t *label, uint32_t v) {
  uint32_t p = ld_coh(label + v);
  while (true) {
    const uint32_t q = ld_coh(label + p);
    if (q == p) return p;
    p = q;
  }
}

// Hook pass: for every edge (u, v), if u's and v's roots differ, the larger
// root is hooked under the smaller with atomicMin; ctl[0] is set if any hook
// lowered a parent.
__device__ bool cc_hook_pass(const uint32_t *row_start, const uint32_t *edge_dst, uint32_t V, uint32_t E,
                             uint32_t *label) {
  const size_t nthreads = (size_t)gridDim.x * blockDim.x;
  bool changed = false;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t * kHookRun < E; t += nthreads) {
    const uint32_t e0 = (uint32_t)(t * kHookRun), e1 = min(E, e0 + kHookRun);
    uint32_t lo = 0, hi = V;  // last u with row_start[u] <= e0
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (row_start[mid] <= e0) lo = mid;
      else hi = mid;
    }
    uint32_t u = lo, next = row_start[u + 1];
    uint32_t ru = find_root(label, u);
    for (uint32_t e = e0; e < e1; ++e) {
      if (next <= e) {
        while (next <= e) next = row_start[++u + 1];
        ru = find_root(label, u);
      }
      const uint32_t rv = find_root(label, edge_dst[e]);
      if (ru == rv) continue;
      const uint32_t hi_r = max(ru, rv), lo_r = min(ru, rv);
      if (atomic_min(label + hi_r, lo_r) > lo_r) changed = true;
      ru = find_root(label, u);
    }
  }
  return changed;
}

// Pointer jumping: every vertex's parent becomes its root.
__device__ void cc_jump_pass(uint32_t *label, uint32_t V) {
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < V; v += (size_t)gridDim.x * blockDim.x) {
    const uint32_t r = find_root(label, (uint32_t)v);
    if (ld_coh(label + v) != r) atomic_min(label + v, r);
  }
}

struct CcArgs {
  const uint32_t *row_start, *edge_dst;
  uint32_t *label;
  Ctrl *ctrl;
  uint32_t V, E;
  uint32_t max_passes;
  uint64_t spin_limit, residency_limit;
};

// Schedule words: round = passes done, done; aux[p & 1] = "pass p hooked".
constexpr uint32_t kCcBlock = 256;

__global__ __launch_bounds__(kCcBlock) void cc_persistent_kernel(CcArgs a) {
  __shared__ uint32_t s_changed, quit;
  Ctrl *c = a.ctrl;
  const uint32_t nb = gridDim.x;
  if (!residency_handshake(c, nb, a.residency_limit)) return;
  for (uint32_t pass = 0;; ++pass) {
    const bool changed = cc_hook_pass(a.row_start, a.edge_dst, a.V, a.E, a.label);
    // one store per block that changed anything (a store per wavefront to
    // the same word serialises like a same-address atomic)
    if (threadIdx.x == 0) s_changed = 0;
    __syncthreads();
    if (__ballot(changed) && lane_id() == 0) s_changed = 1;
    __syncthreads();
    if (threadIdx.x == 0 && s_changed) st_coh(&c->s.aux[pass & 1], 1u);
    grid_barrier_block(c, nb, a.spin_limit, [] {});
    cc_jump_pass(a.label, a.V);
    grid_barrier_block(c, nb, a.spin_limit, [&] {
      if (threadIdx.x == 0) {
        const uint32_t hooked = ld_coh(&c->s.aux[pass & 1]);
        st_coh(&c->s.aux[(pass + 1) & 1], 0u);  // the next pass's flag (nobody writes it before the release)
        st_coh(&c->s.round, pass + 1);
        if (!hooked || pass + 1 >= a.max_passes) st_coh(&c->s.done, 1u);
      }
    });
    if (threadIdx.x == 0) quit = ld_coh(&c->s.done) != 0 || ld_coh(&c->abort) != kAbortNone;
    __syncthreads();
    if (quit) break;
  }
}

__global__ void cc_ctrl_init_kernel(Ctrl *c) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x, nt = (size_t)gridDim.x * blockDim.x;
  ctrl_reset_barrier(c, t, nt);
  if (t == 0) {
    c->abort = kAbortNone;
    c->checkin = 0;
    c->resident_seen = 0;
    c->progress = 0;
    Sched s;
    memset(&s, 0, sizeof s);
    c->s = s;
  }
}

__global__ void cc_count_kernel(const uint32_t *label, uint32_t V, uint32_t *ctl) {
  __shared__ uint32_t s_n;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  uint32_t n = 0;
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < V; v += (size_t)gridDim.x * blockDim.x)
    n += label[v] == v;
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) n += __shfl_xor(n, d);
  if (lane_id() == 0 && n) atomicAdd(&s_n, n);
  __syncthreads();
  if (threadIdx.x == 0 && s_n) atomic_add(ctl + kCtlCount, s_n);
}

}  // namespace lsg

using namespace lsg;

extern "C" {

int lsg_cc(const lsg_graph *g, int32_t *label_dev, uint32_t *ncomponents, lsg_stats *stats) {
  clear_error();
  auto t0 = std::chrono::steady_clock::now();
  if (!g || g->device_graph == 0 || !g->row_start || (g->nedges && !g->edge_dst))
    return fail(LSG_ERR_ARG, "cc: expected a device graph (copy_to_gpu first)");
  if (!label_dev) return fail(LSG_ERR_ARG, "cc: NULL label array");
  Workspace *w;
  int rc = get_workspace(g->device, &w);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(w->mu);
  HIPCHK(hipSetDevice(w->device));
  void *cbuf;
  if ((rc = ws_buffer(w, kBufSeeds, 2048, &cbuf))) return rc;
  uint32_t *ctl = (uint32_t *)cbuf;
  const uint32_t V = g->nnodes, E = g->nedges;
  uint32_t *label = (uint32_t *)label_dev;
  const uint32_t vgrid = (uint32_t)std::min<uint64_t>(((uint64_t)V + 255) / 256, 16ull * w->cus);
  const void *fn = (const void *)cc_persistent_kernel;
  int grid = 0;
  if ((rc = resident_grid(w, fn, kCcBlock, &grid))) return rc;
  grid = std::min<int>(grid, 4 * (int)w->cus);
  CcArgs a;
  a.row_start = g->row_start;
  a.edge_dst = g->edge_dst;
  a.label = label;
  a.ctrl = w->ctrl;
  a.V = V;
  a.E = E;
  // A pass that changes anything lowers a label and labels never rise, so a
  // fixpoint always comes; the bound only guards against a broken invariant.
  a.max_passes = V + 2;
  a.spin_limit = spin_limit_ticks(w);
  a.residency_limit = residency_limit_ticks(w);
  Ctrl hc;
  float ms = 0;
  uint32_t *c = w->pinned;
  for (int attempt = 0;; ++attempt) {
    HIPCHK(hipEventRecord(w->ev0, w->stream));
    hipLaunchKernelGGL(cc_init_kernel, dim3(vgrid), dim3(256), 0, w->stream, label, V, ctl);
    hipLaunchKernelGGL(cc_ctrl_init_kernel, dim3(1), dim3(256), 0, w->stream, w->ctrl);
    hipLaunchKernelGGL(cc_persistent_kernel, dim3(grid), dim3(kCcBlock), 0, w->stream, a);
    hipLaunchKernelGGL(cc_count_kernel, dim3(vgrid), dim3(256), 0, w->stream, label, V, ctl);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(w->ev1, w->stream));
    HIPCHK(hipMemcpyAsync(&hc, w->ctrl, kCtrlHostBytes, hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipMemcpyAsync(c, ctl, 16, hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    HIPCHK(hipEventElapsedTime(&ms, w->ev0, w->ev1));
    if (hc.abort != kAbortResidency) break;
    if (attempt > 0 || hc.resident_seen == 0 || hc.resident_seen >= (uint32_t)grid)
      return fail(LSG_ERR_HIP, "cc: persistent grid of %d x %d blocks is not co-resident (%u were)", grid,
                  (int)kCcBlock, hc.resident_seen);
    grid = (int)hc.resident_seen;
    record_resident_grid(w, fn, kCcBlock, grid);
  }
  if (hc.abort == kAbortTimeout) return fail(LSG_ERR_TIMEOUT, "cc: grid barrier timed out after %u passes", hc.s.round);
  if (!hc.s.done) return fail(LSG_ERR_TIMEOUT, "cc: no fixpoint after %u passes", hc.s.round);
  const uint32_t iters = hc.s.round;
  if (ncomponents) *ncomponents = c[kCtlCount];
  if (stats) {
    memset(stats, 0, sizeof *stats);
    stats->ms_kernel = ms;
    stats->iterations = iters;
    stats->items = c[kCtlCount];
    stats->grid_blocks = grid;
    stats->block_size = kCcBlock;
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return LSG_OK;
}

int lsg_cc_gg_main(lsg_graph *hg, lsg_graph *gg, lsg_stats *stats) {
  (void)hg;
  if (!gg || !gg->node_data) return fail(LSG_ERR_ARG, "cc gg_main: device graph without node_data");
  return lsg_cc(gg, gg->node_data, nullptr, stats);
}

}  // extern "C"
