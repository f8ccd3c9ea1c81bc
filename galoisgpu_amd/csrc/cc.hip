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
// The fixpoint test stays on the device: the host enqueues passes in
// batches, a one-thread step kernel after each pass sets the done flag when
// the pass hooked nothing, and every kernel of a later pass returns at once
// when it finds the flag set.  The host reads the flag once per batch
// instead of once per pass (the reference copies its flag back after every
// pass, apps/cc/kernel.cu:690-716).  The passes stay separate launches: a
// launch boundary is what makes one pass's labels visible to every XCD's L2
// for the next, which a persistent kernel would have to buy with an L2
// invalidate in every block at every pass.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>

#include "engine.h"
#include "lsg_internal.h"
#include "runtime.h"

namespace lsg {

constexpr uint32_t kHookRun = 16;  // edges per thread in a hook pass
constexpr uint32_t kPassBatch = 4;  // passes enqueued per host check

// control words
constexpr int kCtlChanged = 0, kCtlCount = 1, kCtlDone = 2, kCtlPasses = 3;

__global__ void cc_init_kernel(uint32_t *label, uint32_t V, uint32_t *ctl) {
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < V; v += (size_t)gridDim.x * blockDim.x)
    label[v] = (uint32_t)v;
  if (blockIdx.x == 0 && threadIdx.x < 4) ctl[threadIdx.x] = 0;
}

// Whether an earlier pass already reached the fixpoint (written by
// cc_step_kernel in an earlier launch, so a plain load sees it).
__device__ __forceinline__ bool cc_done(const uint32_t *ctl) {
  __shared__ uint32_t s_done;
  if (threadIdx.x == 0) s_done = ctl[kCtlDone];
  __syncthreads();
  return s_done != 0;
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
__global__ void cc_hook_kernel(const uint32_t *row_start, const uint32_t *edge_dst, uint32_t V, uint32_t E,
                               uint32_t *label, uint32_t *ctl) {
  if (cc_done(ctl)) return;
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
  // one store per block that changed anything: a store per wavefront to the
  // same word serialises like a same-address atomic
  __shared__ uint32_t s_changed;
  if (threadIdx.x == 0) s_changed = 0;
  __syncthreads();
  if (__ballot(changed) && lane_id() == 0) s_changed = 1;
  __syncthreads();
  if (threadIdx.x == 0 && s_changed) st_coh(ctl + kCtlChanged, 1u);
}

// Pointer jumping: every vertex's parent becomes its root.
__global__ void cc_jump_kernel(uint32_t *label, uint32_t V, const uint32_t *ctl) {
  if (cc_done(ctl)) return;
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < V; v += (size_t)gridDim.x * blockDim.x) {
    const uint32_t r = find_root(label, (uint32_t)v);
    if (ld_coh(label + v) != r) atomic_min(label + v, r);
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

// After each pass (one thread): count the pass, and either clear the changed
// flag for the next pass or, when the pass hooked nothing, raise done.
__global__ void cc_step_kernel(uint32_t *ctl) {
  if (ctl[kCtlDone]) return;
  ctl[kCtlPasses] += 1;
  if (ctl[kCtlChanged]) ctl[kCtlChanged] = 0;
  else ctl[kCtlDone] = 1;
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
  const uint32_t hgrid = (uint32_t)std::min<uint64_t>(((uint64_t)E / kHookRun + 256) / 256, 16ull * w->cus);
  HIPCHK(hipEventRecord(w->ev0, w->stream));
  hipLaunchKernelGGL(cc_init_kernel, dim3(vgrid), dim3(256), 0, w->stream, label, V, ctl);
  uint32_t *c = w->pinned;
  for (uint64_t enqueued = 0;;) {
    for (uint32_t k = 0; k < kPassBatch; ++k) {
      hipLaunchKernelGGL(cc_hook_kernel, dim3(hgrid), dim3(256), 0, w->stream, g->row_start, g->edge_dst, V, E,
                         label, ctl);
      hipLaunchKernelGGL(cc_jump_kernel, dim3(vgrid), dim3(256), 0, w->stream, label, V, ctl);
      hipLaunchKernelGGL(cc_step_kernel, dim3(1), dim3(1), 0, w->stream, ctl);
    }
    HIPCHK(hipGetLastError());
    enqueued += kPassBatch;
    HIPCHK(hipMemcpyAsync(c, ctl + kCtlDone, 4, hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    if (c[0]) break;
    // A pass that changes anything lowers a label and labels never rise, so a
    // fixpoint always comes; this only guards against a broken invariant.
    if (enqueued > (uint64_t)V + 2) return fail(LSG_ERR_TIMEOUT, "cc: no fixpoint after %llu passes",
                                                (unsigned long long)enqueued);
  }
  hipLaunchKernelGGL(cc_count_kernel, dim3(vgrid), dim3(256), 0, w->stream, label, V, ctl);
  HIPCHK(hipEventRecord(w->ev1, w->stream));
  HIPCHK(hipMemcpyAsync(c, ctl, 16, hipMemcpyDeviceToHost, w->stream));
  HIPCHK(hipStreamSynchronize(w->stream));
  const uint32_t iters = c[kCtlPasses];
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, w->ev0, w->ev1));
  if (ncomponents) *ncomponents = c[kCtlCount];
  if (stats) {
    memset(stats, 0, sizeof *stats);
    stats->ms_kernel = ms;
    stats->iterations = iters;
    stats->items = c[kCtlCount];
    stats->grid_blocks = (int32_t)hgrid;
    stats->block_size = 256;
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
