// PageRank on gfx950: pull-based Jacobi iteration over the reverse CSR.
//
// Reference: apps/pr/kernel.cu -- ALPHA 0.85 and EPSILON 1e-6 (:19-20),
// p = 1 - ALPHA initially (:38), residual push until no residual exceeds
// EPSILON (:110-236), "PR took %d iterations" (:387); apps/pr/support.cu --
// options -n -t -x (:15-42), "sum: %f (%d)" and one "%d %f" line per node
// (:56-97).  Both methods converge to the fixpoint
//     p[v] = (1 - ALPHA) + ALPHA * sum over in-edges (u, v) of p[u] / outdeg(u)
// (vertices without out-edges pass nothing on).  The reference pushes
// residuals with float atomics in a data-driven order; here every iteration
// recomputes every vertex from its in-neighbours' contributions
// c[u] = p[u] / outdeg(u) -- no atomics, a deterministic summation order --
// and stops when no rank moved by more than epsilon relative to itself
// (every rank is at least 1 - ALPHA; an absolute test would wait on the
// float rounding of the largest ranks).
//
// Layout: rank and contribution arrays f32[V], two of each (current and
// next); the reverse CSR from lsg_graph_transpose; a list of the vertices
// with in-degree >= kWaveDeg, which a whole wavefront sums (the others take
// one lane each).
//
// The convergence test stays on the device: the host enqueues iterations in
// batches; after each iteration a one-thread step kernel counts it and sets a
// done flag when no rank moved by more than epsilon, and every kernel of a
// later iteration returns at once when it finds the flag set.  The host reads
// the flag once per batch instead of once per iteration.  Iterations stay
// separate launches: the launch boundary makes one iteration's contributions
// visible to every XCD's L2 for the next iteration's gathers, which a
// persistent kernel would have to buy with an L2 invalidate in every block at
// every iteration.
//
// Why pull Jacobi and not the reference's residual push: a push sweep adds
// to its out-neighbours' residuals with one float atomic per edge (measured
// on gfx950: ~27 G random 4-byte atomics/s, against ~90 G gathers/s of the
// pull sweep here), and float atomics make every sum's order, and so the
// ranks' last bits, depend on the schedule; the pull sweep has no atomics and
// one summation order.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>

#include "engine.h"
#include "lsg_internal.h"
#include "runtime.h"

namespace lsg {

constexpr uint32_t kWaveDeg = 32;  // in-degree from which a wavefront sums a vertex

struct PrArgs {
  const uint32_t *row_start;         // forward CSR (out-degrees)
  const uint32_t *in_row, *in_src;   // reverse CSR
  const float *c_cur;                // contributions of the current ranks
  float *p_next, *c_next;
  const uint32_t *heavy;             // vertices with in-degree >= kWaveDeg
  uint32_t nheavy, nnodes;
  float alpha;
  uint32_t *maxdiff;                 // bits of the largest |p_next - p_cur| / p_next (non-negative floats order as u32)
  const float *p_cur;
  const uint32_t *done;              // set once an earlier iteration converged
};

constexpr uint32_t kIterBatch = 8;  // iterations enqueued per host check

// control words; the largest change is kept per iteration parity, so the step
// kernel can clear the next iteration's word while this one's is read
constexpr int kCtlHeavy = 0, kCtlMax = 1, kCtlDone = 3, kCtlIters = 4;

// Whether an earlier iteration converged (written by pr_step_kernel in an
// earlier launch, so a plain load sees it).
__device__ __forceinline__ bool pr_done(const PrArgs &a) {
  __shared__ uint32_t s_done;
  if (threadIdx.x == 0) s_done = *a.done;
  __syncthreads();
  return s_done != 0;
}

__global__ void pr_init_kernel(const uint32_t *row_start, const uint32_t *in_row, uint32_t V, float alpha, float *p,
                               float *c, uint32_t *heavy, uint32_t *nheavy) {
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < V; v += (size_t)gridDim.x * blockDim.x) {
    const float p0 = 1.0f - alpha;  // apps/pr/kernel.cu:38
    const uint32_t od = row_start[v + 1] - row_start[v];
    p[v] = p0;
    c[v] = od ? p0 / (float)od : 0.0f;
    const bool h = in_row[v + 1] - in_row[v] >= kWaveDeg;
    const uint64_t m = __ballot(h);
    if (m) {
      uint32_t base = 0;
      const uint32_t leader = (uint32_t)__builtin_ctzll(m);
      if (lane_id() == leader) base = atomic_add(nheavy, (uint32_t)__builtin_popcountll(m));
      base = readlane(base, leader);
      if (h) heavy[base + lanes_below(m)] = (uint32_t)v;
    }
  }
}

__device__ __forceinline__ void pr_store(const PrArgs &a, uint32_t v, float sum, float &dmax) {
  const float pn = (1.0f - a.alpha) + a.alpha * sum;
  const uint32_t od = a.row_start[v + 1] - a.row_start[v];
  a.p_next[v] = pn;
  a.c_next[v] = od ? pn / (float)od : 0.0f;
  dmax = fmaxf(dmax, fabsf(pn - a.p_cur[v]) / pn);  // pn >= 1 - alpha > 0
}

// The block's largest change goes to maxdiff with one atomic per block (one
// per wavefront on the same word serialises at the memory-side atomic unit).
__device__ __forceinline__ void pr_reduce_max(const PrArgs &a, float dmax) {
  __shared__ uint32_t s_max;
  if (threadIdx.x == 0) s_max = 0;
  __syncthreads();
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) dmax = fmaxf(dmax, __shfl_xor(dmax, d));
  if (lane_id() == 0 && dmax > 0.0f) atomicMax(&s_max, __float_as_uint(dmax));
  __syncthreads();
  if (threadIdx.x == 0 && s_max)
    __hip_atomic_fetch_max(a.maxdiff, s_max, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One lane per vertex of in-degree < kWaveDeg.
__global__ void pr_light_kernel(PrArgs a) {
  if (pr_done(a)) return;
  float dmax = 0.0f;
  const size_t nt = (size_t)gridDim.x * blockDim.x;
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < a.nnodes; v += nt) {
    const uint32_t e0 = a.in_row[v], e1 = a.in_row[v + 1];
    if (e1 - e0 >= kWaveDeg) continue;
    // four in-edges per step, summed in edge order
    float sum = 0.0f;
    uint32_t e = e0;
    for (; e + 4 <= e1; e += 4) {
      uint32_t u[4];
      float c[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) u[k] = a.in_src[e + k];
#pragma unroll
      for (int k = 0; k < 4; ++k) c[k] = a.c_cur[u[k]];
#pragma unroll
      for (int k = 0; k < 4; ++k) sum += c[k];
    }
    for (; e < e1; ++e) sum += a.c_cur[a.in_src[e]];
    pr_store(a, (uint32_t)v, sum, dmax);
  }
  pr_reduce_max(a, dmax);
}

// One wavefront per vertex of in-degree >= kWaveDeg: 64 in-edges per step,
// then a butterfly sum (the same order every iteration).
__global__ void pr_heavy_kernel(PrArgs a) {
  if (pr_done(a)) return;
  float dmax = 0.0f;
  const uint32_t lane = lane_id();
  const size_t waves = (size_t)gridDim.x * blockDim.x / kWave;
  for (size_t i = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) / kWave; i < a.nheavy; i += waves) {
    const uint32_t v = a.heavy[i];
    const uint32_t e0 = a.in_row[v], e1 = a.in_row[v + 1];
    // lane l sums in-edges e0 + l, e0 + l + 64, ... in that order, four per step
    float sum = 0.0f;
    for (uint32_t e = e0 + lane; e < e1; e += 4 * kWave) {
      uint32_t u[4];
      float c[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) u[k] = e + k * kWave < e1 ? a.in_src[e + k * kWave] : 0u;
#pragma unroll
      for (int k = 0; k < 4; ++k) c[k] = e + k * kWave < e1 ? a.c_cur[u[k]] : 0.0f;
#pragma unroll
      for (int k = 0; k < 4; ++k) sum += c[k];
    }
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) sum += __shfl_xor(sum, d);
    if (lane == 0) pr_store(a, v, sum, dmax);
  }
  pr_reduce_max(a, dmax);
}

__global__ void pr_clear_kernel(uint32_t *w, uint32_t n) {
  if (threadIdx.x < n) w[threadIdx.x] = 0;
}

// After iteration `it` (one thread): count it, raise done when no rank moved
// by more than epsilon, and clear the next iteration's change word.
__global__ void pr_step_kernel(uint32_t *ctl, uint32_t it, float epsilon) {
  if (ctl[kCtlDone]) return;
  ctl[kCtlIters] = it + 1;
  if (__uint_as_float(ctl[kCtlMax + (it & 1)]) <= epsilon) ctl[kCtlDone] = 1;
  ctl[kCtlMax + ((it + 1) & 1)] = 0;
}

}  // namespace lsg

using namespace lsg;

extern "C" {

int lsg_pagerank(const lsg_graph *g, const lsg_graph *gt, float alpha, float epsilon, int32_t max_iterations,
                 float *rank_dev, lsg_stats *stats) {
  clear_error();
  auto t0 = std::chrono::steady_clock::now();
  if (!g || g->device_graph == 0 || !g->row_start || !gt || gt->device_graph == 0 || !gt->row_start)
    return fail(LSG_ERR_ARG, "pagerank: expected a device graph and its reverse (lsg_graph_transpose)");
  if (gt->nnodes != g->nnodes || gt->nedges != g->nedges || gt->device != g->device)
    return fail(LSG_ERR_ARG, "pagerank: the reverse graph does not match the graph");
  if (!rank_dev) return fail(LSG_ERR_ARG, "pagerank: NULL rank array");
  if (!(alpha > 0.0f && alpha < 1.0f)) return fail(LSG_ERR_ARG, "pagerank: alpha must be in (0, 1)");
  if (!(epsilon > 0.0f)) return fail(LSG_ERR_ARG, "pagerank: epsilon must be > 0");
  if (max_iterations <= 0) return fail(LSG_ERR_ARG, "pagerank: max_iterations must be > 0");
  Workspace *w;
  int rc = get_workspace(g->device, &w);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(w->mu);
  HIPCHK(hipSetDevice(w->device));
  const uint32_t V = g->nnodes;
  void *buf, *cbuf;
  // p[2], c[2] (f32[V] each) and the heavy list (u32[V])
  if ((rc = ws_buffer(w, kBufQueues, 5ull * V * 4, &buf)) || (rc = ws_buffer(w, kBufSeeds, 2048, &cbuf))) return rc;
  float *p[2] = {(float *)buf, (float *)buf + V};
  float *c[2] = {(float *)buf + 2ull * V, (float *)buf + 3ull * V};
  uint32_t *heavy = (uint32_t *)buf + 4ull * V;
  uint32_t *ctl = (uint32_t *)cbuf;  // heavy count, max change (x2), done, iterations
  const uint32_t vgrid = (uint32_t)std::min<uint64_t>(((uint64_t)V + 255) / 256, 16ull * w->cus);
  HIPCHK(hipEventRecord(w->ev0, w->stream));
  hipLaunchKernelGGL(pr_clear_kernel, dim3(1), dim3(64), 0, w->stream, ctl, 5u);
  hipLaunchKernelGGL(pr_init_kernel, dim3(vgrid), dim3(256), 0, w->stream, g->row_start, gt->row_start, V, alpha,
                     p[0], c[0], heavy, ctl);
  HIPCHK(hipGetLastError());
  uint32_t *hc = w->pinned;
  HIPCHK(hipMemcpyAsync(hc, ctl, 4, hipMemcpyDeviceToHost, w->stream));
  HIPCHK(hipStreamSynchronize(w->stream));
  const uint32_t nheavy = hc[0];
  PrArgs a;
  a.row_start = g->row_start;
  a.in_row = gt->row_start;
  a.in_src = gt->edge_dst;
  a.heavy = heavy;
  a.nheavy = nheavy;
  a.nnodes = V;
  a.alpha = alpha;
  a.done = ctl + kCtlDone;
  const uint32_t hgrid = (uint32_t)std::min<uint64_t>(((uint64_t)nheavy * kWave + 255) / 256 + 1, 16ull * w->cus);
  // Iteration `it` reads p[it & 1], c[it & 1] and writes the other pair.
  for (uint32_t it = 0; it < (uint32_t)max_iterations;) {
    const uint32_t end = std::min<uint32_t>((uint32_t)max_iterations, it + kIterBatch);
    for (; it < end; ++it) {
      a.p_cur = p[it & 1];
      a.c_cur = c[it & 1];
      a.p_next = p[(it & 1) ^ 1];
      a.c_next = c[(it & 1) ^ 1];
      a.maxdiff = ctl + kCtlMax + (it & 1);
      hipLaunchKernelGGL(pr_light_kernel, dim3(vgrid), dim3(256), 0, w->stream, a);
      if (nheavy) hipLaunchKernelGGL(pr_heavy_kernel, dim3(hgrid), dim3(256), 0, w->stream, a);
      hipLaunchKernelGGL(pr_step_kernel, dim3(1), dim3(1), 0, w->stream, ctl, it, epsilon);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hc, ctl + kCtlDone, 4, hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    if (hc[0]) break;
  }
  HIPCHK(hipMemcpyAsync(hc, ctl + kCtlIters, 4, hipMemcpyDeviceToHost, w->stream));
  HIPCHK(hipStreamSynchronize(w->stream));
  const int iters = (int)hc[0];
  HIPCHK(hipMemcpyAsync(rank_dev, p[iters & 1], (size_t)V * 4, hipMemcpyDeviceToDevice, w->stream));
  HIPCHK(hipEventRecord(w->ev1, w->stream));
  HIPCHK(hipEventSynchronize(w->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, w->ev0, w->ev1));
  if (stats) {
    memset(stats, 0, sizeof *stats);
    stats->ms_kernel = ms;
    stats->iterations = (uint64_t)iters;
    stats->items = nheavy;
    stats->grid_blocks = (int32_t)vgrid;
    stats->block_size = 256;
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return LSG_OK;
}

}  // extern "C"
