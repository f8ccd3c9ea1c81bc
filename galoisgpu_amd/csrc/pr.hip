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
// with in-degree >= kWaveDeg (the others take one lane each), whose in-edges
// are cut into chunks of kChunk: one wavefront sums a chunk into a partial
// (part[]), then one lane per vertex adds its partials in chunk order.  (One
// wavefront per vertex left the whole sweep waiting on the largest in-degree,
// ~97 K edges on RMAT-22: ~0.4 ms of a 0.72 ms iteration.)
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

constexpr uint32_t kChunk = 4096;  // in-edges per wavefront-sized partial sum
constexpr int kUnrollPr = 8;       // in-edges per lane in flight in a partial sum

struct PrArgs {
  const uint32_t *row_start;         // forward CSR (out-degrees)
  const uint32_t *in_row, *in_src;   // reverse CSR
  const float *c_cur;                // contributions of the current ranks
  float *p_next, *c_next;
  const uint32_t *heavy;             // vertices with in-degree >= kWaveDeg
  const uint32_t *hbase;             // heavy vertex i's first partial; its partials are consecutive
  const uint32_t *chunk;             // partial k belongs to heavy vertex chunk[k] (chunk number k - hbase[.])
  float *part;                       // partial sums, kChunk in-edges each
  uint32_t nheavy, nchunks, nnodes;
  float alpha;
  uint32_t *maxdiff;                 // bits of the largest |p_next - p_cur| / p_next (non-negative floats order as u32)
  const float *p_cur;
  const uint32_t *done;              // set once an earlier iteration converged
};

constexpr uint32_t kIterBatch = 8;  // iterations enqueued per host check

// control words; the largest change is kept per iteration parity, so the step
// kernel can clear the next iteration's word while this one's is read
constexpr int kCtlHeavy = 0, kCtlMax = 1, kCtlDone = 3, kCtlIters = 4, kCtlChunks = 5;

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

// After pr_init_kernel: heavy vertex i gets ceil(indeg / kChunk) consecutive
// partials starting at hbase[i] (where they sit depends on the order of the
// atomics; the order in which a vertex's partials are added never does).
__global__ void pr_chunks_kernel(const uint32_t *in_row, const uint32_t *heavy, uint32_t nheavy, uint32_t *hbase,
                                 uint32_t *chunk, uint32_t *nchunks) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const uint32_t lane = lane_id();
  // wavefront-uniform loop: one reservation atomic per wavefront (one per
  // vertex, ~280 K on one word for RMAT-22, would serialise for milliseconds)
  for (size_t i0 = blockIdx.x * (size_t)blockDim.x + (threadIdx.x & ~(uint32_t)(kWave - 1)); i0 < nheavy; i0 += stride) {
    const size_t i = i0 + lane;
    const bool ok = i < nheavy;
    const uint32_t v = ok ? heavy[i] : 0u;
    const uint32_t n = ok ? (in_row[v + 1] - in_row[v] + kChunk - 1) / kChunk : 0u;
    uint32_t incl = n;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const uint32_t y = __shfl_up(incl, d);
      if (lane >= (uint32_t)d) incl += y;
    }
    const uint32_t total = readlane(incl, kWave - 1);
    uint32_t base = 0;
    if (lane == 0 && total) base = atomic_add(nchunks, total);
    base = readlane(base, 0) + incl - n;
    if (ok) {
      hbase[i] = base;
      for (uint32_t k = 0; k < n; ++k) chunk[base + k] = (uint32_t)i;
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

// One wavefront per partial: kChunk in-edges of a heavy vertex, 64 per step
// (four steps' loads in flight), then a butterfly sum -- the same order every
// iteration.
__global__ void pr_chunk_kernel(PrArgs a) {
  if (pr_done(a)) return;
  const uint32_t lane = lane_id();
  const size_t waves = (size_t)gridDim.x * blockDim.x / kWave;
  for (size_t k = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) / kWave; k < a.nchunks; k += waves) {
    const uint32_t i = a.chunk[k];
    const uint32_t v = a.heavy[i];
    const uint32_t e0 = a.in_row[v] + (uint32_t)(k - a.hbase[i]) * kChunk;
    const uint32_t e1 = min(a.in_row[v + 1], e0 + kChunk);
    // lane l sums in-edges e0 + l, e0 + l + 64, ... in that order, eight per
    // step.  Every lane loads (past the chunk's end it reloads the last edge
    // and adds 0): loads under per-lane conditions make the compiler wait for
    // the whole memory queue at each, one round trip per load.
    float sum = 0.0f;
    for (uint32_t b = e0; b < e1; b += kUnrollPr * kWave) {
      uint32_t u[kUnrollPr];
      float c[kUnrollPr];
#pragma unroll
      for (int q = 0; q < kUnrollPr; ++q) u[q] = a.in_src[min(b + q * kWave + lane, e1 - 1)];
#pragma unroll
      for (int q = 0; q < kUnrollPr; ++q) c[q] = a.c_cur[u[q]];
#pragma unroll
      for (int q = 0; q < kUnrollPr; ++q) sum += b + q * kWave + lane < e1 ? c[q] : 0.0f;
    }
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) sum += __shfl_xor(sum, d);
    if (lane == 0) a.part[k] = sum;
  }
}

// One lane per heavy vertex: its partials added in chunk order.
__global__ void pr_finish_kernel(PrArgs a) {
  if (pr_done(a)) return;
  float dmax = 0.0f;
  const size_t nt = (size_t)gridDim.x * blockDim.x;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < a.nheavy; i += nt) {
    const uint32_t v = a.heavy[i];
    const uint32_t n = (a.in_row[v + 1] - a.in_row[v] + kChunk - 1) / kChunk;
    const float *pp = a.part + a.hbase[i];
    float sum = 0.0f;
    for (uint32_t k = 0; k < n; ++k) sum += pp[k];
    pr_store(a, v, sum, dmax);
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
  // p[2], c[2] (f32[V] each), the heavy list and its partials' bases (u32[V]
  // each), and per partial its heavy vertex (u32) and sum (f32): at most
  // V + E / kChunk partials
  const uint64_t maxchunks = (uint64_t)V + g->nedges / kChunk + 1;
  if ((rc = ws_buffer(w, kBufQueues, (6ull * V + 2 * maxchunks) * 4, &buf)) ||
      (rc = ws_buffer(w, kBufSeeds, 2048, &cbuf)))
    return rc;
  float *p[2] = {(float *)buf, (float *)buf + V};
  float *c[2] = {(float *)buf + 2ull * V, (float *)buf + 3ull * V};
  uint32_t *heavy = (uint32_t *)buf + 4ull * V;
  uint32_t *hbase = (uint32_t *)buf + 5ull * V;
  uint32_t *chunk = (uint32_t *)buf + 6ull * V;
  float *part = (float *)buf + 6ull * V + maxchunks;
  uint32_t *ctl = (uint32_t *)cbuf;  // heavy count, max change (x2), done, iterations, partial count
  const uint32_t vgrid = (uint32_t)std::min<uint64_t>(((uint64_t)V + 255) / 256, 16ull * w->cus);
  HIPCHK(hipEventRecord(w->ev0, w->stream));
  hipLaunchKernelGGL(pr_clear_kernel, dim3(1), dim3(64), 0, w->stream, ctl, 6u);
  hipLaunchKernelGGL(pr_init_kernel, dim3(vgrid), dim3(256), 0, w->stream, g->row_start, gt->row_start, V, alpha,
                     p[0], c[0], heavy, ctl);
  HIPCHK(hipGetLastError());
  uint32_t *hc = w->pinned;
  HIPCHK(hipMemcpyAsync(hc, ctl, 4, hipMemcpyDeviceToHost, w->stream));
  HIPCHK(hipStreamSynchronize(w->stream));
  const uint32_t nheavy = hc[0];
  const uint32_t fgrid = (uint32_t)std::min<uint64_t>(((uint64_t)nheavy + 255) / 256 + 1, 16ull * w->cus);
  if (nheavy) {
    hipLaunchKernelGGL(pr_chunks_kernel, dim3(fgrid), dim3(256), 0, w->stream, gt->row_start, heavy, nheavy, hbase,
                       chunk, ctl + kCtlChunks);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hc, ctl + kCtlChunks, 4, hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
  } else {
    hc[0] = 0;
  }
  const uint32_t nchunks = hc[0];
  if (nchunks > maxchunks) return fail(LSG_ERR_CAPACITY, "pagerank: %u partial sums exceed the buffer", nchunks);
  PrArgs a;
  a.row_start = g->row_start;
  a.in_row = gt->row_start;
  a.in_src = gt->edge_dst;
  a.heavy = heavy;
  a.hbase = hbase;
  a.chunk = chunk;
  a.part = part;
  a.nheavy = nheavy;
  a.nchunks = nchunks;
  a.nnodes = V;
  a.alpha = alpha;
  a.done = ctl + kCtlDone;
  const uint32_t cgrid = (uint32_t)std::min<uint64_t>(((uint64_t)nchunks * kWave + 255) / 256 + 1, 16ull * w->cus);
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
      if (nheavy) {
        hipLaunchKernelGGL(pr_chunk_kernel, dim3(cgrid), dim3(256), 0, w->stream, a);
        hipLaunchKernelGGL(pr_finish_kernel, dim3(fgrid), dim3(256), 0, w->stream, a);
      }
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
