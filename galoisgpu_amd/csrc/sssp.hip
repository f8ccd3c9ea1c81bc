// Single-source traversal on gfx950: fixed-delta near/far SSSP and
// level-synchronous BFS in one persistent, cooperatively launched kernel.
//
// Reference path: apps/sssp/kernel.cu (gg_main :455, persistent driver
// gg_main_pipe_1_gpu_gb :355-391, relax operator sssp_kernel_dev :177-306,
// far-pile dedupe remove_dups_dev :36-66) and apps/bfs/kernel.cu (gg_main
// :380, driver :312-334, operator bfs_kernel_dev :121-274).
//
// Layout in HBM (all u32): row_start[V+1], edge_dst[E], edge_data[E],
// dist[V]; three item queues of capacity V (in / near / far roles rotate),
// two heavy-chunk queues, near_stamp[V] and far_stamp[V] used to push a
// vertex at most once per round (near) and once per phase (far), and a
// control block.  The reference's queues hold 2E items each with an equal
// priority array (apps/sssp/kernel.cu:475, rt/include/worklist.h:82-113);
// deduplicating at push time bounds every queue by V instead.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "engine.h"
#include "lsg_internal.h"
#include "runtime.h"
#include "sched_mlp.h"
#include "sssp_dev.h"

namespace lsg {

// ---- host side ------------------------------------------------------------

using KernelFn = void (*)(KArgs);

template <Algo A>
static KernelFn pick_kernel(int bs) {
  switch (bs) {
    case 512: return persistent_kernel<A, 512>;
    case 1024: return persistent_kernel<A, 1024>;
    default: return persistent_kernel<A, 256>;
  }
}

// Runs one traversal from `nseeds` seeds (host array) into dist (device,
// g.nnodes entries).  Synchronous.  With graph_launch the init, seed and
// persistent kernels are replayed as one captured hipGraph.  With `resume`:
// continue from the distances already in dist instead (resume_seed_kernel's
// queue, curdelta resume->curdelta), not captured; the seeds are then not used.
template <Algo A>
static int run_core(Workspace *w, const CsrView &g, const uint32_t *seeds, uint32_t nseeds, const lsg_opts &o,
                    uint32_t *dist, bool graph_launch, lsg_stats *st, const Resume *resume = nullptr) {
  auto t0 = std::chrono::steady_clock::now();
  int bs, rc;
  KArgs a;
  void *seedbuf;
  if ((rc = setup_single<A>(w, g, nseeds, o, dist, &a, &seedbuf, &bs))) return rc;
  HIPCHK(hipMemcpyAsync(seedbuf, seeds, (size_t)nseeds * 4, hipMemcpyHostToDevice, w->stream));

  KernelFn fn = A == Algo::SSSP ? pick_kernel<Algo::SSSP>(bs) : pick_kernel<Algo::BFS>(bs);
  int grid = 0;
  if ((rc = resident_grid(w, (const void *)fn, bs, &grid))) return rc;
  // Default: 2 blocks per CU at most for SSSP, 4 for BFS.  Fewer blocks per
  // grid barrier beat more parallel tiles (256-thread blocks; full resident
  // grid 1280 -> 512): RMAT-20 SSSP 2.24 -> 2.05 ms, road 2048^2 10.2 -> 8.3
  // ms, RMAT-22 6.31 -> 6.23 ms, grid 256^2 12.9 -> 12.8 ms; BFS levels are
  // wider and cheaper, RMAT-20 BFS 0.56 ms at 1280, 0.55 at 1024, 0.60 at 512.
  if (o.grid_blocks > 0) grid = std::min(o.grid_blocks, grid);
  else grid = std::min(grid, (A == Algo::BFS ? 4 : 2) * (int)w->cus);

  Ctrl hc;
  float ms = 0;
  for (int attempt = 0;; ++attempt) {
    const uint32_t init_blocks = std::min<uint32_t>((g.nnodes + 255) / 256, 8u * w->cus);
    const uint32_t *sd = (const uint32_t *)seedbuf;
    auto enqueue = [&]() -> int {
      if (resume) {
        hipLaunchKernelGGL(resume_init_kernel, dim3(init_blocks), dim3(256), 0, w->stream, a, a.delta,
                           resume->curdelta);
        hipLaunchKernelGGL(resume_seed_kernel, dim3(init_blocks), dim3(256), 0, w->stream, a, resume->lo,
                           resume->hi);
      } else {
        hipLaunchKernelGGL(init_kernel, dim3(init_blocks), dim3(256), 0, w->stream, a, a.delta);
        hipLaunchKernelGGL(seed_kernel, dim3(1), dim3(256), 0, w->stream, a, sd, nseeds);
      }
      hipLaunchKernelGGL(fn, dim3(grid), dim3(bs), 0, w->stream, a);
      HIPCHK(hipGetLastError());
      return LSG_OK;
    };
    HIPCHK(hipEventRecord(w->ev0, w->stream));
    if (graph_launch && !resume) {
      std::vector<uint8_t> key;
      key_append(key, (int)A);
      key_append(key, a);
      key_append(key, grid);
      key_append(key, bs);
      key_append(key, nseeds);
      if ((rc = launch_captured(w, key, enqueue))) return rc;
    } else if ((rc = enqueue())) {
      return rc;
    }
    HIPCHK(hipEventRecord(w->ev1, w->stream));
    HIPCHK(hipMemcpyAsync(&hc, w->ctrl, kCtrlHostBytes, hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    HIPCHK(hipEventElapsedTime(&ms, w->ev0, w->ev1));
    if (hc.abort != kAbortResidency) break;
    // Fewer blocks were resident than the occupancy query promised: remember
    // the count that was, and run again on that grid.
    if (attempt > 0 || hc.resident_seen == 0 || hc.resident_seen >= (uint32_t)grid)
      return fail(LSG_ERR_HIP, "persistent grid of %d x %d blocks is not co-resident (%u were)", grid, bs,
                  hc.resident_seen);
    grid = (int)hc.resident_seen;
    record_resident_grid(w, (const void *)fn, bs, grid);
  }
  if (hc.abort == kAbortTimeout)
    return fail(LSG_ERR_TIMEOUT,
                "grid barrier timed out (grid %d x %d) at round %u phase %u done %u with %u of %u arrival "
                "groups complete, generation %u; distances are incomplete",
                grid, bs, hc.s.round, hc.s.phase, hc.s.done, hc.bar_count % ((grid + kBarGroup - 1) / kBarGroup),
                (grid + kBarGroup - 1) / kBarGroup, hc.bar_gen);
  if (hc.abort == kAbortCapacity) return fail(LSG_ERR_CAPACITY, "device queue capacity exceeded");
  if (o.trace) {
    uint32_t n = (uint32_t)std::min<uint64_t>(hc.stat_rounds, kTraceCap);
    w->last_trace.resize(n);
    if (n) HIPCHK(hipMemcpy(w->last_trace.data(), a.trace, n * sizeof(lsg_round_trace), hipMemcpyDeviceToHost));
    // one feature record per phase boundary: phase p -> p + 1 is record p
    uint32_t np = a.feat ? std::min<uint32_t>(hc.s.phase, kMaxForced) : 0u;
    w->last_feat.resize((size_t)np * kFeatRec);
    if (np) HIPCHK(hipMemcpy(w->last_feat.data(), a.feat, w->last_feat.size() * sizeof(float), hipMemcpyDeviceToHost));
  }
  if (st) {
    memset(st, 0, sizeof *st);
    st->ms_kernel = ms;
    st->iterations = A == Algo::SSSP ? hc.stat_phases : hc.stat_rounds;
    st->near_iters = hc.stat_rounds;
    st->items = hc.stat_items;
    st->heavy_chunks = hc.stat_heavy;
    st->last_delta = (int32_t)std::min<int64_t>(hc.s.curdelta(), INT_MAX);
    st->grid_blocks = grid;
    st->block_size = bs;
    st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return LSG_OK;
}

// One source on one device graph into dist (device array of g->nnodes).
template <Algo A>
static int run_single(const lsg_graph *g, const lsg_opts *o_in, uint32_t *dist, lsg_stats *st) {
  auto t0 = std::chrono::steady_clock::now();
  if (!g || !g->device_graph) return fail(LSG_ERR_ARG, "expected a device graph (copy_to_gpu first)");
  if (!dist) return fail(LSG_ERR_ARG, "NULL distance array");
  lsg_opts o;
  if (o_in) o = *o_in;
  else lsg_opts_default(&o);
  if (o.start_node < 0 || (uint32_t)o.start_node >= g->nnodes)
    return fail(LSG_ERR_ARG, "start node %d out of range [0, %u)", o.start_node, g->nnodes);
  Workspace *w;
  int rc = get_workspace(g->device, &w);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(w->mu);
  HIPCHK(hipSetDevice(w->device));
  CsrView v{g->row_start, g->edge_dst, (const uint32_t *)g->edge_data, g->nnodes, g->nedges};
  uint32_t seed = (uint32_t)o.start_node;
  if (o.single_block && o.enable_lb) return fail(LSG_ERR_ARG, "single_block and enable_lb (-l) exclude each other");
  if (o.single_block) {
    lsg_stats bst;
    Resume res;
    rc = run_block_single((int)A, w, v, seed, o, dist, &bst, &res);
    if (rc == kBlockFallback) {
      // The piles overflowed: the persistent kernel continues from the
      // distances the workgroup reached (not captured: a one-off launch).
      rc = run_core<A>(w, v, &seed, 1, o, dist, false, st, &res);
      if (rc == LSG_OK && st) {
        st->ms_kernel += bst.ms_kernel;
        st->items += bst.items;
        st->near_iters += bst.near_iters;
        st->iterations += bst.iterations;
      }
    } else if (rc == LSG_OK && st) {
      *st = bst;
    }
  } else {
    rc = o.enable_lb ? run_lb_single((int)A, w, v, seed, o, dist, st)
                     : run_core<A>(w, v, &seed, 1, o, dist, o.graph_launch != 0, st);
  }
  if (rc == LSG_OK && st)
    st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

// ---- many graphs in one launch -----------------------------------------------

// One member graph of a packed union: its arrays, output, and offsets.
struct UnionMember {
  const uint32_t *row, *dst, *w;
  int32_t *out;
  uint32_t nnodes, voff;
  uint64_t nedges, eoff;
};

__device__ __forceinline__ uint32_t member_of_vertex(const UnionMember *m, uint32_t n, uint32_t v) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (m[mid].voff <= v) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint32_t member_of_edge(const UnionMember *m, uint32_t n, uint64_t e) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (m[mid].eoff <= e) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Disjoint union: vertex ids shifted by voff, edge offsets by eoff.
__global__ void pack_union_kernel(const UnionMember *m, uint32_t n, uint32_t *urow, uint32_t *udst, uint32_t *uw,
                                  uint32_t V, uint64_t E) {
  size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x, nt = (size_t)gridDim.x * blockDim.x;
  for (size_t v = tid; v < V; v += nt) {
    const UnionMember &g = m[member_of_vertex(m, n, (uint32_t)v)];
    urow[v] = (uint32_t)(g.row[v - g.voff] + g.eoff);
  }
  if (tid == 0) urow[V] = (uint32_t)E;
  for (size_t e = tid; e < E; e += nt) {
    const UnionMember &g = m[member_of_edge(m, n, e)];
    udst[e] = g.dst[e - g.eoff] + g.voff;
    uw[e] = g.w[e - g.eoff];
  }
}

__global__ void unpack_union_kernel(const UnionMember *m, uint32_t n, const uint32_t *udist, uint32_t V) {
  size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x, nt = (size_t)gridDim.x * blockDim.x;
  for (size_t v = tid; v < V; v += nt) {
    const UnionMember &g = m[member_of_vertex(m, n, (uint32_t)v)];
    g.out[v - g.voff] = (int32_t)udist[v];
  }
}

// Checks a device CSR in place: row_start[0] == 0, offsets never decrease,
// row_start[V] == E, every destination < V, every weight >= 0.  bad[0] gets
// the first failing check (1 offsets, 2 destination, 3 weight) -- any one, by
// atomicMin over (check, index) -- and bad[1] its index.
__global__ void validate_csr_kernel(const uint32_t *row, const uint32_t *dst, const int32_t *w, uint32_t V,
                                    uint32_t E, unsigned long long *bad) {
  const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x, nt = (size_t)gridDim.x * blockDim.x;
  if (tid == 0 && (row[0] != 0 || row[V] != E)) atomicMin(bad, 1ull << 32);
  for (size_t i = tid; i < V; i += nt)
    if (row[i + 1] < row[i]) atomicMin(bad, (1ull << 32) | i);
  for (size_t e = tid; e < E; e += nt) {
    if (dst[e] >= V) atomicMin(bad, (2ull << 32) | e);
    else if (w && w[e] < 0) atomicMin(bad, (3ull << 32) | e);
  }
}

}  // namespace lsg

using namespace lsg;

extern "C" {

int lsg_set_device(int device) {
  HIPCHK(hipSetDevice(device));
  return LSG_OK;
}

int lsg_device_name(int device, char *buf, size_t len) {
  if (!buf || len == 0) return fail(LSG_ERR_ARG, "device_name: no buffer");
  hipDeviceProp_t p;
  HIPCHK(hipGetDeviceProperties(&p, device));
  snprintf(buf, len, "%s (%s, %d CUs)", p.name[0] ? p.name : "AMD GPU", p.gcnArchName, p.multiProcessorCount);
  return LSG_OK;
}

int lsg_device_sync(int device) {
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipDeviceSynchronize());
  return LSG_OK;
}

int lsg_graph_copy_to_gpu(const lsg_graph *h, lsg_graph *d, int device) {
  if (!h || !d) return fail(LSG_ERR_ARG, "copy_to_gpu: NULL graph");
  int rc = validate_host_csr(h);
  if (rc) return rc;
  memset(d, 0, sizeof *d);
  HIPCHK(hipSetDevice(device));
  size_t ne = h->nedges ? h->nedges : 1;
  if (hipMalloc(&d->row_start, ((size_t)h->nnodes + 1) * 4) != hipSuccess ||
      hipMalloc(&d->edge_dst, ne * 4) != hipSuccess || hipMalloc(&d->edge_data, ne * 4) != hipSuccess ||
      hipMalloc(&d->node_data, (size_t)h->nnodes * 4) != hipSuccess) {
    (void)hipFree(d->row_start);
    (void)hipFree(d->edge_dst);
    (void)hipFree(d->edge_data);
    (void)hipFree(d->node_data);
    memset(d, 0, sizeof *d);
    return fail(LSG_ERR_NOMEM, "copy_to_gpu: device allocation failed");
  }
  d->nnodes = h->nnodes;
  d->nedges = h->nedges;
  d->device_graph = 1;
  d->device = device;
  HIPCHK(hipMemcpy(d->edge_dst, h->edge_dst, (size_t)h->nedges * 4, hipMemcpyHostToDevice));
  if (h->edge_data) HIPCHK(hipMemcpy(d->edge_data, h->edge_data, (size_t)h->nedges * 4, hipMemcpyHostToDevice));
  else HIPCHK(hipMemset(d->edge_data, 0, (size_t)h->nedges * 4));
  if (h->node_data) HIPCHK(hipMemcpy(d->node_data, h->node_data, (size_t)h->nnodes * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d->row_start, h->row_start, ((size_t)h->nnodes + 1) * 4, hipMemcpyHostToDevice));
  return LSG_OK;
}

int lsg_graph_copy_to_cpu(const lsg_graph *d, lsg_graph *h) {
  if (!d || !h || !d->device_graph || h->device_graph) return fail(LSG_ERR_ARG, "copy_to_cpu: need device -> host");
  if (d->nnodes != h->nnodes) return fail(LSG_ERR_ARG, "copy_to_cpu: node counts differ");
  if (!d->node_data || !h->node_data) return fail(LSG_ERR_ARG, "copy_to_cpu: missing node_data");
  HIPCHK(hipSetDevice(d->device));
  HIPCHK(hipMemcpy(h->node_data, d->node_data, (size_t)d->nnodes * 4, hipMemcpyDeviceToHost));
  return LSG_OK;
}

void lsg_graph_dealloc(lsg_graph *g) {
  if (!g) return;
  if (g->device_graph == 1) {
    (void)hipSetDevice(g->device);
    (void)hipFree(g->row_start);
    (void)hipFree(g->edge_dst);
    (void)hipFree(g->edge_data);
    (void)hipFree(g->node_data);
    memset(g, 0, sizeof *g);
  } else if (g->device_graph == 0) {
    free(g->row_start);
    free(g->edge_dst);
    free(g->edge_data);
    free(g->node_data);
    memset(g, 0, sizeof *g);
  } else {
    memset(g, 0, sizeof *g);  // wrapped caller-owned arrays
  }
}

int lsg_graph_wrap_device(lsg_graph *d, uint32_t nnodes, uint32_t nedges, uint32_t *row_start, uint32_t *edge_dst,
                          int32_t *edge_data, int32_t *node_data, int device) {
  clear_error();
  if (!d || !row_start || (nedges && (!edge_dst || !edge_data)) || nnodes == 0 || nnodes == UINT32_MAX)
    return fail(LSG_ERR_ARG, "wrap_device: bad arguments");
  // The kernels trust the CSR (an out-of-range destination would make their
  // atomics write out of bounds), so it is checked on the device first.
  Workspace *w;
  int rc = get_workspace(device, &w);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(w->mu);
    HIPCHK(hipSetDevice(device));
    void *vb;
    if ((rc = ws_buffer(w, kBufValidate, 256, &vb))) return rc;
    unsigned long long *bad = (unsigned long long *)vb;
    const unsigned long long none = ~0ull;
    HIPCHK(hipMemcpyAsync(bad, &none, 8, hipMemcpyHostToDevice, w->stream));
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(((uint64_t)nnodes + nedges + 255) / 256 + 1, 16ull * w->cus);
    hipLaunchKernelGGL(validate_csr_kernel, dim3(blocks), dim3(256), 0, w->stream, (const uint32_t *)row_start,
                       (const uint32_t *)edge_dst, (const int32_t *)edge_data, nnodes, nedges, bad);
    HIPCHK(hipGetLastError());
    unsigned long long hb = 0;
    HIPCHK(hipMemcpyAsync(&hb, bad, 8, hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    const uint32_t what = (uint32_t)(hb >> 32), at = (uint32_t)hb;
    if (what == 1) return fail(LSG_ERR_FORMAT, "wrap_device: row offsets are not a CSR (row_start[0] = 0, "
                                               "non-decreasing, row_start[%u] = %u) near node %u", nnodes, nedges, at);
    if (what == 2) return fail(LSG_ERR_FORMAT, "wrap_device: edge %u has destination >= %u", at, nnodes);
    if (what == 3) return fail(LSG_ERR_FORMAT, "wrap_device: edge %u has a negative weight", at);
  }
  memset(d, 0, sizeof *d);
  d->nnodes = nnodes;
  d->nedges = nedges;
  d->row_start = row_start;
  d->edge_dst = edge_dst;
  d->edge_data = edge_data;
  d->node_data = node_data;
  d->device_graph = 2;  // device arrays owned by the caller
  d->device = device;
  return LSG_OK;
}

int lsg_sssp_dist(const lsg_graph *gg, const lsg_opts *opts, int32_t *dist_dev, lsg_stats *stats) {
  clear_error();
  return run_single<Algo::SSSP>(gg, opts, (uint32_t *)dist_dev, stats);
}

int lsg_bfs_dist(const lsg_graph *gg, const lsg_opts *opts, int32_t *level_dev, lsg_stats *stats) {
  clear_error();
  return run_single<Algo::BFS>(gg, opts, (uint32_t *)level_dev, stats);
}

int lsg_sssp_graphs(const lsg_graph *graphs, uint32_t ngraphs, const uint32_t *srcs, const lsg_opts *opts,
                    int32_t *const *dist_out, lsg_batch_stats *stats) {
  clear_error();
  auto t0 = std::chrono::steady_clock::now();
  if (!graphs || !srcs || !dist_out || ngraphs == 0) return fail(LSG_ERR_ARG, "sssp_graphs: empty or NULL input");
  lsg_opts o;
  if (opts) o = *opts;
  else lsg_opts_default(&o);
  int device = graphs[0].device;
  std::vector<UnionMember> mem(ngraphs);
  uint64_t V = 0, E = 0;
  std::vector<uint32_t> seeds(ngraphs);
  for (uint32_t i = 0; i < ngraphs; ++i) {
    const lsg_graph &g = graphs[i];
    if (!g.device_graph || g.device != device) return fail(LSG_ERR_ARG, "graph %u is not a device graph on device %d", i, device);
    if (srcs[i] >= g.nnodes) return fail(LSG_ERR_ARG, "source of graph %u (%u) out of range [0, %u)", i, srcs[i], g.nnodes);
    if (!dist_out[i]) return fail(LSG_ERR_ARG, "output %u is NULL", i);
    mem[i] = UnionMember{g.row_start, g.edge_dst, (const uint32_t *)g.edge_data, dist_out[i], g.nnodes,
                         (uint32_t)V, g.nedges, E};
    seeds[i] = (uint32_t)V + srcs[i];
    V += g.nnodes;
    E += g.nedges;
    if (V >= UINT32_MAX || E >= UINT32_MAX) return fail(LSG_ERR_ARG, "graphs exceed the u32 index range together");
  }
  Workspace *w;
  int rc = get_workspace(device, &w);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(w->mu);
  HIPCHK(hipSetDevice(device));
  if (o.single_block) {
    // one workgroup per graph; a pile overflow runs the union launch instead
    if (o.sched || o.trace || o.enable_lb || o.graph_launch)
      return fail(LSG_ERR_ARG, "single_block graphs run without sched, trace, enable_lb and graph_launch");
    rc = run_graphs_block(w, graphs, ngraphs, srcs, dist_out, o, stats);
    if (rc != kBlockFallback) {
      if (rc == LSG_OK && stats)
        stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      return rc;
    }
  }
  size_t mem_bytes = ngraphs * sizeof(UnionMember);
  size_t bytes = (V + 1) * 4 + E * 8 + V * 4 + mem_bytes + 256;
  void *ub;
  if ((rc = ws_buffer(w, kBufUnion, bytes, &ub))) return rc;
  uint32_t *urow = (uint32_t *)ub, *udst = urow + V + 1, *uw = udst + E, *udist = uw + E;
  UnionMember *dmem = (UnionMember *)(((uintptr_t)(udist + V) + 255) & ~(uintptr_t)255);
  HIPCHK(hipMemcpyAsync(dmem, mem.data(), mem_bytes, hipMemcpyHostToDevice, w->stream));
  uint32_t blocks = (uint32_t)std::min<uint64_t>((E + V + 255) / 256 + 1, 16ull * w->cus);
  hipLaunchKernelGGL(pack_union_kernel, dim3(blocks), dim3(256), 0, w->stream, dmem, ngraphs, urow, udst, uw,
                     (uint32_t)V, E);
  HIPCHK(hipGetLastError());
  CsrView v{urow, udst, uw, (uint32_t)V, E};
  lsg_stats st;
  if ((rc = run_core<Algo::SSSP>(w, v, seeds.data(), ngraphs, o, udist, o.graph_launch != 0, &st))) return rc;
  hipLaunchKernelGGL(unpack_union_kernel, dim3(blocks), dim3(256), 0, w->stream, dmem, ngraphs, udist, (uint32_t)V);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(w->stream));
  if (stats) {
    memset(stats, 0, sizeof *stats);
    stats->ms_kernel = st.ms_kernel;
    stats->rounds = st.near_iters;
    stats->phases = st.iterations;
    stats->items = st.items;
    stats->heavy_chunks = st.heavy_chunks;
    stats->grid_blocks = st.grid_blocks;
    stats->block_size = st.block_size;
    stats->launches = 1;
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return LSG_OK;
}

int lsg_last_trace(int device, lsg_round_trace *out, uint32_t cap, uint32_t *n, double *tick_hz) {
  Workspace *w;
  int rc = get_workspace(device, &w);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(w->mu);
  uint32_t m = (uint32_t)std::min<size_t>(cap, w->last_trace.size());
  if (out && m) memcpy(out, w->last_trace.data(), m * sizeof(lsg_round_trace));
  if (n) *n = (uint32_t)w->last_trace.size();
  if (tick_hz) *tick_hz = w->wall_khz * 1000.0;
  return LSG_OK;
}

int lsg_sssp_gg_main(lsg_graph *hg, lsg_graph *gg, const lsg_opts *opts, lsg_stats *stats) {
  (void)hg;
  if (!gg) return fail(LSG_ERR_ARG, "gg_main: NULL device graph");
  return lsg_sssp_dist(gg, opts, gg->node_data, stats);
}

int lsg_bfs_gg_main(lsg_graph *hg, lsg_graph *gg, const lsg_opts *opts, lsg_stats *stats) {
  (void)hg;
  if (!gg) return fail(LSG_ERR_ARG, "gg_main: NULL device graph");
  return lsg_bfs_dist(gg, opts, gg->node_data, stats);
}

}  // extern "C"
