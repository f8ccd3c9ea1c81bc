// Single-source SSSP that allocates nothing: lsg_sssp_dist_noalloc.
//
// No counterpart in the reference.  For callers that cannot allocate device
// memory during a call (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc): every buffer the traversal needs beyond
// the caller's graph and distance array -- the control block, the three item
// queues, the two heavy-chunk queues, the push stamps and the seed word --
// lives in one file-scope device array, `scratch`, sized for graphs of at
// most kNoAllocMaxNodes vertices and kNoAllocMaxEdges edges.  The kernels are
// the persistent single-source kernels of lsg_sssp_dist (sssp_dev.h), so the
// distances are the same bit for bit.
//
// Each GPU holds its own copy of `scratch` for the life of the process.  Host
// code never passes the array's name to a kernel (in host code it is a
// host-side stand-in, not a device address): the first call on a device,
// with that device current, looks up the device address with
// hipGetSymbolAddress and keeps it per device; every use goes through that
// pointer.  Calls share the array: overlapping calls on one GPU (from other
// threads without the per-device lock, other streams or other processes'
// launches of this library on the same context) race; each call re-initialises
// what must start cleared (counters, queues' tails, stamps, distances) with
// kernels, not memsets.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <mutex>

#include "engine.h"
#include "lsg_internal.h"
#include "runtime.h"
#include "sssp_dev.h"

namespace lsg {

constexpr uint32_t kNoAllocMaxNodes = 1u << 20;  // 1,048,576 vertices
constexpr uint32_t kNoAllocMaxEdges = 1u << 24;  // 16,777,216 edges
constexpr uint32_t kNoAllocMinHeavy = 256;       // chunk size floor (the smallest block size)
constexpr uint32_t kNoAllocHcap = 2u * kNoAllocMaxEdges / kNoAllocMinHeavy + 64;

// Layout of `scratch` (byte offsets, each 256-aligned).
constexpr size_t align256(size_t x) { return (x + 255) / 256 * 256; }
constexpr size_t kOffCtrl = 0;
constexpr size_t kOffQueues = align256(kOffCtrl + sizeof(Ctrl));
constexpr size_t kOffHeavy = align256(kOffQueues + 3ull * kNoAllocMaxNodes * 4);
constexpr size_t kOffStamps = align256(kOffHeavy + 2ull * kNoAllocHcap * 8);
constexpr size_t kOffSeed = align256(kOffStamps + 2ull * kNoAllocMaxNodes * 4);
constexpr size_t kScratchBytes = align256(kOffSeed + 4);  // about 22.4 MB

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

struct NoAllocDevice {
  int device = -1;
  uint8_t *ptr = nullptr;  // device address of this device's `scratch`
  int cus = 0;
  double wall_khz = 0;
  int grid[3] = {0, 0, 0};  // calibrated persistent grid per block size (256, 512, 1024)
  std::mutex mu;
};

static std::mutex g_na_mu;
static NoAllocDevice g_na[64];

// Set-up of `device` (current): the scratch address and device figures, once.
static int noalloc_device(int device, NoAllocDevice **out) {
  if (device < 0 || device >= 64) return fail(LSG_ERR_ARG, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_na_mu);
  NoAllocDevice &d = g_na[device];
  if (d.device != device) {
    HIPCHK(hipSetDevice(device));
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(LSG_ERR_HIP, "no device address for the scratch array on device %d", device);
    HIPCHK(hipDeviceGetAttribute(&d.cus, hipDeviceAttributeMultiprocessorCount, device));
    int khz = 0;
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
    d.wall_khz = khz > 0 ? khz : 100000.0;
    d.ptr = (uint8_t *)p;
    d.device = device;
  }
  *out = &d;
  return LSG_OK;
}

template <int BS>
static int noalloc_run(NoAllocDevice &nd, hipStream_t stream, const lsg_graph *g, const lsg_opts &o, uint32_t *dist,
                       int slot, lsg_stats *st) {
  KArgs a;
  memset(&a, 0, sizeof a);
  uint8_t *base = nd.ptr;
  a.row_start = g->row_start;
  a.edge_dst = g->edge_dst;
  a.edge_data = (const uint32_t *)g->edge_data;
  a.dist = dist;
  a.ctrl = (Ctrl *)(base + kOffCtrl);
  for (int i = 0; i < 3; ++i) a.q[i] = (uint32_t *)(base + kOffQueues) + (size_t)i * g->nnodes;
  a.hq[0] = (uint32_t *)(base + kOffHeavy);
  a.hq[1] = (uint32_t *)(base + kOffHeavy) + 2ull * kNoAllocHcap;
  a.near_stamp = (uint32_t *)(base + kOffStamps);
  a.far_stamp = a.near_stamp + g->nnodes;
  a.nnodes = g->nnodes;
  a.qcap = g->nnodes;
  a.hcap = kNoAllocHcap;
  a.delta = (uint32_t)o.delta;
  a.thread_deg = o.thread_deg > 0 ? (uint32_t)o.thread_deg : 8u;
  a.heavy_deg = std::max<uint32_t>(o.heavy_deg > 0 ? (uint32_t)o.heavy_deg : 512u, (uint32_t)BS);
  a.expand_mode = (uint32_t)o.expand_mode;
  a.far_mode = (uint32_t)o.far_mode;
  double sec = 20.0;
  if (const char *e = getenv("LSG_BARRIER_TIMEOUT_S")) sec = std::max(0.0, atof(e));
  a.spin_limit = (uint64_t)(sec * nd.wall_khz * 1000.0);
  a.residency_limit = (uint64_t)(0.005 * nd.wall_khz * 1000.0);
  a.nedges = g->nedges;
  a.lb_limit = UINT32_MAX;
  if (o.sched) {
    a.sched = 1;
    a.mlp = sched_model_device(nd.device, LSG_SCHED_PLAIN);
    if (!a.mlp)
      return fail(LSG_ERR_ARG, "learned scheduler requested but no plain-kernel model is loaded (lsg_sched_set_model)");
  }
  uint32_t *seed = (uint32_t *)(base + kOffSeed);
  const uint32_t src = (uint32_t)o.start_node;
  HIPCHK(hipMemcpyAsync(seed, &src, 4, hipMemcpyHostToDevice, stream));

  void (*fn)(KArgs) = persistent_kernel<Algo::SSSP, BS>;
  if (nd.grid[slot] == 0) {
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn, BS, 0));
    if (per_cu <= 0) return fail(LSG_ERR_HIP, "persistent kernel cannot be resident (block %d)", BS);
    nd.grid[slot] = std::min<int>(per_cu * nd.cus, (int)kMaxGridBlocks);
  }
  int grid = o.grid_blocks > 0 ? std::min(o.grid_blocks, nd.grid[slot]) : std::min(nd.grid[slot], 2 * nd.cus);
  auto t0 = std::chrono::steady_clock::now();
  Ctrl hc;
  for (int attempt = 0;; ++attempt) {
    const uint32_t init_blocks = std::min<uint32_t>((g->nnodes + 255) / 256, 8u * nd.cus);
    hipLaunchKernelGGL(init_kernel, dim3(init_blocks), dim3(256), 0, stream, a, a.delta);
    hipLaunchKernelGGL(seed_kernel, dim3(1), dim3(256), 0, stream, a, (const uint32_t *)seed, 1u);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(BS), 0, stream, a);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&hc, a.ctrl, kCtrlHostBytes, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (hc.abort != kAbortResidency) break;
    if (attempt > 0 || hc.resident_seen == 0 || hc.resident_seen >= (uint32_t)grid)
      return fail(LSG_ERR_HIP, "persistent grid of %d x %d blocks is not co-resident (%u were)", grid, BS,
                  hc.resident_seen);
    grid = (int)hc.resident_seen;
    nd.grid[slot] = grid;
  }
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (hc.abort == kAbortTimeout) return fail(LSG_ERR_TIMEOUT, "grid barrier timed out; distances are incomplete");
  if (hc.abort == kAbortCapacity) return fail(LSG_ERR_CAPACITY, "device queue capacity exceeded");
  if (st) {
    memset(st, 0, sizeof *st);
    st->ms_kernel = ms;
    st->ms_total = ms;
    st->iterations = hc.stat_phases;
    st->near_iters = hc.stat_rounds;
    st->items = hc.stat_items;
    st->heavy_chunks = hc.stat_heavy;
    st->last_delta = (int32_t)std::min<int64_t>(hc.s.curdelta(), INT_MAX);
    st->grid_blocks = grid;
    st->block_size = BS;
  }
  return LSG_OK;
}

}  // namespace lsg

using namespace lsg;

extern "C" {

int lsg_sssp_dist_noalloc(const lsg_graph *gg, const lsg_opts *opts, int32_t *dist_dev, lsg_stats *stats) {
  clear_error();
  if (!gg || !gg->device_graph) return fail(LSG_ERR_ARG, "expected a device graph (copy_to_gpu first)");
  if (!dist_dev) return fail(LSG_ERR_ARG, "NULL distance array");
  if (gg->nnodes > kNoAllocMaxNodes || gg->nedges > kNoAllocMaxEdges)
    return fail(LSG_ERR_ARG,
                "graph of %u vertices and %u edges exceeds the no-allocation limit of %u vertices and %u edges",
                gg->nnodes, gg->nedges, kNoAllocMaxNodes, kNoAllocMaxEdges);
  lsg_opts o;
  if (opts) o = *opts;
  else lsg_opts_default(&o);
  if (o.start_node < 0 || (uint32_t)o.start_node >= gg->nnodes)
    return fail(LSG_ERR_ARG, "start node %d out of range [0, %u)", o.start_node, gg->nnodes);
  int bs, expand, rc;
  if ((rc = check_single_opts(o, true, &bs, &expand))) return rc;
  if (o.trace || o.graph_launch || o.enable_lb || o.single_block)
    return fail(LSG_ERR_ARG, "trace, graph_launch, -l and single_block need workspace buffers; use lsg_sssp_dist");
  NoAllocDevice *nd;
  if ((rc = noalloc_device(gg->device, &nd))) return rc;
  std::lock_guard<std::mutex> lk(nd->mu);
  HIPCHK(hipSetDevice(gg->device));
  hipStream_t stream = nullptr;  // the device's null stream: nothing to create
  uint32_t *dist = (uint32_t *)dist_dev;
  switch (bs) {
    case 512: return noalloc_run<512>(*nd, stream, gg, o, dist, 1, stats);
    case 1024: return noalloc_run<1024>(*nd, stream, gg, o, dist, 2, stats);
    default: return noalloc_run<256>(*nd, stream, gg, o, dist, 0, stats);
  }
}

}  // extern "C"
