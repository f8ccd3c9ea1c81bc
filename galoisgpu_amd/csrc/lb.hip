// The -l path of apps/sssp and apps/bfs on gfx950: host-driven rounds with an
// edge-balanced pass for the frontier's heavy vertices.
//
// Reference: gg_main_pipe_1 (apps/sssp/kernel.cu:318-354, apps/bfs/kernel.cu:
// 286-311).  Each near iteration runs Inspect_*_kernel_dev (:150-176), which
// lists the in-worklist vertices of degree >= DEGREE_LIMIT (the sweep's total
// thread count, rt/include/ggcuda.h:19-22) with their degrees; a device-wide
// inclusive prefix sum of those degrees (rt/include/thread_work.h:27-41);
// *_kernel_dev_TB_LB (:78-149), where thread t of the grid takes flattened
// edge t, t + nthreads, ... and finds its vertex and offset by binary search
// over the prefix (compute_src_and_offset, thread_work.h:51-79); then the
// ordinary sweep over the other vertices; the host swaps the piles and loops.
//
// Here the sweep and the edge-balanced pass relax with the same operators as
// the persistent kernel (sssp_dev.h), the pile rotation is the same leader
// step run by a one-block launch, and the host reads back the schedule once
// per round (the reference reads the worklist sizes back as often).  The
// reference bounds its per-thread loop by ceilf((float)total / nthreads),
// which under-covers once the total exceeds 2^24 edges (float has a 24-bit
// significand); the loop here is bounded by the integer total.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>

#include "engine.h"
#include "lsg_internal.h"
#include "runtime.h"
#include "scan_dev.h"
#include "sssp_dev.h"

namespace lsg {

constexpr int kLbBlock = 256;  // TB_SIZE of the reference's -l kernels (apps/sssp/kernel.cu:8)

// Lists of the edge-balanced pass, one entry per heavy vertex of the round.
struct LbLists {
  uint32_t *count;   // [0]: heavy vertices listed this round
  uint32_t *deg;     // their out-degrees (thread_work_wl)
  uint32_t *src;     // their ids (thread_src_wl)
  uint32_t *prefix;  // inclusive prefix sum of deg (thread_prefix_work_wl)
  uint32_t *tiles;   // scan tile totals
};

// Inspect_sssp_kernel_dev / Inspect_bfs_kernel_dev: every in-queue vertex of
// degree >= lb_limit (and, for SSSP, not INF and not deferred by a split or
// bucket far pile) goes to the lists.  One atomic per wavefront.
template <Algo A>
__global__ __launch_bounds__(kLbBlock) void lb_inspect_kernel(KArgs a, LbLists l) {
  __shared__ Sched ss;
  if (threadIdx.x == 0) sched_load(ss, &a.ctrl->s);
  __syncthreads();
  const Sched s = ss;
  const uint32_t nt = gridDim.x * kLbBlock;
  const uint32_t end = (s.in_count + kWave - 1) / kWave * kWave;  // whole wavefronts reach the ballot
  for (uint32_t i = blockIdx.x * kLbBlock + threadIdx.x; i < end; i += nt) {
    bool pick = false;
    uint32_t u = 0, deg = 0;
    if (i < s.in_count) {
      u = ld_coh(a.q[s.in_q] + i);
      if (u < a.nnodes) {
        deg = glb(a.row_start)[u + 1] - glb(a.row_start)[u];
        pick = deg >= a.lb_limit;
        if (pick && A == Algo::SSSP) {
          const uint32_t du = ld_coh(a.dist + u);
          pick = du != kInf && !(s.aux[0] && s.aux[1] != LSG_FAR_REFERENCE && (int64_t)du > s.curdelta());
        }
      }
    }
    const uint64_t m = __ballot(pick);
    if (m == 0) continue;
    uint32_t base = 0;
    if (lane_id() == (uint32_t)__builtin_ctzll(m)) base = atomic_add(l.count, (uint32_t)__builtin_popcountll(m));
    base = readlane(base, (uint32_t)__builtin_ctzll(m));
    if (pick) {
      const uint32_t pos = base + lanes_below(m);
      glb(l.deg)[pos] = deg;
      glb(l.src)[pos] = u;
    }
  }
}

// Inclusive prefix from the tile scan: prefix[i] = exclusive[i] + tile offset + deg[i].
__global__ void lb_prefix_kernel(LbLists l, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    glb(l.prefix)[i] += glb(l.tiles)[i / kScanTile] + glb(l.deg)[i];
}

// compute_src_and_offset (rt/include/thread_work.h:51-79): the list entry
// holding flattened edge w is the first index in [lo, hi] whose inclusive
// prefix exceeds w.
__device__ __forceinline__ uint32_t lb_find(const uint32_t *prefix, uint32_t lo, uint32_t hi, uint32_t w) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (prefix[mid] > w) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// sssp_kernel_dev_TB_LB / bfs_kernel_dev_TB_LB: the heavy vertices' edges,
// flattened, one per thread; a block's kLbBlock consecutive edges first find
// the range of list entries they span (thread 0, two searches over the whole
// list), then each thread searches only that range.
template <Algo A>
__global__ __launch_bounds__(kLbBlock) void lb_relax_kernel(KArgs a, LbLists l) {
  __shared__ TileShared<kLbBlock> sh;
  __shared__ Sched ss;
  __shared__ uint32_t first, last;
  if (threadIdx.x < 2 * (kLbBlock / kWave)) (&sh.ps.n[0][0])[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    sched_load(ss, &a.ctrl->s);
    sh.ps.local = 0;
    sh.ps.far_min = kInf;
  }
  __syncthreads();
  const Sched s = ss;
  const uint32_t n = *glb(l.count);
  const uint32_t total = n ? glb(l.prefix)[n - 1] : 0u;
  const uint32_t nt = gridDim.x * kLbBlock;
  for (uint32_t base = blockIdx.x * kLbBlock; base < total; base += nt) {
    if (threadIdx.x == 0) {
      first = lb_find(l.prefix, 0, n - 1, base);
      last = lb_find(l.prefix, first, n - 1, min(base + kLbBlock, total) - 1);
    }
    __syncthreads();
    const uint32_t w = base + threadIdx.x;
    const bool act = w < total;
    uint32_t e = 0, du = 0;
    if (act) {
      const uint32_t k = lb_find(l.prefix, first, last, w);
      const uint32_t u = glb(l.src)[k];
      e = glb(a.row_start)[u] + (w - (glb(l.prefix)[k] - glb(l.deg)[k]));
      du = ld_coh(a.dist + u);
    }
    relax_edge<A, kLbBlock>(a, s, e, du, act, sh.ps, true);
    __syncthreads();  // first / last are rewritten by the next block step
  }
  stage_drain<kLbBlock>(a, s, sh.ps);
  __syncthreads();
  if (threadIdx.x == 0) flush_far_min<kLbBlock>(a, sh.ps);
}

// The sweep (sssp_kernel / bfs_kernel): one round over the in-queue, vertices
// of degree >= lb_limit left to the edge-balanced pass.  Tiles as in the
// persistent kernel, then dynamic tickets.
template <Algo A, int BS>
__global__ __launch_bounds__(BS) void lb_sweep_kernel(KArgs a) {
  __shared__ TileShared<BS> sh;
  __shared__ Sched ss;
  __shared__ uint32_t tk;
  Ctrl *c = a.ctrl;
  const uint32_t nb = gridDim.x;
  if (threadIdx.x < 2 * (BS / kWave)) (&sh.ps.n[0][0])[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    sched_load(ss, &c->s);
    sh.ps.local = 0;
    sh.ps.far_min = kInf;
    sh.edges = 0;
  }
  __syncthreads();
  const Sched s = ss;
  const uint32_t ti = tile_items<BS>(s.in_count, nb);
  const uint32_t ntiles = (s.in_count + ti - 1) / ti;
  const uint32_t total = ntiles + s.heavy_count;
  uint32_t t = blockIdx.x;
  while (t < total) {
    if (t < ntiles) process_tile<A, BS>(a, s, t, ti, sh);
    else process_heavy<A, BS>(a, s, t - ntiles, sh);
    if (total <= nb) break;
    __syncthreads();
    if (threadIdx.x == 0) tk = nb + atomic_add(&c->ticket, 1u);
    __syncthreads();
    t = tk;
  }
  stage_drain<BS>(a, s, sh.ps);
  __syncthreads();
  if (threadIdx.x == 0) {
    if ((a.sched || a.feat) && sh.edges)
      __hip_atomic_fetch_add(&c->stat_edges, (uint64_t)sh.edges, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    flush_far_min<BS>(a, sh.ps);
  }
}

// The host's pile swaps (retry2 / advance2, curdelta += DELTA) as the
// persistent kernel's leader step, run by one block between rounds.
template <Algo A>
__global__ __launch_bounds__(kLbBlock) void lb_leader_kernel(KArgs a) {
  __shared__ Sched ls;
  __shared__ uint32_t labort, lfar;
  leader_step<A>(a, ls, labort, lfar);
}

// Runs one -l traversal from `seed` into dist.  Synchronous.
template <Algo A>
static int run_lb(Workspace *w, const CsrView &g, uint32_t seed, const lsg_opts &o, uint32_t *dist, lsg_stats *st) {
  int bs, rc;
  KArgs a;
  void *seedbuf;
  if ((rc = setup_single<A>(w, g, 1, o, dist, &a, &seedbuf, &bs))) return rc;
  // The reference launches its -l kernels on kernel_sizing's grid: 8 blocks of
  // 256 threads per CU (skelapp/skel.cu:37-43); DEGREE_LIMIT is that grid's
  // thread count.
  const uint32_t grid = 8u * (uint32_t)w->cus;
  a.lb_limit = o.lb_deg > 0 ? (uint32_t)o.lb_deg : grid * kLbBlock;
  // Vertices below the limit are the sweep's: none is split into chunks.
  a.heavy_deg = std::max(a.heavy_deg, a.lb_limit);
  const uint32_t V = g.nnodes;
  const uint32_t ntiles = (V + kScanTile - 1) / kScanTile;
  void *lbuf;
  if ((rc = ws_buffer(w, kBufLb, 256 + (3ull * V + ntiles + 64) * 4, &lbuf))) return rc;
  LbLists l;
  l.count = (uint32_t *)lbuf;
  l.deg = (uint32_t *)((char *)lbuf + 256);
  l.src = l.deg + V;
  l.prefix = l.src + V;
  l.tiles = l.prefix + V;
  uint32_t *scan_total = l.tiles + ntiles;
  HIPCHK(hipMemcpyAsync(seedbuf, &seed, 4, hipMemcpyHostToDevice, w->stream));

  using SweepFn = void (*)(KArgs);
  SweepFn sweep = bs == 512    ? (SweepFn)lb_sweep_kernel<A, 512>
                  : bs == 1024 ? (SweepFn)lb_sweep_kernel<A, 1024>
                               : (SweepFn)lb_sweep_kernel<A, 256>;
  auto t0 = std::chrono::steady_clock::now();
  const uint32_t init_blocks = std::min<uint32_t>((V + 255) / 256, 8u * w->cus);
  HIPCHK(hipEventRecord(w->ev0, w->stream));
  hipLaunchKernelGGL(init_kernel, dim3(init_blocks), dim3(256), 0, w->stream, a, a.delta);
  hipLaunchKernelGGL(seed_kernel, dim3(1), dim3(256), 0, w->stream, a, (const uint32_t *)seedbuf, 1u);
  HIPCHK(hipGetLastError());
  Ctrl hc;
  uint32_t nheavy = 0;
  uint64_t lb_edges = 0;
  for (;;) {
    // Inspect (t_work.reset_thread_work + Inspect_*_kernel_dev), then read the
    // list size back (thread_work_wl.nitems()).
    HIPCHK(hipMemsetAsync(l.count, 0, 4, w->stream));
    hipLaunchKernelGGL(lb_inspect_kernel<A>, dim3(grid), dim3(kLbBlock), 0, w->stream, a, l);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(w->pinned, l.count, 4, hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    nheavy = w->pinned[0];
    if (nheavy > V) return fail(LSG_ERR_CAPACITY, "-l inspect listed %u vertices of %u", nheavy, V);
    if (nheavy) {
      const uint32_t nt = (nheavy + kScanTile - 1) / kScanTile;
      hipLaunchKernelGGL(scan_tile_kernel, dim3(nt), dim3(kScanBlock), 0, w->stream, (const uint32_t *)l.deg, nheavy,
                         l.prefix, l.tiles);
      hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kScanBlock), 0, w->stream, l.tiles, nt, scan_total);
      hipLaunchKernelGGL(lb_prefix_kernel, dim3(std::min<uint32_t>(nt * 4, 4u * w->cus)), dim3(256), 0, w->stream, l,
                         nheavy);
      hipLaunchKernelGGL(lb_relax_kernel<A>, dim3(grid), dim3(kLbBlock), 0, w->stream, a, l);
      HIPCHK(hipGetLastError());
    }
    // the sweep over the other vertices, then the pile rotation
    hipLaunchKernelGGL(sweep, dim3(grid * kLbBlock / bs), dim3(bs), 0, w->stream, a);
    hipLaunchKernelGGL(lb_leader_kernel<A>, dim3(1), dim3(kLbBlock), 0, w->stream, a);
    HIPCHK(hipGetLastError());
    if (nheavy) {
      uint32_t tot = 0;
      HIPCHK(hipMemcpyAsync(&tot, scan_total, 4, hipMemcpyDeviceToHost, w->stream));
      HIPCHK(hipStreamSynchronize(w->stream));
      lb_edges += tot;
    }
    HIPCHK(hipMemcpyAsync(&hc, w->ctrl, kCtrlHostBytes, hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    if (hc.s.done || hc.abort != kAbortNone) break;
  }
  HIPCHK(hipEventRecord(w->ev1, w->stream));
  HIPCHK(hipEventSynchronize(w->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, w->ev0, w->ev1));
  if (hc.abort == kAbortCapacity) return fail(LSG_ERR_CAPACITY, "device queue capacity exceeded");
  if (hc.abort != kAbortNone) return fail(LSG_ERR_HIP, "-l traversal aborted (code %u)", hc.abort);
  if (st) {
    memset(st, 0, sizeof *st);
    st->ms_kernel = ms;
    st->iterations = A == Algo::SSSP ? hc.stat_phases : hc.stat_rounds;
    st->near_iters = hc.stat_rounds;
    st->items = hc.stat_items;
    st->heavy_chunks = lb_edges;  // -l: edges relaxed by the edge-balanced pass
    st->last_delta = (int32_t)std::min<int64_t>(hc.s.curdelta(), INT_MAX);
    st->grid_blocks = (int32_t)grid;
    st->block_size = bs;
    st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return LSG_OK;
}

int run_lb_single(int algo, Workspace *w, const CsrView &g, uint32_t seed, const lsg_opts &o, uint32_t *dist,
                  lsg_stats *st) {
  return algo == (int)Algo::SSSP ? run_lb<Algo::SSSP>(w, g, seed, o, dist, st)
                                 : run_lb<Algo::BFS>(w, g, seed, o, dist, st);
}

}  // namespace lsg
