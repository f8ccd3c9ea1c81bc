// One-workgroup asynchronous traversal for small, sparse, high-diameter
// graphs (lsg_opts.single_block).
//
// On a graph like a 256x256 grid a near/far round relaxes a few hundred
// vertices and costs a fixed chain of dependent memory round trips and
// barriers (about 6 us), and a traversal is 1000+ rounds.  Here one
// workgroup of 1024 threads (one CU) works through an LDS ring of queued
// vertices with no rounds at all: every wavefront claims up to 64 queued
// vertices, relaxes their edges, and appends the vertices it improved; a
// vertex is queued at most once at a time (a per-vertex flag).  This is
// label-correcting SSSP (asynchronous Bellman-Ford): the order differs from
// the reference's near/far piles (apps/sssp/kernel.cu:355-391), the
// fixpoint -- and so every distance -- is the same.  BFS is the same with
// unit weights.
//
// The ring holds kRing vertices; a traversal whose queue would outgrow it
// stops (kAbortCapacity) and the host runs the persistent kernel instead.
// Every wait is bounded by the wall clock (no progress for the barrier spin
// limit raises kAbortTimeout).

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>

#include "sssp_dev.h"

namespace lsg {

constexpr uint32_t kBlkThreads = 1024;
constexpr uint32_t kRing = 16384;  // queued vertices held in LDS (64 KB)
constexpr uint32_t kSlotEmpty = 0xFFFFFFFFu;
static_assert((kRing & (kRing - 1)) == 0, "kRing is a power of two");

struct BlockShared {
  uint32_t ring[kRing];
  uint32_t head;   // next slot to claim (grows without bound; slot = head % kRing)
  uint32_t tail;   // next slot to fill
  uint32_t busy;   // wavefronts between a claim attempt and the end of its work
  uint32_t done;
  uint32_t abort;
  unsigned long long items;     // vertices expanded
  unsigned long long progress;  // wall clock of the last expansion
};

__device__ __forceinline__ uint32_t lds_ld(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_st(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <Algo A>
__global__ __launch_bounds__(kBlkThreads) void block_kernel(KArgs a) {
  __shared__ BlockShared sh;
  Ctrl *c = a.ctrl;
  const uint32_t tid = threadIdx.x, lane = lane_id();
  // the seed kernel left the seeds in queue 0 and their count in the schedule
  const uint32_t n0 = c->s.in_count;
  for (uint32_t i = tid; i < kRing; i += kBlkThreads) sh.ring[i] = kSlotEmpty;
  if (tid == 0) {
    sh.head = 0;
    sh.tail = n0 <= kRing ? n0 : 0u;
    sh.busy = 0;
    sh.done = 0;
    sh.abort = n0 <= kRing ? kAbortNone : kAbortCapacity;
    sh.items = 0;
    sh.progress = wall_clock64();
  }
  __syncthreads();
  if (n0 <= kRing)
    for (uint32_t i = tid; i < n0; i += kBlkThreads) {
      const uint32_t v = a.q[0][i];
      sh.ring[i] = v;
      glb(a.near_stamp)[v] = 1u;  // queued
    }
  // the queued flags are visible to every wavefront's atomics before work starts
  order_release();
  __syncthreads();

  for (;;) {
    // ---- claim up to 64 queued vertices (none once the run is abandoned) ----
    uint32_t h = 0, n = 0, stop = 0;
    if (lane == 0) stop = lds_ld(&sh.abort) != kAbortNone;
    if (readlane(stop, 0)) break;
    if (lane == 0) {
      atomicAdd(&sh.busy, 1u);
      for (;;) {
        h = lds_ld(&sh.head);
        const uint32_t t = lds_ld(&sh.tail);
        n = min(t - h, (uint32_t)kWave);
        if (n == 0 || atomicCAS(&sh.head, h, h + n) == h) break;
      }
    }
    h = readlane(h, 0);
    n = readlane(n, 0);
    if (n == 0) {
      // Nothing queued.  Done when the ring stays empty while no wavefront
      // holds claimed work (a wavefront raises `busy` before it claims and
      // lowers it only after its appends), else wait, bounded by the clock.
      uint32_t fin = 0;
      if (lane == 0) {
        atomicSub(&sh.busy, 1u);
        const uint32_t h0 = lds_ld(&sh.head), t0 = lds_ld(&sh.tail);
        if (h0 == t0 && lds_ld(&sh.busy) == 0 && lds_ld(&sh.head) == h0 && lds_ld(&sh.tail) == t0)
          lds_st(&sh.done, 1u);
        fin = lds_ld(&sh.done) != 0 || lds_ld(&sh.abort) != kAbortNone;
        if (!fin && wall_clock64() - __hip_atomic_load(&sh.progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >
                        a.spin_limit) {
          lds_st(&sh.abort, kAbortTimeout);
          fin = 1;
        }
        if (!fin) __builtin_amdgcn_s_sleep(2);
      }
      if (readlane(fin, 0)) break;
      continue;
    }
    // ---- take the claimed vertices (an append reserves its slot before it
    // writes it, so a slot may still be on its way) ----
    uint32_t u = kSlotEmpty;
    if (lane < n) {
      uint32_t *slot = &sh.ring[(h + lane) & (kRing - 1)];
      u = lds_ld(slot);
      for (uint32_t spins = 0; u == kSlotEmpty; ++spins) {
        if (spins > (1u << 20)) break;  // never seen: a lost append would be a bug
        __builtin_amdgcn_s_sleep(1);
        u = lds_ld(slot);
      }
      lds_st(slot, kSlotEmpty);
    }
    const bool valid = lane < n && u < a.nnodes;
    if (lane < n && !valid) lds_st(&sh.abort, kAbortTimeout);
    // Un-queue before reading d[u]: an improvement of u that lands after the
    // read finds the flag clear and queues u again.
    uint32_t du = kInf, rs = 0, re = 0;
    if (valid) atomic_exch(a.near_stamp + u, 0u);
    order_release();  // the flag is clear before the distance is read
    if (valid) {
      du = ld_coh(a.dist + u);
      rs = glb(a.row_start)[u];
      re = glb(a.row_start)[u + 1];
    }
    uint32_t deg = valid && du != kInf ? re - rs : 0u;
    uint32_t mx = deg;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, off));
    // ---- relax, kInFlight edges per lane at a time, and append ----
    for (uint32_t j = 0; j < mx; j += kInFlight) {
      uint32_t v[kInFlight], nd[kInFlight], old[kInFlight];
      bool act[kInFlight];
#pragma unroll
      for (int k = 0; k < kInFlight; ++k) {
        act[k] = j + k < deg;
        const uint32_t e = act[k] ? rs + j + k : 0u;
        v[k] = glb(a.edge_dst)[e];
        nd[k] = A == Algo::SSSP ? du + glb(a.edge_data)[e] : du + 1u;
      }
      uint32_t dv[kInFlight];
#pragma unroll
      for (int k = 0; k < kInFlight; ++k) dv[k] = glb(a.dist)[v[k]];  // pre-check through the caches
#pragma unroll
      for (int k = 0; k < kInFlight; ++k)
        old[k] = act[k] && nd[k] < dv[k] ? atomic_min(a.dist + v[k], nd[k]) : 0u;
#pragma unroll
      for (int k = 0; k < kInFlight; ++k) asm volatile("" : "+v"(old[k]));
      uint32_t was[kInFlight];
#pragma unroll
      for (int k = 0; k < kInFlight; ++k) was[k] = nd[k] < old[k] ? atomic_exch(a.near_stamp + v[k], 1u) : 1u;
#pragma unroll
      for (int k = 0; k < kInFlight; ++k) asm volatile("" : "+v"(was[k]));
#pragma unroll
      for (int k = 0; k < kInFlight; ++k) {
        const bool push = was[k] == 0u;
        const uint64_t m = __ballot(push);
        if (m == 0) continue;
        uint32_t base = 0, over = 0;
        if (lane_id() == (uint32_t)__builtin_ctzll(m)) {
          const uint32_t cnt = (uint32_t)__builtin_popcountll(m);
          base = atomicAdd(&sh.tail, cnt);
          if (base + cnt - lds_ld(&sh.head) > kRing) {
            over = 1;
            lds_st(&sh.abort, kAbortCapacity);
          }
        }
        const uint32_t leader = (uint32_t)__builtin_ctzll(m);
        base = readlane(base, leader);
        over = readlane(over, leader);
        if (push && !over) {
          // the slot's last vertex was claimed; its reader empties it at once
          uint32_t *slot = &sh.ring[(base + lanes_below(m)) & (kRing - 1)];
          uint32_t spins = 0;
          while (lds_ld(slot) != kSlotEmpty && spins <= (1u << 20)) {
            __builtin_amdgcn_s_sleep(1);
            ++spins;
          }
          if (spins <= (1u << 20)) lds_st(slot, v[k]);
          else lds_st(&sh.abort, kAbortTimeout);  // never seen: never overwrite a queued vertex
        }
      }
    }
    if (lane == 0) {
      atomicAdd(&sh.items, (unsigned long long)n);
      __hip_atomic_store(&sh.progress, (unsigned long long)wall_clock64(), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
      atomicSub(&sh.busy, 1u);  // after this wavefront's appends
      if (lds_ld(&sh.abort) != kAbortNone) lds_st(&sh.done, 1u);
    }
  }
  __syncthreads();
  if (tid == 0) {
    st_coh(&c->abort, sh.abort);
    st_coh64(&c->stat_items, sh.items);
  }
}

// One traversal from the seeds in `seedbuf` (already copied); returns
// LSG_OK, an error, or kBlockFallback when the ring overflowed (the caller
// runs the persistent kernel instead).
template <Algo A>
static int run_block(Workspace *w, const CsrView &g, const uint32_t *seeds, uint32_t nseeds, const lsg_opts &o,
                     uint32_t *dist, lsg_stats *st) {
  int bs, rc;
  KArgs a;
  void *seedbuf;
  if ((rc = setup_single<A>(w, g, nseeds, o, dist, &a, &seedbuf, &bs))) return rc;
  HIPCHK(hipMemcpyAsync(seedbuf, seeds, (size_t)nseeds * 4, hipMemcpyHostToDevice, w->stream));
  const uint32_t init_blocks = std::min<uint32_t>((g.nnodes + 255) / 256, 8u * w->cus);
  HIPCHK(hipEventRecord(w->ev0, w->stream));
  hipLaunchKernelGGL(init_kernel, dim3(init_blocks), dim3(256), 0, w->stream, a, a.delta);
  hipLaunchKernelGGL(seed_kernel, dim3(1), dim3(256), 0, w->stream, a, (const uint32_t *)seedbuf, nseeds);
  hipLaunchKernelGGL(block_kernel<A>, dim3(1), dim3(kBlkThreads), 0, w->stream, a);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(w->ev1, w->stream));
  Ctrl hc;
  HIPCHK(hipMemcpyAsync(&hc, w->ctrl, kCtrlHostBytes, hipMemcpyDeviceToHost, w->stream));
  HIPCHK(hipStreamSynchronize(w->stream));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, w->ev0, w->ev1));
  if (hc.abort == kAbortCapacity) return kBlockFallback;
  if (hc.abort != kAbortNone)
    return fail(LSG_ERR_TIMEOUT, "single-block traversal made no progress for the spin limit; distances are incomplete");
  if (st) {
    memset(st, 0, sizeof *st);
    st->ms_kernel = ms;
    st->items = hc.stat_items;
    st->grid_blocks = 1;
    st->block_size = (int32_t)kBlkThreads;
  }
  return LSG_OK;
}

int run_block_single(int algo, Workspace *w, const CsrView &g, uint32_t seed, const lsg_opts &o, uint32_t *dist,
                     lsg_stats *st) {
  return algo == (int)Algo::SSSP ? run_block<Algo::SSSP>(w, g, &seed, 1, o, dist, st)
                                 : run_block<Algo::BFS>(w, g, &seed, 1, o, dist, st);
}

}  // namespace lsg
