// One-workgroup near/far traversal for small, sparse, high-diameter graphs
// (lsg_opts.single_block).
//
// On a graph like a 256x256 grid a traversal is 1000+ near/far rounds of a
// few hundred vertices each, and a round of the persistent kernel costs a
// chain of dependent global round trips (in-queue load, d[u], row bounds,
// edges, atomicMin, queue and count stores, the leader's schedule) of about
// 8 us.  Here one workgroup of 1024 threads runs the same near/far rounds
// (apps/sssp/kernel.cu:355-391) with the piles in LDS and workgroup barriers
// between rounds, and shortens each round to two dependent round trips:
//
//   - a pile entry carries the vertex, the distance it was pushed with and its
//     edge range, so a round starts with the edge loads; the vertex's current
//     distance is loaded beside them and an entry whose vertex has since been
//     lowered below the entry's distance is skipped (a later entry carries the
//     lower distance; distances only fall, so a stale read never skips);
//   - the atomicMin on each neighbour is issued together with the loads of the
//     neighbour's row bounds, which the entry pushed for it needs.
//
// Entries are not deduplicated: a vertex lowered twice in one round has two
// entries, and the older one is skipped.  Far entries wait in an LDS far pile;
// when a round pushes nothing near, the threshold moves on by DELTA (or to the
// far pile's minimum) and the far entries at or below it become the next
// round's input.  BFS is the same with unit weights and no threshold.
//
// The piles hold kPileCap and kFarCap entries; a traversal that would outgrow
// them stops (kAbortCapacity) and the host runs the persistent kernel instead.
// There are no spin waits: every wait is a workgroup barrier.

#include <hip/hip_runtime.h>

#include <cstring>

#include "sssp_dev.h"

namespace lsg {

constexpr uint32_t kBlkThreads = 1024;
constexpr uint32_t kPileCap = 2048;  // entries per round pile (two of them: 64 KB)
constexpr uint32_t kFarCap = 4096;   // entries in the far pile (64 KB)
constexpr uint32_t kFarPerThread = kFarCap / kBlkThreads;
// Vertices of more edges are expanded by their whole wavefront, 64 edges a step.
constexpr uint32_t kBlkHeavyDeg = 32;

struct Entry {
  uint32_t v;   // vertex
  uint32_t d;   // distance it was pushed with
  uint32_t rs;  // row_start[v]
  uint32_t re;  // row_start[v + 1]
};

struct BlockShared {
  Entry pile[2][kPileCap];  // this round's input and the next round's
  Entry far[kFarCap];
  uint32_t count[2];
  uint32_t far_count;
  uint32_t far_min[2];  // the far pile's minimum, found at a split: far_min[phase & 1]
  uint32_t abort;
  uint32_t cur;    // pile holding this round's input
  uint32_t thr;    // entries of distance <= thr are near
  uint32_t state;  // after a round: kRun, kSplit, kDone
  uint32_t phases;
  unsigned long long items;   // entries expanded (stale ones not counted)
  unsigned long long rounds;
};

constexpr uint32_t kRun = 0, kSplit = 1, kDone = 2;

// Wave-aggregated append to an LDS pile: one LDS atomic per wavefront.
__device__ __forceinline__ void pile_push(bool pred, const Entry &e, Entry *arr, uint32_t *count, uint32_t cap,
                                          uint32_t *abort) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return;
  const uint32_t leader = (uint32_t)__builtin_ctzll(m);
  uint32_t base = 0;
  if (lane_id() == leader) base = atomicAdd(count, (uint32_t)__builtin_popcountll(m));
  base = readlane(base, leader);
  if (pred) {
    const uint32_t pos = base + lanes_below(m);
    if (pos < cap) arr[pos] = e;
    else *abort = kAbortCapacity;
  }
}

// One workgroup issues every access of a round through one CU's vector memory
// pipeline, which processes a scattered access lane by lane: a round costs in
// proportion to the lane accesses it makes, so consecutive words are fetched
// by one instruction (4-byte alignment is all that multi-word loads need).
__device__ __forceinline__ uint64_t ld_pair(const uint32_t *p) {
  uint64_t x;
  __builtin_memcpy(&x, glb(p), 8);
  return x;
}
__device__ __forceinline__ uint4 ld_quad(const uint32_t *p) {
  uint4 x;
  __builtin_memcpy(&x, glb(p), 16);
  return x;
}

// Relaxes kInFlight edges per lane: the atomicMins and the neighbours' row
// bounds are issued together, then every improved neighbour is pushed near or
// far.  Inactive lanes hold an in-range vertex (their loads are harmless).
template <Algo A>
__device__ __forceinline__ void relax_batch(const KArgs &a, BlockShared &sh, const uint32_t (&v)[kInFlight],
                                            const uint32_t (&nd)[kInFlight], const bool (&act)[kInFlight],
                                            uint32_t nxt, uint32_t thr) {
  uint32_t old[kInFlight];
  uint64_t rb[kInFlight];  // row_start[v], row_start[v + 1]
#pragma unroll
  for (int k = 0; k < kInFlight; ++k) {
    old[k] = act[k] ? atomic_min(a.dist + v[k], nd[k]) : 0u;
    rb[k] = ld_pair(a.row_start + v[k]);
  }
#pragma unroll
  for (int k = 0; k < kInFlight; ++k) asm volatile("" : "+v"(old[k]));
  // one LDS reservation per pile for the whole batch: the wavefront's pushes
  // of edge k follow those of edges 0..k-1
  bool to_near[kInFlight], to_far[kInFlight];
  uint64_t mn[kInFlight], mf[kInFlight];
  uint32_t tn = 0, tf = 0;
#pragma unroll
  for (int k = 0; k < kInFlight; ++k) {
    const bool imp = act[k] && nd[k] < old[k];
    const bool near = A == Algo::BFS || nd[k] <= thr;
    to_near[k] = imp && near;
    to_far[k] = imp && !near;
    mn[k] = __ballot(to_near[k]);
    mf[k] = A == Algo::SSSP ? __ballot(to_far[k]) : 0ull;
    tn += (uint32_t)__builtin_popcountll(mn[k]);
    tf += (uint32_t)__builtin_popcountll(mf[k]);
  }
  if (tn + tf == 0) return;
  uint32_t bn = 0, bf = 0;
  if (lane_id() == 0) {
    if (tn) bn = atomicAdd(&sh.count[nxt], tn);
    if (tf) bf = atomicAdd(&sh.far_count, tf);
  }
  bn = readlane(bn, 0);
  bf = readlane(bf, 0);
#pragma unroll
  for (int k = 0; k < kInFlight; ++k) {
    const Entry e{v[k], nd[k], (uint32_t)rb[k], (uint32_t)(rb[k] >> 32)};
    if (to_near[k]) {
      const uint32_t pos = bn + lanes_below(mn[k]);
      if (pos < kPileCap) sh.pile[nxt][pos] = e;
      else sh.abort = kAbortCapacity;
    }
    if (A == Algo::SSSP && to_far[k]) {
      const uint32_t pos = bf + lanes_below(mf[k]);
      if (pos < kFarCap) sh.far[pos] = e;
      else sh.abort = kAbortCapacity;
    }
    bn += (uint32_t)__builtin_popcountll(mn[k]);
    bf += (uint32_t)__builtin_popcountll(mf[k]);
  }
}

// Expands the entries in[0, n): light vertices one per lane, heavy ones by
// their whole wavefront.
template <Algo A>
__device__ __forceinline__ void expand_round(const KArgs &a, BlockShared &sh, const Entry *in, uint32_t n,
                                             uint32_t nxt, uint32_t thr) {
  const uint32_t lane = lane_id();
  uint32_t expanded = 0;
  for (uint32_t b0 = 0; b0 < n; b0 += kBlkThreads) {
    const uint32_t i = b0 + threadIdx.x;
    const bool valid = i < n;
    Entry e{0u, kInf, 0u, 0u};
    if (valid) e = in[i];
    // issued by every lane (a load under a per-lane condition makes the
    // compiler wait for it before the edge loads); d[0] for idle lanes
    const uint32_t du = ld_coh(a.dist + e.v);
    const uint32_t deg = valid ? e.re - e.rs : 0u;
    const bool heavy = deg > kBlkHeavyDeg;
    const uint32_t ldeg = heavy ? 0u : deg;
    auto light = [&](uint32_t j) {
      uint32_t v[kInFlight], nd[kInFlight];
      bool act[kInFlight];
      // edges base .. base+3 of the row (past its end: other rows' edges, unused)
      const uint32_t base = j < ldeg ? e.rs + j : 0u;
      if (base + kInFlight <= a.nedges) {
        const uint4 dq = ld_quad(a.edge_dst + base);
        v[0] = dq.x, v[1] = dq.y, v[2] = dq.z, v[3] = dq.w;
        if (A == Algo::SSSP) {
          const uint4 wq = ld_quad(a.edge_data + base);
          nd[0] = e.d + wq.x, nd[1] = e.d + wq.y, nd[2] = e.d + wq.z, nd[3] = e.d + wq.w;
        }
      } else {  // the last edges of the array
#pragma unroll
        for (int k = 0; k < kInFlight; ++k) {
          const uint32_t ei = (uint32_t)min<uint64_t>(base + k, a.nedges - 1);
          v[k] = glb(a.edge_dst)[ei];
          if (A == Algo::SSSP) nd[k] = e.d + glb(a.edge_data)[ei];
        }
      }
      if (A == Algo::BFS)
#pragma unroll
        for (int k = 0; k < kInFlight; ++k) nd[k] = e.d + 1u;
      // d[u] is read only after the edge loads are issued (left to itself the
      // compiler hoists the test, and its wait, above them)
      uint32_t dnow = du;
      asm volatile("" : "+v"(dnow));
#pragma unroll
      for (int k = 0; k < kInFlight; ++k) act[k] = j + k < ldeg && !(dnow < e.d);  // a lowered vertex has a newer entry
      relax_batch<A>(a, sh, v, nd, act, nxt, thr);
    };
    // the first batch outside the loop: its edge loads go out beside the load of d[u]
    // (loop bounds by ballot: no cross-lane reduction through LDS)
    if (__ballot(ldeg > 0)) light(0);
    for (uint32_t j = kInFlight; __ballot(j < ldeg); j += kInFlight) light(j);
    const bool live = valid && !(du < e.d);
    expanded += (uint32_t)__builtin_popcountll(__ballot(live));
    uint64_t hm = __ballot(live && heavy);
    while (hm) {
      const uint32_t l = (uint32_t)__builtin_ctzll(hm);
      hm &= hm - 1;
      const uint32_t hd = readlane(e.d, l), hrs = readlane(e.rs, l), hre = readlane(e.re, l);
      for (uint32_t j = hrs; j < hre; j += kWave * kInFlight) {
        uint32_t v[kInFlight], nd[kInFlight];
        bool act[kInFlight];
#pragma unroll
        for (int k = 0; k < kInFlight; ++k) {
          const uint32_t idx = j + (uint32_t)k * kWave + lane;
          act[k] = idx < hre;
          const uint32_t ei = act[k] ? idx : hrs;
          v[k] = glb(a.edge_dst)[ei];
          nd[k] = hd + (A == Algo::SSSP ? glb(a.edge_data)[ei] : 1u);
        }
        relax_batch<A>(a, sh, v, nd, act, nxt, thr);
      }
    }
  }
  if (lane == 0 && expanded) atomicAdd(&sh.items, (unsigned long long)expanded);
}

template <Algo A>
__global__ __launch_bounds__(kBlkThreads) void block_kernel(KArgs a) {
  __shared__ BlockShared sh;
  Ctrl *c = a.ctrl;
  const uint32_t tid = threadIdx.x;
  // the seed kernel left the seeds in queue 0 and their count in the schedule
  const uint32_t n0 = c->s.in_count;
  if (tid == 0) {
    sh.count[0] = n0 <= kPileCap ? n0 : 0u;
    sh.count[1] = 0;
    sh.far_count = 0;
    sh.far_min[0] = kInf;
    sh.far_min[1] = kInf;
    sh.abort = n0 <= kPileCap ? kAbortNone : kAbortCapacity;
    sh.cur = 0;
    sh.thr = A == Algo::SSSP ? 0u : 0xFFFFFFFFu;  // curdelta starts at 0 (apps/sssp/kernel.cu:472)
    sh.state = n0 <= kPileCap ? kRun : kDone;
    sh.phases = 0;
    sh.items = 0;
    sh.rounds = 0;
  }
  if (n0 <= kPileCap)
    for (uint32_t i = tid; i < n0; i += kBlkThreads) {
      const uint32_t v = a.q[0][i];
      sh.pile[0][i] = Entry{v, 0u, glb(a.row_start)[v], glb(a.row_start)[v + 1]};
    }
  __syncthreads();
  // `state` changes only between barriers (`abort` may be raised mid-round)
  for (;;) {
    if (sh.state == kDone) break;
    const uint32_t cur = sh.cur, nxt = cur ^ 1u, n = sh.count[cur];
    expand_round<A>(a, sh, sh.pile[cur], n, nxt, sh.thr);
    __syncthreads();  // every push of the round is in
    if (tid == 0) {
      ++sh.rounds;
      sh.count[cur] = 0;
      sh.cur = nxt;
      if (sh.abort != kAbortNone) sh.state = kDone;
      else if (sh.count[nxt] > 0) sh.state = kRun;
      else if (sh.far_count == 0) sh.state = kDone;
      else {
        sh.state = kSplit;
        ++sh.phases;
      }
    }
    __syncthreads();
    if (sh.state != kSplit) continue;
    // Split the far pile at the new threshold: entries at or below it are the
    // next round's input, the rest stay.  A far pile past half full first
    // drops the entries whose vertex has since been lowered below them.
    const uint32_t nf = sh.far_count, in = sh.cur, ph = sh.phases & 1u;
    Entry fe[kFarPerThread];
    bool has[kFarPerThread];
    uint32_t m = kInf;
#pragma unroll
    for (uint32_t k = 0; k < kFarPerThread; ++k) {
      const uint32_t i = k * kBlkThreads + tid;
      has[k] = i < nf;
      fe[k] = has[k] ? sh.far[i] : Entry{0u, kInf, 0u, 0u};
      m = min(m, fe[k].d);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = min(m, (uint32_t)__shfl_xor(m, off));
    if (lane_id() == 0 && m != kInf) atomicMin(&sh.far_min[ph], m);
    __syncthreads();  // every entry is in registers, far_min[ph] is the pile's minimum
    // apps/sssp/kernel.cu:383: curdelta += DELTA, and straight on to the far
    // pile's minimum when the buckets between are empty
    const uint32_t fmin = sh.far_min[ph];
    uint64_t t = (uint64_t)sh.thr + a.delta;
    if (fmin != kInf && fmin > t) t += ((uint64_t)fmin - t + a.delta - 1) / a.delta * a.delta;
    const uint32_t thr = t > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)t;
    const bool prune = nf > kFarCap / 2;
    uint32_t dnow[kFarPerThread];
#pragma unroll
    for (uint32_t k = 0; k < kFarPerThread; ++k)
      dnow[k] = prune && has[k] && fe[k].d > thr ? ld_coh(a.dist + fe[k].v) : 0u;
    if (tid == 0) {
      sh.far_count = 0;
      sh.far_min[ph ^ 1u] = kInf;  // for the next split (this one's is read by every thread now)
    }
    __syncthreads();  // every thread has the threshold; the far pile is empty
    if (tid == 0) sh.thr = thr;
#pragma unroll
    for (uint32_t k = 0; k < kFarPerThread; ++k) {
      const bool to_in = has[k] && fe[k].d <= thr;
      const bool keep = has[k] && fe[k].d > thr && !(prune && dnow[k] < fe[k].d);
      pile_push(to_in, fe[k], sh.pile[in], &sh.count[in], kPileCap, &sh.abort);
      pile_push(keep, fe[k], sh.far, &sh.far_count, kFarCap, &sh.abort);
    }
    __syncthreads();
    if (tid == 0) {
      // done on an overflow, or when only stale entries were left
      sh.state = sh.abort != kAbortNone || (sh.count[in] == 0 && sh.far_count == 0) ? kDone : kRun;
    }
    __syncthreads();
  }
  if (tid == 0) {
    st_coh(&c->abort, sh.abort);
    st_coh64(&c->stat_items, sh.items);
    st_coh64(&c->stat_rounds, sh.rounds);
    st_coh64(&c->stat_phases, A == Algo::SSSP ? (unsigned long long)sh.phases : sh.rounds);
    c->s.set_curdelta(sh.thr);
  }
}

// One traversal from the seeds in `seeds`; returns LSG_OK, an error, or
// kBlockFallback when a pile overflowed (the caller runs the persistent
// kernel instead).
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
  if (hc.abort == kAbortCapacity) return kBlockFallback;  // the only way it stops early
  if (st) {
    memset(st, 0, sizeof *st);
    st->ms_kernel = ms;
    st->items = hc.stat_items;
    st->near_iters = hc.stat_rounds;
    st->iterations = hc.stat_phases;
    st->last_delta = (int32_t)std::min<int64_t>(hc.s.curdelta(), INT32_MAX);
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
