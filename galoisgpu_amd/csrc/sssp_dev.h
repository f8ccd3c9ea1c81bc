// Device side of the single-source traversal: kernel arguments, the per-edge
// operators, tile and heavy-chunk expansion, the leader step (pile rotation)
// and block-local rounds.  Included by sssp.hip (the persistent launch) and
// lb.hip (the host-driven edge-balanced -l loop), which share every operator
// so that both compute exactly the same relaxations.
//
// Reference path: apps/sssp/kernel.cu (relax operator sssp_kernel_dev
// :177-306, persistent driver :355-391) and apps/bfs/kernel.cu (operator
// bfs_kernel_dev :121-274, driver :312-334).
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>

#include <algorithm>
#include <cstdint>

#include "engine.h"
#include "lsg_internal.h"
#include "runtime.h"
#include "sched_mlp.h"

namespace lsg {


enum class Algo : int { SSSP = 0, BFS = 1 };

constexpr uint32_t kTraceCap = 1u << 16;  // round records kept per traced call

struct KArgs {
  const uint32_t *row_start;
  const uint32_t *edge_dst;
  const uint32_t *edge_data;
  uint32_t *dist;
  uint32_t *q[3];
  uint32_t *hq[2];  // heavy chunk queues: (vertex, chunk index) pairs
  uint32_t *near_stamp;
  uint32_t *far_stamp;
  Ctrl *ctrl;
  uint32_t nnodes;
  uint32_t qcap;   // capacity of each item queue
  uint32_t hcap;   // capacity of each heavy queue in chunks
  uint32_t delta;
  uint32_t thread_deg;
  uint32_t heavy_deg;
  uint32_t expand_mode;
  uint32_t far_mode;
  uint64_t spin_limit;       // grid barrier spin bound, wall-clock ticks
  uint64_t residency_limit;  // start-up handshake bound, wall-clock ticks
  lsg_round_trace *trace;    // per-round records, or null
  uint32_t trace_cap;
  uint32_t sched;            // 1 = learned delta width and expansion mode per phase
  const float *mlp;          // scheduler weights (plain-kernel model), or null
  float *feat;               // per-phase feature records [kMaxForced][kFeatRec], or null
  uint64_t nedges;
  uint32_t lb_limit;         // -l: degrees >= this are expanded by the edge-balanced pass (else UINT32_MAX)
};

// ---- pile pushes -----------------------------------------------------------

// Pushes are staged per wavefront in LDS, kStage items per pile, and written
// out with one queue-tail atomic per flush (and at the end of every round):
// one atomic per wave_push on a tail that every wavefront of the grid shares
// serialises at the memory-side atomic unit (~12 ns each on gfx950).
constexpr uint32_t kStage = 256;

template <int BS>
struct PushStage {
  uint32_t item[BS / kWave][2][kStage];  // [wave][near, far]
  uint32_t n[BS / kWave][2];
  // Bucket far mode: smallest distance this block sent to the far pile this
  // round (pushed candidates and deferred items), folded into the control
  // block's far_min once per round.
  uint32_t far_min;
  // In a block-local round this block is the only one pushing: the piles'
  // tails are kept here and stored once at the end of the round.
  uint32_t local;
  uint32_t tail[2];
  // End-of-round drain: each wavefront's place in the block's reservation.
  uint32_t off[BS / kWave][2];
  uint32_t base[2];
};

// Writes this wavefront's staged items of pile `which` (0 near, 1 far) to the queue.
template <int BS>
__device__ __forceinline__ void stage_flush(const KArgs &a, const Sched &s, PushStage<BS> &ps, int which) {
  const uint32_t w = threadIdx.x / kWave;
  const uint32_t n = __builtin_amdgcn_readfirstlane(ps.n[w][which]);
  if (n == 0) return;
  const uint32_t qi = which ? s.far_q : s.near_q;
  uint32_t base = 0;
  if (lane_id() == 0) base = ps.local ? atomicAdd(&ps.tail[which], n) : atomic_add(&a.ctrl->qcount[qi], n);
  base = readlane(base, 0);
  for (uint32_t i = lane_id(); i < n; i += kWave) {
    const uint32_t pos = base + i;
    if (pos < a.qcap) st_coh(a.q[qi] + pos, ps.item[w][which][i]);
    else st_coh(&a.ctrl->abort, kAbortCapacity);
  }
  ps.n[w][which] = 0;
}

// Pushes `item` from every lane with `pred` onto pile `which`.
template <int BS>
__device__ __forceinline__ void stage_push(const KArgs &a, const Sched &s, PushStage<BS> &ps, int which, bool pred,
                                           uint32_t item) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return;
  const uint32_t w = threadIdx.x / kWave;
  const uint32_t c = (uint32_t)__builtin_popcountll(m);
  uint32_t n = __builtin_amdgcn_readfirstlane(ps.n[w][which]);
  if (n + c > kStage) {
    stage_flush<BS>(a, s, ps, which);
    n = 0;
  }
  if (pred) ps.item[w][which][n + lanes_below(m)] = item;
  ps.n[w][which] = n + c;
}

// End of a round, called by every thread of the block: every wavefront writes
// out what it staged, in places reserved for the whole block with one
// queue-tail atomic per pile.  (One atomic per wavefront, ~2048 of them on
// one tail when every wavefront of the grid pushes, serialised for ~25 us at
// the memory-side atomic unit: most of a round's time on the 2048^2 road graph.)
template <int BS>
__device__ __forceinline__ void stage_drain(const KArgs &a, const Sched &s, PushStage<BS> &ps) {
  constexpr int NW = BS / kWave;
  __syncthreads();  // every wavefront's staged count is final
  if (threadIdx.x < 2) {
    const int which = (int)threadIdx.x;  // lane 0: near pile, lane 1: far pile
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      ps.off[w][which] = tot;
      tot += ps.n[w][which];
    }
    uint32_t base = 0;
    if (tot) {
      if (ps.local) {
        // only this block pushes during a block-local round, and every
        // wavefront is past its last flush
        base = ps.tail[which];
        ps.tail[which] = base + tot;
      } else {
        base = atomic_add(&a.ctrl->qcount[which ? s.far_q : s.near_q], tot);
      }
    }
    ps.base[which] = base;
  }
  __syncthreads();
  const uint32_t w = threadIdx.x / kWave;
#pragma unroll
  for (int which = 0; which < 2; ++which) {
    const uint32_t n = __builtin_amdgcn_readfirstlane(ps.n[w][which]);
    if (n == 0) continue;
    const uint32_t qi = which ? s.far_q : s.near_q;
    const uint32_t base = ps.base[which] + ps.off[w][which];
    for (uint32_t i = lane_id(); i < n; i += kWave) {
      const uint32_t pos = base + i;
      if (pos < a.qcap) st_coh(a.q[qi] + pos, ps.item[w][which][i]);
      else st_coh(&a.ctrl->abort, kAbortCapacity);
    }
  }
  // every lane has read the wavefront's count before lane 0 clears it
  __builtin_amdgcn_wave_barrier();
  if (lane_id() == 0) ps.n[w][0] = ps.n[w][1] = 0;
}

// Bucket far mode: folds the smallest `x` of the lanes with `pred` into the
// block's far minimum (one LDS atomic per wavefront).
template <int BS>
__device__ __forceinline__ void note_far_min(PushStage<BS> &ps, bool pred, uint32_t x) {
  if (__ballot(pred) == 0) return;
  uint32_t m = pred ? x : kInf;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) m = min(m, (uint32_t)__shfl_xor(m, off));
  if (lane_id() == 0) atomicMin(&ps.far_min, m);
}

// End of a round, thread 0 after a block barrier: the block's far minimum
// goes to the control block (one atomic per block and round).
template <int BS>
__device__ __forceinline__ void flush_far_min(const KArgs &a, PushStage<BS> &ps) {
  if (ps.far_min != kInf) atomic_min(&a.ctrl->far_min, ps.far_min);
  ps.far_min = kInf;
}

// ---- per-edge operators ---------------------------------------------------

// Relax u->v with candidate nd = d[u] + w.  Same test and atomicMin as
// apps/sssp/kernel.cu:129-144, with the push decided by whether the atomicMin
// really lowered d[v] (the reference pushes on the pre-check alone), and the
// near/far split at curdelta (:134).  Arithmetic is u32: d[u] < 2^31 and
// w <= INT_MAX never wrap, and a candidate >= INT_MAX never beats INF.
struct SsspRelax {
  template <int BS>
  __device__ __forceinline__ static void edge(const KArgs &a, const Sched &s, uint32_t v, uint32_t nd, bool active,
                                              PushStage<BS> &ps, bool precheck) {
    // The pre-checks read through the cache: a stale d[v] is only ever larger
    // than the current one (costing at most a redundant atomicMin, whose
    // return value is authoritative), and a stale stamp is only ever older.
    // Without `precheck` every edge goes straight to the atomicMin: one
    // round trip fewer, for rounds too small for the extra atomics to matter.
    bool imp = false;
    const bool near = (int64_t)nd <= s.curdelta();
    // select between the pointers as values (selecting between the fields by
    // a per-lane condition loads the chosen pointer from the argument block)
    uint32_t *const near_st = a.near_stamp, *const far_st = a.far_stamp;
    uint32_t *stamps = near ? near_st : far_st;
    const uint32_t stamp = near ? s.round + 1 : s.phase + 1;
    uint32_t seen = stamp;
    if (active) {
      const uint32_t dv = precheck ? glb(a.dist)[v] : kInf;
      if (!precheck || nd < dv) {
        // the stamp is loaded with the atomicMin, and only for candidates:
        // most edges of a low-diameter graph fail the pre-check, and each
        // stamp load is a random line fetch
        seen = glb(stamps)[v];
        imp = nd < atomic_min(a.dist + v, nd);
      }
    }
    // Push once per round (near) or phase (far): the exchange decides between
    // racing improvers, the cached pre-check skips it when already stamped.
    const bool push = imp && seen != stamp && atomic_exch(stamps + v, stamp) != stamp;
    const bool to_near = push && near, to_far = push && !near;
    if (s.aux[1] == LSG_FAR_BUCKET) note_far_min<BS>(ps, imp && !near, nd);
    stage_push<BS>(a, s, ps, 0, to_near, v);
    stage_push<BS>(a, s, ps, 1, to_far, v);
  }
};

// BFS visit.  The reference writes LEVEL when it sees INF and pushes, racing
// benignly (apps/bfs/kernel.cu:194-200).  Here a level is a unit-weight
// distance lowered with atomicMin, so a vertex reached late (heavy chunks run
// one round after their vertex is popped) still ends at its hop count; every
// improvement is pushed to the next level's pile, once per round.
struct BfsVisit {
  template <int BS>
  __device__ __forceinline__ static void edge(const KArgs &a, const Sched &s, uint32_t v, uint32_t nd, bool active,
                                              PushStage<BS> &ps, bool precheck) {
    bool imp = false;
    if (active) {
      const uint32_t dv = precheck ? glb(a.dist)[v] : kInf;
      if (!precheck || nd < dv) imp = nd < atomic_min(a.dist + v, nd);
    }
    if (imp) {
      uint32_t stamp = s.round + 1;
      imp = glb(a.near_stamp)[v] != stamp && atomic_exch(a.near_stamp + v, stamp) != stamp;
    }
    stage_push<BS>(a, s, ps, 0, imp, v);
  }
};

template <Algo A, int BS>
__device__ __forceinline__ void relax_edge(const KArgs &a, const Sched &s, uint32_t e, uint32_t du, bool active,
                                           PushStage<BS> &ps, bool precheck) {
  uint32_t v = 0, nd = 0;
  if (active) {
    v = glb(a.edge_dst)[e];
    nd = A == Algo::SSSP ? du + glb(a.edge_data)[e] : du + 1u;
  }
  if (A == Algo::SSSP) SsspRelax::edge<BS>(a, s, v, nd, active, ps, precheck);
  else BfsVisit::edge<BS>(a, s, v, nd, active, ps, precheck);
}

// Edges a thread relaxes together.  Each step of one edge's relaxation (edge
// loads; pre-check loads; atomicMin; stamp exchange) waits for the previous
// one, so one edge at a time costs about four dependent memory round trips
// per edge per thread; with kInFlight independent edges each step is issued
// for all of them before any result is used, and the round trips overlap.
constexpr int kInFlight = 4;

// The same tests, atomics and pushes as SsspRelax / BfsVisit for the U edges
// e[k] (leaving vertices at distance du[k]) of the lanes with act[k].  The
// loads are issued by every lane (inactive lanes read edge 0 and the vertex it
// leads to, which exist whenever the caller has an edge to relax): loads under
// per-lane branches make the compiler wait for the whole memory queue at each
// branch, which would serialise the edges again.  Callers pass e[k] = 0 for
// inactive lanes.
template <Algo A, int BS, int U>
__device__ __forceinline__ void relax_edges(const KArgs &a, const Sched &s, const uint32_t (&e)[U],
                                            const uint32_t (&du)[U], const bool (&act)[U], PushStage<BS> &ps,
                                            bool precheck) {
  uint32_t v[U], nd[U];
#pragma unroll
  for (int k = 0; k < U; ++k) {
    v[k] = glb(a.edge_dst)[e[k]];
    nd[k] = A == Algo::SSSP ? du[k] + glb(a.edge_data)[e[k]] : du[k] + 1u;
  }
  if (A == Algo::SSSP) {
    const int64_t cd = s.curdelta();
    // the two stamp arrays as values: selecting between the fields by a
    // per-lane condition would load the chosen pointer from the argument block
    uint32_t *const near_st = a.near_stamp, *const far_st = a.far_stamp;
    uint32_t dv[U], seen[U], stamp[U];
    bool near[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      near[k] = (int64_t)nd[k] <= cd;
      stamp[k] = near[k] ? s.round + 1 : s.phase + 1;
      dv[k] = precheck ? glb(a.dist)[v[k]] : kInf;
    }
    // The stamps are loaded with the atomicMins, and only for candidates
    // (other lanes read vertex 0's, one line for the whole wavefront): most
    // edges of a low-diameter graph fail the pre-check, and each stamp load
    // is a random line fetch.  old = 0 where no atomic is issued: nd < 0
    // never holds.
    uint32_t old[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const bool cand = act[k] && (!precheck || nd[k] < dv[k]);
      seen[k] = glb(near[k] ? near_st : far_st)[cand ? v[k] : 0u];
      old[k] = cand ? atomic_min(a.dist + v[k], nd[k]) : 0u;
    }
    // Results are taken only here (left to itself the compiler tests each
    // result inside its atomic's branch, waiting for the whole memory queue).
#pragma unroll
    for (int k = 0; k < U; ++k) asm volatile("" : "+v"(old[k]));
    uint32_t got[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const bool want = nd[k] < old[k] && seen[k] != stamp[k];
      got[k] = want ? atomic_exch((near[k] ? near_st : far_st) + v[k], stamp[k]) : stamp[k];
    }
#pragma unroll
    for (int k = 0; k < U; ++k) asm volatile("" : "+v"(got[k]));
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const bool imp = nd[k] < old[k];
      const bool push = imp && got[k] != stamp[k];
      if (s.aux[1] == LSG_FAR_BUCKET) note_far_min<BS>(ps, imp && !near[k], nd[k]);
      stage_push<BS>(a, s, ps, 0, push && near[k], v[k]);
      stage_push<BS>(a, s, ps, 1, push && !near[k], v[k]);
    }
  } else {
    const uint32_t stamp = s.round + 1;
    uint32_t dv[U], old[U], seen[U];
#pragma unroll
    for (int k = 0; k < U; ++k) dv[k] = precheck ? glb(a.dist)[v[k]] : kInf;
    // stamps loaded with the atomicMins, for candidates only (as for SSSP)
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const bool cand = act[k] && (!precheck || nd[k] < dv[k]);
      seen[k] = glb(a.near_stamp)[cand ? v[k] : 0u];
      old[k] = cand ? atomic_min(a.dist + v[k], nd[k]) : 0u;
    }
#pragma unroll
    for (int k = 0; k < U; ++k) asm volatile("" : "+v"(old[k]));
    uint32_t got[U];
#pragma unroll
    for (int k = 0; k < U; ++k)
      got[k] = nd[k] < old[k] && seen[k] != stamp ? atomic_exch(a.near_stamp + v[k], stamp) : stamp;
#pragma unroll
    for (int k = 0; k < U; ++k) asm volatile("" : "+v"(got[k]));
#pragma unroll
    for (int k = 0; k < U; ++k) stage_push<BS>(a, s, ps, 0, nd[k] < old[k] && got[k] != stamp, v[k]);
  }
}

// ---- block-level helpers ---------------------------------------------------

// Exclusive prefix sum over the block; returns this thread's offset and sets
// *total.  Wave scans with shuffles, wave totals through LDS.
template <int BS>
__device__ __forceinline__ uint32_t block_exclusive_sum(uint32_t x, uint32_t *total, uint32_t *wsum) {
  constexpr int NW = BS / kWave;
  uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
  uint32_t incl = x;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    uint32_t y = __shfl_up(incl, off);
    if (lane >= (uint32_t)off) incl += y;
  }
  if (lane == kWave - 1) wsum[wid] = incl;
  __syncthreads();
  if (wid == 0) {
    uint32_t w = lane < (uint32_t)NW ? wsum[lane] : 0u;
    uint32_t wi = w;
#pragma unroll
    for (int off = 1; off < NW; off <<= 1) {
      uint32_t y = __shfl_up(wi, off);
      if (lane >= (uint32_t)off) wi += y;
    }
    if (lane < (uint32_t)NW) wsum[lane] = wi - w;  // exclusive wave offsets
    if (lane == NW - 1) wsum[NW] = wi;
  }
  __syncthreads();
  uint32_t r = wsum[wid] + incl - x;
  *total = wsum[NW];
  __syncthreads();
  return r;
}

template <int BS>
struct TileShared {
  uint32_t offs[BS];
  uint32_t start[BS];
  uint32_t du[BS];
  uint32_t wsum[BS / 64 + 1];
  unsigned long long edges;  // edges expanded by this block in the current round
  PushStage<BS> ps;
};

// One tile of the in-queue: items [t*ti, t*ti+ti), ti <= BS items, one per
// thread (threads >= ti idle).  Loads the row bounds of
// each popped vertex (the reference's pop + filter, apps/sssp/kernel.cu:201-210),
// queues heavy vertices as chunks for the next round, and expands the rest by
// the selected mode.
template <Algo A, int BS>
__device__ void process_tile(const KArgs &a, const Sched &s, uint32_t t, uint32_t ti, TileShared<BS> &sh,
                             bool precheck = true) {
  uint32_t i = t * ti + threadIdx.x;
  bool valid = threadIdx.x < ti && i < s.in_count;
  uint32_t u = valid ? ld_coh(a.q[s.in_q] + i) : 0u;
  uint32_t du = 0, rs = 0, re = 0, deg = 0;
  valid = valid && u < a.nnodes;
  if (valid) {
    // distance and row bounds in one round trip
    du = ld_coh(a.dist + u);
    rs = glb(a.row_start)[u];
    re = glb(a.row_start)[u + 1];
    if (du == kInf) valid = false;  // INF nodes are never expanded (apps/sssp/kernel.cu:202)
  }
  if (A == Algo::SSSP && s.aux[0] && s.aux[1] != LSG_FAR_REFERENCE) {
    // Split and bucket far piles: an item still above this phase's threshold
    // re-enters the far pile instead of expanding early.
    bool defer = valid && (int64_t)du > s.curdelta();
    if (s.aux[1] == LSG_FAR_BUCKET) note_far_min<BS>(sh.ps, defer, du);
    if (defer) {
      uint32_t stamp = s.phase + 1;
      defer = glb(a.far_stamp)[u] != stamp && atomic_exch(a.far_stamp + u, stamp) != stamp;
      valid = false;
    }
    stage_push<BS>(a, s, sh.ps, 1, defer, u);
  }
  if (valid) deg = re - rs;
  // -l: vertices of degree >= lb_limit are the edge-balanced pass's
  // (apps/sssp/kernel.cu:201: the sweep pops only degrees below DEGREE_LIMIT)
  if (deg >= a.lb_limit) deg = 0;
  // Heavy vertices: split into ceil(deg / heavy_deg) chunks, processed by
  // whole blocks in the next round.
  bool heavy = deg >= a.heavy_deg;
  uint32_t nch = heavy ? (deg + a.heavy_deg - 1) / a.heavy_deg : 0u;
  if (__ballot(heavy)) {
    uint32_t hq = 1u - s.heavy_in;
    uint32_t base = wave_reserve(nch, &a.ctrl->hcount[hq]);
    uint64_t *dst = reinterpret_cast<uint64_t *>(a.hq[hq]);
    for (uint32_t k = 0; k < nch; ++k) {
      uint32_t pos = base + k;
      if (pos < a.hcap) st_coh64(dst + pos, ((uint64_t)k << 32) | u);
      else st_coh(&a.ctrl->abort, kAbortCapacity);
    }
  }
  if (heavy) deg = 0;
  if (a.sched || a.feat) {
    // edges expanded, one of the scheduler's features
    uint32_t ws = deg;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) ws += __shfl_xor(ws, off);
    if (lane_id() == 0 && ws) atomicAdd(&sh.edges, (unsigned long long)ws);
  }

  // Expansion mode of this phase (the learned scheduler may change it).
  if (s.aux[2] == LSG_EXPAND_LB) {
    // Tile-flattened expansion: prefix sum of degrees, then every lane takes
    // one edge of the flattened tile (the reference's fg scheduler,
    // apps/sssp/kernel.cu:266-302, at block granularity).
    // A tile of at most 64 items (all in wavefront 0: most rounds of a
    // high-diameter graph) is scanned by that wavefront alone, with one
    // block barrier instead of four.
    const uint32_t nitems = min(ti, s.in_count - t * ti);
    const uint32_t span = nitems <= (uint32_t)kWave ? (uint32_t)kWave : (uint32_t)BS;
    uint32_t total;
    if (span == (uint32_t)kWave) {
      if (threadIdx.x < (uint32_t)kWave) {
        uint32_t incl = deg;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
          const uint32_t y = __shfl_up(incl, d);
          if (lane_id() >= (uint32_t)d) incl += y;
        }
        sh.offs[threadIdx.x] = incl - deg;
        sh.start[threadIdx.x] = rs;
        sh.du[threadIdx.x] = du;
        if (threadIdx.x == kWave - 1) sh.wsum[0] = incl;
      }
      __syncthreads();
      total = sh.wsum[0];
    } else {
      const uint32_t off = block_exclusive_sum<BS>(deg, &total, sh.wsum);
      sh.offs[threadIdx.x] = off;
      sh.start[threadIdx.x] = rs;
      sh.du[threadIdx.x] = du;
      __syncthreads();
    }
    for (uint32_t base = 0; base < total; base += BS * kInFlight) {
      uint32_t ee[kInFlight], d0[kInFlight];
      bool act[kInFlight];
#pragma unroll
      for (int k = 0; k < kInFlight; ++k) {
        const uint32_t e = base + k * BS + threadIdx.x;
        act[k] = e < total;
        ee[k] = 0;
        d0[k] = 0;
        if (act[k]) {
          uint32_t lo = 0, hi = span;
          while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (sh.offs[mid] <= e) lo = mid;
            else hi = mid;
          }
          ee[k] = sh.start[lo] + (e - sh.offs[lo]);
          d0[k] = sh.du[lo];
        }
      }
      relax_edges<A, BS, kInFlight>(a, s, ee, d0, act, sh.ps, precheck);
    }
    __syncthreads();
  } else {
    // Degree-binned: one lane per small vertex, the whole wavefront for
    // medium vertices (the reference's wp scheduler, :218-264, at 64 lanes).
    bool small = deg < a.thread_deg;
    uint32_t my = small ? deg : 0u;
    uint32_t mx = my;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, off));
    for (uint32_t j = 0; j < mx; ++j) relax_edge<A, BS>(a, s, rs + j, du, j < my, sh.ps, precheck);
    uint64_t med = __ballot(!small && deg > 0);
    while (med) {
      uint32_t l = (uint32_t)__builtin_ctzll(med);
      med &= med - 1;
      uint32_t ws = readlane(rs, l), wd = readlane(deg, l), wu = readlane(du, l);
      for (uint32_t j = 0; j < wd; j += kWave) {
        uint32_t jj = j + lane_id();
        relax_edge<A, BS>(a, s, ws + jj, wu, jj < wd, sh.ps, precheck);
      }
    }
  }
}

// One heavy chunk: heavy_deg consecutive edges of one vertex, spread over
// the block (replaces the -l edge-balanced TB_LB pass, apps/sssp/kernel.cu:78-149).
template <Algo A, int BS>
__device__ void process_heavy(const KArgs &a, const Sched &s, uint32_t c, TileShared<BS> &sh) {
  uint64_t ent = ld_coh64(reinterpret_cast<const uint64_t *>(a.hq[s.heavy_in]) + c);
  uint32_t u = (uint32_t)ent, k = (uint32_t)(ent >> 32);
  if (u >= a.nnodes) return;
  uint32_t du = ld_coh(a.dist + u);
  uint32_t rs = glb(a.row_start)[u], re = glb(a.row_start)[u + 1];
  if ((uint64_t)k * a.heavy_deg >= re - rs) return;
  uint32_t lo = rs + k * a.heavy_deg;
  uint32_t hi = min(re, lo + a.heavy_deg);
  if ((a.sched || a.feat) && threadIdx.x == 0) atomicAdd(&sh.edges, (unsigned long long)(hi - lo));
  for (uint32_t b = lo; b < hi; b += BS * kInFlight) {
    uint32_t ee[kInFlight], d0[kInFlight];
    bool act[kInFlight];
#pragma unroll
    for (int k = 0; k < kInFlight; ++k) {
      act[k] = k * BS + threadIdx.x < hi - b;  // b < hi: no wrap-around near 2^32 edges
      ee[k] = act[k] ? b + k * BS + threadIdx.x : 0u;
      d0[k] = du;
    }
    relax_edges<A, BS, kInFlight>(a, s, ee, d0, act, sh.ps, true);
  }
}

// Leader step between rounds, run by every thread of the last block to reach
// the barrier.  Implements the pile rotation of the persistent driver
// (apps/sssp/kernel.cu:363-384): iterate the near pile to empty (retry2), then
// make the far pile the next phase's input (advance2) with curdelta += DELTA;
// finish when both are empty.  At a phase boundary the learned scheduler (when
// enabled) picks the next phase's delta width and expansion mode; it never
// changes which vertices are relaxed or how, only when.
// `s`, `abort_seen` and `far_tail` are the caller's LDS words: on return they
// hold the schedule just stored, the abort flag loaded with the step's
// counters and the far pile's count for the next round (its near pile starts
// empty), so rounds that the same block runs next need not load them again.
template <Algo A>
__device__ void leader_step(const KArgs &a, Sched &s, uint32_t &abort_seen, uint32_t &far_tail) {
  __shared__ SchedFeatures f;
  __shared__ MlpScratch sc;
  __shared__ uint32_t decide, consumed, far_min;
  __shared__ uint64_t phases_seen;
  Ctrl *c = a.ctrl;
  if (threadIdx.x == 0) {
    decide = 0;
    // Every control word of the step is loaded in one batch: the step costs
    // one round trip to the coherence point instead of one per dependent load
    // (the schedule goes to registers first: loaded straight into LDS, the
    // counters' loads would wait for its stores).
    Sched sr;
    sched_load(sr, &c->s);
    CtrlCounts k;
    counts_load(k, c);
    abort_seen = ld_coh(&c->abort);
    s = sr;
    far_min = k.far_min;
    const uint32_t nnear = k.q(s.near_q), nfar = k.q(s.far_q), nheavy = k.h(1u - s.heavy_in);
    far_tail = 0;
    if (nnear > a.qcap || nfar > a.qcap || nheavy > a.hcap) {
      // A push found its queue full (and raised kAbortCapacity); never hand
      // out a count beyond the buffer.
      st_coh(&c->abort, kAbortCapacity);
      s.done = 1;
    } else {
      st_coh64(&c->stat_items, k.items + s.in_count);
      st_coh64(&c->stat_heavy, k.heavy + s.heavy_count);
      st_coh64(&c->stat_rounds, k.rounds + 1);
      if (a.trace && s.round - 1 < a.trace_cap) {
        lsg_round_trace &r = a.trace[s.round - 1];
        r.round = s.round;
        r.phase = s.phase;
        r.in_items = s.in_count;
        r.heavy_chunks = s.heavy_count;
        r.near_out = nnear;
        r.far_out = nfar;
        r.heavy_out = nheavy;
        r.curdelta_lo = s.curdelta_lo;
        r.curdelta_hi = s.curdelta_hi;
        r.delta = s.delta;
        r.t_end = wall_clock64();
      }
      consumed = s.in_q;
      st_coh(&c->hcount[s.heavy_in], 0u);  // consumed heavy queue becomes the next producer
      s.heavy_in = 1u - s.heavy_in;
      s.heavy_count = nheavy;
      if (nnear > 0 || nheavy > 0) {
        s.in_q = s.near_q;
        s.in_count = nnear;
        s.near_q = consumed;
        st_coh(&c->qcount[consumed], 0u);
        s.aux[0] = 0;
        s.round += 1;
        far_tail = nfar;
      } else if (A == Algo::SSSP && nfar > 0) {
        decide = 1;
        s.in_count = nfar;
        f.in_items = nfar;
        f.phase = s.phase;
        f.round = s.round;
        f.curdelta = (float)s.curdelta();
        f.base_delta = a.delta;
        f.prev_delta = s.delta;
        f.lane_items = (float)(k.items + s.in_count);
        f.edges = (float)k.edges;
        f.nnodes = a.nnodes;
        f.nedges = (float)a.nedges;
        f.nsrc = 1;
      } else {
        s.done = 1;
        st_coh64(&c->stat_phases, k.phases + 1);
      }
    }
    phases_seen = k.phases;
  }
  __syncthreads();
  if (decide) {
    int act = a.sched ? mlp_block_action(a.mlp, f, sc) : -1;
    if (threadIdx.x == 0) {
      if (act >= 0) {
        s.delta = action_delta(a.mlp, act, a.delta);
        s.aux[2] = (uint32_t)a.mlp[kOffFarMode + act];  // plain-kernel models: expansion mode
      }
      if (a.feat && s.phase < (uint32_t)kMaxForced) {
        float *rec = a.feat + (size_t)s.phase * kFeatRec;
        sched_featurize(f, rec);
        rec[kFeat] = (float)act;
        rec[kFeat + 1] = (float)s.delta;
        rec[kFeat + 2] = (float)s.aux[2];
        rec[kFeat + 3] = (float)(s.round + 1);  // first round of the phase being started
      }
      s.in_q = s.far_q;
      s.far_q = consumed;
      st_coh(&c->qcount[consumed], 0u);
      s.aux[0] = 1;  // this round pops the far pile
      // apps/sssp/kernel.cu:383: curdelta += DELTA.  In bucket mode the
      // threshold moves on by whole steps of delta until it reaches the
      // smallest distance waiting in the far pile (delta-stepping's lowest
      // non-empty bucket), so empty buckets cost no phase.
      int64_t cd = s.curdelta() + s.delta;
      if (s.aux[1] == LSG_FAR_BUCKET) {
        if (far_min != kInf && (int64_t)far_min > cd) cd += ((int64_t)far_min - cd + s.delta - 1) / s.delta * s.delta;
        st_coh(&c->far_min, kInf);
      }
      s.set_curdelta(cd);
      s.phase += 1;
      s.round += 1;
      st_coh64(&c->stat_phases, phases_seen + 1);
    }
  }
  if (threadIdx.x == 0) {
    sched_store(&c->s, s);
    st_coh(&c->ticket, 0u);
  }
}

// ---- pile entries: block-local rounds and one-workgroup traversals --------
//
// A pile entry carries the vertex, the distance it was pushed with and its
// edge range, so a round of a single workgroup starts with the edge loads; the
// vertex's current distance is loaded beside them, and an entry whose vertex
// has since been lowered below the entry's distance is skipped (a later entry
// carries the lower distance; distances only fall, so a stale read never
// skips).  The neighbours' row bounds are loaded beside the atomicMins.  A
// round is then two dependent round trips (edges, atomicMins), against the
// in-queue load, d[u], row bounds, edges, atomicMin and stamp exchange of a
// round that pops vertex ids.

// Vertices of more edges are expanded by their whole wavefront, 64 edges a step.
constexpr uint32_t kEntryHeavyDeg = 32;

struct alignas(16) Entry {
  uint32_t v;   // vertex
  uint32_t d;   // distance it was pushed with
  uint32_t rs;  // row_start[v]
  uint32_t re;  // row_start[v + 1]
};

// Pile pointers carry their address space, so that LDS piles are accessed
// with LDS instructions and HBM piles with global ones.
constexpr int kLdsAS = 3, kGlobalAS = 1;
template <int AS>
using EntryPtr = __attribute__((address_space(AS))) Entry *;
template <int AS>
__device__ __forceinline__ Entry ld_entry(EntryPtr<AS> p) {
  return Entry{p->v, p->d, p->rs, p->re};
}
template <int AS>
__device__ __forceinline__ void st_entry(EntryPtr<AS> p, const Entry &e) {
  p->v = e.v;
  p->d = e.d;
  p->rs = e.rs;
  p->re = e.re;
}

// Where a round appends: the next round's pile and the far pile, their counts
// (in LDS) and capacities, and the overflow flag.
template <int AS>
struct Sink {
  EntryPtr<AS> near;
  uint32_t *near_count;
  uint32_t near_cap;
  EntryPtr<AS> far;
  uint32_t *far_count;
  uint32_t far_cap;
  uint32_t *abort;
};

constexpr uint32_t kRun = 0, kSplit = 1, kDone = 2;

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

// Appends entry e[k] to the near pile where pa[k] holds and to the far pile
// where pb[k] holds, for k < K: one reservation per pile per wavefront, the
// wavefront's entries of k following those of 0..k-1.  Wavefront-uniform
// control flow; an overflow raises kAbortCapacity.
template <int K, int AS>
__device__ __forceinline__ void push2(const bool (&pa)[K], const bool (&pb)[K], const Entry (&e)[K],
                                      const Sink<AS> &s) {
  uint64_t ma[K], mb[K];
  uint32_t ta = 0, tb = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    ma[k] = __ballot(pa[k]);
    mb[k] = __ballot(pb[k]);
    ta += (uint32_t)__builtin_popcountll(ma[k]);
    tb += (uint32_t)__builtin_popcountll(mb[k]);
  }
  if (ta + tb == 0) return;
  uint32_t ba = 0, bb = 0;
  if (lane_id() == 0) {
    if (ta) ba = atomicAdd(s.near_count, ta);
    if (tb) bb = atomicAdd(s.far_count, tb);
  }
  ba = readlane(ba, 0);
  bb = readlane(bb, 0);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (pa[k]) {
      const uint32_t pos = ba + lanes_below(ma[k]);
      if (pos < s.near_cap) st_entry<AS>(s.near + pos, e[k]);
      else *s.abort = kAbortCapacity;
    }
    if (pb[k]) {
      const uint32_t pos = bb + lanes_below(mb[k]);
      if (pos < s.far_cap) st_entry<AS>(s.far + pos, e[k]);
      else *s.abort = kAbortCapacity;
    }
    ba += (uint32_t)__builtin_popcountll(ma[k]);
    bb += (uint32_t)__builtin_popcountll(mb[k]);
  }
}

// Relaxes edges j .. j+3 of a light entry e (ldeg edges; du: the vertex's
// distance as loaded) with relax(a, v, nd, act), for the whole wavefront.
template <Algo A, class Relax>
__device__ __forceinline__ void expand_light(const KArgs &a, const Entry &e, uint32_t du, uint32_t ldeg, uint32_t j,
                                             const Relax &relax) {
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
  relax(a, v, nd, act);
}

// Expands the entries in[0, n) with a workgroup of NT threads: light vertices
// one per lane, heavy ones (more than kEntryHeavyDeg edges) by their whole
// wavefront.  relax(a, v, nd, act) relaxes kInFlight edges per lane (edge k:
// to v[k] with candidate nd[k], where act[k]), called with the whole
// wavefront; a heavy vertex's edges stop early once relax.stopped().  Adds
// the live (not stale) entries to *items.  (Functors rather
// than lambdas: a lambda holding a reference to the kernel arguments made the
// persistent kernel copy them to scratch.)
template <Algo A, int AS, int NT, class Relax>
__device__ __forceinline__ void expand_entries(const KArgs &a, EntryPtr<AS> in, uint32_t n, const Relax &relax,
                                               unsigned long long *items) {
  const uint32_t lane = lane_id();
  uint32_t expanded = 0;
  for (uint32_t b0 = 0; b0 < n; b0 += NT) {
    const uint32_t i = b0 + threadIdx.x;
    const bool valid = i < n;
    Entry e{0u, kInf, 0u, 0u};
    if (valid) e = ld_entry<AS>(in + i);
    // issued by every lane (a load under a per-lane condition makes the
    // compiler wait for it before the edge loads); d[0] for idle lanes.  A
    // BFS entry is never stale (its level is exact when it is pushed, and a
    // vertex is pushed once), so BFS takes the entry's level without a load.
    const uint32_t du = A == Algo::BFS ? e.d : ld_coh(a.dist + e.v);
    const uint32_t deg = valid ? e.re - e.rs : 0u;
    const bool heavy = deg > kEntryHeavyDeg;
    const uint32_t ldeg = heavy ? 0u : deg;
    // the first batch outside the loop: its edge loads go out beside the load of d[u]
    // (loop bounds by ballot: no cross-lane reduction through LDS)
    if (__ballot(ldeg > 0)) expand_light<A>(a, e, du, ldeg, 0u, relax);
    for (uint32_t j = kInFlight; __ballot(j < ldeg); j += kInFlight) expand_light<A>(a, e, du, ldeg, j, relax);
    const bool live = valid && !(du < e.d);
    expanded += (uint32_t)__builtin_popcountll(__ballot(live));
    uint64_t hm = __ballot(live && heavy);
    while (hm) {
      const uint32_t l = (uint32_t)__builtin_ctzll(hm);
      hm &= hm - 1;
      const uint32_t hd = readlane(e.d, l), hrs = readlane(e.rs, l), hre = readlane(e.re, l);
      for (uint32_t j = hrs; j < hre && !relax.stopped(); j += kWave * kInFlight) {
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
        relax(a, v, nd, act);
      }
    }
  }
  if (lane == 0 && expanded) atomicAdd(items, (unsigned long long)expanded);
}

template <int BS>
__device__ __forceinline__ uint32_t tile_items(uint32_t in_count, uint32_t nblocks) {
  uint32_t per = (in_count + nblocks - 1) / nblocks;
  per = (per + kWave - 1) / kWave * kWave;
  return per < (uint32_t)kWave ? (uint32_t)kWave : per > (uint32_t)BS ? (uint32_t)BS : per;
}

// Rounds small enough for one block.  While the next round's in-queue fits in
// kLocalTiles tiles and has no heavy chunks, the block that ran the leader
// step runs that round itself, separated by block barriers, while the rest of
// the grid stays parked at the grid barrier; it bumps `progress` each round so
// the parked blocks' spin limit measures lack of progress.  A round of a small
// frontier (high-diameter graphs, the first and last rounds of any traversal)
// then costs a few dependent memory round trips instead of a grid barrier.
// One tile per local round measured fastest (1, 2, 4 tiles x blocks of 256,
// 512, 1024 threads: 12.4-16.4 ms on the 256^2 grid, one tile of 256 best).
// Local rounds skip the d[v] pre-check and go straight to the atomicMin (one
// round trip fewer; for a few hundred edges the extra atomics cost less).
constexpr uint32_t kLocalTiles = 1;

// `ls` and `labort`: the schedule and abort flag the leader step just left in LDS.
template <Algo A, int BS>
__device__ void local_rounds(const KArgs &a, TileShared<BS> &sh, Sched &ls, uint32_t &labort, uint32_t &lfar) {
  __shared__ uint32_t go;
  Ctrl *c = a.ctrl;
  for (;;) {
    if (threadIdx.x == 0) {
      go = !ls.done && ls.heavy_count == 0 && ls.in_count <= kLocalTiles * BS && labort == kAbortNone;
      sh.edges = 0;
      sh.ps.local = go;
      sh.ps.tail[0] = 0;  // the near pile starts empty
      sh.ps.tail[1] = lfar;
    }
    __syncthreads();
    if (!go) return;
    const Sched s = ls;
    for (uint32_t t = 0; t * BS < s.in_count; ++t) {
      process_tile<A, BS>(a, s, t, BS, sh, false);
      __syncthreads();
    }
    stage_drain<BS>(a, s, sh.ps);
    __syncthreads();  // every wavefront has taken its places from the LDS tails
    if (threadIdx.x == 0) {
      st_coh(&c->qcount[s.near_q], sh.ps.tail[0]);
      st_coh(&c->qcount[s.far_q], sh.ps.tail[1]);
      flush_far_min<BS>(a, sh.ps);
      if ((a.sched || a.feat) && sh.edges)
        __hip_atomic_fetch_add(&c->stat_edges, (uint64_t)sh.edges, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // this block's pushes and counts are complete before the leader step reads them
    order_release();
    __syncthreads();
    leader_step<A>(a, ls, labort, lfar);
    if (threadIdx.x == 0) st_coh(&c->progress, s.round);
  }
}

// At least 4 waves per SIMD (at most 128 VGPRs): two blocks of 512 per CU.
template <Algo A, int BS>
__global__ __launch_bounds__(BS, 4) void persistent_kernel(KArgs a) {
  __shared__ TileShared<BS> sh;
  __shared__ Sched ss;
  __shared__ Sched lsched;  // the leader step's schedule, reused by local rounds
  __shared__ uint32_t labort, lfar;
  __shared__ uint32_t tk;
  __shared__ uint32_t quit;
  Ctrl *c = a.ctrl;
  const uint32_t nb = gridDim.x;
  if (!residency_handshake(c, nb, a.residency_limit)) return;
  if (threadIdx.x < 2 * (BS / kWave)) (&sh.ps.n[0][0])[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    sh.ps.local = 0;
    sh.ps.far_min = kInf;
  }
  for (;;) {
    if (threadIdx.x == 0) {
      Sched s;
      sched_load(s, &c->s);
      ss = s;
      quit = s.done || ld_coh(&c->abort) != kAbortNone;
      sh.edges = 0;
    }
    __syncthreads();
    if (quit) break;
    const Sched s = ss;
    // Tile size: BS items, or fewer when that leaves blocks idle -- a small
    // frontier spread over more blocks finishes its round in fewer dependent
    // steps (multiples of 64 items, one wavefront's worth, at least).
    const uint32_t ti = tile_items<BS>(s.in_count, nb);
    const uint32_t ntiles = (s.in_count + ti - 1) / ti;
    const uint32_t total = ntiles + s.heavy_count;
    // Static first pass (ticket = block id), then dynamic tickets.
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
    if (a.sched || a.feat || s.aux[1] == LSG_FAR_BUCKET) {
      __syncthreads();
      if (threadIdx.x == 0 && sh.edges)
        __hip_atomic_fetch_add(&c->stat_edges, (uint64_t)sh.edges, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (threadIdx.x == 0) flush_far_min<BS>(a, sh.ps);
    }
    grid_barrier_block(c, nb, a.spin_limit, [&] {
      leader_step<A>(a, lsched, labort, lfar);
      local_rounds<A, BS>(a, sh, lsched, labort, lfar);
    });
  }
}

// Resets the control block: empty queues, round 1, curdelta cd0 (0 for a new
// traversal, apps/sssp/kernel.cu:472), the barrier words.  Every thread of the
// grid calls it (the barrier words are cleared grid-strided).
__device__ __forceinline__ void ctrl_init(const KArgs &a, uint32_t delta, int64_t cd0, uint32_t tid, uint32_t nt) {
  ctrl_reset_barrier(a.ctrl, tid, nt);
  if (tid == 0) {
    Ctrl *c = a.ctrl;
    c->qcount[0] = 0;
    c->qcount[1] = 0;
    c->qcount[2] = 0;
    c->hcount[0] = 0;
    c->hcount[1] = 0;
    c->ticket = 0;
    c->abort = kAbortNone;
    c->checkin = 0;
    c->resident_seen = 0;
    c->far_min = kInf;
    Sched s;
    memset(&s, 0, sizeof s);
    s.in_q = 0;
    s.near_q = 1;
    s.far_q = 2;
    s.round = 1;
    s.set_curdelta(cd0);
    s.delta = delta;
    s.aux[1] = a.far_mode;
    s.aux[2] = a.expand_mode;
    c->s = s;
    c->stat_items = c->stat_rounds = c->stat_heavy = c->stat_phases = c->stat_edges = 0;
  }
}

// Initialise distances, stamps and the control block (apps/sssp/kernel.cu:23-35,469;
// apps/bfs/kernel.cu:20-32).  The sources are seeded by seed_kernel.
static __global__ void init_kernel(KArgs a, uint32_t delta) {
  uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
  for (uint32_t v = tid; v < a.nnodes; v += nt) {
    glb(a.dist)[v] = kInf;
    glb(a.near_stamp)[v] = 0u;
    glb(a.far_stamp)[v] = 0u;
  }
  ctrl_init(a, delta, 0, tid, nt);
}

// Continuing a traversal that another kernel left part done (the one-workgroup
// kernel whose piles overflowed): distances kept, stamps and control block
// reset, curdelta at the threshold the traversal had reached.
// resume_seed_kernel then queues the vertices to expand.
static __global__ void resume_init_kernel(KArgs a, uint32_t delta, int64_t cd0) {
  uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
  for (uint32_t v = tid; v < a.nnodes; v += nt) {
    glb(a.near_stamp)[v] = 0u;
    glb(a.far_stamp)[v] = 0u;
  }
  ctrl_init(a, delta, cd0, tid, nt);
}

// Where a traversal continues after another kernel left it part done: the
// vertices with lo <= d <= hi form the first in-queue, and distances above hi
// are set back to INF.  SSSP (lo 0, hi INF - 1): every reached vertex is
// queued, so the traversal continues exactly: every distance is the length of
// a path, and every vertex whose distance may not yet have been passed on to
// its neighbours (a pending pile entry, or an improvement whose push found
// the pile full) is expanded again; settled vertices are expanded once more,
// for nothing.  BFS, whose rounds are levels: lo = hi = the level that was
// being expanded, which was complete and exact when its round began; the
// next level, part found, is found again from it (queueing it as well would
// mix two levels in one round, and every vertex first reached at too deep a
// level is expanded twice).
struct Resume {
  int64_t curdelta;
  uint32_t lo, hi;
};

// The queue order is unspecified; the distances do not depend on it.
// 256-thread blocks, grid-strided in whole blocks; one queue reservation per
// block step.
static __global__ __launch_bounds__(256) void resume_seed_kernel(KArgs a, uint32_t lo, uint32_t hi) {
  __shared__ uint32_t cnt, base;
  Ctrl *c = a.ctrl;
  for (uint32_t v0 = blockIdx.x * 256u; v0 < a.nnodes; v0 += gridDim.x * 256u) {
    const uint32_t v = v0 + threadIdx.x;
    const uint32_t d = v < a.nnodes ? glb(a.dist)[v] : kInf;
    const bool queued = d >= lo && d <= hi;
    if (d > hi && d != kInf) glb(a.dist)[v] = kInf;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const uint32_t slot = queued ? atomicAdd(&cnt, 1u) : 0u;
    __syncthreads();
    if (threadIdx.x == 0 && cnt) {
      base = atomicAdd(&c->qcount[0], cnt);
      atomicAdd(&c->s.in_count, cnt);
    }
    __syncthreads();
    if (queued) glb(a.q[0])[base + slot] = v;
    __syncthreads();  // base and cnt are read before the next step resets them
  }
}

// d[seed] = 0 and the seeds form the first in-queue (apps/sssp/kernel.cu:33,476-477).
// One block.  Several seeds make a multi-seed traversal: on a disjoint union of
// graphs with one seed each it computes every graph's own distances.
static __global__ void seed_kernel(KArgs a, const uint32_t *seeds, uint32_t nseeds) {
  __shared__ uint32_t cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nseeds; i += blockDim.x) {
    uint32_t v = seeds[i];
    if (v < a.nnodes && atomicExch(a.dist + v, 0u) != 0u) a.q[0][atomicAdd(&cnt, 1u)] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a.ctrl->qcount[0] = cnt;
    a.ctrl->s.in_count = cnt;
  }
}

// ---- host helpers shared by the persistent and the -l launches ----------

// The arrays of one traversal: a device CSR, possibly a packed union of graphs.
struct CsrView {
  const uint32_t *row, *dst, *w;
  uint32_t nnodes;
  uint64_t nedges;
};

inline int check_single_opts(const lsg_opts &o, bool sssp, int *bs, int *expand) {
  if (sssp && o.delta <= 0) return fail(LSG_ERR_ARG, "delta must be > 0 (got %d)", o.delta);
  *bs = o.block_size ? o.block_size : 256;
  if (*bs != 256 && *bs != 512 && *bs != 1024) return fail(LSG_ERR_ARG, "block_size must be 256, 512 or 1024");
  *expand = o.expand_mode;
  if (*expand != LSG_EXPAND_TWB && *expand != LSG_EXPAND_LB) return fail(LSG_ERR_ARG, "bad expand_mode %d", *expand);
  if (o.far_mode != LSG_FAR_REFERENCE && o.far_mode != LSG_FAR_SPLIT && o.far_mode != LSG_FAR_BUCKET)
    return fail(LSG_ERR_ARG, "bad far_mode %d", o.far_mode);
  if (o.lb_deg < 0) return fail(LSG_ERR_ARG, "lb_deg must be >= 0 (got %d)", o.lb_deg);
  if (o.sched && !sssp) return fail(LSG_ERR_ARG, "the learned scheduler chooses SSSP delta widths; BFS has none");
  return LSG_OK;
}

// Validates the options of a single-source run and fills the kernel
// arguments: workspace buffers (queues, heavy queues, stamps, the seed list
// of `nseeds` entries), schedule knobs, scheduler model, trace buffers.
template <Algo A>
inline int setup_single(Workspace *w, const CsrView &g, uint32_t nseeds, const lsg_opts &o, uint32_t *dist, KArgs *ap,
                        void **seedbuf_out, int *bs_out) {
  int bs, expand, rc;
  if ((rc = check_single_opts(o, A == Algo::SSSP, &bs, &expand))) return rc;
  uint32_t thread_deg = o.thread_deg > 0 ? (uint32_t)o.thread_deg : 8u;
  // 512: vertices of a few hundred edges and up are spread over the grid as
  // chunks (RMAT-20 from the hub: 10.0 ms at 4096, 5.8 ms at 512; 128-1024 alike).
  uint32_t heavy_deg = o.heavy_deg > 0 ? (uint32_t)o.heavy_deg : 512u;
  if (heavy_deg < (uint32_t)bs) heavy_deg = (uint32_t)bs;
  uint64_t hcap64 = 2ull * g.nedges / heavy_deg + 64;
  uint32_t hcap = (uint32_t)std::min<uint64_t>(hcap64, UINT32_MAX / 2);
  void *qbuf, *hbuf, *sbuf, *seedbuf;
  if ((rc = ws_buffer(w, kBufQueues, 3ull * g.nnodes * 4, &qbuf)) ||
      (rc = ws_buffer(w, kBufHeavy, 2ull * hcap * 8, &hbuf)) ||
      (rc = ws_buffer(w, kBufStamps, 2ull * g.nnodes * 4, &sbuf)) ||
      (rc = ws_buffer(w, kBufSeedList, (size_t)nseeds * 4, &seedbuf)))
    return rc;

  KArgs &a = *ap;
  memset(&a, 0, sizeof a);
  a.row_start = g.row;
  a.edge_dst = g.dst;
  a.edge_data = g.w;
  a.dist = dist;
  for (int i = 0; i < 3; ++i) a.q[i] = (uint32_t *)qbuf + (size_t)i * g.nnodes;
  a.hq[0] = (uint32_t *)hbuf;
  a.hq[1] = (uint32_t *)hbuf + 2ull * hcap;
  a.near_stamp = (uint32_t *)sbuf;
  a.far_stamp = (uint32_t *)sbuf + g.nnodes;
  a.ctrl = w->ctrl;
  a.nnodes = g.nnodes;
  a.qcap = g.nnodes;
  a.hcap = hcap;
  a.delta = A == Algo::SSSP ? (uint32_t)o.delta : 1u;
  a.thread_deg = thread_deg;
  a.heavy_deg = heavy_deg;
  a.expand_mode = (uint32_t)expand;
  a.far_mode = (uint32_t)o.far_mode;
  a.spin_limit = spin_limit_ticks(w);
  a.residency_limit = residency_limit_ticks(w);
  a.nedges = g.nedges;
  if (o.sched) {
    a.sched = 1;
    a.mlp = sched_model_device(w->device, LSG_SCHED_PLAIN);
    if (!a.mlp)
      return fail(LSG_ERR_ARG, "learned scheduler requested but no plain-kernel model is loaded (lsg_sched_set_model)");
  }
  if (o.trace) {
    void *tbuf;
    const size_t tbytes = kTraceCap * sizeof(lsg_round_trace);
    if ((rc = ws_buffer(w, kBufTrace, tbytes + kMaxForced * kFeatRec * sizeof(float), &tbuf))) return rc;
    a.trace = (lsg_round_trace *)tbuf;
    a.trace_cap = kTraceCap;
    if (A == Algo::SSSP) a.feat = (float *)((char *)tbuf + tbytes);
  }
  a.lb_limit = UINT32_MAX;
  *seedbuf_out = seedbuf;
  *bs_out = bs;
  return LSG_OK;
}

// The -l loop (lb.hip) for algo = (int)Algo::SSSP or Algo::BFS.
int run_lb_single(int algo, Workspace *w, const CsrView &g, uint32_t seed, const lsg_opts &o, uint32_t *dist,
                  lsg_stats *st);

// The one-workgroup near/far traversals (block.hip).  kBlockFallback: a pile
// overflowed.  A single source then continues in the persistent kernel from
// the distances left in dist, as *resume says (run_core's resume); a batch of
// graphs runs the union launch instead.  *st holds the one-workgroup launch's
// figures either way.
constexpr int kBlockFallback = 1;
int run_block_single(int algo, Workspace *w, const CsrView &g, uint32_t seed, const lsg_opts &o, uint32_t *dist,
                     lsg_stats *st, Resume *resume);
int run_graphs_block(Workspace *w, const lsg_graph *graphs, uint32_t ngraphs, const uint32_t *srcs,
                     int32_t *const *dist_out, const lsg_opts &o, lsg_batch_stats *stats);

}  // namespace lsg
