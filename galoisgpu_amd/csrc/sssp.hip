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

namespace lsg {

// Load the push stamp beside d[v] (one round trip fewer per improving edge;
// C1 12.3 -> 11.9 ms, C3 12.0 -> 11.9 ms, RMAT unchanged).
#ifndef LSG_STAMP_PRELOAD
#define LSG_STAMP_PRELOAD 1
#endif

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
};

// ---- pile pushes -----------------------------------------------------------

// Pushes are staged per wavefront in LDS, kStage items per pile, and written
// out with one queue-tail atomic per flush (and at the end of every round):
// one atomic per wave_push on a tail that every wavefront of the grid shares
// serialises at the memory-side atomic unit (~12 ns each on gfx950).
#ifndef LSG_PLAIN_PUSH_BUF
#define LSG_PLAIN_PUSH_BUF 256
#endif
constexpr uint32_t kStage = LSG_PLAIN_PUSH_BUF > 0 ? LSG_PLAIN_PUSH_BUF : 1;

template <int BS>
struct PushStage {
  uint32_t item[BS / kWave][2][kStage];  // [wave][near, far]
  uint32_t n[BS / kWave][2];
  // In a block-local round this block is the only one pushing: the piles'
  // tails are kept here and stored once at the end of the round.
  uint32_t local;
  uint32_t tail[2];
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
#if LSG_PLAIN_PUSH_BUF
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
#else
  const uint32_t qi = which ? s.far_q : s.near_q;
  wave_push(pred, item, a.q[qi], &a.ctrl->qcount[qi], a.qcap, &a.ctrl->abort);
#endif
}

// End of a round: every wavefront writes out what it staged.
template <int BS>
__device__ __forceinline__ void stage_drain(const KArgs &a, const Sched &s, PushStage<BS> &ps) {
#if LSG_PLAIN_PUSH_BUF
  stage_flush<BS>(a, s, ps, 0);
  stage_flush<BS>(a, s, ps, 1);
#endif
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
    uint32_t *stamps = near ? a.near_stamp : a.far_stamp;
    const uint32_t stamp = near ? s.round + 1 : s.phase + 1;
    uint32_t seen = stamp;
    if (active) {
      const uint32_t dv = precheck ? a.dist[v] : kInf;
#if LSG_STAMP_PRELOAD
      seen = stamps[v];  // loaded beside d[v]: one round trip fewer when the edge improves
#endif
      if (!precheck || nd < dv) imp = nd < atomic_min(a.dist + v, nd);
    }
#if !LSG_STAMP_PRELOAD
    if (imp) seen = stamps[v];
#endif
    // Push once per round (near) or phase (far): the exchange decides between
    // racing improvers, the cached pre-check skips it when already stamped.
    const bool push = imp && seen != stamp && atomic_exch(stamps + v, stamp) != stamp;
    const bool to_near = push && near, to_far = push && !near;
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
      const uint32_t dv = precheck ? a.dist[v] : kInf;
      if (!precheck || nd < dv) imp = nd < atomic_min(a.dist + v, nd);
    }
    if (imp) {
      uint32_t stamp = s.round + 1;
      imp = a.near_stamp[v] != stamp && atomic_exch(a.near_stamp + v, stamp) != stamp;
    }
    stage_push<BS>(a, s, ps, 0, imp, v);
  }
};

template <Algo A, int BS>
__device__ __forceinline__ void relax_edge(const KArgs &a, const Sched &s, uint32_t e, uint32_t du, bool active,
                                           PushStage<BS> &ps, bool precheck) {
  uint32_t v = 0, nd = 0;
  if (active) {
    v = a.edge_dst[e];
    nd = A == Algo::SSSP ? du + a.edge_data[e] : du + 1u;
  }
  if (A == Algo::SSSP) SsspRelax::edge<BS>(a, s, v, nd, active, ps, precheck);
  else BfsVisit::edge<BS>(a, s, v, nd, active, ps, precheck);
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
    rs = a.row_start[u];
    re = a.row_start[u + 1];
    if (du == kInf) valid = false;  // INF nodes are never expanded (apps/sssp/kernel.cu:202)
  }
  if (A == Algo::SSSP && s.aux[0] && s.aux[1] == LSG_FAR_SPLIT) {
    // Split far pile: an item still above this phase's threshold re-enters
    // the far pile instead of expanding early.
    bool defer = valid && (int64_t)du > s.curdelta();
    if (defer) {
      uint32_t stamp = s.phase + 1;
      defer = a.far_stamp[u] != stamp && atomic_exch(a.far_stamp + u, stamp) != stamp;
      valid = false;
    }
    stage_push<BS>(a, s, sh.ps, 1, defer, u);
  }
  if (valid) deg = re - rs;
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
    uint32_t total;
    uint32_t off = block_exclusive_sum<BS>(deg, &total, sh.wsum);
    sh.offs[threadIdx.x] = off;
    sh.start[threadIdx.x] = rs;
    sh.du[threadIdx.x] = du;
    __syncthreads();
    for (uint32_t base = 0; base < total; base += BS) {
      uint32_t e = base + threadIdx.x;
      bool act = e < total;
      uint32_t ee = 0, d0 = 0;
      if (act) {
        uint32_t lo = 0, hi = BS;
        while (hi - lo > 1) {
          uint32_t mid = (lo + hi) >> 1;
          if (sh.offs[mid] <= e) lo = mid;
          else hi = mid;
        }
        ee = sh.start[lo] + (e - sh.offs[lo]);
        d0 = sh.du[lo];
      }
      relax_edge<A, BS>(a, s, ee, d0, act, sh.ps, precheck);
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
  uint32_t rs = a.row_start[u], re = a.row_start[u + 1];
  if ((uint64_t)k * a.heavy_deg >= re - rs) return;
  uint32_t lo = rs + k * a.heavy_deg;
  uint32_t hi = min(re, lo + a.heavy_deg);
  if ((a.sched || a.feat) && threadIdx.x == 0) atomicAdd(&sh.edges, (unsigned long long)(hi - lo));
  for (uint32_t b = lo; b < hi; b += BS) {
    uint32_t e = b + threadIdx.x;
    relax_edge<A, BS>(a, s, e, du, e < hi, sh.ps, true);
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
  __shared__ uint32_t decide, consumed;
  __shared__ uint64_t phases_seen;
  Ctrl *c = a.ctrl;
  if (threadIdx.x == 0) {
    decide = 0;
    // Every control word of the step is loaded in one batch: the step costs
    // one round trip to the coherence point instead of one per dependent load.
    sched_load(s, &c->s);
    CtrlCounts k;
    counts_load(k, c);
    abort_seen = ld_coh(&c->abort);
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
      s.set_curdelta(s.curdelta() + s.delta);
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
#ifndef LSG_LOCAL_TILES
#define LSG_LOCAL_TILES 1
#endif
#ifndef LSG_LOCAL_PRECHECK
#define LSG_LOCAL_PRECHECK 0
#endif
constexpr uint32_t kLocalTiles = LSG_LOCAL_TILES;

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
      process_tile<A, BS>(a, s, t, BS, sh, LSG_LOCAL_PRECHECK != 0);
      __syncthreads();
    }
    stage_drain<BS>(a, s, sh.ps);
    __syncthreads();  // every wavefront has taken its places from the LDS tails
    if (threadIdx.x == 0) {
      st_coh(&c->qcount[s.near_q], sh.ps.tail[0]);
      st_coh(&c->qcount[s.far_q], sh.ps.tail[1]);
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

template <Algo A, int BS>
__global__ __launch_bounds__(BS) void persistent_kernel(KArgs a) {
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
  if (threadIdx.x == 0) sh.ps.local = 0;
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
    if (a.sched || a.feat) {
      __syncthreads();
      if (threadIdx.x == 0 && sh.edges)
        __hip_atomic_fetch_add(&c->stat_edges, (uint64_t)sh.edges, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    grid_barrier_block(c, nb, a.spin_limit, [&] {
      leader_step<A>(a, lsched, labort, lfar);
      local_rounds<A, BS>(a, sh, lsched, labort, lfar);
    });
  }
}

// Initialise distances, stamps and the control block (apps/sssp/kernel.cu:23-35,469;
// apps/bfs/kernel.cu:20-32).  The sources are seeded by seed_kernel.
__global__ void init_kernel(KArgs a, uint32_t delta) {
  uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
  for (uint32_t v = tid; v < a.nnodes; v += nt) {
    a.dist[v] = kInf;
    a.near_stamp[v] = 0u;
    a.far_stamp[v] = 0u;
  }
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
    Sched s;
    memset(&s, 0, sizeof s);
    s.in_q = 0;
    s.near_q = 1;
    s.far_q = 2;
    s.round = 1;
    s.set_curdelta(0);  // apps/sssp/kernel.cu:472
    s.delta = delta;
    s.aux[1] = a.far_mode;
    s.aux[2] = a.expand_mode;
    c->s = s;
    c->stat_items = c->stat_rounds = c->stat_heavy = c->stat_phases = c->stat_edges = 0;
  }
}

// d[seed] = 0 and the seeds form the first in-queue (apps/sssp/kernel.cu:33,476-477).
// One block.  Several seeds make a multi-seed traversal: on a disjoint union of
// graphs with one seed each it computes every graph's own distances.
__global__ void seed_kernel(KArgs a, const uint32_t *seeds, uint32_t nseeds) {
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

// The arrays of one traversal: a device CSR, possibly a packed union of graphs.
struct CsrView {
  const uint32_t *row, *dst, *w;
  uint32_t nnodes;
  uint64_t nedges;
};

static int check_single_opts(const lsg_opts &o, bool sssp, int *bs, int *expand) {
  if (sssp && o.delta <= 0) return fail(LSG_ERR_ARG, "delta must be > 0 (got %d)", o.delta);
  *bs = o.block_size ? o.block_size : 256;
  if (*bs != 256 && *bs != 512 && *bs != 1024) return fail(LSG_ERR_ARG, "block_size must be 256, 512 or 1024");
  *expand = o.enable_lb ? LSG_EXPAND_LB : o.expand_mode;
  if (*expand != LSG_EXPAND_TWB && *expand != LSG_EXPAND_LB) return fail(LSG_ERR_ARG, "bad expand_mode %d", *expand);
  if (o.far_mode != LSG_FAR_REFERENCE && o.far_mode != LSG_FAR_SPLIT)
    return fail(LSG_ERR_ARG, "far_mode %d is not available for single-source runs (reference or split)", o.far_mode);
  if (o.sched && !sssp) return fail(LSG_ERR_ARG, "the learned scheduler chooses SSSP delta widths; BFS has none");
  return LSG_OK;
}

// Runs one traversal from `nseeds` seeds (host array) into dist (device,
// g.nnodes entries).  Synchronous.  With graph_launch the init, seed and
// persistent kernels are replayed as one captured hipGraph.
template <Algo A>
static int run_core(Workspace *w, const CsrView &g, const uint32_t *seeds, uint32_t nseeds, const lsg_opts &o,
                    uint32_t *dist, bool graph_launch, lsg_stats *st) {
  auto t0 = std::chrono::steady_clock::now();
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

  KArgs a;
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
      hipLaunchKernelGGL(init_kernel, dim3(init_blocks), dim3(256), 0, w->stream, a, a.delta);
      hipLaunchKernelGGL(seed_kernel, dim3(1), dim3(256), 0, w->stream, a, sd, nseeds);
      hipLaunchKernelGGL(fn, dim3(grid), dim3(bs), 0, w->stream, a);
      HIPCHK(hipGetLastError());
      return LSG_OK;
    };
    HIPCHK(hipEventRecord(w->ev0, w->stream));
    if (graph_launch) {
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
  rc = run_core<A>(w, v, &seed, 1, o, dist, o.graph_launch != 0, st);
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
  if (!d || !row_start || (nedges && (!edge_dst || !edge_data)) || nnodes == 0)
    return fail(LSG_ERR_ARG, "wrap_device: bad arguments");
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
