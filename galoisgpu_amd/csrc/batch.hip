// Batched SSSP on gfx950: up to 64 sources on one graph in one persistent
// launch, vertex-major.
//
// No reference counterpart: the reference runs one source per process
// (apps/sssp/kernel.cu:455-480).  Each source follows exactly the plain
// method's rules -- fixed-delta near/far piles, curdelta += DELTA per phase
// (apps/sssp/kernel.cu:129-144,355-391) -- so its distances are identical to a
// single-source run; what changes is that all sources share every round.
//
// Layout in HBM: D[v][s] (u32, row stride S <= 64): the 64 lanes of a
// wavefront carry the 64 sources, so relaxing edge u->v for every source is one
// coalesced row load of D[v] and masked atomicMin on the same row, and the
// edge's column and weight are read once for all sources.  Per-vertex 64-bit
// masks say which sources must expand a vertex: nmask[2][V] for the near pile
// (indexed by round parity: written in round r, consumed in round r+1) and
// fmask[2][V] for the far pile (phase parity).  A vertex enters a pile's queue
// when its mask goes from 0 to non-zero, so each queue holds a vertex once.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <vector>

#include "engine.h"
#include "lsg_internal.h"
#include "runtime.h"
#include "sched_mlp.h"

namespace lsg {

// Flattened path (items of degree < kLongDeg): row loads issued ahead per
// wavefront and frontier items per ticket.  Measured on RMAT-22 x 64 sources
// when every item took this path (ms per launch): 16 / 16 at 4 waves/SIMD
// 65-66; 8 / 8 at 5 waves 68; 8 / 8 at 6 waves 68-69; 4 / 8 at 6 waves 69-70.
constexpr int kUnroll = 16;
// Frontier items a wavefront takes per ticket: their masks, row bounds and
// distance rows are loaded together and their edges relaxed as one list.
constexpr int kItems = 16;
static_assert(kItems >= 1 && kItems <= kWave, "items per ticket");
// Waves per SIMD the register allocation must allow: 4 (128 VGPRs).  5 or 6
// spill and were slower (42 vs 34 ms on the item-major path).
constexpr int kMultiWaves = 4;

struct MArgs {
  const uint32_t *row_start;
  const uint32_t *edge_dst;
  const uint32_t *edge_data;
  uint32_t *D;          // [V][S]
  uint64_t *nmask[2];   // near masks by round parity
  uint64_t *fmask[2];   // far masks by phase parity
  uint32_t *q[3];       // item queues (vertex ids)
  uint64_t *hq[2];      // heavy chunks: two u64 words each: (k << 32 | u), mask
  const uint32_t *srcs; // S source vertices
  Ctrl *ctrl;
  uint32_t nnodes, S, qcap, hcap, delta, heavy_deg;
  uint64_t spin_limit, residency_limit;
  lsg_round_trace *trace;
  uint32_t trace_cap;
  uint32_t sched;       // 1 = learned delta per phase
  const float *mlp;     // scheduler weights (sched_mlp.h layout)
  uint64_t nedges;
  uint32_t far_mode;    // initial far-pile mode (LSG_FAR_*)
  float *feat;          // per-phase feature records [kMaxForced][kFeatRec], or null
  uint32_t *fmin;       // [2][64] smallest pending far distance per source, by phase parity
  uint32_t *thr;        // [64] per-source near/far threshold of the current phase (bucket mode)
  uint32_t *qstamp;     // [V] last round whose in-queue holds the vertex
};

// Pushes are staged per wavefront in LDS, kPushBuf items per pile, and
// written out with one queue-tail atomic per flush: one atomic per wave_push
// on a tail that every wavefront of the grid shares serialises at the
// memory-side atomic unit (~12 ns each on gfx950).  512 entries measured
// faster than 256 (26.0 vs 26.5 ms on RMAT-22 x 64 sources).
constexpr uint32_t kPushBuf = 512;

// Per-round, per-lane view of the schedule used by the edge loop.
struct LaneSched {
  uint32_t thr;   // near/far threshold of this lane's source
  bool bucket;    // per-source thresholds (LSG_FAR_BUCKET phase)
  uint32_t fmin;  // smallest distance this lane pushed to the far pile this round
  uint32_t thr32; // threshold of this lane as u32: a candidate nd goes near iff nd <= thr32
  uint32_t *pb;   // this wavefront's LDS push buffers: [0, kPushBuf) near, [kPushBuf, 2 kPushBuf) far
  uint32_t pbn[2];  // items staged per pile (wave-uniform)
};

// Writes pile `which`'s staged items to its queue (one atomic for all).
__device__ __forceinline__ void push_flush(const MArgs &a, const Sched &s, LaneSched &ls, int which) {
  const uint32_t n = ls.pbn[which];
  if (n == 0) return;
  const uint32_t qi = which ? s.far_q : s.near_q;
  uint32_t base = 0;
  if (lane_id() == 0) base = atomic_add(&a.ctrl->qcount[qi], n);
  base = readlane(base, 0);
  const uint32_t *buf = ls.pb + which * kPushBuf;
  for (uint32_t i = lane_id(); i < n; i += kWave) {
    const uint32_t pos = base + i;
    if (pos < a.qcap) st_coh(a.q[qi] + pos, buf[i]);
    else st_coh(&a.ctrl->abort, kAbortCapacity);
  }
  ls.pbn[which] = 0;
}

// Pushes `item` from every lane with `pred` onto pile `which` (0 near, 1 far).
__device__ __forceinline__ void pile_push(const MArgs &a, const Sched &s, LaneSched &ls, int which, bool pred,
                                          uint32_t item) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return;
  const uint32_t n = (uint32_t)__builtin_popcountll(m);
  if (ls.pbn[which] + n > kPushBuf) push_flush(a, s, ls, which);
  if (pred) ls.pb[which * kPushBuf + ls.pbn[which] + lanes_below(m)] = item;
  ls.pbn[which] += n;
}

// Counters the leader turns into trace records and scheduler features.
struct MultiCounters {
  uint64_t lane_items;   // sum over expanded items of active sources
  uint64_t edges;        // edges expanded (per vertex, shared by its sources)
  uint64_t improved;     // rows in which at least one lane improved
};

__device__ __forceinline__ uint64_t bcast64(uint64_t x, uint32_t l) {
  uint32_t lo = readlane((uint32_t)x, l), hi = readlane((uint32_t)(x >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

// Near-pile push admission: a vertex joins the next round's queue at most
// once (a pop that took its next-round bits leaves it queued while its mask
// is empty again).
__device__ __forceinline__ bool near_admit(const MArgs &a, const Sched &s, uint32_t v) {
  return __hip_atomic_fetch_max(glb(a.qstamp + v), s.round + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < s.round + 1;
}

// Distance rows of the items a wavefront is expanding: dl[slot][lane] is the
// item's distance for source `lane`, INF where that source does not expand it.
// Each lane only reads its own column, so this is per-lane indexed storage.
typedef uint32_t DistSlots[kItems][kWave];

// Relaxes a batch of n <= 64 edges held one per lane: lane j has edge j's
// destination ev, weight ew and the slot ek of the item it leaves.  The row
// loads of a group of kUnroll edges are all issued before any is used, then
// the group's atomicMins; lane j collects edge j's near / far source bits, and
// the masks of the n destinations are updated with one vector atomic each.
__device__ __forceinline__ void relax_batch(const MArgs &a, const Sched &s, LaneSched &ls, const DistSlots &dl,
                                            uint32_t n, uint32_t ev, uint32_t ew, uint32_t ek,
                                            uint64_t *improved_rows) {
  const uint32_t lane = lane_id();
  // Near/far threshold of this lane: the shared curdelta (apps/sssp/kernel.cu:134)
  // or, in a bucket phase, the lane's own source threshold.
  const int64_t thr = ls.bucket ? (int64_t)ls.thr : s.curdelta();
  // Column this lane loads: its own source, or for lanes >= S (no source, no
  // candidate) the row's last word, so that every load stays inside row v.
  const uint32_t col = min(lane, a.S - 1);
  uint64_t my_nb = 0, my_fb = 0;
  for (uint32_t j0 = 0; j0 < n; j0 += kUnroll) {
    uint32_t v[kUnroll], nd[kUnroll], cur[kUnroll], old[kUnroll];
    bool want[kUnroll];
    // Candidates first (LDS and register work only) ...
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) {
      const uint32_t j = min(j0 + k, n - 1);  // a tail slot repeats the batch's last edge, with no candidate
      v[k] = readlane(ev, j);
      const uint32_t du = dl[readlane(ek, j)][lane];
      nd[k] = du == kInf || j0 + k >= n ? kInf : du + readlane(ew, j);
    }
    // ... then the group's row loads, unconditional so that none waits on a
    // branch: every lane loads a word of D[v], whose row (S words) is fetched
    // anyway when any lane needs it.
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) cur[k] = ld_coh(a.D + (size_t)v[k] * a.S + col);
    // every atomicMin of the group is issued before any result is used
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) {
      want[k] = nd[k] < cur[k];
      old[k] = want[k] ? atomic_min(a.D + (size_t)v[k] * a.S + lane, nd[k]) : kInf;
    }
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) {
      bool imp = want[k] && nd[k] < old[k];
      bool nearl = imp && (int64_t)nd[k] <= thr;
      if (imp && !nearl) ls.fmin = min(ls.fmin, nd[k]);
      uint64_t nb = __ballot(nearl), fb = __ballot(imp && !nearl);
      if (nb | fb) *improved_rows += 1;
      if (lane == j0 + k) {
        my_nb = nb;
        my_fb = fb;
      }
    }
  }
  // Two edges of a batch may share a destination; only the lane whose OR
  // finds the mask empty queues it.
  bool pn = my_nb != 0 && atomic_or64(a.nmask[s.round & 1] + ev, my_nb) == 0 && near_admit(a, s, ev);
  bool pf = my_fb != 0 && atomic_or64(a.fmask[s.phase & 1] + ev, my_fb) == 0;
  pile_push(a, s, ls, 0, pn, ev);
  pile_push(a, s, ls, 1, pf, ev);
}

// ---- item-major relaxation ------------------------------------------------
//
// Items of degree >= kLongDeg and heavy chunks are relaxed one item at a time:
// the item's distance for this lane's source stays in a register, its edges
// come kGroup at a time through the scalar cache (one s_load_dwordx8 for the
// destinations, one for the weights), and each edge costs one row load, an
// add and a compare per wavefront -- no lane shuffles.  Window width,
// measured on RMAT-22 x 64 sources: 8 wide and software-pipelined 33.6 ms,
// 16 wide 34.2 (pipelined or not; pipelined it spills).  Improvements are
// collected in a per-wavefront list and the destinations' masks are updated
// 64 at a time.
constexpr int kGroup = 8;
// A work ticket covers kSpan units (28.0 -> 27.2 ms from 1 to 2; 4 was
// slower, 29.1 ms, from load imbalance).
constexpr uint32_t kSpan = 2;
constexpr uint32_t kLongDeg = kGroup;  // a relax window must fit inside the item

typedef uint32_t U32xG __attribute__((ext_vector_type(kGroup), aligned(4)));
typedef const __attribute__((address_space(4))) U32xG ScalarXG;
typedef const __attribute__((address_space(4))) uint32_t ScalarU32;

// Words [e, e + kGroup) of a CSR array that the launch never writes, loaded
// into scalar registers with one instruction (e is wave-uniform).
__device__ __forceinline__ U32xG scalar_window(const uint32_t *arr, uint32_t e) {
  return *(ScalarXG *)(arr + e);
}

// Improved edges since the last flush: lane i holds the destination of the
// i-th and the sources for which it must join the near / far pile.
struct Pending {
  uint32_t v, nb_lo, nb_hi, fb_lo, fb_hi;
  uint32_t n;  // wave-uniform
};


// Sets the pending destinations' mask bits; a destination whose mask was
// empty joins the pile's queue (two entries for one destination: only the
// lane whose OR finds the mask empty pushes it).
__device__ __forceinline__ void flush_pending(const MArgs &a, const Sched &s, LaneSched &ls, Pending &p) {
  if (p.n == 0) return;
  const bool act = lane_id() < p.n;
  const uint64_t nb = ((uint64_t)p.nb_hi << 32) | p.nb_lo, fb = ((uint64_t)p.fb_hi << 32) | p.fb_lo;
  const bool pn = act && nb != 0 && atomic_or64(a.nmask[s.round & 1] + p.v, nb) == 0 && near_admit(a, s, p.v);
  const bool pf = act && fb != 0 && atomic_or64(a.fmask[s.phase & 1] + p.v, fb) == 0;
  pile_push(a, s, ls, 0, pn, p.v);
  pile_push(a, s, ls, 1, pf, p.v);
  p.n = 0;
}

// One window of kGroup edges in flight: candidates nd, then the loaded
// distances, then the atomics' results; dst holds destination k in lane k.
struct Window {
  uint32_t nd[kGroup];
  uint32_t val[kGroup];
  uint32_t dst;
};

// Issues the row loads of a window of edges (destinations vv, weights ww)
// leaving an item whose distance for this lane's source is du (INF: no
// candidate); the first `skip` edges carry no candidate (relaxed already).
__device__ __forceinline__ void window_load(const MArgs &a, uint32_t du, const U32xG &vv, const U32xG &ww,
                                            uint32_t skip, uint32_t col, Window &w) {
#pragma unroll
  for (int k = 0; k < kGroup; ++k) {
    // du + w never wraps (du <= INF, w <= INT_MAX); INF + w and a skipped
    // slot's du + 2^31 both exceed every stored distance
    w.nd[k] = du + ((uint32_t)k < skip ? 0x80000000u : ww[k]);
  }
#pragma unroll
  for (int k = 0; k < kGroup; ++k) w.val[k] = ld_coh(a.D + (size_t)vv[k] * a.S + col);
}

// Issues the atomicMins of the lanes whose candidate beats the loaded distance.
__device__ __forceinline__ void window_atomics(const MArgs &a, const U32xG &vv, uint32_t col, Window &w) {
#pragma unroll
  for (int k = 0; k < kGroup; ++k)
    w.val[k] = w.nd[k] < w.val[k] ? atomic_min(a.D + (size_t)vv[k] * a.S + col, w.nd[k]) : 0u;
}

// Takes a window's atomic results (issued earlier, so the wait overlaps the
// loads issued since) and records its improved edges.
__device__ __forceinline__ void window_results(const MArgs &a, LaneSched &ls, Window &w, Pending &p,
                                               uint64_t &improved) {
  const uint32_t lane = lane_id();
  // Results are taken only here (left to itself the compiler tests each
  // result inside its atomic's branch, waiting for the whole memory queue).
#pragma unroll
  for (int k = 0; k < kGroup; ++k) asm volatile("" : "+v"(w.val[k]));
#pragma unroll
  for (int k = 0; k < kGroup; ++k) {
    const bool imp = w.nd[k] < w.val[k];
    const uint64_t im = __ballot(imp);
    if (im == 0) continue;
    const uint32_t v = readlane(w.dst, k);
    const bool nearl = imp && w.nd[k] <= ls.thr32;
    const uint64_t nb = __ballot(nearl), fb = im & ~nb;
    if (fb && imp && !nearl) ls.fmin = min(ls.fmin, w.nd[k]);
    improved += 1;
    if (lane == p.n) {
      p.v = v;
      p.nb_lo = (uint32_t)nb;
      p.nb_hi = (uint32_t)(nb >> 32);
      p.fb_lo = (uint32_t)fb;
      p.fb_hi = (uint32_t)(fb >> 32);
    }
    p.n += 1;
  }
}

// Relaxes edges [lo, hi) of one item with at least kGroup edges (distance du
// for this lane's source), kGroup at a time.  The last window is shifted back to end
// at hi, so every window lies inside the item's edges, and its slots that
// the previous window covered are skipped.  Software-pipelined: while a
// window's row loads are in flight the next window's edges are fetched and
// the previous window's atomic results are taken, so a window costs about one
// memory round trip instead of three.
__device__ __forceinline__ void relax_edges(const MArgs &a, const Sched &s, LaneSched &ls, uint32_t du, uint32_t lo,
                                            uint32_t hi, Pending &p, uint64_t &improved) {
  const uint32_t lane = lane_id();
  // lanes >= S have no source and du = INF: they load the row's last word
  // and never issue an atomic
  const uint32_t col = min(lane, a.S - 1);
  // the bounds are wave-uniform; say so, so that they stay in scalar registers
  lo = __builtin_amdgcn_readfirstlane(lo);
  hi = __builtin_amdgcn_readfirstlane(hi);
  uint32_t e = lo, w0 = min(lo, hi - kGroup);
  U32xG vv = scalar_window(a.edge_dst, w0), ww = scalar_window(a.edge_data, w0);
  Window cur, prev;
  bool have_prev = false;
  for (;;) {
    window_load(a, du, vv, ww, e - w0, col, cur);
    cur.dst = lane < kGroup ? glb(a.edge_dst)[w0 + lane] : 0u;
    const uint32_t en = e + kGroup;
    const bool more = en < hi;
    const uint32_t wn = more ? min(en, hi - kGroup) : w0;
    const U32xG vvn = scalar_window(a.edge_dst, wn), wwn = scalar_window(a.edge_data, wn);
    if (have_prev) window_results(a, ls, prev, p, improved);
    if (p.n > (uint32_t)(kWave - kGroup)) flush_pending(a, s, ls, p);
    window_atomics(a, vv, col, cur);
    prev = cur;
    have_prev = true;
    if (!more) break;
    vv = vvn;
    ww = wwn;
    e = en;
    w0 = wn;
  }
  window_results(a, ls, prev, p, improved);
}

// Queues heavy vertex u (edges [rs, re)) as chunks of heavy_deg edges, each
// expanded by a wavefront in the next round with the sources active now.
__device__ __forceinline__ void queue_heavy(const MArgs &a, const Sched &s, uint32_t u, uint32_t rs, uint32_t re,
                                            uint64_t am) {
  const uint32_t lane = lane_id();
  uint32_t nch = (re - rs + a.heavy_deg - 1) / a.heavy_deg;
  uint32_t hqi = 1u - s.heavy_in;
  uint32_t base = 0;
  if (lane == 0) base = atomic_add(&a.ctrl->hcount[hqi], nch);
  base = readlane(base, 0);
  for (uint32_t k = lane; k < nch; k += kWave) {
    uint32_t pos = base + k;
    if (pos < a.hcap) {
      st_coh64(a.hq[hqi] + 2ull * pos, ((uint64_t)k << 32) | u);
      st_coh64(a.hq[hqi] + 2ull * pos + 1, am);
    } else {
      st_coh(&a.ctrl->abort, kAbortCapacity);
    }
  }
}

// Frontier items [base, base + cnt), cnt <= kItems: lane k consumes item k's
// mask and loads its row bounds; the distance rows of all the items are then
// loaded together and parked in dl, and the items' edges are relaxed as one
// flattened list, 64 per batch.  Heavy items are queued as chunks instead.
__device__ __forceinline__ void process_items(const MArgs &a, const Sched &s, LaneSched &ls, uint32_t base,
                                              DistSlots &dl, MultiCounters &mc, Pending &p) {
  const uint32_t lane = lane_id();
  const uint32_t cnt = min((uint32_t)kItems, s.in_count - base);
  uint64_t *mp = s.aux[0] ? a.fmask[(s.phase - 1) & 1] : a.nmask[(s.round - 1) & 1];
  uint32_t u = a.nnodes;
  if (lane < cnt) u = ld_coh(a.q[s.in_q] + base + lane);
  uint64_t m = 0;
  uint32_t rs = 0, re = 0;
  if (u < a.nnodes) {
    m = atomic_exch64(mp + u, 0ull);
    // Gauss-Seidel pop: the bits set for the next round so far are expanded
    // now, with the distances that set them; u stays queued for the next round.
    m |= atomic_exch64(a.nmask[s.round & 1] + u, 0ull);
    rs = glb(a.row_start)[u];
    re = glb(a.row_start)[u + 1];
  }
  // The row loads depend on the exchanged masks, so each reads a distance at
  // least as new as the improvements that set the consumed bits.
  uint32_t du[kItems];
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const uint64_t mk = bcast64(m, k);
    const uint32_t uk = readlane(u, k);
    du[k] = lane < a.S && ((mk >> lane) & 1) ? ld_coh(a.D + (size_t)uk * a.S + lane) : kInf;
  }
  const bool split = s.aux[0] && s.aux[1];
  uint32_t mydeg = 0;  // lane k: edges of item k to relax in the flattened list
  uint32_t longm = 0;  // items relaxed one by one (degree >= kLongDeg)
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    bool act = du[k] != kInf;
    if (__ballot(act) == 0) continue;
    if (split) {
      // Split / bucket far pile: a source whose distance is still above this
      // phase's threshold (shared, or its own in a bucket phase) waits for a
      // later phase instead of expanding early.
      const bool defer = act && (ls.bucket ? du[k] > ls.thr : (int64_t)du[k] > s.curdelta());
      const uint64_t dm = __ballot(defer);
      if (dm) {
        const uint32_t uk = readlane(u, k);
        bool push = false;
        if (lane == 0) push = atomic_or64(a.fmask[s.phase & 1] + uk, dm) == 0;
        pile_push(a, s, ls, 1, push, uk);
        if (defer) ls.fmin = min(ls.fmin, du[k]);
        act = act && !defer;
      }
    }
    const uint64_t am = __ballot(act);
    if (am == 0) continue;
    const uint32_t rsk = readlane(rs, k), rek = readlane(re, k);
    if (rek - rsk >= a.heavy_deg) {
      queue_heavy(a, s, readlane(u, k), rsk, rek, am);
      continue;
    }
    if (rek == rsk) continue;
    dl[k][lane] = act ? du[k] : kInf;
    mc.lane_items += __builtin_popcountll(am);
    mc.edges += rek - rsk;
    if (rek - rsk >= kLongDeg) longm |= 1u << k;
    else if (lane == (uint32_t)k) mydeg = rek - rsk;
  }
  // Inclusive prefix of the items' edge counts over lanes 0..kItems-1: item k
  // owns flattened edges [end_k - deg_k, end_k).
  uint32_t end = mydeg;
#pragma unroll
  for (int d = 1; d < kItems; d <<= 1) {
    const uint32_t t = __shfl_up(end, d);
    if (lane >= (uint32_t)d) end += t;
  }
  const uint32_t total = readlane(end, kItems - 1);
  const uint32_t first = end - mydeg;
  for (uint32_t b0 = 0; b0 < total; b0 += kWave) {
    const uint32_t f = b0 + lane;
    // the item holding flattened edge f is the number of items ending at or before f
    uint32_t item = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) item += readlane(end, k) <= f;
    const uint32_t e = __shfl(rs, item) + (f - __shfl(first, item));
    uint32_t ev = 0, ew = 0;
    if (f < total) {
      ev = glb(a.edge_dst)[e];
      ew = glb(a.edge_data)[e];
    }
    relax_batch(a, s, ls, dl, min((uint32_t)kWave, total - b0), ev, ew, item, &mc.improved);
  }
  for (uint32_t lm = longm; lm; lm &= lm - 1) {
    const uint32_t k = (uint32_t)__builtin_ctz(lm);
    relax_edges(a, s, ls, dl[k][lane], readlane(rs, k), readlane(re, k), p, mc.improved);
  }
}

// Heavy chunk c: edges [k * heavy_deg, (k + 1) * heavy_deg) of u, for the
// sources recorded with the chunk.
__device__ __forceinline__ void process_chunk(const MArgs &a, const Sched &s, LaneSched &ls, uint32_t c,
                                              MultiCounters &mc, Pending &p) {
  const uint32_t lane = lane_id();
  uint64_t ent = ld_coh64(a.hq[s.heavy_in] + 2ull * c);
  uint64_t m = ld_coh64(a.hq[s.heavy_in] + 2ull * c + 1);
  uint32_t u = (uint32_t)ent, k = (uint32_t)(ent >> 32);
  if (u >= a.nnodes) return;
  uint32_t rs = glb(a.row_start)[u], re = glb(a.row_start)[u + 1];
  if ((uint64_t)k * a.heavy_deg >= re - rs) return;
  uint32_t lo = rs + k * a.heavy_deg, hi = min(re, lo + a.heavy_deg);
  bool act = lane < a.S && ((m >> lane) & 1);
  uint32_t du = act ? ld_coh(a.D + (size_t)u * a.S + lane) : kInf;
  act = act && du != kInf;
  uint64_t am = __ballot(act);
  if (am == 0) return;
  mc.lane_items += __builtin_popcountll(am);
  mc.edges += hi - lo;
  relax_edges(a, s, ls, act ? du : kInf, lo, hi, p, mc.improved);
}

// Leader step, run by every thread of the last block to reach the barrier:
// the pile rotation of apps/sssp/kernel.cu:363-384 for all sources at once,
// plus, at each phase boundary, the learned scheduler's choice of the next
// phase's delta width and far-pile mode.  Schedule words: aux[0] = the
// in-queue is the far pile, aux[1] = far mode of this phase (LSG_FAR_*).
__device__ __forceinline__ void multi_leader(const MArgs &a, MultiCounters *tot) {
  __shared__ Sched s;
  __shared__ SchedFeatures f;
  __shared__ MlpScratch sc;
  __shared__ uint32_t phase_end, consumed;
  __shared__ uint64_t phases_seen;
  Ctrl *c = a.ctrl;
  const uint32_t t = threadIdx.x;
  if (t == 0) {
    // every control word of the step in one batch of loads (the schedule to
    // registers first, so that the other loads need not wait for LDS stores)
    Sched sr;
    sched_load(sr, &c->s);
    CtrlCounts k;
    counts_load(k, c);
    const uint64_t tot_lane_items = ld_coh64(&tot->lane_items), tot_edges = ld_coh64(&tot->edges);
    s = sr;
    // every wavefront has left the round's item loops: restart the tickets
    st_coh(&c->ticket, 0u);
    st_coh(&c->ticket2, 0u);
    const uint32_t nnear = k.q(s.near_q), nfar = k.q(s.far_q), nheavy = k.h(1u - s.heavy_in);
    phase_end = 0;
    phases_seen = k.phases;
    if (nnear > a.qcap || nfar > a.qcap || nheavy > a.hcap) {
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
      st_coh(&c->hcount[s.heavy_in], 0u);
      s.heavy_in = 1u - s.heavy_in;
      s.heavy_count = nheavy;
      if (nnear > 0 || nheavy > 0) {
        s.in_q = s.near_q;
        s.in_count = nnear;
        s.near_q = consumed;
        st_coh(&c->qcount[consumed], 0u);
        s.aux[0] = 0;
        s.round += 1;
      } else if (nfar > 0) {
        phase_end = 1;
        f.in_items = nfar;
        f.phase = s.phase;
        f.round = s.round;
        f.curdelta = (float)s.curdelta();
        f.base_delta = a.delta;
        f.prev_delta = s.delta;
        f.lane_items = (float)tot_lane_items;
        f.edges = (float)tot_edges;
        f.nnodes = a.nnodes;
        f.nedges = (float)a.nedges;
        f.nsrc = a.S;
      } else {
        s.done = 1;
        st_coh64(&c->stat_phases, k.phases + 1);
      }
    }
  }
  __syncthreads();
  if (!phase_end) {
    if (t == 0) sched_store(&c->s, s);
    return;
  }
  __shared__ int act_sh;
  int act = -1;
  if (a.sched) act = mlp_block_action(a.mlp, f, sc);
  if (t == 0) {
    if (act >= 0) {
      s.delta = action_delta(a.mlp, act, a.delta);
      s.aux[1] = (uint32_t)a.mlp[kOffFarMode + act];
    }
    act_sh = act;
  }
  __syncthreads();
  // Per-source thresholds of the next phase: smallest pending far distance
  // of each source + delta - 1; the recorded minima of the phase after start
  // empty.  (Tracked in every phase so any phase can switch to bucket mode.)
  if (t < a.S) {
    uint32_t m = ld_coh(a.fmin + (s.phase & 1) * 64 + t);
    uint64_t th = m == kInf ? (uint64_t)kInf : (uint64_t)m + s.delta - 1;
    st_coh(a.thr + t, (uint32_t)min(th, (uint64_t)kInf - 1));
    st_coh(a.fmin + ((s.phase + 1) & 1) * 64 + t, kInf);
  }
  if (t == 0) {
    act = act_sh;
    if (a.feat && s.phase < (uint32_t)kMaxForced) {
      float *rec = a.feat + (size_t)s.phase * kFeatRec;
      sched_featurize(f, rec);
      rec[kFeat] = (float)act;
      rec[kFeat + 1] = (float)s.delta;
      rec[kFeat + 2] = (float)s.aux[1];
      rec[kFeat + 3] = (float)(s.round + 1);  // first round of the phase being started
    }
    s.in_q = s.far_q;
    s.in_count = f.in_items;
    s.far_q = consumed;
    st_coh(&c->qcount[consumed], 0u);
    s.aux[0] = 1;
    s.set_curdelta(s.curdelta() + s.delta);
    s.phase += 1;
    s.round += 1;
    st_coh64(&c->stat_phases, phases_seen + 1);
    sched_store(&c->s, s);
  }
}

template <int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(kMultiWaves, kMultiWaves)))
void multi_kernel(MArgs a, MultiCounters *tot) {
  __shared__ Sched ss;
  __shared__ uint32_t quit;
  __shared__ unsigned long long blk[3];
  __shared__ DistSlots dls[BS / kWave];
  __shared__ uint32_t pushbuf[BS / kWave][2 * kPushBuf];
  __shared__ uint32_t drain_n[BS / kWave][2], drain_base[2];
  Ctrl *c = a.ctrl;
  const uint32_t nb = gridDim.x;
  if (!residency_handshake(c, nb, a.residency_limit)) return;
  const uint32_t wpb = BS / kWave;
  const uint32_t wave = blockIdx.x * wpb + threadIdx.x / kWave;
  const uint32_t nwaves = nb * wpb;
  for (;;) {
    if (threadIdx.x == 0) {
      Sched s;
      sched_load(s, &c->s);
      ss = s;
      quit = s.done || ld_coh(&c->abort) != kAbortNone;
      blk[0] = blk[1] = blk[2] = 0;
    }
    __syncthreads();
    if (quit) break;
    const Sched s = ss;
    MultiCounters mc = {0, 0, 0};
    LaneSched ls;
    ls.bucket = s.aux[1] == LSG_FAR_BUCKET;
    ls.thr = ls.bucket && lane_id() < a.S ? ld_coh(a.thr + lane_id()) : 0u;
    ls.thr32 = ls.bucket ? ls.thr : (uint32_t)min(s.curdelta(), (int64_t)UINT32_MAX);
    ls.fmin = kInf;
    ls.pb = &pushbuf[threadIdx.x / kWave][0];
    ls.pbn[0] = ls.pbn[1] = 0;
    Pending p;
    p.n = 0;
    p.v = p.nb_lo = p.nb_hi = p.fb_lo = p.fb_hi = 0;
    DistSlots &dl = dls[threadIdx.x / kWave];
    // Static first pass (ticket = wavefront id), then dynamic tickets; the
    // next ticket is fetched while the current one is processed.
    const uint32_t nchunks = (s.in_count + kItems - 1) / kItems;
    // A ticket covers kSpan consecutive work units; the next ticket is taken
    // when the last unit of the current range starts.
    for (uint32_t ck = wave, cend = wave + 1; ck < nchunks;) {
      const bool last = ck + 1 >= cend;
      uint32_t t = 0;
      if (last && lane_id() == 0) t = atomic_add(&c->ticket, 1u);
      process_items(a, s, ls, ck * kItems, dl, mc, p);
      if (last) {
        ck = nwaves + readlane(t, 0) * kSpan;
        cend = ck + kSpan;
      } else {
        ++ck;
      }
    }
    for (uint32_t ck = wave, cend = wave + 1; ck < s.heavy_count;) {
      const bool last = ck + 1 >= cend;
      uint32_t t = 0;
      if (last && lane_id() == 0) t = atomic_add(&c->ticket2, 1u);
      process_chunk(a, s, ls, ck, mc, p);
      if (last) {
        ck = nwaves + readlane(t, 0) * kSpan;
        cend = ck + kSpan;
      } else {
        ++ck;
      }
    }
    flush_pending(a, s, ls, p);
    // this wavefront's smallest far push per source -> phase minimum
    if (__ballot(ls.fmin != kInf) && ls.fmin != kInf) atomic_min(a.fmin + (s.phase & 1) * 64 + lane_id(), ls.fmin);
    const uint32_t wv = threadIdx.x / kWave;
    if (lane_id() == 0) {
      drain_n[wv][0] = ls.pbn[0];
      drain_n[wv][1] = ls.pbn[1];
      if (mc.lane_items | mc.edges) {
        atomicAdd(&blk[0], (unsigned long long)mc.lane_items);
        atomicAdd(&blk[1], (unsigned long long)mc.edges);
        atomicAdd(&blk[2], (unsigned long long)mc.improved);
      }
    }
    __syncthreads();
    // The block's staged pushes get their places with one queue-tail atomic
    // per pile (one per wavefront, 4096 on one tail, serialise for ~50 us at
    // the memory-side atomic unit every round).
    if (threadIdx.x < 2) {
      const uint32_t which = threadIdx.x;
      uint32_t tot_n = 0;
      for (uint32_t w = 0; w < BS / kWave; ++w) {
        const uint32_t n = drain_n[w][which];
        drain_n[w][which] = tot_n;  // becomes the wavefront's offset
        tot_n += n;
      }
      drain_base[which] = tot_n ? atomic_add(&c->qcount[which ? s.far_q : s.near_q], tot_n) : 0u;
    }
    if (threadIdx.x == 0 && (blk[0] | blk[1])) {
      __hip_atomic_fetch_add(&tot->lane_items, (uint64_t)blk[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&tot->edges, (uint64_t)blk[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&tot->improved, (uint64_t)blk[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
#pragma unroll
    for (int which = 0; which < 2; ++which) {
      const uint32_t n = ls.pbn[which];
      if (n == 0) continue;
      const uint32_t qi = which ? s.far_q : s.near_q;
      const uint32_t base = drain_base[which] + drain_n[wv][which];
      const uint32_t *buf = ls.pb + which * kPushBuf;
      for (uint32_t i = lane_id(); i < n; i += kWave) {
        const uint32_t pos = base + i;
        if (pos < a.qcap) st_coh(a.q[qi] + pos, buf[i]);
        else st_coh(&a.ctrl->abort, kAbortCapacity);
      }
    }
    grid_barrier_block(c, nb, a.spin_limit, [&] { multi_leader(a, tot); });
  }
}

// D[v][s] = INF, masks cleared.
__global__ void multi_init(MArgs a, MultiCounters *tot) {
  size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x, nt = (size_t)gridDim.x * blockDim.x;
  size_t cells = (size_t)a.nnodes * a.S;
  for (size_t i = tid; i < cells; i += nt) a.D[i] = kInf;
  for (size_t v = tid; v < a.nnodes; v += nt) {
    a.nmask[0][v] = 0;
    a.nmask[1][v] = 0;
    a.fmask[0][v] = 0;
    a.fmask[1][v] = 0;
    a.qstamp[v] = 0;
  }
  if (tid == 0) *tot = MultiCounters{0, 0, 0};
}

// Seeds every source (apps/sssp/kernel.cu:33,476-477) and the control block.
// One wavefront; lane s seeds source s.
__global__ void multi_seed(MArgs a) {
  const uint32_t lane = threadIdx.x;
  bool act = lane < a.S;
  uint32_t src = act ? a.srcs[lane] : 0u;
  bool push = false;
  if (act) {
    a.D[(size_t)src * a.S + lane] = 0u;
    push = atomic_or64(a.nmask[0] + src, 1ull << lane) == 0;
    a.qstamp[src] = 1u;  // in round 1's queue
  }
  uint64_t pm = __ballot(push);
  if (push) a.q[0][lanes_below(pm)] = src;
  ctrl_reset_barrier(a.ctrl, lane, kWave);
  if (lane == 0) {
    Ctrl *c = a.ctrl;
    c->qcount[0] = (uint32_t)__builtin_popcountll(pm);
    c->qcount[1] = 0;
    c->qcount[2] = 0;
    c->hcount[0] = 0;
    c->hcount[1] = 0;
    c->ticket = 0;
    c->ticket2 = 0;
    c->abort = kAbortNone;
    c->checkin = 0;
    c->resident_seen = 0;
    Sched s;
    memset(&s, 0, sizeof s);
    s.in_q = 0;
    s.near_q = 1;
    s.far_q = 2;
    s.in_count = (uint32_t)__builtin_popcountll(pm);
    s.round = 1;
    s.set_curdelta(0);
    s.delta = a.delta;
    s.aux[1] = a.far_mode;
    c->s = s;
    for (int k = 0; k < 64; ++k) {
      a.fmin[k] = kInf;
      a.fmin[64 + k] = kInf;
      a.thr[k] = 0u;  // phase 0: threshold 0, as curdelta starts at 0 (apps/sssp/kernel.cu:472)
    }
    c->stat_items = c->stat_rounds = c->stat_heavy = c->stat_phases = c->stat_edges = 0;
  }
}

// out[k][v] = D[v][k] for the units of one launch (per-source arrays).
__global__ void multi_scatter(const uint32_t *D, uint32_t S, uint32_t nnodes, int32_t *const *outs) {
  size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x, nt = (size_t)gridDim.x * blockDim.x;
  size_t cells = (size_t)nnodes * S;
  for (size_t i = tid; i < cells; i += nt) {
    uint32_t v = (uint32_t)(i / S), k = (uint32_t)(i % S);
    outs[k][v] = (int32_t)D[i];
  }
}

// Out-degree of each source (deg[i] for srcs[i]).
__global__ void source_degree_kernel(const uint32_t *row_start, const uint32_t *srcs, uint32_t n, uint32_t *deg) {
  if (threadIdx.x < n) deg[threadIdx.x] = row_start[srcs[threadIdx.x] + 1] - row_start[srcs[threadIdx.x]];
}

// D [V][S] from the compact Dc [V][Sc] of the sources that have out-edges:
// column s is Dc's column col[s], or, for a source without out-edges, INF
// except 0 at the source itself.  One wavefront per row, lane = column (8
// rows per wavefront step measured slower: 0.63 vs 0.48 ms on RMAT-22).
__global__ void multi_expand(const uint32_t *Dc, uint32_t Sc, const int32_t *col, const uint32_t *srcs, uint32_t S,
                             uint32_t V, uint32_t *D) {
  const uint32_t lane = lane_id();
  const size_t nwaves = (size_t)gridDim.x * blockDim.x / kWave;
  const int32_t c = lane < S ? col[lane] : -1;
  const uint32_t sv = lane < S ? srcs[lane] : 0u;
  for (size_t v = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) / kWave; v < V; v += nwaves)
    if (lane < S) D[v * S + lane] = c >= 0 ? Dc[v * Sc + (uint32_t)c] : (v == sv ? 0u : kInf);
}

using MultiFn = void (*)(MArgs, MultiCounters *);

static MultiFn pick_multi(int bs) {
  switch (bs) {
    case 256: return multi_kernel<256>;
    case 1024: return multi_kernel<1024>;
    default: return multi_kernel<512>;
  }
}

struct MultiRun {
  double ms_kernel = 0;
  uint64_t rounds = 0, phases = 0, items = 0, heavy = 0, lane_items = 0, edges = 0, improved = 0;
  int grid = 0, bs = 0;
};

// Runs one vertex-major launch: S = nsrc <= 64 sources of `g` into D [V][S].
static int run_multi(Workspace *w, const lsg_graph *g, const uint32_t *srcs_host, uint32_t nsrc, const lsg_opts &o,
                     uint32_t *D, MultiRun *out) {
  int rc;
  int bs = o.block_size ? o.block_size : 512;
  if (bs != 256 && bs != 512 && bs != 1024) return fail(LSG_ERR_ARG, "block_size must be 256, 512 or 1024");
  // default chunk size, measured on RMAT-22 x 64 sources: 512 30.0 ms, 256
  // 34.3, 1024 30.6, 128 63
  uint32_t heavy_deg = o.heavy_deg > 0 ? (uint32_t)o.heavy_deg : 512u;
  if (heavy_deg < kLongDeg) return fail(LSG_ERR_ARG, "heavy_deg must be >= %u for batched runs", kLongDeg);
  uint32_t V = g->nnodes;
  uint64_t hcap64 = 2ull * g->nedges / heavy_deg + 64;
  uint32_t hcap = (uint32_t)std::min<uint64_t>(hcap64, UINT32_MAX / 4);
  void *qbuf, *hbuf, *mbuf, *sbuf;
  if ((rc = ws_buffer(w, kBufQueues, 3ull * V * 4, &qbuf)) || (rc = ws_buffer(w, kBufHeavy, 2ull * hcap * 16, &hbuf)) ||
      (rc = ws_buffer(w, kBufMasks, 4ull * V * 8 + 4ull * V + 256, &mbuf)) || (rc = ws_buffer(w, kBufSeeds, 2048, &sbuf)))
    return rc;
  MArgs a;
  memset(&a, 0, sizeof a);
  a.row_start = g->row_start;
  a.edge_dst = g->edge_dst;
  a.edge_data = (const uint32_t *)g->edge_data;
  a.D = D;
  uint64_t *masks = (uint64_t *)mbuf;
  a.nmask[0] = masks;
  a.nmask[1] = masks + V;
  a.fmask[0] = masks + 2ull * V;
  a.fmask[1] = masks + 3ull * V;
  a.qstamp = (uint32_t *)(masks + 4ull * V);
  for (int i = 0; i < 3; ++i) a.q[i] = (uint32_t *)qbuf + (size_t)i * V;
  a.hq[0] = (uint64_t *)hbuf;
  a.hq[1] = (uint64_t *)hbuf + 2ull * hcap;
  a.srcs = (const uint32_t *)sbuf;
  a.ctrl = w->ctrl;
  a.nnodes = V;
  a.S = nsrc;
  a.qcap = V;
  a.hcap = hcap;
  a.delta = (uint32_t)o.delta;
  a.heavy_deg = heavy_deg;
  a.spin_limit = spin_limit_ticks(w);
  a.residency_limit = residency_limit_ticks(w);
  a.nedges = g->nedges;
  MultiCounters *tot = (MultiCounters *)((char *)sbuf + 256);
  a.fmin = (uint32_t *)((char *)sbuf + 512);
  a.thr = (uint32_t *)((char *)sbuf + 1024);
  if (o.sched) {
    a.sched = 1;
    a.mlp = sched_model_device(w->device, LSG_SCHED_BATCHED);
    if (!a.mlp)
      return fail(LSG_ERR_ARG, "learned scheduler requested but no batched-kernel model is loaded (lsg_sched_set_model)");
  }
  if (o.far_mode != LSG_FAR_REFERENCE && o.far_mode != LSG_FAR_SPLIT && o.far_mode != LSG_FAR_BUCKET)
    return fail(LSG_ERR_ARG, "bad far_mode %d", o.far_mode);
  a.far_mode = (uint32_t)o.far_mode;
  if (o.trace) {
    void *tbuf;
    size_t tbytes = (1u << 16) * sizeof(lsg_round_trace);
    if ((rc = ws_buffer(w, kBufTrace, tbytes + kMaxForced * kFeatRec * sizeof(float), &tbuf))) return rc;
    a.trace = (lsg_round_trace *)tbuf;
    a.trace_cap = 1u << 16;
    a.feat = (float *)((char *)tbuf + tbytes);
  }
  HIPCHK(hipMemcpyAsync(sbuf, srcs_host, nsrc * 4, hipMemcpyHostToDevice, w->stream));
  MultiFn fn = pick_multi(bs);
  int grid = 0;
  if ((rc = resident_grid(w, (const void *)fn, bs, &grid))) return rc;
  if (o.grid_blocks > 0) grid = std::min(o.grid_blocks, grid);
  Ctrl hc;
  float ms = 0;
  for (int attempt = 0;; ++attempt) {
    uint32_t init_blocks = (uint32_t)std::min<uint64_t>(((uint64_t)V * nsrc + 255) / 256, 16ull * w->cus);
    auto enqueue = [&]() -> int {
      hipLaunchKernelGGL(multi_init, dim3(init_blocks), dim3(256), 0, w->stream, a, tot);
      hipLaunchKernelGGL(multi_seed, dim3(1), dim3(64), 0, w->stream, a);
      hipLaunchKernelGGL(fn, dim3(grid), dim3(bs), 0, w->stream, a, tot);
      HIPCHK(hipGetLastError());
      return LSG_OK;
    };
    HIPCHK(hipEventRecord(w->ev0, w->stream));
    if (o.graph_launch) {
      std::vector<uint8_t> key;
      key_append(key, a);
      key_append(key, tot);
      key_append(key, grid);
      key_append(key, bs);
      if ((rc = launch_captured(w, key, enqueue))) return rc;
    } else if ((rc = enqueue())) {
      return rc;
    }
    HIPCHK(hipEventRecord(w->ev1, w->stream));
    HIPCHK(hipMemcpyAsync(&hc, w->ctrl, kCtrlHostBytes, hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    HIPCHK(hipEventElapsedTime(&ms, w->ev0, w->ev1));
    if (hc.abort != kAbortResidency) break;
    if (attempt > 0 || hc.resident_seen == 0 || hc.resident_seen >= (uint32_t)grid)
      return fail(LSG_ERR_HIP, "persistent grid of %d x %d blocks is not co-resident (%u were)", grid, bs,
                  hc.resident_seen);
    grid = (int)hc.resident_seen;
    record_resident_grid(w, (const void *)fn, bs, grid);
  }
  if (hc.abort == kAbortTimeout)
    return fail(LSG_ERR_TIMEOUT, "grid barrier timed out (grid %d x %d) at round %u; distances are incomplete",
                grid, bs, hc.s.round);
  if (hc.abort == kAbortCapacity) return fail(LSG_ERR_CAPACITY, "device queue capacity exceeded");
  MultiCounters ht;
  HIPCHK(hipMemcpy(&ht, tot, sizeof ht, hipMemcpyDeviceToHost));
  if (o.trace) {
    uint32_t n = (uint32_t)std::min<uint64_t>(hc.stat_rounds, a.trace_cap);
    w->last_trace.resize(n);
    if (n) HIPCHK(hipMemcpy(w->last_trace.data(), a.trace, n * sizeof(lsg_round_trace), hipMemcpyDeviceToHost));
    // one record per phase boundary: phase p -> p + 1 is record p
    uint32_t np = std::min<uint32_t>(hc.s.phase, kMaxForced);
    w->last_feat.resize((size_t)np * kFeatRec);
    if (np) HIPCHK(hipMemcpy(w->last_feat.data(), a.feat, w->last_feat.size() * sizeof(float), hipMemcpyDeviceToHost));
  }
  out->ms_kernel = ms;
  out->rounds = hc.stat_rounds;
  out->phases = hc.stat_phases;
  out->items = hc.stat_items;
  out->heavy = hc.stat_heavy;
  out->lane_items = ht.lane_items;
  out->edges = ht.edges;
  out->improved = ht.improved;
  out->grid = grid;
  out->bs = bs;
  return LSG_OK;
}

// Runs nsrc <= 64 sources into D [V][nsrc].  A source without out-edges
// reaches only itself, so it gets no lane: the launch runs the other sources
// into a compact matrix (rows of fewer words: on RMAT-22, where half the
// vertices have no out-edges, 64 random sources leave 33 lanes and rows of
// 132 instead of 256 bytes, 25.5 -> 23.6 ms), which is then expanded.
static int run_multi_live(Workspace *w, const lsg_graph *g, const uint32_t *srcs, uint32_t nsrc, const lsg_opts &o,
                          uint32_t *D, MultiRun *out) {
  int rc;
  void *lbuf;
  if ((rc = ws_buffer(w, kBufSeedList, 4 * 64 * 4, &lbuf))) return rc;
  uint32_t *d_srcs = (uint32_t *)lbuf, *d_deg = d_srcs + 64, *d_col = d_srcs + 128;
  uint32_t deg[64];
  HIPCHK(hipMemcpyAsync(d_srcs, srcs, nsrc * 4, hipMemcpyHostToDevice, w->stream));
  hipLaunchKernelGGL(source_degree_kernel, dim3(1), dim3(64), 0, w->stream, g->row_start, (const uint32_t *)d_srcs,
                     nsrc, d_deg);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(deg, d_deg, nsrc * 4, hipMemcpyDeviceToHost, w->stream));
  HIPCHK(hipStreamSynchronize(w->stream));
  uint32_t live[64];
  int32_t col[64];
  uint32_t nlive = 0;
  for (uint32_t i = 0; i < nsrc; ++i) {
    col[i] = deg[i] ? (int32_t)nlive : -1;
    if (deg[i]) live[nlive++] = srcs[i];
  }
  if (nlive == nsrc) return run_multi(w, g, srcs, nsrc, o, D, out);
  const uint32_t V = g->nnodes;
  void *cbuf = nullptr;
  if (nlive) {
    if ((rc = ws_buffer(w, kBufTranspose, (size_t)V * nlive * 4, &cbuf))) return rc;
    if ((rc = run_multi(w, g, live, nlive, o, (uint32_t *)cbuf, out))) return rc;
  } else {
    *out = MultiRun();
    w->last_trace.clear();
    w->last_feat.clear();
  }
  HIPCHK(hipMemcpyAsync(d_col, col, nsrc * 4, hipMemcpyHostToDevice, w->stream));
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(((uint64_t)V * kWave + 255) / 256, 16ull * w->cus);
  hipLaunchKernelGGL(multi_expand, dim3(blocks), dim3(256), 0, w->stream, (const uint32_t *)cbuf, nlive,
                     (const int32_t *)d_col, (const uint32_t *)d_srcs, nsrc, V, D);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(w->stream));
  return LSG_OK;
}

static int check_multi_args(const lsg_graph *g, const uint32_t *srcs, uint32_t nsrc, const lsg_opts &o) {
  if (!g || !g->device_graph) return fail(LSG_ERR_ARG, "expected a device graph (copy_to_gpu first)");
  if (!srcs || nsrc == 0) return fail(LSG_ERR_ARG, "no sources");
  if (o.delta <= 0) return fail(LSG_ERR_ARG, "delta must be > 0 (got %d)", o.delta);
  for (uint32_t i = 0; i < nsrc; ++i)
    if (srcs[i] >= g->nnodes) return fail(LSG_ERR_ARG, "source %u = %u out of range [0, %u)", i, srcs[i], g->nnodes);
  return LSG_OK;
}

static void fill_stats(lsg_batch_stats *st, const MultiRun &r, int launches) {
  st->ms_kernel += r.ms_kernel;
  st->rounds += r.rounds;
  st->phases += r.phases;
  st->items += r.items;
  st->heavy_chunks += r.heavy;
  st->lane_items += r.lane_items;
  st->edges_expanded += r.edges;
  st->improved_rows += r.improved;
  st->grid_blocks = r.grid;
  st->block_size = r.bs;
  st->launches = launches;
}

}  // namespace lsg

using namespace lsg;

extern "C" {

int lsg_sssp_multi(const lsg_graph *g, const uint32_t *srcs, uint32_t nsrc, const lsg_opts *opts, int32_t *dist_vs,
                   lsg_batch_stats *stats) {
  clear_error();
  auto t0 = std::chrono::steady_clock::now();
  lsg_opts o;
  if (opts) o = *opts;
  else lsg_opts_default(&o);
  int rc = check_multi_args(g, srcs, nsrc, o);
  if (rc) return rc;
  if (nsrc > 64) return fail(LSG_ERR_ARG, "lsg_sssp_multi takes at most 64 sources (got %u)", nsrc);
  if (!dist_vs) return fail(LSG_ERR_ARG, "NULL distance matrix");
  Workspace *w;
  if ((rc = get_workspace(g->device, &w))) return rc;
  std::lock_guard<std::mutex> lk(w->mu);
  HIPCHK(hipSetDevice(w->device));
  MultiRun r;
  if ((rc = run_multi_live(w, g, srcs, nsrc, o, (uint32_t *)dist_vs, &r))) return rc;
  if (stats) {
    memset(stats, 0, sizeof *stats);
    fill_stats(stats, r, 1);
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return LSG_OK;
}

int lsg_last_features(int device, float *out, uint32_t cap_phases, uint32_t *nphases) {
  Workspace *w;
  int rc = get_workspace(device, &w);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(w->mu);
  uint32_t np = (uint32_t)(w->last_feat.size() / kFeatRec);
  uint32_t m = std::min(np, cap_phases);
  if (out && m) memcpy(out, w->last_feat.data(), (size_t)m * kFeatRec * sizeof(float));
  if (nphases) *nphases = np;
  return LSG_OK;
}

int lsg_sssp_batch(const lsg_graph *g, const uint32_t *srcs, uint32_t nsrc, const lsg_opts *opts,
                   int32_t *const *dist_out, lsg_batch_stats *stats) {
  clear_error();
  auto t0 = std::chrono::steady_clock::now();
  lsg_opts o;
  if (opts) o = *opts;
  else lsg_opts_default(&o);
  int rc = check_multi_args(g, srcs, nsrc, o);
  if (rc) return rc;
  if (!dist_out) return fail(LSG_ERR_ARG, "NULL output array list");
  for (uint32_t i = 0; i < nsrc; ++i)
    if (!dist_out[i]) return fail(LSG_ERR_ARG, "output %u is NULL", i);
  Workspace *w;
  if ((rc = get_workspace(g->device, &w))) return rc;
  std::lock_guard<std::mutex> lk(w->mu);
  HIPCHK(hipSetDevice(w->device));
  uint32_t S = std::min<uint32_t>(nsrc, 64);
  void *dbuf, *pbuf;
  if ((rc = ws_buffer(w, kBufUnion, (size_t)g->nnodes * S * 4 + 64 * sizeof(void *), &dbuf))) return rc;
  pbuf = (char *)dbuf + (size_t)g->nnodes * S * 4;
  if (stats) memset(stats, 0, sizeof *stats);
  int launches = 0;
  for (uint32_t base = 0; base < nsrc; base += 64) {
    uint32_t n = std::min<uint32_t>(64, nsrc - base);
    MultiRun r;
    if ((rc = run_multi_live(w, g, srcs + base, n, o, (uint32_t *)dbuf, &r))) return rc;
    ++launches;
    HIPCHK(hipMemcpyAsync(pbuf, dist_out + base, n * sizeof(void *), hipMemcpyHostToDevice, w->stream));
    uint32_t blocks = (uint32_t)std::min<uint64_t>(((uint64_t)g->nnodes * n + 255) / 256, 16ull * w->cus);
    hipLaunchKernelGGL(multi_scatter, dim3(blocks), dim3(256), 0, w->stream, (const uint32_t *)dbuf, n, g->nnodes,
                       (int32_t *const *)pbuf);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(w->stream));
    if (stats) fill_stats(stats, r, launches);
  }
  if (stats)
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return LSG_OK;
}

}  // extern "C"
