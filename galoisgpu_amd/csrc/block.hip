// One-workgroup near/far traversals (lsg_opts.single_block).
//
// On a graph like a 256x256 grid a traversal is 1000+ near/far rounds of a
// few hundred vertices each, and a round of the persistent kernel costs a
// chain of dependent global round trips (in-queue load, d[u], row bounds,
// edges, atomicMin, queue and count stores, the leader's schedule) of about
// 8 us.  Here one workgroup of 1024 threads runs the same near/far rounds
// (apps/sssp/kernel.cu:355-391) with workgroup barriers between rounds, and
// shortens each round to two dependent round trips:
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
// entries, and the older one is skipped.  Far entries wait in a far pile;
// when a round pushes nothing near, the threshold moves on by DELTA (or to the
// far pile's minimum) and the far entries at or below it become the next
// round's input.  BFS is the same with unit weights and no threshold.
//
// Two kernels share the round code:
//   block_kernel   one graph, one workgroup, the piles in LDS (kPileCap near
//                  and kFarCap far entries); a traversal that would outgrow
//                  them stops (kAbortCapacity) and the host runs the
//                  persistent kernel instead;
//   graphs_kernel  a batch of graphs (lsg_sssp_graphs), one workgroup per
//                  graph, the piles in HBM, sized from the graph; no
//                  workgroup waits for another, so no grid barrier and no
//                  co-residency are needed.
// There are no spin waits: every wait is a workgroup barrier.

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "sssp_dev.h"

namespace lsg {

constexpr uint32_t kBlkThreads = 1024;
constexpr uint32_t kPileCap = 2048;  // LDS: entries per round pile (two of them: 64 KB)
constexpr uint32_t kFarCap = 4096;   // LDS: entries in the far pile (64 KB)
constexpr uint32_t kFarPerThread = kFarCap / kBlkThreads;
// A BFS level wider than this continues in the persistent kernel: such a
// level is faster over the whole GPU than in one CU, and levels only widen
// on the graphs that reach it (BFS from source 0, road-like graphs of 256^2
// to 2048^2, switch at 2048 / 1024 / 512 entries: 256^2 0.98 / 0.85 / 0.82
// ms, 1024^2 1.39 / 1.31 / 1.29, persistent kernel alone 0.92 and 1.45; 1024
// keeps the levels of grids up to 1024^2 in the workgroup).  SSSP rounds
// switch only when the pile is full: at 512 entries the 256^2 grid at delta
// 256 took 3.9 instead of 2.6 ms, at 1024 no faster than at 2048.
constexpr uint32_t kBfsLevelCap = 1024;

// Relaxes kInFlight edges per lane: the atomicMins and the neighbours' row
// bounds are issued together, then every improved neighbour is pushed near or
// far.  Inactive lanes hold an in-range vertex (their loads are harmless).
template <Algo A, int AS>
__device__ __forceinline__ void relax_batch(const KArgs &a, const uint32_t (&v)[kInFlight],
                                            const uint32_t (&nd)[kInFlight], const bool (&act)[kInFlight],
                                            uint32_t thr, const Sink<AS> &s) {
  uint32_t old[kInFlight];
  uint64_t rb[kInFlight];  // row_start[v], row_start[v + 1]
#pragma unroll
  for (int k = 0; k < kInFlight; ++k) {
    old[k] = act[k] ? atomic_min(a.dist + v[k], nd[k]) : 0u;
    rb[k] = ld_pair(a.row_start + v[k]);
  }
#pragma unroll
  for (int k = 0; k < kInFlight; ++k) asm volatile("" : "+v"(old[k]));
  bool to_near[kInFlight], to_far[kInFlight];
  Entry e[kInFlight];
#pragma unroll
  for (int k = 0; k < kInFlight; ++k) {
    const bool imp = act[k] && nd[k] < old[k];
    const bool near = A == Algo::BFS || nd[k] <= thr;
    to_near[k] = imp && near;
    to_far[k] = imp && !near;
    e[k] = Entry{v[k], nd[k], (uint32_t)rb[k], (uint32_t)(rb[k] >> 32)};
  }
  push2<kInFlight, AS>(to_near, to_far, e, s);
}

template <Algo A, int AS>
struct BlockRelax {
  uint32_t thr;
  Sink<AS> s;
  __device__ __forceinline__ void operator()(const KArgs &a, const uint32_t (&v)[kInFlight],
                                             const uint32_t (&nd)[kInFlight], const bool (&act)[kInFlight]) const {
    relax_batch<A, AS>(a, v, nd, act, thr, s);
  }
  // Once a push has found a pile full the traversal continues elsewhere, so
  // what is left of the round is wasted work: a heavy vertex's loop stops.
  __device__ __forceinline__ bool stopped() const {
    return __hip_atomic_load(s.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != kAbortNone;
  }
};

// One round of the one-workgroup kernels: expands in[0, n), appending to `s`.
template <Algo A, int AS>
__device__ __forceinline__ void expand_round(const KArgs &a, EntryPtr<AS> in, uint32_t n, uint32_t thr,
                                             const Sink<AS> &s, unsigned long long *items) {
  expand_entries<A, AS, kBlkThreads>(a, in, n, BlockRelax<A, AS>{thr, s}, items);
}

// apps/sssp/kernel.cu:383: curdelta += DELTA, and straight on to the far
// pile's minimum when the buckets between are empty.
__device__ __forceinline__ uint32_t next_threshold(uint32_t thr, uint32_t delta, uint32_t fmin) {
  uint64_t t = (uint64_t)thr + delta;
  if (fmin != kInf && fmin > t) t += ((uint64_t)fmin - t + delta - 1) / delta * delta;
  return t > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)t;
}

// Round state in LDS, changed only between barriers by thread 0 (`abort` may
// be raised mid-round).
struct RoundState {
  uint32_t count[2];
  uint32_t far_count;
  uint32_t far_min[2];  // the far pile's minimum, found at a split: far_min[phase & 1]
  uint32_t abort;
  uint32_t cur;    // pile holding this round's input
  uint32_t thr;    // entries of distance <= thr are near
  uint32_t state;  // after a round: kRun, kSplit, kDone
  uint32_t phases;
  uint32_t fcur;   // graphs_kernel: far buffer holding the far pile
  unsigned long long items;  // entries expanded (stale ones not counted)
  unsigned long long rounds;
};

// After a round (every thread, after the round's barrier; thread 0 decides).
__device__ __forceinline__ void end_round(RoundState &rs) {
  if (threadIdx.x == 0) {
    const uint32_t cur = rs.cur, nxt = cur ^ 1u;
    ++rs.rounds;
    rs.count[cur] = 0;
    rs.cur = nxt;
    if (rs.abort != kAbortNone) rs.state = kDone;
    else if (rs.count[nxt] > 0) rs.state = kRun;
    else if (rs.far_count == 0) rs.state = kDone;
    else {
      rs.state = kSplit;
      ++rs.phases;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void init_round_state(RoundState &rs, uint32_t n0, bool fits, uint32_t thr0) {
  rs.count[0] = fits ? n0 : 0u;
  rs.count[1] = 0;
  rs.far_count = 0;
  rs.far_min[0] = kInf;
  rs.far_min[1] = kInf;
  rs.abort = fits ? kAbortNone : kAbortCapacity;
  rs.cur = 0;
  rs.thr = thr0;
  rs.state = fits ? kRun : kDone;
  rs.phases = 0;
  rs.fcur = 0;
  rs.items = 0;
  rs.rounds = 0;
}

template <Algo A>
__global__ __launch_bounds__(kBlkThreads) void block_kernel(KArgs a) {
  __shared__ Entry pile[2][kPileCap];  // this round's input and the next round's
  __shared__ Entry far[kFarCap];
  __shared__ RoundState rs;
  Ctrl *c = a.ctrl;
  const uint32_t tid = threadIdx.x;
  // the seed kernel left the seeds in queue 0 and their count in the schedule
  const uint32_t n0 = c->s.in_count;
  const bool fits = n0 <= kPileCap;
  // curdelta starts at 0 (apps/sssp/kernel.cu:472)
  if (tid == 0) init_round_state(rs, n0, fits, A == Algo::SSSP ? 0u : 0xFFFFFFFFu);
  if (fits)
    for (uint32_t i = tid; i < n0; i += kBlkThreads) {
      const uint32_t v = a.q[0][i];
      pile[0][i] = Entry{v, 0u, glb(a.row_start)[v], glb(a.row_start)[v + 1]};
    }
  __syncthreads();
  for (;;) {
    if (rs.state == kDone) break;
    const uint32_t cur = rs.cur, nxt = cur ^ 1u;
    const Sink<kLdsAS> s{(EntryPtr<kLdsAS>)pile[nxt], &rs.count[nxt], A == Algo::BFS ? kBfsLevelCap : kPileCap,
                         (EntryPtr<kLdsAS>)far, &rs.far_count, kFarCap, &rs.abort};
    expand_round<A, kLdsAS>(a, (EntryPtr<kLdsAS>)pile[cur], rs.count[cur], rs.thr, s, &rs.items);
    __syncthreads();  // every push of the round is in
    end_round(rs);
    if (rs.state != kSplit) continue;
    // Split the far pile at the new threshold: entries at or below it are the
    // next round's input, the rest stay.  A far pile past half full first
    // drops the entries whose vertex has since been lowered below them.
    const uint32_t nf = rs.far_count, in = rs.cur, ph = rs.phases & 1u;
    Entry fe[kFarPerThread];
    bool has[kFarPerThread];
    uint32_t m = kInf;
#pragma unroll
    for (uint32_t k = 0; k < kFarPerThread; ++k) {
      const uint32_t i = k * kBlkThreads + tid;
      has[k] = i < nf;
      fe[k] = has[k] ? far[i] : Entry{0u, kInf, 0u, 0u};
      m = min(m, fe[k].d);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = min(m, (uint32_t)__shfl_xor(m, off));
    if (lane_id() == 0 && m != kInf) atomicMin(&rs.far_min[ph], m);
    __syncthreads();  // every entry is in registers, far_min[ph] is the pile's minimum
    const uint32_t thr = next_threshold(rs.thr, a.delta, rs.far_min[ph]);
    const bool prune = nf > kFarCap / 2;
    uint32_t dnow[kFarPerThread];
#pragma unroll
    for (uint32_t k = 0; k < kFarPerThread; ++k)
      dnow[k] = prune && has[k] && fe[k].d > thr ? ld_coh(a.dist + fe[k].v) : 0u;
    if (tid == 0) {
      rs.far_count = 0;
      rs.far_min[ph ^ 1u] = kInf;  // for the next split (this one's is read by every thread now)
    }
    __syncthreads();  // every thread has the threshold; the far pile is empty
    if (tid == 0) rs.thr = thr;
    bool to_in[kFarPerThread], keep[kFarPerThread];
#pragma unroll
    for (uint32_t k = 0; k < kFarPerThread; ++k) {
      to_in[k] = has[k] && fe[k].d <= thr;
      keep[k] = has[k] && fe[k].d > thr && !(prune && dnow[k] < fe[k].d);
    }
    const Sink<kLdsAS> sp{(EntryPtr<kLdsAS>)pile[in], &rs.count[in], kPileCap, (EntryPtr<kLdsAS>)far,
                          &rs.far_count, kFarCap, &rs.abort};
    push2<(int)kFarPerThread, kLdsAS>(to_in, keep, fe, sp);
    __syncthreads();
    if (tid == 0) {
      // done on an overflow, or when only stale entries were left
      rs.state = rs.abort != kAbortNone || (rs.count[in] == 0 && rs.far_count == 0) ? kDone : kRun;
    }
    __syncthreads();
  }
  if (tid == 0) {
    st_coh(&c->abort, rs.abort);
    st_coh64(&c->stat_items, rs.items);
    st_coh64(&c->stat_rounds, rs.rounds);
    st_coh64(&c->stat_phases, A == Algo::SSSP ? (unsigned long long)rs.phases : rs.rounds);
    c->s.set_curdelta(rs.thr);
  }
}

// One graph of a graphs_kernel batch: its CSR, the caller's distance array,
// its source, and its piles in the shared pile buffer (entries pile_off ...:
// two round piles of near_cap entries, two far buffers of far_cap).
struct GraphJob {
  const uint32_t *row, *dst, *w;
  uint32_t *dist;
  uint64_t nedges;
  uint64_t pile_off;
  uint32_t nnodes, seed, near_cap, far_cap;
};

struct GraphStat {
  uint32_t abort, rounds, phases, pad;
  unsigned long long items;
};

// SSSP of graph blockIdx.x of a batch, from its source, into its distance
// array (INF everywhere else).  The rounds are block_kernel's, with the piles
// in HBM: a round pile holds the entries of one round (at most one per
// improvement, so nedges + 1 bound it), and the far pile is split into the
// other far buffer.
__global__ __launch_bounds__(kBlkThreads) void graphs_kernel(const GraphJob *jobs, Entry *piles, GraphStat *stats,
                                                            uint32_t delta) {
  __shared__ RoundState rs;
  __shared__ uint32_t low[2];  // at a split: some far entry lies in the next bucket (low[phase & 1])
  const uint32_t tid = threadIdx.x;
  const GraphJob jb = jobs[blockIdx.x];
  KArgs a;
  a.row_start = jb.row;
  a.edge_dst = jb.dst;
  a.edge_data = jb.w;
  a.dist = jb.dist;
  a.nnodes = jb.nnodes;
  a.nedges = jb.nedges;
  a.delta = delta;
  const EntryPtr<kGlobalAS> base = (EntryPtr<kGlobalAS>)(piles + jb.pile_off);
  const EntryPtr<kGlobalAS> rp[2] = {base, base + jb.near_cap};
  const EntryPtr<kGlobalAS> fp[2] = {base + 2ull * jb.near_cap, base + 2ull * jb.near_cap + jb.far_cap};
  // d = INF (apps/sssp/kernel.cu:23-35), then d[source] = 0 and the source is
  // the first round's input (:33,476-477)
  for (uint32_t v = tid; v < jb.nnodes; v += kBlkThreads) st_coh(a.dist + v, kInf);
  if (tid == 0) {
    init_round_state(rs, 1u, true, 0u);
    low[0] = low[1] = 0u;
  }
  __syncthreads();  // every INF is stored
  if (tid == 0) {
    st_coh(a.dist + jb.seed, 0u);
    st_entry<kGlobalAS>(rp[0], Entry{jb.seed, 0u, glb(a.row_start)[jb.seed], glb(a.row_start)[jb.seed + 1]});
  }
  __syncthreads();
  for (;;) {
    if (rs.state == kDone) break;
    const uint32_t cur = rs.cur, nxt = cur ^ 1u, f = rs.fcur;
    const Sink<kGlobalAS> s{rp[nxt], &rs.count[nxt], jb.near_cap, fp[f], &rs.far_count, jb.far_cap, &rs.abort};
    expand_round<Algo::SSSP, kGlobalAS>(a, rp[cur], rs.count[cur], rs.thr, s, &rs.items);
    __syncthreads();  // every push of the round is in
    end_round(rs);
    if (rs.state != kSplit) continue;
    // Split far buffer f at the new threshold into the next round's pile and
    // far buffer f^1.  First pass: does the next bucket hold an entry, and the
    // minimum in case it does not.
    const uint32_t nf = rs.far_count, in = rs.cur, ph = rs.phases & 1u;
    const uint32_t t1 = next_threshold(rs.thr, delta, kInf);
    bool lo = false;
    uint32_t m = kInf;
    for (uint32_t i = tid; i < nf; i += kBlkThreads) {
      const uint32_t d = fp[f][i].d;
      lo = lo || d <= t1;
      m = min(m, d);
    }
    const bool wave_lo = __ballot(lo) != 0;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = min(m, (uint32_t)__shfl_xor(m, off));
    if (lane_id() == 0) {
      if (wave_lo) low[ph] = 1u;
      if (m != kInf) atomicMin(&rs.far_min[ph], m);
    }
    __syncthreads();
    const uint32_t thr = low[ph] ? t1 : next_threshold(rs.thr, delta, rs.far_min[ph]);
    if (tid == 0) {
      rs.far_count = 0;
      // the next split's flags (this one's are read by every thread now)
      rs.far_min[ph ^ 1u] = kInf;
      low[ph ^ 1u] = 0u;
    }
    __syncthreads();  // every thread has the threshold; far buffer f^1 is empty
    if (tid == 0) {
      rs.thr = thr;
      rs.fcur = f ^ 1u;
    }
    const bool prune = nf > jb.far_cap / 2;
    const Sink<kGlobalAS> sp{rp[in], &rs.count[in], jb.near_cap, fp[f ^ 1u], &rs.far_count, jb.far_cap, &rs.abort};
    for (uint32_t b0 = 0; b0 < nf; b0 += kBlkThreads) {
      const uint32_t i = b0 + tid;
      bool to_in[1], keep[1];
      Entry fe[1] = {Entry{0u, kInf, 0u, 0u}};
      if (i < nf) fe[0] = ld_entry<kGlobalAS>(fp[f] + i);
      const uint32_t dnow = prune && i < nf && fe[0].d > thr ? ld_coh(a.dist + fe[0].v) : 0u;
      to_in[0] = i < nf && fe[0].d <= thr;
      keep[0] = i < nf && fe[0].d > thr && !(prune && dnow < fe[0].d);
      push2<1, kGlobalAS>(to_in, keep, fe, sp);
    }
    __syncthreads();
    if (tid == 0) rs.state = rs.abort != kAbortNone || (rs.count[in] == 0 && rs.far_count == 0) ? kDone : kRun;
    __syncthreads();
  }
  if (tid == 0) {
    GraphStat gs;
    gs.abort = rs.abort;
    gs.rounds = (uint32_t)rs.rounds;
    gs.phases = rs.phases;
    gs.pad = 0;
    gs.items = rs.items;
    stats[blockIdx.x] = gs;
  }
}

// One traversal from the seeds in `seeds`; returns LSG_OK, an error, or
// kBlockFallback when a pile overflowed (the caller runs the persistent
// kernel instead).
template <Algo A>
static int run_block(Workspace *w, const CsrView &g, const uint32_t *seeds, uint32_t nseeds, const lsg_opts &o,
                     uint32_t *dist, lsg_stats *st, Resume *resume) {
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
  // The workgroup records no round trace: a traced call's trace holds the
  // persistent kernel's rounds after an overflow, and is empty otherwise.
  if (o.trace) {
    w->last_trace.clear();
    w->last_feat.clear();
  }
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
  // A pile overflow is the only way it stops early, at the end of the round
  // in which it happened; the distances reached so far stay in dist.  SSSP
  // continues with curdelta at the threshold it had reached (entries at or
  // below it are near) from every reached vertex.  BFS continues from the
  // level of that round (round r expands level r).
  // (No round at all: the seeds did not fit, and level 0 is theirs.)
  const uint32_t level = hc.stat_rounds ? (uint32_t)(hc.stat_rounds - 1) : 0u;
  if (A == Algo::SSSP) *resume = Resume{hc.s.curdelta(), 0u, kInf - 1u};
  else *resume = Resume{0, level, level};
  return hc.abort == kAbortCapacity ? kBlockFallback : LSG_OK;
}

int run_block_single(int algo, Workspace *w, const CsrView &g, uint32_t seed, const lsg_opts &o, uint32_t *dist,
                     lsg_stats *st, Resume *resume) {
  return algo == (int)Algo::SSSP ? run_block<Algo::SSSP>(w, g, &seed, 1, o, dist, st, resume)
                                 : run_block<Algo::BFS>(w, g, &seed, 1, o, dist, st, resume);
}

int run_graphs_block(Workspace *w, const lsg_graph *graphs, uint32_t ngraphs, const uint32_t *srcs,
                     int32_t *const *dist_out, const lsg_opts &o, lsg_batch_stats *stats) {
  if (o.delta <= 0) return fail(LSG_ERR_ARG, "delta must be > 0 (got %d)", o.delta);
  // LSG_GRAPHS_PILE_CAP: a smaller pile capacity, so that tests can make the
  // piles overflow and check the fallback
  uint64_t cap_limit = UINT32_MAX / 2;
  if (const char *e = getenv("LSG_GRAPHS_PILE_CAP")) cap_limit = std::max<uint64_t>(strtoull(e, nullptr, 10), 64);
  std::vector<GraphJob> jobs(ngraphs);
  uint64_t off = 0;
  for (uint32_t i = 0; i < ngraphs; ++i) {
    const lsg_graph &g = graphs[i];
    // a round pile holds at most one entry per improvement of the round
    // (<= nedges) plus the source; the far pile is pruned when half full
    const uint64_t cap = std::min<uint64_t>(g.nedges + g.nnodes + 64, cap_limit);
    jobs[i] = GraphJob{g.row_start, g.edge_dst, (const uint32_t *)g.edge_data, (uint32_t *)dist_out[i], g.nedges, off,
                       g.nnodes, srcs[i], (uint32_t)cap, (uint32_t)cap};
    off += 4 * cap;
  }
  const size_t jobs_bytes = ((size_t)ngraphs * sizeof(GraphJob) + 255) & ~(size_t)255;
  const size_t stat_bytes = ((size_t)ngraphs * sizeof(GraphStat) + 255) & ~(size_t)255;
  void *buf;
  int rc;
  if ((rc = ws_buffer(w, kBufGraphPiles, jobs_bytes + stat_bytes + off * sizeof(Entry), &buf))) return rc;
  GraphJob *djobs = (GraphJob *)buf;
  GraphStat *dstats = (GraphStat *)((char *)buf + jobs_bytes);
  Entry *piles = (Entry *)((char *)buf + jobs_bytes + stat_bytes);
  HIPCHK(hipMemcpyAsync(djobs, jobs.data(), (size_t)ngraphs * sizeof(GraphJob), hipMemcpyHostToDevice, w->stream));
  HIPCHK(hipEventRecord(w->ev0, w->stream));
  hipLaunchKernelGGL(graphs_kernel, dim3(ngraphs), dim3(kBlkThreads), 0, w->stream, (const GraphJob *)djobs, piles,
                     dstats, (uint32_t)o.delta);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(w->ev1, w->stream));
  std::vector<GraphStat> hs(ngraphs);
  HIPCHK(hipMemcpyAsync(hs.data(), dstats, (size_t)ngraphs * sizeof(GraphStat), hipMemcpyDeviceToHost, w->stream));
  HIPCHK(hipStreamSynchronize(w->stream));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, w->ev0, w->ev1));
  for (uint32_t i = 0; i < ngraphs; ++i)
    if (hs[i].abort != kAbortNone) return kBlockFallback;
  if (stats) {
    memset(stats, 0, sizeof *stats);
    stats->ms_kernel = ms;
    for (const GraphStat &g : hs) {
      stats->rounds = std::max<uint64_t>(stats->rounds, g.rounds);
      stats->phases = std::max<uint64_t>(stats->phases, g.phases);
      stats->items += g.items;
    }
    stats->grid_blocks = (int32_t)ngraphs;
    stats->block_size = (int32_t)kBlkThreads;
    stats->launches = 1;
  }
  return LSG_OK;
}

}  // namespace lsg
