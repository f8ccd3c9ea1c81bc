// Direction-optimising BFS on gfx950, and the reverse CSR it needs.
//
// The levels are the reference's (apps/bfs/kernel.cu:20-32,312-334,380-393):
// level[src] = 0, first hop LEVEL 1, INF = INT_MAX for unreachable vertices.
// BFS levels are unique, so how a level is discovered does not change them.
// All levels run in one persistent launch with a grid barrier between them.
// Every level marks the vertices it discovers in a bitmap (three bitmaps
// rotate by level mod 3) and is either
//   top-down:  the previous level's vertices -- read from its bitmap, the
//              heavy ones as chunks of edges queued when they were found --
//              have their out-edges flattened over each wavefront, and an
//              unvisited destination is claimed with a compare-and-swap on
//              its level;
//   bottom-up: every unvisited vertex scans its in-edges and stops at the
//              first parent whose bit is set in the previous level's bitmap.
// The barrier leader picks the direction of the next level from the
// frontier's size and edge count (Beamer, Asanovic and Patterson's rule:
// bottom-up once the frontier's edges exceed 1/alpha of the unexplored edges,
// top-down again once the frontier holds fewer than 1/beta of the vertices).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>

#include "engine.h"
#include "lsg_internal.h"
#include "runtime.h"
#include "scan_dev.h"

namespace lsg {

constexpr int kDoBlock = 256;
constexpr uint32_t kScatterRun = 16;  // edges per thread when finding edge sources
constexpr uint32_t kAlpha = 14, kBeta = 24;

// ---- reverse CSR ------------------------------------------------------------

// Adds each tile's offset; writes the row starts and the scatter cursors.
__global__ void scan_add_kernel(uint32_t *row, uint32_t n, const uint32_t *tile_off, uint32_t *cursor,
                                const uint32_t *total) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i <= n; i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t r = i < n ? row[i] + tile_off[i / kScanTile] : *total;
    row[i] = r;
    if (cursor && i < n) cursor[i] = r;
  }
}

// ---- reverse CSR by a stable radix sort ------------------------------------------
//
// The forward CSR lists edges by source, so sorting the (destination, source,
// weight) triples by destination with a stable sort gives each reverse row
// in ascending source order: the reverse CSR is the same on every call (a
// PageRank sum over in-edges has one order) and a bottom-up search meets low
// source ids first.  LSD radix sort, 8 bits per pass; each pass is a per-tile
// digit histogram, a prefix sum over (digit, tile), and a stable scatter that
// ranks equal digits within a wavefront by ballots.  No atomics on global
// memory: the atomic-counter version serialised on the in-degree counters of
// the hubs (RMAT-20: 4.3 ms).

constexpr uint32_t kRadixTile = 16384;                 // elements per block
constexpr uint32_t kRadixIters = kRadixTile / kDoBlock;  // elements per thread, strided by the block

// src[e] = the source of edge e; thread t takes edges [16 t, 16 t + 16).
__global__ void edge_source_kernel(const uint32_t *row_start, uint32_t V, uint32_t E, uint32_t *src) {
  const size_t nthreads = (size_t)gridDim.x * blockDim.x;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t * kScatterRun < E; t += nthreads) {
    const uint32_t e0 = (uint32_t)(t * kScatterRun), e1 = min(E, e0 + kScatterRun);
    uint32_t lo = 0, hi = V;  // last u with row_start[u] <= e0
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (row_start[mid] <= e0) lo = mid;
      else hi = mid;
    }
    uint32_t u = lo, next = row_start[u + 1];
    for (uint32_t e = e0; e < e1; ++e) {
      while (next <= e) next = row_start[++u + 1];
      src[e] = u;
    }
  }
}

// hist[d * ntiles + b] = elements of tile b whose digit is d.
__global__ __launch_bounds__(kDoBlock) void radix_hist_kernel(const uint32_t *key, uint32_t n, uint32_t shift,
                                                              uint32_t ntiles, uint32_t *hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * kRadixTile;
  for (uint32_t it = 0; it < kRadixIters; ++it) {
    const size_t e = base + (size_t)it * kDoBlock + threadIdx.x;
    if (e < n) atomicAdd(&h[(key[e] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of tile b: element e (in tile order: iteration, wavefront,
// lane) goes to off[d * ntiles + b] + the number of earlier elements of the
// tile with the same digit d.
__global__ __launch_bounds__(kDoBlock) void radix_scatter_kernel(const uint32_t *key, const uint32_t *v1,
                                                                 const uint32_t *v2, uint32_t n, uint32_t shift,
                                                                 uint32_t ntiles, const uint32_t *off,
                                                                 uint32_t *key_out, uint32_t *v1_out,
                                                                 uint32_t *v2_out) {
  constexpr uint32_t NW = kDoBlock / kWave;
  __shared__ uint32_t run[256];      // next position per digit
  __shared__ uint32_t wcnt[NW][256];  // this iteration's count per wavefront and digit
  const uint32_t lane = lane_id(), wv = threadIdx.x / kWave;
  run[threadIdx.x] = off[(size_t)threadIdx.x * ntiles + blockIdx.x];
  for (uint32_t k = 0; k < NW; ++k) wcnt[k][threadIdx.x] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * kRadixTile;
  for (uint32_t it = 0; it < kRadixIters; ++it) {
    const size_t e = base + (size_t)it * kDoBlock + threadIdx.x;
    const bool valid = e < n;
    uint32_t k = 0, a = 0, b = 0, d = 0;
    if (valid) {
      k = key[e];
      a = v1[e];
      b = v2 ? v2[e] : 0u;
      d = (k >> shift) & 255u;
    }
    // lanes holding the same digit: AND over the digit's 8 bits of the matching ballots
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (uint32_t bit = 0; bit < 8; ++bit) {
      const uint64_t m = __ballot((d >> bit) & 1u);
      peers &= ((d >> bit) & 1u) ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__builtin_popcountll(peers & ((1ull << lane) - 1));
    if (valid && rank == 0) wcnt[wv][d] = (uint32_t)__builtin_popcountll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + rank;
      for (uint32_t w = 0; w < wv; ++w) pos += wcnt[w][d];
      key_out[pos] = k;
      v1_out[pos] = a;
      v2_out[pos] = b;
    }
    __syncthreads();
    uint32_t sum = 0;
    for (uint32_t w = 0; w < NW; ++w) {
      sum += wcnt[w][threadIdx.x];
      wcnt[w][threadIdx.x] = 0;
    }
    run[threadIdx.x] += sum;
    __syncthreads();
  }
}

// Row starts of the reverse CSR from the sorted destinations: the first
// position of each destination value, E for the vertices after the last.
__global__ void row_bounds_kernel(const uint32_t *dst, uint32_t E, uint32_t V, uint32_t *row) {
  const size_t nt = (size_t)gridDim.x * blockDim.x;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i <= E; i += nt) {
    const uint64_t first = i == 0 ? 0ull : (uint64_t)dst[i - 1] + 1;  // vertices whose row starts here
    const uint64_t last = i < E ? (uint64_t)dst[i] : (uint64_t)V;        // ... through this one
    for (uint64_t v = first; v <= last; ++v) row[v] = (uint32_t)i;
  }
}

constexpr uint32_t kBuRun = 4;  // in-edges a bottom-up lane tests per step

// ---- the persistent traversal -------------------------------------------------
//
// All levels run in one co-resident launch; the barrier leader applies the
// direction rule between levels, so no level waits for the host.  Words that
// other blocks wrote earlier in the launch (levels, bitmaps, chunk queues,
// per-block counts) are read and written with agent-scope atomics or coherent
// loads and stores: the XCDs' L2 caches are not coherent with each other.
// The graph arrays are read-only and read through the caches.

constexpr uint32_t kTdHeavy = 256;  // out-degree from which a frontier vertex is spread as chunks
constexpr uint32_t kTdChunk = 256;  // edges per chunk

struct PdoArgs {
  const uint32_t *row_start, *edge_dst;  // forward CSR
  const uint32_t *in_row, *in_src;       // reverse CSR
  uint32_t *level;
  uint32_t *fb[3];       // fb[L % 3] bit v: level[v] == L
  uint64_t *hq[2];       // chunks (k << 32 | u) of the heavy vertices found at level L, in hq[L & 1]
  uint32_t *partial;     // [block][2]: vertices found by the block this level, their out-edges
  Ctrl *ctrl;            // hcount[L & 1]: chunks queued at level L
  uint32_t nnodes, words, hcap, direction;
  uint64_t spin_limit, residency_limit;
};

// Per-block tallies of one level.
struct PdoShared {
  uint32_t cnt, deg;
};

// Records vertices discovered at level L: counts, out-edges, and chunks for
// the heavy ones (used if the next level is top-down).
__device__ __forceinline__ void pdo_found(const PdoArgs &a, uint32_t L, bool found, uint32_t v, uint32_t &cnt,
                                          uint32_t &deg) {
  uint32_t d = 0;
  if (found) {
    d = a.row_start[v + 1] - a.row_start[v];
    cnt += 1;
    deg += d;
  }
  const bool heavy = found && d >= kTdHeavy;
  if (__ballot(heavy) == 0) return;
  const uint32_t nch = heavy ? (d + kTdChunk - 1) / kTdChunk : 0u;
  const uint32_t base = wave_reserve(nch, &a.ctrl->hcount[L & 1]);
  for (uint32_t k = 0; k < nch; ++k) {
    if (base + k < a.hcap) st_coh64(a.hq[L & 1] + base + k, ((uint64_t)k << 32) | v);
    else st_coh(&a.ctrl->abort, kAbortCapacity);
  }
}

// Top-down visit of v at level L: claims it with a compare-and-swap.
__device__ __forceinline__ void pdo_visit(const PdoArgs &a, uint32_t L, bool valid, uint32_t v, uint32_t &cnt,
                                          uint32_t &deg) {
  bool found = false;
  if (valid && ld_coh(a.level + v) == kInf) {
    uint32_t expect = kInf;
    found = __hip_atomic_compare_exchange_strong(a.level + v, &expect, L, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
  }
  if (found)
    __hip_atomic_fetch_or(a.fb[L % 3] + (v >> 5), 1u << (v & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  pdo_found(a, L, found, v, cnt, deg);
}

// Top-down level L.  The heavy frontier vertices come as chunks of kTdChunk
// edges, one wavefront per chunk; the light ones are read from the previous
// level's bitmap, 64 words per wavefront pass: each step takes one vertex per
// lane, and their edges are flattened over the wavefront (a prefix sum of the
// degrees, each lane finding its edge's owner by binary search over the lanes).
__device__ __forceinline__ void pdo_top_down(const PdoArgs &a, const Sched &s, uint32_t gwave, uint32_t nwaves,
                                             uint32_t &cnt, uint32_t &deg) {
  const uint32_t L = s.round, lane = lane_id();
  const uint64_t *hq = a.hq[(L - 1) & 1];
  for (uint32_t c = gwave; c < s.heavy_count; c += nwaves) {
    const uint64_t ent = ld_coh64(hq + c);
    const uint32_t u = (uint32_t)ent, k = (uint32_t)(ent >> 32);
    const uint32_t rs = a.row_start[u], re = a.row_start[u + 1];
    const uint32_t lo = rs + k * kTdChunk, hi = min(re, lo + kTdChunk);
    for (uint32_t e0 = lo; e0 < hi; e0 += kWave) {
      const uint32_t e = e0 + lane;
      pdo_visit(a, L, e < hi, e < hi ? a.edge_dst[e] : 0u, cnt, deg);
    }
  }
  const uint32_t *fr = a.fb[(L - 1) % 3];
  for (size_t w0 = (size_t)gwave * kWave; w0 < a.words; w0 += (size_t)nwaves * kWave) {
    const size_t wi = w0 + lane;
    uint32_t bits = wi < a.words ? ld_coh(fr + wi) : 0u;
    while (__ballot(bits != 0)) {
      uint32_t rs = 0, dg = 0;
      if (bits) {
        const uint32_t u = (uint32_t)wi * 32 + (uint32_t)__builtin_ctz(bits);
        bits &= bits - 1;
        rs = a.row_start[u];
        dg = a.row_start[u + 1] - rs;
        if (dg >= kTdHeavy) dg = 0;  // expanded through its chunks
      }
      uint32_t end = dg;  // inclusive prefix over the lanes
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t t = __shfl_up(end, d);
        if (lane >= (uint32_t)d) end += t;
      }
      const uint32_t total = readlane(end, kWave - 1);
      for (uint32_t b0 = 0; b0 < total; b0 += kWave) {
        const uint32_t f = b0 + lane;
        // owner: the number of lanes whose range ends at or before f
        uint32_t k = 0;
#pragma unroll
        for (uint32_t step = 32; step >= 1; step >>= 1)
          if ((uint32_t)__shfl(end, (int)(k + step - 1)) <= f) k += step;
        const uint32_t ks = __shfl(rs, (int)k), kend = __shfl(end, (int)k), kdg = __shfl(dg, (int)k);
        const bool valid = f < total;
        pdo_visit(a, L, valid, valid ? a.edge_dst[ks + (f - (kend - kdg))] : 0u, cnt, deg);
      }
    }
  }
}

// Bottom-up level L: one unvisited vertex per lane searches its in-edges for
// a parent in the previous level's bitmap; a wavefront owns 64 consecutive
// vertices and stores their two words of this level's bitmap.
__device__ __forceinline__ void pdo_bottom_up(const PdoArgs &a, const Sched &s, uint32_t &cnt, uint32_t &deg) {
  const uint32_t L = s.round, lane = lane_id();
  const uint32_t *prev = a.fb[(L - 1) % 3];
  uint32_t *cur = a.fb[L % 3];
  const size_t nt = (size_t)gridDim.x * blockDim.x;
  const size_t vend = (a.nnodes + (size_t)kWave - 1) / kWave * kWave;
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < vend; v += nt) {
    bool found = false;
    if (v < a.nnodes && ld_coh(a.level + v) == kInf) {
      const uint32_t e1 = a.in_row[v + 1];
      for (uint32_t e = a.in_row[v]; e < e1 && !found; e += kBuRun) {
        uint32_t u[kBuRun];
#pragma unroll
        for (uint32_t k = 0; k < kBuRun; ++k) u[k] = e + k < e1 ? a.in_src[e + k] : a.nnodes;
#pragma unroll
        for (uint32_t k = 0; k < kBuRun; ++k)
          found = found || (u[k] < a.nnodes && ((ld_coh(prev + (u[k] >> 5)) >> (u[k] & 31)) & 1u));
      }
      if (found) st_coh(a.level + v, L);
    }
    const uint64_t m = __ballot(found);
    if (lane == 0) {
      const uint32_t w0 = (uint32_t)(v / 32);  // the wave's first vertex, < nnodes
      st_coh(cur + w0, (uint32_t)m);
      if (w0 + 1 < a.words) st_coh(cur + w0 + 1, (uint32_t)(m >> 32));
    }
    pdo_found(a, L, found, (uint32_t)v, cnt, deg);
  }
}

// Barrier leader after level L: totals of the level, the unexplored-edge
// estimate, and the direction of level L + 1 (the rule of the host loop it
// replaces: bottom-up once the frontier's edges exceed 1/alpha of the
// unexplored edges, top-down again below 1/beta of the vertices).
__device__ __forceinline__ void pdo_leader(const PdoArgs &a) {
  __shared__ uint32_t tot[2];
  if (threadIdx.x < 2) tot[threadIdx.x] = 0;
  __syncthreads();
  uint32_t c = 0, d = 0;
  for (uint32_t b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
    c += ld_coh(a.partial + 2 * b);
    d += ld_coh(a.partial + 2 * b + 1);
  }
#pragma unroll
  for (int k = kWave / 2; k >= 1; k >>= 1) {
    c += __shfl_xor(c, k);
    d += __shfl_xor(d, k);
  }
  if (lane_id() == 0) {
    atomicAdd(&tot[0], c);
    atomicAdd(&tot[1], d);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Ctrl *ctl = a.ctrl;
    Sched s;
    sched_load(s, &ctl->s);
    const uint32_t L = s.round, n_new = tot[0], mf_new = tot[1];
    s.phase += s.aux[0];  // bottom-up levels so far
    const uint32_t m_u = s.aux[3] > s.aux[2] ? s.aux[3] - s.aux[2] : 0u;
    bool bu = s.aux[0] != 0;
    if (a.direction == 0) {
      if (!bu && (uint64_t)mf_new * kAlpha > m_u) bu = true;
      else if (bu && (uint64_t)n_new * kBeta < a.nnodes) bu = false;
    }
    s.aux[0] = bu ? 1u : 0u;
    s.aux[1] = n_new;
    s.aux[2] = mf_new;
    s.aux[3] = m_u;
    s.heavy_count = min(ld_coh(&ctl->hcount[L & 1]), a.hcap);
    st_coh(&ctl->hcount[(L + 1) & 1], 0u);
    s.round = L + 1;
    s.done = n_new == 0;
    sched_store(&ctl->s, s);
  }
}

template <int BS>
__global__ __launch_bounds__(BS) void pdo_kernel(PdoArgs a) {
  __shared__ Sched ss;
  __shared__ uint32_t quit;
  __shared__ PdoShared sh;
  Ctrl *c = a.ctrl;
  const uint32_t nb = gridDim.x;
  if (!residency_handshake(c, nb, a.residency_limit)) return;
  const uint32_t gwave = (blockIdx.x * BS + threadIdx.x) / kWave, nwaves = nb * (BS / kWave);
  for (;;) {
    if (threadIdx.x == 0) {
      sched_load(ss, &c->s);
      quit = ss.done || ld_coh(&c->abort) != kAbortNone;
      sh.cnt = sh.deg = 0;
    }
    __syncthreads();
    if (quit) break;
    const Sched s = ss;
    // the bitmap of level L + 1 was last read at level L - 1
    uint32_t *nxt = a.fb[(s.round + 1) % 3];
    for (size_t i = blockIdx.x * (size_t)BS + threadIdx.x; i < a.words; i += (size_t)nb * BS) st_coh(nxt + i, 0u);
    uint32_t cnt = 0, deg = 0;
    if (s.aux[0]) pdo_bottom_up(a, s, cnt, deg);
    else pdo_top_down(a, s, gwave, nwaves, cnt, deg);
#pragma unroll
    for (int k = kWave / 2; k >= 1; k >>= 1) {
      cnt += __shfl_xor(cnt, k);
      deg += __shfl_xor(deg, k);
    }
    if (lane_id() == 0 && cnt) {
      atomicAdd(&sh.cnt, cnt);
      atomicAdd(&sh.deg, deg);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      st_coh(a.partial + 2 * blockIdx.x, sh.cnt);
      st_coh(a.partial + 2 * blockIdx.x + 1, sh.deg);
    }
    grid_barrier_block(c, nb, a.spin_limit, [&] { pdo_leader(a); });
  }
}

// level = INF except src = 0; bitmap 0 = {src}, bitmaps 1 and 2 empty; the
// control block and the schedule of level 1; src's chunks if it is heavy.
__global__ void pdo_init_kernel(PdoArgs a, uint32_t src, uint32_t src_deg, uint32_t bottom_up, uint64_t nedges) {
  const size_t nt = (size_t)gridDim.x * blockDim.x, t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  for (size_t v = t; v < a.nnodes; v += nt) a.level[v] = v == src ? 0u : kInf;
  for (size_t i = t; i < a.words; i += nt) {
    a.fb[0][i] = i == src / 32 ? 1u << (src & 31) : 0u;
    a.fb[1][i] = 0;
    a.fb[2][i] = 0;
  }
  const uint32_t nch = src_deg >= kTdHeavy ? (src_deg + kTdChunk - 1) / kTdChunk : 0u;
  for (size_t k = t; k < nch; k += nt) a.hq[0][k] = ((uint64_t)k << 32) | src;
  Ctrl *c = a.ctrl;
  ctrl_reset_barrier(c, t, nt);
  if (t == 0) {
    c->abort = kAbortNone;
    c->checkin = 0;
    c->resident_seen = 0;
    c->progress = 0;
    c->hcount[0] = c->hcount[1] = 0;
    Sched s;
    memset(&s, 0, sizeof s);
    s.round = 1;
    s.heavy_count = nch;
    s.aux[0] = bottom_up;
    s.aux[1] = 1;
    s.aux[2] = src_deg;
    s.aux[3] = (uint32_t)nedges;
    c->s = s;
  }
}

static int check_device_graph(const lsg_graph *g, const char *what) {
  if (!g || g->device_graph == 0 || !g->row_start || (g->nedges && !g->edge_dst))
    return fail(LSG_ERR_ARG, "%s: expected a device graph (copy_to_gpu first)", what);
  return LSG_OK;
}

}  // namespace lsg

using namespace lsg;

extern "C" {

int lsg_graph_transpose(const lsg_graph *g, lsg_graph *gt) {
  clear_error();
  int rc = check_device_graph(g, "transpose");
  if (rc) return rc;
  if (!gt) return fail(LSG_ERR_ARG, "transpose: NULL output graph");
  Workspace *w;
  if ((rc = get_workspace(g->device, &w))) return rc;
  std::lock_guard<std::mutex> lk(w->mu);
  HIPCHK(hipSetDevice(w->device));
  const uint32_t V = g->nnodes, E = g->nedges;
  // (destination, source, weight) triples sorted by destination, 8 bits per
  // pass; pass p writes slot (P - p) & 1, so the last pass lands in slot 1,
  // the output arrays.  Pass 0 reads the forward arrays and the sources,
  // which sit in the slot pass 0 does not write.
  uint32_t bits = 0;
  while (bits < 32 && (V > 1 ? (uint64_t)(V - 1) >> bits : 0) != 0) ++bits;
  const uint32_t P = std::max<uint32_t>(1, (bits + 7) / 8);
  const size_t ne4 = (size_t)(E ? E : 1) * 4;
  const uint32_t ntiles = (uint32_t)(((uint64_t)E + kRadixTile - 1) / kRadixTile);
  const uint64_t nh = 256ull * std::max<uint32_t>(ntiles, 1);
  const uint32_t hscan = (uint32_t)((nh + 1 + kScanTile - 1) / kScanTile);
  void *tbuf, *hbuf;
  if ((rc = ws_buffer(w, kBufTranspose, 4 * ne4, &tbuf)) ||
      (rc = ws_buffer(w, kBufSeedList, (nh + 1 + hscan + 64) * 4, &hbuf)))
    return rc;
  lsg_graph t;
  memset(&t, 0, sizeof t);
  if (hipMalloc(&t.row_start, ((size_t)V + 1) * 4) != hipSuccess || hipMalloc(&t.edge_dst, ne4) != hipSuccess ||
      hipMalloc(&t.edge_data, ne4) != hipSuccess || hipMalloc(&t.node_data, (size_t)V * 4) != hipSuccess) {
    (void)hipFree(t.row_start);
    (void)hipFree(t.edge_dst);
    (void)hipFree(t.edge_data);
    (void)hipFree(t.node_data);
    return fail(LSG_ERR_NOMEM, "transpose: device allocation failed");
  }
  uint32_t *key[2] = {(uint32_t *)tbuf, (uint32_t *)((char *)tbuf + ne4)};
  uint32_t *vs[2] = {(uint32_t *)((char *)tbuf + 2 * ne4), t.edge_dst};
  uint32_t *vw[2] = {(uint32_t *)((char *)tbuf + 3 * ne4), (uint32_t *)t.edge_data};
  uint32_t *hist = (uint32_t *)hbuf, *htiles = hist + nh + 1, *htotal = htiles + hscan;
  const uint32_t sgrid = (uint32_t)std::min<uint64_t>(((uint64_t)E / kScatterRun + 256) / 256, 16ull * w->cus);
  const uint32_t first_out = P & 1;
  uint32_t *src0 = vs[first_out ^ 1];
  if (E) {
    hipLaunchKernelGGL(edge_source_kernel, dim3(sgrid), dim3(256), 0, w->stream, g->row_start, V, E, src0);
    const uint32_t *kin = g->edge_dst, *sin = src0, *win = (const uint32_t *)g->edge_data;
    // (an unweighted forward graph passes win = null: its reverse weights are 0)
    for (uint32_t p = 0; p < P; ++p) {
      const uint32_t o = (P - p) & 1, shift = 8 * p;
      hipLaunchKernelGGL(radix_hist_kernel, dim3(ntiles), dim3(kDoBlock), 0, w->stream, kin, E, shift, ntiles, hist);
      hipLaunchKernelGGL(scan_tile_kernel, dim3(hscan), dim3(kDoBlock), 0, w->stream, hist, (uint32_t)nh, hist,
                         htiles);
      hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kDoBlock), 0, w->stream, htiles, hscan, htotal);
      const uint32_t agrid = (uint32_t)std::min<uint64_t>((nh + 256) / 256, 16ull * w->cus);
      hipLaunchKernelGGL(scan_add_kernel, dim3(agrid), dim3(256), 0, w->stream, hist, (uint32_t)nh,
                         (const uint32_t *)htiles, (uint32_t *)nullptr, (const uint32_t *)htotal);
      hipLaunchKernelGGL(radix_scatter_kernel, dim3(ntiles), dim3(kDoBlock), 0, w->stream, kin, sin, win, E, shift,
                         ntiles, (const uint32_t *)hist, key[o], vs[o], vw[o]);
      kin = key[o];
      sin = vs[o];
      win = vw[o];
    }
  }
  const uint32_t rgrid = (uint32_t)std::min<uint64_t>(((uint64_t)E + 256) / 256, 16ull * w->cus);
  hipLaunchKernelGGL(row_bounds_kernel, dim3(rgrid), dim3(256), 0, w->stream, (const uint32_t *)key[1], E, V,
                     t.row_start);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(w->stream));
  t.nnodes = V;
  t.nedges = E;
  t.device_graph = 1;
  t.device = g->device;
  *gt = t;
  return LSG_OK;
}

int lsg_bfs_do(const lsg_graph *g, const lsg_graph *gt, uint32_t src, int direction, int32_t *level_dev,
               lsg_stats *stats) {
  clear_error();
  auto t0 = std::chrono::steady_clock::now();
  int rc = check_device_graph(g, "bfs_do");
  if (rc) return rc;
  if ((rc = check_device_graph(gt, "bfs_do (reverse graph)"))) return rc;
  if (gt->nnodes != g->nnodes || gt->nedges != g->nedges || gt->device != g->device)
    return fail(LSG_ERR_ARG, "bfs_do: the reverse graph does not match the graph (lsg_graph_transpose)");
  if (src >= g->nnodes) return fail(LSG_ERR_ARG, "source %u out of range [0, %u)", src, g->nnodes);
  if (direction < 0 || direction > 2) return fail(LSG_ERR_ARG, "bad direction %d", direction);
  if (!level_dev) return fail(LSG_ERR_ARG, "NULL level array");
  Workspace *w;
  if ((rc = get_workspace(g->device, &w))) return rc;
  std::lock_guard<std::mutex> lk(w->mu);
  HIPCHK(hipSetDevice(w->device));
  const uint32_t V = g->nnodes;
  const uint32_t words = (V + 31) / 32;
  // every vertex is found once, so all chunks over the whole traversal fit:
  // sum of ceil(deg / kTdChunk) over heavy vertices <= 2 E / kTdChunk
  const uint64_t hcap64 = 2ull * g->nedges / kTdChunk + 64;
  if (hcap64 > UINT32_MAX) return fail(LSG_ERR_ARG, "bfs_do: graph too large");
  const uint32_t hcap = (uint32_t)hcap64;
  constexpr int kBs = 256;
  const void *fn = (const void *)pdo_kernel<kBs>;
  int grid = 0;
  if ((rc = resident_grid(w, fn, kBs, &grid))) return rc;
  // 4 blocks per CU: fewer blocks per barrier and below the 6 per CU that are
  // co-resident (the occupancy figure says 7, so a full grid costs a failed
  // residency handshake on the first call). RMAT-20 from the hub: 0.142 ms
  // (1024 blocks) vs 0.156 ms (1536) vs 0.148 ms (512).
  grid = std::min<int>(grid, 4 * (int)w->cus);
  void *bbuf, *hbuf, *pbuf;
  if ((rc = ws_buffer(w, kBufQueues, 3ull * words * 4, &bbuf)) || (rc = ws_buffer(w, kBufHeavy, 2ull * hcap * 8, &hbuf)) ||
      (rc = ws_buffer(w, kBufSeedList, 2ull * kMaxGridBlocks * 4, &pbuf)))
    return rc;
  uint32_t rs2[2];
  HIPCHK(hipMemcpy(rs2, g->row_start + src, 8, hipMemcpyDeviceToHost));
  const uint32_t src_deg = rs2[1] - rs2[0];
  PdoArgs a;
  memset(&a, 0, sizeof a);
  a.row_start = g->row_start;
  a.edge_dst = g->edge_dst;
  a.in_row = gt->row_start;
  a.in_src = gt->edge_dst;
  a.level = (uint32_t *)level_dev;
  for (int i = 0; i < 3; ++i) a.fb[i] = (uint32_t *)bbuf + (size_t)i * words;
  a.hq[0] = (uint64_t *)hbuf;
  a.hq[1] = (uint64_t *)hbuf + hcap;
  a.partial = (uint32_t *)pbuf;
  a.ctrl = w->ctrl;
  a.nnodes = V;
  a.words = words;
  a.hcap = hcap;
  a.direction = (uint32_t)direction;
  a.spin_limit = spin_limit_ticks(w);
  a.residency_limit = residency_limit_ticks(w);
  // direction of level 1: the frontier is the source alone
  const uint32_t bu1 = direction == 2 || (direction == 0 && (uint64_t)src_deg * kAlpha > g->nedges);
  const uint32_t igrid = (uint32_t)std::min<uint64_t>(((uint64_t)V + 255) / 256, 8ull * w->cus);
  Ctrl hc;
  float ms = 0;
  for (int attempt = 0;; ++attempt) {
    HIPCHK(hipEventRecord(w->ev0, w->stream));
    hipLaunchKernelGGL(pdo_init_kernel, dim3(igrid), dim3(256), 0, w->stream, a, src, src_deg, bu1,
                       (uint64_t)g->nedges);
    hipLaunchKernelGGL(pdo_kernel<kBs>, dim3(grid), dim3(kBs), 0, w->stream, a);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(w->ev1, w->stream));
    HIPCHK(hipMemcpyAsync(&hc, w->ctrl, kCtrlHostBytes, hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    HIPCHK(hipEventElapsedTime(&ms, w->ev0, w->ev1));
    if (hc.abort != kAbortResidency) break;
    if (attempt > 0 || hc.resident_seen == 0 || hc.resident_seen >= (uint32_t)grid)
      return fail(LSG_ERR_HIP, "bfs_do: persistent grid of %d x %d blocks is not co-resident (%u were)", grid, kBs,
                  hc.resident_seen);
    grid = (int)hc.resident_seen;
    record_resident_grid(w, fn, kBs, grid);
  }
  if (hc.abort == kAbortTimeout)
    return fail(LSG_ERR_TIMEOUT, "bfs_do: grid barrier timed out at level %u; levels are incomplete", hc.s.round);
  if (hc.abort == kAbortCapacity) return fail(LSG_ERR_CAPACITY, "bfs_do: chunk queue capacity exceeded");
  if (stats) {
    memset(stats, 0, sizeof *stats);
    stats->ms_kernel = ms;
    stats->iterations = hc.s.round - 2;  // levels with a vertex (the last level finds nothing)
    stats->near_iters = hc.s.phase;      // bottom-up levels
    stats->grid_blocks = grid;
    stats->block_size = kBs;
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return LSG_OK;
}

}  // extern "C"
