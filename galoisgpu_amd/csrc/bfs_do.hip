// Direction-optimising BFS on gfx950, and the reverse CSR it needs.
//
// The levels are the reference's (apps/bfs/kernel.cu:20-32,312-334,380-393):
// level[src] = 0, first hop LEVEL 1, INF = INT_MAX for unreachable vertices.
// BFS levels are unique, so how a level is discovered does not change them.
// Each level is either
//   top-down:  the frontier queue's out-edges are flattened across the grid
//              (device prefix sum of degrees, one edge per thread) and an
//              unvisited destination is claimed with a compare-and-swap on
//              its level;
//   bottom-up: every unvisited vertex scans its in-edges and stops at the
//              first parent whose level is L-1.
// The host picks the direction per level from the frontier's size and edge
// count (Beamer, Asanovic and Patterson's rule: bottom-up once the frontier's
// edges exceed 1/alpha of the unexplored edges, top-down again once the
// frontier holds fewer than 1/beta of the vertices).  Every level marks the
// vertices it discovers in a bitmap; a top-down level also queues them, and
// a top-down level after a bottom-up one compacts that level's bitmap into
// its queue first.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>

#include "engine.h"
#include "lsg_internal.h"
#include "runtime.h"

namespace lsg {

constexpr int kDoBlock = 256;
constexpr uint32_t kScanTile = 1024;  // elements per block of the scan kernels
constexpr uint32_t kScatterRun = 16;  // edges per thread in the reverse-CSR scatter
constexpr uint32_t kAlpha = 14, kBeta = 24;

// ---- reverse CSR ------------------------------------------------------------

__global__ void fill_kernel(uint32_t *p, size_t n, uint32_t val) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = val;
}

__global__ void count_in_kernel(const uint32_t *edge_dst, uint32_t E, uint32_t *cnt) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < E; e += (size_t)gridDim.x * blockDim.x)
    atomic_add(cnt + edge_dst[e], 1u);
}

// Exclusive scan of one tile of kScanTile values per block (4 per thread):
// out[i] = sum of in[tile start .. i), tile_sum[b] = the tile's total.
__global__ __launch_bounds__(kDoBlock) void scan_tile_kernel(const uint32_t *in, uint32_t n, uint32_t *out,
                                                             uint32_t *tile_sum) {
  __shared__ uint32_t wsum[kDoBlock / kWave];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4;
  uint32_t x[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    x[k] = base + k < n ? in[base + k] : 0u;
    s += x[k];
  }
  // inclusive scan of the per-thread sums across the wavefront, then across waves
  uint32_t inc = s;
  const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t t = __shfl_up(inc, d);
    if (lane >= (uint32_t)d) inc += t;
  }
  if (lane == kWave - 1) wsum[w] = inc;
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t i = 0; i < w; ++i) before += wsum[i];
  uint32_t run = before + inc - s;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = run;
    run += x[k];
  }
  if (threadIdx.x == kDoBlock - 1) tile_sum[blockIdx.x] = before + inc;
}

// One block: exclusive scan of the tile totals in place; *total gets the sum.
__global__ __launch_bounds__(kDoBlock) void scan_sums_kernel(uint32_t *sums, uint32_t n, uint32_t *total) {
  __shared__ uint32_t wsum[kDoBlock / kWave];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
  for (uint32_t b0 = 0; b0 < n; b0 += kDoBlock) {
    const uint32_t i = b0 + threadIdx.x;
    const uint32_t s = i < n ? sums[i] : 0u;
    uint32_t inc = s;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const uint32_t t = __shfl_up(inc, d);
      if (lane >= (uint32_t)d) inc += t;
    }
    if (lane == kWave - 1) wsum[w] = inc;
    __syncthreads();
    uint32_t before = carry;
    for (uint32_t k = 0; k < w; ++k) before += wsum[k];
    if (i < n) sums[i] = before + inc - s;
    __syncthreads();
    if (threadIdx.x == kDoBlock - 1) carry = before + inc;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// Adds each tile's offset; writes the row starts and the scatter cursors.
__global__ void scan_add_kernel(uint32_t *row, uint32_t n, const uint32_t *tile_off, uint32_t *cursor,
                                const uint32_t *total) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i <= n; i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t r = i < n ? row[i] + tile_off[i / kScanTile] : *total;
    row[i] = r;
    if (cursor && i < n) cursor[i] = r;
  }
}

// Thread t moves edges [16 t, 16 t + 16) of the forward CSR: it finds their
// source once by binary search and walks forward.  Rows of the reverse CSR
// are unordered (positions come from atomics).
__global__ void scatter_in_kernel(const uint32_t *row_start, const uint32_t *edge_dst, const uint32_t *edge_data,
                                  uint32_t V, uint32_t E, uint32_t *cursor, uint32_t *in_src, uint32_t *in_w) {
  const size_t nthreads = (size_t)gridDim.x * blockDim.x;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t * kScatterRun < E; t += nthreads) {
    const uint32_t e0 = (uint32_t)(t * kScatterRun), e1 = min(E, e0 + kScatterRun);
    // last u with row_start[u] <= e0
    uint32_t lo = 0, hi = V;
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (row_start[mid] <= e0) lo = mid;
      else hi = mid;
    }
    uint32_t u = lo, next = row_start[u + 1];
    for (uint32_t e = e0; e < e1; ++e) {
      while (next <= e) next = row_start[++u + 1];
      const uint32_t v = edge_dst[e];
      const uint32_t pos = atomic_add(cursor + v, 1u);
      in_src[pos] = u;
      if (edge_data) in_w[pos] = edge_data[e];
    }
  }
}

// ---- the two BFS steps ---------------------------------------------------------

struct DoArgs {
  const uint32_t *row_start, *edge_dst;  // forward CSR
  const uint32_t *in_row, *in_src;       // reverse CSR
  uint32_t *level;
  const uint32_t *q_in;
  uint32_t *q_out;
  uint32_t *ctr;  // [2][2]: (pushed, their out-edge count) by level parity
  uint32_t *fb[3];  // frontier bitmaps: fb[L % 3] has bit v set iff level[v] == L
  uint32_t nnodes, n_in, L;
};

// Pushes the vertices with `found` (wave-aggregated), sets their bits in the
// level's frontier bitmap and adds their out-degree to the level's edge count.
__device__ __forceinline__ void do_push(const DoArgs &a, bool found, uint32_t v) {
  const uint64_t m = __ballot(found);
  if (m == 0) return;
  if (found) atomicOr(a.fb[a.L % 3] + (v >> 5), 1u << (v & 31));
  uint32_t *c = a.ctr + 2 * (a.L & 1);
  const uint32_t leader = (uint32_t)__builtin_ctzll(m);
  uint32_t deg = found ? a.row_start[v + 1] - a.row_start[v] : 0u;
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) deg += __shfl_xor(deg, d);
  uint32_t base = 0;
  if (lane_id() == leader) {
    base = atomic_add(c, (uint32_t)__builtin_popcountll(m));
    atomic_add(c + 1, deg);
  }
  base = readlane(base, leader);
  if (found) a.q_out[base + lanes_below(m)] = v;
}

// Run by every level's kernel: the next level's counters start at zero, and
// the bitmap the next level will fill (last used two levels ago) is cleared.
__device__ __forceinline__ void clear_next(const DoArgs &a) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    uint32_t *n = a.ctr + 2 * ((a.L + 1) & 1);
    n[0] = 0;
    n[1] = 0;
  }
  uint32_t *nb = a.fb[(a.L + 1) % 3];
  const uint32_t words = (a.nnodes + 31) / 32;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
    nb[i] = 0;
}

// Out-degrees of the frontier, the input of the top-down prefix sum.
__global__ void frontier_deg_kernel(const uint32_t *q, uint32_t n, const uint32_t *row_start, uint32_t *deg) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t u = q[i];
    deg[i] = row_start[u + 1] - row_start[u];
  }
}

// Top-down level L over the frontier's out-edges flattened across the whole
// grid: pre[i] is the first flattened edge of frontier vertex i (exclusive
// prefix sum of degrees, pre[n_in] = total), so one vertex's edges -- the
// source hub's at level 1 -- spread over every block.  Each thread finds
// the owner of its edge by binary search and claims an unvisited destination
// with a compare-and-swap on its level.
__global__ __launch_bounds__(kDoBlock) void top_down_kernel(DoArgs a, const uint32_t *pre) {
  clear_next(a);
  const uint32_t total = pre[a.n_in];
  const size_t nt = (size_t)gridDim.x * blockDim.x;
  // whole wavefronts iterate together so that the push's ballot sees them all
  const size_t fend = (total + (size_t)kWave - 1) / kWave * kWave;
  for (size_t f = blockIdx.x * (size_t)blockDim.x + threadIdx.x; f < fend; f += nt) {
    bool found = false;
    uint32_t v = 0;
    if (f < total) {
      uint32_t lo = 0, hi = a.n_in;  // last i with pre[i] <= f
      while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (pre[mid] <= f) lo = mid;
        else hi = mid;
      }
      v = a.edge_dst[a.row_start[a.q_in[lo]] + ((uint32_t)f - pre[lo])];
      if (a.level[v] == kInf) {
        uint32_t expect = kInf;
        found = __hip_atomic_compare_exchange_strong(a.level + v, &expect, a.L, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    do_push(a, found, v);
  }
}

// Bottom-up level L: every unvisited vertex (one per lane) looks for a parent
// at level L-1 among its in-edges and stops at the first.  A lane reads 4
// in-edges, then tests their sources' bits in the previous level's bitmap,
// so each step has 4 loads in flight.  Levels written earlier in this launch
// are L, never L-1, so racing reads cannot mislead the search.  (Measured on
// RMAT-20: 8 lanes per vertex with coalesced in-edge reads was 4x slower --
// each wavefront then searches 8 vertices instead of 64.)
//
// A wavefront owns 64 consecutive vertices (64-aligned), so it writes their
// two words of this level's bitmap with plain stores -- every word of the
// bitmap is written, and no atomics are needed.  The level's vertex count and
// out-edge count are summed per block into partial[2 b], partial[2 b + 1]
// (bu_reduce_kernel adds them up): one counter that every wavefront with a
// discovery bumps serialises the whole level (RMAT-20: 2 x 16K same-address
// atomics took ~240 us per level).  Nothing is queued: a top-down level that
// follows compacts the bitmap into its queue (bitmap_queue_kernel).
constexpr uint32_t kBuRun = 4;
__global__ __launch_bounds__(kDoBlock) void bottom_up_kernel(DoArgs a, uint32_t *partial) {
  __shared__ uint32_t s_cnt, s_deg;
  if (threadIdx.x == 0) s_cnt = s_deg = 0;
  clear_next(a);
  __syncthreads();
  const uint32_t *prev = a.fb[(a.L - 1) % 3];
  uint32_t *cur = a.fb[a.L % 3];
  const uint32_t words = (a.nnodes + 31) / 32;
  const uint32_t lane = lane_id();
  const size_t nt = (size_t)gridDim.x * blockDim.x;
  uint32_t cnt = 0, deg = 0;
  // whole wavefronts iterate together: lane 0's vertex is the first of the 64
  const size_t vend = (a.nnodes + (size_t)kWave - 1) / kWave * kWave;
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < vend; v += nt) {
    bool found = false;
    if (v < a.nnodes && a.level[v] == kInf) {
      const uint32_t e1 = a.in_row[v + 1];
      for (uint32_t e = a.in_row[v]; e < e1 && !found; e += kBuRun) {
        uint32_t u[kBuRun];
#pragma unroll
        for (uint32_t k = 0; k < kBuRun; ++k) u[k] = e + k < e1 ? a.in_src[e + k] : a.nnodes;
#pragma unroll
        for (uint32_t k = 0; k < kBuRun; ++k)
          found = found || (u[k] < a.nnodes && ((prev[u[k] >> 5] >> (u[k] & 31)) & 1u));
      }
      if (found) {
        a.level[v] = a.L;
        cnt += 1;
        deg += a.row_start[v + 1] - a.row_start[v];
      }
    }
    const uint64_t m = __ballot(found);
    if (lane == 0) {
      const uint32_t w0 = (uint32_t)(v / 32);  // v = the wave's first vertex, < nnodes
      cur[w0] = (uint32_t)m;
      if (w0 + 1 < words) cur[w0 + 1] = (uint32_t)(m >> 32);
    }
  }
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) {
    cnt += __shfl_xor(cnt, d);
    deg += __shfl_xor(deg, d);
  }
  if (lane == 0 && cnt) {
    atomicAdd(&s_cnt, cnt);
    atomicAdd(&s_deg, deg);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = s_cnt;
    partial[2 * blockIdx.x + 1] = s_deg;
  }
}

// One block: out[0] = sum of partial[2 b], out[1] = sum of partial[2 b + 1]
// over n blocks.
__global__ __launch_bounds__(kDoBlock) void bu_reduce_kernel(const uint32_t *partial, uint32_t n, uint32_t *out) {
  __shared__ uint32_t s[2];
  if (threadIdx.x < 2) s[threadIdx.x] = 0;
  __syncthreads();
  uint32_t c = 0, d = 0;
  for (uint32_t b = threadIdx.x; b < n; b += blockDim.x) {
    c += partial[2 * b];
    d += partial[2 * b + 1];
  }
#pragma unroll
  for (int k = kWave / 2; k >= 1; k >>= 1) {
    c += __shfl_xor(c, k);
    d += __shfl_xor(d, k);
  }
  if (lane_id() == 0) {
    atomicAdd(&s[0], c);
    atomicAdd(&s[1], d);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = s[0];
    out[1] = s[1];
  }
}

// Queue of the vertices whose bit is set in bitmap fb (words 0 .. words-1):
// lane i takes word i of a 64-word run, and one atomic per wavefront reserves
// the run's places in q.
__global__ void bitmap_queue_kernel(const uint32_t *fb, uint32_t words, uint32_t *q, uint32_t *tail) {
  const size_t nt = (size_t)gridDim.x * blockDim.x;
  const size_t iend = ((size_t)words + kWave - 1) / kWave * kWave;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < iend; i += nt) {
    uint32_t bits = i < words ? fb[i] : 0u;
    const uint32_t n = (uint32_t)__builtin_popcount(bits);
    uint32_t pos = wave_reserve(n, tail);
    while (bits) {
      const uint32_t b = (uint32_t)__builtin_ctz(bits);
      q[pos++] = (uint32_t)i * 32 + b;
      bits &= bits - 1;
    }
  }
}

__global__ void zero_words_kernel(uint32_t *p, uint32_t n) {
  if (threadIdx.x < n) p[threadIdx.x] = 0;
}

// level = INF except src = 0; bitmaps 0 and 1 cleared (2 is cleared by
// level 1's kernel) and src's bit set in bitmap 0; queue and counters seeded.
__global__ void do_init_kernel(uint32_t *level, uint32_t V, uint32_t src, uint32_t *q, uint32_t *ctr, uint32_t *fb0,
                               uint32_t *fb1) {
  const size_t nt = (size_t)gridDim.x * blockDim.x, t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  for (size_t v = t; v < V; v += nt) level[v] = v == src ? 0u : kInf;
  for (size_t i = t; i < (V + 31) / 32; i += nt) {
    fb0[i] = i == src / 32 ? 1u << (src & 31) : 0u;
    fb1[i] = 0;
  }
  if (t == 0) {
    q[0] = src;
    ctr[0] = ctr[1] = ctr[2] = ctr[3] = 0;
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
  lsg_graph t;
  memset(&t, 0, sizeof t);
  const size_t ne = E ? E : 1;
  if (hipMalloc(&t.row_start, ((size_t)V + 1) * 4) != hipSuccess || hipMalloc(&t.edge_dst, ne * 4) != hipSuccess ||
      hipMalloc(&t.edge_data, ne * 4) != hipSuccess || hipMalloc(&t.node_data, (size_t)V * 4) != hipSuccess) {
    (void)hipFree(t.row_start);
    (void)hipFree(t.edge_dst);
    (void)hipFree(t.edge_data);
    (void)hipFree(t.node_data);
    return fail(LSG_ERR_NOMEM, "transpose: device allocation failed");
  }
  const uint32_t ntiles = (V + kScanTile - 1) / kScanTile;
  void *buf;
  if ((rc = ws_buffer(w, kBufSeedList, ((size_t)V + ntiles + 64) * 4, &buf))) return rc;
  uint32_t *cnt = (uint32_t *)buf, *tiles = cnt + V, *total = tiles + ntiles;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(((uint64_t)std::max(V, E) + 255) / 256, 16ull * w->cus);
  // cnt = 0 (a fill kernel rather than a memset, so the sequence stays capturable)
  hipLaunchKernelGGL(fill_kernel, dim3(grid), dim3(256), 0, w->stream, cnt, (size_t)V, 0u);
  hipLaunchKernelGGL(count_in_kernel, dim3(grid), dim3(256), 0, w->stream, g->edge_dst, E, cnt);
  hipLaunchKernelGGL(scan_tile_kernel, dim3(ntiles), dim3(kDoBlock), 0, w->stream, cnt, V, t.row_start, tiles);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kDoBlock), 0, w->stream, tiles, ntiles, total);
  hipLaunchKernelGGL(scan_add_kernel, dim3(grid), dim3(256), 0, w->stream, t.row_start, V, tiles, cnt, total);
  const uint32_t sgrid = (uint32_t)std::min<uint64_t>(((uint64_t)E / kScatterRun + 256) / 256, 16ull * w->cus);
  hipLaunchKernelGGL(scatter_in_kernel, dim3(sgrid), dim3(256), 0, w->stream, g->row_start, g->edge_dst,
                     (const uint32_t *)g->edge_data, V, E, cnt, t.edge_dst, (uint32_t *)t.edge_data);
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
  void *qbuf, *cbuf, *pbuf;
  const uint32_t ptiles = (V + kScanTile - 1) / kScanTile;
  const size_t words = ((size_t)V + 31) / 32;
  const uint32_t vgrid = (uint32_t)std::min<uint64_t>(((uint64_t)V + 255) / 256, 16ull * w->cus);
  if ((rc = ws_buffer(w, kBufQueues, 2ull * V * 4 + 3 * words * 4, &qbuf)) || (rc = ws_buffer(w, kBufSeeds, 2048, &cbuf)) ||
      (rc = ws_buffer(w, kBufSeedList, ((size_t)V + ptiles + 64 + 2ull * vgrid) * 4, &pbuf)))
    return rc;
  uint32_t *q[2] = {(uint32_t *)qbuf, (uint32_t *)qbuf + V};
  uint32_t *fbits = (uint32_t *)qbuf + 2ull * V;
  uint32_t *ctr = (uint32_t *)cbuf;  // [0..3] level counts by parity, [4] queue tail of a compaction
  uint32_t *pre = (uint32_t *)pbuf, *ptile = pre + V + 1, *ptotal = ptile + ptiles;
  uint32_t *partial = pre + V + ptiles + 64;  // per-block counts of a bottom-up level
  uint32_t rs2[2];
  HIPCHK(hipMemcpy(rs2, g->row_start + src, 8, hipMemcpyDeviceToHost));
  HIPCHK(hipEventRecord(w->ev0, w->stream));
  hipLaunchKernelGGL(do_init_kernel, dim3(vgrid), dim3(256), 0, w->stream, (uint32_t *)level_dev, V, src, q[0], ctr,
                     fbits, fbits + words);
  DoArgs a;
  a.row_start = g->row_start;
  a.edge_dst = g->edge_dst;
  a.in_row = gt->row_start;
  a.in_src = gt->edge_dst;
  a.level = (uint32_t *)level_dev;
  a.ctr = ctr;
  for (int i = 0; i < 3; ++i) a.fb[i] = fbits + i * words;
  a.nnodes = V;
  uint64_t m_u = g->nedges, m_f = rs2[1] - rs2[0];
  uint32_t n_in = 1, L = 1, bu_levels = 0;
  bool bottom_up = direction == 2, prev_bottom_up = false;
  while (n_in > 0) {
    if (direction == 0) {
      // Beamer's rule: go bottom-up when the frontier's edges exceed the
      // unexplored edges / alpha; return once the frontier is small again
      if (!bottom_up && m_f * kAlpha > m_u) bottom_up = true;
      else if (bottom_up && (uint64_t)n_in * kBeta < V) bottom_up = false;
    }
    a.q_in = q[(L - 1) & 1];
    a.q_out = q[L & 1];
    a.n_in = n_in;
    a.L = L;
    if (bottom_up) {
      hipLaunchKernelGGL(bottom_up_kernel, dim3(vgrid), dim3(kDoBlock), 0, w->stream, a, partial);
      hipLaunchKernelGGL(bu_reduce_kernel, dim3(1), dim3(kDoBlock), 0, w->stream, (const uint32_t *)partial, vgrid,
                         ctr + 2 * (L & 1));
      ++bu_levels;
    } else {
      if (prev_bottom_up) {
        // a bottom-up level leaves its vertices in its bitmap only
        const uint32_t wgrid = (uint32_t)std::min<uint64_t>((words + 255) / 256, 16ull * w->cus);
        hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, w->stream, ctr + 4, 1u);
        hipLaunchKernelGGL(bitmap_queue_kernel, dim3(wgrid), dim3(256), 0, w->stream,
                           (const uint32_t *)a.fb[(L - 1) % 3], (uint32_t)words, q[(L - 1) & 1], ctr + 4);
      }
      const uint32_t nt = (n_in + kScanTile - 1) / kScanTile;
      const uint32_t g1 = (uint32_t)std::min<uint64_t>(((uint64_t)n_in + 255) / 256, 16ull * w->cus);
      hipLaunchKernelGGL(frontier_deg_kernel, dim3(g1), dim3(256), 0, w->stream, a.q_in, n_in, a.row_start, pre);
      hipLaunchKernelGGL(scan_tile_kernel, dim3(nt), dim3(kDoBlock), 0, w->stream, pre, n_in, pre, ptile);
      hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kDoBlock), 0, w->stream, ptile, nt, ptotal);
      hipLaunchKernelGGL(scan_add_kernel, dim3(g1), dim3(256), 0, w->stream, pre, n_in, ptile, (uint32_t *)nullptr,
                         ptotal);
      const uint32_t g2 = (uint32_t)std::min<uint64_t>((m_f + kDoBlock - 1) / kDoBlock + 1, 16ull * w->cus);
      hipLaunchKernelGGL(top_down_kernel, dim3(g2), dim3(kDoBlock), 0, w->stream, a, (const uint32_t *)pre);
    }
    HIPCHK(hipGetLastError());
    uint32_t *c = w->pinned;  // pinned: the per-level readback is one short DMA
    HIPCHK(hipMemcpyAsync(c, ctr + 2 * (L & 1), 8, hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    prev_bottom_up = bottom_up;
    m_u = m_u > m_f ? m_u - m_f : 0;
    n_in = c[0];
    m_f = c[1];
    ++L;
  }
  HIPCHK(hipEventRecord(w->ev1, w->stream));
  HIPCHK(hipEventSynchronize(w->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, w->ev0, w->ev1));
  if (stats) {
    memset(stats, 0, sizeof *stats);
    stats->ms_kernel = ms;
    stats->iterations = L - 2;  // levels with a vertex (the last launch finds nothing)
    stats->near_iters = bu_levels;
    stats->grid_blocks = (int32_t)vgrid;
    stats->block_size = kDoBlock;
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return LSG_OK;
}

}  // extern "C"
