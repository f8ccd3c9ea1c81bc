// Reverse CSR (in-edges) of a device graph, for pull rounds.
//
// No reference counterpart: the reference only pushes along out-edges.  The
// reverse graph is built on the device the first time a pull-enabled call
// sees a graph (count in-degrees, prefix-sum them, scatter), together with the
// list of pull work items (vertex, chunk of kPullChunk in-edges), and kept
// until lsg_graph_dealloc releases the graph.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "reverse.h"

namespace lsg {

__global__ void rev_count_kernel(const uint32_t *dst, uint64_t E, uint32_t *cnt) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x, st = (size_t)gridDim.x * blockDim.x;
  for (; i < E; i += st) atomicAdd(cnt + dst[i] + 1, 1u);
}

constexpr int kScanThreads = 256, kScanPer = 4, kScanTile = kScanThreads * kScanPer;

// Inclusive scan of x[0..n) within tiles of kScanTile; tile totals to sums.
__global__ void scan_tiles_kernel(uint32_t *x, uint64_t n, uint32_t *sums) {
  __shared__ uint32_t part[kScanThreads];
  uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanPer;
  uint32_t v[kScanPer], s = 0;
  for (int k = 0; k < kScanPer; ++k) {
    v[k] = base + k < n ? x[base + k] : 0u;
    s += v[k];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    uint32_t y = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += y;
    __syncthreads();
  }
  uint32_t run = part[threadIdx.x] - s;
  for (int k = 0; k < kScanPer; ++k) {
    run += v[k];
    if (base + k < n) x[base + k] = run;
  }
  if (threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = part[threadIdx.x];
}

// Exclusive scan of the tile totals, one block looping over them.
__global__ void scan_sums_kernel(uint32_t *sums, uint32_t ntiles) {
  __shared__ uint32_t part[kScanThreads];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t b = 0; b < ntiles; b += kScanThreads) {
    uint32_t i = b + threadIdx.x;
    uint32_t s = i < ntiles ? sums[i] : 0u;
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
      uint32_t y = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
      __syncthreads();
      part[threadIdx.x] += y;
      __syncthreads();
    }
    if (i < ntiles) sums[i] = carry + part[threadIdx.x] - s;
    __syncthreads();
    if (threadIdx.x == kScanThreads - 1) carry += part[threadIdx.x];
    __syncthreads();
  }
}

__global__ void add_tile_offsets_kernel(uint32_t *x, uint64_t n, const uint32_t *sums) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) x[i] += sums[i / kScanTile];
}

__device__ __forceinline__ uint32_t row_of_edge(const uint32_t *row, uint32_t V, uint32_t e) {
  uint32_t lo = 0, hi = V;  // largest u with row[u] <= e
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (row[mid] <= e) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ void rev_scatter_kernel(const uint32_t *row, const uint32_t *dst, const uint32_t *w, uint32_t V,
                                   uint64_t E, uint32_t *cursor, uint32_t *rsrc, uint32_t *rw) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x, st = (size_t)gridDim.x * blockDim.x;
  for (; i < E; i += st) {
    uint32_t u = row_of_edge(row, V, (uint32_t)i);
    uint32_t pos = atomicAdd(cursor + dst[i], 1u);
    rsrc[pos] = u;
    rw[pos] = w[i];
  }
}

// Pull items: ceil(indeg / kPullChunk) entries (chunk << 32 | v) per vertex.
__global__ void pull_items_kernel(const uint32_t *rrow, uint32_t V, const uint32_t *item_off, uint64_t *items) {
  size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x, st = (size_t)gridDim.x * blockDim.x;
  for (; v < V; v += st) {
    uint32_t deg = rrow[v + 1] - rrow[v];
    uint32_t n = (deg + kPullChunk - 1) / kPullChunk, o = item_off[v];
    for (uint32_t k = 0; k < n; ++k) items[o + k] = ((uint64_t)k << 32) | v;
  }
}

__global__ void pull_count_kernel(const uint32_t *rrow, uint32_t V, uint32_t *cnt) {
  size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x, st = (size_t)gridDim.x * blockDim.x;
  for (; v < V; v += st) cnt[v + 1] = (rrow[v + 1] - rrow[v] + kPullChunk - 1) / kPullChunk;
}

struct RevEntry {
  const void *row;
  uint32_t nnodes;
  uint64_t nedges;
  int device;
  ReverseCsr r;
  void *mem;
};

static std::mutex g_rev_mu;
static std::vector<RevEntry> g_rev;

// In-place exclusive->inclusive prefix sum helper: x[0] = 0, x[1..n] counts.
static int scan_counts(Workspace *w, uint32_t *x, uint64_t n, uint32_t *sums) {
  uint32_t ntiles = (uint32_t)((n + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(scan_tiles_kernel, dim3(ntiles), dim3(kScanThreads), 0, w->stream, x, n, sums);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kScanThreads), 0, w->stream, sums, ntiles);
  uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 16ull * w->cus);
  hipLaunchKernelGGL(add_tile_offsets_kernel, dim3(blocks), dim3(256), 0, w->stream, x, n, sums);
  HIPCHK(hipGetLastError());
  return LSG_OK;
}

int get_reverse(Workspace *w, const lsg_graph *g, ReverseCsr *out) {
  std::lock_guard<std::mutex> lk(g_rev_mu);
  for (auto &e : g_rev)
    if (e.row == g->row_start && e.nnodes == g->nnodes && e.nedges == g->nedges && e.device == g->device) {
      *out = e.r;
      return LSG_OK;
    }
  const uint32_t V = g->nnodes;
  const uint64_t E = g->nedges;
  // layout: rrow[V+1] | rsrc[E] | rw[E] | item_off[V+1] | sums | cursor[V] | items[nitems]
  uint64_t ntiles = (V + 1 + kScanTile - 1) / kScanTile + 1;
  size_t head = (size_t)(V + 1) * 4 + E * 8 + (size_t)(V + 1) * 4 + ntiles * 4 + (size_t)V * 4;
  size_t items_cap = (size_t)V + E / kPullChunk + 1;
  size_t bytes = ((head + 7) & ~(size_t)7) + items_cap * 8;
  void *mem = nullptr;
  if (hipMalloc(&mem, bytes) != hipSuccess) return fail(LSG_ERR_NOMEM, "reverse graph: %zu bytes unavailable", bytes);
  uint32_t *rrow = (uint32_t *)mem, *rsrc = rrow + V + 1, *rw = rsrc + E, *ioff = rw + E, *sums = ioff + V + 1,
           *cursor = sums + ntiles;
  uint64_t *items = (uint64_t *)(((uintptr_t)(cursor + V) + 7) & ~(uintptr_t)7);
  uint32_t blocks = (uint32_t)std::min<uint64_t>((E + 255) / 256 + 1, 16ull * w->cus);
  int rc;
  HIPCHK(hipMemsetAsync(rrow, 0, (size_t)(V + 1) * 4, w->stream));
  hipLaunchKernelGGL(rev_count_kernel, dim3(blocks), dim3(256), 0, w->stream, g->edge_dst, E, rrow);
  if ((rc = scan_counts(w, rrow, (uint64_t)V + 1, sums))) return rc;
  HIPCHK(hipMemcpyAsync(cursor, rrow, (size_t)V * 4, hipMemcpyDeviceToDevice, w->stream));
  hipLaunchKernelGGL(rev_scatter_kernel, dim3(blocks), dim3(256), 0, w->stream, g->row_start, g->edge_dst,
                     (const uint32_t *)g->edge_data, V, E, cursor, rsrc, rw);
  HIPCHK(hipMemsetAsync(ioff, 0, (size_t)(V + 1) * 4, w->stream));
  uint32_t vblocks = (uint32_t)std::min<uint64_t>((V + 255) / 256 + 1, 16ull * w->cus);
  hipLaunchKernelGGL(pull_count_kernel, dim3(vblocks), dim3(256), 0, w->stream, rrow, V, ioff);
  if ((rc = scan_counts(w, ioff, (uint64_t)V + 1, sums))) return rc;
  hipLaunchKernelGGL(pull_items_kernel, dim3(vblocks), dim3(256), 0, w->stream, rrow, V, ioff, items);
  HIPCHK(hipGetLastError());
  uint32_t nitems = 0;
  HIPCHK(hipMemcpyAsync(&nitems, ioff + V, 4, hipMemcpyDeviceToHost, w->stream));
  HIPCHK(hipStreamSynchronize(w->stream));
  ReverseCsr r{rrow, rsrc, rw, items, nitems};
  g_rev.push_back({g->row_start, V, E, g->device, r, mem});
  *out = r;
  return LSG_OK;
}

void drop_reverse(const lsg_graph *g) {
  std::lock_guard<std::mutex> lk(g_rev_mu);
  for (size_t i = 0; i < g_rev.size(); ++i)
    if (g_rev[i].row == g->row_start && g_rev[i].device == g->device) {
      (void)hipSetDevice(g_rev[i].device);
      (void)hipFree(g_rev[i].mem);
      g_rev.erase(g_rev.begin() + i);
      return;
    }
}

}  // namespace lsg
