// Host side of the CSR graph: .gr v1 reader and writer, host allocation,
// validation, the apps' output() text, and the seeded synthetic generators.
//
// Reference: rt/src/csr_graph.cu (loader, alloc) and apps/sssp/support.cu,
// apps/bfs/support.cu (output).

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "lsg_internal.h"

namespace lsg {

static thread_local std::string t_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  t_err = buf;
  return code;
}

void clear_error() { t_err.clear(); }

unsigned host_threads() {
  if (const char *s = getenv("LSG_THREADS")) {
    int v = atoi(s);
    if (v > 0) return (unsigned)v;
  }
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  return n > 16 ? 16 : n;
}

int validate_host_csr(const lsg_graph *g) {
  if (!g || g->device_graph) return fail(LSG_ERR_ARG, "validate: expected a host graph");
  if (g->nnodes == 0) return fail(LSG_ERR_FORMAT, "graph has no nodes");
  if (g->nnodes == UINT32_MAX) return fail(LSG_ERR_FORMAT, "nnodes must be below 2^32-1");
  if (!g->row_start || (g->nedges && !g->edge_dst))
    return fail(LSG_ERR_ARG, "graph arrays are NULL");
  if (g->row_start[0] != 0) return fail(LSG_ERR_FORMAT, "row_start[0] = %u, expected 0", g->row_start[0]);
  if (g->row_start[g->nnodes] != g->nedges)
    return fail(LSG_ERR_FORMAT, "row_start[nnodes] = %u but nedges = %u", g->row_start[g->nnodes], g->nedges);
  std::atomic<int> bad{0};
  std::atomic<uint64_t> where{0};
  parallel_for(g->nnodes, [&](size_t lo, size_t hi, unsigned) {
    for (size_t i = lo; i < hi && !bad.load(std::memory_order_relaxed); ++i) {
      if (g->row_start[i + 1] < g->row_start[i]) {
        bad = 1;
        where = i;
      }
    }
  });
  if (bad) return fail(LSG_ERR_FORMAT, "row offsets decrease at node %llu", (unsigned long long)where.load());
  parallel_for(g->nedges, [&](size_t lo, size_t hi, unsigned) {
    for (size_t e = lo; e < hi && !bad.load(std::memory_order_relaxed); ++e) {
      if (g->edge_dst[e] >= g->nnodes) {
        bad = 2;
        where = e;
      } else if (g->edge_data && g->edge_data[e] < 0) {
        bad = 3;
        where = e;
      }
    }
  });
  if (bad == 2)
    return fail(LSG_ERR_FORMAT, "edge %llu has destination %u >= nnodes %u", (unsigned long long)where.load(),
                g->edge_dst[where.load()], g->nnodes);
  if (bad == 3)
    return fail(LSG_ERR_FORMAT, "edge %llu has weight %u > INT_MAX", (unsigned long long)where.load(),
                (unsigned)g->edge_data[where.load()]);
  return LSG_OK;
}

// Sorts each row's destinations, drops self loops and duplicates, compacts.
int build_csr(lsg_graph *g, uint32_t nnodes, std::vector<uint32_t> &src, std::vector<uint32_t> &dst,
              bool dedupe, uint32_t wlo, uint32_t whi, uint64_t wseed) {
  const size_t m = src.size();
  if (nnodes == 0) return fail(LSG_ERR_ARG, "generator: zero nodes");
  std::vector<std::atomic<uint32_t>> deg(nnodes);
  parallel_for(nnodes, [&](size_t lo, size_t hi, unsigned) {
    for (size_t i = lo; i < hi; ++i) deg[i].store(0, std::memory_order_relaxed);
  });
  parallel_for(m, [&](size_t lo, size_t hi, unsigned) {
    for (size_t i = lo; i < hi; ++i) deg[src[i]].fetch_add(1, std::memory_order_relaxed);
  });
  std::vector<uint64_t> off(nnodes + 1);
  off[0] = 0;
  for (uint32_t v = 0; v < nnodes; ++v) off[v + 1] = off[v] + deg[v].load(std::memory_order_relaxed);
  std::vector<uint32_t> bucket(m);
  {
    std::vector<std::atomic<uint64_t>> cur(nnodes);
    parallel_for(nnodes, [&](size_t lo, size_t hi, unsigned) {
      for (size_t i = lo; i < hi; ++i) cur[i].store(off[i], std::memory_order_relaxed);
    });
    parallel_for(m, [&](size_t lo, size_t hi, unsigned) {
      for (size_t i = lo; i < hi; ++i) bucket[cur[src[i]].fetch_add(1, std::memory_order_relaxed)] = dst[i];
    });
  }
  std::vector<uint32_t>().swap(src);
  std::vector<uint32_t>().swap(dst);
  std::vector<uint32_t> ndeg(nnodes);
  parallel_for(nnodes, [&](size_t lo, size_t hi, unsigned) {
    for (size_t v = lo; v < hi; ++v) {
      uint32_t *b = bucket.data() + off[v], *e = bucket.data() + off[v + 1];
      std::sort(b, e);
      if (dedupe) {
        uint32_t *w = b;
        for (uint32_t *p = b; p < e; ++p) {
          if (*p == v) continue;
          if (w > b && w[-1] == *p) continue;
          *w++ = *p;
        }
        ndeg[v] = (uint32_t)(w - b);
      } else {
        ndeg[v] = (uint32_t)(e - b);
      }
    }
  }, 1 << 12);
  uint64_t total = 0;
  for (uint32_t v = 0; v < nnodes; ++v) total += ndeg[v];
  if (total >= UINT32_MAX) return fail(LSG_ERR_ARG, "generator: %llu edges exceed the u32 index range",
                                       (unsigned long long)total);
  int rc = lsg_graph_alloc_host(g, nnodes, (uint32_t)total);
  if (rc) return rc;
  g->row_start[0] = 0;
  for (uint32_t v = 0; v < nnodes; ++v) g->row_start[v + 1] = g->row_start[v] + ndeg[v];
  parallel_for(nnodes, [&](size_t lo, size_t hi, unsigned) {
    for (size_t v = lo; v < hi; ++v) {
      uint32_t o = g->row_start[v];
      const uint32_t *b = bucket.data() + off[v];
      for (uint32_t j = 0; j < ndeg[v]; ++j) {
        g->edge_dst[o + j] = b[j];
        g->edge_data[o + j] = (int32_t)edge_weight(wseed, (uint32_t)v, b[j], wlo, whi);
      }
    }
  }, 1 << 12);
  return LSG_OK;
}

// Formats one output line ("%d %d\n" or "%d INF\n") into p; returns the end.
static inline char *fmt_line(char *p, uint32_t i, int32_t d) {
  char tmp[16];
  int n = 0;
  uint32_t x = i;
  do {
    tmp[n++] = char('0' + x % 10);
    x /= 10;
  } while (x);
  while (n) *p++ = tmp[--n];
  *p++ = ' ';
  if (d == INT_MAX) {
    memcpy(p, "INF\n", 4);
    return p + 4;
  }
  uint32_t y;
  if (d < 0) {
    *p++ = '-';
    y = (uint32_t)(-(int64_t)d);
  } else {
    y = (uint32_t)d;
  }
  do {
    tmp[n++] = char('0' + y % 10);
    y /= 10;
  } while (y);
  while (n) *p++ = tmp[--n];
  *p++ = '\n';
  return p;
}

}  // namespace lsg

using namespace lsg;

extern "C" {

int lsg_abi_version(void) { return LSG_ABI_VERSION; }

const char *lsg_last_error(void) { return t_err.c_str(); }

void lsg_opts_default(lsg_opts *o) {
  if (!o) return;
  memset(o, 0, sizeof *o);
  o->delta = 10000;  // apps/sssp/support.cu:10
  o->start_node = 0; // apps/sssp/support.cu:12
  o->enable_lb = 0;
  // Tile-flattened expansion, like the reference's default fine-grained
  // scheduler (apps/sssp/kernel.cu:266-302); 10-25% faster than degree-binned
  // on every benchmark graph.
  o->expand_mode = LSG_EXPAND_LB;
  o->far_mode = LSG_FAR_REFERENCE;
}

int lsg_graph_alloc_host(lsg_graph *g, uint32_t nnodes, uint32_t nedges) {
  if (!g) return fail(LSG_ERR_ARG, "alloc_host: NULL graph");
  if (nnodes == 0) return fail(LSG_ERR_ARG, "alloc_host: nnodes must be > 0");  // csr_graph.cu:29
  memset(g, 0, sizeof *g);
  g->nnodes = nnodes;
  g->nedges = nedges;
  g->row_start = (uint32_t *)calloc((size_t)nnodes + 1, sizeof(uint32_t));
  g->edge_dst = (uint32_t *)calloc(nedges ? nedges : 1, sizeof(uint32_t));
  g->edge_data = (int32_t *)calloc(nedges ? nedges : 1, sizeof(int32_t));
  g->node_data = (int32_t *)calloc(nnodes, sizeof(int32_t));
  if (!g->row_start || !g->edge_dst || !g->edge_data || !g->node_data) {
    free(g->row_start);
    free(g->edge_dst);
    free(g->edge_data);
    free(g->node_data);
    memset(g, 0, sizeof *g);
    return fail(LSG_ERR_NOMEM, "alloc_host: out of host memory for %u nodes / %u edges", nnodes, nedges);
  }
  return LSG_OK;
}

int lsg_graph_read(lsg_graph *g, const char *path, int quiet) {
  if (!g || !path) return fail(LSG_ERR_ARG, "graph_read: NULL argument");
  memset(g, 0, sizeof *g);
  int fd = open(path, O_RDONLY);
  if (fd == -1) return fail(LSG_ERR_IO, "unable to open %s: %s", path, strerror(errno));
  struct stat st;
  if (fstat(fd, &st) == -1) {
    close(fd);
    return fail(LSG_ERR_IO, "unable to stat %s", path);
  }
  size_t len = (size_t)st.st_size;
  if (len < 32) {
    close(fd);
    return fail(LSG_ERR_FORMAT, "%s: %zu bytes is too short for a .gr header", path, len);
  }
  void *m = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (m == MAP_FAILED) return fail(LSG_ERR_IO, "mmap of %s failed", path);
  auto t0 = std::chrono::steady_clock::now();
  const uint64_t *h = (const uint64_t *)m;
  // Header: version, sizeEdgeTy, numNodes, numEdges (rt/src/csr_graph.cu:187-192).
  uint64_t version = h[0], size_edge = h[1], nn = h[2], ne = h[3];
  int rc = LSG_OK;
  if (version != 1) rc = fail(LSG_ERR_FORMAT, "%s: .gr version %llu, expected 1", path, (unsigned long long)version);
  else if (size_edge != 0 && size_edge != 4)
    rc = fail(LSG_ERR_FORMAT, "%s: sizeEdgeTy %llu unsupported (0 or 4)", path, (unsigned long long)size_edge);
  else if (nn == 0 || nn >= UINT32_MAX || ne >= UINT32_MAX)
    rc = fail(LSG_ERR_FORMAT, "%s: %llu nodes / %llu edges outside the u32 index range", path,
              (unsigned long long)nn, (unsigned long long)ne);
  uint64_t need = 32 + 8 * nn + 4 * ne + ((ne & 1) ? 4 : 0) + (size_edge ? 4 * ne : 0);
  if (rc == LSG_OK && len < need)
    rc = fail(LSG_ERR_FORMAT, "%s: truncated (%zu bytes, layout needs %llu)", path, len, (unsigned long long)need);
  if (rc) {
    munmap(m, len);
    return rc;
  }
  const uint64_t *outIdx = h + 4;
  const uint32_t *outs = (const uint32_t *)(outIdx + nn);
  const uint32_t *edata = outs + ne + (ne & 1);  // padding word when numEdges is odd (:198)
  if (!quiet) printf("nnodes=%d, nedges=%d.\n", (int)nn, (int)ne);
  rc = lsg_graph_alloc_host(g, (uint32_t)nn, (uint32_t)ne);
  if (rc) {
    munmap(m, len);
    return rc;
  }
  if (!quiet) {
    size_t mem = ((nn + 1) + ne) * 4 + ne * 4 + nn * 4;
    printf("Host memory for graph: %3u MB\n", (unsigned)(mem / 1048756));  // csr_graph.cu:43 divisor kept
  }
  uint64_t prev = 0;
  for (uint64_t i = 0; i < nn; ++i) {
    uint64_t o = outIdx[i];
    if (o < prev || o > ne) {
      rc = fail(LSG_ERR_FORMAT, "%s: out-index of node %llu is %llu (previous %llu, edges %llu)", path,
                (unsigned long long)i, (unsigned long long)o, (unsigned long long)prev, (unsigned long long)ne);
      break;
    }
    g->row_start[i + 1] = (uint32_t)o;
    prev = o;
  }
  if (rc == LSG_OK && prev != ne)
    rc = fail(LSG_ERR_FORMAT, "%s: last out-index %llu != numEdges %llu", path, (unsigned long long)prev,
              (unsigned long long)ne);
  if (rc == LSG_OK) {
    memcpy(g->edge_dst, outs, ne * 4);
    if (size_edge) memcpy(g->edge_data, edata, ne * 4);  // else weights stay 0 (calloc, :37)
    rc = validate_host_csr(g);
  }
  munmap(m, len);
  if (rc) {
    lsg_graph_dealloc(g);
    return rc;
  }
  auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
  if (!quiet)
    printf("read %lld bytes in %d ms (%0.2f MB/s)\n\r\n", (long long)len, (int)ms,
           ms ? (len / 1000.0) / (double)ms : 0.0);
  return LSG_OK;
}

int lsg_graph_write(const lsg_graph *g, const char *path, int with_weights) {
  if (!g || !path || g->device_graph) return fail(LSG_ERR_ARG, "graph_write: need a host graph and a path");
  FILE *f = fopen(path, "wb");
  if (!f) return fail(LSG_ERR_IO, "unable to create %s: %s", path, strerror(errno));
  uint64_t hdr[4] = {1, with_weights ? 4u : 0u, g->nnodes, g->nedges};
  bool ok = fwrite(hdr, 8, 4, f) == 4;
  std::vector<uint64_t> idx(g->nnodes);
  for (uint32_t i = 0; i < g->nnodes; ++i) idx[i] = g->row_start[i + 1];
  ok = ok && fwrite(idx.data(), 8, g->nnodes, f) == g->nnodes;
  ok = ok && fwrite(g->edge_dst, 4, g->nedges, f) == g->nedges;
  uint32_t pad = 0;
  if (g->nedges & 1) ok = ok && fwrite(&pad, 4, 1, f) == 1;
  if (with_weights) ok = ok && fwrite(g->edge_data, 4, g->nedges, f) == g->nedges;
  ok = (fclose(f) == 0) && ok;
  if (!ok) return fail(LSG_ERR_IO, "write to %s failed", path);
  return LSG_OK;
}

int lsg_output_dist(const int32_t *dist, uint32_t n, const char *path) {
  if (!path) return LSG_OK;  // support.cu:37-38
  if (!dist) return fail(LSG_ERR_ARG, "output: NULL distances");
  FILE *f;
  bool own = strcmp(path, "-") != 0;
  f = own ? fopen(path, "w") : stdout;  // support.cu:40-43
  if (!f) return fail(LSG_ERR_IO, "unable to open %s for writing: %s", path, strerror(errno));
  const size_t CH = 1 << 16;
  std::vector<char> buf(CH * 26);
  bool ok = true;
  for (uint32_t base = 0; base < n && ok; base += CH) {
    uint32_t end = (uint32_t)std::min<uint64_t>((uint64_t)base + CH, n);
    char *p = buf.data();
    for (uint32_t i = base; i < end; ++i) p = fmt_line(p, i, dist[i]);
    size_t bytes = (size_t)(p - buf.data());
    ok = fwrite(buf.data(), 1, bytes, f) == bytes;
  }
  if (own) ok = (fclose(f) == 0) && ok;
  else fflush(f);
  if (!ok) return fail(LSG_ERR_IO, "write to %s failed", path);
  return LSG_OK;
}

int lsg_output(const lsg_graph *g, const char *path) {
  if (!g) return fail(LSG_ERR_ARG, "output: NULL graph");
  if (g->device_graph) return fail(LSG_ERR_ARG, "output: expected a host graph (copy_to_cpu first)");
  return lsg_output_dist(g->node_data, g->nnodes, path);
}

int lsg_output_cc(const lsg_graph *g, const char *path) {
  if (!g) return fail(LSG_ERR_ARG, "output: NULL graph");
  if (g->device_graph) return fail(LSG_ERR_ARG, "output: expected a host graph (copy_to_cpu first)");
  if (!path) return LSG_OK;  // apps/cc/support.cu:20-21
  if (!g->node_data) return fail(LSG_ERR_ARG, "output: NULL labels");
  bool own = strcmp(path, "-") != 0;
  FILE *f = own ? fopen(path, "w") : stdout;  // support.cu:23-28
  if (!f) return fail(LSG_ERR_IO, "unable to open %s for writing: %s", path, strerror(errno));
  bool ok = true;
  // support.cu:30-35: the node's label, then each out-edge with its weight and
  // the destination's label
  for (uint32_t i = 0; i < g->nnodes && ok; ++i) {
    ok = fprintf(f, "%d %d\n", (int)i, g->node_data[i]) > 0;
    for (uint32_t j = g->row_start[i]; j < g->row_start[i + 1] && ok; ++j) {
      const uint32_t dst = g->edge_dst[j];
      ok = fprintf(f, "\te %d: %d %d %d\n", (int)j, (int)dst, g->edge_data ? g->edge_data[j] : 0,
                   g->node_data[dst]) > 0;
    }
  }
  if (own) ok = (fclose(f) == 0) && ok;
  else fflush(f);
  if (!ok) return fail(LSG_ERR_IO, "write to %s failed", path);
  return LSG_OK;
}

int lsg_output_pr(const float *rank, uint32_t n, const char *path, int ids_only) {
  if (!rank) return fail(LSG_ERR_ARG, "output: NULL ranks");
  float sum = 0;  // apps/pr/support.cu:57-62: a float running sum in node order
  for (uint32_t i = 0; i < n; ++i) sum += rank[i];
  fprintf(stdout, "sum: %f (%d)\n", sum, (int)n);
  if (!path) return LSG_OK;
  bool own = strcmp(path, "-") != 0;
  FILE *f = own ? fopen(path, "w") : stdout;
  if (!f) return fail(LSG_ERR_IO, "unable to open %s for writing: %s", path, strerror(errno));
  bool ok = true;
  for (uint32_t i = 0; i < n && ok; ++i)  // support.cu:91-96
    ok = (ids_only ? fprintf(f, "%d\n", (int)i) : fprintf(f, "%d %f\n", (int)i, (double)rank[i])) > 0;
  if (own) ok = (fclose(f) == 0) && ok;
  else fflush(f);
  if (!ok) return fail(LSG_ERR_IO, "write to %s failed", path);
  return LSG_OK;
}

uint32_t lsg_max_degree_vertex(const lsg_graph *g) {
  if (!g || !g->row_start || g->device_graph) return 0;
  uint32_t best = 0, bd = 0;
  for (uint32_t v = 0; v < g->nnodes; ++v) {
    uint32_t d = g->row_start[v + 1] - g->row_start[v];
    if (d > bd) {
      bd = d;
      best = v;
    }
  }
  return best;
}

int lsg_gen_grid(lsg_graph *g, uint32_t rows, uint32_t cols, uint32_t wlo, uint32_t whi, uint64_t seed) {
  if (!g || rows == 0 || cols == 0 || wlo > whi || whi > (uint32_t)INT_MAX)
    return fail(LSG_ERR_ARG, "gen_grid: bad arguments");
  uint64_t n = (uint64_t)rows * cols;
  uint64_t m = 2ull * (rows * (uint64_t)(cols - 1) + cols * (uint64_t)(rows - 1));
  if (n >= UINT32_MAX || m >= UINT32_MAX) return fail(LSG_ERR_ARG, "gen_grid: too large");
  int rc = lsg_graph_alloc_host(g, (uint32_t)n, (uint32_t)m);
  if (rc) return rc;
  auto degree = [&](uint32_t r, uint32_t c) {
    return (uint32_t)(r > 0) + (c > 0) + (c + 1 < cols) + (r + 1 < rows);
  };
  g->row_start[0] = 0;
  for (uint32_t r = 0; r < rows; ++r)
    for (uint32_t c = 0; c < cols; ++c) {
      uint32_t id = r * cols + c;
      g->row_start[id + 1] = g->row_start[id] + degree(r, c);
    }
  parallel_for(n, [&](size_t lo, size_t hi, unsigned) {
    for (size_t id = lo; id < hi; ++id) {
      uint32_t r = (uint32_t)(id / cols), c = (uint32_t)(id % cols);
      uint32_t o = g->row_start[id];
      uint32_t nb[4];
      int k = 0;
      if (r > 0) nb[k++] = (uint32_t)id - cols;
      if (c > 0) nb[k++] = (uint32_t)id - 1;
      if (c + 1 < cols) nb[k++] = (uint32_t)id + 1;
      if (r + 1 < rows) nb[k++] = (uint32_t)id + cols;
      for (int j = 0; j < k; ++j) {
        g->edge_dst[o + j] = nb[j];
        g->edge_data[o + j] = (int32_t)edge_weight(seed, (uint32_t)id, nb[j], wlo, whi);
      }
    }
  });
  return LSG_OK;
}

int lsg_gen_rmat(lsg_graph *g, uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                 uint32_t wlo, uint32_t whi, uint64_t seed) {
  if (!g || scale == 0 || scale > 30 || edge_factor == 0 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0 ||
      wlo > whi || whi > (uint32_t)INT_MAX)
    return fail(LSG_ERR_ARG, "gen_rmat: bad arguments");
  uint64_t n = 1ull << scale, m = n * edge_factor;
  if (m >= UINT32_MAX) return fail(LSG_ERR_ARG, "gen_rmat: %llu pairs exceed the u32 index range",
                                   (unsigned long long)m);
  // Quadrant thresholds on a 32-bit uniform draw (integer compares keep the
  // stream identical on every platform).
  const double two32 = 4294967296.0;
  const uint64_t ta = (uint64_t)std::llround(a * two32), tb = (uint64_t)std::llround((a + b) * two32),
                 tc = (uint64_t)std::llround((a + b + c) * two32);
  std::vector<uint32_t> src(m), dst(m);
  parallel_for(m, [&](size_t lo, size_t hi, unsigned) {
    for (size_t i = lo; i < hi; ++i) {
      uint64_t s = hash2(seed, i);
      uint32_t u = 0, v = 0;
      for (uint32_t k = 0; k < scale; ++k) {
        if ((k & 1) == 0) s = splitmix64(s);
        uint64_t r = (k & 1) ? (s >> 32) : (s & 0xFFFFFFFFull);
        uint32_t bit = 1u << (scale - 1 - k);
        if (r < ta) {
        } else if (r < tb) {
          v |= bit;
        } else if (r < tc) {
          u |= bit;
        } else {
          u |= bit;
          v |= bit;
        }
      }
      src[i] = u;
      dst[i] = v;
    }
  });
  return build_csr(g, (uint32_t)n, src, dst, true, wlo, whi, seed);
}

int lsg_gen_roadlike(lsg_graph *g, uint32_t rows, uint32_t cols, double drop, double shortcut_frac,
                     uint32_t wlo, uint32_t whi, uint64_t seed) {
  if (!g || rows == 0 || cols == 0 || drop < 0 || drop > 1 || shortcut_frac < 0 || wlo > whi ||
      whi > (uint32_t)INT_MAX)
    return fail(LSG_ERR_ARG, "gen_roadlike: bad arguments");
  uint64_t n = (uint64_t)rows * cols;
  if (n >= UINT32_MAX) return fail(LSG_ERR_ARG, "gen_roadlike: too large");
  const uint64_t keep_t = (uint64_t)std::llround((1.0 - drop) * 4294967296.0);
  // Undirected grid edges: id 2*node = edge to the right, 2*node+1 = edge down.
  std::vector<uint8_t> kept(2 * n);
  std::atomic<uint64_t> nkept{0};
  parallel_for(n, [&](size_t lo, size_t hi, unsigned) {
    uint64_t local = 0;
    for (size_t id = lo; id < hi; ++id) {
      uint32_t r = (uint32_t)(id / cols), c = (uint32_t)(id % cols);
      for (int k = 0; k < 2; ++k) {
        bool exists = k == 0 ? (c + 1 < cols) : (r + 1 < rows);
        bool keep = exists && (hash3(seed ^ 0x524f4144ull, id, k) & 0xFFFFFFFFull) < keep_t;
        kept[2 * id + k] = keep;
        local += keep;
      }
    }
    nkept += local;
  });
  uint64_t nshort = (uint64_t)std::llround(shortcut_frac * (double)nkept.load());
  uint64_t m = 2 * (nkept.load() + nshort);
  if (m >= UINT32_MAX) return fail(LSG_ERR_ARG, "gen_roadlike: too many edges");
  std::vector<uint32_t> src, dst;
  src.reserve(m);
  dst.reserve(m);
  for (uint64_t id = 0; id < n; ++id) {
    for (int k = 0; k < 2; ++k) {
      if (!kept[2 * id + k]) continue;
      uint32_t u = (uint32_t)id, v = k == 0 ? (uint32_t)id + 1 : (uint32_t)(id + cols);
      src.push_back(u);
      dst.push_back(v);
      src.push_back(v);
      dst.push_back(u);
    }
  }
  for (uint64_t i = 0; i < nshort; ++i) {
    uint64_t h = hash2(seed ^ 0x53484f5254ull, i);
    uint32_t u = (uint32_t)((h & 0xFFFFFFFFull) * n >> 32), v = (uint32_t)((h >> 32) * n >> 32);
    src.push_back(u);
    dst.push_back(v);
    src.push_back(v);
    dst.push_back(u);
  }
  return build_csr(g, (uint32_t)n, src, dst, true, wlo, whi, seed);
}

}  // extern "C"
