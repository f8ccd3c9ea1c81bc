/*
 * oracle.c -- CPU restatement of LonestarGPU's SSSP/BFS path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker.  The product (galoisgpu_amd/liblsg.so) never links or calls it.
 *
 * The reference path is CUDA-only and has no CPU implementation, so the
 * "reference CPU path" is this restatement of its algorithm, single-threaded:
 *   oracle_nearfar     apps/sssp/kernel.cu:23-35 (init), :129-144 (relax and
 *                      near/far split), :177-306 (pop + filter), :36-66
 *                      (far-pile dedupe), :355-391 (phase loop, pile swaps),
 *                      :455-480 (seeding, curdelta = 0), rt/include/pipe.h:73-87
 *                      (advance2 / retry2), rt/include/worklist.h:224-246 (slots)
 *   oracle_bfs         apps/bfs/kernel.cu:20-32 (init), :121-274 (visit),
 *                      :312-334 (level loop), :380-393 (LEVEL = 1)
 *   oracle_dijkstra    textbook binary-heap Dijkstra over the same int32
 *                      arithmetic (INF = INT_MAX, apps/sssp/kernel.cu:18)
 *   oracle_cc          weakly connected components, labelled by the smallest
 *                      id of each component (apps/cc/kernel.cu:630-723's
 *                      partition; union-find)
 *   oracle_pagerank    the fixpoint apps/pr/kernel.cu converges to, by Jacobi
 *                      iteration in double
 *   oracle_sssp_certificate  O(E) shortest-path certificate: d[src] = 0,
 *                      every edge out of a reached vertex is relaxed, every
 *                      reached vertex other than the source has a tight
 *                      in-edge, and unreached vertices hold INF.
 * Candidates d[u] + w are formed in 64 bits; one >= INT_MAX never improves
 * INF (the reference's int addition is undefined there).
 *
 * Parity pins (tests/test_oracle.py): scipy.sparse.csgraph.dijkstra and
 * networkx golden vectors (tests/golden/), closed forms (grid Manhattan
 * distance, path prefix sums, stars, disconnected parts), brute-force
 * Bellman-Ford on tiny graphs, and the certificate.  Planted mistakes in the
 * oracle must fail them (tests/test_oracle.py::test_pins_catch_planted_bugs).
 */
#include <limits.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_INF INT_MAX

/* Plants a mistake in the oracle so the tests can check that their pins catch
 * it: 0 none, 1 drop the weight term, 2 off-by-one in the row bound,
 * 3 wrong comparison (improve only when better by more than 1), 4 BFS levels
 * start at 0. */
static int g_plant = 0;
void oracle_plant_bug(int which) { g_plant = which; }

typedef struct {
  uint32_t *a;
  size_t n, cap;
} vec_u32;

static int vpush(vec_u32 *v, uint32_t x) {
  if (v->n == v->cap) {
    size_t nc = v->cap ? v->cap * 2 : 1024;
    uint32_t *p = (uint32_t *)realloc(v->a, nc * sizeof(uint32_t));
    if (!p) return -1;
    v->a = p;
    v->cap = nc;
  }
  v->a[v->n++] = x;
  return 0;
}

static uint32_t row_end(const uint32_t *row, uint32_t u) { return row[u + 1] - (g_plant == 2 && row[u + 1] > row[u]); }

static int64_t weight_of(const int32_t *w, uint32_t e) { return g_plant == 1 ? 0 : (int64_t)w[e]; }

/* ---- Dijkstra ------------------------------------------------------------ */
typedef struct {
  int64_t d;
  uint32_t v;
} hitem;

static void hsift_up(hitem *h, size_t i) {
  while (i > 0) {
    size_t p = (i - 1) / 2;
    if (h[p].d <= h[i].d) break;
    hitem t = h[p];
    h[p] = h[i];
    h[i] = t;
    i = p;
  }
}

static void hsift_down(hitem *h, size_t n, size_t i) {
  for (;;) {
    size_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < n && h[l].d < h[m].d) m = l;
    if (r < n && h[r].d < h[m].d) m = r;
    if (m == i) break;
    hitem t = h[m];
    h[m] = h[i];
    h[i] = t;
    i = m;
  }
}

int oracle_dijkstra(uint32_t n, const uint32_t *row, const uint32_t *col, const int32_t *w, uint32_t src,
                    int32_t *dist) {
  if (src >= n) return -1;
  for (uint32_t v = 0; v < n; ++v) dist[v] = ORC_INF;
  size_t cap = 1024, hn = 0;
  hitem *h = (hitem *)malloc(cap * sizeof(hitem));
  if (!h) return -2;
  dist[src] = 0;
  h[hn++] = (hitem){0, src};
  while (hn) {
    hitem top = h[0];
    h[0] = h[--hn];
    hsift_down(h, hn, 0);
    if (top.d != dist[top.v]) continue;
    uint32_t u = top.v;
    for (uint32_t e = row[u]; e < row_end(row, u); ++e) {
      uint32_t v = col[e];
      int64_t nd = top.d + weight_of(w, e);
      int better = g_plant == 3 ? nd + 1 < (int64_t)dist[v] : nd < (int64_t)dist[v];
      if (better && nd < ORC_INF) {
        dist[v] = (int32_t)nd;
        if (hn == cap) {
          cap *= 2;
          hitem *p = (hitem *)realloc(h, cap * sizeof(hitem));
          if (!p) {
            free(h);
            return -2;
          }
          h = p;
        }
        h[hn] = (hitem){nd, v};
        hsift_up(h, hn++);
      }
    }
  }
  free(h);
  return 0;
}

/* ---- near/far restatement ------------------------------------------------- */
/* stats[0] = phases ("iterations: %d"), [1] = near rounds, [2] = items
 * popped, [3] = edges examined, [4] = final curdelta. */
int oracle_nearfar(uint32_t n, const uint32_t *row, const uint32_t *col, const int32_t *w, uint32_t src,
                   int32_t delta, int32_t *dist, uint64_t *stats) {
  if (src >= n || delta <= 0) return -1;
  uint64_t st[5] = {0, 0, 0, 0, 0};
  /* kernel (init), apps/sssp/kernel.cu:23-35 */
  for (uint32_t v = 0; v < n; ++v) dist[v] = v == src ? 0 : ORC_INF;
  /* three piles: in / out (far) / re (near), rt/include/pipe.h:37-43 */
  vec_u32 pile[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  int in = 0, out = 1, re = 2, t;
  int32_t *marks = (int32_t *)malloc((size_t)n * sizeof(int32_t)); /* Shared<int> level(V), :464 */
  if (!marks) return -2;
  int64_t curdelta = 0; /* :472 */
  if (vpush(&pile[in], src)) goto oom;  /* :476-477 */
  while (pile[in].n) {            /* :363 */
    while (pile[in].n) {          /* :365 */
      st[1]++;
      /* sssp_kernel_dev (:177-306): pop every item, skip INF nodes (:202),
       * relax every out-edge (:129-144). */
      for (size_t i = 0; i < pile[in].n; ++i) {
        uint32_t u = pile[in].a[i];
        st[2]++;
        if (dist[u] == ORC_INF) continue;
        for (uint32_t e = row[u]; e < row_end(row, u); ++e) {
          uint32_t v = col[e];
          int64_t nd = (int64_t)dist[u] + weight_of(w, e);
          st[3]++;
          int better = g_plant == 3 ? (int64_t)dist[v] > nd + 1 : (int64_t)dist[v] > nd;
          if (better && nd < ORC_INF) {
            dist[v] = (int32_t)nd; /* atomicMin, :133 */
            if (nd <= curdelta) {  /* :134 */
              if (vpush(&pile[re], v)) goto oom;
            } else {
              if (vpush(&pile[out], v)) goto oom;
            }
          }
        }
      }
      pile[in].n = 0;           /* in.swap_slots(): the fresh slot is empty (:370) */
      t = re; re = in; in = t;  /* retry2 (:372, pipe.h:84-87) */
    }
    t = in; in = out; out = t;  /* advance2 (:375) */
    /* remove_dups_dev (:36-66): marks[node] = index; keep the index that
     * survived the race (here: the last writer). */
    for (size_t i = 0; i < pile[in].n; ++i) marks[pile[in].a[i]] = (int32_t)i;
    for (size_t i = 0; i < pile[in].n; ++i) {
      uint32_t node = pile[in].a[i];
      if (marks[node] == (int32_t)i && vpush(&pile[out], node)) goto oom;
    }
    pile[in].n = 0;             /* in.swap_slots() (:379) */
    t = in; in = out; out = t;  /* advance2 (:381) */
    st[0]++;                    /* i++ (:382) */
    curdelta += delta;          /* :383 */
  }
  st[4] = (uint64_t)curdelta;
  for (int k = 0; k < 3; ++k) free(pile[k].a);
  free(marks);
  if (stats) memcpy(stats, st, sizeof st);
  return 0;
oom:
  for (int k = 0; k < 3; ++k) free(pile[k].a);
  free(marks);
  return -2;
}

/* ---- level-synchronous BFS restatement ------------------------------------ */
/* stats[0] = levels run, [1] = items popped, [2] = edges examined. */
int oracle_bfs(uint32_t n, const uint32_t *row, const uint32_t *col, uint32_t src, int32_t *level,
               uint64_t *stats) {
  if (src >= n) return -1;
  uint64_t st[3] = {0, 0, 0};
  for (uint32_t v = 0; v < n; ++v) level[v] = v == src ? 0 : ORC_INF; /* bfs_init, :20-32 */
  vec_u32 a = {0, 0, 0}, b = {0, 0, 0};
  int32_t LEVEL = g_plant == 4 ? 0 : 1; /* :388 */
  if (vpush(&a, src)) return -2;
  while (a.n) { /* :319 */
    st[0]++;
    for (size_t i = 0; i < a.n; ++i) {
      uint32_t u = a.a[i];
      st[1]++;
      for (uint32_t e = row[u]; e < row_end(row, u); ++e) {
        uint32_t v = col[e];
        st[2]++;
        if (level[v] == ORC_INF) { /* :194-200 */
          level[v] = LEVEL;
          if (vpush(&b, v)) {
            free(a.a);
            free(b.a);
            return -2;
          }
        }
      }
    }
    vec_u32 t = a; /* swap_slots + advance2, :324-326 */
    a = b;
    b = t;
    b.n = 0;
    LEVEL++; /* :327 */
  }
  free(a.a);
  free(b.a);
  if (stats) memcpy(stats, st, sizeof st);
  return 0;
}

/* ---- certificate ---------------------------------------------------------- */
/* Returns the number of violated conditions (0 = dist is the shortest-path
 * distance vector from src), or -1 on bad arguments.  Uses the graph's own
 * weights; pass w = NULL for unit weights (BFS levels). */
int64_t oracle_sssp_certificate(uint32_t n, const uint32_t *row, const uint32_t *col, const int32_t *w,
                                uint32_t src, const int32_t *dist) {
  if (src >= n) return -1;
  int64_t bad = 0;
  if (dist[src] != 0) bad++;
  uint8_t *tight = (uint8_t *)calloc(n, 1);
  if (!tight) return -1;
  tight[src] = 1;
  for (uint32_t u = 0; u < n; ++u) {
    if (dist[u] < 0) bad++;
    if (dist[u] == ORC_INF) continue;
    for (uint32_t e = row[u]; e < row[u + 1]; ++e) {
      uint32_t v = col[e];
      int64_t nd = (int64_t)dist[u] + (w ? (int64_t)w[e] : 1);
      if (nd < ORC_INF && (int64_t)dist[v] > nd) bad++; /* an edge that still relaxes */
      if ((int64_t)dist[v] == nd) tight[v] = 1;
    }
  }
  for (uint32_t v = 0; v < n; ++v)
    if (dist[v] != ORC_INF && !tight[v]) bad++; /* a finite value with no witness path */
  free(tight);
  return bad;
}

/* V_r = vertices with a finite value; E_r = sum of their out-degrees. */
void oracle_reach(uint32_t n, const uint32_t *row, const int32_t *dist, uint64_t *vr, uint64_t *er) {
  uint64_t a = 0, b = 0;
  for (uint32_t v = 0; v < n; ++v)
    if (dist[v] != ORC_INF) {
      a++;
      b += row[v + 1] - row[v];
    }
  *vr = a;
  *er = b;
}

/* Connected components of the graph with edge directions ignored
 * (apps/cc/kernel.cu:630-723 computes the same partition by hooking and
 * pointer jumping).  label[v] = the smallest vertex id of v's component, a
 * canonical form of the reference's schedule-dependent labels.  Union-find
 * with union by smaller root and path halving.  Planted mistake 4 (tests
 * only): the last edge of every row is skipped.  Returns the number of
 * components. */
static uint32_t uf_find(uint32_t *p, uint32_t x) {
  while (p[x] != x) {
    p[x] = p[p[x]];
    x = p[x];
  }
  return x;
}

int64_t oracle_cc(uint32_t n, const uint32_t *row, const uint32_t *col, int32_t *label) {
  uint32_t *p = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  if (!p) return -1;
  for (uint32_t v = 0; v < n; ++v) p[v] = v;
  for (uint32_t u = 0; u < n; ++u) {
    const uint32_t end = row[u + 1] - (g_plant == 4 && row[u + 1] > row[u]);
    for (uint32_t e = row[u]; e < end; ++e) {
      uint32_t a = uf_find(p, u), b = uf_find(p, col[e]);
      if (a == b) continue;
      if (a < b) p[b] = a; /* the smaller id stays the root */
      else p[a] = b;
    }
  }
  int64_t comps = 0;
  for (uint32_t v = 0; v < n; ++v) {
    label[v] = (int32_t)uf_find(p, v);
    comps += label[v] == (int32_t)v;
  }
  free(p);
  return comps;
}

/* PageRank fixpoint (apps/pr/kernel.cu:19-20,38,110-236 reach it by pushing
 * residuals):  p[v] = (1 - alpha) + alpha * sum over edges (u, v) of
 * p[u] / outdeg(u), by Jacobi iteration in double until the largest change
 * falls below tol or max_iter passes.  Returns the iterations taken.
 * Planted mistake 5 (tests only): the out-degree of every vertex is taken
 * one too large. */
int oracle_pagerank(uint32_t n, const uint32_t *row, const uint32_t *col, double alpha, double tol, int max_iter,
                    double *p) {
  double *q = (double *)malloc((size_t)n * sizeof(double));
  double *c = (double *)malloc((size_t)n * sizeof(double));
  if (!q || !c) {
    free(q);
    free(c);
    return -1;
  }
  for (uint32_t v = 0; v < n; ++v) p[v] = 1.0 - alpha;
  int it = 0;
  while (it < max_iter) {
    for (uint32_t u = 0; u < n; ++u) {
      const uint32_t od = row[u + 1] - row[u] + (g_plant == 5);
      c[u] = od ? p[u] / (double)od : 0.0;
    }
    for (uint32_t v = 0; v < n; ++v) q[v] = 0.0;
    for (uint32_t u = 0; u < n; ++u)
      for (uint32_t e = row[u]; e < row[u + 1]; ++e) q[col[e]] += c[u];
    double change = 0.0;
    for (uint32_t v = 0; v < n; ++v) {
      const double pn = (1.0 - alpha) + alpha * q[v];
      const double d = pn > p[v] ? pn - p[v] : p[v] - pn;
      if (d > change) change = d;
      p[v] = pn;
    }
    ++it;
    if (change < tol) break;
  }
  free(q);
  free(c);
  return it;
}
