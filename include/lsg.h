/*
 * lsg.h -- C ABI of the MI355X (gfx950) worklist shortest-path library.
 *
 * This is the drop-in boundary for LonestarGPU's apps/sssp and apps/bfs path.
 * The reference's boundary is a link-time plugin contract between the CLI
 * skeleton and each app (skelapp/skel.cu:12-32): the skeleton reads a .gr file
 * into a host CSRGraph, copies it to the device, times the app's gg_main and
 * prints the result with the app's output().  Every entry point below replaces
 * one piece of that contract and cites the reference function it stands in
 * for.  All entry points are plain C: pointers, sizes and fixed-width ints, no
 * HIP or torch types.  They return LSG_OK (0) or a negative LSG_ERR_* code
 * and never call exit(); lsg_last_error() gives a message for the calling
 * thread.
 *
 * Distances and levels are int32 with INF = INT_MAX, exactly as the reference
 * stores them (apps/sssp/kernel.cu:15-18, apps/bfs/kernel.cu:14-16).
 */
#ifndef LSG_H
#define LSG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSG_INF 0x7FFFFFFF /* INT_MAX, apps/sssp/kernel.cu:18 */
#define LSG_ABI_VERSION 2

enum {
  LSG_OK = 0,
  LSG_ERR_ARG = -1,      /* bad argument (bad source, delta <= 0, NULL pointer) */
  LSG_ERR_IO = -2,       /* open/read/write/mmap failure */
  LSG_ERR_FORMAT = -3,   /* malformed .gr file or graph arrays */
  LSG_ERR_NOMEM = -4,    /* host or device allocation failed */
  LSG_ERR_HIP = -5,      /* HIP runtime error */
  LSG_ERR_CAPACITY = -6, /* a device queue would overflow (never silently dropped) */
  LSG_ERR_TIMEOUT = -7,  /* a grid barrier exceeded its spin limit; results invalid */
  LSG_ERR_NODEV = -8     /* no usable gfx950 device */
};

/* Mirrors struct CSRGraph (rt/include/csr_graph.h:22-24,96-101): index_type is
 * unsigned, edge_data_type and node_data_type are int.  row_start has
 * nnodes+1 entries, row_start[nnodes] == nedges.  When device_graph is 1 the
 * four arrays are device pointers on HIP device `device`. */
typedef struct lsg_graph {
  uint32_t nnodes, nedges;
  uint32_t *row_start;
  uint32_t *edge_dst;
  int32_t *edge_data;
  int32_t *node_data;
  int32_t device_graph;
  int32_t device;
} lsg_graph;

/* Expansion modes: how a frontier vertex's edges are spread over lanes. */
enum {
  LSG_EXPAND_TWB = 0, /* by degree: thread (< thread_deg), wavefront (< heavy_deg),
                         heavy vertices split into block chunks */
  LSG_EXPAND_LB = 1   /* tile-flattened (default): block prefix sum of degrees, every
                         lane takes one edge (the reference's fine-grained np) */
};

/* Far-pile modes.  Every mode gives the same distances; they differ in which
 * vertices a phase expands. */
enum {
  LSG_FAR_REFERENCE = 0, /* whole far pile becomes the next phase (apps/sssp/kernel.cu:374-383) */
  LSG_FAR_SPLIT = 1,     /* far items above the next threshold wait for a later phase */
  LSG_FAR_BUCKET = 2     /* delta-stepping's lowest non-empty bucket: items above the threshold
                            wait as in SPLIT, and the threshold skips empty buckets.  Single
                            source: curdelta moves on by whole steps of delta until it reaches
                            the smallest waiting distance.  Batched: each source's threshold is
                            its own smallest pending distance + delta - 1 */
};

/* Options of the sssp/bfs apps: -d, -s and -l (apps/sssp/support.cu:6-31,
 * apps/bfs/support.cu:5-25) plus the schedule knobs the autotuner and the
 * learned scheduler set.  Zero means "default" for every tuning field. */
typedef struct lsg_opts {
  int32_t delta;       /* -d, sssp only; default 10000 (apps/sssp/support.cu:10) */
  int32_t start_node;  /* -s; default 0 (apps/sssp/support.cu:12) */
  int32_t enable_lb;   /* -l: the reference's host-driven loop with an edge-balanced pass
                          (apps/sssp/kernel.cu:318-354, apps/bfs/kernel.cu:286-311): each round
                          an inspect pass lists the frontier vertices of degree >= lb_deg, a prefix
                          sum of their degrees spreads their edges one per thread over the grid,
                          and the sweep relaxes the other vertices */
  int32_t expand_mode; /* LSG_EXPAND_* */
  int32_t far_mode;    /* LSG_FAR_* */
  int32_t block_size;  /* threads per block: 256 (default), 512 or 1024 */
  int32_t grid_blocks; /* persistent grid; 0 = every block the device can keep resident */
  int32_t thread_deg;  /* TWB: degrees below this are expanded by one lane (default 8) */
  int32_t heavy_deg;   /* degrees >= this are split into chunks of heavy_deg edges, expanded by whole
                          blocks in the next round (default 512; batched runs need >= 8) */
  int32_t sched;       /* 0 = fixed delta; 1 = the learned scheduler picks each phase's delta width
                          and mode (expansion mode for SSSP, far mode for batched runs) from the
                          model loaded with lsg_sched_set_model */
  int32_t trace;       /* 1 = record one lsg_round_trace per round (lsg_last_trace) */
  int32_t graph_launch; /* 1 = launch the kernel sequence as a captured hipGraph (replayed while
                           the arguments stay the same) */
  int32_t lb_deg;      /* -l: degree threshold of the edge-balanced pass; 0 = the sweep's total
                          thread count, the reference's DEGREE_LIMIT (rt/include/ggcuda.h:19-22) */
  int32_t single_block; /* 1 = the near/far rounds run in one workgroup with the piles in LDS and
                           a workgroup barrier between rounds (same distances), for small sparse
                           high-diameter graphs, where a round of the persistent kernel is a chain
                           of global round trips: SSSP on the 256x256 grid 8.0 -> 4.2 ms at delta
                           32, 2.5 ms at delta 256; BFS 3.6 -> 1.5 ms.  When the piles outgrow
                           LDS (SSSP: 2048 entries per round, 4096 far; BFS levels wider than
                           1024 vertices) the persistent kernel continues from the distances
                           reached (road-like 1024x1024 at delta 10000: 3.4 -> 3.0 ms); a
                           frontier that explodes at once (RMAT BFS from a hub) costs that launch
                           extra.  stats then add both launches' rounds, items and kernel time;
                           a trace holds the persistent kernel's rounds only.  Single-source
                           lsg_sssp_dist / lsg_bfs_dist / gg_main only */
  int32_t reserved[2];
} lsg_opts;

/* One record per round of a single-source traversal, written when
 * lsg_opts.trace is set.  No reference counterpart: it is the input of the
 * learned scheduler's trace generator and of the autotuner. */
typedef struct lsg_round_trace {
  uint32_t round, phase;
  uint32_t in_items;     /* frontier items popped this round */
  uint32_t heavy_chunks; /* heavy chunks consumed this round */
  uint32_t near_out;     /* items pushed to the near pile */
  uint32_t far_out;      /* far pile size after this round */
  uint32_t heavy_out;    /* heavy chunks queued for the next round */
  uint32_t curdelta_lo, curdelta_hi;
  uint32_t delta;        /* width in force this round */
  uint64_t t_end;        /* device wall clock at the end of the round */
} lsg_round_trace;

/* Run statistics.  ms_total is host time of the whole call, the reference's
 * timed region around gg_main (skelapp/skel.cu:60-63).  The reference
 * allocates its worklists inside that region on every call
 * (apps/sssp/kernel.cu:459,475); this library keeps its device buffers
 * between calls, so ms_total includes allocation only on a device's first
 * call or after lsg_workspace_release.  ms_kernel is device time of the
 * traversal kernels only (HIP events on the library's stream; for -l runs,
 * the whole host-driven loop).  heavy_chunks: heavy chunks consumed, or for
 * -l runs the edges relaxed by the edge-balanced pass. */
typedef struct lsg_stats {
  double ms_total;
  double ms_kernel;
  uint64_t iterations;  /* sssp: delta phases ("iterations: %d", apps/sssp/kernel.cu:479);
                           bfs: levels */
  uint64_t near_iters;  /* queue rounds over all phases */
  uint64_t items;       /* frontier items popped */
  uint64_t heavy_chunks;
  int32_t last_delta;   /* curdelta at exit */
  int32_t grid_blocks;
  int32_t block_size;
  int32_t reserved[5];
} lsg_stats;

/* ---- version / errors ---------------------------------------------------- */
int lsg_abi_version(void);
const char *lsg_last_error(void);
void lsg_opts_default(lsg_opts *o);

/* ---- graph container (rt/src/csr_graph.cu) -------------------------------- */
/* CSRGraph::read -> readFromGR (rt/src/csr_graph.cu:152-241): reads a .gr v1
 * file into a host graph (arrays malloc'ed, node_data zeroed).  Unlike the
 * reference, which only warns about dst >= nnodes (:219) and asserts on the
 * version (:189), every inconsistency is an LSG_ERR_FORMAT. Weights above
 * INT_MAX are rejected too. `quiet` suppresses the reference's progress prints. */
int lsg_graph_read(lsg_graph *g, const char *path, int quiet);
/* .gr v1 writer, the inverse of lsg_graph_read (byte layout of
 * rt/src/csr_graph.cu:187-199).  with_weights = 0 writes sizeEdgeTy = 0. */
int lsg_graph_write(const lsg_graph *g, const char *path, int with_weights);
/* CSRGraph::allocOnHost (rt/src/csr_graph.cu:28-46): calloc'ed host arrays. */
int lsg_graph_alloc_host(lsg_graph *g, uint32_t nnodes, uint32_t nedges);
/* CSRGraph::copy_to_gpu (rt/src/csr_graph.cu:250-261): allocates `dev` on HIP
 * device `device` and copies the four arrays.  Validates the CSR first. */
int lsg_graph_copy_to_gpu(const lsg_graph *host, lsg_graph *dev, int device);
/* CSRGraph::copy_to_cpu (rt/src/csr_graph.cu:263-275).  Only node_data can
 * change on the device, so only node_data is copied back. */
int lsg_graph_copy_to_cpu(const lsg_graph *dev, lsg_graph *host);
/* CSRGraph::dealloc (rt/src/csr_graph.cu:243-248). */
void lsg_graph_dealloc(lsg_graph *g);
/* Reverse CSR of a device graph, built on its device: gt's row v lists the
 * sources of v's in-edges with their weights, in forward order (ascending
 * source, and for one source in the order of its out-edges), the same on
 * every call.  gt is allocated here and freed with lsg_graph_dealloc.  No
 * reference counterpart; bottom-up BFS steps and PageRank read it. */
int lsg_graph_transpose(const lsg_graph *dev, lsg_graph *gt);
/* Wraps caller-owned device arrays (for example torch tensors) as a device
 * graph without copying.  node_data may be NULL when only *_dist entry points
 * are used. Nothing is freed by lsg_graph_dealloc for such graphs. */
int lsg_graph_wrap_device(lsg_graph *dev, uint32_t nnodes, uint32_t nedges, uint32_t *row_start,
                          uint32_t *edge_dst, int32_t *edge_data, int32_t *node_data, int device);

/* ---- synthetic inputs (seeded, deterministic, multi-threaded) ------------- */
/* R x C 4-neighbour grid, both directions, id = r*C + c; weights uniform in
 * [wlo, whi] drawn per directed edge from `seed`. */
int lsg_gen_grid(lsg_graph *g, uint32_t rows, uint32_t cols, uint32_t wlo, uint32_t whi, uint64_t seed);
/* Directed RMAT: ef * 2^scale pairs with quadrant probabilities a, b, c
 * (d = 1-a-b-c); self loops and duplicate edges removed; rows sorted. */
int lsg_gen_rmat(lsg_graph *g, uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                 uint32_t wlo, uint32_t whi, uint64_t seed);
/* Road-like grid: R x C grid with a fraction `drop` of undirected edges
 * removed and `shortcut_frac` * (#kept undirected edges) random long
 * undirected shortcuts added. */
int lsg_gen_roadlike(lsg_graph *g, uint32_t rows, uint32_t cols, double drop, double shortcut_frac,
                     uint32_t wlo, uint32_t whi, uint64_t seed);
/* Vertex of maximum out-degree, lowest id on ties. */
uint32_t lsg_max_degree_vertex(const lsg_graph *g);

/* ---- the apps (apps/sssp/kernel.cu:455-480, apps/bfs/kernel.cu:380-393) --- */
/* gg_main of apps/sssp: initialises gg->node_data to (v == src ? 0 : INF) and
 * runs fixed-delta near/far relaxation to the fixpoint in one persistent
 * kernel.  `hg` may be NULL (the reference only reads hg.nnodes). */
int lsg_sssp_gg_main(lsg_graph *hg, lsg_graph *gg, const lsg_opts *opts, lsg_stats *stats);
/* gg_main of apps/bfs: level-synchronous BFS; source 0, first hop LEVEL 1. */
int lsg_bfs_gg_main(lsg_graph *hg, lsg_graph *gg, const lsg_opts *opts, lsg_stats *stats);
/* Same as the gg_main entry points but writing into a caller-owned device
 * array dist[gg->nnodes] instead of gg->node_data (for callers that keep
 * several result arrays resident). */
int lsg_sssp_dist(const lsg_graph *gg, const lsg_opts *opts, int32_t *dist_dev, lsg_stats *stats);
int lsg_bfs_dist(const lsg_graph *gg, const lsg_opts *opts, int32_t *level_dev, lsg_stats *stats);
/* No counterpart in the reference: lsg_sssp_dist for callers that cannot
 * allocate device memory during a call (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It allocates nothing:
 * every buffer beyond the caller's graph and dist (control block, item and
 * heavy-chunk queues, push stamps, seed) lives in one static device array of
 * about 22 MB that each GPU holds until the process ends, so graphs are
 * limited to 2^20 vertices and 2^24 edges (LSG_ERR_ARG beyond).  Calls share
 * that array: calls that overlap on one GPU race.  Runs on the device's null
 * stream; trace, graph_launch, enable_lb and single_block are rejected (they need
 * workspace buffers).  Same distances as lsg_sssp_dist, bit for bit. */
int lsg_sssp_dist_noalloc(const lsg_graph *gg, const lsg_opts *opts, int32_t *dist_dev, lsg_stats *stats);
/* Connected components (apps/cc/kernel.cu:630-723), edges taken without
 * direction (weakly connected components).  label_dev[v] gets the smallest
 * vertex id of v's component; *ncomponents (may be NULL) the number of
 * components.  The reference's labels depend on its racing hooks; the
 * partition is the same.  All hook and jump passes run in one persistent
 * launch behind the leader grid barrier (the reference's gg_main_pipe_4
 * loop, apps/cc/kernel.cu:465-707); a barrier that times out returns
 * LSG_ERR_TIMEOUT.  stats: iterations = hook passes, items = components,
 * grid_blocks = the persistent grid. */
int lsg_cc(const lsg_graph *g, int32_t *label_dev, uint32_t *ncomponents, lsg_stats *stats);
/* gg_main of apps/cc: labels into gg->node_data ("components: %d" is
 * stats->items, apps/cc/kernel.cu:719).  `hg` may be NULL. */
int lsg_cc_gg_main(lsg_graph *hg, lsg_graph *gg, lsg_stats *stats);
/* apps/cc/support.cu:17-36 output: "%d %d" per node (id, label), then per
 * edge "\te %d: %d %d %d" (edge id, destination, weight, destination's label);
 * file "-" is stdout. */
int lsg_output_cc(const lsg_graph *host_g, const char *file);
/* PageRank (apps/pr/kernel.cu): the fixpoint of
 *   p[v] = (1 - alpha) + alpha * sum over in-edges (u, v) of p[u] / outdeg(u)
 * by pull-based Jacobi iteration over `gt`, the reverse of `g`
 * (lsg_graph_transpose), until no rank changes by more than epsilon times
 * itself or max_iterations pass.  The reference uses alpha 0.85, epsilon 1e-6
 * (apps/pr/kernel.cu:19-20) and pushes residuals instead; both reach the same
 * fixpoint, within float rounding.  rank_dev is a device array of nnodes
 * floats.  stats: iterations. */
int lsg_pagerank(const lsg_graph *g, const lsg_graph *gt, float alpha, float epsilon, int32_t max_iterations,
                 float *rank_dev, lsg_stats *stats);
/* apps/pr/support.cu:45-97 output: prints "sum: %f (%d)" (float sum of the
 * ranks, node count) on stdout, then, if file is not NULL, one line per node:
 * "%d %f" (id, rank), or "%d" with ids_only (-n); file "-" is stdout. */
int lsg_output_pr(const float *rank, uint32_t n, const char *file, int ids_only);
/* Direction-optimising BFS from `src` (no reference counterpart as an
 * algorithm; the levels are apps/bfs's, as lsg_bfs_dist writes them).  `gt`
 * is the reverse of `g` (lsg_graph_transpose).  direction: 0 = per level by
 * frontier size (top-down, bottom-up once the frontier's edges exceed 1/14 of
 * the unexplored edges, top-down again below V/24 frontier vertices), 1 =
 * top-down only, 2 = bottom-up only.  level_dev is a device array of nnodes.
 * stats: iterations = levels, near_iters = bottom-up levels. */
int lsg_bfs_do(const lsg_graph *g, const lsg_graph *gt, uint32_t src, int direction, int32_t *level_dev,
               lsg_stats *stats);
/* Copies the round records of the last traced call on `device` into out[cap];
 * *n gets the number of records, *tick_hz the device wall-clock rate. */
int lsg_last_trace(int device, lsg_round_trace *out, uint32_t cap, uint32_t *n, double *tick_hz);

/* ---- batched SSSP (no reference counterpart) ------------------------------ */
typedef struct lsg_batch_stats {
  double ms_total;
  double ms_kernel;        /* device time of all launches (HIP events) */
  uint64_t rounds, phases, items, heavy_chunks;
  uint64_t lane_items;     /* (vertex, source) expansions */
  uint64_t edges_expanded; /* edge expansions, each shared by the sources of its vertex */
  uint64_t improved_rows;  /* edge expansions in which at least one source improved */
  int32_t grid_blocks, block_size;
  int32_t launches;        /* persistent launches (one per 64 sources) */
  int32_t reserved[5];
} lsg_batch_stats;

/* Up to 64 sources of one device graph in ONE persistent launch.  dist_vs is a
 * device array [nnodes][nsrc] (row v holds the distances of v from every
 * source, in source order).  Each column equals lsg_sssp_dist with the same
 * opts.delta bit for bit.  srcs is a host array. */
int lsg_sssp_multi(const lsg_graph *g, const uint32_t *srcs, uint32_t nsrc, const lsg_opts *opts, int32_t *dist_vs,
                   lsg_batch_stats *stats);
/* Any number of sources of one device graph, 64 per launch; dist_out[i] is a
 * device array of nnodes int32 for source i (host array of device pointers). */
int lsg_sssp_batch(const lsg_graph *g, const uint32_t *srcs, uint32_t nsrc, const lsg_opts *opts,
                   int32_t *const *dist_out, lsg_batch_stats *stats);
/* Many independent device graphs (same device), one source each, in ONE
 * persistent launch: the graphs are packed into a disjoint union and run as
 * one multi-seed traversal, which gives every graph exactly its own
 * single-source distances.  dist_out[i] is a device array of graphs[i].nnodes.
 * opts.single_block = 1: one launch of one workgroup per graph instead, each
 * running its graph's near/far rounds with workgroup barriers and piles in
 * HBM (no pack, no grid barrier); same distances.  A pile overflow falls back
 * to the union launch.  It suits batches of small or high-diameter graphs
 * (256 grids of 64x64: 4.7 -> 0.47 ms; 256 RMAT graphs of scales 8-11:
 * 0.51 -> 0.35 ms); where a few large graphs dominate, one CU each is slower
 * (256 RMAT graphs of scales 10-16: 2.7 -> 5.3 ms). */
int lsg_sssp_graphs(const lsg_graph *graphs, uint32_t ngraphs, const uint32_t *srcs, const lsg_opts *opts,
                    int32_t *const *dist_out, lsg_batch_stats *stats);
/* Per-phase feature records of the last traced SSSP call (single-source or
 * batched): nphases rows of 20 floats (16 scheduler features, action or -1,
 * delta, mode -- expansion mode single-source, far mode batched -- and the
 * first round of the phase). */
int lsg_last_features(int device, float *out, uint32_t cap_phases, uint32_t *nphases);

/* ---- learned scheduler (no reference counterpart) ------------------------- */
/* Size in floats of a scheduler model (normalisation, 16-64-64-8 MLP weights,
 * per-action delta multiplier and far mode, forced per-phase actions). */
int lsg_sched_model_floats(void);
/* Model kinds: one model per kind and device. */
enum {
  LSG_SCHED_BATCHED = 0, /* lsg_sssp_multi / _batch / _graphs: actions pick delta x far mode */
  LSG_SCHED_PLAIN = 1    /* single-source SSSP: actions pick delta x expansion mode */
};
/* Loads the model of `kind` used when lsg_opts.sched = 1 on `device`. */
int lsg_sched_set_model(int device, int kind, const float *model, uint32_t nfloats);
/* Trains the MLP part of `model` in place on n samples (X: n x 16 raw
 * features, y: action labels, weights: per-sample loss weights or NULL) with
 * minibatch SGD + momentum in a HIP kernel; also sets the feature
 * normalisation.  *loss gets the mean loss over the last 100 steps. */
int lsg_mlp_train(int device, float *model, const float *X, const int32_t *y, const float *weights, uint32_t n,
                  uint32_t steps, float lr, float momentum, uint64_t seed, float *loss);
/* Argmax actions for n feature rows (HIP kernel, one thread per row). */
int lsg_mlp_infer(int device, const float *model, const float *X, uint32_t n, int32_t *labels);

/* output() of apps/sssp and apps/bfs (apps/sssp/support.cu:34-53): one line
 * per node, "%d INF\n" for INT_MAX and "%d %d\n" otherwise.  path NULL writes
 * nothing; "-" writes stdout.  `g` must be a host graph. */
int lsg_output(const lsg_graph *g, const char *path);
/* The same text for an explicit distance array. */
int lsg_output_dist(const int32_t *dist, uint32_t n, const char *path);

/* Frees the device buffers the traversals of `device` keep between calls
 * (queues, stamps, masks, batched distance planes, captured hipGraphs); the
 * next call allocates them again.  The reference allocates its worklists
 * inside every gg_main (apps/sssp/kernel.cu:459,475): a call made right after
 * this one is timed the way the reference's is. */
int lsg_workspace_release(int device);

/* Device memory for callers without a GPU runtime of their own (the Python
 * host side, tests, benchmarks).  kind: 0 host->device, 1 device->host,
 * 2 device->device. */
int lsg_dev_malloc(int device, size_t bytes, void **ptr);
int lsg_dev_free(int device, void *ptr);
int lsg_dev_copy(int device, void *dst, const void *src, size_t bytes, int kind);
/* Times a plain device-to-device copy of `bytes` (iters launches): the HBM
 * reference figure, and the known-size kernel PMC counters are scaled by. */
int lsg_stream_copy(int device, size_t bytes, int iters, double *ms_per_copy);

/* Device synchronisation and name, for the CLI's "Using GPU:" line
 * (skelapp/skel.cu:146-148). */
int lsg_set_device(int device);
int lsg_device_name(int device, char *buf, size_t len);
int lsg_device_sync(int device);

#ifdef __cplusplus
}
#endif
#endif /* LSG_H */
