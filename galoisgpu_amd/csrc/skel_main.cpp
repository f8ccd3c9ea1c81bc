// The sssp / bfs / cc / pr command line, one binary per app (-DLSG_APP_SSSP,
// -DLSG_APP_BFS, -DLSG_APP_CC, -DLSG_APP_PR).
//
// Mirrors the reference's skeleton driver and app options:
//   skelapp/skel.cu:80-151   usage, getopt "g:qo:" + app options, main
//   skelapp/skel.cu:45-78    load graph, copy to GPU, timed gg_main, copy back, output
//   apps/sssp/support.cu:6-31   "ls:d:" options, DELTA 10000, start node 0
//   apps/bfs/support.cu:5-25    "ls:" options
// Invocation: sssp [-q] [-g gpunum] [-o output-file] [-l] [-d delta] [-s startNode] graph-file
// Reference behaviours kept: the same messages on stdout/stderr, "-o -" for
// stdout, no output file when -o is absent, exit 1 with the usage text on bad
// options.  Differences: errors from the library are reported with a message
// and exit status 1 instead of assert()/crash (negative delta, source out of
// range, unreadable or malformed graph, unwritable output).

#include <getopt.h>
#include <time.h>

#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <climits>
#include <vector>

#include "lsg.h"

#if defined(LSG_APP_SSSP)
static const char *prog_opts = "lws:d:";
static const char *prog_usage = "[-l] [-w] [-d delta] [-s startNode]";
static const char *GGC_OPTIONS = "lsg-mi355x sssp: persistent near/far, wave64, gfx950";
#elif defined(LSG_APP_BFS)
static const char *prog_opts = "lwDs:";
static const char *prog_usage = "[-l] [-w] [-D] [-s startNode]";
static const char *GGC_OPTIONS = "lsg-mi355x bfs: persistent level-synchronous, wave64, gfx950";
#elif defined(LSG_APP_CC)
static const char *prog_opts = "";
static const char *prog_usage = "";
static const char *GGC_OPTIONS = "lsg-mi355x cc: hooking + pointer jumping, wave64, gfx950";
#elif defined(LSG_APP_PR)
static const char *prog_opts = "nt:x:";  // apps/pr/support.cu:15-17
static const char *prog_usage = "[-n] [-t top_ranks] [-x max_iterations]";
static const char *GGC_OPTIONS = "lsg-mi355x pr: pull Jacobi over the reverse CSR, wave64, gfx950";
#else
#error "define LSG_APP_SSSP, LSG_APP_BFS, LSG_APP_CC or LSG_APP_PR"
#endif
#if defined(LSG_APP_CC) || defined(LSG_APP_PR)
static const char *prog_args_usage = "";
#elif defined(LSG_APP_BFS)
static const char *prog_args_usage =
    "-l: enable thread block load balancer (by default false)\n"
    " -w: the first levels in one workgroup, until one is wider than it takes (small high-diameter graphs)\n"
    " -D: direction-optimising BFS over a reverse CSR built on the GPU before the timed region";
#else
static const char *prog_args_usage =
    "-l: enable thread block load balancer (by default false)\n"
    " -w: the first rounds in one workgroup, until its piles overflow (small high-diameter graphs)";
#endif
static int DIRECTION_OPT = 0;  // bfs -D
#if defined(LSG_APP_PR)
static int NO_PRINT_PAGERANK = 0;     // pr -n (apps/pr/support.cu:23-24)
static int MAX_ITERATIONS = INT_MAX;  // pr -x (apps/pr/support.cu:25)
#endif

static int QUIET = 0;
static char *INPUT = nullptr, *OUTPUT = nullptr;
static int CUDA_DEVICE = 0;
static lsg_opts OPTS;

static void usage(int, char *argv[]) {
  fprintf(stderr, "usage: %s [-q] [-g gpunum] [-o output-file] %s graph-file \n %s\n", argv[0], prog_usage,
          prog_args_usage);
}

static void die(const char *what, int rc) {
  fprintf(stderr, "%s: %s (error %d)\n", what, lsg_last_error(), rc);
  exit(1);
}

static void process_prog_opt(char c, char *optarg) {
  if (c == 'd') {
    OPTS.delta = atoi(optarg);
    if (OPTS.delta <= 0) {
      fprintf(stderr, "delta must be > 0 (got '%s')\n", optarg);
      exit(EXIT_FAILURE);
    }
  }
  if (c == 'l') OPTS.enable_lb = 1;
  if (c == 'w') OPTS.single_block = 1;  // no reference counterpart: lsg_opts.single_block
  if (c == 'D') DIRECTION_OPT = 1;
#if defined(LSG_APP_PR)
  if (c == 'n') NO_PRINT_PAGERANK = 1;
  if (c == 'x') MAX_ITERATIONS = atoi(optarg);
  if (c == 't') (void)atoi(optarg);  // accepted; the reference's top-ranks listing is disabled (support.cu:83-89)
#endif
  if (c == 's') {
    OPTS.start_node = atoi(optarg);
    if (OPTS.start_node < 0) {
      fprintf(stderr, "start node must be >= 0 (got '%s')\n", optarg);
      exit(EXIT_FAILURE);
    }
  }
}

static void parse_args(int argc, char *argv[]) {
  char opts[64];
  snprintf(opts, sizeof opts, "g:qo:%s", prog_opts);
  int c;
  while ((c = getopt(argc, argv, opts)) != -1) {
    switch (c) {
      case 'q': QUIET = 1; break;
      case 'o': OUTPUT = optarg; break;
      case 'g': {
        char *end;
        errno = 0;
        CUDA_DEVICE = (int)strtol(optarg, &end, 10);
        if (errno != 0 || *end != '\0') {
          fprintf(stderr, "Invalid GPU device '%s'. An integer must be specified.\n", optarg);
          exit(EXIT_FAILURE);
        }
        break;
      }
      case '?': usage(argc, argv); exit(EXIT_FAILURE);
      default: process_prog_opt((char)c, optarg); break;
    }
  }
  if (optind < argc) INPUT = argv[optind];
  else {
    usage(argc, argv);
    exit(EXIT_FAILURE);
  }
  if (OPTS.single_block && (OPTS.enable_lb || DIRECTION_OPT)) {
    fprintf(stderr, "-w excludes -l and -D\n");
    exit(EXIT_FAILURE);
  }
}

static uint64_t now_ns() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC_RAW, &ts);  // rt/include/Timer.h:23-26
  return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

static int load_graph_and_run_kernel(char *graph_file) {
  lsg_graph g, gg;
  fprintf(stderr, "OPTIONS: %s\n", GGC_OPTIONS);
  int rc = lsg_graph_read(&g, graph_file, 0);
  if (rc) die("reading graph", rc);
  rc = lsg_graph_copy_to_gpu(&g, &gg, CUDA_DEVICE);
  if (rc) die("copy_to_gpu", rc);
  lsg_stats st;
#if defined(LSG_APP_SSSP)
  printf("delta: %d\n", OPTS.delta);  // apps/sssp/kernel.cu:473
#endif
#if defined(LSG_APP_BFS)
  lsg_graph gt;
  memset(&gt, 0, sizeof gt);
  if (DIRECTION_OPT) {
    uint64_t p0 = now_ns();
    rc = lsg_graph_transpose(&gg, &gt);
    if (rc) die("transpose", rc);
    fprintf(stderr, "transpose: %llu us\n", (unsigned long long)((now_ns() - p0) / 1000ull));
  }
#endif
  uint64_t t0 = now_ns();
#if defined(LSG_APP_PR)
  // gg_main of apps/pr: ranks in their own float array (P_CURR), the reverse
  // CSR built inside the timed region
  float *rank = nullptr;
  lsg_graph gr;
  memset(&gr, 0, sizeof gr);
  rc = lsg_dev_malloc(CUDA_DEVICE, (size_t)gg.nnodes * sizeof(float), (void **)&rank);
  if (rc) die("allocating ranks", rc);
#endif
#if defined(LSG_APP_SSSP)
  rc = lsg_sssp_gg_main(&g, &gg, &OPTS, &st);
#elif defined(LSG_APP_CC)
  rc = lsg_cc_gg_main(&g, &gg, &st);
#elif defined(LSG_APP_PR)
  rc = lsg_graph_transpose(&gg, &gr);
  if (rc == LSG_OK) rc = lsg_pagerank(&gg, &gr, 0.85f, 0.000001f, MAX_ITERATIONS, rank, &st);  // kernel.cu:19-20
#else
  if (DIRECTION_OPT)  // lsg_bfs_do rejects a source out of range
    rc = lsg_bfs_do(&gg, &gt, (uint32_t)OPTS.start_node, 0, gg.node_data, &st);
  else
    rc = lsg_bfs_gg_main(&g, &gg, &OPTS, &st);
#endif
  if (rc == LSG_OK) rc = lsg_device_sync(CUDA_DEVICE);
  uint64_t t1 = now_ns();
  if (rc) die("gg_main", rc);
#if defined(LSG_APP_SSSP)
  printf("iterations: %d\n", (int)st.iterations);  // apps/sssp/kernel.cu:479
#elif defined(LSG_APP_CC)
  printf("components: %d\n", (int)st.items);  // apps/cc/kernel.cu:719
#elif defined(LSG_APP_PR)
  printf("PR took %d iterations\n", (int)st.iterations);  // apps/pr/kernel.cu:387
#endif
  printf("gg_main %llu %llu\n", (unsigned long long)t0, (unsigned long long)t1);  // Timer.h:77-79
  fprintf(stderr, "Total time: %llu ms\n", (unsigned long long)((t1 - t0) / 1000000ull));
  fprintf(stderr, "Total time: %llu ns\n", (unsigned long long)(t1 - t0));
#if defined(LSG_APP_PR)
  std::vector<float> hrank(gg.nnodes);
  rc = lsg_dev_copy(CUDA_DEVICE, hrank.data(), rank, hrank.size() * sizeof(float), 1);
  if (rc) die("copying ranks", rc);
  lsg_dev_free(CUDA_DEVICE, rank);
  lsg_graph_dealloc(&gr);
#else
  rc = lsg_graph_copy_to_cpu(&gg, &g);
  if (rc) die("copy_to_cpu", rc);
#endif
  if (!QUIET) {
    fflush(stdout);
#if defined(LSG_APP_CC)
    rc = lsg_output_cc(&g, OUTPUT);
#elif defined(LSG_APP_PR)
    rc = lsg_output_pr(hrank.data(), gg.nnodes, OUTPUT, NO_PRINT_PAGERANK);
#else
    rc = lsg_output(&g, OUTPUT);
#endif
    if (rc) die("output", rc);
  }
#if defined(LSG_APP_BFS)
  if (DIRECTION_OPT) lsg_graph_dealloc(&gt);
#endif
  lsg_graph_dealloc(&gg);
  lsg_graph_dealloc(&g);
  return 0;
}

int main(int argc, char *argv[]) {
  if (argc == 1) {
    usage(argc, argv);
    exit(1);
  }
  lsg_opts_default(&OPTS);
  parse_args(argc, argv);
  int rc = lsg_set_device(CUDA_DEVICE);
  if (rc) die("selecting GPU", rc);
  char name[256];
  rc = lsg_device_name(CUDA_DEVICE, name, sizeof name);
  if (rc) die("querying GPU", rc);
  printf("Using GPU: %s\n", name);
  return load_graph_and_run_kernel(INPUT);
}
