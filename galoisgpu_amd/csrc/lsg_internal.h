// Internal helpers shared by the host translation units of liblsg.
#pragma once

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <thread>
#include <vector>

#include "lsg.h"

namespace lsg {

// Records a printf-style message for lsg_last_error() and returns `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Number of host worker threads: LSG_THREADS if set, else the hardware count
// capped at 16.
unsigned host_threads();

// Static-chunked parallel loop over [0, n) on host threads.
template <class F>
void parallel_for(size_t n, F f, size_t min_chunk = 1 << 14) {
  unsigned nt = host_threads();
  if (n < 2 * min_chunk || nt <= 1) {
    f(size_t(0), n, 0u);
    return;
  }
  size_t chunks = (n + min_chunk - 1) / min_chunk;
  if (chunks < nt) nt = (unsigned)chunks;
  std::vector<std::thread> ts;
  ts.reserve(nt);
  for (unsigned t = 0; t < nt; ++t) {
    size_t lo = n * t / nt, hi = n * (t + 1) / nt;
    ts.emplace_back([=, &f] { f(lo, hi, t); });
  }
  for (auto &th : ts) th.join();
}

// Counter-based random streams: every draw is a pure function of
// (seed, stream id, index), so generation parallelises and reproduces exactly.
static inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
static inline uint64_t hash2(uint64_t seed, uint64_t a) { return splitmix64(seed ^ splitmix64(a)); }
static inline uint64_t hash3(uint64_t seed, uint64_t a, uint64_t b) {
  return splitmix64(hash2(seed, a) ^ (b * 0xD6E8FEB86659FD93ull));
}
// Uniform integer in [lo, hi] from the top 32 bits of h (multiply-shift).
static inline uint32_t uniform_u32(uint64_t h, uint32_t lo, uint32_t hi) {
  uint64_t range = (uint64_t)hi - lo + 1;
  return lo + (uint32_t)(((h >> 32) * range) >> 32);
}
// Edge weight as a function of the directed edge's identity.
static inline uint32_t edge_weight(uint64_t seed, uint32_t u, uint32_t v, uint32_t lo, uint32_t hi) {
  return uniform_u32(hash3(seed ^ 0x5745494748545321ull, u, v), lo, hi);
}

// Validates a host CSR: offsets monotone, row_start[0] == 0,
// row_start[n] == nedges, every dst < nnodes, every weight <= INT_MAX.
int validate_host_csr(const lsg_graph *g);

// Builds a CSR from a directed pair list (sorted rows; self loops and
// duplicates removed when asked) and draws weights with edge_weight().
int build_csr(lsg_graph *g, uint32_t nnodes, std::vector<uint32_t> &src, std::vector<uint32_t> &dst,
              bool dedupe, uint32_t wlo, uint32_t whi, uint64_t wseed);

}  // namespace lsg
