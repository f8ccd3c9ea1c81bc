// Reverse CSR (in-edges) of a device graph and its pull work items.
#pragma once

#include <stdint.h>

#include "runtime.h"

namespace lsg {

constexpr uint32_t kPullChunk = 512;  // in-edges per pull work item

struct ReverseCsr {
  const uint32_t *row;   // [V+1]
  const uint32_t *src;   // [E] source of each in-edge
  const uint32_t *w;     // [E] its weight
  const uint64_t *items; // pull work items (chunk << 32 | v)
  uint32_t nitems;
};

// Builds (first call) or returns the reverse graph of device graph g.
int get_reverse(Workspace *w, const lsg_graph *g, ReverseCsr *out);
// Releases the reverse graph of g, if any (called by lsg_graph_dealloc).
void drop_reverse(const lsg_graph *g);

}  // namespace lsg
