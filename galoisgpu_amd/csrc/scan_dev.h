// Device-wide exclusive prefix sum of u32 values in three launches: a scan
// of each tile of kScanTile values (scan_tile_kernel, one block per tile),
// an exclusive scan of the tile totals in one block (scan_sums_kernel), and
// an add of each tile's offset that the caller fuses with its own epilogue.
// Used by the reverse-CSR build (row starts from in-degree histograms) and
// by the -l pass (prefix of heavy vertices' degrees, the reference's
// cub::DeviceScan::InclusiveSum at rt/include/thread_work.h:27-41).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine.h"

namespace lsg {

constexpr int kScanBlock = 256;
constexpr uint32_t kScanTile = 4 * kScanBlock;  // elements per block of the scan kernels

// Exclusive scan of one tile of kScanTile values per block (4 per thread):
// out[i] = sum of in[tile start .. i), tile_sum[b] = the tile's total.
static __global__ __launch_bounds__(kScanBlock) void scan_tile_kernel(const uint32_t *in, uint32_t n, uint32_t *out,
                                                             uint32_t *tile_sum) {
  __shared__ uint32_t wsum[kScanBlock / kWave];
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
  if (threadIdx.x == kScanBlock - 1) tile_sum[blockIdx.x] = before + inc;
}

// One block: exclusive scan of the tile totals in place; *total gets the sum.
static __global__ __launch_bounds__(kScanBlock) void scan_sums_kernel(uint32_t *sums, uint32_t n, uint32_t *total) {
  __shared__ uint32_t wsum[kScanBlock / kWave];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
  for (uint32_t b0 = 0; b0 < n; b0 += kScanBlock) {
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
    if (threadIdx.x == kScanBlock - 1) carry = before + inc;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

}  // namespace lsg
