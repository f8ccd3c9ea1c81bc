// Device-side building blocks shared by the single-source and batched
// traversal kernels: control block layout, wave64 helpers, wave-aggregated
// queue pushes and the leader grid barrier.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsg {

constexpr uint32_t kInf = 0x7FFFFFFFu;  // INT_MAX, apps/sssp/kernel.cu:18
constexpr int kWave = 64;

enum AbortCode : uint32_t { kAbortNone = 0, kAbortTimeout = 1, kAbortCapacity = 2 };

// Schedule for the next round, written by the barrier leader and read by
// every block after the barrier.  Replaces PipeContextT's by-value pile
// indices (rt/include/pipe.h:29-87) with one authoritative copy.
struct Sched {
  uint32_t in_q, near_q, far_q;  // roles of the three item queues
  uint32_t in_count;             // items in the in-queue this round
  uint32_t heavy_in;             // heavy chunk queue consumed this round (0/1)
  uint32_t heavy_count;          // chunks to consume this round
  uint32_t round;                // 1-based round number (BFS: the LEVEL written)
  uint32_t phase;                // 0-based delta phase
  int64_t curdelta;              // near/far threshold (apps/sssp/kernel.cu:134)
  uint32_t done;
  uint32_t ntiles;               // in_count rounded up to blocks
  uint32_t delta;                // width used for the next phase increment
  uint32_t pad[3];
};

// Device control block.  Each hot word sits on its own 128-byte line.
struct Ctrl {
  uint32_t bar_count; uint32_t _p0[31];
  uint32_t bar_gen;   uint32_t _p1[31];
  uint32_t qcount[3]; uint32_t _p2[29];
  uint32_t hcount[2]; uint32_t _p3[30];
  uint32_t ticket;    uint32_t _p4[31];
  uint32_t abort;     uint32_t _p5[31];
  Sched s;            uint32_t _p6[16];
  uint64_t stat_items, stat_rounds, stat_heavy, stat_phases;
  uint32_t _p7[24];
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t ld_acquire(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_add(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_min(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_exch(uint32_t *p, uint32_t v) {
  return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_cas(uint32_t *p, uint32_t expect, uint32_t v) {
  __hip_atomic_compare_exchange_strong(p, &expect, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expect;
}
__device__ __forceinline__ uint64_t atomic_or64(uint64_t *p, uint64_t v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t atomic_exch64(uint64_t *p, uint64_t v) {
  return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Number of set bits of `mask` below this lane (v_mbcnt).
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t readlane(uint32_t v, uint32_t l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}

// Wave-aggregated push: one atomicAdd per wave on the queue tail, every
// active lane with `pred` writes its item at base + rank.  Replaces
// Worklist::setup_push_warp_one + do_push (rt/include/worklist.h:348-368,414-420)
// with 64-lane ballot/mbcnt.  Overflow raises kAbortCapacity instead of
// asserting (worklist.h:356); nothing is silently dropped.
__device__ __forceinline__ void wave_push(bool pred, uint32_t item, uint32_t *q, uint32_t *tail, uint32_t cap,
                                          uint32_t *abort_flag) {
  uint64_t m = __ballot(pred);
  if (m == 0) return;
  uint32_t leader = (uint32_t)__builtin_ctzll(m);
  uint32_t base = 0;
  if (lane_id() == leader) base = atomic_add(tail, (uint32_t)__builtin_popcountll(m));
  base = readlane(base, leader);
  if (pred) {
    uint32_t pos = base + lanes_below(m);
    if (pos < cap) q[pos] = item;
    else st_relaxed(abort_flag, kAbortCapacity);
  }
}

// Reserves `n` consecutive slots per lane (n may differ per lane); returns the
// first slot for this lane.  One atomic per wave.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t n, uint32_t *tail) {
  // inclusive scan of n over the active lanes via shuffles
  uint32_t lane = lane_id();
  uint32_t x = n;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    uint32_t y = __shfl_up(x, off);
    if (lane >= (uint32_t)off) x += y;
  }
  uint32_t total = __shfl(x, kWave - 1);
  uint32_t base = 0;
  if (lane == 0 && total) base = atomic_add(tail, total);
  base = __shfl(base, 0);
  return base + x - n;
}

// Leader grid barrier.  Every block arrives on one counter; the last to
// arrive runs `leader()` (single thread, sees all writes made before the
// barrier), resets the counter and bumps the generation; the others spin on
// the generation with acquire loads.  Replaces GlobalBarrier::Sync
// (rt/include/gbar.cuh:75-121), whose volatile/LOAD_CG idiom is not enough on
// gfx950, where each XCD has its own L2: arrival is a release and departure
// an acquire at agent scope.  The spin is bounded by `spin_limit` ticks of the
// constant-rate wall clock; when exceeded, or when another block has aborted,
// the block raises kAbortTimeout and returns so that every wave exits.
template <class Leader>
__device__ __forceinline__ void grid_barrier(Ctrl *c, uint32_t nblocks, uint64_t spin_limit, Leader leader) {
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t gen = ld_relaxed(&c->bar_gen);
    uint32_t arrived = __hip_atomic_fetch_add(&c->bar_count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == nblocks - 1) {
      leader();
      st_relaxed(&c->bar_count, 0u);
      st_release(&c->bar_gen, gen + 1);
    } else {
      uint64_t t0 = wall_clock64();
      while (ld_acquire(&c->bar_gen) == gen) {
        if (ld_relaxed(&c->abort) != kAbortNone) break;
        if (wall_clock64() - t0 > spin_limit) {
          st_relaxed(&c->abort, kAbortTimeout);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
  }
  __syncthreads();
}

}  // namespace lsg
