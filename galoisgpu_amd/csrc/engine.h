// Device-side building blocks shared by the single-source and batched
// traversal kernels: control block layout, coherent loads/stores, wave64
// helpers, wave-aggregated queue pushes and the leader grid barrier.
//
// Coherence on gfx950.  Each XCD has its own L2, and an agent-scope
// release/acquire fence writes back and invalidates that L2 -- done by every
// block at every grid barrier it costs ~0.25 us per block per barrier.  So the
// kernels never rely on fences for visibility: every word that one block
// writes and another block reads inside the same launch (queue items, the
// distances read at pop time, schedule words, counters) is accessed with
// relaxed agent-scope atomics, which the hardware writes and reads through to
// the coherence point.  Fences are used only for ordering: before a block
// signals the barrier it waits for its own memory operations to complete
// (s_waitcnt vmcnt(0)), and after it leaves it keeps later loads below the
// spin.  The CSR arrays are read-only during a launch and use plain loads.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace lsg {

constexpr uint32_t kInf = 0x7FFFFFFFu;  // INT_MAX, apps/sssp/kernel.cu:18
constexpr int kWave = 64;

enum AbortCode : uint32_t { kAbortNone = 0, kAbortTimeout = 1, kAbortCapacity = 2, kAbortResidency = 3 };

// ---- coherent memory access ----------------------------------------------
//
// Every pointer these helpers take is device (global) memory.  They say so
// with an address-space cast: through a plain pointer the compiler cannot
// tell global from LDS memory and emits flat instructions, which count
// against the LDS wait counter too, so that every LDS wait of a wave also
// waits for its outstanding global loads and atomics.

#define LSG_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ LSG_GLOBAL T *glb(T *p) {
  return (LSG_GLOBAL T *)p;
}
template <class T>
__device__ __forceinline__ const LSG_GLOBAL T *glb(const T *p) {
  return (const LSG_GLOBAL T *)p;
}

__device__ __forceinline__ uint32_t ld_coh(const uint32_t *p) {
  return __hip_atomic_load(glb(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_coh64(const uint64_t *p) {
  return __hip_atomic_load(glb(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh(uint32_t *p, uint32_t v) {
  __hip_atomic_store(glb(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh64(uint64_t *p, uint64_t v) {
  __hip_atomic_store(glb(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_add(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_add(glb(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_min(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_min(glb(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_exch(uint32_t *p, uint32_t v) {
  return __hip_atomic_exchange(glb(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t atomic_or64(uint64_t *p, uint64_t v) {
  return __hip_atomic_fetch_or(glb(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t atomic_exch64(uint64_t *p, uint64_t v) {
  return __hip_atomic_exchange(glb(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Orders this wave's earlier memory operations before its later ones: the
// compiler may not move accesses across it and the wave waits until every
// vector memory operation it issued has completed.
__device__ __forceinline__ void order_release() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));  // vmcnt(0)
}
__device__ __forceinline__ void order_acquire() {
  __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// ---- schedule and control block ------------------------------------------

// Schedule for the next round, written by the barrier leader and read by
// every block after the barrier.  Replaces PipeContextT's by-value pile
// indices (rt/include/pipe.h:29-87) with one authoritative copy.  All words
// are u32 so it can be copied with coherent loads.
struct Sched {
  uint32_t in_q, near_q, far_q;  // roles of the three item queues
  uint32_t in_count;             // items in the in-queue this round
  uint32_t heavy_in;             // heavy chunk queue consumed this round (0/1)
  uint32_t heavy_count;          // chunks to consume this round
  uint32_t round;                // 1-based round number
  uint32_t phase;                // 0-based delta phase
  uint32_t curdelta_lo, curdelta_hi;  // near/far threshold (apps/sssp/kernel.cu:134)
  uint32_t done;
  uint32_t delta;                // width of the next phase step
  uint32_t aux[4];               // kernel-specific words
  __device__ __host__ int64_t curdelta() const { return (int64_t)(((uint64_t)curdelta_hi << 32) | curdelta_lo); }
  __device__ __host__ void set_curdelta(int64_t v) {
    curdelta_lo = (uint32_t)(uint64_t)v;
    curdelta_hi = (uint32_t)((uint64_t)v >> 32);
  }
};
constexpr int kSchedWords = sizeof(Sched) / 4;
static_assert(sizeof(Sched) == 64, "Sched is 16 words");

// All words are loaded before any is stored: `s` is often in LDS, and with a
// store after each load the compiler, unable to rule out that the store
// changes a later word, waited for every load in turn (16 round trips, ~2.5
// us of every leader step, instead of one).
__device__ __forceinline__ void sched_load(Sched &s, const Sched *src) {
  const uint32_t *p = reinterpret_cast<const uint32_t *>(src);
  uint32_t *d = reinterpret_cast<uint32_t *>(&s);
  uint32_t w[kSchedWords];
#pragma unroll
  for (int i = 0; i < kSchedWords; ++i) w[i] = ld_coh(p + i);
#pragma unroll
  for (int i = 0; i < kSchedWords; ++i) d[i] = w[i];
}
__device__ __forceinline__ void sched_store(Sched *dst, const Sched &s) {
  uint32_t *p = reinterpret_cast<uint32_t *>(dst);
  const uint32_t *d = reinterpret_cast<const uint32_t *>(&s);
#pragma unroll
  for (int i = 0; i < kSchedWords; ++i) st_coh(p + i, d[i]);
}

// Grid barrier arrival tree: blocks arrive in groups of kBarGroup on a
// group counter; the last of each group arrives on the top counter.  One
// counter for every block costs ~25 us per barrier at 2048 blocks on gfx950
// (same-address atomics serialise), the two-level tree ~3 us.
constexpr uint32_t kBarGroup = 32;
constexpr uint32_t kMaxBarGroups = 128;
constexpr uint32_t kMaxGridBlocks = kBarGroup * kMaxBarGroups;

// Device control block.  Each hot word sits on its own 128-byte line.  The
// host copies back the part before bar_grp (kCtrlHostBytes) after a launch.
struct Ctrl {
  uint32_t bar_count; uint32_t _p0[31];
  uint32_t bar_gen;   uint32_t _p1[31];
  uint32_t qcount[3]; uint32_t _p2[29];
  uint32_t hcount[2]; uint32_t _p3[30];
  uint32_t ticket;    uint32_t ticket2; uint32_t _p4[30];
  uint32_t abort;     uint32_t checkin; uint32_t resident_seen; uint32_t progress; uint32_t _p5[28];
  Sched s;            uint32_t _p6[16];
  uint64_t stat_items, stat_rounds, stat_heavy, stat_phases, stat_edges;
  uint32_t _p7[22];
  uint32_t far_min;   uint32_t _p8[31];  // bucket far mode: smallest distance sent to the far pile this phase
  uint32_t bar_grp[kMaxBarGroups][32];  // group arrival counters, one line each
};
constexpr size_t kCtrlHostBytes = offsetof(Ctrl, bar_grp);

// Zeroes the barrier counters; run by the init kernels before each launch
// (tid / nthreads: the caller's grid-stride position).
__device__ __forceinline__ void ctrl_reset_barrier(Ctrl *c, size_t tid, size_t nthreads) {
  for (size_t g = tid; g < kMaxBarGroups; g += nthreads) c->bar_grp[g][0] = 0;
  if (tid == 0) {
    c->bar_count = 0;
    c->bar_gen = 0;
  }
}

// Queue counters and statistics words of the control block, loaded together
// by a barrier leader.
struct CtrlCounts {
  uint32_t qc[3], hc[2];
  uint64_t items, rounds, heavy, phases, edges;
  uint32_t far_min;
  __device__ uint32_t q(uint32_t i) const { return i == 0 ? qc[0] : i == 1 ? qc[1] : qc[2]; }
  __device__ uint32_t h(uint32_t i) const { return i == 0 ? hc[0] : hc[1]; }
};
__device__ __forceinline__ void counts_load(CtrlCounts &k, const Ctrl *c) {
  k.qc[0] = ld_coh(&c->qcount[0]);
  k.qc[1] = ld_coh(&c->qcount[1]);
  k.qc[2] = ld_coh(&c->qcount[2]);
  k.hc[0] = ld_coh(&c->hcount[0]);
  k.hc[1] = ld_coh(&c->hcount[1]);
  k.items = ld_coh64(&c->stat_items);
  k.rounds = ld_coh64(&c->stat_rounds);
  k.heavy = ld_coh64(&c->stat_heavy);
  k.phases = ld_coh64(&c->stat_phases);
  k.edges = ld_coh64(&c->stat_edges);
  k.far_min = ld_coh(&c->far_min);
}

// ---- wave64 helpers --------------------------------------------------------

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

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
    if (pos < cap) st_coh(q + pos, item);
    else st_coh(abort_flag, kAbortCapacity);
  }
}

// Reserves `n` consecutive slots per lane (n may differ per lane); returns the
// first slot for this lane.  One atomic per wave; call with the whole wave.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t n, uint32_t *tail) {
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

// ---- grid barrier ----------------------------------------------------------

// One block's arrival (thread 0): true in the block that completes the barrier.
// The counters only ever grow: a group's barrier completes when its count
// reaches a multiple of the group size, the top level when its count reaches
// a multiple of the number of groups.  Arrivals of the next barrier cannot
// start before this one's release, so no reset is needed.
__device__ __forceinline__ bool barrier_arrive(Ctrl *c, uint32_t nblocks) {
  const uint32_t g = blockIdx.x / kBarGroup;
  const uint32_t gsize = min(kBarGroup, nblocks - g * kBarGroup);
  const uint32_t ngroups = (nblocks + kBarGroup - 1) / kBarGroup;
  if ((atomic_add(&c->bar_grp[g][0], 1u) + 1) % gsize != 0) return false;
  return (atomic_add(&c->bar_count, 1u) + 1) % ngroups == 0;
}

// Leader grid barrier.  Blocks arrive through barrier_arrive; every thread of
// the last block to arrive runs `leader()` (which may use __syncthreads and
// LDS, so leader work such as scheduler inference is spread over the block)
// and then bumps the generation; the others spin on the generation.  Replaces
// GlobalBarrier::Sync (rt/include/gbar.cuh:75-121).  The spin is bounded by
// `spin_limit` ticks of the constant-rate wall clock without progress; when
// exceeded, or when another block has aborted, the block raises kAbortTimeout
// and returns so every wave exits.
template <class Leader>
__device__ __forceinline__ void grid_barrier_block(Ctrl *c, uint32_t nblocks, uint64_t spin_limit, Leader leader) {
  __shared__ uint32_t s_lead, s_gen;
  order_release();
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t gen = ld_coh(&c->bar_gen);
    s_lead = barrier_arrive(c, nblocks);
    s_gen = gen;
  }
  __syncthreads();
  if (s_lead) {
    order_acquire();
    leader();
    order_release();
    __syncthreads();
    if (threadIdx.x == 0) st_coh(&c->bar_gen, s_gen + 1);
  } else if (threadIdx.x == 0) {
    uint32_t gen = s_gen;
    uint64_t t0 = wall_clock64();
    uint32_t seen = ld_coh(&c->progress);
    while (ld_coh(&c->bar_gen) == gen) {
      if (ld_coh(&c->abort) != kAbortNone) break;
      if (wall_clock64() - t0 > spin_limit) {
        // The leader block may be running rounds on its own (it bumps
        // `progress` each round): keep waiting while it advances.
        uint32_t p = ld_coh(&c->progress);
        if (p != seen) {
          seen = p;
          t0 = wall_clock64();
          continue;
        }
        st_coh(&c->abort, kAbortTimeout);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  order_acquire();
  __syncthreads();
}

// Residency handshake, run once at kernel start.  A persistent grid is only
// correct if every block is resident at once; the runtime's occupancy figure
// has proven optimistic, so each block checks in and waits (bounded by
// `limit` ticks) for the others.  If they do not all arrive, the first block
// to give up records how many had and raises kAbortResidency; blocks that
// start later see the flag and leave.  Returns false when the launch must
// be abandoned.
__device__ __forceinline__ bool residency_handshake(Ctrl *c, uint32_t nblocks, uint64_t limit) {
  __shared__ uint32_t ok;
  if (threadIdx.x == 0) {
    ok = 1;
    if (ld_coh(&c->abort) != kAbortNone) {
      ok = 0;
    } else {
      uint32_t arrived = atomic_add(&c->checkin, 1u) + 1;
      uint64_t t0 = wall_clock64();
      while (arrived < nblocks) {
        if (ld_coh(&c->abort) != kAbortNone) {
          ok = 0;
          break;
        }
        if (wall_clock64() - t0 > limit) {
          // first to give up records the count it saw
          uint32_t expect = kAbortNone;
          if (__hip_atomic_compare_exchange_strong(&c->abort, &expect, (uint32_t)kAbortResidency,
                                                   __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            st_coh(&c->resident_seen, arrived);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        arrived = ld_coh(&c->checkin);
      }
    }
  }
  __syncthreads();
  return ok != 0;
}

}  // namespace lsg
