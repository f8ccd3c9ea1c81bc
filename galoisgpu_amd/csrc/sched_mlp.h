// Learned phase scheduler: features, model layout and device inference.
//
// At every phase boundary the barrier leader describes the frontier with 16
// features and an FP32 MLP (16 -> 64 -> 64 -> 8, ReLU) picks one of 8 actions:
// a delta width (a multiple of the base delta) and an expansion setting.  The
// choice only changes which vertices are expanded when; every relaxation
// still uses the same u32 arithmetic, so distances are identical whatever it
// picks.  No reference counterpart: the reference fixes delta per run
// (apps/sssp/support.cu:10,20-23).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsg {

constexpr int kFeat = 16, kHid = 64, kAct = 8, kMaxForced = 256;

// Model buffer layout (floats).
constexpr int kOffMean = 0;
constexpr int kOffInvStd = kOffMean + kFeat;
constexpr int kOffW1 = kOffInvStd + kFeat;    // [kFeat][kHid]
constexpr int kOffB1 = kOffW1 + kFeat * kHid;
constexpr int kOffW2 = kOffB1 + kHid;         // [kHid][kHid]
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;         // [kHid][kAct]
constexpr int kOffB3 = kOffW3 + kHid * kAct;
constexpr int kOffDeltaMult = kOffB3 + kAct;  // delta multiplier per action
// Mode per action: the far-pile mode (LSG_FAR_*) in a batched-kernel model,
// the expansion mode (LSG_EXPAND_*) in a plain-kernel model.
constexpr int kOffFarMode = kOffDeltaMult + kAct;
constexpr int kOffForced = kOffFarMode + kAct;       // forced action per phase (< 0: use the MLP)
constexpr int kModelFloats = kOffForced + kMaxForced;
// Per-phase feature record: 16 features, action (-1 fixed), delta, far mode,
// first round of the phase.
constexpr int kFeatRec = kFeat + 4;
constexpr int kNetFloats = kOffDeltaMult - kOffW1;   // trainable parameters

// Raw counters the leader has at a phase boundary.
struct SchedFeatures {
  uint32_t in_items;    // size of the far pile that becomes the next phase
  uint32_t phase, round;
  float curdelta;
  uint32_t base_delta, prev_delta;
  float lane_items;     // (vertex, source) expansions so far
  float edges;          // edge expansions so far
  uint32_t nnodes;
  float nedges;
  uint32_t nsrc;
};

__host__ __device__ inline float lsg_log2p1(float x) { return __builtin_log2f(1.0f + x); }

// The 16 features.  Kept in one place: the trace writer, the trainer's data
// and the leader all call this.
__host__ __device__ inline void sched_featurize(const SchedFeatures &f, float x[kFeat]) {
  float n = (float)(f.nnodes ? f.nnodes : 1), s = (float)(f.nsrc ? f.nsrc : 1);
  float E = f.nedges > 1.0f ? f.nedges : 1.0f;
  x[0] = lsg_log2p1((float)f.in_items);                 // frontier size
  x[1] = (float)f.in_items / n;                         // frontier fraction
  x[2] = f.lane_items / (n * s);                        // settled fraction proxy
  x[3] = f.edges / E;                                   // edge sweeps so far
  x[4] = (float)f.phase;
  x[5] = (float)f.round;
  x[6] = f.curdelta / (float)(f.base_delta ? f.base_delta : 1);
  x[7] = (float)f.prev_delta / (float)(f.base_delta ? f.base_delta : 1);
  x[8] = lsg_log2p1(n);
  x[9] = lsg_log2p1(E);
  x[10] = E / n;                                        // mean degree
  x[11] = s;
  x[12] = f.lane_items > 0 ? f.edges / f.lane_items : 0.0f;  // edges per expansion
  x[13] = lsg_log2p1((float)f.base_delta);
  x[14] = f.round > 0 ? (float)f.phase / (float)f.round : 0.0f;  // rounds per phase
  x[15] = 1.0f;
}

// Forward pass for one sample on one thread; returns the argmax action.
__host__ __device__ inline int mlp_forward_argmax(const float *m, const float xin[kFeat]) {
  float x[kFeat], h1[kHid], h2[kHid];
  for (int i = 0; i < kFeat; ++i) x[i] = (xin[i] - m[kOffMean + i]) * m[kOffInvStd + i];
  for (int j = 0; j < kHid; ++j) {
    float acc = m[kOffB1 + j];
    for (int i = 0; i < kFeat; ++i) acc += x[i] * m[kOffW1 + i * kHid + j];
    h1[j] = acc > 0.0f ? acc : 0.0f;
  }
  for (int j = 0; j < kHid; ++j) {
    float acc = m[kOffB2 + j];
    for (int i = 0; i < kHid; ++i) acc += h1[i] * m[kOffW2 + i * kHid + j];
    h2[j] = acc > 0.0f ? acc : 0.0f;
  }
  int best = 0;
  float bv = -3.0e38f;
  for (int k = 0; k < kAct; ++k) {
    float acc = m[kOffB3 + k];
    for (int i = 0; i < kHid; ++i) acc += h2[i] * m[kOffW3 + i * kAct + k];
    if (acc > bv) {
      bv = acc;
      best = k;
    }
  }
  return best;
}

// LDS scratch for block-cooperative inference.
struct MlpScratch {
  float x[kFeat], h1[kHid], h2[kHid], z[kAct];
  int action;
};

// Forward pass spread over the block (needs >= kHid threads; every thread of
// the block must call it).  Uses the forced schedule when one is set: entry
// min(phase, kMaxForced - 1), so the last entry holds for every later phase.
// Returns the action in every thread.
__device__ inline int mlp_block_action(const float *m, const SchedFeatures &f, MlpScratch &sc) {
  const int t = threadIdx.x;
  const uint32_t fp = f.phase < (uint32_t)kMaxForced ? f.phase : (uint32_t)kMaxForced - 1;
  bool forced = m[kOffForced + fp] >= 0.0f;
  if (t == 0) {
    float x[kFeat];
    sched_featurize(f, x);
#pragma unroll 1
    for (int i = 0; i < kFeat; ++i) sc.x[i] = (x[i] - m[kOffMean + i]) * m[kOffInvStd + i];
    sc.action = forced ? (int)m[kOffForced + fp] : -1;
  }
  __syncthreads();
  if (!forced) {
    if (t < kHid) {
      float acc = m[kOffB1 + t];
#pragma unroll 2
      for (int i = 0; i < kFeat; ++i) acc += sc.x[i] * m[kOffW1 + i * kHid + t];
      sc.h1[t] = acc > 0.0f ? acc : 0.0f;
    }
    __syncthreads();
    if (t < kHid) {
      float acc = m[kOffB2 + t];
#pragma unroll 2
      for (int i = 0; i < kHid; ++i) acc += sc.h1[i] * m[kOffW2 + i * kHid + t];
      sc.h2[t] = acc > 0.0f ? acc : 0.0f;
    }
    __syncthreads();
    if (t < kAct) {
      float acc = m[kOffB3 + t];
#pragma unroll 2
      for (int i = 0; i < kHid; ++i) acc += sc.h2[i] * m[kOffW3 + i * kAct + t];
      sc.z[t] = acc;
    }
    __syncthreads();
    if (t == 0) {
      int best = 0;
      for (int k = 1; k < kAct; ++k)
        if (sc.z[k] > sc.z[best]) best = k;
      sc.action = best;
    }
  }
  __syncthreads();
  return sc.action;
}

__host__ __device__ inline uint32_t action_delta(const float *m, int act, uint32_t base_delta) {
  float d = m[kOffDeltaMult + act] * (float)base_delta;
  if (d < 1.0f) d = 1.0f;
  if (d > 2147483647.0f) d = 2147483647.0f;
  return (uint32_t)d;
}

// Device copy of the scheduler model for `device`, or null if none is set.
// Device copy of the model of `kind` (LSG_SCHED_*) loaded on `device`, or null.
const float *sched_model_device(int device, int kind);

}  // namespace lsg
