// Learned phase scheduler: model storage, batched inference and training
// kernels for the 16 -> 64 -> 64 -> 8 FP32 MLP described in sched_mlp.h.
// No reference counterpart (the reference's tuning is an offline OpenTuner
// search over BLKSIZE, tuning/lsgtuner.py:150-207).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "lsg_internal.h"
#include "runtime.h"
#include "sched_mlp.h"

namespace lsg {

// Loaded models, keyed by (device, kind); each lives until the process ends.
struct ModelSlot {
  int device, kind;
  float *dptr;
};
static std::mutex g_model_mu;
static std::vector<ModelSlot> g_models;

const float *sched_model_device(int device, int kind) {
  std::lock_guard<std::mutex> lk(g_model_mu);
  for (auto &e : g_models)
    if (e.device == device && e.kind == kind) return e.dptr;
  return nullptr;
}

// One thread per sample.
__global__ void mlp_infer_kernel(const float *m, const float *X, uint32_t n, int32_t *out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = mlp_forward_argmax(m, X + (size_t)i * kFeat);
}

constexpr int kTrainThreads = 256;
constexpr int kBatch = 64;

__device__ __forceinline__ uint32_t train_hash(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return (uint32_t)x;
}

// Minibatch SGD with momentum on weighted softmax cross-entropy.  One
// workgroup; parameters, momentum and activations live in LDS.  X holds
// normalised features.  `model` (global) is read at start and written back.
__global__ __launch_bounds__(kTrainThreads) void mlp_train_kernel(float *model, const float *X, const int32_t *y,
                                                                  const float *wt, uint32_t n, uint32_t steps,
                                                                  float lr, float momentum, uint64_t seed,
                                                                  float *loss_out) {
  extern __shared__ float lds[];
  float *P = lds;                        // kNetFloats parameters (W1 b1 W2 b2 W3 b3)
  float *M = P + kNetFloats;             // momentum
  float *xb = M + kNetFloats;            // [kBatch][kFeat]
  float *h1 = xb + kBatch * kFeat;       // [kBatch][kHid]
  float *h2 = h1 + kBatch * kHid;        // [kBatch][kHid]
  float *g2 = h2 + kBatch * kHid;        // [kBatch][kHid] dL/d(pre-act 2), then reused
  float *g1 = g2 + kBatch * kHid;        // [kBatch][kHid]
  float *dl = g1 + kBatch * kHid;        // [kBatch][kAct]
  float *bw = dl + kBatch * kAct;        // [kBatch] sample weight
  int *by = (int *)(bw + kBatch);        // [kBatch] labels
  float *red = (float *)(by + kBatch);   // [kTrainThreads] loss reduction
  const int t = threadIdx.x;
  float *W1 = P, *B1 = W1 + kFeat * kHid, *W2 = B1 + kHid, *B2 = W2 + kHid * kHid, *W3 = B2 + kHid,
        *B3 = W3 + kHid * kAct;
  for (int i = t; i < kNetFloats; i += kTrainThreads) {
    P[i] = model[kOffW1 + i];
    M[i] = 0.0f;
  }
  float loss_acc = 0.0f, wacc = 0.0f;
  __syncthreads();
  for (uint32_t step = 0; step < steps; ++step) {
    // sample a minibatch (with replacement)
    if (t < kBatch) {
      uint32_t idx = train_hash(seed * 0x9E3779B97F4A7C15ull + (uint64_t)step * kBatch + t) % n;
      by[t] = y[idx];
      bw[t] = wt ? wt[idx] : 1.0f;
      for (int f = 0; f < kFeat; ++f) xb[t * kFeat + f] = X[(size_t)idx * kFeat + f];
    }
    __syncthreads();
    // layer 1
    for (int p = t; p < kBatch * kHid; p += kTrainThreads) {
      int b = p / kHid, j = p % kHid;
      float acc = B1[j];
      for (int i = 0; i < kFeat; ++i) acc += xb[b * kFeat + i] * W1[i * kHid + j];
      h1[p] = acc > 0.0f ? acc : 0.0f;
    }
    __syncthreads();
    // layer 2
    for (int p = t; p < kBatch * kHid; p += kTrainThreads) {
      int b = p / kHid, j = p % kHid;
      float acc = B2[j];
      for (int i = 0; i < kHid; ++i) acc += h1[b * kHid + i] * W2[i * kHid + j];
      h2[p] = acc > 0.0f ? acc : 0.0f;
    }
    __syncthreads();
    // logits, softmax, loss, dL/dlogits
    float wsum = 0.0f;
    for (int b = 0; b < kBatch; ++b) wsum += bw[b];
    wsum = wsum > 0.0f ? wsum : 1.0f;
    if (t < kBatch) {
      float z[kAct], mx = -3.0e38f;
      for (int k = 0; k < kAct; ++k) {
        float acc = B3[k];
        for (int i = 0; i < kHid; ++i) acc += h2[t * kHid + i] * W3[i * kAct + k];
        z[k] = acc;
        mx = acc > mx ? acc : mx;
      }
      float se = 0.0f;
      for (int k = 0; k < kAct; ++k) {
        z[k] = expf(z[k] - mx);
        se += z[k];
      }
      float scale = bw[t] / wsum;
      for (int k = 0; k < kAct; ++k) dl[t * kAct + k] = (z[k] / se - (k == by[t] ? 1.0f : 0.0f)) * scale;
      red[t] = -logf(z[by[t]] / se + 1e-30f) * bw[t];
    } else {
      red[t] = 0.0f;
    }
    __syncthreads();
    if (step + 100 >= steps) {
      if (t == 0) {
        float s = 0.0f;
        for (int b = 0; b < kBatch; ++b) s += red[b];
        loss_acc += s;
        wacc += wsum;
      }
    }
    // dh2 (through ReLU) using W3 before its update
    for (int p = t; p < kBatch * kHid; p += kTrainThreads) {
      int b = p / kHid, i = p % kHid;
      float acc = 0.0f;
      for (int k = 0; k < kAct; ++k) acc += dl[b * kAct + k] * W3[i * kAct + k];
      g2[p] = h2[p] > 0.0f ? acc : 0.0f;
    }
    __syncthreads();
    // W3, B3 update
    for (int p = t; p < kHid * kAct + kAct; p += kTrainThreads) {
      float g = 0.0f;
      if (p < kHid * kAct) {
        int i = p / kAct, k = p % kAct;
        for (int b = 0; b < kBatch; ++b) g += h2[b * kHid + i] * dl[b * kAct + k];
      } else {
        int k = p - kHid * kAct;
        for (int b = 0; b < kBatch; ++b) g += dl[b * kAct + k];
      }
      float *w = p < kHid * kAct ? W3 + p : B3 + (p - kHid * kAct);
      float *mm = M + (w - P);
      *mm = momentum * *mm + g;
      *w -= lr * *mm;
    }
    // dh1 using W2 before its update
    for (int p = t; p < kBatch * kHid; p += kTrainThreads) {
      int b = p / kHid, i = p % kHid;
      float acc = 0.0f;
      for (int j = 0; j < kHid; ++j) acc += g2[b * kHid + j] * W2[i * kHid + j];
      g1[p] = h1[p] > 0.0f ? acc : 0.0f;
    }
    __syncthreads();
    // W2, B2 update
    for (int p = t; p < kHid * kHid + kHid; p += kTrainThreads) {
      float g = 0.0f;
      if (p < kHid * kHid) {
        int i = p / kHid, j = p % kHid;
        for (int b = 0; b < kBatch; ++b) g += h1[b * kHid + i] * g2[b * kHid + j];
      } else {
        int j = p - kHid * kHid;
        for (int b = 0; b < kBatch; ++b) g += g2[b * kHid + j];
      }
      float *w = p < kHid * kHid ? W2 + p : B2 + (p - kHid * kHid);
      float *mm = M + (w - P);
      *mm = momentum * *mm + g;
      *w -= lr * *mm;
    }
    // W1, B1 update
    for (int p = t; p < kFeat * kHid + kHid; p += kTrainThreads) {
      float g = 0.0f;
      if (p < kFeat * kHid) {
        int i = p / kHid, j = p % kHid;
        for (int b = 0; b < kBatch; ++b) g += xb[b * kFeat + i] * g1[b * kHid + j];
      } else {
        int j = p - kFeat * kHid;
        for (int b = 0; b < kBatch; ++b) g += g1[b * kHid + j];
      }
      float *w = p < kFeat * kHid ? W1 + p : B1 + (p - kFeat * kHid);
      float *mm = M + (w - P);
      *mm = momentum * *mm + g;
      *w -= lr * *mm;
    }
    __syncthreads();
  }
  for (int i = t; i < kNetFloats; i += kTrainThreads) model[kOffW1 + i] = P[i];
  if (t == 0 && loss_out) *loss_out = wacc > 0.0f ? loss_acc / wacc : 0.0f;
}

constexpr size_t kTrainLds =
    sizeof(float) * (2 * kNetFloats + kBatch * kFeat + 4 * kBatch * kHid + kBatch * kAct + kBatch) +
    sizeof(int) * kBatch + sizeof(float) * kTrainThreads;

}  // namespace lsg

using namespace lsg;

extern "C" {

int lsg_sched_model_floats(void) { return kModelFloats; }

int lsg_sched_set_model(int device, int kind, const float *model, uint32_t nfloats) {
  if (kind != LSG_SCHED_BATCHED && kind != LSG_SCHED_PLAIN) return fail(LSG_ERR_ARG, "bad scheduler model kind %d", kind);
  if (!model || nfloats != (uint32_t)kModelFloats)
    return fail(LSG_ERR_ARG, "scheduler model must have %d floats (got %u)", kModelFloats, nfloats);
  for (uint32_t i = 0; i < nfloats; ++i)
    if (!std::isfinite(model[i])) return fail(LSG_ERR_ARG, "scheduler model has a non-finite value at %u", i);
  Workspace *w;
  int rc = get_workspace(device, &w);
  if (rc) return rc;
  HIPCHK(hipSetDevice(device));
  std::lock_guard<std::mutex> lk(g_model_mu);
  float *dptr = nullptr;
  for (auto &e : g_models)
    if (e.device == device && e.kind == kind) dptr = e.dptr;
  if (!dptr) {
    HIPCHK(hipMalloc(&dptr, kModelFloats * sizeof(float)));
    g_models.push_back({device, kind, dptr});
  }
  HIPCHK(hipMemcpy(dptr, model, kModelFloats * sizeof(float), hipMemcpyHostToDevice));
  return LSG_OK;
}

int lsg_mlp_infer(int device, const float *model, const float *X, uint32_t n, int32_t *labels) {
  if (!model || (!X && n) || (!labels && n)) return fail(LSG_ERR_ARG, "mlp_infer: NULL argument");
  if (n == 0) return LSG_OK;
  Workspace *w;
  int rc = get_workspace(device, &w);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(w->mu);
  HIPCHK(hipSetDevice(device));
  float *dm, *dx;
  int32_t *dy;
  HIPCHK(hipMalloc(&dm, kModelFloats * sizeof(float)));
  HIPCHK(hipMalloc(&dx, (size_t)n * kFeat * sizeof(float)));
  HIPCHK(hipMalloc(&dy, (size_t)n * sizeof(int32_t)));
  HIPCHK(hipMemcpy(dm, model, kModelFloats * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dx, X, (size_t)n * kFeat * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(mlp_infer_kernel, dim3((n + 255) / 256), dim3(256), 0, w->stream, dm, dx, n, dy);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(w->stream));
  HIPCHK(hipMemcpy(labels, dy, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
  HIPCHK(hipFree(dm));
  HIPCHK(hipFree(dx));
  HIPCHK(hipFree(dy));
  return LSG_OK;
}

int lsg_mlp_train(int device, float *model, const float *X, const int32_t *y, const float *weights, uint32_t n,
                  uint32_t steps, float lr, float momentum, uint64_t seed, float *loss) {
  if (!model || !X || !y || n == 0) return fail(LSG_ERR_ARG, "mlp_train: empty or NULL training set");
  for (uint32_t i = 0; i < n; ++i)
    if (y[i] < 0 || y[i] >= kAct) return fail(LSG_ERR_ARG, "mlp_train: label %d out of range at %u", y[i], i);
  // Feature normalisation from the training set, stored in the model.
  for (int f = 0; f < kFeat; ++f) {
    double s = 0, s2 = 0;
    for (uint32_t i = 0; i < n; ++i) {
      double v = X[(size_t)i * kFeat + f];
      s += v;
      s2 += v * v;
    }
    double mean = s / n, var = s2 / n - mean * mean;
    model[kOffMean + f] = (float)mean;
    model[kOffInvStd + f] = var > 1e-12 ? (float)(1.0 / std::sqrt(var)) : 1.0f;
  }
  std::vector<float> Xn((size_t)n * kFeat);
  for (uint32_t i = 0; i < n; ++i)
    for (int f = 0; f < kFeat; ++f)
      Xn[(size_t)i * kFeat + f] = (X[(size_t)i * kFeat + f] - model[kOffMean + f]) * model[kOffInvStd + f];
  Workspace *w;
  int rc = get_workspace(device, &w);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(w->mu);
  HIPCHK(hipSetDevice(device));
  float *dm, *dx, *dw = nullptr, *dl;
  int32_t *dy;
  HIPCHK(hipMalloc(&dm, kModelFloats * sizeof(float)));
  HIPCHK(hipMalloc(&dx, Xn.size() * sizeof(float)));
  HIPCHK(hipMalloc(&dy, (size_t)n * sizeof(int32_t)));
  HIPCHK(hipMalloc(&dl, sizeof(float)));
  if (weights) {
    HIPCHK(hipMalloc(&dw, (size_t)n * sizeof(float)));
    HIPCHK(hipMemcpy(dw, weights, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(dm, model, kModelFloats * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dx, Xn.data(), Xn.size() * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dy, y, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(hipFuncSetAttribute((const void *)mlp_train_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)kTrainLds));
  hipLaunchKernelGGL(mlp_train_kernel, dim3(1), dim3(kTrainThreads), kTrainLds, w->stream, dm, dx, dy, dw, n,
                     steps, lr, momentum, seed, dl);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(w->stream));
  HIPCHK(hipMemcpy(model + kOffW1, dm + kOffW1, kNetFloats * sizeof(float), hipMemcpyDeviceToHost));
  if (loss) HIPCHK(hipMemcpy(loss, dl, sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipFree(dm));
  HIPCHK(hipFree(dx));
  HIPCHK(hipFree(dy));
  HIPCHK(hipFree(dl));
  if (dw) HIPCHK(hipFree(dw));
  return LSG_OK;
}

}  // extern "C"
