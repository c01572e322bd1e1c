// mlp.hip -- the learned tile / schedule selector: a 2-layer MLP trained on
// the autotuner's own timings, with hand-written HIP kernels for inference,
// back-propagation and the Adam update.
//
// It replaces the role of Eigen's blocking heuristic
// (evaluateProductBlockingSizesHeuristic,
// Eigen/src/Core/products/GeneralBlockPanelKernel.h:201-405), which derives
// kc/mc/nc from cache sizes; here the network maps a product shape to one of
// the candidate tile / split-K configurations.
//
//   x  (7)  : log2 M / 14, log2 N / 14, log2 K / 14, [FP32, FP64] one-hot,
//             transA, transB
//   h  (64) : relu(W1 x + b1)
//   z  (24) : W2 h + b2,  p = softmax(z)
//
// Parameters are one flat FP32 array: W1 [64][7] (row-major), b1 [64],
// W2 [24][64], b2 [24] -- 2072 values.  Training minimises the cross entropy
// against soft targets q (rows of a probability matrix built from the
// timings), averaged over the batch; gradients are those of
//     L = -(1/B) sum_b sum_o q[b][o] log p[b][o].
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/eigen_amd.h"
#include "internal.hpp"

namespace eh {

constexpr int kIn = 7, kHid = 64, kOut = kNumConfigs;
constexpr int kOffW1 = 0, kOffB1 = kHid * kIn, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kOut * kHid;
constexpr int kNumParams = kOffB2 + kOut;

// ---- forward: one thread per (sample, hidden unit), then per (sample, output)
__global__ void mlp_hidden_kernel(const float* __restrict__ P, const float* __restrict__ X, int64_t B,
                                  float* __restrict__ H) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= B * kHid) return;
  const int64_t b = idx / kHid;
  const int h = (int)(idx % kHid);
  float s = P[kOffB1 + h];
#pragma unroll
  for (int i = 0; i < kIn; ++i) s = fmaf(P[kOffW1 + h * kIn + i], X[b * kIn + i], s);
  H[idx] = s > 0.0f ? s : 0.0f;
}

// one 64-thread wave per sample: logits, max-subtracted softmax
__global__ void mlp_output_kernel(const float* __restrict__ P, const float* __restrict__ H, int64_t B,
                                  float* __restrict__ Z, float* __restrict__ Prob) {
  const int64_t b = blockIdx.x;
  const int o = threadIdx.x;
  __shared__ float z[kOut];
  __shared__ float mx, sum;
  if (o < kOut) {
    float s = P[kOffB2 + o];
    for (int h = 0; h < kHid; ++h) s = fmaf(P[kOffW2 + o * kHid + h], H[b * kHid + h], s);
    z[o] = s;
    if (Z) Z[b * kOut + o] = s;
  }
  __syncthreads();
  if (o == 0) {
    float m = z[0];
    for (int i = 1; i < kOut; ++i) m = fmaxf(m, z[i]);
    float t = 0.0f;
    for (int i = 0; i < kOut; ++i) t += expf(z[i] - m);
    mx = m;
    sum = t;
  }
  __syncthreads();
  if (o < kOut && Prob) Prob[b * kOut + o] = expf(z[o] - mx) / sum;
}

// dZ = (p - q) / B and the per-element loss terms -q log p / B (summed to the
// batch mean by mlp_sum_kernel)
__global__ void mlp_dz_kernel(const float* __restrict__ Prob, const float* __restrict__ Q, int64_t B,
                              float* __restrict__ dZ, float* __restrict__ loss_parts) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= B * kOut) return;
  const float p = Prob[idx], q = Q[idx];
  dZ[idx] = (p - q) / (float)B;
  loss_parts[idx] = q > 0.0f ? -q * logf(fmaxf(p, 1e-30f)) / (float)B : 0.0f;
}

// dH = (W2^T dZ) * [h > 0]
__global__ void mlp_dh_kernel(const float* __restrict__ P, const float* __restrict__ H, const float* __restrict__ dZ,
                              int64_t B, float* __restrict__ dH) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= B * kHid) return;
  const int64_t b = idx / kHid;
  const int h = (int)(idx % kHid);
  if (H[idx] <= 0.0f) {
    dH[idx] = 0.0f;
    return;
  }
  float s = 0.0f;
  for (int o = 0; o < kOut; ++o) s = fmaf(P[kOffW2 + o * kHid + h], dZ[b * kOut + o], s);
  dH[idx] = s;
}

// Parameter gradients: one thread per parameter, reducing over the batch.
__global__ void mlp_grad_kernel(const float* __restrict__ X, const float* __restrict__ H,
                                const float* __restrict__ dZ, const float* __restrict__ dH, int64_t B,
                                float* __restrict__ G) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= kNumParams) return;
  float s = 0.0f;
  if (j < kOffB1) {  // W1[h][i]
    const int h = j / kIn, i = j % kIn;
    for (int64_t b = 0; b < B; ++b) s = fmaf(dH[b * kHid + h], X[b * kIn + i], s);
  } else if (j < kOffW2) {  // b1[h]
    const int h = j - kOffB1;
    for (int64_t b = 0; b < B; ++b) s += dH[b * kHid + h];
  } else if (j < kOffB2) {  // W2[o][h]
    const int o = (j - kOffW2) / kHid, h = (j - kOffW2) % kHid;
    for (int64_t b = 0; b < B; ++b) s = fmaf(dZ[b * kOut + o], H[b * kHid + h], s);
  } else {  // b2[o]
    const int o = j - kOffB2;
    for (int64_t b = 0; b < B; ++b) s += dZ[b * kOut + o];
  }
  G[j] = s;
}

// Adam with decoupled weight decay on the weights (not the biases).
__global__ void mlp_adam_kernel(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ Mo,
                                float* __restrict__ Ve, float lr, float b1, float b2, float eps, float wd,
                                int step) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= kNumParams) return;
  const float g = G[j];
  const float m = b1 * Mo[j] + (1.0f - b1) * g;
  const float v = b2 * Ve[j] + (1.0f - b2) * g * g;
  Mo[j] = m;
  Ve[j] = v;
  const float mh = m / (1.0f - powf(b1, (float)step));
  const float vh = v / (1.0f - powf(b2, (float)step));
  const bool weight = (j < kOffB1) || (j >= kOffW2 && j < kOffB2);
  P[j] -= lr * (mh / (sqrtf(vh) + eps) + (weight ? wd * P[j] : 0.0f));
}

__global__ void mlp_sum_kernel(const float* __restrict__ x, int64_t n, float* __restrict__ out) {
  __shared__ float part[256];
  float s = 0.0f;
  for (int64_t i = threadIdx.x; i < n; i += 256) s += x[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = part[0];
}

__global__ void mlp_argmax_kernel(const float* __restrict__ Prob, int64_t B, int32_t* __restrict__ out) {
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= B) return;
  int best = 0;
  float v = Prob[b * kOut];
  for (int o = 1; o < kOut; ++o)
    if (Prob[b * kOut + o] > v) {
      v = Prob[b * kOut + o];
      best = o;
    }
  out[b] = best;
}

static unsigned blocks_for(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace eh

using namespace eh;

// A network and its device buffers (parameters, Adam moments, activations for
// up to `capacity` samples).
struct eh_mlp {
  float* P = nullptr;  // parameters
  float* G = nullptr;  // gradients
  float* Mo = nullptr;
  float* Ve = nullptr;
  float* H = nullptr;
  float* Z = nullptr;
  float* Prob = nullptr;
  float* dZ = nullptr;
  float* dH = nullptr;
  float* lossp = nullptr;
  float* loss = nullptr;
  int64_t capacity = 0;
  int step = 0;
  std::vector<float> host;  // copy of the parameters for host-side planning
};

static eh_status mlp_forward_dev(eh_mlp* m, const float* X, int64_t B, hipStream_t s) {
  hipLaunchKernelGGL(mlp_hidden_kernel, dim3(blocks_for(B * kHid)), dim3(256), 0, s, m->P, X, B, m->H);
  hipLaunchKernelGGL(mlp_output_kernel, dim3((unsigned)B), dim3(64), 0, s, m->P, m->H, B, m->Z, m->Prob);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "mlp forward launch");
}

static eh_status mlp_backward_dev(eh_mlp* m, const float* X, const float* Q, int64_t B, hipStream_t s) {
  hipLaunchKernelGGL(mlp_dz_kernel, dim3(blocks_for(B * kOut)), dim3(256), 0, s, m->Prob, Q, B, m->dZ, m->lossp);
  hipLaunchKernelGGL(mlp_dh_kernel, dim3(blocks_for(B * kHid)), dim3(256), 0, s, m->P, m->H, m->dZ, B, m->dH);
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks_for(kNumParams)), dim3(256), 0, s, X, m->H, m->dZ, m->dH, B, m->G);
  hipLaunchKernelGGL(mlp_sum_kernel, dim3(1), dim3(256), 0, s, m->lossp, B * kOut, m->loss);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "mlp backward launch");
}

#if __has_include("selector_weights.inc")
#include "selector_weights.inc"
#define EH_HAVE_DEFAULT_SELECTOR 1
#endif

namespace eh {
// The planner's network (host copy of the parameters); empty = none set.
// Starts as the compiled-in network trained on this device's timing table.
static std::mutex g_mlp_mutex;
#ifdef EH_HAVE_DEFAULT_SELECTOR
static std::vector<float> g_planner_params(kDefaultSelectorParams, kDefaultSelectorParams + kNumParams);
bool mlp_has_default() { return true; }
static bool g_planner_is_default = true;  // the compiled-in network (and its tuned shapes) is in use
#else
static std::vector<float> g_planner_params;
bool mlp_has_default() { return false; }
static bool g_planner_is_default = false;
#endif

// Measured best configuration of a shape that was tuned together with the
// compiled-in network (exact match), or -1.
static int tuned_shape_plan(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k) {
#ifdef EH_HAVE_DEFAULT_SELECTOR
  const long long ta = opa != EH_OP_N, tb = opb != EH_OP_N;
  for (int i = 0; i < kNumTunedShapes; ++i) {
    const long long* r = kTunedShapes[i];
    if (r[0] == dtype && r[1] == ta && r[2] == tb && r[3] == m && r[4] == n && r[5] == k) return (int)r[6];
  }
#else
  (void)dtype, (void)opa, (void)opb, (void)m, (void)n, (void)k;
#endif
  return -1;
}

// Feature vector of a product shape (the network's input).
void mlp_features(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, float* x) {
  x[0] = (float)(std::log2((double)std::max<int64_t>(m, 1)) / 14.0);
  x[1] = (float)(std::log2((double)std::max<int64_t>(n, 1)) / 14.0);
  x[2] = (float)(std::log2((double)std::max<int64_t>(k, 1)) / 14.0);
  x[3] = dtype == EH_F32 ? 1.0f : 0.0f;
  x[4] = dtype == EH_F64 ? 1.0f : 0.0f;
  x[5] = opa != EH_OP_N ? 1.0f : 0.0f;
  x[6] = opb != EH_OP_N ? 1.0f : 0.0f;
}

// Host evaluation of the planner network (same arithmetic as the kernels:
// FP32 fused multiply-adds in the same order); returns -1 if no network set.
// With the compiled-in network, shapes it was tuned on use their measured
// best configuration.
int mlp_plan(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k) {
  std::lock_guard<std::mutex> lk(g_mlp_mutex);
  if (g_planner_params.empty()) return -1;
  if (g_planner_is_default) {
    const int c = tuned_shape_plan(dtype, opa, opb, m, n, k);
    if (c >= 0) return c;
  }
  const float* P = g_planner_params.data();
  float x[kIn], h[kHid];
  mlp_features(dtype, opa, opb, m, n, k, x);
  for (int j = 0; j < kHid; ++j) {
    float s = P[kOffB1 + j];
    for (int i = 0; i < kIn; ++i) s = std::fma(P[kOffW1 + j * kIn + i], x[i], s);
    h[j] = s > 0.0f ? s : 0.0f;
  }
  int best = 0;
  float bv = 0.0f;
  for (int o = 0; o < kOut; ++o) {
    float s = P[kOffB2 + o];
    for (int j = 0; j < kHid; ++j) s = std::fma(P[kOffW2 + o * kHid + j], h[j], s);
    if (o == 0 || s > bv) {
      bv = s;
      best = o;
    }
  }
  return best;
}
}  // namespace eh

extern "C" {

int64_t eh_mlp_num_params(eh_mlp* m) { return m ? kNumParams : kNumParams; }

eh_status eh_mlp_create(int inputs, int hidden, int outputs, int64_t capacity, eh_mlp** out) {
  if (!out) return fail(EH_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (inputs != kIn || hidden != kHid || outputs != kOut)
    return fail(EH_ERR_UNSUPPORTED, "the selector network is 7 -> 64 -> 24");
  if (capacity < 1) return fail(EH_ERR_INVALID, "capacity must be >= 1");
  auto* m = new eh_mlp();
  m->capacity = capacity;
  hipError_t e = hipSuccess;
  float** bufs[] = {&m->P, &m->G, &m->Mo, &m->Ve};
  for (float** b : bufs)
    if (e == hipSuccess) e = hipMalloc(b, kNumParams * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->H, capacity * kHid * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->dH, capacity * kHid * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->Z, capacity * kOut * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->Prob, capacity * kOut * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->dZ, capacity * kOut * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->lossp, capacity * kOut * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->loss, sizeof(float));
  if (e == hipSuccess) e = hipMemset(m->Mo, 0, kNumParams * sizeof(float));
  if (e == hipSuccess) e = hipMemset(m->Ve, 0, kNumParams * sizeof(float));
  if (e != hipSuccess) {
    eh_mlp_destroy(m);
    return fail(EH_ERR_ALLOC, std::string("mlp buffers: ") + hipGetErrorString(e));
  }
  // default initialisation: small deterministic pseudo-random weights
  std::vector<float> init(kNumParams, 0.0f);
  uint64_t z = 0x243F6A8885A308D3ull;
  auto next = [&]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t t = z;
    t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
    t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
    t ^= t >> 31;
    return (double)(t >> 11) * 0x1.0p-53 * 2.0 - 1.0;
  };
  for (int j = 0; j < kOffB1; ++j) init[j] = (float)(next() * std::sqrt(6.0 / kIn));
  for (int j = kOffW2; j < kOffB2; ++j) init[j] = (float)(next() * std::sqrt(6.0 / kHid));
  eh_status st = eh_mlp_set_params(m, init.data());
  if (st != EH_SUCCESS) {
    eh_mlp_destroy(m);
    return st;
  }
  *out = m;
  return EH_SUCCESS;
}

eh_status eh_mlp_destroy(eh_mlp* m) {
  if (!m) return EH_SUCCESS;
  float* bufs[] = {m->P, m->G, m->Mo, m->Ve, m->H, m->Z, m->Prob, m->dZ, m->dH, m->lossp, m->loss};
  for (float* b : bufs)
    if (b) hipFree(b);
  delete m;
  return EH_SUCCESS;
}

eh_status eh_mlp_set_params(eh_mlp* m, const float* host_params) {
  if (!m || !host_params) return fail(EH_ERR_INVALID, "NULL argument");
  hipError_t e = hipMemcpy(m->P, host_params, kNumParams * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "mlp set params");
  m->host.assign(host_params, host_params + kNumParams);
  m->step = 0;
  e = hipMemset(m->Mo, 0, kNumParams * sizeof(float));
  if (e == hipSuccess) e = hipMemset(m->Ve, 0, kNumParams * sizeof(float));
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "mlp reset moments");
}

eh_status eh_mlp_get_params(eh_mlp* m, float* host_params) {
  if (!m || !host_params) return fail(EH_ERR_INVALID, "NULL argument");
  hipError_t e = hipMemcpy(host_params, m->P, kNumParams * sizeof(float), hipMemcpyDeviceToHost);
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "mlp get params");
}

eh_status eh_mlp_features(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, float* x) {
  if (!x) return fail(EH_ERR_INVALID, "x is NULL");
  mlp_features(dtype, opa, opb, m, n, k, x);
  return EH_SUCCESS;
}

eh_status eh_mlp_forward(eh_mlp* m, const float* X, int64_t B, float* probs, void* stream) {
  if (!m || !X) return fail(EH_ERR_INVALID, "NULL argument");
  if (B < 1 || B > m->capacity) return fail(EH_ERR_INVALID, "batch outside [1, capacity]");
  hipStream_t s = static_cast<hipStream_t>(stream);
  eh_status st = mlp_forward_dev(m, X, B, s);
  if (st != EH_SUCCESS) return st;
  if (probs) {
    hipError_t e = hipMemcpyAsync(probs, m->Prob, B * kOut * sizeof(float), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "mlp copy probs");
  }
  return EH_SUCCESS;
}

eh_status eh_mlp_backward(eh_mlp* m, const float* X, const float* Q, int64_t B, float* grads, float* loss,
                          void* stream) {
  if (!m || !X || !Q) return fail(EH_ERR_INVALID, "NULL argument");
  if (B < 1 || B > m->capacity) return fail(EH_ERR_INVALID, "batch outside [1, capacity]");
  hipStream_t s = static_cast<hipStream_t>(stream);
  eh_status st = mlp_forward_dev(m, X, B, s);
  if (st == EH_SUCCESS) st = mlp_backward_dev(m, X, Q, B, s);
  if (st != EH_SUCCESS) return st;
  hipError_t e = hipSuccess;
  if (grads) e = hipMemcpyAsync(grads, m->G, kNumParams * sizeof(float), hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToDevice, s);
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "mlp copy grads");
}

eh_status eh_mlp_adam_step(eh_mlp* m, const float* grads, float lr, float beta1, float beta2, float eps,
                           float weight_decay, void* stream) {
  if (!m) return fail(EH_ERR_INVALID, "NULL handle");
  hipStream_t s = static_cast<hipStream_t>(stream);
  m->step += 1;
  hipLaunchKernelGGL(mlp_adam_kernel, dim3(blocks_for(kNumParams)), dim3(256), 0, s, m->P, grads ? grads : m->G,
                     m->Mo, m->Ve, lr, beta1, beta2, eps, weight_decay, m->step);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "adam launch");
}

// Full-batch training: `epochs` Adam steps on (X, Q), learning rate decayed
// by cosine from lr to lr/20.  loss_curve (device, epochs floats) may be NULL.
eh_status eh_mlp_train(eh_mlp* m, const float* X, const float* Q, int64_t B, int epochs, float lr,
                       float weight_decay, float* loss_curve, void* stream) {
  if (!m || !X || !Q) return fail(EH_ERR_INVALID, "NULL argument");
  if (B < 1 || B > m->capacity) return fail(EH_ERR_INVALID, "batch outside [1, capacity]");
  if (epochs < 1) return fail(EH_ERR_INVALID, "epochs must be >= 1");
  hipStream_t s = static_cast<hipStream_t>(stream);
  for (int ep = 0; ep < epochs; ++ep) {
    eh_status st = mlp_forward_dev(m, X, B, s);
    if (st == EH_SUCCESS) st = mlp_backward_dev(m, X, Q, B, s);
    if (st != EH_SUCCESS) return st;
    if (loss_curve) {
      hipError_t e = hipMemcpyAsync(loss_curve + ep, m->loss, sizeof(float), hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) return hip_fail(e, "loss copy");
    }
    const float cosf_ = 0.5f * (1.0f + std::cos(3.14159265f * ep / (float)epochs));
    const float lr_t = lr * (0.05f + 0.95f * cosf_);
    st = eh_mlp_adam_step(m, nullptr, lr_t, 0.9f, 0.999f, 1e-8f, weight_decay, stream);
    if (st != EH_SUCCESS) return st;
  }
  // keep the host copy in step with the trained parameters
  m->host.resize(kNumParams);
  hipError_t e = hipMemcpyAsync(m->host.data(), m->P, kNumParams * sizeof(float), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "mlp params to host");
}

// Batched inference on the device: argmax configuration per sample.
eh_status eh_mlp_predict(eh_mlp* m, const float* X, int64_t B, int32_t* configs, void* stream) {
  if (!m || !X || !configs) return fail(EH_ERR_INVALID, "NULL argument");
  if (B < 1 || B > m->capacity) return fail(EH_ERR_INVALID, "batch outside [1, capacity]");
  hipStream_t s = static_cast<hipStream_t>(stream);
  eh_status st = mlp_forward_dev(m, X, B, s);
  if (st != EH_SUCCESS) return st;
  hipLaunchKernelGGL(mlp_argmax_kernel, dim3(blocks_for(B)), dim3(256), 0, s, m->Prob, B, configs);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "argmax launch");
}

// Install the network as the planner's selector (its current parameters are
// copied; later training does not change the planner until set again).
// NULL restores the compiled-in default network (if the build has one).
eh_status eh_planner_set_mlp(eh_mlp* m, void* stream) {
  std::vector<float> p(kNumParams);
  if (m) {
    hipError_t e = hipMemcpyAsync(p.data(), m->P, kNumParams * sizeof(float), hipMemcpyDeviceToHost,
                                  static_cast<hipStream_t>(stream));
    if (e == hipSuccess) e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "mlp params to host");
  }
  return eh_planner_set_mlp_params(m ? p.data() : nullptr);
}

// Same from a host parameter array (no device work); NULL restores the
// compiled-in default.
eh_status eh_planner_set_mlp_params(const float* host_params) {
  std::lock_guard<std::mutex> lk(g_mlp_mutex);
  if (host_params) {
    g_planner_params.assign(host_params, host_params + kNumParams);
    g_planner_is_default = false;
  } else {
#ifdef EH_HAVE_DEFAULT_SELECTOR
    g_planner_params.assign(kDefaultSelectorParams, kDefaultSelectorParams + kNumParams);
    g_planner_is_default = true;
#else
    g_planner_params.clear();
#endif
  }
  return EH_SUCCESS;
}

}  // extern "C"
