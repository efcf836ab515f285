// Adam step over many parameter tensors in one launch: the optimiser of the reference's training
// (torch.optim.Adam built by src/optimizers/OptimizerFactory02.py:9-21, amsgrad off, maximize off,
// weight decay added to the gradient).
//
// Per element, in torch.optim.Adam's order of operations (_single_tensor_adam):
//   g  = grad (+ weight_decay * param)
//   m  = m + (1 - beta1) * (g - m)                 (exp_avg.lerp_(g, 1 - beta1), for 1 - beta1 < 0.5)
//   v  = v * beta2 + ((1 - beta2) * g) * g          (exp_avg_sq.mul_(beta2).addcmul_(g, g, 1 - beta2))
//   p  = p + (-step_size) * (m / (sqrt(v) / bc2_sqrt + eps))
// with step_size = lr / (1 - beta1^t) and bc2_sqrt = sqrt(1 - beta2^t) evaluated in double, as
// Python evaluates them for torch, t being the step count after this step.  The step counts live
// on the device (one float per tensor, torch's state['step']) so that a captured step advances them.
//
// Layout.  Parameter, gradient and moments of a tensor share one dense layout, which the kernel
// walks as a flat array; a block handles kAdamChunk consecutive elements of one tensor as float4s
// (the scalar tail is handled by the same threads).  The tensor of a block comes from a table of
// first blocks in the kernel arguments (at most kAdamMaxTensors tensors per launch).
#include <math.h>
#include <string.h>

#include "common.h"

namespace srf {
namespace {

constexpr int kAdamMaxTensors = 32;
constexpr int kAdamThreads = 256;
constexpr int kAdamChunk = kAdamThreads * 4 * 4;  // elements per block: 4 float4 per thread

struct AdamArgs {
  float* p[kAdamMaxTensors];
  const float* g[kAdamMaxTensors];
  float* m[kAdamMaxTensors];
  float* v[kAdamMaxTensors];
  float* step[kAdamMaxTensors];
  int64_t numel[kAdamMaxTensors];
  int first_block[kAdamMaxTensors + 1];
  float lr[kAdamMaxTensors];
  float wd[kAdamMaxTensors];
  int n;
  float beta1, beta2, eps;
};

__device__ __forceinline__ void adam_element(float& p, float g, float& m, float& v, float wd, float w1, float beta2,
                                             float w2, float neg_step, float bc2s, float eps) {
  if (wd != 0.f) g = __fadd_rn(g, __fmul_rn(wd, p));
  // torch's lerp: self + w (end - self) for w < 0.5, else end - (end - self) (1 - w)
  m = w1 < 0.5f ? __fadd_rn(m, __fmul_rn(w1, __fsub_rn(g, m)))
                : __fsub_rn(g, __fmul_rn(__fsub_rn(g, m), __fsub_rn(1.f, w1)));
  v = __fadd_rn(__fmul_rn(v, beta2), __fmul_rn(__fmul_rn(w2, g), g));
  const float denom = __fadd_rn(__fdiv_rn(__fsqrt_rn(v), bc2s), eps);
  p = __fadd_rn(p, __fmul_rn(neg_step, __fdiv_rn(m, denom)));
}

__global__ __launch_bounds__(kAdamThreads) void adam_kernel(AdamArgs a) {
  int t = 0;
#pragma unroll 1
  while (t + 1 < a.n && (int)blockIdx.x >= a.first_block[t + 1]) ++t;
  const int64_t n = a.numel[t];
  const int64_t e0 = (int64_t)(blockIdx.x - a.first_block[t]) * kAdamChunk;
  // the bias corrections of step count s = state step + 1, in double as torch's Python scalars
  const double s = (double)*a.step[t] + 1.0;
  const double bc1 = 1.0 - pow((double)a.beta1, s);
  const double bc2 = 1.0 - pow((double)a.beta2, s);
  const float neg_step = (float)(-((double)a.lr[t] / bc1));
  const float bc2s = (float)sqrt(bc2);
  const float w1 = (float)(1.0 - (double)a.beta1), w2 = (float)(1.0 - (double)a.beta2);
  const float wd = a.wd[t], beta2 = a.beta2, eps = a.eps;
  float* P = a.p[t];
  const float* G = a.g[t];
  float* M = a.m[t];
  float* V = a.v[t];
  const bool vec = ((reinterpret_cast<uintptr_t>(P) | reinterpret_cast<uintptr_t>(G) | reinterpret_cast<uintptr_t>(M) |
                     reinterpret_cast<uintptr_t>(V)) & 15) == 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t e = e0 + ((int64_t)k * kAdamThreads + threadIdx.x) * 4;
    if (vec && e + 4 <= n) {
      float4 p = *reinterpret_cast<const float4*>(P + e);
      const float4 g = *reinterpret_cast<const float4*>(G + e);
      float4 m = *reinterpret_cast<const float4*>(M + e);
      float4 v = *reinterpret_cast<const float4*>(V + e);
      adam_element(p.x, g.x, m.x, v.x, wd, w1, beta2, w2, neg_step, bc2s, eps);
      adam_element(p.y, g.y, m.y, v.y, wd, w1, beta2, w2, neg_step, bc2s, eps);
      adam_element(p.z, g.z, m.z, v.z, wd, w1, beta2, w2, neg_step, bc2s, eps);
      adam_element(p.w, g.w, m.w, v.w, wd, w1, beta2, w2, neg_step, bc2s, eps);
      *reinterpret_cast<float4*>(P + e) = p;
      *reinterpret_cast<float4*>(M + e) = m;
      *reinterpret_cast<float4*>(V + e) = v;
    } else {
      for (int64_t i = e; i < e + 4 && i < n; ++i) {
        float p = P[i], m = M[i], v = V[i];
        adam_element(p, G[i], m, v, wd, w1, beta2, w2, neg_step, bc2s, eps);
        P[i] = p;
        M[i] = m;
        V[i] = v;
      }
    }
  }
}

// After the update: every tensor's step count advances by one.
__global__ void adam_count_kernel(AdamArgs a) {
  if ((int)threadIdx.x < a.n) *a.step[threadIdx.x] = __fadd_rn(*a.step[threadIdx.x], 1.f);
}

}  // namespace
}  // namespace srf

using namespace srf;

extern "C" int srf_adam_step(int n_tensors, float* const* params, const float* const* grads, float* const* exp_avg,
                             float* const* exp_avg_sq, float* const* steps, const int64_t* numel, const float* lr,
                             const float* weight_decay, float beta1, float beta2, float eps, void* stream) {
  if (n_tensors < 0) return SRF_EINVAL;
  if (n_tensors == 0) return SRF_OK;
  if (!params || !grads || !exp_avg || !exp_avg_sq || !steps || !numel || !lr || !weight_decay) return SRF_EINVAL;
  if (!(beta1 >= 0.f && beta1 < 1.f) || !(beta2 >= 0.f && beta2 < 1.f) || !(eps >= 0.f)) return SRF_EINVAL;
  hipStream_t s = as_stream(stream);
  for (int t0 = 0; t0 < n_tensors; t0 += kAdamMaxTensors) {
    AdamArgs a;
    memset(&a, 0, sizeof(a));
    a.beta1 = beta1;
    a.beta2 = beta2;
    a.eps = eps;
    int blocks = 0;
    for (int k = 0; k < kAdamMaxTensors && t0 + k < n_tensors; ++k) {
      const int t = t0 + k;
      if (numel[t] < 0 || (numel[t] > 0 && (!params[t] || !grads[t] || !exp_avg[t] || !exp_avg_sq[t])) || !steps[t])
        return SRF_EINVAL;
      const int64_t nb = (numel[t] + kAdamChunk - 1) / kAdamChunk;
      if ((int64_t)blocks + nb > 0x7fffffffLL) return SRF_EINVAL;
      a.p[k] = params[t];
      a.g[k] = grads[t];
      a.m[k] = exp_avg[t];
      a.v[k] = exp_avg_sq[t];
      a.step[k] = steps[t];
      a.numel[k] = numel[t];
      a.lr[k] = lr[t];
      a.wd[k] = weight_decay[t];
      a.first_block[k] = blocks;
      blocks += (int)nb;
      a.n = k + 1;
    }
    a.first_block[a.n] = blocks;
    if (blocks > 0) hipLaunchKernelGGL(adam_kernel, dim3(blocks), dim3(kAdamThreads), 0, s, a);
    hipLaunchKernelGGL(adam_count_kernel, dim3(1), dim3(kAdamMaxTensors), 0, s, a);
    const int st = launch_status();
    if (st != SRF_OK) return st;
  }
  return SRF_OK;
}
