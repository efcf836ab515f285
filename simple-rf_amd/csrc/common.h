// Shared helpers for the gfx950 kernels of the Simple-RF rendering hot path.
//
// Conventions of the C-ABI (include/srf_hip.h):
//   * every entry point returns an int status (SRF_OK = 0, see the SRF_E* codes);
//     no C++ exception crosses the boundary;
//   * all pointers are device pointers owned by the caller; nothing here allocates;
//   * every launch goes on the caller's stream (a hipStream_t passed as void*);
//   * floating-point expressions that must reproduce PyTorch's op-by-op rounding are
//     written with explicit __fmul_rn/__fadd_rn/__fsub_rn or fmaf, and the sources are
//     compiled with -ffp-contract=off so the compiler does not fuse them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/srf_hip.h"

#define SRF_WAVE 64

namespace srf {

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Current value of a launch-schedule knob (SRF_TUNE_*; 0 = the built-in choice).
int64_t tune_value(int knob);

inline int launch_status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SRF_OK : SRF_EHIP;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Wavefront inclusive prefix sum (Hillis-Steele over 64 lanes).
__device__ __forceinline__ float wave_inclusive_sum(float v) {
  const int l = lane_id();
#pragma unroll
  for (int d = 1; d < SRF_WAVE; d <<= 1) {
    float o = __shfl_up(v, d, SRF_WAVE);
    if (l >= d) v = v + o;
  }
  return v;
}

// Wavefront inclusive suffix sum: lane l gets sum over lanes >= l.
__device__ __forceinline__ float wave_inclusive_suffix_sum(float v) {
  const int l = lane_id();
#pragma unroll
  for (int d = 1; d < SRF_WAVE; d <<= 1) {
    float o = __shfl_down(v, d, SRF_WAVE);
    if (l + d < SRF_WAVE) v = v + o;
  }
  return v;
}

// Butterfly all-reduce: every lane receives the wave total.
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int d = SRF_WAVE / 2; d >= 1; d >>= 1) v = v + __shfl_xor(v, d, SRF_WAVE);
  return v;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int d = SRF_WAVE / 2; d >= 1; d >>= 1) v = v + __shfl_xor(v, d, SRF_WAVE);
  return v;
}

// torch.linspace(start, end, steps)[i] exactly as PyTorch's CPU kernel produces it:
// step = (end - start) / (steps - 1) in float; the first half is fma(step, i, start),
// the second half fma(-step, steps - 1 - i, end).
__device__ __forceinline__ float torch_linspace(float start, float end, int steps, int i) {
  if (steps == 1) return start;
  const float step = (end - start) / (float)(steps - 1);
  const int halfway = steps / 2;
  return (i < halfway) ? fmaf(step, (float)i, start) : fmaf(-step, (float)(steps - 1 - i), end);
}

// torch.norm over a length-3 vector as PyTorch's CPU kernel evaluates it (float fma chain).
__device__ __forceinline__ float torch_norm3(float x, float y, float z) {
  return sqrtf(fmaf(z, z, fmaf(y, y, __fmul_rn(x, x))));
}

}  // namespace srf
