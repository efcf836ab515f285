// Library-level entry points: version, status strings, the launch-schedule knobs and the MFMA
// lane-layout self-test.
#include <atomic>

#include "common.h"
#include "mfma.h"

namespace srf {

namespace {
std::atomic<int64_t> g_tune[SRF_TUNE_KNOBS];

bool tune_in_range(int knob, int64_t v) {
  if (v == 0) return true;
  switch (knob) {
    case SRF_TUNE_DW_CHUNK: return v >= 2048 && v <= 262144 && v % 256 == 0;
    case SRF_TUNE_COLOR_DW_CHUNK: return v >= 2048 && v <= 65536 && v % 256 == 0;
    case SRF_TUNE_GATHER_BLOCKS: return v >= 256 && v <= 65536;
    case SRF_TUNE_GATHER_RUN: return v >= 1 && v <= 256;
    default: return false;
  }
}
}  // namespace

int64_t tune_value(int knob) {
  return knob >= 0 && knob < SRF_TUNE_KNOBS ? g_tune[knob].load(std::memory_order_relaxed) : 0;
}

namespace {

// One wave computes D = A B for a single 16x16 tile with the operand layouts of mfma.h
// (slot (g, i) = logical k index 8 g + i for 16-bit types, g for fp32).
template <int P>
__global__ __launch_bounds__(64) void selftest_mfma_kernel(const float* __restrict__ a,
                                                           const float* __restrict__ b,
                                                           float* __restrict__ d) {
  const int l = threadIdx.x;
  const int g = l / 16, j = l % 16;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  if constexpr (P == kF32) {
    acc = mfma32(a[j * 4 + g], b[g * 16 + j], acc);
  } else {
    u32x4 av, bv;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = 8 * g + 2 * q;
      av[q] = pack2<P>(a[j * 32 + k], a[j * 32 + k + 1]);
      bv[q] = pack2<P>(b[k * 16 + j], b[(k + 1) * 16 + j]);
    }
    acc = mfma16<P>(av, bv, acc);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) d[(4 * g + r) * 16 + j] = acc[r];
}

// Four waves each copy two 1 KB fragments into LDS with lds_dma_1k; the block then writes LDS
// back out in reversed fragment order, so dst fragment f = src fragment 7 - f.
__global__ __launch_bounds__(256) void selftest_lds_dma_kernel(const uint8_t* __restrict__ src,
                                                               uint8_t* __restrict__ dst) {
  __shared__ __attribute__((aligned(16))) uint8_t buf[8 * 1024];
  const int w = threadIdx.x >> 6;
  lds_dma_1k(src + (2 * w) * 1024, buf + (2 * w) * 1024);
  lds_dma_1k(src + (2 * w + 1) * 1024, buf + (2 * w + 1) * 1024);
  wait_vmem();
  __syncthreads();
  for (int i = threadIdx.x; i < 8 * 64; i += 256) {
    const int f = i / 64, l = i % 64;
    *reinterpret_cast<u32x4*>(dst + (7 - f) * 1024 + 16 * l) = *reinterpret_cast<const u32x4*>(buf + f * 1024 + 16 * l);
  }
}

}  // namespace
}  // namespace srf

using namespace srf;

extern "C" int srf_tune_set(int knob, int64_t value) {
  if (knob < 0 || knob >= SRF_TUNE_KNOBS || !tune_in_range(knob, value)) return SRF_EINVAL;
  g_tune[knob].store(value, std::memory_order_relaxed);
  return SRF_OK;
}

extern "C" int64_t srf_tune_get(int knob) {
  if (knob < 0 || knob >= SRF_TUNE_KNOBS) return -1;
  return g_tune[knob].load(std::memory_order_relaxed);
}

extern "C" int srf_selftest_lds_dma(const void* src, void* dst, void* stream) {
  if (!src || !dst) return SRF_EINVAL;
  hipLaunchKernelGGL(selftest_lds_dma_kernel, dim3(1), dim3(256), 0, as_stream(stream),
                     reinterpret_cast<const uint8_t*>(src), reinterpret_cast<uint8_t*>(dst));
  return launch_status();
}

extern "C" int srf_abi_version(void) { return SRF_ABI_VERSION; }

extern "C" const char* srf_status_string(int status) {
  switch (status) {
    case SRF_OK: return "ok";
    case SRF_EHIP: return "HIP runtime or launch error";
    case SRF_EINVAL: return "invalid argument (size, null pointer or option)";
    case SRF_EUNSUPPORTED: return "shape outside what the specialised kernels support";
    case SRF_ETOOLARGE: return "input beyond the fixed scratch of an allocation-free entry point";
    default: return "unknown status";
  }
}

extern "C" int srf_selftest_mfma(int prec, const float* a, const float* b, float* d, void* stream) {
  if (!a || !b || !d) return SRF_EINVAL;
  switch (prec) {
    case kF32:
      hipLaunchKernelGGL(selftest_mfma_kernel<kF32>, dim3(1), dim3(64), 0, as_stream(stream), a, b, d);
      break;
    case kBF16:
      hipLaunchKernelGGL(selftest_mfma_kernel<kBF16>, dim3(1), dim3(64), 0, as_stream(stream), a, b, d);
      break;
    case kF16:
      hipLaunchKernelGGL(selftest_mfma_kernel<kF16>, dim3(1), dim3(64), 0, as_stream(stream), a, b, d);
      break;
    default: return SRF_EINVAL;
  }
  return launch_status();
}
