// MFMA building blocks for gfx950 (CDNA4), 16x16 output tiles.
//
// Lane layout used throughout this library (64-lane wave, g = lane / 16, j = lane % 16):
//   C/D tile (16x16 fp32, 4 VGPRs):  D[4g + r][j] in register r            (r = 0..3)
//   A operand (16 x K):              A[j][slot (g, i)]
//   B operand (K x 16):              B[slot (g, i)][j]
// where a "slot" (g, i) is one of the K reduction positions held by lane group g; A and B
// hold the same K positions in the same registers, so any fixed map from slots to logical
// reduction indices gives the same product.  The kernels choose that map so that a D tile
// produced by one layer can be fed to the next layer as its B operand without moving data:
//   16-bit types (16x16x32, 8 values per lane): slot (g, i) of k-step s is logical feature
//     32 s + 4 g + i          for i < 4   (D tile 2s,   register i)
//     32 s + 16 + 4 g + i - 4 for i >= 4  (D tile 2s+1, register i-4)
//   fp32 (16x16x4, 1 value per lane): k-step s = 4 t + r uses lane group g's value
//     feature 16 t + 4 g + r (D tile t, register r).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace srf {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// kBFS3 / kBFS2: split-BF16 operands.  Activations, stored tiles and accumulators are FP32 (the
// kF32 layouts); each MFMA operand x is split into BF16 planes x = x_0 + x_1 (+ x_2), every plane
// the BF16 rounding of what the previous planes leave, and a product is the sum of the plane
// products x_i y_j with i + j < planes (16x16x32 BF16 MFMAs into one FP32 accumulator).  Three
// planes (six products) carry 24 significand bits, as FP32; two planes (three products) 16.
// kBFS1: one plane, i.e. BF16 operands rounded from FP32 activations and weights (FP32 storage).
enum Prec { kF32 = 0, kBF16 = 1, kF16 = 2, kBFS3 = 3, kBFS2 = 4, kBFS1 = 5 };

template <int P>
__host__ __device__ constexpr bool f32_layout() { return P == kF32 || P == kBFS3 || P == kBFS2 || P == kBFS1; }
// BF16 operand planes of the FP32-layout precisions
template <int P>
__host__ __device__ constexpr bool split_prec() { return P == kBFS3 || P == kBFS2 || P == kBFS1; }
template <int P>
__host__ __device__ constexpr int split_planes() { return P == kBFS3 ? 3 : (P == kBFS2 ? 2 : 1); }

// Round-to-nearest-even float -> bf16 bits (NaN kept quiet).
__device__ __forceinline__ uint32_t f32_to_bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return ((u >> 16) | 0x40u) & 0xffffu;
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

__device__ __forceinline__ float bf16_bits_to_f32(uint32_t b) { return __uint_as_float(b << 16); }

__device__ __forceinline__ uint32_t f32_to_f16_bits(float f) {
  _Float16 h = (_Float16)f;
  return (uint32_t)__builtin_bit_cast(uint16_t, h);
}

__device__ __forceinline__ float f16_bits_to_f32(uint32_t b) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)(b & 0xffffu));
}

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// Two floats -> packed 16-bit pair (round to nearest even; gfx950 v_cvt_pk_bf16_f32 for BF16).
template <int P>
__device__ __forceinline__ uint32_t pack2(float lo, float hi) {
  if constexpr (P == kBF16) {
    const bf16x2_t v = {(__bf16)lo, (__bf16)hi};
    return __builtin_bit_cast(uint32_t, v);
  } else {
    return f32_to_f16_bits(lo) | (f32_to_f16_bits(hi) << 16);
  }
}

template <int P>
__device__ __forceinline__ float unpack_lo(uint32_t v) {
  if constexpr (P == kBF16) return bf16_bits_to_f32(v & 0xffffu);
  else return f16_bits_to_f32(v & 0xffffu);
}

template <int P>
__device__ __forceinline__ float unpack_hi(uint32_t v) {
  if constexpr (P == kBF16) return bf16_bits_to_f32(v >> 16);
  else return f16_bits_to_f32(v >> 16);
}

// One 16x16x32 MFMA on 16-bit operands held as 4 packed dwords.
template <int P>
__device__ __forceinline__ f32x4 mfma16(const u32x4& a, const u32x4& b, const f32x4& c) {
  if constexpr (P == kBF16)
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                  __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
}

__device__ __forceinline__ f32x4 mfma32(float a, float b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// BF16 planes of two floats (a -> low half, b -> high half of each packed dword): plane 0 is the
// round-to-nearest BF16 pair of (a, b), plane i + 1 the BF16 pair of what planes 0..i leave.  Each
// remainder is exact in FP32 (x minus its own BF16 rounding drops only the bits that rounding
// kept), so after three planes |x - sum| <= 2^-27 |x|, after two 2^-18 |x|.
template <int NP>
__device__ __forceinline__ void split_pair(float a, float b, uint32_t (&out)[NP]) {
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    const uint32_t h = pack2<kBF16>(a, b);
    out[i] = h;
    if (i + 1 < NP) {
      a = a - __uint_as_float(h << 16);
      b = b - __uint_as_float(h & 0xffff0000u);
    }
  }
}

// Direct global -> LDS copy of one 1 KB fragment by a whole wave (gfx950
// global_load_lds_dwordx4): lane l's 16 bytes at src + 16 l land at lds_base + 16 l.  src is
// the fragment base, lds_base must be wave-uniform.  Completion is tracked by vmcnt.
__device__ __forceinline__ void lds_dma_1k(const uint8_t* src, uint8_t* lds_base) {
  const int lane = threadIdx.x & 63;
  __builtin_amdgcn_global_load_lds((const void*)(src + 16 * lane),
                                   (__attribute__((address_space(3))) void*)lds_base, 16, 0, 0);
}

// Wait for every outstanding vector-memory operation of this wave (LDS DMA included).
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0); }

}  // namespace srf
