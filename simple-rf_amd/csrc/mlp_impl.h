// Fused positional encoding + NeRF MLP (kernel templates shared by mlp*.hip) (forward, backward, weight gradients) on gfx950 MFMA.
//
// Reference semantics: MLP.forward / get_view_independent_outputs / get_view_dependent_outputs
// and PositionalEncoder (src/models/SimpleNeRF17.py:581-785), driven by run_network/batchify
// (:419-484).  Three variants share the kernels (desc->view_dependent, desc->trunk_pe):
//   main              trunk PE 63,  head: feature(256) + sigma(1), view layer 283 -> 128 -> 3
//   points-augmented  trunk PE 21,  view layer input [feature, PE_xyz[21:63], PE_dir] = 325
//   views-augmented   trunk PE 63,  head: 4 channels (sigma, rgb) straight from the trunk
//
// Design (one 64-lane wave owns 16*CB points, CB = 4 for 16-bit operands, 2 for FP32):
//   * The MLP is evaluated "transposed": each layer computes H_out^T = W H_in^T with the
//     weights as the MFMA A operand and the points as the N dimension.  With the slot map of
//     mfma.h, the D tiles of one layer are, bit for bit, the B operand of the next layer, so
//     all activations stay in VGPRs from the positional encoding to the rgb head.
//   * Positional-encoding features are generated in registers in the same tile layout
//     (Cody-Waite sin/cos of x * 2^m, no HBM traffic) for layer 0 and the view layer; the skip
//     layer re-reads layer 0's PE tiles from LDS.
//   * Weights stream through double-buffered LDS, copied global -> LDS once per block
//     (gfx950 global_load_lds_dwordx4) and shared by its four waves.
//   * Weights are repacked once per parameter update (srf_mlp_pack) into per-lane 16-byte
//     fragments: one 1 KB fragment per (output tile, k-unit).
//   * Training stores the inputs of every layer (16-bit or FP32) for the weight gradients, in
//     16-feature tiles laid out as the GEMM's operand slices (store_tile: each lane writes one
//     coalesced 32 B record straight from its accumulator registers after one pair exchange),
//     plus one 64-bit ReLU mask per lane per layer for the backward chain.
//   * Backward chain: the same transposed product with W^T tiles walks from the heads to
//     layer 0, applying the stored ReLU masks, and writes every layer's pre-activation
//     gradient dZ feature-major.
//   * Weight gradients: one grouped GEMM launch dW = dZ X^T over all layers, K = points; a block
//     of eight waves reduces a whole 256 x 256 layer over a chunk of points (each dZ and X row
//     read once) and adds it with FP32 atomics.
#pragma once

#include <string.h>

#include <type_traits>

#include "common.h"
#include "mfma.h"

// Lambdas in the kernels must inline or the register arrays they touch spill to scratch.
#define SRF_AI __attribute__((always_inline))

namespace srf {
namespace mlp {

constexpr int kLayers = 11;        // kernel layers: trunk 0-7, head 8, view 9, rgb 10
constexpr int kParams = SRF_MLP_NUM_PARAMS;
constexpr int kMaskSlots = 9;      // ReLU masks of layers 0-7 and the view layer
constexpr int kWaves = 4;          // waves per block
constexpr int kMaxJobs = 16;
// padded biases of the largest model: trunk 8 x 256, head 272, view 128, rgb 16
constexpr int kMaxBiasFloats = 8 * 256 + 272 + 128 + 16;

// ----------------------------------------------------------------------------- plan
struct Plan {
  int prec, view_dep, softplus, trunk_pe, pts_deg, dir_deg;
  int extra;  // computed inputs of the view layer: [PE_xyz[trunk_pe:], PE_dir]
  int elem;   // bytes per stored value
  // bytes of the forward and of the transposed (backward) operand image per (output tile, input
  // tile): 16 x 16 values, times the BF16 planes of the split precision
  int fwd_tile_bytes, bwd_tile_bytes;
  int out_rows[kLayers], out_tiles[kLayers], reg_tiles[kLayers], comp_feats[kLayers],
      comp_tiles[kLayers], bwd_tiles[kLayers], present[kLayers];
  int bias_floats;  // padded biases of every present layer, contiguous from bias_off[0]
  size_t fwd_off[kLayers], bwd_off[kLayers], bias_off[kLayers], packed_bytes;
};

__host__ __device__ inline int round_up(int a, int b) { return (a + b - 1) / b * b; }

inline bool make_plan(const srf_mlp_desc* d, Plan* p) {
  if (!d) return false;
  memset(p, 0, sizeof(Plan));
  if (d->prec < SRF_PREC_F32 || d->prec > SRF_PREC_BF16_SPLIT) return false;
  if (d->pts_pe_degree < 0 || d->pts_pe_degree > 16 || d->dirs_pe_degree < 0 || d->dirs_pe_degree > 16)
    return false;
  const int pts_dim = 3 * (2 * d->pts_pe_degree + 1);
  if (d->trunk_pe < 3 || d->trunk_pe > pts_dim || d->trunk_pe > 64) return false;
  p->prec = d->prec;
  p->view_dep = d->view_dependent ? 1 : 0;
  p->softplus = d->sigma_softplus ? 1 : 0;
  p->trunk_pe = d->trunk_pe;
  p->pts_deg = d->pts_pe_degree;
  p->dir_deg = d->dirs_pe_degree;
  const bool split = d->prec == SRF_PREC_BF16_SPLIT;
  p->elem = (d->prec == SRF_PREC_F32 || split) ? 4 : 2;
  // split: forward image in three BF16 planes, transposed image in two
  p->fwd_tile_bytes = split ? 3 * 512 : 256 * p->elem;
  p->bwd_tile_bytes = split ? 2 * 512 : 256 * p->elem;
  p->extra = p->view_dep ? (pts_dim - d->trunk_pe) + 3 * (2 * d->dirs_pe_degree + 1) : 0;
  if (p->extra > 96) return false;
  if (!p->view_dep && pts_dim != d->trunk_pe) return false;
  const int pe_tiles = round_up(round_up(d->trunk_pe, 16) / 16, 2);
  const int ex_tiles = round_up(round_up(p->extra, 16) / 16, 2);
  for (int l = 0; l < kLayers; ++l) p->present[l] = (l <= 8) || p->view_dep;
  for (int l = 0; l < 8; ++l) {
    p->out_rows[l] = 256;
    p->reg_tiles[l] = l == 0 ? 0 : 16;
    p->comp_feats[l] = (l == 0 || l == 5) ? d->trunk_pe : 0;
    p->comp_tiles[l] = (l == 0 || l == 5) ? pe_tiles : 0;
  }
  p->out_rows[8] = p->view_dep ? 257 : 4;
  p->reg_tiles[8] = 16;
  if (p->view_dep) {
    p->out_rows[9] = 128;
    p->reg_tiles[9] = 16;
    p->comp_feats[9] = p->extra;
    p->comp_tiles[9] = ex_tiles;
    p->out_rows[10] = 3;
    p->reg_tiles[10] = 8;
  }
  size_t off = 0;
  for (int l = 0; l < kLayers; ++l) {
    if (!p->present[l]) continue;
    p->out_tiles[l] = (p->out_rows[l] + 15) / 16;
    p->bwd_tiles[l] = round_up(p->out_tiles[l], 2);
    const int in_tiles = p->reg_tiles[l] + p->comp_tiles[l];
    p->fwd_off[l] = off;
    off += (size_t)p->out_tiles[l] * in_tiles * p->fwd_tile_bytes;
  }
  for (int l = 1; l < kLayers; ++l) {
    if (!p->present[l]) continue;
    p->bwd_off[l] = off;
    off += (size_t)p->reg_tiles[l] * p->bwd_tiles[l] * p->bwd_tile_bytes;
  }
  for (int l = 0; l < kLayers; ++l) {
    if (!p->present[l]) continue;
    p->bias_off[l] = off;
    off += (size_t)p->out_tiles[l] * 16 * sizeof(float);
    p->bias_floats += p->out_tiles[l] * 16;
  }
  p->packed_bytes = round_up((int)off, 256);
  return true;
}

// Training workspace: arrays of 16-feature tiles (each the bytes of 16 rows of ldP points; the
// layout inside a tile is store_tile's), ldP a multiple of 256 points.
struct Ws {
  int cols;  // points processed: n rounded up to whole blocks of 256
  int ldP;   // row stride in points (cols, plus one block when that is an even number of blocks)
  size_t x0, h[8], f, ex, v, dz[kLayers], mask, bytes;  // byte offsets
  int dz_rows[kLayers];
};

inline Ws make_ws(const Plan& p, int n) {
  Ws w;
  memset(&w, 0, sizeof(w));
  w.cols = round_up(n > 0 ? n : 1, 256);  // whole blocks of 4 waves (64 or 32 points each)
  // Rows are spaced by an odd number of 256-point blocks, so they are not a multiple of 1 KB x a
  // power of two apart and the 16 rows a wave touches at once do not fall on the same HBM
  // channels.  Only the stride grows; the kernels still process `cols` points.
  w.ldP = (w.cols / 256) % 2 == 0 ? w.cols + 256 : w.cols;
  const size_t row = (size_t)w.ldP * p.elem;
  size_t off = 0;
  auto take = [&](int rows) {
    size_t o = off;
    off += (size_t)rows * row;
    off = (off + 255) / 256 * 256;
    return o;
  };
  w.x0 = take(p.comp_tiles[0] * 16);
  for (int l = 0; l < 8; ++l) w.h[l] = take(256);
  if (p.view_dep) {
    w.f = take(256);
    w.ex = take(p.comp_tiles[9] * 16);
    w.v = take(128);
  }
  for (int l = 0; l < kLayers; ++l) {
    if (!p.present[l]) continue;
    w.dz_rows[l] = p.out_tiles[l] * 16;
    w.dz[l] = take(w.dz_rows[l]);
  }
  w.mask = off;
  off += (size_t)kMaskSlots * w.ldP * 32;
  w.bytes = (off + 255) / 256 * 256;
  return w;
}

// The kernels address every workspace array (one per layer input, dZ and mask slot) with 32-bit
// byte offsets from the array's base, so each must stay below 4 GB: a training launch of up to
// about 3.9 M points (FP32) or 7.8 M points (16-bit operands).
inline bool ws_addressable(const Plan& p, const Ws& w) {
  int rows = 256;
  for (int l = 0; l < kLayers; ++l) rows = w.dz_rows[l] > rows ? w.dz_rows[l] : rows;
  return (uint64_t)rows * (uint64_t)w.ldP * (uint64_t)p.elem < (1ull << 32);
}

// --------------------------------------------------------------------- type traits
template <int P>
struct Tr;
template <>
struct Tr<kBF16> {
  static constexpr int CB = 4;
  using Act = u32x2;
};
template <>
struct Tr<kF16> {
  static constexpr int CB = 4;
  using Act = u32x2;
};
template <>
struct Tr<kF32> {
  static constexpr int CB = 2;
  using Act = f32x4;
};
template <>
struct Tr<kBFS3> : Tr<kF32> {};
template <>
struct Tr<kBFS2> : Tr<kF32> {};
template <>
struct Tr<kBFS1> : Tr<kF32> {};

template <int P>
__device__ __forceinline__ typename Tr<P>::Act to_act(const f32x4& v) {
  if constexpr (f32_layout<P>()) {
    return v;
  } else {
    u32x2 a;
    a.x = pack2<P>(v[0], v[1]);
    a.y = pack2<P>(v[2], v[3]);
    return a;
  }
}

template <int P>
__device__ __forceinline__ typename Tr<P>::Act zero_act() {
  typename Tr<P>::Act a;
  if constexpr (f32_layout<P>()) a = f32x4{0.f, 0.f, 0.f, 0.f};
  else a = u32x2{0u, 0u};
  return a;
}

template <int P>
__device__ __forceinline__ uint32_t act_elem_bits(const typename Tr<P>::Act& a, int r) {
  if constexpr (f32_layout<P>()) return __float_as_uint(a[r]);
  else return (r < 2 ? a.x : a.y) >> (16 * (r & 1)) & 0xffffu;
}

// A wave's 16 activation tiles x CB column blocks in registers, one Act per tile (ActTiles) or,
// for 16-bit operands, tiles 2u and 2u + 1 of a column block in one 128-bit register quad
// (ActQuads), which is exactly the MFMA B operand of k-unit u, so that no register copies
// assemble it.  Measured on gfx950: the quads make the training forward 8 % faster and the
// inference forward 20 % slower (register allocation differs), so each uses its own.  Indices
// are compile-time constants after unrolling (the arrays stay in VGPRs).
template <int P>
struct ActTiles {
  using Act = typename Tr<P>::Act;
  Act q[16][Tr<P>::CB];
  __device__ SRF_AI Act get(int t, int cb) const { return q[t][cb]; }
  __device__ SRF_AI void set(int t, int cb, Act a) { q[t][cb] = a; }
};
template <int P>
struct ActQuads {
  using Act = typename Tr<P>::Act;
  u32x4 q[8][Tr<P>::CB];
  __device__ SRF_AI Act get(int t, int cb) const {
    const u32x4 v = q[t >> 1][cb];
    return (t & 1) ? u32x2{v.z, v.w} : u32x2{v.x, v.y};
  }
  __device__ SRF_AI void set(int t, int cb, Act a) {
    if (t & 1) {
      q[t >> 1][cb].z = a.x;
      q[t >> 1][cb].w = a.y;
    } else {
      q[t >> 1][cb].x = a.x;
      q[t >> 1][cb].y = a.y;
    }
  }
};
template <>
struct ActQuads<kF32> : ActTiles<kF32> {};
template <>
struct ActQuads<kBFS3> : ActTiles<kBFS3> {};
template <>
struct ActQuads<kBFS2> : ActTiles<kBFS2> {};

// sin and cos of x for |x| up to ~1e4 (positional-encoding arguments are |2^m x| < 1e3):
// Cody-Waite reduction x = k pi/2 + r with pi/2 split into three floats applied by fma, then
// Taylor polynomials on |r| <= pi/4 (truncation error below 2e-10, i.e. ~1 ulp overall).
__device__ __forceinline__ void pe_sincos(float x, float* sn, float* cs) {
  const float kf = rintf(x * 0.636619772367581343f);
  float r = fmaf(-kf, 1.57079637050628662f, x);
  r = fmaf(-kf, -4.37113882867379e-08f, r);
  r = fmaf(-kf, -1.77635683940025e-15f, r);
  const float r2 = r * r;
  float sp = fmaf(r2, 2.75573192239859e-06f, -1.98412698412698e-04f);
  sp = fmaf(r2, sp, 8.33333333333333e-03f);
  sp = fmaf(r2, sp, -1.66666666666667e-01f);
  const float sr = fmaf(r * r2, sp, r);
  float cp = fmaf(r2, -2.75573192239859e-07f, 2.48015873015873e-05f);
  cp = fmaf(r2, cp, -1.38888888888889e-03f);
  cp = fmaf(r2, cp, 4.16666666666667e-02f);
  cp = fmaf(r2, cp, -0.5f);
  const float cr = fmaf(r2, cp, 1.0f);
  const int q = ((int)kf) & 3;
  *sn = q == 0 ? sr : (q == 1 ? cr : (q == 2 ? -sr : -cr));
  *cs = q == 0 ? cr : (q == 1 ? -sr : (q == 2 ? -cr : sr));
}

// Positional encoding (degree L, include_input, log sampling) of x = (x0, x1, x2):
// features [x, sin(2^0 x), cos(2^0 x), ..., sin(2^(L-1) x), cos(2^(L-1) x)] in blocks of 3
// coordinates, i.e. feature c is x_c, feature 3 + 6m + c is sin(2^m x_c) and 6 + 6m + c is
// cos(2^m x_c).  Computed point-major: each lane evaluates every feature of one of the wave's
// points into that point's LDS row (PeRows), one pe_sincos call per (m, c) giving both the sin
// and the cos feature; the lanes then gather their B-operand slices from the rows.
template <int P>
struct PeRows {
  static constexpr int kNpt = 16 * Tr<P>::CB;  // points per wave
  static constexpr int kParts = 64 / kNpt;     // lanes per point (FP32: two, levels split)
  static constexpr int kStride = 100;          // elements per row: up to 96 features, and
                                               // conflict-free 8 (16)-byte gathers
  static constexpr int kBytes = kNpt * kStride * (f32_layout<P>() ? 4 : 2);
  using E = std::conditional_t<f32_layout<P>(), float, uint16_t>;

  __device__ static SRF_AI E cvt(float v) {
    if constexpr (f32_layout<P>()) return v;
    else return (uint16_t)(pack2<P>(v, 0.f) & 0xffffu);
  }
  // Features f in [f_lo, f_hi) of PE_L(x), written at row[f - shift]; this lane computes the
  // levels m = part, part + kParts, ...
  __device__ static SRF_AI void write(E* row, const float (&x)[3], int L, int f_lo, int f_hi, int shift, int part) {
    if (part == 0)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        if (c >= f_lo && c < f_hi) row[c - shift] = cvt(x[c]);
    for (int m = part; m < L; m += kParts) {
      if (8 + 6 * m < f_lo || 3 + 6 * m >= f_hi) continue;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float sn, cs;
        pe_sincos(ldexpf(x[c], m), &sn, &cs);
        const int fs = 3 + 6 * m + c, fc = 6 + 6 * m + c;
        if (fs >= f_lo && fs < f_hi) row[fs - shift] = cvt(sn);
        if (fc >= f_lo && fc < f_hi) row[fc - shift] = cvt(cs);
      }
    }
  }
  // B-operand slice of 16-column tile t for column block cb: columns 16 t + 4 g + r (r = 0..3)
  // of point 16 cb + j's row, zero at or beyond n_valid.
  __device__ static SRF_AI typename Tr<P>::Act gather(const E* rows, int t, int cb, int n_valid) {
    const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
    const int c0 = 16 * t + 4 * g;
    const E* src = rows + (16 * cb + j) * kStride + c0;
    if constexpr (f32_layout<P>()) {
      f32x4 v = *reinterpret_cast<const f32x4*>(src);
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = c0 + r < n_valid ? v[r] : 0.f;
      return v;
    } else {
      u32x2 v = *reinterpret_cast<const u32x2*>(src);
      v.x &= (c0 + 0 < n_valid ? 0x0000ffffu : 0u) | (c0 + 1 < n_valid ? 0xffff0000u : 0u);
      v.y &= (c0 + 2 < n_valid ? 0x0000ffffu : 0u) | (c0 + 3 < n_valid ? 0xffff0000u : 0u);
      return v;
    }
  }
};

// Orders a wave's LDS writes before its later LDS reads of other lanes' entries.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------- kernel arguments
struct FwdArgs {
  Plan p;
  Ws w;
  const uint8_t* packed;
  int n;
  const float* pts;
  const float* dirs;
  const float* noise;
  float* sigma;
  float* rgb;
  uint8_t* ws;  // null for inference
  // ray form (z non-null): point p is rays_o[r] + rays_d[r] * z[p] of ray r = p / n_samples, and
  // dirs holds one view direction per ray
  const float* rays_o;
  const float* rays_d;
  const float* z;
  int n_samples;
};

struct BwdArgs {
  Plan p;
  Ws w;
  const uint8_t* packed;
  int n;
  const float* sigma;
  const float* rgb;
  const float* d_sigma;
  const float* d_rgb;
  uint8_t* ws;
};

// -------------------------------------------------------------- generic layer product
// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).  Register arrays are
// only ever indexed with these constants, so they stay in VGPRs.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Weights are streamed through LDS: a "k-unit" is the 1 KB of packed A that one output tile
// needs for 32 reduction features (16-bit operands, one k-step) or 16 features (FP32, four
// k-steps).  Chunks of kKC k-units for all output tiles are copied global -> LDS by the four
// waves of the block (lds_dma_1k), double-buffered so the copy of chunk c+1 overlaps the MFMAs
// of chunk c; every wave then reads the fragments it needs from LDS.  This divides the weight
// traffic out of L2 by the number of waves per block and hides its latency.
constexpr int kKC = 2;
constexpr int kStageBytes = 16 * kKC * 1024;
// Split-BF16 operands: a k-unit's fragment is one 1 KB plane per BF16 plane (3 KB forward, 2 KB
// backward), and a chunk holds one k-unit when there are three planes, so that one buffer of the
// stage is 48 KB (forward) or 64 KB (backward, two k-units).
template <int P>
constexpr int kc_of() { return split_planes<P>() == 3 ? 1 : kKC; }
template <int P>
constexpr int stage_half_bytes() { return 16 * kc_of<P>() * split_planes<P>() * 1024; }

template <int P>
__host__ __device__ constexpr int units(int tiles) { return P == kF32 ? tiles : tiles / 2; }

// s_waitcnt immediate (gfx9 encoding) that waits for at most n outstanding vector-memory
// operations and leaves expcnt / lgkmcnt unconstrained.
constexpr int vmcnt_only(int n) { return 0x0F70 | (n & 0xF) | ((n >> 4) << 14); }

// DMA of weight chunk C of a product (kKC k-units of every one of its OUT output tiles, in the
// stage layout mma_tiles reads) into buf, issued by the block's waves together.
template <int P, int NT, int OUT, int TB, int C>
__device__ __forceinline__ void stage_chunk(const uint8_t* __restrict__ A, int upr, int u0, uint8_t* buf) {
  constexpr int NU = NT / (P == kF32 ? 1 : 2);
  constexpr int KC = kc_of<P>(), NPL = split_planes<P>();
  constexpr int kn = (NU - C * KC) < KC ? (NU - C * KC) : KC;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
  for (int i = 0; i < (OUT * kn * NPL + kWaves - 1) / kWaves; ++i) {
    const int f = w + kWaves * i;
    if (f < OUT * kn * NPL) {
      // fragment f: plane f % NPL of k-unit kc of output tile to
      const int fk = f / NPL, pl = f - fk * NPL;
      const int to = fk / kn, kc = fk - (fk / kn) * kn;
      const int frag = __builtin_amdgcn_readfirstlane(((TB + to) * upr + u0 + C * KC + kc) * NPL + pl);
      const int dst = __builtin_amdgcn_readfirstlane(((to * KC + kc) * NPL + pl) * 1024);
      lds_dma_1k(A + (size_t)frag * 1024, buf + dst);
    }
  }
}

// s_waitcnt vmcnt(n) for a wave-uniform n in 0..15 (the immediate has to be a constant).
__device__ __forceinline__ void wait_vmem_at_most(int n) {
  switch (n) {
    case 0: __builtin_amdgcn_s_waitcnt(vmcnt_only(0)); break;
    case 1: __builtin_amdgcn_s_waitcnt(vmcnt_only(1)); break;
    case 2: __builtin_amdgcn_s_waitcnt(vmcnt_only(2)); break;
    case 3: __builtin_amdgcn_s_waitcnt(vmcnt_only(3)); break;
    case 4: __builtin_amdgcn_s_waitcnt(vmcnt_only(4)); break;
    case 5: __builtin_amdgcn_s_waitcnt(vmcnt_only(5)); break;
    case 6: __builtin_amdgcn_s_waitcnt(vmcnt_only(6)); break;
    case 7: __builtin_amdgcn_s_waitcnt(vmcnt_only(7)); break;
    case 8: __builtin_amdgcn_s_waitcnt(vmcnt_only(8)); break;
    case 9: __builtin_amdgcn_s_waitcnt(vmcnt_only(9)); break;
    case 10: __builtin_amdgcn_s_waitcnt(vmcnt_only(10)); break;
    case 11: __builtin_amdgcn_s_waitcnt(vmcnt_only(11)); break;
    case 12: __builtin_amdgcn_s_waitcnt(vmcnt_only(12)); break;
    case 13: __builtin_amdgcn_s_waitcnt(vmcnt_only(13)); break;
    case 14: __builtin_amdgcn_s_waitcnt(vmcnt_only(14)); break;
    default: __builtin_amdgcn_s_waitcnt(vmcnt_only(15)); break;
  }
}

// No deferred stores (see mma_tiles).
struct NoSide {
  template <typename I>
  __device__ SRF_AI void operator()(I) const {}
};

// acc[OUT][CB] += W[TB .. TB+OUT-1, k-units u0 .. u0+NT/TPU-1] x In for all OUT output tiles at
// once (the accumulators fill the AGPR file while the activations stay in arch VGPRs).  upr is
// the number of k-units per packed row of the layer; fin(t, cb) returns input tile t (counted
// from the first tile of k-unit u0) for column block cb.  stage holds 2 x ST x kKC KB.  Must be
// reached by all waves of the block (it synchronises them).
//
// Deferred stores: side(i) for i in [0, NSIDE) (compile-time indices) issues at least two
// vector-memory stores each (the previous layer's activation tiles and ReLU masks).  The items
// are spread over the weight chunks, issued right after the DMA of the next chunk, and each
// chunk then waits with vmcnt(stores issued after that DMA): vector-memory operations complete
// in issue order, so the DMA has landed while the stores keep draining behind the MFMAs.  Without
// this, a layer's stores would all be waited for at the next chunk boundary, when every wave of
// the chip issues them at once.
template <int P, int NT, int OUT, int TB, int ST = 16, int NSIDE = 0, typename FIn, typename Side = NoSide>
__device__ __forceinline__ void mma_tiles(f32x4 (&acc)[OUT][Tr<P>::CB], const uint8_t* __restrict__ A,
                                          int upr, int u0, FIn&& fin, uint8_t* stage, Side&& side = Side{}) {
  constexpr int CB = Tr<P>::CB;
  constexpr int TPU = P == kF32 ? 1 : 2;
  constexpr int NU = NT / TPU;
  constexpr int KC = kc_of<P>(), NPL = split_planes<P>();
  constexpr int NCH = (NU + KC - 1) / KC;
  constexpr int kHalf = ST * KC * NPL * 1024;  // one buffer of the double-buffered stage (2 kHalf bytes)
  static_assert(OUT <= ST, "the stage buffer holds ST output tiles");
  const int lane = threadIdx.x & 63;
  stage_chunk<P, NT, OUT, TB, 0>(A, upr, u0, stage);
  wait_vmem();
  __syncthreads();
  constexpr int kPer = (NSIDE + NCH - 1) / NCH;  // deferred-store items per chunk
  static_for<0, NCH>([&](auto CI) SRF_AI {
    constexpr int c = decltype(CI)::value;
    const uint8_t* cur = stage + (c & 1) * kHalf;
    if constexpr (c + 1 < NCH) stage_chunk<P, NT, OUT, TB, c + 1>(A, upr, u0, stage + ((c + 1) & 1) * kHalf);
    constexpr int i0 = c * kPer < NSIDE ? c * kPer : NSIDE;
    constexpr int i1 = (c + 1) * kPer < NSIDE ? (c + 1) * kPer : NSIDE;
    if constexpr (i1 > i0) {
      // the stores must stay after the DMA above and before the wait below; one item at a time
      // (scheduling barriers), so that only one item's store data is live at once
      asm volatile("" ::: "memory");
      static_for<i0, i1>([&](auto II) SRF_AI {
        __builtin_amdgcn_sched_barrier(0);
        side(II);
        __builtin_amdgcn_sched_barrier(0);
      });
      asm volatile("" ::: "memory");
    }
    constexpr int kn = (NU - c * KC) < KC ? (NU - c * KC) : KC;
    static_for<0, kn>([&](auto KI) SRF_AI {
      constexpr int kc = decltype(KI)::value;
      constexpr int u = c * KC + kc;
      if constexpr (P == kF32) {
        f32x4 b[CB];
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) b[cb] = fin(u, cb);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
#pragma unroll
          for (int to = 0; to < OUT; ++to) {
            const float a = *reinterpret_cast<const float*>(cur + (to * KC + kc) * 1024 + (r * 64 + lane) * 4);
#pragma unroll
            for (int cb = 0; cb < CB; ++cb) acc[to][cb] = mfma32(a, b[cb][r], acc[to][cb]);
          }
        }
      } else if constexpr (split_prec<P>()) {
        // Split BF16: the FP32 input tiles 2u, 2u + 1 of each column block become NPL BF16 B
        // operands (slot order of the 16-bit map), the weight fragments are already split into
        // NPL planes (stage layout: plane pl of (to, kc) at ((to KC + kc) NPL + pl) KB), and each
        // output tile accumulates the NPL (NPL + 1) / 2 plane products i + j < NPL.
        u32x4 b[NPL][CB];
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
          const f32x4 x0 = fin(2 * u, cb), x1 = fin(2 * u + 1, cb);
          uint32_t q0[NPL], q1[NPL], q2[NPL], q3[NPL];
          split_pair<NPL>(x0[0], x0[1], q0);
          split_pair<NPL>(x0[2], x0[3], q1);
          split_pair<NPL>(x1[0], x1[1], q2);
          split_pair<NPL>(x1[2], x1[3], q3);
#pragma unroll
          for (int i = 0; i < NPL; ++i) b[i][cb] = u32x4{q0[i], q1[i], q2[i], q3[i]};
        }
        // output tiles in groups of two (the A planes of the next group are read while this
        // group's MFMAs run; two tiles x two column blocks keep four independent accumulators
        // between dependent MFMAs)
        constexpr int kG = 2;
        constexpr int NG = (OUT + kG - 1) / kG;
        auto read_a = [&](auto GI, u32x4 (&a)[kG][NPL]) SRF_AI {
          constexpr int gi = decltype(GI)::value;
#pragma unroll
          for (int q = 0; q < kG; ++q)
            if (gi * kG + q < OUT)
#pragma unroll
              for (int pl = 0; pl < NPL; ++pl)
                a[q][pl] = *reinterpret_cast<const u32x4*>(cur + (((gi * kG + q) * KC + kc) * NPL + pl) * 1024 + lane * 16);
        };
        u32x4 a_cur[kG][NPL], a_nxt[kG][NPL];
        read_a(std::integral_constant<int, 0>{}, a_cur);
        static_for<0, NG>([&](auto GI) SRF_AI {
          constexpr int gi = decltype(GI)::value;
          if constexpr (gi + 1 < NG) read_a(std::integral_constant<int, gi + 1>{}, a_nxt);
          // smallest products first: (i, j) ordered by i + j descending
#pragma unroll
          for (int sum = NPL - 1; sum >= 0; --sum)
#pragma unroll
            for (int i = 0; i <= sum; ++i)
#pragma unroll
              for (int q = 0; q < kG; ++q)
                if (gi * kG + q < OUT)
#pragma unroll
                  for (int cb = 0; cb < CB; ++cb)
                    acc[gi * kG + q][cb] = mfma16<kBF16>(a_cur[q][i], b[sum - i][cb], acc[gi * kG + q][cb]);
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int q = 0; q < kG; ++q)
#pragma unroll
            for (int pl = 0; pl < NPL; ++pl) a_cur[q][pl] = a_nxt[q][pl];
        });
      } else {
        u32x4 b[CB];
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
          const u32x2 x0 = fin(2 * u, cb), x1 = fin(2 * u + 1, cb);
          b[cb] = u32x4{x0.x, x0.y, x1.x, x1.y};
        }
        // Output tiles in groups of kG: the A fragments of group g+1 are read from LDS while
        // the MFMAs of group g run; scheduling barriers keep the compiler from hoisting every
        // fragment read of the chunk (which would exhaust the arch VGPRs).
        constexpr int kG = 4;
        constexpr int NG = (OUT + kG - 1) / kG;
        auto read_a = [&](auto GI, u32x4 (&a)[kG]) SRF_AI {
          constexpr int gi = decltype(GI)::value;
#pragma unroll
          for (int q = 0; q < kG; ++q)
            if (gi * kG + q < OUT)
              a[q] = *reinterpret_cast<const u32x4*>(cur + ((gi * kG + q) * KC + kc) * 1024 + lane * 16);
        };
        u32x4 a_cur[kG], a_nxt[kG];
        read_a(std::integral_constant<int, 0>{}, a_cur);
        static_for<0, NG>([&](auto GI) SRF_AI {
          constexpr int gi = decltype(GI)::value;
          if constexpr (gi + 1 < NG) read_a(std::integral_constant<int, gi + 1>{}, a_nxt);
#pragma unroll
          for (int q = 0; q < kG; ++q)
            if (gi * kG + q < OUT)
#pragma unroll
              for (int cb = 0; cb < CB; ++cb)
                acc[gi * kG + q][cb] = mfma16<P>(a_cur[q], b[cb], acc[gi * kG + q][cb]);
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int q = 0; q < kG; ++q) a_cur[q] = a_nxt[q];
        });
      }
    });
    if constexpr (NSIDE == 0) {
      wait_vmem();
      __syncthreads();
    } else {
      // the next chunk's DMA (issued before this chunk's 2 (i1 - i0) or more stores) has landed
      // once at most that many operations are outstanding; the barrier then also frees the
      // buffer this chunk read (its LDS reads were consumed by the MFMAs above)
      if constexpr (c + 1 < NCH) __builtin_amdgcn_s_waitcnt(vmcnt_only(2 * (i1 - i0)));
      asm volatile("" ::: "memory");
      __builtin_amdgcn_s_barrier();
      asm volatile("" ::: "memory");
    }
  });
}

template <int P, int OUT>
__device__ __forceinline__ void zero_acc(f32x4 (&acc)[OUT][Tr<P>::CB]) {
#pragma unroll
  for (int to = 0; to < OUT; ++to)
#pragma unroll
    for (int cb = 0; cb < Tr<P>::CB; ++cb) acc[to][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
}

// Accumulators start at the bias (rows TB.. of the layer's padded FP32 bias): the bias then
// lives in the accumulator registers instead of being loaded in the epilogue.
template <int P, int OUT, int TB>
__device__ __forceinline__ void bias_acc(f32x4 (&acc)[OUT][Tr<P>::CB], const float* bias) {
  const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int to = 0; to < OUT; ++to) {
    const f32x4 b = *reinterpret_cast<const f32x4*>(bias + 16 * (TB + to) + 4 * g);
#pragma unroll
    for (int cb = 0; cb < Tr<P>::CB; ++cb) acc[to][cb] = b;
  }
}

// One whole layer with IN input tiles: accumulate (from the bias if `bias` is non-null), then
// epi(T, acc) for every output tile (after the last k-unit, so epi may overwrite the input
// tiles in place).  side / NSIDE: deferred stores issued during the product (mma_tiles).
template <int P, int IN, int OUT, int TB, int NSIDE = 0, typename FIn, typename Epi, typename Side = NoSide>
__device__ __forceinline__ void run_layer(const uint8_t* __restrict__ A, FIn&& fin, Epi&& epi, uint8_t* stage,
                                          const float* bias = nullptr, Side&& side = Side{}) {
  f32x4 acc[OUT][Tr<P>::CB];
  if (bias) bias_acc<P, OUT, TB>(acc, bias);
  else zero_acc<P, OUT>(acc);
  mma_tiles<P, IN, OUT, TB, 16, NSIDE>(acc, A, units<P>(IN), 0, fin, stage, side);
#pragma unroll
  for (int to = 0; to < OUT; ++to) epi(TB + to, acc[to]);
}

// Value of the same register in the neighbouring lane (lane ^ 1), by a DPP quad permutation
// [1, 0, 3, 2]: no index register and no LDS round trip, unlike a general lane shuffle.
__device__ __forceinline__ uint32_t swap_adjacent(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
}
__device__ __forceinline__ float swap_adjacent(float v) {
  return __builtin_bit_cast(float, swap_adjacent(__builtin_bit_cast(uint32_t, v)));
}

// Lane index recomputed at the point of use (v_mbcnt; volatile so that it is neither hoisted nor
// kept live across the layer loop, where register pressure would spill it to scratch and its
// reload's vmcnt wait would drain the deferred stores).
__device__ __forceinline__ uint32_t lane_here() {
  uint32_t r;
  asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(r));
  return r;
}
// odd lanes ? if_odd : if_even, with the lane parity as a scalar lane mask (no per-lane flag).
__device__ __forceinline__ uint32_t sel_odd(uint32_t if_odd, uint32_t if_even) {
  uint32_t r;
  asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(if_even), "v"(if_odd), "s"(0xaaaaaaaaaaaaaaaaull));
  return r;
}
__device__ __forceinline__ float sel_odd(float if_odd, float if_even) {
  return __builtin_bit_cast(float, sel_odd(__builtin_bit_cast(uint32_t, if_odd), __builtin_bit_cast(uint32_t, if_even)));
}

// Training stores go through raw buffer descriptors: the array base (the same in every lane) in
// scalar registers, the wave's block offset as the 32-bit scalar offset, and only the lane's byte
// offset in a vector register.  Per-lane 64-bit addresses for every tile would otherwise be kept
// live (or spilled) across the layer loop.  Descriptor word 3 0x00020000: 32-bit data format, as
// gfx9 raw buffer accesses require; the record count is the maximum (no clipping).  Offsets are
// 32-bit: the host keeps every workspace array below 4 GB.
using Rsrc = __amdgpu_buffer_rsrc_t;
constexpr int kBufNT = 2;  // cache-policy bit of the non-temporal store
__device__ __forceinline__ Rsrc uniform_rsrc(const void* base) {
  const uint64_t v = reinterpret_cast<uint64_t>(base);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(((uint64_t)hi << 32) | lo), (short)0, -1,
                                           0x00020000);
}
__device__ __forceinline__ void store_nt16(Rsrc r, uint32_t lane_off, uint32_t blk, u32x4 v) {
  __builtin_amdgcn_raw_buffer_store_b128(v, r, lane_off, blk, kBufNT);
}
__device__ __forceinline__ void store_nt8(Rsrc r, uint32_t lane_off, uint32_t blk, u32x2 v) {
  __builtin_amdgcn_raw_buffer_store_b64(v, r, lane_off, blk, kBufNT);
}
__device__ __forceinline__ u32x2 load8(Rsrc r, uint32_t lane_off, uint32_t blk) {
  return __builtin_amdgcn_raw_buffer_load_b64(r, lane_off, blk, 0);
}

// Byte offset, inside an array of 16-feature tiles, of tile t's 2 KB record block written by the
// wave owning points p0 .. p0 + 16 CB - 1 (p0 wave-uniform).
template <int P>
__device__ __forceinline__ uint32_t tile_block(int t, int ldP, int elem, int p0) {
  return (uint32_t)t * 16u * (uint32_t)ldP * (uint32_t)elem + (uint32_t)(p0 / (16 * Tr<P>::CB)) * 2048u;
}

// Writes one 16-feature tile of the wave's points in the paired tile layout read by the weight-
// gradient GEMM: the 2 KB block at byte offset blk (tile_block) of the array described by dst.
// A tile of the whole workspace array occupies the bytes of 16 feature-major rows;
// inside it, the wave owning points p0 .. p0 + 16 CB - 1 writes one 2 KB block of 64 lane records,
// each two 16 B chunks.  The lanes of a pair (j = 2m, 2m + 1) first swap half of their D-tile
// values (swap_adjacent), so that record (g, 2m + e) holds features 4 g + 2 e + c (c = 0, 1), each
// as one 16 B chunk of that feature at the points 16 cb + 2 m and 16 cb + 2 m + 1 (cb = 0 ..
// CB - 1), in point order.  A 16 B chunk is then exactly one lane's MFMA operand slice of one
// feature, so the GEMM reads fragments with plain 16 B copies.  The block is chunk-major: chunk c
// of record r sits at byte 1024 c + 16 r, so each of the two stores writes 1 KB of whole
// contiguous lines (a record-major block would leave every line half written between them).
// The stores are non-temporal: a layer's tiles of all the CUs of an XCD (~4 MB) would otherwise
// displace the streamed weights from that XCD's L2, and they are read back only by the backward.
template <int P>
__device__ __forceinline__ void store_tile(Rsrc dst, uint32_t blk, const typename Tr<P>::Act (&v)[Tr<P>::CB]) {
  constexpr int CB = Tr<P>::CB;
  const uint32_t rec = lane_here() * 16u;
  if constexpr (f32_layout<P>()) {
    // v[cb] = features 4g + 0..3 at point 16 cb + j; even lanes keep r0, r1 and take the
    // partner's; odd lanes keep r2, r3
    float rc[CB][2];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
      rc[cb][0] = swap_adjacent(sel_odd(v[cb][0], v[cb][2]));
      rc[cb][1] = swap_adjacent(sel_odd(v[cb][1], v[cb][3]));
    }
    f32x4 c0, c1;
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
      const float own0 = sel_odd(v[cb][2], v[cb][0]), own1 = sel_odd(v[cb][3], v[cb][1]);
      c0[2 * cb + 0] = sel_odd(rc[cb][0], own0);
      c0[2 * cb + 1] = sel_odd(own0, rc[cb][0]);
      c1[2 * cb + 0] = sel_odd(rc[cb][1], own1);
      c1[2 * cb + 1] = sel_odd(own1, rc[cb][1]);
    }
    store_nt16(dst, rec, blk, __builtin_bit_cast(u32x4, c0));
    store_nt16(dst, rec, blk + 1024, __builtin_bit_cast(u32x4, c1));
  } else {
    // v[cb] = {bf16 (r0, r1), bf16 (r2, r3)} at point 16 cb + j
    u32x4 c0, c1;
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
      const uint32_t recv = swap_adjacent(sel_odd(v[cb][0], v[cb][1]));
      const uint32_t own = sel_odd(v[cb][1], v[cb][0]);
      const uint32_t first = sel_odd(recv, own), second = sel_odd(own, recv);  // points 2m, 2m + 1
      c0[cb] = (first & 0xffffu) | (second << 16);
      c1[cb] = (first >> 16) | (second & 0xffff0000u);
    }
    store_nt16(dst, rec, blk, c0);
    store_nt16(dst, rec, blk + 1024, c1);
  }
}

// Inverse of store_tile for FP32 tiles (CB = 2): the lane's D-tile values (v[cb][r] = feature 4 g + r
// at point 16 cb + j) of one 16-feature tile, read from its 2 KB block.  The lane reads its own
// record (g, j): chunk c holds features 4 g + 2 (j & 1) + c at points j & ~1, j | 1 and 16 more;
// the lanes of a pair exchange the two features of the point each one lacks.
__device__ __forceinline__ void load_tile_f32(Rsrc src, uint32_t blk, f32x4 (&v)[2]) {
  const uint32_t rec = lane_here() * 16u;
  const f32x4 c0 = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(src, rec, blk, 0));
  const f32x4 c1 = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(src, rec, blk + 1024, 0));
  // odd lanes send their first point's values (positions 0, 2), even lanes their second's (1, 3)
  const float ka = sel_odd(c0[1], c0[0]), kb = sel_odd(c0[3], c0[2]);
  const float kc = sel_odd(c1[1], c1[0]), kd = sel_odd(c1[3], c1[2]);
  const float ra = swap_adjacent(sel_odd(c0[0], c0[1])), rb = swap_adjacent(sel_odd(c0[2], c0[3]));
  const float rc = swap_adjacent(sel_odd(c1[0], c1[1])), rd = swap_adjacent(sel_odd(c1[2], c1[3]));
  v[0] = f32x4{sel_odd(ra, ka), sel_odd(rc, kc), sel_odd(ka, ra), sel_odd(kc, rc)};
  v[1] = f32x4{sel_odd(rb, kb), sel_odd(rd, kd), sel_odd(kb, rb), sel_odd(kd, rd)};
}

__device__ __forceinline__ float sigmoidf_(float x) { return 1.0f / (1.0f + expf(-x)); }

__device__ __forceinline__ float softplusf_(float x) { return x > 20.f ? x : log1pf(expf(x)); }

// ------------------------------------------------------------------------ forward
template <int P, bool VD, bool TRAIN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void mlp_fwd_kernel(
    FwdArgs a) {
  constexpr int CB = Tr<P>::CB;
  using Act = typename Tr<P>::Act;
  __shared__ __attribute__((aligned(16))) uint8_t stage[2 * stage_half_bytes<P>()];
  // computed input features of each wave's points (trunk PE for layers 0 and 5, then the view
  // layer's [PE_xyz[trunk_pe:], PE_dir]) as B-operand tiles: 6 tiles x CB column blocks x 64
  // lanes.  Each lane reads back only its own entries, so no barrier is needed.
  __shared__ __attribute__((aligned(16))) Act pe_lds[kWaves][6][CB][64];
  // every layer's padded FP32 bias, copied once per block: the accumulators start from LDS
  // instead of waiting on global loads at each layer boundary
  __shared__ __attribute__((aligned(16))) float bias_lds[kMaxBiasFloats];
  const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: SGPR addresses
  const int wave = blockIdx.x * kWaves + wv;
  const int p0 = wave * 16 * CB;  // the grid covers cols exactly: every wave has points
  const Plan& pl = a.p;
  const int ldP = a.w.ldP;

  auto valid = [&](int cb) SRF_AI { return p0 + 16 * cb + j < a.n; };

  std::conditional_t<TRAIN, ActQuads<P>, ActTiles<P>> h;
  uint32_t mbits[CB][2];  // ReLU mask bits 4T+r of tiles 0-7 (word 0) and 8-15 (word 1)

  // trunk PE tiles -> pe_lds (training stores them for the weight gradients during layer 0)
  // Point-major PE rows (PeRows) in this wave's part of the weight stage, which is free before
  // layer 0 and between the feature and view layers; a block barrier follows the gathers,
  // before the next product's DMA refills the stage.
  using PR = PeRows<P>;
  static_assert(kWaves * PR::kBytes <= 2 * stage_half_bytes<P>(), "PE rows fit in the stage");
  typename PR::E* pe_rows = reinterpret_cast<typename PR::E*>(stage + wv * PR::kBytes);
  const int pe_pt = lane % PR::kNpt, pe_part = lane / PR::kNpt;
  auto load_pt = [&](const float* src, float (&v)[3]) SRF_AI {
    const int p = p0 + pe_pt;
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = p < a.n ? src[3 * p + c] : 0.f;
  };
  // the point's position: stored, or rays_o[..., None, :] + rays_d[..., None, :] * z[..., :, None]
  // of its ray (src/models/SimpleNeRF17.py:213-215, :266-268; two rounded FP32 operations)
  auto load_xyz = [&](float (&v)[3]) SRF_AI {
    if (a.z) {
      const int p = p0 + pe_pt;
      const bool in = p < a.n;
      const int r = in ? p / a.n_samples : 0;
      const float zz = in ? a.z[p] : 0.f;
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c] = in ? __fadd_rn(a.rays_o[3 * r + c], __fmul_rn(a.rays_d[3 * r + c], zz)) : 0.f;
    } else {
      load_pt(a.pts, v);
    }
  };
  // the point's view direction: stored per point, or its ray's
  // (viewdirs[:, None].expand(pts.shape), src/models/SimpleNeRF17.py:429-433)
  auto load_dir = [&](float (&v)[3]) SRF_AI {
    if (a.z) {
      const int p = p0 + pe_pt;
      const int r = p < a.n ? p / a.n_samples : 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c] = p < a.n ? a.dirs[3 * r + c] : 0.f;
    } else {
      load_pt(a.dirs, v);
    }
  };
  auto pe_trunk = [&]() SRF_AI {
    float x[3];
    load_xyz(x);
    PR::write(pe_rows + pe_pt * PR::kStride, x, pl.pts_deg, 0, pl.trunk_pe, 0, pe_part);
    wave_lds_sync();
    static_for<0, 4>([&](auto TI) SRF_AI {
      constexpr int t = decltype(TI)::value;
      if (t < pl.comp_tiles[0])
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) pe_lds[wv][t][cb][lane] = PR::gather(pe_rows, t, cb, pl.trunk_pe);
    });
    __syncthreads();
  };
  auto fin_pe = [&](int t, int cb) SRF_AI { return pe_lds[wv][t][cb][lane]; };
  auto mask_store = [&](Rsrc m) SRF_AI {
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
      store_nt8(m, lane_here() * 8u, (uint32_t)(wave * CB + cb) * 512u, u32x2{mbits[cb][0], mbits[cb][1]});
  };
  {
    const float* src = reinterpret_cast<const float*>(a.packed + pl.bias_off[0]);
    for (int i = threadIdx.x; i < pl.bias_floats; i += 64 * kWaves) bias_lds[i] = src[i];
  }  // (visible after pe_trunk's block barrier)
  auto bias_of = [&](int l) SRF_AI { return bias_lds + (pl.bias_off[l] - pl.bias_off[0]) / 4; };
  // ReLU epilogue: tile T into h[T] (in place) and its mask bits.  Nothing is stored here: the
  // training copies of h and mbits are written during the next layer's product (side_of).
  auto relu_epi = [&](int T, f32x4(&acc)[CB]) SRF_AI {
    if (T == 0) {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) mbits[cb][0] = mbits[cb][1] = 0;
    }
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
      f32x4 v = acc[cb];  // bias already in the accumulator
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const bool pos = v[r] > 0.f;
        v[r] = pos ? v[r] : 0.f;
        mbits[cb][T >> 3] |= pos ? (1u << (4 * (T & 7) + r)) : 0u;
      }
      h.set(T, cb, to_act<P>(v));
    }
  };
  // Deferred training stores (mma_tiles' side items) of the tiles the layer being computed
  // reads: item i < NH stores activation tile i of h at ws offset h_off, the next NPE items store
  // the pe_lds tiles at pe_off, and a last item (when mask_slot >= 0) stores the ReLU mask bits of
  // the layer that produced h.  Every item issues at least two vector-memory stores.
  constexpr int kT = TRAIN ? 1 : 0;
  auto side_of = [&](auto NHI, size_t h_off, auto NPEI, size_t pe_off, int mask_slot) SRF_AI {
    constexpr int nh = decltype(NHI)::value, npe = decltype(NPEI)::value;
    const Rsrc rh = uniform_rsrc(a.ws + h_off), rpe = uniform_rsrc(a.ws + pe_off);
    const Rsrc rm = uniform_rsrc(a.ws + a.w.mask + (size_t)(mask_slot < 0 ? 0 : mask_slot) * ldP * 32);
    return [&, rh, rpe, rm](auto II) SRF_AI {
      constexpr int i = decltype(II)::value;
      if constexpr (i < nh) {
        Act tl[CB];
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) tl[cb] = h.get(i, cb);
        store_tile<P>(rh, tile_block<P>(i, ldP, pl.elem, p0), tl);
      } else if constexpr (i < nh + npe) {
        constexpr int t = i - nh;
        Act tl[CB];
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) tl[cb] = pe_lds[wv][t][cb][lane];
        store_tile<P>(rpe, tile_block<P>(t, ldP, pl.elem, p0), tl);
      } else {
        mask_store(rm);
      }
    };
  };
  using I0 = std::integral_constant<int, 0>;
  using I16 = std::integral_constant<int, 16>;
  auto fin_h = [&](int t, int cb) SRF_AI { return h.get(t, cb); };
  auto fin_hc = [&](int t, int cb) SRF_AI { return t < 16 ? h.get(t, cb) : pe_lds[wv][t - 16][cb][lane]; };

  // ---- layer 0: PE(trunk) -> 256 (stores the PE tiles X0)
  pe_trunk();
  if (pl.comp_tiles[0] == 4)
    run_layer<P, 4, 16, 0, 4 * kT>(a.packed + pl.fwd_off[0], fin_pe, relu_epi, stage, bias_of(0),
                                   side_of(I0{}, 0, std::integral_constant<int, 4>{}, a.w.x0, -1));
  else
    run_layer<P, 2, 16, 0, 2 * kT>(a.packed + pl.fwd_off[0], fin_pe, relu_epi, stage, bias_of(0),
                                   side_of(I0{}, 0, std::integral_constant<int, 2>{}, a.w.x0, -1));

  // ---- layers 1..7, each in place in h; layer 5 adds the trunk PE k-steps after the 16
  // register tiles (one loop body for all seven layers keeps register pressure flat)
  // (stores layer l - 1's output and mask)
  for (int l = 1; l < 8; ++l) {
    f32x4 acc[16][CB];
    bias_acc<P, 16, 0>(acc, bias_of(l));
    const uint8_t* A = a.packed + pl.fwd_off[l];
    const int upr = units<P>(16 + pl.comp_tiles[l]);
    mma_tiles<P, 16, 16, 0, 16, 17 * kT>(acc, A, upr, 0, fin_h, stage, side_of(I16{}, a.w.h[l - 1], I0{}, 0, l - 1));
    if (l == 5) {
      if (pl.comp_tiles[5] == 4)
        mma_tiles<P, 4, 16, 0>(acc, A, upr, units<P>(16), fin_pe, stage);
      else
        mma_tiles<P, 2, 16, 0>(acc, A, upr, units<P>(16), fin_pe, stage);
    }
#pragma unroll
    for (int T = 0; T < 16; ++T) relu_epi(T, acc[T]);
  }

  // ---- heads
  auto head_sigma = [&](int cb, float raw) SRF_AI {
    const int p = p0 + 16 * cb + j;
    if (a.noise) raw = raw + a.noise[p];
    a.sigma[p] = pl.softplus ? softplusf_(raw) : (raw > 0.f ? raw : 0.f);
  };
  if constexpr (!VD) {
    // 4-channel output: sigma, rgb (src/models/SimpleNeRF17.py:735-750); stores layer 7
    run_layer<P, 16, 1, 0, 17 * kT>(a.packed + pl.fwd_off[8], fin_h, [&](int T, f32x4(&acc)[CB]) SRF_AI {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        const int p = p0 + 16 * cb + j;
        if (g == 0 && valid(cb)) {
          const f32x4 v = acc[cb];
          head_sigma(cb, v[0]);
          a.rgb[3 * p + 0] = sigmoidf_(v[1]);
          a.rgb[3 * p + 1] = sigmoidf_(v[2]);
          a.rgb[3 * p + 2] = sigmoidf_(v[3]);
        }
      }
    }, stage, bias_of(8), side_of(I16{}, a.w.h[7], I0{}, 0, 7));
  } else {
    // sigma (row 256, tile 16) first (storing layer 7), then feature rows 0..255 (identity) in
    // place in h
    run_layer<P, 16, 1, 16, 17 * kT>(a.packed + pl.fwd_off[8], fin_h, [&](int T, f32x4(&acc)[CB]) SRF_AI {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
        if (g == 0 && valid(cb)) head_sigma(cb, acc[cb][0]);
    }, stage, bias_of(8), side_of(I16{}, a.w.h[7], I0{}, 0, 7));
    run_layer<P, 16, 16, 0>(a.packed + pl.fwd_off[8], fin_h, [&](int T, f32x4(&acc)[CB]) SRF_AI {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) h.set(T, cb, to_act<P>(acc[cb]));
    }, stage, bias_of(8));
    // view layer input extras: [PE_xyz[trunk_pe:], PE_dir], one row per point (the feature
    // layer's product ended with a block barrier, so the stage is free)
    {
      const int pts_dim = 3 * (2 * pl.pts_deg + 1), n_hi = pts_dim - pl.trunk_pe;
      float x[3], d[3];
      load_xyz(x);
      load_dir(d);
      typename PR::E* row = pe_rows + pe_pt * PR::kStride;
      if (n_hi > 0) PR::write(row, x, pl.pts_deg, pl.trunk_pe, pts_dim, pl.trunk_pe, pe_part);
      PR::write(row, d, pl.dir_deg, 0, pl.extra - n_hi, -n_hi, pe_part);
      wave_lds_sync();
      static_for<0, 6>([&](auto TI) SRF_AI {
        constexpr int t = decltype(TI)::value;
        if (t < pl.comp_tiles[9])
#pragma unroll
          for (int cb = 0; cb < CB; ++cb) pe_lds[wv][t][cb][lane] = PR::gather(pe_rows, t, cb, pl.extra);
      });
      __syncthreads();
    }
    // view layer [feature, extras] -> 128 (ReLU), in place into h[0..7]; stores the feature
    // tiles and the extras
    if (pl.comp_tiles[9] == 2)
      run_layer<P, 18, 8, 0, 18 * kT>(a.packed + pl.fwd_off[9], fin_hc, relu_epi, stage, bias_of(9),
                                      side_of(I16{}, a.w.f, std::integral_constant<int, 2>{}, a.w.ex, -1));
    else if (pl.comp_tiles[9] == 4)
      run_layer<P, 20, 8, 0, 20 * kT>(a.packed + pl.fwd_off[9], fin_hc, relu_epi, stage, bias_of(9),
                                      side_of(I16{}, a.w.f, std::integral_constant<int, 4>{}, a.w.ex, -1));
    else
      run_layer<P, 22, 8, 0, 22 * kT>(a.packed + pl.fwd_off[9], fin_hc, relu_epi, stage, bias_of(9),
                                      side_of(I16{}, a.w.f, std::integral_constant<int, 6>{}, a.w.ex, -1));
    // rgb head 128 -> 3, sigmoid (stores the view layer's output and mask)
    run_layer<P, 8, 1, 0, 9 * kT>(a.packed + pl.fwd_off[10], fin_h, [&](int T, f32x4(&acc)[CB]) SRF_AI {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        const int p = p0 + 16 * cb + j;
        if (g == 0 && valid(cb)) {
          const f32x4 v = acc[cb];
          a.rgb[3 * p + 0] = sigmoidf_(v[0]);
          a.rgb[3 * p + 1] = sigmoidf_(v[1]);
          a.rgb[3 * p + 2] = sigmoidf_(v[2]);
        }
      }
    }, stage, bias_of(10), side_of(std::integral_constant<int, 8>{}, a.w.v, I0{}, 0, 8));
  }
}

// ------------------------------------------------------------------- backward chain
template <int P, bool VD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void mlp_bwd_kernel(
    BwdArgs a) {
  constexpr int CB = Tr<P>::CB;
  using Act = typename Tr<P>::Act;
  __shared__ __attribute__((aligned(16))) uint8_t stage[2 * stage_half_bytes<P>()];
  const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: scalar store offsets
  const int wave = blockIdx.x * kWaves + wv;
  const int p0 = wave * 16 * CB;  // the grid covers cols exactly: every wave has points
  const Plan& pl = a.p;
  const int ldP = a.w.ldP;

  float dsr[CB], drr[CB][3];  // head pre-activation gradients
#pragma unroll
  for (int cb = 0; cb < CB; ++cb) {
    const int p = p0 + 16 * cb + j;
    const bool valid = p < a.n;
    float ds = 0.f, s = 0.f;
    if (valid) {
      ds = a.d_sigma[p];
      s = a.sigma[p];
    }
    dsr[cb] = pl.softplus ? ds * (1.0f - expf(-s)) : (s > 0.f ? ds : 0.f);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float dr = 0.f, y = 0.f;
      if (valid) {
        dr = a.d_rgb[3 * p + c];
        y = a.rgb[3 * p + c];
      }
      drr[cb][c] = dr * (1.0f - y) * y;
    }
  }

  // z holds the current layer's pre-activation gradients; every product writes in place.  Each
  // layer's dZ tiles are stored for the weight gradients during the next product (mma_tiles'
  // deferred stores: dz_side), while z still holds them as that product's input.
  Act z[18][CB];
  uint32_t mb[CB][2];
  auto store_dz = [&](int l, int T, const Act(&t)[CB]) SRF_AI {
    store_tile<P>(uniform_rsrc(a.ws + a.w.dz[l]), tile_block<P>(T, ldP, pl.elem, p0), t);
  };
  // deferred-store items: tile T0 + i of z as tile T0 + i of dZ_l
  auto dz_side = [&](int l, auto T0I) SRF_AI {
    constexpr int t0 = decltype(T0I)::value;
    const Rsrc r = uniform_rsrc(a.ws + a.w.dz[l]);
    return [&, r](auto II) SRF_AI {
      constexpr int t = t0 + decltype(II)::value;
      store_tile<P>(r, tile_block<P>(t, ldP, pl.elem, p0), z[t]);
    };
  };
  using I0 = std::integral_constant<int, 0>;
  // The ReLU masks of one layer (as the forward stored them), for the epilogue of the product in
  // which they are loaded: a deferred item of that product, issued after its first weight chunk
  // has landed and completing behind its MFMAs (a load before the product would hold up the wait
  // for that chunk).
  auto mask_side = [&](int slot) SRF_AI {
    const Rsrc r = uniform_rsrc(a.ws + a.w.mask + (size_t)slot * ldP * 32);
    return [&, r](auto) SRF_AI {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        const u32x2 v = load8(r, (uint32_t)lane * 8u, (uint32_t)(wave * CB + cb) * 512u);
        mb[cb][0] = v.x;
        mb[cb][1] = v.y;
      }
    };
  };
  // item 0: the masks of slot `slot`; items 1..: dz_side(l) items
  auto mask_dz_side = [&](int slot, int l) SRF_AI {
    const auto ms = mask_side(slot);
    const auto ds = dz_side(l, I0{});
    return [&, ms, ds](auto II) SRF_AI {
      constexpr int i = decltype(II)::value;
      if constexpr (i == 0) ms(II);
      else ds(std::integral_constant<int, i - 1>{});
    };
  };
  // epilogue: ReLU'-masked gradient of layer l into z[T] (dZ_l, stored during the next product)
  auto masked_epi = [&](int T, f32x4(&acc)[CB]) SRF_AI {
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
      f32x4 v = acc[cb];
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = (mb[cb][T >> 3] & (1u << (4 * (T & 7) + r))) ? v[r] : 0.f;
      z[T][cb] = to_act<P>(v);
    }
  };
  auto fin_z = [&](int t, int cb) SRF_AI { return z[t][cb]; };

  if constexpr (VD) {
    // rgb head: dZ10 tile 0 (rows 0..2), zero pad tile
    {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        f32x4 v = {0.f, 0.f, 0.f, 0.f};
        if (g == 0) v = f32x4{drr[cb][0], drr[cb][1], drr[cb][2], 0.f};
        z[0][cb] = to_act<P>(v);
        z[1][cb] = zero_act<P>();
      }
    }
    // dV = W10^T dZ10, masked by the view-layer ReLU -> dZ9 (z[0..7]); stores dZ10
    run_layer<P, 2, 8, 0, 2>(a.packed + pl.bwd_off[10], fin_z, masked_epi, stage, nullptr, mask_dz_side(8, 10));
    // dF = W9^T dZ9 (feature part, identity activation) -> dZ8 rows 0..255 (z[0..15]); stores dZ9
    run_layer<P, 8, 16, 0, 8>(a.packed + pl.bwd_off[9], fin_z, [&](int T, f32x4(&acc)[CB]) SRF_AI {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) z[T][cb] = to_act<P>(acc[cb]);
    }, stage, nullptr, dz_side(9, I0{}));
    {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        f32x4 v = {0.f, 0.f, 0.f, 0.f};
        if (g == 0) v[0] = dsr[cb];
        z[16][cb] = to_act<P>(v);
        z[17][cb] = zero_act<P>();
      }
    }
    // dZ8 (tiles 0..16) is stored here rather than during the next product: with 18 input tiles
    // live, deferring it spills registers
#pragma unroll
    for (int T = 0; T < 17; ++T) store_dz(8, T, z[T]);
    run_layer<P, 18, 16, 0, 1>(a.packed + pl.bwd_off[8], fin_z, masked_epi, stage, nullptr, mask_side(7));
  } else {
    {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        f32x4 v = {0.f, 0.f, 0.f, 0.f};
        if (g == 0) v = f32x4{dsr[cb], drr[cb][0], drr[cb][1], drr[cb][2]};
        z[0][cb] = to_act<P>(v);
        z[1][cb] = zero_act<P>();
      }
    }
    run_layer<P, 2, 16, 0, 2>(a.packed + pl.bwd_off[8], fin_z, masked_epi, stage, nullptr, mask_dz_side(7, 8));
  }
  // trunk: z holds dZ_l; dH_{l-1} = W_l^T dZ_l, masked -> dZ_{l-1}, in place (storing dZ_l)
  for (int l = 7; l >= 1; --l) {
    run_layer<P, 16, 16, 0, 17>(a.packed + pl.bwd_off[l], fin_z, masked_epi, stage, nullptr, mask_dz_side(l - 1, l));
  }
  // dZ0: no product follows
#pragma unroll
  for (int T = 0; T < 16; ++T) store_dz(0, T, z[T]);
}

// ----------------------------------------------------------------- weight gradients
struct DwJob {
  const uint8_t* dz;  // first dZ row used
  int dz_rows;        // rows stored from there (multiple of 16)
  const uint8_t* x;   // first input row used
  int x_rows;
  int M, N;           // valid rows (outputs) and columns (inputs)
  float* dw;          // reference gradient, first element of the block
  int ld_dw;
  float* db;          // nullable
  int mblocks, nblocks, work_begin, work_count;
};

// Points per work item of a grouped weight-gradient launch of nj jobs over cols points: whole
// multiples of 256 points, at least 2048.  One block (160 KB of LDS) is resident per CU, so a
// launch costs about (rounds of `cus` blocks) x (per-block time), and a last, partly filled round
// costs as much as a full one.  The per-block time is modelled as the chunk plus a fixed
// overhead worth kDwBlockOverhead points (the tile's FP32 atomics and the ring's fill); the
// chunk minimising rounds x that time over 3..8 rounds is chosen.
constexpr int kDwBlockOverhead = 640;
inline int dw_default_chunk(int cols, int nj, int cus) {
  long long best_cost = -1;
  int best = 2048;
  for (int rounds = 3; rounds <= 8; ++rounds) {
    const int max_chunks = rounds * cus / (nj > 0 ? nj : 1);
    if (max_chunks < 1) continue;
    int chunk = round_up((cols + max_chunks - 1) / max_chunks, 256);
    if (chunk < 2048) chunk = 2048;
    const long long blocks = (long long)nj * ((cols + chunk - 1) / chunk);
    const long long cost = (blocks + cus - 1) / cus * (long long)(chunk + kDwBlockOverhead);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = chunk;
    }
  }
  return best;
}

// Compute units of the current device (the launch-schedule heuristics above; 256 if the query
// fails).
inline int device_compute_units() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    return 256;
  return cus;
}

struct DwArgs {
  DwJob jobs[kMaxJobs];
  int n_jobs, ldP, cols, chunk, n_chunks, total_work;
  // dZ / X layout: 0 plain feature-major rows; 1 the paired tile blocks of store_tile (tile-major
  // for the wide tilings, point-block-major for the narrow one; see mlp_dw_kernel)
  int paired;
  // optional device-side point count: only the columns below it, rounded up to whole 256-point
  // blocks, are reduced (the chunks past it exit); the columns up to that bound must hold finite
  // values with dZ = 0 beyond the count, as the forward and backward kernels leave them
  const int32_t* cols_dev;
};

// Columns of a weight-gradient launch: a.cols, or fewer when a device-side count is given.
__device__ __forceinline__ int dw_cols(const DwArgs& a) {
  return a.cols_dev ? min(a.cols, (*a.cols_dev + 255) / 256 * 256) : a.cols;
}

// One block of 8 waves reduces a 256 x 256 tile of dW = dZ X^T (a whole layer of the MLP) over a
// chunk of points, so every dZ and X row of the chunk is read from HBM once; the reduction over
// points streams at the HBM rate (a 256 x 256 layer does 128 FLOP per byte of dZ and X, below
// the MFMA:HBM balance of the chip).  Per k-unit (32 points for 16-bit operands, 16 for FP32)
// the 16 dZ-row and 16 X-row fragments of the tile (1 KB each, gathered from the feature-major
// arrays by per-lane addresses) are copied global -> LDS by the eight waves into a ring of four
// k-unit slots, three k-units ahead of the one being multiplied (96 KB in flight per CU to cover
// the HBM latency with a single resident block); wave w runs the MFMAs of its 64 x 128 sub-tile
// (row tiles 4 (w >> 1) .. +3, column tiles 8 (w & 1) .. +7) from LDS.  Partial sums are added
// to dW with FP32 atomics; the bias gradient is the row sum of dZ.
constexpr int kDwTile = 256;
constexpr int kDwWaves = 8;
constexpr int kDwFrags = 32;              // 16 dZ + 16 X fragments per k-unit
constexpr int kDwSlot = kDwFrags * 1024;  // one k-unit
constexpr int kDwSlots = 5;               // ring depth (prefetch distance kDwSlots - 1): 160 KB of LDS
constexpr int kDwNarrowRows = 128, kDwNarrowCols = 160;  // tile of the RM = 1, NC = 10 wave tiling


// Wave tiling: RM row tiles x NC column tiles per wave.  NC = 8: the waves split the tile into
// four row groups x two column halves, RM = 4 giving 256 x 256 tiles (the NeRF layers) and RM = 2
// 128 x 256 tiles.  RM = 1, NC = 10: wave w takes row tile w and column tiles 0..9 of a 128 x 160
// tile, which keeps all eight waves (and all four SIMDs) busy on jobs of at most 128 rows and 160
// columns (the TensoRF colour network), where the half-tile split left half of them idle.  The
// DMA is the same for all (16 dZ and 16 X row tiles from m0 / n0).
template <int P, int RM = 4, int NC = 8>
__global__ __launch_bounds__(64 * kDwWaves) void mlp_dw_kernel(DwArgs a) {
  static_assert((NC == 8 && (RM == 4 || RM == 2)) || (RM == 1 && NC == 10), "supported wave tilings");
  constexpr bool kRowWaves = NC != 8;
  constexpr int kTM = kRowWaves ? 16 * kDwWaves : 64 * RM;  // tile rows
  constexpr int kTN = kRowWaves ? 16 * NC : kDwTile;         // tile columns
  __shared__ __attribute__((aligned(16))) uint8_t ring[kDwSlots * kDwSlot];
  const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int item = blockIdx.x;
  int ji = 0;
#pragma unroll 1
  while (ji + 1 < a.n_jobs && item >= a.jobs[ji + 1].work_begin) ++ji;
  const DwJob& jb = a.jobs[ji];
  const int local = item - jb.work_begin;
  const int nblk = jb.mblocks * jb.nblocks;
  const int tile = local % nblk, chunk = local / nblk;
  const int m0 = kTM * (tile % jb.mblocks), n0 = kTN * (tile / jb.mblocks);
  const int pb = chunk * a.chunk;
  const int pe = min(pb + a.chunk, dw_cols(a));
  if (pb >= pe) return;  // the chunk lies past the device-side count (uniform over the block)
  const size_t ldP = a.ldP;
  constexpr bool F32L = f32_layout<P>();
  constexpr int ELEM = F32L ? 4 : 2;
  constexpr int KPTS = F32L ? 16 : 32;  // points per k-unit (split BF16: two units per MFMA k-step)
  const int qm = kRowWaves ? w : RM * (w >> 1), qn = kRowWaves ? 0 : 8 * (w & 1);
  bool mv[RM], nv[NC];
#pragma unroll
  for (int i = 0; i < RM; ++i) mv[i] = m0 + 16 * (qm + i) < jb.dz_rows;
#pragma unroll
  for (int k = 0; k < NC; ++k) nv[k] = n0 + 16 * (qn + k) < jb.x_rows;
  bool any_m = false, any_n = false;
#pragma unroll
  for (int i = 0; i < RM; ++i) any_m |= mv[i];
#pragma unroll
  for (int k = 0; k < NC; ++k) any_n |= nv[k];
  // valid row and column tiles of this wave (prefixes of the tile lists)
  int nmv = 0, nnv = 0;
#pragma unroll
  for (int i = 0; i < RM; ++i) nmv += mv[i] ? 1 : 0;
#pragma unroll
  for (int k = 0; k < NC; ++k) nnv += nv[k] ? 1 : 0;
  nmv = __builtin_amdgcn_readfirstlane(nmv);
  nnv = __builtin_amdgcn_readfirstlane(nnv);
  // DMA of the 32 fragments of the k-unit at point p into a ring slot: fragment f < 16 is dZ row
  // tile f, f >= 16 is X row tile f - 16; wave w copies f = w, w + 8, w + 16, w + 24.  Every wave
  // issues exactly four loads per k-unit (a tile beyond the job's rows re-reads its first row
  // tile into the unused slot), so the vmcnt waits below count whole k-units.
  // Narrow tiling: only the fragments the tile uses are copied, nA dZ row tiles and nB X row
  // tiles (nA <= 8, nB <= 10), fragment f of the list going to wave f % 8; each wave waits for
  // its own count (dma_n per k-unit).
  const int nA = kRowWaves ? min(8, (jb.dz_rows - m0 + 15) / 16) : 16;
  const int nB = kRowWaves ? min(NC, (jb.x_rows - n0 + 15) / 16) : 16;
  const int dma_n = kRowWaves ? (nA + nB - w + kDwWaves - 1) / kDwWaves : 4;
  auto dma_unit = [&](int p, uint8_t* buf) SRF_AI {
#pragma unroll
    for (int q = 0; q < (kRowWaves ? 3 : 4); ++q) {
      int f = w + kDwWaves * q;
      if (kRowWaves) {
        if (f >= nA + nB) break;
        f = f < nA ? f : 16 + (f - nA);  // slot: dZ tile f or X tile 16 + k, as the MFMA loop reads
      }
      const bool is_a = f < 16;
      const int rt = is_a ? f : f - 16;
      int row0 = (is_a ? m0 : n0) + 16 * rt;
      if (row0 >= (is_a ? jb.dz_rows : jb.x_rows)) row0 = is_a ? m0 : n0;
      const uint8_t* base = is_a ? jb.dz : jb.x;
      const uint8_t* src;
      if (a.paired) {
        // lane (g, j) takes feature j's chunk of the record pair m = 4 u + g of the 2 KB block of
        // the points' wave (see store_tile): record (j >> 2, 2 m + ((j >> 1) & 1)), chunk j & 1
        constexpr int PPW = F32L ? 32 : 64;  // points per stored wave block
        const int wb = p / PPW, u = (p % PPW) / KPTS;
        const int rec = 16 * (j >> 2) + 2 * (4 * u + g) + ((j >> 1) & 1);
        // the wide tilings read tile-major blocks (tile t's blocks span 16 rows of ldP points, the MLP
        // workspace); the narrow one point-block-major blocks (the array's tiles of one wave block
        // adjacent, 2 KB each: the colour network's workspace)
        size_t blk;
        if constexpr (kRowWaves) blk = ((size_t)wb * ((is_a ? jb.dz_rows : jb.x_rows) / 16) + row0 / 16) * 2048;
        else blk = (size_t)row0 * ldP * ELEM + (size_t)wb * 2048;
        src = base + blk + (j & 1) * 1024 + rec * 16;
      } else {
        const int pg = F32L ? 4 * g : 8 * g;
        src = base + (((size_t)(row0 + j)) * ldP + p + pg) * ELEM;
      }
      __builtin_amdgcn_global_load_lds((const void*)src,
                                       (__attribute__((address_space(3))) void*)(buf + f * 1024), 16, 0, 0);
    }
  };
  f32x4 acc[RM][NC];
  float rs[RM];
#pragma unroll
  for (int i = 0; i < RM; ++i) {
    rs[i] = 0.f;
#pragma unroll
    for (int k = 0; k < NC; ++k) acc[i][k] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  const bool do_db = jb.db != nullptr && n0 == 0 && qn == 0;
  const bool busy = any_m && (any_n || do_db);
  const int n_units = (pe - pb) / KPTS;
  // prologue: k-units 0 .. kDwSlots - 2 in flight
#pragma unroll
  for (int u = 0; u < kDwSlots - 1; ++u)
    if (u < n_units) dma_unit(pb + u * KPTS, ring + u * kDwSlot);
#pragma unroll 1
  for (int u = 0; u < n_units; ++u) {
    // k-unit u has landed once at most (issued after it) k-units' loads are outstanding
    const int after = min(kDwSlots - 2, n_units - 1 - u);
    static_assert(kDwSlots <= 5, "waits below cover up to three k-units in flight after u");
    if constexpr (kRowWaves) {
      wait_vmem_at_most(after * dma_n);  // at most 3 x 3
    } else {
      if (after >= 3) __builtin_amdgcn_s_waitcnt(vmcnt_only(12));
      else if (after == 2) __builtin_amdgcn_s_waitcnt(vmcnt_only(8));
      else if (after == 1) __builtin_amdgcn_s_waitcnt(vmcnt_only(4));
      else __builtin_amdgcn_s_waitcnt(vmcnt_only(0));
    }
    // After the barrier every wave's part of k-unit u is in LDS (each wave waited for its own
    // copies above) and slot (u - 1) % kDwSlots is free (each wave consumed its LDS reads of it
    // in the MFMAs of iteration u - 1).  A plain s_barrier with compiler-only memory barriers:
    // __syncthreads()'s workgroup fences would drain the prefetches still in flight
    // (vmcnt(0)) and defeat the ring.
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (u + kDwSlots - 1 < n_units)
      dma_unit(pb + (u + kDwSlots - 1) * KPTS, ring + ((u + kDwSlots - 1) % kDwSlots) * kDwSlot);
    if (busy) {
      const uint8_t* fr = ring + (u % kDwSlots) * kDwSlot;
      // No branch between the MFMAs (a branch after every FP32 MFMA doubled the products' time).
      // Narrow tiling: the valid row and column tiles are prefixes (nmv, nnv, uniform over the
      // wave), so the products run as one unguarded block per count of valid columns, chosen once
      // per k-unit.  Wide tilings: every tile's product runs (a tile past the job's rows multiplies
      // the re-read first row tile; the epilogue stores only the valid tiles), since a second,
      // unguarded copy of their products spills.
      if constexpr (P == kF32) {
        f32x4 av[RM], bv[NC];
#pragma unroll
        for (int i = 0; i < RM; ++i) av[i] = *reinterpret_cast<const f32x4*>(fr + (qm + i) * 1024 + lane * 16);
#pragma unroll
        for (int k = 0; k < NC; ++k) bv[k] = *reinterpret_cast<const f32x4*>(fr + (16 + qn + k) * 1024 + lane * 16);
        auto prod = [&](auto NRI, auto NVI) SRF_AI {
          constexpr int NR = decltype(NRI)::value, NV = decltype(NVI)::value;
#pragma unroll
          for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int i = 0; i < NR; ++i)
#pragma unroll
              for (int k = 0; k < NV; ++k) acc[i][k] = mfma32(av[i][q], bv[k][q], acc[i][k]);
        };
        if constexpr (kRowWaves) {
          if (nmv == RM)
            static_for<1, NC + 1>([&](auto VI) SRF_AI {
              if (nnv == decltype(VI)::value) prod(std::integral_constant<int, RM>{}, VI);
            });
        } else {
          prod(std::integral_constant<int, RM>{}, std::integral_constant<int, NC>{});  // as the 16-bit tiling
        }
        if (do_db)
#pragma unroll
          for (int i = 0; i < RM; ++i) rs[i] += (av[i][0] + av[i][1]) + (av[i][2] + av[i][3]);
      } else {
        u32x4 av[RM], bv[NC];
#pragma unroll
        for (int i = 0; i < RM; ++i) av[i] = *reinterpret_cast<const u32x4*>(fr + (qm + i) * 1024 + lane * 16);
#pragma unroll
        for (int k = 0; k < NC; ++k) bv[k] = *reinterpret_cast<const u32x4*>(fr + (16 + qn + k) * 1024 + lane * 16);
        // every tile's product, valid or not (the epilogue stores the valid ones; the slot holds a
        // whole k-unit either way): no branch between the MFMAs
#pragma unroll
        for (int i = 0; i < RM; ++i)
#pragma unroll
          for (int k = 0; k < NC; ++k) acc[i][k] = mfma16<P>(av[i], bv[k], acc[i][k]);
        if (do_db)
#pragma unroll
          for (int i = 0; i < RM; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) rs[i] += unpack_lo<P>(av[i][q]) + unpack_hi<P>(av[i][q]);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < RM; ++i) {
    if (!mv[i]) continue;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      if (!nv[k]) continue;
      const int c = n0 + 16 * (qn + k) + j;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int o = m0 + 16 * (qm + i) + 4 * g + r;
        if (o < jb.M && c < jb.N) unsafeAtomicAdd(jb.dw + (size_t)o * jb.ld_dw + c, acc[i][k][r]);
      }
    }
    if (do_db) {
      float sr = rs[i];
      sr += __shfl_xor(sr, 16, 64);
      sr += __shfl_xor(sr, 32, 64);
      const int o = m0 + 16 * (qm + i) + j;
      if (g == 0 && o < jb.M) unsafeAtomicAdd(jb.db + o, sr);
    }
  }
}

// Split-BF16 weight gradients: the FP32 dZ and X tiles (the FP32 layouts, DMA'd through the same
// five-slot ring as the FP32 GEMM, 16 points per k-unit) are split into two BF16 planes and reduced
// with 32x32x16 BF16 MFMAs, dW += dZ_hi X_hi^T + dZ_hi X_lo^T + dZ_lo X_hi^T.  One 16-point unit is
// exactly one K = 16 step: lane (h, i) of an A operand holds row i of a 32-row block at the eight
// points of lane groups 2h and 2h + 1 of the FP32 fragments (records 16 (2h) + (i mod 16) and
// 16 (2h + 1) + (i mod 16)), and the B operand takes the same points of its columns, so the two
// agree slot for slot.  Wave w owns rows 64 (w >> 1) .. + 63 (two 32-row blocks) and columns
// 128 (w & 1) .. + 127 (four 32-column blocks) of a 256 x 256 tile.  32x32 D layout: lane l, register
// r holds row 8 (r / 4) + 4 (l / 32) + (r mod 4), column l mod 32.
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <int P>
__global__ __launch_bounds__(64 * kDwWaves) void mlp_dw_split_kernel(DwArgs a) {
  static_assert(split_planes<P>() == 2, "two BF16 planes");
  __shared__ __attribute__((aligned(16))) uint8_t ring[kDwSlots * kDwSlot];
  const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
  const int h = lane >> 5, i32 = lane & 31;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int item = blockIdx.x;
  int ji = 0;
#pragma unroll 1
  while (ji + 1 < a.n_jobs && item >= a.jobs[ji + 1].work_begin) ++ji;
  const DwJob& jb = a.jobs[ji];
  const int local = item - jb.work_begin;
  const int nblk = jb.mblocks * jb.nblocks;
  const int tile = local % nblk, chunk = local / nblk;
  const int m0 = kDwTile * (tile % jb.mblocks), n0 = kDwTile * (tile / jb.mblocks);
  const int pb = chunk * a.chunk;
  const int pe = min(pb + a.chunk, dw_cols(a));
  if (pb >= pe) return;  // the chunk lies past the device-side count (uniform over the block)
  const size_t ldP = a.ldP;
  constexpr int KPTS = 16;
  const int qm = 4 * (w >> 1), qn = 8 * (w & 1);  // first 16-row / 16-column tile of the wave
  bool mv[2], nv[4];  // 32-row / 32-column blocks with data (their first 16-row tile inside the job)
#pragma unroll
  for (int b = 0; b < 2; ++b) mv[b] = m0 + 16 * (qm + 2 * b) < jb.dz_rows;
#pragma unroll
  for (int b = 0; b < 4; ++b) nv[b] = n0 + 16 * (qn + 2 * b) < jb.x_rows;
  // DMA of one 16-point unit: the 16 dZ and 16 X fragments of the tile, as in mlp_dw_kernel (a row
  // tile beyond the job's rows re-reads its first tile, so every wave issues four loads per unit)
  auto dma_unit = [&](int p, uint8_t* buf) SRF_AI {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int f = w + kDwWaves * q;
      const bool is_a = f < 16;
      const int rt = is_a ? f : f - 16;
      int row0 = (is_a ? m0 : n0) + 16 * rt;
      if (row0 >= (is_a ? jb.dz_rows : jb.x_rows)) row0 = is_a ? m0 : n0;
      const uint8_t* base = is_a ? jb.dz : jb.x;
      const int wb = p / 32, u = (p % 32) / KPTS;
      const int rec = 16 * (j >> 2) + 2 * (4 * u + g) + ((j >> 1) & 1);
      const uint8_t* src = base + (size_t)row0 * ldP * 4 + (size_t)wb * 2048 + (j & 1) * 1024 + rec * 16;
      __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)(buf + f * 1024),
                                       16, 0, 0);
    }
  };
  f32x16 acc[2][4];
  float rs[2] = {0.f, 0.f};
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[rb][cb][r] = 0.f;
  const bool do_db = jb.db != nullptr && n0 == 0 && qn == 0;
  bool any_m = mv[0] || mv[1], any_n = false;
#pragma unroll
  for (int b = 0; b < 4; ++b) any_n |= nv[b];
  const bool busy = any_m && (any_n || do_db);
  const int n_units = (pe - pb) / KPTS;
  // two BF16 planes of the eight FP32 values lane (h, i) takes from fragment f: records 2h, 2h + 1
  // of its lane group pair
  auto planes = [&](const uint8_t* fr, int f, u32x4& hi, u32x4& lo, float* sum) SRF_AI {
    const int jj = i32 & 15;
    const f32x4 v0 = *reinterpret_cast<const f32x4*>(fr + f * 1024 + ((2 * h) * 16 + jj) * 16);
    const f32x4 v1 = *reinterpret_cast<const f32x4*>(fr + f * 1024 + ((2 * h + 1) * 16 + jj) * 16);
    uint32_t t[4][2];
    split_pair<2>(v0[0], v0[1], t[0]);
    split_pair<2>(v0[2], v0[3], t[1]);
    split_pair<2>(v1[0], v1[1], t[2]);
    split_pair<2>(v1[2], v1[3], t[3]);
    hi = u32x4{t[0][0], t[1][0], t[2][0], t[3][0]};
    lo = u32x4{t[0][1], t[1][1], t[2][1], t[3][1]};
    if (sum) *sum += ((v0[0] + v0[1]) + (v0[2] + v0[3])) + ((v1[0] + v1[1]) + (v1[2] + v1[3]));
  };
#pragma unroll
  for (int u = 0; u < kDwSlots - 1; ++u)
    if (u < n_units) dma_unit(pb + u * KPTS, ring + u * kDwSlot);
#pragma unroll 1
  for (int u = 0; u < n_units; ++u) {
    const int after = min(kDwSlots - 2, n_units - 1 - u);
    if (after >= 3) __builtin_amdgcn_s_waitcnt(vmcnt_only(12));
    else if (after == 2) __builtin_amdgcn_s_waitcnt(vmcnt_only(8));
    else if (after == 1) __builtin_amdgcn_s_waitcnt(vmcnt_only(4));
    else __builtin_amdgcn_s_waitcnt(vmcnt_only(0));
    // every wave's part of unit u is in LDS and slot (u - 1) % kDwSlots is free (see mlp_dw_kernel)
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (u + kDwSlots - 1 < n_units)
      dma_unit(pb + (u + kDwSlots - 1) * KPTS, ring + ((u + kDwSlots - 1) % kDwSlots) * kDwSlot);
    if (busy) {
      const uint8_t* fr = ring + (u % kDwSlots) * kDwSlot;
      u32x4 ah[2], al[2], bh[4], bl[4];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
        planes(fr, qm + 2 * rb + (i32 >> 4), ah[rb], al[rb], do_db ? &rs[rb] : nullptr);
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) planes(fr, 16 + qn + 2 * cb + (i32 >> 4), bh[cb], bl[cb], nullptr);
      // every block's products, valid or not (as in mlp_dw_kernel: the slot holds a whole unit, the
      // epilogue stores the valid blocks), so that no branch separates the MFMAs
#pragma unroll
      for (int pr = 0; pr < 3; ++pr)
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
          for (int cb = 0; cb < 4; ++cb)
            acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                __builtin_bit_cast(bf16x8, pr == 0 ? al[rb] : ah[rb]),
                __builtin_bit_cast(bf16x8, pr == 1 ? bl[cb] : bh[cb]), acc[rb][cb], 0, 0, 0);
    }
  }
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) {
    if (!mv[rb]) continue;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      if (!nv[cb]) continue;
      const int c = n0 + 16 * qn + 32 * cb + i32;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int o = m0 + 16 * qm + 32 * rb + 8 * (r >> 2) + 4 * h + (r & 3);
        if (o < jb.M && c < jb.N) unsafeAtomicAdd(jb.dw + (size_t)o * jb.ld_dw + c, acc[rb][cb][r]);
      }
    }
    if (do_db) {
      const float sr = rs[rb] + __shfl_xor(rs[rb], 32, 64);
      const int o = m0 + 16 * qm + 32 * rb + i32;
      if (h == 0 && o < jb.M) unsafeAtomicAdd(jb.db + o, sr);
    }
  }
}

// --------------------------------------------------------------------- host glue
template <int P, bool VD, bool TRAIN>
int launch_fwd(const FwdArgs& a, hipStream_t s) {
  const int ppw = 16 * Tr<P>::CB;
  const int waves = a.w.cols / ppw;
  const int blocks = (waves + kWaves - 1) / kWaves;
  hipLaunchKernelGGL((mlp_fwd_kernel<P, VD, TRAIN>), dim3(blocks), dim3(64 * kWaves), 0, s, a);
  return launch_status();
}

template <int P, bool VD>
int launch_bwd(const BwdArgs& a, hipStream_t s) {
  const int ppw = 16 * Tr<P>::CB;
  const int waves = a.w.cols / ppw;
  const int blocks = (waves + kWaves - 1) / kWaves;
  hipLaunchKernelGGL((mlp_bwd_kernel<P, VD>), dim3(blocks), dim3(64 * kWaves), 0, s, a);
  return launch_status();
}

template <int P, int RM = 4, int NC = 8>
int launch_dw(const DwArgs& d, hipStream_t s) {
  hipLaunchKernelGGL((mlp_dw_kernel<P, RM, NC>), dim3(d.total_work), dim3(64 * kDwWaves), 0, s, d);
  return launch_status();
}

// Per-precision launchers, one translation unit each (mlp_f32.hip, mlp_bf16.hip, mlp_bf16s.hip;
// mlp_f16.hip holds the inference forward only).
int fwd_launch(int prec, const FwdArgs& a, bool view_dep, bool train, hipStream_t s);
int bwd_launch(int prec, const BwdArgs& a, bool view_dep, hipStream_t s);
int dw_launch(int prec, const DwArgs& d, hipStream_t s);
int fwd_f32(const FwdArgs& a, bool vd, bool tr, hipStream_t s);
int fwd_bf16(const FwdArgs& a, bool vd, bool tr, hipStream_t s);
int fwd_f16(const FwdArgs& a, bool vd, bool tr, hipStream_t s);
int bwd_f32(const BwdArgs& a, bool vd, hipStream_t s);
int bwd_bf16(const BwdArgs& a, bool vd, hipStream_t s);
int dw_f32(const DwArgs& d, hipStream_t s);
int dw_bf16(const DwArgs& d, hipStream_t s);
int dw_f32_narrow(const DwArgs& d, hipStream_t s);  // 128 x 160 tiles (kDwNarrowRows x kDwNarrowCols)
int fwd_bf16s(const FwdArgs& a, bool vd, bool tr, hipStream_t s);  // split BF16 (mlp_bf16s.hip)
int bwd_bf16s(const BwdArgs& a, bool vd, hipStream_t s);
int dw_bf16s(const DwArgs& d, hipStream_t s);  // mlp_dw_split_kernel

}  // namespace mlp
}  // namespace srf
