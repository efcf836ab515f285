// TensoRF colour network on gfx950 MFMA: basis matrix, positional encoding of the features and
// view directions, 128 -> 128 -> 3 MLP with sigmoid, its hand-written backward chain and the
// weight gradients.
//
// Reference semantics: VmDecomposedTensor.get_color after the gather, basis_matrix_color
// (src/models/SimpleTensoRF09.py:1151, :1262) and MlpFeaturesColorPredictor
// (:1370-1421): x = [PE_Lf(F), PE_Lv(d)], h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2),
// rgb = sigmoid(W3 h2 + b3), F = B app (no bias), PE blocks [v, sin(2^0 v), cos(2^0 v), ...].
//
// Layout and mapping follow the NeRF MLP kernels (mfma.h): the network runs transposed,
// H_out^T = W H_in^T with the weights as the MFMA A operand (pre-packed into per-lane fragments
// by srf_color_pack) and 16-point column blocks as N; a wave owns 32 points (2 column blocks), so
// one layer's D tiles are the next layer's B operand without data movement.  The A fragments are
// streamed global -> LDS by the four waves of a block (double-buffered, as in mlp::mma_tiles) and
// shared by them.  The basis output F and the view directions go through a per-wave LDS tile from
// which every lane builds its positional-encoding inputs.  Training stores the layer inputs and
// output gradients as FP32 tile blocks (store_tile_rows) for the weight-gradient GEMMs: the MLP's
// FP32 grouped GEMM (mlp_dw_kernel, 128 x 160 tiles) or color_dw_kernel.
//
// Operand precision (desc->prec).  Activations, accumulators and the stored layer inputs are FP32
// in every mode; the products of the forward and of the backward chain multiply
//   SRF_PREC_F32          FP32 operands (16x16x4 MFMA, one k-unit = 16 features)
//   SRF_PREC_BF16_SPLIT   split BF16 (16x16x32 MFMA, one k-unit = 32 features): three planes in
//                         the forward (six plane products, FP32-equivalent operands), two in the
//                         backward chain (three products), as the NeRF MLP's split mode
//   SRF_PREC_BF16         BF16 operands rounded from the FP32 activations and weights (one plane)
// Reduction depths are padded to whole k-units (32 features for the BF16 modes) with zeros.  The
// weight gradients reduce the stored FP32 tiles with the FP32 GEMM (SRF_PREC_F32) or the split-BF16
// one, two planes per operand (both BF16 modes).
#include <string.h>

#include "mlp_impl.h"

namespace srf {
namespace color {
namespace {


constexpr int kHid = 128;     // num_units_color_predictor of the shipped configurations
constexpr int kHT = kHid / 16;
constexpr int kCB = 2;        // 16-point column blocks per wave
constexpr int kPW = 16 * kCB; // points per wave
constexpr int kWaves = 4;
constexpr int kMaxC = 256;    // appearance components
constexpr int kMaxB = 32;     // basis outputs (features_dimension_color)
constexpr int kMaxD1 = 240;   // first-layer inputs (the backward's per-wave LDS tile fits beside the stage)
constexpr int kST = 8;        // output tiles per weight-stage buffer (the widest layer has kHT = 8)

// Forward / backward-chain operand precisions of a descriptor's prec.
constexpr int fwd_prec_of(int prec) { return prec == SRF_PREC_BF16_SPLIT ? kBFS3 : (prec == SRF_PREC_BF16 ? kBFS1 : kF32); }
constexpr int bwd_prec_of(int prec) { return prec == SRF_PREC_BF16_SPLIT ? kBFS2 : (prec == SRF_PREC_BF16 ? kBFS1 : kF32); }
// Double-buffered weight stage of a product in precision P (bytes): kST output tiles x the
// k-units of one chunk x the BF16 planes, 1 KB each, twice.  F32 and BF16 32 KB, split forward
// 48 KB, split backward 64 KB.
template <int P>
constexpr int stage_bytes() { return 2 * kST * mlp::kc_of<P>() * split_planes<P>() * 1024; }
// Features per k-unit: 16 (FP32), 32 (BF16 operands).
template <int P>
constexpr int kunit() { return P == kF32 ? 16 : 32; }

struct Geo {
  int prec;      // SRF_PREC_* of the descriptor
  int kal;       // reduction padding: 16 (FP32) or 32 (BF16 modes)
  int fpl, bpl;  // BF16 planes of the forward / backward fragments (1 for FP32)
  int C, Cp;     // components, padded to kal
  int NB;        // basis outputs (27)
  int Lf, Lv, view;
  int D1, D1p;   // first-layer inputs, padded to kal
  // fragment offsets (floats) in the packed buffer
  size_t fB, fW1, fW2, fW3, fW3T, fW2T, fW1T, fBT, b1, b2, b3, total;
};

// Floats of the packed fragments of `tiles` output tiles over a reduction of K features: one KB
// per k-unit and BF16 plane (FP32: per 16 features).
inline size_t frag_floats(const Geo& g, int tiles, int K, int planes) {
  return (size_t)tiles * (K / g.kal) * planes * 256;
}


bool make_geo(const srf_color_desc* d, Geo* g) {
  if (!d || d->n_features < 4 || d->n_features % 4 != 0 || d->n_features > kMaxC || d->basis_dim < 1 ||
      d->basis_dim > kMaxB || d->features_pe < 0 || d->views_pe < 0 || d->hidden != kHid)
    return false;
  if (d->prec != SRF_PREC_F32 && d->prec != SRF_PREC_BF16 && d->prec != SRF_PREC_BF16_SPLIT) return false;
  memset(g, 0, sizeof(*g));
  g->prec = d->prec;
  g->kal = d->prec == SRF_PREC_F32 ? 16 : 32;
  g->fpl = d->prec == SRF_PREC_BF16_SPLIT ? 3 : 1;
  g->bpl = d->prec == SRF_PREC_BF16_SPLIT ? 2 : 1;
  g->C = d->n_features;
  g->Cp = (g->C + g->kal - 1) / g->kal * g->kal;
  g->NB = d->basis_dim;
  g->Lf = d->features_pe;
  g->Lv = d->views_pe;
  g->view = d->use_view_dirs != 0;
  g->D1 = g->NB * (2 * g->Lf + 1) + (g->view ? 3 * (2 * g->Lv + 1) : 0);
  g->D1p = (g->D1 + g->kal - 1) / g->kal * g->kal;
  if (g->D1p > kMaxD1) return false;
  size_t o = 0;
  auto take = [&](size_t n) { size_t r = o; o += n; return r; };
  g->fB = take(frag_floats(*g, 2, g->Cp, g->fpl));
  g->fW1 = take(frag_floats(*g, kHT, g->D1p, g->fpl));
  g->fW2 = take(frag_floats(*g, kHT, kHid, g->fpl));
  g->fW3 = take(frag_floats(*g, 1, kHid, g->fpl));
  g->fW3T = take(frag_floats(*g, kHT, g->kal, g->bpl));     // rows 128, k: the 3 (16) dz3 rows
  g->fW2T = take(frag_floats(*g, kHT, kHid, g->bpl));
  g->fW1T = take(frag_floats(*g, g->D1p / 16, kHid, g->bpl));
  g->fBT = take(frag_floats(*g, g->Cp / 16, kMaxB, g->bpl));  // rows Cp, k 32
  g->b1 = take(kHid);
  g->b2 = take(kHid);
  g->b3 = take(16);
  g->total = o;
  return true;
}

// Training workspace: appT, F, x1, h1, h2, dz3, dz2, dz1, dF, each an array of 16-feature tiles of
// FP32 values in the point-block-major tile-block layout (store_tile_rows), an array of R rows taking
// the bytes of R rows of `ld` points.  The stores address every array with 32-bit byte offsets, so each stays below 4 GB
// (ws_addressable).
struct Ws {
  int cols, ld;
  size_t appT, F, x1, h1, h2, dz3, dz2, dz1, dF, mask, bytes;
};

inline bool ws_addressable(const Geo& g, const Ws& w) {
  int rows = g.Cp > g.D1p ? g.Cp : g.D1p;
  rows = rows > kHid ? rows : kHid;
  return (uint64_t)rows * (uint64_t)w.ld * 4u < (1ull << 32);
}

Ws make_ws(const Geo& g, int n) {
  Ws w;
  w.cols = ((n > 0 ? n : 1) + 255) / 256 * 256;
  w.ld = (w.cols / 256) % 2 == 0 ? w.cols + 256 : w.cols;  // rows an odd number of blocks apart
  const size_t row = (size_t)w.ld * 4;
  size_t o = 0;
  auto take = [&](int rows) { size_t r = o; o += (size_t)rows * row; return r; };
  w.appT = take(g.Cp);
  w.F = take(kMaxB);
  w.x1 = take(g.D1p);
  w.h1 = take(kHid);
  w.h2 = take(kHid);
  w.dz3 = take(16);
  w.dz2 = take(kHid);
  w.dz1 = take(kHid);
  w.dF = take(kMaxB);
  w.mask = take(8);  // ReLU bits of h1 and h2: 16 B per lane, 1 KB per wave of 32 points
  w.bytes = o;
  return w;
}

// ------------------------------------------------------------------------ packing
// FP32 fragments (planes 0): fragment f of the A operand of a transposed layer: tile T (16 output
// rows), k-step s (lane group g supplies input feature 16 (s / 4) + 4 g + s % 4), lane l = 16 g + j
// holds W'[16 T + j][k] where W' = W (rows x cols, leading dimension ld) or its transpose.
// BF16 fragments (planes 1-3): per tile T and 32-feature k-unit u, `planes` consecutive 1 KB
// planes; in each, lane l = 16 g + j holds 8 BF16 values of row 16 T + j (dword d: slots 2 d and
// 2 d + 1), slot i being feature 32 u + 4 g + i (i < 4) or 32 u + 16 + 4 g + i - 4 (i >= 4), the
// 16-bit slot map of the MFMA operands; plane p + 1 is the BF16 rounding of what planes 0..p leave.
struct PackJob {
  const float* src;
  int rows, cols, ld, trans, tiles, ksteps;  // ksteps: FP32 k-steps, or k-units for BF16 planes
  int planes;
  size_t off;
};

struct PackArgs {
  PackJob jobs[8];
  int n_jobs;
  const float* b1;
  const float* b2;
  const float* b3;
  size_t ob1, ob2, ob3;
  float* out;
};

__device__ __forceinline__ float pack_src(const PackJob& jb, int r, int k) {
  if (!jb.trans) return (r < jb.rows && k < jb.cols) ? jb.src[(size_t)r * jb.ld + k] : 0.f;
  return (r < jb.cols && k < jb.rows) ? jb.src[(size_t)k * jb.ld + r] : 0.f;
}

__global__ __launch_bounds__(256) void pack_kernel(PackArgs a) {
  const int jid = blockIdx.y;
  if (jid < a.n_jobs) {
    const PackJob& jb = a.jobs[jid];
    if (jb.planes == 0) {
      const int total = jb.tiles * jb.ksteps * 64;
      for (int e = blockIdx.x * 256 + threadIdx.x; e < total; e += gridDim.x * 256) {
        const int l = e & 63, s = (e >> 6) % jb.ksteps, T = (e >> 6) / jb.ksteps;
        const int g = l >> 4, j = l & 15;
        a.out[jb.off + e] = pack_src(jb, 16 * T + j, 16 * (s >> 2) + 4 * g + (s & 3));
      }
    } else {
      // one dword (two BF16 slots) of one lane of one (tile, k-unit) per thread, all its planes
      uint32_t* out = reinterpret_cast<uint32_t*>(a.out + jb.off);
      const int total = jb.tiles * jb.ksteps * 256;
      for (int e = blockIdx.x * 256 + threadIdx.x; e < total; e += gridDim.x * 256) {
        const int d = e & 3, l = (e >> 2) & 63, u = (e >> 8) % jb.ksteps, T = (e >> 8) / jb.ksteps;
        const int g = l >> 4, j = l & 15;
        const int i0 = 2 * d;  // slots i0, i0 + 1 lie on the same side of slot 4
        const int k0 = 32 * u + (i0 < 4 ? 4 * g + i0 : 16 + 4 * g + i0 - 4);
        float v0 = pack_src(jb, 16 * T + j, k0), v1 = pack_src(jb, 16 * T + j, k0 + 1);
        const size_t frag = (size_t)(T * jb.ksteps + u) * jb.planes;
        for (int pl = 0; pl < jb.planes; ++pl) {
          const uint32_t h = pack2<kBF16>(v0, v1);
          out[(frag + pl) * 256 + l * 4 + d] = h;
          v0 = v0 - __uint_as_float(h << 16);
          v1 = v1 - __uint_as_float(h & 0xffff0000u);
        }
      }
    }
  } else if (blockIdx.x == 0) {
    for (int i = threadIdx.x; i < kHid; i += 256) {
      a.out[a.ob1 + i] = a.b1[i];
      a.out[a.ob2 + i] = a.b2[i];
    }
    if (threadIdx.x < 16) a.out[a.ob3 + threadIdx.x] = threadIdx.x < 3 ? a.b3[threadIdx.x] : 0.f;
  }
}

// ------------------------------------------------------------------------ kernels
// acc[T][cb] += sum over the nt input tiles t of A(tile T, t) x fin(t, cb) for T < OUT:
// mlp::mma_tiles for a reduction depth known only at run time (nt 16-feature tiles, a whole number
// of k-units: one tile for FP32 operands, two for BF16).  The A fragments (1 KB per tile, k-unit
// and BF16 plane, upr k-units per packed tile row) are staged through LDS in chunks of
// mlp::kc_of<P>() k-units, double-buffered; fin may load or compute its operand (it is called once
// per input tile and column block, so it must not index register arrays); keep(t, b) then sees
// tile t's operands of all column blocks (the training copies of the layer input are stored
// there).  Block-collective: every wave calls it with the same arguments.
template <int P, int OUT, typename FIn, typename Keep>
__device__ __forceinline__ void staged_layer_rt(f32x4 (&acc)[OUT][kCB], const float* __restrict__ A, int upr,
                                                int nt, FIn&& fin, Keep&& keep, uint8_t* stage) {
  constexpr int KC = mlp::kc_of<P>(), NPL = split_planes<P>();
  constexpr int kHalf = kST * KC * NPL * 1024;
  static_assert(OUT <= kST, "the stage buffer holds kST output tiles");
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const uint8_t* Ab = reinterpret_cast<const uint8_t*>(A);
  const int nu = P == kF32 ? nt : nt / 2;
  const int nch = (nu + KC - 1) / KC;
  auto dma = [&](int c, uint8_t* buf) {
    const int kn = nu - c * KC < KC ? nu - c * KC : KC;
    for (int f = w; f < OUT * kn * NPL; f += kWaves) {
      const int fk = f / NPL, pl = f - fk * NPL;
      const int to = fk / kn, kc = fk - (fk / kn) * kn;
      const int frag = __builtin_amdgcn_readfirstlane((to * upr + c * KC + kc) * NPL + pl);
      const int dst = __builtin_amdgcn_readfirstlane(((to * KC + kc) * NPL + pl) * 1024);
      lds_dma_1k(Ab + (size_t)frag * 1024, buf + dst);
    }
  };
  dma(0, stage);
  wait_vmem();
  __syncthreads();
  for (int c = 0; c < nch; ++c) {
    const uint8_t* cur = stage + (c & 1) * kHalf;
    if (c + 1 < nch) dma(c + 1, stage + ((c + 1) & 1) * kHalf);
    const int kn = nu - c * KC < KC ? nu - c * KC : KC;
    for (int kc = 0; kc < kn; ++kc) {
      const int u = c * KC + kc;
      if constexpr (P == kF32) {
        f32x4 b[kCB];
#pragma unroll
        for (int cb = 0; cb < kCB; ++cb) b[cb] = fin(u, cb);
        keep(u, b);
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int T = 0; T < OUT; ++T) {
            const float a = *reinterpret_cast<const float*>(cur + (T * KC + kc) * 1024 + (r * 64 + lane) * 4);
#pragma unroll
            for (int cb = 0; cb < kCB; ++cb) acc[T][cb] = mfma32(a, b[cb][r], acc[T][cb]);
          }
      } else {
        // input tiles 2u, 2u + 1 -> NPL BF16 B operands per column block (mlp::mma_tiles' split)
        f32x4 x0[kCB], x1[kCB];
#pragma unroll
        for (int cb = 0; cb < kCB; ++cb) {
          x0[cb] = fin(2 * u, cb);
          x1[cb] = fin(2 * u + 1, cb);
        }
        keep(2 * u, x0);
        keep(2 * u + 1, x1);
        u32x4 b[NPL][kCB];
#pragma unroll
        for (int cb = 0; cb < kCB; ++cb) {
          uint32_t q0[NPL], q1[NPL], q2[NPL], q3[NPL];
          split_pair<NPL>(x0[cb][0], x0[cb][1], q0);
          split_pair<NPL>(x0[cb][2], x0[cb][3], q1);
          split_pair<NPL>(x1[cb][0], x1[cb][1], q2);
          split_pair<NPL>(x1[cb][2], x1[cb][3], q3);
#pragma unroll
          for (int i = 0; i < NPL; ++i) b[i][cb] = u32x4{q0[i], q1[i], q2[i], q3[i]};
        }
        // output tiles in pairs (four independent accumulators between dependent MFMAs), the
        // plane products smallest first
#pragma unroll
        for (int T0 = 0; T0 < OUT; T0 += 2) {
          u32x4 av[2][NPL];
#pragma unroll
          for (int q = 0; q < 2; ++q)
            if (T0 + q < OUT)
#pragma unroll
              for (int pl = 0; pl < NPL; ++pl)
                av[q][pl] = *reinterpret_cast<const u32x4*>(cur + (((T0 + q) * KC + kc) * NPL + pl) * 1024 + lane * 16);
#pragma unroll
          for (int sum = NPL - 1; sum >= 0; --sum)
#pragma unroll
            for (int i = 0; i <= sum; ++i)
#pragma unroll
              for (int q = 0; q < 2; ++q)
                if (T0 + q < OUT)
#pragma unroll
                  for (int cb = 0; cb < kCB; ++cb)
                    acc[T0 + q][cb] = mfma16<kBF16>(av[q][i], b[sum - i][cb], acc[T0 + q][cb]);
        }
      }
    }
    wait_vmem();
    __syncthreads();
  }
}

// acc += A(tiles of the packed block at A) x in with a compile-time depth of NT 16-feature input
// tiles (a whole number of k-units), through mlp::mma_tiles on the kST-tile stage.
template <int P, int OUT, int NT, typename FIn>
__device__ __forceinline__ void staged_layer(f32x4 (&acc)[OUT][kCB], const float* __restrict__ A, int upr,
                                             FIn&& fin, uint8_t* stage) {
  static_assert(P == kF32 || NT % 2 == 0, "BF16 k-units hold two input tiles");
  mlp::mma_tiles<P, NT, OUT, 0, kST>(acc, reinterpret_cast<const uint8_t*>(A), upr, 0, fin, stage);
}

// Positional-encoding input i of the first layer for one point, from its basis outputs F (LDS
// column) and view direction (reference order: [F, sin(F), cos(F), sin(2F), ...], then the same
// blocks of the direction).
__device__ __forceinline__ float pe_input(const Geo& g, int i, const float* Fcol, int fstride, const float* dir) {
  const int nf = g.NB * (2 * g.Lf + 1);
  if (i < nf) {
    const int b = i / g.NB, f = i - b * g.NB;
    const float v = Fcol[f * fstride];
    if (b == 0) return v;
    const float x = __fmul_rn(v, (float)(1 << ((b - 1) >> 1)));
    return (b & 1) ? sinf(x) : cosf(x);
  }
  if (g.view && i < g.D1) {
    const int i2 = i - nf;
    const int b = i2 / 3, d = i2 - 3 * b;
    const float v = dir[d];
    if (b == 0) return v;
    const float x = __fmul_rn(v, (float)(1 << ((b - 1) >> 1)));
    return (b & 1) ? sinf(x) : cosf(x);
  }
  return 0.f;
}

struct FwdArgs {
  Geo g;
  Ws w;
  const float* pk;
  int n;                 // capacity (rows of app, idx and the workspace)
  const int32_t* count;  // device-side row count, or null (n rows)
  const float* app;  // [n, C]
  const int32_t* idx;
  int n_samples;
  const float* view_dirs;
  float* rgb;
  int rgb_at_idx;  // rgb row idx[p] of the dense sample buffer instead of row p
  uint8_t* ws;  // null: inference
};

// Rows a launch processes: the capacity, or the device-side count below it.
__device__ __forceinline__ int active_rows(int capacity, const int32_t* count) {
  return count ? min(capacity, *count) : capacity;
}

// Whether block b has work: blocks up to the count rounded to whole 256-point workspace blocks run
// (their points past the count as padding: zero features, finite activations, dZ = 0), so every
// column the weight-gradient GEMM reads (it rounds the count the same way) was written this call.
__device__ __forceinline__ bool block_has_rows(int b, int n) {
  return (int64_t)b * kWaves * kPW < (int64_t)(n + 255) / 256 * 256;
}

// Training rows are kept as 2 KB tile blocks in the record layout of the NeRF MLP's FP32 workspace
// (mlp::store_tile: one 16-feature tile of 32 points), so that the weight-gradient GEMM copies every
// operand fragment as one contiguous 1 KB; the blocks of an array are point-block-major (the n_tiles
// tiles of points 32 q .. 32 q + 31 adjacent, at byte 2048 (q n_tiles + t)), so that one unit of the
// GEMM reads one short run per array rather than one run per 16-row tile spread over the whole
// capacity.  A wave stores tile t of its 32 points (p0) straight from D-tile (or B-operand)
// registers v[cb] = features 16 t + 4 g + r at point p0 + 16 cb + j; all lanes call it (pair
// exchange).
__host__ __device__ inline uint32_t blk_off(int n_tiles, int t, int p0) {
  return ((uint32_t)(p0 >> 5) * (uint32_t)n_tiles + (uint32_t)t) * 2048u;
}

__device__ __forceinline__ void store_tile_rows(uint8_t* ws, size_t arr, int n_tiles, int t, int p0,
                                                const f32x4 (&v)[kCB]) {
  mlp::store_tile<kF32>(mlp::uniform_rsrc(ws + arr), blk_off(n_tiles, t, p0), v);
}

__device__ __forceinline__ void load_tile_rows(const uint8_t* ws, size_t arr, int n_tiles, int t, int p0,
                                               f32x4 (&v)[kCB]) {
  mlp::load_tile_f32(mlp::uniform_rsrc(ws + arr), blk_off(n_tiles, t, p0), v);
}

template <int P>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(2, 2))) void color_fwd_kernel(FwdArgs a) {
  __shared__ __attribute__((aligned(16))) uint8_t stage[stage_bytes<P>()];
  constexpr int KU = kunit<P>();
  __shared__ float sF[kWaves][kMaxB][kPW + 1];
  __shared__ float sD[kWaves][3][kPW];
  const Geo& g = a.g;
  const int n = active_rows(a.n, a.count);
  if (!block_has_rows(blockIdx.x, n)) return;  // uniform over the block
  const int lane = threadIdx.x & 63, gq = lane >> 4, j = lane & 15;
  const int wv = threadIdx.x >> 6;
  const int p0 = (blockIdx.x * kWaves + wv) * kPW;
  const bool train = a.ws != nullptr;
  // 1) basis: F^T = B app^T (2 tiles of 16 outputs)
  f32x4 accF[2][kCB];
#pragma unroll
  for (int T = 0; T < 2; ++T)
#pragma unroll
    for (int cb = 0; cb < kCB; ++cb) accF[T][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
  auto fin_app = [&](int u, int cb) -> f32x4 {
    const int p = p0 + 16 * cb + j;
    const int c = 16 * u + 4 * gq;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (p < n && c < g.C) v = *reinterpret_cast<const float4*>(a.app + (size_t)p * g.C + c);
    return f32x4{v.x, v.y, v.z, v.w};
  };
  // training copy of the appearance features (tile u; zero past the components)
  auto keep_app = [&](int u, const f32x4(&b)[kCB]) {
    if (train) store_tile_rows(a.ws, a.w.appT, g.Cp / 16, u, p0, b);
  };
  staged_layer_rt<P, 2>(accF, a.pk + g.fB, g.Cp / KU, g.Cp / 16, fin_app, keep_app, stage);
  // F and the view directions to LDS (and F feature-major for the backward)
#pragma unroll
  for (int T = 0; T < 2; ++T)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int f = 16 * T + 4 * gq + r;
#pragma unroll
      for (int cb = 0; cb < kCB; ++cb) sF[wv][f][16 * cb + j] = accF[T][cb][r];
    }
  if (train)
#pragma unroll
    for (int T = 0; T < 2; ++T) store_tile_rows(a.ws, a.w.F, kMaxB / 16, T, p0, accF[T]);
  if (g.view && lane < kPW) {
    const int p = p0 + lane;
    float d0 = 0.f, d1 = 0.f, d2 = 0.f;
    if (p < n) {
      const int64_t smp = a.idx ? (int64_t)a.idx[p] : (int64_t)p;
      const int64_t ray = smp / a.n_samples;
      d0 = a.view_dirs[3 * ray + 0];
      d1 = a.view_dirs[3 * ray + 1];
      d2 = a.view_dirs[3 * ray + 2];
    }
    sD[wv][0][lane] = d0;
    sD[wv][1][lane] = d1;
    sD[wv][2][lane] = d2;
  }
  __syncthreads();
  // 2) first layer over the positional-encoding inputs built per k-step from LDS
  f32x4 h1[kHT][kCB];
#pragma unroll
  for (int T = 0; T < kHT; ++T)
#pragma unroll
    for (int cb = 0; cb < kCB; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) h1[T][cb][r] = a.pk[g.b1 + 16 * T + 4 * gq + r];
  auto fin_pe = [&](int u, int cb) -> f32x4 {
    const int col = 16 * cb + j;
    const float dir[3] = {sD[wv][0][col], sD[wv][1][col], sD[wv][2][col]};
    f32x4 b;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * u + 4 * gq + r;
      b[r] = pe_input(g, i, &sF[wv][0][col], kPW + 1, dir);
    }
    return b;
  };
  auto keep_pe = [&](int u, const f32x4(&b)[kCB]) {
    if (train) store_tile_rows(a.ws, a.w.x1, g.D1p / 16, u, p0, b);
  };
  staged_layer_rt<P, kHT>(h1, a.pk + g.fW1, g.D1p / KU, g.D1p / 16, fin_pe, keep_pe, stage);
  // ReLU decisions of the two hidden layers, bit 4 T + r of word cb (h1) and 2 + cb (h2), for the
  // backward chain (it then reads 32 B per point instead of h1 and h2)
  uint32_t relu_bits[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int T = 0; T < kHT; ++T) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int cb = 0; cb < kCB; ++cb) {
        const bool pos = h1[T][cb][r] > 0.f;
        h1[T][cb][r] = pos ? h1[T][cb][r] : 0.f;
        relu_bits[cb] |= pos ? 1u << (4 * T + r) : 0u;
      }
    if (train) store_tile_rows(a.ws, a.w.h1, kHT, T, p0, h1[T]);
  }
  // 3) second layer
  f32x4 h2[kHT][kCB];
#pragma unroll
  for (int T = 0; T < kHT; ++T)
#pragma unroll
    for (int cb = 0; cb < kCB; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) h2[T][cb][r] = a.pk[g.b2 + 16 * T + 4 * gq + r];
  staged_layer<P, kHT, kHT>(h2, a.pk + g.fW2, kHid / KU, [&](int u, int cb) { return h1[u][cb]; }, stage);
#pragma unroll
  for (int T = 0; T < kHT; ++T) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int cb = 0; cb < kCB; ++cb) {
        const bool pos = h2[T][cb][r] > 0.f;
        h2[T][cb][r] = pos ? h2[T][cb][r] : 0.f;
        relu_bits[2 + cb] |= pos ? 1u << (4 * T + r) : 0u;
      }
    if (train) store_tile_rows(a.ws, a.w.h2, kHT, T, p0, h2[T]);
  }
  if (train)
    mlp::store_nt16(mlp::uniform_rsrc(a.ws + a.w.mask), (uint32_t)lane * 16u, (uint32_t)(p0 >> 5) * 1024u,
                    u32x4{relu_bits[0], relu_bits[1], relu_bits[2], relu_bits[3]});
  // 4) output layer + sigmoid
  f32x4 o[1][kCB];
#pragma unroll
  for (int cb = 0; cb < kCB; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) o[0][cb][r] = a.pk[g.b3 + 4 * gq + r];
  staged_layer<P, 1, kHT>(o, a.pk + g.fW3, kHid / KU, [&](int u, int cb) { return h2[u][cb]; }, stage);
  if (gq == 0) {
#pragma unroll
    for (int cb = 0; cb < kCB; ++cb) {
      const int p = p0 + 16 * cb + j;
      if (p < n) {
        const size_t dst = a.rgb_at_idx ? (size_t)a.idx[p] : (size_t)p;
#pragma unroll
        for (int r = 0; r < 3; ++r) a.rgb[3 * dst + r] = 1.f / (1.f + expf(-o[0][cb][r]));
      }
    }
  }
}

struct BwdArgs {
  Geo g;
  Ws w;
  const float* pk;
  int n;                 // capacity
  const int32_t* count;  // device-side row count, or null
  const int32_t* idx;    // rows of rgb / d_rgb when rgb_at_idx
  int rgb_at_idx;
  const float* rgb;
  const float* d_rgb;
  float* d_app;
  uint8_t* ws;
};

// One element (feature f, point p) of a tile-block array of n_tiles tiles (the record of
// mlp::store_tile holding it: chunk f & 1 of record 16 (fl / 4) + 2 (jj / 2) + (fl / 2) % 2, at
// position 2 cb + jj % 2, fl = f mod 16, point p = 32 q + 16 cb + jj).
__device__ __forceinline__ float load_elem(const uint8_t* base, size_t arr, int n_tiles, int f, int p) {
  const int fl = f & 15, q = p & 31, jj = q & 15;
  const int rec = 16 * (fl >> 2) + 2 * (jj >> 1) + ((fl >> 1) & 1);
  const uint32_t off = blk_off(n_tiles, f >> 4, p) + 1024u * (fl & 1) + 16u * rec + 4u * (2 * (q >> 4) + (jj & 1));
  return *reinterpret_cast<const float*>(base + arr + off);
}

template <int P>
__device__ __forceinline__ void color_bwd_body(const BwdArgs& a) {
  __shared__ __attribute__((aligned(16))) uint8_t stage[stage_bytes<P>()];
  // per wave: the basis outputs F of its 32 points and the gradient dF being accumulated
  __shared__ float sF[kWaves][kMaxB][kPW + 1];
  __shared__ float sdF[kWaves][kMaxB][kPW + 1];
  constexpr int KU = kunit<P>();
  constexpr int NPL = split_planes<P>();
  const Geo& g = a.g;
  const int n = active_rows(a.n, a.count);
  if (!block_has_rows(blockIdx.x, n)) return;  // uniform over the block
  const int lane = threadIdx.x & 63, gq = lane >> 4, j = lane & 15;
  const int wv = threadIdx.x >> 6;
  const int p0 = (blockIdx.x * kWaves + wv) * kPW;
  // 1) dz3 = d_rgb * s (1 - s)
  f32x4 dz3[1][kCB];
#pragma unroll
  for (int cb = 0; cb < kCB; ++cb) {
    const int p = p0 + 16 * cb + j;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float v = 0.f;
      if (gq == 0 && r < 3 && p < n) {
        const size_t src = a.rgb_at_idx ? (size_t)a.idx[p] : (size_t)p;
        const float s = a.rgb[3 * src + r];
        v = __fmul_rn(__fmul_rn(a.d_rgb[3 * src + r], __fsub_rn(1.f, s)), s);
      }
      dz3[0][cb][r] = v;
    }
  }
  store_tile_rows(a.ws, a.w.dz3, 1, 0, p0, dz3[0]);
  // 2) dz2 = (W3^T dz3) * [h2 > 0]
  f32x4 dz2[kHT][kCB];
#pragma unroll
  for (int T = 0; T < kHT; ++T)
#pragma unroll
    for (int cb = 0; cb < kCB; ++cb) dz2[T][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
  // k: the 16 dz3 rows (one FP32 k-unit; BF16: one 32-row k-unit, rows 16-31 zero)
  staged_layer<P, kHT, KU / 16>(dz2, a.pk + g.fW3T, 1, [&](int u, int cb) {
    return u == 0 ? dz3[0][cb] : f32x4{0.f, 0.f, 0.f, 0.f};
  }, stage);
  // the forward's ReLU bits (words 0-1: h1, 2-3: h2, per column block)
  const u32x4 relu_bits = __builtin_amdgcn_raw_buffer_load_b128(mlp::uniform_rsrc(a.ws + a.w.mask),
                                                                 (uint32_t)lane * 16u, (uint32_t)(p0 >> 5) * 1024u, 0);
#pragma unroll
  for (int T = 0; T < kHT; ++T) {
#pragma unroll
    for (int cb = 0; cb < kCB; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) dz2[T][cb][r] = (relu_bits[2 + cb] >> (4 * T + r)) & 1u ? dz2[T][cb][r] : 0.f;
    store_tile_rows(a.ws, a.w.dz2, kHT, T, p0, dz2[T]);
  }
  // 3) dz1 = (W2^T dz2) * [h1 > 0]
  f32x4 dz1[kHT][kCB];
#pragma unroll
  for (int T = 0; T < kHT; ++T)
#pragma unroll
    for (int cb = 0; cb < kCB; ++cb) dz1[T][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
  staged_layer<P, kHT, kHT>(dz1, a.pk + g.fW2T, kHid / KU, [&](int u, int cb) { return dz2[u][cb]; }, stage);
#pragma unroll
  for (int T = 0; T < kHT; ++T) {
#pragma unroll
    for (int cb = 0; cb < kCB; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) dz1[T][cb][r] = (relu_bits[cb] >> (4 * T + r)) & 1u ? dz1[T][cb][r] : 0.f;
    store_tile_rows(a.ws, a.w.dz1, kHT, T, p0, dz1[T]);
  }
  // 4) dx1 = W1^T dz1, kST output tiles per staged pass and single tiles for the rest, folded into the
  // positional-encoding backward as its tiles leave the accumulators:
  //   dF[f] = dx[f] + sum_k 2^k (dx[sin_k f] cos(2^k F[f]) - dx[cos_k f] sin(2^k F[f]))
  // accumulated per (f, point) in the wave's LDS rows in increasing input row i = b NB + f (the rows of
  // one f lie in different tiles, 27 apart, and the tiles come in order), so the sums run in the
  // reference's order; the view-direction rows (i >= NB (2 Lf + 1)) have no gradient to keep.
  {
    f32x4 fv[kCB];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      load_tile_rows(a.ws, a.w.F, kMaxB / 16, t, p0, fv);
#pragma unroll
      for (int cb = 0; cb < kCB; ++cb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          sF[wv][16 * t + 4 * gq + r][16 * cb + j] = fv[cb][r];
          sdF[wv][16 * t + 4 * gq + r][16 * cb + j] = 0.f;  // rows past basis_dim stay 0
        }
    }
  }
  const int nf = g.NB * (2 * g.Lf + 1);
  auto fin_dz1 = [&](int u, int cb) { return dz1[u][cb]; };
  auto dx_tiles = [&](auto OUTC, int T0) {
    constexpr int OUT = decltype(OUTC)::value;
    f32x4 dx[OUT][kCB];
#pragma unroll
    for (int T = 0; T < OUT; ++T)
#pragma unroll
      for (int cb = 0; cb < kCB; ++cb) dx[T][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
    staged_layer<P, OUT, kHT>(dx, a.pk + g.fW1T + (size_t)T0 * (kHid / KU) * NPL * 256, kHid / KU, fin_dz1, stage);
#pragma unroll
    for (int T = 0; T < OUT; ++T)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * (T0 + T) + 4 * gq + r;
        if (i < nf) {
          const int b = i / g.NB, f = i - b * g.NB;
#pragma unroll
          for (int cb = 0; cb < kCB; ++cb) {
            const int col = 16 * cb + j;
            const float v = dx[T][cb][r];
            float& d = sdF[wv][f][col];
            if (b == 0) {
              d = v;
            } else {
              const float fr = (float)(1 << ((b - 1) >> 1));
              float sn, cs;
              sincosf(__fmul_rn(sF[wv][f][col], fr), &sn, &cs);
              if (b & 1) d = __fadd_rn(d, __fmul_rn(__fmul_rn(v, cs), fr));
              else d = __fsub_rn(d, __fmul_rn(__fmul_rn(v, sn), fr));
            }
          }
        }
      }
  };
  int T0 = 0;
  for (; T0 + kST <= g.D1p / 16; T0 += kST) dx_tiles(std::integral_constant<int, kST>{}, T0);
  for (; T0 < g.D1p / 16; ++T0) dx_tiles(std::integral_constant<int, 1>{}, T0);
  mlp::wave_lds_sync();  // the wave's dF rows, written by its lanes, are read across lanes below
  // 5) d_app = B^T dF (point-major, float4 per lane)
  f32x4 dFt[2][kCB];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int cb = 0; cb < kCB; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) dFt[t][cb][r] = sdF[wv][16 * t + 4 * gq + r][16 * cb + j];
#pragma unroll
  for (int t = 0; t < 2; ++t) store_tile_rows(a.ws, a.w.dF, kMaxB / 16, t, p0, dFt[t]);  // rows past basis_dim: 0
  auto fin_dF = [&](int u, int cb) { return dFt[u][cb]; };
  auto da_tiles = [&](auto OUTC, int T0) {
    constexpr int OUT = decltype(OUTC)::value;
    f32x4 da[OUT][kCB];
#pragma unroll
    for (int T = 0; T < OUT; ++T)
#pragma unroll
      for (int cb = 0; cb < kCB; ++cb) da[T][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
    staged_layer<P, OUT, 2>(da, a.pk + g.fBT + (size_t)T0 * (kMaxB / KU) * NPL * 256, kMaxB / KU, fin_dF, stage);
#pragma unroll
    for (int T = 0; T < OUT; ++T) {
      const int c = 16 * (T0 + T) + 4 * gq;
      if (c < g.C)
#pragma unroll
        for (int cb = 0; cb < kCB; ++cb) {
          const int p = p0 + 16 * cb + j;
          if (p < n)
            *reinterpret_cast<float4*>(a.d_app + (size_t)p * g.C + c) =
                make_float4(da[T][cb][0], da[T][cb][1], da[T][cb][2], da[T][cb][3]);
        }
    }
  };
  int Ta = 0;
  for (; Ta + kST <= g.Cp / 16; Ta += kST) da_tiles(std::integral_constant<int, kST>{}, Ta);
  for (; Ta < g.Cp / 16; ++Ta) da_tiles(std::integral_constant<int, 1>{}, Ta);
}

// The FP32 and BF16 chains fit two waves per SIMD (two blocks per CU: 66 KB of LDS each); the
// split-BF16 chain, with its 64 KB backward stage and the plane operands, runs one.
template <int P>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(2, 2))) void color_bwd_kernel(BwdArgs a) {
  color_bwd_body<P>(a);
}

__global__ __launch_bounds__(64 * kWaves) void color_bwd_split_kernel(BwdArgs a) { color_bwd_body<kBFS2>(a); }

// ------------------------------------------------------------------- weight gradients
// dW = dZ X^T (+ db = row sums of dZ) of the four colour jobs (W1, W2, W3, basis) over the stored
// FP32 tile blocks, in the BF16 modes (FP32: the MLP's FP32 GEMM, mlp_dw_kernel).  A work item is one
// job over one chunk of points; its block copies, per 16-point unit, only the job's own fragments
// (nA dZ and nB X row tiles, one contiguous 1 KB each) global -> LDS into a ring whose slot holds two
// units, 2 (nA + nB) KB, so a job of 9-24 fragments keeps 4-14 units in flight in the 160 KB of LDS
// with one barrier per two units (10 % faster than one per unit on the 300^2 step).  The
// products are split-BF16 32x32x16 MFMAs (dZ_hi X_hi + dZ_hi X_lo + dZ_lo X_hi, as
// mlp_dw_split_kernel); wave w multiplies row block w & 3 (32 rows) by column blocks 4 (w >> 2) .. + 3
// (32 columns each).  Partial sums go to dW with FP32 atomics.
constexpr int kCdwWaves = 8;
constexpr int kCdwRingKB = 160;
constexpr int kCdwMaxSlots = 8;
constexpr int kCdwMaxA = 8, kCdwMaxB = 16;  // row tiles: dZ <= 128 rows, X <= 256 rows

struct CdwJob {
  const uint8_t* dz;
  const uint8_t* x;
  int nA, nB;  // 16-row tiles of dZ and X (the tiles of their arrays)
  int M, N;    // valid rows and columns of dW
  float* dw;
  int ld_dw;
  float* db;
};

struct CdwArgs {
  CdwJob jobs[4];
  int n_jobs, cols, chunk;
  const int32_t* cols_dev;
};

typedef float f32x16 __attribute__((ext_vector_type(16)));

// s_waitcnt vmcnt(n) for a wave-uniform n in 0..47.
__device__ __forceinline__ void wait_vmem_le(int n) {
  switch (n) {
#define SRF_WAIT_CASE(k) case k: __builtin_amdgcn_s_waitcnt(mlp::vmcnt_only(k)); break;
    SRF_WAIT_CASE(0) SRF_WAIT_CASE(1) SRF_WAIT_CASE(2) SRF_WAIT_CASE(3) SRF_WAIT_CASE(4) SRF_WAIT_CASE(5)
    SRF_WAIT_CASE(6) SRF_WAIT_CASE(7) SRF_WAIT_CASE(8) SRF_WAIT_CASE(9) SRF_WAIT_CASE(10) SRF_WAIT_CASE(11)
    SRF_WAIT_CASE(12) SRF_WAIT_CASE(13) SRF_WAIT_CASE(14) SRF_WAIT_CASE(15) SRF_WAIT_CASE(16) SRF_WAIT_CASE(17)
    SRF_WAIT_CASE(18) SRF_WAIT_CASE(19) SRF_WAIT_CASE(20) SRF_WAIT_CASE(21) SRF_WAIT_CASE(22) SRF_WAIT_CASE(23)
    SRF_WAIT_CASE(24) SRF_WAIT_CASE(25) SRF_WAIT_CASE(26) SRF_WAIT_CASE(27) SRF_WAIT_CASE(28) SRF_WAIT_CASE(29)
    SRF_WAIT_CASE(30) SRF_WAIT_CASE(31) SRF_WAIT_CASE(32) SRF_WAIT_CASE(33) SRF_WAIT_CASE(34) SRF_WAIT_CASE(35)
    SRF_WAIT_CASE(36) SRF_WAIT_CASE(37) SRF_WAIT_CASE(38) SRF_WAIT_CASE(39) SRF_WAIT_CASE(40) SRF_WAIT_CASE(41)
    SRF_WAIT_CASE(42) SRF_WAIT_CASE(43) SRF_WAIT_CASE(44) SRF_WAIT_CASE(45) SRF_WAIT_CASE(46)
#undef SRF_WAIT_CASE
    default: __builtin_amdgcn_s_waitcnt(mlp::vmcnt_only(47)); break;
  }
}

__global__ __launch_bounds__(64 * kCdwWaves) void color_dw_kernel(CdwArgs a) {
  __shared__ __attribute__((aligned(16))) uint8_t ring[kCdwRingKB * 1024];
  const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
  const int h = lane >> 5, i32 = lane & 31;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // items are chunk-major, so that the items with points (the first chunks) of all jobs come first
  const int job = blockIdx.x % a.n_jobs, chunk = blockIdx.x / a.n_jobs;
  const CdwJob& jb = a.jobs[job];
  const int cols = a.cols_dev ? min(a.cols, (*a.cols_dev + 255) / 256 * 256) : a.cols;
  const int pb = chunk * a.chunk;
  const int pe = min(pb + a.chunk, cols);
  if (pb >= pe) return;  // past the device-side count (uniform over the block)
  const int nA = jb.nA, nB = jb.nB, nf = nA + nB;
  const int nslots = min(kCdwMaxSlots, kCdwRingKB / (2 * nf));  // two-unit slots
  const int dma_n = (nf - w + kCdwWaves - 1) / kCdwWaves;  // this wave's fragments per unit (<= 3)
  // unit at point p: fragment f < nA is dZ row tile f, else X row tile f - nA; lane (g, j) copies
  // the 16 B of feature j at the four points of lane group g (the record of mlp::store_tile), to
  // slot byte f KB + 16 lane
  // the job's array bases, read from the kernel arguments once (wave-uniform, scalar registers)
  const uint8_t* const dz_base = jb.dz;
  const uint8_t* const x_base = jb.x;
  auto dma_unit = [&](int p, uint8_t* buf) {
    const int u = (p & 31) >> 4;
    const int rec = 16 * (j >> 2) + 2 * (4 * u + g) + ((j >> 1) & 1);
    const uint32_t in_blk = (j & 1) * 1024 + rec * 16;
    for (int q = 0; q < 3; ++q) {
      const int f = w + kCdwWaves * q;
      if (f >= nf) break;
      const uint8_t* src = f < nA ? dz_base + blk_off(nA, f, p) : x_base + blk_off(nB, f - nA, p);
      __builtin_amdgcn_global_load_lds((const void*)(src + in_blk),
                                       (__attribute__((address_space(3))) void*)(buf + f * 1024), 16, 0, 0);
    }
  };
  // row block rb (32 rows), column blocks cq .. cq + 3
  const int rb = w & 3, cq = 4 * (w >> 2);
  const bool row_ok = 2 * rb < nA;
  bool col_ok[4];
  bool any_col = false;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    col_ok[c] = 2 * (cq + c) < nB;
    any_col |= col_ok[c];
  }
  const bool do_db = jb.db != nullptr && cq == 0;
  const bool busy = row_ok && (any_col || do_db);
  // the fragments this lane's half (i32 >> 4) reads: a 32-row block's second tile past the job's
  // rows repeats the first (rows past M are not stored)
  const int fa = min(2 * rb + (i32 >> 4), nA - 1);
  int fb[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) fb[c] = nA + min(2 * (cq + c) + (i32 >> 4), nB - 1);
  f32x16 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[c][r] = 0.f;
  float rs = 0.f;
  // the two BF16 planes of the eight FP32 values lane (h, i) takes from fragment f: records of lane
  // groups 2h and 2h + 1 of its row i mod 16
  auto planes = [&](const uint8_t* fr, int f, u32x4& hi, u32x4& lo, bool sum) {
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
    if (sum) rs += ((v0[0] + v0[1]) + (v0[2] + v0[3])) + ((v1[0] + v1[1]) + (v1[2] + v1[3]));
  };
  // The ring advances by steps of two units (one 32-point tile block, both halves): one barrier and one
  // wait per step.
  const int n_steps = (pe - pb) / 32;
  const int slot = 2 * nf * 1024;
  auto dma_step = [&](int p, uint8_t* buf) {
    dma_unit(p, buf);
    dma_unit(p + 16, buf + nf * 1024);
  };
  // the valid column blocks are a prefix (ncol of them, uniform over the wave): the products run as
  // straight-line code for that count, one uniform choice per unit instead of a branch per MFMA
  int ncol = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) ncol += col_ok[c] ? 1 : 0;
  ncol = __builtin_amdgcn_readfirstlane(ncol);
  auto unit_products = [&](const uint8_t* fr) {
    u32x4 ah, al, bh[4], bl[4];
    planes(fr, fa, ah, al, do_db);
    mlp::static_for<1, 5>([&](auto NI) {
      constexpr int NCOL = decltype(NI)::value;
      if (ncol != NCOL) return;
#pragma unroll
      for (int c = 0; c < NCOL; ++c) planes(fr, fb[c], bh[c], bl[c], false);
#pragma unroll
      for (int pr = 0; pr < 3; ++pr)
#pragma unroll
        for (int c = 0; c < NCOL; ++c)
          acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, pr == 0 ? al : ah),
                                                           __builtin_bit_cast(bf16x8, pr == 1 ? bl[c] : bh[c]),
                                                           acc[c], 0, 0, 0);
    });
  };
  for (int st = 0; st < nslots - 1 && st < n_steps; ++st) dma_step(pb + 32 * st, ring + st * slot);
  for (int st = 0; st < n_steps; ++st) {
    // step st has landed once at most (steps issued after it) x 2 dma_n loads are outstanding
    wait_vmem_le(min(nslots - 2, n_steps - 1 - st) * 2 * dma_n);
    // every wave's part of step st is in LDS, and slot (st - 1) mod nslots is free (the waves consumed
    // their reads of it in the MFMAs of iteration st - 1); compiler-only fences around a plain
    // s_barrier, which would otherwise drain the prefetches (vmcnt(0))
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (st + nslots - 1 < n_steps)
      dma_step(pb + 32 * (st + nslots - 1), ring + ((st + nslots - 1) % nslots) * slot);
    if (!busy) continue;
    const uint8_t* fr = ring + (st % nslots) * slot;
    unit_products(fr);
    unit_products(fr + nf * 1024);
  }
  if (!row_ok) return;
  // 32x32 D layout: lane l, register r holds row 8 (r / 4) + 4 (l / 32) + (r mod 4), column l mod 32
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if (!col_ok[c]) continue;
    const int col = 32 * (cq + c) + i32;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int o = 32 * rb + 8 * (r >> 2) + 4 * h + (r & 3);
      if (o < jb.M && col < jb.N) unsafeAtomicAdd(jb.dw + (size_t)o * jb.ld_dw + col, acc[c][r]);
    }
  }
  if (do_db) {
    const float sr = rs + __shfl_xor(rs, 32, 64);
    const int o = 32 * rb + i32;
    if (h == 0 && o < jb.M) unsafeAtomicAdd(jb.db + o, sr);
  }
}

}  // namespace
}  // namespace color
}  // namespace srf

using namespace srf;
using namespace srf::color;

extern "C" size_t srf_color_packed_bytes(const srf_color_desc* desc) {
  Geo g;
  return make_geo(desc, &g) ? g.total * sizeof(float) : 0;
}

extern "C" size_t srf_color_workspace_bytes(const srf_color_desc* desc, int n_points) {
  Geo g;
  if (!make_geo(desc, &g) || n_points < 0) return 0;
  return make_ws(g, n_points).bytes;
}

extern "C" int srf_color_pack(const srf_color_desc* desc, const float* basis, const float* w1, const float* b1,
                              const float* w2, const float* b2, const float* w3, const float* b3, void* packed,
                              void* stream) {
  Geo g;
  if (!make_geo(desc, &g)) return SRF_EINVAL;
  if (!basis || !w1 || !b1 || !w2 || !b2 || !w3 || !b3 || !packed) return SRF_EINVAL;
  PackArgs a;
  memset(&a, 0, sizeof(a));
  // ksteps: FP32 k-steps of 4 features, or the 32-feature k-units of a BF16 image
  auto job = [&](const float* src, int rows, int cols, int trans, int tiles, int K, int planes, size_t off) {
    PackJob& jb = a.jobs[a.n_jobs++];
    jb.src = src;
    jb.rows = rows;
    jb.cols = cols;
    jb.ld = cols;
    jb.trans = trans;
    jb.tiles = tiles;
    jb.ksteps = g.prec == SRF_PREC_F32 ? K / 4 : K / 32;
    jb.planes = g.prec == SRF_PREC_F32 ? 0 : planes;
    jb.off = off;
  };
  // forward fragments: W [out, in] as A (rows = out)
  job(basis, g.NB, g.C, 0, 2, g.Cp, g.fpl, g.fB);
  job(w1, kHid, g.D1, 0, kHT, g.D1p, g.fpl, g.fW1);
  job(w2, kHid, kHid, 0, kHT, kHid, g.fpl, g.fW2);
  job(w3, 3, kHid, 0, 1, kHid, g.fpl, g.fW3);
  // backward fragments: W^T as A (rows = in)
  job(w3, 3, kHid, 1, kHT, g.kal, g.bpl, g.fW3T);
  job(w2, kHid, kHid, 1, kHT, kHid, g.bpl, g.fW2T);
  job(w1, kHid, g.D1, 1, g.D1p / 16, kHid, g.bpl, g.fW1T);
  job(basis, g.NB, g.C, 1, g.Cp / 16, kMaxB, g.bpl, g.fBT);
  a.b1 = b1;
  a.b2 = b2;
  a.b3 = b3;
  a.ob1 = g.b1;
  a.ob2 = g.b2;
  a.ob3 = g.b3;
  a.out = reinterpret_cast<float*>(packed);
  hipLaunchKernelGGL(pack_kernel, dim3(32, a.n_jobs + 1), dim3(256), 0, as_stream(stream), a);
  return launch_status();
}

extern "C" int srf_color_forward(const srf_color_desc* desc, const void* packed, int n_points,
                                 const int32_t* point_count, const float* app, const int32_t* idx, int n_samples,
                                 const float* view_dirs, float* rgb, int rgb_at_idx, void* workspace,
                                 void* stream) {
  Geo g;
  if (!make_geo(desc, &g) || n_points < 0 || (g.view && n_samples < 1)) return SRF_EINVAL;
  if (n_points == 0) return SRF_OK;
  if (!packed || !app || !rgb || (g.view && !view_dirs) || (rgb_at_idx && !idx)) return SRF_EINVAL;
  if (reinterpret_cast<uintptr_t>(app) & 15) return SRF_EINVAL;
  FwdArgs a;
  a.g = g;
  a.w = make_ws(g, n_points);
  if (workspace && !ws_addressable(g, a.w)) return SRF_ETOOLARGE;
  a.pk = reinterpret_cast<const float*>(packed);
  a.n = n_points;
  a.count = point_count;
  a.app = app;
  a.idx = idx;
  a.n_samples = n_samples > 0 ? n_samples : 1;
  a.view_dirs = view_dirs;
  a.rgb = rgb;
  a.rgb_at_idx = rgb_at_idx != 0;
  a.ws = reinterpret_cast<uint8_t*>(workspace);
  const int blocks = a.w.cols / (kPW * kWaves);
  switch (fwd_prec_of(g.prec)) {
    case kBFS3: hipLaunchKernelGGL(color_fwd_kernel<kBFS3>, dim3(blocks), dim3(64 * kWaves), 0, as_stream(stream), a); break;
    case kBFS1: hipLaunchKernelGGL(color_fwd_kernel<kBFS1>, dim3(blocks), dim3(64 * kWaves), 0, as_stream(stream), a); break;
    default: hipLaunchKernelGGL(color_fwd_kernel<kF32>, dim3(blocks), dim3(64 * kWaves), 0, as_stream(stream), a); break;
  }
  return launch_status();
}

extern "C" int srf_color_backward(const srf_color_desc* desc, const void* packed, int n_points,
                                  const int32_t* point_count, const int32_t* idx, int rgb_at_idx, const float* rgb,
                                  const float* d_rgb, void* workspace, float* d_app, float* d_basis, float* d_w1,
                                  float* d_b1, float* d_w2, float* d_b2, float* d_w3, float* d_b3, void* stream) {
  Geo g;
  if (!make_geo(desc, &g) || n_points < 0) return SRF_EINVAL;
  if (n_points == 0) return SRF_OK;
  if (!packed || !rgb || !d_rgb || !workspace || !d_app || !d_basis || !d_w1 || !d_b1 || !d_w2 || !d_b2 || !d_w3 ||
      !d_b3 || (rgb_at_idx && !idx))
    return SRF_EINVAL;
  if (reinterpret_cast<uintptr_t>(d_app) & 15) return SRF_EINVAL;
  hipStream_t s = as_stream(stream);
  BwdArgs b;
  b.g = g;
  b.w = make_ws(g, n_points);
  if (!ws_addressable(g, b.w)) return SRF_ETOOLARGE;
  b.pk = reinterpret_cast<const float*>(packed);
  b.n = n_points;
  b.count = point_count;
  b.idx = idx;
  b.rgb_at_idx = rgb_at_idx != 0;
  b.rgb = rgb;
  b.d_rgb = d_rgb;
  b.d_app = d_app;
  b.ws = reinterpret_cast<uint8_t*>(workspace);
  const dim3 grid(b.w.cols / (kPW * kWaves));
  switch (bwd_prec_of(g.prec)) {
    case kBFS2: hipLaunchKernelGGL(color_bwd_split_kernel, grid, dim3(64 * kWaves), 0, s, b); break;
    case kBFS1: hipLaunchKernelGGL(color_bwd_kernel<kBFS1>, grid, dim3(64 * kWaves), 0, s, b); break;
    default: hipLaunchKernelGGL(color_bwd_kernel<kF32>, grid, dim3(64 * kWaves), 0, s, b); break;
  }
  int st = launch_status();
  if (st != SRF_OK) return st;
  // weight gradients over the stored dZ / input tiles (color_dw_kernel)
  const Ws& w = b.w;
  uint8_t* ws = b.ws;
  CdwArgs c;
  memset(&c, 0, sizeof(c));
  auto cjob = [&](size_t dz, int dz_rows, size_t x, int x_rows, int M, int N, float* dw, int ld, float* db) {
    CdwJob& jb = c.jobs[c.n_jobs++];
    jb.dz = ws + dz;
    jb.x = ws + x;
    jb.nA = dz_rows / 16;
    jb.nB = x_rows / 16;
    jb.M = M;
    jb.N = N;
    jb.dw = dw;
    jb.ld_dw = ld;
    jb.db = db;
    return jb.nA <= kCdwMaxA && jb.nB <= kCdwMaxB;
  };
  if (!cjob(w.dz1, kHid, w.x1, g.D1p, kHid, g.D1, d_w1, g.D1, d_b1) ||
      !cjob(w.dz2, kHid, w.h1, kHid, kHid, kHid, d_w2, kHid, d_b2) ||
      !cjob(w.dz3, 16, w.h2, kHid, 3, kHid, d_w3, kHid, d_b3) ||
      !cjob(w.dF, kMaxB, w.appT, g.Cp, g.NB, g.C, d_basis, g.C, nullptr))
    return SRF_EINVAL;
  c.cols = w.cols;
  c.cols_dev = point_count;
  // 3840-point chunks by default (the items of the capacity's chunks past the device-side count
  // exit at once; measured within 2048-8448 on the 300^2 step, whose times differ by their last
  // partial round of blocks), or the autotuned size
  const long long tuned = tune_value(SRF_TUNE_COLOR_DW_CHUNK);
  c.chunk = tuned ? (int)tuned : 3840;
  if (g.prec != SRF_PREC_F32) {
    hipLaunchKernelGGL(color_dw_kernel, dim3(c.n_jobs * ((w.cols + c.chunk - 1) / c.chunk)), dim3(64 * kCdwWaves), 0,
                       s, c);
    return launch_status();
  }
  // FP32: the MLP's grouped FP32 GEMM on 128 x 160 tiles (one tile per job), whose fixed ring of
  // four k-units in flight keeps its MFMA-bound blocks busy
  mlp::DwArgs d;
  memset(&d, 0, sizeof(d));
  d.ldP = w.ld;
  d.cols = w.cols;
  d.cols_dev = point_count;
  d.paired = 1;  // tile blocks (point-block-major, as the narrow tiling reads them)
  d.n_jobs = c.n_jobs;
  for (int i = 0; i < c.n_jobs; ++i) {
    mlp::DwJob& jb = d.jobs[i];
    const CdwJob& cj = c.jobs[i];
    jb.dz = cj.dz;
    jb.dz_rows = 16 * cj.nA;
    jb.x = cj.x;
    jb.x_rows = 16 * cj.nB;
    jb.M = cj.M;
    jb.N = cj.N;
    jb.dw = cj.dw;
    jb.ld_dw = cj.ld_dw;
    jb.db = cj.db;
    jb.mblocks = (cj.M + mlp::kDwNarrowRows - 1) / mlp::kDwNarrowRows;
    jb.nblocks = (cj.N + mlp::kDwNarrowCols - 1) / mlp::kDwNarrowCols;
  }
  long long target = mlp::dw_default_chunk(w.cols, d.n_jobs, mlp::device_compute_units());
  if (tuned) target = tuned;
  d.chunk = (int)(target < 2048 ? 2048 : (target > 65536 ? 65536 : target));
  d.n_chunks = (w.cols + d.chunk - 1) / d.chunk;
  int work = 0;
  for (int i = 0; i < d.n_jobs; ++i) {
    d.jobs[i].work_begin = work;
    d.jobs[i].work_count = d.jobs[i].mblocks * d.jobs[i].nblocks * d.n_chunks;
    work += d.jobs[i].work_count;
  }
  d.total_work = work;
  return mlp::dw_f32_narrow(d, s);
}
