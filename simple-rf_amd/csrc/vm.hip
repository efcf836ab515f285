// TensoRF factor sampling on gfx950: the vector-matrix (VM) and CANDECOMP/PARAFAC (CP)
// decompositions.
//
// Reference semantics: LowRankTensor.forward validity + normalize_points
// (src/models/SimpleTensoRF09.py:701-765), AlphaGridMask.sample_alpha (:1342-1349),
// VmDecomposedTensor.get_volume_density / get_color before the basis matrix (:1214-1263),
// CpDecomposedTensor.get_volume_density / get_color before the basis matrix (:1043-1078),
// the density predictors (:671-679) and the backward of torch.nn.functional.grid_sample
// (bilinear, align_corners=True, zero padding) that autograd runs for them.
//
// Layout.  The factors live texel-major in HBM: plane p is [H_p][W_p][c_p] and line p is
// [L_p][c_p] (the reference's [1, c, H, W] / [1, c, L, 1] tensors viewed with the component
// axis innermost).  A bilinear tap then reads c_p contiguous floats.
//
// Mapping.  The channels of the three planes are cut into 4-channel quads (c_p % 4 == 0) and
// concatenated, Q quads in all.  A group of G lanes (G = next power of two >= Q) owns one
// sample; lane j of the group handles quads j, j + G, ...  Every tap of a group is thus one
// contiguous 16 B x (quads of that plane) read, appearance rows are written as coalesced
// 16 B stores, and the density sum over channels is a butterfly inside the group.
//
// Arithmetic follows grid_sample's op order: source index ((x + 1) / 2) * (size - 1), corner
// weights as products of the opposite distances, taps accumulated nw, ne, sw, se; the file is
// compiled with -ffp-contract=off so no product is fused into a sum.
#include <string.h>

#include <mutex>

#include "common.h"

namespace srf {
namespace {

constexpr int kMaxQuads = 64;  // at most 256 components over the three planes
constexpr int kPlanes = 3;

// matrix_axes = [[0,1],[0,2],[1,2]] and vector_axes = [2,1,0] (SimpleTensoRF09.py:1131-1132):
// plane p is sampled at (x = pts[kA0[p]], y = pts[kA1[p]]), line p at y = pts[kV[p]].
__device__ __forceinline__ int axis_a0(int p) { return p == 2 ? 1 : 0; }
__device__ __forceinline__ int axis_a1(int p) { return p == 0 ? 1 : 2; }
__device__ __forceinline__ int axis_v(int p) { return 2 - p; }

struct VmArgs {
  const float* plane[kPlanes];
  const float* line[kPlanes];
  float* dplane[kPlanes];
  float* dline[kPlanes];
  int ph[kPlanes], pw[kPlanes], ll[kPlanes], comp[kPlanes], qoff[kPlanes];
  int n_quads, total_comp;
  float lo[3], hi[3], size[3];
};

struct AlphaArgs {
  const float* vol;  // [D][H][W] or null
  int d, h, w;
  float lo[3], hi[3], size[3];
};

// Normalised coordinate of pts in the box: ((p - lo) / size) * 2 - 1.
__device__ __forceinline__ float normalize_coord(float p, float lo, float size) {
  return __fsub_rn(__fmul_rn(__fdiv_rn(__fsub_rn(p, lo), size), 2.0f), 1.0f);
}

// grid_sample's align_corners=True unnormalisation.
__device__ __forceinline__ float source_index(float x, int size) {
  return __fmul_rn(__fmul_rn(__fadd_rn(x, 1.0f), 0.5f), (float)(size - 1));
}

// The four bilinear corners of one plane sample; idx < 0 marks a corner outside the plane.
struct PlaneTaps {
  int idx[4];
  float w[4];
};

__device__ __forceinline__ PlaneTaps plane_taps(float x, float y, int W, int H) {
  const float ix = source_index(x, W), iy = source_index(y, H);
  const float x0 = floorf(ix), y0 = floorf(iy);
  const float x1 = __fadd_rn(x0, 1.0f), y1 = __fadd_rn(y0, 1.0f);
  PlaneTaps t;
  t.w[0] = __fmul_rn(__fsub_rn(x1, ix), __fsub_rn(y1, iy));  // nw
  t.w[1] = __fmul_rn(__fsub_rn(ix, x0), __fsub_rn(y1, iy));  // ne
  t.w[2] = __fmul_rn(__fsub_rn(x1, ix), __fsub_rn(iy, y0));  // sw
  t.w[3] = __fmul_rn(__fsub_rn(ix, x0), __fsub_rn(iy, y0));  // se
  const int xi = (int)x0, yi = (int)y0;
  const bool xin0 = xi >= 0 && xi < W, xin1 = xi + 1 >= 0 && xi + 1 < W;
  const bool yin0 = yi >= 0 && yi < H, yin1 = yi + 1 >= 0 && yi + 1 < H;
  t.idx[0] = (xin0 && yin0) ? yi * W + xi : -1;
  t.idx[1] = (xin1 && yin0) ? yi * W + xi + 1 : -1;
  t.idx[2] = (xin0 && yin1) ? (yi + 1) * W + xi : -1;
  t.idx[3] = (xin1 && yin1) ? (yi + 1) * W + xi + 1 : -1;
  return t;
}

// Line sample: grid_sample on [1, c, L, 1] at (x = 0, y): the x weight is exactly 1 and the
// x + 1 column is outside, so only the two y corners remain.
struct LineTaps {
  int idx[2];
  float w[2];
};

__device__ __forceinline__ LineTaps line_taps(float y, int L) {
  const float iy = source_index(y, L);
  const float y0 = floorf(iy);
  const float y1 = __fadd_rn(y0, 1.0f);
  LineTaps t;
  t.w[0] = __fsub_rn(y1, iy);
  t.w[1] = __fsub_rn(iy, y0);
  const int yi = (int)y0;
  t.idx[0] = (yi >= 0 && yi < L) ? yi : -1;
  t.idx[1] = (yi + 1 >= 0 && yi + 1 < L) ? yi + 1 : -1;
  return t;
}

__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }

__device__ __forceinline__ float4 madd4(float4 acc, float4 v, float w) {
  acc.x = __fadd_rn(acc.x, __fmul_rn(v.x, w));
  acc.y = __fadd_rn(acc.y, __fmul_rn(v.y, w));
  acc.z = __fadd_rn(acc.z, __fmul_rn(v.z, w));
  acc.w = __fadd_rn(acc.w, __fmul_rn(v.w, w));
  return acc;
}

__device__ __forceinline__ float4 mul4(float4 a, float4 b) {
  return make_float4(__fmul_rn(a.x, b.x), __fmul_rn(a.y, b.y), __fmul_rn(a.z, b.z), __fmul_rn(a.w, b.w));
}

__device__ __forceinline__ float sum4(float4 v) {
  return __fadd_rn(__fadd_rn(__fadd_rn(v.x, v.y), v.z), v.w);
}

// Plane index of quad q given the quad offsets of planes 1 and 2 (wave-uniform scalars).
__device__ __forceinline__ int quad_plane(const VmArgs& a, int q) {
  return (q >= a.qoff[1]) + (q >= a.qoff[2]);
}

template <typename T>
__device__ __forceinline__ T sel3(int p, T v0, T v1, T v2) {
  return p == 0 ? v0 : (p == 1 ? v1 : v2);
}

// Everything one lane needs to read or scatter its quad of one sample.
struct QuadSite {
  int p, c;  // plane, first component within the plane
  PlaneTaps pt;
  LineTaps lt;
};

__device__ __forceinline__ QuadSite quad_site(const VmArgs& a, int q, const float pn[3]) {
  QuadSite s;
  s.p = quad_plane(a, q);
  s.c = 4 * (q - sel3(s.p, a.qoff[0], a.qoff[1], a.qoff[2]));
  const float x = sel3(s.p, pn[axis_a0(0)], pn[axis_a0(1)], pn[axis_a0(2)]);
  const float y = sel3(s.p, pn[axis_a1(0)], pn[axis_a1(1)], pn[axis_a1(2)]);
  const float v = sel3(s.p, pn[axis_v(0)], pn[axis_v(1)], pn[axis_v(2)]);
  s.pt = plane_taps(x, y, sel3(s.p, a.pw[0], a.pw[1], a.pw[2]), sel3(s.p, a.ph[0], a.ph[1], a.ph[2]));
  s.lt = line_taps(v, sel3(s.p, a.ll[0], a.ll[1], a.ll[2]));
  return s;
}

// Bilinear plane value and linear line value of the 4 components of a quad.
__device__ __forceinline__ void quad_values(const VmArgs& a, const QuadSite& s, float4& pv, float4& lv) {
  const int C = sel3(s.p, a.comp[0], a.comp[1], a.comp[2]);
  const float* P = sel3(s.p, a.plane[0], a.plane[1], a.plane[2]);
  const float* L = sel3(s.p, a.line[0], a.line[1], a.line[2]);
  pv = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int t = 0; t < 4; ++t)
    if (s.pt.idx[t] >= 0) pv = madd4(pv, ld4(P + (int64_t)s.pt.idx[t] * C + s.c), s.pt.w[t]);
  lv = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int t = 0; t < 2; ++t)
    if (s.lt.idx[t] >= 0) lv = madd4(lv, ld4(L + (int64_t)s.lt.idx[t] * C + s.c), s.lt.w[t]);
}

// Trilinear grid_sample of the alpha volume (align_corners=True, zero padding); corners
// accumulated z0 (y0x0, y0x1, y1x0, y1x1) then z1, weights ((wx * wy) * wz).
__device__ float alpha_value(const AlphaArgs& m, const float p[3]) {
  float n[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) n[k] = normalize_coord(p[k], m.lo[k], m.size[k]);
  const float ix = source_index(n[0], m.w), iy = source_index(n[1], m.h), iz = source_index(n[2], m.d);
  const float x0 = floorf(ix), y0 = floorf(iy), z0 = floorf(iz);
  const float x1 = __fadd_rn(x0, 1.f), y1 = __fadd_rn(y0, 1.f), z1 = __fadd_rn(z0, 1.f);
  const float wx[2] = {__fsub_rn(x1, ix), __fsub_rn(ix, x0)};
  const float wy[2] = {__fsub_rn(y1, iy), __fsub_rn(iy, y0)};
  const float wz[2] = {__fsub_rn(z1, iz), __fsub_rn(iz, z0)};
  const int xi = (int)x0, yi = (int)y0, zi = (int)z0;
  float out = 0.f;
#pragma unroll
  for (int dz = 0; dz < 2; ++dz)
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) {
        const int X = xi + dx, Y = yi + dy, Z = zi + dz;
        if (X >= 0 && X < m.w && Y >= 0 && Y < m.h && Z >= 0 && Z < m.d) {
          const float w = __fmul_rn(__fmul_rn(wx[dx], wy[dy]), wz[dz]);
          out = __fadd_rn(out, __fmul_rn(m.vol[((int64_t)Z * m.h + Y) * m.w + X], w));
        }
      }
  return out;
}

__device__ __forceinline__ bool in_box(const float p[3], const float lo[3], const float hi[3]) {
  return lo[0] <= p[0] && p[0] <= hi[0] && lo[1] <= p[1] && p[1] <= hi[1] && lo[2] <= p[2] && p[2] <= hi[2];
}

template <int G>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int d = G / 2; d >= 1; d >>= 1) v = __fadd_rn(v, __shfl_xor(v, d, SRF_WAVE));
  return v;
}

// Density predictors (SimpleTensoRF09.py:671-679): ReLU, or softplus(x + offset) with
// torch's default beta = 1 and threshold = 20.
__device__ __forceinline__ float predict_density(float f, int kind, float offset) {
  if (kind == SRF_DENSITY_RELU) return f > 0.f ? f : 0.f;
  const float x = __fadd_rn(f, offset);
  return x > 20.f ? x : log1pf(expf(x));
}

__device__ __forceinline__ float predict_density_grad(float f, int kind, float offset) {
  if (kind == SRF_DENSITY_RELU) return f > 0.f ? 1.f : 0.f;
  const float x = __fadd_rn(f, offset);
  if (x > 20.f) return 1.f;
  const float e = expf(x);
  return e / __fadd_rn(e, 1.f);
}

struct GroupPos {
  int64_t first, stride;  // first sample slot of this lane's group, stride between passes
  int ql;                 // lane within the group
};

template <int G>
__device__ __forceinline__ GroupPos group_pos() {
  const int lane = threadIdx.x & (SRF_WAVE - 1);
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / SRF_WAVE;
  const int64_t n_waves = (int64_t)gridDim.x * blockDim.x / SRF_WAVE;
  GroupPos g;
  g.first = wave * (SRF_WAVE / G) + lane / G;
  g.stride = n_waves * (SRF_WAVE / G);
  g.ql = lane % G;
  return g;
}

// --------------------------------------------------------------------------- kernels

template <int G>
__global__ __launch_bounds__(256) void density_fwd_kernel(VmArgs a, AlphaArgs m, int64_t n,
                                                          const float* __restrict__ pts, int kind,
                                                          float offset, float* __restrict__ sigma,
                                                          float* __restrict__ feat_out,
                                                          uint8_t* __restrict__ valid_out) {
  const GroupPos gp = group_pos<G>();
  // all lanes of a group share i, so the group stays converged for its butterfly
  for (int64_t i = gp.first; i < n; i += gp.stride) {
    const float p[3] = {pts[3 * i + 0], pts[3 * i + 1], pts[3 * i + 2]};
    bool valid = in_box(p, a.lo, a.hi);
    if (valid && m.vol) valid = alpha_value(m, p) > 0.f;
    float pn[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) pn[k] = normalize_coord(p[k], a.lo[k], a.size[k]);
    float s[kPlanes] = {0.f, 0.f, 0.f};
    if (valid) {
      for (int q = gp.ql; q < a.n_quads; q += G) {
        const QuadSite site = quad_site(a, q, pn);
        float4 pv, lv;
        quad_values(a, site, pv, lv);
        const float d = sum4(mul4(pv, lv));
        s[0] = site.p == 0 ? __fadd_rn(s[0], d) : s[0];
        s[1] = site.p == 1 ? __fadd_rn(s[1], d) : s[1];
        s[2] = site.p == 2 ? __fadd_rn(s[2], d) : s[2];
      }
    }
#pragma unroll
    for (int k = 0; k < kPlanes; ++k) s[k] = group_sum<G>(s[k]);
    if (gp.ql == 0) {
      const float f = __fadd_rn(__fadd_rn(s[0], s[1]), s[2]);
      sigma[i] = valid ? predict_density(f, kind, offset) : 0.f;
      if (feat_out) feat_out[i] = valid ? f : 0.f;
      if (valid_out) valid_out[i] = valid ? 1 : 0;
    }
  }
}

// Gradient of one lane's quad summed over a run of samples that share the same corners, written
// to HBM with atomics only when the corners change.  Consecutive samples of a ray mostly share
// them (the step is half a voxel, and in NDC a ray's x, y barely move), so the atomics to the
// most contended texels drop by up to the run length.
constexpr int kRun = 32;  // samples walked by one lane group (SRF_TUNE_GATHER_RUN overrides it)

struct RunAcc {
  int pidx[4], lidx[2];
  float4 pg[4], lg[2];
};

__device__ __forceinline__ void run_init(RunAcc& r) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    r.pidx[t] = -1;
    r.pg[t] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    r.lidx[t] = -1;
    r.lg[t] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

__device__ __forceinline__ void atomic_add4v(float* p, float4 v) {
  unsafeAtomicAdd(p + 0, v.x);
  unsafeAtomicAdd(p + 1, v.y);
  unsafeAtomicAdd(p + 2, v.z);
  unsafeAtomicAdd(p + 3, v.w);
}

__device__ __forceinline__ void run_flush_plane(const VmArgs& a, int p, int c, RunAcc& r) {
  const int C = sel3(p, a.comp[0], a.comp[1], a.comp[2]);
  float* dP = sel3(p, a.dplane[0], a.dplane[1], a.dplane[2]);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    if (r.pidx[t] >= 0) atomic_add4v(dP + (int64_t)r.pidx[t] * C + c, r.pg[t]);
    r.pg[t] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

__device__ __forceinline__ void run_flush_line(const VmArgs& a, int p, int c, RunAcc& r) {
  const int C = sel3(p, a.comp[0], a.comp[1], a.comp[2]);
  float* dL = sel3(p, a.dline[0], a.dline[1], a.dline[2]);
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    if (r.lidx[t] >= 0) atomic_add4v(dL + (int64_t)r.lidx[t] * C + c, r.lg[t]);
    r.lg[t] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// Per-corner run accumulation.  Slot k keeps corner role k (plane: nw, ne, sw, se; line: lower,
// upper), so a sample whose cell is the previous one moved by one texel along an axis carries
// the shared half of the slots over (the texel row or column both cells touch) and flushes only
// the half that left: along a ray, a texel is flushed once instead of once per cell that uses it.
// A slot is carried only when its texel index equals the new corner's, so any move is handled
// correctly (unmatched slots are flushed); index -1 marks a corner outside the factor, whose
// slot is never flushed.  Static selects only: the slots stay in registers.
__device__ __forceinline__ void flush_slot(float* grad, int idx, float4 acc, int C, int c) {
  if (idx >= 0) atomic_add4v(grad + (int64_t)idx * C + c, acc);
}

// The block's LDS copy of a line (dynamic shared memory of the cell-sorted backward).  Named at
// file scope so that every access is known to be an LDS access (ds_add_f32), not a flat one.
extern __shared__ float g_line_lds[];

__device__ __forceinline__ void flush_slot_lds(int idx, float4 acc, int C, int c) {
  if (idx < 0) return;
  const int o = idx * C + c;
  atomicAdd(&g_line_lds[o + 0], acc.x);
  atomicAdd(&g_line_lds[o + 1], acc.y);
  atomicAdd(&g_line_lds[o + 2], acc.z);
  atomicAdd(&g_line_lds[o + 3], acc.w);
}

__device__ __forceinline__ float4 pick4(bool b, float4 x, float4 y) {
  return make_float4(b ? x.x : y.x, b ? x.y : y.y, b ? x.z : y.z, b ? x.w : y.w);
}

__device__ __forceinline__ void plane_accum(int (&si)[4], float4 (&sa)[4], const int (&ci)[4], const float (&cw)[4],
                                            float4 g, float* grad, int C, int c) {
  const bool same = ci[0] == si[0] && ci[1] == si[1] && ci[2] == si[2] && ci[3] == si[3];
  const bool px = !same && ci[0] == si[1] && ci[2] == si[3];         // cell moved +x: nw<-ne, sw<-se
  const bool nx = !same && !px && ci[1] == si[0] && ci[3] == si[2];  // -x: ne<-nw, se<-sw
  const bool py = !same && !px && !nx && ci[0] == si[2] && ci[1] == si[3];          // +y: nw<-sw, ne<-se
  const bool ny = !same && !px && !nx && !py && ci[2] == si[0] && ci[3] == si[1];  // -y: sw<-nw, se<-ne
  if (!(same || nx || ny)) flush_slot(grad, si[0], sa[0], C, c);
  if (!(same || px || ny)) flush_slot(grad, si[1], sa[1], C, c);
  if (!(same || nx || py)) flush_slot(grad, si[2], sa[2], C, c);
  if (!(same || px || py)) flush_slot(grad, si[3], sa[3], C, c);
  const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 n0 = pick4(same, sa[0], pick4(px, sa[1], pick4(py, sa[2], z)));
  const float4 n1 = pick4(same, sa[1], pick4(nx, sa[0], pick4(py, sa[3], z)));
  const float4 n2 = pick4(same, sa[2], pick4(px, sa[3], pick4(ny, sa[0], z)));
  const float4 n3 = pick4(same, sa[3], pick4(nx, sa[2], pick4(ny, sa[1], z)));
  sa[0] = madd4(n0, g, cw[0]);
  sa[1] = madd4(n1, g, cw[1]);
  sa[2] = madd4(n2, g, cw[2]);
  sa[3] = madd4(n3, g, cw[3]);
#pragma unroll
  for (int k = 0; k < 4; ++k) si[k] = ci[k];
}

template <bool kLds = false>
__device__ __forceinline__ void line_accum(int (&si)[2], float4 (&sa)[2], const int (&ci)[2], const float (&cw)[2],
                                           float4 g, float* grad, int C, int c) {
  const bool same = ci[0] == si[0] && ci[1] == si[1];
  const bool up = !same && ci[0] == si[1];          // lower <- upper
  const bool down = !same && !up && ci[1] == si[0];  // upper <- lower
  if (kLds) {  // grad unused: the slots flush into the LDS line copy
    if (!(same || down)) flush_slot_lds(si[0], sa[0], C, c);
    if (!(same || up)) flush_slot_lds(si[1], sa[1], C, c);
  } else {
    if (!(same || down)) flush_slot(grad, si[0], sa[0], C, c);
    if (!(same || up)) flush_slot(grad, si[1], sa[1], C, c);
  }
  const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 n0 = pick4(same, sa[0], pick4(up, sa[1], z));
  const float4 n1 = pick4(same, sa[1], pick4(down, sa[0], z));
  sa[0] = madd4(n0, g, cw[0]);
  sa[1] = madd4(n1, g, cw[1]);
  si[0] = ci[0];
  si[1] = ci[1];
}

// Adds one sample's plane gradient gp and line gradient gl (before the corner weights).
__device__ __forceinline__ void run_add(const VmArgs& a, const QuadSite& s, RunAcc& r, float4 gp, float4 gl) {
  const int C = sel3(s.p, a.comp[0], a.comp[1], a.comp[2]);
  float* dP = sel3(s.p, a.dplane[0], a.dplane[1], a.dplane[2]);
  float* dL = sel3(s.p, a.dline[0], a.dline[1], a.dline[2]);
  const int c = s.c;
  plane_accum(r.pidx, r.pg, s.pt.idx, s.pt.w, gp, dP, C, c);
  line_accum(r.lidx, r.lg, s.lt.idx, s.lt.w, gl, dL, C, c);
}

// d_feat = d_sigma * predictor'(feat) on valid samples; planes receive d_feat * line and
// lines d_feat * plane, per component and corner.  A lane group walks run_len consecutive samples.
template <int G>
__global__ __launch_bounds__(256) void density_bwd_kernel(VmArgs a, int64_t n, int run_len, const float* __restrict__ pts,
                                                          const uint8_t* __restrict__ valid,
                                                          const float* __restrict__ feat, int kind,
                                                          float offset, const float* __restrict__ d_sigma) {
  const GroupPos gp = group_pos<G>();
  const int64_t n_runs = (n + run_len - 1) / run_len;
  for (int64_t run = gp.first; run < n_runs; run += gp.stride) {
    const int64_t i0 = run * run_len;
    const int64_t i1 = i0 + run_len < n ? i0 + run_len : n;
    for (int q = gp.ql; q < a.n_quads; q += G) {
      RunAcc r;
      run_init(r);
      int p = 0, c = 0;
      for (int64_t i = i0; i < i1; ++i) {
        if (!valid[i]) continue;
        const float g = __fmul_rn(d_sigma[i], predict_density_grad(feat[i], kind, offset));
        if (g == 0.f) continue;
        float pn[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) pn[k] = normalize_coord(pts[3 * i + k], a.lo[k], a.size[k]);
        const QuadSite site = quad_site(a, q, pn);
        float4 pv, lv;
        quad_values(a, site, pv, lv);
        const float4 gg = make_float4(g, g, g, g);
        run_add(a, site, r, mul4(gg, lv), mul4(gg, pv));
        p = site.p;
        c = site.c;
      }
      run_flush_plane(a, p, c, r);
      run_flush_line(a, p, c, r);
    }
  }
}

// Appearance features of the listed samples: out[j, c] = plane_c * line_c for the sample
// idx[j] (or j when idx is null), components concatenated over the planes.
template <int G>
__global__ __launch_bounds__(256) void appearance_fwd_kernel(VmArgs a, int64_t n, const int32_t* __restrict__ n_dev,
                                                             const int32_t* __restrict__ idx,
                                                             const float* __restrict__ pts,
                                                             float* __restrict__ out) {
  const GroupPos gp = group_pos<G>();
  if (n_dev) n = min(n, (int64_t)*n_dev);  // device-side row count (n is the capacity)
  for (int64_t j = gp.first; j < n; j += gp.stride) {
    const int64_t i = idx ? (int64_t)idx[j] : j;
    float pn[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) pn[k] = normalize_coord(pts[3 * i + k], a.lo[k], a.size[k]);
    for (int q = gp.ql; q < a.n_quads; q += G) {
      const QuadSite site = quad_site(a, q, pn);
      float4 pv, lv;
      quad_values(a, site, pv, lv);
      *reinterpret_cast<float4*>(out + j * a.total_comp + 4 * q) = mul4(pv, lv);
    }
  }
}

template <int G>
__global__ __launch_bounds__(256) void appearance_bwd_kernel(VmArgs a, int64_t n, const int32_t* __restrict__ n_dev,
                                                             int run_len, const int32_t* __restrict__ idx,
                                                             const float* __restrict__ pts,
                                                             const float* __restrict__ d_out) {
  const GroupPos gp = group_pos<G>();
  if (n_dev) n = min(n, (int64_t)*n_dev);  // device-side row count (n is the capacity)
  const int64_t n_runs = (n + run_len - 1) / run_len;
  for (int64_t run = gp.first; run < n_runs; run += gp.stride) {
    const int64_t j0 = run * run_len;
    const int64_t j1 = j0 + run_len < n ? j0 + run_len : n;
    for (int q = gp.ql; q < a.n_quads; q += G) {
      RunAcc r;
      run_init(r);
      int p = 0, c = 0;
      for (int64_t j = j0; j < j1; ++j) {
        const int64_t i = idx ? (int64_t)idx[j] : j;
        float pn[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) pn[k] = normalize_coord(pts[3 * i + k], a.lo[k], a.size[k]);
        const float4 g = ld4(d_out + j * a.total_comp + 4 * q);
        const QuadSite site = quad_site(a, q, pn);
        float4 pv, lv;
        quad_values(a, site, pv, lv);
        run_add(a, site, r, mul4(g, lv), mul4(g, pv));
        p = site.p;
        c = site.c;
      }
      run_flush_plane(a, p, c, r);
      run_flush_line(a, p, c, r);
    }
  }
}

// ------------------------------------------------ cell-sorted backward (privatised scatter)
// The run walk above flushes a texel whenever a ray's consecutive samples leave its cell.  A plane
// texel is touched by many rays, so the scatter still issues tens of atomics per texel and
// component.  With a caller workspace the backward instead orders the contributing rows by the
// plane cell of their nw corner (a counting sort per plane: histogram, one scan, a scatter of row
// numbers), so that a lane group walking the sorted rows meets each cell's samples in one run and
// flushes each texel about once per run; the line gradients of the walk (the third coordinate of
// the rows of one cell is arbitrary) go to a copy of the line in LDS, flushed once per block.
// The order within a cell comes from integer atomics and varies between calls, as the order of
// the FP32 atomics does anyway.
struct BwdRows {
  int64_t n;               // capacity
  const int32_t* n_dev;    // device-side row count (appearance) or null
  const int32_t* idx;      // sample of row j (appearance) or null
  const float* pts;
  const uint8_t* valid;    // density: rows are samples, valid and with a non-zero gradient
  const float* feat;
  int kind;
  float offset;
  const float* d_sigma;
  const float* d_out;      // appearance: d_features [rows, total_comp]
};

constexpr int kMaxChunks = 1024;  // scan chunks per plane

struct SortWs {
  int32_t* hist[kPlanes];    // cells + 1: counts, then exclusive offsets
  int32_t* cursor[kPlanes];  // cells: scatter positions
  float4* order[kPlanes];    // capacity: per row in cell order, its normalised point and either its
                             // scalar gradient (density) or its row number as int bits (appearance)
  int32_t* total;            // [kPlanes] rows per plane
  int32_t* partial;          // [kPlanes][kMaxChunks] chunk sums, then chunk offsets
  int64_t chunk_cells;       // cells per scan chunk
  int cells[kPlanes];
  int32_t* zero_begin;       // the histograms, zeroed before each call
  int64_t zero_count;
};

__device__ __forceinline__ int64_t bwd_rows_count(const BwdRows& r) {
  return r.n_dev ? min(r.n, (int64_t)*r.n_dev) : r.n;
}

// Whether row j contributes, its sample i and (density) its scalar gradient.
template <bool kDensity>
__device__ __forceinline__ bool bwd_row(const BwdRows& r, int64_t j, int64_t& i, float& g) {
  if (kDensity) {
    i = j;
    if (!r.valid[j]) return false;
    g = __fmul_rn(r.d_sigma[j], predict_density_grad(r.feat[j], r.kind, r.offset));
    return g != 0.f;
  }
  i = r.idx ? (int64_t)r.idx[j] : j;
  g = 0.f;
  return true;
}

__device__ __forceinline__ void normalized_point(const VmArgs& a, const float* pts, int64_t i, float pn[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) pn[k] = normalize_coord(pts[3 * i + k], a.lo[k], a.size[k]);
}

// Sort key of a sample for plane p: the cell of its nw corner, clamped into the plane.
__device__ __forceinline__ int plane_cell(const VmArgs& a, int p, const float pn[3]) {
  const int W = sel3(p, a.pw[0], a.pw[1], a.pw[2]), H = sel3(p, a.ph[0], a.ph[1], a.ph[2]);
  const float x = sel3(p, pn[axis_a0(0)], pn[axis_a0(1)], pn[axis_a0(2)]);
  const float y = sel3(p, pn[axis_a1(0)], pn[axis_a1(1)], pn[axis_a1(2)]);
  int xi = (int)floorf(source_index(x, W)), yi = (int)floorf(source_index(y, H));
  xi = xi < 0 ? 0 : (xi > W - 1 ? W - 1 : xi);
  yi = yi < 0 ? 0 : (yi > H - 1 ? H - 1 : yi);
  return yi * W + xi;
}

__global__ __launch_bounds__(256) void fill_zero_kernel(int32_t* __restrict__ p, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    p[e] = 0;
}

// Runs of equal keys among the active lanes of a wave (consecutive rows of a ray mostly share a
// cell): the run's first lane adds the run length with one atomic, instead of one per row.
struct WaveRun {
  bool head;      // this lane starts a run
  int head_lane;  // first lane of this lane's run
  int length;     // run length (meaningful on heads)
};

__device__ __forceinline__ WaveRun wave_run(int key, bool active) {
  const int lane = threadIdx.x & (SRF_WAVE - 1);
  const int prev_key = __shfl_up(key, 1, SRF_WAVE);
  const int prev_active = __shfl_up(active ? 1 : 0, 1, SRF_WAVE);
  WaveRun r;
  r.head = active && (lane == 0 || !prev_active || prev_key != key);
  const uint64_t heads = __ballot(r.head);
  const uint64_t breaks = heads | ~__ballot(active);  // a run ends before the next head or inactive lane
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const uint64_t h = heads & upto;
  r.head_lane = h ? 63 - __clzll(h) : lane;
  const uint64_t after = lane == 63 ? 0ull : (breaks >> (lane + 1));
  r.length = after ? __ffsll((long long)after) : 64 - lane;
  return r;
}

// Rows are visited wave by wave (every lane of a wave runs the same iterations, so the wave-wide
// operations of wave_run see the whole wave).
template <bool kDensity, bool kScatter>
__global__ __launch_bounds__(256) void cell_count_kernel(VmArgs a, BwdRows r, SortWs w) {
  const int64_t n = bwd_rows_count(r);
  const int lane = threadIdx.x & (SRF_WAVE - 1);
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / SRF_WAVE;
  const int64_t n_waves = (int64_t)gridDim.x * blockDim.x / SRF_WAVE;
  for (int64_t base = wave * SRF_WAVE; base < n; base += n_waves * SRF_WAVE) {
    const int64_t j = base + lane;
    int64_t i = 0;
    float g = 0.f;
    const bool active = j < n && bwd_row<kDensity>(r, j, i, g);
    float pn[3] = {0.f, 0.f, 0.f};
    if (active) normalized_point(a, r.pts, i, pn);
#pragma unroll
    for (int p = 0; p < kPlanes; ++p) {
      if (a.comp[p] == 0) continue;  // uniform
      const int key = active ? plane_cell(a, p, pn) : -1;
      const WaveRun run = wave_run(key, active);
      if (!kScatter) {
        if (run.head) atomicAdd(w.hist[p] + key, run.length);
      } else {
        int pos = 0;
        if (run.head) pos = atomicAdd(w.cursor[p] + key, run.length);
        const int first = __shfl(pos, run.head_lane, SRF_WAVE);
        if (active)
          w.order[p][first + (lane - run.head_lane)] =
              make_float4(pn[0], pn[1], pn[2], kDensity ? g : __int_as_float((int)j));
      }
    }
  }
}

// Exclusive prefix of v over a block of kScanThreads threads, and the block total.
constexpr int kScanThreads = 1024;

__device__ __forceinline__ int32_t block_exclusive_scan(int32_t v, int32_t* total) {
  __shared__ int32_t wave_tot[kScanThreads / SRF_WAVE];
  const int lane = threadIdx.x & (SRF_WAVE - 1), wv = threadIdx.x / SRF_WAVE;
  int32_t incl = v;
#pragma unroll
  for (int d = 1; d < SRF_WAVE; d <<= 1) {
    const int32_t o = __shfl_up(incl, d, SRF_WAVE);
    if (lane >= d) incl += o;
  }
  if (lane == SRF_WAVE - 1) wave_tot[wv] = incl;
  __syncthreads();
  int32_t before = 0, all = 0;
#pragma unroll
  for (int k = 0; k < kScanThreads / SRF_WAVE; ++k) {
    const int32_t t = wave_tot[k];
    if (k < wv) before += t;
    all += t;
  }
  __syncthreads();  // wave_tot may be rewritten by the next call
  *total = all;
  return before + incl - v;
}

// The cell scan in three launches: per-chunk sums, a scan of the chunk sums (one block per plane,
// which also writes the plane's row total), then each chunk scanned with its carry-in, writing the
// exclusive offsets and the scatter cursors.  Chunks are kScanThreads-cell tiles read coalesced.
__global__ __launch_bounds__(kScanThreads) void cell_chunk_sum_kernel(SortWs w) {
  const int p = blockIdx.y;
  const int64_t c0 = (int64_t)blockIdx.x * w.chunk_cells;
  if (c0 >= w.cells[p]) return;  // uniform
  const int64_t c1 = c0 + w.chunk_cells < w.cells[p] ? c0 + w.chunk_cells : w.cells[p];
  int32_t v = 0;
  for (int64_t c = c0 + threadIdx.x; c < c1; c += kScanThreads) v += w.hist[p][c];
  int32_t total;
  block_exclusive_scan(v, &total);
  if (threadIdx.x == 0) w.partial[p * kMaxChunks + blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void cell_chunk_scan_kernel(SortWs w) {
  const int p = blockIdx.x;
  const int n_chunks = (int)((w.cells[p] + w.chunk_cells - 1) / w.chunk_cells);
  const int t = threadIdx.x;
  const int32_t v = t < n_chunks ? w.partial[p * kMaxChunks + t] : 0;
  int32_t total;
  const int32_t ex = block_exclusive_scan(v, &total);
  if (t < n_chunks) w.partial[p * kMaxChunks + t] = ex;
  if (t == 0) w.total[p] = total;
}

__global__ __launch_bounds__(kScanThreads) void cell_chunk_apply_kernel(SortWs w) {
  const int p = blockIdx.y;
  const int64_t c0 = (int64_t)blockIdx.x * w.chunk_cells;
  if (c0 >= w.cells[p]) return;  // uniform
  const int64_t c1 = c0 + w.chunk_cells < w.cells[p] ? c0 + w.chunk_cells : w.cells[p];
  int32_t carry = w.partial[p * kMaxChunks + blockIdx.x];
  for (int64_t t0 = c0; t0 < c1; t0 += kScanThreads) {  // uniform trip count
    const int64_t c = t0 + threadIdx.x;
    const int32_t v = c < c1 ? w.hist[p][c] : 0;
    int32_t total;
    const int32_t ex = block_exclusive_scan(v, &total);
    if (c < c1) {
      w.hist[p][c] = carry + ex;
      w.cursor[p][c] = carry + ex;
    }
    carry += total;
  }
}

constexpr int kSortedThreads = 512;
constexpr int kMaxLineLds = 160 * 1024;  // a CU's LDS
// Component slices of the sorted backward: a line copy above kLineLdsSplit (at most two blocks per
// CU) is split into slices of at most kLineLdsSlice.  Measured: the 512 x 64 appearance line (128 KB,
// one block per CU) walks faster in four slices (4.42 -> 4.08 ms per uneven-batch step), while
// splitting the 57.6 KB and 48 KB copies of the 300^2 appearance and 512^2 density lines cost more
// (re-walked rows) than the added blocks gained.
constexpr int kLineLdsSplit = 64 * 1024;
constexpr int kLineLdsSlice = 32 * 1024;

// Plane p's gradients from its cell-sorted rows: a lane group walks run_len consecutive sorted
// rows with the corner slots of the run walk (flushing a plane texel when the walk leaves it) and
// adds the line gradients into g_line_lds (the line copy in LDS, line_lds != 0) or, when the line is
// too large for LDS, through the line slots of the run walk.  The plane's components are split into
// gridDim.y slices of whole quads: a block walks its rows for one slice and keeps only that slice's
// columns of the line in LDS (L x Cs floats), so that more blocks fit in a CU's LDS.
template <int G, bool kDensity>
__global__ __launch_bounds__(kSortedThreads) void sorted_bwd_kernel(VmArgs a, BwdRows r, SortWs w, int p,
                                                                    int run_len, int line_lds) {
  const int C = sel3(p, a.comp[0], a.comp[1], a.comp[2]);
  const int L = sel3(p, a.ll[0], a.ll[1], a.ll[2]);
  const int qs = C / 4 / (int)gridDim.y;           // quads of the slice
  const int Cs = 4 * qs, cb = Cs * (int)blockIdx.y;  // its components: cb .. cb + Cs - 1
  const int64_t m = w.total[p];
  const int groups = kSortedThreads / G;
  const int64_t b0 = (int64_t)blockIdx.x * groups * run_len;
  if (b0 >= m) return;  // uniform over the block
  if (line_lds) {
    for (int e = threadIdx.x; e < L * Cs; e += kSortedThreads) g_line_lds[e] = 0.f;
    __syncthreads();
  }
  const int grp = threadIdx.x / G, ql = threadIdx.x % G;
  const int64_t s0 = b0 + (int64_t)grp * run_len;
  const int64_t s1 = s0 + run_len < m ? s0 + run_len : m;
  const int q0 = sel3(p, a.qoff[0], a.qoff[1], a.qoff[2]);
  float* dP = sel3(p, a.dplane[0], a.dplane[1], a.dplane[2]);
  float* dL = sel3(p, a.dline[0], a.dline[1], a.dline[2]);
  const float4* order = sel3(p, w.order[0], w.order[1], w.order[2]);
  for (int q = q0 + cb / 4 + ql; q < q0 + cb / 4 + qs; q += G) {
    RunAcc acc;
    run_init(acc);
    int c = 0;
    // The next row's entry (and its output gradient) are loaded one row ahead, so that a row waits
    // for its taps only; the look-ahead index is clamped to the last entry (always written).
    auto grad_of = [&](const float4& e) -> float4 {
      return kDensity ? make_float4(e.w, e.w, e.w, e.w)
                      : ld4(r.d_out + (int64_t)__float_as_int(e.w) * a.total_comp + 4 * q);
    };
    float4 en = order[s0 < m ? s0 : m - 1];
    float4 gn = grad_of(en);
    for (int64_t s = s0; s < s1; ++s) {
      const float4 e = en, g = gn;
      en = order[s + 1 < m ? s + 1 : m - 1];
      const float pn[3] = {e.x, e.y, e.z};
      const QuadSite site = quad_site(a, q, pn);
      float4 pv, lv;
      quad_values(a, site, pv, lv);
      plane_accum(acc.pidx, acc.pg, site.pt.idx, site.pt.w, mul4(g, lv), dP, C, site.c);
      const float4 gl = mul4(g, pv);
      // the line slots carry a texel while consecutive rows (one ray's samples within the cell) stay
      // on it, and flush into the LDS copy (or HBM) when the walk leaves it
      if (line_lds)
        line_accum<true>(acc.lidx, acc.lg, site.lt.idx, site.lt.w, gl, nullptr, Cs, site.c - cb);
      else
        line_accum<false>(acc.lidx, acc.lg, site.lt.idx, site.lt.w, gl, dL, C, site.c);
      c = site.c;
      gn = grad_of(en);
    }
    run_flush_plane(a, p, c, acc);
    for (int t = 0; t < 2; ++t) {
      if (line_lds)
        flush_slot_lds(acc.lidx[t], acc.lg[t], Cs, c - cb);
      else
        flush_slot(dL, acc.lidx[t], acc.lg[t], C, c);
    }
  }
  if (line_lds) {
    __syncthreads();
    for (int e = threadIdx.x; e < L * Cs; e += kSortedThreads) {
      const float v = g_line_lds[e];
      const int row = e / Cs;
      if (v != 0.f) unsafeAtomicAdd(dL + (int64_t)row * C + cb + (e - row * Cs), v);
    }
  }
}

// ------------------------------------------------------------ CP decomposition (3 lines)
// CpDecomposedTensor (SimpleTensoRF09.py:1043-1089): every component r is the product of the
// three lines at pts[2], pts[1], pts[0], multiplied in that order.  The lines share one rank, so
// a lane's quad reads the same 4 components of all three lines.

struct CpArgs {
  const float* line[kPlanes];
  float* dline[kPlanes];
  int ll[kPlanes];
  int comp, n_quads;
  float lo[3], hi[3], size[3];
};

struct CpTaps {
  LineTaps t[kPlanes];
};

__device__ __forceinline__ CpTaps cp_taps(const CpArgs& a, const float pn[3]) {
  CpTaps s;
#pragma unroll
  for (int k = 0; k < kPlanes; ++k) s.t[k] = line_taps(pn[axis_v(k)], a.ll[k]);
  return s;
}

// Linear line values of the 4 components starting at c, for the three lines.
__device__ __forceinline__ void cp_values(const CpArgs& a, const CpTaps& s, int c, float4 v[kPlanes]) {
#pragma unroll
  for (int k = 0; k < kPlanes; ++k) {
    v[k] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int t = 0; t < 2; ++t)
      if (s.t[k].idx[t] >= 0) v[k] = madd4(v[k], ld4(a.line[k] + (int64_t)s.t[k].idx[t] * a.comp + c), s.t[k].w[t]);
  }
}

__device__ __forceinline__ float4 cp_product(const float4 v[kPlanes]) { return mul4(mul4(v[0], v[1]), v[2]); }

// Run-length gradient accumulation for the three lines of one quad (as RunAcc for VM).
struct CpRun {
  int idx[kPlanes][2];
  float4 g[kPlanes][2];
};

__device__ __forceinline__ void cp_run_init(CpRun& r) {
#pragma unroll
  for (int k = 0; k < kPlanes; ++k)
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      r.idx[k][t] = -1;
      r.g[k][t] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

__device__ __forceinline__ void cp_run_flush(const CpArgs& a, int k, int c, CpRun& r) {
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    if (r.idx[k][t] >= 0) atomic_add4v(a.dline[k] + (int64_t)r.idx[k][t] * a.comp + c, r.g[k][t]);
    r.g[k][t] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// Adds the gradient g of the quad's component products: line k receives g times the other two
// lines' values, spread over its two taps.
__device__ __forceinline__ void cp_run_add(const CpArgs& a, const CpTaps& s, int c, CpRun& r, float4 g,
                                           const float4 v[kPlanes]) {
  const float4 d[kPlanes] = {mul4(mul4(g, v[1]), v[2]), mul4(mul4(g, v[0]), v[2]), mul4(mul4(g, v[0]), v[1])};
#pragma unroll
  for (int k = 0; k < kPlanes; ++k)
    line_accum(r.idx[k], r.g[k], s.t[k].idx, s.t[k].w, d[k], a.dline[k], a.comp, c);
}

template <int G>
__global__ __launch_bounds__(256) void cp_density_fwd_kernel(CpArgs a, AlphaArgs m, int64_t n,
                                                             const float* __restrict__ pts, int kind,
                                                             float offset, float* __restrict__ sigma,
                                                             float* __restrict__ feat_out,
                                                             uint8_t* __restrict__ valid_out) {
  const GroupPos gp = group_pos<G>();
  for (int64_t i = gp.first; i < n; i += gp.stride) {
    const float p[3] = {pts[3 * i + 0], pts[3 * i + 1], pts[3 * i + 2]};
    bool valid = in_box(p, a.lo, a.hi);
    if (valid && m.vol) valid = alpha_value(m, p) > 0.f;
    float pn[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) pn[k] = normalize_coord(p[k], a.lo[k], a.size[k]);
    float s = 0.f;
    if (valid) {
      const CpTaps taps = cp_taps(a, pn);
      for (int q = gp.ql; q < a.n_quads; q += G) {
        float4 v[kPlanes];
        cp_values(a, taps, 4 * q, v);
        s = __fadd_rn(s, sum4(cp_product(v)));
      }
    }
    s = group_sum<G>(s);
    if (gp.ql == 0) {
      sigma[i] = valid ? predict_density(s, kind, offset) : 0.f;
      if (feat_out) feat_out[i] = valid ? s : 0.f;
      if (valid_out) valid_out[i] = valid ? 1 : 0;
    }
  }
}

template <int G>
__global__ __launch_bounds__(256) void cp_density_bwd_kernel(CpArgs a, int64_t n, int run_len, const float* __restrict__ pts,
                                                             const uint8_t* __restrict__ valid,
                                                             const float* __restrict__ feat, int kind,
                                                             float offset, const float* __restrict__ d_sigma) {
  const GroupPos gp = group_pos<G>();
  const int64_t n_runs = (n + run_len - 1) / run_len;
  for (int64_t run = gp.first; run < n_runs; run += gp.stride) {
    const int64_t i0 = run * run_len;
    const int64_t i1 = i0 + run_len < n ? i0 + run_len : n;
    for (int q = gp.ql; q < a.n_quads; q += G) {
      CpRun r;
      cp_run_init(r);
      for (int64_t i = i0; i < i1; ++i) {
        if (!valid[i]) continue;
        const float g = __fmul_rn(d_sigma[i], predict_density_grad(feat[i], kind, offset));
        if (g == 0.f) continue;
        float pn[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) pn[k] = normalize_coord(pts[3 * i + k], a.lo[k], a.size[k]);
        const CpTaps taps = cp_taps(a, pn);
        float4 v[kPlanes];
        cp_values(a, taps, 4 * q, v);
        cp_run_add(a, taps, 4 * q, r, make_float4(g, g, g, g), v);
      }
#pragma unroll
      for (int k = 0; k < kPlanes; ++k) cp_run_flush(a, k, 4 * q, r);
    }
  }
}

template <int G>
__global__ __launch_bounds__(256) void cp_appearance_fwd_kernel(CpArgs a, int64_t n, const int32_t* __restrict__ n_dev,
                                                                const int32_t* __restrict__ idx,
                                                                const float* __restrict__ pts,
                                                                float* __restrict__ out) {
  const GroupPos gp = group_pos<G>();
  if (n_dev) n = min(n, (int64_t)*n_dev);  // device-side row count (n is the capacity)
  for (int64_t j = gp.first; j < n; j += gp.stride) {
    const int64_t i = idx ? (int64_t)idx[j] : j;
    float pn[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) pn[k] = normalize_coord(pts[3 * i + k], a.lo[k], a.size[k]);
    const CpTaps taps = cp_taps(a, pn);
    for (int q = gp.ql; q < a.n_quads; q += G) {
      float4 v[kPlanes];
      cp_values(a, taps, 4 * q, v);
      *reinterpret_cast<float4*>(out + j * a.comp + 4 * q) = cp_product(v);
    }
  }
}

template <int G>
__global__ __launch_bounds__(256) void cp_appearance_bwd_kernel(CpArgs a, int64_t n, const int32_t* __restrict__ n_dev,
                                                                int run_len, const int32_t* __restrict__ idx,
                                                                const float* __restrict__ pts,
                                                                const float* __restrict__ d_out) {
  const GroupPos gp = group_pos<G>();
  if (n_dev) n = min(n, (int64_t)*n_dev);  // device-side row count (n is the capacity)
  const int64_t n_runs = (n + run_len - 1) / run_len;
  for (int64_t run = gp.first; run < n_runs; run += gp.stride) {
    const int64_t j0 = run * run_len;
    const int64_t j1 = j0 + run_len < n ? j0 + run_len : n;
    for (int q = gp.ql; q < a.n_quads; q += G) {
      CpRun r;
      cp_run_init(r);
      for (int64_t j = j0; j < j1; ++j) {
        const int64_t i = idx ? (int64_t)idx[j] : j;
        float pn[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) pn[k] = normalize_coord(pts[3 * i + k], a.lo[k], a.size[k]);
        const CpTaps taps = cp_taps(a, pn);
        float4 v[kPlanes];
        cp_values(a, taps, 4 * q, v);
        cp_run_add(a, taps, 4 * q, r, ld4(d_out + j * a.comp + 4 * q), v);
      }
#pragma unroll
      for (int k = 0; k < kPlanes; ++k) cp_run_flush(a, k, 4 * q, r);
    }
  }
}

__global__ __launch_bounds__(256) void alpha_sample_kernel(AlphaArgs m, int64_t n, const float* __restrict__ pts,
                                                           float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    out[i] = alpha_value(m, p);
  }
}

// ------------------------------------------------------------------- threshold compaction
// Three stream-ordered passes, no host synchronisation: per-tile counts, one exclusive scan
// of the tile counts (which also writes the total), then an ordered scatter with wave
// ballots.  Tiles are kCompactTile consecutive entries; indices come out ascending, as
// boolean-mask indexing returns them.
constexpr int kCompactTile = 1024;

__global__ __launch_bounds__(256) void compact_count_kernel(int64_t n, const float* __restrict__ v, float thr,
                                                            int32_t* __restrict__ tile_count) {
  __shared__ int s_count;
  if (threadIdx.x == 0) s_count = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kCompactTile;
  int c = 0;
#pragma unroll
  for (int it = 0; it < kCompactTile / 256; ++it) {
    const int64_t i = base + it * 256 + threadIdx.x;
    const bool f = i < n && v[i] > thr;
    c += __popcll(__ballot(f));
  }
  if ((threadIdx.x & (SRF_WAVE - 1)) == 0) atomicAdd(&s_count, c);
  __syncthreads();
  if (threadIdx.x == 0) tile_count[blockIdx.x] = s_count;
}

// One block: tile counts -> exclusive tile offsets (in place); *count = total.
__global__ __launch_bounds__(1024) void compact_scan_kernel(int64_t n_tiles, int32_t* __restrict__ tile,
                                                            int32_t* __restrict__ count) {
  __shared__ int32_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (n_tiles + 1023) / 1024;
  const int64_t b0 = t * per;
  const int64_t b1 = b0 + per < n_tiles ? b0 + per : n_tiles;
  int32_t s = 0;
  for (int64_t b = b0; b < b1; ++b) s += tile[b];
  part[t] = s;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const int32_t o = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += o;
    __syncthreads();
  }
  int32_t run = t > 0 ? part[t - 1] : 0;
  for (int64_t b = b0; b < b1; ++b) {
    const int32_t c = tile[b];
    tile[b] = run;
    run += c;
  }
  if (t == 1023) *count = part[1023];
}

__global__ __launch_bounds__(256) void compact_scatter_kernel(int64_t n, const float* __restrict__ v, float thr,
                                                              const int32_t* __restrict__ tile_off,
                                                              int32_t* __restrict__ idx) {
  __shared__ int s_wave[4];
  const int w = threadIdx.x / SRF_WAVE, lane = threadIdx.x & (SRF_WAVE - 1);
  int out = tile_off[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * kCompactTile;
  for (int it = 0; it < kCompactTile / 256; ++it) {
    const int64_t i = base + it * 256 + threadIdx.x;
    const bool f = i < n && v[i] > thr;
    const uint64_t m = __ballot(f);
    const int within = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) s_wave[w] = __popcll(m);
    __syncthreads();
    int off = out;
    for (int j = 0; j < w; ++j) off += s_wave[j];
    if (f) idx[off + within] = (int32_t)i;
    out += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------ host side

int next_pow2(int q) {
  int g = 1;
  while (g < q) g <<= 1;
  return g;
}

void set_box(float lo[3], float hi[3], float size[3], const float* box) {
  for (int k = 0; k < 3; ++k) {
    lo[k] = box[k];
    hi[k] = box[3 + k];
    size[k] = box[3 + k] - box[k];  // float subtraction, as the reference's bounding_box_size
  }
}

int make_vm_args(const srf_vm_factors* f, const float* bbox, VmArgs* a, bool need_values) {
  if (!f || !bbox) return SRF_EINVAL;
  memset(a, 0, sizeof(*a));
  int q = 0;
  for (int p = 0; p < kPlanes; ++p) {
    if (f->n_comp[p] < 0 || f->plane_h[p] < 1 || f->plane_w[p] < 1 || f->line_len[p] < 1) return SRF_EINVAL;
    // components are gathered four at a time (float4 taps)
    if (f->n_comp[p] % 4 != 0) return SRF_EUNSUPPORTED;
    if (need_values && f->n_comp[p] > 0 && (!f->planes[p] || !f->lines[p])) return SRF_EINVAL;
    // taps are read as float4: the factor bases must be 16-byte aligned
    if ((reinterpret_cast<uintptr_t>(f->planes[p]) | reinterpret_cast<uintptr_t>(f->lines[p])) & 15)
      return SRF_EINVAL;
    a->plane[p] = f->planes[p];
    a->line[p] = f->lines[p];
    a->ph[p] = f->plane_h[p];
    a->pw[p] = f->plane_w[p];
    a->ll[p] = f->line_len[p];
    a->comp[p] = f->n_comp[p];
    a->qoff[p] = q;
    q += f->n_comp[p] / 4;
  }
  if (q < 1) return SRF_EINVAL;
  if (q > kMaxQuads) return SRF_EUNSUPPORTED;
  a->n_quads = q;
  a->total_comp = 4 * q;
  set_box(a->lo, a->hi, a->size, bbox);
  return SRF_OK;
}

int make_cp_args(const srf_cp_factors* f, const float* bbox, CpArgs* a) {
  if (!f || !bbox) return SRF_EINVAL;
  memset(a, 0, sizeof(*a));
  if (f->n_comp < 4) return SRF_EINVAL;
  if (f->n_comp % 4 != 0) return SRF_EUNSUPPORTED;
  if (f->n_comp / 4 > kMaxQuads) return SRF_EUNSUPPORTED;
  for (int k = 0; k < kPlanes; ++k) {
    if (f->line_len[k] < 1 || !f->lines[k]) return SRF_EINVAL;
    if (reinterpret_cast<uintptr_t>(f->lines[k]) & 15) return SRF_EINVAL;  // float4 taps
    a->line[k] = f->lines[k];
    a->ll[k] = f->line_len[k];
  }
  a->comp = f->n_comp;
  a->n_quads = f->n_comp / 4;
  set_box(a->lo, a->hi, a->size, bbox);
  return SRF_OK;
}

int set_cp_grads(CpArgs* a, float* const* d_lines) {
  if (!d_lines) return SRF_EINVAL;
  for (int k = 0; k < kPlanes; ++k) {
    if (!d_lines[k] || (reinterpret_cast<uintptr_t>(d_lines[k]) & 3)) return SRF_EINVAL;
    a->dline[k] = d_lines[k];
  }
  return SRF_OK;
}

int make_alpha_args(const srf_alpha_grid* g, AlphaArgs* m) {
  memset(m, 0, sizeof(*m));
  if (!g || !g->volume) return SRF_OK;
  if (g->depth < 1 || g->height < 1 || g->width < 1) return SRF_EINVAL;
  m->vol = g->volume;
  m->d = g->depth;
  m->h = g->height;
  m->w = g->width;
  set_box(m->lo, m->hi, m->size, g->bbox);
  return SRF_OK;
}

unsigned grid_for(int64_t n, int G) {
  const int64_t per_block = (int64_t)(256 / SRF_WAVE) * (SRF_WAVE / G);
  int64_t blocks = (n + per_block - 1) / per_block;
  // 256 CUs x 8 resident blocks of 256 threads, 4 rounds: beyond that the grid-stride loop takes
  // over (SRF_TUNE_GATHER_BLOCKS overrides the cap)
  const int64_t tuned = tune_value(SRF_TUNE_GATHER_BLOCKS);
  const int64_t cap = tuned > 0 ? tuned : 256 * 8 * 4;
  if (blocks > cap) blocks = cap;
  return (unsigned)(blocks < 1 ? 1 : blocks);
}

int run_length() {
  const int64_t v = tune_value(SRF_TUNE_GATHER_RUN);
  return v > 0 ? (int)v : kRun;
}

#define SRF_VM_DISPATCH(G_, KERNEL, ...)                                                   \
  switch (G_) {                                                                           \
    case 1: hipLaunchKernelGGL(KERNEL<1>, __VA_ARGS__); break;                            \
    case 2: hipLaunchKernelGGL(KERNEL<2>, __VA_ARGS__); break;                            \
    case 4: hipLaunchKernelGGL(KERNEL<4>, __VA_ARGS__); break;                            \
    case 8: hipLaunchKernelGGL(KERNEL<8>, __VA_ARGS__); break;                            \
    case 16: hipLaunchKernelGGL(KERNEL<16>, __VA_ARGS__); break;                          \
    case 32: hipLaunchKernelGGL(KERNEL<32>, __VA_ARGS__); break;                          \
    default: hipLaunchKernelGGL(KERNEL<64>, __VA_ARGS__); break;                          \
  }

#define SRF_VM_DISPATCH_KD(G_, KD, KERNEL, ...)                                            \
  switch (G_) {                                                                           \
    case 1: hipLaunchKernelGGL((KERNEL<1, KD>), __VA_ARGS__); break;                      \
    case 2: hipLaunchKernelGGL((KERNEL<2, KD>), __VA_ARGS__); break;                      \
    case 4: hipLaunchKernelGGL((KERNEL<4, KD>), __VA_ARGS__); break;                      \
    case 8: hipLaunchKernelGGL((KERNEL<8, KD>), __VA_ARGS__); break;                      \
    case 16: hipLaunchKernelGGL((KERNEL<16, KD>), __VA_ARGS__); break;                    \
    case 32: hipLaunchKernelGGL((KERNEL<32, KD>), __VA_ARGS__); break;                    \
    default: hipLaunchKernelGGL((KERNEL<64, KD>), __VA_ARGS__); break;                    \
  }

// Workspace of the cell-sorted backward for `rows` rows (capacity), and its carving.
size_t align256(size_t b) { return (b + 255) / 256 * 256; }

size_t sorted_ws_bytes(const VmArgs& a, int64_t rows) {
  size_t b = align256(kPlanes * sizeof(int32_t)) + align256((size_t)kPlanes * kMaxChunks * sizeof(int32_t));
  for (int p = 0; p < kPlanes; ++p) {
    const int64_t cells = (int64_t)a.ph[p] * a.pw[p];
    b += align256((size_t)(cells + 1) * sizeof(int32_t)) + align256((size_t)cells * sizeof(int32_t)) +
         align256((size_t)rows * sizeof(float4));
  }
  return b;
}

SortWs carve_sorted_ws(const VmArgs& a, int64_t rows, void* ws) {
  SortWs w;
  memset(&w, 0, sizeof(w));
  uint8_t* base = reinterpret_cast<uint8_t*>(ws);
  size_t off = 0;
  w.total = reinterpret_cast<int32_t*>(base + off);
  off += align256(kPlanes * sizeof(int32_t));
  w.partial = reinterpret_cast<int32_t*>(base + off);
  off += align256((size_t)kPlanes * kMaxChunks * sizeof(int32_t));
  // the histograms first and contiguous, so one fill zeroes them
  w.zero_begin = reinterpret_cast<int32_t*>(base + off);
  for (int p = 0; p < kPlanes; ++p) {
    w.cells[p] = a.ph[p] * a.pw[p];
    w.hist[p] = reinterpret_cast<int32_t*>(base + off);
    off += align256((size_t)(w.cells[p] + 1) * sizeof(int32_t));
  }
  w.zero_count = (int64_t)((base + off) - reinterpret_cast<uint8_t*>(w.zero_begin)) / (int64_t)sizeof(int32_t);
  for (int p = 0; p < kPlanes; ++p) {
    w.cursor[p] = reinterpret_cast<int32_t*>(base + off);
    off += align256((size_t)w.cells[p] * sizeof(int32_t));
  }
  for (int p = 0; p < kPlanes; ++p) {
    w.order[p] = reinterpret_cast<float4*>(base + off);
    off += align256((size_t)rows * sizeof(float4));
  }
  int64_t most = 1;
  for (int p = 0; p < kPlanes; ++p) most = w.cells[p] > most ? w.cells[p] : most;
  const int64_t per = (most + kMaxChunks - 1) / kMaxChunks;
  w.chunk_cells = (per + kScanThreads - 1) / kScanThreads * kScanThreads;  // whole tiles, <= kMaxChunks chunks
  return w;
}

// The dynamic LDS of the sorted kernels (the line copy) can exceed the 64 KB default: raise the
// limit of every instantiation once per device, before any call that could be under capture.
template <bool kDensity>
int raise_sorted_lds_limit() {
  constexpr int kMaxDevices = 64;
  static std::mutex mutex;
  static bool raised[kMaxDevices];
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return SRF_EHIP;
  std::lock_guard<std::mutex> lock(mutex);
  if (raised[dev]) return SRF_OK;
  const void* fns[] = {
      reinterpret_cast<const void*>(&sorted_bwd_kernel<1, kDensity>),
      reinterpret_cast<const void*>(&sorted_bwd_kernel<2, kDensity>),
      reinterpret_cast<const void*>(&sorted_bwd_kernel<4, kDensity>),
      reinterpret_cast<const void*>(&sorted_bwd_kernel<8, kDensity>),
      reinterpret_cast<const void*>(&sorted_bwd_kernel<16, kDensity>),
      reinterpret_cast<const void*>(&sorted_bwd_kernel<32, kDensity>),
      reinterpret_cast<const void*>(&sorted_bwd_kernel<64, kDensity>)};
  for (const void* f : fns)
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLineLds) != hipSuccess)
      return SRF_EHIP;
  raised[dev] = true;
  return SRF_OK;
}

template <bool kDensity>
int launch_sorted_bwd(const VmArgs& a, const BwdRows& r, void* ws, hipStream_t s) {
  if (const int st = raise_sorted_lds_limit<kDensity>(); st != SRF_OK) return st;
  const SortWs w = carve_sorted_ws(a, r.n, ws);
  const int64_t zb = (w.zero_count + 255) / 256;
  hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)(zb < 2048 ? zb : 2048)), dim3(256), 0, s, w.zero_begin,
                     w.zero_count);
  const int64_t rb = (r.n + 255) / 256;
  const unsigned row_blocks = (unsigned)(rb < 8192 ? rb : 8192);
  hipLaunchKernelGGL((cell_count_kernel<kDensity, false>), dim3(row_blocks), dim3(256), 0, s, a, r, w);
  int64_t most = 1;
  for (int p = 0; p < kPlanes; ++p) most = w.cells[p] > most ? w.cells[p] : most;
  const unsigned chunks = (unsigned)((most + w.chunk_cells - 1) / w.chunk_cells);
  hipLaunchKernelGGL(cell_chunk_sum_kernel, dim3(chunks, kPlanes), dim3(kScanThreads), 0, s, w);
  hipLaunchKernelGGL(cell_chunk_scan_kernel, dim3(kPlanes), dim3(kScanThreads), 0, s, w);
  hipLaunchKernelGGL(cell_chunk_apply_kernel, dim3(chunks, kPlanes), dim3(kScanThreads), 0, s, w);
  hipLaunchKernelGGL((cell_count_kernel<kDensity, true>), dim3(row_blocks), dim3(256), 0, s, a, r, w);
  for (int p = 0; p < kPlanes; ++p) {
    if (a.comp[p] == 0) continue;
    // component slices: halve the slice (whole quads) while its line copy exceeds kLineLdsSlice,
    // when the whole line fits in LDS but above kLineLdsSplit
    const int quads = a.comp[p] / 4;
    const size_t full_line = (size_t)a.ll[p] * a.comp[p] * sizeof(float);
    int ns = 1;
    if (full_line <= (size_t)kMaxLineLds && full_line > (size_t)kLineLdsSplit)
      while ((size_t)a.ll[p] * (a.comp[p] / ns) * sizeof(float) > (size_t)kLineLdsSlice && quads % (2 * ns) == 0)
        ns *= 2;
    const int G = next_pow2(quads / ns);
    const int groups = kSortedThreads / G;
    // about 1024 blocks over the capacity (fewer run when the device count is lower), unless the
    // gather run knob sets the walk length
    int64_t run = (r.n + 1024LL * groups - 1) / (1024LL * groups);
    run = run < 16 ? 16 : (run > 256 ? 256 : run);
    if (const int64_t v = tune_value(SRF_TUNE_GATHER_RUN)) run = v;
    const int64_t blocks = (r.n + (int64_t)groups * run - 1) / ((int64_t)groups * run);
    const size_t line_bytes = (size_t)a.ll[p] * (a.comp[p] / ns) * sizeof(float);
    const int line_lds = line_bytes <= (size_t)kMaxLineLds ? 1 : 0;
    SRF_VM_DISPATCH_KD(G, kDensity, sorted_bwd_kernel, dim3((unsigned)(blocks > 0 ? blocks : 1), ns),
                       dim3(kSortedThreads), line_lds ? line_bytes : 0, s, a, r, w, p, (int)run, line_lds);
  }
  return launch_status();
}


}  // namespace
}  // namespace srf

using namespace srf;

extern "C" int srf_vm_density_fwd(const srf_vm_factors* factors, const float* bbox,
                                  const srf_alpha_grid* alpha, int64_t n_points, const float* pts,
                                  int predictor, float offset, float* sigma, float* feat,
                                  uint8_t* valid, void* stream) {
  if (n_points < 0 || (predictor != SRF_DENSITY_RELU && predictor != SRF_DENSITY_SOFTPLUS)) return SRF_EINVAL;
  if (n_points == 0) return SRF_OK;
  VmArgs a;
  int st = make_vm_args(factors, bbox, &a, true);
  if (st != SRF_OK) return st;
  AlphaArgs m;
  if ((st = make_alpha_args(alpha, &m)) != SRF_OK) return st;
  if (!pts || !sigma) return SRF_EINVAL;
  const int G = next_pow2(a.n_quads);
  SRF_VM_DISPATCH(G, density_fwd_kernel, dim3(grid_for(n_points, G)), dim3(256), 0, as_stream(stream), a, m,
                  n_points, pts, predictor, offset, sigma, feat, valid);
  return launch_status();
}

extern "C" size_t srf_vm_bwd_workspace_bytes(const srf_vm_factors* factors, int64_t n_rows) {
  VmArgs a;
  if (n_rows < 0 || !factors) return 0;
  const float box[6] = {-1.f, -1.f, -1.f, 1.f, 1.f, 1.f};
  if (make_vm_args(factors, box, &a, false) != SRF_OK) return 0;
  return sorted_ws_bytes(a, n_rows);
}

extern "C" int srf_vm_density_bwd(const srf_vm_factors* factors, float* const* d_planes,
                                  float* const* d_lines, const float* bbox, int64_t n_points,
                                  const float* pts, const uint8_t* valid, const float* feat,
                                  int predictor, float offset, const float* d_sigma, void* workspace,
                                  void* stream) {
  if (n_points < 0 || (predictor != SRF_DENSITY_RELU && predictor != SRF_DENSITY_SOFTPLUS)) return SRF_EINVAL;
  if (n_points == 0) return SRF_OK;
  VmArgs a;
  int st = make_vm_args(factors, bbox, &a, true);
  if (st != SRF_OK) return st;
  if (!d_planes || !d_lines || !pts || !valid || !feat || !d_sigma) return SRF_EINVAL;
  for (int p = 0; p < kPlanes; ++p) {
    if (a.comp[p] > 0 && (!d_planes[p] || !d_lines[p])) return SRF_EINVAL;
    a.dplane[p] = d_planes[p];
    a.dline[p] = d_lines[p];
  }
  if (workspace) {
    BwdRows r;
    memset(&r, 0, sizeof(r));
    r.n = n_points;
    r.pts = pts;
    r.valid = valid;
    r.feat = feat;
    r.kind = predictor;
    r.offset = offset;
    r.d_sigma = d_sigma;
    return launch_sorted_bwd<true>(a, r, workspace, as_stream(stream));
  }
  const int G = next_pow2(a.n_quads);
  const int run = run_length();
  SRF_VM_DISPATCH(G, density_bwd_kernel, dim3(grid_for((n_points + run - 1) / run, G)), dim3(256), 0,
                  as_stream(stream), a, n_points, run, pts, valid, feat, predictor, offset, d_sigma);
  return launch_status();
}

extern "C" int srf_vm_appearance_fwd(const srf_vm_factors* factors, const float* bbox, int64_t n_rows,
                                     const int32_t* row_count, const int32_t* idx, const float* pts,
                                     float* features, void* stream) {
  if (n_rows < 0) return SRF_EINVAL;
  if (n_rows == 0) return SRF_OK;
  VmArgs a;
  int st = make_vm_args(factors, bbox, &a, true);
  if (st != SRF_OK) return st;
  if (!pts || !features) return SRF_EINVAL;
  const int G = next_pow2(a.n_quads);
  SRF_VM_DISPATCH(G, appearance_fwd_kernel, dim3(grid_for(n_rows, G)), dim3(256), 0, as_stream(stream), a,
                  n_rows, row_count, idx, pts, features);
  return launch_status();
}

extern "C" int srf_vm_appearance_bwd(const srf_vm_factors* factors, float* const* d_planes,
                                     float* const* d_lines, const float* bbox, int64_t n_rows,
                                     const int32_t* row_count, const int32_t* idx, const float* pts,
                                     const float* d_features, void* workspace, void* stream) {
  if (n_rows < 0) return SRF_EINVAL;
  if (n_rows == 0) return SRF_OK;
  VmArgs a;
  int st = make_vm_args(factors, bbox, &a, true);
  if (st != SRF_OK) return st;
  if (!d_planes || !d_lines || !pts || !d_features) return SRF_EINVAL;
  for (int p = 0; p < kPlanes; ++p) {
    if (a.comp[p] > 0 && (!d_planes[p] || !d_lines[p])) return SRF_EINVAL;
    a.dplane[p] = d_planes[p];
    a.dline[p] = d_lines[p];
  }
  if (workspace) {
    BwdRows r;
    memset(&r, 0, sizeof(r));
    r.n = n_rows;
    r.n_dev = row_count;
    r.idx = idx;
    r.pts = pts;
    r.d_out = d_features;
    return launch_sorted_bwd<false>(a, r, workspace, as_stream(stream));
  }
  const int G = next_pow2(a.n_quads);
  const int run = run_length();
  SRF_VM_DISPATCH(G, appearance_bwd_kernel, dim3(grid_for((n_rows + run - 1) / run, G)), dim3(256), 0,
                  as_stream(stream), a, n_rows, row_count, run, idx, pts, d_features);
  return launch_status();
}

extern "C" int srf_cp_density_fwd(const srf_cp_factors* factors, const float* bbox,
                                  const srf_alpha_grid* alpha, int64_t n_points, const float* pts,
                                  int predictor, float offset, float* sigma, float* feat,
                                  uint8_t* valid, void* stream) {
  if (n_points < 0 || (predictor != SRF_DENSITY_RELU && predictor != SRF_DENSITY_SOFTPLUS)) return SRF_EINVAL;
  if (n_points == 0) return SRF_OK;
  CpArgs a;
  int st = make_cp_args(factors, bbox, &a);
  if (st != SRF_OK) return st;
  AlphaArgs m;
  if ((st = make_alpha_args(alpha, &m)) != SRF_OK) return st;
  if (!pts || !sigma) return SRF_EINVAL;
  const int G = next_pow2(a.n_quads);
  SRF_VM_DISPATCH(G, cp_density_fwd_kernel, dim3(grid_for(n_points, G)), dim3(256), 0, as_stream(stream), a, m,
                  n_points, pts, predictor, offset, sigma, feat, valid);
  return launch_status();
}

extern "C" int srf_cp_density_bwd(const srf_cp_factors* factors, float* const* d_lines, const float* bbox,
                                  int64_t n_points, const float* pts, const uint8_t* valid, const float* feat,
                                  int predictor, float offset, const float* d_sigma, void* stream) {
  if (n_points < 0 || (predictor != SRF_DENSITY_RELU && predictor != SRF_DENSITY_SOFTPLUS)) return SRF_EINVAL;
  if (n_points == 0) return SRF_OK;
  CpArgs a;
  int st = make_cp_args(factors, bbox, &a);
  if (st != SRF_OK) return st;
  if ((st = set_cp_grads(&a, d_lines)) != SRF_OK) return st;
  if (!pts || !valid || !feat || !d_sigma) return SRF_EINVAL;
  const int G = next_pow2(a.n_quads);
  const int run = run_length();
  SRF_VM_DISPATCH(G, cp_density_bwd_kernel, dim3(grid_for((n_points + run - 1) / run, G)), dim3(256), 0,
                  as_stream(stream), a, n_points, run, pts, valid, feat, predictor, offset, d_sigma);
  return launch_status();
}

extern "C" int srf_cp_appearance_fwd(const srf_cp_factors* factors, const float* bbox, int64_t n_rows,
                                     const int32_t* row_count, const int32_t* idx, const float* pts,
                                     float* features, void* stream) {
  if (n_rows < 0) return SRF_EINVAL;
  if (n_rows == 0) return SRF_OK;
  CpArgs a;
  int st = make_cp_args(factors, bbox, &a);
  if (st != SRF_OK) return st;
  if (!pts || !features || (reinterpret_cast<uintptr_t>(features) & 15)) return SRF_EINVAL;
  const int G = next_pow2(a.n_quads);
  SRF_VM_DISPATCH(G, cp_appearance_fwd_kernel, dim3(grid_for(n_rows, G)), dim3(256), 0, as_stream(stream), a,
                  n_rows, row_count, idx, pts, features);
  return launch_status();
}

extern "C" int srf_cp_appearance_bwd(const srf_cp_factors* factors, float* const* d_lines, const float* bbox,
                                     int64_t n_rows, const int32_t* row_count, const int32_t* idx,
                                     const float* pts, const float* d_features, void* stream) {
  if (n_rows < 0) return SRF_EINVAL;
  if (n_rows == 0) return SRF_OK;
  CpArgs a;
  int st = make_cp_args(factors, bbox, &a);
  if (st != SRF_OK) return st;
  if ((st = set_cp_grads(&a, d_lines)) != SRF_OK) return st;
  if (!pts || !d_features || (reinterpret_cast<uintptr_t>(d_features) & 15)) return SRF_EINVAL;
  const int G = next_pow2(a.n_quads);
  const int run = run_length();
  SRF_VM_DISPATCH(G, cp_appearance_bwd_kernel, dim3(grid_for((n_rows + run - 1) / run, G)), dim3(256), 0,
                  as_stream(stream), a, n_rows, row_count, run, idx, pts, d_features);
  return launch_status();
}

extern "C" int srf_alpha_mask_sample(const srf_alpha_grid* alpha, int64_t n_points, const float* pts,
                                     float* out, void* stream) {
  if (n_points < 0 || !alpha || !alpha->volume) return SRF_EINVAL;
  if (n_points == 0) return SRF_OK;
  AlphaArgs m;
  int st = make_alpha_args(alpha, &m);
  if (st != SRF_OK) return st;
  if (!pts || !out) return SRF_EINVAL;
  int64_t blocks = (n_points + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(alpha_sample_kernel, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), m, n_points,
                     pts, out);
  return launch_status();
}

extern "C" size_t srf_compact_workspace_bytes(int64_t n) {
  if (n < 0) return 0;
  const int64_t tiles = (n + kCompactTile - 1) / kCompactTile;
  return (size_t)((tiles * 4 + 255) / 256 * 256 + 256);
}

namespace srf {
namespace {
// The three compaction passes on tile counts at `tile` (tiles = ceil(n / kCompactTile) entries).
void launch_compact(int64_t n, const float* values, float threshold, int32_t* idx, int32_t* count, int32_t* tile,
                    hipStream_t s) {
  const int64_t tiles = (n + kCompactTile - 1) / kCompactTile;
  if (tiles > 0) {
    hipLaunchKernelGGL(compact_count_kernel, dim3((unsigned)tiles), dim3(256), 0, s, n, values, threshold, tile);
  }
  // the scan also writes *count (0 for an empty input), so no memset is needed
  hipLaunchKernelGGL(compact_scan_kernel, dim3(1), dim3(1024), 0, s, tiles, tile, count);
  if (tiles > 0) {
    hipLaunchKernelGGL(compact_scatter_kernel, dim3((unsigned)tiles), dim3(256), 0, s, n, values, threshold, tile,
                       idx);
  }
}
}  // namespace
}  // namespace srf

extern "C" int srf_threshold_compact(int64_t n, const float* values, float threshold, int32_t* idx,
                                     int32_t* count, void* workspace, void* stream) {
  if (n < 0 || n > 0x7fffffffLL || !count) return SRF_EINVAL;
  if (n > 0 && (!values || !idx || !workspace)) return SRF_EINVAL;
  launch_compact(n, values, threshold, idx, count, reinterpret_cast<int32_t*>(workspace), as_stream(stream));
  return launch_status();
}

// ------------------------------------------------- allocation-free compaction (fixed scratch)
// The tile counts of srf_threshold_compact_noalloc.  Each GPU holds its own copy of this array
// at its own address; host code never passes the array's name to a kernel (on the host it
// names a stand-in, not device memory) but the per-device address from hipGetSymbolAddress.
// Every count the passes read is written first by compact_count_kernel in the same call, so
// nothing needs clearing between calls.
constexpr int64_t kNoAllocTiles = SRF_COMPACT_NOALLOC_MAX / srf::kCompactTile;  // 262144 ints = 1 MiB
static __device__ __align__(256) int32_t scratch[kNoAllocTiles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
int32_t* g_scratch_ptr[kMaxDevices];  // device address of `scratch` on each GPU, looked up once

// The scratch address on the current device (looked up on the first call there).
int scratch_on_current_device(hipStream_t stream, int32_t** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return SRF_EHIP;
  if (dev < 0 || dev >= kMaxDevices) return SRF_EINVAL;
  if (stream != nullptr) {
    hipDevice_t sdev = -1;
    if (hipStreamGetDevice(stream, &sdev) != hipSuccess) return SRF_EHIP;
    if (sdev != dev) return SRF_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return SRF_EHIP;
    g_scratch_ptr[dev] = static_cast<int32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return SRF_OK;
}
}  // namespace

extern "C" int srf_threshold_compact_noalloc(int64_t n, const float* values, float threshold, int32_t* idx,
                                             int32_t* count, void* stream) {
  if (n < 0 || !count) return SRF_EINVAL;
  if (n > SRF_COMPACT_NOALLOC_MAX) return SRF_ETOOLARGE;
  if (n > 0 && (!values || !idx)) return SRF_EINVAL;
  int32_t* tile = nullptr;
  const int st = scratch_on_current_device(as_stream(stream), &tile);
  if (st != SRF_OK) return st;
  launch_compact(n, values, threshold, idx, count, tile, as_stream(stream));
  return launch_status();
}

