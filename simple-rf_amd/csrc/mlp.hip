// C-ABI of the fused PE + MLP kernels: planning, weight packing and dispatch.
// Kernel design notes are at the top of mlp_impl.h.
#include <string.h>

#include "mlp_impl.h"

namespace srf {
namespace mlp {

// ----------------------------------------------------------------------- packing
struct PackArgs {
  Plan p;
  const float* w[kParams];
  const float* b[kParams];
  uint8_t* out;
};

__device__ inline int ref_ld(const Plan& p, int l) {
  switch (l) {
    case 0: return p.trunk_pe;
    case 5: return p.trunk_pe + 256;
    case 9: return 256 + p.extra;
    case 10: return 128;
    default: return 256;
  }
}

// Reference tensor index and row for kernel-layer output row o.
__device__ inline void ref_row(const Plan& p, int l, int o, int* tensor, int* row) {
  if (l < 8) {
    *tensor = l;
    *row = o;
  } else if (l == 8) {
    if (p.view_dep) {
      *tensor = o < 256 ? 8 : 9;
      *row = o < 256 ? o : o - 256;
    } else {
      *tensor = 9;
      *row = o;
    }
  } else {
    *tensor = l == 9 ? 10 : 11;
    *row = o;
  }
}

// Reference column of kernel-layer input position kp (-1 for padding).
__device__ inline int ref_col(const Plan& p, int l, int kp) {
  const int reg = 16 * p.reg_tiles[l];
  if (kp < reg) return l == 5 ? p.trunk_pe + kp : kp;
  const int e = kp - reg;
  if (e >= p.comp_feats[l]) return -1;
  return l == 9 ? 256 + e : e;
}

__device__ inline float ref_weight(const PackArgs& a, int l, int o, int col) {
  if (o >= a.p.out_rows[l] || col < 0) return 0.f;
  int t, r;
  ref_row(a.p, l, o, &t, &r);
  return a.w[t][(size_t)r * ref_ld(a.p, l) + col];
}

// logical reduction index of (k-step s, lane group g, slot i)
template <bool F32>
__device__ inline int slot_k(int s, int g, int i) {
  if (F32) return 16 * (s >> 2) + 4 * g + (s & 3);
  return 32 * s + (i < 4 ? 4 * g + i : 16 + 4 * g + i - 4);
}

// Element idx of the forward (bwd = 0) or transposed (bwd = 1) operand image of layer l.
__device__ void pack_element(const PackArgs& a, int l, int bwd, long idx) {
  const Plan& p = a.p;
  const bool f32 = p.prec == SRF_PREC_F32;
  const bool split = p.prec == SRF_PREC_BF16_SPLIT;
  const int kpl = f32 ? 1 : 8;
  int rows_t, k_tiles;
  if (!bwd) {
    rows_t = p.out_tiles[l];
    k_tiles = p.reg_tiles[l] + p.comp_tiles[l];
  } else {
    rows_t = p.reg_tiles[l];
    k_tiles = p.bwd_tiles[l];
  }
  const int ks = f32 ? 4 * k_tiles : k_tiles / 2;
  const long total = (long)rows_t * ks * 64 * kpl;
  if (idx >= total) return;
  // an image holds at most 2^31 elements (the launch's block count is an int): 32-bit index
  // arithmetic, shifts for the power-of-two factors and one 32-bit division by ks
  const unsigned ui = (unsigned)idx;
  const unsigned slot = f32 ? 0u : (ui & 7u);
  const unsigned rest = f32 ? ui : (ui >> 3);  // idx / kpl
  const int lane = (int)(rest & 63u);
  const unsigned fs = rest >> 6;               // idx / (kpl * 64)
  const int T = (int)(fs / (unsigned)ks);
  const int s = (int)(fs - (unsigned)T * (unsigned)ks);
  const int g = lane >> 4, jj = lane & 15;
  const int kp = f32 ? slot_k<true>(s, g, slot) : slot_k<false>(s, g, slot);
  float v;
  if (!bwd) {
    v = ref_weight(a, l, 16 * T + jj, ref_col(p, l, kp));
  } else {
    // A^T: row = input feature (register segment), k = forward output row
    v = ref_weight(a, l, kp, ref_col(p, l, 16 * T + jj));
  }
  uint8_t* dst = a.out + (bwd ? p.bwd_off[l] : p.fwd_off[l]);
  if (split) {
    // BF16 planes of v, each 1 KB fragment (512 values of one (tile, k-step)) followed by its
    // other planes: plane pl of fragment f at byte (f planes + pl) KB
    const int np = bwd ? 2 : 3;
    const long f = idx / 512, within = idx % 512;
    uint16_t* d16 = reinterpret_cast<uint16_t*>(dst);
    float r = v;
    for (int pl = 0; pl < np; ++pl) {
      const uint32_t h = f32_to_bf16_bits(r);
      d16[(f * np + pl) * 512 + within] = (uint16_t)h;
      r = r - bf16_bits_to_f32(h);
    }
  } else if (f32) {
    reinterpret_cast<float*>(dst)[idx] = v;
  } else if (p.prec == SRF_PREC_BF16) {
    reinterpret_cast<uint16_t*>(dst)[idx] = (uint16_t)f32_to_bf16_bits(v);
  } else {
    reinterpret_cast<uint16_t*>(dst)[idx] = (uint16_t)f32_to_f16_bits(v);
  }
}

__device__ void pack_bias_element(const PackArgs& a, int l, int o) {
  if (o >= 16 * a.p.out_tiles[l]) return;
  float v = 0.f;
  if (o < a.p.out_rows[l]) {
    int t, r;
    ref_row(a.p, l, o, &t, &r);
    v = a.b[t][r];
  }
  reinterpret_cast<float*>(a.out + a.p.bias_off[l])[o] = v;
}

// Every operand image and bias row of one model in a single launch: job j covers the blocks
// [first_block[j], first_block[j + 1]) (kind 0: forward image, 1: transposed image, 2: bias).
constexpr int kPackJobs = 3 * kLayers;
struct PackJobs {
  int n_jobs;
  int layer[kPackJobs], kind[kPackJobs];
  unsigned first_block[kPackJobs + 1];
};

__global__ __launch_bounds__(256) void pack_all_kernel(PackArgs a, PackJobs jobs) {
  int j = 0;
  while (j + 1 < jobs.n_jobs && jobs.first_block[j + 1] <= blockIdx.x) ++j;  // block-uniform
  const long idx = (long)(blockIdx.x - jobs.first_block[j]) * blockDim.x + threadIdx.x;
  if (jobs.kind[j] == 2) pack_bias_element(a, jobs.layer[j], (int)idx);
  else pack_element(a, jobs.layer[j], jobs.kind[j], idx);
}


int fwd_launch(int prec, const FwdArgs& a, bool vd, bool tr, hipStream_t s) {
  switch (prec) {
    case SRF_PREC_F32: return fwd_f32(a, vd, tr, s);
    case SRF_PREC_BF16: return fwd_bf16(a, vd, tr, s);
    case SRF_PREC_F16: return fwd_f16(a, vd, tr, s);  // inference only (tr -> SRF_EUNSUPPORTED)
    case SRF_PREC_BF16_SPLIT: return fwd_bf16s(a, vd, tr, s);
    default: return SRF_EUNSUPPORTED;
  }
}

int bwd_launch(int prec, const BwdArgs& a, bool vd, hipStream_t s) {
  switch (prec) {
    case SRF_PREC_F32: return bwd_f32(a, vd, s);
    case SRF_PREC_BF16: return bwd_bf16(a, vd, s);
    case SRF_PREC_BF16_SPLIT: return bwd_bf16s(a, vd, s);
    default: return SRF_EUNSUPPORTED;
  }
}

int dw_launch(int prec, const DwArgs& d, hipStream_t s) {
  switch (prec) {
    case SRF_PREC_F32: return dw_f32(d, s);
    case SRF_PREC_BF16: return dw_bf16(d, s);
    case SRF_PREC_BF16_SPLIT: return dw_bf16s(d, s);
    default: return SRF_EUNSUPPORTED;
  }
}

}  // namespace mlp
}  // namespace srf

using namespace srf;
using namespace srf::mlp;

extern "C" size_t srf_mlp_packed_bytes(const srf_mlp_desc* desc) {
  Plan p;
  return make_plan(desc, &p) ? p.packed_bytes : 0;
}

extern "C" size_t srf_mlp_workspace_bytes(const srf_mlp_desc* desc, int n_points) {
  Plan p;
  if (!make_plan(desc, &p) || n_points < 0) return 0;
  return make_ws(p, n_points).bytes;
}

extern "C" int srf_mlp_pack(const srf_mlp_desc* desc, const float* const* weights,
                            const float* const* biases, void* packed, void* stream) {
  Plan p;
  if (!make_plan(desc, &p) || !weights || !biases || !packed) return SRF_EINVAL;
  PackArgs a;
  a.p = p;
  a.out = reinterpret_cast<uint8_t*>(packed);
  for (int i = 0; i < kParams; ++i) {
    a.w[i] = weights[i];
    a.b[i] = biases[i];
  }
  const bool need[kParams] = {true, true, true, true, true, true, true, true,
                              (bool)p.view_dep, true, (bool)p.view_dep, (bool)p.view_dep};
  for (int i = 0; i < kParams; ++i)
    if (need[i] && (!a.w[i] || !a.b[i])) return SRF_EINVAL;
  const int kpl = p.prec == SRF_PREC_F32 ? 1 : 8;
  PackJobs jobs;
  memset(&jobs, 0, sizeof(jobs));
  unsigned blocks = 0;
  auto add = [&](int l, int kind, long total) {
    jobs.layer[jobs.n_jobs] = l;
    jobs.kind[jobs.n_jobs] = kind;
    jobs.first_block[jobs.n_jobs] = blocks;
    blocks += (unsigned)((total + 255) / 256);
    ++jobs.n_jobs;
  };
  for (int l = 0; l < kLayers; ++l) {
    if (!p.present[l]) continue;
    for (int bwd = 0; bwd < 2; ++bwd) {
      if (bwd && l == 0) continue;
      const int rows_t = bwd ? p.reg_tiles[l] : p.out_tiles[l];
      const int k_tiles = bwd ? p.bwd_tiles[l] : p.reg_tiles[l] + p.comp_tiles[l];
      const int ks = p.prec == SRF_PREC_F32 ? 4 * k_tiles : k_tiles / 2;
      add(l, bwd, (long)rows_t * ks * 64 * kpl);
    }
    add(l, 2, 16L * p.out_tiles[l]);
  }
  jobs.first_block[jobs.n_jobs] = blocks;
  hipLaunchKernelGGL(pack_all_kernel, dim3(blocks), dim3(256), 0, as_stream(stream), a, jobs);
  return launch_status();
}

static int mlp_forward_impl(const srf_mlp_desc* desc, const void* packed, int n_points, const float* pts,
                            const float* rays_o, const float* rays_d, const float* z, int n_samples,
                            const float* dirs, const float* noise, float* sigma, float* rgb, void* workspace,
                            void* stream) {
  Plan p;
  if (!make_plan(desc, &p) || n_points < 0 || !packed) return SRF_EINVAL;
  if (n_points == 0) return SRF_OK;
  if (!sigma || !rgb || (p.view_dep && !dirs)) return SRF_EINVAL;
  if (z ? (!rays_o || !rays_d || n_samples <= 0 || n_points % n_samples != 0) : !pts) return SRF_EINVAL;
  FwdArgs a;
  a.p = p;
  a.w = make_ws(p, n_points);
  a.packed = reinterpret_cast<const uint8_t*>(packed);
  a.n = n_points;
  a.pts = pts;
  a.dirs = dirs;
  a.noise = noise;
  a.sigma = sigma;
  a.rgb = rgb;
  a.ws = reinterpret_cast<uint8_t*>(workspace);
  a.rays_o = rays_o;
  a.rays_d = rays_d;
  a.z = z;
  a.n_samples = n_samples;
  if (workspace && !ws_addressable(p, a.w)) return SRF_EINVAL;
  return fwd_launch(p.prec, a, p.view_dep, workspace != nullptr, as_stream(stream));
}

extern "C" int srf_mlp_forward(const srf_mlp_desc* desc, const void* packed, int n_points,
                               const float* pts, const float* dirs, const float* noise, float* sigma,
                               float* rgb, void* workspace, void* stream) {
  return mlp_forward_impl(desc, packed, n_points, pts, nullptr, nullptr, nullptr, 0, dirs, noise, sigma, rgb,
                          workspace, stream);
}

extern "C" int srf_mlp_forward_rays(const srf_mlp_desc* desc, const void* packed, int n_rays, int n_samples,
                                    const float* rays_o, const float* rays_d, const float* z_vals,
                                    const float* view_dirs, const float* noise, float* sigma, float* rgb,
                                    void* workspace, void* stream) {
  if (n_rays < 0 || n_samples <= 0 || (long long)n_rays * n_samples > 0x7fffffffLL) return SRF_EINVAL;
  if (!z_vals && n_rays > 0) return SRF_EINVAL;
  return mlp_forward_impl(desc, packed, n_rays * n_samples, nullptr, rays_o, rays_d, z_vals, n_samples, view_dirs,
                          noise, sigma, rgb, workspace, stream);
}

static int mlp_backward_impl(const srf_mlp_desc* desc, const void* packed, int n_points,
                             const float* sigma, const float* rgb, const float* d_sigma,
                             const float* d_rgb, void* workspace, float* const* d_weights,
                             float* const* d_biases, void* stream, bool chain, bool weights) {
  Plan p;
  if (!make_plan(desc, &p) || n_points < 0) return SRF_EINVAL;
  if (n_points == 0) return SRF_OK;
  if (!workspace) return SRF_EINVAL;
  if (chain && (!packed || !sigma || !rgb || !d_sigma || !d_rgb)) return SRF_EINVAL;
  const bool need[kParams] = {true, true, true, true, true, true, true, true,
                              (bool)p.view_dep, true, (bool)p.view_dep, (bool)p.view_dep};
  if (weights) {
    if (!d_weights || !d_biases) return SRF_EINVAL;
    for (int i = 0; i < kParams; ++i)
      if (need[i] && (!d_weights[i] || !d_biases[i])) return SRF_EINVAL;
  }
  hipStream_t s = as_stream(stream);
  const Ws w = make_ws(p, n_points);
  if (!ws_addressable(p, w)) return SRF_EINVAL;
  uint8_t* ws = reinterpret_cast<uint8_t*>(workspace);
  if (chain) {
    BwdArgs b;
    b.p = p;
    b.w = w;
    b.packed = reinterpret_cast<const uint8_t*>(packed);
    b.n = n_points;
    b.sigma = sigma;
    b.rgb = rgb;
    b.d_sigma = d_sigma;
    b.d_rgb = d_rgb;
    b.ws = ws;
    const int st = bwd_launch(p.prec, b, p.view_dep, s);
    if (st != SRF_OK || !weights) return st;
  }
  // grouped weight-gradient GEMM
  const size_t rb = (size_t)w.ldP * p.elem;  // bytes per stored row
  DwArgs d;
  memset(&d, 0, sizeof(d));
  d.ldP = w.ldP;
  d.cols = w.cols;
  d.paired = 1;

  int nj = 0;
  auto job = [&](size_t dz_off, int dz_row0, int dz_rows, size_t x_off, int x_rows, int M, int N, float* dw,
                 int ld, float* db) {
    DwJob& jb = d.jobs[nj++];
    jb.dz = ws + dz_off + (size_t)dz_row0 * rb;
    jb.dz_rows = dz_rows;
    jb.x = ws + x_off;
    jb.x_rows = x_rows;
    jb.M = M;
    jb.N = N;
    jb.dw = dw;
    jb.ld_dw = ld;
    jb.db = db;
    jb.mblocks = (M + kDwTile - 1) / kDwTile;
    jb.nblocks = (N + kDwTile - 1) / kDwTile;
  };
  const int D = p.trunk_pe, pe_rows = p.comp_tiles[0] * 16;
  job(w.dz[0], 0, 256, w.x0, pe_rows, 256, D, d_weights[0], D, d_biases[0]);
  for (int l = 1; l < 8; ++l) {
    if (l == 5) {
      job(w.dz[5], 0, 256, w.h[4], 256, 256, 256, d_weights[5] + D, D + 256, d_biases[5]);
      job(w.dz[5], 0, 256, w.x0, pe_rows, 256, D, d_weights[5], D + 256, nullptr);
    } else {
      job(w.dz[l], 0, 256, w.h[l - 1], 256, 256, 256, d_weights[l], 256, d_biases[l]);
    }
  }
  if (p.view_dep) {
    job(w.dz[8], 0, 256, w.h[7], 256, 256, 256, d_weights[8], 256, d_biases[8]);
    job(w.dz[8], 256, 16, w.h[7], 256, 1, 256, d_weights[9], 256, d_biases[9]);
    job(w.dz[9], 0, 128, w.f, 256, 128, 256, d_weights[10], 256 + p.extra, d_biases[10]);
    job(w.dz[9], 0, 128, w.ex, p.comp_tiles[9] * 16, 128, p.extra, d_weights[10] + 256, 256 + p.extra,
        nullptr);
    job(w.dz[10], 0, 16, w.v, 128, 3, 128, d_weights[11], 128, d_biases[11]);
  } else {
    job(w.dz[8], 0, 16, w.h[7], 256, 4, 256, d_weights[9], 256, d_biases[9]);
  }
  d.n_jobs = nj;
  // points per work item: whole rounds of resident blocks (dw_default_chunk), whole multiples of
  // 256 points (so every item holds an even number of k-units), between 2048 and 262144 points
  long long target = dw_default_chunk(w.cols, nj, device_compute_units());
  if (const int64_t v = tune_value(SRF_TUNE_DW_CHUNK)) target = v;  // autotuned schedule
  d.chunk = (int)(target < 2048 ? 2048 : (target > 262144 ? 262144 : target));
  d.n_chunks = (w.cols + d.chunk - 1) / d.chunk;
  int work = 0;
  for (int i = 0; i < nj; ++i) {
    d.jobs[i].work_begin = work;
    d.jobs[i].work_count = d.jobs[i].mblocks * d.jobs[i].nblocks * d.n_chunks;
    work += d.jobs[i].work_count;
  }
  d.total_work = work;
  return dw_launch(p.prec, d, s);
}

extern "C" int srf_mlp_backward(const srf_mlp_desc* desc, const void* packed, int n_points,
                                const float* sigma, const float* rgb, const float* d_sigma,
                                const float* d_rgb, void* workspace, float* const* d_weights,
                                float* const* d_biases, void* stream) {
  return mlp_backward_impl(desc, packed, n_points, sigma, rgb, d_sigma, d_rgb, workspace, d_weights,
                           d_biases, stream, true, true);
}

extern "C" int srf_mlp_backward_data(const srf_mlp_desc* desc, const void* packed, int n_points,
                                     const float* sigma, const float* rgb, const float* d_sigma,
                                     const float* d_rgb, void* workspace, void* stream) {
  return mlp_backward_impl(desc, packed, n_points, sigma, rgb, d_sigma, d_rgb, workspace, nullptr,
                           nullptr, stream, true, false);
}

extern "C" int srf_mlp_weight_grads(const srf_mlp_desc* desc, int n_points, void* workspace,
                                    float* const* d_weights, float* const* d_biases, void* stream) {
  return mlp_backward_impl(desc, nullptr, n_points, nullptr, nullptr, nullptr, nullptr, workspace,
                           d_weights, d_biases, stream, false, true);
}
