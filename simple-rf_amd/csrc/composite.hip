// Alpha compositing (forward and backward) for NeRF and TensoRF rays on gfx950.
//
// Reference semantics: SimpleNeRF.volume_rendering + convert_depth_from_ndc
// (src/models/SimpleNeRF17.py:486-558) and LowRankTensor.get_volume_rendering_weights +
// volume_render (src/models/SimpleTensoRF09.py:767-819).
//
// Mapping: one 64-lane wavefront per ray, samples walked in chunks of 64 (lane l owns
// sample c*64+l), so every load of sigma/z/rgb is a contiguous 256-768 B run.  The
// transmittance T_k = prod_{j<k}(1 - alpha_j + 1e-10) is evaluated in log space:
// a wave prefix sum of log(q_j) plus a running carry between chunks, T = exp(carry+excl).
// The forward makes two passes over the ray (the second recomputes T to form the depth
// variances around the final depth, as the reference does); the backward makes a forward
// pass that records the log-T carry at every chunk start in LDS and then a reverse pass
// that builds the suffix sums sum_{j>k} T_j (dL/dT_j) chunk by chunk, so no per-sample
// intermediate (alpha, T, weights) is stored between forward and backward.
#include "common.h"

namespace srf {
namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kMaxChunks = 64;  // samples per ray <= 4096

struct RaySpan {
  int64_t begin;
  int count;
};

__device__ __forceinline__ RaySpan ray_span(int r, int n_samples, const int32_t* off) {
  if (off) return {(int64_t)off[r], off[r + 1] - off[r]};
  return {(int64_t)r * n_samples, n_samples};
}

struct RayConst {
  float dnorm;  // |d| used for delta
  float tn;     // metric conversion: t_near
  float cz;     // (oz + tn dz) / dz
  float inf_depth;
};

__device__ __forceinline__ RayConst ray_const(int r, const float* rays_o, const float* rays_d,
                                              const float* rays_d_ndc, int ndc) {
  RayConst rc;
  const float* dv = ndc ? rays_d_ndc + 3 * r : rays_d + 3 * r;
  rc.dnorm = torch_norm3(dv[0], dv[1], dv[2]);
  rc.inf_depth = ndc ? 1.0f : 1e10f;
  rc.tn = 0.f;
  rc.cz = 0.f;
  if (ndc) {
    // convert_depth_from_ndc, src/utils/CommonUtils04.py:217-223 (near = 1)
    const float oz = rays_o[3 * r + 2];
    const float dz = rays_d[3 * r + 2];
    rc.tn = (-(__fadd_rn(1.0f, oz))) / dz;
    rc.cz = __fadd_rn(oz, __fmul_rn(rc.tn, dz)) / dz;
  }
  return rc;
}

__device__ __forceinline__ float metric_depth(const RayConst& rc, float zn) {
  const float c = (zn == 1.0f) ? 1e-3f : 0.0f;
  const float inv = 1.0f / __fadd_rn(__fsub_rn(1.0f, zn), c);
  return __fadd_rn(__fmul_rn(rc.cz, __fsub_rn(inv, 1.0f)), rc.tn);
}

struct SampleTerms {
  float a;      // exp(-sigma delta ds)
  float alpha;  // 1 - a
  float q;      // 1 - alpha + 1e-10
  float lq;     // log q
  float delta;
};

__device__ __forceinline__ SampleTerms sample_terms(float sig, float zk, float zk1, float dnorm,
                                                    float ds) {
  SampleTerms t;
  t.delta = __fmul_rn(__fsub_rn(zk1, zk), dnorm);
  const float x = __fmul_rn(__fmul_rn(-sig, t.delta), ds);
  t.a = expf(x);
  t.alpha = __fsub_rn(1.0f, t.a);
  t.q = __fadd_rn(__fsub_rn(1.0f, t.alpha), 1e-10f);
  t.lq = logf(t.q);
  return t;
}

// Exclusive prefix of v inside the wave given its inclusive prefix.
__device__ __forceinline__ float exclusive_from_inclusive(float incl) {
  float e = __shfl_up(incl, 1, SRF_WAVE);
  return lane_id() == 0 ? 0.0f : e;
}

__global__ __launch_bounds__(256) void composite_fwd_kernel(
    int n_rays, int n_samples, const int32_t* __restrict__ off, const float* __restrict__ sigma,
    const float* __restrict__ rgb, const float* __restrict__ z, const float* __restrict__ rays_o,
    const float* __restrict__ rays_d, const float* __restrict__ rays_d_ndc, int ndc, int white_bkgd,
    float ds, float* __restrict__ rgb_map, float* __restrict__ acc_out,
    float* __restrict__ depth_out, float* __restrict__ depth_var_out,
    float* __restrict__ depth_ndc_out, float* __restrict__ depth_var_ndc_out,
    float* __restrict__ alpha_out, float* __restrict__ vis_out, float* __restrict__ w_out) {
  const int r = blockIdx.x * kWavesPerBlock + threadIdx.x / SRF_WAVE;
  if (r >= n_rays) return;
  const int lane = lane_id();
  const RaySpan sp = ray_span(r, n_samples, off);
  const RayConst rc = ray_const(r, rays_o, rays_d, rays_d_ndc, ndc);

  // pass 1: weights, maps
  float carry = 0.f;
  float s_acc = 0.f, s_r = 0.f, s_g = 0.f, s_b = 0.f, s_zn = 0.f, s_zm = 0.f;
  for (int c0 = 0; c0 < sp.count; c0 += SRF_WAVE) {
    const int k = c0 + lane;
    const bool valid = k < sp.count;
    const int64_t i = sp.begin + k;
    float sig = 0.f, zk = 0.f, zk1 = 0.f;
    if (valid) {
      sig = sigma[i];
      zk = z[i];
      zk1 = (k + 1 < sp.count) ? z[i + 1] : rc.inf_depth;
    }
    const SampleTerms t = sample_terms(sig, zk, zk1, rc.dnorm, ds);
    const float lq = valid ? t.lq : 0.f;
    const float incl = wave_inclusive_sum(lq);
    const float T = expf(carry + exclusive_from_inclusive(incl));
    carry = carry + __shfl(incl, SRF_WAVE - 1, SRF_WAVE);
    if (valid) {
      const float w = __fmul_rn(t.alpha, T);
      if (alpha_out) alpha_out[i] = t.alpha;
      if (vis_out) vis_out[i] = T;
      if (w_out) w_out[i] = w;
      s_acc += w;
      if (rgb) {
        s_r += __fmul_rn(w, rgb[3 * i + 0]);
        s_g += __fmul_rn(w, rgb[3 * i + 1]);
        s_b += __fmul_rn(w, rgb[3 * i + 2]);
      }
      if (ndc) {
        s_zn += __fmul_rn(w, zk);
        s_zm += __fmul_rn(w, metric_depth(rc, zk));
      } else {
        s_zm += __fmul_rn(w, zk);
      }
    }
  }
  s_acc = wave_sum(s_acc);
  s_r = wave_sum(s_r);
  s_g = wave_sum(s_g);
  s_b = wave_sum(s_b);
  s_zn = wave_sum(s_zn);
  s_zm = wave_sum(s_zm);
  const float den = __fadd_rn(s_acc, 1e-6f);
  const float D = s_zm / den;
  const float Dn = s_zn / den;

  // pass 2: variances around the final depths (recompute T; same op sequence as pass 1)
  carry = 0.f;
  float s_v = 0.f, s_vn = 0.f;
  for (int c0 = 0; c0 < sp.count; c0 += SRF_WAVE) {
    const int k = c0 + lane;
    const bool valid = k < sp.count;
    const int64_t i = sp.begin + k;
    float sig = 0.f, zk = 0.f, zk1 = 0.f;
    if (valid) {
      sig = sigma[i];
      zk = z[i];
      zk1 = (k + 1 < sp.count) ? z[i + 1] : rc.inf_depth;
    }
    const SampleTerms t = sample_terms(sig, zk, zk1, rc.dnorm, ds);
    const float lq = valid ? t.lq : 0.f;
    const float incl = wave_inclusive_sum(lq);
    const float T = expf(carry + exclusive_from_inclusive(incl));
    carry = carry + __shfl(incl, SRF_WAVE - 1, SRF_WAVE);
    if (valid) {
      const float w = __fmul_rn(t.alpha, T);
      const float zm = ndc ? metric_depth(rc, zk) : zk;
      const float dm = __fsub_rn(zm, D);
      s_v += __fmul_rn(w, __fmul_rn(dm, dm));
      if (ndc) {
        const float dn = __fsub_rn(zk, Dn);
        s_vn += __fmul_rn(w, __fmul_rn(dn, dn));
      }
    }
  }
  s_v = wave_sum(s_v);
  s_vn = wave_sum(s_vn);
  if (lane == 0) {
    const float bg = white_bkgd ? __fsub_rn(1.0f, s_acc) : 0.0f;
    if (rgb_map) {
      rgb_map[3 * r + 0] = white_bkgd ? __fadd_rn(s_r, bg) : s_r;
      rgb_map[3 * r + 1] = white_bkgd ? __fadd_rn(s_g, bg) : s_g;
      rgb_map[3 * r + 2] = white_bkgd ? __fadd_rn(s_b, bg) : s_b;
    }
    acc_out[r] = s_acc;
    depth_out[r] = D;
    depth_var_out[r] = s_v;
    if (ndc) {
      if (depth_ndc_out) depth_ndc_out[r] = Dn;
      if (depth_var_ndc_out) depth_var_ndc_out[r] = s_vn;
    }
  }
}

__device__ __forceinline__ float ld_or0(const float* p, int64_t i) { return p ? p[i] : 0.f; }

__global__ __launch_bounds__(256) void composite_bwd_kernel(
    int n_rays, int n_samples, const int32_t* __restrict__ off, const float* __restrict__ sigma,
    const float* __restrict__ rgb, const float* __restrict__ z, const float* __restrict__ rays_o,
    const float* __restrict__ rays_d, const float* __restrict__ rays_d_ndc, int ndc, int white_bkgd,
    float ds, const float* __restrict__ acc_in, const float* __restrict__ depth_in,
    const float* __restrict__ depth_ndc_in, const float* __restrict__ g_rgb,
    const float* __restrict__ g_acc, const float* __restrict__ g_depth,
    const float* __restrict__ g_depth_var, const float* __restrict__ g_depth_ndc,
    const float* __restrict__ g_depth_var_ndc, const float* __restrict__ g_alpha,
    const float* __restrict__ g_vis, const float* __restrict__ g_w, float* __restrict__ d_sigma,
    float* __restrict__ d_rgb) {
  __shared__ float carries[kWavesPerBlock][kMaxChunks];
  const int wv = threadIdx.x / SRF_WAVE;
  const int r = blockIdx.x * kWavesPerBlock + wv;
  if (r >= n_rays) return;
  const int lane = lane_id();
  const RaySpan sp = ray_span(r, n_samples, off);
  const RayConst rc = ray_const(r, rays_o, rays_d, rays_d_ndc, ndc);

  const float eps = 1e-6f;
  const float A = acc_in[r];
  const float inv = 1.0f / (A + eps);
  const float D = depth_in[r];
  const float Dn = (ndc && depth_ndc_in) ? depth_ndc_in[r] : 0.f;
  const float gC0 = ld_or0(g_rgb, 3 * r + 0), gC1 = ld_or0(g_rgb, 3 * r + 1),
              gC2 = ld_or0(g_rgb, 3 * r + 2);
  float gA = ld_or0(g_acc, r);
  if (white_bkgd) gA -= (gC0 + gC1 + gC2);
  const float gD = ld_or0(g_depth, r), gV = ld_or0(g_depth_var, r);
  const float gDn = ndc ? ld_or0(g_depth_ndc, r) : 0.f;
  const float gVn = ndc ? ld_or0(g_depth_var_ndc, r) : 0.f;

  // forward pass: log-T carry at each chunk start
  const int n_chunks = (sp.count + SRF_WAVE - 1) / SRF_WAVE;
  if (n_chunks > kMaxChunks) {
    // A ragged ray longer than the documented maximum: flag it instead of overrunning LDS.
    for (int k = lane; k < sp.count; k += SRF_WAVE) d_sigma[sp.begin + k] = __builtin_nanf("");
    return;
  }
  float carry = 0.f;
  for (int c = 0; c < n_chunks; ++c) {
    if (lane == 0) carries[wv][c] = carry;
    const int k = c * SRF_WAVE + lane;
    const bool valid = k < sp.count;
    const int64_t i = sp.begin + k;
    float sig = 0.f, zk = 0.f, zk1 = 0.f;
    if (valid) {
      sig = sigma[i];
      zk = z[i];
      zk1 = (k + 1 < sp.count) ? z[i + 1] : rc.inf_depth;
    }
    const SampleTerms t = sample_terms(sig, zk, zk1, rc.dnorm, ds);
    carry = carry + wave_sum(valid ? t.lq : 0.f);
  }
  __builtin_amdgcn_wave_barrier();

  // reverse pass: suffix sums of x_j = T_j (g_vis_j + g_w_j alpha_j)
  float suffix = 0.f;
  for (int c = n_chunks - 1; c >= 0; --c) {
    const int k = c * SRF_WAVE + lane;
    const bool valid = k < sp.count;
    const int64_t i = sp.begin + k;
    float sig = 0.f, zk = 0.f, zk1 = 0.f, c0 = 0.f, c1 = 0.f, c2 = 0.f;
    if (valid) {
      sig = sigma[i];
      zk = z[i];
      zk1 = (k + 1 < sp.count) ? z[i + 1] : rc.inf_depth;
      if (rgb) {
        c0 = rgb[3 * i + 0];
        c1 = rgb[3 * i + 1];
        c2 = rgb[3 * i + 2];
      }
    }
    const SampleTerms t = sample_terms(sig, zk, zk1, rc.dnorm, ds);
    const float lq = valid ? t.lq : 0.f;
    const float incl = wave_inclusive_sum(lq);
    const float T = expf(carries[wv][c] + exclusive_from_inclusive(incl));
    const float w = t.alpha * T;
    // dL/dw_k (all per-ray maps + the direct weights gradient)
    float g = gC0 * c0 + gC1 * c1 + gC2 * c2 + gA;
    const float zm = ndc ? metric_depth(rc, zk) : zk;
    const float dm = zm - D;
    g += gD * dm * inv + gV * (dm * dm - 2.f * eps * D * dm * inv);
    if (ndc) {
      const float dn = zk - Dn;
      g += gDn * dn * inv + gVn * (dn * dn - 2.f * eps * Dn * dn * inv);
    }
    if (valid) g += ld_or0(g_w, i);
    const float x = valid ? T * (ld_or0(g_vis, i) + g * t.alpha) : 0.f;
    const float sincl = wave_inclusive_suffix_sum(x);
    float sexcl = __shfl_down(sincl, 1, SRF_WAVE);
    if (lane == SRF_WAVE - 1) sexcl = 0.f;
    const float S = sexcl + suffix;
    suffix = suffix + __shfl(sincl, 0, SRF_WAVE);
    if (valid) {
      const float dalpha = (g_alpha ? g_alpha[i] : 0.f) + g * T;
      d_sigma[i] = t.delta * ds * t.a * (dalpha - S / t.q);
      if (d_rgb) {
        d_rgb[3 * i + 0] = w * gC0;
        d_rgb[3 * i + 1] = w * gC1;
        d_rgb[3 * i + 2] = w * gC2;
      }
    }
  }
}

// Weighted sums of per-sample values along each ray (LowRankTensor.volume_render,
// src/models/SimpleTensoRF09.py:814-817): one wave per ray, values read channel-interleaved.
__global__ __launch_bounds__(256) void volume_render_fwd_kernel(int n_rays, int n_samples,
                                                                const int32_t* __restrict__ off,
                                                                const float* __restrict__ w,
                                                                const float* __restrict__ v, int C,
                                                                float* __restrict__ out) {
  const int r = blockIdx.x * kWavesPerBlock + threadIdx.x / SRF_WAVE;
  if (r >= n_rays) return;
  const int lane = lane_id();
  const RaySpan sp = ray_span(r, n_samples, off);
  for (int c = 0; c < C; ++c) {
    float s = 0.f;
    for (int k = lane; k < sp.count; k += SRF_WAVE) {
      const int64_t i = sp.begin + k;
      s += __fmul_rn(w[i], v[i * C + c]);
    }
    s = wave_sum(s);
    if (lane == 0) out[(int64_t)r * C + c] = s;
  }
}

__global__ __launch_bounds__(256) void volume_render_bwd_kernel(int n_rays, int n_samples,
                                                                const int32_t* __restrict__ off,
                                                                const float* __restrict__ w,
                                                                const float* __restrict__ v, int C,
                                                                const float* __restrict__ g_out,
                                                                float* __restrict__ g_w,
                                                                float* __restrict__ g_v) {
  const int r = blockIdx.x * kWavesPerBlock + threadIdx.x / SRF_WAVE;
  if (r >= n_rays) return;
  const int lane = lane_id();
  const RaySpan sp = ray_span(r, n_samples, off);
  const float* go = g_out + (int64_t)r * C;
  for (int k = lane; k < sp.count; k += SRF_WAVE) {
    const int64_t i = sp.begin + k;
    if (g_w) {
      float s = 0.f;
      for (int c = 0; c < C; ++c) s += __fmul_rn(go[c], v[i * C + c]);
      g_w[i] = s;
    }
    if (g_v) {
      const float wi = w[i];
      for (int c = 0; c < C; ++c) g_v[i * C + c] = __fmul_rn(wi, go[c]);
    }
  }
}

}  // namespace
}  // namespace srf

using namespace srf;

extern "C" int srf_composite_fwd(int n_rays, int n_samples, const int32_t* ray_offsets,
                                 const float* sigma, const float* rgb, const float* z,
                                 const float* rays_o, const float* rays_d, const float* rays_d_ndc,
                                 int ndc, int white_bkgd, float distance_scale, float* rgb_map,
                                 float* acc, float* depth, float* depth_var, float* depth_ndc,
                                 float* depth_var_ndc, float* alpha, float* visibility,
                                 float* weights, void* stream) {
  if (n_rays < 0 || n_samples < 0) return SRF_EINVAL;
  if (n_rays == 0) return SRF_OK;
  if (!sigma || !z || !acc || !depth || !depth_var) return SRF_EINVAL;
  if (ndc && (!rays_o || !rays_d || !rays_d_ndc)) return SRF_EINVAL;
  if (!ndc && !rays_d) return SRF_EINVAL;
  const int blocks = (n_rays + kWavesPerBlock - 1) / kWavesPerBlock;
  hipLaunchKernelGGL(composite_fwd_kernel, dim3(blocks), dim3(kWavesPerBlock * SRF_WAVE), 0,
                     as_stream(stream), n_rays, n_samples, ray_offsets, sigma, rgb, z, rays_o,
                     rays_d, rays_d_ndc, ndc, white_bkgd, distance_scale, rgb_map, acc, depth,
                     depth_var, depth_ndc, depth_var_ndc, alpha, visibility, weights);
  return launch_status();
}

extern "C" int srf_composite_bwd(int n_rays, int n_samples, const int32_t* ray_offsets,
                                 const float* sigma, const float* rgb, const float* z,
                                 const float* rays_o, const float* rays_d, const float* rays_d_ndc,
                                 int ndc, int white_bkgd, float distance_scale, const float* acc,
                                 const float* depth, const float* depth_ndc, const float* g_rgb,
                                 const float* g_acc, const float* g_depth,
                                 const float* g_depth_var, const float* g_depth_ndc,
                                 const float* g_depth_var_ndc, const float* g_alpha,
                                 const float* g_visibility, const float* g_weights, float* d_sigma,
                                 float* d_rgb, void* stream) {
  if (n_rays < 0 || n_samples < 0) return SRF_EINVAL;
  if (n_rays == 0) return SRF_OK;
  if (!ray_offsets && n_samples > kMaxChunks * SRF_WAVE) return SRF_EUNSUPPORTED;
  if (!sigma || !z || !acc || !depth || !d_sigma) return SRF_EINVAL;
  if (ndc && (!rays_o || !rays_d || !rays_d_ndc)) return SRF_EINVAL;
  if (!ndc && !rays_d) return SRF_EINVAL;
  // Ragged rays: the caller guarantees max count <= n_samples (the documented max).
  if (n_samples > kMaxChunks * SRF_WAVE) return SRF_EUNSUPPORTED;
  const int blocks = (n_rays + kWavesPerBlock - 1) / kWavesPerBlock;
  hipLaunchKernelGGL(composite_bwd_kernel, dim3(blocks), dim3(kWavesPerBlock * SRF_WAVE), 0,
                     as_stream(stream), n_rays, n_samples, ray_offsets, sigma, rgb, z, rays_o,
                     rays_d, rays_d_ndc, ndc, white_bkgd, distance_scale, acc, depth, depth_ndc,
                     g_rgb, g_acc, g_depth, g_depth_var, g_depth_ndc, g_depth_var_ndc, g_alpha,
                     g_visibility, g_weights, d_sigma, d_rgb);
  return launch_status();
}

extern "C" int srf_volume_render_fwd(int n_rays, int n_samples, const int32_t* ray_offsets,
                                     const float* weights, const float* values, int n_channels,
                                     float* out, void* stream) {
  if (n_rays < 0 || n_channels < 1 || (!ray_offsets && n_samples < 1)) return SRF_EINVAL;
  if (n_rays == 0) return SRF_OK;
  if (!weights || !values || !out) return SRF_EINVAL;
  const int blocks = (n_rays + kWavesPerBlock - 1) / kWavesPerBlock;
  hipLaunchKernelGGL(volume_render_fwd_kernel, dim3(blocks), dim3(kWavesPerBlock * SRF_WAVE), 0,
                     as_stream(stream), n_rays, n_samples, ray_offsets, weights, values, n_channels, out);
  return launch_status();
}

extern "C" int srf_volume_render_bwd(int n_rays, int n_samples, const int32_t* ray_offsets,
                                     const float* weights, const float* values, int n_channels,
                                     const float* g_out, float* g_weights, float* g_values,
                                     void* stream) {
  if (n_rays < 0 || n_channels < 1 || (!ray_offsets && n_samples < 1)) return SRF_EINVAL;
  if (n_rays == 0) return SRF_OK;
  if (!g_out || (g_weights && !values) || (g_values && !weights)) return SRF_EINVAL;
  const int blocks = (n_rays + kWavesPerBlock - 1) / kWavesPerBlock;
  hipLaunchKernelGGL(volume_render_bwd_kernel, dim3(blocks), dim3(kWavesPerBlock * SRF_WAVE), 0,
                     as_stream(stream), n_rays, n_samples, ray_offsets, weights, values, n_channels, g_out,
                     g_weights, g_values);
  return launch_status();
}
