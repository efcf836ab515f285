// Ray sampling on gfx950: stratified depths, inverse-CDF (hierarchical) sampling and the
// coarse+fine sorted merge.
//
// Reference semantics: SimpleNeRF.get_z_vals_coarse / get_z_vals_fine / sample_pdf
// (src/models/SimpleNeRF17.py:330-417) and their TensoRF copies
// (src/models/SimpleTensoRF09.py:353-460).
//
// Bit-exactness: every arithmetic step that the reference performs as a separate PyTorch
// op is written as its own correctly rounded operation (the file is compiled with
// -ffp-contract=off), torch.linspace is reproduced exactly (common.h), and the cumulative
// sum of the pdf accumulates in double as PyTorch's CPU cumsum does.  Given the same CDF
// and u, searchsorted(right=True) indices and the interpolated samples are bit-identical.
//
// Mapping: one wavefront per ray.  The CDF and the bins live in LDS; each lane resolves
// its u values by a binary search over LDS (upper_bound).  The merge sorts the union of the
// coarse and fine depths: up to 256 of them with a bitonic network in registers (four slots
// per lane), more with a stable counting merge (every element written at its rank).  Both
// only move values, so the output is torch.sort's values bit for bit.
#include "common.h"

namespace srf {
namespace {

constexpr int kMaxBins = 1024;

__global__ __launch_bounds__(256) void stratified_kernel(int n_rays, int n_samples, float near,
                                                         float far, int lindisp, float inv_near,
                                                         float inv_far, const float* __restrict__ u,
                                                         float* __restrict__ z) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)n_rays * n_samples;
  if (idx >= total) return;
  const int k = (int)(idx % n_samples);
  auto zval = [&](int j) -> float {
    const float t = torch_linspace(0.f, 1.f, n_samples, j);
    const float omt = __fsub_rn(1.f, t);
    if (!lindisp) return __fadd_rn(__fmul_rn(near, omt), __fmul_rn(far, t));
    return 1.f / __fadd_rn(__fmul_rn(inv_near, omt), __fmul_rn(inv_far, t));
  };
  const float zk = zval(k);
  if (!u) {
    z[idx] = zk;
    return;
  }
  // mids = .5 (z[1:] + z[:-1]); upper = [mids, z[-1]]; lower = [z[0], mids]
  const float upper = (k + 1 < n_samples) ? __fmul_rn(0.5f, __fadd_rn(zval(k + 1), zk)) : zk;
  const float lower = (k > 0) ? __fmul_rn(0.5f, __fadd_rn(zk, zval(k - 1))) : zk;
  z[idx] = __fadd_rn(lower, __fmul_rn(__fsub_rn(upper, lower), u[idx]));
}

// Builds the CDF of one ray into LDS: w = weights + 1e-5, pdf = w / sum(w),
// cdf = [0, cumsum(pdf)] with the cumulative sum accumulated in double.
__device__ void build_cdf(const float* __restrict__ w_row, int n_w, float* s_cdf) {
  const int lane = lane_id();
  double tot = 0.0;
  for (int j = lane; j < n_w; j += SRF_WAVE) tot += (double)__fadd_rn(w_row[j], 1e-5f);
  const float total = (float)wave_sum_d(tot);
  double carry = 0.0;
  if (lane == 0) s_cdf[0] = 0.f;
  for (int c0 = 0; c0 < n_w; c0 += SRF_WAVE) {
    const int j = c0 + lane;
    double p = 0.0;
    if (j < n_w) p = (double)(__fadd_rn(w_row[j], 1e-5f) / total);
    double incl = p;
#pragma unroll
    for (int d = 1; d < SRF_WAVE; d <<= 1) {
      double o = __shfl_up(incl, d, SRF_WAVE);
      if (lane >= d) incl += o;
    }
    if (j < n_w) s_cdf[j + 1] = (float)(carry + incl);
    carry += __shfl(incl, SRF_WAVE - 1, SRF_WAVE);
  }
}

// Number of cdf entries <= u (torch.searchsorted(cdf, u, right=True)).
__device__ __forceinline__ int upper_bound(const float* s_cdf, int n, float u) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (s_cdf[mid] <= u) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ float invert_cdf(const float* s_cdf, const float* s_bins, int n_bins,
                                            float uv, int* ind_out) {
  const int ind = upper_bound(s_cdf, n_bins, uv);
  const int below = ind - 1 > 0 ? ind - 1 : 0;
  const int above = ind < n_bins - 1 ? ind : n_bins - 1;
  const float cb = s_cdf[below], ca = s_cdf[above];
  float denom = __fsub_rn(ca, cb);
  if (denom < 1e-5f) denom = 1.f;
  const float t = __fsub_rn(uv, cb) / denom;
  const float bb = s_bins[below], ba = s_bins[above];
  *ind_out = ind;
  return __fadd_rn(bb, __fmul_rn(t, __fsub_rn(ba, bb)));
}

__global__ __launch_bounds__(64) void sample_pdf_kernel(int n_rays, int n_bins, int n_fine,
                                                        const float* __restrict__ bins,
                                                        const float* __restrict__ weights,
                                                        const float* __restrict__ cdf_in,
                                                        const float* __restrict__ u,
                                                        float* __restrict__ samples,
                                                        int64_t* __restrict__ inds,
                                                        float* __restrict__ cdf_out) {
  // dynamic LDS sized to the call (2 n_bins floats), so that short rows keep many waves resident
  extern __shared__ float s_dyn[];
  float* s_cdf = s_dyn;
  float* s_bins = s_dyn + n_bins;
  const int r = blockIdx.x;
  if (r >= n_rays) return;
  const int lane = lane_id();
  const int64_t rb = (int64_t)r * n_bins;
  for (int j = lane; j < n_bins; j += SRF_WAVE) s_bins[j] = bins[rb + j];
  if (cdf_in) {
    for (int j = lane; j < n_bins; j += SRF_WAVE) s_cdf[j] = cdf_in[rb + j];
  } else {
    build_cdf(weights + (int64_t)r * (n_bins - 1), n_bins - 1, s_cdf);
  }
  __syncthreads();
  if (cdf_out)
    for (int j = lane; j < n_bins; j += SRF_WAVE) cdf_out[rb + j] = s_cdf[j];
  const int64_t ub = (int64_t)r * n_fine;
  for (int i = lane; i < n_fine; i += SRF_WAVE) {
    const float uv = u ? u[ub + i] : torch_linspace(0.f, 1.f, n_fine, i);
    int ind;
    samples[ub + i] = invert_cdf(s_cdf, s_bins, n_bins, uv, &ind);
    if (inds) inds[ub + i] = (int64_t)ind;
  }
}

// Ascending bitonic sort of the wave's 256 slots, slot 4 lane + q in v[q] (padding: +inf).  Each
// compare-exchange either keeps or swaps the two values (a select on one comparison, the same
// decision on both sides of a cross-lane pair), so the result is a permutation of the input:
// equal values of different bits (+0, -0) and NaN stay in the multiset.
__device__ __forceinline__ void wave_bitonic_sort256(float (&v)[4]) {
  const int lane = lane_id();
#pragma unroll
  for (int k = 2; k <= 256; k <<= 1) {
#pragma unroll
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      if (jj >= 4) {
        // partner slot in lane ^ (jj / 4), same q; the lane's four slots share both bits
        const bool asc = ((4 * lane) & k) == 0, lower = ((4 * lane) & jj) == 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float o = __shfl_xor(v[q], jj >> 2, SRF_WAVE);
          const bool take = (lower == asc) ? (o < v[q]) : (v[q] < o);
          v[q] = take ? o : v[q];
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (q & jj) continue;
          const bool asc = ((4 * lane + q) & k) == 0;
          const float x = v[q], y = v[q | jj];
          const bool swap = asc ? (y < x) : (x < y);
          v[q] = swap ? y : x;
          v[q | jj] = swap ? x : y;
        }
      }
    }
  }
}

__global__ __launch_bounds__(64) void fine_z_kernel(int n_rays, int n_coarse, int n_fine,
                                                    const float* __restrict__ z_coarse,
                                                    const float* __restrict__ w_coarse,
                                                    const float* __restrict__ u,
                                                    float* __restrict__ z_merged,
                                                    float* __restrict__ z_fine_out) {
  const int n_bins = n_coarse - 1;
  const int n_all = n_coarse + n_fine;
  // dynamic LDS sized to the call (2 n_bins + n_all floats)
  extern __shared__ float s_dyn[];
  float* s_cdf = s_dyn;
  float* s_bins = s_dyn + n_bins;
  float* s_all = s_dyn + 2 * n_bins;
  const int r = blockIdx.x;
  if (r >= n_rays) return;
  const int lane = lane_id();
  const float* zc = z_coarse + (int64_t)r * n_coarse;
  // z_vals_mid = .5 (z[1:] + z[:-1]) are the bins; weights[..., 1:-1] build the CDF
  for (int j = lane; j < n_coarse; j += SRF_WAVE) {
    const float zj = zc[j];
    s_all[j] = zj;
    if (j < n_bins) s_bins[j] = __fmul_rn(0.5f, __fadd_rn(zc[j + 1], zj));
  }
  build_cdf(w_coarse + (int64_t)r * n_coarse + 1, n_coarse - 2, s_cdf);
  __syncthreads();
  const int64_t ub = (int64_t)r * n_fine;
  for (int i = lane; i < n_fine; i += SRF_WAVE) {
    const float uv = u ? u[ub + i] : torch_linspace(0.f, 1.f, n_fine, i);
    int ind;
    const float zf = invert_cdf(s_cdf, s_bins, n_bins, uv, &ind);
    s_all[n_coarse + i] = zf;
    if (z_fine_out) z_fine_out[ub + i] = zf;
  }
  __syncthreads();
  float* out = z_merged + (int64_t)r * n_all;
  if (n_all <= 4 * SRF_WAVE) {
    float v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = 4 * lane + q;
      v[q] = e < n_all ? s_all[e] : __builtin_inff();
    }
    wave_bitonic_sort256(v);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (4 * lane + q < n_all) out[4 * lane + q] = v[q];
    return;
  }
  // stable counting merge: element e goes to #{v < v_e} + #{j < e : v_j == v_e}
  for (int e = lane; e < n_all; e += SRF_WAVE) {
    const float ve = s_all[e];
    int rank = 0;
    for (int j = 0; j < n_all; ++j) {
      const float vj = s_all[j];
      rank += (vj < ve) || (vj == ve && j < e);
    }
    out[rank] = ve;
  }
}

// Non-NDC TensoRF depths (SimpleTensoRF09.py:382-400): the AABB entry of each ray plus a
// fixed step.  One thread per (ray, sample); the per-ray entry is recomputed per thread
// (6 divisions) rather than staged, which keeps the kernel a single pass.
struct Box {
  float lo[3], hi[3];
};

__global__ __launch_bounds__(256) void aabb_kernel(int n_rays, int n_samples, const float* __restrict__ o,
                                                   const float* __restrict__ d, Box box, float near,
                                                   float far, float step, const float* __restrict__ u,
                                                   float* __restrict__ z) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n_rays * n_samples) return;
  const int r = (int)(idx / n_samples), k = (int)(idx % n_samples);
  float tmin = -__builtin_inff();
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float da = d[3 * r + a], oa = o[3 * r + a];
    const float v = da == 0.f ? 1e-6f : da;
    const float ra = __fdiv_rn(__fsub_rn(box.hi[a], oa), v);
    const float rb = __fdiv_rn(__fsub_rn(box.lo[a], oa), v);
    tmin = fmaxf(tmin, fminf(ra, rb));
  }
  tmin = fminf(fmaxf(tmin, near), far);
  const float rng = u ? __fadd_rn((float)k, u[r]) : (float)k;
  z[idx] = __fadd_rn(tmin, __fmul_rn(step, rng));
}

}  // namespace
}  // namespace srf

using namespace srf;

extern "C" int srf_stratified_samples(int n_rays, int n_samples, float near, float far, int lindisp,
                                      const float* u, float* z, void* stream) {
  if (n_rays < 0 || n_samples < 1) return SRF_EINVAL;
  if (n_rays == 0) return SRF_OK;
  if (!z) return SRF_EINVAL;
  const int64_t total = (int64_t)n_rays * n_samples;
  const int threads = 256;
  const int64_t blocks = (total + threads - 1) / threads;
  if (blocks > 0x7fffffff) return SRF_EUNSUPPORTED;
  // 1./near and 1./far are Python double divisions in the reference (:345).
  const float inv_near = (float)(1.0 / (double)near);
  const float inv_far = (float)(1.0 / (double)far);
  hipLaunchKernelGGL(stratified_kernel, dim3((unsigned)blocks), dim3(threads), 0, as_stream(stream),
                     n_rays, n_samples, near, far, lindisp, inv_near, inv_far, u, z);
  return launch_status();
}

extern "C" int srf_sample_pdf(int n_rays, int n_bins, int n_fine, const float* bins,
                              const float* weights, const float* cdf_in, const float* u,
                              float* samples, int64_t* inds, float* cdf_out, void* stream) {
  if (n_rays < 0 || n_bins < 2 || n_bins > kMaxBins || n_fine < 1) return SRF_EINVAL;
  if (n_rays == 0) return SRF_OK;
  if (!bins || !samples || (!weights && !cdf_in)) return SRF_EINVAL;
  const size_t lds = (size_t)2 * n_bins * sizeof(float);
  hipLaunchKernelGGL(sample_pdf_kernel, dim3(n_rays), dim3(SRF_WAVE), lds, as_stream(stream), n_rays, n_bins,
                     n_fine, bins, weights, cdf_in, u, samples, inds, cdf_out);
  return launch_status();
}

extern "C" int srf_fine_z_vals(int n_rays, int n_coarse, int n_fine, const float* z_coarse,
                               const float* weights_coarse, const float* u, float* z_merged,
                               float* z_fine, void* stream) {
  if (n_rays < 0 || n_coarse < 3 || n_fine < 1 || n_coarse + n_fine > kMaxBins) return SRF_EINVAL;
  if (n_rays == 0) return SRF_OK;
  if (!z_coarse || !weights_coarse || !z_merged) return SRF_EINVAL;
  const size_t lds = (size_t)(2 * (n_coarse - 1) + n_coarse + n_fine) * sizeof(float);
  hipLaunchKernelGGL(fine_z_kernel, dim3(n_rays), dim3(SRF_WAVE), lds, as_stream(stream), n_rays,
                     n_coarse, n_fine, z_coarse, weights_coarse, u, z_merged, z_fine);
  return launch_status();
}

extern "C" int srf_aabb_samples(int n_rays, int n_samples, const float* rays_o, const float* rays_d,
                                const float* bbox, float near, float far, float step, const float* u,
                                float* z, void* stream) {
  if (n_rays < 0 || n_samples < 1 || !bbox) return SRF_EINVAL;
  if (n_rays == 0) return SRF_OK;
  if (!rays_o || !rays_d || !z) return SRF_EINVAL;
  Box b;
  for (int a = 0; a < 3; ++a) {
    b.lo[a] = bbox[a];
    b.hi[a] = bbox[3 + a];
  }
  const int64_t total = (int64_t)n_rays * n_samples;
  const int64_t blocks = (total + 255) / 256;
  if (blocks > 0x7fffffff) return SRF_EUNSUPPORTED;
  hipLaunchKernelGGL(aabb_kernel, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), n_rays, n_samples,
                     rays_o, rays_d, b, near, far, step, u, z);
  return launch_status();
}
