// TensoRF model modifications and their regularisers on gfx950.
//
// Reference semantics (src/models/SimpleTensoRF09.py):
//   * LowRankTensor.update_alpha_mask / compute_alpha (:849-897): the dense grid points
//     (srf_dense_grid_points) go through the density gather (srf_vm_density_fwd), then
//     srf_alpha_mask_build turns the densities into alpha = 1 - exp(-sigma step), clamps, transposes to
//     (z, y, x), max-pools 3x3x3 (stride 1, padding 1), thresholds to 0 / 1 and finds the occupied box;
//   * VmDecomposedTensor.upsample_vectors_and_matrices (:1281-1302): srf_factor_upsample, bilinear with
//     align_corners = True on the texel-major factors;
//   * TotalVariationLoss.compute_tv_loss (src/loss_functions/TotalVariationLoss04.py:97-116):
//     srf_tv_fwd / _bwd over texel-major planes;
//   * MassConcentrationLoss.compute_entropy_loss (src/loss_functions/MassConcentrationLoss06.py:63-83):
//     srf_entropy_fwd / _bwd on the rays' weights.
// Reductions are fixed-order (per-block partials, then one block), so values do not depend on the
// schedule, and nothing here synchronises the host.
#include "common.h"

namespace srf {
namespace {

constexpr int kRed = 1024;

__device__ float block_sum(float v, float* sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  float t = 0.f;
  const int nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nw; ++i) t += sh[i];
  __syncthreads();
  return t;
}

// ------------------------------------------------------------------ dense grid
// pts[(i gy + j) gz + k] = bb0 (1 - s) + bb1 s with s the linspace(0, 1) position of (i, j, k) on each
// axis (torch.meshgrid indexing 'ij', :851-857).
__global__ void dense_grid_points_kernel(int gx, int gy, int gz, float x0, float y0, float z0, float x1, float y1,
                                         float z1, float* pts) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)gx * gy * gz;
  if (e >= n) return;
  const int k = (int)(e % gz), j = (int)((e / gz) % gy), i = (int)(e / ((int64_t)gy * gz));
  const float s[3] = {torch_linspace(0.f, 1.f, gx, i), torch_linspace(0.f, 1.f, gy, j), torch_linspace(0.f, 1.f, gz, k)};
  const float lo[3] = {x0, y0, z0}, hi[3] = {x1, y1, z1};
#pragma unroll
  for (int c = 0; c < 3; ++c)
    pts[3 * e + c] = __fadd_rn(__fmul_rn(lo[c], __fsub_rn(1.0f, s[c])), __fmul_rn(hi[c], s[c]));
}

// ---------------------------------------------------------------- alpha mask
// alpha_out[z][y][x] = (max over the 3x3x3 neighbourhood of clamp(1 - exp(-sigma length), 0, 1)) >=
// threshold ? 1 : 0, with sigma [gx][gy][gz] as the dense grid lists it; bounds (pre-set by
// alpha_bounds_init) receive the min / max x, y, z index of the occupied voxels.
__global__ void alpha_bounds_init_kernel(int* bounds) {
  if (threadIdx.x < 3) bounds[threadIdx.x] = 0x7fffffff;
  else if (threadIdx.x < 6) bounds[threadIdx.x] = -1;
}

__device__ __forceinline__ float alpha_at(const float* sigma, int gy, int gz, int x, int y, int z, float length) {
  const float s = sigma[((int64_t)x * gy + y) * gz + z];
  const float a = __fsub_rn(1.0f, expf(-__fmul_rn(s, length)));
  return fminf(fmaxf(a, 0.f), 1.f);
}

__global__ void alpha_mask_build_kernel(int gx, int gy, int gz, const float* sigma, float length, float threshold,
                                        float* alpha_out, int* bounds) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)gx * gy * gz;
  if (e >= n) return;
  const int x = (int)(e % gx), y = (int)((e / gx) % gy), z = (int)(e / ((int64_t)gx * gy));
  float m = -INFINITY;  // max_pool3d pads with -inf: out-of-grid neighbours never win
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int xx = x + dx, yy = y + dy, zz = z + dz;
        if (xx < 0 || xx >= gx || yy < 0 || yy >= gy || zz < 0 || zz >= gz) continue;
        m = fmaxf(m, alpha_at(sigma, gy, gz, xx, yy, zz, length));
      }
  const bool on = m >= threshold;
  alpha_out[e] = on ? 1.f : 0.f;
  if (on) {
    atomicMin(bounds + 0, x);
    atomicMin(bounds + 1, y);
    atomicMin(bounds + 2, z);
    atomicMax(bounds + 3, x);
    atomicMax(bounds + 4, y);
    atomicMax(bounds + 5, z);
  }
}

// -------------------------------------------------------------- factor upsample
// dst[y'][x'][c] from src[y][x][c] (texel-major), bilinear with align_corners = True as
// F.interpolate(mode='bilinear', align_corners=True) computes it: source index = scale * dst index
// with scale = (in - 1) / (out - 1), the lower texel its integer part, the weights lambda and 1 - lambda.
__global__ void factor_upsample_kernel(int c, int h, int w, const float* __restrict__ src, int h2, int w2,
                                       float* __restrict__ dst) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)h2 * w2 * c;
  if (e >= n) return;
  const int ch = (int)(e % c), x = (int)((e / c) % w2), y = (int)(e / ((int64_t)c * w2));
  const float sh = h2 > 1 ? (float)(h - 1) / (float)(h2 - 1) : 0.f;
  const float sw = w2 > 1 ? (float)(w - 1) / (float)(w2 - 1) : 0.f;
  const float yr = __fmul_rn(sh, (float)y), xr = __fmul_rn(sw, (float)x);
  const int y0 = (int)yr, x0 = (int)xr;
  const int yp = y0 < h - 1 ? 1 : 0, xp = x0 < w - 1 ? 1 : 0;
  const float ly1 = __fsub_rn(yr, (float)y0), lx1 = __fsub_rn(xr, (float)x0);
  const float ly0 = __fsub_rn(1.f, ly1), lx0 = __fsub_rn(1.f, lx1);
  auto at = [&](int yy, int xx) { return src[((int64_t)yy * w + xx) * c + ch]; };
  const float top = __fadd_rn(__fmul_rn(lx0, at(y0, x0)), __fmul_rn(lx1, at(y0, x0 + xp)));
  const float bot = __fadd_rn(__fmul_rn(lx0, at(y0 + yp, x0)), __fmul_rn(lx1, at(y0 + yp, x0 + xp)));
  dst[e] = __fadd_rn(__fmul_rn(ly0, top), __fmul_rn(ly1, bot));
}

// ------------------------------------------------------------------------- TV
constexpr int kMaxTvPlanes = 8;
struct TvPlanes {
  int n;
  int c[kMaxTvPlanes], h[kMaxTvPlanes], w[kMaxTvPlanes];
  const float* p[kMaxTvPlanes];  // texel-major [h][w][c]
  float* d[kMaxTvPlanes];        // gradients (backward), same layout
};

// partial[(plane, block)][2]: sums of squared differences along H and along W of the block's texels
__global__ void tv_partial_kernel(TvPlanes tp, int blocks_per_plane, float* partial) {
  __shared__ float sh[1024 / 64];
  const int pl = blockIdx.x / blocks_per_plane, b = blockIdx.x % blocks_per_plane;
  const int c = tp.c[pl], h = tp.h[pl], w = tp.w[pl];
  const float* p = tp.p[pl];
  const int64_t n = (int64_t)h * w * c;
  float sh_h = 0.f, sh_w = 0.f;
  for (int64_t e = (int64_t)b * blockDim.x + threadIdx.x; e < n; e += (int64_t)blocks_per_plane * blockDim.x) {
    const int x = (int)((e / c) % w), y = (int)(e / ((int64_t)c * w));
    const float v = p[e];
    if (y + 1 < h) {
      const float d = __fsub_rn(p[e + (int64_t)w * c], v);
      sh_h += __fmul_rn(d, d);
    }
    if (x + 1 < w) {
      const float d = __fsub_rn(p[e + c], v);
      sh_w += __fmul_rn(d, d);
    }
  }
  const float th = block_sum(sh_h, sh);
  const float tw = block_sum(sh_w, sh);
  if (threadIdx.x == 0) {
    partial[2 * (size_t)blockIdx.x] = th;
    partial[2 * (size_t)blockIdx.x + 1] = tw;
  }
}

// loss = scale * sum_p 2 (S_h / max(numel_h, 1) + S_w / max(numel_w, 1))
__global__ void tv_final_kernel(TvPlanes tp, int blocks_per_plane, const float* partial, float scale, float* loss) {
  __shared__ float sh[kRed / 64];
  float total = 0.f;
  for (int pl = 0; pl < tp.n; ++pl) {
    float a = 0.f, b = 0.f;
    for (int i = threadIdx.x; i < blocks_per_plane; i += blockDim.x) {
      a += partial[2 * ((size_t)pl * blocks_per_plane + i)];
      b += partial[2 * ((size_t)pl * blocks_per_plane + i) + 1];
    }
    const float sa = block_sum(a, sh), sb = block_sum(b, sh);
    const float nh = fmaxf((float)tp.c[pl] * (float)(tp.h[pl] - 1) * (float)tp.w[pl], 1.f);
    const float nw = fmaxf((float)tp.c[pl] * (float)tp.h[pl] * (float)(tp.w[pl] - 1), 1.f);
    total += 2.f * (sa / nh + sb / nw);
  }
  if (threadIdx.x == 0) *loss = total * scale;
}

// d p = g scale 2 [2 / numel_h ((p - p_up) [y > 0] - (p_down - p) [y < h - 1]) + the same along W]
__global__ void tv_bwd_kernel(TvPlanes tp, int blocks_per_plane, float scale, const float* g_loss) {
  const int pl = blockIdx.x / blocks_per_plane, b = blockIdx.x % blocks_per_plane;
  const int c = tp.c[pl], h = tp.h[pl], w = tp.w[pl];
  const float* p = tp.p[pl];
  float* d = tp.d[pl];
  if (!d) return;
  const int64_t n = (int64_t)h * w * c;
  const float nh = fmaxf((float)c * (float)(h - 1) * (float)w, 1.f);
  const float nw = fmaxf((float)c * (float)h * (float)(w - 1), 1.f);
  const float g = *g_loss * scale * 2.f;
  const float kh = 2.f / nh, kw = 2.f / nw;
  for (int64_t e = (int64_t)b * blockDim.x + threadIdx.x; e < n; e += (int64_t)blocks_per_plane * blockDim.x) {
    const int x = (int)((e / c) % w), y = (int)(e / ((int64_t)c * w));
    const float v = p[e];
    float gh = 0.f, gw = 0.f;
    if (y > 0) gh += __fsub_rn(v, p[e - (int64_t)w * c]);
    if (y + 1 < h) gh -= __fsub_rn(p[e + (int64_t)w * c], v);
    if (x > 0) gw += __fsub_rn(v, p[e - c]);
    if (x + 1 < w) gw -= __fsub_rn(p[e + c], v);
    d[e] = g * (kh * gh + kw * gw);
  }
}

// ------------------------------------------------------------ mass concentration
constexpr int kMaxBins = 64;
struct Bins {
  int n;
  int start[kMaxBins + 1];  // bin b covers samples [start[b], start[b + 1])
};

// One wave per ray: binned[r][b] = sum of the ray's weights in bin b; ray_loss[r] = sum_b entr(binned + 1e-3)
// (0 for rays not selected).
__global__ void entropy_rays_kernel(int n_rays, int n_samples, const float* w, const uint8_t* select, Bins bins,
                                    float* binned, float* ray_loss) {
  const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= n_rays) return;
  if (select && !select[r]) {
    if (lane == 0) ray_loss[r] = 0.f;
    return;
  }
  float total = 0.f;
  for (int b = 0; b < bins.n; ++b) {
    float s = 0.f;
    for (int k = bins.start[b] + lane; k < bins.start[b + 1]; k += 64) s += w[(int64_t)r * n_samples + k];
    s = wave_sum(s);
    const float x = __fadd_rn(s, 1e-3f);
    if (lane == 0) binned[(int64_t)r * bins.n + b] = x;
    total = __fadd_rn(total, -__fmul_rn(x, logf(x)));
  }
  if (lane == 0) ray_loss[r] = total;
}

__global__ __launch_bounds__(kRed) void entropy_final_kernel(int n_rays, const float* ray_loss, const uint8_t* select,
                                                             float* loss, int* count) {
  __shared__ float sh[kRed / 64];
  float s = 0.f, c = 0.f;
  for (int r = threadIdx.x; r < n_rays; r += blockDim.x) {
    if (select && !select[r]) continue;
    s += ray_loss[r];
    c += 1.f;
  }
  const float ts = block_sum(s, sh), tc = block_sum(c, sh);
  if (threadIdx.x == 0) {
    *loss = tc > 0.f ? ts / tc : 0.f;
    *count = (int)tc;
  }
}

// d w[r][k] = g / count * -(log(binned[r][bin(k)]) + 1) on selected rays, 0 elsewhere
__global__ void entropy_bwd_kernel(int n_rays, int n_samples, const uint8_t* select, Bins bins, const float* binned,
                                   const int* count, const float* g_loss, float* d_w) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n_rays * n_samples) return;
  const int r = (int)(e / n_samples), k = (int)(e % n_samples);
  const int m = *count;
  if ((select && !select[r]) || m == 0) {
    d_w[e] = 0.f;
    return;
  }
  int b = 0;
  while (b + 1 < bins.n && k >= bins.start[b + 1]) ++b;
  const float x = binned[(int64_t)r * bins.n + b];
  d_w[e] = (*g_loss / (float)m) * -(logf(x) + 1.f);
}

bool make_tv(int n, const int* c, const int* h, const int* w, const float* const* p, float* const* d, TvPlanes* tp) {
  if (n < 1 || n > kMaxTvPlanes || !c || !h || !w || !p) return false;
  tp->n = n;
  for (int i = 0; i < kMaxTvPlanes; ++i) {
    const bool in = i < n;
    tp->c[i] = in ? c[i] : 1;
    tp->h[i] = in ? h[i] : 1;
    tp->w[i] = in ? w[i] : 1;
    tp->p[i] = in ? p[i] : nullptr;
    tp->d[i] = (in && d) ? d[i] : nullptr;
    if (in && (c[i] < 1 || h[i] < 1 || w[i] < 1 || !p[i])) return false;
  }
  return true;
}

int tv_blocks_per_plane(const TvPlanes& tp) {
  int64_t most = 1;
  for (int i = 0; i < tp.n; ++i) {
    const int64_t n = (int64_t)tp.c[i] * tp.h[i] * tp.w[i];
    most = n > most ? n : most;
  }
  const int64_t b = (most + 256 * 8 - 1) / (256 * 8);  // about eight texels per thread
  return (int)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

}  // namespace
}  // namespace srf

using namespace srf;

extern "C" int srf_dense_grid_points(int gx, int gy, int gz, const float* bbox, float* pts, void* stream) {
  if (gx < 1 || gy < 1 || gz < 1 || !bbox || !pts) return SRF_EINVAL;
  const int64_t n = (int64_t)gx * gy * gz;
  const int threads = 256;
  hipLaunchKernelGGL(dense_grid_points_kernel, dim3((unsigned)((n + threads - 1) / threads)), dim3(threads), 0,
                     as_stream(stream), gx, gy, gz, bbox[0], bbox[1], bbox[2], bbox[3], bbox[4], bbox[5], pts);
  return launch_status();
}

extern "C" int srf_alpha_mask_build(int gx, int gy, int gz, const float* sigma, float length, float threshold,
                                    float* alpha_out, int* bounds, void* stream) {
  if (gx < 1 || gy < 1 || gz < 1 || !sigma || !alpha_out || !bounds) return SRF_EINVAL;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(alpha_bounds_init_kernel, dim3(1), dim3(64), 0, s, bounds);
  const int64_t n = (int64_t)gx * gy * gz;
  const int threads = 256;
  hipLaunchKernelGGL(alpha_mask_build_kernel, dim3((unsigned)((n + threads - 1) / threads)), dim3(threads), 0, s, gx,
                     gy, gz, sigma, length, threshold, alpha_out, bounds);
  return launch_status();
}

extern "C" int srf_factor_upsample(int channels, int height, int width, const float* src, int height_new,
                                   int width_new, float* dst, void* stream) {
  if (channels < 1 || height < 1 || width < 1 || height_new < 1 || width_new < 1 || !src || !dst) return SRF_EINVAL;
  const int64_t n = (int64_t)height_new * width_new * channels;
  const int threads = 256;
  hipLaunchKernelGGL(factor_upsample_kernel, dim3((unsigned)((n + threads - 1) / threads)), dim3(threads), 0,
                     as_stream(stream), channels, height, width, src, height_new, width_new, dst);
  return launch_status();
}

extern "C" size_t srf_tv_workspace_bytes(int n_planes) {
  return n_planes < 1 || n_planes > kMaxTvPlanes ? 0 : (size_t)n_planes * 1024 * 2 * sizeof(float);
}

extern "C" int srf_tv_fwd(int n_planes, const int* channels, const int* heights, const int* widths,
                          const float* const* planes, float scale, void* workspace, float* loss, void* stream) {
  TvPlanes tp;
  if (!make_tv(n_planes, channels, heights, widths, planes, nullptr, &tp) || !workspace || !loss) return SRF_EINVAL;
  const int bpp = tv_blocks_per_plane(tp);
  hipStream_t s = as_stream(stream);
  float* partial = reinterpret_cast<float*>(workspace);
  hipLaunchKernelGGL(tv_partial_kernel, dim3(bpp * n_planes), dim3(256), 0, s, tp, bpp, partial);
  const int st = launch_status();
  if (st != SRF_OK) return st;
  hipLaunchKernelGGL(tv_final_kernel, dim3(1), dim3(kRed), 0, s, tp, bpp, partial, scale, loss);
  return launch_status();
}

extern "C" int srf_tv_bwd(int n_planes, const int* channels, const int* heights, const int* widths,
                          const float* const* planes, float scale, const float* g_loss, float* const* d_planes,
                          void* stream) {
  TvPlanes tp;
  if (!make_tv(n_planes, channels, heights, widths, planes, d_planes, &tp) || !d_planes || !g_loss) return SRF_EINVAL;
  const int bpp = tv_blocks_per_plane(tp);
  hipLaunchKernelGGL(tv_bwd_kernel, dim3(bpp * n_planes), dim3(256), 0, as_stream(stream), tp, bpp, scale, g_loss);
  return launch_status();
}

extern "C" int srf_entropy_fwd(int n_rays, int n_samples, const float* weights, const uint8_t* select, int n_bins,
                               const int* bin_sizes, float* binned, float* ray_loss, float* loss, int* count,
                               void* stream) {
  if (n_rays < 0 || n_samples < 1 || n_bins < 1 || n_bins > kMaxBins || !bin_sizes || !weights || !binned ||
      !ray_loss || !loss || !count)
    return SRF_EINVAL;
  Bins bins;
  bins.n = n_bins;
  bins.start[0] = 0;
  for (int b = 0; b < n_bins; ++b) {
    if (bin_sizes[b] < 0) return SRF_EINVAL;
    bins.start[b + 1] = bins.start[b] + bin_sizes[b];
  }
  if (bins.start[n_bins] != n_samples) return SRF_EINVAL;
  hipStream_t s = as_stream(stream);
  if (n_rays > 0) {
    const int waves = 4;
    hipLaunchKernelGGL(entropy_rays_kernel, dim3((n_rays + waves - 1) / waves), dim3(64 * waves), 0, s, n_rays,
                       n_samples, weights, select, bins, binned, ray_loss);
    const int st = launch_status();
    if (st != SRF_OK) return st;
  }
  hipLaunchKernelGGL(entropy_final_kernel, dim3(1), dim3(kRed), 0, s, n_rays, ray_loss, select, loss, count);
  return launch_status();
}

extern "C" int srf_entropy_bwd(int n_rays, int n_samples, const uint8_t* select, int n_bins, const int* bin_sizes,
                               const float* binned, const int* count, const float* g_loss, float* d_weights,
                               void* stream) {
  if (n_rays < 0 || n_samples < 1 || n_bins < 1 || n_bins > kMaxBins || !bin_sizes || !binned || !count || !g_loss ||
      !d_weights)
    return SRF_EINVAL;
  Bins bins;
  bins.n = n_bins;
  bins.start[0] = 0;
  for (int b = 0; b < n_bins; ++b) bins.start[b + 1] = bins.start[b] + bin_sizes[b];
  if (bins.start[n_bins] != n_samples) return SRF_EINVAL;
  const int64_t n = (int64_t)n_rays * n_samples;
  if (n == 0) return SRF_OK;
  const int threads = 256;
  hipLaunchKernelGGL(entropy_bwd_kernel, dim3((unsigned)((n + threads - 1) / threads)), dim3(threads), 0,
                     as_stream(stream), n_rays, n_samples, select, bins, binned, count, g_loss, d_weights);
  return launch_status();
}
