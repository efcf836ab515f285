// Camera rays of a batch of pixels on gfx950, in one pass: pinhole rays, view directions and the
// forward-facing NDC warp.
//
// Reference semantics: get_rays_tr (nerf_synthetic=False), get_view_dirs_tr and
// get_ndc_rays_tr (src/utils/CommonUtils04.py:73-95, :147-149, :120-138).  The reference runs
// them as some forty small PyTorch ops per batch plus a batched LU inverse of the intrinsics;
// here one thread evaluates every output of its ray.
//
// Numerics: each arithmetic step the reference performs as a separate op is a separately
// rounded FP32 operation in the same order (the file is compiled with -ffp-contract=off), so the
// NDC warp and the rotation reproduce the reference's elementwise results.  The inverse
// intrinsics come from the adjugate (3 x 3 closed form) instead of an LU factorisation, which
// can differ from torch.linalg.inv in the last bit.
//
// No backward: the training step of a model that learns its camera differentiates the PyTorch
// restatement instead (the gradient reduces over all rays of a frame).
#include "common.h"

namespace srf {
namespace {

__global__ __launch_bounds__(256) void rays_kernel(int n, const int32_t* __restrict__ pixel_id, int n_frames,
                                                   const float* __restrict__ K, const float* __restrict__ E,
                                                   int mip, int ndc, float h, float w, float near,
                                                   float* __restrict__ rays_o, float* __restrict__ rays_d,
                                                   float* __restrict__ view_dirs, float* __restrict__ o_ndc,
                                                   float* __restrict__ d_ndc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int f = pixel_id[3 * i];
  if (f < 0 || f >= n_frames) {
    const float q = __builtin_nanf("");
    for (int c = 0; c < 3; ++c) {
      rays_o[3 * i + c] = rays_d[3 * i + c] = view_dirs[3 * i + c] = q;
      if (ndc) o_ndc[3 * i + c] = d_ndc[3 * i + c] = q;
    }
    return;
  }
  float x = (float)pixel_id[3 * i + 1], y = (float)pixel_id[3 * i + 2];
  if (mip) {
    x = __fadd_rn(x, 0.5f);
    y = __fadd_rn(y, 0.5f);
  }
  // inverse intrinsics (adjugate / determinant)
  const float* k = K + 9 * f;
  const float a = k[0], b = k[1], c = k[2], d = k[3], e = k[4], g = k[5], p = k[6], q = k[7], r = k[8];
  const float c00 = __fsub_rn(__fmul_rn(e, r), __fmul_rn(g, q));
  const float c01 = __fsub_rn(__fmul_rn(c, q), __fmul_rn(b, r));
  const float c02 = __fsub_rn(__fmul_rn(b, g), __fmul_rn(c, e));
  const float c10 = __fsub_rn(__fmul_rn(g, p), __fmul_rn(d, r));
  const float c11 = __fsub_rn(__fmul_rn(a, r), __fmul_rn(c, p));
  const float c12 = __fsub_rn(__fmul_rn(c, d), __fmul_rn(a, g));
  const float c20 = __fsub_rn(__fmul_rn(d, q), __fmul_rn(e, p));
  const float c21 = __fsub_rn(__fmul_rn(b, p), __fmul_rn(a, q));
  const float c22 = __fsub_rn(__fmul_rn(a, e), __fmul_rn(b, d));
  const float det = __fadd_rn(__fadd_rn(__fmul_rn(a, c00), __fmul_rn(b, c10)), __fmul_rn(c, c20));
  const float inv_det = __fdiv_rn(1.0f, det);
  const float ki[3][3] = {{__fmul_rn(c00, inv_det), __fmul_rn(c01, inv_det), __fmul_rn(c02, inv_det)},
                          {__fmul_rn(c10, inv_det), __fmul_rn(c11, inv_det), __fmul_rn(c12, inv_det)},
                          {__fmul_rn(c20, inv_det), __fmul_rn(c21, inv_det), __fmul_rn(c22, inv_det)}};
  // dirs = K^-1 [x, y, 1], then y and z negated
  float dir[3];
  for (int t = 0; t < 3; ++t) dir[t] = __fadd_rn(__fadd_rn(__fmul_rn(ki[t][0], x), __fmul_rn(ki[t][1], y)), ki[t][2]);
  dir[1] = -dir[1];
  dir[2] = -dir[2];
  // rays_d = sum(dirs[None] * c2w[:3, :3], -1), rays_o = c2w[:3, 3]
  const float* m = E + 16 * f;
  float o[3], dv[3];
  for (int t = 0; t < 3; ++t) {
    dv[t] = __fadd_rn(__fadd_rn(__fmul_rn(dir[0], m[4 * t + 0]), __fmul_rn(dir[1], m[4 * t + 1])),
                      __fmul_rn(dir[2], m[4 * t + 2]));
    o[t] = m[4 * t + 3];
  }
  const float nrm = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(dv[0], dv[0]), __fmul_rn(dv[1], dv[1])),
                                         __fmul_rn(dv[2], dv[2])));
  for (int t = 0; t < 3; ++t) {
    rays_o[3 * i + t] = o[t];
    rays_d[3 * i + t] = dv[t];
    view_dirs[3 * i + t] = __fdiv_rn(dv[t], nrm);
  }
  if (!ndc) return;
  // get_ndc_rays_tr: shift the origins to the near plane, then project
  const float fx = a, fy = e;
  const float tn = __fdiv_rn(-__fadd_rn(near, o[2]), dv[2]);
  float os[3];
  for (int t = 0; t < 3; ++t) os[t] = __fadd_rn(o[t], __fmul_rn(tn, dv[t]));
  const float ax = __fdiv_rn(-1.0f, __fdiv_rn(w, __fmul_rn(2.0f, fx)));
  const float ay = __fdiv_rn(-1.0f, __fdiv_rn(h, __fmul_rn(2.0f, fy)));
  const float ox_oz = __fdiv_rn(os[0], os[2]), oy_oz = __fdiv_rn(os[1], os[2]);
  o_ndc[3 * i + 0] = __fdiv_rn(__fmul_rn(ax, os[0]), os[2]);
  o_ndc[3 * i + 1] = __fdiv_rn(__fmul_rn(ay, os[1]), os[2]);
  o_ndc[3 * i + 2] = __fadd_rn(1.0f, __fdiv_rn(__fmul_rn(2.0f, near), os[2]));
  d_ndc[3 * i + 0] = __fmul_rn(ax, __fsub_rn(__fdiv_rn(dv[0], dv[2]), ox_oz));
  d_ndc[3 * i + 1] = __fmul_rn(ay, __fsub_rn(__fdiv_rn(dv[1], dv[2]), oy_oz));
  d_ndc[3 * i + 2] = __fdiv_rn(__fmul_rn(-2.0f, near), os[2]);
}

}  // namespace
}  // namespace srf

extern "C" int srf_rays(int n, const int32_t* pixel_id, int n_frames, const float* intrinsics,
                        const float* extrinsics, int mip_nerf, int ndc, int height, int width, float near,
                        float* rays_o, float* rays_d, float* view_dirs, float* rays_o_ndc, float* rays_d_ndc,
                        void* stream) {
  if (n < 0 || n_frames <= 0) return SRF_EINVAL;
  if (n == 0) return SRF_OK;
  if (!pixel_id || !intrinsics || !extrinsics || !rays_o || !rays_d || !view_dirs) return SRF_EINVAL;
  if (ndc && (!rays_o_ndc || !rays_d_ndc || height <= 0 || width <= 0)) return SRF_EINVAL;
  const int threads = 256;
  hipLaunchKernelGGL(srf::rays_kernel, dim3((n + threads - 1) / threads), dim3(threads), 0,
                     srf::as_stream(stream), n, pixel_id, n_frames, intrinsics, extrinsics, mip_nerf, ndc,
                     (float)height, (float)width, near, rays_o, rays_d, view_dirs, rays_o_ndc, rays_d_ndc);
  return srf::launch_status();
}
