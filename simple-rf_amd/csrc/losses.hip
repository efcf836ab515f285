// Fused training losses on the rendered rays (forward and backward) on gfx950.
//
// Reference semantics:
//   * MSE.compute_mse (src/loss_functions/MSE14.py:81-94) and SparseDepthMSE.compute_depth_loss
//     (src/loss_functions/SparseDepthMSE14.py:82-95), summed over the regularised models with their
//     weights (:29-79): srf_mse_terms_fwd / _bwd;
//   * compute_loss_nerf of AugmentationsDepthLoss (src/loss_functions/AugmentationsDepthLoss11.py:108-181)
//     and CoarseFineConsistencyLoss (src/loss_functions/CoarseFineConsistencyLoss34.py:85-164), with the
//     reprojection of src/utils/CommonUtils04.py:227-253: srf_depth_consistency_fwd / _bwd.
//
// The reference evaluates these as dozens of small launches per call (25 patch gathers per model
// pair, the masks, the means).  Here the per-ray work of a depth-consistency term (reprojection of
// both depths into the nearest view, three 5x5 ground-truth patches, both RMSEs, the gate and the
// loss maps) is one thread per ray, and every mean is a fixed-order reduction in one block, so the
// loss value does not depend on the schedule.  Selections (the reference's boolean indices_mask_*)
// are byte masks over the batch, and their counts stay on the device, so no call synchronises the
// host.
#include "common.h"

namespace srf {
namespace {

constexpr int kRedThreads = 1024;

// Fixed-order block sum (every thread returns the total).
__device__ float block_sum(float v, float* sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  float t = 0.f;
  const int nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nw; ++i) t += sh[i];
  return t;
}

// ------------------------------------------------------------------ MSE terms
constexpr int kMaxTerms = 8;  // regularised models of one loss: main and augmented, coarse and fine

// The terms of one loss, passed by value in the launch arguments (host arrays of device pointers).
struct Terms {
  int n_terms;
  const float* pred[kMaxTerms];
  float weight[kMaxTerms];
  float* d_pred[kMaxTerms];  // backward outputs (entries may be NULL)
};

// loss = sum_t weight[t] * mean over selected rays i of (mean over c of (pred_t[i,c] - target[i,c])^2);
// the count of selected rays goes to *count (the backward reads it).  A thread takes rays
// threadIdx.x + k blockDim.x; the loads of kMseRays of them are issued before their map stores, and
// CC > 0 fixes the channel count at compile time (CC = 0: c of the arguments).  The sums run in ray
// order either way.
constexpr int kMseRays = 4;

template <int CC>
__global__ __launch_bounds__(kRedThreads) void mse_terms_fwd_kernel(Terms tm, const float* target,
                                                                    const uint8_t* select, int n, int c,
                                                                    float* loss, float* term_values, float* maps,
                                                                    int* count) {
  __shared__ float sh[kRedThreads / 64];
  if (CC > 0) c = CC;
  float cnt = 0.f;
  for (int i = threadIdx.x; i < n; i += blockDim.x) cnt += (!select || select[i]) ? 1.f : 0.f;
  const float m = block_sum(cnt, sh);
  float total = 0.f;
  for (int t = 0; t < tm.n_terms; ++t) {
    const float* p = tm.pred[t];
    float s = 0.f;
    for (int i0 = threadIdx.x; i0 < n; i0 += kMseRays * blockDim.x) {
      float r[kMseRays];
      bool on[kMseRays];
#pragma unroll
      for (int u = 0; u < kMseRays; ++u) {
        const int i = i0 + u * blockDim.x;
        on[u] = i < n && (!select || select[i]);
        r[u] = 0.f;
        if (on[u]) {
          if constexpr (CC > 0) {
#pragma unroll
            for (int k = 0; k < CC; ++k) {
              const float e = __fsub_rn(p[(int64_t)i * CC + k], target[(int64_t)i * CC + k]);
              r[u] = __fadd_rn(r[u], __fmul_rn(e, e));
            }
          } else {
            for (int k = 0; k < c; ++k) {
              const float e = __fsub_rn(p[(int64_t)i * c + k], target[(int64_t)i * c + k]);
              r[u] = __fadd_rn(r[u], __fmul_rn(e, e));
            }
          }
          r[u] = r[u] / (float)c;
        }
      }
#pragma unroll
      for (int u = 0; u < kMseRays; ++u) {
        const int i = i0 + u * blockDim.x;
        if (i >= n) continue;
        if (maps) maps[(int64_t)t * n + i] = on[u] ? r[u] : 0.f;
        if (on[u]) s += r[u];
      }
    }
    const float st = block_sum(s, sh);
    const float mean = m > 0.f ? st / m : 0.f;  // an empty selection contributes 0 (MSE14.py:90)
    if (threadIdx.x == 0 && term_values) term_values[t] = mean;
    total = __fadd_rn(total, __fmul_rn(tm.weight[t], mean));
  }
  if (threadIdx.x == 0) {
    *loss = total;
    *count = (int)m;
  }
}

// d pred_t[i,c] = g * weight[t] * 2 (pred_t[i,c] - target[i,c]) / (c * count) on selected rays, 0 elsewhere.
__global__ void mse_terms_bwd_kernel(Terms tm, const float* target, const uint8_t* select, int n, int c,
                                     const int* count, const float* g_loss) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * c) return;
  const int i = (int)(e / c);
  const int m = *count;
  const bool on = (!select || select[i]) && m > 0;
  const float scale = on ? *g_loss * 2.0f / ((float)c * (float)m) : 0.f;
  for (int t = 0; t < tm.n_terms; ++t) {
    if (!tm.d_pred[t]) continue;
    tm.d_pred[t][e] = on ? tm.weight[t] * scale * __fsub_rn(tm.pred[t][e], target[e]) : 0.f;
  }
}

// ---------------------------------------------------------- depth consistency
struct DcArgs {
  int n, n_views, h, w, c, px, py;
  float rmse_threshold;
  int both_invalid_rule;
  const float* depth1;
  const float* depth2;
  const float* rays_o;
  const float* rays_d;
  const int32_t* pixel_id;  // [n, 3]: image, x, y
  const uint8_t* select;    // [n] or NULL (all)
  const float* poses;       // [n_views, 4, 4] camera-to-world
  const float* intrinsic;   // [3, 3]: the first view's (CommonUtils04.py:249)
  const float* images;      // [n_views, h, w, c]
  float* map1;              // [n] (0 where not selected)
  float* map2;
  uint8_t* flags;           // [n]: bit 0 mask1, bit 1 mask2, bit 2 selected
};

// The view whose centre is nearest to view v's (the second smallest distance, the smallest being v's
// own zero; ties go to the lower index), AugmentationsDepthLoss11.py:138-141.
__device__ int nearest_view(const DcArgs& a, int v) {
  const float* ov = a.poses + 16 * v;
  float best0 = INFINITY, best1 = INFINITY;
  int i0 = -1, i1 = -1;
  for (int u = 0; u < a.n_views; ++u) {
    const float* ou = a.poses + 16 * u;
    const float dx = __fsub_rn(ov[3], ou[3]), dy = __fsub_rn(ov[7], ou[7]), dz = __fsub_rn(ov[11], ou[11]);
    const float d = sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz)));
    if (d < best0) {
      best1 = best0;
      i1 = i0;
      best0 = d;
      i0 = u;
    } else if (d < best1) {
      best1 = d;
      i1 = u;
    }
  }
  return i1 < 0 ? i0 : i1;
}

// Pixel position of point p in view b (CommonUtils04.py:240-252): K diag(1,-1,-1) R_b^T (p - o_b),
// evaluated as ((K P) R^T) (p - o), then divided by its third component and rounded half to even.
__device__ void reproject_round(const DcArgs& a, int b, const float (&p)[3], int* xi, int* yi) {
  const float* T = a.poses + 16 * b;
  const float rel[3] = {__fsub_rn(p[0], T[3]), __fsub_rn(p[1], T[7]), __fsub_rn(p[2], T[11])};
  float kp[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    kp[i][0] = a.intrinsic[3 * i + 0];
    kp[i][1] = -a.intrinsic[3 * i + 1];
    kp[i][2] = -a.intrinsic[3 * i + 2];
  }
  float pos[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float m[3];
#pragma unroll
    for (int j = 0; j < 3; ++j)  // (K P R^T)[i][j] = sum_k (K P)[i][k] R[j][k]
      m[j] = __fadd_rn(__fadd_rn(__fmul_rn(kp[i][0], T[4 * j + 0]), __fmul_rn(kp[i][1], T[4 * j + 1])),
                       __fmul_rn(kp[i][2], T[4 * j + 2]));
    pos[i] = __fadd_rn(__fadd_rn(__fmul_rn(m[0], rel[0]), __fmul_rn(m[1], rel[1])), __fmul_rn(m[2], rel[2]));
  }
  // .round().long(): half to even; an out-of-range or non-finite position is invalid either way
  const float u = rintf(pos[0] / pos[2]), v = rintf(pos[1] / pos[2]);
  *xi = isfinite(u) ? (int)fmaxf(fminf(u, 1e9f), -1e9f) : -1000000000;
  *yi = isfinite(v) ? (int)fmaxf(fminf(v, 1e9f), -1e9f) : -1000000000;
}

// Ground-truth value at (row, col) as the reference reads it from images zero-padded by hpx rows
// and hpy columns at the far ends (F.pad(images, (0, 0, 0, hpy, 0, hpx))), where a negative index
// counts from the padded end.
__device__ __forceinline__ float gt_at(const DcArgs& a, int view, int row, int col, int ch) {
  const int hpx = a.px / 2, hpy = a.py / 2;
  if (row < 0) row += a.h + hpx;
  if (col < 0) col += a.w + hpy;
  if (row < 0 || row >= a.h || col < 0 || col >= a.w) return 0.f;
  return a.images[(((int64_t)view * a.h + row) * a.w + col) * a.c + ch];
}

// RMSE between the py x px x c patches centred at (xa, ya) of view va and at (xb, yb) of view vb
// (compute_patch_rmse, AugmentationsDepthLoss11.py:183-197).
__device__ float patch_rmse(const DcArgs& a, int va, int xa, int ya, int vb, int xb, int yb) {
  const int hpx = a.px / 2, hpy = a.py / 2;
  float s = 0.f;
  for (int i = -hpy; i <= hpy; ++i)
    for (int j = -hpx; j <= hpx; ++j)
      for (int ch = 0; ch < a.c; ++ch) {
        const float e = __fsub_rn(gt_at(a, va, ya + i, xa + j, ch), gt_at(a, vb, yb + i, xb + j, ch));
        s = __fadd_rn(s, __fmul_rn(e, e));
      }
  return sqrtf(s / (float)(a.px * a.py * a.c));
}

// gt_at of all CH channels of a pixel (CH = a.c), without branches: the address is clamped into the
// image and the values zeroed outside it, so that all loads of a patch row can be in flight before
// the first is used, and a pixel's channels are one contiguous load.
template <int CH>
__device__ __forceinline__ void gt_px_sel(const DcArgs& a, int view, int row, int col, int hpx, int hpy,
                                          float (&v)[CH]) {
  row = row < 0 ? row + a.h + hpx : row;
  col = col < 0 ? col + a.w + hpy : col;
  const bool in = row >= 0 && row < a.h && col >= 0 && col < a.w;
  const float* px = a.images + (in ? (((int64_t)view * a.h + row) * a.w + col) * CH : 0);
#pragma unroll
  for (int ch = 0; ch < CH; ++ch) v[ch] = in ? px[ch] : 0.f;
}

// patch_rmse of the patch at (xa, ya) of view va against the patches at (x1, y1) and (x2, y2) of view
// vb, for a compile-time patch shape (the reference's 5 x 5 x 3): each patch row's values of the three
// patches are loaded before any is summed, and the sums run in patch_rmse's order (the same bits).
template <int PX, int PY, int CH>
__device__ void patch_rmse2(const DcArgs& a, int va, int xa, int ya, int vb, int x1, int y1, int x2, int y2,
                            float& r1, float& r2) {
  constexpr int hpx = PX / 2, hpy = PY / 2;
  float s1 = 0.f, s2 = 0.f;
#pragma unroll
  for (int i = -hpy; i <= hpy; ++i) {
    float ga[PX][CH], g1[PX][CH], g2[PX][CH];
#pragma unroll
    for (int j = -hpx; j <= hpx; ++j) {
      gt_px_sel<CH>(a, va, ya + i, xa + j, hpx, hpy, ga[j + hpx]);
      gt_px_sel<CH>(a, vb, y1 + i, x1 + j, hpx, hpy, g1[j + hpx]);
      gt_px_sel<CH>(a, vb, y2 + i, x2 + j, hpx, hpy, g2[j + hpx]);
    }
#pragma unroll
    for (int j = 0; j < PX; ++j)
#pragma unroll
      for (int ch = 0; ch < CH; ++ch) {
        const float e1 = __fsub_rn(ga[j][ch], g1[j][ch]);
        s1 = __fadd_rn(s1, __fmul_rn(e1, e1));
        const float e2 = __fsub_rn(ga[j][ch], g2[j][ch]);
        s2 = __fadd_rn(s2, __fmul_rn(e2, e2));
      }
  }
  r1 = sqrtf(s1 / (float)(PX * PY * CH));
  r2 = sqrtf(s2 / (float)(PX * PY * CH));
}

// One thread per ray.  PX = 0: the patch shape of the arguments (patch_rmse); otherwise the
// compile-time shape PX x PY x CH (patch_rmse2).
constexpr int kDcThreads = 64;  // a train batch is a few thousand rays: spread them over many CUs

template <int PX, int PY, int CH>
__global__ __launch_bounds__(kDcThreads) void depth_consistency_kernel(DcArgs a) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.n) return;
  if (a.select && !a.select[r]) {
    a.map1[r] = 0.f;
    a.map2[r] = 0.f;
    a.flags[r] = 0;
    return;
  }
  const int va = a.pixel_id[3 * r], xa = a.pixel_id[3 * r + 1], ya = a.pixel_id[3 * r + 2];
  if (va < 0 || va >= a.n_views) {  // no such image: no patch to compare, the ray is not gated in
    a.map1[r] = 0.f;
    a.map2[r] = 0.f;
    a.flags[r] = 4;
    return;
  }
  const int vb = nearest_view(a, va);
  const float d1 = a.depth1[r], d2 = a.depth2[r];
  float p1[3], p2[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float o = a.rays_o[3 * r + k], d = a.rays_d[3 * r + k];
    p1[k] = __fadd_rn(o, __fmul_rn(d, d1));
    p2[k] = __fadd_rn(o, __fmul_rn(d, d2));
  }
  int x1, y1, x2, y2;
  reproject_round(a, vb, p1, &x1, &y1);
  reproject_round(a, vb, p2, &x2, &y2);
  const int hpx = a.px / 2, hpy = a.py / 2;
  auto inside = [&](int x, int y) { return x >= hpx && x < a.w - hpx && y >= hpy && y < a.h - hpy; };
  const bool valid_a = inside(xa, ya), valid_1 = inside(x1, y1), valid_2 = inside(x2, y2);
  const int x1c = min(max(x1, 0), a.w - 1), y1c = min(max(y1, 0), a.h - 1);
  const int x2c = min(max(x2, 0), a.w - 1), y2c = min(max(y2, 0), a.h - 1);
  float rmse1, rmse2;
  if constexpr (PX > 0) {
    patch_rmse2<PX, PY, CH>(a, va, xa, ya, vb, x1c, y1c, x2c, y2c, rmse1, rmse2);
  } else {
    rmse1 = patch_rmse(a, va, xa, ya, vb, x1c, y1c);
    rmse2 = patch_rmse(a, va, xa, ya, vb, x2c, y2c);
  }
  bool m1 = (rmse1 < rmse2 || !valid_2) && rmse1 < a.rmse_threshold && valid_1 && valid_a;
  bool m2 = (rmse2 < rmse1 || !valid_1) && rmse2 < a.rmse_threshold && valid_2 && valid_a;
  if (a.both_invalid_rule && !valid_1 && !valid_2) {
    m1 = m1 || d1 > d2;
    m2 = m2 || d2 > d1;
  }
  // compute_depth_mse(depth1, depth2.detach(), mask2) zeroes both where mask2 is false; the second
  // call's inputs alias the first's, so its map is non-zero only where both masks hold
  const float e1 = __fsub_rn(d1, d2);
  a.map1[r] = m2 ? __fmul_rn(e1, e1) : 0.f;
  const float e2 = __fsub_rn(d2, d1);
  a.map2[r] = (m1 && m2) ? __fmul_rn(e2, e2) : 0.f;
  a.flags[r] = (uint8_t)((m1 ? 1 : 0) | (m2 ? 2 : 0) | 4);
}

// loss = mean(map1) + mean(map2) over the selected rays (their count to *count).
__global__ __launch_bounds__(kRedThreads) void depth_consistency_reduce_kernel(int n, const float* map1,
                                                                               const float* map2,
                                                                               const uint8_t* flags, float* loss,
                                                                               int* count) {
  __shared__ float sh[kRedThreads / 64];
  float c = 0.f, s1 = 0.f, s2 = 0.f;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (!(flags[i] & 4)) continue;
    c += 1.f;
    s1 += map1[i];
    s2 += map2[i];
  }
  const float m = block_sum(c, sh);
  const float t1 = block_sum(s1, sh);
  const float t2 = block_sum(s2, sh);
  if (threadIdx.x == 0) {
    *loss = m > 0.f ? __fadd_rn(t1 / m, t2 / m) : 0.f;
    *count = (int)m;
  }
}

// d depth1 = g 2 (d1 - d2) / count where mask2; d depth2 = g 2 (d2 - d1) / count where mask1 and mask2.
__global__ void depth_consistency_bwd_kernel(int n, const float* depth1, const float* depth2, const uint8_t* flags,
                                             const int* count, const float* g_loss, float* d_depth1,
                                             float* d_depth2) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int m = *count;
  const uint8_t f = flags[r];
  const float s = m > 0 ? *g_loss * 2.0f / (float)m : 0.f;
  if (d_depth1) d_depth1[r] = (f & 2) ? s * __fsub_rn(depth1[r], depth2[r]) : 0.f;
  if (d_depth2) d_depth2[r] = ((f & 3) == 3) ? s * __fsub_rn(depth2[r], depth1[r]) : 0.f;
}

}  // namespace
}  // namespace srf

using namespace srf;

static bool make_terms(int n_terms, const float* const* preds, const float* weights, float* const* d_preds,
                       Terms* tm) {
  if (n_terms < 1 || n_terms > kMaxTerms || !preds || !weights) return false;
  tm->n_terms = n_terms;
  for (int t = 0; t < kMaxTerms; ++t) {
    tm->pred[t] = t < n_terms ? preds[t] : nullptr;
    tm->weight[t] = t < n_terms ? weights[t] : 0.f;
    tm->d_pred[t] = (t < n_terms && d_preds) ? d_preds[t] : nullptr;
    if (t < n_terms && !tm->pred[t]) return false;
  }
  return true;
}

extern "C" int srf_mse_terms_fwd(int n_terms, const float* const* preds, const float* weights, const float* target,
                                 const uint8_t* select, int n, int channels, float* loss, float* term_values,
                                 float* maps, int* count, void* stream) {
  Terms tm;
  if (!make_terms(n_terms, preds, weights, nullptr, &tm) || !target || n < 0 || channels < 1 || !loss || !count)
    return SRF_EINVAL;
  if (channels == 3)
    hipLaunchKernelGGL(mse_terms_fwd_kernel<3>, dim3(1), dim3(kRedThreads), 0, as_stream(stream), tm, target,
                       select, n, channels, loss, term_values, maps, count);
  else
    hipLaunchKernelGGL(mse_terms_fwd_kernel<0>, dim3(1), dim3(kRedThreads), 0, as_stream(stream), tm, target,
                       select, n, channels, loss, term_values, maps, count);
  return launch_status();
}

extern "C" int srf_mse_terms_bwd(int n_terms, const float* const* preds, const float* weights, const float* target,
                                 const uint8_t* select, int n, int channels, const int* count, const float* g_loss,
                                 float* const* d_preds, void* stream) {
  Terms tm;
  if (!make_terms(n_terms, preds, weights, d_preds, &tm) || !d_preds || !target || n < 0 || channels < 1 || !count ||
      !g_loss)
    return SRF_EINVAL;
  const int64_t total = (int64_t)n * channels;
  if (total == 0) return SRF_OK;
  const int threads = 256;
  hipLaunchKernelGGL(mse_terms_bwd_kernel, dim3((unsigned)((total + threads - 1) / threads)), dim3(threads), 0,
                     as_stream(stream), tm, target, select, n, channels, count, g_loss);
  return launch_status();
}

extern "C" int srf_depth_consistency_fwd(int n, const float* depth1, const float* depth2, const float* rays_o,
                                         const float* rays_d, const int32_t* pixel_id, const uint8_t* select,
                                         int n_views, const float* poses, const float* intrinsic,
                                         const float* images, int height, int width, int channels, int patch_x,
                                         int patch_y, float rmse_threshold, int both_invalid_rule, float* map1,
                                         float* map2, uint8_t* flags, float* loss, int* count, void* stream) {
  if (n < 0 || n_views < 1 || height < 1 || width < 1 || channels < 1 || patch_x < 1 || patch_y < 1 ||
      patch_x > 31 || patch_y > 31)
    return SRF_EINVAL;
  if (!depth1 || !depth2 || !rays_o || !rays_d || !pixel_id || !poses || !intrinsic || !images || !map1 || !map2 ||
      !flags || !loss || !count)
    return SRF_EINVAL;
  DcArgs a;
  a.n = n;
  a.n_views = n_views;
  a.h = height;
  a.w = width;
  a.c = channels;
  a.px = patch_x;
  a.py = patch_y;
  a.rmse_threshold = rmse_threshold;
  a.both_invalid_rule = both_invalid_rule;
  a.depth1 = depth1;
  a.depth2 = depth2;
  a.rays_o = rays_o;
  a.rays_d = rays_d;
  a.pixel_id = pixel_id;
  a.select = select;
  a.poses = poses;
  a.intrinsic = intrinsic;
  a.images = images;
  a.map1 = map1;
  a.map2 = map2;
  a.flags = flags;
  hipStream_t s = as_stream(stream);
  if (n > 0) {
    const int threads = kDcThreads;
    if (patch_x == 5 && patch_y == 5 && channels == 3)
      hipLaunchKernelGGL((depth_consistency_kernel<5, 5, 3>), dim3((n + threads - 1) / threads), dim3(threads), 0, s,
                         a);
    else
      hipLaunchKernelGGL((depth_consistency_kernel<0, 0, 0>), dim3((n + threads - 1) / threads), dim3(threads), 0, s,
                         a);
    const int st = launch_status();
    if (st != SRF_OK) return st;
  }
  hipLaunchKernelGGL(depth_consistency_reduce_kernel, dim3(1), dim3(kRedThreads), 0, s, n, map1, map2, flags, loss,
                     count);
  return launch_status();
}

extern "C" int srf_depth_consistency_bwd(int n, const float* depth1, const float* depth2, const uint8_t* flags,
                                         const int* count, const float* g_loss, float* d_depth1, float* d_depth2,
                                         void* stream) {
  if (n < 0 || !depth1 || !depth2 || !flags || !count || !g_loss) return SRF_EINVAL;
  if (n == 0) return SRF_OK;
  const int threads = 256;
  hipLaunchKernelGGL(depth_consistency_bwd_kernel, dim3((n + threads - 1) / threads), dim3(threads), 0,
                     as_stream(stream), n, depth1, depth2, flags, count, g_loss, d_depth1, d_depth2);
  return launch_status();
}
