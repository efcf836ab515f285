/*
 * srf_hip.h -- C-ABI of the MI355X (gfx950) kernels for Simple-RF's rendering hot path.
 *
 * The reference (NagabhushanSN95/Simple-RF) has no FFI: its hot path is PyTorch ops
 * called from the model methods cited below.  Each entry point here replaces the
 * reference function named in its comment, with the same argument meaning; the
 * Python drop-in (simple-rf_amd/srf_hip, simple-rf_amd/models) binds them with ctypes.
 *
 * Conventions
 *   - Return value: SRF_OK (0) or an SRF_E* code.  Nothing throws across the ABI.
 *   - Every pointer is a device pointer owned by the caller; the library never
 *     allocates.  Tensors are contiguous row-major float32 unless stated otherwise.
 *   - `stream` is a hipStream_t (passed as void*); all work is stream-ordered, there
 *     are no host synchronisations, so every entry point is hipGraph-capturable.
 *   - Entry points are re-entrant.  The only process-global state is the table of
 *     launch-schedule knobs (srf_tune_set), read atomically when a kernel is launched.
 */
#ifndef SRF_HIP_H_
#define SRF_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SRF_ABI_VERSION 3

enum srf_status {
  SRF_OK = 0,
  SRF_EHIP = 1,        /* a HIP launch/runtime error */
  SRF_EINVAL = 2,      /* bad size, null required pointer, unsupported option */
  SRF_EUNSUPPORTED = 3, /* shape outside what the specialised kernels support */
  SRF_ETOOLARGE = 4     /* input beyond the fixed scratch of an allocation-free entry point */
};

/* ABI version of the loaded library (SRF_ABI_VERSION). */
int srf_abi_version(void);

/* Human-readable text for a status code. */
const char* srf_status_string(int status);

/* ------------------------------------------------------------------ sampling (a) */

/* Camera rays of n pixels in one pass.  Replaces get_rays_tr (nerf_synthetic=False),
 * get_view_dirs_tr and get_ndc_rays_tr (src/utils/CommonUtils04.py:73-95, :147-149, :120-138)
 * as called by SimpleNeRF.render_rays (src/models/SimpleNeRF17.py:159-199).
 *   pixel_id [n,3] int32 (frame, x, y); intrinsics [n_frames,3,3] and camera-to-world extrinsics
 *   [n_frames,4,4] per frame (FP32, row-major); mip_nerf: pixel centres at x + 0.5, y + 0.5.
 * Outputs rays_o, rays_d, view_dirs [n,3]; with ndc, rays_o_ndc / rays_d_ndc [n,3] (resolution
 * height x width and near as get_ndc_rays_tr; NULL otherwise).  A frame index outside
 * [0, n_frames) gives NaN rays.  No backward: a model that learns its camera differentiates the
 * PyTorch restatement. */
int srf_rays(int n, const int32_t* pixel_id, int n_frames, const float* intrinsics,
             const float* extrinsics, int mip_nerf, int ndc, int height, int width, float near,
             float* rays_o, float* rays_d, float* view_dirs, float* rays_o_ndc, float* rays_d_ndc,
             void* stream);

/* Coarse depths z[n_rays, n_samples].
 * Replaces SimpleNeRF.get_z_vals_coarse (src/models/SimpleNeRF17.py:330-358) and the NDC
 * branch of SimpleTensoRF.get_z_vals_coarse (src/models/SimpleTensoRF09.py:364-381).
 * t = torch.linspace(0,1,n_samples); z = near(1-t)+far t (lindisp: 1/(1/near(1-t)+1/far t));
 * if u != NULL (training, perturb=True) z = lower + (upper-lower) u with u[n_rays,n_samples]
 * taking the place of the CPU torch.rand draw at :355. */
int srf_stratified_samples(int n_rays, int n_samples, float near, float far, int lindisp,
                           const float* u, float* z, void* stream);

/* Inverse-CDF sampling of n_fine depths per ray.
 * Replaces SimpleNeRF.sample_pdf (src/models/SimpleNeRF17.py:384-417; identical copy at
 * src/models/SimpleTensoRF09.py:427-460).
 *   bins    [n_rays, n_bins]        (z_vals_mid)
 *   weights [n_rays, n_bins-1]      (weights[..., 1:-1]); ignored when cdf_in != NULL
 *   cdf_in  [n_rays, n_bins] or NULL: a precomputed CDF, bypassing :386-390
 *   u       [n_rays, n_fine] or NULL: NULL is det=True (u = torch.linspace(0,1,n_fine))
 * Outputs: samples [n_rays, n_fine]; inds [n_rays, n_fine] int64 or NULL
 * (= torch.searchsorted(cdf, u, right=True)); cdf_out [n_rays, n_bins] or NULL.
 * n_bins <= 1024. */
int srf_sample_pdf(int n_rays, int n_bins, int n_fine, const float* bins, const float* weights,
                   const float* cdf_in, const float* u, float* samples, int64_t* inds,
                   float* cdf_out, void* stream);

/* Hierarchical depths: sample_pdf on the coarse weights followed by the sorted merge
 * torch.sort(torch.cat([z_coarse, z_fine], -1), -1).
 * Replaces SimpleNeRF.get_z_vals_fine (src/models/SimpleNeRF17.py:360-371; TensoRF copy
 * src/models/SimpleTensoRF09.py:403-414).
 *   z_coarse [n_rays, n_coarse], weights_coarse [n_rays, n_coarse], u [n_rays, n_fine] or NULL.
 * Outputs: z_merged [n_rays, n_coarse + n_fine] (sorted); z_fine [n_rays, n_fine] or NULL.
 * n_coarse + n_fine <= 1024. */
int srf_fine_z_vals(int n_rays, int n_coarse, int n_fine, const float* z_coarse,
                    const float* weights_coarse, const float* u, float* z_merged, float* z_fine,
                    void* stream);

/* ----------------------------------------------------------------- compositing (d) */

/* Alpha compositing of one ray per wavefront with a log-space FP32 transmittance scan.
 * Replaces SimpleNeRF.volume_rendering + convert_depth_from_ndc
 * (src/models/SimpleNeRF17.py:486-558) and LowRankTensor.get_volume_rendering_weights +
 * volume_render (src/models/SimpleTensoRF09.py:767-819; distance_scale at :787).
 *   Dense layout when ray_offsets == NULL: sample k of ray r at r*n_samples + k.
 *   Ragged layout: ray r owns samples [ray_offsets[r], ray_offsets[r+1]) (int32 CSR).
 *   sigma [P], rgb [P,3], z [P] (NDC depths when ndc != 0, metric otherwise);
 *   rays_o, rays_d [n_rays,3] (world, used for the metric depth and, when !ndc, delta);
 *   rays_d_ndc [n_rays,3] (NDC delta norm, required when ndc != 0).
 *   distance_scale multiplies sigma*delta (1 for NeRF, 25 for TensoRF).
 *   rgb may be NULL (TensoRF composites colours afterwards with srf_volume_render_fwd);
 *   rgb_map is then not written and may be NULL too.
 * Outputs (per ray): rgb_map [n_rays,3], acc, depth, depth_var, and when ndc != 0
 * depth_ndc, depth_var_ndc (may be NULL otherwise); (per sample, each may be NULL):
 * alpha, visibility, weights [P]. */
int srf_composite_fwd(int n_rays, int n_samples, const int32_t* ray_offsets, const float* sigma,
                      const float* rgb, const float* z, const float* rays_o, const float* rays_d,
                      const float* rays_d_ndc, int ndc, int white_bkgd, float distance_scale,
                      float* rgb_map, float* acc, float* depth, float* depth_var, float* depth_ndc,
                      float* depth_var_ndc, float* alpha, float* visibility, float* weights,
                      void* stream);

/* Backward of srf_composite_fwd: a forward scan for the chunk transmittances followed by
 * a reverse (suffix) scan, recomputing alpha/T from sigma and z instead of storing them.
 *   acc, depth, depth_ndc: the forward's per-ray outputs (depth_ndc may be NULL if !ndc).
 *   Upstream gradients, each may be NULL (= zero): g_rgb [n_rays,3], g_acc, g_depth,
 *   g_depth_var, g_depth_ndc, g_depth_var_ndc [n_rays], g_alpha, g_visibility, g_weights [P].
 * Outputs: d_sigma [P], d_rgb [P,3] (overwritten; d_rgb may be NULL, and rgb NULL means
 * zero colours as in the forward). */
int srf_composite_bwd(int n_rays, int n_samples, const int32_t* ray_offsets, const float* sigma,
                      const float* rgb, const float* z, const float* rays_o, const float* rays_d,
                      const float* rays_d_ndc, int ndc, int white_bkgd, float distance_scale,
                      const float* acc, const float* depth, const float* depth_ndc,
                      const float* g_rgb, const float* g_acc, const float* g_depth,
                      const float* g_depth_var, const float* g_depth_ndc,
                      const float* g_depth_var_ndc, const float* g_alpha,
                      const float* g_visibility, const float* g_weights, float* d_sigma,
                      float* d_rgb, void* stream);

/* ------------------------------------------------------------------- PE + MLP (b) */

/* Structure of one NeRF MLP (class MLP, src/models/SimpleNeRF17.py:616-785).  The kernels are
 * specialised to the shipped geometry: 8 trunk layers of width 256 with the skip after layer 4,
 * one view layer of width 128 (src/NerfLlffTrainerTester10.py:293-358).
 *   trunk_pe:       xyz-PE features fed to the trunk: 63 (L=10) for the main and views-augmented
 *                   models, 21 (points_sigma_positional_encoding_degree=3) for the
 *                   points-augmented model, whose remaining 42 features join the view head
 *                   (:634-636, :700-703);
 *   pts_pe_degree:  L of the xyz PE (10);
 *   dirs_pe_degree: L of the view-direction PE (4);
 *   view_dependent: 1 -> feature_linear + views_linears + views_output_linear (rgb = sigmoid of
 *                   the view head); 0 -> 4-channel pts_output_linear (views-augmented model);
 *   sigma_softplus: 0 -> relu(raw + noise) (reference, :738-743); 1 -> softplus(raw + noise);
 *   prec:           SRF_PREC_F32 / SRF_PREC_BF16 / SRF_PREC_F16 / SRF_PREC_BF16_SPLIT operands of
 *                   the MFMA tiles (accumulation is always FP32).  SRF_PREC_F16 covers the
 *                   inference forward only (srf_mlp_forward with workspace NULL); its training
 *                   entry points return SRF_EUNSUPPORTED.  SRF_PREC_BF16_SPLIT keeps FP32
 *                   activations and stored tiles (the SRF_PREC_F32 workspace) and multiplies on
 *                   BF16 MFMAs over split operands x = x_0 + x_1 (+ x_2), each plane the BF16
 *                   rounding of what the previous planes leave: three planes (six plane
 *                   products, FP32-level rounding) in the forward, two (three products) in the
 *                   backward chain and the weight gradients. */
typedef struct srf_mlp_desc {
  int trunk_pe;
  int pts_pe_degree;
  int dirs_pe_degree;
  int view_dependent;
  int sigma_softplus;
  int prec;
} srf_mlp_desc;

#define SRF_PREC_F32 0
#define SRF_PREC_BF16 1
#define SRF_PREC_F16 2
#define SRF_PREC_BF16_SPLIT 3

/* Reference parameter order for srf_mlp_pack / srf_mlp_backward (12 entries; NULL where the
 * variant has no such layer): pts_linears.0..7, feature_linear, pts_output_linear,
 * views_linears.0, views_output_linear.  Each weight is [out, in] row-major FP32 as in
 * torch.nn.Linear (src/models/SimpleNeRF17.py:644-666). */
#define SRF_MLP_NUM_PARAMS 12

/* Bytes of the packed operand image built by srf_mlp_pack (0 for an unsupported desc). */
size_t srf_mlp_packed_bytes(const srf_mlp_desc* desc);

/* Repacks the FP32 reference parameters into MFMA operand order (forward A tiles, transposed
 * backward tiles and padded biases).  weights/biases: host arrays of SRF_MLP_NUM_PARAMS device
 * pointers.  Must be re-run after every parameter update. */
int srf_mlp_pack(const srf_mlp_desc* desc, const float* const* weights, const float* const* biases,
                 void* packed, void* stream);

/* Bytes of the training workspace for n_points (saved activations, ReLU bit masks and
 * activation gradients, all feature-major).  Training launches (forward with a workspace,
 * backward) take up to about 3.9 M points (FP32, split BF16) or 7.8 M points (16-bit operands) per call, so
 * that each workspace array stays below 4 GB; larger calls return SRF_EINVAL. */
size_t srf_mlp_workspace_bytes(const srf_mlp_desc* desc, int n_points);

/* Fused PE + MLP forward of n_points points.
 * Replaces MLP.forward / run_network / batchify (src/models/SimpleNeRF17.py:419-484, :696-785).
 *   pts [n,3]; dirs [n,3] (view directions, needed when view_dependent);
 *   noise [n] or NULL: already scaled sigma noise (randn * raw_noise_std, :740), NULL = none;
 * Outputs: sigma [n] and rgb [n,3] (FP32).  workspace: NULL for inference, else the training
 * workspace (srf_mlp_workspace_bytes) that srf_mlp_backward reads. */
int srf_mlp_forward(const srf_mlp_desc* desc, const void* packed, int n_points, const float* pts,
                    const float* dirs, const float* noise, float* sigma, float* rgb,
                    void* workspace, void* stream);

/* srf_mlp_forward with the sample points generated in the kernel instead of read from a
 * materialised [n_rays, n_samples, 3] array.  Replaces the point generation
 * pts = rays_o[..., None, :] + rays_d[..., None, :] * z_vals[..., :, None]
 * (src/models/SimpleNeRF17.py:213-215, :266-268; the product and the sum are separately rounded
 * FP32 operations, so the points equal the reference's) and the view-direction expansion
 * viewdirs[:, None].expand(pts.shape) of run_network (:429-433), then runs the same forward.
 *   rays_o, rays_d [n_rays,3]; z_vals [n_rays, n_samples]; view_dirs [n_rays,3] (per ray, needed
 *   when view_dependent); noise [n_rays * n_samples] or NULL.
 * Outputs sigma [n_rays * n_samples], rgb [n_rays * n_samples, 3]; point p belongs to ray
 * p / n_samples.  The workspace is the one of n_rays * n_samples points, and srf_mlp_backward /
 * _backward_data / _weight_grads take it with n_points = n_rays * n_samples. */
int srf_mlp_forward_rays(const srf_mlp_desc* desc, const void* packed, int n_rays, int n_samples,
                         const float* rays_o, const float* rays_d, const float* z_vals,
                         const float* view_dirs, const float* noise, float* sigma, float* rgb,
                         void* workspace, void* stream);

/* Backward of srf_mlp_forward: d_sigma [n], d_rgb [n,3] upstream gradients; sigma, rgb the
 * forward outputs.  ACCUMULATES (+=) the FP32 parameter gradients into d_weights / d_biases
 * (host arrays of SRF_MLP_NUM_PARAMS device pointers in the reference order and shapes). */
int srf_mlp_backward(const srf_mlp_desc* desc, const void* packed, int n_points,
                     const float* sigma, const float* rgb, const float* d_sigma,
                     const float* d_rgb, void* workspace, float* const* d_weights,
                     float* const* d_biases, void* stream);

/* The two halves of srf_mlp_backward, for callers that schedule or time them separately:
 * srf_mlp_backward_data runs the backward chain (writes every layer's pre-activation gradient
 * into the workspace); srf_mlp_weight_grads then reduces them against the saved layer inputs
 * into d_weights / d_biases (accumulating, as srf_mlp_backward does). */
int srf_mlp_backward_data(const srf_mlp_desc* desc, const void* packed, int n_points,
                          const float* sigma, const float* rgb, const float* d_sigma,
                          const float* d_rgb, void* workspace, void* stream);
int srf_mlp_weight_grads(const srf_mlp_desc* desc, int n_points, void* workspace,
                         float* const* d_weights, float* const* d_biases, void* stream);

/* ------------------------------------------------------- TensoRF VM factors (c) */

/* Density predictors of LowRankTensor.get_density_predictor (src/models/SimpleTensoRF09.py:671-679). */
#define SRF_DENSITY_RELU 0
#define SRF_DENSITY_SOFTPLUS 1 /* softplus(x + offset), beta 1, threshold 20 */

/* The vector-matrix factors of one VmDecomposedTensor field (density or appearance;
 * src/models/SimpleTensoRF09.py:1126-1165).  Plane p is the reference's matrices_*[p] of shape
 * [1, n_comp[p], plane_h[p], plane_w[p]] and line p its vectors_*[p] of shape
 * [1, n_comp[p], line_len[p], 1], stored texel-major: plane p as [plane_h][plane_w][n_comp]
 * and line p as [line_len][n_comp] (the component axis innermost, 16-byte aligned bases).
 * Plane p is sampled at (x = pts[a0], y = pts[a1]) with (a0, a1) = (0,1), (0,2), (1,2) and
 * line p at pts[2 - p] (matrix_axes / vector_axes at :1131-1132).  n_comp[p] % 4 == 0 (else
 * SRF_EUNSUPPORTED) and sum n_comp <= 256. */
typedef struct srf_vm_factors {
  int n_comp[3];
  int plane_h[3];
  int plane_w[3];
  int line_len[3];
  const float* planes[3];
  const float* lines[3];
} srf_vm_factors;

/* Occupancy volume of AlphaGridMask (src/models/SimpleTensoRF09.py:1323-1349): float
 * [depth][height][width] (z, y, x), with its own bounding box {lo xyz, hi xyz} (host floats). */
typedef struct srf_alpha_grid {
  const float* volume; /* NULL = no alpha mask */
  int depth, height, width;
  float bbox[6];
} srf_alpha_grid;

/* Validity, normalisation and density of n_points samples: valid = pts inside bbox (host
 * floats {lo xyz, hi xyz}) and, with an alpha grid, its trilinear occupancy > 0; the density
 * feature sum_p sum_c plane_p * line_p on valid samples, then the predictor; invalid samples
 * get sigma 0.  Replaces LowRankTensor.forward up to the density (src/models/SimpleTensoRF09.py:
 * 701-721), normalize_points (:763-765), AlphaGridMask.sample_alpha (:1342-1345) and
 * VmDecomposedTensor.get_volume_density (:1214-1239).
 * pts [n,3]; outputs sigma [n]; feat [n] (pre-activation, 0 when invalid) and valid [n]
 * (uint8) may be NULL and are what srf_vm_density_bwd needs. */
int srf_vm_density_fwd(const srf_vm_factors* factors, const float* bbox,
                       const srf_alpha_grid* alpha, int64_t n_points, const float* pts,
                       int predictor, float offset, float* sigma, float* feat, uint8_t* valid,
                       void* stream);

/* Bytes of the workspace of the cell-sorted factor backwards (srf_vm_density_bwd,
 * srf_vm_appearance_bwd) for n_rows rows (samples or appearance rows, the capacity). */
size_t srf_vm_bwd_workspace_bytes(const srf_vm_factors* factors, int64_t n_rows);

/* Backward of srf_vm_density_fwd into the factor gradients d_planes[3] / d_lines[3] (same
 * texel-major layout as the factors; accumulated with FP32 atomics, so the caller zeroes them
 * first).  Coordinates carry no gradient (the reference detaches them, :1224-1228) -- the
 * scatter of grid_sample's backward.  With a workspace (srf_vm_bwd_workspace_bytes) the
 * contributing samples are first ordered by plane cell, so that each texel is flushed about once
 * per run of samples and the line gradients are summed in LDS; workspace NULL walks the samples in
 * their own order (more atomics).  Both give the same sums up to the order of FP32 additions. */
int srf_vm_density_bwd(const srf_vm_factors* factors, float* const* d_planes,
                       float* const* d_lines, const float* bbox, int64_t n_points,
                       const float* pts, const uint8_t* valid, const float* feat, int predictor,
                       float offset, const float* d_sigma, void* workspace, void* stream);

/* Appearance features of the listed samples: features[j, :] = cat_p(plane_p * line_p) at sample
 * idx[j] (idx NULL: sample j), [n_rows, sum n_comp] row-major.  Replaces the gather part of
 * VmDecomposedTensor.get_color (src/models/SimpleTensoRF09.py:1241-1263, before
 * basis_matrix_color).  n_rows is the capacity of idx and features; row_count (device int32,
 * NULL: n_rows) is the number of rows processed, e.g. the count srf_threshold_compact wrote, so
 * that a caller needs no host synchronisation (the reference's boolean indexing at :1249 has
 * one). */
int srf_vm_appearance_fwd(const srf_vm_factors* factors, const float* bbox, int64_t n_rows,
                          const int32_t* row_count, const int32_t* idx, const float* pts,
                          float* features, void* stream);

/* Backward of srf_vm_appearance_fwd: d_features [n_rows, sum n_comp] (rows below the count)
 * scattered into d_planes / d_lines with FP32 atomics (accumulating); workspace as for
 * srf_vm_density_bwd. */
int srf_vm_appearance_bwd(const srf_vm_factors* factors, float* const* d_planes,
                          float* const* d_lines, const float* bbox, int64_t n_rows,
                          const int32_t* row_count, const int32_t* idx, const float* pts,
                          const float* d_features, void* workspace, void* stream);

/* The CANDECOMP/PARAFAC factors of one CpDecomposedTensor field (src/models/SimpleTensoRF09.py:
 * 964-1127): three lines of the same rank n_comp, line i being the reference's vectors_*[i] of
 * shape [1, n_comp, line_len[i], 1] (built at :989-996), stored texel-major as
 * [line_len[i]][n_comp] (16-byte aligned bases).  Line i is sampled at pts[2 - i] (vector_axes
 * at :968).  n_comp % 4 == 0 and n_comp <= 256. */
typedef struct srf_cp_factors {
  int n_comp;
  int line_len[3];
  const float* lines[3];
} srf_cp_factors;

/* srf_vm_density_fwd for the CP decomposition: the density feature is
 * sum_r line_0[r] * line_1[r] * line_2[r] (CpDecomposedTensor.get_volume_density,
 * src/models/SimpleTensoRF09.py:1043-1062), with the same validity, alpha mask, predictor and
 * outputs. */
int srf_cp_density_fwd(const srf_cp_factors* factors, const float* bbox,
                       const srf_alpha_grid* alpha, int64_t n_points, const float* pts,
                       int predictor, float offset, float* sigma, float* feat, uint8_t* valid,
                       void* stream);

/* Backward of srf_cp_density_fwd into d_lines[3] (texel-major, FP32 atomics, accumulating). */
int srf_cp_density_bwd(const srf_cp_factors* factors, float* const* d_lines, const float* bbox,
                       int64_t n_points, const float* pts, const uint8_t* valid, const float* feat,
                       int predictor, float offset, const float* d_sigma, void* stream);

/* CP appearance features: features[j, r] = line_0[r] * line_1[r] * line_2[r] at sample idx[j]
 * (idx NULL: sample j), [n_rows, n_comp] row-major.  Replaces the gather part of
 * CpDecomposedTensor.get_color (src/models/SimpleTensoRF09.py:1064-1078, before
 * basis_matrix_color).  n_rows and row_count as for srf_vm_appearance_fwd. */
int srf_cp_appearance_fwd(const srf_cp_factors* factors, const float* bbox, int64_t n_rows,
                          const int32_t* row_count, const int32_t* idx, const float* pts,
                          float* features, void* stream);

/* Backward of srf_cp_appearance_fwd: d_features [n_rows, n_comp] (rows below the count)
 * scattered into d_lines[3] with FP32 atomics (accumulating). */
int srf_cp_appearance_bwd(const srf_cp_factors* factors, float* const* d_lines, const float* bbox,
                          int64_t n_rows, const int32_t* row_count, const int32_t* idx,
                          const float* pts, const float* d_features, void* stream);

/* Trilinear occupancy value of each point (AlphaGridMask.sample_alpha,
 * src/models/SimpleTensoRF09.py:1342-1345): out [n_points]. */
int srf_alpha_mask_sample(const srf_alpha_grid* alpha, int64_t n_points, const float* pts,
                          float* out, void* stream);

/* Indices (ascending, int32) of the entries with values[i] > threshold, and their number in
 * *count (device int32); no host synchronisation.  Replaces the surface test
 * weights > ray_marching_weight_threshold and the boolean indexing that follows it
 * (src/models/SimpleTensoRF09.py:726, :1249).  idx must hold n entries; workspace holds
 * srf_compact_workspace_bytes(n) bytes. */
size_t srf_compact_workspace_bytes(int64_t n);
int srf_threshold_compact(int64_t n, const float* values, float threshold, int32_t* idx,
                          int32_t* count, void* workspace, void* stream);

/* srf_threshold_compact without a workspace argument and without any allocation: the per-tile
 * counts live in a fixed device array of the library (1 MiB per GPU, held until the process
 * ends), so n <= SRF_COMPACT_NOALLOC_MAX (else SRF_ETOOLARGE).  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Call it with the stream's device current (else
 * SRF_EINVAL).  Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (on different streams, or from different host threads) race and must be serialised by the
 * caller.  Same results as srf_threshold_compact. */
#define SRF_COMPACT_NOALLOC_MAX (1LL << 28)
int srf_threshold_compact_noalloc(int64_t n, const float* values, float threshold, int32_t* idx,
                                  int32_t* count, void* stream);

/* out[r, c] = sum_k weights[k] * values[k, c] over the samples k of ray r (dense or ragged,
 * as srf_composite_fwd).  Replaces LowRankTensor.volume_render (src/models/SimpleTensoRF09.py:
 * 814-817). */
int srf_volume_render_fwd(int n_rays, int n_samples, const int32_t* ray_offsets,
                          const float* weights, const float* values, int n_channels, float* out,
                          void* stream);

/* Backward of srf_volume_render_fwd: g_weights[k] = sum_c g_out[r, c] values[k, c] and
 * g_values[k, c] = weights[k] g_out[r, c] (both overwritten; either may be NULL). */
int srf_volume_render_bwd(int n_rays, int n_samples, const int32_t* ray_offsets,
                          const float* weights, const float* values, int n_channels,
                          const float* g_out, float* g_weights, float* g_values, void* stream);

/* Depths of the non-NDC TensoRF sampler (src/models/SimpleTensoRF09.py:382-400): the ray's
 * entry into bbox (host floats), t_min = max_k min((hi_k - o_k) / v_k, (lo_k - o_k) / v_k)
 * with v = d (d == 0 -> 1e-6), clamped to [near, far]; z[r, k] = t_min + step * (k + u[r])
 * (u NULL: k).  rays_o, rays_d [n_rays,3]; z [n_rays, n_samples]. */
int srf_aabb_samples(int n_rays, int n_samples, const float* rays_o, const float* rays_d,
                     const float* bbox, float near, float far, float step, const float* u,
                     float* z, void* stream);

/* ----------------------------------------------------------- TensoRF colour network */

/* Structure of the colour path after the appearance gather: basis_matrix_color
 * (Linear(n_features -> basis_dim), no bias; src/models/SimpleTensoRF09.py:1151) and
 * MlpFeaturesColorPredictor (:1370-1421): [PE_Lf(features), PE_Lv(view dir)] -> hidden ReLU ->
 * hidden ReLU -> 3 sigmoid.  n_features % 4 == 0 and <= 256, basis_dim <= 32, hidden == 128,
 * first-layer width basis_dim (2 features_pe + 1) + 3 (2 views_pe + 1) <= 240.
 * prec: operands of the MFMA products of the forward and of the backward chain.  SRF_PREC_F32
 * (FP32 operands); SRF_PREC_BF16_SPLIT (split BF16: three planes in the forward, two in the
 * backward chain, as srf_mlp_desc's); SRF_PREC_BF16 (BF16 operands rounded from the FP32 values).
 * Activations, the stored layer inputs and accumulation are FP32 in every mode; the weight
 * gradients reduce the FP32 rows with FP32 operands (SRF_PREC_F32) or split BF16 (two planes, both
 * BF16 modes).  The reference computes in FP32 throughout. */
typedef struct srf_color_desc {
  int n_features;    /* sum of the appearance components */
  int basis_dim;     /* features_dimension_color (27) */
  int features_pe;   /* features_positional_encoding_degree */
  int use_view_dirs; /* view_dependent_color */
  int views_pe;      /* views_positional_encoding_degree */
  int hidden;        /* num_units_color_predictor */
  int prec;          /* SRF_PREC_F32 / SRF_PREC_BF16 / SRF_PREC_BF16_SPLIT */
} srf_color_desc;

/* Size of the MFMA-fragment copy of the weights written by srf_color_pack (bytes; 0 if the
 * descriptor is unsupported). */
size_t srf_color_packed_bytes(const srf_color_desc* desc);

/* Training workspace for n_points rows (bytes). */
size_t srf_color_workspace_bytes(const srf_color_desc* desc, int n_points);

/* Repacks the reference-layout parameters (basis [basis_dim, n_features]; mlp.0 / mlp.2 / mlp.4
 * weights [out, in] and biases) into the kernels' fragments, once per parameter update. */
int srf_color_pack(const srf_color_desc* desc, const float* basis, const float* w1, const float* b1,
                   const float* w2, const float* b2, const float* w3, const float* b3, void* packed,
                   void* stream);

/* rgb of the appearance rows app [n, n_features] (16-byte aligned).  View direction of row j:
 * view_dirs[r] with r = (idx ? idx[j] : j) / n_samples (the sample's ray in the dense
 * [rays, n_samples] layout).  n_points is the capacity (rows of app and idx, and the workspace
 * size); point_count (device int32, NULL: n_points) the rows processed.  rgb_at_idx 0: rgb row j
 * of [n, 3]; 1: rgb row idx[j] of the dense [rays * n_samples, 3] sample buffer, whose other rows
 * the caller zero-fills (the reference's rgb = zeros(...), rgb[validity_mask] = surface_rgb at
 * :1246, :1271).  Replaces
 * basis_matrix_color + color_predictor in VmDecomposedTensor.get_color
 * (src/models/SimpleTensoRF09.py:1262-1270).  workspace NULL: inference; otherwise the layer
 * inputs are kept for srf_color_backward. */
int srf_color_forward(const srf_color_desc* desc, const void* packed, int n_points,
                      const int32_t* point_count, const float* app, const int32_t* idx, int n_samples,
                      const float* view_dirs, float* rgb, int rgb_at_idx, void* workspace, void* stream);

/* Backward of srf_color_forward (same workspace, capacity, count and rgb addressing): d_app
 * [n, n_features] (rows below the count overwritten, 16-byte aligned) for
 * srf_vm_appearance_bwd, and the parameter gradients in the reference layouts, accumulated with
 * FP32 atomics (zero them first). */
int srf_color_backward(const srf_color_desc* desc, const void* packed, int n_points,
                       const int32_t* point_count, const int32_t* idx, int rgb_at_idx, const float* rgb,
                       const float* d_rgb, void* workspace, float* d_app, float* d_basis, float* d_w1,
                       float* d_b1, float* d_w2, float* d_b2, float* d_w3, float* d_b3, void* stream);

/* ------------------------------------------------------------------ optimiser (f4) */

/* One Adam step (torch.optim.Adam as src/optimizers/OptimizerFactory02.py:9-21 builds it: amsgrad
 * and maximize off, weight decay added to the gradient) over n_tensors parameter tensors in one
 * launch per 32 tensors.  Tensor t: params[t], grads[t], exp_avg[t], exp_avg_sq[t] (FP32, numel[t]
 * elements in one shared dense layout), steps[t] (device float, torch's state['step'], advanced
 * by one), lr[t] and weight_decay[t] (its parameter group's, host values).  Same operation order
 * as torch's single-tensor Adam; the bias corrections in double. */
int srf_adam_step(int n_tensors, float* const* params, const float* const* grads, float* const* exp_avg,
                  float* const* exp_avg_sq, float* const* steps, const int64_t* numel, const float* lr,
                  const float* weight_decay, float beta1, float beta2, float eps, void* stream);

/* ------------------------------------------------------- launch-schedule tuning */

/* Knobs of the launch schedule, picked by an autotuner (srf_hip.autotune) for a workload.  No
 * counterpart in the reference, whose PyTorch ops have no schedule.  Value 0 restores the
 * built-in choice; results are the same for every value up to the order of FP32 atomic
 * additions.  Launches read the table when they are issued, so a captured graph keeps the
 * values it was captured with. */
#define SRF_TUNE_DW_CHUNK 0       /* points per work item of the MLP weight-gradient GEMM (srf_mlp_weight_grads):
                                     0 = about 1024 work items; else 2048..262144, a multiple of 256 */
#define SRF_TUNE_COLOR_DW_CHUNK 1 /* the same for the colour network's weight gradients (srf_color_backward):
                                     0 = adaptive; else 2048..65536, a multiple of 256 */
#define SRF_TUNE_GATHER_BLOCKS 2  /* grid cap of the factor-gather kernels (srf_vm_*, srf_cp_*): 0 = 8192;
                                     else 256..65536 blocks of 256 threads */
#define SRF_TUNE_GATHER_RUN 3     /* consecutive samples one lane group walks in the gather backwards, summing
                                     equal-corner gradients before the atomics: 0 = 32; else 1..256 */
#define SRF_TUNE_KNOBS 4

/* Sets knob to value (SRF_EINVAL for an unknown knob or a value outside its range). */
int srf_tune_set(int knob, int64_t value);

/* Current value of knob (0 = built-in choice), -1 for an unknown knob. */
int64_t srf_tune_get(int knob);

/* ---------------------------------------------------------------- training losses (f1) */

/* Weighted sum of MSE terms over the selected rays:
 *   loss = sum_t weights[t] * mean_{i selected} mean_c (preds[t][i,c] - target[i,c])^2.
 * Replaces MSE.compute_mse summed over models_to_regularize (src/loss_functions/MSE14.py:29-94; c = 3,
 * target_rgb, indices_mask_nerf) and SparseDepthMSE.compute_depth_loss (src/loss_functions/
 * SparseDepthMSE14.py:29-95; c = 1, the sparse depths, indices_mask_sparse_depth), and
 * CoarseFineConsistencyLoss.compute_loss_sd (src/loss_functions/CoarseFineConsistencyLoss34.py:166-181;
 * target = the detached fine depth).
 *   preds: host array of n_terms (<= 8) device pointers to [n, channels]; weights: host array of n_terms;
 *   select [n] bytes (nonzero = selected) or NULL for all rays.
 * Outputs (device): loss [1], term_values [n_terms] or NULL (each term's mean), maps [n_terms, n] or NULL
 * (each ray's channel mean, the reference's loss maps; 0 on rays not selected), count [1] (the number of
 * selected rays; an empty selection gives 0, as the reference's `if numel() > 0` does).  One block,
 * fixed-order sums: no host synchronisation, the same value on every run. */
int srf_mse_terms_fwd(int n_terms, const float* const* preds, const float* weights, const float* target,
                      const uint8_t* select, int n, int channels, float* loss, float* term_values, float* maps,
                      int* count, void* stream);

/* Backward of srf_mse_terms_fwd: d_preds[t][i,c] = g_loss weights[t] 2 (preds[t][i,c] - target[i,c]) /
 * (channels count) on selected rays, 0 elsewhere.  g_loss [1] and count [1] are device scalars;
 * d_preds: host array of n_terms device pointers (NULL entries are skipped). */
int srf_mse_terms_bwd(int n_terms, const float* const* preds, const float* weights, const float* target,
                      const uint8_t* select, int n, int channels, const int* count, const float* g_loss,
                      float* const* d_preds, void* stream);

/* Reprojection-gated depth consistency between two depth estimates of the same rays.  Replaces
 * compute_loss_nerf of AugmentationsDepthLoss (src/loss_functions/AugmentationsDepthLoss11.py:108-181,
 * both_invalid_rule = 1) and of CoarseFineConsistencyLoss (src/loss_functions/
 * CoarseFineConsistencyLoss34.py:85-164, both_invalid_rule = 0), with CommonUtils.reproject
 * (src/utils/CommonUtils04.py:227-253).  Per selected ray: both depths' points o + d z reprojected into
 * the view with the nearest camera centre (rounded half to even), the patch_y x patch_x ground-truth
 * patches around the source pixel and the two reprojections (zero padded as the reference pads), their
 * RMSEs, the gate masks, and the loss maps map1 = mask2 (d1 - d2)^2, map2 = mask1 mask2 (d2 - d1)^2 (the
 * reference zeroes its inputs in place, so its second term only counts rays where both masks hold).
 *   depth1, depth2 [n]; rays_o, rays_d [n,3] (world); pixel_id [n,3] int32 (image, x, y);
 *   select [n] bytes or NULL (indices_mask_nerf); poses [n_views,4,4] camera-to-world (extrinsics_all);
 *   intrinsic [3,3] (the first view's, as the reference uses); images [n_views,height,width,channels].
 * Outputs: map1, map2 [n] (0 on rays not selected), flags [n] (bit 0 mask1, bit 1 mask2, bit 2
 * selected), loss [1] = mean(map1) + mean(map2) over the selected rays, count [1]. */
int srf_depth_consistency_fwd(int n, const float* depth1, const float* depth2, const float* rays_o,
                              const float* rays_d, const int32_t* pixel_id, const uint8_t* select, int n_views,
                              const float* poses, const float* intrinsic, const float* images, int height,
                              int width, int channels, int patch_x, int patch_y, float rmse_threshold,
                              int both_invalid_rule, float* map1, float* map2, uint8_t* flags, float* loss,
                              int* count, void* stream);

/* Backward of srf_depth_consistency_fwd: d_depth1 = g 2 (d1 - d2) / count where mask2, d_depth2 =
 * g 2 (d2 - d1) / count where mask1 and mask2, 0 elsewhere (either output may be NULL). */
int srf_depth_consistency_bwd(int n, const float* depth1, const float* depth2, const uint8_t* flags,
                              const int* count, const float* g_loss, float* d_depth1, float* d_depth2,
                              void* stream);

/* ------------------------------------------------ TensoRF model modifications (f3) */

/* Points of the dense evaluation grid of LowRankTensor.update_alpha_mask (src/models/SimpleTensoRF09.py:
 * 851-857): pts [gx*gy*gz, 3], point (i, j, k) at row (i gy + j) gz + k (meshgrid 'ij'), each coordinate
 * bb0 (1 - s) + bb1 s with s the torch.linspace(0, 1, g) position on its axis.  bbox: host array of 6
 * floats (min xyz, max xyz). */
int srf_dense_grid_points(int gx, int gy, int gz, const float* bbox, float* pts, void* stream);

/* The new occupancy volume of update_alpha_mask (:858-868) from the dense grid's densities: alpha =
 * clamp(1 - exp(-sigma length), 0, 1), transposed to (z, y, x), max-pooled over 3x3x3 (stride 1, padding
 * 1), then 1 where >= threshold and 0 elsewhere.  sigma [gx*gy*gz] in srf_dense_grid_points' order;
 * alpha_out [gz, gy, gx]; bounds int32 [6]: min x, y, z and max x, y, z index of the occupied voxels (the
 * box of :870-875; min > max when no voxel is occupied). */
int srf_alpha_mask_build(int gx, int gy, int gz, const float* sigma, float length, float threshold,
                         float* alpha_out, int* bounds, void* stream);

/* F.interpolate(mode='bilinear', align_corners=True) of one texel-major factor [height, width, channels]
 * to [height_new, width_new, channels] (VmDecomposedTensor.upsample_vectors_and_matrices, :1281-1302;
 * lines are factors of width 1). */
int srf_factor_upsample(int channels, int height, int width, const float* src, int height_new, int width_new,
                        float* dst, void* stream);

/* Workspace bytes of srf_tv_fwd for n_planes (<= 8) planes (0 otherwise). */
size_t srf_tv_workspace_bytes(int n_planes);

/* TotalVariationLoss.compute_tv_loss (src/loss_functions/TotalVariationLoss04.py:97-116) over n_planes
 * texel-major components [heights[i], widths[i], channels[i]]: loss [1] = scale sum_i 2 (sum of squared
 * differences along H / (c (H - 1) W) + along W / (c H (W - 1))), each count at least 1.  planes,
 * channels, heights, widths: host arrays; scale = the iteration weight times the loss weight. */
int srf_tv_fwd(int n_planes, const int* channels, const int* heights, const int* widths,
               const float* const* planes, float scale, void* workspace, float* loss, void* stream);

/* Backward of srf_tv_fwd: d_planes[i] (texel-major, overwritten) = g_loss times the derivative. */
int srf_tv_bwd(int n_planes, const int* channels, const int* heights, const int* widths,
               const float* const* planes, float scale, const float* g_loss, float* const* d_planes,
               void* stream);

/* MassConcentrationLoss.compute_entropy_loss (src/loss_functions/MassConcentrationLoss06.py:63-83) on
 * weights [n_rays, n_samples]: per selected ray, the weights summed over n_bins (<= 64) consecutive bins of
 * bin_sizes[b] samples (host array, summing to n_samples), entr(bin + 1e-3) = -x log x summed over the
 * bins; loss [1] = their mean over the selected rays (select [n_rays] bytes or NULL), count [1].
 * binned [n_rays, n_bins] (bin + 1e-3) and ray_loss [n_rays] are outputs the backward reads. */
int srf_entropy_fwd(int n_rays, int n_samples, const float* weights, const uint8_t* select, int n_bins,
                    const int* bin_sizes, float* binned, float* ray_loss, float* loss, int* count, void* stream);

/* Backward of srf_entropy_fwd: d_weights [n_rays, n_samples] = g_loss / count * -(log(binned) + 1) of
 * the sample's bin on selected rays, 0 elsewhere. */
int srf_entropy_bwd(int n_rays, int n_samples, const uint8_t* select, int n_bins, const int* bin_sizes,
                    const float* binned, const int* count, const float* g_loss, float* d_weights, void* stream);

/* Lane-layout self-test of the MFMA building blocks (one 16x16 tile): d = a b with
 * a [16,K], b [K,16] row-major FP32 (rounded to the operand type), K = 4 for SRF_PREC_F32 and
 * 32 for SRF_PREC_BF16 / SRF_PREC_F16.  No counterpart in the reference: a unit check of the
 * operand layouts the MLP kernels are built on. */
int srf_selftest_mfma(int prec, const float* a, const float* b, float* d, void* stream);

/* Self-test of the direct global -> LDS fragment copy the MLP kernels stream weights with:
 * src and dst are 8 KB; dst fragment f (1 KB) = src fragment 7 - f.  No counterpart in the
 * reference. */
int srf_selftest_lds_dma(const void* src, void* dst, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SRF_HIP_H_ */
