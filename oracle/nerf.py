"""Oracle: Simple-NeRF hot path restated in plain PyTorch (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use this
module.  Each function follows the reference function cited in its docstring,
operation for operation, so that in float32 it reproduces the reference's
rounding and in float64 it is the high-precision reference.

Random draws that the reference takes from the CPU generator inside the path
(t_rand at SimpleNeRF17.py:355, u at :397, sigma noise at :740) are explicit
arguments here, exactly as they are in the HIP kernels.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F


# --------------------------------------------------------------------------- rays
def get_rays(pixel_id: torch.Tensor, intrinsics: torch.Tensor, extrinsics: torch.Tensor,
             mip_nerf_used: bool = False):
    """Pinhole rays. src/utils/CommonUtils04.py:73-95 (nerf_synthetic=False).

    pixel_id: [N,3] (image_id, x, y); intrinsics [N,3,3]; extrinsics (c2w) [N,4,4].
    """
    dtype = intrinsics.dtype
    pid = pixel_id.to(dtype)
    x, y = pid[:, 1].clone(), pid[:, 2].clone()
    if mip_nerf_used:
        x += 0.5
        y += 0.5
    pts_h = torch.stack([x, y, torch.ones_like(x)], dim=1)
    dirs = (torch.linalg.inv(intrinsics) @ pts_h[:, :, None])[:, :, 0]
    dirs[:, 1:] *= -1
    rays_d = torch.sum(dirs[:, None, :] * extrinsics[:, :3, :3], dim=-1)
    rays_o = extrinsics[:, :3, -1]
    return rays_o, rays_d


def get_ndc_rays(rays_o, rays_d, resolution, intrinsics, near: float):
    """Forward-facing NDC warp. src/utils/CommonUtils04.py:120-138."""
    h, w = resolution
    fx, fy = intrinsics[:, 0, 0], intrinsics[:, 1, 1]
    t = -(near + rays_o[:, 2]) / rays_d[:, 2]
    rays_o = rays_o + t[:, None] * rays_d
    o0 = -1. / (w / (2. * fx)) * rays_o[:, 0] / rays_o[:, 2]
    o1 = -1. / (h / (2. * fy)) * rays_o[:, 1] / rays_o[:, 2]
    o2 = 1. + 2. * near / rays_o[:, 2]
    d0 = -1. / (w / (2. * fx)) * (rays_d[:, 0] / rays_d[:, 2] - rays_o[:, 0] / rays_o[:, 2])
    d1 = -1. / (h / (2. * fy)) * (rays_d[:, 1] / rays_d[:, 2] - rays_o[:, 1] / rays_o[:, 2])
    d2 = -2. * near / rays_o[:, 2]
    return torch.stack([o0, o1, o2], -1), torch.stack([d0, d1, d2], -1)


def get_view_dirs(rays_d):
    """src/utils/CommonUtils04.py:147-149."""
    return rays_d / torch.linalg.norm(rays_d, ord=2, dim=-1, keepdim=True)


def convert_depth_to_ndc(depths, rays_o, rays_d, near):
    """src/utils/CommonUtils04.py:198-205."""
    oz = rays_o[:, 2:]
    dz = rays_d[:, 2:]
    tn = -(near + oz) / dz
    oz_prime = oz + tn * dz
    return 1 - oz_prime / (oz_prime + (depths - tn) * dz)


def convert_depth_from_ndc(depths_ndc, rays_o, rays_d):
    """NDC -> metric depth, near hard-coded to 1. src/utils/CommonUtils04.py:208-224
    (identical copy at src/models/SimpleNeRF17.py:541-558)."""
    near = 1
    oz = rays_o[..., 2:3]
    dz = rays_d[..., 2:3]
    tn = -(near + oz) / dz
    constant = torch.where(depths_ndc == 1., 1e-3, 0.).to(depths_ndc.dtype)
    return (oz + tn * dz) / dz * (1 / (1 - depths_ndc + constant) - 1) + tn


# ----------------------------------------------------------------------- sampling
def linspace(start: float, end: float, steps: int, dtype=torch.float32, device='cpu'):
    """torch.linspace, used at src/models/SimpleNeRF17.py:341 and :394."""
    return torch.linspace(start, end, steps=steps, dtype=dtype, device=device)


def stratified_z_vals(n_rays: int, n_samples: int, near: float, far: float, lindisp: bool = False,
                      u: torch.Tensor | None = None, dtype=torch.float32, device='cpu'):
    """Coarse depths. src/models/SimpleNeRF17.py:330-358 (NeRF), and the NDC branch of
    src/models/SimpleTensoRF09.py:364-381.  ``u`` [n_rays, n_samples] replaces the CPU
    ``torch.rand`` at :355; ``u=None`` is the deterministic (eval, perturb=False) path."""
    t_vals = linspace(0., 1., n_samples, dtype=dtype, device=device)
    if not lindisp:
        z = near * (1. - t_vals) + far * t_vals
    else:
        z = 1. / (1. / near * (1. - t_vals) + 1. / far * t_vals)
    z = z.expand([n_rays, n_samples])
    if u is not None:
        mids = .5 * (z[..., 1:] + z[..., :-1])
        upper = torch.cat([mids, z[..., -1:]], -1)
        lower = torch.cat([z[..., :1], mids], -1)
        z = lower + (upper - lower) * u
    return z


def pdf_cdf(weights: torch.Tensor):
    """CDF construction of sample_pdf. src/models/SimpleNeRF17.py:386-390."""
    weights = weights + 1e-5
    pdf = weights / torch.sum(weights, -1, keepdim=True)
    cdf = torch.cumsum(pdf, -1)
    return torch.cat([torch.zeros_like(cdf[..., :1]), cdf], -1)


def sample_pdf(bins: torch.Tensor, weights: torch.Tensor, n_samples: int, u: torch.Tensor | None = None,
               cdf: torch.Tensor | None = None):
    """Inverse-CDF sampling. src/models/SimpleNeRF17.py:384-417 (copy at
    src/models/SimpleTensoRF09.py:427-460).

    ``u=None`` is det=True (linspace, :393-395); otherwise ``u`` [N, n_samples] replaces
    the CPU ``torch.rand`` at :397.  ``cdf`` may be supplied to bypass :386-390.
    Returns (samples, inds, cdf); ``inds`` is torch.searchsorted(cdf, u, right=True).
    """
    if cdf is None:
        cdf = pdf_cdf(weights)
    if u is None:
        u = linspace(0., 1., n_samples, dtype=cdf.dtype, device=cdf.device)
        u = u.expand(list(cdf.shape[:-1]) + [n_samples])
    u = u.contiguous()
    inds = torch.searchsorted(cdf, u, right=True)
    below = torch.max(torch.zeros_like(inds - 1), inds - 1)
    above = torch.min((cdf.shape[-1] - 1) * torch.ones_like(inds), inds)
    inds_g = torch.stack([below, above], -1)
    matched_shape = [inds_g.shape[0], inds_g.shape[1], cdf.shape[-1]]
    cdf_g = torch.gather(cdf.unsqueeze(1).expand(matched_shape), 2, inds_g)
    bins_g = torch.gather(bins.unsqueeze(1).expand(matched_shape), 2, inds_g)
    denom = (cdf_g[..., 1] - cdf_g[..., 0])
    denom = torch.where(denom < 1e-5, torch.ones_like(denom), denom)
    t = (u - cdf_g[..., 0]) / denom
    samples = bins_g[..., 0] + t * (bins_g[..., 1] - bins_g[..., 0])
    return samples, inds, cdf


def fine_z_vals(z_coarse: torch.Tensor, weights_coarse: torch.Tensor, n_fine: int,
                u: torch.Tensor | None = None):
    """Hierarchical depths. src/models/SimpleNeRF17.py:360-371.  Returns (z_merged, z_samples)."""
    z_mid = .5 * (z_coarse[..., 1:] + z_coarse[..., :-1])
    z_samples, _, _ = sample_pdf(z_mid, weights_coarse[..., 1:-1], n_fine, u=u)
    z_samples = z_samples.detach()
    z_merged, _ = torch.sort(torch.cat([z_coarse, z_samples], -1), -1)
    return z_merged, z_samples


def sample_points(rays_o, rays_d, z_vals):
    """o + d z. src/models/SimpleNeRF17.py:213-215, :266-268."""
    return rays_o[..., None, :] + rays_d[..., None, :] * z_vals[..., :, None]


# ---------------------------------------------------------------- encoding + MLP
def positional_encoding(x: torch.Tensor, degree: int):
    """[x, sin(2^0 x), cos(2^0 x), ..., sin(2^(L-1) x), cos(2^(L-1) x)].
    src/models/SimpleNeRF17.py:581-613 configured by MLP.get_positional_encoder :669-682
    (include_input, log_sampling, max_freq_log2 = L-1, num_freqs = L)."""
    outs = [x]
    if degree > 0:
        freq_bands = 2. ** torch.linspace(0., degree - 1, steps=degree)
        for freq in freq_bands:
            f = freq.to(x.dtype).to(x.device)
            outs.append(torch.sin(x * f))
            outs.append(torch.cos(x * f))
    return torch.cat(outs, -1)


def mlp_dims(pe_degree_pts=10, pe_degree_dirs=4, sigma_pe_degree=None, use_view_dirs=True,
             view_dependent_rgb=True, depth=8, width=256, views_width=128):
    """Input widths of MLP (src/models/SimpleNeRF17.py:625-636)."""
    pts_dim = (2 * pe_degree_pts + 1) * 3
    views_dim = (2 * pe_degree_dirs + 1) * 3 if use_view_dirs else 0
    if sigma_pe_degree is not None:
        trunk_dim = (2 * sigma_pe_degree + 1) * 3
        views_dim += pts_dim - trunk_dim
    else:
        trunk_dim = pts_dim
    return dict(pts_dim=pts_dim, trunk_dim=trunk_dim, views_dim=views_dim)


def mlp_forward(params: dict, pts: torch.Tensor, dirs: torch.Tensor | None, *, pe_degree_pts=10,
                pe_degree_dirs=4, sigma_pe_degree=None, use_view_dirs=True, view_dependent_rgb=True,
                depth=8, skips=(4,), noise: torch.Tensor | None = None, sigma_softplus: bool = False):
    """NeRF MLP. src/models/SimpleNeRF17.py:696-785 (forward, get_view_independent_outputs,
    get_view_dependent_outputs) with layer shapes from :644-666.

    ``params`` uses the reference state_dict names (``pts_linears.{i}.weight`` ...).
    ``noise`` is the already-scaled sigma noise (randn * raw_noise_std, :740) or None.
    ``sigma_softplus`` selects softplus for the density head instead of relu (synthetic
    benchmark option; the reference NeRF uses relu at :742).
    Returns dict(sigma [P,1], rgb [P,3]).
    """
    dims = mlp_dims(pe_degree_pts, pe_degree_dirs, sigma_pe_degree, use_view_dirs, view_dependent_rgb)
    view_dep_outputs = view_dependent_rgb  # predict_visibility is False on this path
    enc = positional_encoding(pts, pe_degree_pts)
    x_in = enc[:, :dims['trunk_dim']]
    h = x_in
    for i in range(depth):
        h = F.linear(h, params[f'pts_linears.{i}.weight'], params[f'pts_linears.{i}.bias'])
        h = F.relu(h)
        if i in skips:
            h = torch.cat([x_in, h], -1)
    out = F.linear(h, params['pts_output_linear.weight'], params['pts_output_linear.bias'])
    raw_sigma = out[..., 0:1]
    if noise is not None:
        raw_sigma = raw_sigma + noise
    sigma = F.softplus(raw_sigma) if sigma_softplus else F.relu(raw_sigma)
    result = {'sigma': sigma}
    if not view_dependent_rgb:
        result['rgb'] = torch.sigmoid(out[..., 1:4])
        return result
    feature = F.linear(h, params['feature_linear.weight'], params['feature_linear.bias'])
    feature = torch.cat([feature, enc[:, dims['trunk_dim']:]], dim=1)
    enc_dirs = positional_encoding(dirs, pe_degree_dirs)
    hv = torch.cat([feature, enc_dirs], -1)
    hv = F.relu(F.linear(hv, params['views_linears.0.weight'], params['views_linears.0.bias']))
    vo = F.linear(hv, params['views_output_linear.weight'], params['views_output_linear.bias'])
    result['rgb'] = torch.sigmoid(vo[..., 0:3])
    return result


def mlp_param_shapes(pe_degree_pts=10, pe_degree_dirs=4, sigma_pe_degree=None, use_view_dirs=True,
                     view_dependent_rgb=True, depth=8, width=256, views_width=128, skips=(4,)):
    """Parameter names/shapes of MLP (src/models/SimpleNeRF17.py:644-666)."""
    dims = mlp_dims(pe_degree_pts, pe_degree_dirs, sigma_pe_degree, use_view_dirs, view_dependent_rgb)
    shapes = {}
    din = dims['trunk_dim']
    for i in range(depth):
        k = din if i == 0 else (width + dims['trunk_dim'] if (i - 1) in skips else width)
        shapes[f'pts_linears.{i}.weight'] = (width, k)
        shapes[f'pts_linears.{i}.bias'] = (width,)
    if view_dependent_rgb:
        shapes['views_linears.0.weight'] = (views_width, dims['views_dim'] + width)
        shapes['views_linears.0.bias'] = (views_width,)
        shapes['pts_output_linear.weight'] = (1, width)
        shapes['pts_output_linear.bias'] = (1,)
        shapes['feature_linear.weight'] = (width, width)
        shapes['feature_linear.bias'] = (width,)
        shapes['views_output_linear.weight'] = (3, views_width)
        shapes['views_output_linear.bias'] = (3,)
    else:
        shapes['pts_output_linear.weight'] = (4, width)
        shapes['pts_output_linear.bias'] = (4,)
    return shapes


# ------------------------------------------------------------------- compositing
def volume_rendering(sigma: torch.Tensor, rgb: torch.Tensor, *, z_vals_ndc=None, rays_o=None, rays_d=None,
                     rays_d_ndc=None, z_vals=None, ndc: bool = True, white_bkgd: bool = False,
                     distance_scale: float | None = None):
    """Alpha compositing. src/models/SimpleNeRF17.py:486-539 (NeRF) and
    src/models/SimpleTensoRF09.py:767-819 (TensoRF: ``distance_scale`` multiplies
    sigma*delta at :787; rgb composited by volume_render :816-819).

    sigma [N,S], rgb [N,S,3].  Returns the reference's output dict.
    """
    if not ndc:
        z1 = torch.cat([z_vals, torch.full_like(z_vals[..., :1], 1e10)], -1)
        delta = (z1[..., 1:] - z1[..., :-1]) * torch.norm(rays_d[..., None, :], dim=-1)
    else:
        z1 = torch.cat([z_vals_ndc, torch.ones_like(z_vals_ndc[..., :1])], -1)
        delta = (z1[..., 1:] - z1[..., :-1]) * torch.norm(rays_d_ndc[..., None, :], dim=-1)
    if distance_scale is None:
        alpha = 1. - torch.exp(-sigma * delta)
    else:
        alpha = 1. - torch.exp(-sigma * delta * distance_scale)
    visibility = torch.cumprod(torch.cat([torch.ones_like(alpha[:, :1]), 1. - alpha + 1e-10], -1), -1)[:, :-1]
    weights = alpha * visibility
    rgb_map = torch.sum(weights[..., None] * rgb, dim=-2)
    acc_map = torch.sum(weights, dim=-1)
    out = {}
    if not ndc:
        depth_map = torch.sum(weights * z_vals, dim=-1) / (acc_map + 1e-6)
        depth_var_map = torch.sum(weights * torch.square(z_vals - depth_map[..., None]), dim=-1)
    else:
        depth_map_ndc = torch.sum(weights * z_vals_ndc, dim=-1) / (acc_map + 1e-6)
        depth_var_map_ndc = torch.sum(weights * torch.square(z_vals_ndc - depth_map_ndc[..., None]), dim=-1)
        zm = convert_depth_from_ndc(z_vals_ndc, rays_o, rays_d)
        depth_map = torch.sum(weights * zm, dim=-1) / (acc_map + 1e-6)
        depth_var_map = torch.sum(weights * torch.square(zm - depth_map[..., None]), dim=-1)
        out['depth_ndc'] = depth_map_ndc
        out['depth_var_ndc'] = depth_var_map_ndc
    if white_bkgd:
        rgb_map = rgb_map + (1. - acc_map[..., None])
    out.update({'rgb': rgb_map, 'acc': acc_map, 'alpha': alpha, 'visibility': visibility,
                'weights': weights, 'depth': depth_map, 'depth_var': depth_var_map})
    return out
