"""Alpha compositing on the HIP kernels, with a hand-written backward.

Mirrors SimpleNeRF.volume_rendering (src/models/SimpleNeRF17.py:486-539, depth conversion
:541-558) and LowRankTensor.get_volume_rendering_weights / volume_render
(src/models/SimpleTensoRF09.py:767-819).
"""
from __future__ import annotations

import torch

from . import _lib
from ._checks import dev_tensor, opt_tensor

_KEYS = ('rgb', 'acc', 'depth', 'depth_var', 'depth_ndc', 'depth_var_ndc', 'alpha', 'visibility', 'weights')


class _Composite(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sigma, rgb, z, rays_o, rays_d, rays_d_ndc, ray_offsets, n_samples, ndc, white_bkgd,
                distance_scale):
        n_rays = rays_d.shape[0]
        P = sigma.shape[0]
        dev = sigma.device
        f = dict(dtype=torch.float32, device=dev)
        rgb_map = torch.empty((n_rays, 3), **f) if rgb is not None else torch.zeros((n_rays, 3), **f)
        acc, depth, depth_var = (torch.empty((n_rays,), **f) for _ in range(3))
        depth_ndc = torch.empty((n_rays,), **f) if ndc else None
        depth_var_ndc = torch.empty((n_rays,), **f) if ndc else None
        alpha, vis, w = (torch.empty((P,), **f) for _ in range(3))
        _lib.call('srf_composite_fwd', n_rays, n_samples, _lib.ptr(ray_offsets), _lib.ptr(sigma), _lib.ptr(rgb),
                  _lib.ptr(z), _lib.ptr(rays_o), _lib.ptr(rays_d), _lib.ptr(rays_d_ndc), int(ndc), int(white_bkgd),
                  float(distance_scale), _lib.ptr(rgb_map), _lib.ptr(acc), _lib.ptr(depth), _lib.ptr(depth_var),
                  _lib.ptr(depth_ndc), _lib.ptr(depth_var_ndc), _lib.ptr(alpha), _lib.ptr(vis), _lib.ptr(w),
                  _lib.stream_ptr(dev))
        ctx.save_for_backward(sigma, rgb, z, rays_o, rays_d, rays_d_ndc, ray_offsets, acc, depth, depth_ndc)
        ctx.cfg = (n_rays, n_samples, ndc, white_bkgd, distance_scale)
        # outputs the caller does not differentiate reach the backward as None (the kernel reads a
        # null upstream gradient as zero) instead of as zero-filled tensors, one fill per output
        ctx.set_materialize_grads(False)
        if rgb is None:
            ctx.mark_non_differentiable(rgb_map)
        if not ndc:
            depth_ndc = torch.zeros((n_rays,), **f)
            depth_var_ndc = torch.zeros((n_rays,), **f)
            ctx.mark_non_differentiable(depth_ndc, depth_var_ndc)
        return rgb_map, acc, depth, depth_var, depth_ndc, depth_var_ndc, alpha, vis, w

    @staticmethod
    def backward(ctx, g_rgb, g_acc, g_depth, g_depth_var, g_depth_ndc, g_depth_var_ndc, g_alpha, g_vis, g_w):
        sigma, rgb, z, rays_o, rays_d, rays_d_ndc, ray_offsets, acc, depth, depth_ndc = ctx.saved_tensors
        n_rays, n_samples, ndc, white_bkgd, distance_scale = ctx.cfg
        c = lambda t: None if t is None else t.contiguous()
        d_sigma = torch.empty_like(sigma)
        d_rgb = torch.empty_like(rgb) if rgb is not None and ctx.needs_input_grad[1] else None
        _lib.call('srf_composite_bwd', n_rays, n_samples, _lib.ptr(ray_offsets), _lib.ptr(sigma), _lib.ptr(rgb),
                  _lib.ptr(z), _lib.ptr(rays_o), _lib.ptr(rays_d), _lib.ptr(rays_d_ndc), int(ndc), int(white_bkgd),
                  float(distance_scale), _lib.ptr(acc), _lib.ptr(depth), _lib.ptr(depth_ndc),
                  _lib.ptr(c(g_rgb)), _lib.ptr(c(g_acc)), _lib.ptr(c(g_depth)), _lib.ptr(c(g_depth_var)),
                  _lib.ptr(c(g_depth_ndc) if ndc else None), _lib.ptr(c(g_depth_var_ndc) if ndc else None),
                  _lib.ptr(c(g_alpha)), _lib.ptr(c(g_vis)), _lib.ptr(c(g_w)), _lib.ptr(d_sigma), _lib.ptr(d_rgb),
                  _lib.stream_ptr(sigma.device))
        return d_sigma, d_rgb, None, None, None, None, None, None, None, None, None


def composite(sigma: torch.Tensor, rgb: torch.Tensor, z: torch.Tensor, rays_o: torch.Tensor, rays_d: torch.Tensor,
              rays_d_ndc: torch.Tensor | None = None, *, ndc: bool = True, white_bkgd: bool = False,
              distance_scale: float = 1.0, ray_offsets: torch.Tensor | None = None, max_samples: int | None = None):
    """Volume rendering of dense [N,S] (or ragged, with ``ray_offsets`` [N+1] int32) samples.

    sigma [N,S] (or [N,S,1]), rgb [N,S,3] or None (weights and depths only; the returned rgb
    is then zeros), z [N,S] (NDC depths if ndc else metric).
    Returns the reference's output dict: rgb, acc, depth, depth_var, (depth_ndc, depth_var_ndc),
    alpha, visibility, weights.
    """
    n_rays = rays_d.shape[0]
    if ray_offsets is None:
        if sigma.dim() == 3:
            sigma = sigma[..., 0]
        n_samples = sigma.shape[1]
        shape_s = sigma.shape
        sigma_f = dev_tensor(sigma, 'sigma', (n_rays, n_samples)).reshape(-1)
        rgb_f = None if rgb is None else dev_tensor(rgb, 'rgb', (n_rays, n_samples, 3)).reshape(-1, 3)
        z_f = dev_tensor(z, 'z', (n_rays, n_samples)).reshape(-1)
    else:
        ray_offsets = dev_tensor(ray_offsets, 'ray_offsets', (n_rays + 1,), torch.int32)
        if max_samples is None:
            raise _lib.SrfError('ragged compositing needs max_samples (the longest ray)')
        n_samples = int(max_samples)
        shape_s = sigma.shape
        sigma_f = dev_tensor(sigma, 'sigma').reshape(-1)
        P = sigma_f.shape[0]
        rgb_f = None if rgb is None else dev_tensor(rgb, 'rgb', (P, 3))
        z_f = dev_tensor(z, 'z', (P,))
    rays_o = dev_tensor(rays_o, 'rays_o', (n_rays, 3))
    rays_d = dev_tensor(rays_d, 'rays_d', (n_rays, 3))
    rays_d_ndc = opt_tensor(rays_d_ndc, 'rays_d_ndc', (n_rays, 3))
    if ndc and rays_d_ndc is None:
        raise _lib.SrfError('ndc compositing needs rays_d_ndc')
    outs = _Composite.apply(sigma_f, rgb_f, z_f, rays_o, rays_d, rays_d_ndc, ray_offsets, n_samples, bool(ndc),
                            bool(white_bkgd), float(distance_scale))
    res = dict(zip(_KEYS, outs))
    for k in ('alpha', 'visibility', 'weights'):
        res[k] = res[k].view(shape_s if ray_offsets is None else (-1,))
    if not ndc:
        del res['depth_ndc'], res['depth_var_ndc']
    return res


class _VolumeRender(torch.autograd.Function):
    @staticmethod
    def forward(ctx, weights, values, ray_offsets, n_rays, n_samples):
        C = values.shape[1]
        out = torch.empty((n_rays, C), dtype=torch.float32, device=values.device)
        _lib.call('srf_volume_render_fwd', n_rays, n_samples, _lib.ptr(ray_offsets), _lib.ptr(weights),
                  _lib.ptr(values), C, _lib.ptr(out), _lib.stream_ptr(values.device))
        ctx.save_for_backward(weights, values, ray_offsets)
        ctx.cfg = (n_rays, n_samples)
        return out

    @staticmethod
    def backward(ctx, g_out):
        weights, values, ray_offsets = ctx.saved_tensors
        n_rays, n_samples = ctx.cfg
        g_w = torch.empty_like(weights) if ctx.needs_input_grad[0] else None
        g_v = torch.empty_like(values) if ctx.needs_input_grad[1] else None
        if g_w is not None or g_v is not None:
            _lib.call('srf_volume_render_bwd', n_rays, n_samples, _lib.ptr(ray_offsets), _lib.ptr(weights),
                      _lib.ptr(values), values.shape[1], _lib.ptr(g_out.contiguous()), _lib.ptr(g_w),
                      _lib.ptr(g_v), _lib.stream_ptr(values.device))
        return g_w, g_v, None, None, None


def volume_render(values: torch.Tensor, weights: torch.Tensor, ray_offsets: torch.Tensor | None = None):
    """sum_k weights[..., k] * values[..., k, :] (LowRankTensor.volume_render,
    src/models/SimpleTensoRF09.py:814-817).  values [N,S,C], weights [N,S] (dense) or, with
    ray_offsets [N+1] int32, values [P,C] and weights [P]."""
    if ray_offsets is None:
        n_rays, n_samples, C = values.shape
        w = dev_tensor(weights, 'weights', (n_rays, n_samples)).reshape(-1)
        v = dev_tensor(values, 'values').reshape(-1, C)
    else:
        n_rays = ray_offsets.shape[0] - 1
        n_samples = 0
        ray_offsets = dev_tensor(ray_offsets, 'ray_offsets', (n_rays + 1,), torch.int32)
        v = dev_tensor(values, 'values')
        w = dev_tensor(weights, 'weights', (v.shape[0],))
    return _VolumeRender.apply(w, v, ray_offsets, n_rays, n_samples)
