"""TensoRF model modifications and their regularisers on the HIP kernels (csrc/tensorf_mods.hip).

  dense_grid_points / alpha_mask_build   LowRankTensor.update_alpha_mask (src/models/SimpleTensoRF09.py:849-897)
  upsample_factor                        VmDecomposedTensor.upsample_vectors_and_matrices (:1281-1302)
  tv_loss                                TotalVariationLoss.compute_tv_loss (src/loss_functions/TotalVariationLoss04.py:97-116)
  entropy_loss                           MassConcentrationLoss.compute_entropy_loss (src/loss_functions/MassConcentrationLoss06.py:63-83)

Factors are the texel-major parameters of srf_hip.tensorf (a [1, c, H, W] view of [H, W, c] storage).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._checks import dev_tensor
from .losses import _select_bytes
from .tensorf import texel_major_param


def _ints(vs):
    arr = (C.c_int * len(vs))()
    for i, v in enumerate(vs):
        arr[i] = int(v)
    return arr


def _ptrs(ts):
    arr = (C.c_void_p * len(ts))()
    for i, t in enumerate(ts):
        arr[i] = None if t is None else t.data_ptr()
    return arr


def _storage(p: torch.Tensor) -> torch.Tensor:
    """The texel-major [H, W, c] storage of a [1, c, H, W] factor (a copy if it is not texel-major)."""
    return p.detach().permute(0, 2, 3, 1)[0].contiguous()


def dense_grid_points(grid, bbox, device) -> torch.Tensor:
    """[gx * gy * gz, 3] points of the dense evaluation grid (meshgrid 'ij' of linspace(0, 1) per axis,
    bb0 (1 - s) + bb1 s; :851-857).  bbox: 6 host floats (min xyz, max xyz)."""
    gx, gy, gz = (int(v) for v in grid)
    pts = torch.empty((gx * gy * gz, 3), dtype=torch.float32, device=device)
    b = (C.c_float * 6)(*[float(v) for v in bbox])
    _lib.call('srf_dense_grid_points', gx, gy, gz, b, pts.data_ptr(), _lib.stream_ptr(device))
    return pts


def alpha_mask_build(sigma: torch.Tensor, grid, length: float, threshold: float):
    """From the densities of the dense grid (sigma [gx * gy * gz] in grid order): the new occupancy volume
    [gz, gy, gx] (0 / 1) and the occupied index bounds int32 [6] (min x, y, z, max x, y, z; min > max when
    nothing is occupied)."""
    gx, gy, gz = (int(v) for v in grid)
    sigma = dev_tensor(sigma.detach().reshape(-1), 'sigma', (gx * gy * gz,))
    alpha = torch.empty((gz, gy, gx), dtype=torch.float32, device=sigma.device)
    bounds = torch.empty(6, dtype=torch.int32, device=sigma.device)
    _lib.call('srf_alpha_mask_build', gx, gy, gz, sigma.data_ptr(), float(length), float(threshold), alpha.data_ptr(),
              bounds.data_ptr(), _lib.stream_ptr(sigma.device))
    return alpha, bounds


def upsample_factor(p: torch.Tensor, rows: int, cols: int) -> torch.nn.Parameter:
    """A [1, c, H, W] factor resampled to [1, c, rows, cols] (bilinear, align_corners = True) as a new
    texel-major parameter."""
    src = dev_tensor(_storage(p), 'factor')
    h, w, c = src.shape
    out = texel_major_param(c, int(rows), int(cols), device=src.device)
    dst = out.detach().permute(0, 2, 3, 1)
    _lib.call('srf_factor_upsample', c, h, w, src.data_ptr(), int(rows), int(cols), dst.data_ptr(),
              _lib.stream_ptr(src.device))
    return out


def crop_factor(p: torch.Tensor, rows: slice, cols: slice) -> torch.nn.Parameter:
    """p[..., rows, cols] as a new texel-major parameter (the shrink of :1304-1321)."""
    part = p.detach()[..., rows, cols]
    return texel_major_param(part.shape[1], part.shape[2], part.shape[3], part, device=p.device)


class _TV(torch.autograd.Function):
    @staticmethod
    def forward(ctx, scale, *planes):
        stores = [dev_tensor(_storage(p), f'planes[{i}]') for i, p in enumerate(planes)]
        dims = [s.shape for s in stores]
        dev = stores[0].device
        n = len(stores)
        ws = torch.empty(_lib.lib().srf_tv_workspace_bytes(n), dtype=torch.uint8, device=dev)
        loss = torch.empty((), dtype=torch.float32, device=dev)
        _lib.call('srf_tv_fwd', n, _ints([d[2] for d in dims]), _ints([d[0] for d in dims]),
                  _ints([d[1] for d in dims]), _ptrs(stores), float(scale), ws.data_ptr(), loss.data_ptr(),
                  _lib.stream_ptr(dev))
        ctx.scale = scale
        ctx.save_for_backward(*stores)
        return loss

    @staticmethod
    def backward(ctx, g_loss):
        stores = ctx.saved_tensors
        dims = [s.shape for s in stores]
        grads = [torch.empty((1, d[0], d[1], d[2]), dtype=torch.float32, device=s.device) for d, s in zip(dims, stores)]
        _lib.call('srf_tv_bwd', len(stores), _ints([d[2] for d in dims]), _ints([d[0] for d in dims]),
                  _ints([d[1] for d in dims]), _ptrs(stores), float(ctx.scale), g_loss.contiguous().data_ptr(),
                  _ptrs(grads), _lib.stream_ptr(stores[0].device))
        return (None, *[g.permute(0, 3, 1, 2) for g in grads])


def tv_loss(components, scale: float = 1.0) -> torch.Tensor:
    """scale * sum over the [1, c, H, W] components of 2 (mean squared difference along H + along W):
    compute_tv_loss with scale = the iteration weight (times the loss weight).  One launch for the
    forward of all components and one for the backward."""
    comps = list(components)
    if not comps or len(comps) > 8:
        raise _lib.SrfError('tv_loss takes 1 to 8 components')
    return _TV.apply(float(scale), *comps)


class _Entropy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, weights, select, sizes):
        n, s = weights.shape
        dev = weights.device
        nb = len(sizes)
        binned = torch.empty((n, nb), dtype=torch.float32, device=dev)
        ray_loss = torch.empty(n, dtype=torch.float32, device=dev)
        loss = torch.empty((), dtype=torch.float32, device=dev)
        count = torch.empty(1, dtype=torch.int32, device=dev)
        _lib.call('srf_entropy_fwd', n, s, weights.data_ptr(), _lib.ptr(select), nb, _ints(sizes), binned.data_ptr(),
                  ray_loss.data_ptr(), loss.data_ptr(), count.data_ptr(), _lib.stream_ptr(dev))
        ctx.sizes = sizes
        ctx.shape = (n, s)
        ctx.has_select = select is not None
        ctx.save_for_backward(binned, count, select if select is not None else torch.empty(0))
        ctx.mark_non_differentiable(ray_loss)
        return loss, ray_loss

    @staticmethod
    def backward(ctx, g_loss, _g_map):
        binned, count, select = ctx.saved_tensors
        n, s = ctx.shape
        d = torch.empty((n, s), dtype=torch.float32, device=binned.device)
        _lib.call('srf_entropy_bwd', n, s, _lib.ptr(select if ctx.has_select else None), len(ctx.sizes),
                  _ints(ctx.sizes), binned.data_ptr(), count.data_ptr(), g_loss.contiguous().data_ptr(), d.data_ptr(),
                  _lib.stream_ptr(binned.device))
        return d, None, None


def entropy_bin_sizes(n_samples: int, num_bins: int) -> list[int]:
    """Bin sizes of compute_entropy_loss (MassConcentrationLoss06.py:67-71): n_samples // num_bins each, the
    remainder one each to the bins linspace(0, num_bins - 1, remainder) picks."""
    per, rem = n_samples // num_bins, n_samples % num_bins
    bigger = set(torch.linspace(0, num_bins - 1, rem).long().tolist())
    return [per + 1 if i in bigger else per for i in range(num_bins)]


def entropy_loss(weights: torch.Tensor, indices_mask, num_bins: int, return_maps: bool = False):
    """Mean over the selected rays of sum over bins of entr(bin weight + 1e-3): weights [n_rays, n_samples].
    With return_maps also the per-ray values [n_rays] (0 on rays not selected)."""
    weights = dev_tensor(weights, 'weights')
    if weights.dim() != 2:
        raise _lib.SrfError(f'weights must be [rays, samples], got {tuple(weights.shape)}')
    sizes = entropy_bin_sizes(weights.shape[1], int(num_bins))
    sel = _select_bytes(indices_mask, weights.shape[0], weights.device)
    loss, ray_loss = _Entropy.apply(weights, sel, sizes)
    return (loss, ray_loss) if return_maps else loss
