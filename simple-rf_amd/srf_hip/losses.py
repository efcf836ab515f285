"""Fused training losses on the HIP kernels (srf_mse_terms_*, srf_depth_consistency_*), each a
torch.autograd.Function with a hand-written backward.

Mirrors the reference's loss arithmetic:
  mse_terms            MSE.compute_mse summed over models_to_regularize (src/loss_functions/MSE14.py:29-94),
                       SparseDepthMSE.compute_depth_loss (src/loss_functions/SparseDepthMSE14.py:29-95),
                       CoarseFineConsistencyLoss.compute_loss_sd (src/loss_functions/CoarseFineConsistencyLoss34.py:166-181)
  depth_consistency    compute_loss_nerf of AugmentationsDepthLoss11 (:108-181) and
                       CoarseFineConsistencyLoss34 (:85-164)

Selections (the reference's indices_mask_*: a boolean mask or an index tensor) become byte masks over
the batch; the selected-ray counts stay on the device, so nothing here synchronises the host.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._checks import dev_tensor


def _select_bytes(mask, n: int, device) -> torch.Tensor | None:
    """The reference's indices_mask (bool [n], or int indices) as a byte mask on the device (None = all)."""
    if mask is None:
        return None
    if not isinstance(mask, torch.Tensor):
        mask = torch.as_tensor(mask)
    mask = mask.to(device)
    if mask.dtype == torch.bool:
        if mask.shape != (n,):
            raise _lib.SrfError(f'selection mask must be [{n}], got {tuple(mask.shape)}')
        return mask.contiguous().view(torch.uint8)
    out = torch.zeros(n, dtype=torch.uint8, device=device)
    out[mask.long()] = 1
    return out


def _ptrs(ts):
    arr = (C.c_void_p * len(ts))()
    for i, t in enumerate(ts):
        arr[i] = None if t is None else t.data_ptr()
    return arr


def _floats(vs):
    arr = (C.c_float * len(vs))()
    for i, v in enumerate(vs):
        arr[i] = float(v)
    return arr


class _MSETerms(torch.autograd.Function):
    @staticmethod
    def forward(ctx, target, select, weights, maps_wanted, *preds):
        n, c = target.shape[0], (target.shape[1] if target.dim() == 2 else 1)
        dev = target.device
        loss = torch.empty((), dtype=torch.float32, device=dev)
        count = torch.empty(1, dtype=torch.int32, device=dev)
        maps = torch.empty((len(preds), n), dtype=torch.float32, device=dev) if maps_wanted else None
        _lib.call('srf_mse_terms_fwd', len(preds), _ptrs(preds), _floats(weights), target.data_ptr(),
                  _lib.ptr(select), n, c, loss.data_ptr(), None, _lib.ptr(maps), count.data_ptr(),
                  _lib.stream_ptr(dev))
        ctx.weights = weights
        ctx.n, ctx.c = n, c
        ctx.save_for_backward(target, select if select is not None else torch.empty(0), count, *preds)
        ctx.has_select = select is not None
        if maps is not None:
            ctx.mark_non_differentiable(maps)
        return loss, maps

    @staticmethod
    def backward(ctx, g_loss, _g_maps):
        target, select, count, *preds = ctx.saved_tensors
        select = select if ctx.has_select else None
        need_t = ctx.needs_input_grad[0]
        d = [torch.empty_like(p) if (ctx.needs_input_grad[4 + i] or need_t) else None for i, p in enumerate(preds)]
        if any(x is not None for x in d):
            _lib.call('srf_mse_terms_bwd', len(preds), _ptrs(preds), _floats(ctx.weights), target.data_ptr(),
                      _lib.ptr(select), ctx.n, ctx.c, count.data_ptr(), g_loss.contiguous().data_ptr(), _ptrs(d),
                      _lib.stream_ptr(target.device))
        # the target enters every term as (pred - target): its gradient is minus the sum of the terms'
        d_t = None
        if need_t:
            d_t = -d[0] if len(d) == 1 else -torch.stack(d).sum(0)
        d = [x if ctx.needs_input_grad[4 + i] else None for i, x in enumerate(d)]
        return (d_t, None, None, None, *d)


def mse_terms(preds, target, indices_mask=None, weights=None, return_maps: bool = False, target_grad: bool = False):
    """sum_t weights[t] * mean over the selected rays of the per-ray channel mean of (preds[t] - target)^2.

    preds: up to 8 tensors shaped like target ([n, 3] colours or [n] depths).  The target gets no
    gradient (it is data, or a detached depth) unless target_grad.  Returns the loss (0-dim) and, with
    return_maps, the per-ray maps [n_terms, n] (0 on rays not selected)."""
    if not preds:
        raise _lib.SrfError('mse_terms needs at least one prediction')
    target = dev_tensor(target if target_grad else target.detach(), 'target')
    shape = target.shape
    t2 = target.reshape(shape[0], -1)
    ps = [dev_tensor(p, f'preds[{i}]', tuple(shape)).reshape(t2.shape) for i, p in enumerate(preds)]
    weights = [1.0] * len(ps) if weights is None else list(weights)
    sel = _select_bytes(indices_mask, shape[0], target.device)
    loss, maps = _MSETerms.apply(t2, sel, weights, return_maps, *ps)
    return (loss, maps) if return_maps else loss


class _DepthConsistency(torch.autograd.Function):
    @staticmethod
    def forward(ctx, depth1, depth2, rays_o, rays_d, pixel_id, select, poses, intrinsic, images, patch, thr, rule):
        n = depth1.shape[0]
        dev = depth1.device
        map1 = torch.empty(n, dtype=torch.float32, device=dev)
        map2 = torch.empty(n, dtype=torch.float32, device=dev)
        flags = torch.empty(n, dtype=torch.uint8, device=dev)
        loss = torch.empty((), dtype=torch.float32, device=dev)
        count = torch.empty(1, dtype=torch.int32, device=dev)
        v, h, w, c = images.shape
        _lib.call('srf_depth_consistency_fwd', n, depth1.data_ptr(), depth2.data_ptr(), rays_o.data_ptr(),
                  rays_d.data_ptr(), pixel_id.data_ptr(), _lib.ptr(select), v, poses.data_ptr(), intrinsic.data_ptr(),
                  images.data_ptr(), h, w, c, int(patch[0]), int(patch[1]), float(thr), int(rule), map1.data_ptr(),
                  map2.data_ptr(), flags.data_ptr(), loss.data_ptr(), count.data_ptr(), _lib.stream_ptr(dev))
        ctx.save_for_backward(depth1, depth2, flags, count)
        ctx.mark_non_differentiable(map1, map2, flags)
        return loss, map1, map2, flags

    @staticmethod
    def backward(ctx, g_loss, *_):
        depth1, depth2, flags, count = ctx.saved_tensors
        d1 = torch.empty_like(depth1) if ctx.needs_input_grad[0] else None
        d2 = torch.empty_like(depth2) if ctx.needs_input_grad[1] else None
        if d1 is not None or d2 is not None:
            _lib.call('srf_depth_consistency_bwd', depth1.shape[0], depth1.data_ptr(), depth2.data_ptr(),
                      flags.data_ptr(), count.data_ptr(), g_loss.contiguous().data_ptr(), _lib.ptr(d1), _lib.ptr(d2),
                      _lib.stream_ptr(depth1.device))
        return (d1, d2) + (None,) * 10


def depth_consistency(depth1, depth2, indices_mask_nerf, rays_o, rays_d, extrinsics, gt_images, pixel_ids,
                      intrinsics, resolution, patch_size=(5, 5), rmse_threshold=0.1, both_invalid_rule=True):
    """compute_loss_nerf of AugmentationsDepthLoss11 (both_invalid_rule True) or CoarseFineConsistencyLoss34
    (False) on the fused kernel.  Returns (loss, map1, map2, flags): the maps are over the whole batch
    (0 on rays not selected; the reference returns them over the selected rays), flags bit 0 / 1 are the
    gate masks mask1 / mask2.  intrinsics: per-ray [n,3,3] as the model outputs them (the first entry is
    used, as CommonUtils04.reproject does) or one [3,3]."""
    n = depth1.shape[0]
    depth1 = dev_tensor(depth1, 'depth1', (n,))
    dev = depth1.device
    depth2 = dev_tensor(depth2, 'depth2', (n,))
    rays_o = dev_tensor(rays_o.detach(), 'rays_o', (n, 3))
    rays_d = dev_tensor(rays_d.detach(), 'rays_d', (n, 3))
    pid = pixel_ids.to(dev).to(torch.int32).contiguous()
    if pid.shape != (n, 3):
        raise _lib.SrfError(f'pixel_ids must be [{n},3], got {tuple(pid.shape)}')
    poses = dev_tensor(extrinsics.detach(), 'extrinsics')
    if poses.dim() != 3 or poses.shape[1:] != (4, 4):
        raise _lib.SrfError(f'extrinsics must be [views,4,4], got {tuple(poses.shape)}')
    k = intrinsics.detach()
    k = (k[0] if k.dim() == 3 else k).to(dev).to(torch.float32).contiguous()
    images = dev_tensor(gt_images.detach(), 'gt_images')
    if images.dim() != 4 or images.shape[0] != poses.shape[0]:
        raise _lib.SrfError(f'gt_images must be [views,h,w,c] with one image per pose, got {tuple(images.shape)}')
    if tuple(int(x) for x in resolution) != tuple(images.shape[1:3]):
        raise _lib.SrfError(f'resolution {tuple(resolution)} does not match the images {tuple(images.shape[1:3])}')
    sel = _select_bytes(indices_mask_nerf, n, dev)
    return _DepthConsistency.apply(depth1, depth2, rays_o, rays_d, pid, sel, poses, k, images, tuple(patch_size),
                                   rmse_threshold, both_invalid_rule)
