"""Oracle: Simple-RF's training losses on the rendered rays, restated in plain PyTorch
(TEST INFRASTRUCTURE ONLY: tests/ and bench.py's baseline legs).

Each function follows the reference's loss code operation for operation (cited per function), so
that in float32 it reproduces the reference's rounding and in float64 it is the high-precision
reference.

The depth-consistency losses (AugmentationsDepthLoss11, CoarseFineConsistencyLoss34) zero the
masked entries of their inputs in place (compute_depth_mse, AugmentationsDepthLoss11.py:213-216),
and the second call's ground truth is a detached alias of the first call's prediction, so the first
call's zeroing reaches it.  The restatement keeps that data flow: the second term only counts rays
where both masks hold, and its gradient is zero elsewhere (test_oracle_losses pins this against
vectors written by the reference's own classes).
"""
from __future__ import annotations

import torch
import torch.nn.functional as F


# ---------------------------------------------------------------------- photometric
def mse(pred_rgb: torch.Tensor, target_rgb: torch.Tensor, indices_mask) -> torch.Tensor:
    """MSE.compute_mse, src/loss_functions/MSE14.py:81-94: mean over rays of the channel mean."""
    p = pred_rgb[indices_mask]
    t = target_rgb[indices_mask]
    m = torch.mean(torch.square(p - t), dim=1)
    return torch.mean(m) if p.numel() > 0 else torch.zeros((), dtype=pred_rgb.dtype, device=pred_rgb.device)


def sparse_depth_mse(pred_depth: torch.Tensor, gt_depth: torch.Tensor, indices_mask) -> torch.Tensor:
    """SparseDepthMSE.compute_depth_loss, src/loss_functions/SparseDepthMSE14.py:82-95."""
    p = pred_depth[indices_mask]
    t = gt_depth[indices_mask]
    return torch.mean(torch.square(p - t)) if p.numel() > 0 else torch.zeros((), dtype=pred_depth.dtype, device=pred_depth.device)


# -------------------------------------------------------------------- reprojection
def reproject(points: torch.Tensor, poses: torch.Tensor, intrinsics: torch.Tensor) -> torch.Tensor:
    """CommonUtils.reproject, src/utils/CommonUtils04.py:227-253: pixel position of each point in its
    view (poses [N,4,4] c2w, one per point; the first intrinsic matrix for every point)."""
    origins = poses[:, :3, 3]
    rotations = poses[:, :3, :3]
    rel = points - origins
    permuter = torch.eye(3, dtype=points.dtype, device=points.device)
    permuter[1:] *= -1
    k = intrinsics[:1]
    pos = (k @ permuter[None] @ rotations.transpose(1, 2) @ rel[..., None]).squeeze()
    return pos[:, :2] / pos[:, 2:]


def closest_view_ids(extrinsics: torch.Tensor, image_ids: torch.Tensor) -> torch.Tensor:
    """The view whose camera centre is nearest (second smallest distance, the smallest being the view
    itself), per ray: AugmentationsDepthLoss11.py:138-141 (kthvalue(distances, 2))."""
    origins = extrinsics[:, :3, 3]
    d = torch.sqrt(torch.sum(torch.square(origins[image_ids].unsqueeze(1).repeat([1, origins.shape[0], 1]) - origins),
                             dim=2))
    return torch.kthvalue(d, 2, dim=1)[1]


def patch_rmse(p1: torch.Tensor, p2: torch.Tensor) -> torch.Tensor:
    """compute_patch_rmse, AugmentationsDepthLoss11.py:183-197."""
    return torch.sqrt(torch.mean(torch.square(p1 - p2), dim=(1, 2, 3)))


def depth_consistency_nerf(depth1, depth2, indices_mask_nerf, rays_o, rays_d, extrinsics, gt_images, pixel_ids,
                           intrinsics, resolution, patch_size=(5, 5), rmse_threshold=0.1,
                           both_invalid_rule=True):
    """compute_loss_nerf of AugmentationsDepthLoss11 (src/loss_functions/AugmentationsDepthLoss11.py:108-181,
    both_invalid_rule=True) and of CoarseFineConsistencyLoss34 (CoarseFineConsistencyLoss34.py:85-164,
    both_invalid_rule=False): reproject each model's depth into the nearest view, gate it by the RMSE of
    the 5x5 ground-truth patches around the source pixel and the two reprojections, and pull the less
    accurate model's depth towards the more accurate one's.  Returns (loss, map1, map2, mask1, mask2)."""
    h, w = resolution
    px, py = patch_size
    hpx, hpy = px // 2, py // 2
    pixel_ids = pixel_ids.long()
    image_ids = pixel_ids[:, 0]
    rays_o = rays_o[indices_mask_nerf]
    rays_d = rays_d[indices_mask_nerf]
    depth1 = depth1[indices_mask_nerf]
    depth2 = depth2[indices_mask_nerf]
    closest = closest_view_ids(extrinsics, image_ids)
    image_ids_a = image_ids[indices_mask_nerf]
    pixel_ids_a = pixel_ids[indices_mask_nerf]
    image_ids_b = closest[indices_mask_nerf]
    poses_b = extrinsics[image_ids_b]
    points1a = rays_o + rays_d * depth1.unsqueeze(-1)
    points2a = rays_o + rays_d * depth2.unsqueeze(-1)
    pos1b = reproject(points1a.detach(), poses_b, intrinsics).round().long()
    pos2b = reproject(points2a.detach(), poses_b, intrinsics).round().long()
    x_a, y_a = pixel_ids_a[:, 1], pixel_ids_a[:, 2]
    x1b, y1b = pos1b[:, 0], pos1b[:, 1]
    x2b, y2b = pos2b[:, 0], pos2b[:, 1]
    valid_a = (x_a >= hpx) & (x_a < w - hpx) & (y_a >= hpy) & (y_a < h - hpy)
    valid_1b = (x1b >= hpx) & (x1b < w - hpx) & (y1b >= hpy) & (y1b < h - hpy)
    valid_2b = (x2b >= hpx) & (x2b < w - hpx) & (y2b >= hpy) & (y2b < h - hpy)
    x1c, y1c = torch.clip(x1b, 0, w - 1), torch.clip(y1b, 0, h - 1)
    x2c, y2c = torch.clip(x2b, 0, w - 1), torch.clip(y2b, 0, h - 1)
    n = image_ids_a.shape[0]
    c = gt_images.shape[3]
    pa = torch.zeros(n, py, px, c, dtype=gt_images.dtype, device=gt_images.device)
    p1 = torch.zeros(n, py, px, c, dtype=gt_images.dtype, device=gt_images.device)
    p2 = torch.zeros(n, py, px, c, dtype=gt_images.dtype, device=gt_images.device)
    # zero padding after the last row and column only (negative offsets wrap to the far end,
    # whose padding is zero, as F.pad(..., (0, 0, 0, hpy, 0, hpx)) then negative indexing does)
    padded = F.pad(gt_images, (0, 0, 0, hpy, 0, hpx), mode='constant', value=0)
    for i, yo in enumerate(range(-hpy, hpy + 1)):
        for j, xo in enumerate(range(-hpx, hpx + 1)):
            pa[:, i, j, :] = padded[image_ids_a, y_a + yo, x_a + xo]
            p1[:, i, j, :] = padded[image_ids_b, y1c + yo, x1c + xo]
            p2[:, i, j, :] = padded[image_ids_b, y2c + yo, x2c + xo]
    rmse1 = patch_rmse(pa, p1)
    rmse2 = patch_rmse(pa, p2)
    mask1 = ((rmse1 < rmse2) | (~valid_2b)) & (rmse1 < rmse_threshold) & valid_1b & valid_a
    mask2 = ((rmse2 < rmse1) | (~valid_1b)) & (rmse2 < rmse_threshold) & valid_2b & valid_a
    if both_invalid_rule:
        both_invalid = ~(torch.stack([valid_1b, valid_2b], dim=0).any(dim=0))
        mask1 = mask1 | (both_invalid & (depth1 > depth2))
        mask2 = mask2 | (both_invalid & (depth2 > depth1))
    # compute_depth_mse(depth1, depth2.detach(), mask2), then compute_depth_mse(depth2, depth1.detach(), mask1),
    # each zeroing its (aliased) inputs where its mask is false
    d1 = torch.where(mask2, depth1, torch.zeros_like(depth1))
    d2_gt = torch.where(mask2, depth2.detach(), torch.zeros_like(depth2))
    map1 = torch.square(d1 - d2_gt)
    both = mask1 & mask2
    d2 = torch.where(both, depth2, torch.zeros_like(depth2))
    d1_gt = torch.where(both, depth1.detach(), torch.zeros_like(depth1))
    map2 = torch.square(d2 - d1_gt)
    zero = torch.zeros((), dtype=depth1.dtype, device=depth1.device)
    loss = (torch.mean(map1) if n > 0 else zero) + (torch.mean(map2) if n > 0 else zero)
    return loss, map1, map2, mask1, mask2


def coarse_fine_sparse_depth(depth_coarse, depth_fine, indices_mask_sd) -> torch.Tensor:
    """CoarseFineConsistencyLoss.compute_loss_sd, src/loss_functions/CoarseFineConsistencyLoss34.py:166-181:
    coarse depth pulled towards the detached fine depth on the sparse-depth rays."""
    dc = depth_coarse[indices_mask_sd]
    df = depth_fine[indices_mask_sd].detach()
    return torch.mean(torch.square(dc - df)) if dc.numel() > 0 else torch.zeros((), dtype=depth_coarse.dtype, device=depth_coarse.device)
