"""Oracle: the TensoRF model modifications and their regularisers, restated in plain PyTorch
(TEST INFRASTRUCTURE ONLY: tests/ and bench.py's baseline legs).

Follows, per function:
  update_alpha_mask      LowRankTensor.update_alpha_mask / compute_alpha, src/models/SimpleTensoRF09.py:849-897
  shrink_bounds          LowRankTensor.shrink_tensor, :899-915
  shrink_vm_factors      VmDecomposedTensor.shrink_tensor, :1304-1321
  upsample_vm_factors    VmDecomposedTensor.upsample_vectors_and_matrices, :1281-1302
  new_num_voxels         LowRankTensor.get_new_num_voxels, :837-847
  tv_loss                TotalVariationLoss.compute_tv_loss, src/loss_functions/TotalVariationLoss04.py:97-116
  entropy_loss           MassConcentrationLoss.compute_entropy_loss, src/loss_functions/MassConcentrationLoss06.py:63-83
Pinned by tests/test_oracle_tensorf_mods.py against vectors the reference's own classes wrote.
"""
from __future__ import annotations

import numpy
import torch
import torch.nn.functional as F

from . import tensorf as ot


def dense_grid_points(resolution, bounding_box: torch.Tensor) -> torch.Tensor:
    """[gx, gy, gz, 3] points bb0 (1 - s) + bb1 s on the linspace(0, 1) grid of each axis (:851-857)."""
    gx, gy, gz = (int(v) for v in resolution)
    samples = torch.stack(torch.meshgrid(torch.linspace(0, 1, gx), torch.linspace(0, 1, gy),
                                         torch.linspace(0, 1, gz), indexing='ij'), -1).to(bounding_box.dtype)
    return bounding_box[0] * (1 - samples) + bounding_box[1] * samples


def compute_alpha(planes, lines, bounding_box, xyz, length, alpha_volume=None, alpha_bbox=None,
                  density='ReLU', offset=-10.0):
    """compute_alpha (:872-890): 1 - exp(-sigma length), sigma 0 where the existing mask is 0."""
    if alpha_volume is not None:
        valid = ot.alpha_mask_sample(alpha_volume, alpha_bbox, xyz) > 0
    else:
        valid = torch.ones_like(xyz[:, 0], dtype=torch.bool)
    sigma = torch.zeros(xyz.shape[:-1], dtype=xyz.dtype)
    if valid.any():
        pn = ot.normalize_points(xyz, bounding_box)
        feat = ot.vm_density_features(planes, lines, pn[valid])
        sigma[valid] = ot.density_predictor(feat, density, offset)
    return 1 - torch.exp(-sigma * length).view(xyz.shape[:-1])


def update_alpha_mask(planes, lines, bounding_box, resolution, step_size, threshold, alpha_volume=None,
                      alpha_bbox=None, density='ReLU', offset=-10.0):
    """update_alpha_mask (:849-870): returns (alpha volume [gz, gy, gx] of 0 / 1, new bounding box [2,3])."""
    grid = tuple(int(v) for v in resolution)
    xyz = dense_grid_points(grid, bounding_box)
    alpha = compute_alpha(planes, lines, bounding_box, xyz.view(-1, 3), step_size, alpha_volume, alpha_bbox,
                          density, offset).view(grid)
    xyz = xyz.transpose(0, 2).contiguous()
    alpha = alpha.clamp(0, 1).transpose(0, 2).contiguous()[None, None]
    alpha = F.max_pool3d(alpha, kernel_size=3, padding=1, stride=1).view(grid[::-1])
    alpha[alpha >= threshold] = 1
    alpha[alpha < threshold] = 0
    valid = xyz[alpha > 0.5]
    return alpha, torch.stack((valid.amin(0), valid.amax(0)))


def shrink_bounds(bounding_box_new, bounding_box, resolution, alpha_resolution):
    """LowRankTensor.shrink_tensor (:899-915): voxel range [t_l, b_r) kept on each axis, the new
    resolution and the new bounding box (snapped to voxel positions when the mask's resolution differs)."""
    voxel_length = (bounding_box[1] - bounding_box[0]) / (resolution - 1)
    xyz_min, xyz_max = bounding_box_new
    t_l, b_r = (xyz_min - bounding_box[0]) / voxel_length, (xyz_max - bounding_box[0]) / voxel_length
    t_l, b_r = torch.round(torch.round(t_l)).long(), torch.round(b_r).long() + 1
    b_r = torch.stack([b_r, resolution]).amin(0)
    if not torch.equal(alpha_resolution, resolution):
        t_l_r, b_r_r = t_l / (resolution - 1), (b_r - 1) / (resolution - 1)
        corrected = torch.zeros_like(bounding_box_new)
        corrected[0] = (1 - t_l_r) * bounding_box[0] + t_l_r * bounding_box[1]
        corrected[1] = (1 - b_r_r) * bounding_box[0] + b_r_r * bounding_box[1]
        bounding_box_new = corrected
    return t_l, b_r, b_r - t_l, bounding_box_new


def shrink_vm_factors(planes, lines, t_l, b_r):
    """VmDecomposedTensor.shrink_tensor (:1304-1321): crop every plane and line to the kept range."""
    new_planes, new_lines = [], []
    for i in range(3):
        m0 = ot.VECTOR_AXES[i]
        new_lines.append(lines[i][..., t_l[m0]:b_r[m0], :])
        a0, a1 = ot.MATRIX_AXES[i]
        new_planes.append(planes[i][..., t_l[a1]:b_r[a1], t_l[a0]:b_r[a0]])
    return new_planes, new_lines


def upsample_vm_factors(planes, lines, resolution_new):
    """upsample_vectors_and_matrices (:1281-1302): bilinear, align_corners=True, to the new resolution."""
    new_planes, new_lines = [], []
    for i in range(3):
        a0, a1 = ot.MATRIX_AXES[i]
        v = ot.VECTOR_AXES[i]
        new_planes.append(F.interpolate(planes[i], size=(int(resolution_new[a1]), int(resolution_new[a0])),
                                        mode='bilinear', align_corners=True))
        new_lines.append(F.interpolate(lines[i], size=(int(resolution_new[v]), 1), mode='bilinear',
                                       align_corners=True))
    return new_planes, new_lines


def new_num_voxels(iter_num, upsampling_iters, num_voxels_initial, num_voxels_final) -> int:
    """get_new_num_voxels (:837-847): geometric interpolation of the voxel count."""
    k = upsampling_iters.index(iter_num) + 1
    lo, hi = numpy.log(num_voxels_initial), numpy.log(num_voxels_final)
    return int(numpy.round(numpy.exp(lo + (hi - lo) * k / len(upsampling_iters))))


def tv_loss(components, iter_weight):
    """compute_tv_loss (TotalVariationLoss04.py:97-116): per component tensor [1, c, H, W],
    2 (mean squared difference along H + along W), times the iteration weight, summed."""
    total = 0
    for comp in components:
        dh = torch.pow(comp[:, :, 1:, :] - comp[:, :, :-1, :], 2)
        dw = torch.pow(comp[:, :, :, 1:] - comp[:, :, :, :-1], 2)
        tv_h = dh.sum() / max(torch.numel(dh), 1)
        tv_w = dw.sum() / max(torch.numel(dw), 1)
        total = total + 2 * (tv_h + tv_w) * iter_weight
    return total


def entropy_bins(n_samples: int, num_bins: int):
    """Bin sizes of compute_entropy_loss (MassConcentrationLoss06.py:67-71): the remainder samples go,
    one each, to the bins linspace(0, num_bins - 1, remainder) selects."""
    per, rem = n_samples // num_bins, n_samples % num_bins
    bigger = torch.linspace(0, num_bins - 1, rem).long()
    return [(per + 1) if i in bigger else per for i in range(num_bins)]


def entropy_loss(weights, indices_mask, num_bins):
    """compute_entropy_loss (MassConcentrationLoss06.py:63-83): mean over the selected rays of
    sum over bins of entr(bin weight + 1e-3)."""
    w = weights[indices_mask]
    sizes = entropy_bins(w.shape[1], num_bins)
    binned = torch.stack([b.sum(dim=1) for b in torch.split(w, sizes, dim=1)], dim=1)
    per_ray = torch.sum(torch.special.entr(binned + 1e-3), dim=1)
    return torch.mean(per_ray) if per_ray.numel() > 0 else torch.zeros((), dtype=weights.dtype)
