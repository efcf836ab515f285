"""Reprojection-gated consistency between the coarse and fine depths, plus coarse-to-fine supervision on
the sparse-depth rays, on the fused HIP kernels: class CoarseFineConsistencyLoss of
src/loss_functions/CoarseFineConsistencyLoss34.py:12-199 (no rule for rays whose reprojections both leave
the image; the sparse-depth term pulls the coarse depth towards the detached fine depth)."""
from __future__ import annotations

import torch

from srf_hip import losses

REF_NAME = 'CoarseFineConsistencyLoss34'


class CoarseFineConsistencyLossHip:
    def __init__(self, configs: dict, loss_configs: dict):
        self.configs = configs
        self.loss_configs = loss_configs
        self.coarse_model_needed = 'coarse_model' in configs['model']
        self.fine_model_needed = 'fine_model' in configs['model']
        self.sparse_depth_needed = 'sparse_depth' in configs['data_loader']
        self.patch_size = tuple(loss_configs['patch_size'])
        self.rmse_threshold = loss_configs['rmse_threshold']

    def compute_loss(self, input_dict: dict, output_dict: dict, model, return_loss_maps: bool = False):
        """src/loss_functions/CoarseFineConsistencyLoss34.py:26-83."""
        total = torch.zeros((), dtype=torch.float32, device=input_dict['target_rgb'].device)
        if not self.coarse_model_needed or not self.fine_model_needed:
            return {'loss_value': total}
        cd = input_dict['common_data']
        sel = input_dict['indices_mask_nerf']
        loss, m1, m2, _ = losses.depth_consistency(
            output_dict['depth_coarse'], output_dict['depth_fine'], sel, output_dict['rays_o'],
            output_dict['rays_d'], output_dict['extrinsics_all'], cd['images'], input_dict['pixel_id'],
            output_dict['intrinsics'], cd['resolution'], self.patch_size, self.rmse_threshold, False)
        total = total + loss
        sd = input_dict.get('indices_mask_sparse_depth', None)
        if self.sparse_depth_needed and sd is not None:
            total = total + losses.mse_terms([output_dict['depth_coarse']], output_dict['depth_fine'].detach(), sd)
        out = {'loss_value': total}
        if return_loss_maps:
            out['loss_maps'] = {f'{REF_NAME}_coarse': m1[sel], f'{REF_NAME}_fine': m2[sel]}
        return out
