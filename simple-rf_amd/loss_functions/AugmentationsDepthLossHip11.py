"""Reprojection-gated depth consistency between the main and each augmented model on the fused HIP
kernel: class AugmentationsDepthLoss of src/loss_functions/AugmentationsDepthLoss11.py:18-199.

Per model pair one srf_depth_consistency launch does what the reference does in ~100 launches: the
nearest view, both reprojections, the three 5x5 ground-truth patches, both RMSEs, the gate (with the
rule for rays whose reprojections both leave the image) and the loss maps."""
from __future__ import annotations

import torch

from srf_hip import losses

REF_NAME = 'AugmentationsDepthLoss11'


class AugmentationsDepthLossHip:
    both_invalid_rule = True

    def __init__(self, configs: dict, loss_configs: dict):
        self.configs = configs
        self.loss_configs = loss_configs
        self.coarse_model_needed = 'coarse_model' in configs['model']
        self.fine_model_needed = 'fine_model' in configs['model']
        self.augmentations_needed = 'augmentations' in configs['model']
        self.patch_size = tuple(loss_configs['patch_size'])
        self.rmse_threshold = loss_configs['rmse_threshold']

    def _pair(self, depth1, depth2, input_dict, output_dict):
        cd = input_dict['common_data']
        return losses.depth_consistency(
            depth1, depth2, input_dict['indices_mask_nerf'], output_dict['rays_o'], output_dict['rays_d'],
            output_dict['extrinsics_all'], cd['images'], input_dict['pixel_id'], output_dict['intrinsics'],
            cd['resolution'], self.patch_size, self.rmse_threshold, self.both_invalid_rule)

    def compute_loss(self, input_dict: dict, output_dict: dict, model, return_loss_maps: bool = False):
        """src/loss_functions/AugmentationsDepthLoss11.py:29-75."""
        total = torch.zeros((), dtype=torch.float32, device=input_dict['target_rgb'].device)
        maps = {}
        sel = input_dict['indices_mask_nerf']
        if self.augmentations_needed:
            for stage, needed in (('coarse', self.coarse_model_needed), ('fine', self.fine_model_needed)):
                if not needed:
                    continue
                for aug in self.configs['model']['augmentations']:
                    if f'{stage}_model' not in aug:
                        continue
                    name = aug['name']
                    loss, m1, m2, _ = self._pair(output_dict[f'depth_{stage}'], output_dict[f'{name}_depth_{stage}'],
                                                 input_dict, output_dict)
                    total = total + loss
                    if return_loss_maps:
                        maps[f'{REF_NAME}_{name}_{stage}_main'] = m1[sel]
                        maps[f'{REF_NAME}_{name}_{stage}_augmented'] = m2[sel]
        out = {'loss_value': total}
        if return_loss_maps:
            out['loss_maps'] = maps
        return out
