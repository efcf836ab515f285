"""Squared difference of the coarse and fine depths on the fused HIP loss kernel: class
CoarseFineConsistencyLoss of src/loss_functions/CoarseFineConsistencyLoss33.py:16-57 (NDC depths when
the data loader is NDC, all rays, both depths differentiated)."""
from __future__ import annotations

import torch

from srf_hip import losses

REF_NAME = 'CoarseFineConsistencyLoss33'  # loss-map keys keep the reference module's name


class CoarseFineConsistencyLossHip:
    def __init__(self, configs: dict, loss_configs: dict):
        self.configs = configs
        self.loss_configs = loss_configs
        self.coarse_model_needed = 'coarse_model' in configs['model']
        self.fine_model_needed = 'fine_model' in configs['model']
        self.ndc = configs['data_loader']['ndc']

    def compute_loss(self, input_dict: dict, output_dict: dict, model, return_loss_maps: bool = True):
        """CoarseFineConsistencyLoss33.py:24-57: mean over the rays of (depth_coarse - depth_fine)^2."""
        if not (self.coarse_model_needed and self.fine_model_needed):
            return {'loss_value': torch.zeros((), dtype=torch.float32, device=input_dict['target_rgb'].device)}
        key = 'depth_ndc' if self.ndc else 'depth'
        coarse, fine = output_dict[f'{key}_coarse'], output_dict[f'{key}_fine']
        res = losses.mse_terms([coarse], fine, None, return_maps=return_loss_maps, target_grad=True)
        if not return_loss_maps:
            return {'loss_value': res}
        loss, maps = res
        return {'loss_value': loss, 'loss_maps': {REF_NAME: maps[0]}}
