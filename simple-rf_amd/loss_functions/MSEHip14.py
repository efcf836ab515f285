"""Photometric MSE on the fused HIP loss kernel: class MSE of src/loss_functions/MSE14.py:19-94.

All regularised renders (main and augmented, coarse and fine) go to one srf_mse_terms launch (and one
backward launch) instead of the reference's gather, subtraction, square and two means per model."""
from __future__ import annotations

import torch

from loss_functions._regularized import regularized_outputs
from srf_hip import losses

REF_NAME = 'MSE14'  # loss-map keys keep the reference module's name


class MSEHip:
    def __init__(self, configs: dict, loss_configs: dict):
        self.configs = configs
        self.loss_configs = loss_configs
        self.models_to_regularize = loss_configs['models_to_regularize']

    def compute_loss(self, input_dict: dict, output_dict: dict, model, return_loss_maps: bool = False):
        """src/loss_functions/MSE14.py:29-79."""
        terms = regularized_outputs(self.configs, self.models_to_regularize, 'rgb')
        target = input_dict['target_rgb']
        if not terms:
            return {'loss_value': torch.zeros((), dtype=torch.float32, device=target.device)}
        preds = [output_dict[k] for _, k, _ in terms]
        res = losses.mse_terms(preds, target, input_dict['indices_mask_nerf'], [w for _, _, w in terms],
                               return_maps=return_loss_maps)
        if not return_loss_maps:
            return {'loss_value': res}
        loss, maps = res
        sel = input_dict['indices_mask_nerf']
        return {'loss_value': loss,
                'loss_maps': {f'{REF_NAME}_{s}': maps[i][sel] for i, (s, _, _) in enumerate(terms)}}
