"""Sparse-depth MSE on the fused HIP loss kernel: class SparseDepthMSE of
src/loss_functions/SparseDepthMSE14.py:18-95 (metric depths of every regularised model against the
sparse depths, on the sparse-depth rays), one srf_mse_terms launch for all models."""
from __future__ import annotations

import torch

from loss_functions._regularized import regularized_outputs
from srf_hip import losses


class SparseDepthMSEHip:
    def __init__(self, configs: dict, loss_configs: dict):
        self.configs = configs
        self.loss_configs = loss_configs
        self.models_to_regularize = loss_configs['models_to_regularize']
        self.ndc = configs['data_loader']['ndc']

    def compute_loss(self, input_dict: dict, output_dict: dict, model, return_loss_maps: bool = False):
        """src/loss_functions/SparseDepthMSE14.py:30-80 (no loss maps, as the reference)."""
        zero = torch.zeros((), dtype=torch.float32, device=input_dict['target_rgb'].device)
        # computed only for batches with sparse depth, not for full images (:36-37)
        if 'indices_mask_sparse_depth' not in input_dict:
            return {'loss_value': zero}
        terms = regularized_outputs(self.configs, self.models_to_regularize, 'depth')
        if not terms:
            return {'loss_value': zero}
        gt = input_dict['sparse_depth_values'][:, 0]
        loss = losses.mse_terms([output_dict[k] for _, k, _ in terms], gt, input_dict['indices_mask_sparse_depth'],
                                [w for _, _, w in terms])
        out = {'loss_value': loss}
        if return_loss_maps:
            out['loss_maps'] = {}
        return out
