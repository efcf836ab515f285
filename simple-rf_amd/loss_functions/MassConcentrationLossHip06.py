"""Binned-entropy mass concentration of the augmented models' weights on the HIP kernels: class
MassConcentrationLoss of src/loss_functions/MassConcentrationLoss06.py:22-88.

Per augmented model and stage one srf_entropy_fwd launch bins, scores and averages the selected
rays (one srf_entropy_bwd launch gives the weights' gradient), in place of the reference's
selection, split, per-bin sums, entr and means."""
from __future__ import annotations

import torch

from srf_hip import tensorf_mods

REF_NAME = 'MassConcentrationLoss06'  # loss-map keys keep the reference module's name


class MassConcentrationLossHip:
    def __init__(self, configs: dict, loss_configs: dict):
        self.configs = configs
        self.loss_configs = loss_configs
        self.augmentations_needed = 'augmentations' in self.configs['model']

    def compute_loss(self, input_dict: dict, output_dict: dict, model, return_loss_maps: bool = False):
        """MassConcentrationLoss06.py:29-61."""
        total = torch.zeros((), dtype=torch.float32, device=input_dict['target_rgb'].device)
        maps = {}
        sel = input_dict['indices_mask_nerf']
        if self.augmentations_needed:
            module = getattr(model, 'module', model)
            for aug_cfg in self.configs['model']['augmentations']:
                name = aug_cfg['name']
                if len([a for a in module.augmented_models if a['name'] == name]) != 1:
                    raise RuntimeError(f'augmented model {name} is not unique in the model')
                for stage in ('coarse', 'fine'):
                    if f'{stage}_model' not in aug_cfg:
                        continue
                    weights = output_dict[f'{name}_weights_{stage}']
                    res = tensorf_mods.entropy_loss(weights, sel, self.loss_configs['num_bins'],
                                                    return_maps=return_loss_maps)
                    if return_loss_maps:
                        loss, per_ray = res
                        # the reference's map holds the selected rays only, in ray order
                        maps[f'{REF_NAME}_{name}_{stage}'] = per_ray[sel]
                    else:
                        loss = res
                    total = total + loss
        out = {'loss_value': total}
        if return_loss_maps:
            out['loss_maps'] = maps
        return out
