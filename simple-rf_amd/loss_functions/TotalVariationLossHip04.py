"""Total variation of the augmented TensoRF models' factors on the HIP kernels: class TotalVariationLoss
of src/loss_functions/TotalVariationLoss04.py:20-123.

Per augmented model and factor kind (density, colour) one srf_tv_fwd launch covers the three
components (and one srf_tv_bwd launch their gradients), in place of the reference's slices,
differences, squares and two sums per component.  As in the reference, only the augmented models'
factors are regularised, VM models on their matrices and CP models on their vectors."""
from __future__ import annotations

import torch

from srf_hip import tensorf_mods

REF_NAME = 'TotalVariationLoss04'  # loss-map keys keep the reference module's name


class TotalVariationLossHip:
    def __init__(self, configs: dict, loss_configs: dict):
        self.configs = configs
        self.loss_configs = loss_configs
        self.augmentations_needed = 'augmentations' in self.configs['model']
        self.lr_decay_ratio, self.lr_decay_iters = self.get_lr_decay_ratio_and_iters(configs)

    def get_lr_decay_ratio_and_iters(self, configs):
        """TotalVariationLoss04.py:28-37: the main optimiser's decay ratio and horizon."""
        ratio, iters = None, None
        for opt in configs['optimizers']:
            if opt['name'] == 'optimizer_main':
                ratio = opt['lr_decay_ratio']
                iters = opt['lr_decay_iters'] if opt['lr_decay_iters'] is not None else self.configs['num_iterations']
        return ratio, iters

    def get_iter_weight(self, iter_num):
        """TotalVariationLoss04.py:120-122."""
        return self.lr_decay_ratio ** ((iter_num + 1) / self.lr_decay_iters)

    @staticmethod
    def get_components(model):
        """TotalVariationLoss04.py:84-94."""
        name = model.__class__.__name__
        if name == 'CpDecomposedTensor':
            return model.vectors_density, model.vectors_color
        if name == 'VmDecomposedTensor':
            return model.matrices_density, model.matrices_color
        raise RuntimeError(f'total variation of a {name} is not defined')

    def compute_loss(self, input_dict: dict, output_dict: dict, model, return_loss_maps: bool = False):
        """TotalVariationLoss04.py:39-82."""
        total = torch.zeros((), dtype=torch.float32, device=input_dict['target_rgb'].device)
        maps = {}
        if self.augmentations_needed:
            iter_weight = self.get_iter_weight(input_dict['iter_num'])
            module = getattr(model, 'module', model)
            for aug_cfg in self.configs['model']['augmentations']:
                name = aug_cfg['name']
                found = [a for a in module.augmented_models if a['name'] == name]
                if len(found) != 1:
                    raise RuntimeError(f'augmented model {name} is not unique in the model')
                for stage in ('coarse', 'fine'):
                    if f'{stage}_model' not in aug_cfg:
                        continue
                    density, color = self.get_components(found[0][f'{stage}_model'])
                    total = total + tensorf_mods.tv_loss(density, iter_weight * self.loss_configs['weight_density'])
                    total = total + tensorf_mods.tv_loss(color, iter_weight * self.loss_configs['weight_color'])
        out = {'loss_value': total}
        if return_loss_maps:
            out['loss_maps'] = maps
        return out
