"""get_optimizer of src/optimizers/OptimizerFactory02.py:9-21 returning the fused HIP Adam
(srf_hip.adam.AdamHip) in place of torch.optim.Adam: same config keys (name, lr_initial, beta1,
beta2), same parameter groups and state, one launch per 32 tensors per step."""
from __future__ import annotations

from srf_hip.adam import AdamHip


def get_optimizer(configs: dict, optimizer_name: str, model_params: list):
    optimizer = None
    for optimizer_configs in configs['optimizers']:
        if optimizer_configs['name'] != optimizer_name:
            continue
        betas = (optimizer_configs['beta1'], optimizer_configs['beta2'])
        if 'lr_initial' in optimizer_configs:
            optimizer = AdamHip(model_params, lr=optimizer_configs['lr_initial'], betas=betas)
        else:
            optimizer = AdamHip(model_params, betas=betas)
    return optimizer
