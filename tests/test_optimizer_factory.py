"""The drop-in optimiser factory builds the fused Adam from the reference's optimiser configs
(src/optimizers/OptimizerFactory02.py:9-21), without a GPU."""
import pytest
import torch

from optimizers.OptimizerFactoryHip02 import get_optimizer
from srf_hip import SrfError
from srf_hip.adam import AdamHip

CONFIGS = {'optimizers': [{'name': 'optimizer_main', 'beta1': 0.9, 'beta2': 0.99, 'lr_initial': 5e-4},
                          {'name': 'optimizer_extrinsics', 'beta1': 0.8, 'beta2': 0.999}]}


def test_factory_builds_the_fused_adam_from_the_reference_configs():
    w = torch.nn.Parameter(torch.zeros(3))
    opt = get_optimizer(CONFIGS, 'optimizer_main', [w])
    assert isinstance(opt, AdamHip)
    assert opt.param_groups[0]['lr'] == 5e-4 and opt.param_groups[0]['betas'] == (0.9, 0.99)
    opt2 = get_optimizer(CONFIGS, 'optimizer_extrinsics', [{'params': [w], 'lr': 1e-3}])
    assert opt2.param_groups[0]['lr'] == 1e-3 and opt2.param_groups[0]['betas'] == (0.8, 0.999)
    assert get_optimizer(CONFIGS, 'optimizer_intrinsics', [w]) is None


def test_cpu_parameters_are_refused():
    w = torch.nn.Parameter(torch.zeros(3))
    w.grad = torch.ones(3)
    with pytest.raises(SrfError):
        AdamHip([w]).step()
