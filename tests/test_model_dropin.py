"""Drop-in model surface (CPU: construction only; the kernels are not called)."""
import pytest
import torch

import oracle.nerf as on
from models.ModelFactory02 import get_model
from srf_hip import configs as cfgs


def test_factory_loads_by_name_and_state_dict_matches_reference_layout():
    model = get_model(cfgs.simple_nerf_configs(), model_configs=cfgs.model_configs())
    sd = model.state_dict()
    ref = {f'coarse_model.{k}': v for k, v in on.mlp_param_shapes().items()}
    ref.update({f'fine_model.{k}': v for k, v in on.mlp_param_shapes().items()})
    ref.update({f'augmented_models_nn.0.coarse_model.{k}': v for k, v in on.mlp_param_shapes(sigma_pe_degree=3).items()})
    ref.update({f'augmented_models_nn.1.coarse_model.{k}': v
                for k, v in on.mlp_param_shapes(use_view_dirs=False, view_dependent_rgb=False).items()})
    for k, shape in ref.items():
        assert tuple(sd[k].shape) == tuple(shape), k
    mlp_keys = {k for k in sd if 'linear' in k}
    assert mlp_keys == set(ref)
    assert {'intrinsics_learner.initial_intrinsics', 'extrinsics_learner.initial_extrinsics',
            'extrinsics_learner.r', 'extrinsics_learner.t'} <= set(sd)


def test_unknown_model_name_raises():
    c = cfgs.simple_nerf_configs(name='SimpleNeRFHip17')
    c['model']['name'] = 'NoSuchModel17'
    with pytest.raises(ModuleNotFoundError):
        get_model(c, model_configs=cfgs.model_configs())


def test_unsupported_geometry_is_rejected():
    c = cfgs.simple_nerf_configs()
    c['model']['coarse_model']['points_net_width'] = 128
    with pytest.raises(NotImplementedError):
        get_model(c, model_configs=cfgs.model_configs())
