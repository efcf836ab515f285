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


def _tensorf_expected_shapes(prefix, comps_d, comps_c, res, feat=27, units=128, in_dim=30):
    """Parameter shapes VmDecomposedTensor creates (src/models/SimpleTensoRF09.py:1144-1165, 1389-1391)."""
    axes_m, axes_v = [[0, 1], [0, 2], [1, 2]], [2, 1, 0]
    out = {}
    for kind, comps in (('density', comps_d), ('color', comps_c)):
        for i in range(3):
            a0, a1 = axes_m[i]
            out[f'{prefix}matrices_{kind}.{i}'] = (1, comps[i], res[a1], res[a0])
            out[f'{prefix}vectors_{kind}.{i}'] = (1, comps[i], res[axes_v[i]], 1)
    out[f'{prefix}basis_matrix_color.weight'] = (feat, sum(comps_c))
    out[f'{prefix}color_predictor.mlp.0.weight'] = (units, in_dim)
    out[f'{prefix}color_predictor.mlp.2.weight'] = (units, units)
    out[f'{prefix}color_predictor.mlp.4.weight'] = (3, units)
    return out


def test_tensorf_dropin_state_dict_matches_reference_layout():
    import oracle.tensorf as ot
    c = cfgs.simple_tensorf_configs(coarse=dict(num_voxels_initial=32 ** 3), augmented=dict(num_voxels_initial=16 ** 3))
    model = get_model(c, model_configs=cfgs.model_configs())
    sd = model.state_dict()
    tc = c['model']['coarse_model']
    res = ot.compute_resolution_in_voxels(tc['num_voxels_initial'], torch.tensor(tc['bounding_box'])).tolist()
    assert model.coarse_model.resolution.tolist() == res
    assert int(model.coarse_model.num_samples) == ot.compute_num_samples(torch.tensor(res), 0.5, 1e6)
    exp = _tensorf_expected_shapes('coarse_model.', [16, 4, 4], [48, 12, 12], res)
    ac = c['model']['augmentations'][0]['coarse_model']
    ares = ot.compute_resolution_in_voxels(ac['num_voxels_initial'], torch.tensor(ac['bounding_box'])).tolist()
    exp.update(_tensorf_expected_shapes('augmented_models_nn.0.coarse_model.', [4, 4, 4], [48, 12, 12], ares))
    for k, shape in exp.items():
        assert tuple(sd[k].shape) == shape, k
    for k in ('bounding_box', 'resolution', 'num_samples', 'bounding_box_size', 'voxel_length', 'step_size'):
        assert f'coarse_model.{k}' in sd
    # factors are stored texel-major (component axis innermost) under the reference's shape
    p = model.coarse_model.matrices_density[0]
    assert p.permute(0, 2, 3, 1).is_contiguous()


def test_tensorf_checkpoint_round_trip_keeps_values_and_layout():
    c = cfgs.simple_tensorf_configs(augmentations=False, coarse=dict(num_voxels_initial=16 ** 3))
    a = get_model(c, model_configs=cfgs.model_configs())
    b = get_model(c, model_configs=cfgs.model_configs())
    # a checkpoint in the reference's contiguous layout loads into the texel-major parameters
    sd = {k: v.contiguous() for k, v in a.state_dict().items()}
    b.load_state_dict(sd)
    for k, v in a.state_dict().items():
        assert torch.equal(b.state_dict()[k], v), k
    assert b.coarse_model.matrices_color[1].permute(0, 2, 3, 1).is_contiguous()


def test_tensorf_unsupported_options_raise():
    c = cfgs.simple_tensorf_configs(augmentations=False, coarse=dict(decomposition_type='Tucker'))
    with pytest.raises(RuntimeError):   # SimpleTensoRF09.py:542-543
        get_model(c, model_configs=cfgs.model_configs())
    c = cfgs.simple_tensorf_configs(augmentations=False, coarse=dict(num_voxels_initial=16 ** 3, predict_visibility=True))
    with pytest.raises(NotImplementedError):
        get_model(c, model_configs=cfgs.model_configs())
    c = cfgs.simple_tensorf_configs(augmentations=False, coarse=dict(num_voxels_initial=16 ** 3,
                                                                      num_components_density=(4, 2, 2)))
    with pytest.raises(NotImplementedError, match='groups of 4'):
        get_model(c, model_configs=cfgs.model_configs())


def _cp_configs(**kw):
    coarse = dict(decomposition_type='CandecompParafac', num_components_density=(16,), num_components_color=(48,),
                  num_voxels_initial=32 ** 3)
    coarse.update(kw)
    return cfgs.simple_tensorf_configs(augmentations=False, coarse=coarse)


def test_tensorf_cp_dropin_state_dict_matches_reference_layout():
    """CpDecomposedTensor (src/models/SimpleTensoRF09.py:979-996): three lines of rank
    num_components[0] per field along z, y, x, no planes, basis Linear(sum(comps) -> 27)."""
    import oracle.tensorf as ot
    c = _cp_configs()
    model = get_model(c, model_configs=cfgs.model_configs())
    sd = model.state_dict()
    tc = c['model']['coarse_model']
    res = ot.compute_resolution_in_voxels(tc['num_voxels_initial'], torch.tensor(tc['bounding_box'])).tolist()
    for kind, r in (('density', 16), ('color', 48)):
        for i in range(3):
            assert tuple(sd[f'coarse_model.vectors_{kind}.{i}'].shape) == (1, r, res[ot.VECTOR_AXES[i]], 1)
    assert not any('matrices_' in k for k in sd)
    assert tuple(sd['coarse_model.basis_matrix_color.weight'].shape) == (27, 48)
    assert model.coarse_model.vectors_color[2].permute(0, 2, 3, 1).is_contiguous()
    groups = model.coarse_model.get_trainable_parameters({'lr_initial_tensor': 0.02, 'lr_initial_network': 1e-3})
    assert [g['name'] for g in groups] == ['coarse_model_tensor_params', 'coarse_model_network_params']
    assert [tuple(p.shape) for p in groups[0]['params']] == \
        [(1, 16, res[v], 1) for v in ot.VECTOR_AXES] + [(1, 48, res[v], 1) for v in ot.VECTOR_AXES]
    b = get_model(_cp_configs(), model_configs=cfgs.model_configs())
    b.load_state_dict(sd)
    for k, v in sd.items():
        assert torch.equal(b.state_dict()[k], v), k
    assert b.coarse_model.vectors_density[1].permute(0, 2, 3, 1).is_contiguous()


@pytest.mark.parametrize('family', ['nerf', 'tensorf'])
def test_learned_camera_raises_instead_of_dropping_its_gradient(family):
    """A camera learner (learn_camera_*, src/models/SimpleNeRF17.py:831-842) makes the rays carry a
    gradient; the kernels give rays none, so the drop-in refuses before any kernel runs (on the CPU
    here) rather than training the camera with a zero gradient."""
    from srf_hip import synthetic
    if family == 'nerf':
        c = cfgs.simple_nerf_configs(precision='fp32')
    else:
        c = cfgs.simple_tensorf_configs(augmentations=False, coarse=dict(num_voxels_initial=16 ** 3))
    c['model']['learn_camera_rotation'] = True
    m = get_model(c, model_configs=cfgs.model_configs())
    m.train()
    b = synthetic.train_batch(16, seed=0)
    with pytest.raises(NotImplementedError, match='camera gradients'):
        m({'pixel_id': b['pixel_id'], 'num_frames': 3, 'iter_num': 0, 'sub_batch_index': 0})
