"""Configuration dicts in the reference's schema for the benchmark workloads.

simple_nerf_configs() follows demo1a of src/NerfLlffTrainerTester10.py:249-457 (coarse 64 +
fine 128 samples, points- and views-augmented coarse models) with the model name switched to
the HIP drop-in.  model_configs() builds the per-scene dict (intrinsics, extrinsics,
resolution, near/far) that DataPreprocessor10 would write, from the synthetic camera.
"""
from __future__ import annotations

import copy

import numpy

from . import synthetic


def _mlp(num_samples=None, **kw):
    d = {'points_net_depth': 8, 'views_net_depth': 1, 'points_net_width': 256, 'views_net_width': 128,
         'points_positional_encoding_degree': 10, 'views_positional_encoding_degree': 4, 'use_view_dirs': True,
         'view_dependent_rgb': True, 'predict_visibility': False}
    if num_samples is not None:
        d['num_samples'] = num_samples
    d.update(kw)
    return d


def simple_nerf_configs(name='SimpleNeRFHip17', precision='bf16', augmentations=True, fine=True,
                        raw_noise_std=1.0, white_bkgd=False, perturb=True):
    model = {
        'name': name,
        'coarse_model': _mlp(64),
        'learn_camera_focal_length': False, 'learn_camera_rotation': False, 'learn_camera_translation': False,
        'chunk': 1 << 30, 'lindisp': False, 'netchunk': 16 * 1024, 'perturb': perturb,
        'raw_noise_std': raw_noise_std, 'white_bkgd': white_bkgd, 'mlp_precision': precision,
    }
    if fine:
        model['fine_model'] = _mlp(128)
    if augmentations:
        model['augmentations'] = [
            {'name': 'points_augmentation', 'coarse_model': _mlp(points_sigma_positional_encoding_degree=3)},
            {'name': 'views_augmentation', 'coarse_model': _mlp(use_view_dirs=False, view_dependent_rgb=False)},
        ]
        # the view-independent model has no view encoding (src/NerfLlffTrainerTester10.py:347-358)
        del model['augmentations'][1]['coarse_model']['views_positional_encoding_degree']
    return {'data_loader': {'ndc': True}, 'model': model, 'database': 'NeRF_LLFF'}


def model_configs(width=504, height=378, n_frames=3, seed=0):
    cam = synthetic.camera(width, height, n_frames, seed)
    return {'intrinsics': numpy.repeat(cam['K'].numpy()[None], n_frames, 0),
            'extrinsics': cam['poses'].numpy(), 'resolution': [height, width],
            'near': 1.0, 'far': 100.0, 'near_ndc': 0.0, 'far_ndc': 1.0}


def with_precision(configs, precision):
    c = copy.deepcopy(configs)
    c['model']['mlp_precision'] = precision
    return c


def _tensor(num_components_density=(16, 4, 4), num_components_color=(48, 12, 12),
            bounding_box=((-1.5, -1.67, -1.0), (1.5, 1.67, 1.0)), num_voxels_initial=128 ** 3 + 4, **kw):
    d = {'decomposition_type': 'VectorMatrix', 'num_samples_max': 1e6,
         'num_components_density': list(num_components_density), 'num_components_color': list(num_components_color),
         'bounding_box': [list(bounding_box[0]), list(bounding_box[1])], 'num_voxels_initial': num_voxels_initial,
         'num_voxels_final': 640 ** 3, 'tensor_upsampling_iters': [2000, 3000, 4000, 5500],
         'num_voxels_per_sample': 0.5, 'alpha_mask_update_iters': [2500], 'alpha_mask_threshold': 0.0001,
         'ray_marching_weight_threshold': 0.0001, 'use_view_dirs': True, 'view_dependent_color': True,
         'views_positional_encoding_degree': 0, 'features_positional_encoding_degree': 0,
         'features_dimension_color': 27, 'density_offset': -10, 'distance_scale': 25, 'density_predictor': 'ReLU',
         'color_predictor': 'MLP_Features', 'num_units_color_predictor': 128, 'predict_visibility': False}
    d.update(kw)
    return d


def simple_tensorf_configs(name='SimpleTensoRFHip09', augmentations=True, white_bkgd=False, perturb=True, ndc=True,
                           coarse=None, augmented=None):
    """The Simple-TensoRF demo of src/NerfLlffTrainerTester10.py:938-1011 (VM [16,4,4] / [48,12,12],
    a low-rank points-augmented model) with the model name switched to the HIP drop-in; ``coarse``
    and ``augmented`` override tensor-config keys."""
    model = {
        'name': name,
        'coarse_model': _tensor(**(coarse or {})),
        'learn_camera_focal_length': False, 'learn_camera_rotation': False, 'learn_camera_translation': False,
        'chunk': 1 << 30, 'lindisp': False, 'perturb': perturb, 'white_bkgd': white_bkgd,
    }
    if augmentations:
        aug = dict(num_components_density=(4, 4, 4), bounding_box=((-1.5, -1.67, -0.5), (1.5, 1.67, 1.0)),
                   num_voxels_initial=64 ** 3, num_voxels_final=160 ** 3)
        aug.update(augmented or {})
        model['augmentations'] = [{'name': 'points_augmentation', 'coarse_model': _tensor(**aug)}]
    return {'data_loader': {'ndc': ndc}, 'model': model, 'database': 'NeRF_LLFF'}
