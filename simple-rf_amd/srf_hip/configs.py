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
